"""GPU parity of the differentiable VBD rollout (IFT adjoint) against the FP64 oracle.

The device adjoint is compared with oracle/adjoint.py, which assembles the
step Hessian in FP64 and solves it directly (SuperLU), on the same rollout:
the oracle's gradient is evaluated along the positions the device forward
produced (read back after each step), which isolates the adjoint from the
forward's FP32-vs-FP64 drift that the forward parity tests bound separately.

Tolerances:
  Hessian-vector product vs FP64:                 rel L2 <= 1e-5
  adjoint solve (PCG / Gauss-Seidel) vs FP64:     rel L2 <= 1e-4
  rollout gradients (log Y, v0, x0) vs FP64 IFT:  rel L2 <= 1e-4 each
  directional central FD of the FP32 forward:     rel 2e-2 (FD noise of FP32)
"""
import numpy as np
import pytest

from physicsbasedanimationtoolkit_amd import meshes, pbat
from oracle import adjoint as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
GRAD_TOL = 1e-4


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def diff_scene(n=8, seed=0, lo=1e5, hi=1e7):
    X, E = meshes.cube_tet_grid(n, jitter=0.1, seed=seed)
    rng = np.random.default_rng(seed + 1)
    nT = E.shape[1]
    Y = meshes.log_uniform(rng, lo, hi, nT)
    mu, lam = meshes.lame_coefficients(Y, 0.45)
    v0 = rng.normal(size=X.shape) * 0.5
    return X, E, mu, lam, v0


def make_gpu(X, E, mu, lam, v0, graph=True):
    nT = E.shape[1]
    data = (pbat.sim.vbd.Data().with_volume_mesh(X, E).with_material(np.full(nT, 1e3), mu, lam)
            .with_acceleration(meshes.gravity(X.shape[1])).with_dirichlet_vertices(meshes.bottom_vertices(X))
            .with_velocity(v0).construct())
    g = pbat.gpu.vbd.Integrator(data)
    g.graph_capture = graph
    return g


def make_oracle(X, E, mu, lam, v0):
    nT = E.shape[1]
    return O.OracleVbd(X, E, v=v0, aext=meshes.gravity(X.shape[1]), rhoe=np.full(nT, 1e3), mue=mu, lambdae=lam,
                       dbc=meshes.bottom_vertices(X), nthreads=8)


def deformed(X, seed=4, amp=0.02):
    return X + np.random.default_rng(seed).normal(size=X.shape) * amp


def test_hessian_vector_matches_fp64(gpu_device):
    X, E, mu, lam, v0 = diff_scene(8)
    g = make_gpu(X, E, mu, lam, v0)
    o = make_oracle(X, E, mu, lam, v0)
    x = deformed(X)
    p = np.random.default_rng(9).normal(size=X.shape)
    free = A.free_mask(o)
    pf = p * free[None, :]
    hv = g.hessian_vector(x, DT, p)
    ref = A.hessian_vector(o, x.T, DT, pf.T).T * free[None, :]
    e = rel(hv, ref)
    print(f"HVP rel err {e:.3e}")
    assert e <= 1e-5
    assert np.all(hv[:, ~free] == 0)


@pytest.mark.parametrize("solver,iters", [("pcg", 300), ("pcg_jacobi", 600), ("pcg_two_level", 600),
                                          ("gauss_seidel", 20000)])
def test_adjoint_solve_matches_fp64(gpu_device, solver, iters):
    X, E, mu, lam, v0 = diff_scene(6)
    g = make_gpu(X, E, mu, lam, v0)
    o = make_oracle(X, E, mu, lam, v0)
    # Small random deformation: the step Hessian stays positive definite, as it is
    # at the strict minimisers the adjoint differentiates (random moves of 6% strain
    # invert elements and make it indefinite).
    x = deformed(X, amp=0.002)
    b = np.random.default_rng(3).normal(size=X.shape)
    y, st = g.adjoint_solve(x, DT, b, solver=solver, max_iterations=iters, tolerance=1e-6, check_every=100)
    H, idx = A.hessian(o, x.T, DT)
    yr = np.zeros(3 * o.nV)
    yr[idx] = A.spla.spsolve(H, b.T.reshape(-1)[idx])
    yr = yr.reshape(o.nV, 3).T
    e = rel(y, yr)
    print(f"{solver}: {st} rel err {e:.3e}")
    assert st["relative_residual"] <= 1e-5
    assert e <= 1e-4


def test_adjoint_solvers_alternate_on_one_handle(gpu_device):
    """The fused solvers number the Hessian's rows aggregate-major, the coloured
    Gauss-Seidel ones colour-major: alternating them on one integrator rebuilds
    the pattern each time, and every solve still matches the FP64 direct solve."""
    X, E, mu, lam, v0 = diff_scene(6)
    g = make_gpu(X, E, mu, lam, v0)
    o = make_oracle(X, E, mu, lam, v0)
    x = deformed(X, amp=0.002)
    b = np.random.default_rng(3).normal(size=X.shape)
    H, idx = A.hessian(o, x.T, DT)
    yr = np.zeros(3 * o.nV)
    yr[idx] = A.spla.spsolve(H, b.T.reshape(-1)[idx])
    yr = yr.reshape(o.nV, 3).T
    for solver in ("pcg_two_level", "pcg", "pcg_jacobi", "pcg", "pcg_two_level"):
        y, st = g.adjoint_solve(x, DT, b, solver=solver, max_iterations=600, tolerance=1e-6, check_every=10)
        e = rel(y, yr)
        print(f"{solver}: {st} rel err {e:.3e}")
        assert st["relative_residual"] <= 1e-5
        assert e <= 1e-4


def fan_scene(n_surface, seed=0):
    """A ball of tets around one centre vertex: n_surface points on the unit
    sphere (golden-angle spiral), their convex hull's triangles each joined to
    the centre. The centre is adjacent to every surface vertex; the surface
    vertices with y < -0.8 are fixed."""
    from scipy.spatial import ConvexHull

    k = np.arange(n_surface) + 0.5
    y = 1.0 - 2.0 * k / n_surface
    r = np.sqrt(1.0 - y * y)
    phi = k * np.pi * (3.0 - np.sqrt(5.0))
    P = np.stack([r * np.cos(phi), y, r * np.sin(phi)], axis=1)
    tris = ConvexHull(P).simplices
    X = np.concatenate([np.zeros((1, 3)), P]).T.copy()
    E = np.concatenate([np.zeros((len(tris), 1), np.int64), tris + 1], axis=1)
    for t in E:
        a, b, c, d = X[:, t].T
        if np.dot(np.cross(b - a, c - a), d - a) < 0:
            t[2], t[3] = t[3], t[2]
    E = E.T.copy()
    rng = np.random.default_rng(seed)
    nT = E.shape[1]
    mu, lam = meshes.lame_coefficients(meshes.log_uniform(rng, 1e5, 1e7, nT), 0.45)
    dbc = np.flatnonzero(X[1] < -0.8).astype(np.int64)
    return X, E, mu, lam, dbc


@pytest.mark.parametrize("n_surface", [40, 90, 120, 200])
@pytest.mark.parametrize("solver", ["pcg_jacobi", "pcg_two_level"])
def test_adjoint_solve_long_rows_matches_fp64(gpu_device, n_surface, solver):
    """The centre's row holds 37, 82, 109 and 181 blocks (its updated
    neighbours and itself): the row-major assembly with 2, 3 and 4 blocks per
    lane, and the per-block assembly that takes rows beyond 128 blocks."""
    X, E, mu, lam, dbc = fan_scene(n_surface)
    nT = E.shape[1]
    v0 = np.zeros_like(X)
    data = (pbat.sim.vbd.Data().with_volume_mesh(X, E).with_material(np.full(nT, 1e3), mu, lam)
            .with_acceleration(meshes.gravity(X.shape[1])).with_dirichlet_vertices(dbc)
            .with_velocity(v0).construct())
    g = pbat.gpu.vbd.Integrator(data)
    o = O.OracleVbd(X, E, v=v0, aext=meshes.gravity(X.shape[1]), rhoe=np.full(nT, 1e3), mue=mu, lambdae=lam,
                    dbc=dbc, nthreads=8)
    x = deformed(X, amp=0.002)
    b = np.random.default_rng(3).normal(size=X.shape)
    y, st = g.adjoint_solve(x, DT, b, solver=solver, max_iterations=600, tolerance=1e-6, check_every=10)
    H, idx = A.hessian(o, x.T, DT)
    yr = np.zeros(3 * o.nV)
    yr[idx] = A.spla.spsolve(H, b.T.reshape(-1)[idx])
    yr = yr.reshape(o.nV, 3).T
    e = rel(y, yr)
    print(f"fan {n_surface} {solver}: {st} rel err {e:.3e}")
    assert st["relative_residual"] <= 1e-5
    assert e <= 1e-4


def cube_with_isolated_vertices():
    """A jittered 6^3 cube (bottom fixed) plus two small tet groups far from it:
    one free vertex alone (a one-vertex aggregate) and two free vertices 0.06 apart
    (one aggregate of two, i.e. collinear, or two of one). Such aggregates have no
    rotation modes: their coarse rows keep translations only."""
    X, E = meshes.cube_tet_grid(6, jitter=0.1, seed=5)
    dbc = meshes.bottom_vertices(X)
    extra = np.array([[10.0, 0.0, 0.0], [10.5, 0.0, 0.0], [10.0, 0.5, 0.0],   # fixed
                      [10.1, 0.1, 0.4], [10.15, 0.12, 0.42],                 # free pair
                      [20.0, 0.0, 0.0], [20.5, 0.0, 0.0], [20.0, 0.5, 0.0],   # fixed
                      [20.1, 0.1, 0.4]]).T                                   # free, alone
    n0 = X.shape[1]
    tets = np.array([[0, 1, 2, 3], [0, 1, 3, 4], [5, 6, 7, 8]]) + n0
    X = np.concatenate([X, extra], axis=1)
    for t in tets:
        a, b, c, d = X[:, t].T
        if np.dot(np.cross(b - a, c - a), d - a) < 0:
            t[2], t[3] = t[3], t[2]
    E = np.concatenate([E, tets.T], axis=1)
    dbc = np.concatenate([dbc, n0 + np.array([0, 1, 2, 5, 6, 7])]).astype(np.int64)
    nT = E.shape[1]
    mu, lam = meshes.lame_coefficients(meshes.log_uniform(np.random.default_rng(6), 1e5, 1e7, nT), 0.45)
    return X, E, mu, lam, dbc


def test_adjoint_solve_with_rotation_free_aggregates_matches_fp64(gpu_device):
    """The two-level solve on a mesh whose far-away free vertices form aggregates
    without rotation modes (identity rotation diagonal in the coarse operator)
    converges to the FP64 solution."""
    X, E, mu, lam, dbc = cube_with_isolated_vertices()
    nT = E.shape[1]
    v0 = np.zeros_like(X)
    data = (pbat.sim.vbd.Data().with_volume_mesh(X, E).with_material(np.full(nT, 1e3), mu, lam)
            .with_acceleration(meshes.gravity(X.shape[1])).with_dirichlet_vertices(dbc)
            .with_velocity(v0).construct())
    g = pbat.gpu.vbd.Integrator(data)
    o = O.OracleVbd(X, E, v=v0, aext=meshes.gravity(X.shape[1]), rhoe=np.full(nT, 1e3), mue=mu, lambdae=lam,
                    dbc=dbc, nthreads=8)
    x = deformed(X, amp=0.002)
    b = np.random.default_rng(3).normal(size=X.shape)
    y, st = g.adjoint_solve(x, DT, b, solver="pcg_two_level", max_iterations=600, tolerance=1e-6,
                            check_every=10)
    H, idx = A.hessian(o, x.T, DT)
    yr = np.zeros(3 * o.nV)
    yr[idx] = A.spla.spsolve(H, b.T.reshape(-1)[idx])
    yr = yr.reshape(o.nV, 3).T
    e = rel(y, yr)
    print(f"cube + isolated vertices: {st} rel err {e:.3e}")
    assert st["relative_residual"] <= 1e-5
    assert e <= 1e-4


def test_adjoint_solve_with_more_than_512_aggregates(gpu_device):
    """A 48^3 cube: the cubes of 1/8 its side leave a ninth layer of face, edge
    and corner vertices on the upper faces, 7 x 7 x 8 + those aggregates (more
    than 512, so the coarse product takes several passes over its 64-column
    chunks), the thin ones planar, collinear (no rotation modes) or a single
    vertex. The two-level solution satisfies H y = b with H applied by the
    matrix-free product and agrees with the block-Jacobi solution."""
    X, E = meshes.cube_tet_grid(48, jitter=0.1, seed=7)
    nT = E.shape[1]
    mu, lam = meshes.lame_coefficients(meshes.log_uniform(np.random.default_rng(8), 1e5, 1e7, nT), 0.45)
    dbc = meshes.bottom_vertices(X)
    v0 = np.zeros_like(X)
    data = (pbat.sim.vbd.Data().with_volume_mesh(X, E).with_material(np.full(nT, 1e3), mu, lam)
            .with_acceleration(meshes.gravity(X.shape[1])).with_dirichlet_vertices(dbc)
            .with_velocity(v0).construct())
    g = pbat.gpu.vbd.Integrator(data)
    # a perturbation small against the cell (1/48), where H stays positive definite
    x = deformed(X, amp=0.0002)
    b = np.random.default_rng(9).normal(size=X.shape)
    b[:, dbc] = 0.0
    y2, st2 = g.adjoint_solve(x, DT, b, solver="pcg_two_level", max_iterations=2000, tolerance=1e-6,
                              check_every=10)
    yj, stj = g.adjoint_solve(x, DT, b, solver="pcg_jacobi", max_iterations=4000, tolerance=1e-6,
                              check_every=10)
    res = rel(g.hessian_vector(x, DT, y2), b)
    print(f"48^3: two-level {st2}, block Jacobi {stj}, |Hy - b|/|b| {res:.3e}, vs Jacobi {rel(y2, yj):.3e}")
    assert st2["relative_residual"] <= 1e-6 and stj["relative_residual"] <= 1e-6
    assert st2["iterations"] < stj["iterations"]
    assert res <= 1e-4
    assert rel(y2, yj) <= 1e-4


def gpu_trajectory(g, x0, v0, steps, iterations):
    """Positions x_0..x_T of stepping g one step at a time (device forward)."""
    g.x = x0
    g.v = v0
    xs = [np.asarray(x0, np.float64).T.copy()]
    for _ in range(steps):
        g.step(DT, iterations, 1)
        xs.append(g.x.astype(np.float64).T.copy())
    return xs


@pytest.mark.parametrize("solver", ["pcg", "pcg_jacobi", "pcg_two_level", "gauss_seidel"])
def test_rollout_gradient_matches_fp64_ift(gpu_device, solver):
    X, E, mu, lam, v0 = diff_scene(8, seed=2)
    steps, iters = 5, 20
    g = make_gpu(X, E, mu, lam, v0)
    o = make_oracle(X, E, mu, lam, v0)
    v0f = o.v.T.copy()  # Dirichlet velocities zeroed
    xs = gpu_trajectory(g, X, v0f, steps, iters)
    rng = np.random.default_rng(21)
    w = rng.normal(size=X.shape)
    u = rng.normal(size=X.shape)
    ref = A.ift_gradient(o, xs, DT, w.T, u.T)
    g.x = X
    g.v = v0f
    g.rollout_forward(DT, iters, steps, checkpoint_every=2)
    np.testing.assert_array_equal(g.x, xs[-1].T.astype(np.float32))
    gs = solver == "gauss_seidel"
    kw = dict(solver=solver, tolerance=1e-6, max_iterations=20000 if gs else 600, check_every=100 if gs else 10)
    r = g.rollout_backward(w, u, **kw)
    print(solver, r["stats"])
    for key, refk in (("log_young_modulus", ref[0]), ("v0", ref[1].T), ("x0", ref[2].T)):
        e = rel(r[key], refk)
        print(f"{key}: rel err {e:.3e}")
        assert e <= GRAD_TOL, key


@pytest.mark.parametrize("history", [0, 1, 3, 8])
@pytest.mark.parametrize("solver", ["pcg_jacobi", "pcg_two_level"])
def test_projected_initial_guess_matches_fp64_ift(gpu_device, solver, history):
    """Each solve of a backward pass started from the last solution (history 0)
    or from its projection onto the span of the last 1, 3 or 8 solutions: 12
    steps, so the longest history is reached, and the gradients stay within
    tolerance of the FP64 IFT gradient; the projected starts take at most one
    CG iteration per solve more than the plain ones (fewer in practice)."""
    X, E, mu, lam, v0 = diff_scene(8, seed=3)
    steps, iters = 12, 20
    g = make_gpu(X, E, mu, lam, v0)
    o = make_oracle(X, E, mu, lam, v0)
    v0f = o.v.T.copy()
    xs = gpu_trajectory(g, X, v0f, steps, iters)
    rng = np.random.default_rng(22)
    w = rng.normal(size=X.shape)
    ref = A.ift_gradient(o, xs, DT, w.T, np.zeros_like(w.T))
    g.x = X
    g.v = v0f
    g.rollout_forward(DT, iters, steps, checkpoint_every=4)
    kw = dict(solver=solver, tolerance=1e-6, max_iterations=600, check_every=10)
    r = g.rollout_backward(w, None, history=history, **kw)
    r0 = g.rollout_backward(w, None, history=0, **kw)
    print(solver, history, "stats", r["stats"], r0["stats"])
    for key, refk in (("log_young_modulus", ref[0]), ("v0", ref[1].T), ("x0", ref[2].T)):
        e = rel(r[key], refk)
        print(f"{key}: rel err {e:.3e}")
        assert e <= GRAD_TOL, key
    assert r["stats"]["total_iterations"] <= r0["stats"]["total_iterations"] + steps


def test_differentiable_workload_gradient_matches_fp64(gpu_device):
    """The 32^3 differentiable workload's settings on a 12^3 grid (the FP64 direct
    solve of the 32^3 step Hessian takes minutes): hidden per-tet log Y, v0 = 0.5 x
    random unit vectors, 20 steps x 20 iterations, L = |x_T - target|^2 from
    log Y = ln 1e6; gradients wrt log Y and v0 vs the FP64 IFT gradient along the
    device trajectory."""
    s = meshes.differentiable_scene(12, seed=3)
    data = (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_material(s.rhoe, s.mue, s.lambdae)
            .with_acceleration(s.aext).with_dirichlet_vertices(s.dbc).with_velocity(s.v0).construct())
    g = pbat.gpu.vbd.Integrator(data)
    v0 = g.v.copy()
    g.set_log_young_modulus(s.log_y_hidden)
    g.rollout_forward(s.dt, s.iterations, s.steps, 10)
    target = g.x.astype(np.float64)
    g.set_log_young_modulus(s.log_y_guess)
    xs = gpu_trajectory(g, s.X, v0, s.steps, s.iterations)
    dLdx = 2 * (xs[-1].T - target)
    g.x = s.X
    g.v = v0
    g.rollout_forward(s.dt, s.iterations, s.steps, 10)
    r = g.rollout_backward(dLdx, None, solver="pcg_two_level", tolerance=1e-6, max_iterations=3000)
    o = O.OracleVbd(s.X, s.E, v=s.v0, aext=s.aext, rhoe=s.rhoe, mue=s.mue, lambdae=s.lambdae, dbc=s.dbc,
                    nthreads=8)
    A.set_log_young_modulus(o, s.log_y_guess)
    ref = A.ift_gradient(o, xs, s.dt, dLdx.T)
    e_y = rel(r["log_young_modulus"], ref[0])
    e_v = rel(r["v0"], ref[1].T)
    print(f"12^3 workload: {r['stats']} rel err log Y {e_y:.3e} v0 {e_v:.3e}")
    assert e_y <= GRAD_TOL
    assert e_v <= GRAD_TOL


def test_gradient_matches_fd_of_device_forward(gpu_device):
    """Directional central differences of the FP32 device rollout on a soft
    material where 200 iterations converge every step."""
    X, E, mu, lam, v0 = diff_scene(5, seed=3, lo=1e4, hi=1e5)
    steps, iters = 3, 200
    g = make_gpu(X, E, mu, lam, v0)
    v0f = make_oracle(X, E, mu, lam, v0).v.T.copy()
    rng = np.random.default_rng(8)
    w = rng.normal(size=X.shape)
    logy0 = g.log_young_modulus().astype(np.float64)

    def loss(logy=logy0, v=v0f):
        g.set_log_young_modulus(logy)
        g.x = X
        g.v = v
        g.rollout_forward(DT, iters, steps, checkpoint_every=steps)
        return float((w * g.x.astype(np.float64)).sum())

    loss()
    r = g.rollout_backward(w, None, tolerance=1e-6, max_iterations=300)
    d = rng.normal(size=logy0.size)
    eps = 2e-2
    fd = (loss(logy0 + eps * d) - loss(logy0 - eps * d)) / (2 * eps)
    an = float(r["log_young_modulus"].astype(np.float64) @ d)
    print(f"log Y: fd {fd:.6e} adjoint {an:.6e}")
    assert abs(fd - an) <= 2e-2 * abs(fd)
    dv = rng.normal(size=X.shape) * (np.abs(v0f).sum(axis=0) > 0)[None, :]
    fd = (loss(v=v0f + eps * dv) - loss(v=v0f - eps * dv)) / (2 * eps)
    an = float((r["v0"].astype(np.float64) * dv).sum())
    print(f"v0: fd {fd:.6e} adjoint {an:.6e}")
    assert abs(fd - an) <= 2e-2 * abs(fd)


def test_checkpoint_interval_does_not_change_gradients(gpu_device):
    X, E, mu, lam, v0 = diff_scene(6, seed=5)
    g = make_gpu(X, E, mu, lam, v0)
    w = np.random.default_rng(1).normal(size=X.shape)
    out = []
    for every in (1, 3, 7):
        g.x = X
        g.v = v0
        g.rollout_forward(DT, 10, 7, checkpoint_every=every)
        out.append(g.rollout_backward(w, None, tolerance=1e-6))
    # The recomputed segments are bitwise equal to the forward; the solves'
    # FP64 atomic dot products may round differently from run to run.
    for r in out[1:]:
        for k in ("log_young_modulus", "v0", "x0"):
            assert rel(r[k], out[0][k]) <= 1e-5, k


def test_backward_reads_the_rollout_not_the_current_state(gpu_device):
    """The backward pass takes each step's minimiser from the rollout (the last one from
    the forward's own copy of its end positions), so stepping the integrator between the
    forward and the backward pass (which keeps the rollout) changes no gradient, at every
    checkpoint interval."""
    X, E, mu, lam, v0 = diff_scene(6, seed=7)
    g = make_gpu(X, E, mu, lam, v0)
    w = np.random.default_rng(2).normal(size=X.shape)
    for every in (1, 2, 5):
        g.x = X
        g.v = v0
        g.rollout_forward(DT, 10, 5, checkpoint_every=every)
        ref = g.rollout_backward(w, None, tolerance=1e-6)
        g.x = X
        g.v = v0
        g.rollout_forward(DT, 10, 5, checkpoint_every=every)
        g.step(DT, 10, 1)
        moved = g.x
        out = g.rollout_backward(w, None, tolerance=1e-6)
        np.testing.assert_array_equal(g.x, moved)
        for k in ("log_young_modulus", "v0", "x0"):
            assert rel(out[k], ref[k]) <= 1e-5, (every, k)


def test_backward_restores_end_state(gpu_device):
    X, E, mu, lam, v0 = diff_scene(6, seed=6)
    a = make_gpu(X, E, mu, lam, v0)
    b = make_gpu(X, E, mu, lam, v0)
    a.rollout_forward(DT, 10, 4, checkpoint_every=2)
    b.rollout_forward(DT, 10, 4, checkpoint_every=2)
    xa, va = a.x, a.v
    a.rollout_backward(np.ones_like(X), np.ones_like(X))
    np.testing.assert_array_equal(a.x, xa)
    np.testing.assert_array_equal(a.v, va)
    a.step(DT, 10, 1)
    b.step(DT, 10, 1)
    np.testing.assert_array_equal(a.x, b.x)


def test_log_young_modulus_setter_matches_oracle(gpu_device):
    X, E, mu, lam, v0 = diff_scene(6, seed=7)
    g = make_gpu(X, E, mu, lam, v0)
    o = make_oracle(X, E, mu, lam, v0)
    logy = g.log_young_modulus()
    assert rel(logy, np.log(A.young_modulus(o))) <= 1e-6
    new = logy + np.random.default_rng(2).uniform(-1.0, 1.0, size=logy.size)
    g.set_log_young_modulus(new)
    np.testing.assert_array_equal(g.log_young_modulus(), new.astype(np.float32))
    A.set_log_young_modulus(o, new.astype(np.float64))
    x0 = o.x.copy()
    g.step(DT, 20, 1)
    o.step(DT, 20, 1)
    e = rel(g.x.T, o.x)
    ed = rel(g.x.T - x0, o.x - x0)
    print(f"rel pos err {e:.3e} rel disp err {ed:.3e}")
    assert e <= 1e-5
    assert ed <= 2e-3


def test_torch_autograd_function(gpu_device):
    torch = pytest.importorskip("torch")
    from physicsbasedanimationtoolkit_amd.gpu import diff

    X, E, mu, lam, v0 = diff_scene(5, seed=9)
    g = make_gpu(X, E, mu, lam, v0)
    o = make_oracle(X, E, mu, lam, v0)
    v0f = o.v.T.copy()
    logy = torch.tensor(g.log_young_modulus(), dtype=torch.float32, requires_grad=True)
    x0 = torch.tensor(X, dtype=torch.float32, requires_grad=True)
    v = torch.tensor(v0f, dtype=torch.float32, requires_grad=True)
    target = torch.tensor(X, dtype=torch.float32) + 0.05
    xT, vT = diff.rollout(g, logy, x0, v, DT, 10, 4, checkpoint_every=2)
    loss = ((xT - target) ** 2).sum() + 0.1 * (vT ** 2).sum()
    loss.backward()
    # same gradients straight from the C ABI
    g.x = X
    g.v = v0f
    g.rollout_forward(DT, 10, 4, checkpoint_every=2)
    r = g.rollout_backward(2 * (g.x - target.numpy()), 0.2 * g.v)
    np.testing.assert_allclose(logy.grad.numpy(), r["log_young_modulus"], rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(v.grad.numpy(), r["v0"], rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(x0.grad.numpy(), r["x0"], rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("solver", ["pcg", "pcg_jacobi", "pcg_two_level"])
def test_batched_gradients_match_per_mesh(gpu_device, solver):
    """Each mesh of a batch runs its own CG, so the batch reproduces the per-mesh
    loop up to rounding."""
    scenes = meshes.batched_scenes(3, 3, 6, seed=4, per_tet_log_y=True)
    datas, singles, ws = [], [], []
    rng = np.random.default_rng(0)
    for s in scenes:
        d = (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_material(s.rhoe, s.mue, s.lambdae)
             .with_acceleration(s.aext).with_dirichlet_vertices(s.dbc).construct())
        datas.append(d)
        singles.append(pbat.gpu.vbd.Integrator(d))
        ws.append(rng.normal(size=s.X.shape))
    batch = pbat.gpu.vbd.BatchIntegrator(datas)
    batch.rollout_forward(DT, 10, 4, checkpoint_every=2)
    rb = batch.rollout_backward(np.concatenate(ws, axis=1), None, solver=solver, tolerance=1e-6, check_every=1)
    t0 = 0
    for b, (g, w, s) in enumerate(zip(singles, ws, scenes)):
        g.rollout_forward(DT, 10, 4, checkpoint_every=2)
        r = g.rollout_backward(w, None, solver=solver, tolerance=1e-6, check_every=1)
        assert r["stats"]["max_relative_residual"] <= 1e-6
        sl = slice(batch.offsets[b], batch.offsets[b + 1])
        nT = s.E.shape[1]
        assert rel(rb["v0"][:, sl], r["v0"]) <= 1e-5
        assert rel(rb["x0"][:, sl], r["x0"]) <= 1e-5
        assert rel(rb["log_young_modulus"][t0:t0 + nT], r["log_young_modulus"]) <= 1e-5
        t0 += nT


def test_adjoint_errors(gpu_device):
    X, E, mu, lam, v0 = diff_scene(4, seed=1)
    g = make_gpu(X, E, mu, lam, v0)
    with pytest.raises(ValueError):
        g.rollout_backward(np.ones_like(X))
    g.rollout_forward(DT, 5, 2)
    g.kD = 1e-3
    with pytest.raises(ValueError):
        g.rollout_backward(np.ones_like(X))
    g.kD = 0.0
    g.rollout_forward(DT, 5, 2)
    g.a = meshes.gravity(X.shape[1]) * 2.0  # changes the steps the backward pass would recompute
    with pytest.raises(ValueError):
        g.rollout_backward(np.ones_like(X))
    g.rollout_forward(DT, 5, 2)
    g.set_sweep_config(8, 0, True, False)
    with pytest.raises(ValueError):
        g.rollout_backward(np.ones_like(X))
    with pytest.raises(ValueError):
        g.set_log_young_modulus(np.zeros(3))
    with pytest.raises(ValueError):
        g.set_log_young_modulus(np.full(E.shape[1], np.nan))
    g.set_sweep_config(0, 0, True, False)
    g.rollout_forward(DT, 5, 2)
    for history in (-1, 9):
        with pytest.raises(ValueError):
            g.rollout_backward(np.ones_like(X), solver="pcg_two_level", history=history)


def test_differentiable_workload_checkpoint_intervals_agree(gpu_device):
    """The 32^3 differentiable workload at full size (hidden per-tet log Y, 20 steps x 20
    iterations, two-level PCG to 1e-6): gradients with the state kept after every step
    equal those with a checkpoint every 7 steps (ragged last segment, recomputed
    steps) bit for bit: the recomputed steps equal the forward's and the solves take
    every sum in a fixed order. Every solve converges."""
    s = meshes.differentiable_scene(32, seed=0)
    data = (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_material(s.rhoe, s.mue, s.lambdae)
            .with_acceleration(s.aext).with_dirichlet_vertices(s.dbc).with_velocity(s.v0).construct())
    g = pbat.gpu.vbd.Integrator(data)
    v0 = g.v.copy()
    w = np.random.default_rng(5).normal(size=s.X.shape)
    g.set_log_young_modulus(s.log_y_hidden)
    out = []
    for every in (1, 7):
        g.x = s.X
        g.v = v0
        g.rollout_forward(s.dt, s.iterations, s.steps, every)
        r = g.rollout_backward(w, None, solver="pcg_two_level", tolerance=1e-6, max_iterations=2000)
        assert r["stats"]["solves"] == s.steps and r["stats"]["max_relative_residual"] <= 1e-6
        out.append(r)
    for k in ("log_young_modulus", "v0", "x0"):
        assert np.array_equal(out[1][k], out[0][k]), (k, rel(out[1][k], out[0][k]))


def test_full_batched_differentiable_workload_matches_per_mesh(gpu_device):
    """The batched differentiable workload's 64 meshes at full size (grid dims U{8..20}
    per axis, hidden per-tet log Y, h = 1/120 x 10 iterations) over 20 of its 200 steps,
    checkpoint every 10: each mesh of the batch runs its own CG over the same row
    chunks, aggregates and coarse rows as alone, so the batched gradients of a spread
    of meshes equal their own integrators' bit for bit."""
    scenes = meshes.batched_differentiable_scenes(64, 8, 20, seed=0)
    datas = [(pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_material(s.rhoe, s.mue, s.lambdae)
              .with_acceleration(s.aext).with_dirichlet_vertices(s.dbc).with_velocity(s.v0).construct())
             for s in scenes]
    b = pbat.gpu.vbd.BatchIntegrator(datas)
    b.set_sweep_config(16, 2, xcd_remap=True, fixed_stride=False)
    dt, it, T, every = scenes[0].dt, scenes[0].iterations, 20, 10
    solver = dict(solver="pcg_two_level", tolerance=1e-6, max_iterations=2000)
    vb = b.v.copy()
    b.set_log_young_modulus(np.concatenate([s.log_y_hidden for s in scenes]))
    b.rollout_forward(dt, it, T, every)
    w = np.random.default_rng(11).normal(size=vb.shape)
    rb = b.rollout_backward(w, None, **solver)
    assert rb["stats"]["max_relative_residual"] <= 1e-6
    tet0 = np.concatenate([[0], np.cumsum([s.n_tets for s in scenes])])
    for k in (0, 21, 42, 63):
        s, sl = scenes[k], slice(b.offsets[k], b.offsets[k + 1])
        g = pbat.gpu.vbd.Integrator(datas[k])
        g.set_sweep_config(16, 2, xcd_remap=False, fixed_stride=False)
        g.set_log_young_modulus(s.log_y_hidden)
        g.x = s.X
        g.v = vb[:, sl]
        g.rollout_forward(dt, it, T, every)
        r = g.rollout_backward(w[:, sl], None, **solver)
        for key, a, o in (("log_young_modulus", rb["log_young_modulus"][tet0[k]:tet0[k + 1]],
                           r["log_young_modulus"]), ("v0", rb["v0"][:, sl], r["v0"]),
                          ("x0", rb["x0"][:, sl], r["x0"])):
            assert np.array_equal(a, o), (k, key, rel(a, o))


def test_batched_gradients_match_fp64_ift(gpu_device):
    """The batched differentiable workload's settings (h = 1/120, 10 iterations, a
    checkpoint every 10 steps, recomputation in between) on four small jittered meshes
    with hidden per-tet log Y, over 25 steps (two full segments and a partial one):
    each mesh's block of the batched gradients (log Y, v0, x0) against the FP64 IFT
    gradient of oracle/adjoint.py along that mesh's device trajectory."""
    scenes = meshes.batched_differentiable_scenes(4, 4, 7, seed=5, steps=25)
    datas = [(pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_material(s.rhoe, s.mue, s.lambdae)
              .with_acceleration(s.aext).with_dirichlet_vertices(s.dbc).with_velocity(s.v0).construct())
             for s in scenes]
    b = pbat.gpu.vbd.BatchIntegrator(datas)
    dt, it, T, every = scenes[0].dt, scenes[0].iterations, scenes[0].steps, 10
    logy = np.concatenate([s.log_y_hidden for s in scenes])
    b.set_log_young_modulus(logy)
    Xb = np.concatenate([s.X for s in scenes], axis=1)
    vb = b.v.copy()
    # device trajectory, one step at a time
    b.x = Xb
    b.v = vb
    traj = [[np.asarray(s.X, np.float64).T.copy()] for s in scenes]
    for _ in range(T):
        b.step(dt, it, 1)
        for k, xk in enumerate(b.mesh_positions()):
            traj[k].append(xk.astype(np.float64).T.copy())
    b.x = Xb
    b.v = vb
    b.rollout_forward(dt, it, T, every)
    np.testing.assert_array_equal(b.x, np.concatenate([t[-1].T for t in traj], axis=1).astype(np.float32))
    rng = np.random.default_rng(17)
    w = rng.normal(size=Xb.shape)
    r = b.rollout_backward(w, None, solver="pcg_two_level", tolerance=1e-6, max_iterations=2000, check_every=10)
    print(r["stats"])
    t0 = 0
    for k, s in enumerate(scenes):
        o = O.OracleVbd(s.X, s.E, v=s.v0, aext=s.aext, rhoe=s.rhoe, mue=s.mue, lambdae=s.lambdae, dbc=s.dbc,
                        nthreads=8)
        A.set_log_young_modulus(o, s.log_y_hidden)
        sl = slice(b.offsets[k], b.offsets[k + 1])
        ref = A.ift_gradient(o, traj[k], dt, w[:, sl].T)
        nT = s.E.shape[1]
        errs = (rel(r["log_young_modulus"][t0:t0 + nT], ref[0]), rel(r["v0"][:, sl], ref[1].T),
                rel(r["x0"][:, sl], ref[2].T))
        print(f"mesh {k} ({s.X.shape[1]} verts): rel err log Y {errs[0]:.2e} v0 {errs[1]:.2e} x0 {errs[2]:.2e}")
        assert max(errs) <= GRAD_TOL
        t0 += nT


def test_two_level_backward_replays_after_reallocation(gpu_device):
    """Two successive two-level backward passes on one handle, with other device
    allocations in between (another integrator's buffers, a torch tensor), give the
    gradients of a run without graph capture: the captured CG chunk must never read
    a buffer the coarse set-up freed and reallocated."""
    torch = pytest.importorskip("torch")
    X, E, mu, lam, v0 = diff_scene(7, seed=12)
    w = np.random.default_rng(3).normal(size=X.shape)
    kw = dict(solver="pcg_two_level", tolerance=1e-6, max_iterations=1000, check_every=5)
    ref = make_gpu(X, E, mu, lam, v0, graph=False)
    ref.rollout_forward(DT, 10, 6, checkpoint_every=3)
    r0 = ref.rollout_backward(w, None, **kw)
    g = make_gpu(X, E, mu, lam, v0, graph=True)
    g.rollout_forward(DT, 10, 6, checkpoint_every=3)
    keep = []
    for k in range(3):
        r = g.rollout_backward(w, None, **kw)
        for key in ("log_young_modulus", "v0", "x0"):
            np.testing.assert_allclose(r[key], r0[key], rtol=1e-5, atol=1e-7 * np.abs(r0[key]).max(),
                                       err_msg=f"pass {k}: {key}")
        keep.append(make_gpu(X, E, mu, lam, v0))
        keep.append(torch.empty(1 << 20, device="cuda"))


def test_autograd_rejects_a_replaced_rollout(gpu_device):
    """A second rollout on the same integrator before the first graph's backward pass
    replaces the rollout the library holds: that backward pass raises instead of
    returning the second rollout's gradients."""
    torch = pytest.importorskip("torch")
    from physicsbasedanimationtoolkit_amd.gpu import diff

    X, E, mu, lam, v0 = diff_scene(4, seed=2)
    g = make_gpu(X, E, mu, lam, v0)
    logy = torch.tensor(g.log_young_modulus(), dtype=torch.float32, requires_grad=True)
    x0 = torch.tensor(X, dtype=torch.float32)
    v = torch.zeros(X.shape, dtype=torch.float32)
    xa, _ = diff.rollout(g, logy, x0, v, DT, 5, 2)
    gen, held = g.rollout_generation()
    assert held
    xb, _ = diff.rollout(g, logy * 1.01, x0, v, DT, 5, 2)
    assert g.rollout_generation()[0] != gen
    with pytest.raises(RuntimeError, match="rollout changed"):
        (xa ** 2).sum().backward()
    (xb ** 2).sum().backward()  # the rollout the integrator holds still differentiates
    assert torch.isfinite(logy.grad).all()
    g.x = X
    assert not g.rollout_generation()[1]
