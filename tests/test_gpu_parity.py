"""GPU parity of the MI355X VBD integrator against the FP64 oracle.

Tolerances (FP32 device vs FP64 oracle, same mesh, colouring and iterations):
  positions after one step: ||x_gpu - x_ref|| / ||x_ref|| <= 1e-5
  displacements after one step: ||dx_gpu - dx_ref|| / ||dx_ref|| <= 2e-3
  total-energy trajectory over 100 steps: max |E_gpu - E_ref| / max |E_ref| <= 1e-2
"""
import os

import numpy as np
import pytest

from physicsbasedanimationtoolkit_amd import meshes, pbat
from oracle import oracle as O

pytestmark = pytest.mark.gpu

POS_TOL = 1e-5
DISP_TOL = 2e-3
NTHREADS = min(16, os.cpu_count() or 1)


def build_pair(s, *, v=None, strategy=pbat.sim.vbd.InitializationStrategy.AdaptivePbat, kD=0.0,
               gpu_velocity_semantics=False, chebyshev_rho=None):
    data = (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_material(s.rhoe, s.mue, s.lambdae)
            .with_acceleration(s.aext).with_dirichlet_vertices(s.dbc)
            .with_initialization_strategy(strategy).with_rayleigh_damping(kD))
    if v is not None:
        data = data.with_velocity(v)
    if chebyshev_rho is not None:
        data = data.with_chebyshev_acceleration(chebyshev_rho)
    data = data.construct()
    gpu = pbat.gpu.vbd.Integrator(data)
    ref = O.OracleVbd(s.X, s.E, v=v, aext=s.aext, rhoe=s.rhoe, mue=s.mue, lambdae=s.lambdae, dbc=s.dbc,
                      strategy=int(strategy), kD=kD, nthreads=NTHREADS,
                      gpu_velocity_semantics=gpu_velocity_semantics, chebyshev_rho=chebyshev_rho)
    return data, gpu, ref


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def check_step(gpu, ref, x0, dt, iterations, substeps=1, disp_tol=DISP_TOL):
    x0 = np.array(x0, dtype=np.float64, copy=True)
    gpu.step(dt, iterations, substeps)
    ref.step(dt, iterations, substeps)
    xg = gpu.x.astype(np.float64)
    xr = ref.positions.copy()
    e_pos = rel(xg, xr)
    e_disp = rel(xg - x0, xr - x0)
    print(f"rel pos err {e_pos:.3e}  rel disp err {e_disp:.3e}")
    assert e_pos <= POS_TOL
    assert e_disp <= disp_tol
    return xg, xr


def test_reference_cube_falls(gpu_device):
    """gpu/impl/vbd/Integrator.cu:384-432: 8-vertex cube, dt=1e-2, 10 iterations, no Dirichlet."""
    P = np.array([[0., 1., 0., 1., 0., 1., 0., 1.], [0., 0., 1., 1., 0., 0., 1., 1.],
                  [0., 0., 0., 0., 1., 1., 1., 1.]])
    T = np.array([[0, 3, 5, 6, 0], [1, 2, 4, 7, 5], [3, 0, 6, 5, 3], [5, 6, 0, 3, 6]])
    data = pbat.sim.vbd.Data().with_volume_mesh(P, T).with_bodies(np.ones(8, dtype=np.int64)).construct()
    vbd = pbat.gpu.vbd.Integrator(data)
    vbd.scene_bounding_box = (P.min(axis=1), P.max(axis=1))
    vbd.step(1e-2, 10, 1)
    dx = vbd.x - P.astype(np.float32)
    assert (dx[2] < 0).all()
    assert (np.abs(dx[:2]) < 1e-4).all()


def test_one_step_plain_10(gpu_device):
    s = meshes.plain_cpu_scene(10)
    _, gpu, ref = build_pair(s)
    check_step(gpu, ref, s.X, s.dt, s.iterations)


def test_one_step_jittered_64(gpu_device):
    s = meshes.plain_gpu_scene(64, seed=0)
    _, gpu, ref = build_pair(s)
    # From rest with stiff tets (Y up to 1e7) 20 iterations move vertices by ~1e-5,
    # i.e. ~100 FP32 ulps of their coordinates: the displacement check can only
    # hold to a few percent, the position check is the 1e-5 bar.
    check_step(gpu, ref, s.X, s.dt, s.iterations, disp_tol=5e-2)


def test_one_step_jittered_64_moving(gpu_device):
    """Same mesh with an initial velocity field, so displacements are O(h |v|)."""
    s = meshes.plain_gpu_scene(64, seed=0)
    rng = np.random.default_rng(21)
    v0 = rng.normal(size=(3, s.n_vertices)) * 0.5
    _, gpu, ref = build_pair(s, v=v0, strategy=pbat.sim.vbd.InitializationStrategy.KineticEnergyMinimum)
    check_step(gpu, ref, s.X, s.dt, s.iterations)


SHAPES = [(1, 0), (2, 0), (4, 0), (8, 0), (16, 0), (32, 0), (64, 0),
          # prefetch depth below ceil(degree / lanes): the remaining incidences take the tail loop
          (8, 1), (4, 2), (2, 4), (1, 4)]


@pytest.mark.parametrize("fixed_stride", [False, True])
@pytest.mark.parametrize("lanes,per_lane", SHAPES)
def test_sweep_shapes(gpu_device, lanes, per_lane, fixed_stride):
    s = meshes.plain_gpu_scene(12, seed=3)
    _, gpu, ref = build_pair(s)
    gpu.set_sweep_config(lanes, per_lane, xcd_remap=True, fixed_stride=fixed_stride)
    cfg = gpu.sweep_config()
    assert cfg["lanes_per_vertex"] == lanes and cfg["fixed_stride"] == fixed_stride
    check_step(gpu, ref, s.X, s.dt, s.iterations)


def test_nontemporal_and_no_xcd(gpu_device):
    s = meshes.plain_gpu_scene(20, seed=8)
    _, gpu, ref = build_pair(s)
    gpu.set_sweep_config(32, 0, xcd_remap=False, nontemporal=True)
    check_step(gpu, ref, s.X, s.dt, s.iterations)


def test_autotune_keeps_state(gpu_device):
    s = meshes.plain_gpu_scene(16, seed=6)
    _, gpu, ref = build_pair(s)
    rows = gpu.autotune(s.dt, sweeps=2)
    assert len(rows) >= 4 and all(r[-1] > 0 for r in rows)
    assert np.array_equal(gpu.x, s.X.astype(np.float32))
    check_step(gpu, ref, s.X, s.dt, s.iterations)


@pytest.mark.parametrize("strategy", list(pbat.sim.vbd.InitializationStrategy))
def test_initialization_strategies(gpu_device, strategy):
    s = meshes.plain_gpu_scene(10, seed=5)
    rng = np.random.default_rng(11)
    v0 = rng.normal(size=(3, s.n_vertices)) * 0.3
    _, gpu, ref = build_pair(s, v=v0, strategy=strategy, gpu_velocity_semantics=True)
    x0 = s.X
    for _ in range(3):
        xg, xr = check_step(gpu, ref, x0, s.dt, 10)
        x0 = xr


def test_damping_and_substeps(gpu_device):
    s = meshes.plain_gpu_scene(10, seed=7)
    _, gpu, ref = build_pair(s, kD=1e-3)
    gpu.kD = 1e-3
    check_step(gpu, ref, s.X, s.dt, 8, substeps=3)


def test_dirichlet_vertices_never_move(gpu_device):
    s = meshes.plain_gpu_scene(10, seed=2)
    _, gpu, _ = build_pair(s)
    for _ in range(5):
        gpu.step(s.dt, 10, 1)
    x = gpu.x
    assert np.array_equal(x[:, s.dbc], s.X[:, s.dbc].astype(np.float32))
    assert np.array_equal(gpu.v[:, s.dbc], np.zeros((3, s.dbc.size), np.float32))


def test_graph_capture_matches_eager(gpu_device):
    s = meshes.plain_gpu_scene(10, seed=4)
    _, g1, _ = build_pair(s)
    _, g2, _ = build_pair(s)
    g2.graph_capture = False
    for _ in range(4):
        g1.step(s.dt, 10, 1)
        g2.step(s.dt, 10, 1)
        assert np.array_equal(g1.x, g2.x)


def test_energy_trajectory_100_steps(gpu_device):
    s = meshes.plain_cpu_scene(10)
    _, gpu, ref = build_pair(s)
    eg, er = [], []
    for _ in range(s.steps):
        gpu.step(s.dt, s.iterations, 1)
        ref.step(s.dt, s.iterations, 1)
        eg.append(gpu.energy().sum())
        er.append(ref.energy().sum())
    eg, er = np.array(eg), np.array(er)
    err = np.abs(eg - er).max() / np.abs(er).max()
    print(f"energy trajectory max rel err {err:.3e}")
    assert err <= 1e-2


def test_energy_reduction_matches_oracle(gpu_device):
    s = meshes.plain_gpu_scene(16, seed=9)
    _, gpu, ref = build_pair(s)
    gpu.step(s.dt, 5, 1)
    ref.x[:] = gpu.x.T.astype(np.float64)
    ref.v[:] = gpu.v.T.astype(np.float64)
    e_gpu, e_ref = gpu.energy(), ref.energy()
    assert np.allclose(e_gpu, e_ref, rtol=1e-5, atol=1e-9 * np.abs(e_ref).max())


def test_set_get_state_round_trip(gpu_device):
    s = meshes.plain_gpu_scene(6, seed=1)
    _, gpu, _ = build_pair(s)
    rng = np.random.default_rng(0)
    x = rng.normal(size=(3, s.n_vertices)).astype(np.float32)
    v = rng.normal(size=(3, s.n_vertices)).astype(np.float32)
    gpu.x = x
    gpu.v = v
    assert np.array_equal(gpu.x, x)
    assert np.array_equal(gpu.v, v)
    with pytest.raises(ValueError):
        gpu.x = np.zeros((3, s.n_vertices + 1))
    with pytest.raises(ValueError):
        gpu.v = np.zeros((2, s.n_vertices))
    with pytest.raises(AttributeError):
        _ = gpu.kD


# ---- Chebyshev-accelerated VBD (gpu/impl/vbd/ChebyshevIntegrator.cu) ---------

def test_chebyshev_reference_cube_falls(gpu_device):
    """sim/vbd/ChebyshevIntegrator.cpp:37-85 on the device: rho 0.9, dt 1e-2, 10 iterations."""
    P = np.array([[0., 1., 0., 1., 0., 1., 0., 1.], [0., 0., 1., 1., 0., 0., 1., 1.],
                  [0., 0., 0., 0., 1., 1., 1., 1.]])
    T = np.array([[0, 3, 5, 6, 0], [1, 2, 4, 7, 5], [3, 0, 6, 5, 3], [5, 6, 0, 3, 6]])
    data = pbat.sim.vbd.Data().with_volume_mesh(P, T).with_chebyshev_acceleration(0.9).construct()
    vbd = pbat.gpu.vbd.Integrator(data)
    vbd.step(1e-2, 10, 1)
    dx = vbd.x - P.astype(np.float32)
    assert (dx[2] < 0).all()
    assert (np.abs(dx[:2]) < 1e-4).all()


@pytest.mark.parametrize("rho", [0.5, 0.9])
def test_chebyshev_one_step_plain_10(gpu_device, rho):
    s = meshes.plain_cpu_scene(10)
    _, gpu, ref = build_pair(s, chebyshev_rho=rho)
    check_step(gpu, ref, s.X, s.dt, s.iterations)


def test_chebyshev_one_step_jittered_64_moving(gpu_device):
    s = meshes.plain_gpu_scene(64, seed=0)
    v0 = np.random.default_rng(21).normal(size=(3, s.n_vertices)) * 0.5
    _, gpu, ref = build_pair(s, v=v0, strategy=pbat.sim.vbd.InitializationStrategy.KineticEnergyMinimum,
                             chebyshev_rho=0.8)
    check_step(gpu, ref, s.X, s.dt, s.iterations)


def test_chebyshev_trajectory_and_graph_replay(gpu_device):
    """10 steps of Chebyshev VBD with substeps: graph replay bitwise equal to eager
    launches, positions within 1e-5 of the oracle throughout."""
    s = meshes.plain_cpu_scene(10)
    _, g1, ref = build_pair(s, chebyshev_rho=0.7)
    _, g2, _ = build_pair(s, chebyshev_rho=0.7)
    g2.graph_capture = False
    for _ in range(10):
        g1.step(s.dt, 8, 2)
        g2.step(s.dt, 8, 2)
        ref.step(s.dt, 8, 2)
        assert np.array_equal(g1.x, g2.x)
        assert rel(g1.x, ref.positions) <= POS_TOL


def test_unavailable_accelerators_raise(gpu_device):
    s = meshes.plain_cpu_scene(3)
    data = (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_nesterov_acceleration(1.0).construct())
    with pytest.raises(ValueError):
        pbat.gpu.vbd.Integrator(data)
    data = (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_broyden_acceleration(3).construct())
    with pytest.raises(ValueError):
        pbat.gpu.vbd.Integrator(data)
    data = (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_trust_region_acceleration(0.2, 2.0).construct())
    with pytest.raises(ValueError):
        pbat.gpu.vbd.BatchIntegrator([data, data])
    data = (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_anderson_acceleration(17).construct())
    with pytest.raises(ValueError):
        pbat.gpu.vbd.Integrator(data)
    data = (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_anderson_acceleration(3).construct())
    with pytest.raises(ValueError):
        pbat.gpu.vbd.BatchIntegrator([data, data])


# ---- Anderson-accelerated VBD (gpu/impl/vbd/AndersonIntegrator.cu) -----------
# Anderson's least-squares coefficients are computed from differences of
# successive iterates, which FP32 and FP64 trajectories resolve differently
# (the iterates differ by ~1e-7 while the differences are ~1e-5), and the
# extrapolation x = G - DG alpha carries that into the positions: one step
# lands ~0.5% of the displacement (~1e-5 of the positions) from the FP64
# oracle. The reference's GPU path computes in FP32 as well.
ANDERSON_POS_TOL = 5e-5

def build_anderson_pair(s, window, v=None):
    data = (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_material(s.rhoe, s.mue, s.lambdae)
            .with_acceleration(s.aext).with_dirichlet_vertices(s.dbc).with_anderson_acceleration(window))
    if v is not None:
        data = data.with_velocity(v)
    data = data.construct()
    gpu = pbat.gpu.vbd.Integrator(data)
    ref = O.OracleVbd(s.X, s.E, v=v, aext=s.aext, rhoe=s.rhoe, mue=s.mue, lambdae=s.lambdae, dbc=s.dbc)
    return gpu, ref


def test_anderson_reference_cube(gpu_device):
    """gpu/impl/vbd/AndersonIntegrator.cu test: m = 5, dt 1e-2, 10 iterations on the
    8-vertex cube: it falls along z only and the objective gradient vanishes."""
    P = np.array([[0., 1., 0., 1., 0., 1., 0., 1.], [0., 0., 1., 1., 0., 0., 1., 1.],
                  [0., 0., 0., 0., 1., 1., 1., 1.]])
    T = np.array([[0, 3, 5, 6, 0], [1, 2, 4, 7, 5], [3, 0, 6, 5, 3], [5, 6, 0, 3, 6]])
    data = pbat.sim.vbd.Data().with_volume_mesh(P, T).with_anderson_acceleration(5).construct()
    vbd = pbat.gpu.vbd.Integrator(data)
    vbd.step(1e-2, 10, 1)
    dx = vbd.x - P.astype(np.float32)
    assert (dx[2] < 0).all()
    assert (np.abs(dx[:2]) < 1e-4).all()
    ref = O.OracleVbd(P, T)
    xtilde = (ref.x + 1e-2 * ref.v + 1e-4 * ref.aext).T
    assert np.linalg.norm(ref.objective_gradient(vbd.x.astype(np.float64), xtilde, 1e-2)) < 1e-4


@pytest.mark.parametrize("window", [1, 3, 5])
def test_anderson_one_step_plain_10(gpu_device, window):
    s = meshes.plain_cpu_scene(10)
    gpu, ref = build_anderson_pair(s, window)
    gpu.step(s.dt, s.iterations, 1)
    ref.step_anderson(s.dt, s.iterations, 1, window=window)
    e = rel(gpu.x, ref.positions)
    ed = rel(gpu.x - s.X, ref.positions - s.X)
    print(f"window {window}: rel pos err {e:.3e} rel disp err {ed:.3e}")
    assert e <= ANDERSON_POS_TOL
    assert ed <= 2e-2


def test_anderson_moving_and_graph_replay(gpu_device):
    """A jittered 16^3 mesh with an initial velocity field, 3 steps with 2 substeps:
    graph replay bitwise equal to eager launches; each step within ANDERSON_POS_TOL of
    the oracle stepping from the device's state."""
    s = meshes.plain_gpu_scene(16, seed=3)
    v0 = np.random.default_rng(5).normal(size=(3, s.n_vertices)) * 0.3
    g1, ref = build_anderson_pair(s, 5, v=v0)
    g2, _ = build_anderson_pair(s, 5, v=v0)
    g2.graph_capture = False
    for _ in range(3):
        ref.x[:] = g1.x.T
        ref.v[:] = g1.v.T
        g1.step(s.dt, 12, 2)
        g2.step(s.dt, 12, 2)
        ref.step_anderson(s.dt, 12, 2, window=5)
        assert np.array_equal(g1.x, g2.x)
        e = rel(g1.x, ref.positions)
        print(f"rel pos err {e:.3e}")
        assert e <= ANDERSON_POS_TOL


# ---- trust-region VBD (gpu/impl/vbd/TrustRegionIntegrator.cu) ------------------
# The linear path extrapolates along the last VBD step by up to ~15x, and its
# accept/reject test compares objective differences that FP32 positions resolve
# to ~1e-7 of f; one step stays within TR_POS_TOL of the FP64 oracle stepping
# from the same state when both take the same decisions.
TR_POS_TOL = 5e-5


def build_tr_pair(s, eta=0.2, tau=2.0, curved=False, v=None):
    data = (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_material(s.rhoe, s.mue, s.lambdae)
            .with_acceleration(s.aext).with_dirichlet_vertices(s.dbc)
            .with_trust_region_acceleration(eta, tau, curved))
    if v is not None:
        data = data.with_velocity(v)
    data = data.construct()
    gpu = pbat.gpu.vbd.Integrator(data)
    ref = O.OracleVbd(s.X, s.E, v=v, aext=s.aext, rhoe=s.rhoe, mue=s.mue, lambdae=s.lambdae, dbc=s.dbc,
                      nthreads=NTHREADS)
    return gpu, ref


@pytest.mark.parametrize("curved", [False, True])
def test_trust_region_reference_cube(gpu_device, curved):
    """The free-fall case of the (commented-out) test at
    gpu/impl/vbd/TrustRegionIntegrator.cu:787-851: eta 0.1, tau 2, Y 1e7, nu 0.45,
    dt 1e-2, 10 iterations on the 8-vertex cube: it falls along z only."""
    P = np.array([[0., 1., 0., 1., 0., 1., 0., 1.], [0., 0., 1., 1., 0., 0., 1., 1.],
                  [0., 0., 0., 0., 1., 1., 1., 1.]])
    T = np.array([[0, 3, 5, 6, 0], [1, 2, 4, 7, 5], [3, 0, 6, 5, 3], [5, 6, 0, 3, 6]])
    mu, lam = meshes.lame_coefficients(np.full(5, 1e7), 0.45)
    data = (pbat.sim.vbd.Data().with_volume_mesh(P, T).with_material(np.full(5, 1e3), mu, lam)
            .with_trust_region_acceleration(0.1, 2.0, curved).construct())
    vbd = pbat.gpu.vbd.Integrator(data)
    vbd.step(1e-2, 10, 1)
    dx = vbd.x - P.astype(np.float32)
    assert (dx[2] < 0).all()
    assert (np.abs(dx[:2]) < 1e-4).all()


def test_trust_region_curved_path_is_plain_vbd_bitwise(gpu_device):
    """The curved path never takes an accelerated step (its constraint solve returns
    0, TrustRegionIntegrator.cu:700-703): bitwise the plain integrator, graph or not."""
    s = meshes.plain_gpu_scene(16, seed=3)
    v0 = np.random.default_rng(5).normal(size=(3, s.n_vertices)) * 0.3
    g1, _ = build_tr_pair(s, curved=True, v=v0)
    _, g2, _ = build_pair(s, v=v0)
    g3, _ = build_tr_pair(s, curved=True, v=v0)
    g3.graph_capture = False
    for _ in range(3):
        g1.step(s.dt, 12, 2)
        g2.step(s.dt, 12, 2)
        g3.step(s.dt, 12, 2)
        assert np.array_equal(g1.x, g2.x) and np.array_equal(g1.v, g2.v)
        assert np.array_equal(g3.x, g2.x)


@pytest.mark.parametrize("iterations", [10, 20])
def test_trust_region_linear_one_step_plain_10(gpu_device, iterations):
    """The 10^3 workload: extrapolated steps up to t ~ 18 are tried, some kept, some
    rolled back; the device keeps and rejects the same steps as the oracle, and the
    positions are within TR_POS_TOL of the FP64 oracle."""
    s = meshes.plain_cpu_scene(10)
    gpu, ref = build_tr_pair(s)
    log = []
    gpu.step(s.dt, iterations, 1)
    ref.step_trust_region(s.dt, iterations, 1, eta=0.2, tau=2.0, curved=False, log=log)
    assert any(acc for _, _, acc, _ in log) and not all(acc for _, _, acc, _ in log)
    dev = gpu.trust_region_log()
    for (td, ad), (k, t, acc, _) in zip(dev, log):
        print(f"  k={k} device t={td:.5f} kept={ad}  oracle t={t:.5f} kept={acc}")
    # every verdict agrees; the model's t agrees to 1e-3 on the first tried steps
    # (later models fit objective differences near FP32 resolution and move more)
    assert [a for _, a in dev] == [acc for _, _, acc, _ in log]
    tried = [(td, t) for (td, _), (_, t, _, _) in zip(dev, log) if t != 1.0][:3]
    assert len(tried) == 3 and all(abs(td - t) <= 1e-3 * t for td, t in tried)
    e = rel(gpu.x, ref.positions)
    ed = rel(gpu.x - s.X, ref.positions - s.X)
    print(f"{iterations} iterations: rel pos err {e:.3e} rel disp err {ed:.3e}")
    assert e <= TR_POS_TOL


def test_trust_region_rejecting_every_step_is_plain_vbd(gpu_device):
    """eta = 1e30: every tried step is rolled back to the VBD iterate; the positions
    are plain VBD's up to the FP32 rounding of x_k + ((x_k + t dx) - x_k) / t."""
    s = meshes.plain_cpu_scene(10)
    gpu, _ = build_tr_pair(s, eta=1e30)
    _, plain, _ = build_pair(s)
    gpu.step(s.dt, 20, 1)
    plain.step(s.dt, 20, 1)
    assert rel(gpu.x, plain.x.astype(np.float64)) <= 1e-6


def test_trust_region_moving_16(gpu_device):
    """A jittered 16^3 mesh with an initial velocity field, 3 steps of 2 substeps x 12
    iterations, each from the device's state. The device's decisions (t, kept) replayed
    by the oracle give positions within POS_TOL. The oracle's own decisions at the
    same states: the first tried step of each substep has the device's t to 1% and
    the same verdict; later ones rest on objective differences of ~1e-5 of f and on
    second differences below that, which the FP32 trajectory (displacements ~1e-3
    off the FP64 ones) moves, so only 75% of all verdicts must agree. The captured
    step (decisions on the device) replays the eager launches bit for bit."""
    s = meshes.plain_gpu_scene(16, seed=3)
    v0 = np.random.default_rng(5).normal(size=(3, s.n_vertices)) * 0.3
    g1, ref = build_tr_pair(s, v=v0)
    g2, _ = build_tr_pair(s, v=v0)
    g2.graph_capture = False
    for _ in range(3):
        ref.x[:] = g1.x.T
        ref.v[:] = g1.v.T
        g1.step(s.dt, 12, 2)
        g2.step(s.dt, 12, 2)
        assert np.array_equal(g1.x, g2.x)
        dev = g1.trust_region_log()
        assert len(dev) == 2 * (12 - 2)
        own = []
        ref.step_trust_region(s.dt, 12, 2, eta=0.2, tau=2.0, curved=False, log=own, forced=dev)
        e = rel(g1.x, ref.positions)
        print(f"rel pos err {e:.3e} (device decisions replayed)")
        for (t_dev, acc_dev), (k, t_own, acc_own, ratio) in zip(dev, own):
            print(f"  k={k} device t={t_dev:.4f} kept={acc_dev}  oracle t={t_own:.4f} kept={acc_own} ratio={ratio}")
        assert e <= POS_TOL
        pairs = list(zip(dev, own))
        for half in (pairs[:10], pairs[10:]):
            first = next(((d, o) for d, o in half if d[0] != 1.0 or o[1] != 1.0), None)
            if first is not None:
                (t_dev, acc_dev), (_, t_own, acc_own, _) = first
                assert abs(t_dev - t_own) <= 1e-2 * t_own and acc_dev == acc_own
        assert sum(d[1] == o[2] for d, o in pairs) >= 0.75 * len(pairs)


def test_trust_region_invalid_parameters_raise(gpu_device):
    """sim/vbd/Data.cpp:294-302: eta >= 0 and tau > 1."""
    s = meshes.plain_cpu_scene(3)
    for eta, tau in ((-0.1, 2.0), (0.2, 1.0)):
        data = (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_trust_region_acceleration(eta, tau)
                .construct(validate=False))
        with pytest.raises(ValueError):
            pbat.gpu.vbd.Integrator(data)


def test_traced_step_runs_plain_iterations_when_accelerated(gpu_device, tmp_path):
    """TracedSolve (gpu/impl/vbd/Integrator.cu:284-301) is not overridden by the
    accelerated integrators: a traced Chebyshev step equals a plain step bitwise."""
    s = meshes.plain_gpu_scene(6, seed=2)
    _, g1, _ = build_pair(s, chebyshev_rho=0.9)
    _, g2, _ = build_pair(s)
    g1.traced_step(s.dt, 5, 1, t=0, dir=str(tmp_path))
    g2.step(s.dt, 5, 1)
    np.testing.assert_array_equal(g1.x, g2.x)


def test_traced_step_writes_reference_files(gpu_device, tmp_path):
    """gpu/impl/vbd/Integrator.cu:105-148: MatrixMarket dumps of one traced step;
    the last iterate equals the untraced step's result."""
    import scipy.io

    s = meshes.plain_gpu_scene(4, seed=2)
    data, g1, _ = build_pair(s)
    _, g2, _ = build_pair(s)
    g1.traced_step(s.dt, 3, 2, t=0, dir=str(tmp_path))
    g2.graph_capture = False
    g2.step(s.dt, 3, 2)
    read = lambda name: np.asarray(scipy.io.mmread(str(tmp_path / name)))
    for sub in range(2):
        assert (tmp_path / f"xtilde.t.0.s.{sub}.mtx").exists()
        for k in range(4):
            assert read(f"x.t.0.s.{sub}.k.{k}.mtx").shape == (s.n_vertices, 3)
    np.testing.assert_array_equal(read("x.t.0.s.1.k.3.mtx").astype(np.float32), g2.x.T)
    np.testing.assert_array_equal(g1.x, g2.x)
    np.testing.assert_array_equal(read("T.mtx").astype(np.int64), data.E.T)
    np.testing.assert_allclose(read("M.mtx").reshape(-1), data.m, rtol=1e-6)
    np.testing.assert_allclose(read("lame.mtx"), data.lame.T, rtol=1e-6)
    np.testing.assert_allclose(read("wg.mtx").reshape(-1), data.wg, rtol=1e-6)


def test_one_step_jittered_128(gpu_device):
    """A 128^3 mesh (2,146,689 vertices, 12,582,912 tets, 50 M incidences): one step
    within 1e-5 of the oracle, Dirichlet vertices unmoved, energies finite."""
    s = meshes.plain_gpu_scene(128, seed=1)
    v0 = np.random.default_rng(8).normal(size=(3, s.n_vertices)) * 0.5
    _, gpu, ref = build_pair(s, v=v0, strategy=pbat.sim.vbd.InitializationStrategy.KineticEnergyMinimum)
    check_step(gpu, ref, s.X, s.dt, s.iterations)
    x = gpu.x
    assert np.array_equal(x[:, s.dbc], s.X[:, s.dbc].astype(np.float32))
    assert np.all(np.isfinite(gpu.energy()))
