"""``pbat.sim.vbd.Integrator`` (the reference's CPU integrator, bindings/pypbat/sim/vbd/
Integrator.cpp:30-90) on the HIP path, against the FP64 oracle with the CPU
integrator's semantics (vt = v before the initial guess, sim/vbd/Integrator.cpp:31-32),
and its FP64 objective / gradient reductions and trace files
(sim/vbd/Integrator.cpp:138-235).

Tolerances: positions 1e-5 relative after each step (FP32 device vs FP64 oracle);
objective 1e-6 and gradient 1e-5 relative at the same FP32 state (the device keeps
shape-function gradients and w mu, w lambda in FP32)."""
import os

import numpy as np
import pytest
import scipy.io

from oracle import oracle as O
from physicsbasedanimationtoolkit_amd import meshes, pbat

pytestmark = pytest.mark.gpu

CUBE_P = np.array([[0., 1., 0., 1., 0., 1., 0., 1.], [0., 0., 1., 1., 0., 0., 1., 1.],
                   [0., 0., 0., 0., 1., 1., 1., 1.]])
CUBE_T = np.array([[0, 3, 5, 6, 0], [1, 2, 4, 7, 5], [3, 0, 6, 5, 3], [5, 6, 0, 3, 6]])
CUBE_F = np.array([[0, 1, 1, 3, 3, 2, 2, 0, 0, 0, 4, 5], [1, 5, 3, 7, 2, 6, 0, 4, 3, 2, 5, 7],
                   [4, 4, 5, 5, 7, 7, 6, 6, 1, 3, 6, 6]])
DT = 1.0 / 60.0


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def scene(n=8, seed=0, strategy=pbat.sim.vbd.AdaptiveVbd):
    s = meshes.plain_cpu_scene(n)
    v = 0.5 * np.random.default_rng(seed).normal(size=s.X.shape)
    data = (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_material(s.rhoe, s.mue, s.lambdae)
            .with_acceleration(s.aext).with_dirichlet_vertices(s.dbc).with_velocity(v)
            .with_initialization_strategy(strategy).construct())
    return s, v, data


@pytest.mark.parametrize("strategy", [pbat.sim.vbd.AdaptiveVbd, pbat.sim.vbd.AdaptivePbat,
                                      pbat.sim.vbd.KineticEnergyMinimum])
def test_sim_integrator_matches_cpu_semantics_oracle(gpu_device, strategy):
    s, v, data = scene(8, 1, strategy)
    sim = pbat.sim.vbd.Integrator(data)
    assert sim.strategy == strategy
    o = O.OracleVbd(s.X, s.E, v=v, aext=s.aext, rhoe=s.rhoe, mue=s.mue, lambdae=s.lambdae, dbc=s.dbc,
                    strategy=int(strategy), gpu_velocity_semantics=False)
    for _ in range(4):
        sim.step(DT, 10, 2)
        o.step(DT, 10, 2)
        assert rel(sim.x, o.positions) <= 1e-5
        assert rel(sim.v, o.velocities) <= 1e-4
    d = sim.data
    np.testing.assert_array_equal(d.x, sim.x)


def test_adaptive_vbd_semantics_differ_between_cpu_and_gpu_paths(gpu_device):
    """The same data under both names: the GPU integrator keeps the previous step's
    velocity for AdaptiveVbd (gpu/impl/vbd/Integrator.cu:333), the CPU one does not;
    each matches the oracle run with its semantics."""
    s, v, data = scene(6, 2, pbat.sim.vbd.AdaptiveVbd)
    sim = pbat.sim.vbd.Integrator(data)
    gpu = pbat.gpu.vbd.Integrator(data)
    kw = dict(v=v, aext=s.aext, rhoe=s.rhoe, mue=s.mue, lambdae=s.lambdae, dbc=s.dbc,
              strategy=O.INIT_ADAPTIVE_VBD)
    o_cpu = O.OracleVbd(s.X, s.E, gpu_velocity_semantics=False, **kw)
    o_gpu = O.OracleVbd(s.X, s.E, gpu_velocity_semantics=True, **kw)
    for _ in range(3):
        for i in (sim, gpu):
            i.step(DT, 10, 1)
        for i in (o_cpu, o_gpu):
            i.step(DT, 10, 1)
    assert rel(sim.x, o_cpu.positions) <= 1e-5
    assert rel(gpu.x, o_gpu.positions) <= 1e-5
    assert rel(o_cpu.positions, o_gpu.positions) > 1e-4


def test_sim_integrator_accelerators(gpu_device):
    """The CPU binding's factory: Chebyshev and Anderson integrators, trust region ->
    plain VBD; Nesterov and Broyden are not built."""
    s, v, _ = scene(6, 3, pbat.sim.vbd.KineticEnergyMinimum)

    def data():
        return (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_material(s.rhoe, s.mue, s.lambdae)
                .with_acceleration(s.aext).with_dirichlet_vertices(s.dbc).with_velocity(v)
                .with_initialization_strategy(pbat.sim.vbd.KineticEnergyMinimum))

    kw = dict(v=v, aext=s.aext, rhoe=s.rhoe, mue=s.mue, lambdae=s.lambdae, dbc=s.dbc,
              strategy=O.INIT_KINETIC_ENERGY_MINIMUM)
    cheb = pbat.sim.vbd.Integrator(data().with_chebyshev_acceleration(0.5).construct())
    oc = O.OracleVbd(s.X, s.E, chebyshev_rho=0.5, **kw)
    cheb.step(DT, 10, 1)
    oc.step(DT, 10, 1)
    assert rel(cheb.x, oc.positions) <= 1e-5
    tr = pbat.sim.vbd.Integrator(data().with_trust_region_acceleration(0.2, 2.0, False).construct())
    op = O.OracleVbd(s.X, s.E, **kw)
    tr.step(DT, 10, 1)
    op.step(DT, 10, 1)
    assert rel(tr.x, op.positions) <= 1e-5
    for bad in (data().with_nesterov_acceleration(1.0), data().with_broyden_acceleration(3)):
        with pytest.raises(ValueError, match="not available"):
            pbat.sim.vbd.Integrator(bad.construct())


def test_objective_and_gradient_match_oracle(gpu_device):
    s, v, data = scene(8, 4)
    sim = pbat.sim.vbd.Integrator(data)
    rng = np.random.default_rng(5)
    x = (s.X + 0.02 * rng.normal(size=s.X.shape)).astype(np.float32).astype(np.float64)
    xt = (x + 0.01 * rng.normal(size=s.X.shape)).astype(np.float32).astype(np.float64)
    o = O.OracleVbd(s.X, s.E, aext=s.aext, rhoe=s.rhoe, mue=s.mue, lambdae=s.lambdae, dbc=s.dbc)
    f = sim.objective_function(x, xt, DT)
    g = sim.objective_function_gradient(x, xt, DT)
    f_ref = o.objective(x, xt, DT)
    g_ref = o.objective_gradient(x, xt, DT)
    print(f"f rel err {abs(f - f_ref) / abs(f_ref):.2e}, grad rel err {rel(g, g_ref):.2e}")
    assert abs(f - f_ref) <= 1e-6 * abs(f_ref)
    assert rel(g, g_ref) <= 1e-5
    # the squared residual is the gradient's norm over the non-Dirichlet vertices
    _, r2 = sim._dev.objective(DT, x, xt)
    free = np.setdiff1d(np.arange(s.X.shape[1]), s.dbc)
    gf = g.reshape(-1, 3)[free]
    assert r2 == pytest.approx(float((gf * gf).sum()), rel=1e-10)


def test_trace_next_step_writes_reference_files(gpu_device, tmp_path):
    s, v, data = scene(6, 6)
    sim = pbat.sim.vbd.Integrator(data)
    sim.step(DT, 10, 1)
    sim.trace_next_step(str(tmp_path), 3)
    it, ss = 10, 2
    sim.step(DT, it, ss)
    nV = s.X.shape[1]
    for sub in range(ss):
        f = scipy.io.mmread(os.path.join(tmp_path, f"3.{sub}.f.mtx"))
        G = scipy.io.mmread(os.path.join(tmp_path, f"3.{sub}.grad.mtx"))
        X = scipy.io.mmread(os.path.join(tmp_path, f"3.{sub}.x.mtx"))
        assert f.shape == (it + 1, 1) and G.shape == (3 * nV, it + 1) and X.shape == (3 * nV, it + 1)
        assert f[-1, 0] < f[0, 0]
        if sub == ss - 1:
            # the last column is the step's result; its objective is what the device
            # evaluates at the end state with this substep's inertial target
            np.testing.assert_array_equal(X[:, -1], sim.x.T.reshape(-1))
            fe, _, ge = sim._dev.objective(DT / ss, gradient=True)
            assert fe == f[-1, 0]
            np.testing.assert_array_equal(ge, G[:, -1])
    # only the marked step is traced
    sim.step(DT, it, 1)
    assert not any(n.startswith("-1.") for n in os.listdir(tmp_path))
    assert len(os.listdir(tmp_path)) == 3 * ss


def test_reference_cube_test(gpu_device):
    """The reference's own CPU integrator test (sim/vbd/Integrator.cpp:245-293): the
    8-vertex cube under default gravity (-z) falls, moves only along z, and after 10
    iterations the objective has decreased and its gradient norm is below 1e-4."""
    data = (pbat.sim.vbd.Data().with_volume_mesh(CUBE_P, CUBE_T)
            .with_surface_mesh(np.arange(8), CUBE_F).construct())
    vbd = pbat.sim.vbd.Integrator(data)
    dt = 1e-2
    xtilde = data.x + dt * data.v + dt * dt * data.aext
    f0 = vbd.objective_function(data.x, xtilde, dt)
    vbd.step(dt, 10, 1)
    dx = vbd.x - CUBE_P
    assert (dx[2] < 0).all()
    assert (np.abs(dx[:2]) < 1e-4).all()
    g = vbd.objective_function_gradient(vbd.x, xtilde, dt)
    print(f"|grad f| = {np.linalg.norm(g):.3e}")
    assert np.linalg.norm(g) < 1e-4
    assert vbd.objective_function(vbd.x, xtilde, dt) < f0


def test_sim_integrator_properties(gpu_device):
    s, v, data = scene(4, 7)
    sim = pbat.sim.vbd.Integrator(data)
    sim.kD = 0.01
    sim.detH_residual = 1e-9
    sim.strategy = pbat.sim.vbd.Position
    assert sim.kD == 0.01 and sim.detH_residual == 1e-9 and sim.strategy == pbat.sim.vbd.Position
    sim.x = s.X
    sim.v = np.zeros_like(s.X)
    np.testing.assert_array_equal(sim.x, s.X.astype(np.float32))
    sim.step(DT, 5)
    d = sim.data
    assert d.kD == 0.01 and d.strategy == pbat.sim.vbd.Position
    sim.data = d  # rebuilds from the data (positions included)
    np.testing.assert_allclose(sim.x, d.x, rtol=0, atol=0)
