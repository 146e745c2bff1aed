"""Status reporting of the device step: non-finite states and the FP64 path of the
3x3 block solve, against the FP64 oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from physicsbasedanimationtoolkit_amd import _native, meshes, pbat

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def scene(n, nu, Y=1e6):
    X, E = meshes.cube_tet_grid(n)
    nT = E.shape[1]
    mu, lam = meshes.lame_coefficients(np.full(nT, Y), nu)
    a = meshes.gravity(X.shape[1])
    dbc = meshes.bottom_vertices(X)
    v = 0.2 * np.random.default_rng(3).normal(size=X.shape)
    data = (pbat.sim.vbd.Data().with_volume_mesh(X, E).with_material(np.full(nT, 1e3), mu, lam)
            .with_acceleration(a).with_dirichlet_vertices(dbc).with_velocity(v).construct())
    o = O.OracleVbd(X, E, v=v, aext=a, rhoe=np.full(nT, 1e3), mue=mu, lambdae=lam, dbc=dbc)
    return X, data, o


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("nu", [0.499995, 0.4999975])
def test_fp64_block_solve_fires_and_matches_oracle(gpu_device, nu):
    """Near-incompressible tets (lambda / mu = 1e5, 2e5): the blocks of 24 vertices of
    the 4^3 grid have |det H| / |H00 H11 H22| ~ 1e-5 at rest (CPU check: the oracle's
    vertex blocks plus inertia), below the 1e-4 threshold of the FP32 adjugate, so the
    sweep solves them in FP64. The counter shows the path ran, and positions stay within
    1e-5 of the FP64 oracle (sim/vbd/Kernels.h:324-339). The solve takes the block as
    accumulated in FP32, whose entries carry rounding of order eps32 * lambda: at
    lambda / mu = 5e5 (nu = 0.499999) that alone moves the soft direction's step enough
    to reach 1.4e-5 after one step."""
    X, data, o = scene(4, nu)
    ratios = []
    for i in range(X.shape[1]):
        H, _ = o.vertex_elastic_block(i)
        H32 = (H + o.m[i] / DT / DT * np.eye(3)).astype(np.float32).astype(np.float64)
        ratios.append(abs(np.linalg.det(H32)) / abs(H32[0, 0] * H32[1, 1] * H32[2, 2]))
    assert np.sum(np.array(ratios) < 1e-4) >= 8
    g = pbat.gpu.vbd.Integrator(data)
    assert g.fp64_solves() == 0
    errs = []
    for _ in range(3):
        g.step(DT, 10, 1)
        o.step(DT, 10, 1)
        errs.append(rel(g.x, o.positions))
    n64 = g.fp64_solves(reset=True)
    print(f"nu {nu}: {n64} FP64 block solves over 3 steps; rel position errors {errs}")
    assert n64 > 0
    assert g.fp64_solves() == 0
    assert max(errs) <= 1e-5


@pytest.mark.parametrize("where", ["free", "dirichlet"])
@pytest.mark.parametrize("graph", [True, False])
def test_nonfinite_state_raises(gpu_device, where, graph):
    """A NaN position (set by the caller) makes the step fail with the non-finite status
    (RuntimeError in Python), whether the vertex is swept or fixed; a finite state steps
    again afterwards."""
    X, data, o = scene(3, 0.45)
    g = pbat.gpu.vbd.Integrator(data)
    g.graph_capture = graph
    g.step(DT, 5, 1)
    x = g.x.copy()
    dbc = meshes.bottom_vertices(X)
    free = np.setdiff1d(np.arange(X.shape[1]), dbc)
    bad = x.copy()
    bad[1, free[3] if where == "free" else dbc[2]] = np.nan
    g.x = bad
    with pytest.raises(RuntimeError, match="non-finite") as err:
        g.step(DT, 5, 1)
    assert f"status {_native.ERROR_NONFINITE}" in str(err.value)
    g.x = x
    g.v = np.zeros_like(x)
    g.step(DT, 5, 1)
    assert np.isfinite(g.x).all()


def test_diverging_velocity_raises(gpu_device):
    """An infinite initial velocity on a free vertex overflows the inertial target:
    the step reports it instead of returning NaN positions silently."""
    X, data, o = scene(3, 0.45)
    g = pbat.gpu.vbd.Integrator(data)
    v = np.zeros(X.shape)
    free = np.setdiff1d(np.arange(X.shape[1]), meshes.bottom_vertices(X))
    v[0, free[0]] = np.inf
    g.v = v
    with pytest.raises(RuntimeError, match="non-finite"):
        g.step(DT, 5, 1)
