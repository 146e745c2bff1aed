"""Golden fixtures of the plain 10^3 scene (tests/golden/make_golden.py).

CPU: the oracle reproduces them (regression pin; the oracle is pinned by
test_oracle.py). GPU: the MI355X integrator matches them within the FP32
tolerances (positions 1e-5 relative after 1 and 10 steps, energy trajectory 1%).
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from physicsbasedanimationtoolkit_amd import meshes, pbat

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "plain_10cube.npz")


@pytest.fixture(scope="module")
def golden():
    return dict(np.load(GOLDEN, allow_pickle=False))


def test_oracle_reproduces_golden(golden):
    s = meshes.plain_cpu_scene(10)
    o = O.OracleVbd(s.X, s.E, aext=s.aext, rhoe=s.rhoe, mue=s.mue, lambdae=s.lambdae, dbc=s.dbc, nthreads=2)
    assert np.array_equal(o.colors, golden["colors"])
    assert np.array_equal(o.m, golden["mass"])
    o.step(s.dt, s.iterations, 1)
    assert np.allclose(o.positions, golden["x1"], rtol=0, atol=1e-13)
    for _ in range(9):
        o.step(s.dt, s.iterations, 1)
    assert np.allclose(o.positions, golden["x10"], rtol=0, atol=1e-12)
    assert np.allclose(o.velocities, golden["v10"], rtol=0, atol=1e-9)


def test_native_precompute_matches_golden(golden):
    s = meshes.plain_cpu_scene(10)
    d = (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_material(s.rhoe, s.mue, s.lambdae)
         .with_acceleration(s.aext).with_dirichlet_vertices(s.dbc).construct())
    assert np.array_equal(d.colors, golden["colors"])
    assert np.allclose(d.m, golden["mass"], rtol=1e-14, atol=0)


@pytest.mark.gpu
def test_gpu_matches_golden(gpu_device, golden):
    s = meshes.plain_cpu_scene(10)
    d = (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_material(s.rhoe, s.mue, s.lambdae)
         .with_acceleration(s.aext).with_dirichlet_vertices(s.dbc).construct())
    vbd = pbat.gpu.vbd.Integrator(d)
    energy = []
    for step in range(1, s.steps + 1):
        vbd.step(s.dt, s.iterations, 1)
        energy.append(vbd.energy().sum())
        x = vbd.x.astype(np.float64)
        for k in (1, 10, 100):
            if step == k:
                ref = golden[f"x{k}"]
                err = np.linalg.norm(x - ref) / np.linalg.norm(ref)
                print(f"step {k}: rel position error {err:.3e}")
                assert err <= 1e-5
    e_ref = golden["energy"].sum(axis=1)
    err = np.abs(np.array(energy) - e_ref).max() / np.abs(e_ref).max()
    print(f"energy trajectory rel error {err:.3e}")
    assert err <= 1e-2
