"""VBD benchmark: vertex updates per second of the plain VBD step on one MI355X.

Workload: 64^3 cube grid, 6 tets per cell (274,625 vertices, 1,572,864 tets),
interior vertices jittered +-0.1 cell, per-tet Young's modulus log-uniform in
[1e5, 1e7], nu = 0.45, rho = 1000, gravity -y, bottom face fixed, h = 1/60,
20 VBD iterations per step, colouring on the host. Inputs are synthetic and
generated from a fixed seed. One "step" = one Integrator.step(h, 20, 1).

value = vertex updates / s = (#non-Dirichlet vertices x iterations x steps) / time,
with the mesh resident in HBM before the timed region starts.

The JSON line also carries
  roofline: algorithmic bytes of the sweep kernel per launch (60 B per updated
            vertex + 120 B per vertex-tet incidence, FP32/int32) divided by the
            kernel's average launch time measured with HIP events on the
            integrator's stream, against 8 TB/s;
  cpu_baseline: the FP64 CPU oracle (same mesh, colouring and iterations) on a
            bounded sample (one step) using the host's cores;
  batched: 256 meshes stepped together vs the plain per-mesh loop;
  adjoint: fwd+adjoint steps/s of a differentiable 20-step rollout on a 32^3 mesh;
  batched_adjoint: fwd+adjoint steps/s of 64 meshes over 200 steps (second
            pass; the first, which also builds the handle's solver set-up,
            is reported beside it and against the plain per-mesh loop), with
            the peak device memory;
  variants: Chebyshev, Anderson and trust-region steps on the 64^3 workload
            (time and distance to the step's minimiser), and contact on 8
            cubes over a slab (time per step with and without contact).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_VERTEX = 60
BYTES_PER_INCIDENCE = 120
# CG tolerance |r| <= tol |b| of the adjoint benches: on the 12^3 stand-in of the
# differentiable workload the gradients stay at the FP32 floor (4-9e-6 against
# the FP64 IFT gradient, as at 1e-6) with 14% fewer iterations than at 1e-6
ADJOINT_TOL = 1e-5


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _kernel_source_sha256():
    import hashlib

    path = os.path.join(ROOT, "physicsbasedanimationtoolkit_amd", "csrc", "vbd_kernels.hpp")
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def _counter_traffic(name, shape):
    """HBM bytes per sweep launch from a committed rocprofv3 counter summary
    (profiles/<name>: FETCH_SIZE and WRITE_SIZE runs per sweep shape, each scaled by
    what the counter reads for a device copy of known size). Used only when the
    summary was measured on the current sweep kernel source and holds the sweep
    shape now in effect; returns (entry or None, provenance note)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, f"none: profiles/{name} absent"
    with open(path) as f:
        t = json.load(f)
    if t.get("kernel_source_sha256") != _kernel_source_sha256():
        return None, f"none: profiles/{name} was measured on another version of the sweep kernel source"
    want = {k: shape[k] for k in ("lanes_per_vertex", "incidences_per_lane", "xcd_remap", "nontemporal",
                                  "fixed_stride")}
    entry = next((e for e in t.get("shapes", []) if e.get("sweep_shape") == want), None)
    if entry is None:
        return None, f"none: profiles/{name} holds no measurement of sweep shape {want}"
    return entry, (f"profiles/{name} (rocprofv3 --pmc FETCH_SIZE x {t['fetch_scale']:.3f} + WRITE_SIZE x "
                   f"{t['write_scale']:.3f}, per launch, same kernel source and sweep shape)")


def bench_batched(args, torch):
    """Batched VBD: 256 jittered cube meshes with grid dims U{4..24} per axis, per-mesh
    Y log-uniform in [1e5, 1e7] and nu uniform in [0.3, 0.48], h = 1/60, 10 iterations.
    One batched step advances all meshes (one launch per colour for the batch); the
    baseline is the plain per-mesh loop (one Integrator.step per mesh)."""
    from physicsbasedanimationtoolkit_amd import meshes, pbat

    count = args.batch_meshes
    scenes = meshes.batched_scenes(count, 4, 24, seed=0)
    t0 = time.time()
    datas = [(pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_material(s.rhoe, s.mue, s.lambdae)
              .with_acceleration(s.aext).with_dirichlet_vertices(s.dbc).construct()) for s in scenes]
    dt, iters = scenes[0].dt, scenes[0].iterations
    batch = pbat.gpu.vbd.BatchIntegrator(datas)
    nColors, nUpd, nInc = batch.counts()
    log(f"[bench] batch: {count} meshes, {batch.offsets[-1]} vertices, {sum(s.n_tets for s in scenes)} tets, "
        f"{nColors} colours, built in {time.time() - t0:.1f}s")
    if args.lanes:
        batch.set_sweep_config(args.lanes, args.per_lane, not args.no_xcd, args.nontemporal,
                               None if args.stride == "auto" else args.stride == "fixed")
        log(f"[bench] batch sweep shape: {batch.sweep_config()} (set)")
    else:
        tune = batch.autotune(dt, sweeps=5)
        log(f"[bench] batch sweep shape: {batch.sweep_config()} ({min(r[-1] for r in tune):.1f} us/sweep)")
    if args.no_graph:
        batch.graph_capture = False
    for _ in range(2):
        batch.step(dt, iters, 1)
    batch.synchronize()
    steps = args.batch_steps
    batch.set_timing(True)
    batch.sweep_timing(reset=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.step(dt, iters, 1)
    batch.synchronize()
    tb = (time.perf_counter() - t0) / steps
    ms, launches = batch.sweep_timing(reset=True)
    batch.set_timing(False)
    if args.batch_no_loop:
        return {"steps_per_s": 1 / tb, "sweep_launches": launches, "sweep_ms": ms}
    # plain per-mesh loop on the same meshes
    singles = [pbat.gpu.vbd.Integrator(d) for d in datas]
    for g in singles:
        g.step(dt, iters, 1)
    torch.cuda.synchronize()
    loop_steps = max(1, steps // 10)
    t0 = time.perf_counter()
    for _ in range(loop_steps):
        for g in singles:
            g.step(dt, iters, 1)
    torch.cuda.synchronize()
    tl = (time.perf_counter() - t0) / loop_steps
    log(f"[bench] batched {1 / tb:.1f} steps/s ({1e3 * tb:.2f} ms) vs per-mesh loop {1 / tl:.2f} steps/s "
        f"({1e3 * tl:.1f} ms): {tl / tb:.1f}x")
    # the sweep kernel: algorithmic bytes per launch (the 64^3 model, 60 B per
    # vertex + 120 B per incidence, averaged over the colours) over the average
    # launch time from HIP events around the replayed step graphs, and the
    # HBM bytes the counters measured for this shape
    shape = batch.sweep_config()
    avg_launch_s = ms * 1e-3 / max(launches, 1)
    bytes_per_launch = (BYTES_PER_VERTEX * nUpd + BYTES_PER_INCIDENCE * nInc) / nColors
    achieved = bytes_per_launch / avg_launch_s / 1e9
    traffic, traffic_note = _counter_traffic("batched_sweep_256meshes_traffic.json", shape)
    tgbs = None if traffic is None else traffic["hbm_bytes_per_launch"] / avg_launch_s / 1e9
    log(f"[bench] batched sweep kernel: {launches} launches, avg {avg_launch_s * 1e6:.2f} us, "
        f"{bytes_per_launch / 1e6:.2f} MB/launch -> {achieved:.0f} GB/s algorithmic; counter traffic "
        f"{traffic_note}" + ("" if tgbs is None else f": {tgbs:.0f} GB/s"))
    return {"workload": f"{count} jittered cube meshes, grid dims U{{4..24}} per axis, per-mesh Y, nu; "
                        f"h=1/60, {iters} iterations",
            "roofline": {"bound": "hbm", "kernel": "VbdSweep", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "avg_launch_us": avg_launch_s * 1e6,
                         "algorithmic_bytes_per_launch": bytes_per_launch,
                         "traffic": None if traffic is None else traffic["hbm_bytes_per_launch"],
                         "traffic_source": traffic_note, "traffic_gbs": tgbs,
                         "traffic_frac": None if tgbs is None else tgbs / HBM_PEAK_GBS},
            "meshes": count, "vertices": int(batch.offsets[-1]), "updated_vertices": nUpd,
            "steps_per_s": 1 / tb, "per_mesh_loop_steps_per_s": 1 / tl, "speedup_vs_loop": tl / tb,
            "vertex_updates_per_s": nUpd * iters / tb, "steps": steps, "loop_steps": loop_steps,
            "sweep_shape": batch.sweep_config()}


def _diff_data(pbat, s):
    return (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_material(s.rhoe, s.mue, s.lambdae)
            .with_acceleration(s.aext).with_dirichlet_vertices(s.dbc).with_velocity(s.v0).construct())


def _rel(a, b):
    import numpy as np
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _adjoint_solver(args):
    """The differentiable measurements' solver: two-level PCG at ADJOINT_TOL per
    mesh, each solve started from the projection onto the last
    args.adjoint_history solutions."""
    return dict(solver="pcg_two_level", tolerance=ADJOINT_TOL, max_iterations=2000,
                check_every=args.adjoint_check_every, history=args.adjoint_history)


def bench_adjoint(args, torch):
    """Differentiable VBD on one mesh: 32^3 grid (35,937 vertices, 196,608 tets),
    hidden per-tet log Y log-uniform in [1e5, 1e7], target = 20-step rollout
    (h = 1/60, 20 iterations) from v0 = 0.5 x random unit vectors; loss
    L = sum |x_T - x_T^target|^2 from log Y = ln 1e6. One fwd+adjoint pass = the
    20-step forward (state kept after every step: 34 MB, so the backward pass
    recomputes nothing) + the IFT backward (one
    two-level PCG solve per step, tolerance ADJOINT_TOL per mesh) giving dL/dlogY,
    dL/dv0, dL/dx0. The gradient checks (against the FP64 IFT gradient and FP64
    finite differences) run in the CPU leg: adjoint_gradient_checks."""
    import numpy as np
    from physicsbasedanimationtoolkit_amd import meshes, pbat, _native

    s = meshes.differentiable_scene(args.adjoint_grid, seed=0)
    g = pbat.gpu.vbd.Integrator(_diff_data(pbat, s))
    g.autotune(s.dt, sweeps=40)  # 8 launches of ~4 us per sweep: a short sample picks noisily
    v0 = g.v.copy()
    dt, it, T, every = s.dt, s.iterations, s.steps, 1
    solver = _adjoint_solver(args)

    def forward(logy):
        g.set_log_young_modulus(logy)
        g.x = s.X
        g.v = v0
        g.rollout_forward(dt, it, T, every)

    forward(s.log_y_hidden)
    target = g.x.astype(np.float64)
    forward(s.log_y_guess)
    g.rollout_backward(2 * (g.x - target), None, **solver)  # warm-up
    g.synchronize()
    reps = args.adjoint_reps
    tf = tb = 0.0
    stats = None
    for _ in range(reps):
        t0 = time.perf_counter()
        forward(s.log_y_guess)
        g.synchronize()
        t1 = time.perf_counter()
        xT = g.x.astype(np.float64)
        r = g.rollout_backward(2 * (xT - target), None, **solver)
        tb += time.perf_counter() - t1
        tf += t1 - t0
        stats = r["stats"]
    loss = float(((xT - target) ** 2).sum())
    out = {"workload": f"differentiable VBD, {args.adjoint_grid}^3 jittered grid ({s.n_vertices} verts, "
                       f"{s.n_tets} tets), hidden per-tet log Y, {T} steps h=1/60 x {it} iterations, "
                       f"v0 = 0.5 random unit vectors, L = |x_T - target|^2 from log Y = ln 1e6",
           "fwd_adjoint_steps_per_s": T * reps / (tf + tb), "forward_steps_per_s": T * reps / tf,
           "forward_ms": 1e3 * tf / reps, "adjoint_ms": 1e3 * tb / reps, "loss": loss,
           "solver": solver, "checkpoint_every": every, "adjoint_stats": stats,
           "device_bytes": _native.device_memory()[0]}
    log(f"[bench] adjoint 32^3: fwd {out['forward_ms']:.1f} ms + adjoint {out['adjoint_ms']:.1f} ms per "
        f"{T}-step rollout -> {out['fwd_adjoint_steps_per_s']:.1f} fwd+adj steps/s; {stats}")
    del g
    return out


def bench_batched_adjoint(args, torch):
    """Batched + differentiable VBD: 64 meshes with grid dims U{8..20} per axis
    (236,010 vertices, 1,158,930 tets), hidden per-tet log Y, 200 steps at
    h = 1/120 x 10 iterations, state kept every 10 steps and recomputed in the
    backward pass, L = sum |x_T - x_T^target|^2 over the batch, gradient wrt
    per-tet log Y (and v0, x0). Two passes: the first also builds the handle's
    Hessian pattern, coarse pattern and captured CG iteration; the second is
    the steady state of an optimisation loop (fwd_adjoint_steps_per_s).
    Compared with the plain per-mesh loop: one Integrator per mesh doing one
    fwd+adjoint rollout (its set-up included), same sweep shape, against the
    batch's first pass."""
    import numpy as np
    from physicsbasedanimationtoolkit_amd import meshes, pbat, _native

    scenes = meshes.batched_differentiable_scenes(args.batch_adjoint_meshes, 8, 20, seed=0)
    datas = [_diff_data(pbat, s) for s in scenes]
    b = pbat.gpu.vbd.BatchIntegrator(datas)
    b.autotune(scenes[0].dt, sweeps=10)
    shape = b.sweep_config()
    dt, it, T, every = scenes[0].dt, scenes[0].iterations, args.batch_adjoint_steps, 10
    solver = _adjoint_solver(args)
    hidden = np.concatenate([s.log_y_hidden for s in scenes])
    guess = np.concatenate([s.log_y_guess for s in scenes])
    Xb = np.concatenate([s.X for s in scenes], axis=1)
    vb = b.v.copy()

    def forward(logy):
        b.set_log_young_modulus(logy)
        b.x = Xb
        b.v = vb
        b.rollout_forward(dt, it, T, every)

    forward(hidden)
    target = b.x.astype(np.float64)
    _native.device_memory(reset_peak=True)

    def rollout():
        t0 = time.perf_counter()
        forward(guess)
        b.synchronize()
        t1 = time.perf_counter()
        r = b.rollout_backward(2 * (b.x - target), None, **solver)
        return r, t1 - t0, time.perf_counter() - t1

    # the first pass also builds the handle's Hessian pattern, coarse operator
    # pattern and captured CG graph (once per handle); the second is the
    # steady state of an optimisation loop
    rb, tf0, tb0 = rollout()
    rb, tf, tb = rollout()
    mem_now, mem_peak = _native.device_memory()
    log(f"[bench] batched adjoint: {len(scenes)} meshes, fwd {tf:.2f}s + adjoint {tb:.2f}s per {T}-step rollout "
        f"-> {T / (tf + tb):.2f} steps/s (first pass {tf0:.2f}s + {tb0:.2f}s -> {T / (tf0 + tb0):.2f}); "
        f"peak device memory {mem_peak / 2**20:.0f} MiB; {rb['stats']}")
    if args.batch_no_loop:
        return {"fwd_adjoint_steps_per_s": T / (tf + tb), "forward_s": tf, "adjoint_s": tb,
                "first_pass_steps_per_s": T / (tf0 + tb0), "adjoint_stats": rb["stats"]}
    # plain per-mesh loop, same sweep shape so that each forward is bitwise the batch's
    loop_t = 0.0
    errs = []
    t0e = 0
    for k, (s, d) in enumerate(zip(scenes, datas)):
        gi = pbat.gpu.vbd.Integrator(d)
        gi.set_sweep_config(shape["lanes_per_vertex"], shape["incidences_per_lane"], shape["xcd_remap"],
                            shape["nontemporal"], shape["fixed_stride"])
        sl = slice(b.offsets[k], b.offsets[k + 1])
        nT = s.n_tets
        tk = target[:, sl]
        gi.rollout_forward(dt, 2, 2, every)  # build the step graph outside the timed region
        gi.synchronize()
        t0 = time.perf_counter()
        gi.set_log_young_modulus(s.log_y_guess)
        gi.x = s.X
        gi.v = vb[:, sl]
        gi.rollout_forward(dt, it, T, every)
        r = gi.rollout_backward(2 * (gi.x - tk), None, **solver)
        loop_t += time.perf_counter() - t0
        errs.append(_rel(rb["log_young_modulus"][t0e:t0e + nT], r["log_young_modulus"]))
        t0e += nT
        del gi
    log(f"[bench] per-mesh loop: {loop_t:.2f}s -> {T / loop_t:.2f} steps/s; batched first pass is "
        f"{loop_t / (tf0 + tb0):.1f}x; max rel grad diff batch vs loop {max(errs):.2e}")
    return {"workload": f"{len(scenes)} jittered meshes, grid dims U{{8..20}} per axis ({int(b.offsets[-1])} verts, "
                        f"{sum(s.n_tets for s in scenes)} tets), hidden per-tet log Y, {T} steps h=1/120 x {it} "
                        f"iterations, checkpoint every {every}, L = |x_T - target|^2, grad wrt per-tet log Y",
            "fwd_adjoint_steps_per_s": T / (tf + tb), "first_pass_steps_per_s": T / (tf0 + tb0),
            "per_mesh_loop_steps_per_s": T / loop_t,
            "speedup_vs_loop": loop_t / (tf0 + tb0),
            "speedup_basis": "first batched pass (handle set-up included) vs one pass of each per-mesh "
                             "integrator (its set-up included)",
            "forward_s": tf, "adjoint_s": tb, "first_pass_s": tf0 + tb0, "loop_s": loop_t,
            "max_rel_grad_diff_vs_loop": max(errs), "peak_device_bytes": mem_peak,
            "solver": solver, "adjoint_stats": rb["stats"], "sweep_shape": shape}


def _time_steps(g, dt, iters, substeps, steps, torch):
    for _ in range(2):
        g.step(dt, iters, substeps)
    g.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.step(dt, iters, substeps)
    g.synchronize()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def bench_variants(args, torch):
    """Accelerated integrators next to plain VBD, 20 iterations per step: time per step
    on the 64^3 workload (one step from rest), and on a 16^3 mesh with an initial
    velocity field the distance of one step's result to that step's minimiser (plain
    VBD with 2000 iterations). Vertex-triangle contact on 8 jittered 12^3 cubes
    falling onto a fixed slab: time per step with and without contact."""
    import numpy as np

    from physicsbasedanimationtoolkit_amd import meshes, pbat

    s = meshes.plain_gpu_scene(args.grid, seed=0)
    small = meshes.plain_gpu_scene(16, seed=3)
    v16 = np.random.default_rng(5).normal(size=(3, small.n_vertices)) * 0.3
    dt, iters = s.dt, args.iterations

    def data(sc, v=None, **acc):
        d = (pbat.sim.vbd.Data().with_volume_mesh(sc.X, sc.E).with_material(sc.rhoe, sc.mue, sc.lambdae)
             .with_acceleration(sc.aext).with_dirichlet_vertices(sc.dbc))
        if v is not None:
            d = d.with_velocity(v)
        if "rho" in acc:
            d = d.with_chebyshev_acceleration(acc["rho"])
        if "window" in acc:
            d = d.with_anderson_acceleration(acc["window"])
        if "eta" in acc:
            d = d.with_trust_region_acceleration(acc["eta"], 2.0, False)
        return d.construct()

    g = pbat.gpu.vbd.Integrator(data(small, v16))
    g.step(dt, 2000, 1)
    x_star = g.x.astype(np.float64)
    g = pbat.gpu.vbd.Integrator(data(small, v16))
    g.step(dt, 0, 1)  # the step's initial iterate
    e0 = np.linalg.norm(g.x.astype(np.float64) - x_star)
    counts = (10, 20, 40, 80)
    out = {"workload": f"time: {args.grid}^3 plain workload, one step from rest, {iters} iterations; error: 16^3 "
                       "jittered mesh with v0 ~ 0.3 N(0,1), one step of k iterations, |x_k - x*| / |x_0 - x*| "
                       "(x_0 the step's initial iterate, x* plain VBD with 2000 iterations) for k = "
                       + ", ".join(map(str, counts))}
    for name, acc in (("plain", {}), ("chebyshev_rho_0.9", {"rho": 0.9}), ("anderson_m5", {"window": 5}),
                      ("trust_region_linear", {"eta": 0.2})):
        err = []
        for k in counts:
            e = pbat.gpu.vbd.Integrator(data(small, v16, **acc))
            e.step(dt, k, 1)
            err.append(float(np.linalg.norm(e.x - x_star) / e0))
        g = pbat.gpu.vbd.Integrator(data(s, **acc))
        g.step(dt, iters, 1)
        reps = 10
        t = 0.0
        for _ in range(reps):
            g.x = s.X.astype(np.float32)
            g.v = np.zeros_like(s.X, dtype=np.float32)
            g.synchronize()
            t0 = time.perf_counter()
            g.step(dt, iters, 1)
            g.synchronize()
            t += time.perf_counter() - t0
        out[name] = {"ms_per_step": 1e3 * t / reps, "error_vs_minimiser": dict(zip(map(str, counts), err))}
        log(f"[bench] {name}: {1e3 * t / reps:.2f} ms/step ({args.grid}^3), error vs minimiser (16^3) "
            + " ".join(f"{k}:{v:.2e}" for k, v in zip(counts, err)))

    # contact: 8 cubes in a 2 x 2 x 2 arrangement above a fixed slab
    n = 12
    parts, elems, bodies = [], [], []
    Xs, Es = meshes.cube_tet_grid(8, 1, 8)
    Xs = Xs * np.array([[1.2], [0.1], [1.2]]) - np.array([[0.1], [0.1], [0.1]])
    parts.append(Xs)
    elems.append(Es)
    bodies.append(np.zeros(Xs.shape[1], np.int64))
    off = Xs.shape[1]
    for b, (i, j, k) in enumerate(np.ndindex(2, 2, 2)):
        Xc, Ec = meshes.cube_tet_grid(n, jitter=0.1, seed=b, size=0.45)
        Xc = Xc + np.array([[0.5 * i + 0.013 * b], [0.004 + 0.47 * j], [0.5 * k + 0.007 * b]])
        parts.append(Xc)
        elems.append(Ec + off)
        bodies.append(np.full(Xc.shape[1], b + 1, np.int64))
        off += Xc.shape[1]
    X = np.concatenate(parts, axis=1)
    E = np.concatenate(elems, axis=1)
    B = np.concatenate(bodies)
    F = meshes.boundary_triangles(E)
    V = np.unique(F)
    v = np.zeros(X.shape)
    v[1, B > 0] = -0.5
    nT = E.shape[1]
    mu, lam = meshes.lame_coefficients(np.full(nT, 1e6), 0.45)
    base = (pbat.sim.vbd.Data().with_volume_mesh(X, E).with_material(np.full(nT, 1e3), mu, lam)
            .with_acceleration(meshes.gravity(X.shape[1])).with_velocity(v)
            .with_dirichlet_vertices(np.arange(Xs.shape[1])))
    with_contact = pbat.gpu.vbd.Integrator(base.with_surface_mesh(V, F).with_bodies(B).construct())
    base2 = (pbat.sim.vbd.Data().with_volume_mesh(X, E).with_material(np.full(nT, 1e3), mu, lam)
             .with_acceleration(meshes.gravity(X.shape[1])).with_velocity(v)
             .with_dirichlet_vertices(np.arange(Xs.shape[1])))
    without = pbat.gpu.vbd.Integrator(base2.construct())
    steps = 30
    tc = _time_steps(with_contact, 1.0 / 60.0, iters, 1, steps, torch)
    tp = _time_steps(without, 1.0 / 60.0, iters, 1, steps, torch)
    active, _ = with_contact.contacts()
    out["contact"] = {"workload": f"8 jittered {n}^3 cubes (size 0.45) over a fixed slab, {X.shape[1]} vertices, "
                                  f"{nT} tets, {V.size} collision vertices, {F.shape[1]} collision triangles, "
                                  f"h=1/60, {iters} iterations, {steps + 2} steps from 0.5 m/s",
                      "ms_per_step": 1e3 * tc, "ms_per_step_without_contact": 1e3 * tp,
                      "active_vertices_at_end": int(active.sum())}
    log(f"[bench] contact: {1e3 * tc:.2f} ms/step vs {1e3 * tp:.2f} ms/step without contact "
        f"({V.size} collision vertices, {F.shape[1]} triangles, {int(active.sum())} active at the end)")
    return out


def shape_sweep(args, torch, workload):
    """Counter-profiling driver: every candidate shape of the sweep autotuner (the
    list its own run returns) steps the workload once with eager launches, in that
    order, so that a rocprofv3 --pmc run of this command attributes each group of
    colours x iterations consecutive VbdSweep dispatches to one shape (the
    autotuner's own 2 x colours dispatches per candidate come first). Prints the
    order as JSON."""
    from physicsbasedanimationtoolkit_amd import meshes, pbat

    if workload == "plain":
        s = meshes.plain_gpu_scene(args.grid, seed=0)
        data = (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_material(s.rhoe, s.mue, s.lambdae)
                .with_acceleration(s.aext).with_dirichlet_vertices(s.dbc).construct())
        target = pbat.gpu.vbd.Integrator(data)
        dt, iters = s.dt, args.iterations
    else:
        scenes = meshes.batched_scenes(args.batch_meshes, 4, 24, seed=0)
        datas = [(pbat.sim.vbd.Data().with_volume_mesh(c.X, c.E).with_material(c.rhoe, c.mue, c.lambdae)
                  .with_acceleration(c.aext).with_dirichlet_vertices(c.dbc).construct()) for c in scenes]
        target = pbat.gpu.vbd.BatchIntegrator(datas)
        dt, iters = scenes[0].dt, scenes[0].iterations
    nColors = target.counts()[0]
    rows = target.autotune(dt, sweeps=1)
    target.graph_capture = False
    shapes = []
    for L, K, X, N, S, _ in rows:
        target.set_sweep_config(int(L), int(K), bool(X), bool(N), bool(S))
        target.step(dt, iters, 1)
        shapes.append(target.sweep_config())
    target.synchronize()
    return {"workload": workload, "autotune_sweep_dispatches": 2 * nColors * len(rows),
            "sweep_dispatches_per_shape": nColors * iters, "shapes": shapes}


def adjoint_gradient_checks(args, solver):
    """CPU leg (oracle as the checker, after every timed region):
    (1) the 32^3 differentiable workload's settings on a 12^3 grid (the FP64 direct
        solve at 32^3 takes minutes per step): device gradients wrt log Y and v0
        against the FP64 IFT gradient along the device trajectory, with the bench's
        solver settings;
    (2) 64 random log Y entries of a soft 4^3 rollout whose 400 iterations per step
        converge to round-off: the device gradient against FP64 central differences
        of oracle rollouts (step 1e-4), which differentiate the minimiser sequence the
        implicit-function theorem describes."""
    import numpy as np
    from oracle import adjoint as A
    from oracle import oracle as O
    from physicsbasedanimationtoolkit_amd import meshes, pbat

    out = {}
    s = meshes.differentiable_scene(12, seed=3)
    g = pbat.gpu.vbd.Integrator(_diff_data(pbat, s))
    v0 = g.v.copy()
    g.set_log_young_modulus(s.log_y_hidden)
    g.rollout_forward(s.dt, s.iterations, s.steps, 10)
    target = g.x.astype(np.float64)
    g.set_log_young_modulus(s.log_y_guess)
    g.x, g.v = s.X, v0
    xs = [np.asarray(s.X, np.float64).T.copy()]
    for _ in range(s.steps):
        g.step(s.dt, s.iterations, 1)
        xs.append(g.x.astype(np.float64).T.copy())
    dLdx = 2 * (xs[-1].T - target)
    g.x, g.v = s.X, v0
    g.rollout_forward(s.dt, s.iterations, s.steps, 10)
    r = g.rollout_backward(dLdx, None, **solver)
    o = O.OracleVbd(s.X, s.E, v=s.v0, aext=s.aext, rhoe=s.rhoe, mue=s.mue, lambdae=s.lambdae, dbc=s.dbc,
                    nthreads=8)
    A.set_log_young_modulus(o, s.log_y_guess)
    ref = A.ift_gradient(o, xs, s.dt, dLdx.T)
    out["vs_fp64_ift"] = {"workload": "12^3 jittered grid, hidden per-tet log Y, v0 = 0.5 random unit vectors, "
                                      f"{s.steps} steps h=1/60 x {s.iterations} iterations, "
                                      "L = |x_T - target|^2 from log Y = ln 1e6",
                          "rel_err_log_young_modulus": _rel(r["log_young_modulus"], ref[0]),
                          "rel_err_v0": _rel(r["v0"], ref[1].T), "solver": solver, "stats": r["stats"]}
    # (2) finite differences on a converged soft rollout
    X, E = meshes.cube_tet_grid(4, jitter=0.1, seed=3)
    rng = np.random.default_rng(3)
    nT = E.shape[1]
    mu, lam = meshes.lame_coefficients(meshes.log_uniform(rng, 1e4, 1e5, nT), 0.45)
    v = rng.normal(size=X.shape) * 0.5
    dbc, a, rho = meshes.bottom_vertices(X), meshes.gravity(X.shape[1]), np.full(nT, 1e3)
    dt, iters, steps = 1.0 / 60.0, 400, 3
    data = (pbat.sim.vbd.Data().with_volume_mesh(X, E).with_material(rho, mu, lam).with_acceleration(a)
            .with_dirichlet_vertices(dbc).with_velocity(v).construct())
    gs = pbat.gpu.vbd.Integrator(data)
    w = np.random.default_rng(11).normal(size=X.shape)
    gs.rollout_forward(dt, iters, steps, 1)
    rs = gs.rollout_backward(w, None, **dict(solver, tolerance=1e-7))

    def make():
        return O.OracleVbd(X, E, v=v, aext=a, rhoe=rho, mue=mu, lambdae=lam, dbc=dbc, nthreads=4)

    logy0 = np.log(A.young_modulus(make()))
    idx = np.random.default_rng(1).choice(nT, 64, replace=False)
    eps = 1e-4
    fd = []
    for e in idx:
        lv = []
        for sgn in (1.0, -1.0):
            oe = make()
            ly = logy0.copy()
            ly[e] += sgn * eps
            A.set_log_young_modulus(oe, ly)
            lv.append(float((w.T * A.rollout(oe, dt, iters, steps)[-1]).sum()))
        fd.append((lv[0] - lv[1]) / (2 * eps))
    fd = np.array(fd)
    dev = np.asarray(rs["log_young_modulus"], np.float64)[idx]
    err = np.abs(dev - fd) / np.abs(fd).max()
    out["fd_check"] = {"params": 64, "workload": f"soft 4^3 jittered grid (Y log-uniform in [1e4, 1e5]), "
                                                 f"{steps} steps x {iters} iterations (converged), L = w . x_T",
                       "eps": eps, "fd": "FP64 central differences of oracle rollouts",
                       "rel_err_vector": _rel(dev, fd), "max_err_rel_to_max": float(err.max()),
                       "median_err_rel_to_max": float(np.median(err))}
    log(f"[bench] gradient checks: 12^3 vs FP64 IFT log Y {out['vs_fp64_ift']['rel_err_log_young_modulus']:.2e} "
        f"v0 {out['vs_fp64_ift']['rel_err_v0']:.2e}; 64-entry FD (4^3 converged) "
        f"rel err {out['fd_check']['rel_err_vector']:.2e}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--iterations", type=int, default=20)
    ap.add_argument("--lanes", type=int, default=0, help="lanes per vertex (0: autotune)")
    ap.add_argument("--per-lane", type=int, default=0, help="incidences prefetched per lane (0: auto)")
    ap.add_argument("--no-xcd", action="store_true", help="disable the XCD-aware workgroup order")
    ap.add_argument("--nontemporal", action="store_true", help="non-temporal loads of the incidence stream")
    ap.add_argument("--stride", choices=["auto", "fixed", "csr"], default="auto",
                    help="incidence rows: fixed-stride (ELL) or CSR (with --lanes)")
    ap.add_argument("--no-graph", action="store_true", help="launch eagerly instead of replaying a hipGraph")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-mesh measurement")
    ap.add_argument("--batch-meshes", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=30)
    ap.add_argument("--batch-no-loop", action="store_true",
                    help="batched measurements without their per-mesh loop baselines (profiling runs)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-rollout measurements")
    ap.add_argument("--adjoint-grid", type=int, default=32)
    ap.add_argument("--adjoint-reps", type=int, default=3)
    ap.add_argument("--adjoint-history", type=int, default=6,
                    help="previous solutions each backward solve's initial guess is projected onto (0..8)")
    ap.add_argument("--adjoint-check-every", type=int, default=10,
                    help="CG iterations per captured chunk between live-mesh polls")
    ap.add_argument("--batch-adjoint-meshes", type=int, default=64)
    ap.add_argument("--batch-adjoint-steps", type=int, default=200)
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the accelerated-integrator and contact measurements")
    ap.add_argument("--only", choices=["batched", "adjoint", "batched_adjoint", "variants", "gradient_checks"],
                    default=None,
                    help="run one secondary measurement alone and print its JSON (profiling runs)")
    ap.add_argument("--shape-sweep", choices=["plain", "batched"], default=None,
                    help="step the workload once per autotuner candidate shape, eagerly, and print the order "
                         "(counter profiling of every shape in one run)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only run a plain device copy of this many MB (counter calibration) and exit")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the VBD path has no multi-GPU form here "
              "(a batch of meshes would shard by mesh, a single mesh by spatial subdomain with a halo "
              "exchange per colour); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from physicsbasedanimationtoolkit_amd import meshes, pbat

    torch.cuda.set_device(0)
    if args.copy_probe_mb:
        n = args.copy_probe_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 4 * n, "launches": 5}), flush=True)
        return
    if args.shape_sweep:
        print(json.dumps(shape_sweep(args, torch, args.shape_sweep)), flush=True)
        return
    if args.only:
        run = {"batched": bench_batched, "adjoint": bench_adjoint, "batched_adjoint": bench_batched_adjoint,
               "variants": bench_variants,
               "gradient_checks": lambda a, t: adjoint_gradient_checks(
                   a, _adjoint_solver(args))}
        print(json.dumps({args.only: run[args.only](args, torch)}), flush=True)
        return
    s = meshes.plain_gpu_scene(args.grid, seed=0)
    dt, iters = s.dt, args.iterations
    log(f"[bench] scene {s.name}: {s.n_vertices} vertices, {s.n_tets} tets")
    t0 = time.time()
    data = (pbat.sim.vbd.Data().with_volume_mesh(s.X, s.E).with_material(s.rhoe, s.mue, s.lambdae)
            .with_acceleration(s.aext).with_dirichlet_vertices(s.dbc).construct())
    vbd = pbat.gpu.vbd.Integrator(data)
    nColors, nUpd, nInc = vbd.counts()
    log(f"[bench] construct+upload {time.time() - t0:.2f}s: {nColors} colours, {nUpd} updated vertices, "
        f"{nInc} incidences")
    if args.lanes:
        vbd.set_sweep_config(args.lanes, args.per_lane, not args.no_xcd, args.nontemporal,
                             None if args.stride == "auto" else args.stride == "fixed")
    else:
        for L, K, X, N, S, us in vbd.autotune(dt, sweeps=10):
            log(f"[bench] autotune lanes={int(L)} per_lane={int(K)} xcd={int(X)} nt={int(N)} "
                f"fixed_stride={int(S)}: {us:.2f} us/sweep")
    if args.no_graph:
        vbd.graph_capture = False
    shape = vbd.sweep_config()
    log(f"[bench] sweep shape: {shape}")

    # HIP events recorded on the integrator's stream around each replayed step
    # graph time the timed region itself; per sweep launch this is an upper
    # bound (the bracket also holds the two elementwise substep kernels).
    vbd.set_timing(True)
    for _ in range(max(1, args.warmup)):
        vbd.step(dt, iters, 1)
    vbd.synchronize()
    torch.cuda.synchronize()
    vbd.sweep_timing(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        vbd.step(dt, iters, 1)
    vbd.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms, launches = vbd.sweep_timing(reset=True)
    vbd.set_timing(False)
    vu = nUpd * iters * args.steps
    value = vu / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    log(f"[bench] {args.steps} steps in {elapsed:.4f}s -> {value:.4e} vertex-updates/s")
    avg_launch_s = ms * 1e-3 / launches
    bytes_per_sweep = BYTES_PER_VERTEX * nUpd + BYTES_PER_INCIDENCE * nInc
    bytes_per_launch = bytes_per_sweep / nColors
    achieved = bytes_per_launch / avg_launch_s / 1e9
    traffic, traffic_note = _counter_traffic("vbd_sweep_64cube_traffic.json", shape)
    log(f"[bench] sweep kernel: {launches} launches, avg {avg_launch_s * 1e6:.2f} us, "
        f"{bytes_per_launch / 1e6:.2f} MB/launch -> {achieved:.1f} GB/s; counter traffic {traffic_note}")

    batched = None if args.no_batched else bench_batched(args, torch)
    variants = None if args.no_variants else bench_variants(args, torch)
    adjoint = batched_adjoint = None
    if not args.no_adjoint:
        adjoint = bench_adjoint(args, torch)
        batched_adjoint = bench_batched_adjoint(args, torch)

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        # every core this process may run on (its affinity mask), within the job's
        # CPU share when the environment states one (OMP_NUM_THREADS)
        affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        threads = args.cpu_threads or (min(affinity, share) if share > 0 else affinity)
        ref = O.OracleVbd(s.X, s.E, aext=s.aext, rhoe=s.rhoe, mue=s.mue, lambdae=s.lambdae, dbc=s.dbc,
                          nthreads=threads)
        ref.step(dt, iters, 1)  # warm-up step
        cpu_steps = 3
        t0 = time.perf_counter()
        for _ in range(cpu_steps):
            ref.step(dt, iters, 1)
        cpu_s = time.perf_counter() - t0
        model = ""
        try:
            with open("/proc/cpuinfo") as f:
                model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
        except OSError:
            pass
        cpu = {"value": nUpd * iters * cpu_steps / cpu_s, "unit": "vertex-updates/s", "cores": threads,
               "kind": "port",
               "sample": f"1 warm-up + {cpu_steps} timed steps x {iters} iterations of the same 64^3 workload "
                         f"(from rest), FP64 oracle, {threads} threads, {cpu_s:.2f}s timed",
               "host": {"cpu_model": model, "os_cpu_count": os.cpu_count(), "affinity_cpus": affinity,
                        "omp_num_threads": share or None}}
        log(f"[bench] cpu baseline {cpu['value']:.4e} vu/s on {threads} threads")
        # one thread, on a shorter sample: a warm-up step of 1 iteration, then one
        # step of 4 iterations
        iters1 = 4
        ref1 = O.OracleVbd(s.X, s.E, aext=s.aext, rhoe=s.rhoe, mue=s.mue, lambdae=s.lambdae, dbc=s.dbc,
                           nthreads=1)
        ref1.step(dt, 1, 1)
        t0 = time.perf_counter()
        ref1.step(dt, iters1, 1)
        cpu1_s = time.perf_counter() - t0
        cpu["single_thread"] = {"value": nUpd * iters1 / cpu1_s, "unit": "vertex-updates/s", "cores": 1,
                                "sample": f"1 warm-up step x 1 iteration, then 1 step x {iters1} iterations of "
                                          f"the same workload, {cpu1_s:.2f}s"}
        log(f"[bench] cpu baseline {cpu['single_thread']['value']:.4e} vu/s on 1 thread")
        if adjoint is not None:
            adjoint["gradient_checks"] = adjoint_gradient_checks(args, adjoint["solver"])

    out = {
        "metric": "VBD vertex-updates/s + fwd+adjoint steps/s vs plain, 1 GPU, \u22641e-5 rel err",
        "value": value,
        "unit": "vertex-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain VBD, {args.grid}^3 jittered cube grid ({s.n_vertices} verts, "
                               f"{s.n_tets} tets), SNH per-tet log-uniform Y, h=1/60, {iters} iterations, "
                               f"bottom fixed", "colours": nColors, "updated_vertices": nUpd,
                   "incidences": nInc, "parallelism": "single GPU",
                   "sweep_shape": dict(shape, hip_graph=not args.no_graph)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None if traffic is None else traffic["hbm_bytes_per_launch"],
                     "traffic_source": traffic_note,
                     "traffic_gbs": None if traffic is None else traffic["hbm_bytes_per_launch"] / avg_launch_s / 1e9,
                     "traffic_frac": None if traffic is None
                     else traffic["hbm_bytes_per_launch"] / avg_launch_s / 1e9 / HBM_PEAK_GBS,
                     "kernel": "VbdSweep", "avg_launch_us": avg_launch_s * 1e6,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": cpu,
        "batched": batched,
        "variants": variants,
        "adjoint": adjoint,
        "batched_adjoint": batched_adjoint,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
