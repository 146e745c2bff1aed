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
            bounded sample (one step) using the host's cores.
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


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
    tune = batch.autotune(dt, sweeps=5)
    log(f"[bench] batch sweep shape: {batch.sweep_config()} ({min(r[-1] for r in tune):.1f} us/sweep)")
    for _ in range(2):
        batch.step(dt, iters, 1)
    batch.synchronize()
    steps = args.batch_steps
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.step(dt, iters, 1)
    batch.synchronize()
    tb = (time.perf_counter() - t0) / steps
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
    return {"workload": f"{count} jittered cube meshes, grid dims U{{4..24}} per axis, per-mesh Y, nu; "
                        f"h=1/60, {iters} iterations",
            "meshes": count, "vertices": int(batch.offsets[-1]), "updated_vertices": nUpd,
            "steps_per_s": 1 / tb, "per_mesh_loop_steps_per_s": 1 / tl, "speedup_vs_loop": tl / tb,
            "vertex_updates_per_s": nUpd * iters / tb, "steps": steps, "loop_steps": loop_steps,
            "sweep_shape": batch.sweep_config()}


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
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only run a plain device copy of this many MB (counter calibration) and exit")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the VBD path has no multi-GPU form in this build "
              "(see DESIGN.md for how it would shard); run with --gpus 1", file=sys.stderr)
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
    traffic = None
    pmc = os.path.join(ROOT, "profiles", "vbd_sweep_64cube_traffic.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    log(f"[bench] sweep kernel: {launches} launches, avg {avg_launch_s * 1e6:.2f} us, "
        f"{bytes_per_launch / 1e6:.2f} MB/launch -> {achieved:.1f} GB/s")

    batched = None if args.no_batched else bench_batched(args, torch)

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        ref = O.OracleVbd(s.X, s.E, aext=s.aext, rhoe=s.rhoe, mue=s.mue, lambdae=s.lambdae, dbc=s.dbc,
                          nthreads=threads)
        t0 = time.perf_counter()
        ref.step(dt, iters, 1)
        cpu_s = time.perf_counter() - t0
        cpu = {"value": nUpd * iters / cpu_s, "unit": "vertex-updates/s", "cores": threads, "kind": "port",
               "sample": f"1 step x {iters} iterations of the same 64^3 workload, FP64 oracle, "
                         f"{threads} threads, {cpu_s:.2f}s"}
        log(f"[bench] cpu baseline {cpu['value']:.4e} vu/s on {threads} threads")

    out = {
        "metric": "VBD vertex-updates/s (plain, 1 GPU, <=1e-5 rel err)",
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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "VbdSweep", "avg_launch_us": avg_launch_s * 1e6,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": cpu,
        "batched": batched,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
