"""Benchmark of the transport hot path on one MI355X.

Workload (one "step"): trace 4,194,304 rays -- half closest-hit, half any-hit
(occlusion) -- through the two-level BVH of the 262,144-triangle procedural
colonnade (30 x 12 x 15 m, 8 emissive quads); rays have origins uniform in
the scene box, directions uniform on the sphere and tmax log-uniform in
0.1-20 m.  Rays and hits stay resident in HBM; the timed region is K
back-to-back trace launches.

Prints one JSON line: interval Mrays/s (value), the roofline of the trace
kernel (HIP-event duration on the library's stream, algorithmic bytes from
the instrumented kernel's node/triangle/instance counts) and a CPU baseline
(the C restatement of the reference CPU algorithms -- SAH build_tree and the
reference traversal -- on a bounded ray sample, timed on this host's cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--rays", type=int, default=4 * 1024 * 1024)
    p.add_argument("--cpu-sample", type=int, default=1 << 20)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--waves-per-cu", type=int, default=0)
    p.add_argument("--regroup", type=int, default=0)
    p.add_argument("--gi-steps", type=int, default=5)
    p.add_argument("--no-gi", action="store_true")
    return p.parse_args()


def colonnade_gi_params(lux):
    """Full 5-cascade split trace on the 64 x 24 x 32 probe grid at 0.5 m
    (32 level-0 directions, interval ends 0.5 * 4^L, terminal tail)."""
    return lux.gi_params((-0.5, -0.5, -0.5), 0.5, (64, 24, 32), 5, base_phi=8, base_y=4,
                         interval_mode=0, interval_length=0.5, normal_bias=0.005,
                         emissive_samples=16, seed=0, terminal_tail=1)


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        return 2
    from bevy_penumbra_amd import lux, scenes
    from bevy_penumbra_amd.lux import DeviceBuffer

    scene, _ = scenes.colonnade()
    rays_np = scenes.microbench_rays(scene.bounds, n=args.rays)
    ctx = lux.Context(0)
    ctx.upload_scene(scene)
    t0 = time.perf_counter()
    ctx.build_bvh()
    ctx.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    ctx.set_trace_tuning(args.waves_per_cu, args.regroup)

    n = rays_np.shape[0]
    rays = DeviceBuffer.from_array(ctx, rays_np)
    hits = DeviceBuffer.empty(ctx, (n, 4), np.uint32)
    counters = DeviceBuffer.empty(ctx, 3, np.uint64)

    # algorithmic work of one launch (instrumented kernel, outside timing)
    ctx.trace_stats_device(rays, n, hits, counters)
    ctx.synchronize()
    nodes, tris, insts = (int(x) for x in counters.numpy().tolist())
    algo_bytes = 32 * n + 16 * n + 32 * nodes + 36 * tris + 48 * insts

    for _ in range(args.warmup):
        ctx.trace_device(rays, n, hits)
    ctx.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.trace_device(rays, n, hits)
        kernel_ms.append(ctx.last_kernel_ms())
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    avg_kernel_ms = float(np.mean(kernel_ms))
    mrays = n / (ms_per_step * 1e-3) / 1e6
    achieved = algo_bytes / (avg_kernel_ms * 1e-3) / 1e9

    gi = None
    if not args.no_gi:
        gp = colonnade_gi_params(lux)
        ctx.gi_configure(gp)
        n_int, n_probes, n_lights = ctx.gi_sizes()
        ctx.gi_update()
        ctx.synchronize()
        gi_ms = [ctx.gi_update(timed=True) for _ in range(args.gi_steps)]
        ctx.synchronize()
        gi = {
            "ms_per_gi_update": round(float(np.median(gi_ms)), 4),
            "intervals": n_int,
            "interval_mrays_s": round(n_int / (float(np.median(gi_ms)) * 1e-3) / 1e6, 3),
            "probes": n_probes,
            "emitter_triangles": n_lights,
            "grid": "64x24x32 at 0.5 m, 5 levels, 32 base directions, NEE 8 samples",
        }

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        threads = min(16, os.cpu_count() or 1)
        built = O.build_scene(scene, sah=True)
        m = min(args.cpu_sample, n)
        half = n // 2
        sample = np.concatenate([rays_np[: m // 2], rays_np[half: half + m // 2]])
        t0 = time.perf_counter()
        built.trace(sample, threads=threads)
        dt = time.perf_counter() - t0
        cpu = {
            "value": round(sample.shape[0] / dt / 1e6, 4),
            "unit": "Mrays/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{sample.shape[0]} rays of the same set (first {m // 2} closest-hit and "
                      f"first {m // 2} any-hit), SAH build_tree BVH, C restatement of the "
                      f"reference CPU traversal, {threads} threads",
        }

    out = {
        "metric": "interval Mrays/s + ms/GI update; surrogate vs plain trace at equal RMSE",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "traversal microbench: 262,144-triangle colonnade, 4,194,304 rays "
                        "(half closest-hit, half any-hit), Morton-30 LBVH",
            "rays": n,
            "triangles": scene.triangle_count,
            "bvh_build_ms": round(build_ms, 3),
            "kernel_ms": round(avg_kernel_ms, 4),
            "node_fetches_per_ray": round(nodes / n, 3),
            "tri_tests_per_ray": round(tris / n, 3),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": None,
        },
        "cpu_baseline": cpu,
        "gi_update": gi,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
