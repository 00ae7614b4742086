"""Benchmark of the transport hot path on one MI355X.

Workload (one "step"): trace 4,194,304 rays -- half closest-hit, half any-hit
(occlusion) -- through the two-level BVH of the 262,144-triangle procedural
colonnade (30 x 12 x 15 m, 8 emissive quads); rays have origins uniform in
the scene box, directions uniform on the sphere and tmax log-uniform in
0.1-20 m.  Rays and hits stay resident in HBM; the timed region is K
back-to-back trace launches.  Before it, the library's autotuner picks the
traversal launch configuration on these rays (results do not depend on it).

Prints one JSON line: interval Mrays/s (value), the roofline of the trace
kernel (HIP-event duration on the library's stream, algorithmic bytes from
the instrumented kernel's node/triangle/instance counts) and a CPU baseline
(the C restatement of the reference CPU algorithms -- SAH build_tree and the
reference traversal -- on a bounded ray sample, timed on this host's cores).
Further sections: the GI update (plain and surrogate-assisted at equal RMSE),
the dynamic 1,048,576-triangle frame (p50/p99) and 64 uneven scenes updated
as one batch captured in a HIP graph.
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
    p.add_argument("--no-autotune", action="store_true",
                   help="keep the context's default traversal configuration")
    p.add_argument("--gi-steps", type=int, default=5)
    p.add_argument("--no-gi", action="store_true")
    p.add_argument("--no-surrogate", action="store_true")
    p.add_argument("--train-frames", type=int, default=1600)
    p.add_argument("--surrogate-levels", default="2,1")
    p.add_argument("--reference-frames", type=int, default=64,
                   help="plain updates averaged into the converged reference of the surrogate section")
    p.add_argument("--reference-eval", type=int, default=8,
                   help="held-out frames scored against the converged reference")
    p.add_argument("--no-dynamic", action="store_true")
    p.add_argument("--dynamic-frames", type=int, default=120)
    p.add_argument("--dynamic-pretrain", type=int, default=400,
                   help="frames of 262,144 traced samples the level-1 surrogate of the dynamic frame trains on "
                        "before the timed frames")
    p.add_argument("--no-batch", action="store_true")
    p.add_argument("--batch-scenes", type=int, default=64)
    p.add_argument("--batch-train", type=int, default=400,
                   help="online training steps of each batch scene's surrogate")
    p.add_argument("--no-receivers", action="store_true")
    p.add_argument("--no-screen", action="store_true")
    p.add_argument("--blas", choices=("sah", "morton"), default="sah",
                   help="BLAS builder of the traversal scene: the reference's binned SAH built on the GPU "
                        "(LUX_BVH_SAH) or Morton-30 on the GPU")
    p.add_argument("--receiver-steps", type=int, default=10)
    p.add_argument("--pmc-workload", action="store_true",
                   help="run only the counter-profiling workload (1 GiB device copy x2, three "
                        "traversal launches, two GI updates; default launch configurations unless "
                        "--trace-config / --gi-config name others)")
    p.add_argument("--trace-config", default="",
                   help="traversal launch configuration of the counter workload: "
                        "waves_per_cu,lds_stack_depth,regroup_threshold,steps_per_refill")
    p.add_argument("--gi-config", default="",
                   help="GI stage launch configuration of the counter workload, same four numbers")
    return p.parse_args()


TRAFFIC_PROFILE = ROOT / "profiles" / "pmc_traffic_colonnade_trace_and_gi.md"


def measured_traffic(kernel: str, cfg: dict, gi: bool = False):
    """HBM-side bytes per launch of `kernel` from the committed counter profile,
    and the launch configuration they were measured with.  The profile's
    blocks each start with a `config` (traversal) and a `gi_config` (GI
    stages) line giving the launch configuration and BLAS builder they were
    measured with (`blas` absent means Morton-30).  The kernel's `traffic`
    line is taken from the first block whose configuration of that kind (gi)
    equals cfg, else from the first block with the same kernel and BLAS
    builder (the same LDS stack depth: the other launch parameters move the
    traffic by a few percent).  (None, None) when no block has the kernel."""
    if not TRAFFIC_PROFILE.exists():
        return None, None
    tag = "gi_config" if gi else "config"
    exact = fallback = None
    want = {}
    for l in TRAFFIC_PROFILE.read_text().splitlines():
        parts = l.split()
        if not parts:
            continue
        if parts[0] == tag:
            want = dict(kv.split("=") for kv in parts[1:])
            want.setdefault("blas", "morton")
        elif parts[0] == "traffic" and len(parts) == 4 and parts[1] == kernel and want:
            vals = dict(kv.split("=") for kv in parts[2:])
            got = (int(vals["fetch_bytes"]) + int(vals["write_bytes"]),
                   {k: (v if k == "blas" else int(v)) for k, v in want.items()})
            if exact is None and all(str(want.get(k)) == str(v) for k, v in cfg.items()):
                exact = got
            if fallback is None and str(want.get("blas")) == str(cfg.get("blas")):
                fallback = got
    return exact or fallback or (None, None)


def gi_roofline(work: dict, intervals: int, probes: int, gp, update_ms: float, stage_ms: dict, cfg: dict,
                args) -> dict:
    """Roofline of the GI update's dominant kernel, the primary walks
    (k_gi_primary): algorithmic bytes = per walk its 16-byte hit record plus
    32 B per node fetch, 36 B per triangle test and 48 B per instance visit
    (the walk generates its ray from the interval index, so no ray is read),
    over the primary stage's HIP-event duration.  The whole update's
    algorithmic bytes add the occlusion walks (32 B ray + 4 B visibility + the
    same traversal bytes), 32 B per interval deposit, 544 B per merged interval
    with children (J, beta; 32 child (I, support); out), 32 B per top-level
    interval and 1,088 B per packed level-0 probe."""
    prim = (16 * work["primary_walks"] + 32 * work["primary_nodes"] + 36 * work["primary_tris"]
            + 48 * work["primary_insts"])
    occl = (36 * work["occlusion_rays"] + 32 * work["occlusion_nodes"] + 36 * work["occlusion_tris"]
            + 48 * work["occlusion_insts"])
    dims = list(gp.dims)
    top = 1
    for a in range(3):
        d = dims[a]
        for _ in range(gp.levels - 1):
            d = (d + 1) // 2
        top *= d
    top *= (gp.base_phi << (gp.levels - 1)) * (gp.base_y << (gp.levels - 1))
    total = prim + occl + 32 * intervals + 544 * (intervals - top) + 32 * top + 1088 * probes
    achieved = prim / (stage_ms["primary"] * 1e-3) / 1e9
    kernel = "k_gi_primary<%d>" % cfg["lds_stack_depth"]
    traffic, traffic_cfg = measured_traffic(kernel, dict(cfg, blas=args.blas), gi=True)
    return {
        "kernel": kernel, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": traffic,
        "traffic_config": traffic_cfg,
        "algorithmic_bytes": prim,
        "occlusion_walks": {"kernel": "k_gi_shadow<%d>" % cfg["lds_stack_depth"], "algorithmic_bytes": occl,
                            "achieved": round(occl / (stage_ms["occlusion"] * 1e-3) / 1e9, 2)},
        "update": {"algorithmic_bytes": total, "achieved": round(total / (update_ms * 1e-3) / 1e9, 2),
                   "frac": round(total / (update_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)},
    }


def blas_mode(lux, args) -> int:
    return lux.LUX_BVH_SAH if args.blas == "sah" else lux.LUX_BVH_MORTON30


def cpu_baseline(args, rays_np) -> dict:
    """The oracle (C restatement of the reference CPU algorithms) timed on
    this host in a child process (oracle/cpu_baseline.py), built for this CPU
    (`make -C oracle native`: -O3 -march=native -ffp-contract=off); its hits
    of the first rays must equal the portable build's bit for bit.  Falls
    back to the portable build if the native one cannot be built."""
    import subprocess
    import tempfile
    from oracle import oracle as O
    tmp = Path(tempfile.mkdtemp(prefix="lux_cpu_"))
    env = dict(os.environ)
    build = "native (-O3 -march=native -ffp-contract=off)"
    r = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "native", f"NATIVE_DIR={tmp}"],
                       capture_output=True, text=True)
    if r.returncode == 0:
        env["LUX_ORACLE_LIB"] = str(tmp / "liblux_oracle.so")
    else:
        build = "portable (-O2; the native build failed)"
    m = min(args.cpu_sample, rays_np.shape[0])
    hits_path = tmp / "hits.npy"
    cmd = [sys.executable, str(ROOT / "oracle" / "cpu_baseline.py"), "--rays", str(m),
           "--rays-one-thread", str(max(m // 16, 2)), "--hits", str(hits_path)]
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600)
    if r.returncode != 0:
        return {"error": r.stderr[-400:]}
    res = json.loads(r.stdout.strip().splitlines()[-1])
    half = rays_np.shape[0] // 2
    sample = np.concatenate([rays_np[: m // 2], rays_np[half: half + m // 2]])[:65536]
    same = bool(np.array_equal(np.load(hits_path), O.build_scene(scenes_colonnade(), sah=True).trace(sample)))
    lscpu = ""
    try:
        lscpu = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
    except OSError:
        pass
    model = next((l.split(":", 1)[1].strip() for l in lscpu.splitlines() if l.startswith("Model name")), "")
    cpus = next((l.split(":", 1)[1].strip() for l in lscpu.splitlines() if l.startswith("CPU(s)")), "")
    tr, gc, g2 = res["traversal"], res["gi_cornell"], res["gi_colonnade"]
    return {
        "value": tr["all_cores"]["mrays_s"],
        "unit": "Mrays/s",
        "cores": tr["all_cores"]["threads"],
        "kind": "port",
        "sample": f"{tr['all_cores']['rays']} rays of the same set (first half closest-hit, first half any-hit), "
                  f"SAH build_tree BVH, C restatement of the reference CPU traversal, {build}, "
                  f"{tr['all_cores']['threads']} threads = every core this process may use "
                  f"(lscpu: {cpus} CPUs, {model})",
        "one_thread": tr["one_thread"],
        "native_hits_equal_portable": same,
        "gi_cornell": gc,
        "gi_colonnade": g2,
        "gi_colonnade_ms_per_update": g2["all_cores"]["ms_per_update_est"],
    }


def scenes_colonnade():
    from bevy_penumbra_amd import scenes
    return scenes.colonnade()[0]


def pmc_workload(lux, scenes, args) -> int:
    from bevy_penumbra_amd.lux import DeviceBuffer
    ctx = lux.Context(0)
    a = DeviceBuffer.empty(ctx, (1 << 28,), np.float32)
    b = DeviceBuffer.empty(ctx, (1 << 28,), np.float32)
    for _ in range(2):
        b.copy_from(a)
    ctx.synchronize()
    sc, _ = scenes.colonnade()
    rays_np = scenes.microbench_rays(sc.bounds)
    ctx.upload_scene(sc)
    ctx.build_bvh(blas_mode(lux, args))
    n = rays_np.shape[0]
    rays = DeviceBuffer.from_array(ctx, rays_np)
    hits = DeviceBuffer.empty(ctx, (n, 4), np.uint32)
    if args.trace_config:
        ctx.set_trace_config(*[int(v) for v in args.trace_config.split(",")])
    for _ in range(3):
        ctx.trace_device(rays, n, hits)
    ctx.gi_configure(colonnade_gi_params(lux))
    if args.gi_config:
        ctx.gi_set_trace_config(*[int(v) for v in args.gi_config.split(",")])
    for _ in range(2):
        ctx.gi_update()
    ctx.synchronize()
    print(json.dumps({"pmc_workload": "done", "trace_config": ctx.trace_config(),
                      "gi_trace_config": ctx.gi_trace_config()}), flush=True)
    ctx.close()
    return 0


def colonnade_gi_params(lux):
    """Full 5-cascade split trace on the 64 x 24 x 32 probe grid at 0.5 m
    (32 level-0 directions, interval ends 0.5 * 4^L, terminal tail)."""
    return lux.gi_params((-0.5, -0.5, -0.5), 0.5, (64, 24, 32), 5, base_phi=8, base_y=4,
                         interval_mode=0, interval_length=0.5, normal_bias=0.005,
                         emissive_samples=16, seed=0, terminal_tail=1)


def probe_rmse(a, b) -> float:
    """Relative RMSE of packed level-0 probe radiance (all texels, RGB)."""
    return float(np.sqrt(np.mean((a[..., :3].astype(np.float64) - b[..., :3]) ** 2) /
                         np.mean(b[..., :3].astype(np.float64) ** 2)))


def converged_reference(ctx, frames: int, first: int):
    """Mean packed level-0 probe radiance (RGB, f64) of `frames` plain updates
    seeded first .. first + frames - 1, and the relative RMS standard error of
    that mean (its own noise, which every RMSE against it carries in
    quadrature)."""
    acc = acc2 = None
    for f in range(first, first + frames):
        ctx.gi_set_frame(f, f)
        ctx.gi_update()
        p = ctx.gi_download_probes()[..., :3].astype(np.float64)
        acc = p if acc is None else acc + p
        acc2 = p * p if acc2 is None else acc2 + p * p
    mu = acc / frames
    var = np.maximum(acc2 / frames - mu * mu, 0.0) * frames / (frames - 1)
    return mu, float(np.sqrt(np.mean(var) / frames / np.mean(mu ** 2)))


def surrogate_section(ctx, lux, gp, args) -> dict:
    """Equal-RMSE comparison of the surrogate-assisted update with the plain
    update on the same grid.  Per replaced level S: online training over
    --train-frames frames (262,144 traced terminal samples of level S and one
    step per frame), fallback bound calibrated on held-out frame 998 (the
    largest predicted-error quantile whose probe RMSE against that frame's
    plain update is <= 2 %), evaluated on held-out frame 999 against its plain
    update.  Also reported: cascades truncated at S (levels 0 .. S-1 only,
    without a terminal tail: the same traced work without the network and
    without any far field), the noise floor (levels >= S taken from another
    frame's trace), and the error curve of training: the network's
    no-fallback update of held-out frame 997 against that frame's plain
    update after 1, 5, 10, 25, 50, 100 ... frames (outside the timed
    training).

    Second view ("converged"): every method's probe RMSE against the mean of
    --reference-frames plain updates (the expected radiance; the plain update
    equals the oracle frame by frame), on --reference-eval held-out frames,
    as the RMS over those frames.  There one frame of the plain update has
    its own sampling noise, and a method is at equal RMSE when its RMS is no
    higher than the plain update's."""
    import ctypes
    import time
    n = 262144
    buf = lux.DeviceBuffer.empty(ctx, (n, 12), np.float32)
    ctx.gi_configure(gp)
    refs, ms_plain = {}, []
    for f in (997, 998, 999):
        ctx.gi_set_frame(f, f)
        ms_plain.append(ctx.gi_update(timed=True))
        refs[f] = ctx.gi_download()["probes"]
    for _ in range(2):
        ms_plain.append(ctx.gi_update(timed=True))
    ms_plain = float(np.median(ms_plain))
    mu, mu_noise = converged_reference(ctx, args.reference_frames, 10000)
    eval_frames = list(range(20000, 20000 + args.reference_eval))

    def conv(run) -> dict:
        e, ms = [], []
        for f in eval_frames:
            ctx.gi_set_frame(f, f)
            ms.append(run())
            e.append(probe_rmse(ctx.gi_download_probes(), mu))
        return {"rmse_rms": round(float(np.sqrt(np.mean(np.square(e)))), 5),
                "rmse_frames": [round(x, 5) for x in e], "ms": round(float(np.median(ms)), 3)}

    conv_plain = conv(lambda: ctx.gi_update(timed=True))
    origin = np.array(gp.origin, np.float64)
    hi = origin + gp.spacing * np.array(gp.dims, np.float64)
    out = {"protocol": "probe RMSE vs the plain update of the same held-out frame (which equals "
                       "the oracle within the parity tests' bounds); bound calibrated on frame 998, "
                       "evaluated on frame 999",
           "plain_ms": round(ms_plain, 3), "rmse_target": 0.02, "levels": {},
           "converged": {"protocol": "probe RMSE against the mean of %d plain updates (seeds 10000+), "
                                     "RMS over held-out frames %d-%d" % (args.reference_frames, eval_frames[0],
                                                                        eval_frames[-1]),
                         "reference_noise": round(mu_noise, 5), "plain": conv_plain}}
    for level in [int(x) for x in args.surrogate_levels.split(",") if x]:
        res = {}
        tp = lux.GiParams()
        ctypes.memmove(ctypes.byref(tp), ctypes.byref(gp), ctypes.sizeof(gp))
        tp.levels = level
        tp.terminal_tail = 0
        ctx.gi_configure(tp)
        ctx.gi_set_frame(999, 999)
        t_ms = float(np.median([ctx.gi_update(timed=True) for _ in range(3)]))
        res["truncated"] = {"ms": round(t_ms, 3), "rmse": round(probe_rmse(ctx.gi_download()["probes"], refs[999]), 5)}
        conv_trunc = conv(lambda: ctx.gi_update(timed=True))
        ctx.gi_configure(gp)
        ctx.gi_set_frame(998, 998)
        ctx.gi_trace_levels(level, gp.levels - 1)
        ctx.gi_set_frame(999, 999)
        ctx.gi_trace_levels(0, level - 1)
        ctx.gi_merge()
        ctx.gi_pack()
        res["noise_floor_rmse"] = round(probe_rmse(ctx.gi_download()["probes"], refs[999]), 5)
        s = lux.Surrogate(ctx, lux.SurrogateParams.default(origin, hi))
        t_samp = t_train = 0.0
        loss, curve = [], []
        checkpoints = {1, 5, 10, 25, 50, 100, 200, 400, 800, 1600, args.train_frames}
        for f in range(args.train_frames):
            ctx.gi_set_frame(f, f)
            ctx.synchronize()
            t0 = time.perf_counter()
            ctx.gi_surrogate_samples(level, n, f, lux.LUX_SAMPLES_TRACE, buf)
            ctx.synchronize()
            t1 = time.perf_counter()
            loss.append(s.train_step(buf, n))
            t2 = time.perf_counter()
            t_samp += t1 - t0
            t_train += t2 - t1
            if f + 1 in checkpoints:
                ctx.gi_set_frame(997, 997)
                ctx.gi_update_surrogate(s, level, float("inf"))
                curve.append([f + 1, round(probe_rmse(ctx.gi_download()["probes"], refs[997]), 5)])
        res["train"] = {"frames": args.train_frames, "samples_per_frame": n,
                        "heldout_rmse_curve": curve,
                        "sample_trace_ms": round(t_samp / args.train_frames * 1e3, 3),
                        "train_step_ms": round(t_train / args.train_frames * 1e3, 3),
                        "loss_first_last10": [round(float(np.mean(loss[:10])), 4), round(float(np.mean(loss[-10:])), 4)]}
        ctx.gi_set_frame(998, 998)
        st = ctx.gi_update_surrogate(s, level, float("inf"))
        err = ctx.gi_surrogate_probe_error(st.probes)
        chosen, calib = None, []
        for q in (1.0, 0.9, 0.75, 0.5, 0.25, 0.0):
            bound = float("inf") if q >= 1.0 else (-1.0 if q <= 0.0 else float(np.quantile(err, q)))
            st = ctx.gi_update_surrogate(s, level, bound)
            e = probe_rmse(ctx.gi_download()["probes"], refs[998])
            calib.append({"quantile": q, "fallback_probes": st.fallback_probes, "rmse": round(e, 5)})
            if chosen is None and e <= 0.02:
                chosen = bound
        if chosen is None:
            chosen = -1.0
        ctx.gi_set_frame(999, 999)
        nf = [ctx.gi_update_surrogate(s, level, float("inf")) for _ in range(3)]
        res["no_fallback"] = {"rmse": round(probe_rmse(ctx.gi_download()["probes"], refs[999]), 5),
                              "ms": round(float(np.median([x.ms_total for x in nf])), 3)}
        sts = [ctx.gi_update_surrogate(s, level, chosen) for _ in range(3)]
        e = probe_rmse(ctx.gi_download()["probes"], refs[999])
        ms = float(np.median([x.ms_total for x in sts]))
        st = sts[-1]
        res.update({
            "calibration": calib,
            "bound": chosen if np.isfinite(chosen) else "inf",
            "eval": {"rmse": round(e, 5), "ms": round(ms, 3), "fallback_probes": st.fallback_probes,
                     "probes": st.probes, "ms_trace": round(st.ms_trace, 3), "ms_infer": round(st.ms_infer, 3),
                     "ms_fallback": round(st.ms_fallback, 3), "ms_merge_pack": round(st.ms_merge_pack, 3)},
            "speedup": round(ms_plain / ms, 3),
            "speedup_with_online_training": round(
                ms_plain / (ms + res["train"]["sample_trace_ms"] + res["train"]["train_step_ms"]), 3),
        })
        cv = {"truncated": conv_trunc,
              "no_fallback": conv(lambda: ctx.gi_update_surrogate(s, level, float("inf")).ms_total),
              "calibrated": conv(lambda: ctx.gi_update_surrogate(s, level, chosen).ms_total)}
        equal = [cv[k]["ms"] for k in ("no_fallback", "calibrated")
                 if cv[k]["rmse_rms"] <= conv_plain["rmse_rms"]]
        cv["equal_rmse_speedup"] = round(conv_plain["ms"] / min(equal), 3) if equal else None
        cv["truncated_at_equal_rmse"] = conv_trunc["rmse_rms"] <= conv_plain["rmse_rms"]
        res["converged"] = cv
        s.close()
        out["levels"][f"replace_from_{level}"] = res
    # every method of the converged view by cost: the cheapest at or below
    # the plain update's RMSE is the equal-RMSE choice
    methods = [{"method": "plain", "ms": conv_plain["ms"], "rmse_rms": conv_plain["rmse_rms"]}]
    for key, res in out["levels"].items():
        for m in ("truncated", "no_fallback", "calibrated"):
            cv = res["converged"][m]
            methods.append({"method": f"{key} {m}", "ms": cv["ms"], "rmse_rms": cv["rmse_rms"]})
    methods.sort(key=lambda m: m["ms"])
    out["converged"]["by_cost"] = methods
    ok = [m for m in methods if m["rmse_rms"] <= conv_plain["rmse_rms"]]
    out["converged"]["cheapest_at_plain_rmse"] = dict(ok[0], speedup=round(conv_plain["ms"] / ok[0]["ms"], 3))
    ctx.gi_configure(gp)
    return out


def dynamic_section(lux, scenes, args) -> dict:
    """Dynamic large frame: 1,048,576 triangles in 256 rigid instances under
    seeded per-frame motion (static meshes: their BLASes from the chosen
    builder, built once).  One frame = new transforms + Morton-30 TLAS
    rebuild (or, in the refit run, a refit of the first frame's TLAS,
    LUX_BVH_REFIT), a 1920 x 1080 camera pass producing the receivers, the cascade
    update (plain; or levels 0-1 traced with the surrogate answering level 2,
    trained online on 65,536 traced samples per frame; or level 0 traced with
    a surrogate answering level 1, trained beforehand on --dynamic-pretrain
    other frames of the animation and online as above; both scored every
    10th frame against that frame's plain update, outside the timed frame),
    and the world resolve of all 2,073,600 receivers.  Wall time per frame (host waits for
    the frame's last kernel), p50/p99 over the frames.  With the temporal
    history on (blends 0.88 / 0.98, persistent), the share of intervals whose
    history each frame reuses -- every instance moving, and only every 32nd
    (8 instances) moving."""
    import time
    from bevy_penumbra_amd.lux import DeviceBuffer
    sc, motion = scenes.dynamic()
    ctx = lux.Context(0)
    ctx.upload_scene(sc)
    t0 = time.perf_counter()
    ctx.build_bvh()
    ctx.synchronize()
    first_ms = (time.perf_counter() - t0) * 1e3
    full = []
    for _ in range(3):
        t0 = time.perf_counter()
        ctx.build_bvh(lux.LUX_BVH_MORTON30 | lux.LUX_BVH_FORCE_BLAS)
        ctx.synchronize()
        full.append((time.perf_counter() - t0) * 1e3)
    # the frames: static meshes keep the BLASes of the chosen builder (built
    # once here, outside the frames); only the TLAS follows the motion
    bmode = blas_mode(lux, args)
    t0 = time.perf_counter()
    ctx.build_bvh(bmode)
    ctx.synchronize()
    blas_ms = (time.perf_counter() - t0) * 1e3
    gp = colonnade_gi_params(lux)
    ctx.gi_configure(gp)
    if not args.no_autotune:
        res_tune = ctx.autotune_gi(repeats=1)
    else:
        res_tune = None
    W, H = 1920, 1080
    recv = DeviceBuffer.empty(ctx, (W * H, 6), np.float32)
    out = DeviceBuffer.empty(ctx, (W * H, 4), np.float32)
    cam = scenes.camera_rays((15.0, 6.0, -9.0), (15.0, 6.0, 7.5), W, H)
    origin = np.array(gp.origin, np.float64)
    hi = origin + gp.spacing * np.array(gp.dims, np.float64)
    sur = lux.Surrogate(ctx, lux.SurrogateParams.default(origin, hi))
    ns = 65536
    buf = DeviceBuffer.empty(ctx, (ns, 12), np.float32)
    res = {"workload": "1,048,576 triangles (256 meshes x 4,096) in 256 rigid instances, 64x24x32 "
                       "probe grid at 0.5 m, 5 levels, 1920x1080 receivers",
           "first_build_ms": round(first_ms, 3),
           "gi_trace_config": ctx.gi_trace_config(),
           "full_rebuild_ms": round(float(np.median(full)), 3),
           "frame_blas": ("binned SAH on the GPU" if args.blas == "sah" else "Morton-30") +
                         " (%.1f ms, once)" % blas_ms}
    inv0, fwd0 = scenes.dynamic_transforms(motion, 0)
    n_int = ctx.gi_sizes()[0]
    # the history variant with 8 moving instances runs on the same geometry
    # with emissive bands on every 32nd mesh only (held still): the moving
    # instances (every 32nd, offset 16) carry no emitter
    sc8, _ = scenes.dynamic(emissive_every=32)
    ctx8 = lux.Context(0)
    ctx8.upload_scene(sc8)
    ctx8.build_bvh(bmode)
    ctx8.gi_configure(gp)
    if res_tune:
        ctx8.gi_set_trace_config(**{k: res_tune[k] for k in ("waves_per_cu", "lds_stack_depth",
                                                               "regroup_threshold", "steps_per_refill")})
    sur1 = lux.Surrogate(ctx, lux.SurrogateParams.default(origin, hi))
    for mode in ("plain", "plain_refit", "surrogate", "surrogate_level1", "history", "history_8_moving"):
        times, tlas, reused, errs, noise = [], [], [], [], []
        c = ctx8 if mode == "history_8_moving" else ctx
        if mode.startswith("history"):
            c.gi_set_history(0.88, 0.98, 24, 1)
        sur_mode = {"surrogate": (sur, 2), "surrogate_level1": (sur1, 1)}.get(mode)
        if mode == "surrogate_level1":
            # trained beforehand on other frames of the same animation (seeds
            # and frames 10000 + k), as by the frames a running application
            # rendered before these; then online on each frame as it goes
            big = DeviceBuffer.empty(ctx, (262144, 12), np.float32)
            t_pre = time.perf_counter()
            for k in range(args.dynamic_pretrain):
                inv, fwd = scenes.dynamic_transforms(motion, 10000 + k)
                ctx.set_transforms(inv, fwd)
                ctx.build_bvh(bmode)
                ctx.gi_set_frame(10000 + k, 10000 + k)
                ctx.gi_surrogate_samples(1, 262144, 10000 + k, lux.LUX_SAMPLES_TRACE, big)
                sur1.train_step(big, 262144, want_loss=False)
            ctx.synchronize()
            res["surrogate_level1_pretrain"] = {"frames": args.dynamic_pretrain, "samples_per_frame": 262144,
                                                "s": round(time.perf_counter() - t_pre, 2)}
            big.free()
        for f in range(-5, args.dynamic_frames):
            fr = max(f, 0)
            t0 = time.perf_counter()
            inv, fwd = scenes.dynamic_transforms(motion, fr)
            if mode == "history_8_moving":
                # only every 32nd instance (offset 16) moves; the others stay
                # where frame 0 put them
                keep = np.arange(inv.shape[0]) % 32 != 16
                inv[keep], fwd[keep] = inv0[keep], fwd0[keep]
            c.set_transforms(inv, fwd)
            tb = time.perf_counter()
            c.build_bvh(bmode | (lux.LUX_BVH_REFIT if mode == "plain_refit" else 0))
            c.synchronize()
            te = time.perf_counter()
            c.camera_receivers(cam, W, H, recv)
            c.gi_set_frame(fr, fr)
            if sur_mode is None:
                c.gi_update()
            else:
                s_net, s_level = sur_mode
                ctx.gi_update_surrogate(s_net, s_level, float("inf"), stats=False)
                ctx.gi_surrogate_samples(s_level, ns, fr, lux.LUX_SAMPLES_TRACE, buf)
                s_net.train_step(buf, ns, want_loss=False)
            c.world_resolve_device(recv, W * H, out)
            c.synchronize()
            if f >= 0:
                times.append((time.perf_counter() - t0) * 1e3)
                tlas.append((te - tb) * 1e3)
                if mode.startswith("history"):
                    reused.append(c.gi_history_reused() / n_int)
                if sur_mode is not None and f % 10 == 0:
                    # outside the timed frame: this frame's plain update as
                    # the reference of the surrogate-assisted probes
                    got = ctx.gi_download_probes()
                    ctx.gi_update()
                    ref = ctx.gi_download_probes()
                    errs.append(probe_rmse(got, ref))
                    # the plain update's own noise: another seed of the same
                    # frame differs from it by about sqrt(2) of it
                    ctx.gi_set_frame(fr, fr + 50000)
                    ctx.gi_update()
                    noise.append(probe_rmse(ctx.gi_download_probes(), ref) / np.sqrt(2.0))
        r = out.numpy()
        res[mode] = {"frames": len(times), "p50_ms": round(float(np.percentile(times, 50)), 3),
                     "p99_ms": round(float(np.percentile(times, 99)), 3),
                     "tlas_p50_ms": round(float(np.percentile(tlas, 50)), 3),
                     "supported_receivers": int((r[:, 3] >= 1.0).sum())}
        if reused:
            res[mode]["history_reused_median"] = round(float(np.median(reused)), 4)
            res[mode]["history_reused_min"] = round(float(np.min(reused)), 4)
            c.gi_history_off()
        if errs:
            res[mode]["probe_rmse_vs_plain_median"] = round(float(np.median(errs)), 5)
            res[mode]["probe_rmse_vs_plain_max"] = round(float(np.max(errs)), 5)
            res[mode]["probe_rmse_frames"] = len(errs)
            res[mode]["plain_noise_rmse_median"] = round(float(np.median(noise)), 5)
    sur.close()
    sur1.close()
    ctx8.close()
    ctx.close()
    return res


def batch_section(lux, scenes, args) -> dict:
    """Batched uneven scenes: --batch-scenes clutter rooms (triangles
    log-uniform 1k-200k, cubic probe grids log-uniform 8^3-48^3, 3-5 levels,
    16 base directions, 4 emissive samples, seeded), updated as ONE batch: one
    persistent closest-hit launch and one persistent occlusion launch over
    every scene's intervals, then every scene's merge and pack, captured as a
    HIP graph and replayed.  Reported: graph replay and eager ms of the plain
    batch; the same scenes updated by lux_gi_update each, issued back to back
    (every context has its own stream, so these run concurrently); the batch
    with every scene's levels >= 2 predicted by its own surrogate (trained
    online for --batch-train steps of 32,768 traced samples, no fallback),
    each scene's probe RMSE against its plain update of the same frame, and
    for context the cascades truncated at level 2 and the noise floor."""
    import ctypes
    import time
    from bevy_penumbra_amd.lux import DeviceBuffer
    specs = scenes.batch_set(seed=0, count=args.batch_scenes)
    ctxs = []
    t0 = time.perf_counter()
    for s in specs:
        sc, _ = scenes.clutter_room(s["seed"], s["triangles"], s["size"])
        c = lux.Context(0)
        c.upload_scene(sc)
        c.build_bvh(blas_mode(lux, args))
        sp = s["spacing"]
        c.gi_configure(lux.gi_params((0.5 * sp, 0.5 * sp, 0.5 * sp), sp, s["dims"], s["levels"], base_phi=4,
                                     base_y=4, interval_length=sp, emissive_samples=4, terminal_tail=1))
        ctxs.append(c)
    for c in ctxs:
        c.synchronize()
    setup_s = time.perf_counter() - t0
    b = lux.Batch(ctxs)
    sizes = b.sizes()
    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        tune = b.autotune(repeats=1)
        tune["tuning_s"] = round(time.perf_counter() - t0, 2)
        tune["ms"] = round(tune["ms"], 3)
    frame = 1
    b.set_frame(frame, frame)
    b.update()
    eager = float(np.median([b.update(timed=True) for _ in range(3)]))
    b.capture()
    b.launch()
    graph = float(np.median([b.launch(timed=True) for _ in range(5)]))
    b.synchronize()
    plain = [c.gi_download()["probes"] for c in ctxs]
    seq = []
    for _ in range(3):
        for c in ctxs:
            c.gi_set_frame(frame, frame)
        t0 = time.perf_counter()
        for c in ctxs:
            c.gi_update()
        for c in ctxs:
            c.synchronize()
        seq.append((time.perf_counter() - t0) * 1e3)
    res = {
        "workload": f"{len(specs)} clutter rooms, triangles log-uniform 1k-200k "
                    f"({sum(s['triangles'] for s in specs)} total), cubic probe grids log-uniform 8^3-48^3, "
                    "3-5 levels, 16 base directions, 4 emissive samples",
        "scenes": sizes["scenes"], "intervals": sizes["intervals"], "items": sizes["items"],
        "probes": sizes["probes"], "setup_s": round(setup_s, 2),
        "plain": {"graph_ms": round(graph, 3), "eager_ms": round(eager, 3),
                  "per_scene_updates_ms": round(float(np.median(seq)), 3),
                  "interval_mrays_s": round(sizes["items"] / (graph * 1e-3) / 1e6, 2)},
        "config": b.config(), "autotune": tune,
    }
    b.close()
    level = 2
    # context for the surrogate's error: levels >= 2 dropped (truncated) and
    # levels >= 2 taken from another frame's trace (Monte Carlo noise floor)
    floor = []
    for c, p0 in zip(ctxs, plain):
        c.gi_set_frame(frame + 1, frame + 1)
        c.gi_trace_levels(level, c._gi.levels - 1)
        c.gi_set_frame(frame, frame)
        c.gi_trace_levels(0, level - 1)
        c.gi_merge()
        c.gi_pack()
        floor.append(probe_rmse(c.gi_download()["probes"], p0))
    saved = []
    for c in ctxs:
        tp = lux.GiParams()
        ctypes.memmove(ctypes.byref(tp), ctypes.byref(c._gi), ctypes.sizeof(tp))
        saved.append(c._gi)
        tp.levels = level
        tp.terminal_tail = 0
        c.gi_configure(tp)
    bt = lux.Batch(ctxs)
    bt.set_config(**res["config"])
    bt.set_frame(frame, frame)
    bt.update()
    t_ms = float(np.median([bt.update(timed=True) for _ in range(3)]))
    bt.synchronize()
    trunc = [probe_rmse(c.gi_download()["probes"], p0) for c, p0 in zip(ctxs, plain)]
    bt.close()
    for c, gp in zip(ctxs, saved):
        c.gi_configure(gp)
    surs, bufs = [], []
    ns = 32768
    for s, c in zip(specs, ctxs):
        L = s["size"]
        surs.append(lux.Surrogate(c, lux.SurrogateParams.default((0.0, 0.0, 0.0), (L, L, L))))
        bufs.append(DeviceBuffer.empty(c, (ns, 12), np.float32))
    t0 = time.perf_counter()
    for f in range(args.batch_train):
        for c, sur, buf in zip(ctxs, surs, bufs):
            c.gi_set_frame(100 + f, 100 + f)
            c.gi_surrogate_samples(level, ns, f, lux.LUX_SAMPLES_TRACE, buf)
            sur.train_step(buf, ns, want_loss=False)
    for c in ctxs:
        c.synchronize()
    train_s = time.perf_counter() - t0
    bs = lux.Batch(ctxs, surrogates={i: (surs[i], level) for i in range(len(ctxs))})
    bs.set_config(**res["config"])
    bs.set_frame(frame, frame)
    bs.update()
    bs.capture()
    bs.launch()
    sgraph = float(np.median([bs.launch(timed=True) for _ in range(5)]))
    sz2 = bs.sizes()
    bs.synchronize()
    rmse = np.array([probe_rmse(c.gi_download()["probes"], p0) for c, p0 in zip(ctxs, plain)])
    res["surrogate"] = {
        "level": level, "train_steps": args.batch_train, "samples_per_step": ns, "train_s": round(train_s, 2),
        "graph_ms": round(sgraph, 3), "traced_items": sz2["items"],
        "speedup_vs_plain_graph": round(graph / sgraph, 3),
        "probe_rmse_median": round(float(np.median(rmse)), 5),
        "probe_rmse_max": round(float(np.max(rmse)), 5),
        "scenes_within_2pct": int((rmse <= 0.02).sum()),
        "truncated": {"ms": round(t_ms, 3), "probe_rmse_median": round(float(np.median(trunc)), 5),
                      "probe_rmse_max": round(float(np.max(trunc)), 5)},
        "noise_floor_probe_rmse_median": round(float(np.median(floor)), 5),
    }
    bs.close()
    for sur in surs:
        sur.close()
    for c in ctxs:
        c.close()
    return res


def receivers_section(lux, scenes, args) -> dict:
    """Receiver-driven updates (seed / rank / parents / prefix /
    generateDirections, receiver trace, reduce, then merge and pack of each
    LOD grid) on the colonnade: 1920 x 1080 camera receivers (2,073,600) over
    a three-LOD pyramid at 0.5 / 1 / 2 m covering the scene, every receiver
    tracing, and the default component's update budget with the fair window
    advancing -- without history, with the temporal history (warm), and
    cold-priority (the history reset before each update)."""
    from bevy_penumbra_amd.lux import DeviceBuffer
    scene, _ = scenes.colonnade()
    # grid k (LOD k): spacing and interval length 0.5 * 2^k m, its origin on its
    # coarsest level's cells (8 * 2^k m, as the world-anchored keys require),
    # covering the scene's triangles (x -1.1 .. 31.1, z -1.3 .. 16.2)
    lo_t, hi_t = np.array([-1.2, 0.0, -1.4]), np.array([31.2, 12.0, 16.3])
    grids = []
    for k in range(3):
        sk, top = 0.5 * 2 ** k, 0.5 * 2 ** k * 16
        org = np.floor(lo_t / top) * top
        org[1] = 0.0
        dims = [int(np.ceil((hi_t[a] - org[a]) / sk)) for a in range(3)]
        grids.append(lux.gi_params(tuple(org), sk, tuple(dims), 5, base_phi=8, base_y=4, interval_mode=0,
                                   interval_length=0.5 * 2 ** k, normal_bias=0.005, emissive_samples=16, seed=0,
                                   terminal_tail=1))
    p0 = grids[0]
    ctxs = []
    for k in range(3):
        c = lux.Context(0)
        c.upload_scene(scene)
        c.build_bvh(blas_mode(lux, args))
        c.gi_configure(grids[k])
        ctxs.append(c)
    # the clutter fills the volume (every interior eye sits inside a box): the
    # camera looks along the rows from 12 m outside the x = 0 end; 92 target
    # pixels put the visible receivers in LOD 0, the overlap and LOD 1
    W, H = 1920, 1080
    cam = scenes.camera_rays((-12.0, 6.0, 7.5), (15.0, 6.0, 7.5), W, H, 50.0)
    recv = DeviceBuffer.empty(ctxs[0], (W * H, 6), np.float32)
    ctxs[0].camera_receivers(cam, W, H, recv)
    ctxs[0].synchronize()
    f = 1.0 / np.tan(np.radians(25.0))
    up = cam[9:12] / np.linalg.norm(cam[9:12])
    terms = dict(camera=tuple(cam[:3]), probe_spacing=0.5, clip_row1=tuple(up * f), viewport_height=H,
                 clip_row3=tuple(cam[3:6]), target_pixels=92)
    view = lux.LodView.make(**terms)
    rc = lux.RadianceCascades.default()
    budget = rc.update_budget()
    out = {"receivers": W * H,
           "grids": ["%dx%dx%d at %.1f m from (%g, %g, %g)" % (tuple(g.dims) + (g.spacing,) + tuple(g.origin))
                     for g in grids],
           "levels": 5, "base_directions": 32, "intervals": [int(c.gi_sizes()[0]) for c in ctxs]}
    # with the temporal history on (persistent, 24 frames): warm updates, and
    # cold-priority ones (the history reset before each)
    for name, b, hist in (("all_receivers", 0, None), ("default_budget", budget, None),
                          ("default_budget_history", budget, "warm"), ("cold_default_budget", budget, "cold")):
        for c in ctxs:
            if hist:
                c.gi_set_history(0.88, 0.98, 24, 1)
            else:
                c.gi_history_off()
        sched = lux.CandidateSchedule()
        ms, wall = [], []
        st = None
        for i in range(args.receiver_steps + 2):
            cursor, quota = sched.phase_for(W, H, W * H, b)
            rp = lux.receiver_params(W, H, b, cursor, quota, frame=i)
            if hist == "cold":
                for c in ctxs:
                    c.gi_reset_history()
            t0 = time.perf_counter()
            st = lux.update_receivers(ctxs, view, recv, rp)
            for c in ctxs:
                c.synchronize()
            if i >= 2:
                wall.append((time.perf_counter() - t0) * 1e3)
                ms.append(st.ms)
            assert st.cold == (hist == "cold" or (hist == "warm" and i == 0))
            sched.advance_after_successful_update(not st.cold)
        med = float(np.median(ms))
        sel_view = lux.receivers_download(ctxs[0], W * H, p0.levels, directory=False)
        seeded = (sel_view["flags"][: W * H * p0.levels: p0.levels] & 1) == 1    # variant 0, level 0
        lod_hist = np.bincount(sel_view["lods"][0][seeded], minlength=3)[:4]
        out[name] = {
            "budget": b,
            "ms_per_update": round(med, 4),
            "wall_ms_per_update": round(float(np.median(wall)), 4),
            "rays": st.rays,
            "high_lod_rays": int((sel_view["sequence"][1] != 0xFFFFFFFF).sum()),
            "receiver_mrays_s": round(st.rays / (med * 1e-3) / 1e6, 3),
            "eligible_receivers": st.receivers,
            "receivers_by_low_lod": lod_hist.tolist(),
            "tickets": st.tickets,
            "deposits": st.deposits,
            "outside": st.outside,
            "unsupported": st.unsupported,
            "cold": st.cold,
            "cold_nominees": st.nominees,
        }
    for c in ctxs:
        c.gi_history_off()
    if not args.no_cpu:
        from oracle import oracle as O
        rec = recv.numpy()
        cpu_budget = 16384
        rp = O.receiver_params(W, H, cpu_budget, 0, cpu_budget, frame=0)
        oview = O.lod_view(**terms)
        built = O.build_scene(scene)
        lights = O.build_lights(scene)
        ogrids = []
        for q in grids:
            g = O.GiParams()
            for fname, _ in O.GiParams._fields_:
                setattr(g, fname, getattr(q, fname))
            ogrids.append(g)
        t0 = time.perf_counter()
        keys, flags, lods = O.rx_seed(rec, rp, oview, p0.spacing, p0.levels)
        sel = O.rx_select(W * H, p0.levels, keys, flags)
        t1 = time.perf_counter()
        _, _, ost = O.rx_trace(built, lights, scene.materials, ogrids, rec, lods, sel["sequence"], sel["tickets"], 0)
        t2 = time.perf_counter()
        out["cpu_baseline"] = {
            "select_ms": round((t1 - t0) * 1e3, 1),
            "trace_ms": round((t2 - t1) * 1e3, 1),
            "receiver_mrays_s": round(int(ost[0]) / (t2 - t1) / 1e6, 4),
            "cores": 1,
            "kind": "port",
            "sample": f"one update of budget {cpu_budget} over the same 2,073,600 receivers: C restatement "
                      "of seed/rank/parents/prefix/generateDirections and of the receiver trace, deposit "
                      "and reduce (merge and pack not included), one thread",
        }
    for c in ctxs:
        c.close()
    return out


def screen_section(lux, args) -> dict:
    """The screen passes after the resolve at 1920 x 1080 (resolve scale 1:
    2,073,600 texels): one temporal pass and three ping-ponged denoise
    passes over a synthetic frame -- two depth regions, two raster
    identities, a few tilted normals, 80 % supported samples, previous
    history with a small camera shift (fractional reprojection).  Bytes per
    texel are the algorithmic reads and writes: temporal 196 (current,
    background, depth, G-buffer and raster words, previous irradiance,
    metadata, receiver and background; four float4 written), a denoise pass
    84 (irradiance, moments, receiver words, depth; two float4 written)."""
    from bevy_penumbra_amd.lux import DeviceBuffer
    ctx = lux.Context(0)
    rng = np.random.default_rng(1)
    W, H = 1920, 1080
    P = lux.ScreenParams()
    for m in ("world_from_clip", "previous_clip_from_world", "previous_world_from_clip"):
        getattr(P, m)[:] = np.eye(4, dtype=np.float32).ravel()
    P.previous_clip_from_world[3], P.previous_world_from_clip[3] = 0.0013, -0.0013
    P.viewport[:] = [0, 0, W, H]
    P.target[:] = [W, H]
    P.size[:] = [W // 8, H // 8]
    P.resolve_scale = 1
    P.persistent = 1
    P.probe_spacing, P.normal_bias = 0.25, 0.005
    P.history_blend, P.history_frames = 0.9, 24.0
    P.stability[:] = [1.0, 1.0, 0.0, 0.0]
    RW, RH = P.extent()
    yy, xx = np.mgrid[0:H, 0:W]
    depth = np.where(xx + yy < 1500, 0.4, 0.7).astype(np.float32)
    raster = np.zeros((H, W, 4), np.uint32)
    raster[..., 0] = np.where(xx < 960, 0x3F800000, 0x40400000)
    raster[..., 1] = 0x40000000
    raster[..., 2] = np.where(rng.random((H, W)) < 0.05, 0x80000000 | 19384 | (18384 << 15), 0)
    gbuffer = np.zeros((H, W, 4), np.uint32)
    gbuffer[..., 3] = 0x0080_0800
    f4 = lambda: np.zeros((RH, RW, 4), np.float32)
    cur, cbg, prev, pmeta, precv, pbg = f4(), f4(), f4(), f4(), f4(), f4()
    cur[..., :3] = rng.uniform(0, 4, (RH, RW, 3))
    cur[..., 3] = np.where(rng.random((RH, RW)) < 0.8, 1.25, -1.0)
    cbg[..., :3], cbg[..., 3] = 0.5, 1.25
    prev[..., :3], prev[..., 3] = rng.uniform(0, 4, (RH, RW, 3)), depth[:RH, :RW]
    pmeta[..., :2], pmeta[..., 2], pmeta[..., 3] = 0.5, 8.0, 0.05
    precv[..., 0] = np.where(xx[:RH, :RW] < 960, 1.0, 3.0)
    precv[..., 1], precv[..., 2], precv[..., 3] = 2.0, 8_390_656.0, 65_537.0
    pbg[..., :3], pbg[..., 3] = 0.5, 8.0
    up = lambda a: DeviceBuffer.from_array(ctx, np.ascontiguousarray(a))
    d_in = [up(cur), up(cbg), up(depth), up(gbuffer), up(raster), None, up(prev), up(pmeta), up(precv), up(pbg)]
    d_out = [DeviceBuffer.empty(ctx, (RH, RW, 4), np.float32) for _ in range(4)]
    ping = [DeviceBuffer.empty(ctx, (RH, RW, 4), np.float32) for _ in range(4)]
    steps = max(args.receiver_steps, 3)

    def temporal():
        lux.screen_temporal(ctx, P, *d_in, *d_out)

    def denoise3():
        src_i, src_r = d_out[0], d_out[2]
        for k in range(3):
            dst_i, dst_r = ping[2 * (k % 2)], ping[2 * (k % 2) + 1]
            lux.screen_denoise(ctx, P, k, src_i, d_out[1], src_r, d_in[2], dst_i, dst_r)
            src_i, src_r = dst_i, dst_r

    res = {"texels": RW * RH, "resolve_scale": 1, "denoise_passes": 3}
    for name, fn, per_texel in (("temporal", temporal, 196), ("denoise_3_passes", denoise3, 3 * 84)):
        fn()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        res[name] = {"ms": round(ms, 4), "algorithmic_gb_s": round(per_texel * RW * RH / (ms * 1e-3) / 1e9, 1)}
    lit = d_out[0].numpy()[..., 3]
    res["temporal_lit_fraction"] = round(float((lit > 0).mean()), 4)
    ctx.close()
    return res


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        return 2
    from bevy_penumbra_amd import lux, scenes
    from bevy_penumbra_amd.lux import DeviceBuffer

    if args.pmc_workload:
        return pmc_workload(lux, scenes, args)
    scene, _ = scenes.colonnade()
    rays_np = scenes.microbench_rays(scene.bounds, n=args.rays)
    ctx = lux.Context(0)
    ctx.upload_scene(scene)
    t0 = time.perf_counter()
    ctx.build_bvh(blas_mode(lux, args))
    ctx.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    n = rays_np.shape[0]
    rays = DeviceBuffer.from_array(ctx, rays_np)
    hits = DeviceBuffer.empty(ctx, (n, 4), np.uint32)
    counters = DeviceBuffer.empty(ctx, 3, np.uint64)
    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        tune = ctx.autotune_trace(rays, n, hits, repeats=3)
        tune["tuning_s"] = round(time.perf_counter() - t0, 2)
        tune["ms"] = round(tune["ms"], 4)
    trace_cfg = ctx.trace_config()

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

    traffic, traffic_cfg = measured_traffic("k_trace_regroup<%d>" % trace_cfg["lds_stack_depth"],
                                            dict(trace_cfg, blas=args.blas))
    gi = None
    if not args.no_gi:
        gp = colonnade_gi_params(lux)
        ctx.gi_configure(gp)
        n_int, n_probes, n_lights = ctx.gi_sizes()
        gi_tune = None
        if not args.no_autotune:
            t0 = time.perf_counter()
            gi_tune = ctx.autotune_gi(repeats=2)
            gi_tune["tuning_s"] = round(time.perf_counter() - t0, 2)
            gi_tune["ms"] = round(gi_tune["ms"], 4)
        ctx.gi_update()
        ctx.synchronize()
        gi_ms = [ctx.gi_update(timed=True) for _ in range(args.gi_steps)]
        ctx.synchronize()
        # stage durations (HIP events between the stages, separate updates so
        # the timed ones above carry no extra synchronisation) and the
        # traversal work of one instrumented trace of the same grid
        ctx.gi_set_stage_timing(True)
        stages = []
        for _ in range(args.gi_steps):
            ctx.gi_update()
            ctx.synchronize()
            stages.append(ctx.gi_stage_times())
        ctx.gi_set_stage_timing(False)
        stage_ms = {k: round(float(np.median([x[k] for x in stages])), 4) for k in stages[0]}
        work = ctx.gi_trace_stats()
        gi_roof = gi_roofline(work, n_int, n_probes, gp, float(np.median(gi_ms)), stage_ms,
                              ctx.gi_trace_config(), args)
        gi = {
            "ms_per_gi_update": round(float(np.median(gi_ms)), 4),
            "stage_ms": stage_ms,
            "traversal_work": work,
            "roofline": gi_roof,
            "intervals": n_int,
            "interval_mrays_s": round(n_int / (float(np.median(gi_ms)) * 1e-3) / 1e6, 3),
            "probes": n_probes,
            "emitter_triangles": n_lights,
            "grid": "64x24x32 at 0.5 m, 5 levels, 32 base directions, NEE 8 samples",
            "trace": "wavefront (regrouped primary and NEE occlusion walks)",
            "trace_config": ctx.gi_trace_config(),
            "autotune": gi_tune,
        }
        if not args.no_surrogate:
            gi["surrogate"] = surrogate_section(ctx, lux, gp, args)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, rays_np)
        if gi is not None and cpu.get("gi_colonnade_ms_per_update"):
            gi["cpu_baseline"] = {k: cpu.pop(k) for k in list(cpu) if k.startswith("gi_")}

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
                        "(half closest-hit, half any-hit), "
                        + ("binned-SAH BLAS (GPU build), Morton-30 TLAS" if args.blas == "sah" else "Morton-30 LBVH"),
            "rays": n,
            "triangles": scene.triangle_count,
            "bvh_build_ms": round(build_ms, 3),
            "kernel_ms": round(avg_kernel_ms, 4),
            "node_fetches_per_ray": round(nodes / n, 3),
            "tri_tests_per_ray": round(tris / n, 3),
            "trace_config": trace_cfg,
            "autotune": tune,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "traffic_config": traffic_cfg,
        },
        "cpu_baseline": cpu,
        "gi_update": gi,
    }
    ctx.close()
    if not args.no_dynamic:
        out["dynamic_frame"] = dynamic_section(lux, scenes, args)
    if not args.no_batch:
        out["batched_scenes"] = batch_section(lux, scenes, args)
    if not args.no_receivers:
        out["receiver_updates"] = receivers_section(lux, scenes, args)
    if not args.no_screen:
        out["screen_filters"] = screen_section(lux, args)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
