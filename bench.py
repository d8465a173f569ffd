"""Benchmark of the wavefront hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload sorting_stress]

A step is one full frame of the plain renderer (WavefrontRenderer::renderScene): every sample
pass of every pixel through every bounce (ray sort, closest hit + split, shading-queue policy
sort, shade with NEE shadow rays, misses, adaptive bookkeeping). The default workload is the
material/texture sorting stress scene: a 1.31M-triangle noise-displaced icosphere with 64
materials and 32 1024^2 checker textures, 1920x1080, 64 spp, depth 12, rendered with each of
the four scheduling policies. Scene upload and BVH build happen before timing; the timed region
is bracketed by device synchronisation. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def workload(name: str, wf, scenes, quick: bool):
    if name == "sorting_stress":
        data = scenes.sorting_stress(subdiv=6 if quick else 8, tex_size=256 if quick else 1024)
        W, H = (480, 270) if quick else (1920, 1080)
        spp, depth = (4 if quick else 64), 12
        desc = (f"sorting_stress: noise-displaced icosphere subdiv {8 if not quick else 6} "
                f"({data.triangle_count} tris), 64 materials (40% Lambertian / 40% GGX / 20% "
                f"dielectric), 32 checker textures, 2 area lights, {W}x{H}, {spp} spp, depth {depth}")
        return data, W, H, spp, depth, desc
    if name == "cornell_box":
        data = scenes.cornell_box()
        return data, 256, 256, 16, 8, "cornell_box: 5132 tris, 256x256, 16 spp, depth 8"
    if name == "sphere_cloud":
        data = scenes.sphere_cloud(copies=256 if quick else 4096)
        W, H = (480, 270) if quick else (1920, 1080)
        spp = 64 if quick else 1024
        desc = (f"sphere_cloud: {data.triangle_count} tris (noise-displaced spheres), 256 "
                f"materials, 1024 emissive triangles, 2048x1024 HDRI, {W}x{H}, adaptive 16..{spp} "
                f"spp to relative std-err 0.01, depth 12, costBenefit")
        return data, W, H, spp, 12, desc
    raise SystemExit(f"unknown workload {name}")


# Per-workload render settings beyond size and depth: physics, policy, adaptive sampling.
def workload_settings(name: str, wf) -> dict:
    if name == "sphere_cloud":
        return {"physics": wf.PHYSICS_UNBIASED, "policy": "costBenefit", "adaptive": True,
                "adaptive_threshold": 0.01, "adaptive_min_samples": 16, "adaptive_interval": 16}
    if name == "sorting_stress":
        return {"physics": wf.PHYSICS_UNBIASED, "policy": "none"}
    return {"physics": wf.PHYSICS_REFERENCE, "policy": "none"}


# The kernels of a shading policy in the engine (hit counts per material, the range plan, and
# the split that lays the order out), and those of the sort-and-gather form used by the C-ABI.
POLICY_KERNELS = ("policy_count", "policy_plan", "split", "policy_keys", "shading_sort",
                  "gather_shading")
# Kernels whose bound is HBM bandwidth (sort, compaction, gathers, shading); the traversal is
# bound by the vector-memory data return of its loads, with VALU issue close behind (see
# measured_pipes).
HBM_KERNELS = ("ray_keys", "ray_sort", "gather_rays", "split", "compact", "policy_count",
               "policy_keys", "shading_sort", "gather_shading", "shade", "next_rays", "miss",
               "raygen", "regen", "finish")


def stream_copy_gbs(wf, ctx, nbytes: int, reps: int) -> float:
    """Achievable HBM bandwidth: the library's streaming copy over nbytes, HIP-event timed."""
    src = ctx.alloc(nbytes)
    dst = ctx.alloc(nbytes)
    try:
        ctx.fill_u32(src, 0x3F800000, nbytes // 4)
        ctx.stream_copy(dst, src, nbytes)            # untimed first touch
        ctx.synchronize()
        ctx.profile(True)
        for _ in range(reps):
            ctx.stream_copy(dst, src, nbytes)
        ms, launches, moved = ctx.profile_query("stream_copy")
        ctx.profile(False)
        return moved / (ms * 1e-3) / 1e9
    finally:
        ctx.free(src)
        ctx.free(dst)


def camera_rays(camera, W, H, rng):
    """Jittered pinhole rays of the workload's camera (numpy; for traversal statistics)."""
    eye = np.array(camera["eye"], np.float64)
    at = np.array(camera["target"], np.float64)
    up = np.array(camera.get("up", (0.0, 1.0, 0.0)), np.float64)
    vh = 2.0 * np.tan(np.radians(camera["vfov"]) / 2.0)
    vw = vh * W / H
    fwd = (eye - at) / np.linalg.norm(eye - at)
    right = np.cross(up, fwd)
    right /= np.linalg.norm(right)
    upv = np.cross(fwd, right)
    y, x = np.mgrid[0:H, 0:W]
    u = (x.ravel() + rng.random(W * H)) / (W - 1)
    v = 1.0 - (y.ravel() + rng.random(W * H)) / (H - 1)
    tl = eye - right * vw / 2 - upv * vh / 2 - fwd
    d = tl + np.outer(u, right * vw) + np.outer(v, upv * vh) - eye
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.broadcast_to(eye, d.shape)
    return o.astype(np.float32), d.astype(np.float32)


def traversal_stats(wf, ctx, scene, data, W, H, rng) -> dict:
    """Per ray class: BVH nodes visited and triangles tested per ray, the SIMD efficiency of a
    64-lane wave that loops until its slowest lane ends, and device rays/s of the trace."""
    def queue(o, d):
        n = len(o)
        q = wf.RayQueue(ctx, n)
        z = np.zeros(n, np.float32)
        q.upload({"ox": o[:, 0], "oy": o[:, 1], "oz": o[:, 2], "dx": d[:, 0], "dy": d[:, 1],
                  "dz": d[:, 2], "tr": z, "tg": z, "tb": z, "pixel": np.arange(n, dtype=np.int32),
                  "depth": np.zeros(n, np.int32), "pdf": z})
        return q

    def timed(kernel, fn):
        ctx.synchronize()
        ctx.profile_filter(kernel)
        ctx.profile(True)
        fn()
        ms, launches, _ = ctx.profile_query(kernel)
        ctx.profile(False)
        ctx.profile_filter(None)
        return ms

    out = {}
    o, d = camera_rays(data.camera, W, H, rng)
    q = queue(o, d)
    n = len(o)
    ms = timed("traverse_closest", lambda: scene.intersect(q, n))
    hits = scene.intersect(q, n)
    nodes, tris = scene.trace_stats(q, n)
    out["primary"] = {"rays": n, "mrays_s": round(n / ms / 1e3, 1),
                      "nodes_per_ray": round(float(nodes.mean()), 2),
                      "tris_per_ray": round(float(tris.mean()), 2),
                      "simd_efficiency": round(wf.simd_efficiency(nodes), 3)}
    h = hits["prim"] >= 0
    P = o[h] + d[h] * hits["t"][h][:, None]
    m = len(P)
    if m == 0:
        return out
    # diffuse-like bounce: uniform directions leaving the hit point
    w = rng.normal(size=(m, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    w = np.where(np.sum(w * d[h], axis=1, keepdims=True) > 0, -w, w).astype(np.float32)
    q2 = queue(P.astype(np.float32), w)
    ms = timed("traverse_closest", lambda: scene.intersect(q2, m))
    nodes, tris = scene.trace_stats(q2, m)
    out["bounce"] = {"rays": m, "mrays_s": round(m / ms / 1e3, 1),
                     "nodes_per_ray": round(float(nodes.mean()), 2),
                     "tris_per_ray": round(float(tris.mean()), 2),
                     "simd_efficiency": round(wf.simd_efficiency(nodes), 3)}
    # shadow rays toward random points of the emitters (any hit)
    lights = [mesh for mesh in data.meshes
              if data.materials[mesh.material_id]["type"] in (wf.MAT_EMISSIVE, wf.MAT_SPOTLIGHT)]
    if lights:
        tri = np.concatenate([mesh.positions[mesh.indices] for mesh in lights])
        k = rng.integers(0, len(tri), m)
        r1, r2 = rng.random(m), rng.random(m)
        flip = r1 + r2 > 1
        r1[flip], r2[flip] = 1 - r1[flip], 1 - r2[flip]
        lp = (tri[k, 0] * (1 - r1 - r2)[:, None] + tri[k, 1] * r1[:, None] +
              tri[k, 2] * r2[:, None])
        tl = lp - P
        dist = np.linalg.norm(tl, axis=1)
        sd = (tl / dist[:, None]).astype(np.float32)
        q3 = queue(P.astype(np.float32), sd)
        nodes, tris = scene.trace_stats(q3, m, 0.001, 1e9, any_hit=True)
        md = (dist - 0.001).astype(np.float32)
        ms = timed("traverse_any", lambda: scene.occluded(P.astype(np.float32), sd, md))
        out["shadow"] = {"rays": m, "mrays_s": round(m / ms / 1e3, 1),
                         "nodes_per_ray": round(float(nodes.mean()), 2),
                         "tris_per_ray": round(float(tris.mean()), 2),
                         "simd_efficiency": round(wf.simd_efficiency(nodes), 3),
                         "note": "any hit up to tfar 1e9 for the counts"}
    return out


# Library kernel names (per-launch HIP-event timing) -> device kernel symbols in rocprofv3 output
KERNEL_SYMBOLS = {"traverse_closest": "k_trace_closest", "shadow": "k_shadow",
                  "traverse_closest_shadow": "k_closest_shadow",
                  "split": "k_split_index", "shade": "k_shade_hits",
                  "next_rays": "k_shade_append", "gather_rays": "k_gather_rays",
                  "miss": "k_miss", "regen": "k_regen", "ray_keys": "k_ray_keys"}


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of this workload (profiles/traffic_<workload>_1080p_64spp.json), or None."""
    path = ROOT / "profiles" / f"traffic_{workload}_1080p_64spp.json"
    sym = KERNEL_SYMBOLS.get(kernel)
    if not sym or not path.exists():
        return None
    t = json.loads(path.read_text())
    for name, v in t["kernels"].items():
        if name.split("<")[0] == sym:
            return {"bytes_per_launch": v["bytes_per_launch"], "source": str(path.relative_to(ROOT)),
                    "fetch_scale": t["fetch_scale"], "write_scale": t["write_scale"]}
    return None


# MI355X: 256 CUs x 4 SIMDs; a wave64 FP32 VALU instruction occupies a 16-lane SIMD for 4 cycles
# (packed FP32 is not used), at the 2.4 GHz peak engine clock
VALU_PEAK_GINST = 256 * 4 * 2.4 / 4.0
# one texture-data (TD) unit per CU returns vector-memory load data to the waves: busy cycles per
# second of all 256, at the same clock
TD_PEAK_GCYCLES = 256 * 2.4


def measured_pipes(workload: str, kernel: str):
    """Per-launch hardware counters of `kernel` (VALU instructions, TD busy cycles, ...) from the
    committed rocprofv3 counter summary of this workload
    (profiles/pipes_<workload>_1080p_64spp.json), or None."""
    path = ROOT / "profiles" / f"pipes_{workload}_1080p_64spp.json"
    sym = KERNEL_SYMBOLS.get(kernel)
    if not sym or not path.exists():
        return None
    t = json.loads(path.read_text())
    for name, v in t["kernels"].items():
        if name.split("<")[0] == sym and "SQ_INSTS_VALU" in v:
            return {**v, "source": str(path.relative_to(ROOT))}
    return None


def host_cpu() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def autotune_batched(wf, ctx, scs, datas, physics, jobs, args) -> list:
    """Engine capacity and sample lanes per scene, and the context's traversal block shape,
    refill threshold and radix digit width, picked by the autotuner on a calibration subset of
    each scene's jobs (its four largest and two others, budgets capped at 64 spp). The context
    knobs are tuned on the scene with the most samples and kept for both engines."""
    autotune = importlib.import_module("wavefront-renderer_amd.autotune")
    forms = importlib.import_module("wavefront-renderer_amd.forms")
    from dataclasses import replace
    work = [sum(j.pixels * j.spp for j in jobs if j.scene == k) for k in range(2)]
    order = sorted(range(2), key=lambda k: -work[k])
    out = [None, None]
    t0 = time.perf_counter()
    for n, k in enumerate(order):
        mine = [j for j in jobs if j.scene == k]
        if not mine:
            out[k] = {"capacity": 1 << 20, "sample_lanes": 0}
            continue
        # the jobs with the most samples set the makespan: calibrate on the four largest (budgets
        # capped at 64 spp) and two spread over the rest
        big = sorted(mine, key=lambda j: -j.pixels * j.spp)
        rest = big[4:]
        picked = big[:4] + (rest[::max(1, len(rest) // 2)][:2] if rest else [])
        calib = [replace(j, spp=min(j.spp, 64)) for j in picked]
        budget = sum(j.pixels for j in calib)
        space = dict(autotune.ENGINE_SPACE)
        if n > 0:
            for knob in autotune.CONTEXT_KNOBS:
                space.pop(knob)
        res = autotune.tune_engine(
            ctx, scs[k], lambda: [forms._desc(j, datas[k].camera) for j in calib], budget,
            len(calib), {"physics": physics[k], "firefly": 1e30, "ray_sort": False}, space=space,
            rounds=1)
        out[k] = {"capacity": res["best"]["capacity"],
                  "sample_lanes": res["best"]["sample_lanes"]}
        log(f"[bench] autotune scene {k}: {res['best']} ({res['trials']} trials)")
    log(f"[bench] autotune: {time.perf_counter() - t0:.1f}s")
    return out


def run_batched(wf, ctx, scenes, args) -> dict:
    """Batched form: 64 jobs from the Cornell-box and sorting-stress generators, sequential wf_render
    against one job-tagged engine per scene (all jobs admitted at once)."""
    from dataclasses import replace
    forms = importlib.import_module("wavefront-renderer_amd.forms")
    datas = [scenes.cornell_box(), scenes.sorting_stress()]
    scs = [wf.Scene(ctx, d) for d in datas]
    physics = [wf.PHYSICS_REFERENCE, wf.PHYSICS_UNBIASED]
    jobs = forms.batched_jobs(args.jobs)
    for k in range(2):   # size wf_render's cached engines before timing
        mine = [j for j in jobs if j.scene == k]
        if mine:
            forms.sequential(ctx, scs, datas, physics, [replace(max(mine, key=lambda j: j.pixels), spp=1)])
    # both legs run three times and report their medians: single runs on the same box have
    # measured up to ~2x apart on either leg
    seq_runs = []
    for _ in range(3):
        secs, rays_seq = forms.sequential(ctx, scs, datas, physics, jobs)
        seq_runs.append(sum(secs))
    t_seq = float(np.median(seq_runs))
    log(f"[bench] batched: sequential {t_seq:.2f}s for {len(jobs)} jobs (runs {seq_runs})")
    pix = [max(1, sum(j.pixels for j in jobs if j.scene == k)) for k in range(2)]
    tuned = autotune_batched(wf, ctx, scs, datas, physics, jobs, args) if args.autotune_batch \
        else [{"capacity": min(pix[k], 4 << 20), "sample_lanes": 0} for k in range(2)]
    runs = []
    for _ in range(3):
        engines = [wf.Engine(ctx, scs[k], capacity=tuned[k]["capacity"], pixel_budget=pix[k],
                             max_jobs=max(1, len(jobs)), physics=physics[k], firefly=1e30,
                             use_graph=True, sample_lanes=tuned[k]["sample_lanes"],
                             ray_sort=False)
                   for k in range(2)]
        r = forms.serve(engines, datas, jobs, realtime=False)
        runs.append((r["makespan_s"], r, [j.done for j in jobs]))
        for e in engines:
            e.close()
    _, res, done = sorted(runs, key=lambda x: x[0])[1]   # the median run
    out = {"jobs": len(jobs), "scenes": ["cornell_box", "sorting_stress"],
           "pixels": int(sum(j.pixels for j in jobs)),
           "samples": int(sum(j.pixels * j.spp for j in jobs)),
           "sequential_s": round(t_seq, 3), "batched_s": round(res["makespan_s"], 3),
           "sequential_jobs_s": round(len(jobs) / t_seq, 2),
           "batched_jobs_s": round(len(jobs) / res["makespan_s"], 2),
           "speedup": round(t_seq / res["makespan_s"], 2),
           "sequential_mrays_s": round(rays_seq / t_seq / 1e6, 1),
           "batched_mrays_s": round(res["rays"] / res["makespan_s"] / 1e6, 1),
           "batched_p50_s": round(forms.pct(done, 50), 3),
           "batched_p99_s": round(forms.pct(done, 99), 3),
           "runs_s": {"sequential": [round(t, 3) for t in seq_runs],
                      "batched": [round(r[0], 3) for r in runs]},
           "engines": tuned, "tuning": ctx.tuning()}
    log(f"[bench] batched: {json.dumps(out)}")
    return out


def run_served(wf, ctx, scenes, args) -> dict:
    """Served form: Poisson requests against three resident scenes at ~0.7 of the measured
    capacity (the same requests admitted at once), adaptive sampling to a relMSE target;
    p50/p99 time-to-converged-image, against the same requests rendered one after another."""
    forms = importlib.import_module("wavefront-renderer_amd.forms")
    datas = [scenes.cornell_box(), scenes.sorting_stress(), scenes.sphere_cloud()]
    scs = [wf.Scene(ctx, d) for d in datas]
    physics = [wf.PHYSICS_REFERENCE, wf.PHYSICS_UNBIASED, wf.PHYSICS_UNBIASED]
    engines = forms.make_engines(ctx, scs, physics, capacity=args.served_capacity,
                                 pixel_budget=8 << 20, max_jobs=64)
    # capacity: the stream's requests admitted at once (a burst); the same requests then arrive
    # as a Poisson stream at 0.7 of that rate, and are rendered one after another with wf_render
    burst = forms.served_jobs(args.jobs, rate=1.0, max_spp=args.served_spp)
    forms.serve(engines, datas, burst[:3], realtime=False)   # warm the engines' graphs
    bursts = []
    for _ in range(3):   # the median of three bursts (single runs have measured ~2x apart)
        burst = forms.served_jobs(args.jobs, rate=1.0, max_spp=args.served_spp)
        bursts.append(forms.serve(engines, datas, burst, realtime=False)["makespan_s"])
    capacity = len(burst) / float(np.median(bursts))
    rate = 0.7 * capacity
    jobs = forms.served_jobs(args.jobs, rate=rate, max_spp=args.served_spp)
    res = forms.serve(engines, datas, jobs, realtime=True, collect=True)
    lat = [j.done - j.arrival for j in jobs]
    for e in engines:
        e.close()
    conv = forms.check_converged(scs, datas, physics, jobs, res["images"])
    del res["images"]
    secs, rays_seq = forms.sequential(ctx, scs, datas, physics, jobs)
    seq_lat = forms.fcfs_latency([j.arrival for j in jobs], secs)
    out = {"requests": len(jobs), "scenes": ["cornell_box", "sorting_stress", "sphere_cloud"],
           "rate_jobs_s": round(rate, 2), "capacity_jobs_s": round(capacity, 2),
           "capacity_runs_s": [round(t, 3) for t in bursts],
           "rel_mse_target": 1e-3, "max_spp": args.served_spp,
           "served_jobs_s": round(len(jobs) / res["makespan_s"], 2),
           "served_mrays_s": round(res["rays"] / res["makespan_s"] / 1e6, 1),
           "p50_s": round(forms.pct(lat, 50), 3), "p99_s": round(forms.pct(lat, 99), 3),
           "sequential_capacity_jobs_s": round(len(jobs) / sum(secs), 2),
           "sequential_p50_s": round(forms.pct(seq_lat, 50), 3),
           "sequential_p99_s": round(forms.pct(seq_lat, 99), 3),
           "sequential_mrays_s": round(rays_seq / sum(secs) / 1e6, 1),
           "converged": conv}
    log(f"[bench] served: {json.dumps(out)}")
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sorting_stress")
    ap.add_argument("--policy", default=None,
                    help="shading-queue policy (default: the workload's own)")
    ap.add_argument("--schedule", default="regenerate", choices=["regenerate", "passes"],
                    help="regenerate: full wavefront, pixels start their next sample as soon as "
                         "their path ends; passes: the reference's pass-by-pass loop (same image)")
    ap.add_argument("--lanes", type=int, default=64, choices=[1, 2, 4, 8, 16, 32, 64],
                    help="regenerate: samples of a pixel in flight at once (1 = the pass loop's "
                         "accumulation order bit for bit)")
    ap.add_argument("--wavefront", type=int, default=0,
                    help="regenerate: ray slots of the wavefront (0 = automatic)")
    ap.add_argument("--ray-key-bits", type=int, default=0,
                    help="regenerate: order each bounce's rays by the top N bits of "
                         "RayQueue::sortKey (1..33, 33 = the reference's full key); 0 = no "
                         "ordering (the queue order there changes no result; the frame is also "
                         "timed with the full key for comparison)")
    ap.add_argument("--quick", action="store_true", help="reduced sizes (CPU-side checks)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-policies", action="store_true")
    ap.add_argument("--copy-only", action="store_true",
                    help="only the streaming-copy kernel (calibrates hardware byte counters)")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--cpu-spp", type=int, default=16,
                    help="samples per pixel of the CPU baseline sample (same image size)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--jobs", type=int, default=64, help="jobs of the batched/served workloads")
    ap.add_argument("--served-spp", type=int, default=2048,
                    help="per-pixel sample cap of served requests (adaptive sampling stops each "
                         "pixel at relative standard error sqrt(1e-3) / 1.25 before it)")
    ap.add_argument("--served-capacity", type=int, default=8 << 20,
                    help="ray slots of each served scene's engine")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the batched and served legs")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the plain frame's wavefront size, sample lanes, traversal block "
                         "shape and radix digit width with the autotuner before timing")
    ap.add_argument("--no-autotune-batch", dest="autotune_batch", action="store_false",
                    help="batched leg: fixed engine settings instead of the autotuner's")
    ap.add_argument("--profile-run", action="store_true",
                    help="only the timed frames (no comparison legs): for profiler runs")
    ap.add_argument("--ref-spp", type=int, default=1024,
                    help="samples per pixel of the FP64 oracle image the timed frame's relMSE is "
                         "measured against (a 128x72 window at the image centre)")
    ap.add_argument("--no-image-check", action="store_true",
                    help="skip the relMSE of the timed frame against the oracle")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    wf = importlib.import_module("wavefront-renderer_amd")
    scenes = importlib.import_module("wavefront-renderer_amd.scenes")
    ctx = wf.Context(0)
    copy_bytes = args.copy_mib << 20
    if args.copy_only:
        gbs = stream_copy_gbs(wf, ctx, copy_bytes, max(args.steps, 1))
        print(json.dumps({"stream_copy_bytes": copy_bytes, "launches": max(args.steps, 1),
                          "GB_s": round(gbs, 1)}))
        return
    if args.workload in ("batched", "served"):
        t0 = time.perf_counter()
        out = run_batched(wf, ctx, scenes, args) if args.workload == "batched" else \
            run_served(wf, ctx, scenes, args)
        key = "batched_jobs_s" if args.workload == "batched" else "served_jobs_s"
        print(json.dumps({"metric": "jobs/s", "value": out[key], "unit": "jobs/s", "n_gpus": 1,
                          "steps": 1, "warmup": 0,
                          "ms_per_step": round((time.perf_counter() - t0) * 1e3, 1),
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                          "dtype": "f32", "data": "synthetic (seeded generator)",
                          "config": {"workload": args.workload, **out}}))
        return
    t0 = time.time()
    data, W, H, spp, depth, desc = workload(args.workload, wf, scenes, args.quick)
    log(f"[bench] scene generated in {time.time() - t0:.1f}s: {desc}")
    t0 = time.time()
    scene = wf.Scene(ctx, data)
    log(f"[bench] uploaded + BVH built in {time.time() - t0:.1f}s "
        f"(device build {scene.stats.build_ms:.1f} ms, binary depth {scene.stats.max_depth}, "
        f"{scene.stats.bvh_width}-wide nodes {scene.stats.node_count}, stack bound "
        f"{scene.stats.stack_bound})")
    ws = workload_settings(args.workload, wf)
    physics = ws["physics"]
    if args.policy is None:
        args.policy = ws["policy"]
    adaptive = {k: ws[k] for k in ("adaptive", "adaptive_threshold", "adaptive_min_samples",
                                   "adaptive_interval") if k in ws}

    if args.ray_key_bits:
        ctx.set_tuning(ray_key_bits=args.ray_key_bits)

    def params(policy, seed=1, w=W, h=H, s=spp, schedule=args.schedule, ray_sort=None):
        rs = args.ray_key_bits > 0 if ray_sort is None else ray_sort
        return wf.render_params(width=w, height=h, spp=s, max_depth=depth,
                                camera=data.camera, policy=policy, physics=physics,
                                firefly=1e30, seed=seed, schedule=schedule, ray_sort=rs,
                                sample_lanes=args.lanes, wavefront=args.wavefront, **adaptive)

    tuned = None
    if args.autotune:
        autotune = importlib.import_module("wavefront-renderer_amd.autotune")
        t0 = time.perf_counter()

        def params_for(lanes, wave):
            return wf.render_params(width=W, height=H, spp=spp, max_depth=depth,
                                    camera=data.camera, policy=args.policy, physics=physics,
                                    firefly=1e30, seed=7, sample_lanes=lanes, wavefront=wave,
                                    **adaptive)
        tuned = autotune.tune_render(ctx, scene, params_for, W * H, rounds=1, repeats=1)
        args.lanes = tuned["best"]["sample_lanes"]
        args.wavefront = tuned["best"]["wavefront_per_pixel"] * W * H
        log(f"[bench] autotune ({time.perf_counter() - t0:.1f}s, {tuned['trials']} trials): "
            f"{tuned['best']}")

    def frame(policy, seed=1, schedule=args.schedule, ray_sort=None):
        img, cnt, st = wf.render(scene, params(policy, seed, schedule=schedule,
                                               ray_sort=ray_sort), keep_on_device=True)
        return st

    for w in range(args.warmup):
        st = frame(args.policy, seed=100 + w)
        log(f"[bench] warmup {w}: {st.frame_ms:.1f} ms")
    # one frame with events around every kernel: the per-kernel table and the dominant kernel
    ctx.profile_filter(None)
    ctx.profile(True)
    frame(args.policy, seed=99)
    kernels = {}
    for name in ctx.profile_names():
        ms, launches, nbytes = ctx.profile_query(name)
        if launches == 0:
            continue
        gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        kernels[name] = {"ms_per_frame": round(ms, 3), "launches_per_frame": launches,
                         "avg_us": round(ms * 1e3 / launches, 2), "GB_s": round(gbs, 1),
                         "frac": round(gbs / HBM_PEAK_GBS, 4)}
    ctx.profile(False)
    # beside the algorithmic rate, the rate of the HBM traffic the kernel actually moves (its
    # calibrated counter bytes per launch from the committed profile, over this frame's launches)
    for name, row in kernels.items():
        tr = measured_traffic(args.workload, name)
        if tr and row["avg_us"] > 0:
            row["traffic_GB_s"] = round(tr["bytes_per_launch"] / (row["avg_us"] * 1e-6) / 1e9, 1)
    log(f"[bench] kernels (profiled frame): {json.dumps(kernels)}")
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_frame"]) if kernels else None

    # timed region: HIP events (on the library's stream) around the dominant kernel only
    ctx.synchronize()
    ctx.profile_filter(dom)
    ctx.profile(True)
    rays = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = frame(args.policy, seed=1 + k)
        rays += st.rays_traced + st.shadow_rays
        log(f"[bench] step {k}: {st.frame_ms:.1f} ms device, {st.rays_traced} rays + "
            f"{st.shadow_rays} shadow rays")
    ctx.synchronize()
    wall = time.perf_counter() - t0
    mrays = rays / wall / 1e6
    ms_per_step = wall * 1e3 / args.steps
    roofline = None
    if dom:
        ms, launches, nbytes = ctx.profile_query(dom)
        gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        avg_us = ms * 1e3 / max(launches, 1)
        hbm_view = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)}
        roofline = {"kernel": dom, **hbm_view, "traffic": None,
                    "avg_launch_us": round(avg_us, 2), "launches": launches,
                    "bytes_per_launch": round(nbytes / max(launches, 1))}
        pipes = measured_pipes(args.workload, dom)
        if pipes:
            # the traversal is bound by the vector-memory data return (TD) of its node and
            # triangle loads, with VALU issue close behind; not by HBM (its algorithmic bytes are
            # rays in and hits out): both views are per launch from the counters over the live
            # launch time
            ginst = pipes["SQ_INSTS_VALU"] / (avg_us * 1e-6) / 1e9
            valu_view = {"bound": "valu", "achieved": round(ginst, 1), "peak": VALU_PEAK_GINST,
                         "unit": "G VALU wave-instructions/s",
                         "frac": round(ginst / VALU_PEAK_GINST, 4),
                         "valu_per_launch": round(pipes["SQ_INSTS_VALU"])}
            roofline.update({"valu": valu_view, "hbm": hbm_view,
                             "counters_source": pipes["source"]})
            if "TD_TD_BUSY_sum" in pipes:
                gtd = pipes["TD_TD_BUSY_sum"] / (avg_us * 1e-6) / 1e9
                roofline.update({"bound": "td", "achieved": round(gtd, 1),
                                 "peak": TD_PEAK_GCYCLES, "unit": "G TD-busy cycles/s",
                                 "frac": round(gtd / TD_PEAK_GCYCLES, 4),
                                 "td_busy_per_launch": round(pipes["TD_TD_BUSY_sum"])})
            else:
                roofline.update({k: valu_view[k] for k in ("bound", "achieved", "peak", "unit",
                                                          "frac")})
        tr = measured_traffic(args.workload, dom)
        if tr:
            roofline["traffic"] = tr["bytes_per_launch"]
            roofline["traffic_source"] = tr["source"]
    ctx.profile(False)
    ctx.profile_filter(None)
    hbm = [k for k in kernels if k in HBM_KERNELS]
    if hbm and roofline is not None:
        ms_sum = sum(kernels[k]["ms_per_frame"] for k in hbm)
        bytes_sum = sum(kernels[k]["GB_s"] * kernels[k]["ms_per_frame"] * 1e6 for k in hbm)
        agg = bytes_sum / (ms_sum * 1e-3) / 1e9 if ms_sum > 0 else 0.0
        roofline["hbm_kernels"] = {"kernels": hbm, "ms_per_frame": round(ms_sum, 2),
                                   "GB_s": round(agg, 1), "frac": round(agg / HBM_PEAK_GBS, 4)}
        withtr = [k for k in hbm if "traffic_GB_s" in kernels[k]]
        if withtr:
            # the same group by the traffic it moves (kernels with a committed traffic profile)
            tms = sum(kernels[k]["ms_per_frame"] for k in withtr)
            tby = sum(kernels[k]["traffic_GB_s"] * kernels[k]["ms_per_frame"] * 1e6 for k in withtr)
            tg = tby / (tms * 1e-3) / 1e9 if tms > 0 else 0.0
            roofline["hbm_kernels"]["by_traffic"] = {"kernels": withtr, "GB_s": round(tg, 1),
                                                     "frac": round(tg / HBM_PEAK_GBS, 4)}
    if not args.profile_run:
        copy = round(stream_copy_gbs(wf, ctx, copy_bytes, 5), 1)
        log(f"[bench] streaming copy {args.copy_mib} MiB: {copy} GB/s")
        if roofline is not None:
            roofline["stream_copy_GB_s"] = copy

    tstats, other_sched, full_key = None, None, None
    if not args.profile_run:
        tstats = traversal_stats(wf, ctx, scene, data, W, H, np.random.default_rng(5))
        log(f"[bench] traversal: {json.dumps(tstats)}")
        # the other schedule of the same frame, for comparison (same paths; the same image up to
        # the order of the float additions, bit for bit with one sample lane)
        other = "passes" if args.schedule == "regenerate" else "regenerate"
        t1 = time.perf_counter()
        st = frame(args.policy, seed=5, schedule=other)
        ctx.synchronize()
        dt = time.perf_counter() - t1
        other_sched = {"schedule": other, "ms": round(dt * 1e3, 1),
                       "mrays_s": round((st.rays_traced + st.shadow_rays) / dt / 1e6, 2)}
        log(f"[bench] schedule {other}: {other_sched}")
        # the same frame with every bounce's rays ordered by the reference's full 33-bit key
        saved_bits = ctx.tuning()["ray_key_bits"]
        ctx.set_tuning(ray_key_bits=33)
        # its first frame with events around every kernel: the rows of the ray-order kernels
        # (keys, sort passes), which the frame without ray ordering does not run
        ctx.synchronize()
        ctx.profile_filter(None)
        ctx.profile(True)
        frame(args.policy, seed=6, ray_sort=True)
        sort_rows = {}
        for name in ctx.profile_names():
            ms, launches, nbytes = ctx.profile_query(name)
            if launches == 0 or name not in ("ray_keys", "ray_sort"):
                continue
            gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
            sort_rows[name] = {"ms_per_frame": round(ms, 3), "launches_per_frame": launches,
                               "avg_us": round(ms * 1e3 / launches, 2), "GB_s": round(gbs, 1),
                               "frac": round(gbs / HBM_PEAK_GBS, 4)}
        ctx.profile(False)
        ctx.synchronize()
        t1 = time.perf_counter()
        st = frame(args.policy, seed=5, ray_sort=True)
        ctx.synchronize()
        dt = time.perf_counter() - t1
        ctx.set_tuning(ray_key_bits=saved_bits)
        full_key = {"ray_key_bits": 33, "ms": round(dt * 1e3, 1),
                    "mrays_s": round((st.rays_traced + st.shadow_rays) / dt / 1e6, 2),
                    "kernels": sort_rows}
        log(f"[bench] full ray key: {full_key}")

    policies = {}
    if not args.no_policies and not args.profile_run:
        # phase times need events between the phases: time no kernel, but keep timing on
        ctx.profile_filter("phases")
        ctx.profile(True)
        for pol in ("none", "material", "texture", "costBenefit"):
            t1 = time.perf_counter()
            st = frame(pol, seed=7)
            ctx.synchronize()
            dt = time.perf_counter() - t1
            policies[pol] = {"mrays_s": round((st.rays_traced + st.shadow_rays) / dt / 1e6, 2),
                             "ms": round(dt * 1e3, 1), "sort_ms": round(st.sort_ms, 1),
                             "intersect_ms": round(st.intersect_ms, 1),
                             "schedule_ms": round(st.schedule_ms, 1),
                             "shade_ms": round(st.shade_ms, 1)}
            log(f"[bench] policy {pol}: {policies[pol]}")
        ctx.profile(False)
        # one more frame per ordering policy with events around every kernel: the policy's
        # kernels against the HBM roof
        for pol in ("material", "texture", "costBenefit"):
            ctx.profile_filter(None)
            ctx.profile(True)
            frame(pol, seed=8)
            rows = {}
            for name in ctx.profile_names():
                ms, launches, nbytes = ctx.profile_query(name)
                if launches == 0 or name not in POLICY_KERNELS:
                    continue
                gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
                rows[name] = {"ms_per_frame": round(ms, 3), "launches_per_frame": launches,
                              "avg_us": round(ms * 1e3 / launches, 2), "GB_s": round(gbs, 1),
                              "frac": round(gbs / HBM_PEAK_GBS, 4)}
            policies[pol]["kernels"] = rows
            ctx.profile(False)
            log(f"[bench] policy {pol} kernels: {json.dumps(rows)}")
        ctx.profile_filter(None)

    cpu = None
    if not args.no_cpu_baseline and not args.profile_run:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_lib  # CPU baseline leg only

        threads = max(1, args.cpu_threads)
        t1 = time.perf_counter()
        osc = oracle_lib.Oracle().scene(data)
        build_s = time.perf_counter() - t1
        cs = max(1, min(args.cpu_spp, spp))
        cp = params(args.policy, seed=1, s=cs)
        t1 = time.perf_counter()
        _, _, _, cst = osc.render(cp, threads=threads)
        dt = time.perf_counter() - t1
        crays = cst.rays_traced + cst.shadow_rays
        cpu = {"value": round(crays / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads,
               "kind": "port",
               "sample": f"same scene and camera at {W}x{H}, {cs} of {spp} spp, depth {depth}, "
                         f"policy {args.policy}: {crays} rays in {dt:.1f}s on {threads} threads "
                         f"(reference-semantics CPU restatement, not Embree; its SAH BVH build "
                         f"{build_s:.1f}s excluded); host {host_cpu()}, "
                         f"{os.cpu_count()} logical CPUs visible"}
        log(f"[bench] cpu baseline: {cpu}")

    image = None
    if not args.no_image_check and not args.profile_run:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_lib  # the checker

        t1 = time.perf_counter()
        osc_img = oracle_lib.Oracle().scene(data)
        ww, wh = min(128, W), min(72, H)
        win = ((W - ww) // 2, (H - wh) // 2, ww, wh)
        ref_p = wf.render_params(width=W, height=H, spp=args.ref_spp, max_depth=depth,
                                 camera=data.camera, policy="none", physics=physics,
                                 firefly=1e30, seed=4242)
        _, _, ref, _ = osc_img.render_window(ref_p, win, threads=max(1, args.cpu_threads))
        # the first timed frame again (seed 1), its window
        img, cnt, _ = wf.render(scene, params(args.policy, seed=1))
        x0, y0 = win[0], win[1]
        got = img[y0:y0 + wh, x0:x0 + ww].astype(np.float64)
        n_spp = float(np.mean(cnt[y0:y0 + wh, x0:x0 + ww]))
        relmse = float(np.mean((got - ref) ** 2 / (ref ** 2 + 1e-2)))
        image = {"window": list(win), "reference": f"oracle, FP64 accumulation, {args.ref_spp} spp",
                 "frame_spp": round(n_spp, 1), "relmse": round(relmse, 6),
                 "reference_noise_share": round(n_spp / args.ref_spp, 4),
                 "spp_for_relmse_1e-3": round(n_spp * relmse / 1e-3),
                 "seconds": round(time.perf_counter() - t1, 1)}
        log(f"[bench] image check: {image}")

    forms_out = None
    if not args.no_forms and not args.profile_run and not args.quick:
        forms_out = {"batched": run_batched(wf, ctx, scenes, args),
                     "served": run_served(wf, ctx, scenes, argparse.Namespace(
                         jobs=max(3, args.jobs * 9 // 16), served_spp=args.served_spp,
                         served_capacity=args.served_capacity))}

    out = {
        "metric": "Mrays/s @1080p depth 12",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator)",
        "config": {"workload": desc, "policy": args.policy, "schedule": args.schedule,
                   "sample_lanes": args.lanes, "wavefront": args.wavefront or "auto",
                   "ray_order": (f"top {args.ray_key_bits} bits of RayQueue::sortKey"
                                 if args.ray_key_bits else "none (queue order)"),
                   "full_ray_key": full_key,
                   "tuning": ctx.tuning(), "autotuned": tuned is not None,
                   "bvh_width": scene.stats.bvh_width,
                   "other_schedule": other_sched, "physics":
                   "unbiased" if physics else "reference", "triangles": data.triangle_count,
                   "bvh_build_ms": round(scene.stats.build_ms, 1), "policies": policies},
        "roofline": roofline,
        "traversal": tstats,
        "image": image,
        "forms": forms_out,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
