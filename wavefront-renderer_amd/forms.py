"""Batched and served forms over device-resident scenes (one wavefront engine per scene).

Batched: independent render jobs drawn by seed from the Cornell-box and sorting-stress
generators (resolution log-uniform 128^2..1024^2, spp in {4, 16, 64, 256}, depth 8..12), all
admitted at once and merged into each scene's job-tagged wavefront; compared with the same jobs
rendered one after another with wf_render.

Served: a Poisson stream of requests against the three resident scenes (Cornell box,
sorting stress, sphere cloud), resolution lognormal (median 512^2, clamped to 64^2..1920x1080),
adaptive sampling (at most 2048 spp) towards a per-pixel relative standard error of
sqrt(target relMSE) / 1.25;
requests are admitted as job slots and framebuffer ranges free up and their paths share the
scene's wavefront while other requests run. Reports jobs/s, Mrays/s and p50/p99
time-to-converged-image, next to the same requests served by sequential wf_render calls (their
measured service times through a first-come-first-served queue with the same arrivals).

Every image is identical to the job rendered alone with the same sample lanes
(tests/test_gpu_engine.py).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import PHYSICS_REFERENCE, PHYSICS_UNBIASED, Engine, job_desc, render, render_params


@dataclass
class Job:
    scene: int
    width: int
    height: int
    spp: int
    max_depth: int
    seed: int
    tag: int
    adaptive: bool = False
    threshold: float = 0.05
    min_samples: int = 16
    interval: int = 16
    arrival: float = 0.0
    done: float = field(default=-1.0)

    @property
    def pixels(self) -> int:
        return self.width * self.height


def batched_jobs(n: int = 64, seed: int = 0x5EED0003) -> list:
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        side = int(round(128.0 * 8.0 ** rng.random()))
        jobs.append(Job(scene=int(rng.integers(0, 2)), width=side, height=side,
                        spp=int(rng.choice([4, 16, 64, 256])), max_depth=int(rng.integers(8, 13)),
                        seed=1000 + i, tag=i))
    return jobs


def served_jobs(n: int, rate: float, n_scenes: int = 3, seed: int = 0x5EED0004,
                rel_mse: float = 1e-3, max_spp: int = 2048) -> list:
    rng = np.random.default_rng(seed)
    t = 0.0
    # AdaptiveSampler's per-pixel rule (relative standard error below the threshold) at
    # sqrt(rel_mse) leaves the image at 1.3-1.5 times the target (pixels stopped by the sample
    # cap, the noisy variance estimate, the +0.01 in the rule's denominator): the threshold is
    # sqrt(rel_mse) / 1.25, which brings the Cornell-box and sorting-stress requests below it
    thr = float(np.sqrt(rel_mse)) / 1.25
    jobs = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        side = 512.0 * np.exp(0.5 * rng.normal())
        jobs.append(Job(scene=int(rng.integers(0, n_scenes)),
                        width=int(min(1920, max(64, round(side)))),
                        height=int(min(1080, max(64, round(side)))), spp=max_spp, max_depth=12,
                        seed=2000 + i, tag=i, adaptive=True, threshold=thr, min_samples=16,
                        interval=16, arrival=t))
    return jobs


def _desc(job: Job, camera):
    return job_desc(job.width, job.height, job.spp, job.max_depth, camera, seed=job.seed,
                    job_tag=job.tag, adaptive=job.adaptive, adaptive_threshold=job.threshold,
                    adaptive_min_samples=job.min_samples, adaptive_interval=job.interval)


def _params(job: Job, camera, physics, ray_sort=False):
    return render_params(job.width, job.height, job.spp, job.max_depth, camera, seed=job.seed,
                         job=job.tag, adaptive=job.adaptive, adaptive_threshold=job.threshold,
                         adaptive_min_samples=job.min_samples, adaptive_interval=job.interval,
                         physics=physics, firefly=1e30, ray_sort=ray_sort)


def sequential(ctx, scenes, datas, physics, jobs, ray_sort=False) -> tuple:
    """wf_render of every job in turn: (per-job seconds, rays)."""
    secs, rays = [], 0
    for j in jobs:
        ctx.synchronize()
        t0 = time.perf_counter()
        img, cnt, st = render(scenes[j.scene],
                              _params(j, datas[j.scene].camera, physics[j.scene], ray_sort),
                              keep_on_device=True)
        ctx.synchronize()
        secs.append(time.perf_counter() - t0)
        rays += st.rays_traced + st.shadow_rays
    return secs, rays


def make_engines(ctx, scenes, physics, capacity, pixel_budget, max_jobs=64, ray_sort=False):
    """One engine per resident scene. Without ray ordering by default: the queue order of the
    wavefront changes no result, and ordering each bounce costs more than it saves."""
    return [Engine(ctx, sc, capacity=capacity, pixel_budget=pixel_budget, max_jobs=max_jobs,
                   physics=ph, firefly=1e30, use_graph=True, ray_sort=ray_sort)
            for sc, ph in zip(scenes, physics)]


def rel_mse(img, ref, trim: float = 0.0) -> float:
    """Mean over pixels of the channel mean of (I - R)^2 / (R^2 + 0.01); with trim > 0 the
    largest `trim` fraction of the per-pixel errors is left out (a view robust to the few
    firefly pixels of either image)."""
    a = np.asarray(img, np.float64)
    r = np.asarray(ref, np.float64)
    e = ((a - r) ** 2 / (r * r + 1e-2)).reshape(-1, a.shape[-1]).mean(axis=1)
    if trim > 0.0:
        e = np.sort(e)[:max(1, int(round(len(e) * (1.0 - trim))))]
    return float(e.mean())


def check_converged(scenes, datas, physics, jobs, images, per_scene: int = 2,
                    ref_spp: int = 4096) -> dict:
    """relMSE of served results against a fixed-budget render of the same request with another
    seed at ref_spp samples per pixel (whose own noise adds about spp/ref_spp of the value), and
    the same with the 0.1% largest per-pixel errors left out."""
    picked, seen = [], {}
    for k, j in enumerate(jobs):
        if seen.get(j.scene, 0) < per_scene:
            seen[j.scene] = seen.get(j.scene, 0) + 1
            picked.append(k)
    vals, trimmed, spps = [], [], []
    for k in picked:
        j = jobs[k]
        p = render_params(j.width, j.height, ref_spp, j.max_depth, datas[j.scene].camera,
                          seed=j.seed + 7919, job=j.tag, physics=physics[j.scene], firefly=1e30)
        ref, _, _ = render(scenes[j.scene], p)
        img, cnt, _ = images[k]
        vals.append(rel_mse(img, ref))
        trimmed.append(rel_mse(img, ref, trim=1e-3))
        spps.append(float(np.mean(cnt)))
    return {"checked": len(vals), "scenes": [jobs[k].scene for k in picked],
            "reference_spp": ref_spp, "relmse": [float(f"{v:.4g}") for v in vals],
            "relmse_trim_0.1pct": [float(f"{v:.4g}") for v in trimmed],
            "mean_spp": [round(s, 1) for s in spps],
            "within_target": sum(v <= 1e-3 for v in vals),
            "within_target_trimmed": sum(v <= 1e-3 for v in trimmed)}


def serve(engines, datas, jobs, iterations: int = 2, realtime: bool = True,
          collect: bool = False) -> dict:
    """Admits jobs at their arrival times (all at once when realtime is False), runs every engine
    with work, collects finished jobs. Returns makespan, rays and per-job completion times (and,
    with collect, every job's (image, counts, result) on the host, in job order).

    Each engine is driven on its own: once its launch has been waited for, its finished jobs are
    collected, waiting requests of its scene admitted and its next launch issued, while the other
    engines' launches keep the GPU busy."""
    order = sorted(range(len(jobs)), key=lambda k: jobs[k].arrival)
    pending = [[] for _ in engines]   # arrived, not admitted yet (FIFO per scene)
    slot = [dict() for _ in engines]
    inflight = [False] * len(engines)
    images = {}
    done = 0
    nxt = 0
    r0 = sum(e.stats().rays_traced + e.stats().shadow_rays for e in engines)
    t0 = time.perf_counter()
    while done < len(jobs):
        now = time.perf_counter() - t0
        while nxt < len(order) and (not realtime or jobs[order[nxt]].arrival <= now):
            pending[jobs[order[nxt]].scene].append(order[nxt])
            nxt += 1
        for i, e in enumerate(engines):
            if inflight[i]:
                for jid in e.wait():
                    k = slot[i].pop(jid)
                    out = e.result(jid, keep_on_device=not collect)
                    if collect:
                        images[k] = out
                    jobs[k].done = time.perf_counter() - t0
                    done += 1
                inflight[i] = False
            still = []
            for k in pending[i]:
                j = jobs[k]
                jid = e.try_admit(_desc(j, datas[i].camera))
                if jid is None:
                    still.append(k)
                else:
                    slot[i][jid] = k
            pending[i] = still
            if slot[i]:
                e.launch(iterations)
                inflight[i] = True
        if not any(inflight) and done < len(jobs) and nxt < len(order):
            time.sleep(max(0.0, jobs[order[nxt]].arrival - (time.perf_counter() - t0)))
    makespan = time.perf_counter() - t0
    rays = sum(e.stats().rays_traced + e.stats().shadow_rays for e in engines) - r0
    res = {"makespan_s": makespan, "rays": rays}
    if collect:
        res["images"] = [images[k] for k in range(len(jobs))]
    return res


def fcfs_latency(arrivals, service) -> np.ndarray:
    """Completion minus arrival for a single first-come-first-served server."""
    t, out = 0.0, []
    for a, s in sorted(zip(arrivals, service)):
        t = max(t, a) + s
        out.append(t - a)
    return np.array(out)


def pct(x, q) -> float:
    return float(np.percentile(np.asarray(x), q)) if len(x) else 0.0
