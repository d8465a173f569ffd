"""MI355X wavefront path-tracing hot path: Python host mirror of the C-ABI.

The compute path is libwavefront_hip.so (hand-written HIP kernels for gfx950); this module only
moves data and calls it. Names follow the reference renderer's surfaces: `RayQueue` and
`ShadingQueue` (include/scheduling/*.h) as device structure-of-arrays, `Scene` + accelerator
(src/core/Scene.cpp, src/accelerator/Embree.cpp), and `render` (WavefrontRenderer::renderScene).
"""
from __future__ import annotations

import ctypes as C
import weakref
from typing import Dict, Optional

import numpy as np

from . import _abi
from ._abi import (  # noqa: F401  (re-exported constants)
    MAT_DIFFUSE, MAT_EMISSIVE, MAT_GLASS, MAT_METAL, MAT_PLASTIC, MAT_SPOTLIGHT,
    MODEL_MICROFACET, MODEL_REFERENCE, PHYSICS_REFERENCE, PHYSICS_UNBIASED, POLICIES,
    POLICY_COST_BENEFIT, POLICY_MATERIAL, POLICY_NONE, POLICY_TEXTURE, RAY_DTYPES, RAY_FIELDS,
    SHADE_DTYPES, SHADE_FIELDS, HIT_DTYPES, HIT_FIELDS, TEX_RGBA8_SRGB, TEX_RGB_F32,
    WavefrontError, check, load_library,
)


def lib():
    return load_library()


class Context:
    """wf_context: one device, one HIP stream, grow-only workspaces."""

    def __init__(self, device: int = 0):
        self._lib = lib()
        h = C.c_void_p()
        check(self._lib.wf_context_create(device, C.byref(h)))
        self.handle = h
        self.device = device
        self._scenes = weakref.WeakSet()
        self._engines = weakref.WeakSet()

    def close(self):
        if self.handle:
            for en in list(self._engines):  # engines first: they run over the scenes
                en.close()
            for sc in list(self._scenes):   # scenes next: they were created on this context
                sc.close()
            self._lib.wf_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        check(self._lib.wf_device_alloc(self.handle, max(int(nbytes), 1), C.byref(p)))
        return p.value

    def free(self, ptr: int):
        check(self._lib.wf_device_free(self.handle, C.c_void_p(ptr)))

    def htod(self, ptr: int, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        check(self._lib.wf_memcpy_htod(self.handle, C.c_void_p(ptr),
                                       arr.ctypes.data_as(C.c_void_p), arr.nbytes))

    def dtoh(self, ptr: int, count: int, dtype) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if count:
            check(self._lib.wf_memcpy_dtoh(self.handle, out.ctypes.data_as(C.c_void_p),
                                           C.c_void_p(ptr), out.nbytes))
        return out

    def fill_u32(self, ptr: int, value: int, count: int):
        check(self._lib.wf_fill_u32(self.handle, C.c_void_p(ptr), value, count))

    def tuning(self) -> dict:
        t = _abi.Tuning()
        check(self._lib.wf_context_get_tuning(self.handle, C.byref(t)))
        return {"refill_lanes": t.refill_lanes, "trace_blocks_per_cu": t.trace_blocks_per_cu,
                "bvh_builder": t.bvh_builder, "ploc_radius": t.ploc_radius,
                "bvh_width": t.bvh_width, "radix_bits": t.radix_bits,
                "shadow_blocks_per_cu": t.shadow_blocks_per_cu,
                "shadow_refill_lanes": t.shadow_refill_lanes, "ray_key_bits": t.ray_key_bits}

    def set_tuning(self, **kw):
        cur = self.tuning()
        cur.update(kw)
        t = _abi.Tuning(cur["refill_lanes"], cur["trace_blocks_per_cu"], cur["bvh_builder"],
                        cur["ploc_radius"], cur["bvh_width"], cur["radix_bits"],
                        cur["shadow_blocks_per_cu"], cur["shadow_refill_lanes"],
                        cur["ray_key_bits"])
        check(self._lib.wf_context_set_tuning(self.handle, C.byref(t)))

    def stream_copy(self, dst: int, src: int, nbytes: int):
        check(self._lib.wf_stream_copy(self.handle, C.c_void_p(dst), C.c_void_p(src), nbytes))

    def synchronize(self):
        check(self._lib.wf_synchronize(self.handle))

    def profile(self, enable: bool):
        check(self._lib.wf_profile_enable(self.handle, 1 if enable else 0))

    def profile_filter(self, name: Optional[str]):
        check(self._lib.wf_profile_filter(self.handle, name.encode() if name else None))

    def profile_names(self):
        buf = C.create_string_buffer(1 << 16)
        check(self._lib.wf_profile_names(self.handle, buf, len(buf)))
        return [n for n in buf.value.decode().split("\n") if n]

    def profile_query(self, name: str):
        ms, launches, nbytes = C.c_double(), C.c_int64(), C.c_double()
        check(self._lib.wf_profile_query(self.handle, name.encode(), C.byref(ms),
                                         C.byref(launches), C.byref(nbytes)))
        return ms.value, launches.value, nbytes.value


class DeviceBuffer:
    def __init__(self, ctx: Context, count: int, dtype):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.ptr = ctx.alloc(self.count * self.dtype.itemsize)

    @classmethod
    def from_array(cls, ctx: Context, arr: np.ndarray) -> "DeviceBuffer":
        b = cls(ctx, arr.size, arr.dtype)
        if arr.size:
            ctx.htod(b.ptr, arr.ravel())
        return b

    def upload(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr, dtype=self.dtype).ravel()
        assert arr.size <= self.count, "upload larger than the buffer"
        if arr.size:
            self.ctx.htod(self.ptr, arr)

    def download(self, count: Optional[int] = None) -> np.ndarray:
        return self.ctx.dtoh(self.ptr, self.count if count is None else int(count), self.dtype)

    def free(self):
        if self.ptr:
            self.ctx.free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _DeviceSoA:
    FIELDS: list = []
    DTYPES: dict = {}
    STRUCT = None

    def __init__(self, ctx: Context, capacity: int):
        self.ctx = ctx
        self.capacity = int(capacity)
        self.arrays: Dict[str, DeviceBuffer] = {
            k: DeviceBuffer(ctx, max(self.capacity, 1), self.DTYPES[k]) for k in self.FIELDS}

    def soa(self):
        return self.STRUCT(*[C.c_void_p(self.arrays[k].ptr) for k in self.FIELDS])

    def upload(self, host: Dict[str, np.ndarray]):
        for k in self.FIELDS:
            if k in host:
                self.arrays[k].upload(np.asarray(host[k]))

    def download(self, n: int) -> Dict[str, np.ndarray]:
        return {k: self.arrays[k].download(n) for k in self.FIELDS}


class RayQueue(_DeviceSoA):
    """Device RayQueue (include/scheduling/RayQueue.h:63-132) without the dead accumLight/alive."""
    FIELDS = _abi.RAY_FIELDS
    DTYPES = _abi.RAY_DTYPES
    STRUCT = _abi.RaySoA


class ShadingQueue(_DeviceSoA):
    """Device ShadingQueue (include/scheduling/ShadingQueue.h:27-53) plus geomID and frontFace."""
    FIELDS = _abi.SHADE_FIELDS
    DTYPES = _abi.SHADE_DTYPES
    STRUCT = _abi.ShadeSoA


class HitBuffer(_DeviceSoA):
    FIELDS = _abi.HIT_FIELDS
    DTYPES = _abi.HIT_DTYPES
    STRUCT = _abi.HitSoA


class Scene:
    """wf_scene: device copy of a scene + its accelerator (Scene::buildAccelerator)."""

    def __init__(self, ctx: Context, data, build: bool = True):
        self.ctx = ctx
        self.data = data
        desc, self._keep = data.desc()
        h = C.c_void_p()
        check(ctx._lib.wf_scene_create(ctx.handle, C.byref(desc), C.byref(h)))
        self.handle = h
        ctx._scenes.add(self)
        self.stats = None
        if build:
            self.build()

    def build(self):
        st = _abi.BuildStats()
        check(self.ctx._lib.wf_accel_build(self.handle, C.byref(st)))
        self.stats = st
        return st

    def device_bytes(self) -> int:
        n = C.c_size_t()
        check(self.ctx._lib.wf_scene_device_bytes(self.handle, C.byref(n)))
        return n.value

    def close(self):
        if getattr(self, "handle", None):
            self.ctx._lib.wf_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_costs(self, material_rel_cost=None, texture_rel_cost=None):
        mc = None if material_rel_cost is None else np.ascontiguousarray(material_rel_cost, np.float64)
        tc = None if texture_rel_cost is None else np.ascontiguousarray(texture_rel_cost, np.float64)
        check(self.ctx._lib.wf_scene_set_costs(
            self.handle, None if mc is None else mc.ctypes.data_as(C.c_void_p),
            None if tc is None else tc.ctypes.data_as(C.c_void_p)))

    def track_costs(self, materials: "CostTracker | None", textures: "CostTracker | None"):
        """wf_scene_track_costs: live shading-cost tracking of the pass loop into these trackers
        (None, None detaches)."""
        self._trackers = (materials, textures)   # the library keeps pointers to them
        check(self.ctx._lib.wf_scene_track_costs(
            self.handle, None if materials is None else materials.handle,
            None if textures is None else textures.handle))

    # EmbreeAccelerator::intersect4 over a batch
    def intersect(self, rays: RayQueue, n: int, tnear: float = 0.001, tfar: float = 1e9):
        hits = HitBuffer(self.ctx, n)
        r, h = rays.soa(), hits.soa()
        check(self.ctx._lib.wf_intersect(self.handle, C.byref(r), n, tnear, tfar, C.byref(h)))
        return hits.download(n)

    def trace_stats(self, rays: RayQueue, n: int, tnear: float = 0.001, tfar: float = 1e9,
                    any_hit: bool = False):
        """(nodes visited, triangles tested) per ray; lanes of a wave are rays 64w..64w+63."""
        nodes = DeviceBuffer(self.ctx, max(n, 1), np.int32)
        tris = DeviceBuffer(self.ctx, max(n, 1), np.int32)
        r = rays.soa()
        check(self.ctx._lib.wf_trace_stats(self.handle, C.byref(r), n, tnear, tfar,
                                           1 if any_hit else 0, C.c_void_p(nodes.ptr),
                                           C.c_void_p(tris.ptr)))
        return nodes.download(n), tris.download(n)

    # EmbreeAccelerator::intersect4 with the HitRecord mapping, one record per input ray
    def intersect_records(self, rays: RayQueue, n: int, tnear: float = 0.001, tfar: float = 1e9):
        rec = ShadingQueue(self.ctx, n)
        flags = DeviceBuffer(self.ctx, max(n, 1), np.uint8)
        r, q = rays.soa(), rec.soa()
        check(self.ctx._lib.wf_intersect_records(self.handle, C.byref(r), n, tnear, tfar,
                                                 C.byref(q), C.c_void_p(flags.ptr)))
        return rec.download(n), flags.download(n).astype(bool)

    # EmbreeAccelerator::occluded over a batch
    def occluded(self, origins: np.ndarray, dirs: np.ndarray, max_dist: np.ndarray) -> np.ndarray:
        n = len(max_dist)
        bufs = [DeviceBuffer.from_array(self.ctx, np.ascontiguousarray(a, np.float32))
                for a in (origins[:, 0], origins[:, 1], origins[:, 2], dirs[:, 0], dirs[:, 1],
                          dirs[:, 2], max_dist)]
        out = DeviceBuffer(self.ctx, max(n, 1), np.uint8)
        check(self.ctx._lib.wf_occluded(self.handle, *[C.c_void_p(b.ptr) for b in bufs], n,
                                        C.c_void_p(out.ptr)))
        return out.download(n).astype(bool)

    # WavefrontRenderer::intersectAll
    def intersect_split(self, rays: RayQueue, n: int):
        hits = ShadingQueue(self.ctx, n)
        misses = RayQueue(self.ctx, n)
        nh, nm = C.c_int32(), C.c_int32()
        r, hs, ms = rays.soa(), hits.soa(), misses.soa()
        check(self.ctx._lib.wf_intersect_split(self.handle, C.byref(r), n, C.byref(hs),
                                               C.byref(nh), C.byref(ms), C.byref(nm)))
        return hits, nh.value, misses, nm.value


def ray_schedule(ctx: Context, rays: RayQueue, n: int, bmin, bmax):
    """RayQueue::schedule: returns (sortedIndices, keys) as numpy arrays."""
    idx = DeviceBuffer(ctx, max(n, 1), np.int32)
    keys = DeviceBuffer(ctx, max(n, 1), np.uint64)
    bmin = (C.c_float * 3)(*[float(x) for x in bmin])
    bmax = (C.c_float * 3)(*[float(x) for x in bmax])
    r = rays.soa()
    check(ctx._lib.wf_ray_schedule(ctx.handle, C.byref(r), n, bmin, bmax, C.c_void_p(idx.ptr),
                                   C.c_void_p(keys.ptr)))
    return idx.download(n), keys.download(n)


def materialize_rays(ctx: Context, rays: RayQueue, idx: np.ndarray, n: int) -> RayQueue:
    out = RayQueue(ctx, n)
    d_idx = DeviceBuffer.from_array(ctx, np.ascontiguousarray(idx, np.int32))
    s, d = rays.soa(), out.soa()
    check(ctx._lib.wf_materialize_rays(ctx.handle, C.byref(s), C.c_void_p(d_idx.ptr), n, C.byref(d)))
    return out


def shading_schedule(scene: Scene, q: ShadingQueue, n: int, policy) -> np.ndarray:
    """ShadingQueue::schedule: returns sortedIndices."""
    pol = POLICIES[policy] if isinstance(policy, str) else int(policy)
    idx = DeviceBuffer(scene.ctx, max(n, 1), np.int32)
    s = q.soa()
    check(scene.ctx._lib.wf_shading_schedule(scene.handle, C.byref(s), n, pol, C.c_void_p(idx.ptr)))
    return idx.download(n)


def materialize_shading(ctx: Context, q: ShadingQueue, idx: np.ndarray, n: int) -> ShadingQueue:
    out = ShadingQueue(ctx, n)
    d_idx = DeviceBuffer.from_array(ctx, np.ascontiguousarray(idx, np.int32))
    s, d = q.soa(), out.soa()
    check(ctx._lib.wf_materialize_shading(ctx.handle, C.byref(s), C.c_void_p(d_idx.ptr), n,
                                          C.byref(d)))
    return out


class CostTracker:
    """CostTracker (include/render/CostTracker.h:8-63) held by the library: per item an EMA of
    the shading cost in ns; relative() is relativeCost. Attach a material and a texture tracker
    to a scene with Scene.track_costs to have wf_render's pass loop time its shading and refresh
    the scene's cost tables every bounce."""

    def __init__(self, items: int, library=None):
        self._lib = library if library is not None else lib()
        self.items = int(items)
        h = C.c_void_p()
        check(self._lib.wf_cost_tracker_create(self.items, C.byref(h)))
        self.handle = h

    def record(self, item: int, cost_ns: float):
        check(self._lib.wf_cost_tracker_record(self.handle, int(item), float(cost_ns)))

    def query(self):
        """(average_ns, samples, relative) arrays."""
        avg = np.zeros(max(self.items, 1), np.float64)
        cnt = np.zeros(max(self.items, 1), np.int32)
        rel = np.zeros(max(self.items, 1), np.float64)
        check(self._lib.wf_cost_tracker_query(self.handle, avg.ctypes.data_as(C.c_void_p),
                                              cnt.ctypes.data_as(C.c_void_p),
                                              rel.ctypes.data_as(C.c_void_p)))
        return avg[:self.items], cnt[:self.items], rel[:self.items]

    def relative(self) -> np.ndarray:
        return self.query()[2]

    def close(self):
        if self.handle:
            self._lib.wf_cost_tracker_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


COHERENCE_OFF, COHERENCE_RUN_LENGTH, COHERENCE_CACHE_LINES = 0, 1, 2


def coherence_stats(ctx: Context, material_ids: np.ndarray, texture_ids: np.ndarray,
                    cache_lines: bool = True) -> tuple:
    """The coherence metrics of one shading queue in its order (renderScene's batch statistics,
    src/render/WavefrontRenderer.cpp:67-92, 201-246): (material run length, texture run length,
    material cache-line homogeneity, texture cache-line homogeneity)."""
    n = int(len(material_ids))
    out = (C.c_double * 4)()
    if n == 0:
        check(ctx._lib.wf_coherence_stats(ctx.handle, None, None, 0, 1 if cache_lines else 0,
                                          out))
        return tuple(out)
    m = DeviceBuffer.from_array(ctx, np.ascontiguousarray(material_ids, np.int32))
    t = DeviceBuffer.from_array(ctx, np.ascontiguousarray(texture_ids, np.int32))
    check(ctx._lib.wf_coherence_stats(ctx.handle, C.c_void_p(m.ptr), C.c_void_p(t.ptr), n,
                                      1 if cache_lines else 0, out))
    return tuple(out)


def shade_params(physics=PHYSICS_REFERENCE, rr_min_depth=3, cost_aware_rr=False,
                 firefly=10.0, seed=0, sample=0, job=0) -> _abi.ShadeParams:
    return _abi.ShadeParams(physics, rr_min_depth, 1 if cost_aware_rr else 0, firefly, seed,
                            sample, job)


def shade(scene: Scene, q: ShadingQueue, n: int, params: _abi.ShadeParams, accum: DeviceBuffer):
    """shadeAll for one bounce. Returns (next RayQueue, n_next)."""
    nxt = RayQueue(scene.ctx, max(n, 1))
    nn = C.c_int32()
    s, d = q.soa(), nxt.soa()
    check(scene.ctx._lib.wf_shade(scene.handle, C.byref(s), n, C.byref(params),
                                  C.c_void_p(accum.ptr), C.byref(d), C.byref(nn)))
    return nxt, nn.value


def shade_misses(scene: Scene, q: RayQueue, n: int, accum: DeviceBuffer, physics=PHYSICS_REFERENCE):
    s = q.soa()
    check(scene.ctx._lib.wf_shade_misses(scene.handle, C.byref(s), n, physics,
                                         C.c_void_p(accum.ptr)))


SCHEDULES = {"regenerate": _abi.SCHEDULE_REGENERATE, "passes": _abi.SCHEDULE_PASSES}


def camera_struct(camera) -> _abi.Camera:
    return _abi.Camera(_abi.f3(*camera["eye"]), _abi.f3(*camera["target"]),
                       _abi.f3(*camera.get("up", (0.0, 1.0, 0.0))), float(camera["vfov"]))


def render_params(width, height, spp, max_depth, camera, policy="none", ray_sort=True,
                  adaptive=False, adaptive_threshold=0.05, adaptive_min_samples=32,
                  adaptive_interval=16, physics=PHYSICS_REFERENCE, rr_min_depth=3,
                  cost_aware_rr=False, firefly=10.0, seed=0, job=0,
                  schedule="regenerate", sample_lanes=0, wavefront=0,
                  coherence_stats=0, first_sample=0) -> _abi.RenderParams:
    """Regenerating schedule: sample_lanes = samples of a pixel in flight at once (a power of
    two <= 64, 0 for the default 64; 1 reproduces the pass loop's accumulation order bit for
    bit); wavefront = ray slots (0 = automatic). coherence_stats: 0 off, 1 material/texture run
    lengths, 2 also cache-line homogeneity (COHERENCE_*), returned in the stats."""
    pol = POLICIES[policy] if isinstance(policy, str) else int(policy)
    sched = SCHEDULES[schedule] if isinstance(schedule, str) else int(schedule)
    sp = shade_params(physics, rr_min_depth, cost_aware_rr, firefly, seed, 0, job)
    return _abi.RenderParams(width, height, spp, max_depth, pol, 1 if ray_sort else 0,
                             1 if adaptive else 0, adaptive_threshold, adaptive_min_samples,
                             adaptive_interval, sched, sample_lanes, wavefront,
                             int(coherence_stats), int(first_sample), sp, camera_struct(camera))


def render(scene: Scene, params: _abi.RenderParams, keep_on_device: bool = False):
    """WavefrontRenderer::renderScene. Returns (image HxWx3 float32, counts HxW, stats)."""
    P = params.width * params.height
    img = DeviceBuffer(scene.ctx, P * 3, np.float32)
    cnt = DeviceBuffer(scene.ctx, P, np.int32)
    st = _abi.RenderStats()
    check(scene.ctx._lib.wf_render(scene.handle, C.byref(params), C.c_void_p(img.ptr),
                                   C.c_void_p(cnt.ptr), C.byref(st)))
    if keep_on_device:
        return img, cnt, st
    image = img.download().reshape(params.height, params.width, 3)
    counts = cnt.download().reshape(params.height, params.width)
    return image, counts, st


# ---- batched / served forms -----------------------------------------------------------------
def job_desc(width, height, spp, max_depth, camera, seed=0, job_tag=0, adaptive=False,
             adaptive_threshold=0.05, adaptive_min_samples=32,
             adaptive_interval=16, first_sample=0) -> _abi.JobDesc:
    return _abi.JobDesc(width, height, spp, max_depth, seed, job_tag, 1 if adaptive else 0,
                        adaptive_threshold, adaptive_min_samples, adaptive_interval,
                        int(first_sample), camera_struct(camera))


class Engine:
    """Job-tagged regenerating wavefront over one resident scene (wf_engine_*)."""

    def __init__(self, ctx: Context, scene: Scene, capacity: int, pixel_budget: int,
                 max_jobs: int = 64, policy="none", ray_sort=True, use_graph=False,
                 physics=PHYSICS_REFERENCE, rr_min_depth=3, cost_aware_rr=False,
                 firefly=10.0, sample_lanes=0):
        self.ctx = ctx
        self.scene = scene
        self._lib = ctx._lib
        pol = POLICIES[policy] if isinstance(policy, str) else int(policy)
        cfg = _abi.EngineConfig(capacity, max_jobs, pixel_budget, pol, 1 if ray_sort else 0,
                                1 if use_graph else 0, sample_lanes,
                                shade_params(physics, rr_min_depth, cost_aware_rr, firefly,
                                             0, 0, 0))
        h = C.c_void_p()
        check(self._lib.wf_engine_create(ctx.handle, scene.handle, C.byref(cfg), C.byref(h)))
        self.handle = h.value
        self.jobs = {}
        self.lanes = sample_lanes or 64     # the lanes a lone render needs to match its jobs
        ctx._engines.add(self)

    def admit(self, job: _abi.JobDesc) -> int:
        jid = C.c_int32()
        check(self._lib.wf_engine_admit(self.handle, C.byref(job), C.byref(jid)))
        self.jobs[jid.value] = (job.width, job.height)
        return jid.value

    def try_admit(self, job: _abi.JobDesc) -> Optional[int]:
        jid = C.c_int32()
        st = self._lib.wf_engine_admit(self.handle, C.byref(job), C.byref(jid))
        if st == _abi.WF_ERR_ADMISSION:
            return None
        check(st)
        self.jobs[jid.value] = (job.width, job.height)
        return jid.value

    def run(self, max_iterations: int = 0, cap: int = 64) -> list:
        buf = (C.c_int32 * cap)()
        n = C.c_int32()
        check(self._lib.wf_engine_run(self.handle, max_iterations, buf, cap, C.byref(n)))
        return list(buf[:n.value])

    def launch(self, iterations: int = 1):
        check(self._lib.wf_engine_launch(self.handle, iterations))

    def wait(self, cap: int = 256) -> list:
        buf = (C.c_int32 * cap)()
        n = C.c_int32()
        check(self._lib.wf_engine_wait(self.handle, buf, cap, C.byref(n)))
        return list(buf[:n.value])

    def result(self, job_id: int, keep_on_device: bool = False):
        w, h = self.jobs.pop(job_id)
        img = DeviceBuffer(self.ctx, w * h * 3, np.float32)
        cnt = DeviceBuffer(self.ctx, w * h, np.int32)
        res = _abi.JobResult()
        check(self._lib.wf_engine_result(self.handle, job_id, C.c_void_p(img.ptr),
                                         C.c_void_p(cnt.ptr), C.byref(res)))
        if keep_on_device:
            return img, cnt, res
        return img.download().reshape(h, w, 3), cnt.download().reshape(h, w), res

    def stats(self) -> _abi.EngineStats:
        st = _abi.EngineStats()
        check(self._lib.wf_engine_query(self.handle, C.byref(st)))
        return st

    def device_bytes(self) -> int:
        out = C.c_size_t()
        check(self._lib.wf_engine_device_bytes(self.handle, C.byref(out)))
        return out.value

    def phase_timing(self, enable: bool):
        check(self._lib.wf_engine_phase_timing(self.handle, 1 if enable else 0))

    def shading_queue(self, capacity: int) -> Dict[str, np.ndarray]:
        """The last iteration's shading queue in shading order (wf_engine_shading_queue)."""
        q = ShadingQueue(self.ctx, max(1, capacity))
        n = C.c_int32()
        soa = q.soa()
        check(self._lib.wf_engine_shading_queue(self.handle, C.byref(soa), capacity,
                                                C.byref(n)))
        return q.download(min(n.value, capacity))

    def close(self):
        if getattr(self, "handle", None):
            self._lib.wf_engine_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def engine_bytes(capacity: int, pixel_budget: int, sample_lanes: int = 0) -> int:
    out = C.c_size_t()
    check(_abi.load_library().wf_engine_bytes(capacity, pixel_budget, sample_lanes,
                                              C.byref(out)))
    return out.value


def simd_efficiency(per_ray: np.ndarray, wave: int = 64) -> float:
    """Useful lane-iterations / issued lane-iterations when a wave loops until its slowest lane
    finishes (per_ray: loop trip count of each ray, waves = consecutive groups of 64)."""
    n = len(per_ray)
    if n == 0:
        return 1.0
    pad = (-n) % wave
    x = np.concatenate([per_ray.astype(np.int64), np.zeros(pad, np.int64)]).reshape(-1, wave)
    issued = x.max(axis=1).sum() * wave
    return float(x.sum() / issued) if issued else 1.0
