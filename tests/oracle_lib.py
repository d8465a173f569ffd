"""ctypes bindings of the CPU oracle (oracle/liborc.so) and of the compiled reference harness
(oracle/_ref/libref.so). Test infrastructure only: the checker, never the thing measured."""
from __future__ import annotations

import ctypes as C
import importlib
import os
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
# WF_ORACLE_LIB points the tests at another build of the oracle (the planted-mistake checks)
ORC_PATH = Path(os.environ.get("WF_ORACLE_LIB", ROOT / "oracle" / "liborc.so"))
REF_PATH = ROOT / "oracle" / "_ref" / "libref.so"
abi = importlib.import_module("wavefront-renderer_amd._abi")

P = C.c_void_p
FP = C.POINTER(C.c_float)
IP = C.POINTER(C.c_int32)


def fptr(a):
    return a.ctypes.data_as(P)


class Oracle:
    def __init__(self, path: Path = ORC_PATH):
        if not path.exists():
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        self.lib = L = C.CDLL(str(path))
        L.orc_philox.argtypes = [P, P, P]
        L.orc_uniform.restype = C.c_float
        L.orc_uniform.argtypes = [C.c_uint32] * 6
        L.orc_ray_sort_keys.argtypes = [C.c_int32] + [P] * 8 + [P]
        L.orc_ray_schedule.argtypes = [C.c_int32] + [P] * 8 + [P]
        L.orc_hit_morton.restype = C.c_uint64
        L.orc_hit_morton.argtypes = [C.c_float] * 9
        L.orc_shading_schedule.restype = C.c_int32
        L.orc_shading_schedule.argtypes = ([C.c_int32, C.c_int32] + [P] * 8 + [C.c_int32, P, P,
                                           C.c_int32, P, P, P, P])
        L.orc_scene_create.restype = P
        L.orc_scene_create.argtypes = [C.POINTER(abi.SceneDesc)]
        L.orc_scene_destroy.argtypes = [P]
        L.orc_scene_info.argtypes = [P, P, P, P, P]
        L.orc_intersect_triangle.restype = C.c_int32
        L.orc_intersect_triangle.argtypes = [P] * 5 + [C.c_float, C.c_float, P, P, P]
        L.orc_closest_hit.argtypes = ([P, C.c_int32] + [P] * 6 + [C.c_float, C.c_float, C.c_int32]
                                      + [P] * 5)
        L.orc_occluded.argtypes = [P, C.c_int32] + [P] * 7 + [C.c_int32, P]
        L.orc_raygen.argtypes = ([C.POINTER(abi.Camera), C.c_int32, C.c_int32, C.c_int32, P,
                                  C.c_uint32, C.c_int32, C.c_int32] + [P] * 6)
        L.orc_intersect_split.argtypes = [P, C.c_int32, C.POINTER(abi.RaySoA),
                                          C.POINTER(abi.ShadeSoA), P, C.POINTER(abi.RaySoA), P,
                                          C.c_int32]
        L.orc_shade.argtypes = [P, C.POINTER(abi.ShadeSoA), C.c_int32,
                                C.POINTER(abi.ShadeParams), P, P, P, C.POINTER(abi.RaySoA), P,
                                P, C.c_int32]
        L.orc_shade_misses.argtypes = [P, C.POINTER(abi.RaySoA), C.c_int32, C.c_int32, P, P]
        L.orc_adaptive_add.argtypes = [P, P, P, C.c_int32, C.c_float]
        L.orc_adaptive_update.argtypes = [C.c_int32, P, P, P, P, C.c_float, C.c_int32]
        L.orc_render.restype = C.c_int32
        L.orc_render.argtypes = [P, C.POINTER(abi.RenderParams), P, P, P, P, P,
                                 C.POINTER(abi.RenderStats), C.c_int32]
        L.orc_render_window.restype = C.c_int32
        L.orc_render_window.argtypes = [P, C.POINTER(abi.RenderParams)] + [C.c_int32] * 4 + \
            [P, P, P, C.POINTER(abi.RenderStats), C.c_int32]
        L.orc_scatter_replay.restype = C.c_int32
        L.orc_scatter_replay.argtypes = [P, C.POINTER(abi.Material), P, P, P, C.c_float,
                                         C.c_float, C.c_int32, P, C.c_int32, P, P, P, P]
        L.orc_emitted.argtypes = [C.POINTER(abi.Material), P, P, P]
        L.orc_surface_color.argtypes = [P, C.POINTER(abi.Material), P, C.c_float, C.c_float, P]
        L.orc_hash_grid_cell.restype = C.c_uint32
        L.orc_hash_grid_cell.argtypes = [C.c_int32] * 3
        L.orc_gather_u32.argtypes = [C.c_int32, P, P, P]
        L.orc_run_length.restype = C.c_double
        L.orc_run_length.argtypes = [C.c_int32, P, P]
        L.orc_line_homogeneity.restype = C.c_double
        L.orc_line_homogeneity.argtypes = [C.c_int32, P]
        L.orc_fm_sincos.argtypes = [C.c_int32, P, P, P]
        for name in ("orc_mis_balance", "orc_mis_power"):
            getattr(L, name).restype = C.c_float
            getattr(L, name).argtypes = [C.c_float, C.c_float]
        L.orc_reciprocal_correction_failures.restype = C.c_int64
        L.orc_reciprocal_correction_failures.argtypes = [C.c_int32, P]
        L.orc_area_to_angle.restype = C.c_float
        L.orc_area_to_angle.argtypes = [C.c_float] * 3
        L.orc_fm_inverse.argtypes = [C.c_int32, P, P, P, P, P]

    # -- RNG
    def philox(self, ctr, key):
        c = np.asarray(ctr, np.uint32)
        k = np.asarray(key, np.uint32)
        out = np.zeros(4, np.uint32)
        self.lib.orc_philox(fptr(c), fptr(k), fptr(out))
        return out

    def fm_sincos(self, x):
        """The shared single-precision sin/cos (wf_fmath.h) over an array."""
        x = np.ascontiguousarray(x, np.float32)
        s, c = np.zeros_like(x), np.zeros_like(x)
        self.lib.orc_fm_sincos(len(x), fptr(x), fptr(s), fptr(c))
        return s, c

    def fm_inverse(self, x, y):
        """The shared asin(x), acos(x) and atan2(y, x) (wf_fmath.h) over arrays."""
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        a, b, t = np.zeros_like(x), np.zeros_like(x), np.zeros_like(x)
        self.lib.orc_fm_inverse(len(x), fptr(x), fptr(y), fptr(a), fptr(b), fptr(t))
        return a, b, t

    def uniform(self, seed, job, pixel, sample, depth, dim):
        return self.lib.orc_uniform(seed, job, pixel, sample, depth, dim)

    # -- scheduling
    def ray_keys(self, rays, bmin, bmax):
        n = len(rays["ox"])
        cols = [np.ascontiguousarray(rays[k], np.float32) for k in ("ox", "oy", "oz", "dx", "dy", "dz")]
        bmin = np.asarray(bmin, np.float32)
        bmax = np.asarray(bmax, np.float32)
        keys = np.zeros(n, np.uint64)
        self.lib.orc_ray_sort_keys(n, *[fptr(c) for c in cols], fptr(bmin), fptr(bmax), fptr(keys))
        return keys

    def ray_schedule(self, rays, bmin, bmax):
        n = len(rays["ox"])
        cols = [np.ascontiguousarray(rays[k], np.float32) for k in ("ox", "oy", "oz", "dx", "dy", "dz")]
        bmin = np.asarray(bmin, np.float32)
        bmax = np.asarray(bmax, np.float32)
        out = np.zeros(n, np.int32)
        self.lib.orc_ray_schedule(n, *[fptr(c) for c in cols], fptr(bmin), fptr(bmax), fptr(out))
        return out

    def coherence(self, mat, tex):
        """computeAverageRunLength of material and texture IDs, cacheLineHomogeneity of both."""
        m = np.ascontiguousarray(mat, np.int32)
        t = np.ascontiguousarray(tex, np.int32)
        n = len(m)
        return (self.lib.orc_run_length(n, fptr(m), fptr(t)),
                self.lib.orc_run_length(n, fptr(t), fptr(t)),
                self.lib.orc_line_homogeneity(n, fptr(m)),
                self.lib.orc_line_homogeneity(n, fptr(t)))

    def hit_morton(self, p, lo, hi):
        return self.lib.orc_hit_morton(*[float(x) for x in (*p, *lo, *hi)])

    def shading_schedule(self, policy, q, mat_type, mat_tex, tex_mem, mat_cost=None, tex_cost=None):
        n = len(q["mat"])
        ints = {k: np.ascontiguousarray(q[k], np.int32) for k in ("mat", "tex")}
        flts = {k: np.ascontiguousarray(q.get(k, np.zeros(n)), np.float32)
                for k in ("tr", "tg", "tb", "px", "py", "pz")}
        mt = np.ascontiguousarray(mat_type, np.int32)
        mx = np.ascontiguousarray(mat_tex, np.int32)
        tm = np.ascontiguousarray(tex_mem, np.uint64)
        mc = None if mat_cost is None else np.ascontiguousarray(mat_cost, np.float64)
        tc = None if tex_cost is None else np.ascontiguousarray(tex_cost, np.float64)
        out = np.zeros(n, np.int32)
        pol = abi.POLICIES[policy] if isinstance(policy, str) else int(policy)
        rc = self.lib.orc_shading_schedule(
            pol, n, fptr(ints["mat"]), fptr(ints["tex"]), *[fptr(flts[k]) for k in
                                                            ("tr", "tg", "tb", "px", "py", "pz")],
            len(mt), fptr(mt), fptr(mx), len(tm), fptr(tm) if len(tm) else None,
            None if mc is None else fptr(mc), None if tc is None else fptr(tc), fptr(out))
        assert rc == 0
        return out

    # -- scene
    def scene(self, data):
        return OracleScene(self, data)

    def intersect_triangle(self, o, d, a, b, c, tmin=0.001, tmax=1e9):
        arrs = [np.asarray(x, np.float32) for x in (o, d, a, b, c)]
        t, u, v = C.c_float(), C.c_float(), C.c_float()
        hit = self.lib.orc_intersect_triangle(*[fptr(x) for x in arrs], tmin, tmax, C.byref(t),
                                              C.byref(u), C.byref(v))
        return bool(hit), t.value, u.value, v.value

    def raygen(self, camera, width, height, pixels, seed, job, sample):
        pixels = np.ascontiguousarray(pixels, np.int32)
        n = len(pixels)
        cols = [np.zeros(n, np.float32) for _ in range(6)]
        cam = abi.Camera(abi.f3(*camera["eye"]), abi.f3(*camera["target"]),
                         abi.f3(*camera.get("up", (0, 1, 0))), float(camera["vfov"]))
        self.lib.orc_raygen(C.byref(cam), width, height, n, fptr(pixels), seed, job, sample,
                            *[fptr(c) for c in cols])
        out = dict(zip(["ox", "oy", "oz", "dx", "dy", "dz"], cols))
        out["tr"] = np.ones(n, np.float32)
        out["tg"] = np.ones(n, np.float32)
        out["tb"] = np.ones(n, np.float32)
        out["pixel"] = pixels.copy()
        out["depth"] = np.zeros(n, np.int32)
        out["pdf"] = np.full(n, -1.0, np.float32)
        return out

    def scatter_replay(self, oscene, material: dict, d, n, p, u, v, front, draws):
        m = oscene.data.material_struct(material)
        draws = np.ascontiguousarray(draws, np.float32)
        att, dr = np.zeros(3, np.float32), np.zeros(3, np.float32)
        pdf, used = C.c_float(), C.c_int32()
        arrs = [np.asarray(x, np.float32) for x in (d, n, p)]
        ok = self.lib.orc_scatter_replay(oscene.h, C.byref(m), *[fptr(a) for a in arrs], u, v,
                                         int(front), fptr(draws), len(draws), fptr(att), fptr(dr),
                                         C.byref(pdf), C.byref(used))
        return bool(ok), att, dr, pdf.value, used.value

    def emitted(self, data, material, ln, dfl):
        m = data.material_struct(material)
        out = np.zeros(3, np.float32)
        self.lib.orc_emitted(C.byref(m), fptr(np.asarray(ln, np.float32)),
                             fptr(np.asarray(dfl, np.float32)), fptr(out))
        return out

    def hash_grid_cell(self, x, y, z):
        return self.lib.orc_hash_grid_cell(x, y, z)


def _soa(fields, dtypes, n, src=None):
    cols = {}
    for k in fields:
        if src is not None and k in src:
            cols[k] = np.ascontiguousarray(src[k], dtypes[k]).copy()
        else:
            cols[k] = np.zeros(max(n, 1), dtypes[k])
    return cols


class OracleScene:
    def __init__(self, orc: Oracle, data):
        self.orc = orc
        self.data = data
        desc, self._keep = data.desc()
        self.h = orc.lib.orc_scene_create(C.byref(desc))

    def __del__(self):
        try:
            self.orc.lib.orc_scene_destroy(self.h)
        except Exception:
            pass

    def info(self):
        bmin, bmax = np.zeros(3, np.float32), np.zeros(3, np.float32)
        area, ntri = C.c_float(), C.c_int32()
        self.orc.lib.orc_scene_info(self.h, fptr(bmin), fptr(bmax), C.byref(area), C.byref(ntri))
        return bmin, bmax, area.value, ntri.value

    def closest_hit(self, rays, tnear=0.001, tfar=1e9, use_bvh=True):
        n = len(rays["ox"])
        cols = [np.ascontiguousarray(rays[k], np.float32) for k in ("ox", "oy", "oz", "dx", "dy", "dz")]
        t = np.zeros(n, np.float32)
        g = np.zeros(n, np.int32)
        p = np.zeros(n, np.int32)
        bu = np.zeros(n, np.float32)
        bv = np.zeros(n, np.float32)
        self.orc.lib.orc_closest_hit(self.h, n, *[fptr(c) for c in cols], tnear, tfar,
                                     1 if use_bvh else 0, fptr(t), fptr(g), fptr(p), fptr(bu),
                                     fptr(bv))
        return {"t": t, "geom": g, "prim": p, "bu": bu, "bv": bv}

    def occluded(self, o, d, maxd, use_bvh=True):
        n = len(maxd)
        cols = [np.ascontiguousarray(x, np.float32) for x in (o[:, 0], o[:, 1], o[:, 2], d[:, 0],
                                                               d[:, 1], d[:, 2], maxd)]
        out = np.zeros(n, np.uint8)
        self.orc.lib.orc_occluded(self.h, n, *[fptr(c) for c in cols], 1 if use_bvh else 0, fptr(out))
        return out.astype(bool)

    def intersect_split(self, rays, threads=8):
        n = len(rays["ox"])
        rin = _soa(abi.RAY_FIELDS, abi.RAY_DTYPES, n, rays)
        hits = _soa(abi.SHADE_FIELDS, abi.SHADE_DTYPES, n)
        miss = _soa(abi.RAY_FIELDS, abi.RAY_DTYPES, n)
        nh, nm = C.c_int32(), C.c_int32()
        r_s = abi.RaySoA(*[fptr(rin[k]) for k in abi.RAY_FIELDS])
        h_s = abi.ShadeSoA(*[fptr(hits[k]) for k in abi.SHADE_FIELDS])
        m_s = abi.RaySoA(*[fptr(miss[k]) for k in abi.RAY_FIELDS])
        self.orc.lib.orc_intersect_split(self.h, n, C.byref(r_s), C.byref(h_s), C.byref(nh),
                                         C.byref(m_s), C.byref(nm), threads)
        return ({k: v[:nh.value] for k, v in hits.items()}, nh.value,
                {k: v[:nm.value] for k, v in miss.items()}, nm.value)

    def shade(self, q, params, accum, mat_cost=None, threads=8):
        n = len(q["mat"])
        qs = _soa(abi.SHADE_FIELDS, abi.SHADE_DTYPES, n, q)
        nxt = _soa(abi.RAY_FIELDS, abi.RAY_DTYPES, n)
        q_s = abi.ShadeSoA(*[fptr(qs[k]) for k in abi.SHADE_FIELDS])
        n_s = abi.RaySoA(*[fptr(nxt[k]) for k in abi.RAY_FIELDS])
        nn = C.c_int32()
        shadows = C.c_int64()
        mc = None if mat_cost is None else np.ascontiguousarray(mat_cost, np.float64)
        self.orc.lib.orc_shade(self.h, C.byref(q_s), n, C.byref(params),
                               None if mc is None else fptr(mc), fptr(accum), None, C.byref(n_s),
                               C.byref(nn), C.byref(shadows), threads)
        return {k: v[:nn.value] for k, v in nxt.items()}, nn.value, shadows.value

    def shade_misses(self, q, physics, accum):
        n = len(q["pixel"])
        qs = _soa(abi.RAY_FIELDS, abi.RAY_DTYPES, n, q)
        q_s = abi.RaySoA(*[fptr(qs[k]) for k in abi.RAY_FIELDS])
        self.orc.lib.orc_shade_misses(self.h, C.byref(q_s), n, physics, fptr(accum), None)

    def render(self, params, mat_cost=None, tex_cost=None, fp64=False, threads=8):
        P = params.width * params.height
        img = np.zeros(P * 3, np.float32)
        cnt = np.zeros(P, np.int32)
        img64 = np.zeros(P * 3, np.float64) if fp64 else None
        st = abi.RenderStats()
        mc = None if mat_cost is None else np.ascontiguousarray(mat_cost, np.float64)
        tc = None if tex_cost is None else np.ascontiguousarray(tex_cost, np.float64)
        self.orc.lib.orc_render(self.h, C.byref(params), None if mc is None else fptr(mc),
                                None if tc is None else fptr(tc), fptr(img), fptr(cnt),
                                None if img64 is None else fptr(img64), C.byref(st), threads)
        shape = (params.height, params.width)
        return (img.reshape(*shape, 3), cnt.reshape(shape),
                None if img64 is None else img64.reshape(*shape, 3), st)


def _render_window(osc, params, window, threads=8):
    """orc_render_window: (float image, counts, FP64 image, stats) of the frame's pixels inside
    window = (x0, y0, w, h), cropped to the window."""
    x0, y0, w, h = window
    P = params.width * params.height
    img = np.zeros(P * 3, np.float32)
    cnt = np.zeros(P, np.int32)
    img64 = np.zeros(P * 3, np.float64)
    st = abi.RenderStats()
    osc.orc.lib.orc_render_window(osc.h, C.byref(params), x0, y0, w, h, fptr(img), fptr(cnt),
                                  fptr(img64), C.byref(st), threads)
    shape = (params.height, params.width)
    crop = (slice(y0, y0 + h), slice(x0, x0 + w))
    return (img.reshape(*shape, 3)[crop], cnt.reshape(shape)[crop],
            img64.reshape(*shape, 3)[crop], st)


OracleScene.render_window = lambda self, params, window, threads=8: _render_window(
    self, params, window, threads)


class Reference:
    """Entry points over the reference's own compiled sources (oracle/_ref)."""

    def __init__(self, path: Path = REF_PATH):
        self.lib = L = C.CDLL(str(path))
        L.ref_intersect_triangle.restype = C.c_int32
        L.ref_intersect_triangle.argtypes = [P] * 5 + [C.c_float, C.c_float, P, P, P]
        L.ref_interpolate_uv.argtypes = [P, P, C.c_float, C.c_float, P, P]
        L.ref_mesh_hit.restype = C.c_int32
        L.ref_mesh_hit.argtypes = [P, C.c_int32, P, C.c_int32, P, P, C.c_float, C.c_float,
                                   P, P, P, P, P, P, P]
        L.ref_hash_grid_cell.restype = C.c_uint32
        L.ref_hash_grid_cell.argtypes = [C.c_int32] * 3
        L.ref_hash_to_uv.argtypes = [C.c_uint32, P, P]
        L.ref_texture_sample.argtypes = [C.c_int32, C.c_int32, P, C.c_float, C.c_float, P]
        L.ref_texture_memory_bytes.restype = C.c_uint64
        L.ref_texture_memory_bytes.argtypes = [C.c_int32, C.c_int32]
        L.ref_camera_ray.argtypes = [P, P, P, C.c_float, C.c_int32, C.c_int32, C.c_float,
                                     C.c_float, P, P]
        L.ref_draws.argtypes = [C.c_uint32, C.c_int32, P]
        L.ref_scatter.restype = C.c_int32
        L.ref_scatter.argtypes = ([C.c_uint32, C.c_int32, P, C.c_float, C.c_float, C.c_int32,
                                   C.c_float, C.c_int32, C.c_float, C.c_int32, P, P, P, P, P, P,
                                   P, C.c_int32, C.c_float, C.c_float, P, P, P, P])
        L.ref_emitted.argtypes = [C.c_int32, P, C.c_float, C.c_float, C.c_float, P, P, P]
        L.ref_surface_color.argtypes = [C.c_int32, P, C.c_int32, C.c_int32, C.c_float,
                                        C.c_int32, P, P, P, P, C.c_float, C.c_float, P]
        L.ref_set_face_normal.argtypes = [P, P, P, P]
        L.ref_luminance.restype = C.c_float
        L.ref_luminance.argtypes = [C.c_float] * 3
        L.ref_adaptive_run.argtypes = [C.c_int32, C.c_float, C.c_int32, P, P, C.c_int32, P, P,
                                       P, P, P]
        L.ref_cost_tracker_run.argtypes = [C.c_int32, C.c_int32, P, P, P]

    def intersect_triangle(self, o, d, a, b, c, tmin=0.001, tmax=1e9):
        arrs = [np.asarray(x, np.float32) for x in (o, d, a, b, c)]
        t, u, v = C.c_float(), C.c_float(), C.c_float()
        hit = self.lib.ref_intersect_triangle(*[fptr(x) for x in arrs], tmin, tmax, C.byref(t),
                                              C.byref(u), C.byref(v))
        return bool(hit), t.value, u.value, v.value

    def draws(self, seed, k):
        out = np.zeros(k, np.float32)
        self.lib.ref_draws(seed, k, fptr(out))
        return out

    def scatter(self, seed, m: dict, o, d, point, normal, front, u, v, textures=()):
        alb = np.asarray(m.get("albedo", (1, 1, 1)), np.float32)
        ws = np.array([t[0] for t in textures] or [0], np.int32)
        hs = np.array([t[1] for t in textures] or [0], np.int32)
        rgb = (np.concatenate([t[2].reshape(-1) for t in textures]).astype(np.float32)
               if textures else np.zeros(3, np.float32))
        att, dr = np.zeros(3, np.float32), np.zeros(3, np.float32)
        pdf, used = C.c_float(), C.c_int32()
        arrs = [np.asarray(x, np.float32) for x in (o, d, point, normal)]
        ok = self.lib.ref_scatter(seed, int(m["type"]), fptr(alb), float(m.get("roughness", 0.0)),
                                  float(m.get("emission", 0.0)), int(m.get("texture_id", -1)),
                                  float(m.get("ior", 1.5)), int(m.get("use_spatial_checker", 0)),
                                  float(m.get("checker_cell_size", 1.0)), len(textures),
                                  fptr(ws), fptr(hs), fptr(rgb), *[fptr(a) for a in arrs],
                                  int(front), u, v, fptr(att), fptr(dr), C.byref(pdf),
                                  C.byref(used))
        return bool(ok), att, dr, pdf.value, used.value


def reference():
    if not REF_PATH.exists():
        return None
    return Reference()
