"""Seeded synthetic scenes for the benchmark workloads and the parity tests.

Every generator is deterministic in its seed (default 0x5EED0000 + workload index) and records
every parameter the workload description leaves open in `SceneData.params`.

  cornell_box()        256x256 Cornell box: 5 wall quads + 0.25x0.25 ceiling light + icosphere(4)
  sorting_stress()     1.3M-triangle noise-displaced icosphere(8), 64 materials, 32 checker textures
  sphere_cloud()       4096 displaced UV spheres (10.5M triangles), 256 materials, 1024 emitters, HDRI
  floor_and_light()    the floor + light-quad scene of the reference's renderer tests
  material_variants()  the reference's material set (fuzz Metal, Schlick Glass, Plastic, SpotLight,
                       spatial checker, textures) with a directional light and an environment map
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import _abi

SEED_BASE = 0x5EED0000


@dataclass
class Mesh:
    positions: np.ndarray            # (V,3) float32
    indices: np.ndarray              # (T,3) int32
    material_id: int
    uv_u: Optional[np.ndarray] = None  # (T,3) float32 face-varying
    uv_v: Optional[np.ndarray] = None
    name: str = ""


@dataclass
class SceneData:
    meshes: List[Mesh]
    materials: List[dict]
    textures: List[dict] = field(default_factory=list)   # {width,height,format,data}
    directional: Optional[dict] = None
    environment: Optional[dict] = None                 # {width,height,rgb,intensity,tint}
    camera: Optional[dict] = None
    params: Dict = field(default_factory=dict)
    name: str = ""

    @property
    def triangle_count(self) -> int:
        return int(sum(m.indices.shape[0] for m in self.meshes))

    def material_struct(self, m: dict) -> _abi.Material:
        return _abi.Material(
            int(m["type"]), _abi.f3(*m.get("albedo", (1.0, 1.0, 1.0))),
            float(m.get("roughness", 0.0)), float(m.get("emission", 0.0)),
            int(m.get("texture_id", -1)), float(m.get("spot_outer_deg", 0.0)),
            float(m.get("spot_falloff_deg", 0.0)), float(m.get("ior", 1.5)),
            int(m.get("use_spatial_checker", 0)), float(m.get("checker_cell_size", 1.0)),
            int(m.get("model", _abi.MODEL_REFERENCE)))

    def desc(self):
        """(wf_scene_desc, keepalive) -- host pointers stay valid while keepalive lives."""
        keep = []
        mesh_arr = (_abi.MeshDesc * max(len(self.meshes), 1))()
        for i, m in enumerate(self.meshes):
            pos = np.ascontiguousarray(m.positions, np.float32)
            idx = np.ascontiguousarray(m.indices, np.int32)
            uu = None if m.uv_u is None else np.ascontiguousarray(m.uv_u, np.float32)
            vv = None if m.uv_v is None else np.ascontiguousarray(m.uv_v, np.float32)
            keep += [pos, idx, uu, vv]
            mesh_arr[i] = _abi.MeshDesc(
                pos.ctypes.data_as(C.c_void_p), pos.shape[0], idx.ctypes.data_as(C.c_void_p),
                idx.shape[0], None if uu is None else uu.ctypes.data_as(C.c_void_p),
                None if vv is None else vv.ctypes.data_as(C.c_void_p), int(m.material_id))
        mat_arr = (_abi.Material * max(len(self.materials), 1))()
        for i, m in enumerate(self.materials):
            mat_arr[i] = self.material_struct(m)
        tex_arr = (_abi.TextureDesc * max(len(self.textures), 1))()
        for i, t in enumerate(self.textures):
            data = np.ascontiguousarray(t["data"])
            keep.append(data)
            tex_arr[i] = _abi.TextureDesc(int(t["width"]), int(t["height"]), int(t["format"]),
                                          data.ctypes.data_as(C.c_void_p))
        dl = _abi.DirectionalLight()
        if self.directional:
            dl = _abi.DirectionalLight(_abi.f3(*self.directional["direction"]),
                                       _abi.f3(*self.directional.get("color", (1, 1, 1))),
                                       float(self.directional["intensity"]))
        env = _abi.EnvironmentDesc()
        if self.environment:
            rgb = np.ascontiguousarray(self.environment["rgb"], np.float32)
            keep.append(rgb)
            env = _abi.EnvironmentDesc(int(self.environment["width"]),
                                       int(self.environment["height"]),
                                       rgb.ctypes.data_as(C.c_void_p),
                                       float(self.environment.get("intensity", 1.0)),
                                       _abi.f3(*self.environment.get("tint", (1, 1, 1))))
        keep += [mesh_arr, mat_arr, tex_arr]
        d = _abi.SceneDesc(mesh_arr, len(self.meshes), mat_arr, len(self.materials), tex_arr,
                           len(self.textures), dl, env)
        return d, keep


# ---- geometry helpers ------------------------------------------------------------------------

def quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3); normal = (p1-p0) x (p2-p0)."""
    pos = np.array([p0, p1, p2, p3], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    return pos, idx


def icosphere(subdiv: int):
    """Unit icosphere: 20 * 4^subdiv faces (midpoints shared through an edge table)."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]],
                 np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9],
                  [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2],
                  [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
                  [8, 6, 7], [9, 8, 1]], np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(subdiv):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid])
        nf = f.shape[0]
        a = base + inv[:nf]
        b = base + inv[nf:2 * nf]
        c = base + inv[2 * nf:]
        f = np.concatenate([
            np.stack([f[:, 0], a, c], 1), np.stack([f[:, 1], b, a], 1),
            np.stack([f[:, 2], c, b], 1), np.stack([a, b, c], 1)])
    return v, f.astype(np.int32)


def uv_sphere(segments: int, rings: int):
    """UV sphere with `segments` longitudes and `rings` latitude bands (2 poles):
    segments*(rings-2)*2 + 2*segments triangles."""
    verts = [[0.0, 1.0, 0.0]]
    for r in range(1, rings):
        th = math.pi * r / rings
        for s in range(segments):
            ph = 2.0 * math.pi * s / segments
            verts.append([math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)])
    verts.append([0.0, -1.0, 0.0])
    v = np.array(verts, np.float64)
    faces = []
    for s in range(segments):
        faces.append([0, 1 + (s + 1) % segments, 1 + s])
    for r in range(rings - 2):
        b0 = 1 + r * segments
        b1 = 1 + (r + 1) * segments
        for s in range(segments):
            s1 = (s + 1) % segments
            faces.append([b0 + s, b0 + s1, b1 + s])
            faces.append([b0 + s1, b1 + s1, b1 + s])
    last = v.shape[0] - 1
    b = 1 + (rings - 2) * segments
    for s in range(segments):
        faces.append([last, b + s, b + (s + 1) % segments])
    return v, np.array(faces, np.int32)


def _lattice(seed: int, ix, iy, iz):
    h = (ix.astype(np.int64) * 73856093) ^ (iy.astype(np.int64) * 19349663) ^ \
        (iz.astype(np.int64) * 83492791) ^ (seed * 2654435761)
    h = (h ^ (h >> 13)) * 1274126177
    h = h ^ (h >> 16)
    return (h & 0xFFFFFF).astype(np.float64) / float(0xFFFFFF) * 2.0 - 1.0


def value_noise(p: np.ndarray, seed: int) -> np.ndarray:
    """Trilinear value noise with smoothstep weights, values in [-1, 1]."""
    i = np.floor(p).astype(np.int64)
    f = p - i
    w = f * f * (3.0 - 2.0 * f)
    out = np.zeros(p.shape[0])
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                val = _lattice(seed, i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz)
                wx = w[:, 0] if dx else 1.0 - w[:, 0]
                wy = w[:, 1] if dy else 1.0 - w[:, 1]
                wz = w[:, 2] if dz else 1.0 - w[:, 2]
                out += val * wx * wy * wz
    return out


def fbm(p: np.ndarray, seed: int, octaves: int = 4, base_freq: float = 2.0) -> np.ndarray:
    total = np.zeros(p.shape[0])
    amp, freq, norm = 1.0, base_freq, 0.0
    for o in range(octaves):
        total += amp * value_noise(p * freq, seed + o)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def spherical_uv(p: np.ndarray):
    """Per-vertex spherical mapping u = 0.5 + atan2(z,x)/2pi, v = 0.5 - asin(y)/pi."""
    n = p / np.linalg.norm(p, axis=1, keepdims=True)
    u = 0.5 + np.arctan2(n[:, 2], n[:, 0]) / (2 * math.pi)
    v = 0.5 - np.arcsin(np.clip(n[:, 1], -1, 1)) / math.pi
    return u.astype(np.float32), v.astype(np.float32)


def checker_texture(rng: np.random.Generator, size: int) -> dict:
    cell = int(rng.integers(8, 129))
    c0 = rng.integers(20, 236, size=3)
    c1 = rng.integers(20, 236, size=3)
    y, x = np.mgrid[0:size, 0:size]
    sel = ((x // cell) + (y // cell)) % 2
    img = np.zeros((size, size, 4), np.uint8)
    img[..., :3] = np.where(sel[..., None] == 0, c0, c1)
    img[..., 3] = 255
    return {"width": size, "height": size, "format": _abi.TEX_RGBA8_SRGB,
            "data": img.reshape(-1).view(np.uint32).copy(), "cell": cell,
            "colors": (c0.tolist(), c1.tolist())}


def _face_mix_materials(rng: np.random.Generator, count: int, n_tex: int):
    """Material mix of the stress workloads: 40% Lambertian, 40% GGX conductor with roughness
    U[0.05,0.8], 20% dielectric with IOR U[1.3,1.7]."""
    mats = []
    for m in range(count):
        u = rng.random()
        if u < 0.4:
            mats.append({"type": _abi.MAT_DIFFUSE, "albedo": tuple(rng.uniform(0.2, 0.85, 3)),
                         "texture_id": int(rng.integers(0, n_tex)) if n_tex else -1,
                         "model": _abi.MODEL_REFERENCE})
        elif u < 0.8:
            tex = int(rng.integers(0, n_tex)) if (n_tex and rng.random() < 0.5) else -1
            mats.append({"type": _abi.MAT_METAL, "albedo": tuple(rng.uniform(0.5, 0.98, 3)),
                         "roughness": float(rng.uniform(0.05, 0.8)), "texture_id": tex,
                         "model": _abi.MODEL_MICROFACET})
        else:
            mats.append({"type": _abi.MAT_GLASS, "albedo": (1.0, 1.0, 1.0),
                         "ior": float(rng.uniform(1.3, 1.7)), "texture_id": -1,
                         "model": _abi.MODEL_MICROFACET})
    return mats


def _split_by_material(pos: np.ndarray, faces: np.ndarray, face_mat: np.ndarray, n_mat: int,
                       uv=None):
    """One mesh per material id (geomID = material id); vertices re-indexed per mesh."""
    meshes = []
    order = np.argsort(face_mat, kind="stable")
    bounds = np.searchsorted(face_mat[order], np.arange(n_mat + 1))
    for m in range(n_mat):
        sel = order[bounds[m]:bounds[m + 1]]
        if sel.size == 0:
            continue
        f = faces[sel]
        used, inv = np.unique(f.reshape(-1), return_inverse=True)
        mesh = Mesh(pos[used].astype(np.float32), inv.reshape(-1, 3).astype(np.int32), m,
                    name=f"/mat{m}")
        if uv is not None:
            mesh.uv_u = uv[0][f].astype(np.float32)
            mesh.uv_v = uv[1][f].astype(np.float32)
        meshes.append(mesh)
    return meshes


# ---- workloads ----------------------------------------------------------------------------

def cornell_box(seed: int = SEED_BASE + 0, subdiv: int = 4) -> SceneData:
    """Cornell box: floor/ceiling/back walls white 0.73, left red (0.65,0.05,0.05), right green
    (0.12,0.45,0.15), a 0.25x0.25 ceiling light of radiance (17,12,4) facing down, and an
    icosphere (Lambertian 0.8) of radius 0.2 at the box centre. Box [-0.5,0.5]x[0,1]x[-0.5,0.5]."""
    mats = [
        {"type": _abi.MAT_DIFFUSE, "albedo": (0.73, 0.73, 0.73)},
        {"type": _abi.MAT_DIFFUSE, "albedo": (0.65, 0.05, 0.05)},
        {"type": _abi.MAT_DIFFUSE, "albedo": (0.12, 0.45, 0.15)},
        {"type": _abi.MAT_DIFFUSE, "albedo": (0.8, 0.8, 0.8)},
        {"type": _abi.MAT_EMISSIVE, "albedo": (17.0, 12.0, 4.0), "emission": 1.0},
    ]
    a, b = -0.5, 0.5
    meshes = []
    for name, corners, mid in [
        ("floor", [(a, 0, a), (a, 0, b), (b, 0, b), (b, 0, a)], 0),
        ("ceiling", [(a, 1, a), (b, 1, a), (b, 1, b), (a, 1, b)], 0),
        ("back", [(a, 0, a), (b, 0, a), (b, 1, a), (a, 1, a)], 0),
        ("left", [(a, 0, a), (a, 1, a), (a, 1, b), (a, 0, b)], 1),
        ("right", [(b, 0, a), (b, 0, b), (b, 1, b), (b, 1, a)], 2),
    ]:
        p, i = quad(*corners)
        meshes.append(Mesh(p, i, mid, name=f"/CornellBox/{name}"))
    h = 0.125
    y = 0.999
    p, i = quad((-h, y, -h), (h, y, -h), (h, y, h), (-h, y, h))
    meshes.append(Mesh(p, i, 4, name="/CornellBox/light"))
    v, f = icosphere(subdiv)
    meshes.append(Mesh((v * 0.2 + np.array([0.0, 0.5, 0.0])).astype(np.float32), f, 3,
                       name="/CornellBox/sphere"))
    cam = {"eye": (0.0, 0.5, 1.9), "target": (0.0, 0.5, 0.0), "up": (0.0, 1.0, 0.0), "vfov": 40.0}
    return SceneData(meshes, mats, camera=cam, name="cornell_box",
                     params={"seed": seed, "box": [[a, 0, a], [b, 1, b]], "light_y": y,
                             "sphere_radius": 0.2, "sphere_center": [0, 0.5, 0],
                             "icosphere_subdiv": subdiv, "width": 256, "height": 256, "spp": 16,
                             "max_depth": 8, "rr_min_depth": 3, "policy": "none"})


def sorting_stress(seed: int = SEED_BASE + 1, subdiv: int = 8, n_materials: int = 64,
                   n_textures: int = 32, tex_size: int = 1024) -> SceneData:
    """Noise-displaced icosphere (4-octave value noise, amplitude 0.1), faces assigned uniformly
    to n_materials ids (40% Lambertian / 40% GGX / 20% dielectric), n_textures checker textures
    tex_size^2 RGBA8 with random cell size and colours, spherical UVs, two rectangular area
    lights facing the sphere."""
    rng = np.random.default_rng(seed)
    v, f = icosphere(subdiv)
    disp = 1.0 + 0.1 * fbm(v, seed, octaves=4, base_freq=2.0)
    pos = (v * disp[:, None]).astype(np.float32)
    uv = spherical_uv(v)
    textures = [checker_texture(rng, tex_size) for _ in range(n_textures)]
    mats = _face_mix_materials(rng, n_materials, n_textures)
    face_mat = rng.integers(0, n_materials, size=f.shape[0])
    meshes = _split_by_material(pos, f, face_mat, n_materials, uv)
    # two rectangular area lights, normals toward the sphere centre
    lights = [((-3.0, 3.0, 3.0), (1.5, 1.0), (14.0, 13.0, 12.0)),
              ((3.2, 2.0, 1.5), (1.2, 1.2), (10.0, 11.0, 14.0))]
    for k, (c, (w, h), rad) in enumerate(lights):
        c = np.array(c)
        n = -c / np.linalg.norm(c)
        t = np.cross(n, [0.0, 1.0, 0.0])
        t /= np.linalg.norm(t)
        b = np.cross(t, n)
        p0, p1, p2, p3 = (c - t * w / 2 - b * h / 2, c + t * w / 2 - b * h / 2,
                          c + t * w / 2 + b * h / 2, c - t * w / 2 + b * h / 2)
        # winding so that (p1-p0)x(p2-p0) points along n
        if np.dot(np.cross(p1 - p0, p2 - p0), n) < 0:
            p1, p3 = p3, p1
        p, i = quad(p0, p1, p2, p3)
        mats.append({"type": _abi.MAT_EMISSIVE, "albedo": rad, "emission": 1.0})
        meshes.append(Mesh(p, i, len(mats) - 1, name=f"/light{k}"))
    cam = {"eye": (0.0, 0.35, 3.3), "target": (0.0, 0.0, 0.0), "up": (0.0, 1.0, 0.0), "vfov": 42.0}
    return SceneData(meshes, mats, textures, camera=cam, name="sorting_stress",
                     params={"seed": seed, "subdiv": subdiv, "noise_octaves": 4,
                             "noise_amplitude": 0.1, "noise_base_freq": 2.0,
                             "materials": n_materials, "textures": n_textures,
                             "texture_size": tex_size, "lights": [l[0] for l in lights],
                             "width": 1920, "height": 1080, "spp": 64, "max_depth": 12})


def hdri_sky(width: int = 2048, height: int = 1024, peak_over_mean: float = 1e4,
             sun_dir=(0.45, 0.6, -0.66), sun_sigma: float = 0.02) -> dict:
    """Sky gradient + Gaussian sun lobe, the sun scaled so that the brightest texel is exactly
    peak_over_mean times the mean texel luminance. With sky s and lobe l (luminances per texel)
    and the lobe's brightest texel j, (s_j + k l_j) / mean(s + k l) = R gives
    k = (R mean(s) - s_j) / (l_j - R mean(l)); a lobe too wide for the ratio (l_j <= R mean(l):
    its texels carry more than 1/R of the peak on average) is narrowed by halves until it fits."""
    y, x = np.mgrid[0:height, 0:width]
    u = (x + 0.5) / width
    v = (y + 0.5) / height
    phi = (u - 0.5) * 2 * math.pi
    lat = (0.5 - v) * math.pi
    d = np.stack([np.cos(lat) * np.cos(phi), np.sin(lat), np.cos(lat) * np.sin(phi)], -1)
    sky = np.where(d[..., 1:2] > 0,
                   np.array([0.25, 0.45, 0.9]) * (0.4 + 0.6 * d[..., 1:2]),
                   np.array([0.12, 0.1, 0.08]) * (1 + d[..., 1:2]))
    s = np.array(sun_dir, np.float64)
    s /= np.linalg.norm(s)
    ang = np.arccos(np.clip((d * s).sum(-1), -1, 1))
    lum = lambda c: 0.2126 * c[..., 0] + 0.7152 * c[..., 1] + 0.0722 * c[..., 2]
    R = float(peak_over_mean)
    ls = lum(sky)
    sigma = float(sun_sigma)
    for _ in range(40):
        lobe = np.exp(-0.5 * (ang / sigma) ** 2)[..., None] * np.array([1.0, 0.95, 0.85])
        ll = lum(lobe)
        j = np.unravel_index(np.argmax(ll), ll.shape)
        den = ll[j] - R * ll.mean()
        if den > 0:
            k = (R * ls.mean() - ls[j]) / den
            tot = ls + k * ll
            if k > 0 and abs(tot.max() / tot.mean() - R) <= 1e-6 * R:
                break
        sigma *= 0.5
    else:
        raise ValueError(f"no sun lobe reaches peak/mean {R} on a {width}x{height} map")
    rgb = (sky + k * lobe).astype(np.float32)
    return {"width": width, "height": height, "rgb": rgb.reshape(-1), "intensity": 1.0,
            "tint": (1.0, 1.0, 1.0), "sun_scale": float(k), "sun_sigma": sigma,
            "sun_dir": tuple(s)}


def sphere_cloud(seed: int = SEED_BASE + 2, copies: int = 4096, n_materials: int = 256,
                 n_emitters: int = 1024, box: float = 100.0, env_size=(2048, 1024)) -> SceneData:
    """copies of a 2560-triangle noise-displaced UV sphere (64 segments x 21 rings) with uniform
    scale U[0.2,1], uniform random rotation and position in a box^3 cube; n_materials ids with the
    stress mix; n_emitters emissive triangles; an HDRI sky importance-sampled by alias table."""
    rng = np.random.default_rng(seed)
    v, f = uv_sphere(64, 21)
    disp = 1.0 + 0.1 * fbm(v, seed, octaves=4, base_freq=2.0)
    base = v * disp[:, None]
    n_tex = 32
    textures = [checker_texture(rng, 256) for _ in range(n_tex)]
    mats = _face_mix_materials(rng, n_materials, n_tex)
    q = rng.normal(size=(copies, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], 1)
    scale = rng.uniform(0.2, 1.0, copies)
    centre = rng.uniform(-box / 2, box / 2, (copies, 3))
    copy_mat = rng.integers(0, n_materials, copies)
    meshes = []
    order = np.argsort(copy_mat, kind="stable")
    for m in range(n_materials):
        sel = order[copy_mat[order] == m]
        if sel.size == 0:
            continue
        P = np.einsum("cij,vj->cvi", R[sel], base) * scale[sel, None, None] + centre[sel, None, :]
        nv = base.shape[0]
        F = (f[None, :, :] + (np.arange(sel.size) * nv)[:, None, None]).reshape(-1, 3)
        meshes.append(Mesh(P.reshape(-1, 3).astype(np.float32), F.astype(np.int32), m,
                           name=f"/cloud/mat{m}"))
    # emitters: small triangles spread through the upper half of the box
    c = rng.uniform([-box / 2, 0, -box / 2], [box / 2, box / 2, box / 2], (n_emitters, 3))
    e1 = rng.normal(size=(n_emitters, 3))
    e2 = rng.normal(size=(n_emitters, 3))
    size = rng.uniform(0.3, 1.2, n_emitters)[:, None]
    tri = np.stack([c, c + e1 * size, c + e2 * size], 1).astype(np.float32)
    mats.append({"type": _abi.MAT_EMISSIVE, "albedo": (8.0, 7.0, 6.0), "emission": 1.0})
    meshes.append(Mesh(tri.reshape(-1, 3), np.arange(n_emitters * 3, dtype=np.int32).reshape(-1, 3),
                       len(mats) - 1, name="/cloud/emitters"))
    env = hdri_sky(*env_size)
    cam = {"eye": (0.0, 10.0, box * 0.75), "target": (0.0, 0.0, 0.0), "up": (0.0, 1.0, 0.0),
           "vfov": 50.0}
    return SceneData(meshes, mats, textures, environment=env, camera=cam, name="sphere_cloud",
                     params={"seed": seed, "copies": copies, "uv_sphere": [64, 21],
                             "materials": n_materials, "emitters": n_emitters, "box": box,
                             "env": list(env_size), "sun_scale": env["sun_scale"],
                             "sun_sigma": env["sun_sigma"],
                             "width": 1920, "height": 1080, "spp_min": 16, "spp_max": 1024,
                             "adaptive_threshold": 0.01, "max_depth": 12,
                             "policy": "costBenefit"})


def floor_and_light() -> SceneData:
    """buildTestScene of the reference's renderer tests (tests/render/WavefrontRendererTest.cpp:9-35)."""
    mats = [{"type": _abi.MAT_DIFFUSE, "albedo": (0.8, 0.8, 0.8)},
            {"type": _abi.MAT_EMISSIVE, "albedo": (1.0, 1.0, 1.0), "emission": 10.0}]
    p0, i0 = quad((-2.0, -1.0, -3.0), (2.0, -1.0, -3.0), (2.0, -1.0, 1.0), (-2.0, -1.0, 1.0))
    p1, i1 = quad((-0.5, 2.0, -2.0), (0.5, 2.0, -2.0), (0.5, 2.0, -1.0), (-0.5, 2.0, -1.0))
    cam = {"eye": (0.0, 0.5, 2.0), "target": (0.0, 0.0, -1.0), "up": (0.0, 1.0, 0.0), "vfov": 45.0}
    return SceneData([Mesh(p0, i0, 0, name="floor"), Mesh(p1, i1, 1, name="light")], mats,
                     camera=cam, name="floor_and_light")


def material_variants(seed: int = SEED_BASE + 5, subdiv: int = 2, env_size=(32, 16),
                      directional: bool = True, environment: bool = True) -> SceneData:
    """Every material of the reference's own material set on the reference model, in one small
    scene, after scenes/materialVariants.usda and cornellBoxDragon.usda: a box [-1,1]x[0,2]x[-1,1]
    open at the camera side (z = +1) with a UV-textured RGBA8 checker floor, a fuzz Metal left
    wall (roughness 0.3), a Plastic back wall (roughness 0.15), a spatial-checker Diffuse right
    wall (cell 0.3 on a 64^2 palette), a Schlick Glass icosphere (IOR 1.5; refracted paths hit
    its faces from inside), a mirror Metal icosphere (roughness 0), a Plastic icosphere with no
    fuzz and a textured Metal panel; a SpotLight quad on the ceiling (outer 60 deg, falloff 30
    deg), a small Emissive quad, a directional light and an environment map seen through the
    open side."""
    from .usd import spatial_palette

    rng = np.random.default_rng(seed)
    floor_tex = checker_texture(rng, 64)
    palette = spatial_palette(64, 0.4)
    textures = [floor_tex, palette]
    R = _abi.MODEL_REFERENCE
    mats = [
        {"type": _abi.MAT_DIFFUSE, "albedo": (0.8, 0.8, 0.8), "texture_id": 0, "model": R},
        {"type": _abi.MAT_METAL, "albedo": (0.9, 0.75, 0.5), "roughness": 0.3, "model": R},
        {"type": _abi.MAT_PLASTIC, "albedo": (0.85, 0.08, 0.08), "roughness": 0.15, "model": R},
        {"type": _abi.MAT_DIFFUSE, "albedo": (0.5, 0.7, 0.85), "texture_id": 1, "model": R,
         "use_spatial_checker": 1, "checker_cell_size": 0.3},
        {"type": _abi.MAT_DIFFUSE, "albedo": (0.73, 0.73, 0.73), "model": R},
        {"type": _abi.MAT_GLASS, "albedo": (1.0, 1.0, 1.0), "ior": 1.5, "model": R},
        {"type": _abi.MAT_METAL, "albedo": (0.95, 0.95, 0.95), "roughness": 0.0, "model": R},
        {"type": _abi.MAT_PLASTIC, "albedo": (0.12, 0.45, 0.15), "roughness": 0.0, "model": R},
        {"type": _abi.MAT_METAL, "albedo": (0.8, 0.8, 0.8), "roughness": 0.05, "texture_id": 0,
         "model": R},
        {"type": _abi.MAT_SPOTLIGHT, "albedo": (30.0, 28.0, 24.0), "emission": 1.0,
         "spot_outer_deg": 60.0, "spot_falloff_deg": 30.0, "model": R},
        {"type": _abi.MAT_EMISSIVE, "albedo": (6.0, 5.5, 7.5), "emission": 1.0, "model": R},
    ]
    a, b, top = -1.0, 1.0, 2.0
    meshes = []
    # floor with face-varying UVs: two checker periods across the floor
    p, i = quad((a, 0, a), (a, 0, b), (b, 0, b), (b, 0, a))
    uvc = np.array([[0.0, 0.0], [0.0, 2.0], [2.0, 2.0], [2.0, 0.0]], np.float32)
    fl = Mesh(p, i, 0, name="/Variants/Floor")
    fl.uv_u = uvc[i, 0].astype(np.float32)
    fl.uv_v = uvc[i, 1].astype(np.float32)
    meshes.append(fl)
    for name, corners, mid in [
        ("Ceiling", [(a, top, a), (b, top, a), (b, top, b), (a, top, b)], 4),
        ("BackWall", [(a, 0, a), (b, 0, a), (b, top, a), (a, top, a)], 2),
        ("LeftWall", [(a, 0, a), (a, top, a), (a, top, b), (a, 0, b)], 1),
        ("RightWall", [(b, 0, a), (b, 0, b), (b, top, b), (b, top, a)], 3),
    ]:
        p, i = quad(*corners)
        meshes.append(Mesh(p, i, mid, name=f"/Variants/{name}"))
    v, f = icosphere(subdiv)
    for name, centre, radius, mid in [("GlassBall", (-0.45, 0.4, 0.1), 0.4, 5),
                                      ("MirrorBall", (0.5, 0.35, -0.35), 0.35, 6),
                                      ("PlasticBall", (0.35, 0.25, 0.45), 0.25, 7)]:
        meshes.append(Mesh((v * radius + np.array(centre)).astype(np.float32), f, mid,
                           name=f"/Variants/{name}"))
    p, i = quad((-0.2, 0.05, -0.95), (0.3, 0.05, -0.95), (0.3, 0.9, -0.95), (-0.2, 0.9, -0.95))
    panel = Mesh(p, i, 8, name="/Variants/TexturedMetalPanel")
    panel.uv_u = uvc[i, 0].astype(np.float32)
    panel.uv_v = uvc[i, 1].astype(np.float32)
    meshes.append(panel)
    # SpotLight facing down (normal (p1-p0)x(p2-p0) = -y), Emissive accent quad on the back wall
    s, y = 0.3, top - 0.001
    p, i = quad((-s, y, -s), (-s, y, s), (s, y, s), (s, y, -s))
    meshes.append(Mesh(p, i, 9, name="/Variants/Spot"))
    p, i = quad((-0.9, 1.2, -0.999), (-0.6, 1.2, -0.999), (-0.6, 1.5, -0.999),
                (-0.9, 1.5, -0.999))
    meshes.append(Mesh(p, i, 10, name="/Variants/Accent"))
    dl = None
    if directional:
        d = np.array([-0.3, -1.0, -0.45])
        dl = {"direction": tuple((d / np.linalg.norm(d)).tolist()), "color": (1.0, 0.95, 0.85),
              "intensity": 1.5}
    env = hdri_sky(*env_size, peak_over_mean=20.0) if environment else None
    cam = {"eye": (0.0, 1.0, 3.4), "target": (0.0, 0.8, 0.0), "up": (0.0, 1.0, 0.0), "vfov": 45.0}
    return SceneData(meshes, mats, textures, directional=dl, environment=env, camera=cam,
                     name="material_variants",
                     params={"seed": seed, "subdiv": subdiv, "env": list(env_size),
                             "box": [[a, 0, a], [b, top, b]], "open_side": "+z"})
