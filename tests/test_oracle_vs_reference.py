"""The oracle against the reference's own sources compiled by oracle/Makefile (`make ref`):
src/shading/Material.cpp, src/geometry/Mesh.cpp and the self-contained headers. Bit-exact
comparisons on random inputs; skipped where oracle/_ref was not built.

Draw order: the reference draws uniforms from its thread RNG in call order. Reseeding that RNG
(`getThreadRNG().seed`) and replaying the same uniforms into the oracle pins every
uniform->direction map. g++ evaluates the three `randomFloat()` arguments of
`Vec3(x, y, z)` in randomInUnitSphere (src/shading/Material.cpp:10-18) right to left, so each
replayed triple is reversed to (x, y, z) for the oracle, which draws x, y, z in order.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib

abi = oracle_lib.abi
f32 = np.float32


def _replay_draws(draws: np.ndarray, material_type: int, first_lobe: bool) -> np.ndarray:
    """Reference draw sequence -> oracle request order (sphere triples reversed)."""
    d = draws.copy()
    start = 1 if first_lobe else 0
    if material_type in (abi.MAT_METAL, abi.MAT_PLASTIC):
        body = d[start:start + 3 * ((len(d) - start) // 3)].reshape(-1, 3)[:, ::-1].reshape(-1)
        d = np.concatenate([d[:start], body, d[start + body.size:]])
    return d


@pytest.fixture(scope="module")
def tex_scene(scenes):
    """A scene whose only purpose is to hold two float textures for getSurfaceColor."""
    rng = np.random.default_rng(3)
    t0 = rng.random((4, 5, 3)).astype(f32)
    t1 = rng.random((8, 8, 3)).astype(f32)
    data = scenes.floor_and_light()
    data.textures = [{"width": 5, "height": 4, "format": abi.TEX_RGB_F32, "data": t0.reshape(-1)},
                     {"width": 8, "height": 8, "format": abi.TEX_RGB_F32, "data": t1.reshape(-1)}]
    return data, [(5, 4, t0), (8, 8, t1)]


def test_intersect_triangle_bit_exact(orc, ref, rng):
    for _ in range(4000):
        tri = rng.normal(size=(3, 3)).astype(f32)
        o = rng.normal(size=3).astype(f32) * f32(3)
        target = (tri.mean(0) + rng.normal(size=3).astype(f32) * f32(0.7)).astype(f32)
        d = (target - o).astype(f32)
        a = orc.intersect_triangle(o, d, *tri)
        b = ref.intersect_triangle(o, d, *tri)
        assert a[0] == b[0]
        if a[0]:
            assert (a[1], a[2], a[3]) == (b[1], b[2], b[3])


def test_mesh_bruteforce_closest_hit_matches(orc, ref, scenes, rng):
    """Mesh::hit (src/geometry/Mesh.cpp:9-44) vs the oracle's closest hit (brute force and BVH)."""
    for trial in range(6):
        nv, nt = 60, 80
        pos = rng.normal(size=(nv, 3)).astype(f32)
        idx = rng.integers(0, nv, (nt, 3)).astype(np.int32)
        idx = idx[(idx[:, 0] != idx[:, 1]) & (idx[:, 1] != idx[:, 2]) & (idx[:, 0] != idx[:, 2])]
        data = scenes.SceneData([scenes.Mesh(pos, idx, 0)], [{"type": 0, "albedo": (0.5, 0.5, 0.5)}])
        osc = orc.scene(data)
        n = 300
        o = (rng.normal(size=(n, 3)) * 4).astype(f32)
        d = (rng.normal(size=(n, 3)) * 0.3 - o * 0.25).astype(f32)
        rays = {"ox": o[:, 0], "oy": o[:, 1], "oz": o[:, 2], "dx": d[:, 0], "dy": d[:, 1], "dz": d[:, 2]}
        brute = osc.closest_hit(rays, use_bvh=False)
        bvh = osc.closest_hit(rays, use_bvh=True)
        for k in brute:
            assert np.array_equal(brute[k], bvh[k]), k
        for i in range(n):
            t, prim, u, v = C.c_float(), C.c_int32(), C.c_float(), C.c_float()
            pt, nn, front = np.zeros(3, f32), np.zeros(3, f32), C.c_int32()
            hit = ref.lib.ref_mesh_hit(pos.ctypes.data_as(C.c_void_p), nv,
                                       idx.ctypes.data_as(C.c_void_p), idx.shape[0],
                                       o[i].ctypes.data_as(C.c_void_p), d[i].ctypes.data_as(C.c_void_p),
                                       0.001, 1e9, C.byref(t), C.byref(prim), C.byref(u), C.byref(v),
                                       pt.ctypes.data_as(C.c_void_p), nn.ctypes.data_as(C.c_void_p),
                                       C.byref(front))
            assert bool(hit) == (brute["prim"][i] >= 0)
            if hit:
                assert t.value == brute["t"][i] and prim.value == brute["prim"][i]
                assert (u.value, v.value) == (brute["bu"][i], brute["bv"][i])


@pytest.mark.parametrize("mtype,lobe", [(abi.MAT_DIFFUSE, False), (abi.MAT_METAL, False),
                                        (abi.MAT_GLASS, True), (abi.MAT_PLASTIC, True),
                                        (abi.MAT_EMISSIVE, False)])
def test_scatter_maps_bit_exact(orc, ref, tex_scene, mtype, lobe, rng):
    data, textures = tex_scene
    osc = orc.scene(data)
    checked = 0
    for seed in range(1, 400):
        d = rng.normal(size=3).astype(f32)
        d /= f32(np.linalg.norm(d))
        nrm = rng.normal(size=3).astype(f32)
        nrm = (nrm / f32(np.linalg.norm(nrm))).astype(f32)
        if float(np.dot(d, nrm)) > 0:
            nrm = (-nrm).astype(f32)
        front = int(rng.integers(0, 2))
        m = {"type": mtype, "albedo": tuple(rng.uniform(0.1, 0.9, 3)),
             "roughness": float(rng.uniform(0, 1)), "ior": float(rng.uniform(1.2, 2.0)),
             "emission": 3.0, "texture_id": int(rng.integers(-1, 2)),
             "use_spatial_checker": int(rng.integers(0, 2)), "checker_cell_size": 0.37}
        p = rng.normal(size=3).astype(f32)
        u, v = float(rng.uniform(-2, 2)), float(rng.uniform(-2, 2))
        ok_r, att_r, dir_r, pdf_r, used_r = ref.scatter(seed, m, (0, 0, 0), d, p, nrm, front, u, v,
                                                        textures)
        draws = ref.draws(seed, 4096)
        ok_o, att_o, dir_o, pdf_o, used_o = orc.scatter_replay(osc, m, d, nrm, p, u, v, front, draws)
        # the plastic coin (first draw) picks the fuzz-sphere branch iff the pdf is -1
        if mtype == abi.MAT_METAL or (mtype == abi.MAT_PLASTIC and pdf_o == -1.0):
            ok_o, att_o, dir_o, pdf_o, used_o = orc.scatter_replay(
                osc, m, d, nrm, p, u, v, front, _replay_draws(draws, mtype, lobe))
        assert ok_r == ok_o
        assert used_r == used_o
        assert np.array_equal(att_r, att_o)
        assert pdf_r == pdf_o
        if ok_r:
            assert np.array_equal(dir_r, dir_o)
        checked += 1
    assert checked > 300


def test_emitted_bit_exact(orc, ref, scenes, rng):
    data = scenes.floor_and_light()
    for _ in range(500):
        mtype = int(rng.choice([abi.MAT_EMISSIVE, abi.MAT_SPOTLIGHT, abi.MAT_DIFFUSE]))
        alb = rng.uniform(0, 2, 3).astype(f32)
        outer = float(rng.uniform(10, 80))
        fall = float(rng.uniform(0, outer))
        m = {"type": mtype, "albedo": tuple(alb), "emission": float(rng.uniform(0, 20)),
             "spot_outer_deg": outer, "spot_falloff_deg": fall}
        ln = rng.normal(size=3).astype(f32)
        ln /= f32(np.linalg.norm(ln))
        dfl = rng.normal(size=3).astype(f32)
        dfl /= f32(np.linalg.norm(dfl))
        a = orc.emitted(data, m, ln, dfl)
        b = np.zeros(3, f32)
        ref.lib.ref_emitted(mtype, alb.ctypes.data_as(C.c_void_p), m["emission"], outer, fall,
                            ln.ctypes.data_as(C.c_void_p), dfl.ctypes.data_as(C.c_void_p),
                            b.ctypes.data_as(C.c_void_p))
        assert np.array_equal(a, b)


def test_camera_ray_bit_exact(orc, ref, rng):
    for _ in range(300):
        eye = rng.normal(size=3).astype(f32) * f32(3)
        tgt = rng.normal(size=3).astype(f32)
        up = np.array([0, 1, 0], f32)
        fov = float(rng.uniform(20, 90))
        w, h = int(rng.integers(2, 4000)), int(rng.integers(2, 3000))
        u, v = float(rng.uniform(0, 1)), float(rng.uniform(0, 1))
        cam = abi.Camera(abi.f3(*eye), abi.f3(*tgt), abi.f3(*up), fov)
        o1, d1, o2, d2 = (np.zeros(3, f32) for _ in range(4))
        orc.lib.orc_camera_ray(C.byref(cam), w, h, C.c_float(u), C.c_float(v),
                               o1.ctypes.data_as(C.c_void_p), d1.ctypes.data_as(C.c_void_p))
        ref.lib.ref_camera_ray(eye.ctypes.data_as(C.c_void_p), tgt.ctypes.data_as(C.c_void_p),
                               up.ctypes.data_as(C.c_void_p), fov, w, h, u, v,
                               o2.ctypes.data_as(C.c_void_p), d2.ctypes.data_as(C.c_void_p))
        assert np.array_equal(o1, o2) and np.array_equal(d1, d2)


def test_hash_and_texture_bit_exact(orc, ref, rng):
    for _ in range(1000):
        x, y, z = (int(v) for v in rng.integers(-2**20, 2**20, 3))
        h = orc.hash_grid_cell(x, y, z)
        assert h == ref.lib.ref_hash_grid_cell(x, y, z)


def test_adaptive_sampler_matches_reference(orc, ref, rng):
    """AdaptiveSampler (include/render/AdaptiveSampler.h:40-80) on random sample streams."""
    P, S = 16, 2000
    pix = rng.integers(0, P, S).astype(np.int32)
    vals = np.where(rng.random(S) < 0.5, rng.uniform(0, 0.02, S),
                    rng.uniform(0, 10, S)).astype(f32)
    vals[pix < 6] = f32(0.5)
    checks = np.array([255, 999, 1999], np.int32)
    rm, rm2 = np.zeros(P, f32), np.zeros(P, f32)
    rc, ra = np.zeros(P, np.int32), np.zeros(P, np.uint8)
    ref.lib.ref_adaptive_run(P, C.c_float(0.05), S, pix.ctypes.data_as(C.c_void_p),
                             vals.ctypes.data_as(C.c_void_p), len(checks),
                             checks.ctypes.data_as(C.c_void_p), rm.ctypes.data_as(C.c_void_p),
                             rm2.ctypes.data_as(C.c_void_p), rc.ctypes.data_as(C.c_void_p),
                             ra.ctypes.data_as(C.c_void_p))
    om, om2 = np.zeros(P, f32), np.zeros(P, f32)
    oc, oa = np.zeros(P, np.int32), np.ones(P, np.uint8)
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    c = 0
    for i in range(S):
        orc.lib.orc_adaptive_add(ptr(om), ptr(om2), ptr(oc), int(pix[i]), C.c_float(vals[i]))
        while c < len(checks) and checks[c] == i:
            orc.lib.orc_adaptive_update(P, ptr(om), ptr(om2), ptr(oc), ptr(oa), C.c_float(0.05), 32)
            c += 1
    assert np.array_equal(rm, om) and np.array_equal(rm2, om2)
    assert np.array_equal(rc, oc) and np.array_equal(ra, oa)
    assert ra[:6].sum() < 6   # constant pixels with >= 32 samples converge


def test_reference_texture_memory_bytes(ref):
    # Texture::memoryBytes = pixels * sizeof(Color) (include/shading/Texture.h:51-54)
    for w, h in [(1, 1), (2, 2), (4, 4), (1024, 1024), (17, 3)]:
        assert ref.lib.ref_texture_memory_bytes(w, h) == w * h * 12
