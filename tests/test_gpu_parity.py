"""HIP path vs the CPU oracle, stage by stage, through the C-ABI (libwavefront_hip.so).

Integer/index results (sort keys, permutations, geomID/primID, hit/miss split order, queue
contents) are bit-exact. Float results are bit-exact too: the device and the oracle run the same
IEEE operations with contraction off, and the transcendentals of the samplers (sin/cos, asin,
acos, atan2, x^5) are the shared single-precision implementations of wf_fmath.h, so a path's
directions, lobe choices and weights are the same bits on both sides. Whole images differ from
the oracle's only where the regenerating schedule sums a pixel's sample lanes in another order
(float reassociation, bounded relative to the pixel value)."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib

pytestmark = pytest.mark.gpu
abi = oracle_lib.abi
f32 = np.float32


def random_rays(rng, n, box=1.0, center=(0.0, 0.5, 0.0)):
    o = (rng.uniform(-box, box, (n, 3)) + np.array(center)).astype(f32)
    d = rng.normal(size=(n, 3)).astype(f32)
    d /= np.linalg.norm(d, axis=1, keepdims=True).astype(f32)
    return {"ox": o[:, 0], "oy": o[:, 1], "oz": o[:, 2], "dx": d[:, 0], "dy": d[:, 1],
            "dz": d[:, 2], "tr": rng.uniform(0.1, 1, n).astype(f32),
            "tg": rng.uniform(0.1, 1, n).astype(f32), "tb": rng.uniform(0.1, 1, n).astype(f32),
            "pixel": rng.permutation(n).astype(np.int32),
            "depth": rng.integers(0, 4, n).astype(np.int32),
            "pdf": rng.uniform(-1, 2, n).astype(f32)}


def upload_rays(wf, ctx, rays):
    n = len(rays["ox"])
    q = wf.RayQueue(ctx, n)
    q.upload(rays)
    return q


@pytest.fixture(scope="module")
def cornell(wf, ctx, scenes, orc):
    data = scenes.cornell_box()
    return data, wf.Scene(ctx, data), orc.scene(data)


def camera_rays(orc, data, w, h, sample=0, seed=7):
    return orc.raygen(data.camera, w, h, np.arange(w * h), seed, 0, sample)


@pytest.fixture(params=[8, 6, 11], ids=lambda b: f"radix{b}")
def radix_bits(request, ctx):
    """Runs a test with each digit width of the radix sort passes."""
    saved = ctx.tuning()
    ctx.set_tuning(radix_bits=request.param)
    yield request.param
    ctx.set_tuning(**saved)


# ---- RayQueue::schedule / materialize ---------------------------------------------------------
def test_ray_schedule_bit_exact(wf, ctx, orc, rng, radix_bits):
    for n in [1, 7, 2048, 2049, 100_003]:
        rays = random_rays(rng, n, box=12.0, center=(0, 0, 0))
        rays["ox"] = np.round(rays["ox"], 0).astype(f32)          # ties in the key
        q = upload_rays(wf, ctx, rays)
        idx, keys = wf.ray_schedule(ctx, q, n, (-10, -10, -10), (10, 10, 10))
        assert np.array_equal(keys, orc.ray_keys(rays, (-10, -10, -10), (10, 10, 10)))
        assert np.array_equal(idx, orc.ray_schedule(rays, (-10, -10, -10), (10, 10, 10)))
        out = wf.materialize_rays(ctx, q, idx, n).download(n)
        for k in abi.RAY_FIELDS:
            assert np.array_equal(out[k], rays[k][idx]), k


def test_ray_schedule_reference_kats(wf, ctx):
    # RayQueueTest.cpp:156-171 and 132-154 through the device path
    rays = {"ox": np.array([5, -5, 0], f32), "oy": np.zeros(3, f32), "oz": np.zeros(3, f32),
            "dx": np.zeros(3, f32), "dy": np.zeros(3, f32), "dz": np.ones(3, f32)}
    q = upload_rays(wf, ctx, rays)
    idx, _ = wf.ray_schedule(ctx, q, 3, (-10, -10, -10), (10, 10, 10))
    assert np.array([100, 200, 300])[idx].tolist() == [200, 300, 100]
    d = np.array([[1, 1, 1], [-1, -1, -1], [1, 1, 1], [-1, -1, -1]], f32)
    rays = {"ox": np.zeros(4, f32), "oy": np.zeros(4, f32), "oz": np.zeros(4, f32),
            "dx": d[:, 0], "dy": d[:, 1], "dz": d[:, 2]}
    idx, _ = wf.ray_schedule(ctx, upload_rays(wf, ctx, rays), 4, (-10, -10, -10), (10, 10, 10))
    assert np.array([10, 20, 11, 21])[idx].tolist() == [10, 11, 20, 21]


# ---- accelerator: closest hit, occlusion, hit/miss split --------------------------------------
def test_build_stats(cornell):
    data, scene, osc = cornell
    st = scene.stats
    assert st.triangle_count == 5132 and st.mesh_count == 7
    bmin, bmax, area, ntri = osc.info()
    assert list(st.bounds_min) == bmin.tolist() and list(st.bounds_max) == bmax.tolist()
    assert 0 < st.max_depth
    assert st.bvh_width in (4, 8)
    assert 0 < st.stack_bound <= 64
    assert st.node_count < st.triangle_count


@pytest.mark.parametrize("width", [4, 8])
def test_build_stats_width(wf, ctx, cornell, width):
    data = cornell[0]
    sc = build_with(wf, ctx, data, bvh_width=width)
    st = sc.stats
    assert st.bvh_width == width
    # a wide node has at most width-1 pushed siblings and covers at least one binary level
    assert 0 < st.stack_bound <= min(64, (width - 1) * st.max_depth)
    # every wide node but the root is the child of another and holds >= 2 children
    assert st.node_count <= st.triangle_count - 1
    if width == 4:
        assert 0 < st.stack_bound <= 3 * ((st.max_depth + 1) // 2)


def build_with(wf, ctx, data, **tuning):
    saved = ctx.tuning()
    ctx.set_tuning(**tuning)
    try:
        return wf.Scene(ctx, data)
    finally:
        ctx.set_tuning(**saved)


def chain_scene(scenes):
    """Triangles centred on (2^-k, 0, 0), (0, 2^-k, 0) and (0, 0, 2^-k), k = 1..15, inside the
    centroid bounds [0, 1]^3: each splits the 48-bit Morton code one bit further from the
    origin's, so the binary LBVH is about 45 levels deep; a dense cluster adds a balanced
    subtree. The four-wide collapse sees both one-sided and balanced parts."""
    a = 0.01
    centres = [(0.0, 0.0, 0.0), (1.0, 1.0, 1.0)]
    for k in range(1, 16):
        x = 2.0 ** -k
        centres += [(x, 0.0, 0.0), (0.0, x, 0.0), (0.0, 0.0, x)]
    rng = np.random.default_rng(7)
    centres += [tuple(p) for p in rng.uniform(0.2, 0.3, (600, 3))]
    c = np.array(centres, np.float64)
    tris = np.stack([c + [-a, -a, 0], c + [a, -a, 0], c + [0, 2 * a, 0]], 1)
    pos = tris.astype(f32).reshape(-1, 3)
    idx = np.arange(len(pos), dtype=np.int32).reshape(-1, 3)
    return scenes.SceneData([scenes.Mesh(pos, idx, 0)], [{"type": 0}]), c


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("builder", ["lbvh", "ploc"])
def test_deep_tree_closest_hit_matches_bruteforce(wf, ctx, scenes, orc, rng, builder, width):
    data, centres = chain_scene(scenes)
    sc = build_with(wf, ctx, data, bvh_width=width,
                    bvh_builder=wf._abi.BVH_LBVH if builder == "lbvh" else wf._abi.BVH_PLOC)
    print("chain scene:", builder, "binary depth", sc.stats.max_depth, "stack bound",
          sc.stats.stack_bound, "width", sc.stats.bvh_width)
    if builder == "lbvh":
        assert sc.stats.max_depth > 32     # the Karras tree follows the code bits one by one
    assert sc.stats.stack_bound <= 64
    osc = orc.scene(data)
    n = 6000
    o = rng.uniform(-0.2, 1.2, (n, 3))
    o[:, 2] = rng.uniform(0.5, 1.5, n)             # in front of the z-facing triangles
    tgt = centres[rng.integers(0, len(centres), n)]
    d = (tgt - o).astype(f32)
    o = o.astype(f32)
    rays = {"ox": o[:, 0].copy(), "oy": o[:, 1].copy(), "oz": o[:, 2].copy(),
            "dx": d[:, 0].copy(), "dy": d[:, 1].copy(), "dz": d[:, 2].copy()}
    hits = sc.intersect(upload_rays(wf, ctx, rays), n)
    ref = osc.closest_hit(rays, use_bvh=False)
    assert (ref["prim"] >= 0).sum() > n // 2
    assert np.array_equal(hits["prim"], ref["prim"])
    h = ref["prim"] >= 0
    assert np.array_equal(hits["t"][h], ref["t"][h])


@pytest.mark.parametrize("builder,width", [("lbvh", 4), ("ploc", 4), ("ploc", 8)])
def test_large_scene_closest_hit_matches_oracle(wf, ctx, scenes, orc, rng, builder, width):
    """~330 k triangles: the oracle's SAH BVH (itself pinned against brute force) against the
    device BVH of either builder and node width, over camera rays and random rays."""
    data = scenes.sorting_stress(subdiv=6, tex_size=8)
    sc = build_with(wf, ctx, data, bvh_width=width,
                    bvh_builder=wf._abi.BVH_LBVH if builder == "lbvh" else wf._abi.BVH_PLOC)
    assert sc.stats.bvh_width == width
    osc = orc.scene(data)
    rays = camera_rays(orc, data, 128, 96)
    extra = random_rays(rng, 20000)
    for k in ("ox", "oy", "oz", "dx", "dy", "dz"):
        rays[k] = np.concatenate([rays[k], extra[k]])
    n = len(rays["ox"])
    hits = sc.intersect(upload_rays(wf, ctx, rays), n)
    ref = osc.closest_hit(rays, use_bvh=True)
    assert np.array_equal(hits["geom"], ref["geom"])
    assert np.array_equal(hits["prim"], ref["prim"])
    h = ref["prim"] >= 0
    assert np.array_equal(hits["t"][h], ref["t"][h])


@pytest.mark.parametrize("width", [4, 8])
def test_closest_hit_matches_bruteforce(wf, ctx, cornell, orc, rng, width):
    data, _, osc = cornell
    scene = build_with(wf, ctx, data, bvh_width=width)
    rays = camera_rays(orc, data, 96, 96)
    extra = random_rays(rng, 8000)
    for k in ("ox", "oy", "oz", "dx", "dy", "dz"):
        rays[k] = np.concatenate([rays[k], extra[k]])
    n = len(rays["ox"])
    hits = scene.intersect(upload_rays(wf, ctx, rays), n)
    ref = osc.closest_hit(rays, use_bvh=False)
    assert np.array_equal(hits["geom"], ref["geom"])
    assert np.array_equal(hits["prim"], ref["prim"])
    h = ref["prim"] >= 0
    assert h.sum() > n // 2
    assert np.array_equal(hits["t"][h], ref["t"][h])
    assert np.array_equal(hits["bu"][h], ref["bu"][h]) and np.array_equal(hits["bv"][h], ref["bv"][h])


def test_reference_embree_kats(wf, ctx, scenes):
    # tests/accelerator/EmbreeTest.cpp:35-156 through the device accelerator
    quad = scenes.quad((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1))
    data = scenes.SceneData([scenes.Mesh(*quad, 0)], [{"type": 0, "albedo": (0.8, 0.8, 0.8)}])
    sc = wf.Scene(ctx, data)

    def one(o, d, tmax=1e9):
        r = {"ox": [o[0]], "oy": [o[1]], "oz": [o[2]], "dx": [d[0]], "dy": [d[1]], "dz": [d[2]]}
        return sc.intersect(upload_rays(wf, ctx, {k: np.asarray(v, f32) for k, v in r.items()}), 1,
                            0.001, tmax)
    h = one((0, 0, 0), (0, 0, -1))
    assert h["geom"][0] == 0 and abs(h["t"][0] - 1.0) < 1e-3
    assert one((0, 0, 0), (0, 0, 1))["geom"][0] == -1
    assert one((5, 5, 0), (0, 0, -1))["geom"][0] == -1
    assert one((0, 0, 0), (0, 0, -1), 0.5)["geom"][0] == -1
    # multi-mesh: nearer mesh B wins (EmbreeTest.cpp:126-156)
    a = scenes.Mesh(np.array([[-1, -1, -2], [1, -1, -2], [0, 1, -2]], f32), np.array([[0, 1, 2]], np.int32), 0)
    b = scenes.Mesh(np.array([[-1, -1, -1], [1, -1, -1], [0, 1, -1]], f32), np.array([[0, 1, 2]], np.int32), 1)
    two = wf.Scene(ctx, scenes.SceneData([a, b], [{"type": 0}, {"type": 0}]))
    r = {k: np.asarray(v, f32) for k, v in {"ox": [0], "oy": [0], "oz": [0], "dx": [0], "dy": [0], "dz": [-1]}.items()}
    hh = two.intersect(upload_rays(wf, ctx, r), 1)
    assert hh["geom"][0] == 1 and abs(hh["t"][0] - 1.0) < 1e-2


def test_empty_scene_misses(wf, ctx, scenes, rng):
    # EmbreeTest.cpp:113-124
    sc = wf.Scene(ctx, scenes.SceneData([], [{"type": 0}]))
    rays = random_rays(rng, 100)
    h = sc.intersect(upload_rays(wf, ctx, rays), 100)
    assert np.all(h["geom"] == -1) and np.all(h["prim"] == -1)


@pytest.mark.parametrize("width", [4, 8])
def test_occluded_matches_bruteforce(wf, ctx, cornell, rng, width):
    data, _, osc = cornell
    scene = build_with(wf, ctx, data, bvh_width=width)
    n = 20000
    o = (rng.uniform(-0.45, 0.45, (n, 3)) + [0, 0.5, 0]).astype(f32)
    tgt = (rng.uniform(-0.45, 0.45, (n, 3)) + [0, 0.5, 0]).astype(f32)
    v = (tgt - o).astype(f32)
    dist = np.linalg.norm(v, axis=1).astype(f32)
    d = (v / dist[:, None]).astype(f32)
    dist[:50] = rng.uniform(0, 0.001, 50).astype(f32)           # tfar < 0 reports occluded
    got = scene.occluded(o, d, dist)
    want = osc.occluded(o, d, dist, use_bvh=False)
    assert np.array_equal(got, want)
    assert 0 < got.sum() < n
    assert got[:50].all()


def test_intersect_split_matches_oracle(wf, ctx, cornell, orc):
    data, scene, osc = cornell
    rays = camera_rays(orc, data, 128, 128, sample=3)
    n = len(rays["ox"])
    hits, nh, misses, nm = scene.intersect_split(upload_rays(wf, ctx, rays), n)
    ohits, onh, omiss, onm = osc.intersect_split(rays)
    assert (nh, nm) == (onh, onm) and nh > 0 and nm > 0
    gh, gm = hits.download(nh), misses.download(nm)
    for k in abi.SHADE_FIELDS:
        assert np.array_equal(gh[k], ohits[k]), k
    for k in abi.RAY_FIELDS:
        assert np.array_equal(gm[k], omiss[k]), k


# ---- ShadingQueue::schedule ------------------------------------------------------------------
def stress_queue(rng, n, n_mat):
    q = {k: np.zeros(n, abi.SHADE_DTYPES[k]) for k in abi.SHADE_FIELDS}
    q["mat"] = rng.integers(0, n_mat, n).astype(np.int32)
    q["px"], q["py"], q["pz"] = (rng.uniform(-2, 2, (3, n))).astype(f32)
    q["tr"], q["tg"], q["tb"] = (rng.uniform(0.05, 1.5, (3, n))).astype(f32)
    q["pixel"] = np.arange(n, dtype=np.int32)
    return q


@pytest.fixture(scope="module")
def policy_scene(wf, ctx, scenes):
    rng = np.random.default_rng(11)
    types = [0, 1, 4, 5, 0, 1, 0, 3, 2, 0, 1, 4]
    mats = []
    for m, t in enumerate(types):
        mats.append({"type": t, "albedo": (0.5, 0.5, 0.5), "texture_id": int(m % 5) - 1})
    tex = [{"width": w, "height": h, "format": abi.TEX_RGBA8_SRGB,
            "data": rng.integers(0, 2**32, w * h, dtype=np.uint64).astype(np.uint32)}
           for w, h in [(4, 4), (16, 8), (2, 2), (32, 32)]]
    tri = scenes.quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0))
    data = scenes.SceneData([scenes.Mesh(*tri, 0)], mats, tex)
    return data, wf.Scene(ctx, data), types, [m["texture_id"] for m in mats], [w * h * 12 for w, h in [(4, 4), (16, 8), (2, 2), (32, 32)]]


@pytest.mark.parametrize("policy", ["none", "material", "texture", "costBenefit"])
@pytest.mark.parametrize("n", [1, 5, 3000, 70_001])
def test_shading_schedule_bit_exact(wf, ctx, orc, policy_scene, policy, n, rng, radix_bits):
    data, scene, types, mtex, tmem = policy_scene
    q = stress_queue(rng, n, len(types))
    q["tex"] = np.asarray(mtex, np.int32)[q["mat"]]
    q["tex"][::11] = 7          # out-of-range texture id counts as 0 bytes
    dq = wf.ShadingQueue(ctx, n)
    dq.upload(q)
    got = wf.shading_schedule(scene, dq, n, policy)
    want = orc.shading_schedule(policy, q, types, mtex, tmem)
    assert np.array_equal(got, want)
    if policy != "none":
        out = wf.materialize_shading(ctx, dq, got, n).download(n)
        for k in ("mat", "tex", "px", "tr", "pixel"):
            assert np.array_equal(out[k], q[k][got]), k


@pytest.mark.parametrize("policy", ["material", "costBenefit"])
def test_shading_schedule_with_cost_tables(wf, ctx, orc, policy_scene, policy, rng):
    data, scene, types, mtex, tmem = policy_scene
    n = 20000
    q = stress_queue(rng, n, len(types))
    q["tex"] = np.asarray(mtex, np.int32)[q["mat"]]
    mc = rng.choice([0.5, 1.0, 1.7, 3.0], len(types))
    tc = rng.choice([0.8, 1.0, 2.5], len(tmem))
    scene.set_costs(mc, tc)
    try:
        dq = wf.ShadingQueue(ctx, n)
        dq.upload(q)
        got = wf.shading_schedule(scene, dq, n, policy)
        want = orc.shading_schedule(policy, q, types, mtex, tmem, mat_cost=mc, tex_cost=tc)
        assert np.array_equal(got, want)
    finally:
        scene.set_costs(None, None)


def test_shading_schedule_reference_kats(wf, ctx, scenes):
    # ShadingQueueTest.cpp:92-150 on the device
    mats = [{"type": t} for t in range(6)]
    tex = [{"width": w, "height": w, "format": abi.TEX_RGB_F32, "data": np.ones(w * w * 3, f32)}
           for w in (1, 2, 4)]
    tri = scenes.quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0))
    sc = wf.Scene(ctx, scenes.SceneData([scenes.Mesh(*tri, 0)], mats, tex))

    def run(policy, mat, texid):
        n = len(mat)
        q = {k: np.zeros(n, abi.SHADE_DTYPES[k]) for k in abi.SHADE_FIELDS}
        q["mat"] = np.array(mat, np.int32)
        q["tex"] = np.array(texid, np.int32)
        dq = wf.ShadingQueue(ctx, n)
        dq.upload(q)
        return wf.shading_schedule(sc, dq, n, policy)
    assert np.array([2, 0, 1])[run("material", [2, 0, 1], [0, 0, 0])].tolist() == [2, 1, 0]
    assert np.array([1, 0, 1, 0, 1])[run("material", [1, 0, 1, 0, 1], [0] * 5)].tolist() == [1, 1, 1, 0, 0]
    assert np.array([0, 2, 1])[run("texture", [0, 1, 0], [0, 2, 1])].tolist() == [2, 1, 0]
    assert run("none", [2, 0, 1], [0, 0, 0]).tolist() == [0, 1, 2]


# ---- shadeAll + miss loop -----------------------------------------------------------------
def close(a, b, rtol=2e-5, atol=1e-6):
    return np.allclose(a, b, rtol=rtol, atol=atol)


@pytest.fixture(scope="module")
def variants(wf, ctx, scenes, orc):
    data = scenes.material_variants()
    return data, wf.Scene(ctx, data), orc.scene(data)


@pytest.fixture(scope="module")
def stress_small(wf, ctx, scenes, orc):
    data = scenes.sorting_stress(subdiv=4, n_materials=16, n_textures=4, tex_size=32)
    return data, wf.Scene(ctx, data), orc.scene(data)


def shade_both(wf, ctx, scene, osc, q, n, pixels, params):
    """One shadeAll over the same queue on the device and in the oracle: next rays, shadow rays
    resolved against the scene, accumulators. Returns (oracle next rays, count, device next
    rays, device accumulators, oracle accumulators)."""
    acc_o = np.zeros(pixels * 3, f32)
    onext, onn, _ = osc.shade(q, params, acc_o)
    dq = wf.ShadingQueue(ctx, n)
    dq.upload(q)
    acc = wf.DeviceBuffer(ctx, pixels * 3, f32)
    ctx.fill_u32(acc.ptr, 0, pixels * 3)
    nxt, nn = wf.shade(scene, dq, n, params, acc)
    assert nn == onn
    return onext, onn, nxt.download(nn), acc, acc_o


def ulp_diff(a, b):
    """Largest distance in float32 units in the last place between two arrays (0 = same bits)."""
    a = np.ascontiguousarray(a, f32).view(np.int32).astype(np.int64)
    b = np.ascontiguousarray(b, f32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return int(np.abs(a - b).max()) if len(a) else 0


def assert_same_bits(got, want, keys, what=""):
    diffs = {k: ulp_diff(got[k], want[k]) for k in keys}
    bad = {k: v for k, v in diffs.items() if v}
    assert not bad, f"{what}: fields differing (max ulp) {bad}"


def assert_next_rays(g, want):
    assert np.array_equal(g["pixel"], want["pixel"]) and np.array_equal(g["depth"], want["depth"])
    assert_same_bits(g, want, ("ox", "oy", "oz", "dx", "dy", "dz", "tr", "tg", "tb", "pdf"),
                     "next rays")


SHADE_SCENES = {
    # Cornell box: Diffuse + Emissive (reference model)
    "cornell": dict(fixture="cornell", W=96, H=96, bounces=4),
    # fuzz Metal, Schlick Glass from both sides, Plastic, SpotLight, spatial checker, RGBA8
    # textures, directional light, environment map (reference model)
    "variants": dict(fixture="variants", W=96, H=96, bounces=4),
    # GGX conductor, exact dielectric, textured Lambertian, two area lights (microfacet model)
    "stress": dict(fixture="stress_small", W=128, H=96, bounces=4),
}


@pytest.mark.parametrize("physics", [0, 1], ids=["reference", "unbiased"])
@pytest.mark.parametrize("case", list(SHADE_SCENES))
def test_shade_bounce_chain_matches_oracle(wf, ctx, orc, request, case, physics):
    """shadeAll (cpp:569-877) and the miss loop (cpp:256-265) on the device against the oracle,
    bounce after bounce: every bounce's queue is the oracle's intersectAll of the oracle's next
    rays, and the device's own intersectAll of those rays must equal it (the rays that refract
    into the glass hit its faces from inside, frontFace false). Next rays, shadow-ray
    contributions and miss radiance are compared per entry and per pixel."""
    cfg = SHADE_SCENES[case]
    data, scene, osc = request.getfixturevalue(cfg["fixture"])
    W, H = cfg["W"], cfg["H"]
    rays = camera_rays(orc, data, W, H, sample=1)
    shaded = np.zeros(len(data.materials), np.int64)
    back_faces = 0
    for bounce in range(cfg["bounces"]):
        n = len(rays["ox"])
        ohits, onh, omiss, onm = osc.intersect_split(rays)
        hits, nh, misses, nm = scene.intersect_split(upload_rays(wf, ctx, rays), n)
        assert (nh, nm) == (onh, onm)
        gh, gm = hits.download(nh), misses.download(nm)
        for k in abi.SHADE_FIELDS:
            assert np.array_equal(gh[k], ohits[k]), (bounce, k)
        for k in abi.RAY_FIELDS:
            assert np.array_equal(gm[k], omiss[k]), (bounce, k)
        if bounce == 0:
            # primary hits: force the MIS and Russian-roulette branches
            ohits["depth"] = (np.arange(onh) % 5).astype(np.int32)
            ohits["pdf"] = np.where(np.arange(onh) % 3 == 0, f32(-1.0), f32(0.4)).astype(f32)
        shaded += np.bincount(ohits["mat"], minlength=len(shaded))
        back_faces += int((ohits["front"] == 0).sum())
        params = wf.shade_params(physics=physics, firefly=1e30, seed=99, sample=1)
        onext, onn, g, acc, acc_o = shade_both(wf, ctx, scene, osc, ohits, onh, W * H, params)
        assert_next_rays(g, onext)
        assert ulp_diff(acc.download(), acc_o) == 0, (bounce, "shade accumulators")
        mq = upload_rays(wf, ctx, omiss)
        wf.shade_misses(scene, mq, onm, acc, physics)
        osc.shade_misses(omiss, physics, acc_o)
        assert ulp_diff(acc.download(), acc_o) == 0, (bounce, "miss accumulators")
        rays = onext
        if onn == 0:
            break
    assert (shaded > 0).all(), shaded
    if case == "variants":
        assert back_faces > 100     # Schlick Glass entered and left (eta = ior on the way out)


def test_shade_variants_directional_term(wf, ctx, variants, orc):
    """The directional light's NEE (cpp:712-737, every material type) is a term the per-bounce
    comparison sees: without it the accumulators change, and device and oracle still agree."""
    data, scene, osc = variants
    W = H = 64
    rays = camera_rays(orc, data, W, H, sample=2)
    ohits, onh, _, _ = osc.intersect_split(rays)
    params = wf.shade_params(physics=0, firefly=1e30, seed=5, sample=2)
    _, _, _, acc, acc_o = shade_both(wf, ctx, scene, osc, ohits, onh, W * H, params)
    assert ulp_diff(acc.download(), acc_o) == 0
    d2 = type(data)(**{**data.__dict__})
    d2.directional = None
    sc2 = wf.Scene(ctx, d2)
    _, _, _, acc2, acc2_o = shade_both(wf, ctx, sc2, orc.scene(d2), ohits, onh, W * H, params)
    assert ulp_diff(acc2.download(), acc2_o) == 0
    lit = acc_o - acc2_o
    assert (lit >= -1e-6).all() and (lit > 1e-3).sum() > 1000
    sc2.close()


# ---- renderScene ------------------------------------------------------------------------------
def render_both(wf, orc_scene, scene, data, **kw):
    params = wf.render_params(camera=data.camera, **kw)
    img, cnt, st = wf.render(scene, params)
    oimg, ocnt, _, ost = orc_scene.render(params)
    return img, cnt, st, oimg, ocnt, ost


def assert_image_matches_oracle(img, cnt, st, oimg, ocnt, ost, lanes):
    """Same paths on both sides: equal sample counts and ray statistics; with one sample lane the
    pixel sums are in the oracle's order (same bits), with L lanes each pixel is a sum of at most
    spp * (depth + 2) non-negative terms in another order (relative difference <= n * 2^-24)."""
    assert np.array_equal(cnt, ocnt)
    assert (st.rays_traced, st.shadow_rays, st.shaded_hits, st.misses) == \
        (ost.rays_traced, ost.shadow_rays, ost.shaded_hits, ost.misses)
    if lanes == 1:
        assert ulp_diff(img, oimg) == 0
        return
    rel = np.abs(img - oimg) / np.maximum(np.abs(oimg), 1e-30)
    worst = np.argsort(rel.reshape(-1))[::-1][:5]
    report = [(int(i // 3), float(img.reshape(-1)[i]), float(oimg.reshape(-1)[i])) for i in worst]
    assert rel.max() <= 2e-5, f"worst pixels (index, device, oracle): {report}"


@pytest.mark.parametrize("lanes", [1, 64])
def test_render_matches_oracle_cornell(wf, cornell, lanes):
    data, scene, osc = cornell
    out = render_both(wf, osc, scene, data, width=64, height=64, spp=4, max_depth=8, firefly=1e30,
                      seed=5, sample_lanes=lanes)
    assert_image_matches_oracle(*out, lanes=lanes)


@pytest.mark.parametrize("lanes", [1, 64])
@pytest.mark.parametrize("physics", [0, 1], ids=["reference", "unbiased"])
def test_render_matches_oracle_variants(wf, variants, physics, lanes):
    """Fuzz Metal, Schlick Glass, Plastic, SpotLight, spatial checker, textures, a directional
    light and an environment map through the whole frame loop."""
    data, scene, osc = variants
    out = render_both(wf, osc, scene, data, width=64, height=48, spp=4, max_depth=8,
                      firefly=1e30, seed=13, physics=physics, sample_lanes=lanes)
    assert_image_matches_oracle(*out, lanes=lanes)
    assert out[2].shadow_rays > 0 and out[2].misses > 0


@pytest.mark.parametrize("policy", ["none", "costBenefit"])
def test_render_matches_oracle_stress(wf, stress_small, policy):
    """GGX conductor, exact dielectric and textured Lambertian through the frame loop (unbiased
    track: power-heuristic MIS, alias-table light sampling)."""
    data, scene, osc = stress_small
    out = render_both(wf, osc, scene, data, width=64, height=48, spp=2, max_depth=12,
                      firefly=1e30, seed=17, physics=wf.PHYSICS_UNBIASED, policy=policy,
                      sample_lanes=1)
    assert_image_matches_oracle(*out, lanes=1)


@pytest.mark.parametrize("policy", ["material", "texture", "costBenefit"])
def test_render_independent_of_policy_and_ray_sort(wf, cornell, policy):
    """Scheduling only reorders work: every policy and ray sort on/off give the identical image."""
    data, scene, osc = cornell
    base = wf.render_params(camera=data.camera, width=48, height=40, spp=2, max_depth=6, seed=3,
                            firefly=1e30)
    img0, _, st0 = wf.render(scene, base)
    p = wf.render_params(camera=data.camera, width=48, height=40, spp=2, max_depth=6, seed=3,
                         firefly=1e30, policy=policy, ray_sort=False)
    img1, _, st1 = wf.render(scene, p)
    assert np.array_equal(img0, img1)
    assert st0.rays_traced == st1.rays_traced


def test_render_independent_of_bvh_width(wf, ctx, cornell):
    """The closest hit is the (t, triangle) minimum over every accepted triangle and the
    quantised eight-wide boxes contain the four-wide ones: both node widths render the same
    image, ray for ray."""
    data = cornell[0]
    p = wf.render_params(camera=data.camera, width=64, height=48, spp=4, max_depth=8, seed=9,
                         firefly=1e30)
    out = []
    for width in (4, 8):
        sc = build_with(wf, ctx, data, bvh_width=width)
        assert sc.stats.bvh_width == width
        img, cnt, st = wf.render(sc, p)
        out.append((img, st.rays_traced, st.shadow_rays))
        sc.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1:] == out[1][1:]


@pytest.mark.parametrize("lanes", [1, 0])
def test_render_adaptive_matches_oracle(wf, cornell, lanes):
    """AdaptiveSampler in the frame loop. One sample lane feeds each pixel's Welford mean in the
    oracle's order: the same counts and image bits. Several lanes (0: the default) feed a round of
    lanes' samples at once, in sample order, and test at the same counts, so the means differ by
    reassociation only and a pixel on the edge of the threshold may stop one round apart."""
    data, scene, osc = cornell
    img, cnt, st, oimg, ocnt, ost = render_both(
        wf, osc, scene, data, width=40, height=32, spp=64, max_depth=6, adaptive=True,
        adaptive_threshold=0.2, adaptive_min_samples=8, adaptive_interval=8, seed=11, firefly=10.0,
        sample_lanes=lanes)
    assert cnt.min() >= 1 and cnt.max() <= 64
    if lanes == 1:
        assert np.array_equal(cnt, ocnt) and st.samples_run == ost.samples_run
        assert ulp_diff(img, oimg) == 0
    else:
        assert np.mean(cnt == ocnt) > 0.97


def test_render_rejects_bad_arguments(wf, ctx, cornell):
    data, scene, osc = cornell
    p = wf.render_params(camera=data.camera, width=1, height=4, spp=1, max_depth=2)
    with pytest.raises(wf.WavefrontError):
        wf.render(scene, p)


# ---- coherence metrics (src/render/WavefrontRenderer.cpp:67-92, 187-246) -----------------------
@pytest.mark.parametrize("n", [1, 15, 16, 17, 1000, 70_001])
@pytest.mark.parametrize("layout", ["random", "runs"])
def test_coherence_stats_matches_oracle(wf, ctx, orc, rng, n, layout):
    mat = rng.integers(0, 6, n).astype(np.int32)
    tex = np.where(rng.random(n) < 0.3, -1, rng.integers(0, 4, n)).astype(np.int32)
    if layout == "runs":   # long runs, as after a material-aware sort
        mat = np.sort(mat)
        tex = np.repeat(rng.integers(-1, 4, (n + 99) // 100), 100)[:n].astype(np.int32)
    got = wf.coherence_stats(ctx, mat, tex)
    assert got == orc.coherence(mat, tex), (got, orc.coherence(mat, tex))
    assert wf.coherence_stats(ctx, mat, tex, cache_lines=False)[2:] == (0.0, 0.0)


def test_coherence_stats_empty(wf, ctx):
    assert wf.coherence_stats(ctx, np.zeros(0, np.int32), np.zeros(0, np.int32)) == (0.0,) * 4


@pytest.fixture(scope="module")
def textured(wf, ctx, scenes, orc):
    data = scenes.sorting_stress(subdiv=3, n_materials=8, n_textures=4, tex_size=16)
    return data, wf.Scene(ctx, data), orc.scene(data)


@pytest.mark.parametrize("policy", ["none", "material", "texture"])
def test_render_coherence_matches_oracle(wf, textured, policy):
    # depth 1: the shading queues are the primary hits, bit-identical on both sides
    data, scene, osc = textured
    kw = dict(width=48, height=40, spp=3, max_depth=1, firefly=1e30, seed=9, policy=policy,
              schedule="passes", physics=wf.PHYSICS_UNBIASED)
    img, _, st, _, _, ost = render_both(wf, osc, scene, data, coherence_stats=2, **kw)
    got = (st.material_run_length, st.texture_run_length, st.material_homogeneity,
           st.texture_homogeneity)
    want = (ost.material_run_length, ost.texture_run_length, ost.material_homogeneity,
            ost.texture_homogeneity)
    assert st.shaded_hits == ost.shaded_hits > 0
    assert got == want and all(v > 0 for v in got), (got, want)
    # statistics change no result, and are zero when not asked for
    img0, _, st0 = wf.render(scene, wf.render_params(camera=data.camera, **kw))
    assert np.array_equal(img, img0)
    assert st0.material_run_length == 0.0 and st0.texture_homogeneity == 0.0


def test_render_coherence_regenerate(wf, textured):
    data, scene, _ = textured
    kw = dict(width=48, height=40, spp=4, max_depth=4, firefly=1e30, seed=3, policy="material",
              physics=wf.PHYSICS_UNBIASED)
    img1, _, st1 = wf.render(scene, wf.render_params(camera=data.camera, coherence_stats=1, **kw))
    img0, _, st0 = wf.render(scene, wf.render_params(camera=data.camera, **kw))
    assert np.array_equal(img0, img1)
    assert st1.material_run_length >= 1.0 and st1.texture_run_length >= 1.0
    assert st1.material_homogeneity == 0.0 and st0.material_run_length == 0.0
