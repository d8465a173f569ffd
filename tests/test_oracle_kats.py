"""The oracle against the reference's own known answers (its GoogleTest suite) and closed forms.

Each test cites the reference test it restates. These run on the CPU only."""
import math

import numpy as np
import pytest

import oracle_lib

abi = oracle_lib.abi


# ---- counter RNG: Philox4x32-10 known-answer vectors (Salmon et al. 2011, Random123 kat) ------
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(orc, ctr, key, expect):
    assert orc.philox(ctr, key).tolist() == expect


def test_uniform_is_24bit_grid_in_unit_interval(orc):
    xs = np.array([orc.uniform(7, 0, p, s, d, dim) for p in range(20) for s in range(3)
                   for d in range(3) for dim in range(12)])
    assert xs.min() >= 0.0 and xs.max() < 1.0
    assert np.all(xs * 16777216.0 == np.floor(xs * 16777216.0))
    # one Philox block serves dims 4k..4k+3: the lane is the block output >> 8
    blk = orc.philox([5, 2, (1 << 10) | 0, 0], [9, 0])
    for lane in range(4):
        assert orc.uniform(9, 0, 5, 2, 1, lane) == np.float32((blk[lane] >> 8) / 16777216.0)


# ---- RayQueue (tests/scheduling/RayQueueTest.cpp) --------------------------------------------
BMIN, BMAX = (-10.0, -10.0, -10.0), (10.0, 10.0, 10.0)


def rays_of(origins, dirs):
    o = np.asarray(origins, np.float32).reshape(-1, 3)
    d = np.asarray(dirs, np.float32).reshape(-1, 3)
    return {"ox": o[:, 0], "oy": o[:, 1], "oz": o[:, 2], "dx": d[:, 0], "dy": d[:, 1], "dz": d[:, 2]}


def test_ray_schedule_spatial_order_kat(orc):
    # RayQueueTest.cpp:156-171 -> pixel order 200, 300, 100
    r = rays_of([(5, 0, 0), (-5, 0, 0), (0, 0, 0)], [(0, 0, 1)] * 3)
    pix = np.array([100, 200, 300])
    assert pix[orc.ray_schedule(r, BMIN, BMAX)].tolist() == [200, 300, 100]


def test_ray_schedule_groups_octants_kat(orc):
    # RayQueueTest.cpp:132-154: the two +++ rays and the two --- rays are grouped
    r = rays_of([(0, 0, 0)] * 4, [(1, 1, 1), (-1, -1, -1), (1, 1, 1), (-1, -1, -1)])
    pix = np.array([10, 20, 11, 21])
    order = pix[orc.ray_schedule(r, BMIN, BMAX)].tolist()
    assert sorted(order[:2]) in ([10, 11], [20, 21]) and sorted(order[2:]) in ([10, 11], [20, 21])
    assert sorted(order[:2]) != sorted(order[2:])
    # canonical tie order: stable by queue index
    assert order == [10, 11, 20, 21]


def test_ray_schedule_clamps_and_single(orc):
    # RayQueueTest.cpp:205-215 and 119-130
    r = rays_of([(9999, 9999, 9999), (0, 0, 0)], [(0, 0, 1)] * 2)
    keys = orc.ray_keys(r, BMIN, BMAX)
    assert keys[0] == (1 << 30) - 1      # every axis clamps to cell 1023, octant 0
    assert sorted(orc.ray_schedule(r, BMIN, BMAX).tolist()) == [0, 1]
    one = rays_of([(3, 4, 5)], [(0, 0, -1)])
    assert orc.ray_schedule(one, BMIN, BMAX).tolist() == [0]


def test_ray_schedule_flat_scene_guard(orc):
    # SceneTest.cpp:67-80: zero extent on z must not divide by zero (extent clamps to 1e-6)
    r = rays_of([(0, 0, 0), (0.5, 0.5, 0)], [(0, 0, -1)] * 2)
    keys = orc.ray_keys(r, (-1, -1, 0), (1, 1, 0))
    assert np.all(keys >> 30 == 4)   # dz < 0 -> octant 4
    assert keys[0] != keys[1]


def _spread_bruteforce(q):
    out = 0
    for b in range(10):
        out |= ((q >> b) & 1) << (3 * b)
    return out


def test_ray_keys_closed_form(orc, rng):
    # key = octant << 30 | interleave(qx, qy, qz), q = trunc(clamp((v-min)/ext,0,1) * 1023)
    n = 2000
    o = rng.uniform(-12, 12, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d[::7, 0] = -0.0   # -0.0 is not negative: octant bit stays 0
    keys = orc.ray_keys(rays_of(o, d), BMIN, BMAX)
    ext = np.float32(20.0)
    for i in range(0, n, 37):
        q = [int(np.float32(min(max(np.float32(np.float32(o[i, a] - np.float32(-10.0)) / ext), 0.0), 1.0))
                 * np.float32(1023.0)) for a in range(3)]
        octant = int(d[i, 0] < 0) | (int(d[i, 1] < 0) << 1) | (int(d[i, 2] < 0) << 2)
        morton = _spread_bruteforce(q[0]) | (_spread_bruteforce(q[1]) << 1) | (_spread_bruteforce(q[2]) << 2)
        assert keys[i] == (octant << 30) | morton


def test_ray_schedule_is_stable_sort_of_keys(orc, rng):
    n = 5000
    o = np.round(rng.uniform(-10, 10, (n, 3)), 0).astype(np.float32)   # many ties
    d = rng.normal(size=(n, 3)).astype(np.float32)
    r = rays_of(o, d)
    keys = orc.ray_keys(r, BMIN, BMAX)
    idx = orc.ray_schedule(r, BMIN, BMAX)
    assert np.array_equal(idx, np.argsort(keys, kind="stable"))


# ---- ShadingQueue (tests/scheduling/ShadingQueueTest.cpp) ------------------------------------
MAT_TYPES = [0, 1, 2, 3, 4, 5]          # makeTestMaterials: Diffuse, Metal, Emissive, Spot, Glass, Plastic
MAT_TEX = [-1] * 6
TEX_MEM = [12, 48, 192]                 # Texture(1,1), (2,2), (4,4): memoryBytes = w*h*12


def sq(mat, tex=None):
    mat = np.asarray(mat, np.int32)
    return {"mat": mat, "tex": np.zeros_like(mat) if tex is None else np.asarray(tex, np.int32)}


def test_none_keeps_arrival_order(orc):
    # ShadingQueueTest.cpp:73-90
    q = sq([2, 0, 1])
    assert orc.shading_schedule("none", q, MAT_TYPES, MAT_TEX, TEX_MEM).tolist() == [0, 1, 2]


def test_material_aware_type_rank_kat(orc):
    # ShadingQueueTest.cpp:92-109: equal counts -> type index descending: mats 2, 1, 0
    q = sq([2, 0, 1])
    order = orc.shading_schedule("material", q, MAT_TYPES, MAT_TEX, TEX_MEM)
    assert q["mat"][order].tolist() == [2, 1, 0]


def test_material_aware_groups_by_count_kat(orc):
    # ShadingQueueTest.cpp:111-130: 1,0,1,0,1 -> 1,1,1,0,0
    q = sq([1, 0, 1, 0, 1])
    order = orc.shading_schedule("material", q, MAT_TYPES, MAT_TEX, TEX_MEM)
    assert q["mat"][order].tolist() == [1, 1, 1, 0, 0]


def test_texture_aware_kat(orc):
    # ShadingQueueTest.cpp:132-150: (mat0,tex0), (mat1,tex2), (mat0,tex1) -> textures 2, 1, 0
    q = sq([0, 1, 0], [0, 2, 1])
    order = orc.shading_schedule("texture", q, MAT_TYPES, MAT_TEX, TEX_MEM)
    assert q["tex"][order].tolist() == [2, 1, 0]


def test_material_aware_misc_kats(orc):
    # ShadingQueueTest.cpp:152-165 single entry, 167-183 already sorted, 185-200 pixels preserved
    assert orc.shading_schedule("material", sq([5]), MAT_TYPES, MAT_TEX, TEX_MEM).tolist() == [0]
    q = sq([2, 1, 0])
    assert q["mat"][orc.shading_schedule("material", q, MAT_TYPES, MAT_TEX, TEX_MEM)].tolist() == [2, 1, 0]
    q = sq([1, 0])
    pix = np.array([42, 99])
    assert pix[orc.shading_schedule("material", q, MAT_TYPES, MAT_TEX, TEX_MEM)].tolist() == [42, 99]
    # ShadingQueueTest.cpp:240-270: Plastic (type 5) before Metal (type 1) at equal counts
    q = sq([5, 1])
    assert q["mat"][orc.shading_schedule("material", q, MAT_TYPES, MAT_TEX, TEX_MEM)].tolist() == [5, 1]


def test_material_aware_cost_subsort(orc):
    # within a type rank, higher relativeCost first (ShadingQueue.h:136-139); ties stable
    types = [0, 0, 0]
    q = sq([0, 1, 2, 0, 1, 2])
    cost = [1.0, 3.0, 2.0]
    order = orc.shading_schedule("material", q, types, [-1] * 3, [], mat_cost=cost)
    assert order.tolist() == [1, 4, 2, 5, 0, 3]


def test_cost_benefit_orders_type_priority_morton(orc, rng):
    # ShadingQueue.h:437-510: type asc, priority desc (benefit/cost), Morton asc
    n = 400
    types = [0, 1, 0, 4]
    mat = rng.integers(0, 4, n).astype(np.int32)
    T = rng.uniform(0.1, 1.0, (n, 3)).astype(np.float32)
    p = rng.uniform(-3, 3, (n, 3)).astype(np.float32)
    q = {"mat": mat, "tex": np.full(n, -1, np.int32), "tr": T[:, 0], "tg": T[:, 1], "tb": T[:, 2],
         "px": p[:, 0], "py": p[:, 1], "pz": p[:, 2]}
    order = orc.shading_schedule("costBenefit", q, types, [-1] * 4, [])
    lum = (np.float32(0.2126) * T[:, 0] + np.float32(0.7152) * T[:, 1]) + np.float32(0.0722) * T[:, 2]
    prio = np.array([lum[mat == m].astype(np.float64).sum() for m in range(4)])
    lo, hi = p.min(0), p.max(0)
    morton = np.array([orc.hit_morton(p[i], lo, hi) for i in range(n)], np.uint64)
    key = sorted(range(n), key=lambda i: (types[mat[i]], -prio[mat[i]], int(morton[i]), i))
    assert order.tolist() == key


def test_cost_benefit_single_material_is_pure_morton(orc, rng):
    n = 300
    p = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    q = {"mat": np.zeros(n, np.int32), "tex": np.full(n, -1, np.int32),
         "tr": np.ones(n, np.float32), "tg": np.ones(n, np.float32), "tb": np.ones(n, np.float32),
         "px": p[:, 0], "py": p[:, 1], "pz": p[:, 2]}
    order = orc.shading_schedule("costBenefit", q, [0], [-1], [])
    lo, hi = p.min(0), p.max(0)
    morton = np.array([orc.hit_morton(p[i], lo, hi) for i in range(n)], np.uint64)
    assert np.array_equal(order, np.argsort(morton, kind="stable"))
    # computeMortonCode closed form: corner points quantise to 0 and 1023 on every axis
    assert orc.hit_morton(lo, lo, hi) == 0
    assert orc.hit_morton(hi, lo, hi) == (1 << 30) - 1


# ---- geometry (tests/geometry/TriangleTest.cpp, tests/shading/MaterialFeaturesTest.cpp) -------
def test_triangle_kats(orc):
    v0, v1, v2 = (-1, 0, 0), (1, 0, 0), (0, 1, 0)
    hit, t, _, _ = orc.intersect_triangle((0, 0.3, 1), (0, 0, -1), v0, v1, v2)
    assert hit and abs(t - 1.0) < 1e-4                                   # TriangleTest.cpp:4-19
    assert not orc.intersect_triangle((5, 5, 1), (0, 0, -1), v0, v1, v2)[0]   # 21-32
    assert not orc.intersect_triangle((0, 0.3, 1), (1, 0, 0), v0, v1, v2)[0]  # 34-45
    assert not orc.intersect_triangle((0, 0.3, 5), (0, 0, -1), v0, v1, v2, 0.001, 2.0)[0]  # 47-54


def test_spatial_hash_kats(orc):
    # MaterialFeaturesTest.cpp:27-39
    assert orc.hash_grid_cell(3, -2, 7) == orc.hash_grid_cell(3, -2, 7)
    assert orc.hash_grid_cell(0, 0, 0) != orc.hash_grid_cell(1, 0, 0)
    # closed form: ((x*p1) ^ (y*p2) ^ (z*p3)) * 2654435761 mod 2^32
    for x, y, z in [(3, -2, 7), (0, 0, 0), (-5, 11, 2**20)]:
        h = ((x * 73856093) ^ (y * 19349663) ^ (z * 83492791)) & 0xFFFFFFFF
        assert orc.hash_grid_cell(x, y, z) == (h * 2654435761) & 0xFFFFFFFF


def test_luminance_and_srgb_constants():
    # AdaptiveSamplerTest.cpp:4-20 and TextureTest.cpp:104-146 (sRGB decode within 0.01)
    lum = lambda c: np.float32(0.2126) * c[0] + np.float32(0.7152) * c[1] + np.float32(0.0722) * c[2]
    assert lum(np.zeros(3, np.float32)) == 0.0
    assert abs(lum(np.ones(3, np.float32)) - 1.0) < 1e-4
    v = 191 / 255.0
    lin = ((v + 0.055) / 1.055) ** 2.4
    assert abs(lin - 0.52) < 0.01


# ---- coherence metrics (src/render/WavefrontRenderer.cpp:67-92, 201-246) ---------------------
@pytest.mark.parametrize("mat, tex, expect", [
    # hand-counted: kept entries / runs of equal IDs among them, for material then texture IDs
    ([1, 1, 2, 2, 2, 3], [0, 0, 0, 0, 0, 0], (6 / 3, 6 / 1)),
    ([5, 7, 5, 5], [0, -1, 0, 1], (3 / 1, 3 / 2)),
    # a skipped entry does not end a run: the kept materials are 1, 1
    ([1, 2, 1], [0, -1, 0], (2 / 1, 2 / 1)),
    ([4, 4], [-1, -1], (0.0, 0.0)),   # nothing kept
    ([], [], (0.0, 0.0)),
    ([0, 1, 0, 1, 0], [3, 3, 3, 3, 3], (5 / 5, 5 / 1)),
])
def test_run_length_kats(orc, mat, tex, expect):
    got = orc.coherence(mat, tex)
    assert got[0] == expect[0] and got[1] == expect[1]


def test_line_homogeneity_kats(orc):
    ids = np.full(32, 7, np.int32)
    assert orc.coherence(ids, ids)[2:] == (1.0, 1.0)
    one_off = ids.copy()
    one_off[17] = 8   # second line mixed
    assert orc.coherence(one_off, ids)[2] == 0.5
    tail = np.concatenate([ids, [9]]).astype(np.int32)   # a one-entry last line counts as one ID
    assert orc.coherence(tail, tail)[2] == 1.0
    assert orc.coherence(np.arange(20, dtype=np.int32), ids[:20])[2:] == (0.0, 1.0)
    assert orc.coherence([], [])[2:] == (0.0, 0.0)
