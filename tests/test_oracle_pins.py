"""Pins of the oracle's light transport against mathematics and the reference's definitions, for the
functions whose reference sources cannot be compiled here (they include Embree): Scene::sampleLight
and the NEE assembly (src/core/Scene.cpp:162-237, src/render/WavefrontRenderer.cpp:660-710), the MIS
helpers (include/shading/MIS.h:6-31), getSkyColor (cpp:879-889), EnvironmentMap::sample
(src/core/EnvironmentMap.cpp:30-48) and Russian roulette (cpp:765-789).

Direct lighting is checked through the oracle's own per-bounce steps (shade, intersectAll, shade
of the light hits, miss shading) at one receiver point, many independent paths: the mean radiance
leaving a Lambertian floor lit by a uniformly emitting rectangle is (albedo / pi) E with E from
Lambert's polygon formula. When the light's triangles have unequal areas the reference's NEE pdf
(1 / total area, Scene.cpp:234) is not the sampling density, and its expectation is the biased
value computed here by quadrature; a plausible slip (area-weighted triangle choice, a swapped MIS
weight in one of the two estimators, a missing cosine) moves the mean by many standard errors."""
import numpy as np
import pytest

import oracle_lib

abi = oracle_lib.abi
f32 = np.float32
ALBEDO = 0.5
LE = 10.0


def lambert_irradiance(p, n, verts):
    """Irradiance at p (normal n) from a polygon of unit radiance (Lambert 1760): half the sum over
    edges of the edge's subtended angle times the cosine of its plane normal with n."""
    r = [np.asarray(v, float) - p for v in verts]
    e = 0.0
    for i in range(len(r)):
        a, b = r[i], r[(i + 1) % len(r)]
        c = np.cross(a, b)
        ang = np.arctan2(np.linalg.norm(c), np.dot(a, b))
        e += ang * np.dot(n, c / np.linalg.norm(c))
    return 0.5 * abs(e)


def light_scene(scenes, tris, env=True):
    """A Lambertian floor (y = 0, albedo 0.5, normal +y), emitting triangles at y = 1 facing down
    (radiance 10) and, with env, an all-black environment so that misses add nothing."""
    floor = scenes.quad((-20, 0, -20), (-20, 0, 20), (20, 0, 20), (20, 0, -20))
    pos = np.asarray(tris, f32).reshape(-1, 3)
    idx = np.arange(len(pos), dtype=np.int32).reshape(-1, 3)
    mats = [{"type": abi.MAT_DIFFUSE, "albedo": (ALBEDO,) * 3},
            {"type": abi.MAT_EMISSIVE, "albedo": (LE,) * 3, "emission": 1.0}]
    data = scenes.SceneData([scenes.Mesh(*floor, 0), scenes.Mesh(pos, idx, 1)], mats)
    if env:
        data.environment = {"width": 4, "height": 2, "rgb": np.zeros(24, f32)}
    return data


def receiver_paths(osc, wf, p, n_paths, physics, seed=3):
    """Radiance leaving point p on the floor towards the camera, one independent path per pixel:
    NEE at p, then the BSDF-sampled ray's hit (emission with its MIS weight) or miss."""
    q = {k: np.zeros(n_paths, abi.SHADE_DTYPES[k]) for k in abi.SHADE_FIELDS}
    q["px"][:], q["py"][:], q["pz"][:] = p
    q["ny"][:] = 1.0
    q["dy"][:] = -1.0
    q["t"][:] = 2.0
    q["tr"][:] = q["tg"][:] = q["tb"][:] = 1.0
    q["pixel"] = np.arange(n_paths, dtype=np.int32)
    q["pdf"][:] = -1.0
    q["front"][:] = 1
    q["geom"][:] = 0
    q["tex"][:] = -1
    acc = np.zeros(n_paths * 3, f32)
    params = wf.shade_params(physics=physics, firefly=1e30, seed=seed, sample=0)
    nxt, nn, _ = osc.shade(q, params, acc)
    hits, nh, miss, nm = osc.intersect_split(nxt)
    osc.shade(hits, params, acc)
    osc.shade_misses(miss, physics, acc)
    return acc.reshape(-1, 3)[:, 0].astype(np.float64)


def tri_quadrature(p, tri, g, k=400):
    """Integral of g(x) over a triangle by the midpoint rule on k^2 congruent sub-triangles."""
    a, b, c = (np.asarray(v, float) for v in tri)
    i, j = np.meshgrid(np.arange(k), np.arange(k), indexing="ij")
    keep = i + j < k
    u, v = (i[keep] + 1 / 3) / k, (j[keep] + 1 / 3) / k          # upright sub-triangles
    u2, v2 = (i[i + j < k - 1] + 2 / 3) / k, (j[i + j < k - 1] + 2 / 3) / k   # inverted ones
    uu, vv = np.concatenate([u, u2]), np.concatenate([v, v2])
    x = a + uu[:, None] * (b - a) + vv[:, None] * (c - a)
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a))
    return g(x).mean() * area


def integrand_parts(p, a_total):
    """f(x) = (albedo/pi) Le cosS cosL / d^2 and the reference's balance weight of NEE,
    w = pL / (pL + pB) with pL = (1/A_total) d^2 / cosL, pB = cosS / pi (MIS.h:6-31)."""
    def f(x):
        d = x - p
        d2 = (d * d).sum(1)
        cs = d[:, 1] / np.sqrt(d2)
        return ALBEDO / np.pi * LE * cs * cs / d2          # light faces down: cosL = cosS
    def w(x):
        d = x - p
        d2 = (d * d).sum(1)
        cs = d[:, 1] / np.sqrt(d2)
        pl = d2 / cs / a_total
        pb = cs / np.pi
        return pl / (pl + pb)
    return f, w


RECT = [(-0.25, 1.0, -0.25), (0.25, 1.0, -0.25), (0.25, 1.0, 0.25), (-0.25, 1.0, 0.25)]


def rect_tris(order):
    a, b, c, d = (RECT[i] for i in order)
    return [[a, b, c], [a, c, d]]


@pytest.mark.parametrize("physics", [0, 1], ids=["reference", "unbiased"])
def test_direct_light_matches_lambert(orc, wf, scenes, physics):
    """Equal-area triangles: the reference's pdf 1/A is exact and both tracks are unbiased."""
    data = light_scene(scenes, rect_tris((0, 1, 2, 3)))        # (b - a) x (c - a) along -y
    osc = orc.scene(data)
    assert osc.info()[2] == pytest.approx(0.25, rel=1e-6)
    for p in [(0.0, 0.0, 0.0), (0.6, 0.0, -0.3), (-1.2, 0.0, 0.9)]:
        p = np.array(p)
        want = ALBEDO / np.pi * LE * lambert_irradiance(p, np.array([0, 1, 0.0]),
                                                      [RECT[i] for i in (0, 3, 2, 1)])
        x = receiver_paths(osc, wf, p, 120_000, physics)
        se = x.std(ddof=1) / np.sqrt(len(x))
        assert abs(x.mean() - want) <= 4 * se + 1e-6 * want, (p, x.mean(), want, se)
        assert se < 0.01 * want


def test_direct_light_unequal_triangles_reference_bias(orc, wf, scenes):
    """Three triangles of 50%, 10% and 40% of the rectangle: sampleLight picks a triangle
    uniformly and a point uniformly in it, yet divides by 1/A_total (Scene.cpp:208-234), so NEE
    over-weights the small triangle; the BSDF-sampled hits carry the complementary weight with the
    same 1/A_total. The mean is the biased expectation, not Lambert's value."""
    A, B, C, D = (np.array(v) for v in RECT)
    E = C + 0.2 * (D - C)                                    # on edge CD: |CE| = 0.2 |CD|
    tris = [[A, B, C], [A, C, E], [A, E, D]]                 # areas 0.5, 0.1, 0.4 of 0.25; normal -y
    data = light_scene(scenes, [np.array(t, f32) for t in tris])
    osc = orc.scene(data)
    for t in tris:
        assert np.cross(t[1] - t[0], t[2] - t[0])[1] < 0     # NEE sees the front (Scene.cpp:224)
    a_total = 0.25
    p = np.array([0.3, 0.0, 0.5])                            # under the small triangle's side
    f, w = integrand_parts(p, a_total)
    areas = [0.5 * np.linalg.norm(np.cross(t[1] - t[0], t[2] - t[0])) for t in tris]
    assert np.allclose(np.array(areas) / a_total, [0.5, 0.1, 0.4])
    nee = sum(a_total / (3 * ak) * tri_quadrature(p, t, lambda x: f(x) * w(x))
              for ak, t in zip(areas, tris))
    bsdf = sum(tri_quadrature(p, t, lambda x: f(x) * (1 - w(x))) for t in tris)
    unbiased = sum(tri_quadrature(p, t, f) for t in tris)
    want = nee + bsdf
    x = receiver_paths(osc, wf, p, 200_000, 0)
    se = x.std(ddof=1) / np.sqrt(len(x))
    assert abs(want - unbiased) > 10 * se                     # the test can tell them apart
    assert abs(x.mean() - want) <= 4 * se, (x.mean(), want, unbiased, se)
    # the unbiased track samples triangles by area: Lambert's value
    y = receiver_paths(osc, wf, p, 200_000, 1)
    se1 = y.std(ddof=1) / np.sqrt(len(y))
    assert abs(y.mean() - unbiased) <= 4 * se1, (y.mean(), unbiased, se1)


def test_total_emissive_area_counts_emissive_and_spot_meshes(orc, scenes):
    """Scene::totalEmissiveArea over the meshes of Emissive and SpotLight materials
    (Scene.cpp:85-90, 162-190)."""
    q1 = scenes.quad((0, 1, 0), (0.25, 1, 0), (0.25, 1, 0.25), (0, 1, 0.25))
    q2 = scenes.quad((1, 1, 0), (1.5, 1, 0), (1.5, 1, 0.5), (1, 1, 0.5))
    q3 = scenes.quad((2, 1, 0), (4, 1, 0), (4, 1, 2), (2, 1, 2))
    mats = [{"type": abi.MAT_EMISSIVE, "albedo": (1, 1, 1), "emission": 1.0},
            {"type": abi.MAT_SPOTLIGHT, "albedo": (1, 1, 1), "emission": 1.0,
             "spot_outer_deg": 40.0, "spot_falloff_deg": 20.0},
            {"type": abi.MAT_DIFFUSE, "albedo": (0.5, 0.5, 0.5)}]
    data = scenes.SceneData([scenes.Mesh(*q1, 0), scenes.Mesh(*q2, 1), scenes.Mesh(*q3, 2)], mats)
    assert orc.scene(data).info()[2] == f32(0.0625) + f32(0.25)


def test_mis_helpers_known_answers(orc):
    """MIS.h:6-31 with asymmetric arguments: the first argument's share; area pdf times d^2 over
    |cos| floored at 1e-4."""
    L = orc.lib
    assert L.orc_mis_balance(1.0, 3.0) == 0.25 and L.orc_mis_balance(3.0, 1.0) == 0.75
    assert L.orc_mis_balance(0.0, 0.0) == 0.0 and L.orc_mis_balance(-1.0, 0.5) == 0.0
    assert L.orc_mis_power(1.0, 3.0) == f32(0.1) and L.orc_mis_power(3.0, 1.0) == f32(0.9)
    assert L.orc_area_to_angle(0.5, 2.0, 0.25) == 8.0
    assert L.orc_area_to_angle(0.5, 2.0, -0.25) == 8.0
    assert L.orc_area_to_angle(0.5, 2.0, 0.0) == f32(2.0) / f32(1e-4)
    assert L.orc_area_to_angle(2.0, 0.5, 1.0) == 0.5


def miss_radiance(orc, wf, scenes, dirs, env=None, physics=0):
    data = scenes.floor_and_light()
    if env is not None:
        data.environment = env
    osc = orc.scene(data)
    n = len(dirs)
    q = {k: np.zeros(n, abi.RAY_DTYPES[k]) for k in abi.RAY_FIELDS}
    d = np.asarray(dirs, f32)
    q["dx"], q["dy"], q["dz"] = d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy()
    q["tr"][:] = q["tg"][:] = q["tb"][:] = 1.0
    q["pixel"] = np.arange(n, dtype=np.int32)
    q["pdf"][:] = -1.0
    acc = np.zeros(n * 3, f32)
    osc.shade_misses(q, physics, acc)
    return acc.reshape(-1, 3)


def test_sky_gradient_known_answers(orc, wf, scenes):
    """getSkyColor (cpp:879-889): dark sky (0.02, 0.02, 0.05) straight up, horizon colour
    (0.05, 0.04, 0.03) straight down, their mean at the horizon; the direction is normalised
    first, so its length does not matter."""
    c = miss_radiance(orc, wf, scenes, [(0, 1, 0), (0, -1, 0), (1, 0, 0), (0, 7, 0), (0, 0, -3)])
    dark, horizon = np.array([0.02, 0.02, 0.05], f32), np.array([0.05, 0.04, 0.03], f32)
    assert np.allclose(c[0], dark, rtol=1e-6) and np.allclose(c[3], dark, rtol=1e-6)
    assert np.allclose(c[1], horizon, rtol=1e-6)
    mid = (horizon + dark) / 2
    assert np.allclose(c[2], mid, rtol=1e-6) and np.allclose(c[4], mid, rtol=1e-6)


@pytest.mark.parametrize("physics", [0, 1])
def test_environment_texel_round_trip(orc, wf, scenes, physics):
    """EnvironmentMap::sample (EnvironmentMap.cpp:30-48): the direction through the centre of
    texel (x, y) -- longitude (u - 1/2) 2 pi from +x towards +z, latitude (1/2 - v) pi -- looks up
    texel (x, y) itself, scaled by intensity and tint."""
    W, H = 24, 12
    ids = np.arange(W * H, dtype=np.float64)
    rgb = np.stack([ids, ids + 0.25, ids + 0.5], -1).astype(f32).reshape(-1)
    env = {"width": W, "height": H, "rgb": rgb, "intensity": 2.0, "tint": (1.0, 0.5, 0.25)}
    y, x = np.mgrid[0:H, 0:W]
    phi = ((x + 0.5) / W - 0.5) * 2 * np.pi
    lat = (0.5 - (y + 0.5) / H) * np.pi
    d = np.stack([np.cos(lat) * np.cos(phi), np.sin(lat), np.cos(lat) * np.sin(phi)], -1)
    c = miss_radiance(orc, wf, scenes, d.reshape(-1, 3) * 3.0, env, physics)
    want = np.stack([ids * 2.0, (ids + 0.25) * 2.0 * 0.5, (ids + 0.5) * 2.0 * 0.25], -1)
    assert np.array_equal(c.astype(np.float64), want.astype(f32).astype(np.float64))


def rr_queue(n, depth, t):
    q = {k: np.zeros(n, abi.SHADE_DTYPES[k]) for k in abi.SHADE_FIELDS}
    q["ny"][:] = 1.0
    q["dy"][:] = -1.0
    q["t"][:] = 1.0
    q["tr"][:], q["tg"][:], q["tb"][:] = t
    q["pixel"] = np.arange(n, dtype=np.int32)
    q["depth"][:] = depth
    q["pdf"][:] = -1.0
    q["front"][:] = 1
    q["tex"][:] = -1
    return q


@pytest.mark.parametrize("cost", [None, 2.0, 0.25])
def test_russian_roulette_unbiased(orc, wf, scenes, cost):
    """cpp:765-789: from depth + 1 >= 3 a path survives with p = min(0.95, max(T * albedo)) (with
    cost-aware RR clamp(p / relativeCost, 0.05, 0.95)) and carries T * albedo / p, so the expected
    next throughput is T * albedo; below that depth every path continues with T * albedo."""
    floor = scenes.quad((-5, 0, -5), (-5, 0, 5), (5, 0, 5), (5, 0, -5))
    data = scenes.SceneData([scenes.Mesh(*floor, 0)],
                            [{"type": abi.MAT_DIFFUSE, "albedo": (ALBEDO,) * 3}])
    osc = orc.scene(data)
    n = 200_000
    t = (0.6, 0.4, 0.2)
    want = np.array(t) * ALBEDO
    params = wf.shade_params(physics=0, firefly=1e30, seed=7, sample=0,
                             cost_aware_rr=cost is not None)
    mc = None if cost is None else np.array([cost])
    nxt, nn, _ = osc.shade(rr_queue(n, 1, t), params, np.zeros(n * 3, f32), mat_cost=mc)
    assert nn == n                                            # depth + 1 = 2 < 3: no roulette
    assert np.array_equal(nxt["tr"], np.full(n, f32(0.6) * f32(ALBEDO)))
    nxt, nn, _ = osc.shade(rr_queue(n, 2, t), params, np.zeros(n * 3, f32), mat_cost=mc)
    p = min(0.95, want.max())
    if cost is not None:
        p = min(0.95, max(0.05, p / cost))
    frac = nn / n
    assert abs(frac - p) <= 4 * np.sqrt(p * (1 - p) / n), (frac, p)
    for k, ch in zip(("tr", "tg", "tb"), range(3)):
        mean = nxt[k].astype(np.float64).sum() / n            # killed paths carry 0
        se = np.sqrt(frac * (1 - frac) / n) * want[ch] / p
        assert abs(mean - want[ch]) <= 4 * se + 1e-6, (k, mean, want[ch], se)
