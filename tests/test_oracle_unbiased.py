"""Closed-form pins for the unbiased track, which has no reference implementation: GGX
(visible-normal sampling vs evaluation, energy bound, pdf normalisation), exact dielectric
Fresnel (limits, Snell/TIR threshold), power/balance MIS weights, cosine-hemisphere density and
the alias tables (they must reproduce their input probabilities exactly)."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib

f32 = np.float32
P = C.c_void_p


def ptr(a):
    return a.ctypes.data_as(P)


@pytest.fixture(scope="module")
def L(orc):
    lib = orc.lib
    lib.orc_ggx_sample.restype = C.c_int32
    lib.orc_ggx_sample.argtypes = [P, C.c_float, P, P, C.c_float, C.c_float, P, P, P]
    lib.orc_ggx_eval.argtypes = [P, C.c_float, P, P, P, P, P]
    lib.orc_fresnel_dielectric.restype = C.c_float
    lib.orc_fresnel_dielectric.argtypes = [C.c_float, C.c_float]
    lib.orc_mis_power.restype = C.c_float
    lib.orc_mis_power.argtypes = [C.c_float, C.c_float]
    lib.orc_mis_balance.restype = C.c_float
    lib.orc_mis_balance.argtypes = [C.c_float, C.c_float]
    lib.orc_cosine_hemisphere.argtypes = [P, C.c_float, C.c_float, P]
    lib.orc_light_tables.restype = C.c_int32
    lib.orc_light_tables.argtypes = [P, P, P, P, P, C.c_int32]
    lib.orc_env_tables.restype = C.c_int32
    lib.orc_env_tables.argtypes = [P, P, P, P, P]
    return lib


def unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v)).astype(f32)


def ggx_sample(L, f0, r, n, wo, u1, u2):
    wi, w, pdf = np.zeros(3, f32), np.zeros(3, f32), C.c_float()
    ok = L.orc_ggx_sample(ptr(np.asarray(f0, f32)), r, ptr(n), ptr(wo), u1, u2, ptr(wi), ptr(w),
                          C.byref(pdf))
    return bool(ok), wi, w, pdf.value


def ggx_eval(L, f0, r, n, wo, wi):
    fc, pdf = np.zeros(3, f32), C.c_float()
    L.orc_ggx_eval(ptr(np.asarray(f0, f32)), r, ptr(n), ptr(wo), ptr(np.asarray(wi, f32)), ptr(fc),
                   C.byref(pdf))
    return fc, pdf.value


@pytest.mark.parametrize("rough", [0.1, 0.3, 0.6, 0.9])
def test_ggx_sample_consistent_with_eval(L, rough, rng):
    n = unit([0.2, 1.0, -0.3])
    for _ in range(300):
        wo = unit(rng.normal(size=3))
        if np.dot(wo, n) <= 0.05:
            wo = unit(wo + 2 * n)
        ok, wi, w, pdf = ggx_sample(L, (0.9, 0.6, 0.3), rough, n, wo, rng.random(), rng.random())
        if not ok:
            continue
        fc, epdf = ggx_eval(L, (0.9, 0.6, 0.3), rough, n, wo, wi)
        assert epdf == pytest.approx(pdf, rel=2e-3)
        # sampling weight = f*cos/pdf
        assert np.allclose(fc / epdf, w, rtol=3e-3, atol=1e-5)


@pytest.mark.parametrize("rough", [0.2, 0.5, 0.9])
def test_ggx_white_furnace_bounded(L, rough, rng):
    """F0 = 1: the single-scattering albedo is <= 1 and close to 1 for smooth surfaces."""
    n = np.array([0, 0, 1], f32)
    for cos_o in (0.9, 0.5, 0.2):
        wo = unit([np.sqrt(1 - cos_o ** 2), 0, cos_o])
        acc = []
        for _ in range(4000):
            ok, wi, w, pdf = ggx_sample(L, (1, 1, 1), rough, n, wo, rng.random(), rng.random())
            acc.append(w[0] if ok else 0.0)
        albedo = float(np.mean(acc))
        assert albedo <= 1.0 + 3 * np.std(acc) / np.sqrt(len(acc))
        if rough <= 0.2:
            assert albedo > 0.9


def test_ggx_pdf_integrates_to_visible_fraction(L, rng):
    """Integral of the sampling pdf over the upper hemisphere = P(sampled direction above the
    surface), estimated by uniform hemisphere sampling."""
    n = np.array([0, 0, 1], f32)
    wo = unit([0.6, 0.0, 0.8])
    rough = 0.5
    m = 60000
    z = rng.random(m)
    phi = rng.random(m) * 2 * np.pi
    r = np.sqrt(1 - z * z)
    dirs = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1).astype(f32)
    vals = np.array([ggx_eval(L, (1, 1, 1), rough, n, wo, d)[1] for d in dirs])
    integral = vals.mean() * 2 * np.pi
    above = np.mean([ggx_sample(L, (1, 1, 1), rough, n, wo, rng.random(), rng.random())[0]
                     for _ in range(20000)])
    assert integral == pytest.approx(above, abs=0.03)


def test_dielectric_fresnel_limits(L):
    for eta in (1 / 1.5, 1.5, 1 / 1.33):
        r0 = ((1 - eta) / (1 + eta)) ** 2
        assert L.orc_fresnel_dielectric(1.0, eta) == pytest.approx(r0, rel=1e-5)
        assert L.orc_fresnel_dielectric(1e-4, eta) == pytest.approx(1.0, abs=2e-3)
    # total internal reflection beyond the critical angle (sin > 1/eta when eta > 1)
    eta = 1.5
    crit = np.sqrt(1 - 1 / eta ** 2)
    assert L.orc_fresnel_dielectric(float(crit) - 1e-3, eta) == 1.0
    assert L.orc_fresnel_dielectric(float(crit) + 1e-2, eta) < 1.0


def test_mis_weights_partition_unity(L, rng):
    for _ in range(1000):
        a, b = rng.uniform(0, 10, 2)
        assert L.orc_mis_power(a, b) + L.orc_mis_power(b, a) == pytest.approx(1.0, abs=1e-6)
        assert L.orc_mis_balance(a, b) + L.orc_mis_balance(b, a) == pytest.approx(1.0, abs=1e-6)
    assert L.orc_mis_power(0.0, 0.0) == 0.0 and L.orc_mis_balance(0.0, 0.0) == 0.0


def test_cosine_hemisphere_density(L, rng):
    """Cosine sampling: E[cos theta] = 2/3 and the directions stay in the hemisphere."""
    n = unit([0.3, -0.5, 0.8])
    out = np.zeros(3, f32)
    cs = []
    for _ in range(20000):
        L.orc_cosine_hemisphere(ptr(n), rng.random(), rng.random(), ptr(out))
        cs.append(float(np.dot(out, n)))
    cs = np.array(cs)
    assert cs.min() >= -1e-6
    assert cs.mean() == pytest.approx(2 / 3, abs=0.01)


def implied(prob, alias):
    n = len(prob)
    p = np.array(prob, np.float64) / n
    for j in range(n):
        p[alias[j]] += (1.0 - prob[j]) / n
    return p


def test_light_alias_table_reproduces_area_power(L, orc, scenes):
    data = scenes.sorting_stress(subdiv=2, tex_size=8)
    osc = orc.scene(data)
    cap = 64
    prob, alias, tris = np.zeros(cap, f32), np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    mat_pdf = np.zeros(len(data.materials), f32)
    n = L.orc_light_tables(osc.h, ptr(prob), ptr(alias), ptr(tris), ptr(mat_pdf), cap)
    assert n == 4   # two quads
    p = implied(prob[:n], alias[:n])
    # expected: area * luminance(radiance) per emitter triangle
    w = []
    for m in data.meshes:
        mt = data.materials[m.material_id]
        if mt["type"] != 2:
            continue
        rad = np.array(mt["albedo"]) * mt["emission"]
        lum = 0.2126 * rad[0] + 0.7152 * rad[1] + 0.0722 * rad[2]
        for t in m.indices:
            a, b, c = m.positions[t].astype(np.float64)
            w.append(0.5 * np.linalg.norm(np.cross(b - a, c - a)) * lum)
    w = np.array(w) / np.sum(w)
    assert np.allclose(p, w, rtol=1e-5, atol=1e-7)
    # per-material area pdf: selection probability per unit area = luminance / sum(area*luminance)
    total = 0.0
    lums = {}
    for m in data.meshes:
        mt = data.materials[m.material_id]
        if mt["type"] != 2:
            continue
        rad = np.array(mt["albedo"]) * mt["emission"]
        lums[m.material_id] = 0.2126 * rad[0] + 0.7152 * rad[1] + 0.0722 * rad[2]
        for t in m.indices:
            a, b, c = m.positions[t].astype(np.float64)
            total += 0.5 * np.linalg.norm(np.cross(b - a, c - a)) * lums[m.material_id]
    for mid, lum in lums.items():
        assert mat_pdf[mid] == pytest.approx(lum / total, rel=1e-5)


def test_env_alias_table_reproduces_texel_weights(L, orc, scenes):
    data = scenes.cornell_box()
    env = scenes.hdri_sky(64, 32, peak_over_mean=100.0)
    data.environment = env
    osc = orc.scene(data)
    n = 64 * 32
    prob, alias, pdf, choice = (np.zeros(n, f32), np.zeros(n, np.int32), np.zeros(n, f32),
                                C.c_float())
    assert L.orc_env_tables(osc.h, ptr(prob), ptr(alias), ptr(pdf), C.byref(choice)) == n
    assert choice.value == 0.5       # area lights and an environment: NEE picks each half the time
    rgb = env["rgb"].reshape(32, 64, 3).astype(np.float64)
    lum = 0.2126 * rgb[..., 0] + 0.7152 * rgb[..., 1] + 0.0722 * rgb[..., 2]
    lat = (0.5 - (np.arange(32) + 0.5) / 32) * np.pi
    w = (lum * np.cos(lat)[:, None]).reshape(-1)
    w /= w.sum()
    assert np.allclose(implied(prob, alias), w, rtol=1e-4, atol=1e-9)
    assert np.allclose(pdf, w, rtol=1e-5, atol=1e-12)
