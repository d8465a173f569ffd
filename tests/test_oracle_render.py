"""Oracle-level properties of the restated renderer and its BVH (CPU only)."""
import numpy as np
import pytest

import oracle_lib

abi = oracle_lib.abi
f32 = np.float32


@pytest.fixture(scope="module")
def cornell_orc(orc, scenes):
    data = scenes.cornell_box()
    return data, orc.scene(data)


def test_scene_constants(cornell_orc):
    data, osc = cornell_orc
    bmin, bmax, area, ntri = osc.info()
    assert ntri == 5132
    assert bmin.tolist() == [-0.5, 0.0, -0.5] and bmax.tolist() == [0.5, 1.0, 0.5]
    assert abs(area - 0.0625) < 1e-6          # 0.25 x 0.25 ceiling light


def test_bvh_equals_bruteforce(cornell_orc, orc, rng):
    data, osc = cornell_orc
    rays = orc.raygen(data.camera, 64, 64, np.arange(64 * 64), 3, 0, 0)
    n = 3000
    o = (rng.uniform(-0.6, 0.6, (n, 3)) + [0, 0.5, 0]).astype(f32)
    d = rng.normal(size=(n, 3)).astype(f32)
    for k, col in zip(("ox", "oy", "oz", "dx", "dy", "dz"), (*o.T, *d.T)):
        rays[k] = np.concatenate([rays[k], col.astype(f32)])
    a = osc.closest_hit(rays, use_bvh=False)
    b = osc.closest_hit(rays, use_bvh=True)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_render_is_deterministic_and_policy_invariant(cornell_orc, wf):
    data, osc = cornell_orc
    kw = dict(width=32, height=32, spp=2, max_depth=6, camera=data.camera, seed=9, firefly=1e30)
    imgs = []
    for pol in ("none", "material", "texture", "costBenefit"):
        img, cnt, _, st = osc.render(wf.render_params(policy=pol, **kw), threads=4)
        imgs.append(img)
    for im in imgs[1:]:
        assert np.array_equal(im, imgs[0])
    img2, _, _, _ = osc.render(wf.render_params(ray_sort=False, **kw), threads=1)
    assert np.array_equal(img2, imgs[0])
    assert np.isfinite(imgs[0]).all() and (imgs[0] >= 0).all()


def test_render_reference_test_scene_lights_pixels(orc, scenes, wf):
    # WavefrontRendererTest.cpp:68-95: fewer than half the pixels are black
    data = scenes.floor_and_light()
    osc = orc.scene(data)
    img, cnt, _, _ = osc.render(wf.render_params(width=16, height=16, spp=2, max_depth=2,
                                                 camera=data.camera, seed=42), threads=2)
    black = np.all(img == 0, axis=-1).sum()
    assert black < 16 * 16 / 2


def test_fp64_image_close_to_fp32(cornell_orc, wf):
    data, osc = cornell_orc
    img, _, img64, _ = osc.render(wf.render_params(width=24, height=24, spp=4, max_depth=5,
                                                   camera=data.camera, seed=1, firefly=1e30),
                                  fp64=True, threads=4)
    assert np.allclose(img, img64, rtol=1e-5, atol=1e-6)
