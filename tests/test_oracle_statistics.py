"""The statistical image test itself, on the oracle (CPU): two independent estimates of the same
estimator pass it, and estimates with a planted mistake fail it. This pins the test that
compares device images with the FP64 oracle reference (test_gpu_statistics.py)."""
import numpy as np
import pytest

import stat_image


@pytest.fixture(scope="module")
def cornell_orc(orc, scenes):
    data = scenes.cornell_box()
    return data, orc.scene(data)


def renders(wf, osc, data, seeds, spp, **kw):
    out = []
    for s in seeds:
        img, _, _, _ = osc.render(wf.render_params(camera=data.camera, spp=spp, seed=s, **kw),
                                  threads=8)
        out.append(img)
    return np.stack(out)


@pytest.fixture(scope="module")
def reference(wf, cornell_orc):
    data, osc = cornell_orc
    kw = dict(width=24, height=18, max_depth=8, firefly=1e30)
    _, _, ref64, _ = osc.render(wf.render_params(camera=data.camera, spp=1024, seed=777, **kw),
                                fp64=True, threads=8)
    return kw, ref64


def test_same_estimator_passes(wf, cornell_orc, reference):
    data, osc = cornell_orc
    kw, ref = reference
    st = stat_image.compare(renders(wf, osc, data, range(1, 17), 32, **kw), 32, ref, 1024)
    print(st)
    assert stat_image.passes(st), st


def test_missing_bounces_fail(wf, cornell_orc, reference):
    # planted mistake: paths cut at depth 3 lose the light of the deeper bounces
    data, osc = cornell_orc
    kw, ref = reference
    cut = dict(kw, max_depth=3)
    st = stat_image.compare(renders(wf, osc, data, range(1, 17), 32, **cut), 32, ref, 1024)
    print(st)
    assert not stat_image.passes(st), st


def test_small_gain_error_fails(wf, cornell_orc, reference):
    # planted mistake: a 3% gain error (e.g. a wrong normalisation) over the whole image
    data, osc = cornell_orc
    kw, ref = reference
    x = renders(wf, osc, data, range(1, 17), 32, **kw) * 1.03
    st = stat_image.compare(x, 32, ref, 1024)
    print(st)
    assert not stat_image.passes(st), st
