"""Device images against a high-spp FP64-accumulated oracle image, statistically: relMSE next to
the value noise alone explains, and a per-pixel 3-sigma test (tests/stat_image.py; the test's
power is pinned on the CPU in test_oracle_statistics.py). The device estimate is the mean of 32
independent renders (seeds 1..32) of the default regenerating schedule (64 sample lanes);
the oracle image uses another seed."""
import numpy as np
import pytest

import stat_image

pytestmark = pytest.mark.gpu

CASES = {
    # scene, physics, firefly threshold, oracle samples, policy
    "cornell_reference": ("cornell", 0, 10.0, 4096, "none"),
    "stress_unbiased": ("stress", 1, 1e30, 2048, "none"),
    # an HDRI environment (alias-table NEE, MIS against BSDF hits of the sky) beside the area
    # lights, cost-benefit shading order
    "stress_env_costbenefit": ("stress_env", 1, 1e30, 2048, "costBenefit"),
}


@pytest.mark.parametrize("case", list(CASES))
def test_image_statistics_vs_fp64_oracle(wf, ctx, scenes, orc, case):
    name, physics, firefly, ref_spp, policy = CASES[case]
    if name == "cornell":
        data = scenes.cornell_box()
    else:
        data = scenes.sorting_stress(subdiv=5, tex_size=64)
        if name == "stress_env":
            data.environment = scenes.hdri_sky(64, 32, peak_over_mean=20.0)
    scene = wf.Scene(ctx, data)
    osc = orc.scene(data)
    kw = dict(width=24, height=18, max_depth=12, camera=data.camera, firefly=firefly,
              physics=physics, policy=policy)
    _, _, ref, _ = osc.render(wf.render_params(spp=ref_spp, seed=777, **kw), fp64=True,
                              threads=16)
    spp = 64
    imgs = np.stack([wf.render(scene, wf.render_params(spp=spp, seed=s, **kw))[0]
                     for s in range(1, 33)])
    st = stat_image.compare(imgs, spp, ref, ref_spp)
    print(case, st)
    assert np.isfinite(imgs).all()
    assert stat_image.passes(st), st
    scene.close()
