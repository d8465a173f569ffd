"""Progress protocol (src/render/WavefrontRenderer.cpp:297-311, src/main.cpp:25-33) and the
first_sample offset it is built on: consecutive chunks trace exactly the paths of one render."""
import importlib

import numpy as np
import pytest

import oracle_lib

wf = importlib.import_module("wavefront-renderer_amd")
progress = importlib.import_module("wavefront-renderer_amd.progress")


def test_preview_path_kats():
    # derivePreviewPath: the extension replaced, or the suffix appended
    assert progress.preview_path_for("output/cornellBox_256_none.exr") == \
        "output/cornellBox_256_none_preview.png"
    assert progress.preview_path_for("render") == "render_preview.png"


@pytest.fixture(scope="module")
def cornell(ctx, scenes):
    data = scenes.cornell_box()
    sc = wf.Scene(ctx, data)
    yield data, sc
    sc.close()


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["passes", "regenerate"])
def test_chunks_equal_one_render(cornell, schedule):
    data, scene = cornell
    kw = dict(width=40, height=32, max_depth=6, firefly=1e30, seed=4, schedule=schedule,
              sample_lanes=1)
    one, _, st = wf.render(scene, wf.render_params(spp=6, camera=data.camera, **kw))
    a, _, sa = wf.render(scene, wf.render_params(spp=4, camera=data.camera, **kw))
    b, _, sb = wf.render(scene, wf.render_params(spp=2, first_sample=4, camera=data.camera, **kw))
    assert sa.rays_traced + sb.rays_traced == st.rays_traced
    assert sa.shadow_rays + sb.shadow_rays == st.shadow_rays
    mix = (a.astype(np.float64) * 4 + b.astype(np.float64) * 2) / 6
    assert np.allclose(mix, one, rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
def test_first_sample_matches_oracle(cornell, orc):
    data, scene = cornell
    osc = orc.scene(data)
    p = wf.render_params(32, 24, 2, 6, data.camera, firefly=1e30, seed=8, schedule="passes",
                         first_sample=5)
    img, cnt, st = wf.render(scene, p)
    oimg, ocnt, _, ost = osc.render(p)
    assert st.rays_traced == ost.rays_traced and st.shaded_hits == ost.shaded_hits
    rel = np.abs(img - oimg) / np.maximum(np.abs(oimg), 1e-3)
    assert np.mean(rel < 1e-4) > 0.97


@pytest.mark.gpu
def test_render_progressive(cornell, tmp_path):
    data, scene = cornell
    params = wf.render_params(40, 32, 8, 6, data.camera, firefly=1e30, seed=1, sample_lanes=1)
    lines, seen = [], []
    preview = str(tmp_path / "out_preview.png")

    def cb(done, total):
        seen.append((done, total, (tmp_path / "out_preview.png").exists()))

    img, cnt, stats = progress.render_progressive(scene, params, preview, 3, cb, lines.append)
    assert lines == ["Sample: 3/8", "Sample: 6/8", "Sample: 8/8"]
    assert [s[:2] for s in seen] == [(3, 8), (6, 8), (8, 8)] and all(s[2] for s in seen)
    assert not (tmp_path / "out_preview.png").exists()     # removed at the end
    one, _, _ = wf.render(scene, params)
    assert np.allclose(img, one, rtol=1e-5, atol=1e-6)
    assert int(cnt.min()) == 8 and len(stats) == 3
