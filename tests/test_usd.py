"""Scenes and cameras in the reference's USD text format (src/scene/UsdSceneLoader.cpp:341-880,
src/scene/CameraLoader.cpp:31-100): round trips of generated scenes, the reference's own Cornell
box scene and camera files read under the reference loader's rules, closed-form checks of the
transform, triangulation, light and fallback rules, and (GPU) identical renders of a scene and
its round trip."""
import importlib
import math
from pathlib import Path

import numpy as np
import pytest

usd = importlib.import_module("wavefront-renderer_amd.usd")
scenes = importlib.import_module("wavefront-renderer_amd.scenes")
abi = importlib.import_module("wavefront-renderer_amd._abi")
REF = Path("/root/reference/scenes")


def _small_stress():
    return scenes.sorting_stress(subdiv=2, n_materials=6, n_textures=2, tex_size=16)


def _texel_data(data, tid):
    t = data.textures[tid]
    return int(t["width"]), int(t["height"]), np.asarray(t["data"]).tobytes()


@pytest.mark.parametrize("make", [scenes.cornell_box, _small_stress])
def test_round_trip(tmp_path, make):
    data = make()
    usd.write_usda(data, tmp_path / "s.usda")
    back = usd.read_usda(tmp_path / "s.usda")
    assert len(back.meshes) == len(data.meshes)
    for a, b in zip(data.meshes, back.meshes):
        assert np.array_equal(np.asarray(a.positions, np.float32), b.positions)
        assert np.array_equal(a.indices, b.indices)
        if a.uv_u is None:
            assert b.uv_u is None
        else:
            assert np.array_equal(np.asarray(a.uv_u, np.float32), b.uv_u)
            assert np.array_equal(np.asarray(a.uv_v, np.float32), b.uv_v)
    for a, b in zip([data.materials[m.material_id] for m in data.meshes],
                    [back.materials[m.material_id] for m in back.meshes]):
        assert a["type"] == b["type"]
        if a["type"] == abi.MAT_EMISSIVE:
            assert np.allclose(np.float32(a["albedo"]) * np.float32(a.get("emission", 1.0)),
                               np.float32(b["albedo"]) * np.float32(b["emission"]), rtol=0)
            continue
        for k in ("roughness", "ior"):
            if k in a and a["type"] != abi.MAT_DIFFUSE:
                assert np.float32(a[k]) == np.float32(b[k])
        assert a.get("model", 0) == b.get("model", 0)
        if a.get("texture_id", -1) >= 0:   # the loader adds one texture per material binding
            assert _texel_data(data, a["texture_id"]) == _texel_data(back, b["texture_id"])
        else:
            assert b.get("texture_id", -1) == -1
            assert np.array_equal(np.float32(a["albedo"]), np.float32(b["albedo"]))


def test_camera_round_trip(tmp_path):
    cam = {"eye": (0.0, 0.35, 3.3), "target": (0.1, 0.0, -0.2), "up": (0.0, 1.0, 0.0),
           "vfov": 42.5}
    usd.write_camera_usda(cam, tmp_path / "c.usda")
    back = usd.read_camera_usda(tmp_path / "c.usda")
    for k in ("eye", "target", "up"):
        assert np.array_equal(np.float32(back[k]), np.float32(cam[k]))
    assert np.float32(back["vfov"]) == np.float32(cam["vfov"])


def test_lights_and_environment_round_trip(tmp_path):
    data = scenes.cornell_box()
    d = np.array([0.3, -1.0, 0.2])
    data.directional = {"direction": tuple(d / np.linalg.norm(d)), "color": (1.0, 0.9, 0.8),
                        "intensity": 2.5}
    rgb = np.random.default_rng(1).uniform(0, 4, (8, 16, 3)).astype(np.float32)
    data.environment = {"width": 16, "height": 8, "rgb": rgb.reshape(-1), "intensity": 1.5,
                        "tint": (1.0, 0.5, 0.25)}
    usd.write_usda(data, tmp_path / "l.usda")
    back = usd.read_usda(tmp_path / "l.usda")
    assert np.allclose(back.directional["direction"], data.directional["direction"], atol=1e-6)
    assert back.directional["intensity"] == 2.5
    assert np.array_equal(back.environment["rgb"].reshape(8, 16, 3), rgb)
    assert back.environment["intensity"] == 1.5 and back.environment["tint"] == (1.0, 0.5, 0.25)


USDA = '''#usda 1.0
(
    upAxis = "Z"
)
def Xform "World"
{
    float3 xformOp:translate = (1, 2, 3)
    float3 xformOp:scale = (2, 2, 2)
    uniform token[] xformOpOrder = ["xformOp:translate", "xformOp:scale"]
    def Mesh "Pent"
    {
        point3f[] points = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (-1, 0.5, 0)]
        int[] faceVertexCounts = [5]
        int[] faceVertexIndices = [0, 1, 2, 3, 4]
        color3f[] primvars:displayColor = [(0.5, 0.25, 1)]
    }
    def RectLight "Panel"
    {
        float inputs:width = 2
        float inputs:height = 1
        float inputs:intensity = 10
        color3f inputs:color = (1, 0.5, 0.25)
        float3 xformOp:rotateXYZ = (90, 0, 0)
        uniform token[] xformOpOrder = ["xformOp:rotateXYZ"]
    }
}
'''


def test_loader_rules_closed_form(tmp_path):
    p = tmp_path / "r.usda"
    p.write_text(USDA)
    s = usd.read_usda(p)
    pent, panel = s.meshes
    # fan from the first vertex (UsdSceneLoader.cpp:58-67)
    assert pent.indices.tolist() == [[0, 1, 2], [0, 2, 3], [0, 3, 4]]
    # scale then translate (row vectors), then Z-up turned to Y-up: (x, y, z) -> (x, z, -y)
    def world(v):
        x, y, z = np.array(v) * 2 + np.array([1, 2, 3])
        return (x, z, -y)
    assert np.allclose(pent.positions[2], world((1, 1, 0)), atol=1e-6)
    # displayColor fallback material (UsdSceneLoader.cpp:852-870)
    m = s.materials[pent.material_id]
    assert m["type"] == abi.MAT_DIFFUSE and m["albedo"] == (0.5, 0.25, 1.0)
    # RectLight: a width x height quad in its local XY plane, emissive colour x intensity
    lm = s.materials[panel.material_id]
    assert lm["type"] == abi.MAT_EMISSIVE and lm["albedo"] == (1.0, 0.5, 0.25)
    assert lm["emission"] == 10.0 and panel.indices.shape == (2, 3)
    ext = panel.positions.max(0) - panel.positions.min(0)
    assert np.allclose(sorted(ext), [0.0, 2.0, 4.0], atol=1e-5)   # x2 scale of the parent


def test_spatial_palette_closed_form(orc):
    pal = usd.spatial_palette(8, 0.25)["data"].reshape(8, 8, 4)
    # cell (0, 0, 0) hashes to 0: every channel 0 + (1 - 0) * 0.25
    assert pal[0, 0, :3].tolist() == [0.25, 0.25, 0.25]
    for x, y in [(1, 0), (3, 5), (7, 7)]:
        h = orc.hash_grid_cell(x, y, 0)
        for c, s in enumerate((16, 8, 0)):
            v = np.float32((h >> s) & 0xFF) / np.float32(255)
            assert pal[y, x, c] == v + (np.float32(1) - v) * np.float32(0.25)


@pytest.mark.skipif(not REF.is_dir(), reason="reference tree absent")
def test_reference_cornell_box():
    s = usd.read_usda(REF / "cornellBox.usda")
    assert len(s.meshes) == 18 and s.triangle_count == 36
    names = [m.name.split("/")[-1] for m in s.meshes]
    assert names[:6] == ["Floor", "Ceiling", "BackWall", "LeftWall", "RightWall", "CeilingLight"]
    types = [m["type"] for m in s.materials]
    assert types == [abi.MAT_DIFFUSE] * 3 + [abi.MAT_EMISSIVE] + [abi.MAT_DIFFUSE] * 2
    light = s.materials[s.meshes[5].material_id]
    assert light["albedo"] == (12.0, 12.0, 12.0) and light["emission"] == 1.0
    assert s.materials[s.meshes[3].material_id]["albedo"] == (0.65, 0.05, 0.05)
    assert s.meshes[0].positions.tolist()[:2] == [[-5, 0, -5], [5, 0, -5]]
    cam = usd.read_camera_usda(REF / "cameras" / "cornellBoxCamera.usda")
    assert cam == {"eye": (0, 5, 15), "target": (0, 5, -2), "up": (0, 1, 0), "vfov": 53.0}


@pytest.mark.gpu
def test_round_trip_renders_identically(tmp_path, ctx):
    wf = importlib.import_module("wavefront-renderer_amd")
    data = _small_stress()
    usd.write_usda(data, tmp_path / "s.usda")
    back = usd.read_usda(tmp_path / "s.usda")
    p = wf.render_params(40, 32, 2, 4, data.camera, firefly=1e30, seed=6,
                         physics=wf.PHYSICS_UNBIASED)
    a = wf.Scene(ctx, data)
    b = wf.Scene(ctx, back)
    ia, _, _ = wf.render(a, p)
    ib, _, _ = wf.render(b, p)
    assert np.array_equal(ia, ib)
    a.close()
    b.close()
