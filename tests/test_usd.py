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


@pytest.mark.skipif(not REF.is_dir(), reason="reference tree absent")
def test_reference_material_variants():
    """scenes/materialVariants.usda under UsdSceneLoader.cpp:441-616: emissive colour with
    spotOuterAngle -> SpotLight with both angles; metallic > 0.5 -> Metal (keeping a connected
    UsdUVTexture); plastic -> Plastic; spatialChecker -> the 512^2 palette at the loader's fixed
    contrast 0.25 (the file's spatialCheckerReduceContrast is not read) and the cell size;
    roughness alone leaves a Diffuse. The dragon meshes are references to absent assets."""
    warnings = []
    s = usd.read_usda(REF / "materialVariants.usda", warn=warnings.append)
    by = {m.name.split("/")[-1]: s.materials[m.material_id] for m in s.meshes}
    spot = by["LampDisc"]
    assert spot["type"] == abi.MAT_SPOTLIGHT and spot["albedo"] == (120.0, 120.0, 120.0)
    assert (spot["spot_outer_deg"], spot["spot_falloff_deg"], spot["emission"]) == (120.0, 90.0, 1.0)
    assert by["LampDisc"] is by["LampDisc"] and s.meshes[5].indices.shape[0] > 2
    mirror = by["MirrorLeft"]
    assert mirror["type"] == abi.MAT_METAL and mirror["roughness"] == 0.0
    assert mirror["albedo"] == (1.0, 1.0, 1.0) and mirror["texture_id"] == -1
    floor = by["Floor"]
    assert floor["type"] == abi.MAT_METAL and floor["texture_id"] >= 0
    tex = s.textures[floor["texture_id"]]
    assert tex["format"] == abi.TEX_RGBA8_SRGB and tex["width"] > 1 and tex["height"] > 1
    assert by["AccentLight_A"]["type"] == abi.MAT_EMISSIVE
    assert by["AccentLight_A"]["albedo"] == (6.0, 5.5, 7.5)
    assert by["AccentLightBacking_B"] == {"type": abi.MAT_DIFFUSE, "albedo": (0.02, 0.02, 0.02),
                                          "texture_id": -1, "model": abi.MODEL_REFERENCE}
    # Material prims in traversal order; the four dragon materials have no mesh here
    none, plastic, checker, cb = s.materials[6:10]
    assert none["type"] == abi.MAT_DIFFUSE and none["albedo"] == (0.75, 0.75, 0.75)
    assert plastic["type"] == abi.MAT_PLASTIC and plastic["albedo"] == (0.85, 0.08, 0.08)
    assert plastic["roughness"] == 0.15
    assert checker["type"] == abi.MAT_DIFFUSE and checker["use_spatial_checker"] == 1
    assert checker["checker_cell_size"] == 0.3
    pal = s.textures[checker["texture_id"]]
    assert pal["width"] == pal["height"] == 512 and pal["format"] == abi.TEX_RGB_F32
    assert np.asarray(pal["data"]).reshape(512, 512, 4)[0, 0, :3].tolist() == [0.25] * 3
    assert cb["type"] == abi.MAT_METAL and cb["albedo"] == (0.72, 0.35, 0.15)
    assert cb["roughness"] == 0.25
    assert sum("xyzrgb_dragon" in w for w in warnings) == 4


@pytest.mark.skipif(not REF.is_dir(), reason="reference tree absent")
def test_reference_cornell_box_dragon_composition():
    """scenes/cornellBoxDragon.usda: the Cornell box comes in through a reference with its own
    paths re-rooted; `over` prims deactivate the tall box, re-bind and move the short box and
    lower the light; the teapot asset resolves from the repository root and binds Glass; the
    dragon asset is absent."""
    warnings = []
    s = usd.read_usda(REF / "cornellBoxDragon.usda", warn=warnings.append)
    names = [m.name.split("/")[-1] for m in s.meshes]
    assert not any(n.startswith("TallBox") for n in names)
    assert names[:6] == ["Floor", "Ceiling", "BackWall", "LeftWall", "RightWall", "CeilingLight"]
    mat = lambda n: s.materials[s.meshes[names.index(n)].material_id]
    assert mat("CeilingLight")["type"] == abi.MAT_EMISSIVE
    assert mat("CeilingLight")["albedo"] == (18.0, 18.0, 18.0)
    short = [n for n in names if n.startswith("ShortBox")]
    assert len(short) == 6
    for n in short:
        m = mat(n)
        assert m["type"] == abi.MAT_METAL and m["albedo"] == (0.75, 0.6, 0.15)
        assert m["roughness"] == 0.9
    top = s.meshes[names.index("ShortBox_Top")].positions
    assert np.allclose(top[:, 1], top[0, 1]) and abs(float(top[0, 1]) - 5.13) < 1e-5
    teapot = s.meshes[names.index("Teapot")]
    assert teapot.indices.shape[0] > 10_000 and mat("Teapot")["type"] == abi.MAT_GLASS
    assert abs(float(teapot.positions[:, 1].min()) - 5.13) < 1e-3     # sits on the short box
    checker = next(m for m in s.materials if m.get("use_spatial_checker"))
    assert checker["checker_cell_size"] == 0.5
    assert any("dragon.usda" in w for w in warnings)


def test_composition_rules_closed_form(tmp_path):
    """A reference with an over that deactivates a child, overrides an attribute and re-binds a
    mesh; the referenced layer's own binding targets follow it to its new path."""
    (tmp_path / "sub").mkdir()
    (tmp_path / "sub" / "part.usda").write_text('''#usda 1.0
(
    defaultPrim = "Part"
)
def Xform "Part"
{
    def Scope "Looks"
    {
        def Material "Red"
        {
            token outputs:surface.connect = </Part/Looks/Red/S.outputs:surface>
            def Shader "S"
            {
                uniform token info:id = "UsdPreviewSurface"
                color3f inputs:diffuseColor = (1, 0, 0)
            }
        }
    }
    def Mesh "Keep" (apiSchemas = ["MaterialBindingAPI"])
    {
        point3f[] points = [(0, 0, 0), (1, 0, 0), (0, 1, 0)]
        int[] faceVertexCounts = [3]
        int[] faceVertexIndices = [0, 1, 2]
        rel material:binding = </Part/Looks/Red>
    }
    def Mesh "Drop"
    {
        point3f[] points = [(0, 0, 0), (1, 0, 0), (0, 1, 0)]
        int[] faceVertexCounts = [3]
        int[] faceVertexIndices = [0, 1, 2]
    }
}
''')
    (tmp_path / "top.usda").write_text('''#usda 1.0
def Xform "World"
{
    def "A" (
        prepend references = @sub/part.usda@
    )
    {
        double3 xformOp:translate = (10, 0, 0)
        uniform token[] xformOpOrder = ["xformOp:translate"]
        over "Drop" (active = false) {}
        over "Looks"
        {
            over "Red"
            {
                over "S"
                {
                    color3f inputs:diffuseColor = (0, 0, 1)
                }
            }
        }
    }
    def "B" (
        prepend references = @sub/part.usda@
    )
    {
    }
}
''')
    s = usd.read_usda(tmp_path / "top.usda")
    assert [m.name for m in s.meshes] == ["/World/A/Keep", "/World/B/Keep", "/World/B/Drop"]
    a, b = s.meshes[0], s.meshes[1]
    assert s.materials[a.material_id]["albedo"] == (0.0, 0.0, 1.0)    # A's over wins
    assert s.materials[b.material_id]["albedo"] == (1.0, 0.0, 0.0)    # B keeps the layer's value
    assert a.positions[:, 0].tolist() == [10.0, 11.0, 10.0]
    assert b.positions[:, 0].tolist() == [0.0, 1.0, 0.0]


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


@pytest.mark.gpu
def test_read_back_scene_renders_as_the_oracle(tmp_path, ctx, orc):
    """A scene written to .usda and read back, rendered on the device, against the oracle's
    render of the scene as generated: the same image bits (one sample lane)."""
    wf = importlib.import_module("wavefront-renderer_amd")
    data = scenes.material_variants()
    data.textures[1] = usd.spatial_palette(512, 0.25)   # what the loader builds for a checker
    usd.write_usda(data, tmp_path / "v.usda")
    back = usd.read_usda(tmp_path / "v.usda")
    back.directional, back.environment = data.directional, data.environment
    for physics in (wf.PHYSICS_REFERENCE, wf.PHYSICS_UNBIASED):
        p = wf.render_params(48, 40, 2, 8, data.camera, firefly=1e30, seed=8, physics=physics,
                             sample_lanes=1)
        sc = wf.Scene(ctx, back)
        img, cnt, st = wf.render(sc, p)
        oimg, ocnt, _, ost = orc.scene(data).render(p)
        assert np.array_equal(cnt, ocnt) and st.rays_traced == ost.rays_traced
        assert np.array_equal(img, oimg)
        sc.close()
