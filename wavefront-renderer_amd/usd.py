"""Scenes and cameras in the reference's on-disk format: USD text (.usda).

`write_usda` emits a SceneData as the UsdPreviewSurface subset the reference's loader reads
(src/scene/UsdSceneLoader.cpp:341-880): one Material per scene material with a UsdPreviewSurface
shader (`diffuseColor`, `emissiveColor`, `roughness`, `metallic`, `ior`, the custom inputs
`plastic`, `spatialChecker`, `spatialCheckerCellSize`, `spotOuterAngle`, `spotFalloffAngle`),
diffuse textures as UsdUVTexture shaders over PNG (sRGB RGBA8) or OpenEXR (float) files, Mesh
prims with face-varying `primvars:st`, a DistantLight and a DomeLight (OpenEXR) when the scene
has them. The microfacet model of a material (GGX conductor, exact dielectric) has no input in
the reference; it is kept in a custom `wavefrontModel` token the reference ignores.
`write_camera_usda` writes the camera file of `CameraLoader::load` (src/scene/CameraLoader.cpp:
31-100): a Camera default prim with `custom:eye`, `custom:target`, `custom:up`, `custom:fov`.

`read_usda` / `read_camera_usda` read that subset back with the reference's rules: materials and
meshes in prim traversal order; a material is emissive when its emissive colour sums above 0
(spot light when it has `spotOuterAngle`), else glass when it has `ior`, metal when `metallic`
> 0.5, plastic when `plastic` is set, else diffuse; `spatialChecker` adds the 512x512 hash
palette (TextureGenerator::generateSpatialPalette, include/shading/Texture.h:76-89); faces are
fanned from their first vertex (triangulateFace, UsdSceneLoader.cpp:58-78); vertices go through
the prim's local-to-world transform (translate, scale, rotateX/Y/Z/XYZ, transform ops); RectLight
and DiskLight prims become emissive meshes; `upAxis = "Z"` is turned to Y up; an unbound mesh
gets a diffuse material of its displayColor (default 0.8 grey). Internal references
(`references = @file.usda@`) are composed: the referenced layer's default prim lands under the
referencing prim with its paths re-rooted there, weaker than the referencing prim's own opinions
and `over` prims, and prims with `active = false` are dropped; an asset path is resolved against
the referencing layer's directory, then its parents. Not read: payloads, variants, instancing,
sublayers, Cube/Sphere/Cylinder gprims and sphere/cylinder lights.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Optional

import numpy as np

from . import _abi
from .image_io import read_exr, write_exr, write_png
from .scenes import Mesh, SceneData

# ---- writer ------------------------------------------------------------------------------------
_TYPE_NAMES = {_abi.MAT_DIFFUSE: "Diffuse", _abi.MAT_METAL: "Metal", _abi.MAT_GLASS: "Glass",
               _abi.MAT_PLASTIC: "Plastic", _abi.MAT_EMISSIVE: "Emissive",
               _abi.MAT_SPOTLIGHT: "SpotLight"}


def _f(x) -> str:
    """A float that reads back to the same float32."""
    return "%.9g" % float(np.float32(x))


def _v3(v) -> str:
    return "(" + ", ".join(_f(x) for x in v) + ")"


def _ident(name: str, fallback: str) -> str:
    s = re.sub(r"[^A-Za-z0-9_]", "_", name.strip("/").split("/")[-1] if name else "")
    if not s or not (s[0].isalpha() or s[0] == "_"):
        s = fallback
    return s


def _array(values, fmt, per_line=8) -> str:
    items = [fmt(v) for v in values]
    lines = [", ".join(items[i:i + per_line]) for i in range(0, len(items), per_line)]
    return "[" + ",\n            ".join(lines) + "]"


def write_usda(data: SceneData, path, root: str = "Scene") -> None:
    """Writes data as a .usda file; textures and the environment map go to files beside it
    (textures/<root>_tex<k>.png|exr, <root>_env.exr)."""
    path = Path(path)
    base = path.parent
    (base / "textures").mkdir(parents=True, exist_ok=True)
    root = _ident(root, "Scene")
    out = ['#usda 1.0', '(', f'    defaultPrim = "{root}"', '    upAxis = "Y"',
           f'    doc = "{data.name or root} - written by the MI355X wavefront renderer"', ')', '',
           f'def Xform "{root}"', '{']
    # textures: sRGB RGBA8 as PNG (the reference decodes them through srgbToLinear), float as EXR
    tex_files = []
    for k, t in enumerate(data.textures):
        w, h = int(t["width"]), int(t["height"])
        if int(t["format"]) == _abi.TEX_RGBA8_SRGB:
            rgba = np.asarray(t["data"]).view(np.uint8).reshape(h, w, 4)
            name = f"textures/{root}_tex{k}.png"
            write_png(base / name, rgba[:, :, :3])
        else:
            rgb = np.asarray(t["data"], np.float32).reshape(h, w, -1)[:, :, :3]
            name = f"textures/{root}_tex{k}.exr"
            write_exr(base / name, rgb)
        tex_files.append(name)
    out.append('    def Scope "Materials"')
    out.append('    {')
    mat_names = []
    for i, m in enumerate(data.materials):
        mname = f"M{i}_{_TYPE_NAMES.get(int(m['type']), 'Material')}"
        mat_names.append(mname)
        mpath = f"/{root}/Materials/{mname}"
        typ = int(m["type"])
        alb = m.get("albedo", (1.0, 1.0, 1.0))
        tex = int(m.get("texture_id", -1))
        checker = int(m.get("use_spatial_checker", 0))
        out += [f'        def Material "{mname}"', '        {',
                f'            token outputs:surface.connect = <{mpath}/Shader.outputs:surface>',
                '            def Shader "Shader"', '            {',
                '                uniform token info:id = "UsdPreviewSurface"']
        if typ in (_abi.MAT_EMISSIVE, _abi.MAT_SPOTLIGHT):
            e = float(m.get("emission", 1.0))
            out.append('                color3f inputs:diffuseColor = (1, 1, 1)')
            out.append('                color3f inputs:emissiveColor = '
                       + _v3([np.float32(c) * np.float32(e) for c in alb]))
            if typ == _abi.MAT_SPOTLIGHT:
                out.append('                float inputs:spotOuterAngle = ' + _f(m.get("spot_outer_deg", 0.0)))
                out.append('                float inputs:spotFalloffAngle = ' + _f(m.get("spot_falloff_deg", 0.0)))
        else:
            if tex >= 0 and not checker:
                out.append(f'                color3f inputs:diffuseColor.connect = <{mpath}/Texture.outputs:rgb>')
            else:
                out.append('                color3f inputs:diffuseColor = ' + _v3(alb))
            out.append('                float inputs:roughness = ' + _f(m.get("roughness", 0.0)))
            if typ == _abi.MAT_METAL:
                out.append('                float inputs:metallic = 1')
            if typ == _abi.MAT_GLASS:
                out.append('                float inputs:ior = ' + _f(m.get("ior", 1.5)))
            if typ == _abi.MAT_PLASTIC:
                out.append('                bool inputs:plastic = 1')
            if checker:
                out.append('                bool inputs:spatialChecker = 1')
                out.append('                float inputs:spatialCheckerCellSize = '
                           + _f(m.get("checker_cell_size", 1.0)))
        if int(m.get("model", _abi.MODEL_REFERENCE)) == _abi.MODEL_MICROFACET:
            out.append('                token inputs:wavefrontModel = "microfacet"')
        out += ['                token outputs:surface', '            }']
        if typ not in (_abi.MAT_EMISSIVE, _abi.MAT_SPOTLIGHT) and tex >= 0 and not checker:
            out += ['            def Shader "Texture"', '            {',
                    '                uniform token info:id = "UsdUVTexture"',
                    f'                asset inputs:file = @{tex_files[tex]}@',
                    '                float3 outputs:rgb', '            }']
        out.append('        }')
    out.append('    }')
    used = set()
    for k, mesh in enumerate(data.meshes):
        name = _ident(mesh.name, f"mesh{k}")
        while name in used:
            name = f"{name}_{k}"
        used.add(name)
        pos = np.asarray(mesh.positions, np.float32).reshape(-1, 3)
        idx = np.asarray(mesh.indices, np.int64).reshape(-1, 3)
        out += [f'    def Mesh "{name}" (apiSchemas = ["MaterialBindingAPI"])', '    {',
                '        point3f[] points = ' + _array(pos, _v3, 4),
                '        int[] faceVertexCounts = ' + _array([3] * len(idx), str, 32),
                '        int[] faceVertexIndices = ' + _array(idx.reshape(-1), str, 30)]
        if mesh.uv_u is not None:
            uu = np.asarray(mesh.uv_u, np.float32).reshape(-1)
            vv = np.asarray(mesh.uv_v, np.float32).reshape(-1)
            out.append('        texCoord2f[] primvars:st = '
                       + _array(list(zip(uu, vv)), _v3, 6) + ' (')
            out += ['            interpolation = "faceVarying"', '        )']
        out += ['        uniform token orientation = "rightHanded"',
                f'        rel material:binding = </{root}/Materials/{mat_names[mesh.material_id]}>',
                '    }']
    if data.directional and float(data.directional.get("intensity", 0.0)) > 0.0:
        d = np.asarray(data.directional["direction"], np.float64)
        d = d / np.linalg.norm(d)
        # local -Z onto d: rows of the row-vector matrix are the images of X, Y, Z
        a = np.array([0.0, 1.0, 0.0]) if abs(d[1]) < 0.9 else np.array([1.0, 0.0, 0.0])
        x = np.cross(a, -d)
        x /= np.linalg.norm(x)
        y = np.cross(-d, x)
        rows = [x, y, -d]
        mat = "(" + ", ".join("(" + ", ".join(repr(float(c)) for c in r) + ", 0)" for r in rows) \
            + ", (0, 0, 0, 1))"
        out += ['    def DistantLight "Sun"', '    {',
                '        float inputs:intensity = ' + _f(data.directional["intensity"]),
                '        color3f inputs:color = ' + _v3(data.directional.get("color", (1, 1, 1))),
                f'        matrix4d xformOp:transform = {mat}',
                '        uniform token[] xformOpOrder = ["xformOp:transform"]', '    }']
    if data.environment:
        env = data.environment
        w, h = int(env["width"]), int(env["height"])
        name = f"textures/{root}_env.exr"
        write_exr(base / name, np.asarray(env["rgb"], np.float32).reshape(h, w, 3))
        out += ['    def DomeLight "Environment"', '    {',
                f'        asset inputs:texture:file = @{name}@',
                '        float inputs:intensity = ' + _f(env.get("intensity", 1.0)),
                '        color3f inputs:color = ' + _v3(env.get("tint", (1, 1, 1))), '    }']
    out.append('}')
    path.write_text("\n".join(out) + "\n")


def write_camera_usda(camera: dict, path) -> None:
    """CameraLoader's camera file (src/scene/CameraLoader.cpp:66-69)."""
    e, t = camera["eye"], camera["target"]
    up = camera.get("up", (0.0, 1.0, 0.0))
    Path(path).write_text("\n".join([
        '#usda 1.0', '(', '    defaultPrim = "Camera"', '    upAxis = "Y"', ')', '',
        'def Camera "Camera"', '{',
        f'    float3 xformOp:translate = {_v3(e)}',
        '    uniform token[] xformOpOrder = ["xformOp:translate"]',
        f'    float3 custom:eye = {_v3(e)}',
        f'    float3 custom:target = {_v3(t)}',
        f'    float3 custom:up = {_v3(up)}',
        f'    float custom:fov = {_f(camera["vfov"])}', '}', '']))


# ---- parser ------------------------------------------------------------------------------------
_TOKEN = re.compile(r'''
    (?P<ws>\s+|\#[^\n]*)
  | (?P<str>"""(?:.|\n)*?"""|"(?:\\.|[^"\\])*"|'(?:\\.|[^'\\])*')
  | (?P<asset>@[^@]*@)
  | (?P<path><[^>]*>)
  | (?P<num>[-+]?(?:\d+\.?\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?|inf|nan)(?![A-Za-z_]))
  | (?P<id>[A-Za-z_][A-Za-z0-9_:.]*)
  | (?P<punct>[()\[\]{}=,;])
''', re.VERBOSE)


def _tokens(text: str):
    pos, out = 0, []
    if text.startswith("#usda"):
        pos = text.index("\n")
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise ValueError(f"usda: cannot read {text[pos:pos + 30]!r}")
        kind = m.lastgroup
        if kind != "ws":
            out.append((kind, m.group()))
        pos = m.end()
    return out


@dataclass
class Prim:
    kind: str
    type: str
    name: str
    path: str
    meta: dict = field(default_factory=dict)
    attrs: Dict[str, tuple] = field(default_factory=dict)     # name -> (type, value, meta)
    rels: Dict[str, object] = field(default_factory=dict)
    children: List["Prim"] = field(default_factory=list)
    parent: Optional["Prim"] = None

    def get(self, name, default=None):
        a = self.attrs.get(name)
        return default if a is None else a[1]


class _Parser:
    def __init__(self, text):
        self.t = _tokens(text)
        self.i = 0

    def peek(self, k=0):
        j = self.i + k
        return self.t[j] if j < len(self.t) else (None, None)

    def take(self, value=None):
        tok = self.t[self.i]
        if value is not None and tok[1] != value:
            raise ValueError(f"usda: expected {value!r}, found {tok[1]!r}")
        self.i += 1
        return tok

    def value(self):
        kind, v = self.peek()
        if v == "(":
            self.take("(")
            items = []
            while self.peek()[1] != ")":
                items.append(self.value())
                if self.peek()[1] == ",":
                    self.take(",")
            self.take(")")
            return tuple(items)
        if v == "[":
            self.take("[")
            items = []
            while self.peek()[1] != "]":
                items.append(self.value())
                if self.peek()[1] == ",":
                    self.take(",")
            self.take("]")
            return items
        self.take()
        if kind == "num":
            return float(v) if any(c in v for c in ".eEn") else int(v)
        if kind == "str":
            return v[3:-3] if v.startswith('"""') else v[1:-1]
        if kind == "asset":
            return ("asset", v[1:-1])
        if kind == "path":
            return ("path", v[1:-1])
        if v in ("true", "false"):
            return v == "true"
        return v   # token / None

    def metadata(self):
        meta = {}
        self.take("(")
        while self.peek()[1] != ")":
            kind, v = self.peek()
            if kind == "str":
                self.take()
                meta.setdefault("doc", v.strip('"'))
                continue
            if v in ("prepend", "append", "add", "delete", "reorder"):
                self.take()
                continue
            key = self.take()[1]
            if self.peek()[1] == "=":
                self.take("=")
                meta[key] = self.value()
            if self.peek()[1] in (",", ";"):
                self.take()
        self.take(")")
        return meta

    def prim(self, parent):
        kind = self.take()[1]                       # def / over / class
        typ = ""
        if self.peek()[0] == "id":
            typ = self.take()[1]
        name = self.take()[1].strip('"')
        p = Prim(kind, typ, name, (parent.path if parent.path != "/" else "") + "/" + name,
                 parent=parent)
        if self.peek()[1] == "(":
            p.meta = self.metadata()
        self.take("{")
        self.body(p)
        self.take("}")
        return p

    def body(self, p):
        while self.peek()[1] not in ("}", None):
            kind, v = self.peek()
            if v in ("def", "over", "class"):
                p.children.append(self.prim(p))
                continue
            if v == "variantSet":
                self.skip_block()
                continue
            if v == "rel":
                self.take()
                name = self.take()[1]
                val = None
                if self.peek()[1] == "=":
                    self.take("=")
                    val = self.value()
                if self.peek()[1] == "(":
                    self.metadata()
                p.rels[name] = val
                continue
            # attribute: [custom] [uniform|varying] type[[]] name [= value] [(meta)]
            while self.peek()[1] in ("custom", "uniform", "varying", "config"):
                self.take()
            typ = self.take()[1]
            if self.peek()[1] == "[":
                self.take("[")
                self.take("]")
                typ += "[]"
            name = self.take()[1]
            val = None
            if self.peek()[1] == "=":
                self.take("=")
                val = self.value()
            meta = self.metadata() if self.peek()[1] == "(" else {}
            p.attrs[name] = (typ, val, meta)

    def skip_block(self):
        while self.peek()[1] != "{":
            self.take()
        depth = 0
        while True:
            v = self.take()[1]
            depth += v == "{"
            depth -= v == "}"
            if depth == 0:
                return

    def stage(self):
        root = Prim("root", "", "", "/")
        if self.peek()[1] == "(":
            root.meta = self.metadata()
        while self.peek()[1] is not None:
            root.children.append(self.prim(root))
        return root


def parse_usda(text: str) -> Prim:
    """The prim tree of a .usda text (stage metadata on the root)."""
    return _Parser(text).stage()


# ---- composition: internal references and overs ---------------------------------------------------
def _resolve_asset(layer_dir: Path, asset: str) -> Optional[Path]:
    """An asset path relative to the referencing layer's directory, else to one of its parents (the
    reference's scenes name some assets from the repository root, e.g. @scenes/objects/x.usda@)."""
    for base in (layer_dir, *layer_dir.parents):
        f = base / asset
        if f.is_file():
            return f
    return None


def _fix_path(v, old: str, new: str):
    if isinstance(v, tuple) and len(v) == 2 and v[0] == "path":
        t = v[1]
        if t == old or t.startswith(old + "/") or t.startswith(old + "."):
            return ("path", new + t[len(old):])
        return v
    if isinstance(v, list):
        return [_fix_path(x, old, new) for x in v]
    return v


def _reroot(p: Prim, old: str, new: str, parent: Optional[Prim], name: str) -> Prim:
    """A copy of p's subtree moved from namespace path `old` to `new` (targets inside follow)."""
    q = Prim(p.kind, p.type, name, new + p.path[len(old):], dict(p.meta),
             {k: (t, _fix_path(v, old, new), m) for k, (t, v, m) in p.attrs.items()},
             {k: _fix_path(v, old, new) for k, v in p.rels.items()}, [], parent)
    q.children = [_reroot(c, old, new, q, c.name) for c in p.children]
    return q


def _merge(strong: Prim, weak: Prim) -> None:
    """strong's opinions over weak's, in place: attributes, relationships and metadata of strong win;
    children are matched by name (weak's order, then strong's new children)."""
    strong.attrs = {**weak.attrs, **strong.attrs}
    strong.rels = {**weak.rels, **strong.rels}
    strong.meta = {**weak.meta, **strong.meta}
    if not strong.type:
        strong.type = weak.type
    if strong.kind == "over" and weak.kind == "def":
        strong.kind = "def"
    local = {c.name: c for c in strong.children}
    merged = []
    for wc in weak.children:
        sc = local.pop(wc.name, None)
        if sc is None:
            wc.parent = strong
            merged.append(wc)
        else:
            _merge(sc, wc)
            merged.append(sc)
    merged += [c for c in strong.children if c.name in local]
    strong.children = merged


def _compose(p: Prim, layer_dir: Path, warn, depth: int) -> None:
    """Composes internal references (`references = @file@`: the file's default prim, or its first
    root prim, under p, weaker than p's own opinions and overs) and drops inactive prims."""
    ref = p.meta.pop("references", None)
    for item in (ref if isinstance(ref, list) else [ref] if ref is not None else []):
        if not (isinstance(item, tuple) and item[0] == "asset"):
            continue
        f = _resolve_asset(layer_dir, item[1])
        if f is None or depth > 8:
            if warn:
                warn(f"{p.path}: reference {item[1]} not found")
            continue
        sub = parse_usda(f.read_text())
        for c in sub.children:
            _compose(c, f.parent, warn, depth + 1)
        name = sub.meta.get("defaultPrim")
        target = next((c for c in sub.children if c.name == name), None) if name else None
        target = target or (sub.children[0] if sub.children else None)
        if target is None:
            continue
        _merge(p, _reroot(target, target.path, p.path, p.parent, p.name))
    for c in p.children:
        _compose(c, layer_dir, warn, depth)
    p.children = [c for c in p.children if c.meta.get("active", True) is not False]


def _traverse(p: Prim):
    for c in p.children:
        if c.kind == "def":
            yield c
            yield from _traverse(c)


def _find(root: Prim, path: str) -> Optional[Prim]:
    node = root
    for part in path.strip("/").split("/"):
        node = next((c for c in node.children if c.name == part), None)
        if node is None:
            return None
    return node


# ---- transforms (row-vector convention: p' = p M) ------------------------------------------------
def _rot(axis: int, deg: float) -> np.ndarray:
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    m = np.eye(4)
    i, j = [(1, 2), (2, 0), (0, 1)][axis]
    m[i, i], m[i, j], m[j, i], m[j, j] = c, s, -s, c
    return m


def _local(p: Prim) -> np.ndarray:
    order = p.get("xformOpOrder") or []
    m = np.eye(4)
    for op in order:            # ops are listed outermost first: p' = p (op_n ... op_1)
        v = p.get(op)
        kind = op.split(":")[1]
        t = np.eye(4)
        if kind == "translate":
            t[3, :3] = v
        elif kind == "scale":
            t[0, 0], t[1, 1], t[2, 2] = v
        elif kind in ("rotateX", "rotateY", "rotateZ"):
            t = _rot("XYZ".index(kind[-1]), float(v))
        elif kind == "rotateXYZ":
            t = _rot(0, v[0]) @ _rot(1, v[1]) @ _rot(2, v[2])
        elif kind == "transform":
            t = np.array(v, np.float64)
        else:
            raise ValueError(f"usda: transform op {op} not supported")
        m = t @ m
    return m


def _world(p: Prim) -> np.ndarray:
    m = np.eye(4)
    node = p
    while node is not None and node.kind != "root":
        m = m @ _local(node)
        node = node.parent
    return m


# ---- scene extraction ----------------------------------------------------------------------------
def spatial_palette(size: int = 512, reduce_contrast: float = 0.25) -> dict:
    """TextureGenerator::generateSpatialPalette (include/shading/Texture.h:76-89) with
    hashGridCell / colorFromHash (include/math/SpatialHash.h:18-37): texel (x, y) is the colour of
    the hash of cell (x, y, 0), float RGB."""
    y, x = np.mgrid[0:size, 0:size].astype(np.uint32)
    h = (x * np.uint32(73856093)) ^ (y * np.uint32(19349663))   # cell z = 0
    h = (h * np.uint32(2654435761)).astype(np.uint32)
    rc = np.float32(reduce_contrast)
    one = np.float32(1.0)
    rgb = np.stack([((h >> s) & 0xFF).astype(np.float32) / np.float32(255.0)
                    for s in (16, 8, 0)], -1)
    rgb = rgb + (one - rgb) * rc
    data = np.zeros((size, size, 4), np.float32)
    data[:, :, :3] = rgb
    return {"width": size, "height": size, "format": _abi.TEX_RGB_F32,
            "data": data.reshape(-1)}


def _load_texture(file: Path) -> Optional[dict]:
    if not file.exists():
        return None
    if file.suffix.lower() == ".exr":
        rgb = read_exr(file)[:, :, :3]
        h, w = rgb.shape[:2]
        data = np.zeros((h, w, 4), np.float32)
        data[:, :, :3] = rgb
        return {"width": w, "height": h, "format": _abi.TEX_RGB_F32, "data": data.reshape(-1)}
    from PIL import Image
    img = np.asarray(Image.open(file).convert("RGBA"), np.uint8)
    h, w = img.shape[:2]
    return {"width": w, "height": h, "format": _abi.TEX_RGBA8_SRGB,
            "data": np.ascontiguousarray(img).reshape(-1).view(np.uint32).copy()}


def _shader_of(root: Prim, target) -> Optional[Prim]:
    if not isinstance(target, tuple) or target[0] != "path":
        return None
    return _find(root, target[1].split(".")[0])


def read_usda(path, warn=None) -> SceneData:
    """The scene of a .usda file under the reference loader's rules (module docstring)."""
    path = Path(path)
    root = parse_usda(path.read_text())
    for c in root.children:
        _compose(c, path.parent, warn, 0)
    root.children = [c for c in root.children if c.meta.get("active", True) is not False]
    z_up = root.meta.get("upAxis") == "Z"
    fix = _rot(0, -90.0) if z_up else np.eye(4)
    materials, textures, mat_index = [], [], {}
    directional = environment = None
    for p in _traverse(root):
        if p.type == "DistantLight" and directional is None:
            d = np.array([0.0, 0.0, -1.0, 0.0]) @ (_world(p) @ fix)
            d = d[:3].astype(np.float32)
            d = d / np.float32(np.sqrt(np.float32(d @ d)))
            directional = {"direction": tuple(d.tolist()),
                           "color": tuple(p.get("inputs:color", (1, 1, 1))),
                           "intensity": float(p.get("inputs:intensity", 1.0))}
        if p.type == "DomeLight" and environment is None:
            f = p.get("inputs:texture:file")
            if f:
                tex = path.parent / f[1]
                if tex.exists() and tex.suffix.lower() == ".exr":
                    rgb = read_exr(tex)[:, :, :3]
                    environment = {"width": rgb.shape[1], "height": rgb.shape[0],
                                   "rgb": rgb.reshape(-1),
                                   "intensity": float(p.get("inputs:intensity", 1.0)),
                                   "tint": tuple(p.get("inputs:color", (1, 1, 1)))}
    for p in _traverse(root):
        if p.type != "Material":
            continue
        diffuse, rough, emission = (0.8, 0.8, 0.8), 0.0, (0.0, 0.0, 0.0)
        tex_id, emissive, spot, glass, metal, plastic, checker = -1, False, False, False, False, False, False
        outer = falloff = 180.0
        ior, cell, model = 1.5, 1.0, _abi.MODEL_REFERENCE
        for name, (typ, val, _) in p.attrs.items():
            if not name.startswith("outputs:") or not name.endswith(".connect"):
                continue
            sh = _shader_of(root, val)
            if sh is None or sh.get("info:id") != "UsdPreviewSurface":
                continue
            if "inputs:diffuseColor" in sh.attrs:
                diffuse = tuple(sh.get("inputs:diffuseColor"))
            src = sh.get("inputs:diffuseColor.connect")
            tsh = _shader_of(root, src)
            if tsh is not None and tsh.get("info:id") == "UsdUVTexture" and tsh.get("inputs:file"):
                t = _load_texture(path.parent / tsh.get("inputs:file")[1])
                if t is not None:
                    textures.append(t)
                    tex_id = len(textures) - 1
                elif warn:
                    warn(f"texture not found: {tsh.get('inputs:file')[1]}")
            rough = float(sh.get("inputs:roughness", rough))
            if "inputs:emissiveColor" in sh.attrs:
                emission = tuple(sh.get("inputs:emissiveColor"))
                emissive = sum(np.float32(c) for c in emission) > 0.0
            if "inputs:spotOuterAngle" in sh.attrs:
                spot, outer = True, float(sh.get("inputs:spotOuterAngle"))
            falloff = float(sh.get("inputs:spotFalloffAngle", falloff))
            if "inputs:ior" in sh.attrs:
                glass, ior = True, float(sh.get("inputs:ior"))
            if "inputs:metallic" in sh.attrs:
                metal = float(sh.get("inputs:metallic")) > 0.5
            plastic = bool(sh.get("inputs:plastic", plastic))
            checker = bool(sh.get("inputs:spatialChecker", checker))
            cell = float(sh.get("inputs:spatialCheckerCellSize", cell))
            if sh.get("inputs:wavefrontModel") == "microfacet":
                model = _abi.MODEL_MICROFACET
        if emissive:
            m = {"type": _abi.MAT_SPOTLIGHT if spot else _abi.MAT_EMISSIVE, "albedo": emission,
                 "emission": 1.0, "spot_outer_deg": outer, "spot_falloff_deg": falloff}
        elif glass:
            m = {"type": _abi.MAT_GLASS, "albedo": (1.0, 1.0, 1.0), "ior": ior}
        elif metal:
            m = {"type": _abi.MAT_METAL, "albedo": diffuse, "roughness": rough, "texture_id": tex_id}
        elif plastic:
            m = {"type": _abi.MAT_PLASTIC, "albedo": diffuse, "roughness": rough,
                 "texture_id": tex_id}
        else:
            m = {"type": _abi.MAT_DIFFUSE, "albedo": diffuse, "texture_id": tex_id}
        m["model"] = model
        if checker:
            textures.append(spatial_palette(512, 0.25))
            m["use_spatial_checker"] = 1
            m["checker_cell_size"] = cell
            m["texture_id"] = len(textures) - 1
        mat_index[p.path] = len(materials)
        materials.append(m)
    if not materials:
        materials.append({"type": _abi.MAT_DIFFUSE, "albedo": (0.8, 0.8, 0.8)})
        mat_index["__default__"] = 0
    meshes = []
    for p in _traverse(root):
        uv = None
        if p.type == "Mesh":
            pts, counts, ids = p.get("points"), p.get("faceVertexCounts"), p.get("faceVertexIndices")
            if not pts or not counts or not ids:
                if warn:
                    warn(f"{p.path}: no points or faces, skipped")
                continue
            st = p.get("primvars:st")
            if st is not None and len(st) == len(ids):
                uv = np.asarray(st, np.float32)
        elif p.type == "RectLight":
            hw = np.float32(p.get("inputs:width", 1.0)) * np.float32(0.5)
            hh = np.float32(p.get("inputs:height", 1.0)) * np.float32(0.5)
            pts = [(-hw, -hh, 0.0), (hw, -hh, 0.0), (hw, hh, 0.0), (-hw, hh, 0.0)]
            counts, ids = [4], [0, 1, 2, 3]
        elif p.type == "DiskLight":
            r, seg = np.float32(p.get("inputs:radius", 0.5)), 24
            ang = [np.float32(i) / np.float32(seg) * np.float32(2.0) * np.float32(math.pi)
                   for i in range(seg)]
            pts = [(r * np.cos(a), r * np.sin(a), 0.0) for a in ang] + [(0.0, 0.0, 0.0)]
            counts = [3] * seg
            ids = [x for i in range(seg) for x in (seg, i, (i + 1) % seg)]
        else:
            if p.type in ("Cube", "Sphere", "Cylinder", "SphereLight", "CylinderLight") and warn:
                warn(f"{p.path}: {p.type} prims are not read")
            continue
        w = _world(p) @ fix
        P = np.asarray(pts, np.float64).reshape(-1, 3)
        Pw = (np.concatenate([P, np.ones((len(P), 1))], 1) @ w)[:, :3].astype(np.float32)
        tri, tuv = [], []
        k = 0
        for c in counts:
            for j in range(1, c - 1):
                tri.append((ids[k], ids[k + j], ids[k + j + 1]))
                if uv is not None:
                    tuv.append((uv[k], uv[k + j], uv[k + j + 1]))
            k += c
        if p.type in ("RectLight", "DiskLight"):
            color = tuple(p.get("inputs:color", (1.0, 1.0, 1.0)))
            materials.append({"type": _abi.MAT_EMISSIVE, "albedo": color,
                              "emission": float(p.get("inputs:intensity", 1.0))})
            mid = len(materials) - 1
        else:
            mid = _binding(root, p, mat_index, materials)
        mesh = Mesh(Pw, np.asarray(tri, np.int32).reshape(-1, 3), mid, name=p.path)
        if uv is not None:
            a = np.asarray(tuv, np.float32).reshape(-1, 3, 2)
            mesh.uv_u, mesh.uv_v = a[:, :, 0].copy(), a[:, :, 1].copy()
        meshes.append(mesh)
    return SceneData(meshes, materials, textures, directional=directional,
                     environment=environment, name=path.stem)


def _binding(root, p, mat_index, materials) -> int:
    node = p
    while node is not None and node.kind != "root":   # ComputeBoundMaterial: nearest binding
        b = node.rels.get("material:binding")
        if b is not None:
            target = b[0] if isinstance(b, list) and b else b
            if isinstance(target, tuple) and target[1] in mat_index:
                return mat_index[target[1]]
            return 0
        node = node.parent
    dc = p.get("primvars:displayColor")
    color = tuple(dc[0]) if dc else (0.8, 0.8, 0.8)
    key = "displayColor_%d_%d_%d" % tuple(int(np.float32(c) * np.float32(255)) for c in color)
    if key not in mat_index:
        materials.append({"type": _abi.MAT_DIFFUSE, "albedo": color})
        mat_index[key] = len(materials) - 1
    return mat_index[key]


def read_camera_usda(path) -> dict:
    """CameraLoader::load (src/scene/CameraLoader.cpp:31-100): the default prim's custom
    attributes, with the loader's fallbacks."""
    root = parse_usda(Path(path).read_text())
    name = root.meta.get("defaultPrim")
    prim = next((c for c in root.children if c.name == name), None) if name else None
    get = (lambda k, d: prim.get(k, d)) if prim else (lambda k, d: d)
    return {"eye": tuple(get("custom:eye", (0, 5, 15))),
            "target": tuple(get("custom:target", (0, 5, -2))),
            "up": tuple(get("custom:up", (0, 1, 0))), "vfov": float(get("custom:fov", 53.0))}
