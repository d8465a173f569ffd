"""The C-ABI library loads without a GPU, exports every function include/wavefront_hip.h declares,
and the ctypes mirrors have the C layout (checked against the header compiled by gcc)."""
import ctypes as C
import subprocess
import textwrap
from pathlib import Path

import pytest

import oracle_lib

abi = oracle_lib.abi
ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_header_function():
    lib = abi.load_library()
    names = abi.header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_prototype_table_covers_header():
    assert sorted(abi._PROTOS) == abi.header_functions()


def test_version_and_error_without_device():
    lib = abi.load_library()
    assert b"gfx950" in lib.wf_version()
    h = C.c_void_p()
    st = lib.wf_context_create(0, C.byref(h))
    if st != 0:   # no GPU here: a clean error, not a crash
        assert st == abi.WF_ERR_HIP
        assert lib.wf_last_error()


@pytest.mark.parametrize("name", ["Material", "TextureDesc", "MeshDesc", "SceneDesc", "RaySoA",
                                  "ShadeSoA", "HitSoA", "BuildStats", "ShadeParams",
                                  "RenderParams", "RenderStats", "Camera", "EngineConfig",
                                  "JobDesc", "JobResult", "EngineStats", "Tuning"])
def test_struct_layout_matches_header(tmp_path, name):
    cname = {"Material": "wf_material", "TextureDesc": "wf_texture_desc",
             "MeshDesc": "wf_mesh_desc", "SceneDesc": "wf_scene_desc", "RaySoA": "wf_ray_soa",
             "ShadeSoA": "wf_shade_soa", "HitSoA": "wf_hit_soa", "BuildStats": "wf_build_stats",
             "ShadeParams": "wf_shade_params", "RenderParams": "wf_render_params",
             "RenderStats": "wf_render_stats", "Camera": "wf_camera",
             "EngineConfig": "wf_engine_config", "JobDesc": "wf_job_desc",
             "JobResult": "wf_job_result", "EngineStats": "wf_engine_stats",
             "Tuning": "wf_tuning"}[name]
    cls = getattr(abi, name)
    fields = [f[0] for f in cls._fields_]
    nl = "\\n"
    body = "".join(f'printf("%zu{nl}", offsetof({cname}, {f}));' for f in fields)
    src = ("#include <stdio.h>\n#include <stddef.h>\n#include \"wavefront_hip.h\"\n"
           f'int main(void) {{ printf("%zu{nl}", sizeof({cname})); {body} return 0; }}\n')
    c = tmp_path / "layout.c"
    c.write_text(src)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", f"-I{ROOT / 'include'}", str(c), "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert out[0] == C.sizeof(cls)
    assert out[1:] == [getattr(cls, f).offset for f in fields]


def test_header_compiles_as_c_and_cpp(tmp_path):
    for compiler, ext in (("gcc", "c"), ("g++", "cpp")):
        f = tmp_path / f"h.{ext}"
        f.write_text('#include "wavefront_hip.h"\nint main(void){return 0;}\n')
        subprocess.run([compiler, "-Wall", "-Werror", f"-I{ROOT / 'include'}", "-fsyntax-only",
                        str(f)], check=True)
