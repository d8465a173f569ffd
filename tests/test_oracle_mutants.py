"""The oracle pins have teeth: each plausible mistake below, planted in a copy of the oracle built
in a temporary directory, makes tests/test_oracle_pins.py fail. CPU only; the oracle under
oracle/ is not touched."""
import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]

# name -> (text in oracle/orc_shade.cpp, replacement)
MUTANTS = {
    "nee_mis_arguments_swapped": ("float w = mis_balance(lps, bps);",
                                  "float w = mis_balance(bps, lps);"),
    "nee_light_cosine_dropped": ("float geometry = (cos_s * cos_l) / (dist * dist);",
                                 "float geometry = (cos_s) / (dist * dist);"),
    "roulette_one_bounce_early": ("if (depth + 1 >= sp.rr_min_depth) {",
                                  "if (depth + 2 >= sp.rr_min_depth) {"),
    "roulette_survival_from_incoming_throughput": (
        "float base = std::min(0.95f, std::max({nt.x, nt.y, nt.z}));",
        "float base = std::min(0.95f, std::max({T.x, T.y, T.z}));"),
    "environment_longitude_axes_swapped": ("m_atan2(d.z, d.x)", "m_atan2(d.x, d.z)"),
    "light_triangle_picked_by_area": (
        "int k = static_cast<int>(rng(DIM_LIGHT_TRI) * tc) % tc;",
        "int k = (rng(DIM_LIGHT_TRI) < 0.5f ? 0 : (rng(DIM_LIGHT_TRI) < 0.2f ? 1 : 2)) % tc;"),
    "sky_colours_swapped": ("V3 dark(0.02f, 0.02f, 0.05f), horizon(0.05f, 0.04f, 0.03f);",
                            "V3 horizon(0.02f, 0.02f, 0.05f), dark(0.05f, 0.04f, 0.03f);"),
}


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    if shutil.which("make") is None or shutil.which("g++") is None:
        pytest.skip("no host compiler")
    base = tmp_path_factory.mktemp("mutants")

    def build(name):
        old, new = MUTANTS[name]
        d = base / name / "oracle"
        shutil.copytree(ROOT / "oracle", d, ignore=shutil.ignore_patterns("_ref", "*.so"))
        src = (d / "orc_shade.cpp").read_text()
        assert src.count(old) == 1, name
        (d / "orc_shade.cpp").write_text(src.replace(old, new))
        (base / name / "include").symlink_to(ROOT / "include")
        (base / name / "wavefront-renderer_amd").symlink_to(ROOT / "wavefront-renderer_amd")
        subprocess.run(["make", "-C", str(d), "-s", "liborc.so"], check=True,
                       capture_output=True)
        return name, d / "liborc.so"

    with concurrent.futures.ThreadPoolExecutor(max_workers=8) as ex:
        return dict(ex.map(build, MUTANTS))


@pytest.mark.parametrize("name", list(MUTANTS))
def test_planted_mistake_fails_a_pin(mutant_libs, name):
    env = dict(os.environ, WF_ORACLE_LIB=str(mutant_libs[name]))
    r = subprocess.run([sys.executable, "-m", "pytest", str(ROOT / "tests" / "test_oracle_pins.py"),
                        "-q", "-x", "-p", "no:cacheprovider"], env=env, cwd=ROOT,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 1 and "failed" in r.stdout, (name, r.stdout[-500:])
