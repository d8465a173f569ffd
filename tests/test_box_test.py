"""The traversal's box test is conservative against the true ray (CPU only).

`box_test_check.c` restates the box test of `csrc/wf_traverse.h` (`boxes4` of the four-wide walk,
`slab` of the eight-wide walk: one fma per plane with the origin term rounded outward, exit
distance times 1 + 2^-21) and compares it with exact 113-bit arithmetic on the ray o + t d, over
rays and boxes built to graze each other within a few ulps. The device rule must accept every box
the exact interval accepts; the three plausible shortcuts (no exit factor, no outward rounding,
plain (x - o) inv planes) must not, which shows the check can fail."""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("no host C compiler")
    exe = tmp_path_factory.mktemp("boxchk") / "box_test_check"
    subprocess.run(["gcc", "-O2", "-o", str(exe), str(HERE / "box_test_check.c"), "-lm"],
                   check=True, capture_output=True)
    return exe


def run(exe, cases, mode, seed):
    out = subprocess.run([str(exe), str(cases), str(mode), str(seed)], check=True,
                         capture_output=True, text=True, timeout=300).stdout.split()
    n, exact, accepted, misses = map(int, out)
    assert n == cases
    return exact, accepted, misses


@pytest.mark.parametrize("seed", [1, 2])
def test_device_rule_never_drops_a_box(checker, seed):
    exact, accepted, misses = run(checker, 2_000_000, 0, seed)
    assert exact > 200_000          # the grazing construction meets many boxes exactly
    assert misses == 0
    # the extra boxes are the ones within a few ulps of the ray
    assert accepted < 3 * exact


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_shortcuts_drop_boxes(checker, mode):
    _, _, misses = run(checker, 2_000_000, mode, 3)
    assert misses > 100
