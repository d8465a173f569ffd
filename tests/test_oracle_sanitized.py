"""The oracle built with AddressSanitizer and UndefinedBehaviorSanitizer (`make -C oracle san`) runs
the oracle's own tests without a report: no out-of-bounds access, use after free or undefined
behaviour in the restatement (CPU only)."""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
FILES = ["tests/test_oracle_kats.py", "tests/test_oracle_pins.py", "tests/test_fmath.py",
         "tests/test_oracle_render.py", "tests/test_oracle_unbiased.py"]


def test_oracle_under_sanitizers(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no host compiler")
    san = tmp_path / "san"
    subprocess.run(["make", "-C", str(ROOT / "oracle"), "-s", "san", f"SAN_DIR={san}"],
                   check=True, capture_output=True)
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True,
                          text=True).stdout.strip()
    ubsan = subprocess.run(["g++", "-print-file-name=libubsan.so"], capture_output=True,
                           text=True).stdout.strip()
    if not Path(asan).exists():
        pytest.skip("no ASan runtime")
    syms = subprocess.run(["nm", "-D", str(san / "liborc.so")], capture_output=True,
                          text=True).stdout
    assert "__asan_" in syms and "__ubsan_" in syms       # instrumented
    env = dict(os.environ, WF_ORACLE_LIB=str(san / "liborc.so"),
               LD_PRELOAD=" ".join(p for p in (asan, ubsan) if Path(p).exists()),
               ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    probe = subprocess.run([sys.executable, "-c", "import sys; sys.path.insert(0, 'tests'); "
                            "import oracle_lib; print(oracle_lib.Oracle().lib._name)"],
                           env=env, cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert probe.stdout.strip() == str(san / "liborc.so"), probe.stdout + probe.stderr
    r = subprocess.run([sys.executable, "-m", "pytest", *FILES, "-q", "-x", "-p",
                        "no:cacheprovider", "-m", "not gpu"], env=env, cwd=ROOT,
                       capture_output=True, text=True, timeout=1200)
    out = r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out[-3000:]
    assert r.returncode == 0, out[-3000:]
