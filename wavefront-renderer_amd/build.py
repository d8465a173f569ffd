"""Build recipe for libwavefront_hip.so (gfx950) and, for tests, the CPU oracle.

    python wavefront-renderer_amd/build.py            # library + oracle (+ reference harness)
    python wavefront-renderer_amd/build.py --lib-only

Every HIP translation unit is compiled for gfx950 only, with FP contraction off (the key,
intersection and shading arithmetic is compared bit-for-bit against a CPU restatement),
correctly rounded fp32 division/sqrt, and packed-fp32 instructions disabled.
"""
from __future__ import annotations

import argparse
import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB_DIR = PKG / "lib"
LIB = LIB_DIR / "libwavefront_hip.so"
OBJ_DIR = PKG / "_build"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    f"-I{ROOT / 'include'}",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}")
    noisy = [l for l in proc.stderr.splitlines()
             if l.strip() and "packed-fp32-ops" not in l and "not a recognized feature" not in l]
    if noisy:
        sys.stderr.write("\n".join(noisy) + "\n")


def build_library(jobs: int = 8) -> Path:
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError(f"no HIP sources in {CSRC}")
    OBJ_DIR.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)
    objs = [OBJ_DIR / (src.stem + ".o") for src in sources]

    def compile_one(pair):
        src, obj = pair
        _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
        return obj

    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(compile_one, zip(sources, objs)))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> None:
    _run(["make", "-C", str(ROOT / "oracle"), "-s"])
    if Path("/root/reference/include").is_dir():
        _run(["make", "-C", str(ROOT / "oracle"), "-s", "ref"])


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    args = ap.parse_args()
    lib = build_library(args.j)
    print(f"built {lib}")
    if not args.lib_only:
        build_oracle()
        print("built oracle")


if __name__ == "__main__":
    main()
