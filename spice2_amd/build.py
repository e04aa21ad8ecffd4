"""Builds the native parts of spice2_amd in-tree (no JIT cache, no install).

  spice2_amd/lib/libspice_hip.so        runtime + kernels behind include/spice_hip.h
  spice2_amd/lib/libspice_workloads.so  test/bench networks written against include/spice/
  spice2_amd/lib/brunel_sample          a user program (spice2_amd/samples/brunel.cpp)
  oracle/liboracle.so (+ oracle/_ref/*.so when the reference tree is present), via make

Usage: python -m spice2_amd.build [--jobs N]
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "lib"
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"

ARCH = "gfx950"
# Strict IEEE semantics (bit-exact models), no packed FP32 instructions.
HIPFLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++20", "-fPIC", "-ffp-contract=off",
            "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-result"]
USERFLAGS = ["-I" + str(INCLUDE), "-DSPICE_ASSERT_PRECONDITIONS", "-DSPICE_ASSERT_INVARIANTS"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (Path(c).exists() or c == "hipcc"):
            return c
    raise RuntimeError("hipcc not found")


def run(cmd: list[str], cwd: Path = ROOT) -> None:
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_runtime() -> Path:
    out = LIB / "libspice_hip.so"
    srcs = [CSRC / "runtime.hip", CSRC / "csr_gen.hip", CSRC / "blocked.hip", CSRC / "jump.cpp"]
    run([hipcc(), *HIPFLAGS, "-shared", *map(str, srcs), "-o", str(out)])
    return out


def build_user(src: Path, out: Path, shared: bool) -> Path:
    cmd = [hipcc(), *HIPFLAGS, *USERFLAGS, str(src), "-L" + str(LIB), "-lspice_hip",
           "-Wl,-rpath," + str(LIB), "-Wl,-rpath,$ORIGIN"]
    if shared:
        cmd.insert(1, "-shared")
    run(cmd + ["-o", str(out)])
    return out


def build_oracle() -> None:
    run(["make", "-s", "-C", str(ROOT / "oracle")])


def build(jobs: int = 4) -> None:
    LIB.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        oracle = ex.submit(build_oracle)
        build_runtime()
        tests_bin = ROOT / "tests" / "native" / "bin"
        tests_bin.mkdir(exist_ok=True)
        users = [ex.submit(build_user, PKG / "workloads" / "networks.cpp", LIB / "libspice_workloads.so", True),
                 ex.submit(build_user, PKG / "samples" / "brunel.cpp", LIB / "brunel_sample", False),
                 ex.submit(build_user, ROOT / "tests" / "native" / "api_host_tests.cpp",
                           tests_bin / "api_host_tests", False)]
        for f in users + [oracle]:
            f.result()


if __name__ == "__main__":
    j = 4
    if "--jobs" in sys.argv:
        j = int(sys.argv[sys.argv.index("--jobs") + 1])
    build(j)
    print("built:", ", ".join(p.name for p in sorted(LIB.iterdir())))
