"""Build liblux_hip.so (gfx950 HIP kernels + C ABI) and the CPU oracle.

Run as ``python -m bevy_penumbra_amd.build`` or through
``__graft_entry__.build()``.  The library is built in-tree so that it travels
with the source snapshot.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "liblux_hip.so"

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-pthread",  # the host SAH build spreads meshes over threads
    # IEEE f32 with no contraction: the traversal and merge arithmetic must
    # match a scalar CPU evaluation bit for bit.
    "-ffp-contract=off",
    "-fno-fast-math",
    # Packed FP32 instructions are kept out of all generated code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = Path(rocm) / "bin" / "hipcc"
    return str(cand) if cand.exists() else "hipcc"


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip")) + sorted((PKG / "host").glob("*.cpp"))


def build_hip(verbose: bool = False) -> Path:
    srcs = [str(s) for s in sources()]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIP_FLAGS, "-I", str(ROOT / "include"), "-o", str(tmp), *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = False) -> Path:
    cmd = ["make", "-s", "-C", str(ROOT / "oracle")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return ROOT / "oracle" / "liblux_oracle.so"


def main(argv: list[str] | None = None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    verbose = "-v" in argv
    build_hip(verbose)
    build_oracle(verbose)
    print(f"built {LIB}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
