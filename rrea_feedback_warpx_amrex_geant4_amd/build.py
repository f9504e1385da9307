"""Build the HIP library in-tree for gfx950 (no JIT, no torch extension).

    python -m rrea_feedback_warpx_amrex_geant4_amd.build
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "librrea_hip.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("rrea_transport.hip", "rrea_maxwell.hip", "rrea_population.hip")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # every product of the step must round like the reference: no contraction
    "-ffp-contract=off",
    # packed FP32 instructions stay off
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # no two-address FP64 fmac: its addend must sit in the destination VGPR,
    # so every Horner step of the inlined libm polynomials paid two VALU
    # moves for its constant; the three-address fma reads it from SGPRs
    # (the same fused operation, rounded once: results do not change)
    "-Xclang", "-target-feature", "-Xclang", "-fmacf64-inst",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"),
]


# A compiled C++ caller of the C-ABI (tests/cpp/rrea_caller.cpp), linked
# against the in-tree library by a relative run path.
CALLER_SRC = os.path.join(ROOT, "tests", "cpp", "rrea_caller.cpp")
CALLER = os.path.join(ROOT, "tests", "cpp", "rrea_caller")


def build_caller(verbose: bool = False) -> str:
    cmd = [HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-Wno-unused-command-line-argument",
           "-I" + os.path.join(ROOT, "include"), "-o", CALLER + ".tmp", CALLER_SRC, "-L" + HERE, "-lrrea_hip",
           "-Wl,-rpath,$ORIGIN/../../rrea_feedback_warpx_amrex_geant4_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building the C++ caller")
    os.replace(CALLER + ".tmp", CALLER)
    return CALLER


def build(verbose: bool = False) -> str:
    cmd = [HIPCC, *FLAGS, "-o", LIB + ".tmp", *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building librrea_hip.so")
    os.replace(LIB + ".tmp", LIB)
    build_caller(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
