"""Build recipe for libnlhip.so (gfx950, in-tree).

    python -m nlhip.build          (with neighbourlists.jl_amd on sys.path)
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)                       # neighbourlists.jl_amd/
CSRC = os.path.join(ROOT, "csrc")
REPO = os.path.dirname(ROOT)
INCLUDE = os.path.join(REPO, "include")
OUT = os.path.join(PKG_DIR, "libnlhip.so")

SOURCES = ["api.hip", "kernels.hip", "pairs.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # bit-exact FP64: keep the reference's non-fused evaluation order
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


USER_KERNEL_DIR = os.path.join(REPO, "tests", "user_kernel")
USER_KERNEL_SRC = os.path.join(USER_KERNEL_DIR, "user_kernel.hip")
USER_KERNEL_OUT = os.path.join(USER_KERNEL_DIR, "libnlh_user_kernel.so")


def _compile(srcs, out, verbose):
    tmp = out + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", INCLUDE, "-o", tmp] + srcs
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if noise not in l and "ignoring feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + err)
    if verbose and err.strip():
        print(err)
    os.replace(tmp, out)
    return out


def build(verbose: bool = False) -> str:
    """libnlhip.so, then the test-only user kernel that exercises the
    device-side header include/nlhip_device.h."""
    out = _compile([os.path.join(CSRC, s) for s in SOURCES], OUT, verbose)
    if os.path.exists(USER_KERNEL_SRC):
        _compile([USER_KERNEL_SRC], USER_KERNEL_OUT, verbose)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
