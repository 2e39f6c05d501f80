"""Build libvcphip.so (all HIP sources, gfx950) in-tree with hipcc.

    python -m vcp_amd.build          # or vcp_amd.build.build()
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvcphip.so")
SOURCES = ["kern_drgemm.hip", "kern_ozaki.hip", "kern_dd.hip", "kern_tune.hip", "kern_lss.hip", "kern_galerkin.hip", "vcphip_host.hip", "vcphip_io.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # every rounding in the kernels is explicit: no FMA contraction
    "-ffp-contract=off",
    # keep hipcc from forming packed FP32 ops on its own
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
# the approximate inverse of the verified solve: rocSOLVER getrf/getrs
LIBS = ["-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocsolver", "-lrocblas"]


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(verbose=False):
    """Compile every HIP source of the library into vcp_amd/libvcphip.so."""
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + sources() + LIBS + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libvcphip.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
