"""Builds the HIP library (gfx950) and the host-side math library in-tree.

    python -m morphompm_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# one shared library: the step kernels and the batched-inversion kernels
HIP_SOURCES = ["mpm_kernels.hip", "mpm_inverse.hip"]
HIP_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
             # multiply-adds fused per source expression only, so a kernel's
             # rounding does not depend on which template instance it is in
             # (a batch of mixed models gives each instance the bits it gets alone)
             "-ffp-contract=on",
             # no packed-FP32 instruction selection
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd[:2])} ... ({r.returncode})")


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    hip_out = os.path.join(OUT, "libmorphompm_hip.so")
    tmp = hip_out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-o", tmp] + [os.path.join(CSRC, f) for f in HIP_SOURCES],
         verbose)
    os.replace(tmp, hip_out)
    host_out = os.path.join(OUT, "libmpm_host.so")
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall", "-o", host_out + ".tmp",
          os.path.join(CSRC, "mpm_host.cpp")], verbose)
    os.replace(host_out + ".tmp", host_out)
    return hip_out, host_out


if __name__ == "__main__":
    for p in build(verbose=True):
        print("built", p)
