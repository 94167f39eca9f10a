"""Build libvoxelproj_hip.so (gfx950) in-tree with hipcc.

Usage: python voxelproj_amd/csrc/build.py   (or voxelproj_amd._lib.build())
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(PKG, "libvoxelproj_hip.so")
SOURCES = ["vp_capi.hip", "vp_backward.hip", "vp_forward.hip", "vp_f64.hip", "vp_shift.hip"]
ARCH = "gfx950"

FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    f"--offload-arch={ARCH}",
    # Packed FP32 (v_pk_*_f32) stays off on this hardware.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden",
    "-Wall",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    srcs = [os.path.join(HERE, s) for s in SOURCES]
    tmp = OUT + ".tmp"
    cmd = [hipcc()] + FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
