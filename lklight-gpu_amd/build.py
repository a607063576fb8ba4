"""Build the gfx950 shared library of the pose-scoring path.

    python lklight-gpu_amd/build.py

Compiles csrc/*.hip (device + host) and csrc/*.cpp (host only) with hipcc for
gfx950 into lklight/liblklight_gpu.so.  Packed-FP32 instruction selection is
disabled for all device code.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "lklight", "liblklight_gpu.so")

SOURCES = ["lkl_score.hip", "lkl_field.hip", "lkl_sasa.hip", "lkl_surrogate.hip", "lkl_gsod.hip", "lkl_synth.cpp", "lkl_gso.cpp"]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: cannot build the gfx950 library")


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-fvisibility=hidden", "-o", tmp] + srcs
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = [ln for ln in res.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(noise))
    if verbose and noise:
        print("\n".join(noise), file=sys.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
