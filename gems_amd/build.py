"""Builds the gfx950 HIP library in-tree: gems_amd/lib/libgems_hip.so."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(OUT_DIR, "libgems_hip.so")
SOURCES = ["gh_kernels.hip", "gh_tile.hip", "gh_batch.hip", "gh_api.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # packed FP32 instructions stay disabled on these machines
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = [HIPCC] + FLAGS + [os.path.join(SRC, s) for s in SOURCES] + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    errs = [ln for ln in (r.stdout + r.stderr).splitlines()
            if "warning" in ln and "packed-fp32-ops" not in ln and "target-feature" not in ln]
    if verbose and errs:
        print("\n".join(errs))
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
