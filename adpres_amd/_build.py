"""Build the gfx950 engine library in-tree (adpres_amd/libadpres_gpu.so)."""

from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libadpres_gpu.so")
SOURCES = [os.path.join(HERE, "csrc", "adg_engine.hip")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 VALU forms are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose: bool = False, out: str = LIB, defines=()) -> str:
    """Compile the engine; `defines` (e.g. ["ADG_NTF=512"]) override the
    launch-shape defaults for tuning builds written to `out`."""
    cmd = [HIPCC, *HIP_FLAGS, *("-D" + d for d in defines), "-o", out + ".tmp", *SOURCES]
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose:
        print(r.stderr)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
