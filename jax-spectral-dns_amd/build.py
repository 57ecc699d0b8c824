"""Build libjsd.so in-tree with hipcc for gfx950 (no JIT cache, no pip)."""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.realpath(__file__))
REPO = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops off: no v_pk_* FP32 forms are generated anywhere
FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-fPIC",
    "-shared",
    "-std=c++17",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def build(verbose: bool = False) -> str:
    src = os.path.join(HERE, "csrc", "jsd.hip")
    out_dir = os.path.join(HERE, "lib")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libjsd.so")
    tmp = out + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-I", os.path.join(REPO, "include"), "-o", tmp, src]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libjsd.so")
    if verbose:
        sys.stderr.write(res.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
