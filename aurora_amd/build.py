"""Builds the gfx950 render-core library in-tree: aurora_amd/libaurora_amd.so.

    python -m aurora_amd.build [--arch gfx950] [--save-temps DIR]

Every HIP source is compiled for gfx950 with:
  -ffp-contract=off   no fused multiply-add is formed, so float results are
                      reproducible and identical to a scalar IEEE restatement;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions are generated.
HIP's default correctly rounded f32 division and square root are kept.
"""
from __future__ import annotations

import argparse
import glob
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libaurora_amd.so")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build aurora_amd)")


def flags(arch: str) -> list[str]:
    return [
        "-O3",
        "-std=c++17",
        f"--offload-arch={arch}",
        "-fPIC",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall",
        "-Wno-unused-function",
    ]


def build(arch: str = "gfx950", save_temps: str | None = None, out: str = LIB) -> str:
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    tmp = os.path.abspath(out + ".tmp")
    extra, cwd = [], None
    if save_temps:
        os.makedirs(save_temps, exist_ok=True)
        extra, cwd = ["-save-temps"], save_temps
    cmd = [hipcc(), *flags(arch), *extra, "-shared", "-o", tmp, *srcs]
    res = subprocess.run(cmd, capture_output=True, text=True, cwd=cwd)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode})")
    noise = [ln for ln in res.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")
    os.replace(tmp, out)
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--arch", default="gfx950")
    ap.add_argument("--save-temps", default=None, help="directory for compiler temporaries")
    a = ap.parse_args()
    print(build(a.arch, a.save_temps))


if __name__ == "__main__":
    main()
