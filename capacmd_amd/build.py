"""In-tree build of the native parts (no GPU needed; hipcc cross-compiles gfx950).

    python -m capacmd_amd.build           # host lib, GPU lib, CLI
    python -m capacmd_amd.build --oracle  # also the oracle (test infrastructure)

Outputs (git-ignored, shipped with the tree):
    capacmd_amd/libcapac_host.so   C deck reader/writers           (gcc)
    capacmd_amd/libcapac_gpu.so    HIP kernels + C ABI             (hipcc, gfx950)
    capacmd_amd/capacmd            CLI, -gro/-par/-mds              (gcc, links both)
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
ARCH = os.environ.get("CAPAC_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def run(cmd, cwd=None):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_host():
    out = os.path.join(HERE, "libcapac_host.so")
    run(["gcc", "-O2", "-fPIC", "-Wall", "-shared", f"-I{INC}", "-o", out,
         os.path.join(HERE, "host", "deck.c"), "-lm"])
    return out


def build_gpu():
    out = os.path.join(HERE, "libcapac_gpu.so")
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for s in srcs:
        o = os.path.join(objdir, os.path.basename(s)[:-4] + ".o")
        run([HIPCC] + HIP_FLAGS + [f"-I{INC}", "-c", "-o", o, s])
        objs.append(o)
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out] + objs)
    return out


def build_cli():
    out = os.path.join(HERE, "capacmd")
    run(["gcc", "-O2", "-Wall", f"-I{INC}", "-o", out, os.path.join(HERE, "host", "main.c"),
         f"-L{HERE}", "-lcapac_host", "-lcapac_gpu", "-Wl,-rpath,$ORIGIN", "-lm"])
    return out


def build_oracle(with_ref: bool = True):
    run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if with_ref and os.path.isdir("/root/reference"):
        run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    build_host()
    build_gpu()
    if os.path.exists(os.path.join(HERE, "host", "main.c")):
        build_cli()
    if "--oracle" in argv:
        build_oracle()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
