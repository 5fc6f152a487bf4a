"""Build libmpmx.so (the gfx950 hot path) in-tree with hipcc.

    python -m mpmwarpydw_amd.build          # compile + link
    python -m mpmwarpydw_amd.build --check  # also list the exported mpmx_* symbols

Objects go to ``mpmwarpydw_amd/_build/`` and the library to
``mpmwarpydw_amd/libmpmx.so`` (both git-ignored; the .so travels with the tree).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmpmx.so")
ARCH = os.environ.get("MPMX_OFFLOAD_ARCH", "gfx950")

SOURCES = ["mpmx_runtime.hip", "mpm_kernels.hip", "p2g.hip", "substep.hip", "reduce.hip", "xpbd.hip", "mlp.hip",
           "learned.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -packed-fp32-ops: no v_pk_* FP32 instructions (their results are not trusted
# on this hardware); no fast-math: NaN/inf handling follows IEEE as the
# reference's guards (isnan, divisions by zero mass) rely on it.
CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    noise = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def exported_symbols(path: str = LIB) -> list[str]:
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return sorted({ln.split()[-1] for ln in out.splitlines() if ln.split()[-1].startswith("mpmx_")})


if __name__ == "__main__":
    build(verbose=True)
    if "--check" in sys.argv:
        print("\n".join(exported_symbols()))
