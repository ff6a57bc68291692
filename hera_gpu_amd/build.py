"""Build ``libheravis.so`` (gfx950 HIP kernels + C ABI) in-tree with hipcc.

    python -m hera_gpu_amd.build          # or hera_gpu_amd.build.build()

Every translation unit is compiled for gfx950 only, with packed-FP32 VALU
instructions disabled, then linked into ``hera_gpu_amd/libheravis.so``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
LIB = PKG / "libheravis.so"
OBJ_DIR = PKG / "_build"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}",
]


def sources():
    return sorted(CSRC.glob("*.hip"))


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass warns about the device-only target feature; drop that noise
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    return r.returncode, out


def build(verbose: bool = False, jobs: int = 8) -> Path:
    """Compile every HIP source for gfx950 and link the shared library."""
    OBJ_DIR.mkdir(exist_ok=True)
    srcs = sources()
    objs = [OBJ_DIR / (s.stem + ".o") for s in srcs]

    def compile_one(pair):
        src, obj = pair
        return _run([HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]), src

    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        results = list(ex.map(compile_one, zip(srcs, objs)))
    for (rc, out), src in results:
        if verbose and out.strip():
            print(out)
        if rc != 0:
            raise RuntimeError(f"hipcc failed on {src.name}:\n{out}")
    tmp = LIB.with_suffix(".so.tmp")
    rc, out = _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    if rc != 0:
        raise RuntimeError(f"link failed:\n{out}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
