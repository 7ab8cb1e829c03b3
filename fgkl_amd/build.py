"""Builds the in-tree engine library ``fgkl_amd/libfgkl.so`` for gfx950 with hipcc.

Every source is compiled with ``-Xclang -target-feature -Xclang -packed-fp32-ops`` so the
compiler never emits packed FP32 instructions; the host pass warns that it does not know the
feature, which is harmless.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OBJ = HERE / "_build"
LIB = HERE / "libfgkl.so"
SOURCES = ["sweep_f32.hip", "sweep_f32_uniform.hip", "sweep_f32_pd_uniform.hip", "sweep_f64.hip", "engine.hip", "noalloc.hip", "peak.hip", "adjoint.hip", "sw.hip", "plan.cpp",
           "synthetic.cpp", "replay.cpp", "aggregate.cpp", "capi.cpp", "jni_shim.cpp"]
ARCH = os.environ.get("FGKL_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the fgkl engine needs ROCm to build")


def _flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-fPIC",
        "-std=c++17",
        "-Wall",
        "-Wno-unused-command-line-argument",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        f"-I{HERE.parent / 'include'}",
    ]


def build(verbose: bool = False) -> Path:
    """Compiles every source and links the shared library; always rebuilds."""
    OBJ.mkdir(exist_ok=True)
    cc = hipcc()
    srcs = [s for s in SOURCES if (CSRC / s).exists()]

    def compile_one(src: str) -> Path:
        out = OBJ / (src.replace(".", "_") + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [cc, *_flags(), *lang, "-c", str(CSRC / src), "-o", str(out)]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{res.stdout}\n{res.stderr}")
        return out

    with ThreadPoolExecutor(max_workers=min(len(srcs), 8)) as pool:
        objs = list(pool.map(compile_one, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
