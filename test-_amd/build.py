"""Build recipe of the native library (gfx950) and of the CPU oracle.

``build_library()`` compiles ``csrc/*`` with hipcc for gfx950 into
``_build/libfemrt.so`` inside this package; ``build_oracle()`` compiles the
FP64 oracle (test infrastructure) into ``oracle/_build/liboracle.so``.
Both always recompile (the build check must exercise the compiler).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "_build"
LIB = OUT / "libfemrt.so"
RENDER = OUT / "femrt_render"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "_build" / "liboracle.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FEMRT_ARCH", "gfx950")

# Device-code flags.  packed-fp32-ops is switched off: hand-free packed FP32
# (v_pk_*_f32) forms have produced wrong results on these parts.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wno-unused-command-line-argument",
    "-munsafe-fp-atomics",
]


def _run(cmd, cwd=None):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, cwd=cwd)


def build_library(verbose: bool = True) -> Path:
    OUT.mkdir(parents=True, exist_ok=True)
    objs, jobs = [], []
    # translation units compile concurrently
    for src in sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp")):
        obj = OUT / (src.stem + ".o")
        if src.suffix == ".hip":
            cmd = [HIPCC, *HIP_FLAGS, "-c", src, "-o", obj]
        else:
            cmd = [HIPCC, "-x", "c++", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-pthread",
                   "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c", src, "-o", obj]
        print("+", " ".join(str(c) for c in cmd), flush=True)
        jobs.append((src, subprocess.Popen([str(c) for c in cmd])))
        objs.append(obj)
    failed = [str(src) for src, p in jobs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, "hipcc " + " ".join(failed))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", *objs, "-o", tmp, "-lz", "-pthread"])
    os.replace(tmp, LIB)
    # command-line renderer over the C++ host layer (links the library)
    tmp = RENDER.with_suffix(".tmp")
    _run([HIPCC, "-x", "c++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
          PKG / "femrt_render.cpp", "-o", tmp, f"-L{OUT}", "-lfemrt", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, RENDER)
    return LIB


def build_oracle() -> Path:
    ORACLE_LIB.parent.mkdir(parents=True, exist_ok=True)
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp",
          ORACLE_DIR / "femrt_oracle.c", ORACLE_DIR / "femrt_refport.c", "-o", tmp, "-lm"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


if __name__ == "__main__":
    what = sys.argv[1:] or ["lib", "oracle"]
    if "oracle" in what:
        build_oracle()
    if "lib" in what:
        build_library()
