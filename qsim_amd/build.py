"""Builds the gfx950 backend library (and, where present, its C++ drivers).

``python -m qsim_amd.build`` compiles every ``csrc/*.hip`` with hipcc for
gfx950 into ``qsim_amd/lib/libqsim_hip.so``.  Compilation always runs (it is
the build check); objects are rebuilt in parallel.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(LIBDIR, "obj")
LIB = os.path.join(LIBDIR, "libqsim_hip.so")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("QSIM_HIP_ARCH", "gfx950")

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 in generated code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE, "-I" + CSRC, "-Wall", "-Wno-unused-function",
    "-Wno-unused-variable",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build_lib(verbose: bool = True) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith(".hip"))
    workers = min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


APPS = os.path.join(HERE, "apps")
BINDIR = os.path.join(HERE, "bin")
REFERENCE = os.environ.get("QSIM_REFERENCE", "/root/reference")


def build_apps(verbose: bool = True) -> None:
    """Builds the C++ drivers of qsim_amd/apps (reference drivers + this
    backend's Factory) when qsim's headers are present; skipped otherwise,
    leaving any prebuilt binaries in place."""
    if not os.path.isdir(APPS):
        return
    mk = os.path.join(APPS, "Makefile")
    if not os.path.exists(mk):
        return
    if not os.path.exists(os.path.join(REFERENCE, "lib", "run_qsim.h")):
        if verbose:
            print("qsim headers absent: C++ drivers not rebuilt")
        return
    subprocess.run(["make", "-s", "-C", APPS, f"REF={REFERENCE}"], check=True)


def main() -> int:
    build_lib()
    build_apps()
    return 0


if __name__ == "__main__":
    sys.exit(main())
