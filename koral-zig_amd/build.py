"""In-tree build of the koral-zig_amd libraries.

  lib/libkz_hip.so   HIP kernels + C-ABI for gfx950 (csrc/*.hip)
  lib/libkz_host.so  caller-side host library (host/kz_host.cpp)

Both are compiled with -ffp-contract=off: no fused multiply-adds anywhere,
so device and host arithmetic round exactly as written.
"""
import os
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
LIB = HERE / "lib"
HIP_SOURCES = [HERE / "csrc" / "kz_step.hip", HERE / "csrc" / "kz_options.hip", HERE / "csrc" / "kz_cellops.hip"]
HOST_SOURCES = [HERE / "host" / "kz_host.cpp"]
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    # keep hipcc from emitting packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libkz_hip.so)")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(map(str, cmd))}")


def compile_hip(out, defines=None):
    """Compile every HIP source to an object of its own, in parallel, and
    link them into the shared library `out`.  `defines` maps macro names to
    values (the tuner's knobs)."""
    out = Path(out)
    dflags = [f"-D{k}={v}" for k, v in (defines or {}).items()]
    cflags = [f for f in HIPCC_FLAGS if f != "-shared"]
    with tempfile.TemporaryDirectory(prefix="kz_build_") as td:
        objs, procs = [], []
        for src in HIP_SOURCES:
            obj = Path(td) / (src.stem + ".o")
            objs.append(obj)
            cmd = [_hipcc(), *cflags, *dflags, "-c", "-o", str(obj), str(src)]
            procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        failed = []
        for cmd, p in procs:
            text, _ = p.communicate()
            if p.returncode != 0:
                sys.stderr.write(text)
                failed.append(" ".join(map(str, cmd)))
        if failed:
            raise RuntimeError("build failed: " + "; ".join(failed))
        tmp = out.with_name(out.name + ".tmp")
        _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
        os.replace(tmp, out)
    return out


def build_hip():
    LIB.mkdir(exist_ok=True)
    return compile_hip(LIB / "libkz_hip.so")


def build_host():
    LIB.mkdir(exist_ok=True)
    out = LIB / "libkz_host.so"
    tmp = LIB / "libkz_host.so.tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-mfma", "-fopenmp", "-o", str(tmp),
          *map(str, HOST_SOURCES)])
    os.replace(tmp, out)
    return out


def build_all():
    build_host()
    build_hip()


if __name__ == "__main__":
    build_all()
