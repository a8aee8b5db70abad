"""Build libnfft2d.so in-tree with hipcc for gfx950.

Every HIP source under csrc/ is compiled into one shared library next to this
file.  The packed-FP32 target feature is disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_PATH = os.path.join(PKG_DIR, "libnfft2d.so")
SOURCES = ["nfft2d_api.hip", "ndft2d.hip", "nufft3.hip", "knab.hip"]
ARCH = os.environ.get("NFFT2D_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def compile_cmd(out: str) -> list[str]:
    return [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            # MFMA accumulators in VGPRs: the gathers read them back with VALU
            # FMAs, which AGPRs would cost a v_accvgpr_read each
            "-mllvm", "-amdgpu-mfma-vgpr-form",
            "-Wno-unused-command-line-argument",
            "-o", out] + [os.path.join(CSRC, s) for s in SOURCES]


def build_library(verbose: bool = False) -> str:
    """Compile (always) and atomically install miscellaneous_amd/libnfft2d.so."""
    tmp = LIB_PATH + f".tmp{os.getpid()}"
    cmd = compile_cmd(tmp)
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        if os.path.exists(tmp):
            os.unlink(tmp)
        raise RuntimeError("hipcc failed:\n" + res.stdout[-8000:])
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


ROOT = os.path.dirname(PKG_DIR)
REFCALLS_SRC = os.path.join(ROOT, "tests", "cpp", "reference_calls.cpp")
REFCALLS_BIN = os.path.join(ROOT, "tests", "cpp", "reference_calls")


def build_reference_calls(verbose: bool = False) -> str:
    """Compile tests/cpp/reference_calls.cpp (the reference's call syntax through
    include/nfft2d_overloads.hpp) against libnfft2d.so, run-path set to the package."""
    tmp = REFCALLS_BIN + f".tmp{os.getpid()}"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
           "-Wno-unused-command-line-argument", "-o", tmp, REFCALLS_SRC, "-L", PKG_DIR, "-lnfft2d",
           "-Wl,-rpath," + PKG_DIR, "-Wl,-rpath,$ORIGIN/../../miscellaneous_amd"]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        if os.path.exists(tmp):
            os.unlink(tmp)
        raise RuntimeError("hipcc failed on reference_calls.cpp:\n" + res.stdout[-8000:])
    os.replace(tmp, REFCALLS_BIN)
    return REFCALLS_BIN


HOSTCHK_SRC = os.path.join(ROOT, "tests", "cpp", "host_checks.cpp")
HOSTCHK_BIN = os.path.join(ROOT, "tests", "cpp", "host_checks")


def build_host_checks(verbose: bool = False) -> str:
    """Compile tests/cpp/host_checks.cpp together with the library sources, the host
    code under AddressSanitizer (device code is built as for the library)."""
    tmp = HOSTCHK_BIN + f".tmp{os.getpid()}"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O1", "-gline-tables-only", "-std=c++17", "-fPIC",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-mllvm", "-amdgpu-mfma-vgpr-form",
           "-I", os.path.join(ROOT, "include"), "-Wno-unused-command-line-argument", "-o", tmp, HOSTCHK_SRC]
    cmd += [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        if os.path.exists(tmp):
            os.unlink(tmp)
        raise RuntimeError("hipcc failed on host_checks.cpp:\n" + res.stdout[-8000:])
    os.replace(tmp, HOSTCHK_BIN)
    return HOSTCHK_BIN


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_reference_calls(verbose=True))
    print(build_host_checks(verbose=True))
