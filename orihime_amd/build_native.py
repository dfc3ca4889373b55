"""Build orihime_amd/libohm_sw.so: the gfx950 HIP kernels and the host (CPU)
implementation behind the C ABI declared in include/ohm_sw.h.

    python orihime_amd/build_native.py          # or build() from Python

(run it as a script: importing the package loads the library it builds).

The library is built in-tree so it travels with the source tree.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libohm_sw.so")
PROBE_LIB = os.path.join(PKG, "libohm_probe.so")

ARCH = os.environ.get("OHM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions are kept out of generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # device code: no IEEE-mode NaN quieting (fmax/fmin need no canonicalize);
    # the kernels never see NaN inputs (scores are validated on the host)
    "-Xarch_device", "-mno-amdgpu-ieee", "-Xarch_device", "-fno-honor-nans",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-pthread", f"-I{INCLUDE}", f"-I{CSRC}"]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the ROCm toolchain is required to build libohm_sw.so")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in res.stdout.splitlines()
        if "packed-fp32-ops" not in line and "is not a recognized feature" not in line
    )
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{out}")


def _start(cmd: list[str]):
    return cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)


def _finish(job) -> None:
    cmd, proc = job
    out, _ = proc.communicate()
    out = "\n".join(line for line in out.splitlines()
                    if "packed-fp32-ops" not in line and "is not a recognized feature" not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def build(verbose: bool = False) -> str:
    """Compile every native source and link libohm_sw.so; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    # the device sources compile in parallel (sw_kernels.hip dominates)
    jobs, objs = [], []
    for src in ("sw_kernels.hip", "embed_mfma.hip", "scan_inputs.hip"):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        jobs.append(_start([hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]))
        objs.append(obj)
    for src in ("sw_cpu.cpp", "embed_cpu.cpp"):
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for job in jobs:
        _finish(job)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    # issue-rate probes (measured roofline constants), a library of their own
    pobj = os.path.join(BUILD, "peak_probe.o")
    _run([hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, "peak_probe.hip"), "-o", pobj])
    tmp = PROBE_LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "-o", tmp, pobj])
    os.replace(tmp, PROBE_LIB)
    if verbose:
        print(f"built {LIB}")
        print(f"built {PROBE_LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
