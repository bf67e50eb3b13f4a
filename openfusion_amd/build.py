"""Builds the native libraries in-tree.

* ``openfusion_amd/libkinfu.so``  — HIP kernels + C ABI (hipcc, gfx950)
* ``openfusion_amd/libkfsynth.so`` — host frame sources: synthetic streams and the
  recorded-stream format (g++, OpenMP)
* ``oracle/libkforacle.so``        — the CPU oracle (g++, OpenMP), via oracle/Makefile

Every call recompiles.  Usage: ``python -m openfusion_amd.build [--no-oracle]``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}", f"-I{CSRC}",
]

LIB_SOURCES = [
    "kf_capi.cpp",
    "kf_pipeline.hip",
    "kf_preprocess.hip",
    "kf_volume.hip",
    "kf_raycast.hip",
    "kf_icp.hip",
    "kf_prior.hip",
    "kf_export.hip",
    "kf_conditioning.hip",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    # Surface real warnings (the host pass does not know -packed-fp32-ops: harmless).
    for line in (r.stderr or "").splitlines():
        if "warning" in line and "packed-fp32-ops" not in line:
            sys.stderr.write(line + "\n")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", "-o", obj, os.path.join(CSRC, src)])
    return obj


def build_kinfu() -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    out = os.path.join(PKG, "libkinfu.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_synth() -> str:
    out = os.path.join(PKG, "libkfsynth.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          f"-I{INCLUDE}", "-o", tmp, os.path.join(CSRC, "kf_synth.cpp"), os.path.join(CSRC, "kf_recording.cpp")])
    os.replace(tmp, out)
    return out


def build_oracle() -> str:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-B"])
    return os.path.join(ROOT, "oracle", "libkforacle.so")


def build_all(oracle: bool = True) -> None:
    with ThreadPoolExecutor(max_workers=3) as ex:
        futs = [ex.submit(build_kinfu), ex.submit(build_synth)]
        if oracle:
            futs.append(ex.submit(build_oracle))
        for f in futs:
            f.result()


if __name__ == "__main__":
    build_all(oracle="--no-oracle" not in sys.argv)
    print("built")
