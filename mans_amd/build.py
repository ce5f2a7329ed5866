"""Build libmans.so and the `compress` driver for gfx950 with hipcc.

    python -m mans_amd.build            # library + CLI
Outputs (git-ignored, built in-tree so they travel with the source tree):
    mans_amd/libmans.so
    mans_amd/bin/compress
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmans.so")
CLI = os.path.join(HERE, "bin", "compress")
PROFILER = os.path.join(HERE, "bin", "mans_profile")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MANS_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: hipcc must not form v_pk_*_f32 (none of the codec's
# arithmetic is floating point except the one fp32 division in the
# normalisation, which must stay a plain correctly rounded v_div sequence).
# -fno-strict-aliasing: the coders write 16-bit words into LDS stages and
# spill areas and move them as uint2 / uint4; those accesses must stay in
# program order.
DEVICE_FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-fno-strict-aliasing",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall"]

HIP_SOURCES = ["hist.hip", "encode.hip", "decode.hip", "adm.hip", "batch.hip", "batch_p9.hip", "batch_p10.hip",
               "batch_p11.hip", "archive.hip", "lorenzo.hip", "capi.hip"]
HOST_SOURCES = ["gpu_ans_shim.cpp", "serve.cpp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    noise = [ln for ln in (r.stdout + r.stderr).splitlines()
             if ln.strip() and "packed-fp32-ops" not in ln and "not a recognized feature" not in ln]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        jobs.append([HIPCC, *DEVICE_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        jobs.append([HIPCC, *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    cli_tmp = CLI + ".tmp"
    _run([HIPCC, *HOST_FLAGS, os.path.join(CSRC, "compress_main.cpp"), "-o", cli_tmp,
          f"-L{HERE}", "-lmans", "-Wl,-rpath,$ORIGIN/.."])
    os.replace(cli_tmp, CLI)
    prof_tmp = PROFILER + ".tmp"
    _run([HIPCC, *DEVICE_FLAGS, "-x", "hip", os.path.join(CSRC, "profile_main.cpp"), "-o", prof_tmp,
          f"-L{HERE}", "-lmans", "-Wl,-rpath,$ORIGIN/.."])
    os.replace(prof_tmp, PROFILER)
    if verbose:
        print(f"built {LIB} and {CLI}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
