"""Builds the in-tree native libraries of jse_amd:

* ``libjsenn_hip.so``   the HIP kernels and the C-ABI (hipcc, gfx950)
* ``libjse_workload.so`` the synthetic-geometry helper (gcc)

Run ``python -m jse_amd.build``; the build always recompiles.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("JSE_AMD_ARCH", "gfx950")

HIP_SOURCES = ["nl.hip", "desc.hip", "desc_gen.hip", "desc_train.hip", "mlp.hip", "mlp_mfma.hip", "force.hip", "serve.hip",
               "lammps.hip", "gpumd.hip", "capi.cpp", "server.cpp", "lammps.cpp", "gpumd.cpp", "model_host.cpp", "sh_coupling.cpp",
               "yaml_json.cpp"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    # packed FP32 VALU ops are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# per-source additions: the matrix-core accumulators of the descriptor forward
# and the MLP live in ordinary VGPRs, so the code that reads them (FP64 totals,
# SiLU epilogues) needs no AGPR-to-VGPR copies
FILE_FLAGS = {"desc.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
              "mlp_mfma.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r.stderr


def build(verbose: bool = False) -> None:
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src.replace(".", "_") + ".o")
        lang = ["-x", "hip"]
        jobs.append(([HIPCC] + HIP_FLAGS + FILE_FLAGS.get(src, []) + lang + ["-c", os.path.join(CSRC, src), "-o", obj],
                     obj))
    workers = min(len(jobs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        for err in ex.map(lambda j: _run(j[0]), jobs):
            if verbose and err.strip():
                sys.stderr.write(err)
    objs = [o for _, o in jobs]
    out = os.path.join(HERE, "libjsenn_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs)
    os.replace(tmp, out)
    wl = os.path.join(HERE, "libjse_workload.so")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-o", wl + ".tmp", os.path.join(CSRC, "workload.c"), "-lm"])
    os.replace(wl + ".tmp", wl)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", os.path.join(HERE, "libjsenn_hip.so"))
