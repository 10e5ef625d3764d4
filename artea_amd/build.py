"""Builds artea_amd/libartea_hip.so with hipcc for gfx950 (in-tree).

Every HIP source is compiled with:
  -ffp-contract=off                            no FMA contraction: FP32 rounds
                                               where the source says, fmaf()
                                               where the reference uses FMA;
  -fhip-fp32-correctly-rounded-divide-sqrt     IEEE sqrt / divide (cosine);
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 VALU ops.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libartea_hip.so")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ARTEA_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["index.hip", "router.hip", "router_noalloc.hip", "refine.hip", "util.hip", "bruteforce.hip", "server.hip",
               "estimators.hip", "routing.hip", "rnet.hip", "hgra_device.hip"]
CXX_SOURCES = ["hgra.cpp", "datagen.cpp"]
# router_inst.hip is compiled once per (metric family, candidate-list blocks)
# pair; router_kernels.hpp lists the same pairs.
ROUTER_INST = [(mf, k) for mf in (0, 1, 2) for k in (1, 2, 4, 8, 13, 16)]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-I", os.path.join(ROOT, "include")]
HIP_FLAGS = COMMON + [
    "--offload-arch=" + ARCH,
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fopenmp",
]
CXX_FLAGS = COMMON + ["-x", "c++", "-fopenmp", "-march=x86-64-v3"]


def _compile(job) -> str:
    src, tag, defs = job if isinstance(job, tuple) else (job, "", [])
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + tag + ".o")
    flags = HIP_FLAGS if src.endswith(".hip") else CXX_FLAGS
    cmd = [HIPCC] + flags + defs + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = [("router_inst.hip", "_mf%d_k%d" % (mf, k), ["-DARTEA_RK_MF=%d" % mf, "-DARTEA_RK_K=%d" % k])
            for mf, k in ROUTER_INST]
    jobs += [s for s in HIP_SOURCES + CXX_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(int(os.environ.get("MAX_JOBS", "8")), len(jobs))) as ex:
        objs = list(ex.map(_compile, jobs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-fopenmp", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
