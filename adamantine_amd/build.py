"""Build the HIP library libadm.so in-tree for gfx950 (and the CPU oracle).

Usage: python -m adamantine_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libadm.so")

# translation units: the API plus one per cell-kernel configuration
# (fe_degree, number of material states, property evaluator)
CELL_CONFIGS = [(p, ns, k) for p in (1, 2, 3, 4, 5) for ns in (1, 2, 3)
                for k in (0, 1, 2)]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # keep the written operation order (FMAs are explicit in the kernels)
    "-ffp-contract=off",
    # hardware FP64 atomics
    "-munsafe-fp-atomics",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _compile(args):
    cmd, verbose = args
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise subprocess.CalledProcessError(r.returncode, cmd)
    return cmd[-1]


def build_library(verbose: bool = True, jobs: int = 0) -> str:
    """Compile every translation unit for gfx950 and link libadm.so."""
    from concurrent.futures import ThreadPoolExecutor
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(ROOT, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include")]
    jobs_list = []
    objs = []
    api_obj = os.path.join(objdir, "adm_api.o")
    jobs_list.append(([hipcc] + HIPCC_FLAGS + inc + ["-c", "-o", api_obj,
                      os.path.join(CSRC, "adm_api.hip")], verbose))
    objs.append(api_obj)
    da_obj = os.path.join(objdir, "adm_da.o")
    jobs_list.append(([hipcc] + HIPCC_FLAGS + inc + ["-c", "-o", da_obj,
                      os.path.join(CSRC, "adm_da.hip")], verbose))
    objs.append(da_obj)
    for unit in ("adm_adjoint", "adm_activation"):
        obj = os.path.join(objdir, unit + ".o")
        jobs_list.append(([hipcc] + HIPCC_FLAGS + inc + ["-c", "-o", obj,
                          os.path.join(CSRC, unit + ".hip")], verbose))
        objs.append(obj)
    for (p, ns, t) in CELL_CONFIGS:
        obj = os.path.join(objdir, "adm_cell_p%d_s%d_t%d.o" % (p, ns, t))
        jobs_list.append(([hipcc] + HIPCC_FLAGS + inc + [
            "-DADM_P=%d" % p, "-DADM_NS=%d" % ns, "-DADM_PK=%d" % t,
            "-c", "-o", obj, os.path.join(CSRC, "adm_cell_inst.hip")],
            verbose))
        objs.append(obj)
    n = jobs or min(len(jobs_list), os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=n) as ex:
        list(ex.map(_compile, jobs_list))
    tmp = LIB + ".tmp"
    link = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_cxx_demo(verbose: bool = True) -> str:
    """The host C++ program on include/adm_cxx.hpp (tests/cxx/), linked
    against the in-tree libadm.so."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    src = os.path.join(ROOT, "tests", "cxx", "adapter_demo.cpp")
    out = os.path.join(ROOT, "tests", "cxx", "adapter_demo")
    cmd = [hipcc, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
           src, "-L", os.path.dirname(LIB), "-ladm",
           "-Wl,-rpath,$ORIGIN/../../adamantine_amd", "-o", out]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return out


def build_oracle(verbose: bool = True):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")],
                   check=True)


def main():
    build_library()
    build_cxx_demo()
    build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main())
