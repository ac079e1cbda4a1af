"""Build libatsat_gpu.so in-tree: every HIP translation unit of atsat_amd/csrc for gfx950.

    python -m atsat_amd.build          (or __graft_entry__.build())

Flags: -ffp-contract=off keeps the FP64 arithmetic in the reference's operand order (no fused
multiply-adds the reference build does not have); packed FP32 instructions are disabled.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libatsat_gpu.so")
# the device's HBM streaming rates for the stage kernel's roofline (bench.py --hbm-probe); not
# part of the library
PROBE_SRC = "hbm_probe.hip"
PROBE = os.path.join(HERE, "hbm_probe")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["k_rk.hip", "k_press.hip", "k_phys.hip", "k_glue.hip", "k_solve.hip", "k_turb.hip", "k_chem.hip", "k_mca.hip", "k_pmodes.hip",
           "atsat_gpu.cpp", "host_setup.cpp", "restart.cpp"]

FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-I" + os.path.join(REPO, "include"), "-I" + CSRC]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC"] + objs + ["-o", tmp, "-lrocblas"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, LIB)
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", os.path.join(CSRC, PROBE_SRC), "-o", PROBE + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (PROBE_SRC, r.stderr[-4000:]))
    os.replace(PROBE + ".tmp", PROBE)
    if verbose:
        print("built", LIB, "and", PROBE)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
