"""Builds libdmrg_hip.so in-tree with hipcc for gfx950.

    python -m dmrgpy_amd.build          # or __graft_entry__.build()

Objects go to dmrgpy_amd/_build/, the shared library to
dmrgpy_amd/libdmrg_hip.so.  Every HIP source is compiled with packed-FP32
instructions disabled (-packed-fp32-ops) and linked against rocBLAS /
rocSOLVER (the latter provides the dense symmetric eigensolver of the SVD).
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdmrg_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
SOURCES = ["gemm.hip", "contract.hip", "krylov.hip", "svd.hip", "sytrd.hip", "capi.hip", "grad.hip"]
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build failed: {}\n{}".format(" ".join(cmd), p.stdout))
    return p.stdout


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(BUILD, s.replace(".hip", ".o")) for s in srcs]

    def compile_one(pair):
        src, obj = pair
        cmd = [HIPCC] + FLAGS + ["-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj]
        out = _run(cmd)
        if verbose and out.strip():
            print(out)
        return obj

    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(compile_one, zip(srcs, objs)))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp] + objs +
         ["-L", os.path.join(ROCM, "lib"), "-lrocsolver", "-lrocblas",
          "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
