"""Build libpssfss_hip.so for gfx950 with hipcc (in-tree, next to this file).

Every HIP source is compiled with packed-FP32 instructions disabled and linked
into one shared library whose only interface is include/pssfss_hip.h.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpssfss_hip.so")
SOURCES = ["pgf.hip", "fill.hip", "linalg.hip", "gsmk.hip", "adjoint.hip", "ufp.hip",
           "capi.hip"]
CALIB_SRC = "calib.hip"
CALIB_LIB = os.path.join(HERE, "libpssfss_calib.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-munsafe-fp-atomics", "-Wno-unused-command-line-argument"]


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES + [CALIB_SRC]))
    calib_obj = objs.pop()
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", LIB]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    # counter-calibration kernels (bench.py --calibrate), a library of their own
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", calib_obj, "-o",
                           CALIB_LIB])
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
