"""Build recipe of libmetriko_hip.so (hipcc, gfx950), in-tree next to this file."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmetriko_hip.so")
SOURCES = ["kernels.hip", "block.hip", "mg.hip", "capi.hip", "igm.hip", "pipeline.hip", "batch.hip", "noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("METRIKO_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # host-side setup (sparse products, transposes, triplet assembly) runs on OpenMP threads
    "-fopenmp",
    # no packed FP32 instructions (v_pk_*_f32) in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-result", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]


def build(verbose=False, jobs=4):
    os.makedirs(OBJ, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(OBJ, src + ".o")
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fopenmp", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_seams_test(verbose)
    return LIB


SEAMS_SRC = os.path.join(ROOT, "tests", "native", "seams_test.cpp")
SEAMS_BIN = os.path.join(ROOT, "tests", "native", "bin", "seams_test")


def build_seams_test(verbose=False):
    """The C++ test of the drop-in solver seams, linked with the library (run by a GPU test)."""
    if not os.path.exists(SEAMS_SRC):
        return None
    os.makedirs(os.path.dirname(SEAMS_BIN), exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), SEAMS_SRC, "-L", HERE, "-lmetriko_hip",
           "-Wl,-rpath,$ORIGIN/../../../metriko_amd", "-o", SEAMS_BIN]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed on seams_test.cpp:\n{r.stdout}\n{r.stderr}")
    return SEAMS_BIN


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
