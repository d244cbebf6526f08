"""Builds the HIP library of this package in-tree (no cmake, no JIT cache).

    python -m orkan_amd.build          # or orkan_amd.build.build()

Products (orkan_amd/lib/):
    liborkan.so      LOG_TILE_DIM=5 (32 x 32 tiles, the reference's Release layout)
    liborkan_t3.so   LOG_TILE_DIM=3 (8 x 8 tiles, the reference's Debug/test layout)

The gfx950 kernels (csrc/*.hip) are compiled once and take the tile size as a
run-time argument; the C17 host layer (csrc/*.c) is compiled per tile size,
because LOG_TILE_DIM is part of the C ABI (reference include/q_types.h:52-57).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-I", CSRC, "-I", os.path.join(ROOT, "include")]
C_FLAGS = ["-std=gnu17", "-O2", "-fPIC", "-Wall", "-Wno-unused-function", "-D__HIP_PLATFORM_AMD__",
           "-I", CSRC, "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROCM, "include")]

TILES = {"liborkan.so": 5, "liborkan_t3.so": 3}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def sources(ext):
    return sorted(f for f in os.listdir(CSRC) if f.endswith(ext))


def build(verbose=False, jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    steps = []
    kobj = []
    for f in sources(".hip"):
        o = os.path.join(BUILD, f[:-4] + ".o")
        kobj.append(o)
        steps.append([hipcc] + HIP_FLAGS + ["-c", os.path.join(CSRC, f), "-o", o])
    hobj = {}
    for lib, lt in TILES.items():
        hobj[lib] = []
        for f in sources(".c"):
            o = os.path.join(BUILD, f"{f[:-2]}_t{lt}.o")
            hobj[lib].append(o)
            steps.append(["gcc"] + C_FLAGS + [f"-DLOG_TILE_DIM={lt}", "-c", os.path.join(CSRC, f), "-o", o])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, steps):
            if verbose and out.strip():
                print(out)
    for lib in TILES:
        tmp = os.path.join(LIB, lib + ".tmp")
        _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-Bsymbolic", "-o", tmp]
             + kobj + hobj[lib] + ["-lpthread", "-lm"])
        os.replace(tmp, os.path.join(LIB, lib))
    return [os.path.join(LIB, lib) for lib in TILES]


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print(p)
