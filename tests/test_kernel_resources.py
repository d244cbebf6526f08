"""Every gfx950 kernel of the library keeps its working set in registers and LDS: no kernel may use
scratch (per-lane private memory in HBM).  The state-update kernels are HBM-bound streams, and scratch
traffic lands on the same HBM: a select of a whole complex value in the whole-tile permutation kernel
once sent 144 B per lane through scratch and doubled that kernel's measured write bytes.  The check
compiles each HIP source for the device only and reads the compiler's per-kernel resource remarks.

One exception, bounded: the 1024-thread shapes of the batched-QAOA kernel run 16 waves per CU and are
held to 128 VGPRs; they spill a few per-evaluation values (loaded once per evaluation or per tile,
never inside the rx or load loops), and the 2^12-amplitude one is the autotuner's fastest shape."""
import glob
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "orkan_amd", "csrc")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# the sources that define kernels (k_noalloc.hip holds a device array only)
SOURCES = sorted(s for s in glob.glob(os.path.join(CSRC, "k_*.hip")) if "__global__" in open(s).read())
# mangled-name fragment -> largest scratch (bytes per lane) accepted
ALLOWED = {"qaoa_persistentILi12ELi1024E": 32, "qaoa_persistentILi13ELi1024E": 96}


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
@pytest.mark.parametrize("src", SOURCES, ids=[os.path.basename(s) for s in SOURCES])
def test_no_kernel_uses_scratch(src, tmp_path):
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-Wno-unused-result", "-I", CSRC, "-I", os.path.join(ROOT, "include"), "-c", "-x",
           "hip", src, "-o", str(tmp_path / "k.o"), "--offload-device-only", "-Rpass-analysis=kernel-resource-usage"]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    kernels, spilled = 0, []
    name = None
    for line in out.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            kernels += 1
            continue
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and int(m.group(1)) != 0:
            limit = next((v for k, v in ALLOWED.items() if k in name), 0)
            if int(m.group(1)) > limit:
                spilled.append((name, int(m.group(1))))
    assert kernels > 0, "no resource remarks: the compiler output format changed"
    assert not spilled, spilled
