"""Build libmappraiser_gpu.so (gfx950) in-tree with hipcc.

    python -m midapack_amd.build          # compile every HIP source and link

Every source is compiled with floating-point contraction off (products and
sums rounded separately, as in the reference's C build; kernels that want FMA
call fma() explicitly) and with packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmappraiser_gpu.so")
SOURCES = ["api.hip", "pointing.hip", "toeplitz.hip", "precond.hip", "pcg.hip", "batch.hip", "server.hip", "noise.hip", "twolvl.hip", "mlmap.hip", "templates.hip", "hostio.hip"]
ARCH = os.environ.get("MGPU_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-label",
    "-I", os.path.join(HERE, "..", "include"),
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in p.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in line)
    return p.returncode, out


def build(verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = [os.path.join(BUILD, s.replace(".hip", ".o")) for s in SOURCES]
    cmds = [[HIPCC, *CFLAGS, "-c", os.path.join(CSRC, s), "-o", o] for s, o in zip(SOURCES, objs)]
    jobs = jobs or min(len(cmds), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_run, cmds))
    failed = False
    for s, (rc, out) in zip(SOURCES, results):
        if out.strip() and (verbose or rc):
            print(f"--- {s}\n{out}", file=sys.stderr)
        failed |= rc != 0
    if failed:
        raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    rc, out = _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
                    "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-lrocprofiler-sdk-roctx",
                    "-Wl,-rpath,/opt/rocm/lib"])
    if rc:
        print(out, file=sys.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
