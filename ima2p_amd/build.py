"""Build libima_gpu.so (HIP, gfx950) in-tree.

``python -m ima2p_amd.build`` compiles every source under ima2p_amd/csrc
with hipcc for gfx950, one object per source in parallel (objects under
ima2p_amd/_build/), and links them into ima2p_amd/libima_gpu.so.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libima_gpu.so")
PLAIN = os.path.join(HERE, "libima_plain.so")
CXX = os.environ.get("CXX", "g++")
SOURCES = ["ctx.hip", "kern_hky.hip", "kern_lik.hip", "kern_lmode.hip", "kern_swap.hip", "kern_mc3.hip",
           "kern_prior.hip", "kern_urate.hip", "kern_topo.hip", "kern_nw.hip", "serve.hip", "ti_io.cpp", "mcf_io.cpp",
           "lmode_drivers.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    return r.returncode, r.stdout + r.stderr


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(OBJ, os.path.splitext(s)[0] + ".o") for s in srcs]
    cmds = [[HIPCC, *FLAGS, "-c", "-o", o, os.path.join(CSRC, s)] for s, o in zip(srcs, objs)]
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)), 16))
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(lambda c: _run(c, verbose), cmds))
    for (rc, out), s in zip(results, srcs):
        if rc != 0:
            sys.stderr.write(out)
            raise RuntimeError(f"hipcc failed compiling {s}")
    tmp = LIB + ".tmp"
    rc, out = _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs], verbose)
    if rc != 0:
        sys.stderr.write(out)
        raise RuntimeError("hipcc failed linking libima_gpu.so")
    os.replace(tmp, LIB)
    # the plain per-locus path (a host program over the ABI, for measurement)
    tmp = PLAIN + ".tmp"
    rc, out = _run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-o", tmp,
                    os.path.join(HERE, "plain", "plain_mc3.cpp"), f"-L{HERE}", "-lima_gpu",
                    "-Wl,-rpath,$ORIGIN"], verbose)
    if rc != 0:
        sys.stderr.write(out)
        raise RuntimeError("building libima_plain.so failed")
    os.replace(tmp, PLAIN)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
