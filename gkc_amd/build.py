"""Build the package's native libraries in-tree (gfx950).

    python -m gkc_amd.build            # everything
Outputs (git-ignored, shipped with the tree):
    gkc_amd/lib/libgkc_hip.so      HIP kernels + C ABI (include/gkc_hip.h)
    gkc_amd/lib/libgkc_dropin.so   C++ drop-in wrappers (include/gkc/*.hpp)
    gkc_amd/lib/gkc_count          C++ driver using the drop-in entry points
    gkc_amd/lib/libgkc_synth.so    synthetic workload generator (host C)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "_build")
INC = os.path.join(ROOT, "include")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("GKC_ARCH", "gfx950")
HIP_SOURCES = ["partition.hip", "sort.hip", "mlp.hip", "ingest.hip", "calib.hip", "positions.hip", "vendor_sort.hip",
               "capi.hip"]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-result",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", INC, "-I", CSRC]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and "warning: feature" not in l)
    if "error" in err:
        sys.stderr.write(err + "\n")


def build_hip() -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(OBJ, s.replace(".hip", ".o")) for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=min(4, len(srcs))) as ex:
        list(ex.map(lambda so: _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, so[0]), "-o", so[1]]),
                    zip(srcs, objs)))
    out = os.path.join(LIB, "libgkc_hip.so")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", out + ".tmp", *objs])
    os.replace(out + ".tmp", out)
    return out


def build_synth() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libgkc_synth.so")
    _run(["gcc", "-O3", "-fPIC", "-shared", "-std=gnu11", "-Wall", "-o", out, os.path.join(CSRC, "synth.c"),
          "-lm", "-lpthread"])
    return out


def build_dropin() -> list[str]:
    """C++ drop-in wrappers (host C++ only, talking to libgkc_hip.so through the C ABI)."""
    src = os.path.join(CSRC, "dropin.cpp")
    if not os.path.exists(src):
        return []
    lib = os.path.join(LIB, "libgkc_dropin.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", INC, "-o", lib, src,
          "-L", LIB, "-lgkc_hip", "-Wl,-rpath,$ORIGIN", "-lpthread"])
    drv = os.path.join(CSRC, "gkc_count.cpp")
    outs = [lib]
    if os.path.exists(drv):
        exe = os.path.join(LIB, "gkc_count")
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INC, "-o", exe, drv, "-L", LIB, "-lgkc_dropin",
              "-lgkc_hip", "-Wl,-rpath,$ORIGIN", "-lpthread"])
        outs.append(exe)
    return outs


def build_all() -> None:
    build_synth()
    build_hip()
    build_dropin()


if __name__ == "__main__":
    build_all()
    print("built:", ", ".join(sorted(os.listdir(LIB))))
