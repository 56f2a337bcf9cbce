"""The drop-in boundary, built from the reference's own caller.

The reference's driver (experiment.cpp + utilities.cpp, unchanged) is
compiled in a temporary directory that holds the reference's sources minus
the four headers this build replaces (gkc_cuda.hpp, kmer_counting.hpp,
superkmers.hpp, types.h) and the two CUDA files, against include/gkc/, and
linked with libgkc_dropin.so + libgkc_hip.so.  Quote-includes search the
source's own directory first, so the replaced headers must be absent there.

The guard: compiled against the reference's own headers instead, the same
driver must fail to link (its SKMStore and T_h_data have other layouts; the
drop-in's types live in an inline namespace, so the symbols do not match)
rather than link and corrupt memory.

Skipped when the reference tree or g++ is absent (machines that only run
the built library have no reference tree).  Nothing is copied into the repository.
"""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
REPLACED = {"gkc_cuda.hpp", "kmer_counting.hpp", "superkmers.hpp", "types.h", "gkc_cuda.cu", "kmer_counting.cu"}
LIB = os.path.join(ROOT, "gkc_amd", "lib")

pytestmark = pytest.mark.skipif(not os.path.isdir(REF) or shutil.which("g++") is None,
                                reason="needs the reference tree and g++")


def _libs_built():
    if not (os.path.exists(os.path.join(LIB, "libgkc_dropin.so")) and os.path.exists(os.path.join(LIB, "libgkc_hip.so"))):
        pytest.skip("drop-in libraries not built")


def _gxx(cwd, srcs, extra=()):
    cmd = ["g++", "-std=gnu++14", "-O1", "-w", "-I", os.path.join(ROOT, "include", "gkc"), "-I",
           os.path.join(ROOT, "include"), *srcs, "-L", LIB, "-lgkc_dropin", "-lgkc_hip", f"-Wl,-rpath,{LIB}",
           "-lpthread", "-o", "gkc", *extra]
    return subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=600)


def test_reference_driver_links_against_dropin(tmp_path):
    _libs_built()
    for name in os.listdir(REF):
        src = os.path.join(REF, name)
        if os.path.isfile(src) and name not in REPLACED and name.endswith((".cpp", ".hpp", ".h")):
            shutil.copy(src, tmp_path / name)
    assert not any((tmp_path / n).exists() for n in REPLACED)
    r = _gxx(tmp_path, ["experiment.cpp", "utilities.cpp"])
    assert r.returncode == 0, r.stderr[-3000:]
    # every header the driver pulled in for the replaced names came from include/gkc
    dep = subprocess.run(["g++", "-std=gnu++14", "-MM", "-I", os.path.join(ROOT, "include", "gkc"), "experiment.cpp"],
                         cwd=tmp_path, capture_output=True, text=True, timeout=600)
    assert dep.returncode == 0, dep.stderr
    for n in ("gkc_cuda.hpp", "kmer_counting.hpp", "superkmers.hpp", "types.h"):
        assert os.path.join(ROOT, "include", "gkc", n) in dep.stdout
    assert os.access(tmp_path / "gkc", os.X_OK)


def test_reference_headers_do_not_link_against_dropin(tmp_path):
    """experiment.cpp compiled beside the reference's own headers (its
    moodycamel SKMStore), with include/gkc on the include path as before, cannot
    reach the drop-in's entry points."""
    _libs_built()
    for name in os.listdir(REF):
        src = os.path.join(REF, name)
        if os.path.isfile(src) and name.endswith((".cpp", ".hpp", ".h")):
            shutil.copy(src, tmp_path / name)
    r = _gxx(tmp_path, ["experiment.cpp", "utilities.cpp"])
    # it compiles against the reference's headers, and only the link fails, on the drop-in's symbols
    assert r.returncode != 0
    assert "undefined reference" in r.stderr and "kmc_counting_GPU" in r.stderr, r.stderr[-3000:]
