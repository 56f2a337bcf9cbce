"""The host code under sanitizers, on the CPU: the oracle (oracle/gkc_oracle.c)
and the drop-in's host side (gkc_amd/csrc/dropin.cpp) built with gcc/g++
-fsanitize=address,undefined (leak checking on) and -fsanitize=thread, driven
by tests/sanitize/*: ragged reads with other bytes, k <= 32 and k > 32, HPC,
multi-threaded counting and concurrent SKMStore::add_skms.  Any sanitizer
report fails the run (-fno-sanitize-recover, nonzero exit)."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
LIB = os.path.join(ROOT, "gkc_amd", "lib")
ORACLE = os.path.join(ROOT, "oracle", "gkc_oracle.c")

MODES = {"asan_ubsan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"],
         "tsan": ["-fsanitize=thread"]}


def _run(cmd, env=None):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, (cmd[0], r.stdout[-2000:], r.stderr[-4000:])
    return r


@pytest.fixture(scope="module")
def tools():
    if shutil.which("gcc") is None or shutil.which("g++") is None:
        pytest.skip("needs gcc and g++")


@pytest.mark.parametrize("mode", sorted(MODES))
def test_oracle_sanitized(tools, tmp_path, mode):
    exe = str(tmp_path / "oracle_drv")
    _run(["gcc", "-g", "-O1", "-std=gnu11", *MODES[mode], os.path.join(SAN, "oracle_driver.c"), ORACLE, "-lpthread",
          "-o", exe])
    n_cfg = 1 if mode == "tsan" else 3             # thread checks are slow: the threaded count once
    r = _run([exe, str(n_cfg)], env={**os.environ, "ASAN_OPTIONS": "detect_leaks=1"})
    assert r.stdout.count("entries=") == 2 * n_cfg


@pytest.mark.parametrize("mode", sorted(MODES))
def test_dropin_host_sanitized(tools, tmp_path, mode):
    if not os.path.exists(os.path.join(LIB, "libgkc_hip.so")):
        pytest.skip("needs the built library (the drop-in's device calls link against it)")
    obj = str(tmp_path / "oracle.o")
    exe = str(tmp_path / "dropin_drv")
    _run(["gcc", "-g", "-O1", "-std=gnu11", *MODES[mode], "-c", ORACLE, "-o", obj])
    _run(["g++", "-g", "-O1", "-std=c++17", *MODES[mode], "-I", os.path.join(ROOT, "include"), "-I",
          os.path.join(ROOT, "gkc_amd", "csrc"), os.path.join(SAN, "dropin_driver.cpp"), obj, "-L", LIB, "-lgkc_hip",
          f"-Wl,-rpath,{LIB}", "-lpthread", "-o", exe])
    r = _run([exe], env={**os.environ, "ASAN_OPTIONS": "detect_leaks=1"})
    assert "stored bases" in r.stdout
