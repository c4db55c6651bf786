"""The DLPack argument/ownership layer of the C ABI, built with g++ under AddressSanitizer.

caustica_amd/csrc/dl_abi.cpp is plain C++ (no HIP): it checks the tensors the DLPack
entry points borrow and builds the managed tensors they export. Here it is compiled
with the host compiler and ``-fsanitize=address,undefined`` together with the checks in
tests/native/dl_abi_check.cpp, and run: every accepted and refused tensor form, and the
export lifecycle (the deleter releases the data once, nothing leaks).
"""

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_dl_abi_under_asan(tmp_path):
    exe = tmp_path / "dl_abi_check"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I", str(ROOT / "include"),
           str(ROOT / "caustica_amd" / "csrc" / "dl_abi.cpp"),
           str(ROOT / "tests" / "native" / "dl_abi_check.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60,
                         env={"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0"})
    assert run.returncode == 0, run.stdout + run.stderr
    assert "dl_abi ok" in run.stdout
