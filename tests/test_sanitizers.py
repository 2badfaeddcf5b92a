"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C libssa_amd asan` builds the host sources with g++ and both sanitizers
(the kernels' objects come from the regular build, uninstrumented); the host
test files then run in a child process against that library, with the
sanitizer runtimes preloaded: the replay self-check (random per-pair results,
2,000 cases), the C-ABI and host logic (matrices, FASTA provider, query
reading, fatal paths), the packed DB files (including corrupt ones) and the
admission policy.  Any sanitizer report fails the child (UBSan without
recovery, ASan by default).  Leak detection is off: CPython's own allocations
would be reported.
"""
import os
import subprocess
import sys

import pytest

import _oracle as O

ROOT = O.ROOT
LIB = os.path.join(ROOT, "libssa_amd", "_build_asan", "libssa_amd_asan.so")


def runtime(name):
    p = subprocess.run(["gcc", "-print-file-name=" + name], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_host_code_under_sanitizers():
    if not os.path.exists(os.path.join(ROOT, "libssa_amd", "_build", "score_kernels.o")):
        pytest.skip("the regular build's kernel objects are not there (run __graft_entry__.build())")
    asan, ubsan = runtime("libasan.so"), runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("gcc's sanitizer runtimes are not installed")
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "libssa_amd"), "asan", "-j8"])
    env = dict(os.environ, SSA_AMD_LIB=LIB, LD_PRELOAD=asan + ":" + ubsan,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=24")
    files = ["tests/test_replay_filter.py", "tests/test_host_abi.py", "tests/test_dbpack.py",
             "tests/test_admission.py"]
    p = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu"] + files,
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = p.stdout + p.stderr
    assert p.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    assert " passed" in out
