"""The oracle and the host side of both C-ABI libraries under AddressSanitizer
and UndefinedBehaviorSanitizer: `make san` builds instrumented copies
(oracle/_san, le2d_amd/lib_san; the device code is unchanged), and the CPU
tests of the oracle, the ABI and the I/O paths (material, set_ball, VTK,
dump/restore, CLI, allocator hook arguments) run against them with the clang
sanitizer runtime preloaded.  Any report fails the run."""
import glob
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RT = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))


@pytest.fixture(scope="module")
def san_build():
    if not RT or not shutil.which("/opt/rocm/bin/hipcc"):
        pytest.skip("clang sanitizer runtime or hipcc not present")
    jobs = str(min(8, os.cpu_count() or 1))
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    subprocess.run(["make", "-s", "-j", jobs, "-C", os.path.join(ROOT, "le2d_amd", "csrc"), "san"],
                   check=True)
    env = dict(os.environ, LD_PRELOAD=RT[-1], ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               LE2D_LIBDIR=os.path.join(ROOT, "le2d_amd", "lib_san"),
               LE2D_ORACLE_LIBDIR=os.path.join(ROOT, "oracle", "_san"))
    return env


def test_instrumented_libraries_are_the_ones_loaded(san_build):
    for lib in ("le2d_amd/lib_san/liblecore_d.so", "le2d_amd/lib_san/liblecore_f.so",
                "oracle/_san/liboracle_d.so", "oracle/_san/liboracle_f.so"):
        syms = subprocess.run(["nm", "-D", os.path.join(ROOT, lib)], capture_output=True,
                              text=True, check=True).stdout
        assert "__asan_report" in syms and "__ubsan_handle" in syms, lib
    code = ("import sys, numpy as np; sys.path.insert(0, %r); import le2d_amd; "
            "from oracle import Oracle; "
            "print(le2d_amd.load(np.float64)._name); print(Oracle(np.float32).lib._name)" % ROOT)
    out = subprocess.run([sys.executable, "-c", code], env=san_build, capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    assert "lib_san/liblecore_d.so" in out.stdout and "_san/liboracle_f.so" in out.stdout


def test_cpu_suites_pass_under_sanitizers(san_build):
    tests = [os.path.join(ROOT, "tests", f) for f in ("test_oracle.py", "test_abi.py",
                                                       "test_io.py")]
    out = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu",
                          "-p", "no:cacheprovider"] + tests, env=san_build, cwd=ROOT,
                         capture_output=True, text=True, timeout=1200)
    text = out.stdout + out.stderr
    assert out.returncode == 0, text[-4000:]
    assert "AddressSanitizer" not in text and "runtime error" not in text, text[-4000:]
