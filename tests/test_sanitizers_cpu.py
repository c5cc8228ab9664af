"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

Builds the library with sanitized host code (build.build_sanitized: the option parser and field
writer are compiled into the test program of test_dropin_options_cpu.py, the PETSc binary reader and
writer, the C-ABI and the set-up into the library) and the oracle (oracle/Makefile, target
_asan/liboracle.so), then runs the CPU tests that exercise that code in a child process with clang's
sanitizer runtime preloaded: any invalid access, leak-free use-after-free, overflow or undefined
behaviour aborts the child and fails this test.  (Leak checking is off: the Python interpreter itself
does not release everything at exit.)
"""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "topopt_in_petsc_transient_amd")
HIPCC = "/opt/rocm/bin/hipcc"
RT = glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
TESTS = ["tests/test_petsc_io_cpu.py", "tests/test_capi_cpu.py", "tests/test_dropin_options_cpu.py",
         "tests/test_oracle_pinning.py", "tests/test_oracle_pins_filter_mma.py", "tests/test_sommg_oracle_cpu.py",
         "tests/test_golden.py"]


def _stale(target, sources):
    return not os.path.exists(target) or any(os.path.getmtime(s) > os.path.getmtime(target) for s in sources)


@pytest.mark.skipif(not (os.path.exists(HIPCC) and RT), reason="hipcc or clang's sanitizer runtime absent")
def test_host_code_under_address_and_undefined_behaviour_sanitizers():
    sys.path.insert(0, ROOT)
    from topopt_in_petsc_transient_amd import build as B
    out = os.path.join(PKG, "_build_asan")
    lib = os.path.join(out, "libtopopt_gpu.so")
    srcs = glob.glob(os.path.join(PKG, "csrc", "*")) + glob.glob(os.path.join(ROOT, "include", "*.h"))
    if _stale(lib, srcs):
        B.build_sanitized(out)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "_asan/liboracle.so"], check=True)
    orc = os.path.join(ROOT, "oracle", "_asan", "liboracle.so")
    for so in (lib, orc):  # both builds carry the sanitizers' checks
        syms = subprocess.run(["nm", "-D", so], capture_output=True, text=True, check=True).stdout
        assert "__asan_report_load8" in syms and "__ubsan_handle_" in syms, so
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": RT[0],
        "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=1:detect_odr_violation=0",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
        "TT_LIB": lib,
        "TT_ORACLE_LIB": orc,
        "TT_SANITIZE_CXXFLAGS": " ".join(B.SAN_HOST_FLAGS),
        "OMP_NUM_THREADS": "1",
    })
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-m", "not gpu", "-p", "no:cacheprovider",
                        *TESTS], cwd=ROOT, env=env, capture_output=True, text=True, timeout=1500)
    print(r.stdout[-4000:], r.stderr[-4000:])
    assert r.returncode == 0, "sanitized CPU tests failed"
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr
