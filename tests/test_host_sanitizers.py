"""The library's host layer under AddressSanitizer and UndefinedBehaviorSanitizer.

The CPU tests that drive the host code behind the C ABI (Y-bus build, symbolic phase, update and
solve schedule builders, islands, PGM ingest) run again in a child process against the host-sanitized
build of the library (gridoxide_amd.build.build_sanitized: host sources compiled with
-fsanitize=address,undefined, the kernels as in the normal build), with the sanitizer runtime
preloaded; any report fails the run.  CPU only: no kernel is launched.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_TESTS = ["tests/test_abi_cpu.py", "tests/test_symbolic_cpu.py", "tests/test_schedule_cpu.py",
              "tests/test_pgm_cpu.py"]


def _sanitized_lib():
    from gridoxide_amd import build
    srcs = [os.path.join(build.CSRC, s) for s in build.SOURCES] + [os.path.join(ROOT, "include", "gridox_hip.h")]
    stale = not os.path.exists(build.SAN_OUT) or any(os.path.getmtime(s) > os.path.getmtime(build.SAN_OUT) for s in srcs)
    if stale:
        build.build_sanitized()
    return build.SAN_OUT, build.sanitizer_runtime()


def test_host_layer_clean_under_asan_and_ubsan():
    lib, runtime = _sanitized_lib()
    # the sanitized build carries the instrumentation (ASan checks on loads, UBSan handlers)
    syms = subprocess.run(["nm", "-D", "--undefined-only", lib], capture_output=True, text=True, check=True).stdout
    assert "__asan_report_load" in syms and "__ubsan_handle_" in syms
    env = dict(os.environ, GRIDOX_HIP_LIB=lib, LD_PRELOAD=runtime,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    probe = ("import sys; sys.path.insert(0, %r); from gridoxide_amd import _lib; _lib.lib(); "
             "maps = open('/proc/self/maps').read(); "
             "assert 'libclang_rt.asan' in maps and 'libgridox_hip_san.so' in maps, 'sanitized build not loaded'" % ROOT)
    subprocess.run([sys.executable, "-c", probe], env=env, cwd=ROOT, check=True, timeout=300)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu"] + HOST_TESTS,
                       env=env, cwd=ROOT, capture_output=True, text=True, timeout=1500)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
