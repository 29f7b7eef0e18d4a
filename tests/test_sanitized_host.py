"""The host code that replaces the reference's unchecked parsers (CSV reader
with fixed fgets buffers, src/functions.cu:100-158; grid strings copied into
char[1024], :244-245 and :322-323; RA reads and writes, src/ra.cu:23-105),
the mrf CLI's argument handling and the FP64 oracle, run under
AddressSanitizer + UndefinedBehaviorSanitizer.

tests/test_host.py, tests/test_oracle.py and tests/test_oracle_diff.py rerun in
a subprocess against the sanitized builds (snapmrf_amd.build.build_sanitized:
host.cpp and mrf_main.cpp instrumented by g++; make -C oracle SANITIZE=1), with
the sanitizer runtimes preloaded into Python.  Any report fails the run
(UBSan without recovery, ASan aborts on error).  No GPU calls."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    r = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True)
    path = r.stdout.strip()
    if r.returncode or not os.path.isabs(path) or not os.path.exists(path):
        pytest.fail(f"{name} not found (g++ -print-file-name)")
    return path


def test_host_code_and_oracle_under_asan_and_ubsan():
    from snapmrf_amd import build as b
    b.build_sanitized()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "SANITIZE=1"], check=True)
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": " ".join(filter(None, [_runtime("libasan.so"), _runtime("libubsan.so"),
                                             os.environ.get("LD_PRELOAD", "")])),
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1",   # the interpreter's own leaks
        "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
        "SNAPMRF_LIB": b.SAN_LIB,
        "SNAPMRF_CLI": b.SAN_CLI,
        "ORACLE_LIB": os.path.join(ROOT, "oracle", "liboracle_san.so"),
    })
    r = subprocess.run([sys.executable, "-m", "pytest", "tests/test_host.py", "tests/test_oracle.py",
                        "tests/test_oracle_diff.py", "-x", "-q", "-p", "no:cacheprovider",
                        "-m", "not gpu"], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert " passed" in out
