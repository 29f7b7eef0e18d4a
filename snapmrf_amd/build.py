"""Builds the in-tree gfx950 library ``snapmrf_amd/libsnapmrf.so`` and the
``mrf`` command-line tool ``snapmrf_amd/bin/mrf``.

    python -m snapmrf_amd.build            # or snapmrf_amd.build.build()
    python -m snapmrf_amd.build --sanitize # also snapmrf_amd/_san/: host code
                                           # under ASan + UBSan (see build_sanitized)

Every HIP translation unit is compiled by hipcc for gfx950 only, with packed
FP32 instructions disabled; host-only C++ is compiled with -ffp-contract=off
so its float32 arithmetic matches the reference's host code bit for bit.
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
OBJ = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libsnapmrf.so")
CLI = os.path.join(HERE, "bin", "mrf")

HIP_SOURCES = ["epg.hip", "match.hip", "lowp.hip", "grad.hip", "svd.hip", "order.hip", "capi.hip"]
HOST_SOURCES = ["host.cpp"]
ARCH = "gfx950"
# rocBLAS / rocSOLVER: the plain GEMMs and the eigensolver of the SVD basis
LIBS = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def _hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _hip_flags() -> list[str]:
    return ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
            f"-I{INCLUDE}", f"-I{CSRC}"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    err = "\n".join(l for l in r.stderr.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if err.strip():
        sys.stderr.write(err + "\n")


def _compile(src: str) -> str:
    os.makedirs(OBJ, exist_ok=True)
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        cmd = [_hipcc(), *_hip_flags(), "-c", "-o", out, path]
    else:
        cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall",
               f"-I{INCLUDE}", f"-I{CSRC}", "-c", "-o", out, path]
    _run(cmd)
    return out


def build(verbose: bool = False) -> str:
    srcs = HIP_SOURCES + HOST_SOURCES
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-Wl,-soname,libsnapmrf.so", *LIBS])
    os.replace(tmp, LIB)
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    cli_obj = _compile_cli()
    _run(["g++", "-o", CLI, cli_obj, f"-L{HERE}", "-lsnapmrf",
          "-Wl,-rpath,$ORIGIN/..", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    if verbose:
        print(f"built {LIB} and {CLI}")
    return LIB


SAN = os.path.join(HERE, "_san")
SAN_LIB = os.path.join(SAN, "libsnapmrf.so")
SAN_CLI = os.path.join(SAN, "mrf")
SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
             "-fno-omit-frame-pointer", "-g"]


def build_sanitized(verbose: bool = False) -> str:
    """The library and CLI with their host-only C++ (host.cpp: CSV, grid
    strings, RA I/O; mrf_main.cpp: the CLI) compiled by g++ under
    AddressSanitizer + UndefinedBehaviorSanitizer, into snapmrf_amd/_san/.
    The HIP translation units are linked in from the regular build without
    instrumentation (no GPU code is sanitized).  Load the library with
    SNAPMRF_LIB / SNAPMRF_CLI pointing here and libasan + libubsan preloaded."""
    hip_objs = [os.path.join(OBJ, src + ".o") for src in HIP_SOURCES]
    if not all(os.path.exists(o) for o in hip_objs):
        build()
    os.makedirs(SAN, exist_ok=True)
    host_o = os.path.join(SAN, "host.o")
    cli_o = os.path.join(SAN, "mrf_main.o")
    _run(["g++", "-std=c++17", "-O1", "-fPIC", "-ffp-contract=off", "-Wall", *SAN_FLAGS,
          f"-I{INCLUDE}", f"-I{CSRC}", "-c", "-o", host_o, os.path.join(CSRC, "host.cpp")])
    _run(["g++", "-std=c++17", "-O1", "-Wall", *SAN_FLAGS, f"-I{INCLUDE}",
          "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c", "-o", cli_o,
          os.path.join(CSRC, "mrf_main.cpp")])
    _run(["g++", "-shared", *SAN_FLAGS, "-o", SAN_LIB, host_o, *hip_objs,
          "-Wl,-soname,libsnapmrf.so", "-L/opt/rocm/lib", "-lamdhip64", *LIBS])
    _run(["g++", *SAN_FLAGS, "-o", SAN_CLI, cli_o, f"-L{SAN}", "-lsnapmrf",
          "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    if verbose:
        print(f"built {SAN_LIB} and {SAN_CLI} (host code under ASan + UBSan)")
    return SAN_LIB


def _compile_cli() -> str:
    os.makedirs(OBJ, exist_ok=True)
    out = os.path.join(OBJ, "mrf_main.o")
    _run(["g++", "-std=c++17", "-O2", "-Wall", f"-I{INCLUDE}", "-D__HIP_PLATFORM_AMD__",
          "-I/opt/rocm/include", "-c", "-o", out, os.path.join(CSRC, "mrf_main.cpp")])
    return out


if __name__ == "__main__":
    build(verbose=True)
    if "--sanitize" in sys.argv[1:]:
        build_sanitized(verbose=True)
