"""CPU tests of the host-side boundary in libsnapmrf.so: symbol exports, grid
parser (src/functions.cu:196-414), rfpulse CSV reader (:100-158), RA I/O
(src/ra.cu) and the mrf CLI's argument handling.  No GPU calls."""
import os
import re
import struct
import subprocess
import sys

import numpy as np
import pytest

import ra_util
import snapmrf_amd as sm

HEADER_H = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "include", "snapmrf.h")


def test_library_exports_every_declared_symbol():
    L = sm.lib()
    decl = set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(mrf_\w+)\s*\(", open(HEADER_H).read(), re.M))
    assert decl == set(sm.EXPORTS)
    for name in decl:
        assert hasattr(L, name), name


@pytest.mark.parametrize("s,expected", [
    ("100:100:200", [100, 200]),                 # src/test.cu:167
    ("20+20:20:40+1000", [20, 40, 1000]),        # src/test.cu:168, README.md:156-160
    ("600:600:600", [600]),
    ("1000:1000", [1000]),                       # "a:b" = min a, max b, step 1
    ("1:5", [1, 2, 3, 4, 5]),
    ("0:1:0", [0]),
    ("0", [0]),
    ("0.8:0.1:1.2", [0.8, 0.9, 1.0, 1.1, 1.2]),
    # a non-deduplicated range leaves the running end one step too far, so the
    # next range's first value 5 is taken for a duplicate and dropped
    ("0:1:2+3:1:4+5:1:6", [0, 1, 2, 3, 4, 6]),
])
def test_grid_strings(s, expected):
    got = sm.parse_grid(s)
    assert len(got) == len(expected)
    assert np.allclose(got, np.array(expected, np.float32), rtol=0, atol=1e-6)


def test_grid_values_are_float32_linspace():
    # min + i*step in float32 (src/functions.cu:269-274)
    got = sm.parse_grid("0.8:0.1:1.2")
    i = np.arange(5, dtype=np.float32)
    assert np.array_equal(got, np.float32(0.8) + i * np.float32(0.1))


def test_config_grids():
    p, A, a = sm.build_params("100:20:2000", "20:10:300", "0", "1")
    assert (A, a) == (2663, 2663)
    p, A, a = sm.build_params("100:20:2000+2300:300:5000", "20:5:100+110:10:300+400:200:2000",
                              "-40:2:40", "0.8:0.1:1.2")
    assert (A, a) == (871455, 174291)
    # atom order: T1 fastest, then T2, B0, B1 (src/functions.cu:355-391)
    assert p[0].tolist() == pytest.approx([100, 20, -40, 0.8])
    assert p[1][0] == 120 and p[1][1] == 20
    assert np.all(p[:, 0] > p[:, 1])
    assert np.all(p[:a, 3] == p[0, 3]) and np.all(p[4 * a:, 3] == p[4 * a, 3])


def test_fixture_grid():
    p, A, a = sm.build_params("100:100:200", "20+20:20:40+1000", "0", "1")
    assert A == 4
    assert p.tolist() == [[100, 20, 0, 1], [200, 20, 0, 1], [100, 40, 0, 1], [200, 40, 0, 1]]


def test_malformed_grid():
    with pytest.raises(sm.MRFError):
        sm.parse_grid("300:10:100")     # max < min (reference: assert)


def test_logspace_symmetric_branch():
    v = sm.logspace(-150.0, 150.0, 21)
    assert v[10] == 0.0
    assert v[0] == pytest.approx(-150.0, rel=1e-5) and v[20] == pytest.approx(150.0, rel=1e-5)
    assert np.allclose(v[:10], -v[11:][::-1])
    w = sm.logspace(0.5, 1.5, 11)
    assert w[0] == pytest.approx(0.5) and w[-1] == pytest.approx(1.5, rel=1e-6)
    assert np.all(np.diff(np.log10(w)) > 0)


def test_natoms_mode_grid():
    t1, t2, b0, b1 = sm.grid_natoms_mode(10000)
    # l_t1 = round(sqrt(10000/21/11)) with integer divisions = round(sqrt(43)) = 7
    assert len(t1) == len(t2) == 7 and len(b0) == 21 and len(b1) == 11
    assert t1[0] == pytest.approx(100.0) and t1[-1] == pytest.approx(3000.0, rel=1e-5)


def test_csv_reader_fixture(golden):
    path = f"{golden}/MRF_5.csv"
    assert sm.csv_dim(path) == 6
    cols, warn = sm.csv_read(path, 5)
    assert warn == ""
    assert np.allclose(cols[0], [0.084857, 0.091571, 0.098286, 0.105, 0.111714])
    assert np.all(cols[1:] == 0)
    big = f"{golden}/MRF001.csv"
    assert sm.csv_dim(big) == 1001
    cols, warn = sm.csv_read(big, 1000)
    assert warn == "" and cols[2, 0] == np.float32(1.50) and cols[0, 3] == np.float32(0.105)


def test_csv_short_header_drops_first_row(tmp_path):
    # fgets(100)-chunk semantics: a header shorter than 99 characters makes
    # the first data row the second skipped chunk (src/functions.cu:138-155)
    p = tmp_path / "short.csv"
    p.write_text("fa,ph,tr,te\n1,2,3,4\n5,6,7,8\n9,10,11,12\n")
    assert sm.csv_dim(str(p)) == 4
    cols, warn = sm.csv_read(str(p), 3)
    assert "zero-filled" in warn
    assert cols[:, 0].tolist() == [5, 6, 7, 8] and cols[:, 1].tolist() == [9, 10, 11, 12]
    assert cols[:, 2].tolist() == [0, 0, 0, 0]


def test_csv_errors(tmp_path):
    p = tmp_path / "bad.csv"
    p.write_text("x" * 99 + "\n1,2\n")
    with pytest.raises(sm.MRFError):
        sm.csv_read(str(p), 1)
    with pytest.raises(sm.MRFError):
        sm.csv_dim(str(tmp_path / "missing.csv"))


@pytest.mark.parametrize("name", ["atoms_epg_ssfp.ra", "atoms_mrf.ra", "params_mrf.ra",
                                  "B1_phantom.ra"])
def test_ra_roundtrip_is_byte_identical(golden, tmp_path, name):
    src = f"{golden}/{name}"
    arr, hdr = sm.ra_read(src)
    assert np.array_equal(arr, ra_util.read(src))
    out = str(tmp_path / name)
    sm.ra_write(out, arr, hdr["dims"])
    assert open(out, "rb").read() == open(src, "rb").read()


def test_ra_file_larger_than_2gib(tmp_path):
    """A (T, A) complex64 atoms file past 2 GiB: every byte written (the
    reference's writer over-reads there, src/ra.cu:83-88) and read back."""
    T, A = 1000, 270_000                      # 2.16e9 bytes of data
    arr = np.empty((A, T), dtype=np.complex64)
    arr.view(np.uint32).ravel()[:] = np.arange(2 * A * T, dtype=np.uint32)
    path = str(tmp_path / "big.ra")
    sm.ra_write(path, arr)
    assert os.path.getsize(path) == 6 * 8 + 2 * 8 + arr.nbytes
    disk = np.memmap(path, dtype=np.uint32, mode="r", offset=64)
    assert np.array_equal(disk[-1000:], arr.view(np.uint32).ravel()[-1000:])
    del disk
    back, hdr = sm.ra_read(path)
    assert hdr["dims"] == [T, A] and hdr["size"] == arr.nbytes
    assert np.array_equal(back.view(np.uint32), arr.view(np.uint32))


def test_ra_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.ra"
    p.write_bytes(b"\0" * 64)
    with pytest.raises(sm.MRFError):
        sm.ra_read(str(p))
    good = np.arange(10, dtype=np.float32)
    q = str(tmp_path / "t.ra")
    sm.ra_write(q, good)
    data = open(q, "rb").read()
    open(q, "wb").write(data[:-8])          # truncated data
    with pytest.raises(sm.MRFError):
        sm.ra_read(q)


def _ra_bytes(eltype, elbyte, size, dims, payload):
    return struct.pack("<6Q", ra_util.MAGIC, 0, eltype, elbyte, size, len(dims)) + \
        struct.pack("<%dQ" % len(dims), *dims) + payload


@pytest.mark.parametrize("size,dims", [
    (16, [4, 4]),                       # header size smaller than its dims imply
    (80, [4, 4]),                       # larger
    (8, [1 << 62, 8]),                  # dims product overflows 64 bits
    (0, [0, 5]),                        # zero-sized: fine (reads nothing)
])
def test_ra_header_size_must_match_dims(tmp_path, size, dims):
    """The readers of images and B1 maps size their copies from dims, so a
    header whose size disagrees with elbyte * prod(dims) is rejected (the
    reference's ra_read trusts both, src/ra.cu:23-55)."""
    p = str(tmp_path / "h.ra")
    open(p, "wb").write(_ra_bytes(3, 4, size, dims, b"\0" * size))
    if size == 4 * dims[0] * dims[1] and size < (1 << 40):
        arr, hdr = sm.ra_read(p)
        assert hdr["size"] == size
    else:
        with pytest.raises(sm.MRFError, match="prod"):
            sm.ra_read(p)


def test_cli_usage_and_errors(tmp_path):
    r = subprocess.run([sm.CLI_PATH, "-h"], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage: mrf [OPTION] <rfpulse.csv> <imgstack.ra>" in r.stderr
    # no atoms with T1 > T2 -> EX_SOFTWARE (src/mrf.cu:240-241), before any GPU work
    r = subprocess.run([sm.CLI_PATH, "-t", "10", "-s", "20"], capture_output=True, text=True)
    assert r.returncode == 70 and "Number of atoms is 0" in r.stderr
    r = subprocess.run([sm.CLI_PATH, "--precision", "fp8"], capture_output=True, text=True)
    assert r.returncode == 64


def test_bench_refuses_more_than_one_gpu():
    """The measured path runs on one GPU; --gpus above 1 exits with a message
    before touching a device."""
    root = os.path.dirname(os.path.dirname(HEADER_H))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, cwd=root, timeout=300)
    assert r.returncode == 2 and "one GPU" in (r.stdout + r.stderr)
