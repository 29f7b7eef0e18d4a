"""FP64 CPU oracle for snapMRF's dictionary-generation and matching path.

TEST INFRASTRUCTURE ONLY: the checker that tests/, ``__graft_entry__.smoke()``
and bench.py's ``cpu_baseline`` leg compare the HIP product path against.  The
product package (``snapmrf_amd``) never imports this module.

Thin ctypes wrappers over ``oracle/liboracle.so`` (built by ``oracle/Makefile``
from oracle.c, a C restatement of src/functions.cu:456-1416 of the reference).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ECHO = {"epg_ssfp": 0, "roa_ssfp": 1, "epg_sr": 2, "epg_se": 3,
        "epg_tse1": 4, "epg_tse2": 4, "epg_fisp": 5}


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        S = ctypes.c_size_t
        L.orc_seq_prepare.argtypes = [ctypes.c_int, P, S, P]
        L.orc_epg.argtypes = [ctypes.c_int, P, S, P, S, S, ctypes.c_int, P]
        L.orc_roa.argtypes = [P, S, P, S, ctypes.c_int, ctypes.c_double, P]
        L.orc_normalize.argtypes = [P, S, S]
        L.orc_match_slice.argtypes = [P, S, S, P, S, P, P, P]
        L.orc_b1_slice.argtypes = [ctypes.c_float, P, ctypes.c_int]
        L.orc_b1_slice.restype = ctypes.c_int
        L.orc_wall.restype = ctypes.c_double
        L.orc_num_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def seq_prepare(echo: str, raw: np.ndarray) -> np.ndarray:
    """raw: float32 [4, T] csv columns -> prepared float32 [4, T]."""
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    out = np.empty_like(raw)
    lib().orc_seq_prepare(ECHO[echo], _p(raw), raw.shape[1], _p(out))
    return out


def epg(echo: str, seq: np.ndarray, params: np.ndarray, natoms: int, w: int) -> np.ndarray:
    """seq: prepared float32 [4, T]; params float32 [A, 4] -> complex128 [A, T]."""
    seq = np.ascontiguousarray(seq, dtype=np.float32)
    params = np.ascontiguousarray(params, dtype=np.float32)
    T = seq.shape[1]
    A = params.shape[0]
    out = np.zeros((A, T, 2), dtype=np.float64)
    rc = lib().orc_epg(ECHO[echo], _p(seq), T, _p(params), A, natoms, w, _p(out))
    if rc != 0:
        raise ValueError("orc_epg failed")
    return out[..., 0] + 1j * out[..., 1]


def roa(seq: np.ndarray, params: np.ndarray, natoms: int, n_iso: int = 1,
        spread_hz: float = 0.0) -> np.ndarray:
    seq = np.ascontiguousarray(seq, dtype=np.float32)
    params = np.ascontiguousarray(params, dtype=np.float32)
    T = seq.shape[1]
    out = np.zeros((natoms, T, 2), dtype=np.float64)
    rc = lib().orc_roa(_p(seq), T, _p(params), natoms, n_iso, spread_hz, _p(out))
    if rc != 0:
        raise ValueError("orc_roa failed")
    return out[..., 0] + 1j * out[..., 1]


def match_slice(atoms_c64: np.ndarray, sig_c64: np.ndarray):
    """atoms [na, T] complex64, signals [V, T] complex64 -> (idx int64[V], pd f64[V], second f64[V])."""
    atoms = np.ascontiguousarray(atoms_c64, dtype=np.complex64)
    sig = np.ascontiguousarray(sig_c64, dtype=np.complex64)
    na, T = atoms.shape
    V = sig.shape[0]
    idx = np.zeros(V, dtype=np.int64)
    pd = np.zeros(V, dtype=np.float64)
    sec = np.zeros(V, dtype=np.float64)
    lib().orc_match_slice(_p(atoms), T, na, _p(sig), V, _p(idx), _p(pd), _p(sec))
    return idx, pd, sec


def b1_slice(b1v: float, slice_b1: np.ndarray) -> int:
    sb = np.ascontiguousarray(slice_b1, dtype=np.float32)
    return int(lib().orc_b1_slice(ctypes.c_float(b1v), _p(sb), len(sb)))


def wall() -> float:
    return float(lib().orc_wall())


def num_threads() -> int:
    return int(lib().orc_num_threads())
