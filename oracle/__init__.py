"""FP64 CPU oracle for snapMRF's dictionary-generation and matching path.

TEST INFRASTRUCTURE ONLY: the FP64 reference that tests/,
``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg compare the HIP
product path against.  The
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
        # ORACLE_LIB selects another build of the same sources (the
        # sanitizer build liboracle_san.so)
        path = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
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
        L.orc_dm_jacobian.argtypes = [P, S, P, ctypes.c_int, P, P]
        L.orc_dm_adjoint.argtypes = [P, S, P, ctypes.c_int, P, P, P, P]
        L.orc_dm_crb.argtypes = [P, S, P, S, ctypes.c_int, P, P]
        L.orc_dm_crb.restype = ctypes.c_int
        L.orc_inv4.argtypes = [P, P]
        L.orc_inv4.restype = ctypes.c_int
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


# ---- differentiable EPG-SSFP model (diff.c) -------------------------------

def dm_jacobian(raw: np.ndarray, theta, w: int, jac: bool = True):
    """Continuous FP64 EPG-SSFP signal s [T] (complex128) and, if jac, the
    Jacobian dS/d(T1, T2, B0, B1) [4, T] (complex128)."""
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    T = raw.shape[1]
    th = np.ascontiguousarray(theta, dtype=np.float64)
    s = np.empty(T, np.complex128)
    J = np.empty((4, T), np.complex128) if jac else None
    lib().orc_dm_jacobian(_p(raw), T, _p(th), w, _p(s), _p(J) if jac else None)
    return (s, J) if jac else s


def dm_adjoint(raw: np.ndarray, theta, w: int, v: np.ndarray):
    """Gradients of Re sum conj(v) s: (g_theta [4], g_fa [T], g_tr [T])."""
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    T = raw.shape[1]
    th = np.ascontiguousarray(theta, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.complex128)
    g = np.empty(4)
    gfa = np.empty(T)
    gtr = np.empty(T)
    lib().orc_dm_adjoint(_p(raw), T, _p(th), w, _p(v), _p(g), _p(gfa), _p(gtr))
    return g, gfa, gtr


def dm_crb(raw: np.ndarray, tissues: np.ndarray, w: int, wts) -> float:
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    tis = np.ascontiguousarray(tissues, dtype=np.float64)
    wt = np.ascontiguousarray(wts, dtype=np.float64)
    out = np.zeros(1)
    rc = lib().orc_dm_crb(_p(raw), raw.shape[1], _p(tis), tis.shape[0], w, _p(wt), _p(out))
    if rc:
        raise ValueError("singular Fisher matrix")
    return float(out[0])


def inv4(A: np.ndarray) -> np.ndarray:
    A = np.ascontiguousarray(A, dtype=np.float64)
    X = np.empty((4, 4))
    if lib().orc_inv4(_p(A), _p(X)):
        raise ValueError("singular")
    return X
