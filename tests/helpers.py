"""Shared test helpers: the reference's unit-test sequences and grids, and
error measures."""
import numpy as np


def const_seq(T, alpha, phi, tr, te):
    """generate_echo (src/functions.cu:1587-1598): constant [4][T] columns."""
    return np.array([[alpha] * T, [phi] * T, [tr] * T, [te] * T], dtype=np.float32)


def fixture_params():
    """T1 "100:100:200", T2 "20+20:20:40+1000", B0 "0", B1 "1" (src/test.cu:167-170)
    -> 4 atoms, T1 fastest."""
    return np.array([[100, 20, 0, 1], [200, 20, 0, 1], [100, 40, 0, 1], [200, 40, 0, 1]],
                    dtype=np.float32)


def mrf5_raw(golden):
    """data/MRF_5.csv as float32 [4, 5] csv columns."""
    rows = open(f"{golden}/MRF_5.csv").read().splitlines()[1:]
    return np.ascontiguousarray(np.array([r.split(",") for r in rows], dtype=np.float32).T)


def rel_err(a, b):
    """max over atoms of max_t |a - b| / max_t |b| (infinity-norm scaled)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.abs(b).max(axis=-1)
    den = np.where(den > 0, den, 1.0)
    return float((np.abs(a - b).max(axis=-1) / den).max())
