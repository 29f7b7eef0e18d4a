"""Deterministic synthetic inputs for tests and benchmarks: the MRF sequence
(flip train, TR, TE), rfpulse-CSV columns in the reference's units, voxel
signals drawn from a dictionary with proton density and complex Gaussian
noise, and a smooth B1 map.

Sequence (per timepoint t = 0..T-1, seeded):
    alpha_t = 10 deg + 50 deg * |sin(pi t / 250)| + U(0, 5 deg),  phase 0,
    TR_t ~ U(10.5, 14) ms,  TE_t = TR_t / 2.
"""
from __future__ import annotations

import numpy as np

HEADER = ("flip_angle_(fraction_of_nominal),rf_phase_(deg),extension_of_base_TR_(ms),"
          "extension_of_base_TE_(ms)")   # 99 characters, as data/MRF001.csv


def sequence(T: int = 1000, seed: int = 0):
    rng = np.random.default_rng(seed)
    t = np.arange(T)
    alpha = 10.0 + 50.0 * np.abs(np.sin(np.pi * t / 250.0)) + rng.uniform(0.0, 5.0, T)
    tr = rng.uniform(10.5, 14.0, T)
    te = tr / 2.0
    return dict(alpha_deg=alpha, phi_deg=np.zeros(T), tr=tr, te=te)


def csv_columns(seq, echo: str) -> np.ndarray:
    """float32 [4, T] as csv_read returns them for this echo type: the
    ssfp/roa programs get (alpha/60, phase deg, TR-16, TE-3.5) which the unit
    transform maps back; the other programs read raw radians and ms."""
    a, p, tr, te = seq["alpha_deg"], seq["phi_deg"], seq["tr"], seq["te"]
    if echo in ("epg_ssfp", "roa_ssfp"):
        cols = [a / 60.0, p, tr - 16.0, te - 3.5]
    else:
        cols = [np.deg2rad(a), np.deg2rad(p), tr, te]
    return np.ascontiguousarray(np.array(cols, dtype=np.float32))


def write_csv(path: str, cols: np.ndarray) -> None:
    with open(path, "w") as f:
        f.write(HEADER + "\n")
        for i in range(cols.shape[1]):
            f.write("%.9g,%.9g,%.9g,%.9g\n" % tuple(float(c) for c in cols[:, i]))


def b1_map(nx: int, ny: int, seed: int = 0, lo: float = 0.8, hi: float = 1.2) -> np.ndarray:
    """Smooth 2-D low-order polynomial mapped to [lo, hi], float32 [ny, nx]."""
    rng = np.random.default_rng(seed)
    y, x = np.meshgrid(np.linspace(-1, 1, ny), np.linspace(-1, 1, nx), indexing="ij")
    c = rng.normal(size=6)
    f = c[0] + c[1] * x + c[2] * y + c[3] * x * x + c[4] * x * y + c[5] * y * y
    f = (f - f.min()) / max(f.max() - f.min(), 1e-12)
    return (lo + (hi - lo) * f).astype(np.float32)


def signals_np(atoms: np.ndarray, V: int, seed: int = 0, pd_range=(0.5, 1.0), snr: float = 30.0,
               candidates: np.ndarray | None = None):
    """Voxels = PD * atom[idx] + complex Gaussian noise.  Noise sigma per
    real component = RMS|PD*atom| / (snr * sqrt(2)).  Returns (sig complex64
    [V, T], idx int64 [V], pd float32 [V])."""
    rng = np.random.default_rng(seed)
    A = atoms.shape[0]
    idx = (rng.integers(0, A, V) if candidates is None
           else candidates[rng.integers(0, len(candidates), V)])
    pd = rng.uniform(*pd_range, V).astype(np.float32)
    clean = atoms[idx] * pd[:, None]
    rms = np.sqrt(np.mean(np.abs(clean) ** 2, axis=1))
    sigma = rms / (snr * np.sqrt(2.0))
    noise = (rng.normal(size=clean.shape) + 1j * rng.normal(size=clean.shape)) * sigma[:, None]
    return (clean + noise).astype(np.complex64), idx, pd
