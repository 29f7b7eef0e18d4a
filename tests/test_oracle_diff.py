"""Pins the FP64 differentiable EPG-SSFP model of the oracle (oracle/diff.c):
its forward against the reference's epg_ssfp fixture, its Jacobian against
FP64 central finite differences (h = 1e-4 relative), its adjoint against the
Jacobian (dot-product test) and against finite differences over flip angles
and repetition times, and its CRB loss against numpy."""
import numpy as np
import pytest

import oracle
import ra_util
from helpers import fixture_params, mrf5_raw
from snapmrf_amd import synth

THETA = np.array([900.0, 80.0, 7.0, 0.93])


def _raw(T, seed=3):
    return synth.csv_columns(synth.sequence(T, seed=seed), "epg_ssfp").astype(np.float32)


def test_forward_matches_reference_fixture(golden):
    """Normalised continuous-model signals equal atoms_epg_ssfp.ra (MRF_5.csv,
    src/test.cu:160-201) within 1e-5."""
    raw = mrf5_raw(golden)
    gt = ra_util.read(f"{golden}/atoms_epg_ssfp.ra")
    for a, p in enumerate(fixture_params()):
        s = oracle.dm_jacobian(raw, p.astype(np.float64), 101, jac=False)
        s = s / np.linalg.norm(s)
        assert np.abs(s - gt[a]).max() / np.abs(gt[a]).max() < 1e-5


def test_forward_matches_dictionary_oracle():
    """The continuous model equals the fixture-pinned dictionary oracle
    (orc_epg) on a 120-TR random sequence, w = 25, three B1 slices."""
    raw = _raw(120)
    seq = oracle.seq_prepare("epg_ssfp", raw)
    params = np.array([[900, 80, 7, 0.8], [900, 80, 7, 1.0], [900, 80, 7, 1.2]], np.float32)
    ref = oracle.epg("epg_ssfp", seq, params, 1, 25)
    for a, p in enumerate(params):
        s = oracle.dm_jacobian(raw, p.astype(np.float64), 25, jac=False)
        s = s / np.linalg.norm(s)
        assert np.abs(s - ref[a]).max() < 2e-6


@pytest.mark.parametrize("theta", [THETA, np.array([300.0, 250.0, -31.0, 1.12]),
                                   np.array([2500.0, 30.0, 0.0, 0.85])])
def test_jacobian_central_differences(theta):
    raw = _raw(150)
    w = 20
    s, J = oracle.dm_jacobian(raw, theta, w)

    def cd(d, h):
        tp, tm = theta.copy(), theta.copy()
        tp[d] += h
        tm[d] -= h
        return (oracle.dm_jacobian(raw, tp, w, False) - oracle.dm_jacobian(raw, tm, w, False)) / (2 * h)

    for d in range(4):
        h = 1e-4 * (abs(theta[d]) if theta[d] != 0 else 1.0)
        scale = np.abs(J[d]).max()
        assert np.abs(cd(d, h) - J[d]).max() / scale < 1e-5, d
        # Richardson-extrapolated differences (O(h^4)) pin it much tighter
        rich = (4 * cd(d, h / 2) - cd(d, h)) / 3
        assert np.abs(rich - J[d]).max() / scale < 1e-8, d


def test_adjoint_dot_product():
    """<J u, v> = <u, J^T v>: the theta gradient of Re(v^H s) equals Re(v^H J)."""
    raw = _raw(130)
    rng = np.random.default_rng(0)
    v = rng.normal(size=130) + 1j * rng.normal(size=130)
    s, J = oracle.dm_jacobian(raw, THETA, 24)
    g, gfa, gtr = oracle.dm_adjoint(raw, THETA, 24, v)
    want = np.array([np.real(np.vdot(v, J[d])) for d in range(4)])
    assert np.allclose(g, want, rtol=1e-10, atol=1e-14 * np.abs(want).max())


def test_adjoint_sequence_gradients_central_differences():
    T, w = 110, 22
    raw = _raw(T, seed=5)
    rng = np.random.default_rng(1)
    v = rng.normal(size=T) + 1j * rng.normal(size=T)
    g, gfa, gtr = oracle.dm_adjoint(raw, THETA, w, v)

    def loss(r):
        return float(np.real(np.vdot(v, oracle.dm_jacobian(r, THETA, w, False))))

    for col, grad in ((0, gfa), (2, gtr)):
        for n in (0, 1, 49, 50, 77, T - 2):
            r1, r2 = raw.copy(), raw.copy()
            h = 2e-3 * max(1.0, abs(float(raw[col, n])))
            r1[col, n] = np.float32(raw[col, n] + h)
            r2[col, n] = np.float32(raw[col, n] - h)
            fd = (loss(r1) - loss(r2)) / (float(r1[col, n]) - float(r2[col, n]))
            assert abs(fd - grad[n]) <= 1e-5 * np.abs(grad).max(), (col, n)
    # the last TR's relaxation after its readout reaches no output
    assert gtr[T - 1] == 0.0


def test_inv4_and_crb_loss():
    rng = np.random.default_rng(2)
    M = rng.normal(size=(4, 4))
    A = M @ M.T + 4 * np.eye(4)
    assert np.allclose(oracle.inv4(A), np.linalg.inv(A), rtol=1e-12, atol=1e-14)
    raw = _raw(90)
    tis = np.array([[900, 80, 7, 0.93], [1500, 120, -5, 1.05]], np.float64)
    wts = np.array([1.0, 1.0, 0.1, 0.5])
    want = 0.0
    for th in tis:
        _, J = oracle.dm_jacobian(raw, th, 20)
        F = np.real(J.conj() @ J.T)
        Fi = np.linalg.inv(F)
        sc = np.array([th[0], th[1], 1.0, 1.0])
        want += float(np.sum(wts * np.diag(Fi) / sc ** 2))
    got = oracle.dm_crb(raw, tis, 20, wts)
    assert abs(got - want) <= 1e-10 * abs(want)
