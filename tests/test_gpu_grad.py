"""Differentiation form on the GPU through the C ABI: Jacobians, adjoints
(checkpointed every 4 TRs), Gauss-Newton refinement and the CRB design-loss
gradient, against the FP64 oracle (oracle/diff.c, itself pinned by
tests/test_oracle_diff.py) and FP64 central finite differences.  Tolerance:
1e-4 relative (FP32 recursion over up to 1000 TRs)."""
import numpy as np
import pytest

import oracle
import snapmrf_amd as sm
from snapmrf_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-4
THETAS = np.array([[900, 80, 7, 0.93], [300, 250, -31, 1.12], [2500, 30, 0, 0.85],
                   [1200, 60, 18.5, 1.0], [150, 40, -2, 0.8], [4000, 1500, 33, 1.2]], np.float32)


def _raw(T, seed=3):
    return synth.csv_columns(synth.sequence(T, seed=seed), "epg_ssfp").astype(np.float32)


def _jac(torch, raw, theta, w):
    n, T = len(theta), raw.shape[1]
    d_th = torch.from_numpy(np.ascontiguousarray(theta, np.float32)).cuda()
    s = torch.empty((n, T), dtype=torch.complex64, device="cuda")
    J = torch.empty((n, 4, T), dtype=torch.complex64, device="cuda")
    sm.epg_jacobian(raw, d_th, n, w, d_sig=s, d_jac=J)
    torch.cuda.synchronize()
    return s.cpu().numpy(), J.cpu().numpy()


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("T,w", [(137, 17), (1000, 30), (60, 2), (250, 64)])
def test_jacobian_matches_oracle(gpu, T, w):
    raw = _raw(T)
    s, J = _jac(gpu, raw, THETAS, w)
    for a, th in enumerate(THETAS):
        s64, J64 = oracle.dm_jacobian(raw, th.astype(np.float64), w)
        assert _rel(s[a], s64) < TOL, a
        for d in range(4):
            assert _rel(J[a, d], J64[d]) < TOL, (a, d)


def test_jacobian_central_differences(gpu):
    """GPU dS/dtheta against FP64 central differences (h = 1e-4 relative) of
    the oracle forward, T = 1000."""
    raw = _raw(1000, seed=8)
    w = 30
    _, J = _jac(gpu, raw, THETAS[:3], w)
    for a, th in enumerate(THETAS[:3].astype(np.float64)):
        for d in range(4):
            h = 1e-4 * (abs(th[d]) if th[d] != 0 else 1.0)
            tp, tm = th.copy(), th.copy()
            tp[d] += h
            tm[d] -= h
            fd = (oracle.dm_jacobian(raw, tp, w, False) - oracle.dm_jacobian(raw, tm, w, False)) / (2 * h)
            assert _rel(J[a, d], fd) < TOL, (a, d)


@pytest.mark.parametrize("T,w", [(137, 17), (50, 8), (49, 8), (3, 2), (1000, 30), (300, 32),
                                 (300, 33), (201, 64), (8192, 4), (8193, 4)])
def test_adjoint_matches_oracle(gpu, T, w):
    """Theta and sequence gradients of L = sum_atoms Re(v^H s).  Checkpoints
    every 4 TRs: T = 137, 49 and 201 end on a one-TR segment, 50 and 8193 on
    two TRs, 3 is a single partial segment.  w <= 32 puts an atom on 8
    lanes (w = 32 fills them), w = 33 .. 64 on 16; the six atoms leave lanes
    of the wave idle, which run on the last atom and store nothing."""
    raw = _raw(T, seed=4)
    n = len(THETAS)
    rng = np.random.default_rng(T)
    v = (rng.normal(size=(n, T)) + 1j * rng.normal(size=(n, T))).astype(np.complex64)
    d_th = gpu.from_numpy(THETAS).cuda()
    d_v = gpu.from_numpy(v).cuda()
    g_th = gpu.empty((n, 4), device="cuda")
    g_seq = gpu.zeros((2, T), device="cuda")
    sm.epg_adjoint(raw, d_th, n, w, d_v, d_grad_theta=g_th, d_grad_seq=g_seq)
    gpu.cuda.synchronize()
    g_th, g_seq = g_th.cpu().numpy(), g_seq.cpu().numpy()
    sfa = np.zeros(T)
    str_ = np.zeros(T)
    for a, th in enumerate(THETAS.astype(np.float64)):
        g, gfa, gtr = oracle.dm_adjoint(raw, th, w, v[a].astype(np.complex128))
        assert np.all(np.abs(g_th[a] - g) <= TOL * np.abs(g) + 1e-6 * np.abs(g).max()), (a, g_th[a], g)
        sfa += gfa
        str_ += gtr
    assert np.abs(g_seq[0] - sfa).max() <= TOL * np.abs(sfa).max()
    assert np.abs(g_seq[1] - str_).max() <= TOL * np.abs(str_).max()


@pytest.mark.parametrize("w", [30, 40])
def test_adjoint_many_atoms(gpu, w):
    """333 random atoms over many work-groups (the last one partly idle):
    every atom's theta gradient and the summed sequence gradients against
    the FP64 oracle.  Random cotangents make some gradients sums of terms
    that cancel (down to 5e-4 of sum_t |v_t| |ds_t/dtheta_d| here), so the
    bound adds 1e-6 of that unsigned sum (from the oracle Jacobian) to the
    1e-4 relative: FP32 rounding of the terms stays near 1e-7 of it."""
    T, n = 120, 333
    raw = _raw(T, seed=12)
    rng = np.random.default_rng(w)
    th = np.stack([rng.uniform(200, 3000, n), rng.uniform(20, 400, n), rng.uniform(-40, 40, n),
                   rng.uniform(0.8, 1.2, n)], 1).astype(np.float32)
    v = (rng.normal(size=(n, T)) + 1j * rng.normal(size=(n, T))).astype(np.complex64)
    g_th = gpu.empty((n, 4), device="cuda")
    g_seq = gpu.zeros((2, T), device="cuda")
    sm.epg_adjoint(raw, gpu.from_numpy(th).cuda(), n, w, gpu.from_numpy(v).cuda(),
                   d_grad_theta=g_th, d_grad_seq=g_seq)
    gpu.cuda.synchronize()
    g_th, g_seq = g_th.cpu().numpy(), g_seq.cpu().numpy()
    sfa, str_ = np.zeros(T), np.zeros(T)
    for a in range(n):
        t64 = th[a].astype(np.float64)
        g, gfa, gtr = oracle.dm_adjoint(raw, t64, w, v[a].astype(np.complex128))
        _, J = oracle.dm_jacobian(raw, t64, w)
        scale = np.abs(J) @ np.abs(v[a]).astype(np.float64)
        assert np.all(np.abs(g_th[a] - g) <= TOL * np.abs(g) + 1e-6 * scale), (a, g_th[a], g)
        sfa += gfa
        str_ += gtr
    assert np.abs(g_seq[0] - sfa).max() <= TOL * np.abs(sfa).max()
    assert np.abs(g_seq[1] - str_).max() <= TOL * np.abs(str_).max()


def test_adjoint_dot_product_on_gpu(gpu):
    """<J u, v> = <u, J^T v> between the GPU's own forward and reverse modes.
    With random cotangents a component can be a sum of cancelling terms (the
    B1 component of atom 1 is 1.8e-3 of sum_t |v_t| |ds_t/dB1|), so the bound
    adds 2e-6 of that unsigned sum (1e-6 per side) to the 1e-4 relative."""
    T, w = 300, 24
    raw = _raw(T, seed=9)
    n = len(THETAS)
    _, J = _jac(gpu, raw, THETAS, w)
    rng = np.random.default_rng(5)
    v = (rng.normal(size=(n, T)) + 1j * rng.normal(size=(n, T))).astype(np.complex64)
    g_th = gpu.empty((n, 4), device="cuda")
    sm.epg_adjoint(raw, gpu.from_numpy(THETAS).cuda(), n, w, gpu.from_numpy(v).cuda(),
                   d_grad_theta=g_th)
    gpu.cuda.synchronize()
    g_th = g_th.cpu().numpy()
    for a in range(n):
        want = np.array([np.real(np.vdot(v[a].astype(np.complex128), J[a, d])) for d in range(4)])
        scale = np.abs(J[a]).astype(np.float64) @ np.abs(v[a]).astype(np.float64)
        assert np.all(np.abs(g_th[a] - want) <= TOL * np.abs(want).max() + 2e-6 * scale), a


def test_adjoint_flip_angle_central_differences(gpu):
    """dL/dFA_raw and dL/dTR_raw against FP64 central differences of the
    oracle forward on 32 entries, T = 1000, w = 30."""
    T, w = 1000, 30
    raw = _raw(T, seed=11)
    th = THETAS[0].astype(np.float64)
    rng = np.random.default_rng(3)
    v = (rng.normal(size=T) + 1j * rng.normal(size=T)).astype(np.complex64)
    g_seq = gpu.zeros((2, T), device="cuda")
    sm.epg_adjoint(raw, gpu.from_numpy(THETAS[:1]).cuda(), 1, w, gpu.from_numpy(v[None]).cuda(),
                   d_grad_seq=g_seq)
    gpu.cuda.synchronize()
    g_seq = g_seq.cpu().numpy()
    vd = v.astype(np.complex128)

    def loss(r):
        return float(np.real(np.vdot(vd, oracle.dm_jacobian(r, th, w, False))))

    for col, row in ((0, 0), (2, 1)):
        scale = np.abs(g_seq[row]).max()
        for n in np.linspace(0, T - 2, 16).astype(int):
            r1, r2 = raw.copy(), raw.copy()
            h = 2e-3 * max(1.0, abs(float(raw[col, n])))
            r1[col, n] = np.float32(raw[col, n] + h)
            r2[col, n] = np.float32(raw[col, n] - h)
            fd = (loss(r1) - loss(r2)) / (float(r1[col, n]) - float(r2[col, n]))
            assert abs(g_seq[row, n] - fd) <= TOL * scale, (col, n)


@pytest.mark.parametrize("V", [64, 1024])
def test_gauss_newton_recovers_off_grid_truths(gpu, V):
    """Noise-free signals of off-grid truths: from a start one grid step away
    the refinement drives the cost to ~0 and recovers (T1, T2, B0, B1); the
    cost history never increases.  1,024 voxels fill 128 work-groups whose
    waves take different accept / reject paths (passes with and without
    tangents run side by side in one work-group's LDS)."""
    T, w, iters = 400, 24, 25
    raw = _raw(T, seed=6)
    rng = np.random.default_rng(7)
    truth = np.stack([rng.uniform(300, 2500, V), rng.uniform(30, 300, V),
                      rng.uniform(-30, 30, V), rng.uniform(0.85, 1.15, V)], 1).astype(np.float32)
    s, _ = _jac(gpu, raw, truth, w)
    rho = (rng.uniform(0.5, 2.0, V) * np.exp(1j * rng.uniform(0, 2 * np.pi, V))).astype(np.complex64)
    x = (s * rho[:, None]).astype(np.complex64)
    start = truth * np.array([1.06, 0.94, 1.0, 1.0], np.float32) + np.array([0, 0, 1.5, 0.04], np.float32)
    d_th = gpu.from_numpy(start.copy()).cuda()
    d_pd = gpu.empty(V, device="cuda")
    d_cost = gpu.empty((V, iters + 1), device="cuda")
    sm.gn_refine(raw, gpu.from_numpy(x).cuda(), V, w, iters, d_th, d_pd=d_pd, d_cost=d_cost)
    gpu.cuda.synchronize()
    th, pd, cost = d_th.cpu().numpy(), d_pd.cpu().numpy(), d_cost.cpu().numpy()
    assert np.all(np.diff(cost, axis=1) <= 0.0)
    xx = np.sum(np.abs(x) ** 2, axis=1)
    # the converged cost is |x|^2 - |s^H x|^2 / |s|^2 of two FP32 sums over T
    # terms of size |x|^2: its floor is rounding, about sqrt(T) eps |x|^2
    # (= 1.2e-6 |x|^2 here, and converged costs come out negative as often as
    # positive), so the bound is 1e-5 |x|^2, four orders below the start
    assert np.all(cost[:, -1] <= 1e-5 * xx)
    assert np.all(cost[:, 0] > 100 * cost[:, -1])
    err = np.abs(th - truth)
    assert np.all(err[:, 0] <= 1e-3 * truth[:, 0])
    assert np.all(err[:, 1] <= 1e-3 * truth[:, 1])
    # B0 of long-T2 voxels (T2 ~ 250-300 ms, |B0| ~ 15-25 Hz) is resolved to
    # about 0.02 Hz once the cost reaches its FP32 floor
    assert np.all(err[:, 2] <= 5e-2)
    assert np.all(err[:, 3] <= 1e-3)
    want_pd = np.abs(rho) * np.linalg.norm(s, axis=1)
    assert np.allclose(pd, want_pd, rtol=1e-3)


@pytest.mark.parametrize("w", [10, 16, 20, 30, 32, 33])
def test_gauss_newton_cost_is_that_of_the_returned_theta(gpu, w):
    """Noisy signals, three iterations: the last recorded cost and the proton
    density are those of the returned theta, recomputed from the Jacobian
    kernel's signal there in FP64 (|x|^2 - |s^H x|^2 / |s|^2 and
    |s^H x| / |s|).  The state counts cover each form of the trial pass: lane
    pairs of 8, 12 and 16 states with and without truncation, and w = 33 on
    the generic evaluation."""
    T, V, iters = 300, 200, 3
    raw = _raw(T, seed=13)
    rng = np.random.default_rng(w)
    truth = np.stack([rng.uniform(300, 2500, V), rng.uniform(30, 300, V),
                      rng.uniform(-30, 30, V), rng.uniform(0.85, 1.15, V)], 1).astype(np.float32)
    s, _ = _jac(gpu, raw, truth, w)
    scale = np.linalg.norm(s, axis=1, keepdims=True) / np.sqrt(T)
    noise = (rng.standard_normal((V, T)) + 1j * rng.standard_normal((V, T))) * 0.05 * scale
    x = (s + noise).astype(np.complex64)
    start = truth * np.array([1.04, 0.96, 1.0, 1.0], np.float32) + np.array([0, 0, 1.0, 0.02], np.float32)
    d_th = gpu.from_numpy(start.copy()).cuda()
    d_pd = gpu.empty(V, device="cuda")
    d_cost = gpu.empty((V, iters + 1), device="cuda")
    sm.gn_refine(raw, gpu.from_numpy(x).cuda(), V, w, iters, d_th, d_pd=d_pd, d_cost=d_cost)
    gpu.cuda.synchronize()
    th, pd, cost = d_th.cpu().numpy(), d_pd.cpu().numpy(), d_cost.cpu().numpy()
    assert np.all(np.diff(cost, axis=1) <= 0.0)
    assert np.all(cost[:, -1] < cost[:, 0])
    s1, _ = _jac(gpu, raw, th, w)
    s64, x64 = s1.astype(np.complex128), x.astype(np.complex128)
    sx = np.abs(np.sum(np.conj(s64) * x64, axis=1))
    ss = np.sum(np.abs(s64) ** 2, axis=1)
    xx = np.sum(np.abs(x64) ** 2, axis=1)
    want = xx - sx ** 2 / ss
    assert np.all(np.abs(cost[:, -1] - want) <= 1e-5 * xx + 1e-3 * want)
    assert np.allclose(pd, sx / np.sqrt(ss), rtol=1e-4)


def test_gauss_newton_edge_voxels(gpu):
    """A zero signal keeps its start with cost 0 and proton density 0; starts
    outside the domain (T1 < 0, T2 NaN, B1 = 0) come back unchanged with NaN
    cost and proton density; the other voxels' results are bit for bit those
    of a launch without the edge voxels."""
    T, w, iters, V = 200, 24, 4, 12
    raw = _raw(T, seed=6)
    rng = np.random.default_rng(9)
    truth = np.stack([rng.uniform(300, 2500, V), rng.uniform(30, 300, V),
                      rng.uniform(-30, 30, V), rng.uniform(0.85, 1.15, V)], 1).astype(np.float32)
    s, _ = _jac(gpu, raw, truth, w)
    x = (s * 0.8).astype(np.complex64)
    start = (truth * np.float32(1.03)).astype(np.float32)

    def run(xv, th0):
        n = len(th0)
        d_th = gpu.from_numpy(th0.copy()).cuda()
        d_pd = gpu.empty(n, device="cuda")
        d_cost = gpu.empty((n, iters + 1), device="cuda")
        sm.gn_refine(raw, gpu.from_numpy(xv).cuda(), n, w, iters, d_th, d_pd=d_pd, d_cost=d_cost)
        gpu.cuda.synchronize()
        return d_th.cpu().numpy(), d_pd.cpu().numpy(), d_cost.cpu().numpy()

    edge = [1, 4, 6, 9]
    keep = [i for i in range(V) if i not in edge]
    th_ref, pd_ref, cost_ref = run(np.ascontiguousarray(x[keep]), np.ascontiguousarray(start[keep]))
    x_e, start_e = x.copy(), start.copy()
    x_e[1] = 0
    start_e[4, 0] = -5.0
    start_e[6, 1] = np.nan
    start_e[9, 3] = 0.0
    th, pd, cost = run(x_e, start_e)
    assert np.array_equal(th[1], start_e[1]) and pd[1] == 0.0 and np.all(cost[1] == 0.0)
    for i in (4, 6, 9):
        assert np.array_equal(th[i], start_e[i], equal_nan=True)
        assert np.isnan(pd[i]) and np.all(np.isnan(cost[i]))
    assert np.array_equal(th[keep], th_ref)
    assert np.array_equal(pd[keep], pd_ref)
    assert np.array_equal(cost[keep], cost_ref)
    assert np.all(cost_ref[:, -1] < cost_ref[:, 0])


def test_gauss_newton_zero_voxels(gpu):
    """V = 0 returns without launching (empty signal and theta tensors)."""
    raw = _raw(50)
    d_sig = gpu.empty((0, 50), dtype=gpu.complex64, device="cuda")
    d_th = gpu.empty((0, 4), device="cuda")
    sm.gn_refine(raw, d_sig, 0, 16, 3, d_th)
    gpu.cuda.synchronize()


def test_crb_loss_and_gradient(gpu):
    """CRB loss per tissue against the FP64 oracle, and its flip-angle and TR
    gradients against FP64 central differences of the oracle loss on 32
    entries (64 tissues, w = 60, T = 300)."""
    T, w = 300, 60
    raw = _raw(T, seed=12)
    rng = np.random.default_rng(9)
    n = 64
    tis = np.stack([rng.uniform(300, 2500, n), rng.uniform(30, 300, n), rng.uniform(-30, 30, n),
                    rng.uniform(0.85, 1.15, n)], 1).astype(np.float32)
    wts = np.array([1.0, 1.0, 0.01, 0.1], np.float32)
    d_loss = gpu.empty(n, dtype=gpu.float64, device="cuda")
    g_seq = gpu.zeros((2, T), device="cuda")
    sm.crb_loss_grad(raw, gpu.from_numpy(tis).cuda(), n, w, wts, d_loss, d_grad_seq=g_seq)
    gpu.cuda.synchronize()
    loss, g_seq = d_loss.cpu().numpy(), g_seq.cpu().numpy()
    t64 = tis.astype(np.float64)
    w64 = wts.astype(np.float64)
    for i in range(0, n, 9):
        want = oracle.dm_crb(raw, t64[i:i + 1], w, w64)
        assert abs(loss[i] - want) <= 1e-3 * want, i
    total = oracle.dm_crb(raw, t64, w, w64)
    assert abs(loss.sum() - total) <= 1e-3 * total
    sub = t64[:8]   # FD over a tissue subset keeps the CPU cost small
    d_loss8 = gpu.empty(8, dtype=gpu.float64, device="cuda")
    g8 = gpu.zeros((2, T), device="cuda")
    sm.crb_loss_grad(raw, gpu.from_numpy(tis[:8]).cuda(), 8, w, wts, d_loss8, d_grad_seq=g8)
    gpu.cuda.synchronize()
    g8 = g8.cpu().numpy()
    for col, row, idx in ((0, 0, np.linspace(0, T - 1, 24).astype(int)),
                          (2, 1, np.linspace(0, T - 2, 8).astype(int))):
        scale = np.abs(g8[row]).max()
        for t in idx:
            r1, r2 = raw.copy(), raw.copy()
            h = 1e-3 * max(1.0, abs(float(raw[col, t])))
            r1[col, t] = np.float32(raw[col, t] + h)
            r2[col, t] = np.float32(raw[col, t] - h)
            fd = (oracle.dm_crb(r1, sub, w, w64) - oracle.dm_crb(r2, sub, w, w64)) / (
                float(r1[col, t]) - float(r2[col, t]))
            assert abs(g8[row, t] - fd) <= 1e-3 * scale, (col, t, g8[row, t], fd)


def test_autograd_function(gpu):
    """torch.autograd binding: forward = the C-ABI signal, backward = the
    adjoint, checked against the FP64 oracle adjoint."""
    from snapmrf_amd.autograd import epg_ssfp_signal
    torch = gpu
    T, w = 120, 20
    raw = _raw(T, seed=2)
    theta = torch.from_numpy(THETAS[:3].copy()).cuda().requires_grad_(True)
    fa = torch.from_numpy(raw[0].copy()).requires_grad_(True)
    ph = torch.from_numpy(raw[1].copy())
    tr = torch.from_numpy(raw[2].copy()).requires_grad_(True)
    te = torch.from_numpy(raw[3].copy())
    s = epg_ssfp_signal(theta, fa, tr, ph, te, w)
    rng = np.random.default_rng(1)
    v = (rng.normal(size=(3, T)) + 1j * rng.normal(size=(3, T))).astype(np.complex64)
    L = (torch.conj(torch.from_numpy(v).cuda()) * s).real.sum()
    L.backward()
    gfa_sum = np.zeros(T)
    for a, th in enumerate(THETAS[:3].astype(np.float64)):
        g, gfa, gtr = oracle.dm_adjoint(raw, th, w, v[a].astype(np.complex128))
        gt = theta.grad[a].cpu().numpy()
        assert np.all(np.abs(gt - g) <= TOL * np.abs(g) + 1e-6 * np.abs(g).max())
        gfa_sum += gfa
    assert np.abs(fa.grad.numpy() - gfa_sum).max() <= TOL * np.abs(gfa_sum).max()
    assert fa.grad.device.type == "cpu" and tr.grad is not None


def test_bad_arguments(gpu):
    raw = _raw(10)
    d_th = gpu.from_numpy(THETAS[:1]).cuda()
    with pytest.raises(RuntimeError, match="nstates"):
        sm.epg_jacobian(raw, d_th, 1, 1, d_sig=gpu.empty((1, 10), dtype=gpu.complex64, device="cuda"))
    with pytest.raises(RuntimeError, match="nstates"):
        sm.epg_jacobian(raw, d_th, 1, 65, d_sig=gpu.empty((1, 10), dtype=gpu.complex64, device="cuda"))
    with pytest.raises(RuntimeError, match="iters"):
        sm.gn_refine(raw, gpu.empty((1, 10), dtype=gpu.complex64, device="cuda"), 1, 8, -1, d_th)


def test_adjoint_and_crb_batches(gpu, monkeypatch):
    """A scratch budget of two atoms' worth (MRF_DIFF_SCRATCH_BYTES) splits a
    call into batches: theta gradients and CRB losses equal one launch bit for
    bit, sequence gradients to FP32 summation order."""
    T, w = 137, 17
    raw = _raw(T, seed=4)
    n = len(THETAS)
    rng = np.random.default_rng(21)
    v = (rng.normal(size=(n, T)) + 1j * rng.normal(size=(n, T))).astype(np.complex64)
    d_th = gpu.from_numpy(THETAS).cuda()
    d_v = gpu.from_numpy(v).cuda()
    wts = np.array([1.0, 1.0, 0.01, 0.1], np.float32)

    def run():
        g_th = gpu.empty((n, 4), device="cuda")
        g_seq = gpu.zeros((2, T), device="cuda")
        sm.epg_adjoint(raw, d_th, n, w, d_v, d_grad_theta=g_th, d_grad_seq=g_seq)
        loss = gpu.empty(n, dtype=gpu.float64, device="cuda")
        c_seq = gpu.zeros((2, T), device="cuda")
        sm.crb_loss_grad(raw, d_th, n, w, wts, loss, d_grad_seq=c_seq)
        gpu.cuda.synchronize()
        return [a.cpu().numpy() for a in (g_th, g_seq, loss, c_seq)]

    one = run()
    per = 4 * (96 * 8 * ((T + 3) // 4) + 2 * T)      # adjoint scratch bytes per atom
    monkeypatch.setenv("MRF_DIFF_SCRATCH_BYTES", str(2 * per))
    many = run()
    assert np.array_equal(one[0], many[0])
    assert np.array_equal(one[2], many[2])
    for a, b in ((one[1], many[1]), (one[3], many[3])):
        assert np.abs(a - b).max() <= 1e-5 * np.abs(a).max()


def test_gauss_newton_zero_iterations(gpu):
    """iters = 0: theta is returned unchanged, the proton density and the
    cost are those of the start, |s^H x| / |s| and |x|^2 - |s^H x|^2 / |s|^2
    with s the GPU's own unnormalised signal there."""
    T, w, V = 200, 16, 24
    raw = _raw(T, seed=6)
    rng = np.random.default_rng(8)
    truth = np.stack([rng.uniform(300, 2500, V), rng.uniform(30, 300, V),
                      rng.uniform(-30, 30, V), rng.uniform(0.85, 1.15, V)], 1).astype(np.float32)
    s_true, _ = _jac(gpu, raw, truth, w)
    x = (s_true * 0.7).astype(np.complex64)
    start = (truth * np.float32(1.05)).astype(np.float32)
    s0, _ = _jac(gpu, raw, start, w)
    d_th = gpu.from_numpy(start.copy()).cuda()
    d_pd = gpu.empty(V, device="cuda")
    d_cost = gpu.empty((V, 1), device="cuda")
    sm.gn_refine(raw, gpu.from_numpy(x).cuda(), V, w, 0, d_th, d_pd=d_pd, d_cost=d_cost)
    gpu.cuda.synchronize()
    assert np.array_equal(d_th.cpu().numpy(), start)
    s64, x64 = s0.astype(np.complex128), x.astype(np.complex128)
    sx = np.abs(np.sum(np.conj(s64) * x64, axis=1))
    ss = np.sum(np.abs(s64) ** 2, axis=1)
    xx = np.sum(np.abs(x64) ** 2, axis=1)
    assert np.allclose(d_pd.cpu().numpy(), sx / np.sqrt(ss), rtol=1e-5)
    assert np.all(np.abs(d_cost.cpu().numpy()[:, 0] - (xx - sx ** 2 / ss)) <= 1e-5 * xx)


def _raw_phase(T, kind, seed=17):
    """Sequences with RF phase: 'ramp' (0 -> 137 degrees plus noise, every TR
    different) or 'const' (30 degrees throughout)."""
    sq = synth.sequence(T, seed=seed)
    rng = np.random.default_rng(seed)
    if kind == "ramp":
        sq["phi_deg"] = np.linspace(0.0, 137.0, T) + rng.uniform(-20.0, 20.0, T)
    else:
        sq["phi_deg"] = np.full(T, 30.0)
    return synth.csv_columns(sq, "epg_ssfp").astype(np.float32)


@pytest.mark.parametrize("kind", ["ramp", "const"])
def test_rf_phase_jacobian_and_adjoint_match_oracle(gpu, kind):
    """The kernels run each TR in the frame of its RF phase (zero-phase RF,
    frame changes folded into the TR - TE precession, readouts and
    cotangents rotated by e^{+/- i phi_n}, the initial state by e^{-/+ i phi_0}):
    signals, Jacobians, theta and sequence adjoints against the FP64 oracle,
    which applies the RF with its phase as written (src/functions.cu:483-528)."""
    T, w = 300, 24
    raw = _raw_phase(T, kind)
    s, J = _jac(gpu, raw, THETAS, w)
    n = len(THETAS)
    rng = np.random.default_rng(21)
    v = (rng.normal(size=(n, T)) + 1j * rng.normal(size=(n, T))).astype(np.complex64)
    g_th = gpu.empty((n, 4), device="cuda")
    g_seq = gpu.zeros((2, T), device="cuda")
    sm.epg_adjoint(raw, gpu.from_numpy(THETAS).cuda(), n, w, gpu.from_numpy(v).cuda(),
                   d_grad_theta=g_th, d_grad_seq=g_seq)
    gpu.cuda.synchronize()
    g_th, g_seq = g_th.cpu().numpy(), g_seq.cpu().numpy()
    sfa, str_ = np.zeros(T), np.zeros(T)
    for a, th in enumerate(THETAS.astype(np.float64)):
        s64, J64 = oracle.dm_jacobian(raw, th, w)
        assert _rel(s[a], s64) < TOL, a
        for d in range(4):
            assert _rel(J[a, d], J64[d]) < TOL, (a, d)
        g, gfa, gtr = oracle.dm_adjoint(raw, th, w, v[a].astype(np.complex128))
        assert np.all(np.abs(g_th[a] - g) <= TOL * np.abs(g) + 1e-6 * np.abs(g).max()), (a, g_th[a], g)
        sfa += gfa
        str_ += gtr
    assert np.abs(g_seq[0] - sfa).max() <= TOL * np.abs(sfa).max()
    assert np.abs(g_seq[1] - str_).max() <= TOL * np.abs(str_).max()


def test_rf_phase_crb_and_gauss_newton(gpu):
    """CRB loss with a varying RF phase against the FP64 oracle, and
    Gauss-Newton on noise-free signals of that sequence recovering off-grid
    truths (its sums use readouts in true coordinates)."""
    T, w = 300, 24
    raw = _raw_phase(T, "ramp", seed=23)
    rng = np.random.default_rng(25)
    n = 16
    tis = np.stack([rng.uniform(300, 2500, n), rng.uniform(30, 300, n), rng.uniform(-30, 30, n),
                    rng.uniform(0.85, 1.15, n)], 1).astype(np.float32)
    wts = np.array([1.0, 1.0, 0.01, 0.1], np.float32)
    d_loss = gpu.empty(n, dtype=gpu.float64, device="cuda")
    sm.crb_loss_grad(raw, gpu.from_numpy(tis).cuda(), n, w, wts, d_loss)
    gpu.cuda.synchronize()
    loss = d_loss.cpu().numpy()
    for i in range(n):
        want = oracle.dm_crb(raw, tis[i:i + 1].astype(np.float64), w, wts.astype(np.float64))
        assert abs(loss[i] - want) <= 1e-3 * want, i
    V, iters = 64, 25
    truth = tis[rng.integers(0, n, V)]
    s, _ = _jac(gpu, raw, truth, w)
    x = (s * 0.7).astype(np.complex64)
    start = truth * np.array([1.05, 0.95, 1.0, 1.0], np.float32) + np.array([0, 0, 1.0, 0.03], np.float32)
    d_th = gpu.from_numpy(start.copy()).cuda()
    d_cost = gpu.empty((V, iters + 1), device="cuda")
    sm.gn_refine(raw, gpu.from_numpy(x).cuda(), V, w, iters, d_th, d_cost=d_cost)
    gpu.cuda.synchronize()
    th, cost = d_th.cpu().numpy(), d_cost.cpu().numpy()
    assert np.all(np.diff(cost, axis=1) <= 0.0)
    xx = np.sum(np.abs(x) ** 2, axis=1)
    assert np.all(cost[:, -1] <= 1e-5 * xx)
    err = np.abs(th - truth)
    assert np.all(err[:, 0] <= 1e-3 * truth[:, 0]) and np.all(err[:, 1] <= 1e-3 * truth[:, 1])
    assert np.all(err[:, 3] <= 1e-3)
