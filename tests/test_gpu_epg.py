"""GPU parity of the dictionary kernels (through the C ABI) against the
reference fixtures and the FP64 oracle.  Tolerance: 1e-5 of each atom's
infinity norm (the reference's own fixtures carry FP32 noise of ~5e-6)."""
import numpy as np
import pytest

import oracle
import ra_util
import snapmrf_amd as sm
from snapmrf_amd import synth
from helpers import const_seq, fixture_params, mrf5_raw, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-5

SIMPLE = {
    "epg_se": ((np.pi, 0.0, 1000.0, 25.0, 1), (600, 100)),
    "epg_sr": ((np.pi / 3, np.pi / 2, 500.0, 1.0, 10), (600, 100)),
    "epg_tse1": ((2 * np.pi / 3, 0.0, 0.0, 0.0, 3), (1000, 100)),
    "epg_tse2": ((np.pi, 0.0, 50.0, 25.0, 10), (600, 100)),
    "epg_fisp": ((np.pi / 6, 0.0, 10.0, 5.0, 3), (1000, 100)),
}


def gen(torch, echo, raw, params, natoms, w, T=None, **kw):
    """Runs the dictionary on the GPU; returns complex128 numpy [A, T]."""
    T = raw.shape[1] if T is None else T
    A = params.shape[0]
    d_params = torch.from_numpy(np.ascontiguousarray(params)).cuda()
    d_atoms = torch.zeros((A, T), dtype=torch.complex64, device="cuda")
    if kw:
        sm.dict_ex(echo, raw, T, d_params, A, natoms, w, d_atoms, **kw)
        torch.cuda.synchronize()
    else:
        sm.MRF_dict(d_atoms, d_params, raw, T, raw.shape[1], A, natoms, 4, w, echo)
    return d_atoms.cpu().numpy().astype(np.complex128)


@pytest.mark.parametrize("echo", sorted(SIMPLE))
def test_reference_unit_tests(gpu, golden, echo):
    """src/test.cu:43-156 with nstates = 101 (4 lanes per atom)."""
    (a, p, tr, te, T), (t1, t2) = SIMPLE[echo]
    raw = const_seq(T, a, p, tr, te)
    params = np.array([[t1, t2, 0, 1]], np.float32)
    got = gen(gpu, echo, raw, params, 1, 101)
    gt = ra_util.read(f"{golden}/atoms_{echo}.ra")
    assert np.abs(got - gt).max() < 1e-6
    ref = oracle.epg(echo, oracle.seq_prepare(echo, raw), params, 1, 101)
    assert rel_err(got, ref) < TOL


@pytest.mark.parametrize("echo", ["epg_ssfp", "roa_ssfp"])
def test_reference_ssfp_unit_tests(gpu, golden, echo):
    raw = mrf5_raw(golden)
    params = fixture_params()
    got = gen(gpu, echo, raw, params, 4, 101)
    gt = ra_util.read(f"{golden}/atoms_{echo}.ra")
    assert rel_err(got, gt) < TOL
    seq = oracle.seq_prepare(echo, raw)
    ref = oracle.epg(echo, seq, params, 4, 101) if echo == "epg_ssfp" else oracle.roa(seq, params, 4)
    assert rel_err(got, ref) < TOL


def test_fisp_config1_vs_oracle(gpu):
    """FISP, T1 100:20:2000, T2 20:10:300 (2,663 atoms), T = 1000, w = 20."""
    s = synth.sequence(1000, seed=1)
    raw = synth.csv_columns(s, "epg_fisp")
    params, A, natoms = sm.build_params("100:20:2000", "20:10:300", "0", "1")
    got = gen(gpu, "epg_fisp", raw, params, natoms, 20)
    sub = np.arange(0, A, 7)
    ref = oracle.epg("epg_fisp", oracle.seq_prepare("epg_fisp", raw), params[sub], len(sub), 20)
    assert rel_err(got[sub], ref) < TOL


@pytest.mark.parametrize("w", [30, 60, 101, 200])
def test_ssfp_b0_b1_grid_vs_oracle(gpu, w):
    """B1-sliced SSFP with off-resonance on 1,530 atoms.  A launch this small
    is spread over more lanes per atom (until it has 2^17 lanes, at least 8
    states per lane): w = 30 runs 4 lanes of 8 states, 60 runs 8 x 8, 101
    and 200 run 16 x 8 and 16 x 15 (padded, truncated at w - 1).  The
    one-lane form is tested by test_lane_forms_vs_oracle and at full size by
    test_full_grid_one_lane_form_vs_oracle."""
    T = 400 if w > 60 else 1000
    s = synth.sequence(T, seed=2)
    raw = synth.csv_columns(s, "epg_ssfp")
    params, A, natoms = sm.build_params("100:300:2000+2300:900:5000", "20:40:300+400:800:2000",
                                        "-40:20:40", "0.8:0.2:1.2")
    got = gen(gpu, "epg_ssfp", raw, params, natoms, w)
    ref = oracle.epg("epg_ssfp", oracle.seq_prepare("epg_ssfp", raw), params, natoms, w)
    assert rel_err(got, ref) < TOL
    assert np.allclose(np.linalg.norm(got, axis=1), 1.0, atol=1e-5)


def test_fused_and_reference_order_paths_agree(gpu):
    s = synth.sequence(300, seed=4)
    s["phi_deg"] = np.linspace(0, 90, 300)     # varying RF phase exercises H's phase ratio
    for echo in ("epg_ssfp", "epg_fisp"):
        raw = synth.csv_columns(s, echo)
        params, A, natoms = sm.build_params("200:400:2000", "30:60:300", "-25:25:25", "0.9:0.2:1.1")
        if echo == "epg_fisp":
            params, A, natoms = params[:natoms], natoms, natoms
        fused = gen(gpu, echo, raw, params, natoms, 24, normalize=0)
        plain = gen(gpu, echo, raw, params, natoms, 24, normalize=0, generic=True)
        assert rel_err(fused, plain) < TOL
        ref = oracle.epg(echo, oracle.seq_prepare(echo, raw), params, natoms, 24)
        if echo == "epg_ssfp":
            plain = plain / np.linalg.norm(plain, axis=1, keepdims=True)
        assert rel_err(plain, ref) < TOL


def test_roa_isochromats_vs_oracle(gpu):
    s = synth.sequence(500, seed=5)
    raw = synth.csv_columns(s, "roa_ssfp")
    params, A, natoms = sm.build_params("200:300:2000", "30:50:300", "-30:15:30", "1")
    seq = oracle.seq_prepare("roa_ssfp", raw)
    # 13: a partial lane group; 230: more isochromats than one pass holds;
    # T = 500 ends on a partial group of TRs
    for n_iso, spread in ((1, 0.0), (100, 60.0), (13, 25.0), (230, 80.0)):
        got = gen(gpu, "roa_ssfp", raw, params, natoms, 2, n_iso=n_iso, iso_spread_hz=spread)
        ref = oracle.roa(seq, params, natoms, n_iso=n_iso, spread_hz=spread)
        assert rel_err(got, ref) < TOL, (n_iso, spread)


def test_ssfp_b0_conjugate_symmetry_full_size(gpu):
    """Size-independent property at T = 1000, w = 30 on the full T1/T2 axes of
    the 871,455-atom grid (one B1 slice, B0 = +-40 Hz)."""
    s = synth.sequence(1000, seed=0)
    raw = synth.csv_columns(s, "epg_ssfp")
    params, A, natoms = sm.build_params("100:20:2000+2300:300:5000",
                                        "20:5:100+110:10:300+400:200:2000", "-40:80:40", "1")
    got = gen(gpu, "epg_ssfp", raw, params, natoms, 30)
    half = A // 2
    assert np.abs(got[:half] + np.conj(got[half:])).max() < 1e-6
    assert np.allclose(np.linalg.norm(got, axis=1), 1.0, atol=1e-5)


def test_bad_arguments_raise(gpu):
    d_params = gpu.zeros((4, 4), device="cuda")
    d_atoms = gpu.zeros((4, 5), dtype=gpu.complex64, device="cuda")
    raw = np.zeros((4, 5), np.float32)
    with pytest.raises(sm.MRFError):
        sm.MRF_dict(d_atoms, d_params, raw, 5, 5, 4, 4, 4, 101, "epg_nope")
    with pytest.raises(sm.MRFError):
        sm.MRF_dict(d_atoms, d_params, raw, 5, 5, 4, 3, 4, 101, "epg_ssfp")   # NATOMS % natoms
    with pytest.raises(sm.MRFError):
        sm.MRF_dict(d_atoms, d_params, raw, 5, 5, 4, 4, 4, 1, "epg_ssfp")     # w < 2


@pytest.mark.parametrize("echo,w,T", [("epg_ssfp", 17, 333), ("epg_fisp", 7, 101),
                                      ("epg_ssfp", 2, 9), ("epg_ssfp", 33, 64)])
def test_truncated_and_odd_lengths(gpu, echo, w, T):
    """w not equal to a compiled lane width (padded states, truncation at w-1)
    and odd T (the TR loop's single-step tail)."""
    s = synth.sequence(T, seed=w)
    raw = synth.csv_columns(s, echo)
    params, A, natoms = sm.build_params("150:350:2000", "25:45:300", "-35:35:35", "1")
    got = gen(gpu, echo, raw, params, natoms, w)
    ref = oracle.epg(echo, oracle.seq_prepare(echo, raw), params, natoms, w)
    assert rel_err(got, ref) < TOL


# 1,530 atoms in 3 B1 slices of 510: the last wave of each slice holds 62
# atoms, so the one-lane form's staged writer and in-kernel normalisation run
# a partial wave
LANE_GRID = ("100:300:2000+2300:900:5000", "20:40:300+400:800:2000", "-40:20:40", "0.8:0.2:1.2")


@pytest.mark.parametrize("echo,w,T,lanes,normalize", [
    ("epg_ssfp", 30, 1000, 1, -1),   # the full-size form: M = 30, in-kernel normalisation
    ("epg_ssfp", 30, 333, 1, 0),     # odd T, unnormalised (norms returned)
    ("epg_ssfp", 20, 999, 1, 1),     # M = 20, odd T
    ("epg_ssfp", 20, 1000, 1, 0),
    ("epg_ssfp", 17, 64, 1, -1),     # M = 20 padded: truncation at w - 1 in the one-lane form
    ("epg_ssfp", 7, 30, 1, -1),      # T below two 16-TR output segments
    ("epg_ssfp", 30, 5, 1, -1),      # T shorter than one segment
    ("epg_fisp", 20, 1000, 1, 0),    # constant-phase FISP branch
    ("epg_ssfp", 30, 500, 2, 1),     # lane pairs of 15 states, in-kernel normalisation
    ("epg_ssfp", 30, 400, 16, 1),    # 16 lanes of 4 states (64, truncated at w - 1)
    ("epg_ssfp", 20, 999, 4, 1),     # 4 lanes of 8 states (truncated), odd T, in-kernel normalisation
    ("epg_ssfp", 32, 1000, 2, 0),    # lane pairs of 16 states, norms returned
])
def test_lane_forms_vs_oracle(gpu, echo, w, T, lanes, normalize):
    """Each lane form of the fused kernel, forced through
    mrf_dict_opts.lanes_per_atom, against the FP64 oracle (1e-5 of each
    atom's infinity norm); normalised atoms have unit norm, and the norms
    returned in d_norm2 are those of the unnormalised rows."""
    s = synth.sequence(T, seed=w + T)
    raw = synth.csv_columns(s, echo)
    params, A, natoms = sm.build_params(*LANE_GRID)
    if echo != "epg_ssfp":
        params, A = params[:natoms], natoms
    d_params = gpu.from_numpy(params).cuda()
    d_atoms = gpu.zeros((A, T), dtype=gpu.complex64, device="cuda")
    d_n2 = gpu.zeros(A, dtype=gpu.float64, device="cuda")
    sm.dict_ex(echo, raw, T, d_params, A, natoms, w, d_atoms, d_norm2=d_n2, normalize=normalize,
               lanes=lanes)
    gpu.cuda.synchronize()
    got = d_atoms.cpu().numpy().astype(np.complex128)
    n2 = d_n2.cpu().numpy()
    ref = oracle.epg(echo, oracle.seq_prepare(echo, raw), params, natoms, w)
    norms = np.linalg.norm(got, axis=1)
    if echo == "epg_ssfp" and normalize != 0:
        assert np.allclose(norms, 1.0, atol=1e-5)
        assert rel_err(got, ref) < TOL
    else:
        assert np.allclose(n2, norms ** 2, rtol=1e-5)
        unit = got / norms[:, None] if echo == "epg_ssfp" else got
        assert rel_err(unit, ref) < TOL


@pytest.mark.parametrize("groups,T", [(1, 1000), (3, 131), (13, 57), (64, 37)])
def test_one_lane_form_group_counts(gpu, groups, T):
    """The one-lane form forms each TR's T2/B0 factors once per wave group of
    equal (T2, B0), a chunk of TRs at a time (64 TRs for up to 4 groups, 16
    for 13, 4 for 64): waves whose 64 atoms hold 1, 3, 13 and 64 distinct
    (T2, B0) pairs (random T1 in each) against the FP64 oracle, odd T included."""
    rng = np.random.default_rng(groups)
    A = 640
    key = (np.arange(A) % 64) * groups // 64 + 1000 * (np.arange(A) // 64)
    _, inv = np.unique(key, return_inverse=True)
    t2 = rng.uniform(20, 300, inv.max() + 1)[inv]
    b0 = rng.uniform(-40, 40, inv.max() + 1)[inv]
    params = np.stack([rng.uniform(200, 3000, A), t2, b0, np.ones(A)], 1).astype(np.float32)
    s = synth.sequence(T, seed=groups)
    raw = synth.csv_columns(s, "epg_ssfp")
    got = gen(gpu, "epg_ssfp", raw, params, A, 30, lanes=1)
    ref = oracle.epg("epg_ssfp", oracle.seq_prepare("epg_ssfp", raw), params, A, 30)
    assert rel_err(got, ref) < TOL


def test_lanes_per_atom_rejects_bad_values(gpu):
    raw = synth.csv_columns(synth.sequence(20, seed=0), "epg_ssfp")
    params, A, natoms = sm.build_params(*LANE_GRID)
    d_params = gpu.from_numpy(params).cuda()
    d_atoms = gpu.zeros((A, 20), dtype=gpu.complex64, device="cuda")
    for lanes, w in ((3, 30), (32, 30), (1, 200)):   # not a power of two, > 16, M > 32 at L = 1
        with pytest.raises(sm.MRFError):
            sm.dict_ex("epg_ssfp", raw, 20, d_params, A, natoms, w, d_atoms, lanes=lanes)


C2_GRID = ("100:20:2000+2300:300:5000", "20:5:100+110:10:300+400:200:2000", "-40:2:40",
           "0.8:0.1:1.2")


@pytest.mark.parametrize("w", [30, 20])
def test_full_grid_one_lane_form_vs_oracle(gpu, w):
    """The 871,455-atom SSFP grid (5 B1 slices of 174,291, B0 -40:2:40 Hz),
    T = 1000: a launch this size runs one atom per lane with the LDS-staged
    row writer and in-kernel normalisation.  It equals the forced one-lane
    launch bit for bit; every atom has unit norm; a seeded subset of 2,100
    atoms (420 grid points in each of the 5 slices) equals the FP64 oracle
    within 1e-5 of each atom's infinity norm."""
    import torch
    T = 1000
    raw = synth.csv_columns(synth.sequence(T, seed=0), "epg_ssfp")
    params, A, natoms = sm.build_params(*C2_GRID)
    assert (A, natoms) == (871455, 174291)
    d_params = torch.from_numpy(params).cuda()
    d_atoms = torch.empty((A, T), dtype=torch.complex64, device="cuda")
    sm.dict_ex("epg_ssfp", raw, T, d_params, A, natoms, w, d_atoms)
    d_one = torch.empty_like(d_atoms)
    sm.dict_ex("epg_ssfp", raw, T, d_params, A, natoms, w, d_one, lanes=1)
    torch.cuda.synchronize()
    assert torch.equal(d_atoms, d_one)
    del d_one
    norms = torch.linalg.vector_norm(d_atoms, dim=1)
    assert float((norms - 1.0).abs().max()) < 1e-5
    rng = np.random.default_rng(100 + w)
    pick = np.sort(rng.choice(natoms, 420, replace=False))
    rows = (np.arange(5)[:, None] * natoms + pick[None, :]).ravel()
    got = d_atoms[torch.from_numpy(rows).cuda()].cpu().numpy().astype(np.complex128)
    sub = params[rows]
    ref = oracle.epg("epg_ssfp", oracle.seq_prepare("epg_ssfp", raw), sub, len(pick), w)
    assert rel_err(got, ref) < TOL
