"""GPU parity of the fused matcher (through the C ABI) against the
reference's matching KAT and the FP64 oracle.

The reference FP32 matcher is cuBLAS CGEMM whose summation order is not
reproducible, so FP32 argmax parity is stated against the FP64 oracle:
indices must agree exactly on every voxel whose FP64 top-2 score gap exceeds
the FP32 error bound (relative gap > 1e-5 here); PD agrees to 1e-5 relative."""
import ctypes

import numpy as np
import pytest

import oracle
import ra_util
import snapmrf_amd as sm
from snapmrf_amd import synth

pytestmark = pytest.mark.gpu
GAP = 1e-5


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def match(torch, atoms, params, sig, natoms, b1map=None, mode=sm.MRF_MATCH_FP32):
    """atoms complex64 [A, T] (numpy), signals complex64 [V, T] -> (maps [V, 5], idx [V])."""
    A, T = atoms.shape
    V = sig.shape[0]
    d_atoms, d_params, d_sig = dev(torch, atoms), dev(torch, params), dev(torch, sig)
    d_maps = torch.zeros((V, 5), device="cuda")
    d_idx = torch.zeros(V, dtype=torch.int32, device="cuda")
    d_b1 = None if b1map is None else dev(torch, b1map.astype(np.float32))
    sm.match_ex(mode, d_atoms, d_params, T, A, natoms, d_sig, V, d_maps, d_b1map=d_b1, d_idx=d_idx)
    torch.cuda.synchronize()
    return d_maps.cpu().numpy(), d_idx.cpu().numpy()


def test_reference_matching_kat(gpu, golden):
    """unittest_matching (src/functions.cu:1720-1824) through the drop-in
    MRF_match with a host image."""
    atoms = ra_util.read(f"{golden}/atoms_mrf.ra")
    params = ra_util.read(f"{golden}/params_mrf.ra").copy()
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    idx = [libc.rand() % 109 for _ in range(10)]
    img = atoms[idx].copy()
    d_atoms, d_params = dev(gpu, atoms), dev(gpu, params)
    d_maps = gpu.zeros((10, 5), device="cuda")
    sm.MRF_match(d_maps, img, d_atoms, d_params, 100, 100, 109, 109, 10, 4, 5, None)
    maps = d_maps.cpu().numpy()
    assert np.abs(maps[:, :4] - params[idx]).sum() < 1e-6
    assert np.allclose(maps[:, 4], 1.0, atol=1e-5)


def _fisp_dict(torch, T=1000, seed=1):
    raw = synth.csv_columns(synth.sequence(T, seed), "epg_fisp")
    params, A, natoms = sm.build_params("100:20:2000", "20:10:300", "0", "1")
    d_params = dev(torch, params)
    d_atoms = torch.zeros((A, T), dtype=torch.complex64, device="cuda")
    sm.dict_ex("epg_fisp", raw, T, d_params, A, natoms, 20, d_atoms, normalize=1)
    torch.cuda.synchronize()
    return d_atoms.cpu().numpy(), params, natoms


def fp64_score(atoms, sig, idx):
    """|<a_idx, s>| in FP64 for each voxel's chosen atom."""
    a = atoms[idx].astype(np.complex128)
    return np.abs(np.sum(np.conj(a) * sig.astype(np.complex128), axis=1))


def _check_against_oracle(maps, idx, atoms, params, sig, base=0):
    """Exact index agreement on voxels whose FP64 top-2 gap exceeds the FP32
    bound; on every voxel the GPU's pick scores within that bound of the
    FP64 best.  Returns the number of near-tie voxels."""
    ref_idx, ref_pd, ref_sec = oracle.match_slice(atoms, sig)
    gap = (ref_pd - ref_sec) / np.maximum(ref_pd, 1e-30)
    clear = gap > GAP
    assert np.array_equal(idx[clear], ref_idx[clear] + base)
    assert np.array_equal(maps[clear, :4], params[ref_idx[clear] + base])
    picked = fp64_score(atoms, sig, idx - base)
    assert np.all(picked >= ref_pd * (1.0 - GAP))
    assert np.array_equal(maps[:, :4], params[idx])
    assert np.allclose(maps[:, 4], ref_pd, rtol=1e-5, atol=1e-7)
    return int((~clear).sum())


def test_config1_fp32_vs_fp64_oracle(gpu):
    """2,663 FISP atoms, 4,096 voxels at SNR 30."""
    atoms, params, natoms = _fisp_dict(gpu)
    sig, truth, pd = synth.signals_np(atoms, 4096, seed=7)
    maps, idx = match(gpu, atoms, params, sig, natoms)
    _check_against_oracle(maps, idx, atoms, params, sig)


@pytest.mark.parametrize("mode", [sm.MRF_MATCH_FP32, sm.MRF_MATCH_BF16_TOP32, sm.MRF_MATCH_SPLIT2])
def test_zero_voxels_is_a_no_op(gpu, mode):
    """V = 0 returns without launching (empty signal and map tensors)."""
    T, A = 37, 300
    rng = np.random.default_rng(5)
    atoms = (rng.normal(size=(A, T)) + 1j * rng.normal(size=(A, T))).astype(np.complex64)
    params = np.ones((A, 4), np.float32)
    d_atoms, d_params = dev(gpu, atoms), dev(gpu, params)
    d_hi = d_lo = None
    if mode != sm.MRF_MATCH_FP32:
        ld = sm.lowp_ld(T)
        d_hi = gpu.empty((A, ld), dtype=gpu.int32, device="cuda")
        d_lo = gpu.empty((A, ld), dtype=gpu.int32, device="cuda")
        sm.lowp_prepare(d_atoms, T, A, d_hi, d_lo)
    d_sig = gpu.empty((0, T), dtype=gpu.complex64, device="cuda")
    d_maps = gpu.empty((0, 5), device="cuda")
    sm.match_ex(mode, d_atoms, d_params, T, A, A, d_sig, 0, d_maps, d_lowp=d_hi, d_lowp_lo=d_lo)
    gpu.cuda.synchronize()


@pytest.mark.parametrize("T,V", [(1, 7), (3, 64), (15, 33), (37, 1), (37, 129), (16, 300), (1000, 257)])
def test_ragged_shapes(gpu, T, V):
    rng = np.random.default_rng(T * 1000 + V)
    A = 300
    atoms = (rng.normal(size=(A, T)) + 1j * rng.normal(size=(A, T))).astype(np.complex64)
    atoms /= np.linalg.norm(atoms, axis=1, keepdims=True)
    params = np.stack([np.arange(A) + 1000, np.arange(A) % 17 + 1, np.zeros(A), np.ones(A)], 1)
    params = params.astype(np.float32)
    sig, truth, pd = synth.signals_np(atoms, V, seed=3, snr=5.0)
    maps, idx = match(gpu, atoms, params, sig, A)
    _check_against_oracle(maps, idx, atoms, params, sig)


def test_ties_go_to_first_index(gpu):
    rng = np.random.default_rng(11)
    T, A = 64, 9000           # spans two atom chunks of the kernel
    atoms = (rng.normal(size=(A, T)) + 1j * rng.normal(size=(A, T))).astype(np.complex64)
    atoms[8500] = atoms[17]   # exact duplicates in different work-groups
    atoms[8999] = atoms[4100]
    atoms[130] = atoms[129]   # and inside one 128-atom sub-tile boundary
    params = np.zeros((A, 4), np.float32)
    params[:, 0] = np.arange(A)
    sig = np.stack([atoms[17], atoms[4100], atoms[129], np.zeros(T, np.complex64),
                    2.0 * atoms[8500]])
    maps, idx = match(gpu, atoms, params, sig, A)
    assert idx.tolist() == [17, 4100, 129, 0, 17]
    assert maps[3, 4] == 0.0


def _ssfp_dict(torch, T=300, seed=2):
    raw = synth.csv_columns(synth.sequence(T, seed), "epg_ssfp")
    params, A, natoms = sm.build_params("100:100:2000", "20:20:300", "-30:15:30", "0.8:0.1:1.2")
    d_params = dev(torch, params)
    d_atoms = torch.zeros((A, T), dtype=torch.complex64, device="cuda")
    sm.dict_ex("epg_ssfp", raw, T, d_params, A, natoms, 30, d_atoms)
    torch.cuda.synchronize()
    return d_atoms.cpu().numpy(), params, natoms


def test_b1_map_selection_vs_oracle(gpu):
    """merge_maps(b1map) semantics: each voxel takes the best atom of the
    slice whose B1 is nearest its map value (src/functions.cu:1351-1384)."""
    atoms, params, natoms = _ssfp_dict(gpu)
    l_b1 = atoms.shape[0] // natoms
    b1 = synth.b1_map(32, 40, seed=3).ravel()
    b1[:5] = [0.85, 0.95, 0.0, 7.0, 1.2]          # ties between slices, out-of-range values
    slice_b1 = params[::natoms, 3]
    ks = np.array([oracle.b1_slice(v, slice_b1) for v in b1])
    # each voxel's signal comes from its own slice
    rng = np.random.default_rng(5)
    pick = ks * natoms + rng.integers(0, natoms, len(b1))
    sig, _, _ = synth.signals_np(atoms, len(b1), seed=9, candidates=None)
    sig = (atoms[pick] * 0.7 + (sig - sig.mean()) * 0.01).astype(np.complex64)
    maps, idx = match(gpu, atoms, params, sig, natoms, b1map=b1)
    for k in range(l_b1):
        sel = ks == k
        if not sel.any():
            continue
        sl = slice(k * natoms, (k + 1) * natoms)
        _check_against_oracle(maps[sel], idx[sel], atoms[sl], params, sig[sel], base=k * natoms)


def test_pd_merge_is_global_first_index_argmax(gpu):
    """Without a B1 map: merge_maps(PD) over per-slice argmaxes
    (src/functions.cu:1386-1416) = global argmax, ties to the lowest index."""
    atoms, params, natoms = _ssfp_dict(gpu)
    sig, truth, pd = synth.signals_np(atoms, 1000, seed=4, snr=20.0)
    maps, idx = match(gpu, atoms, params, sig, natoms)
    _check_against_oracle(maps, idx, atoms, params, sig)


# --------------------------------------------------------------- BF16 ----

LOWP_MODES = (sm.MRF_MATCH_BF16_TOP8, sm.MRF_MATCH_BF16_TOP16, sm.MRF_MATCH_BF16_TOP32,
              sm.MRF_MATCH_BF16_TOP64, sm.MRF_MATCH_SPLIT2, sm.MRF_MATCH_SPLIT3)
SPLIT_MODES = (sm.MRF_MATCH_SPLIT2, sm.MRF_MATCH_SPLIT3)


def match_lowp(torch, atoms, params, sig, natoms, mode, b1map=None):
    """BF16 (top-8 + FP32 rescore) or split-BF16 matching through the C ABI."""
    A, T = atoms.shape
    d_atoms = dev(torch, atoms)
    ld = sm.lowp_ld(T)
    d_hi = torch.empty((A, ld), dtype=torch.int32, device="cuda")
    d_lo = torch.empty((A, ld), dtype=torch.int32, device="cuda") if mode in SPLIT_MODES else None
    sm.lowp_prepare(d_atoms, T, A, d_hi, d_lo)
    V = sig.shape[0]
    d_params, d_sig = dev(torch, params), dev(torch, sig)
    d_maps = torch.zeros((V, 5), device="cuda")
    d_idx = torch.zeros(V, dtype=torch.int32, device="cuda")
    d_b1 = None if b1map is None else dev(torch, b1map.astype(np.float32))
    sm.match_ex(mode, d_atoms, d_params, T, A, natoms, d_sig, V, d_maps, d_b1map=d_b1,
                d_idx=d_idx, d_lowp=d_hi, d_lowp_lo=d_lo)
    torch.cuda.synchronize()
    return d_maps.cpu().numpy(), d_idx.cpu().numpy()


def test_lowp_prepare_rounding(gpu):
    """hi = bf16 round-to-nearest-even of x, lo = bf16(x - hi): checked
    against numpy's bit arithmetic."""
    rng = np.random.default_rng(0)
    x = (rng.normal(size=(7, 33)) + 1j * rng.normal(size=(7, 33))).astype(np.complex64)
    x[0, :4] = [0, 1e-30, -2.5, 3.0000002]
    d = dev(gpu, x)
    ld = sm.lowp_ld(33)
    assert ld == 64 and sm.lowp_ld(32) == 32 and sm.lowp_ld(1000) == 1024
    hi = gpu.full((7, ld), -1, dtype=gpu.int32, device="cuda")
    lo = gpu.full((7, ld), -1, dtype=gpu.int32, device="cuda")
    sm.lowp_prepare(d, 33, 7, hi, lo)
    h = hi.cpu().numpy().view(np.uint32)
    l = lo.cpu().numpy().view(np.uint32)
    assert not h[:, 33:].any() and not l[:, 33:].any()   # zero pad
    h, l = h[:, :33], l[:, :33]

    def rne(f):
        u = f.view(np.uint32).astype(np.uint64)
        return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32)

    def tof(b):
        return (b.astype(np.uint32) << 16).view(np.float32)

    re = x.real.astype(np.float32)
    im = x.imag.astype(np.float32)
    assert np.array_equal(h & 0xFFFF, rne(re))
    assert np.array_equal(h >> 16, rne(im))
    assert np.array_equal(l & 0xFFFF, rne((re - tof(rne(re))).astype(np.float32)))
    # hi + lo reproduces x to ~2^-16 relative
    rec = tof(h & 0xFFFF) + tof(l & 0xFFFF)
    assert np.all(np.abs(rec - re) <= np.abs(re) * 2.0 ** -15 + 1e-38)


@pytest.mark.parametrize("mode", LOWP_MODES)
def test_lowp_agrees_with_fp32_and_oracle(gpu, mode):
    """B1-sliced SSFP dictionary, noisy voxels: the rescored argmax agrees
    with the FP32 matcher on >= 99.9% of voxels and equals the FP64 oracle on
    every voxel whose top-2 gap exceeds the FP32 bound (the top-8 recall)."""
    atoms, params, natoms = _ssfp_dict(gpu, T=400)
    sig, truth, pd = synth.signals_np(atoms, 3000, seed=12, snr=25.0)
    m32, i32 = match(gpu, atoms, params, sig, natoms)
    mlo, ilo = match_lowp(gpu, atoms, params, sig, natoms, mode)
    assert np.mean(ilo == i32) >= 0.999
    ref_idx, ref_pd, ref_sec = oracle.match_slice(atoms, sig)
    clear = (ref_pd - ref_sec) / ref_pd > GAP
    assert np.mean(ilo[clear] == ref_idx[clear]) >= 0.999
    picked = fp64_score(atoms, sig, ilo)
    assert np.mean(picked >= ref_pd * (1 - GAP)) >= 0.999
    assert np.array_equal(mlo[:, :4], params[ilo])
    ok = ilo == i32
    assert np.allclose(mlo[ok, 4], m32[ok, 4], rtol=2e-5)


def test_lowp_b1_map_mode(gpu):
    atoms, params, natoms = _ssfp_dict(gpu)
    b1 = synth.b1_map(30, 40, seed=8).ravel()
    slice_b1 = params[::natoms, 3]
    ks = np.array([oracle.b1_slice(v, slice_b1) for v in b1])
    rng = np.random.default_rng(2)
    pick = ks * natoms + rng.integers(0, natoms, len(b1))
    sig = (atoms[pick] * 0.6).astype(np.complex64)
    m32, i32 = match(gpu, atoms, params, sig, natoms, b1map=b1)
    mlo, ilo = match_lowp(gpu, atoms, params, sig, natoms, sm.MRF_MATCH_BF16_TOP8, b1map=b1)
    assert np.mean(ilo == i32) >= 0.999
    assert np.all(ilo // natoms == ks)


def test_bf16_top32_candidates_contain_top16(gpu):
    """BF16_TOP32 runs TOP16's 3-entry lane lists over TOP16's atom chunks
    split in two (here, at most 64 voxel tiles, TOP16 already takes the
    maximum of 32 chunks and TOP32 takes the same ones) and rescores the best
    32 (TOP16: 16), so every candidate TOP16 rescores is rescored by TOP32 and
    its FP64-rescored pick scores at least as high on every voxel; on a dense
    grid (FP64 top-2 gaps below the BF16 rounding) it finds the FP64 argmax
    at least as often as TOP16 and TOP8.  BF16_TOP64 runs the same lane lists
    over the same 32 atom chunks and rescores the best 64 of them, so the
    same two properties hold against TOP16."""
    T = 400
    raw = synth.csv_columns(synth.sequence(T, 4), "epg_ssfp")
    params, A, natoms = sm.build_params("300:10:900", "40:4:120", "-10:2:10", "1")
    d_atoms = gpu.zeros((A, T), dtype=gpu.complex64, device="cuda")
    sm.dict_ex("epg_ssfp", raw, T, dev(gpu, params), A, natoms, 30, d_atoms)
    atoms = d_atoms.cpu().numpy()
    sig, truth, pd = synth.signals_np(atoms, 2000, seed=21, snr=25.0)
    ref_idx, ref_pd, ref_sec = oracle.match_slice(atoms, sig)
    picks = {}
    for mode in (sm.MRF_MATCH_BF16_TOP8, sm.MRF_MATCH_BF16_TOP16, sm.MRF_MATCH_BF16_TOP32,
                 sm.MRF_MATCH_BF16_TOP64):
        picks[mode] = match_lowp(gpu, atoms, params, sig, natoms, mode)[1]
    right = {m: np.mean(i == ref_idx) for m, i in picks.items()}
    assert right[sm.MRF_MATCH_BF16_TOP32] >= right[sm.MRF_MATCH_BF16_TOP16] >= right[sm.MRF_MATCH_BF16_TOP8]
    assert right[sm.MRF_MATCH_BF16_TOP64] >= right[sm.MRF_MATCH_BF16_TOP16]
    s16 = fp64_score(atoms, sig, picks[sm.MRF_MATCH_BF16_TOP16])
    s32 = fp64_score(atoms, sig, picks[sm.MRF_MATCH_BF16_TOP32])
    s64 = fp64_score(atoms, sig, picks[sm.MRF_MATCH_BF16_TOP64])
    assert np.all(s32 >= s16 * (1 - 1e-12))
    assert np.all(s64 >= s16 * (1 - 1e-12))
    print("\nFP64 argmax found: " + ", ".join(f"mode {m} {r:.4f}" for m, r in right.items()))


def test_split_bf16_reference_kat(gpu, golden):
    """The near-degenerate matching fixture (max off-diagonal |<a_i,a_j>| =
    0.9999941) through the split-BF16 path: exact indices."""
    atoms = ra_util.read(f"{golden}/atoms_mrf.ra")
    params = ra_util.read(f"{golden}/params_mrf.ra").copy()
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    idx = [libc.rand() % 109 for _ in range(10)]
    maps, got = match_lowp(gpu, atoms, params, atoms[idx].copy(), 109, sm.MRF_MATCH_SPLIT3)
    assert got.tolist() == idx
    # plain BF16 cannot rank this fixture: more than 8 atoms lie within BF16
    # rounding (~4e-4 relative at T = 100) of each voxel's own atom, so the
    # top-8 may miss it; its pick is still one of those near-duplicates
    sig = atoms[idx].copy()
    maps, got16 = match_lowp(gpu, atoms, params, sig, 109, sm.MRF_MATCH_BF16_TOP8)
    assert np.all(fp64_score(atoms, sig, got16) >= 0.999)


@pytest.mark.parametrize("T,V,A", [(3, 7, 300), (15, 33, 500), (37, 5, 300), (16, 130, 1000), (1000, 257, 700)])
def test_lowp_ragged(gpu, T, V, A):
    rng = np.random.default_rng(T + V)
    atoms = (rng.normal(size=(A, T)) + 1j * rng.normal(size=(A, T))).astype(np.complex64)
    atoms /= np.linalg.norm(atoms, axis=1, keepdims=True)
    params = np.stack([np.arange(A), np.zeros(A), np.zeros(A), np.ones(A)], 1).astype(np.float32)
    sig, truth, pd = synth.signals_np(atoms, V, seed=5, snr=8.0)
    for mode in LOWP_MODES:
        maps, idx = match_lowp(gpu, atoms, params, sig, A, mode)
        _check_against_oracle(maps, idx, atoms, params, sig)
