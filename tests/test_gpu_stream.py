"""Dictionary generation streamed through the matcher (mrf_dict_match_ex): the
dictionary exists one block of atoms per slice at a time.  FP32: the maps and
indices must equal generating the whole dictionary (mrf_dict_ex) and matching
it (mrf_match_ex) bit for bit, for any block size, with and without a B1 map;
every block is generated in the lane form of the whole dictionary.
Low-precision modes: the streamed pick must score at least as high in FP64 as
the whole-dictionary call's.  The whole-dictionary path is pinned to the
reference's fixtures and the FP64 oracle by test_gpu_epg.py and
test_gpu_match.py."""
import numpy as np
import pytest

import snapmrf_amd as sm
from snapmrf_amd import synth

pytestmark = pytest.mark.gpu

T, W = 200, 12
GRID = ("100:40:2000+2300:300:5000", "20:10:100+110:30:300+400:400:2000", "-30:6:30", "0.8:0.2:1.2")


def _inputs(torch, echo="epg_ssfp", grid=GRID, V=3000, seed=0):
    raw = synth.csv_columns(synth.sequence(T, seed=seed), echo)
    params, A, natoms = sm.build_params(*grid)
    d_params = torch.from_numpy(params).cuda()
    d_atoms = torch.empty((A, T), dtype=torch.complex64, device="cuda")
    sm.dict_ex(echo, raw, T, d_params, A, natoms, W, d_atoms)
    rng = np.random.default_rng(seed + 1)
    ns = A // natoms
    b1 = rng.uniform(0.75, 1.25, V).astype(np.float32)
    slice_b1 = params[::natoms, 3]
    ks = np.argmin(np.abs(b1[:, None] - slice_b1[None, :]), axis=1)
    truth = ks * natoms + rng.integers(0, natoms, V)
    sig = d_atoms[torch.from_numpy(truth).cuda()] * torch.from_numpy(
        rng.uniform(0.3, 1.0, V).astype(np.float32)).cuda()[:, None]
    g = torch.Generator(device="cuda").manual_seed(seed + 2)
    sig = (sig + 0.01 * torch.complex(torch.randn(V, T, device="cuda", generator=g),
                                      torch.randn(V, T, device="cuda", generator=g))).contiguous()
    torch.cuda.synchronize()
    return raw, params, d_params, A, natoms, d_atoms, sig, torch.from_numpy(b1).cuda(), ns


def _whole(torch, d_atoms, d_params, A, natoms, sig, d_b1):
    V = sig.shape[0]
    d_maps = torch.empty((V, 5), device="cuda")
    d_idx = torch.empty(V, dtype=torch.int32, device="cuda")
    sm.match_ex(sm.MRF_MATCH_FP32, d_atoms, d_params, T, A, natoms, sig, V, d_maps,
                d_b1map=d_b1, d_idx=d_idx)
    torch.cuda.synchronize()
    return d_maps.cpu().numpy(), d_idx.cpu().numpy()


def _streamed(torch, echo, raw, d_params, A, natoms, sig, d_b1, block):
    V = sig.shape[0]
    d_maps = torch.full((V, 5), -7.0, device="cuda")
    d_idx = torch.full((V,), -7, dtype=torch.int32, device="cuda")
    sm.dict_match_ex(echo, raw, T, d_params, A, natoms, W, sig, V, d_maps, d_b1map=d_b1,
                     d_idx=d_idx, block_atoms=block)
    torch.cuda.synchronize()
    return d_maps.cpu().numpy(), d_idx.cpu().numpy()


@pytest.mark.parametrize("use_b1", [True, False])
def test_streamed_equals_whole_dictionary(gpu, use_b1):
    raw, params, d_params, A, natoms, d_atoms, sig, d_b1, ns = _inputs(gpu)
    assert ns == 3 and natoms > 1000
    b1 = d_b1 if use_b1 else None
    m_ref, i_ref = _whole(gpu, d_atoms, d_params, A, natoms, sig, b1)
    # 128: many blocks; 300: blocks not a multiple of the 128-atom tile;
    # natoms: one block; 0: the automatic size
    for block in (128, 300, natoms, 0):
        m, i = _streamed(gpu, "epg_ssfp", raw, d_params, A, natoms, sig, b1, block)
        assert np.array_equal(i, i_ref), block
        assert np.array_equal(m, m_ref), block


def test_streamed_fisp_one_slice(gpu):
    grid = ("100:20:2000", "20:10:300", "0", "1")
    raw, params, d_params, A, natoms, d_atoms, sig, d_b1, ns = _inputs(gpu, "epg_fisp", grid, V=1500)
    m_ref, i_ref = _whole(gpu, d_atoms, d_params, A, natoms, sig, None)
    m, i = _streamed(gpu, "epg_fisp", raw, d_params, A, natoms, sig, None, 512)
    assert np.array_equal(i, i_ref) and np.array_equal(m, m_ref)


def test_streamed_rejects_partial_dictionaries(gpu):
    raw = synth.csv_columns(synth.sequence(T, seed=0), "roa_ssfp")
    params, A, natoms = sm.build_params(*GRID)
    d_params = gpu.from_numpy(params).cuda()
    sig = gpu.zeros((4, T), dtype=gpu.complex64, device="cuda")
    d_maps = gpu.empty((4, 5), device="cuda")
    # roa_ssfp fills one slice only, so a 3-slice dictionary cannot stream
    with pytest.raises(sm.MRFError):
        sm.dict_match_ex("roa_ssfp", raw, T, d_params, A, natoms, W, sig, 4, d_maps)


def _fp64_score(atoms, sig, idx):
    a = atoms[idx].astype(np.complex128)
    return np.abs(np.sum(np.conj(a) * sig.astype(np.complex128), axis=1))


@pytest.mark.parametrize("use_b1", [True, False])
def test_streamed_low_precision(gpu, use_b1):
    """Low-precision modes streamed block by block: each block's candidates
    are rescored against its own complex64 rows and merged into a running
    best.  A block's sub-tiles keep the atom chunks and lane lists they have
    in the whole-dictionary call, so every candidate that call rescores is
    among its block's, and on every voxel the streamed pick's FP64 score is
    at least the whole call's."""
    import oracle
    raw, params, d_params, A, natoms, d_atoms, sig, d_b1, ns = _inputs(gpu, V=2000)
    b1 = d_b1 if use_b1 else None
    atoms = d_atoms.cpu().numpy()
    s_np = sig.cpu().numpy()
    ld = sm.lowp_ld(T)
    d_hi = gpu.empty((A, ld), dtype=gpu.int32, device="cuda")
    d_lo = gpu.empty((A, ld), dtype=gpu.int32, device="cuda")
    sm.lowp_prepare(d_atoms, T, A, d_hi, d_lo)
    V = s_np.shape[0]
    for mode in (sm.MRF_MATCH_BF16_TOP16, sm.MRF_MATCH_BF16_TOP32, sm.MRF_MATCH_BF16_TOP64,
                 sm.MRF_MATCH_SPLIT2):
        d_maps = gpu.empty((V, 5), device="cuda")
        d_idx = gpu.empty(V, dtype=gpu.int32, device="cuda")
        sm.match_ex(mode, d_atoms, d_params, T, A, natoms, sig, V, d_maps, d_b1map=b1,
                    d_idx=d_idx, d_lowp=d_hi, d_lowp_lo=d_lo)
        gpu.cuda.synchronize()
        whole = d_idx.cpu().numpy()
        m, i = _streamed_mode(gpu, raw, d_params, A, natoms, sig, b1, 300, mode)
        assert np.mean(i == whole) >= 0.995, mode
        assert np.all(_fp64_score(atoms, s_np, i) >= _fp64_score(atoms, s_np, whole) * (1 - 1e-12))
        assert np.array_equal(m[:, :4], params[i])
        if use_b1:
            assert np.all(i // natoms == whole // natoms)     # the B1 map's slice
    # SPLIT2 streamed: the FP64 argmax wherever the top-2 gap is clear
    if not use_b1:
        ridx, rpd, rsec = oracle.match_slice(atoms, s_np)
        clear = (rpd - rsec) / rpd > 1e-5
        assert np.array_equal(i[clear], ridx[clear])


def _streamed_mode(torch, raw, d_params, A, natoms, sig, d_b1, block, mode):
    V = sig.shape[0]
    d_maps = torch.full((V, 5), -7.0, device="cuda")
    d_idx = torch.full((V,), -7, dtype=torch.int32, device="cuda")
    sm.dict_match_ex("epg_ssfp", raw, T, d_params, A, natoms, W, sig, V, d_maps, d_b1map=d_b1,
                     d_idx=d_idx, block_atoms=block, mode=mode)
    torch.cuda.synchronize()
    return d_maps.cpu().numpy(), d_idx.cpu().numpy()


@pytest.mark.parametrize("T_, V_, block", [(37, 5, 64), (16, 130, 1000), (101, 257, 300)])
def test_streamed_ragged(gpu, T_, V_, block):
    """Odd and tiny T, few voxels, blocks larger than a slice or not a
    multiple of the tile (the low-precision modes round them up to 128
    atoms), slices that are not a multiple of 128 atoms: FP32 streamed equals
    the whole dictionary; the low-precision modes' streamed picks score at
    least as high in FP64 as the whole-dictionary call's."""
    raw = synth.csv_columns(synth.sequence(T_, seed=9), "epg_ssfp")
    params, A, natoms = sm.build_params("100:100:2000", "20:20:300", "-20:10:20", "0.9:0.1:1.1")
    d_params = gpu.from_numpy(params).cuda()
    d_atoms = gpu.empty((A, T_), dtype=gpu.complex64, device="cuda")
    sm.dict_ex("epg_ssfp", raw, T_, d_params, A, natoms, 8, d_atoms)
    atoms = d_atoms.cpu().numpy()
    rng = np.random.default_rng(T_)
    pick = rng.integers(0, A, V_)
    s_np = (atoms[pick] * 0.7 + 0.02 * (rng.normal(size=(V_, T_)) + 1j * rng.normal(size=(V_, T_)))
            ).astype(np.complex64)
    sig = gpu.from_numpy(s_np).cuda()
    b1 = gpu.from_numpy(rng.uniform(0.85, 1.15, V_).astype(np.float32)).cuda()
    for d_b1 in (b1, None):
        d_maps = gpu.empty((V_, 5), device="cuda")
        d_idx = gpu.empty(V_, dtype=gpu.int32, device="cuda")
        sm.match_ex(sm.MRF_MATCH_FP32, d_atoms, d_params, T_, A, natoms, sig, V_, d_maps,
                    d_b1map=d_b1, d_idx=d_idx)
        gpu.cuda.synchronize()
        i32 = d_idx.cpu().numpy()
        m, i = _streamed_generic(gpu, raw, T_, d_params, A, natoms, sig, d_b1, block, 0)
        assert np.array_equal(i, i32) and np.array_equal(m, d_maps.cpu().numpy())
        ld = sm.lowp_ld(T_)
        d_hi = gpu.empty((A, ld), dtype=gpu.int32, device="cuda")
        d_lo = gpu.empty((A, ld), dtype=gpu.int32, device="cuda")
        sm.lowp_prepare(d_atoms, T_, A, d_hi, d_lo)
        for mode in (sm.MRF_MATCH_BF16_TOP32, sm.MRF_MATCH_BF16_TOP64, sm.MRF_MATCH_SPLIT3):
            sm.match_ex(mode, d_atoms, d_params, T_, A, natoms, sig, V_, d_maps, d_b1map=d_b1,
                        d_idx=d_idx, d_lowp=d_hi, d_lowp_lo=d_lo)
            gpu.cuda.synchronize()
            whole = d_idx.cpu().numpy()
            m, i = _streamed_generic(gpu, raw, T_, d_params, A, natoms, sig, d_b1, block, mode)
            assert np.all(_fp64_score(atoms, s_np, i) >= _fp64_score(atoms, s_np, whole) * (1 - 1e-12)), mode


def _streamed_generic(torch, raw, T_, d_params, A, natoms, sig, d_b1, block, mode):
    V = sig.shape[0]
    d_maps = torch.full((V, 5), -7.0, device="cuda")
    d_idx = torch.full((V,), -7, dtype=torch.int32, device="cuda")
    sm.dict_match_ex("epg_ssfp", raw, T_, d_params, A, natoms, 8, sig, V, d_maps, d_b1map=d_b1,
                     d_idx=d_idx, block_atoms=block, mode=mode)
    torch.cuda.synchronize()
    return d_maps.cpu().numpy(), d_idx.cpu().numpy()


def test_streamed_blocks_keep_the_whole_dictionarys_lane_form(gpu):
    """w = 30, T = 200 on 267,813 atoms (3 B1 slices of 89,271): the whole
    dictionary has more than 2^17 atoms and runs one atom per lane, blocks of
    50,000 atoms per slice leave a last block of 39,271 x 3 = 117,813 atoms,
    which on its own would be spread over lane pairs.  The blocks must take
    the whole dictionary's form, so the maps -- proton density included,
    which moves with any rounding difference in the atoms -- equal the whole
    dictionary's bit for bit.  The whole dictionary generated in the two-lane
    form instead gives different proton densities, which shows the comparison
    can see a change of form."""
    import torch
    Ts, Ws = 200, 30
    grid = ("100:20:2000+2300:300:5000", "20:5:100+110:10:300+400:200:2000", "-40:4:40",
            "0.8:0.2:1.2")
    raw = synth.csv_columns(synth.sequence(Ts, seed=3), "epg_ssfp")
    params, A, natoms = sm.build_params(*grid)
    assert (A, natoms) == (267813, 89271)
    d_params = torch.from_numpy(params).cuda()
    V = 4096
    rng = np.random.default_rng(33)
    slice_b1 = params[::natoms, 3]
    b1 = rng.uniform(0.75, 1.25, V).astype(np.float32)
    ks = np.argmin(np.abs(b1[:, None] - slice_b1[None, :]), axis=1)
    truth = ks * natoms + rng.integers(0, natoms, V)
    d_b1 = torch.from_numpy(b1).cuda()

    def whole(lanes):
        d_atoms = torch.empty((A, Ts), dtype=torch.complex64, device="cuda")
        sm.dict_ex("epg_ssfp", raw, Ts, d_params, A, natoms, Ws, d_atoms, lanes=lanes)
        return d_atoms
    d_atoms = whole(0)
    sig = d_atoms[torch.from_numpy(truth).cuda()] * 0.8
    g = torch.Generator(device="cuda").manual_seed(34)
    sig = (sig + 0.01 * torch.complex(torch.randn(V, Ts, device="cuda", generator=g),
                                      torch.randn(V, Ts, device="cuda", generator=g))).contiguous()

    def match_whole(atoms):
        d_maps = torch.empty((V, 5), device="cuda")
        d_idx = torch.empty(V, dtype=torch.int32, device="cuda")
        sm.match_ex(sm.MRF_MATCH_FP32, atoms, d_params, Ts, A, natoms, sig, V, d_maps,
                    d_b1map=d_b1, d_idx=d_idx)
        torch.cuda.synchronize()
        return d_maps.cpu().numpy(), d_idx.cpu().numpy()
    m_ref, i_ref = match_whole(d_atoms)
    d_maps = torch.full((V, 5), -7.0, device="cuda")
    d_idx = torch.full((V,), -7, dtype=torch.int32, device="cuda")
    sm.dict_match_ex("epg_ssfp", raw, Ts, d_params, A, natoms, Ws, sig, V, d_maps, d_b1map=d_b1,
                     d_idx=d_idx, block_atoms=50000)
    torch.cuda.synchronize()
    assert np.array_equal(d_idx.cpu().numpy(), i_ref)
    assert np.array_equal(d_maps.cpu().numpy(), m_ref)
    del d_atoms
    m2, _ = match_whole(whole(2))
    assert not np.array_equal(m2[:, 4], m_ref[:, 4])
