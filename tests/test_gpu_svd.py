"""Rank-K (SVD) compressed matching (matlab/MRF_dict_B1.m:125-159, McGivney
et al. 2014) on the GPU, pinned by LAPACK: numpy's FP64 SVD of the same
dictionary fixes the singular values, the subspace and the MATLAB rank rule;
an FP64 restatement of the compressed matching (projection with the GPU basis,
first-index argmax of |<C a, U^H s>|) fixes the matches on voxels whose
top-2 gap is clear."""
import numpy as np
import pytest

import oracle
import snapmrf_amd as sm
from snapmrf_amd import synth

pytestmark = pytest.mark.gpu

T, W, KMAX = 300, 12, 40
GRID = ("100:60:2000+2300:300:5000", "20:10:100+110:40:300+400:400:2000", "-30:10:30", "0.8:0.2:1.2")


@pytest.fixture(scope="module")
def problem(gpu):
    torch = gpu
    raw = synth.csv_columns(synth.sequence(T, seed=3), "epg_ssfp")
    params, A, natoms = sm.build_params(*GRID)
    d_params = torch.from_numpy(params).cuda()
    d_atoms = torch.empty((A, T), dtype=torch.complex64, device="cuda")
    sm.dict_ex("epg_ssfp", raw, T, d_params, A, natoms, W, d_atoms)
    atoms = d_atoms.cpu().numpy()
    rng = np.random.default_rng(5)
    V = 2500
    b1 = rng.uniform(0.75, 1.25, V).astype(np.float32)
    slice_b1 = params[::natoms, 3]
    ks = np.array([oracle.b1_slice(float(v), slice_b1) for v in b1])
    truth = ks * natoms + rng.integers(0, natoms, V)
    sig = atoms[truth] * rng.uniform(0.3, 1.0, V)[:, None]
    sig = (sig + 0.003 * (rng.normal(size=sig.shape) + 1j * rng.normal(size=sig.shape))).astype(np.complex64)
    return dict(params=params, A=A, natoms=natoms, d_params=d_params, d_atoms=d_atoms,
                atoms=atoms, sig=sig, d_sig=torch.from_numpy(sig).cuda(), b1=b1, ks=ks,
                d_b1=torch.from_numpy(b1).cuda(), ns=A // natoms)


def _basis(torch, p, per_slice=True, energy=0.0):
    nb = p["ns"] if per_slice else 1
    d_basis = torch.zeros((nb * KMAX, T), dtype=torch.complex64, device="cuda")
    K, sv = sm.svd_basis(p["d_atoms"], T, p["A"], p["natoms"], d_basis, KMAX, per_slice, energy)
    return K, sv, d_basis[: nb * K].reshape(nb, K, T)


def _lapack(p, b, per_slice=True):
    na = p["natoms"]
    D = p["atoms"][b * na:(b + 1) * na] if per_slice else p["atoms"]
    U, s, _ = np.linalg.svd(D.T.astype(np.complex128), full_matrices=False)
    return U, s


def test_basis_against_lapack_svd(gpu, problem):
    p = problem
    K, sv, basis = _basis(gpu, p)
    assert K == KMAX and basis.shape[0] == 3
    B = basis.cpu().numpy().astype(np.complex128)
    for b in range(3):
        U, s = _lapack(p, b)
        assert np.max(np.abs(sv[b, :K] - s[:K])) < 1e-9 * s[0]
        # orthonormal rows
        assert np.abs(B[b] @ B[b].conj().T - np.eye(K)).max() < 1e-5
        # the leading 20 singular directions lie in the GPU subspace
        Ul = U[:, :20]
        resid = Ul - B[b].T @ (B[b].conj() @ Ul)
        assert np.linalg.norm(resid) < 1e-4


def test_rank_rule_and_compressed_dictionary(gpu, problem):
    p = problem
    frac = 0.9999
    K, sv, basis = _basis(gpu, p, energy=frac)
    _, s = _lapack(p, 0)
    f = np.cumsum(s ** 2) / np.sum(s ** 2)
    assert K == max(1, min(KMAX, int(np.sum(f <= frac))))
    d_c = gpu.empty((p["A"], K), dtype=gpu.complex64, device="cuda")
    sm.svd_compress(basis, K, T, p["d_atoms"], p["A"], p["natoms"], d_c)
    C = d_c.cpu().numpy()
    B = basis.cpu().numpy().astype(np.complex128)
    na = p["natoms"]
    for b in range(3):
        ref = p["atoms"][b * na:(b + 1) * na].astype(np.complex128) @ B[b].conj().T
        assert np.abs(C[b * na:(b + 1) * na] - ref).max() < 1e-5 * np.abs(ref).max()


def _fp64_compressed_match(C, B, sig, ks, natoms, per_slice):
    """first-index argmax of |<c_a, U_k^H s>| over the voxel's slice (or all)."""
    V = sig.shape[0]
    idx = np.zeros(V, np.int64)
    gap = np.zeros(V)
    pd = np.zeros(V)
    s128 = sig.astype(np.complex128)
    for v in range(V):
        k = ks[v] if ks is not None else 0
        Ub = B[k] if per_slice else B[0]
        sp = Ub.conj() @ s128[v]
        rows = slice(k * natoms, (k + 1) * natoms) if ks is not None else slice(0, C.shape[0])
        sc = np.abs(C[rows].conj() @ sp)
        o = np.argsort(-sc, kind="stable")
        idx[v] = o[0] + (rows.start or 0)
        pd[v] = sc[o[0]]
        gap[v] = (sc[o[0]] - sc[o[1]]) / sc[o[0]]
    return idx, pd, gap


@pytest.mark.parametrize("per_slice", [True, False])
def test_compressed_matching(gpu, problem, per_slice):
    p = problem
    K, sv, basis = _basis(gpu, p, per_slice=per_slice, energy=0.99999)
    d_c = gpu.empty((p["A"], K), dtype=gpu.complex64, device="cuda")
    sm.svd_compress(basis, K, T, p["d_atoms"], p["A"], p["natoms"], d_c, per_slice=per_slice)
    V = p["sig"].shape[0]
    d_maps = gpu.empty((V, 5), device="cuda")
    d_idx = gpu.empty(V, dtype=gpu.int32, device="cuda")
    d_b1 = p["d_b1"] if per_slice else None
    sm.svd_match_ex(basis, K, T, d_c, p["d_params"], p["A"], p["natoms"], p["d_sig"], V, d_maps,
                    d_b1map=d_b1, d_idx=d_idx, per_slice=per_slice)
    gpu.cuda.synchronize()
    idx = d_idx.cpu().numpy()
    maps = d_maps.cpu().numpy()
    C = d_c.cpu().numpy().astype(np.complex128)
    B = basis.cpu().numpy().astype(np.complex128)
    ridx, rpd, gap = _fp64_compressed_match(C, B, p["sig"], p["ks"] if per_slice else None,
                                            p["natoms"], per_slice)
    # FP32 sums over K values err by ~1e-6 relative: voxels whose FP64 top-2
    # gap exceeds 1e-5 must match exactly
    clear = gap > 1e-5
    assert clear.mean() > 0.8
    assert np.array_equal(idx[clear], ridx[clear])
    assert np.allclose(maps[clear, 4], rpd[clear], rtol=1e-4)
    assert np.array_equal(maps[:, :4], p["params"][idx])
    # against the uncompressed FP32 matcher on the same voxels
    d_maps32 = gpu.empty((V, 5), device="cuda")
    d_idx32 = gpu.empty(V, dtype=gpu.int32, device="cuda")
    sm.match_ex(sm.MRF_MATCH_FP32, p["d_atoms"], p["d_params"], T, p["A"], p["natoms"],
                p["d_sig"], V, d_maps32, d_b1map=d_b1, d_idx=d_idx32)
    gpu.cuda.synchronize()
    assert np.mean(idx == d_idx32.cpu().numpy()) > 0.95


def test_per_slice_bases_need_a_b1_map(gpu, problem):
    p = problem
    K, sv, basis = _basis(gpu, p)
    d_c = gpu.empty((p["A"], K), dtype=gpu.complex64, device="cuda")
    d_maps = gpu.empty((4, 5), device="cuda")
    with pytest.raises(sm.MRFError):
        sm.svd_match_ex(basis, K, T, d_c, p["d_params"], p["A"], p["natoms"], p["d_sig"], 4,
                        d_maps)
