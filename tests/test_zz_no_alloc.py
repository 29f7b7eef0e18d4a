"""mrf_match_ex_noalloc: FP32 matching with the library's static per-GPU
scratch instead of device allocations.  Its maps and indices must equal
mrf_match_ex (MRF_MATCH_FP32) bit for bit -- the same kernels on the same
inputs -- on consecutive calls sharing the one scratch array, the second
with fewer voxels and a different dictionary, with and without a B1 map;
sizes beyond the scratch are rejected before any launch."""
import numpy as np
import pytest

import snapmrf_amd as sm
from snapmrf_amd import synth

pytestmark = pytest.mark.gpu


def _case(torch, T, grid, V, seed):
    raw = synth.csv_columns(synth.sequence(T, seed=seed), "epg_ssfp")
    params, A, natoms = sm.build_params(*grid)
    d_params = torch.from_numpy(params).cuda()
    d_atoms = torch.empty((A, T), dtype=torch.complex64, device="cuda")
    sm.dict_ex("epg_ssfp", raw, T, d_params, A, natoms, 16, d_atoms)
    rng = np.random.default_rng(seed)
    truth = rng.integers(0, A, V)
    sig = d_atoms[torch.from_numpy(truth).cuda()] * 0.7
    g = torch.Generator(device="cuda").manual_seed(seed)
    sig = (sig + 0.02 * torch.complex(torch.randn(V, T, device="cuda", generator=g),
                                      torch.randn(V, T, device="cuda", generator=g))).contiguous()
    b1 = torch.from_numpy(rng.uniform(0.8, 1.2, V).astype(np.float32)).cuda()
    return params, A, natoms, d_params, d_atoms, sig, b1


def _both(torch, T, A, natoms, d_params, d_atoms, sig, b1):
    V = sig.shape[0]
    out = []
    for fn in ("ex", "noalloc"):
        d_maps = torch.full((V, 5), -3.0, device="cuda")
        d_idx = torch.full((V,), -3, dtype=torch.int32, device="cuda")
        if fn == "ex":
            sm.match_ex(sm.MRF_MATCH_FP32, d_atoms, d_params, T, A, natoms, sig, V, d_maps,
                        d_b1map=b1, d_idx=d_idx)
        else:
            sm.match_ex_noalloc(d_atoms, d_params, T, A, natoms, sig, V, d_maps, d_b1map=b1,
                                d_idx=d_idx)
        torch.cuda.synchronize()
        out.append((d_maps.cpu().numpy(), d_idx.cpu().numpy()))
    return out


@pytest.mark.parametrize("use_b1", [True, False])
def test_noalloc_equals_match_ex_on_consecutive_calls(gpu, use_b1):
    torch = gpu
    big = ("100:40:2000+2300:300:5000", "20:10:100+110:30:300+400:400:2000", "-30:6:30",
           "0.8:0.2:1.2")
    small = ("200:100:2000", "30:30:300", "-20:20:20", "0.9:0.1:1.1")
    for T, grid, V, seed in ((300, big, 5000, 1), (77, small, 333, 2)):
        params, A, natoms, d_params, d_atoms, sig, b1 = _case(torch, T, grid, V, seed)
        (m_ex, i_ex), (m_na, i_na) = _both(torch, T, A, natoms, d_params, d_atoms, sig,
                                           b1 if use_b1 else None)
        assert np.array_equal(i_na, i_ex), (T, V)
        assert np.array_equal(m_na, m_ex), (T, V)
        assert np.array_equal(m_na[:, :4], params[i_na])


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu):
    torch = gpu
    T = 8
    d_atoms = torch.zeros((128, T), dtype=torch.complex64, device="cuda")
    d_params = torch.ones((128, 4), device="cuda")
    # 2^21 voxels need 12 bytes each (key, permutation, tile share) without a
    # B1 map: 24 MiB, more than the 20 MiB scratch
    V = 1 << 21
    d_sig = torch.zeros((V, T), dtype=torch.complex64, device="cuda")
    d_maps = torch.empty((V, 5), device="cuda")
    with pytest.raises(sm.MRFError, match="scratch"):
        sm.match_ex_noalloc(d_atoms, d_params, T, 128, 128, d_sig, V, d_maps)
    # the limit: 2^20 voxels over 1,024 B1 slices of one atom each fit
    V = 1 << 20
    d_b1 = torch.ones(V, device="cuda")
    p = torch.ones((1024, 4), device="cuda")
    a = torch.zeros((1024, T), dtype=torch.complex64, device="cuda")
    d_idx = torch.empty(V, dtype=torch.int32, device="cuda")
    sm.match_ex_noalloc(a, p, T, 1024, 1, d_sig[:V], V, d_maps[:V], d_b1map=d_b1, d_idx=d_idx)
    torch.cuda.synchronize()
    assert int(d_idx.min()) == 0 and int(d_idx.max()) == 0
