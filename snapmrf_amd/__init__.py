"""snapmrf_amd -- MI355X (gfx950) dictionary generation and template matching
for MR fingerprinting, a drop-in for snapMRF's hot path.

The compute path is the in-tree C-ABI library ``libsnapmrf.so`` (HIP kernels +
C++ host code, declared in ``include/snapmrf.h``).  This module is a thin
ctypes binding over it that mirrors the reference's interface
(``MRF_dict`` / ``MRF_match`` of src/functions.h:115-134, same argument names
and meaning) for torch device tensors.  There is no CPU fallback: if the
library is missing, importing the binding raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# SNAPMRF_LIB / SNAPMRF_CLI select another build of the same sources (the
# sanitizer build in snapmrf_amd/_san/)
LIB_PATH = os.environ.get("SNAPMRF_LIB") or os.path.join(_HERE, "libsnapmrf.so")
CLI_PATH = os.environ.get("SNAPMRF_CLI") or os.path.join(_HERE, "bin", "mrf")

MRF_MATCH_FP32 = 0
MRF_MATCH_BF16_TOP8 = 1
MRF_MATCH_SPLIT3 = 2
MRF_MATCH_BF16_TOP16 = 3
MRF_MATCH_SPLIT2 = 4
MRF_MATCH_BF16_TOP32 = 5
MRF_MATCH_BF16_TOP64 = 6
MRF_MATCH_RESCORE_FP32 = 0x100   # OR into a low-precision mode: FP32 rescore sums

ECHO_TYPES = ("epg_ssfp", "roa_ssfp", "epg_sr", "epg_se", "epg_tse1", "epg_tse2", "epg_fisp")

# every symbol include/snapmrf.h declares
EXPORTS = (
    "mrf_last_error", "mrf_csv_dim", "mrf_csv_read", "mrf_parse_length", "mrf_parse_params",
    "mrf_logspace", "mrf_compute_natoms", "mrf_trans_params", "mrf_grid_natoms_mode",
    "mrf_ra_read", "mrf_ra_write", "mrf_ra_free", "mrf_dict", "mrf_dict_ex", "mrf_match",
    "mrf_match_ex", "mrf_match_ex_noalloc", "mrf_dict_match_ex", "mrf_svd_basis", "mrf_svd_compress", "mrf_svd_match_ex",
    "mrf_lowp_prepare", "mrf_lowp_ld", "mrf_epg_jacobian", "mrf_epg_adjoint",
    "mrf_gn_refine", "mrf_crb_loss_grad", "mrf_device_count",
)


class MRFError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class MrfRa(ctypes.Structure):
    _fields_ = [("flags", ctypes.c_uint64), ("eltype", ctypes.c_uint64),
                ("elbyte", ctypes.c_uint64), ("size", ctypes.c_uint64),
                ("ndims", ctypes.c_uint64), ("dims", ctypes.POINTER(ctypes.c_uint64)),
                ("data", ctypes.POINTER(ctypes.c_uint8))]


class MrfDictOpts(ctypes.Structure):
    _fields_ = [("normalize", ctypes.c_int), ("n_iso", ctypes.c_int),
                ("iso_spread_hz", ctypes.c_float), ("kernel_form", ctypes.c_int),
                ("lanes_per_atom", ctypes.c_int), ("reserved", ctypes.c_int * 6)]


_lib = None


def lib() -> ctypes.CDLL:
    """Loads libsnapmrf.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m snapmrf_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    P, S, I, F = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_float
    cs = ctypes.c_char_p
    sig = {
        "mrf_last_error": ([], cs),
        "mrf_csv_dim": ([cs], S),
        "mrf_csv_read": ([P, S, cs], I),
        "mrf_parse_length": ([cs], S),
        "mrf_parse_params": ([P, cs, S], I),
        "mrf_logspace": ([P, F, F, S], None),
        "mrf_compute_natoms": ([P, P, S, S, S, S], S),
        "mrf_trans_params": ([P, P, P, P, P, S, S, S, S, S, I], I),
        "mrf_grid_natoms_mode": ([S, P, P, P, P, P, P, P, P], I),
        "mrf_ra_read": ([ctypes.POINTER(MrfRa), cs], I),
        "mrf_ra_write": ([ctypes.POINTER(MrfRa), cs], I),
        "mrf_ra_free": ([ctypes.POINTER(MrfRa)], None),
        "mrf_dict": ([P, P, P, S, S, S, S, I, I, cs], I),
        "mrf_dict_ex": ([P, cs, P, S, S, P, S, S, I, ctypes.POINTER(MrfDictOpts), P, P], I),
        "mrf_match": ([P, P, P, P, S, S, S, S, S, I, I, cs], I),
        "mrf_match_ex": ([P, I, P, P, P, P, S, S, S, P, S, P, P, P], I),
        "mrf_match_ex_noalloc": ([P, P, P, S, S, S, P, S, P, P, P], I),
        "mrf_dict_match_ex": ([P, cs, P, S, S, P, S, S, I, ctypes.POINTER(MrfDictOpts), I, P, S,
                               P, S, P, P], I),
        "mrf_svd_basis": ([P, P, S, S, S, I, ctypes.c_double, S, P, P, ctypes.POINTER(S)], I),
        "mrf_svd_compress": ([P, P, S, S, I, P, S, S, P], I),
        "mrf_svd_match_ex": ([P, P, S, S, I, P, P, S, S, P, S, P, P, P], I),
        "mrf_lowp_prepare": ([P, P, S, S, P, P], I),
        "mrf_lowp_ld": ([S], S),
        "mrf_epg_jacobian": ([P, P, S, S, P, S, I, P, P], I),
        "mrf_epg_adjoint": ([P, P, S, S, P, S, I, P, P, P], I),
        "mrf_gn_refine": ([P, P, S, S, P, S, I, I, P, P, P], I),
        "mrf_crb_loss_grad": ([P, P, S, S, P, S, I, P, P, P], I),
        "mrf_device_count": ([], I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def last_error() -> str:
    return lib().mrf_last_error().decode(errors="replace")


def check(rc: int) -> int:
    if rc < 0:
        raise MRFError(rc, last_error())
    return rc


def _ptr(t) -> int | None:
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    if hasattr(t, "ctypes"):
        return t.ctypes.data
    return int(t)


def _stream_ptr(stream) -> int | None:
    if stream is None:
        return None
    return getattr(stream, "cuda_stream", stream)


def _enc(s: str | None):
    return None if s is None else s.encode()


# ---------------------------------------------------------------- host ----

def csv_dim(path: str) -> int:
    n = lib().mrf_csv_dim(_enc(path))
    if n == ctypes.c_size_t(-1).value:
        raise MRFError(-2, last_error())
    return n


def csv_read(path: str, nreps: int):
    """Returns float32 [4, nreps] (FA | phase | TR | TE) and the warning text
    ('' when every row was read)."""
    import numpy as np
    out = np.zeros((4, nreps), dtype=np.float32)
    rc = check(lib().mrf_csv_read(out.ctypes.data, nreps, _enc(path)))
    return out, (last_error() if rc > 0 else "")


def parse_grid(s: str):
    import numpy as np
    n = lib().mrf_parse_length(_enc(s))
    if n == 0:
        raise MRFError(-1, last_error())
    out = np.zeros(n, dtype=np.float32)
    check(lib().mrf_parse_params(out.ctypes.data, _enc(s), n))
    return out


def logspace(mn: float, mx: float, n: int):
    import numpy as np
    out = np.zeros(n + 1, dtype=np.float32)
    lib().mrf_logspace(out.ctypes.data, mn, mx, n)
    return out[:n]


def build_params(T1: str, T2: str, B0: str, B1: str):
    """Grid strings -> (params float32 [NATOMS, 4], NATOMS, natoms), reference
    order (T1 fastest, then T2, B0, B1)."""
    import numpy as np
    t1, t2, b0, b1 = (parse_grid(x) for x in (T1, T2, B0, B1))
    return params_from_axes(t1, t2, b0, b1)


def params_from_axes(t1, t2, b0, b1):
    import numpy as np
    t1, t2, b0, b1 = (np.ascontiguousarray(x, dtype=np.float32) for x in (t1, t2, b0, b1))
    L = lib()
    NATOMS = L.mrf_compute_natoms(t1.ctypes.data, t2.ctypes.data, len(t1), len(t2), len(b0), len(b1))
    if NATOMS == 0:
        raise MRFError(-1, "Number of atoms is 0")
    params = np.zeros((NATOMS, 4), dtype=np.float32)
    check(L.mrf_trans_params(params.ctypes.data, t1.ctypes.data, t2.ctypes.data, b0.ctypes.data,
                             b1.ctypes.data, len(t1), len(t2), len(b0), len(b1), NATOMS, 4))
    return params, NATOMS, NATOMS // len(b1)


def grid_natoms_mode(N: int):
    """The -N grid of the reference CLI -> (t1, t2, b0, b1) float32 arrays."""
    import numpy as np
    L = lib()
    lens = [ctypes.c_size_t() for _ in range(4)]
    check(L.mrf_grid_natoms_mode(N, *[ctypes.byref(x) for x in lens], None, None, None, None))
    arrs = [np.zeros(x.value + 1, dtype=np.float32) for x in lens]
    check(L.mrf_grid_natoms_mode(N, *[ctypes.byref(x) for x in lens], *[a.ctypes.data for a in arrs]))
    return tuple(a[:x.value] for a, x in zip(arrs, lens))


def ra_read(path: str):
    """Reads an RA file through the library -> (numpy array shaped dims[::-1], header dict)."""
    import numpy as np
    ra = MrfRa()
    check(lib().mrf_ra_read(ctypes.byref(ra), _enc(path)))
    try:
        dims = [ra.dims[i] for i in range(ra.ndims)]
        dt = {(3, 4): np.float32, (3, 8): np.float64, (4, 8): np.complex64,
              (4, 16): np.complex128, (1, 4): np.int32, (1, 8): np.int64,
              (2, 4): np.uint32, (2, 8): np.uint64}.get((ra.eltype, ra.elbyte), np.uint8)
        raw = (np.ctypeslib.as_array(ra.data, shape=(ra.size,)) if ra.size
               else np.zeros(0, np.uint8))
        arr = raw.view(dt).copy()
        if dt is not np.uint8:
            arr = arr.reshape(dims[::-1])
        hdr = dict(flags=ra.flags, eltype=ra.eltype, elbyte=ra.elbyte, size=ra.size, dims=dims)
    finally:
        lib().mrf_ra_free(ctypes.byref(ra))
    return arr, hdr


def ra_write(path: str, arr, dims=None) -> None:
    import numpy as np
    arr = np.ascontiguousarray(arr)
    kind = {np.dtype(np.float32): (3, 4), np.dtype(np.float64): (3, 8),
            np.dtype(np.complex64): (4, 8), np.dtype(np.complex128): (4, 16),
            np.dtype(np.int32): (1, 4), np.dtype(np.int64): (1, 8)}[arr.dtype]
    dims = list(arr.shape[::-1]) if dims is None else list(dims)
    d = (ctypes.c_uint64 * len(dims))(*dims)
    ra = MrfRa(0, kind[0], kind[1], arr.nbytes, len(dims),
               ctypes.cast(d, ctypes.POINTER(ctypes.c_uint64)),
               ctypes.cast(arr.ctypes.data, ctypes.POINTER(ctypes.c_uint8)))
    check(lib().mrf_ra_write(ctypes.byref(ra), _enc(path)))


# ---------------------------------------------------------------- GPU ----

def device_count() -> int:
    return int(lib().mrf_device_count())


def MRF_dict(d_atoms, d_params, h_mrf, nreps: int, NREPS: int, NATOMS: int, natoms: int,
             nparams: int, nstates: int, echo_type: str) -> None:
    """Mirror of MRF_dict (src/functions.cu:1224).  d_atoms: device complex64
    [NATOMS, nreps]; d_params device float32 [NATOMS, 4]; h_mrf host float32
    [4, NREPS] (numpy).  Synchronous."""
    import numpy as np
    h = np.ascontiguousarray(h_mrf, dtype=np.float32)
    check(lib().mrf_dict(_ptr(d_atoms), _ptr(d_params), h.ctypes.data, nreps, NREPS, NATOMS,
                         natoms, nparams, nstates, _enc(echo_type)))


def dict_ex(echo_type: str, h_mrf, nreps: int, d_params, NATOMS: int, natoms: int, nstates: int,
            d_atoms, d_norm2=None, normalize: int = -1, n_iso: int = 1, iso_spread_hz: float = 0.0,
            generic: bool = False, lanes: int = 0, stream=None) -> None:
    """Extended dictionary generation on a stream (asynchronous).  generic:
    the unfused reference-order kernel; lanes: EPG lanes per atom (1, 2, 4,
    8, 16; 0 = chosen from the dictionary size)."""
    import numpy as np
    h = np.ascontiguousarray(h_mrf, dtype=np.float32)
    opts = MrfDictOpts()
    opts.normalize = normalize
    opts.n_iso = n_iso
    opts.iso_spread_hz = iso_spread_hz
    opts.kernel_form = 1 if generic else 0
    opts.lanes_per_atom = lanes
    check(lib().mrf_dict_ex(_stream_ptr(stream), _enc(echo_type), h.ctypes.data, nreps, h.shape[1],
                            _ptr(d_params), NATOMS, natoms, nstates, ctypes.byref(opts),
                            _ptr(d_atoms), _ptr(d_norm2)))


def MRF_match(d_maps, h_img, d_atoms, d_params, NREPS: int, nreps: int, NATOMS: int,
              natoms: int, nvoxels: int, nparams: int, nmaps: int, in_b1map: str | None) -> None:
    """Mirror of MRF_match (src/functions.cu:1493).  h_img host complex64
    [nvoxels, NREPS] (numpy); d_maps device float32 [nvoxels, 5]."""
    import numpy as np
    h = np.ascontiguousarray(h_img, dtype=np.complex64)
    check(lib().mrf_match(_ptr(d_maps), h.ctypes.data, _ptr(d_atoms), _ptr(d_params), NREPS,
                          nreps, NATOMS, natoms, nvoxels, nparams, nmaps, _enc(in_b1map)))


def match_ex(mode: int, d_atoms, d_params, T: int, NATOMS: int, natoms: int, d_sig, V: int,
             d_maps, d_b1map=None, d_idx=None, d_lowp=None, d_lowp_lo=None, stream=None) -> None:
    check(lib().mrf_match_ex(_stream_ptr(stream), mode, _ptr(d_atoms), _ptr(d_lowp),
                             _ptr(d_lowp_lo), _ptr(d_params), T, NATOMS, natoms, _ptr(d_sig), V,
                             _ptr(d_b1map), _ptr(d_maps), _ptr(d_idx)))


def match_ex_noalloc(d_atoms, d_params, T: int, NATOMS: int, natoms: int, d_sig, V: int,
                     d_maps, d_b1map=None, d_idx=None, stream=None) -> None:
    """FP32 matching as match_ex(MRF_MATCH_FP32) with no device allocation
    (static per-GPU scratch: at most 2^20 voxels; overlapping calls on one GPU
    race)."""
    check(lib().mrf_match_ex_noalloc(_stream_ptr(stream), _ptr(d_atoms), _ptr(d_params), T,
                                     NATOMS, natoms, _ptr(d_sig), V, _ptr(d_b1map), _ptr(d_maps),
                                     _ptr(d_idx)))


def dict_match_ex(echo_type: str, h_mrf, nreps: int, d_params, NATOMS: int, natoms: int,
                  nstates: int, d_sig, V: int, d_maps, d_b1map=None, d_idx=None,
                  block_atoms: int = 0, normalize: int = -1, n_iso: int = 1,
                  iso_spread_hz: float = 0.0, mode: int = 0, lanes: int = 0,
                  stream=None) -> None:
    """Dictionary generation streamed through the matcher (mode MRF_MATCH_*),
    block_atoms atoms per slice at a time (0: blocks of at most 1 GiB); with
    FP32 the maps equal dict_ex + match_ex(MRF_MATCH_FP32) with the same
    lanes (0 on both: each block takes the whole dictionary's lane form)."""
    import numpy as np
    h = np.ascontiguousarray(h_mrf, dtype=np.float32)
    opts = MrfDictOpts()
    opts.normalize = normalize
    opts.n_iso = n_iso
    opts.iso_spread_hz = iso_spread_hz
    opts.lanes_per_atom = lanes
    check(lib().mrf_dict_match_ex(_stream_ptr(stream), _enc(echo_type), h.ctypes.data, nreps,
                                  h.shape[1], _ptr(d_params), NATOMS, natoms, nstates,
                                  ctypes.byref(opts), mode, _ptr(d_sig), V, _ptr(d_b1map),
                                  block_atoms, _ptr(d_maps), _ptr(d_idx)))


def svd_basis(d_atoms, T: int, NATOMS: int, natoms: int, d_basis, K_max: int,
              per_slice: bool = True, energy_frac: float = 0.0, stream=None):
    """Rank-K basis (McGivney SVD compression); d_basis device complex64 with
    room for bases * K_max * T values.  Returns (K, singular values
    [bases, K_max] float64, zero past K)."""
    import numpy as np
    nb = NATOMS // natoms if per_slice else 1
    sv = np.zeros((nb, K_max), dtype=np.float64)
    K = ctypes.c_size_t()
    check(lib().mrf_svd_basis(_stream_ptr(stream), _ptr(d_atoms), T, NATOMS, natoms,
                              1 if per_slice else 0, energy_frac, K_max, _ptr(d_basis),
                              sv.ctypes.data, ctypes.byref(K)))
    return int(K.value), sv


def svd_compress(d_basis, K: int, T: int, d_atoms, NATOMS: int, natoms: int, d_catoms,
                 per_slice: bool = True, stream=None) -> None:
    check(lib().mrf_svd_compress(_stream_ptr(stream), _ptr(d_basis), K, T, 1 if per_slice else 0,
                                 _ptr(d_atoms), NATOMS, natoms, _ptr(d_catoms)))


def svd_match_ex(d_basis, K: int, T: int, d_catoms, d_params, NATOMS: int, natoms: int, d_sig,
                 V: int, d_maps, d_b1map=None, d_idx=None, per_slice: bool = True,
                 stream=None) -> None:
    check(lib().mrf_svd_match_ex(_stream_ptr(stream), _ptr(d_basis), K, T, 1 if per_slice else 0,
                                 _ptr(d_catoms), _ptr(d_params), NATOMS, natoms, _ptr(d_sig), V,
                                 _ptr(d_b1map), _ptr(d_maps), _ptr(d_idx)))


def lowp_ld(T: int) -> int:
    """Row stride (32-bit words) of the BF16 dictionary copies."""
    return int(lib().mrf_lowp_ld(T))


def lowp_prepare(d_atoms, T: int, NATOMS: int, d_hi, d_lo=None, stream=None) -> None:
    """d_hi/d_lo: device buffers of NATOMS * lowp_ld(T) 32-bit words."""
    check(lib().mrf_lowp_prepare(_stream_ptr(stream), _ptr(d_atoms), T, NATOMS, _ptr(d_hi),
                                 _ptr(d_lo)))


# ---- differentiation form ---------------------------------------------------

def _seq_table(h_mrf):
    """Host [4][T] float32 CSV table -> (array, T)."""
    import numpy as np
    h = np.ascontiguousarray(h_mrf, dtype=np.float32)
    if h.ndim != 2 or h.shape[0] != 4:
        raise ValueError("h_mrf must be [4][T]")
    return h, h.shape[1]


def epg_jacobian(h_mrf, d_theta, n: int, nstates: int, d_sig=None, d_jac=None,
                 stream=None) -> None:
    """Unnormalised EPG-SSFP signals d_sig [n][T] and ds/d(T1,T2,B0,B1)
    d_jac [n][4][T] (complex64 device buffers, either optional)."""
    h, T = _seq_table(h_mrf)
    check(lib().mrf_epg_jacobian(_stream_ptr(stream), h.ctypes.data, T, T, _ptr(d_theta), n,
                                 nstates, _ptr(d_sig), _ptr(d_jac)))


def epg_adjoint(h_mrf, d_theta, n: int, nstates: int, d_cot, d_grad_theta=None,
                d_grad_seq=None, stream=None) -> None:
    """Reverse mode of L = sum Re(conj(cot) s): d_grad_theta [n][4] and
    d_grad_seq [2][T] += (dL/dFA_raw, dL/dTR_raw)."""
    h, T = _seq_table(h_mrf)
    check(lib().mrf_epg_adjoint(_stream_ptr(stream), h.ctypes.data, T, T, _ptr(d_theta), n,
                                nstates, _ptr(d_cot), _ptr(d_grad_theta), _ptr(d_grad_seq)))


def gn_refine(h_mrf, d_sig, V: int, nstates: int, iters: int, d_theta, d_pd=None,
              d_cost=None, stream=None) -> None:
    h, T = _seq_table(h_mrf)
    check(lib().mrf_gn_refine(_stream_ptr(stream), h.ctypes.data, T, T, _ptr(d_sig), V, nstates,
                              iters, _ptr(d_theta), _ptr(d_pd), _ptr(d_cost)))


def crb_loss_grad(h_mrf, d_tissues, n: int, nstates: int, wts, d_loss, d_grad_seq=None,
                  stream=None) -> None:
    import numpy as np
    h, T = _seq_table(h_mrf)
    w = np.ascontiguousarray(wts, dtype=np.float32)
    check(lib().mrf_crb_loss_grad(_stream_ptr(stream), h.ctypes.data, T, T, _ptr(d_tissues), n,
                                  nstates, w.ctypes.data, _ptr(d_loss), _ptr(d_grad_seq)))
