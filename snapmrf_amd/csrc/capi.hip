// capi.hip -- C ABI of the snapMRF dictionary + matching path (host
// orchestration of the gfx950 kernels).  See include/snapmrf.h.
#include "snapmrf.h"
#include "internal.h"
#include "kernels.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

namespace mrf {

int echo_id(const char *e)
{
    // src/functions.cu:1238-1311
    if (!e) return -1;
    if (!std::strcmp(e, "epg_ssfp")) return kEpgSsfp;
    if (!std::strcmp(e, "roa_ssfp")) return kRoaSsfp;
    if (!std::strcmp(e, "epg_sr")) return kEpgSr;
    if (!std::strcmp(e, "epg_se")) return kEpgSe;
    if (!std::strcmp(e, "epg_tse1") || !std::strcmp(e, "epg_tse2")) return kEpgTse;
    if (!std::strcmp(e, "epg_fisp")) return kEpgFisp;
    return -1;
}

namespace {

#define MRF_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) return mrf_fail(MRF_E_GPU, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// Per-TR coefficient table, one thread per TR.  For B1-sliced SSFP the flip
// angles follow the reference's in-place float scaling: FA *= b1_j before
// slice j and FA *= 1/b1_j after it (src/functions.cu:970-973, 1012-1013).
__global__ void coef_kernel(const float *seq, int T, const float *params, size_t natoms,
                            int n_slices, int b1_scale, TrCoef *coef)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= T) return;
    float fa = seq[n];
    const float ph = seq[T + n];
    const float ph_next = (n + 1 < T) ? seq[T + n + 1] : ph;
    const float tr = seq[2 * T + n], te = seq[3 * T + n];
    for (int j = 0; j < n_slices; j++) {
        float b1 = 1.0f;
        if (b1_scale) {
            b1 = params[(size_t)j * natoms * 4 + 3];
            fa = fa * b1;
        }
        TrCoef c;
        const double a = (double)fa;
        const double ca = cos(a), sa = sin(a);
        c.ca2 = (float)((1.0 + ca) * 0.5);
        c.sa2 = (float)((1.0 - ca) * 0.5);
        c.sa = (float)sa;
        c.ca = (float)ca;
        c.zr = (float)cos((double)ph);
        c.zi = (float)sin((double)ph);
        const double d = (double)ph - (double)ph_next;
        c.rr = (float)cos(d);
        c.ri = (float)sin(d);
        c.te = te;
        c.tr = tr;
        c.tmte = tr - te;
        c.s2p = (float)sin(2.0 * (double)ph);
        c.c2p = (float)cos(2.0 * (double)ph);
        c.pad[0] = fa;   // the angle alpha * B1 itself (per-voxel B1 paths)
        c.pad[1] = c.pad[2] = 0.f;
        coef[(size_t)j * T + n] = c;
        if (b1_scale) fa = fa * (1.0f / b1);
    }
}

// CSV unit transform (src/functions.cu:1240-1261, 1271-1291), float32.
void prepare_sequence(int echo, const float *h_mrf, size_t nreps, size_t NREPS,
                      std::vector<float> &out)
{
    out.resize(4 * nreps);
    for (int c = 0; c < 4; c++)
        std::memcpy(&out[c * nreps], h_mrf + (size_t)c * NREPS, nreps * sizeof(float));
    if (echo != kEpgSsfp && echo != kRoaSsfp) return;
    const float fa_scale = (float)(60.0f * M_PI / 180.0f);
    const float ph_scale = (float)(M_PI / 180.0f);
    for (size_t i = 0; i < nreps; i++) {
        out[i] = out[i] * fa_scale;
        out[nreps + i] = out[nreps + i] * ph_scale;
        out[2 * nreps + i] = out[2 * nreps + i] + 16.0f;
        out[3 * nreps + i] = out[3 * nreps + i] + 3.5f;
    }
}

// Atoms of `natoms` parameter rows per slice (n_slices slices of params,
// rows natoms apart) with the per-slice coefficient table `coef`; `lanes`
// fixes the EPG lane form (0: chosen from the launch size).
int gen_atoms(int echo, const TrCoef *coef, const float *params, size_t natoms, int n_slices,
              int T, int nstates, const mrf_dict_opts *opts, bool h_real, int normalize,
              int lanes, float2 *atoms, double *n2, hipStream_t s)
{
    int rc = 0;
    if (echo == kRoaSsfp) {
        RoaLaunch l{};
        l.coef = coef;
        l.params = params;
        l.atoms = atoms;
        l.norm2 = n2;
        l.natoms = natoms;
        l.T = T;
        l.n_iso = opts ? opts->n_iso : 1;
        l.iso_spread_hz = opts ? opts->iso_spread_hz : 0.f;
        rc = launch_roa(l, s);
        if (!rc && normalize) rc = launch_normalize(atoms, n2, natoms, T, s);
    } else {
        EpgLaunch l{};
        l.coef = coef;
        l.params = params;
        l.atoms = atoms;
        l.norm2 = n2;
        l.natoms = natoms;
        l.n_slices = n_slices;
        l.T = T;
        l.w = nstates;
        l.echo = echo;
        l.generic = opts && opts->kernel_form == 1;
        l.lanes = lanes;
        l.h_real = h_real && echo == kEpgFisp;
        l.normalize = normalize != 0;
        rc = launch_epg(l, s);
    }
    return rc;
}

int check_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
        return mrf_fail(MRF_E_GPU, "no HIP device available");
    return 0;
}

}  // namespace
}  // namespace mrf

using namespace mrf;

extern "C" int mrf_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int mrf_dict_ex(void *stream, const char *echo_type, const float *h_mrf,
                           size_t nreps, size_t NREPS, const float *d_params, size_t NATOMS,
                           size_t natoms, int nstates, const mrf_dict_opts *opts,
                           mrf_complex *d_atoms, double *d_norm2)
{
    mrf_clear_error();
    const int echo = echo_id(echo_type);
    if (echo < 0)
        return mrf_fail(MRF_E_ECHO, "unknown echo type '%s' (expected epg_ssfp, roa_ssfp, "
                        "epg_sr, epg_se, epg_tse1, epg_tse2 or epg_fisp)",
                        echo_type ? echo_type : "(null)");
    if (!h_mrf || !d_params || !d_atoms) return mrf_fail(MRF_E_ARG, "null pointer argument");
    if (nreps == 0 || nreps > NREPS) return mrf_fail(MRF_E_ARG, "need 0 < nreps <= NREPS");
    if (nreps > 0x7fffffff) return mrf_fail(MRF_E_ARG, "nreps too large");
    if (natoms == 0 || NATOMS == 0 || NATOMS % natoms)
        return mrf_fail(MRF_E_ARG, "NATOMS (%zu) must be a positive multiple of natoms (%zu)",
                        NATOMS, natoms);
    if (int rc = check_device()) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int T = (int)nreps;
    const int n_slices = (echo == kEpgSsfp) ? (int)(NATOMS / natoms) : 1;
    const size_t n_gen = (size_t)n_slices * natoms;
    const int normalize = (opts && opts->normalize >= 0)
                              ? opts->normalize
                              : (echo == kEpgSsfp || echo == kRoaSsfp);

    std::vector<float> seq;
    prepare_sequence(echo, h_mrf, nreps, NREPS, seq);
    bool h_real = true;
    for (size_t i = 1; i < nreps; i++)
        if (seq[nreps + i] != seq[nreps]) h_real = false;

    float *d_seq = nullptr;
    TrCoef *d_coef = nullptr;
    double *d_n2 = d_norm2, *own_n2 = nullptr;
    int rc = 0;
    auto alloc = [&](void **p, size_t bytes) {
        if (rc) return;
        hipError_t e = hipMallocAsync(p, bytes, s);
        if (e != hipSuccess)
            rc = mrf_fail(MRF_E_NOMEM, "device allocation of %zu bytes: %s", bytes,
                          hipGetErrorString(e));
    };
    alloc((void **)&d_seq, 4 * nreps * sizeof(float));
    alloc((void **)&d_coef, (size_t)n_slices * T * sizeof(TrCoef));
    if (!d_n2 && normalize) {
        alloc((void **)&own_n2, n_gen * sizeof(double));
        d_n2 = own_n2;
    }
    if (!rc) {
        hipError_t e = hipMemcpyAsync(d_seq, seq.data(), 4 * nreps * sizeof(float),
                                      hipMemcpyHostToDevice, s);
        if (e != hipSuccess) rc = mrf_fail(MRF_E_GPU, "sequence upload: %s", hipGetErrorString(e));
    }
    if (!rc) {
        coef_kernel<<<dim3((T + 127) / 128), 128, 0, s>>>(d_seq, T, d_params, natoms, n_slices,
                                                           echo == kEpgSsfp ? 1 : 0, d_coef);
        rc = gen_atoms(echo, d_coef, d_params, natoms, n_slices, T, nstates, opts, h_real,
                       normalize, opts ? opts->lanes_per_atom : 0,
                       reinterpret_cast<float2 *>(d_atoms), d_n2, s);
    }
    for (void *p : {(void *)d_seq, (void *)d_coef, (void *)own_n2})
        if (p) (void)hipFreeAsync(p, s);
    return rc;
}

extern "C" int mrf_dict(mrf_complex *d_atoms, const float *d_params, const float *h_mrf,
                        size_t nreps, size_t NREPS, size_t NATOMS, size_t natoms, int nparams,
                        int nstates, const char *echo_type)
{
    if (nparams != 4) return mrf_fail(MRF_E_ARG, "nparams must be 4 (T1, T2, B0, B1)");
    int rc = mrf_dict_ex(nullptr, echo_type, h_mrf, nreps, NREPS, d_params, NATOMS, natoms,
                         nstates, nullptr, d_atoms, nullptr);
    if (rc) return rc;
    MRF_HIP(hipStreamSynchronize(nullptr));
    MRF_HIP(hipGetLastError());
    return MRF_OK;
}

extern "C" int mrf_match_ex(void *stream, int mode, const mrf_complex *d_atoms,
                            const void *d_lowp, const void *d_lowp_lo, const float *d_params,
                            size_t T, size_t NATOMS, size_t natoms, const mrf_complex *d_sig,
                            size_t V, const float *d_b1map, float *d_maps, int32_t *d_idx)
{
    mrf_clear_error();
    const bool rescore_fp32 = (mode & MRF_MATCH_RESCORE_FP32) != 0;
    mode &= ~MRF_MATCH_RESCORE_FP32;
    if (rescore_fp32 && mode == MRF_MATCH_FP32)
        return mrf_fail(MRF_E_ARG, "the FP32 rescore flag applies to low-precision modes");
    if (!d_atoms || !d_params || (V && (!d_sig || !d_maps)))
        return mrf_fail(MRF_E_ARG, "null pointer argument");
    if (T == 0 || T > 0x7fffffff) return mrf_fail(MRF_E_ARG, "bad T");
    if (natoms == 0 || NATOMS == 0 || NATOMS % natoms)
        return mrf_fail(MRF_E_ARG, "NATOMS (%zu) must be a positive multiple of natoms (%zu)",
                        NATOMS, natoms);
    if (mode != MRF_MATCH_FP32 && !d_lowp)
        return mrf_fail(MRF_E_ARG, "low-precision mode needs the bf16 dictionary");
    if ((mode == MRF_MATCH_SPLIT3 || mode == MRF_MATCH_SPLIT2) && !d_lowp_lo)
        return mrf_fail(MRF_E_ARG, "split mode needs the bf16 lo dictionary");
    if (int rc = check_device()) return rc;
    if (V == 0) return MRF_OK;
    hipStream_t s = (hipStream_t)stream;
    const int n_slices = (int)(NATOMS / natoms);
    MatchLaunch l{};
    l.atoms = reinterpret_cast<const float2 *>(d_atoms);
    l.lowp_hi = (const uint32_t *)d_lowp;
    l.lowp_lo = (const uint32_t *)d_lowp_lo;
    l.params = d_params;
    l.sig = reinterpret_cast<const float2 *>(d_sig);
    l.natoms = natoms;
    l.n_slices = n_slices;
    l.T = (int)T;
    l.V = V;
    l.mode = mode;
    l.rescore_fp32 = rescore_fp32;
    l.maps = d_maps;
    l.idx = d_idx;
    int32_t *d_vs = nullptr;
    void *scratch = nullptr;
    int rc = 0;
    auto alloc = [&](void **p, size_t bytes) {
        if (rc) return;
        hipError_t e = hipMallocAsync(p, bytes, s);
        if (e != hipSuccess)
            rc = mrf_fail(MRF_E_NOMEM, "device allocation of %zu bytes: %s", bytes,
                          hipGetErrorString(e));
    };
    if (d_b1map) alloc((void **)&d_vs, V * sizeof(int32_t));
    if (!rc) {
        l.vox_slice = d_vs;   // scratch size depends on the slice bucketing
        alloc(&scratch, match_scratch_bytes(l));
    }
    if (!rc && d_b1map) rc = launch_b1_select(d_b1map, d_params, natoms, n_slices, V, d_vs, s);
    if (!rc) rc = launch_match(l, scratch, s);
    for (void *p : {(void *)d_vs, scratch})
        if (p) (void)hipFreeAsync(p, s);
    return rc;
}

// ---------------------------------------------------------------------------
// FP32 matching without device allocation (mrf_match_ex_noalloc)

// Scratch of mrf_match_ex_noalloc, one copy per GPU for the life of the
// process: the B1 slice of each voxel, the per-voxel (score, index) keys, the
// voxel permutation, the slice counts and the tile list of one FP32 matching
// call.  2^20 voxels over 1,024 B1 slices need about 16.2 MiB of it.
static constexpr size_t kNoAllocScratchBytes = (size_t)20 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace mrf {
namespace {

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kNoAllocMaxDevices];

// Device address of the current device's copy of `scratch`, looked up on the
// first call on each device (the array's host-side name is not a device
// address and must never reach a kernel).
int noalloc_scratch(void **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return mrf_fail(MRF_E_GPU, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kNoAllocMaxDevices)
        return mrf_fail(MRF_E_UNSUPPORTED, "device %d: the static scratch is tracked for devices "
                        "0..%d", dev, kNoAllocMaxDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return mrf_fail(MRF_E_GPU, "address of the static scratch: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace
}  // namespace mrf

extern "C" int mrf_match_ex_noalloc(void *stream, const mrf_complex *d_atoms,
                                    const float *d_params, size_t T, size_t NATOMS,
                                    size_t natoms, const mrf_complex *d_sig, size_t V,
                                    const float *d_b1map, float *d_maps, int32_t *d_idx)
{
    mrf_clear_error();
    if (!d_atoms || !d_params || (V && (!d_sig || !d_maps)))
        return mrf_fail(MRF_E_ARG, "null pointer argument");
    if (T == 0 || T > 0x7fffffff) return mrf_fail(MRF_E_ARG, "bad T");
    if (natoms == 0 || NATOMS == 0 || NATOMS % natoms)
        return mrf_fail(MRF_E_ARG, "NATOMS (%zu) must be a positive multiple of natoms (%zu)",
                        NATOMS, natoms);
    if (int rc = check_device()) return rc;
    if (V == 0) return MRF_OK;
    void *base = nullptr;
    if (int rc = noalloc_scratch(&base)) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int n_slices = (int)(NATOMS / natoms);
    MatchLaunch l{};
    l.atoms = reinterpret_cast<const float2 *>(d_atoms);
    l.params = d_params;
    l.sig = reinterpret_cast<const float2 *>(d_sig);
    l.natoms = natoms;
    l.n_slices = n_slices;
    l.T = (int)T;
    l.V = V;
    l.mode = MRF_MATCH_FP32;
    l.maps = d_maps;
    l.idx = d_idx;
    // [voxel slices | match scratch], both 256-byte aligned
    const size_t vs_bytes = d_b1map ? (V * sizeof(int32_t) + 255) / 256 * 256 : 0;
    int32_t *d_vs = d_b1map ? (int32_t *)base : nullptr;
    l.vox_slice = d_vs;
    const size_t need = vs_bytes + match_scratch_bytes(l);
    if (need > kNoAllocScratchBytes)
        return mrf_fail(MRF_E_ARG, "%zu voxels over %d slices need %zu bytes of scratch; the "
                        "static scratch holds %zu (2^20 voxels over 1,024 slices fit)", V,
                        d_b1map ? n_slices : 1, need, kNoAllocScratchBytes);
    if (d_b1map)
        if (int rc = launch_b1_select(d_b1map, d_params, natoms, n_slices, V, d_vs, s)) return rc;
    return launch_match(l, (unsigned char *)base + vs_bytes, s);
}

extern "C" int mrf_dict_match_ex(void *stream, const char *echo_type, const float *h_mrf,
                                 size_t nreps, size_t NREPS, const float *d_params,
                                 size_t NATOMS, size_t natoms, int nstates,
                                 const mrf_dict_opts *opts, int mode, const mrf_complex *d_sig,
                                 size_t V, const float *d_b1map, size_t block_atoms,
                                 float *d_maps, int32_t *d_idx)
{
    mrf_clear_error();
    const bool rescore_fp32 = (mode & MRF_MATCH_RESCORE_FP32) != 0;
    mode &= ~MRF_MATCH_RESCORE_FP32;
    if (rescore_fp32 && mode == MRF_MATCH_FP32)
        return mrf_fail(MRF_E_ARG, "the FP32 rescore flag applies to low-precision modes");
    const int echo = echo_id(echo_type);
    if (echo < 0)
        return mrf_fail(MRF_E_ECHO, "unknown echo type '%s'", echo_type ? echo_type : "(null)");
    if (!h_mrf || !d_params || (V && (!d_sig || !d_maps)))
        return mrf_fail(MRF_E_ARG, "null pointer argument");
    if (nreps == 0 || nreps > NREPS) return mrf_fail(MRF_E_ARG, "need 0 < nreps <= NREPS");
    if (nreps > 0x7fffffff) return mrf_fail(MRF_E_ARG, "nreps too large");
    if (natoms == 0 || NATOMS == 0 || NATOMS % natoms)
        return mrf_fail(MRF_E_ARG, "NATOMS (%zu) must be a positive multiple of natoms (%zu)",
                        NATOMS, natoms);
    const int n_slices = (int)(NATOMS / natoms);
    if (echo != kEpgSsfp && n_slices != 1)
        return mrf_fail(MRF_E_ARG, "%s fills one B1 slice only; streamed matching needs "
                        "NATOMS == natoms", echo_type);
    if (int rc = check_device()) return rc;
    if (V == 0) return MRF_OK;
    hipStream_t s = (hipStream_t)stream;
    const int T = (int)nreps;
    // block: whole multiples of 128 atoms per slice, at most 1 GiB of atoms
    size_t block = block_atoms;
    if (block == 0) {
        const size_t per = (size_t)n_slices * T * sizeof(float2);
        block = ((size_t)1 << 30) / per / 128 * 128;
        if (block < 128) block = 128;
    }
    // low-precision modes: blocks start on the matcher's 128-atom sub-tiles
    if (mode != MRF_MATCH_FP32) block = (block + 127) / 128 * 128;
    if (block > natoms) block = natoms;
    const int normalize = (opts && opts->normalize >= 0)
                              ? opts->normalize
                              : (echo == kEpgSsfp || echo == kRoaSsfp);
    std::vector<float> seq;
    prepare_sequence(echo, h_mrf, nreps, NREPS, seq);
    bool h_real = true;
    for (size_t i = 1; i < nreps; i++)
        if (seq[nreps + i] != seq[nreps]) h_real = false;
    // Every block is generated in the lane form mrf_dict_ex picks for the
    // whole dictionary: the form sets the FP32 rounding of each atom, so the
    // blocks then hold exactly the rows the whole dictionary would.
    int lanes = opts ? opts->lanes_per_atom : 0;
    if (echo != kRoaSsfp && lanes == 0) {
        EpgLaunch whole{};
        whole.natoms = natoms;
        whole.n_slices = n_slices;
        whole.w = nstates;
        whole.echo = echo;
        whole.generic = opts && opts->kernel_form == 1;
        int M = 0;
        if (int rc = epg_form(whole, &lanes, &M)) return rc;
    }

    MatchLaunch l{};
    l.params = d_params;
    l.sig = reinterpret_cast<const float2 *>(d_sig);
    l.natoms = natoms;
    l.n_slices = n_slices;
    l.T = T;
    l.V = V;
    l.mode = mode;
    l.rescore_fp32 = rescore_fp32;
    l.maps = d_maps;
    l.idx = d_idx;
    if (mode < MRF_MATCH_FP32 || mode > MRF_MATCH_BF16_TOP64)
        return mrf_fail(MRF_E_UNSUPPORTED, "matching mode %d not available", mode);
    if (mode != MRF_MATCH_FP32 && T > kLowpMaxT)
        return mrf_fail(MRF_E_ARG, "low-precision matching supports T <= %d", kLowpMaxT);
    const size_t nblk = (size_t)n_slices * block;
    float *d_seq = nullptr, *d_pblk = nullptr;
    TrCoef *d_coef = nullptr;
    float2 *d_buf = nullptr;
    double *d_n2 = nullptr;
    int32_t *d_vs = nullptr;
    void *scratch = nullptr;
    LowpBlockBufs lb{};
    int rc = 0;
    auto alloc = [&](void **p, size_t bytes) {
        if (rc) return;
        hipError_t e = hipMallocAsync(p, bytes, s);
        if (e != hipSuccess)
            rc = mrf_fail(MRF_E_NOMEM, "device allocation of %zu bytes: %s", bytes,
                          hipGetErrorString(e));
    };
    alloc((void **)&d_seq, 4 * nreps * sizeof(float));
    alloc((void **)&d_coef, (size_t)n_slices * T * sizeof(TrCoef));
    alloc((void **)&d_buf, nblk * T * sizeof(float2));
    alloc((void **)&d_pblk, nblk * 4 * sizeof(float));
    if (normalize) alloc((void **)&d_n2, nblk * sizeof(double));
    if (d_b1map) alloc((void **)&d_vs, V * sizeof(int32_t));
    if (mode != MRF_MATCH_FP32) {
        const size_t lw = nblk * lowp_ld(nreps) * sizeof(uint32_t);
        alloc((void **)&lb.hi, lw);
        if (mode == MRF_MATCH_SPLIT2 || mode == MRF_MATCH_SPLIT3) alloc((void **)&lb.lo, lw);
        alloc((void **)&lb.best_s, V * sizeof(double));
        alloc((void **)&lb.best_i, V * sizeof(uint32_t));
    }
    if (!rc) {
        l.vox_slice = d_vs;
        alloc(&scratch, match_scratch_bytes(l));
    }
    if (!rc) {
        hipError_t e = hipMemcpyAsync(d_seq, seq.data(), 4 * nreps * sizeof(float),
                                      hipMemcpyHostToDevice, s);
        if (e != hipSuccess) rc = mrf_fail(MRF_E_GPU, "sequence upload: %s", hipGetErrorString(e));
    }
    if (!rc) {
        coef_kernel<<<dim3((T + 127) / 128), 128, 0, s>>>(d_seq, T, d_params, natoms, n_slices,
                                                           echo == kEpgSsfp ? 1 : 0, d_coef);
        if (d_b1map) rc = launch_b1_select(d_b1map, d_params, natoms, n_slices, V, d_vs, s);
    }
    if (!rc) {
        // block a0: parameter rows a0 .. a0+nb-1 of every slice, then their atoms
        auto gen = [&](size_t a0, size_t nb) -> int {
            hipError_t e = hipMemcpy2DAsync(d_pblk, nb * 4 * sizeof(float), d_params + a0 * 4,
                                            natoms * 4 * sizeof(float), nb * 4 * sizeof(float),
                                            (size_t)n_slices, hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess)
                return mrf_fail(MRF_E_GPU, "parameter gather: %s", hipGetErrorString(e));
            return gen_atoms(echo, d_coef, d_pblk, nb, n_slices, T, nstates, opts, h_real,
                             normalize, lanes, d_buf, d_n2, s);
        };
        rc = launch_match_blocks(l, block, d_buf, mode != MRF_MATCH_FP32 ? &lb : nullptr, scratch, s,
                                 gen);
    }
    for (void *p : {(void *)d_seq, (void *)d_coef, (void *)d_buf, (void *)d_pblk, (void *)d_n2,
                    (void *)d_vs, scratch, (void *)lb.hi, (void *)lb.lo, (void *)lb.best_s,
                    (void *)lb.best_i})
        if (p) (void)hipFreeAsync(p, s);
    return rc;
}

extern "C" int mrf_match(float *d_maps, const mrf_complex *h_img, const mrf_complex *d_atoms,
                         const float *d_params, size_t NREPS, size_t nreps, size_t NATOMS,
                         size_t natoms, size_t nvoxels, int nparams, int nmaps,
                         const char *in_b1map)
{
    mrf_clear_error();
    if (nparams != 4 || nmaps != 5) return mrf_fail(MRF_E_ARG, "need nparams = 4 and nmaps = 5");
    if (!h_img && nvoxels) return mrf_fail(MRF_E_ARG, "null image");
    // The reference strides voxels by NREPS but indexes atoms with nreps
    // (src/functions.cu:1524-1535); it is only meaningful when they agree.
    if (nreps != NREPS)
        return mrf_fail(MRF_E_ARG, "image frames (%zu) must equal the dictionary length (%zu)",
                        nreps, NREPS);
    if (int rc = check_device()) return rc;
    std::vector<float> b1;
    if (in_b1map) {
        mrf_ra_t ra;
        if (int rc = mrf_ra_read(&ra, in_b1map)) return rc;
        size_t n = 1;
        for (uint64_t d = 0; d < ra.ndims; d++) n *= ra.dims[d];
        if (ra.eltype != 3 || ra.elbyte != 4 || n < nvoxels) {
            mrf_ra_free(&ra);
            return mrf_fail(MRF_E_FORMAT, "%s: need a float32 B1 map with %zu voxels", in_b1map,
                            nvoxels);
        }
        b1.assign((const float *)ra.data, (const float *)ra.data + nvoxels);
        mrf_ra_free(&ra);
    }
    mrf_complex *d_sig = nullptr;
    float *d_b1 = nullptr;
    MRF_HIP(hipMalloc((void **)&d_sig, nvoxels * nreps * sizeof(mrf_complex)));
    int rc = 0;
    if (hipMemcpy(d_sig, h_img, nvoxels * nreps * sizeof(mrf_complex), hipMemcpyHostToDevice) !=
        hipSuccess)
        rc = mrf_fail(MRF_E_GPU, "image upload failed");
    if (!rc && in_b1map) {
        if (hipMalloc((void **)&d_b1, nvoxels * sizeof(float)) != hipSuccess ||
            hipMemcpy(d_b1, b1.data(), nvoxels * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            rc = mrf_fail(MRF_E_GPU, "B1 map upload failed");
    }
    if (!rc)
        rc = mrf_match_ex(nullptr, MRF_MATCH_FP32, d_atoms, nullptr, nullptr, d_params, nreps,
                          NATOMS, natoms, d_sig, nvoxels, d_b1, d_maps, nullptr);
    if (!rc && hipStreamSynchronize(nullptr) != hipSuccess)
        rc = mrf_fail(MRF_E_GPU, "matching failed: %s", hipGetErrorString(hipGetLastError()));
    (void)hipFree(d_sig);
    if (d_b1) (void)hipFree(d_b1);
    return rc;
}

extern "C" int mrf_lowp_prepare(void *stream, const mrf_complex *d_atoms, size_t T,
                                size_t NATOMS, void *d_hi, void *d_lo)
{
    mrf_clear_error();
    if (!d_atoms || !d_hi) return mrf_fail(MRF_E_ARG, "null pointer argument");
    if (T == 0 || T > (size_t)kLowpMaxT)
        return mrf_fail(MRF_E_ARG, "low-precision copies support 1 <= T <= %d", kLowpMaxT);
    if (int rc = check_device()) return rc;
    return launch_lowp_prepare(reinterpret_cast<const float2 *>(d_atoms), nullptr, NATOMS, (int)T,
                               (int)lowp_ld(T), (uint32_t *)d_hi, (uint32_t *)d_lo,
                               (hipStream_t)stream);
}

extern "C" size_t mrf_lowp_ld(size_t T)
{
    return lowp_ld(T);
}

// ---------------------------------------------------------------------------
// Differentiation form

namespace mrf {
namespace {

// Per-TR table of the continuous model: CSV unit transform
// (src/functions.cu:1242-1261) in double.
void build_dtr(const float *h_mrf, size_t nreps, size_t NREPS, std::vector<DTr> &out)
{
    out.resize(nreps);
    for (size_t i = 0; i < nreps; i++) {
        const double fa = (double)h_mrf[i] * (60.0 * M_PI / 180.0);
        const double ph = (double)h_mrf[NREPS + i] * (M_PI / 180.0);
        const double ph_next =
            i + 1 < nreps ? (double)h_mrf[NREPS + i + 1] * (M_PI / 180.0) : ph;
        const double tr = (double)h_mrf[2 * NREPS + i] + 16.0;
        const double te = (double)h_mrf[3 * NREPS + i] + 3.5;
        DTr d;
        d.al = (float)fa;
        d.cph = (float)cos(ph);
        d.sph = (float)sin(ph);
        d.rr = (float)cos(ph - ph_next);
        d.ri = (float)sin(ph - ph_next);
        d.te = (float)te;
        d.t2 = (float)(tr - te);
        d.pad = 0.f;
        out[i] = d;
    }
}

// The same sequence as build_dtr in the fused kernels' per-TR record, B1 = 1,
// with the flip angle in pad[0] (the Gauss-Newton trial passes form their own
// RF angle alpha_n * B1 per voxel).
void build_gn_coef(const float *h_mrf, size_t nreps, size_t NREPS, std::vector<TrCoef> &out)
{
    std::vector<DTr> d;
    build_dtr(h_mrf, nreps, NREPS, d);
    out.resize(nreps);
    for (size_t i = 0; i < nreps; i++) {
        const double ph = (double)h_mrf[NREPS + i] * (M_PI / 180.0);
        const double ph_next = i + 1 < nreps ? (double)h_mrf[NREPS + i + 1] * (M_PI / 180.0) : ph;
        TrCoef c{};
        c.ca2 = 1.f;
        c.zr = d[i].cph;
        c.zi = d[i].sph;
        c.rr = (float)cos(ph - ph_next);
        c.ri = (float)sin(ph - ph_next);
        c.te = d[i].te;
        c.tmte = d[i].t2;
        c.tr = d[i].te + d[i].t2;
        c.s2p = (float)sin(2.0 * ph);
        c.c2p = (float)cos(2.0 * ph);
        c.pad[0] = d[i].al;
        out[i] = c;
    }
}

constexpr float kFaUnit = (float)(60.0 * M_PI / 180.0);
constexpr size_t kDiffScratchBudget = (size_t)16 << 30;   // bytes per launch batch

// scratch bytes per launch batch: MRF_DIFF_SCRATCH_BYTES when set (read at
// every call), else 16 GiB
size_t diff_budget()
{
    if (const char *e = std::getenv("MRF_DIFF_SCRATCH_BYTES")) {
        char *end = nullptr;
        const unsigned long long v = std::strtoull(e, &end, 10);
        if (end != e && v > 0) return (size_t)v;
    }
    return kDiffScratchBudget;
}

// atoms per launch batch: at most the scratch budget allows, the batches of
// one call equal in size (multiples of 8 atoms, a work-group's worth) so the
// last launch is not a short tail
size_t diff_batch(size_t n, size_t per_atom)
{
    size_t cap = diff_budget() / per_atom;
    if (cap < 1) cap = 1;
    if (cap >= n) return n;
    const size_t nb = (n + cap - 1) / cap;
    size_t b = (n + nb - 1) / nb;
    b = (b + 7) / 8 * 8;
    return b <= cap ? b : cap;
}

int diff_check(const float *h_mrf, size_t nreps, size_t NREPS, int nstates)
{
    if (!h_mrf) return mrf_fail(MRF_E_ARG, "null sequence table");
    if (nreps == 0 || nreps > NREPS || nreps > 0x7fffffff)
        return mrf_fail(MRF_E_ARG, "need 0 < nreps <= NREPS");
    if (nstates < 2 || nstates > 64)
        return mrf_fail(MRF_E_ARG, "differentiation needs 2 <= nstates <= 64 (one lane per state)");
    return check_device();
}

// Uploads the per-TR table; *d must be freed by the caller on s.
int upload_dtr(const float *h_mrf, size_t nreps, size_t NREPS, hipStream_t s, DTr **d)
{
    std::vector<DTr> tab;
    build_dtr(h_mrf, nreps, NREPS, tab);
    MRF_HIP(hipMallocAsync((void **)d, nreps * sizeof(DTr), s));
    MRF_HIP(hipMemcpyAsync(*d, tab.data(), nreps * sizeof(DTr), hipMemcpyHostToDevice, s));
    return 0;
}

}  // namespace
}  // namespace mrf

extern "C" int mrf_epg_jacobian(void *stream, const float *h_mrf, size_t nreps, size_t NREPS,
                                const float *d_theta, size_t n, int nstates, mrf_complex *d_sig,
                                mrf_complex *d_jac)
{
    mrf_clear_error();
    if (int rc = diff_check(h_mrf, nreps, NREPS, nstates)) return rc;
    if (n && !d_theta) return mrf_fail(MRF_E_ARG, "null theta");
    if (n == 0 || (!d_sig && !d_jac)) return MRF_OK;
    hipStream_t s = (hipStream_t)stream;
    DTr *d_seq = nullptr;
    if (int rc = upload_dtr(h_mrf, nreps, NREPS, s, &d_seq)) return rc;
    int rc = launch_jacobian(d_seq, (int)nreps, nstates, d_theta, n,
                             reinterpret_cast<float2 *>(d_sig), reinterpret_cast<float2 *>(d_jac), s);
    (void)hipFreeAsync(d_seq, s);
    return rc;
}

extern "C" int mrf_epg_adjoint(void *stream, const float *h_mrf, size_t nreps, size_t NREPS,
                               const float *d_theta, size_t n, int nstates,
                               const mrf_complex *d_cot, float *d_grad_theta, float *d_grad_seq)
{
    mrf_clear_error();
    if (int rc = diff_check(h_mrf, nreps, NREPS, nstates)) return rc;
    if (n && (!d_theta || !d_cot)) return mrf_fail(MRF_E_ARG, "null theta or cotangent");
    if (n == 0 || (!d_grad_theta && !d_grad_seq)) return MRF_OK;
    hipStream_t s = (hipStream_t)stream;
    const int T = (int)nreps;
    const size_t per = diff_scratch_floats(0, T, nstates) * sizeof(float);
    const size_t batch = diff_batch(n, per);
    DTr *d_seq = nullptr;
    if (int rc = upload_dtr(h_mrf, nreps, NREPS, s, &d_seq)) return rc;
    float *scratch = nullptr;
    if (hipMallocAsync((void **)&scratch, batch * per, s) != hipSuccess) {
        (void)hipFreeAsync(d_seq, s);
        return mrf_fail(MRF_E_NOMEM, "adjoint scratch (%zu bytes)", batch * per);
    }
    int rc = 0;
    for (size_t a0 = 0; a0 < n && !rc; a0 += batch) {
        const size_t m = (n - a0 < batch) ? n - a0 : batch;
        rc = launch_adjoint(d_seq, T, nstates, d_theta + 4 * a0, m,
                            reinterpret_cast<const float2 *>(d_cot) + a0 * nreps,
                            d_grad_theta ? d_grad_theta + 4 * a0 : nullptr, d_grad_seq, kFaUnit,
                            scratch, s);
    }
    (void)hipFreeAsync(scratch, s);
    (void)hipFreeAsync(d_seq, s);
    return rc;
}

extern "C" int mrf_gn_refine(void *stream, const float *h_mrf, size_t nreps, size_t NREPS,
                             const mrf_complex *d_sig, size_t V, int nstates, int iters,
                             float *d_theta, float *d_pd, float *d_cost)
{
    mrf_clear_error();
    if (int rc = diff_check(h_mrf, nreps, NREPS, nstates)) return rc;
    if (V && (!d_sig || !d_theta)) return mrf_fail(MRF_E_ARG, "null signal or theta");
    if (iters < 0 || iters > 1000) return mrf_fail(MRF_E_ARG, "iters must be in [0, 1000]");
    if (V == 0) return MRF_OK;
    hipStream_t s = (hipStream_t)stream;
    DTr *d_seq = nullptr;
    if (int rc = upload_dtr(h_mrf, nreps, NREPS, s, &d_seq)) return rc;
    std::vector<TrCoef> coef;
    build_gn_coef(h_mrf, nreps, NREPS, coef);
    TrCoef *d_coef = nullptr;
    void *scratch = nullptr;
    if (hipMallocAsync((void **)&d_coef, nreps * sizeof(TrCoef), s) != hipSuccess ||
        hipMallocAsync(&scratch, gn_scratch_bytes(V), s) != hipSuccess) {
        (void)hipFreeAsync(d_coef, s);
        (void)hipFreeAsync(d_seq, s);
        return mrf_fail(MRF_E_NOMEM, "Gauss-Newton state (%zu bytes)", gn_scratch_bytes(V));
    }
    int rc = 0;
    if (hipMemcpyAsync(d_coef, coef.data(), nreps * sizeof(TrCoef), hipMemcpyHostToDevice, s) !=
        hipSuccess)
        rc = mrf_fail(MRF_E_GPU, "Gauss-Newton coefficient upload failed");
    if (!rc)
        rc = launch_gn(d_seq, d_coef, (int)nreps, nstates, reinterpret_cast<const float2 *>(d_sig),
                       V, d_theta, iters, d_pd, d_cost, scratch, s);
    (void)hipFreeAsync(scratch, s);
    (void)hipFreeAsync(d_coef, s);
    (void)hipFreeAsync(d_seq, s);
    return rc;
}

extern "C" int mrf_crb_loss_grad(void *stream, const float *h_mrf, size_t nreps, size_t NREPS,
                                 const float *d_tissues, size_t n, int nstates, const float *wts,
                                 double *d_loss, float *d_grad_seq)
{
    mrf_clear_error();
    if (int rc = diff_check(h_mrf, nreps, NREPS, nstates)) return rc;
    if (n && (!d_tissues || !d_loss || !wts)) return mrf_fail(MRF_E_ARG, "null argument");
    if (n == 0) return MRF_OK;
    hipStream_t s = (hipStream_t)stream;
    const int T = (int)nreps;
    const size_t per = diff_scratch_floats(4, T, nstates) * sizeof(float);
    const size_t batch = diff_batch(n, per);
    DTr *d_seq = nullptr;
    if (int rc = upload_dtr(h_mrf, nreps, NREPS, s, &d_seq)) return rc;
    float *scratch = nullptr;
    if (hipMallocAsync((void **)&scratch, batch * per, s) != hipSuccess) {
        (void)hipFreeAsync(d_seq, s);
        return mrf_fail(MRF_E_NOMEM, "crb scratch (%zu bytes)", batch * per);
    }
    int rc = 0;
    for (size_t a0 = 0; a0 < n && !rc; a0 += batch) {
        const size_t m = (n - a0 < batch) ? n - a0 : batch;
        rc = launch_crb(d_seq, T, nstates, d_tissues + 4 * a0, m, wts, d_loss + a0, d_grad_seq,
                        kFaUnit, scratch, s);
    }
    (void)hipFreeAsync(scratch, s);
    (void)hipFreeAsync(d_seq, s);
    return rc;
}
