// kernels.h -- host-side launchers of the gfx950 kernels (implemented in the
// .hip translation units), called by the C-ABI layer in capi.hip.
#ifndef SNAPMRF_KERNELS_H
#define SNAPMRF_KERNELS_H

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>

#include "internal.h"

namespace mrf {

struct EpgLaunch {
    const TrCoef *coef;    // device [n_slices][T]
    const float *params;   // device [A][4]
    float2 *atoms;         // device [n_slices * natoms][T]
    double *norm2;         // device [n_slices * natoms] or nullptr
    size_t natoms;         // atoms per slice
    int n_slices;
    int T;
    int w;                 // EPG states
    int echo;              // Echo
    bool generic;          // force the unfused reference-order path
    bool h_real;           // constant RF phase and no B0 rotation (FISP)
    bool normalize;        // unit L2 norm per atom (norm2 must be non-null)
    int lanes;             // lanes per atom (1, 2, 4, 8, 16); 0 = chosen from the launch size
};

// Lane form of a launch: lanes per atom *L and states per lane *M.  With
// lanes == 0 the fused path takes the fewest lanes whose compiled M covers
// w, then spreads launches of fewer than 2^17 lanes over more lanes per
// atom (at least 8 states each).  Two launches of the same form compute
// every atom with the same FP32 operations.
int epg_form(const EpgLaunch &a, int *L, int *M);

// EPG dictionary: SSFP/FISP fused path or SR/SE/TSE reference-order path.
int launch_epg(const EpgLaunch &a, hipStream_t s);

struct RoaLaunch {
    const TrCoef *coef;    // device [T] (unscaled FA: the reference ignores B1)
    const float *params;   // device [natoms][4]
    float2 *atoms;         // device [natoms][T]
    double *norm2;         // device [natoms] or nullptr
    size_t natoms;
    int T;
    int n_iso;
    float iso_spread_hz;
};
int launch_roa(const RoaLaunch &a, hipStream_t s);

// atoms[a][t] *= 1/sqrt(norm2[a]) for norm2 > 0 (functions.cu:670-685).
int launch_normalize(float2 *atoms, const double *norm2, size_t n, int T, hipStream_t s);

// Matching ---------------------------------------------------------------

struct MatchLaunch {
    const float2 *atoms;        // [NATOMS][T]
    const uint32_t *lowp_hi;    // [NATOMS][lowp_ld(T)] bf16x2 or nullptr
    const uint32_t *lowp_lo;    // [NATOMS][lowp_ld(T)] bf16x2 or nullptr
    const float *params;        // [NATOMS][4]
    const float2 *sig;          // [V][T]
    const int32_t *vox_slice;   // [V] slice index per voxel, or nullptr = all slices
    size_t natoms;              // atoms per slice
    int n_slices;
    int T;
    size_t V;
    int mode;                   // MRF_MATCH_* (without MRF_MATCH_RESCORE_FP32)
    bool rescore_fp32;          // low-precision modes: rescore sums in FP32 (default FP64)
    float *maps;                // [V][5]
    int32_t *idx;               // [V] or nullptr
};
// Scratch size in bytes a matching call needs.
size_t match_scratch_bytes(const MatchLaunch &a);
int launch_match(const MatchLaunch &a, void *scratch, hipStream_t s);
// Matching of a dictionary that exists one block at a time: for each block,
// gen(a0, nb) fills `buf` [n_slices][nb][T] with atoms a0 .. a0+nb-1 of every
// slice (nb <= block), which is then matched; a.atoms, a.lowp_hi/lo unused.
// FP32: the keys merge across blocks, so the maps equal launch_match's on
// the whole dictionary.  Low-precision modes: each block is converted to
// BF16 (lb->hi, lb->lo: [n_slices * block][lowp_ld(T)] words), its candidates
// are rescored against its complex64 rows and merged into a running best
// per voxel (lb->best_s / best_i, [V]); the candidate set is a superset of
// the whole-dictionary call's, so the pick scores at least as high.  Scratch
// as match_scratch_bytes(a).
struct LowpBlockBufs {
    uint32_t *hi, *lo;
    double *best_s;
    uint32_t *best_i;
};
int launch_match_blocks(const MatchLaunch &a, size_t block, float2 *buf,
                        const LowpBlockBufs *lb, void *scratch, hipStream_t s,
                        const std::function<int(size_t, size_t)> &gen);

// Row stride (bf16x2 words) of the low-precision copies: T rounded up to a
// multiple of 32, the pad zero-filled.
constexpr int kLowpMaxT = 1 << 20;
inline size_t lowp_ld(size_t T) { return (T + 31) / 32 * 32; }

// hi[r][t] = bf16(src[perm ? perm[r] : r][t]), lo = bf16(src - hi) (optional),
// rows of lowp_ld(T) words, t >= T zero.
int launch_lowp_prepare(const float2 *src, const int32_t *perm, size_t rows, int T, int ld,
                        uint32_t *hi, uint32_t *lo, hipStream_t s);

// Voxel order of the candidate rescore (order.hip): temporary bytes of a
// radix sort of V (key, voxel) pairs, and the sort by the low `bits` bits.
size_t voxel_sort_tmp_bytes(size_t V);
int sort_voxels(void *tmp, size_t tmp_bytes, const uint32_t *keys_in, uint32_t *keys_out,
                const uint32_t *vox_in, uint32_t *vox_out, size_t V, int bits, hipStream_t s);

// Differentiation form (grad.hip) --------------------------------------

// per-TR table of the continuous EPG-SSFP model (unit transform in double).
// The kernels run TR n in the frame of its RF phase (F+ e^{-i phi_n},
// F- e^{i phi_n}), where the RF has zero phase.
struct DTr {
    float al;        // flip angle (rad) before B1
    float cph, sph;  // e^{i phi_n}: readouts of TR n back to true coordinates
    float rr, ri;    // e^{i (phi_n - phi_{n+1})} (phi_T := phi_{T-1}): frame n -> n+1
    float te, t2;    // TE and TR - TE (ms)
    float pad;
};
// scratch floats per atom of launch_adjoint (nd = 0) / launch_crb (nd = 4)
size_t diff_scratch_floats(int nd, int T, int w);
int launch_jacobian(const DTr *seq, int T, int w, const float *theta, size_t n, float2 *sig,
                    float2 *jac, hipStream_t s);
int launch_adjoint(const DTr *seq, int T, int w, const float *theta, size_t n, const float2 *cot,
                   float *grad_theta, float *grad_seq, float fa_unit, float *scratch, hipStream_t s);
int launch_crb(const DTr *seq, int T, int w, const float *theta, size_t n, const float *wts,
               double *loss, float *grad_seq, float fa_unit, float *scratch, hipStream_t s);
// Gauss-Newton state record: float offsets of the trial step (T1, T2, B0,
// B1), its cost, |x|^2 and its sums s^H s, s^H x (re, im).
constexpr int kGnRecTrial = 4, kGnRecCt = 9, kGnRecXx = 10, kGnRecHs = 14, kGnRecHsx = 15;
// Trial cost of every voxel by the one-voxel-per-lane EPG-SSFP recursion
// (coef: [T] coefficients at B1 = 1 with the flip angle in pad[0]); returns 1
// when nstates is not a state count it is compiled for.
int launch_gn_trial(const TrCoef *coef, int T, int w, const float2 *sig, size_t V, float *rec,
                    int stride, hipStream_t s);
// scratch bytes of launch_gn for V voxels
size_t gn_scratch_bytes(size_t V);
int launch_gn(const DTr *seq, const TrCoef *coef, int T, int w, const float2 *sig, size_t V,
              float *theta, int iters, float *pd, float *cost, void *scratch, hipStream_t s);

// B1 slice per voxel: argmin_k |b1map[v] - slice_b1[k]| (strict <, init 10).
int launch_b1_select(const float *b1map, const float *params, size_t natoms, int n_slices,
                     size_t V, int32_t *vox_slice, hipStream_t s);

}  // namespace mrf

#endif
