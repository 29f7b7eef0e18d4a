// epg.hip -- EPG and ROA dictionary kernels for gfx950.
//
// Layout and mapping
//   * One atom per lane group of L lanes (L = 1, 2, 4, 8 or 16, inside one
//     16-lane DPP row).  Lane j of a group holds EPG states k = j*M .. j*M+M-1
//     (M is a compile-time constant) as F+_k, F-_k, Z_k complex values in
//     VGPRs.  The gradient shift crosses lanes only at the group-internal
//     boundaries, with one DPP row_shr / row_shl per value.
//   * The TR loop runs inside the kernel; the per-TR, per-B1-slice RF
//     coefficients come from a 64-byte TrCoef record.  One-lane-per-atom
//     waves read each TR's factors from an LDS table that the wave's lanes
//     fill together for each group of atoms sharing (T2, B0); multi-lane
//     groups form them per lane from TrCoef records read with uniform
//     (scalar) loads.
//   * Output atoms are [slice][atom][t] complex64 (RA dims (T, A), time
//     fastest), the reference's layout (src/functions.cu:609-617).
//
// Fused SSFP/FISP step ("frame" form).  Write the RF of TR n as
// P_n M0_n P_n^-1 with P_n = diag(e^{i phi_n}, e^{-i phi_n}, 1) and M0_n the
// Weigel matrix at zero phase (src/functions.cu:483-508).  Relaxation, B0
// precession and the gradient shift all commute with P_n, so in coordinates
// x~ = P_n^-1 x the step of TR n is
//     (P+, P-, Q) = M0_n (F~+, F~-, Z)                       (per state)
//     F~+_k <- H P+_{k-1},  F~+_0 <- H conj(P-_1),  F~-_k <- conj(H) P-_{k+1},
//     F~-_{w-1} <- 0,  Z_k <- E1 Q_k (+ 1 - E1 for k = 0)
//     sample = E2(TE) e^{i theta_TE} e^{i phi_n} P+_0
// with H = E2(TE) E2(TR-TE) e^{i(theta_TE + theta_TR-TE)} e^{i(phi_n - phi_{n+1})}.
// This is algebraically the reference's RF, decay(TE), phase(TE), sample,
// phase(TR-TE), dephase, decay(TR-TE) sequence (src/functions.cu:975-1009),
// at 24 VALU operations per state instead of the 102 flops written.
//
// The SR/SE/TSE programs (and SSFP/FISP when `generic` is requested) run the
// reference's step order unfused (src/functions.cu:687-933).
#include "kernels.h"
#include "snapmrf.h"

#include <hip/hip_runtime.h>

#include <cmath>

namespace mrf {
namespace {

constexpr int kBlock = 256;

struct EpgArgs {
    const TrCoef *coef;
    const float *params;
    float2 *atoms;
    double *norm2;
    size_t natoms;
    int T;
    int L;          // lanes per atom (power of two, <= 16)
    int lshift;     // log2(L)
    int istar;      // truncation index of state w-1 inside its lane, -1 if none
    int jstar;      // lane of state w-1 inside the group
    float init_alpha, init_phi, TI;   // initial pulse (float32 radians / ms)
    int normalize;  // scale each atom to unit L2 norm inside the kernel
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// conj(a) * b
__device__ __forceinline__ float2 cmulc(float2 a, float2 b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// Lane i receives lane i-1 of its 16-lane row (row_shr:1); the first lane of
// each row receives 0.
__device__ __forceinline__ float shr1(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                 0x111, 0xF, 0xF, true));
}

// Lane i receives lane i+1 of its 16-lane row (row_shl:1); the last lane of
// each row receives 0.
__device__ __forceinline__ float shl1(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                 0x101, 0xF, 0xF, true));
}

// B0 precession angle as the reference forms it: 2*M_PI*B0*t/1000 evaluated
// in double and handed to a float (src/functions.cu:598-600).
__device__ __forceinline__ float phase_angle(float B0, float t)
{
    return (float)(2.0 * M_PI * (double)B0 * (double)t / 1000.0);
}

// sin and cos of 2*pi*r for r in revolutions.  The range reduction runs in
// double (r - rint(r), quadrant split, scaling by 2*pi), so the float
// argument of the polynomials carries only its own final rounding; the
// polynomials are the Cephes single-precision minimax fits on [-pi/4, pi/4].
__device__ __forceinline__ void sincos_rev(double r, float *s, float *c)
{
    r -= rint(r);                                   // [-1/2, 1/2]
    const double q = rint(4.0 * r);                 // quadrant, -2..2
    const float x = (float)((r - 0.25 * q) * 6.283185307179586);   // [-pi/4, pi/4]
    const float x2 = x * x;
    const float sp = x + x * x2 * fmaf(x2, fmaf(x2, -1.9515295891e-4f, 8.3321608736e-3f),
                                       -1.6666654611e-1f);
    const float cp = fmaf(x2 * x2, fmaf(x2, fmaf(x2, 2.443315711809948e-5f,
                                                 -1.388731625493765e-3f),
                                        4.166664568298827e-2f),
                          fmaf(x2, -0.5f, 1.0f));
    const int iq = (int)q & 3;
    const float ss = (iq & 1) ? cp : sp;
    const float cc = (iq & 1) ? sp : cp;
    *s = (iq & 2) ? -ss : ss;
    *c = ((iq + 1) & 2) ? -cc : cc;
}

// exp(-t/T) as 2^(t*k) with k = -log2(e)/T formed once per atom in double.
__device__ __forceinline__ float decay_k(float T)
{
    return (float)(-1.4426950408889634 / (double)T);
}

__device__ __forceinline__ float decay(float t, float k)
{
    return __builtin_amdgcn_exp2f(t * k);
}

template <int M>
struct EpgState {
    float2 fp[M], fm[M], z[M];
};

// Initial pulse of init_rf_pulse (src/functions.cu:456-481), written to
// state 0 (lane j == 0) in true coordinates.
template <int M>
__device__ __forceinline__ void epg_init(EpgState<M> &s, const EpgArgs &a, int j, float T1,
                                         float T2)
{
#pragma unroll
    for (int i = 0; i < M; i++) {
        s.fp[i] = make_float2(0.f, 0.f);
        s.fm[i] = make_float2(0.f, 0.f);
        s.z[i] = make_float2(0.f, 0.f);
    }
    if (j == 0) {
        float e1 = expf(-a.TI / T1), e2 = expf(-a.TI / T2);
        float sa, ca, sph, cph;
        sincosf(a.init_alpha, &sa, &ca);
        sincosf(a.init_phi, &sph, &cph);
        float2 tmp = make_float2(e2 * sa * sph, e2 * sa * cph);
        s.fp[0] = make_float2(tmp.x, -tmp.y);
        s.fm[0] = tmp;
        s.z[0] = make_float2(e1 * ca + 1.f - e1, 0.f);
    }
}

// Zero F-_{w-1} when the lane grid holds more than w states.
template <int M>
__device__ __forceinline__ void epg_truncate(EpgState<M> &s, const EpgArgs &a, int j)
{
    if (a.istar >= 0) {
#pragma unroll
        for (int i = 0; i < M; i++)
            if (i == a.istar && j == a.jstar) s.fm[i] = make_float2(0.f, 0.f);
    }
}

// ------------------------------------------------------------------------
// Fused SSFP / FISP kernel
// ------------------------------------------------------------------------

// TR loop unroll factor.  Within one TR the states are swept from the top
// (k = M-1) down: new F+_{k+1} overwrites a register whose old value was
// consumed one step earlier, and new F-_{k-1} overwrites the register of the
// old F-_k read in the same step.  The F- array therefore drifts up by one
// index per TR (slot k lives at fm[k + r] after r TRs) and is copied back
// once per kUnroll TRs, so neither shift needs per-TR register moves.
constexpr int kUnroll = 2;

template <int M>
struct FusedState {
    float2 fp[M];
    float2 fm[M + kUnroll];
    float2 z[M];
};

struct FusedCtx {
    float k1, k2;     // -log2(e)/T1, -log2(e)/T2
    double b0r;       // B0 in revolutions per ms
    int j;            // lane within the atom's group
};

// Factors of one TR that depend on the atom's T2 and B0 but not on T1:
// (H.x, H.y, csave.x, csave.y) with
//     H     = E2(TE) E2(TR-TE) e^{i(theta_TE + theta_TR-TE)} e^{i(phi_n - phi_{n+1})} / 2,
//     csave = E2(TE) e^{i theta_TE} e^{i phi_n} / 2
// (k2 = -log2(e)/T2, b0r = B0 in revolutions per ms).
template <bool PHASE, bool HREAL>
__device__ __forceinline__ float4 group_factors(const TrCoef &c, float k2, double b0r)
{
    const float ea = decay(c.te, k2);
    const float eb = decay(c.tmte, k2);
    float2 H, csave;
    if (PHASE) {
        // B0 precession over TE and TR-TE (src/functions.cu:585-607)
        float sa_, ca_, sb_, cb_;
        sincos_rev(b0r * (double)c.te, &sa_, &ca_);
        sincos_rev(b0r * (double)c.tmte, &sb_, &cb_);
        const float2 e = cmul(make_float2(ca_, sa_), make_float2(cb_, sb_));
        const float g = 0.5f * ea * eb;
        const float2 er = cmul(e, make_float2(c.rr, c.ri));
        H = make_float2(g * er.x, g * er.y);
        const float2 ez = cmul(make_float2(ca_, sa_), make_float2(c.zr, c.zi));
        const float h = 0.5f * ea;
        csave = make_float2(h * ez.x, h * ez.y);
    } else {
        const float g = 0.5f * ea * eb;
        H = HREAL ? make_float2(g, 0.f) : make_float2(g * c.rr, g * c.ri);
        const float h = 0.5f * ea;
        csave = make_float2(h * c.zr, h * c.zi);
    }
    return make_float4(H.x, H.y, csave.x, csave.y);
}

// Factors of one TR that depend on the B1 slice only: (cos a, 2 sin a,
// sin(a)/2, TR).
__device__ __forceinline__ float4 slice_factors(const TrCoef &c)
{
    return make_float4(c.ca, 2.0f * c.sa, 0.5f * c.sa, c.tr);
}

// One TR of the fused frame-form step with the F- array offset R, given the
// group factors gf (group_factors) and the slice factors sf (slice_factors);
// returns the sample E2(TE) e^{i theta_TE} e^{i phi_n} P+_0.
//
// The zero-phase RF is evaluated through S = A + B and D = A - B (A, B, C =
// F~+, F~-, Z of a state): since (1+cos a)/2 + (1-cos a)/2 = 1 and their
// difference is cos a,
//     P+ = (S + W)/2,  P- = (S - W)/2,  W = cos(a) D + 2 sin(a) (C.y, -C.x),
//     Q  = cos(a) C + sin(a)/2 (D.y, -D.x),
// and the exact factor 1/2 is folded into H and into the sample factor.  That
// is 24 operations per state (20 with a real H), 10 of them additions.
template <int M, int R, bool MULTI, bool HREAL, bool TRUNC>
__device__ __forceinline__ float2 fused_core(FusedState<M> &s, float4 gf, float4 sf,
                                             const FusedCtx &x, const EpgArgs &a)
{
    const float2 H = make_float2(gf.x, HREAL ? 0.f : gf.y);
    const float2 csave = make_float2(gf.z, gf.w);
    const float E1 = decay(sf.w, x.k1);
    const float ca = sf.x, s2 = sf.y;
    const float eca = E1 * sf.x, eh = E1 * sf.z;

    float2 top = make_float2(0.f, 0.f), d0 = top, fp0 = top, u0 = top;
#pragma unroll
    for (int i = M - 1; i >= 0; i--) {
        const float2 A = s.fp[i], B = s.fm[i + R], C = s.z[i];
        const float sx = A.x + B.x, sy = A.y + B.y;
        const float dx = A.x - B.x, dy = A.y - B.y;
        const float wx = ca * dx + s2 * C.y, wy = ca * dy - s2 * C.x;
        const float2 U = make_float2(sx + wx, sy + wy);    // 2 P+
        const float2 V = make_float2(sx - wx, sy - wy);    // 2 P-
        s.z[i] = make_float2(eca * C.x + eh * dy, eca * C.y - eh * dx);
        const float2 hp = HREAL ? make_float2(H.x * U.x, H.x * U.y) : cmul(H, U);
        const float2 hm = HREAL ? make_float2(H.x * V.x, H.x * V.y) : cmulc(H, V);
        if (i == M - 1) top = hp;
        else s.fp[i + 1] = hp;            // new F+_{i+1}
        if (i >= 1) s.fm[i + R] = hm;     // new F-_{i-1}, at index (i-1) + (R+1)
        else d0 = hm;
        if (i == 1)
            fp0 = HREAL ? make_float2(H.x * V.x, -H.x * V.y) : cmul(H, make_float2(V.x, -V.y));
        if (i == 0) u0 = U;
    }
    if (x.j == 0) s.z[0].x += 1.0f - E1;
    if (MULTI) {
        const float2 up = make_float2(shr1(top.x), shr1(top.y));
        const float2 dn = make_float2(shl1(d0.x), shl1(d0.y));
        s.fp[0] = (x.j == 0) ? fp0 : up;
        s.fm[M + R] = (x.j == a.L - 1) ? make_float2(0.f, 0.f) : dn;
    } else {
        s.fp[0] = fp0;
        s.fm[M + R] = make_float2(0.f, 0.f);
    }
    if (TRUNC) {
#pragma unroll
        for (int i = 0; i < M; i++)
            if (i == a.istar && x.j == a.jstar) s.fm[i + R + 1] = make_float2(0.f, 0.f);
    }
    return cmul(csave, u0);
}

// One TR with the factors formed by the lane itself.
template <int M, int R, bool MULTI, bool PHASE, bool HREAL, bool TRUNC>
__device__ __forceinline__ float2 fused_step(FusedState<M> &s, const TrCoef &c,
                                             const FusedCtx &x, const EpgArgs &a)
{
    return fused_core<M, R, MULTI, HREAL, TRUNC>(s, group_factors<PHASE, HREAL>(c, x.k2, x.b0r),
                                                 slice_factors(c), x, a);
}

template <int M, int SH>
__device__ __forceinline__ void fm_copy_down(FusedState<M> &s)
{
#pragma unroll
    for (int k = 0; k < M; k++) s.fm[k] = s.fm[k + SH];
}

// Output staging.  Each atom's samples go to its row of a per-wave LDS tile
// (64 / L rows for L lanes per atom; the group's lane 0 holds the sample);
// every kSeg TRs the wave writes its rows as whole 128-byte segments (8 lanes
// x 16 bytes per row), so every HBM line is written once and completely.
constexpr int kSeg = 16;
constexpr int kSegStride = 2 * kSeg + 2;   // floats per LDS row (padding spreads banks)

// Atom rows are written once (and, with in-kernel normalisation, read back
// once by the same wave): non-temporal stores keep them from displacing the
// TR coefficient records in the L2.  Measured on the 871,455-atom w = 30
// launch: HBM traffic 8.49 -> 7.06 GB per launch (6.97 GB algorithmic; the
// coefficient re-fetches went from 0.53 to 0.02 GB), 1.7 % faster.
typedef float ntf4 __attribute__((ext_vector_type(4)));
typedef float ntf2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void nt_store(float4 v, float4 *p)
{
    __builtin_nontemporal_store(ntf4{v.x, v.y, v.z, v.w}, reinterpret_cast<ntf4 *>(p));
}

__device__ __forceinline__ void nt_store(float2 v, float2 *p)
{
    __builtin_nontemporal_store(ntf2{v.x, v.y}, reinterpret_cast<ntf2 *>(p));
}

struct WaveStage {
    float *lds;        // this wave's [rows][kSegStride] floats
    float2 *rows;      // first atom row of this wave
    int T;
    int nvalid;        // atom rows of this wave that exist
    int lane;
};

__device__ __forceinline__ void stage_put(const WaveStage &w, int row, int slot, float2 v)
{
    *reinterpret_cast<float2 *>(&w.lds[row * kSegStride + 2 * slot]) = v;
}

// Writes TRs [t0, t0 + count) of the wave's rows.
__device__ __forceinline__ void stage_flush(const WaveStage &w, int t0, int count, bool even_t)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int p = w.lane & 7;
#pragma unroll
    for (int it = 0; it < 8; it++) {
        const int row = it * 8 + (w.lane >> 3);
        if (row < w.nvalid) {
            const float *src = &w.lds[row * kSegStride + 4 * p];
            float2 *dst = w.rows + (size_t)row * w.T + t0 + 2 * p;
            const float2 a = *reinterpret_cast<const float2 *>(src);
            const float2 b = *reinterpret_cast<const float2 *>(src + 2);
            if (2 * p + 1 < count) {
                if (even_t) {
                    nt_store(make_float4(a.x, a.y, b.x, b.y), reinterpret_cast<float4 *>(dst));
                } else {
                    nt_store(a, dst);
                    nt_store(b, dst + 1);
                }
            } else if (2 * p < count) {
                nt_store(a, dst);
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float4 scale_v(float4 v, float f)
{
    return make_float4(v.x * f, v.y * f, v.z * f, v.w * f);
}
__device__ __forceinline__ float2 scale_v(float2 v, float f) { return make_float2(v.x * f, v.y * f); }

// Scales the wave's nvalid contiguous rows by the factors in w.lds[row];
// VEC elements, per_row of them per row, kDepth loads in flight per lane.
template <typename VEC, int kDepth>
__device__ __forceinline__ void normalize_rows(const WaveStage &w, int per_row)
{
    VEC *base = reinterpret_cast<VEC *>(w.rows);
    const int total = w.nvalid * per_row;
    const float inv = 1.0f / (float)per_row;
    for (int i0 = w.lane; i0 < total; i0 += 64 * kDepth) {
        VEC v[kDepth];
#pragma unroll
        for (int u = 0; u < kDepth; u++) {
            const int idx = i0 + 64 * u;
            if (idx < total) v[u] = base[idx];
        }
#pragma unroll
        for (int u = 0; u < kDepth; u++) {
            const int idx = i0 + 64 * u;
            if (idx < total) {
                int r = (int)((float)idx * inv);
                r += (r + 1) * per_row <= idx;
                r -= r * per_row > idx;
                nt_store(scale_v(v[u], w.lds[r]), &base[idx]);
            }
        }
    }
}

// Per-TR factor tables.  The per-TR factors that do not depend on T1 (two
// exp2, two sincos with their FP64 range reduction, the complex products of
// group_factors) are the same for every atom of a wave with the same
// (T2, B0): in the reference's atom order (T1 fastest) a wave holds one or
// two such groups.  The wave finds its G groups, and its lanes form the group
// factors of a chunk of TRs together (G x chunk entries, chunk = 32 TRs for
// G <= 4, fewer for more groups, so a chunk never has more than 128
// entries), together with the slice factors of the chunk's TRs; the TR loop
// then reads each TR's factors from LDS with two broadcast loads instead of
// forming them per lane.  Two buffers: the next chunk is filled while the
// current one is read.
constexpr int kTabGroupEntries = 128;
constexpr int kTabChunkMax = 32;

struct TabBuf {
    float4 grp[kTabGroupEntries];   // [group][TR of the chunk] group factors
    float4 trc[kTabChunkMax];       // [TR of the chunk] slice factors
};

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Groups of equal (T2, B0) bit patterns among the 64 lanes of a wave.
// Returns G (1..64); *my_g is the lane's group and lane g < G holds in
// *lead the lowest lane of group g.
__device__ __forceinline__ int wave_groups(float T2, float B0, int lane, int *my_g, int *lead)
{
    const int kt = __float_as_int(T2), kb = __float_as_int(B0);
    unsigned long long todo = ~0ull;
    int G = 0;
    *my_g = 0;
    *lead = 0;
    while (todo) {
        // the lowest unassigned lane founds a group; a lane that matches it
        // is unassigned, since earlier groups hold other keys
        const int ld = __builtin_ctzll(todo);
        const int lt = __shfl(kt, ld), lb = __shfl(kb, ld);
        const bool same = kt == lt && kb == lb;
        const unsigned long long m = __ballot(same);
        if (same) *my_g = G;
        if (lane == G) *lead = ld;
        todo &= ~m;
        G++;
    }
    return G;
}

// Fills b with the factors of TRs [t0, t0 + 2^cs) (clamped to T - 1 past the
// end) for the wave's G groups; called by all 64 lanes.
template <bool PHASE, bool HREAL>
__device__ __forceinline__ void fill_table(TabBuf &b, const TrCoef *__restrict__ cf, int T, int t0,
                                           int cs, int G, int lead, float T2, float B0, int lane)
{
    const int chunk = 1 << cs;
    const int ne = G << cs;
    for (int e0 = 0; e0 < ne; e0 += 64) {
        const int e = e0 + lane;
        const int g = min(e >> cs, G - 1);
        const int ld = __shfl(lead, g);
        const float t2 = __shfl(T2, ld), b0 = __shfl(B0, ld);
        const int t = min(t0 + (e & (chunk - 1)), T - 1);
        const float4 f = group_factors<PHASE, HREAL>(cf[t], decay_k(t2), (double)b0 * 1e-3);
        if (e < ne) b.grp[e] = f;
    }
    if (lane < chunk) b.trc[lane] = slice_factors(cf[min(t0 + lane, T - 1)]);
}

template <int M, bool MULTI, bool PHASE, bool HREAL, bool TRUNC>
__device__ __forceinline__ void epg_fused_body(const EpgArgs &a, const TrCoef *__restrict__ coef,
                                               const float *__restrict__ params,
                                               float2 *__restrict__ atoms)
{
    const size_t gl = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t atom = MULTI ? (gl >> a.lshift) : gl;
    FusedCtx x;
    x.j = MULTI ? (int)(gl & (size_t)(a.L - 1)) : 0;
    const int slice = blockIdx.y;
    const bool valid = atom < a.natoms;
    const size_t pa = valid ? atom : 0;
    // SSFP kernels index the slice-0 parameter rows (src/functions.cu:462-463,
    // 537-538, 596): T1/T2/B0 repeat identically in every B1 slice.
    const float T1 = params[pa * 4 + 0];
    const float T2 = params[pa * 4 + 1];
    const float B0 = params[pa * 4 + 2];
    x.k1 = decay_k(T1);
    x.k2 = decay_k(T2);
    x.b0r = (double)B0 * 1e-3;

    const TrCoef *__restrict__ cf = coef + (size_t)slice * a.T;
    FusedState<M> s;
    {
        EpgState<M> s0;
        epg_init(s0, a, x.j, T1, T2);
        // to frame coordinates of TR 0
        const float2 z0 = make_float2(cf[0].zr, cf[0].zi);
#pragma unroll
        for (int i = 0; i < M; i++) {
            s.fp[i] = s0.fp[i];
            s.fm[i] = s0.fm[i];
            s.z[i] = s0.z[i];
        }
#pragma unroll
        for (int i = M; i < M + kUnroll; i++) s.fm[i] = make_float2(0.f, 0.f);
        s.fp[0] = cmulc(z0, s.fp[0]);
        s.fm[0] = cmul(z0, s.fm[0]);
    }

    const bool store = valid && x.j == 0;
    const bool even_t = (a.T & 1) == 0;
    double nrm = 0.0;
    int n = 0;
    // the wave's atoms are contiguous rows: 64 / L of them
    constexpr int kRows = MULTI ? 32 : 64;
    const int lshift = MULTI ? a.lshift : 0;
    __shared__ __attribute__((aligned(16))) float stage_lds[kBlock / 64][kRows * kSegStride];
    const size_t first = ((size_t)blockIdx.x * kBlock + (threadIdx.x & ~63u)) >> lshift;
    const int rows = 64 >> lshift;
    WaveStage w;
    w.lds = stage_lds[threadIdx.x >> 6];
    w.rows = atoms + ((size_t)slice * a.natoms + first) * (size_t)a.T;
    w.T = a.T;
    w.nvalid = first < a.natoms ? (int)((a.natoms - first) < (size_t)rows ? (a.natoms - first) : rows)
                                : 0;
    w.lane = threadIdx.x & 63;
    const int row = w.lane >> lshift;
    __shared__ TabBuf tab_lds[kBlock / 64][2];
    TabBuf *tb = tab_lds[threadIdx.x >> 6];
    int my_g, lead;
    const int G = wave_groups(T2, B0, w.lane, &my_g, &lead);
    // chunk = 2^cs TRs: 32 for G <= 4, halved per doubling of G (2 for 64)
    const int cs = min(5, 7 - (32 - __builtin_clz((unsigned)(2 * G - 1)) - 1));
    const int chunk = 1 << cs;
    fill_table<PHASE, HREAL>(tb[0], cf, a.T, 0, cs, G, lead, T2, B0, w.lane);
    // one flat TR loop (a nested chunk loop makes the register allocator move
    // the state between the loops); at the first TR of each chunk the next
    // chunk's table goes into the other buffer
    int t0 = 0;    // first TR of the staged segment
    int kk = 0;    // TR index within the chunk
    int cb = 0;    // buffer of the current chunk
    for (; n + kUnroll <= a.T; n += kUnroll) {
        if (kk == 0) {
            wave_lds_sync();
            if (n + chunk < a.T)
                fill_table<PHASE, HREAL>(tb[cb ^ 1], cf, a.T, n + chunk, cs, G, lead, T2, B0,
                                         w.lane);
            wave_lds_sync();
        }
        const TabBuf &tab = tb[cb];
        const float4 *gp = &tab.grp[my_g << cs];
        const float2 v0 = fused_core<M, 0, MULTI, HREAL, TRUNC>(s, gp[kk], tab.trc[kk], x, a);
        const float2 v1 = fused_core<M, 1, MULTI, HREAL, TRUNC>(s, gp[kk + 1], tab.trc[kk + 1], x,
                                                                a);
        fm_copy_down<M, 2>(s);
        if (x.j == 0) {
            stage_put(w, row, n - t0, v0);
            stage_put(w, row, n + 1 - t0, v1);
        }
        nrm += (double)v0.x * (double)v0.x + (double)v0.y * (double)v0.y;
        nrm += (double)v1.x * (double)v1.x + (double)v1.y * (double)v1.y;
        kk += kUnroll;
        if (kk == chunk) {
            kk = 0;
            cb ^= 1;
        }
        if (n + 2 - t0 == kSeg) {
            stage_flush(w, t0, kSeg, even_t);
            t0 += kSeg;
        }
    }
    if (n < a.T) {   // odd T: the last TR (its chunk's table is filled)
        wave_lds_sync();
        const float2 v = fused_core<M, 0, MULTI, HREAL, TRUNC>(s, tb[cb].grp[(my_g << cs) + kk],
                                                               tb[cb].trc[kk], x, a);
        fm_copy_down<M, 1>(s);
        if (x.j == 0) stage_put(w, row, n - t0, v);
        nrm += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
        n++;
    }
    if (a.T > t0) stage_flush(w, t0, a.T - t0, even_t);
    if (a.normalize) {
        // normalize (src/functions.cu:670-685) on the wave's own rows, which
        // are contiguous: per-row factors go to LDS, then the wave streams the
        // block with 32 independent loads in flight per lane.  (the wave
        // re-reads its own stores through its CU's caches: workgroup scope
        // suffices and avoids an L2 write-back)
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        if (x.j == 0) w.lds[row] = nrm > 0.0 ? (float)(1.0 / sqrt(nrm)) : 1.0f;
        wave_lds_sync();
        // (16 loads in flight for the lane-group forms, whose 3 waves per
        // SIMD leave 168 registers)
        constexpr int kDepth = MULTI ? 16 : 32;
        if (even_t) normalize_rows<float4, kDepth>(w, a.T / 2);
        else normalize_rows<float2, kDepth>(w, a.T);
    }
    if (store && a.norm2) a.norm2[(size_t)slice * a.natoms + atom] = nrm;
}

// Occupancy classes: without truncation the state fits 256 VGPRs (2 waves
// per SIMD), or 168 (3 waves) for up to 16 states per lane; the truncating
// variants get the full register file.
template <int M, bool MULTI, bool PHASE, bool HREAL>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(M <= 16 ? 3 : 2)))
void epg_fused_kernel(EpgArgs a, const TrCoef *__restrict__ coef,
                      const float *__restrict__ params, float2 *__restrict__ atoms)
{
    epg_fused_body<M, MULTI, PHASE, HREAL, false>(a, coef, params, atoms);
}

template <int M, bool MULTI, bool PHASE, bool HREAL>
__global__ __launch_bounds__(kBlock) void epg_fused_trunc_kernel(EpgArgs a, const TrCoef *__restrict__ coef,
                            const float *__restrict__ params, float2 *__restrict__ atoms)
{
    epg_fused_body<M, MULTI, PHASE, HREAL, true>(a, coef, params, atoms);
}

// ------------------------------------------------------------------------
// Reference-order kernel (SR, SE, TSE; SSFP/FISP cross-check)
// ------------------------------------------------------------------------

// fill_transition_matrix + apply_rf_pulse (src/functions.cu:483-528)
template <int M>
__device__ __forceinline__ void g_rf(EpgState<M> &s, const TrCoef &c)
{
    const float2 zz = make_float2(c.zr, c.zi);
    const float2 zi = make_float2(c.zr, -c.zi);
    const float2 z2 = cmul(zz, zz), zi2 = cmul(zi, zi);
    const float2 m01 = make_float2(z2.x * c.sa2, z2.y * c.sa2);
    const float2 m02 = make_float2(zz.y * c.sa, -zz.x * c.sa);          // -i z sa
    const float2 m10 = make_float2(zi2.x * c.sa2, zi2.y * c.sa2);
    const float2 m12 = make_float2(-zi.y * c.sa, zi.x * c.sa);          // i zbar sa
    const float2 m20 = make_float2(0.5f * zi.y * c.sa, -0.5f * zi.x * c.sa);   // -i/2 zbar sa
    const float2 m21 = make_float2(-0.5f * zz.y * c.sa, 0.5f * zz.x * c.sa);   // i/2 z sa
#pragma unroll
    for (int i = 0; i < M; i++) {
        const float2 A = s.fp[i], B = s.fm[i], C = s.z[i];
        float2 p = cmul(m01, B), q = cmul(m02, C);
        s.fp[i] = make_float2(c.ca2 * A.x + p.x + q.x, c.ca2 * A.y + p.y + q.y);
        p = cmul(m10, A);
        q = cmul(m12, C);
        s.fm[i] = make_float2(p.x + c.ca2 * B.x + q.x, p.y + c.ca2 * B.y + q.y);
        p = cmul(m20, A);
        q = cmul(m21, B);
        s.z[i] = make_float2(p.x + q.x + c.ca * C.x, p.y + q.y + c.ca * C.y);
    }
}

// compute_exp + decay_signal (src/functions.cu:530-583)
template <int M>
__device__ __forceinline__ void g_relax(EpgState<M> &s, float t, float T1, float T2, int j)
{
    const float e1 = expf(-t / T1), e2 = expf(-t / T2);
#pragma unroll
    for (int i = 0; i < M; i++) {
        s.fp[i] = make_float2(e2 * s.fp[i].x, e2 * s.fp[i].y);
        s.fm[i] = make_float2(e2 * s.fm[i].x, e2 * s.fm[i].y);
        s.z[i] = make_float2(e1 * s.z[i].x, e1 * s.z[i].y);
    }
    if (j == 0) s.z[0].x += 1.f - e1;
}

// shift_phase (src/functions.cu:585-607)
template <int M>
__device__ __forceinline__ void g_phase(EpgState<M> &s, float t, float B0)
{
    float sn, cs;
    sincosf(phase_angle(B0, t), &sn, &cs);
    const float2 e = make_float2(cs, sn);
#pragma unroll
    for (int i = 0; i < M; i++) {
        s.fp[i] = cmul(e, s.fp[i]);
        s.fm[i] = cmulc(e, s.fm[i]);
    }
}

// dephase_gradients_stage1/2 (src/functions.cu:619-668)
template <int M, bool MULTI>
__device__ __forceinline__ void g_dephase(EpgState<M> &s, const EpgArgs &a, int j)
{
    float2 up = make_float2(0.f, 0.f), dn = up;
    if (MULTI) {
        up = make_float2(shr1(s.fp[M - 1].x), shr1(s.fp[M - 1].y));
        dn = make_float2(shl1(s.fm[0].x), shl1(s.fm[0].y));
    }
#pragma unroll
    for (int i = M - 1; i >= 1; i--) s.fp[i] = s.fp[i - 1];
#pragma unroll
    for (int i = 0; i < M - 1; i++) s.fm[i] = s.fm[i + 1];
    s.fm[M - 1] = (MULTI && j != a.L - 1) ? dn : make_float2(0.f, 0.f);
    s.fp[0] = (j == 0) ? make_float2(s.fm[0].x, -s.fm[0].y) : up;
    epg_truncate(s, a, j);
}

template <int M, bool MULTI, int PROG>
__global__ __launch_bounds__(kBlock) void epg_generic_kernel(EpgArgs a)
{
    const size_t gl = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t atom = MULTI ? (gl >> a.lshift) : gl;
    const int j = MULTI ? (int)(gl & (size_t)(a.L - 1)) : 0;
    const int slice = blockIdx.y;
    const bool valid = atom < a.natoms;
    const size_t pa = valid ? atom : 0;
    const float T1 = a.params[pa * 4 + 0];
    const float T2 = a.params[pa * 4 + 1];
    const float B0 = a.params[pa * 4 + 2];
    const TrCoef *cf = a.coef + (size_t)slice * a.T;

    EpgState<M> s;
    epg_init(s, a, j, T1, T2);
    if (PROG == kEpgSe || PROG == kEpgTse) {
        g_dephase<M, MULTI>(s, a, j);
        g_relax(s, cf[0].te, T1, T2, j);
    }
    float2 *out = a.atoms + ((size_t)slice * a.natoms + pa) * (size_t)a.T;
    const bool store = valid && j == 0;
    double nrm = 0.0;
    for (int n = 0; n < a.T; n++) {
        const TrCoef c = cf[n];
        g_rf(s, c);
        if (PROG == kEpgSe || PROG == kEpgTse) g_dephase<M, MULTI>(s, a, j);
        g_relax(s, c.te, T1, T2, j);
        if (PROG == kEpgSsfp) g_phase(s, c.te, B0);
        const float2 v = s.fp[0];
        if (store) {
            out[n] = v;
            nrm += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
        }
        if (PROG == kEpgSsfp) g_phase(s, c.tmte, B0);
        if (PROG != kEpgSr) g_dephase<M, MULTI>(s, a, j);
        g_relax(s, c.tmte, T1, T2, j);
    }
    if (store && a.norm2) a.norm2[(size_t)slice * a.natoms + atom] = nrm;
}

// ------------------------------------------------------------------------
// ROA (rotation operators, src/functions.cu:1022-1222) with isochromats
// ------------------------------------------------------------------------
struct RoaArgs {
    const TrCoef *coef;
    const float *params;
    float2 *atoms;
    double *norm2;
    size_t natoms;
    int T;
    int n_iso;
    float spread;
};

// One isochromat (the reference's roa_ssfp): one atom per lane.
__global__ __launch_bounds__(kBlock) void roa_kernel(RoaArgs a)
{
    const size_t atom = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (atom >= a.natoms) return;
    const float T1 = a.params[atom * 4 + 0];
    const float T2 = a.params[atom * 4 + 1];
    const float B0 = a.params[atom * 4 + 2];
    const float k1 = decay_k(T1), k2 = decay_k(T2);
    float2 *out = a.atoms + atom * (size_t)a.T;
    double nrm = 0.0;
    // initial magnetisation of init_rf_pulse(180, 0, TI = 0) (:1022-1047)
    float sa0, ca0;
    sincosf((float)(180.0 * M_PI / 180.0), &sa0, &ca0);
    float mx = 0.f, my = sa0, mz = ca0;
    const double b0r = (double)B0 / 2000.0;   // revolutions per ms of TR, half-TR precession
    for (int n = 0; n < a.T; n++) {
        const TrCoef c = a.coef[n];
        const float e1 = decay(c.tr, k1), e2 = decay(c.tr, k2);
        // fill_transition_matrix + Sgemm (:1049-1092)
        const float nx = c.c2p * mx + c.s2p * c.ca * my + c.s2p * c.sa * mz;
        const float ny = -c.s2p * mx + c.c2p * c.ca * my + c.c2p * c.sa * mz;
        const float nz = -c.sa * my + c.ca * mz;
        // decay over TR (:1094-1105)
        const float x = e2 * nx, y = e2 * ny;
        mz = e1 * nz + (1.f - e1);
        // half-TR precession pi*B0*TR/1000 rad = B0*TR/2000 rev (:1107-1143)
        float sb, cb;
        sincos_rev(b0r * (double)c.tr, &sb, &cb);
        const float x2 = cb * x + sb * y, y2 = cb * y - sb * x;
        out[n] = make_float2(x2, y2);
        nrm += (double)x2 * (double)x2 + (double)y2 * (double)y2;
        mx = cb * x2 + sb * y2;
        my = cb * y2 - sb * x2;
    }
    if (a.norm2) a.norm2[atom] = nrm;
}

// Isochromat form (n_iso > 1): G lanes per atom, lane l holding the
// isochromats q = p0 + l + G*u (u < NI) of pass p0.  Isochromat q precesses
// at f0 + q*D Hz, f0 = B0 + spread*(1/(2n) - 1/2), D = spread/n, both in
// double; a lane's first isochromat takes one sincos per TR, the next ones
// follow by rotation steps of 2*pi*G*D*TR/2000.  The G lanes sum their
// isochromats each TR (xor shuffles); lane (n mod G) keeps TR n, so every
// G TRs the group writes 8G contiguous bytes of the atom's row.
constexpr int kIsoLanes = 8;   // 4 and 16 measured 3 % and 12 % slower at 100 isochromats

template <int NI>
__global__ __launch_bounds__(kBlock) void roa_iso_kernel(RoaArgs a)
{
    constexpr int G = kIsoLanes;
    const size_t atom = ((size_t)blockIdx.x * kBlock + threadIdx.x) / G;
    const int l = (int)(threadIdx.x & (G - 1));
    // lanes of a missing atom run on the last atom and store nothing, so
    // that every lane of a group takes part in the shuffles
    const bool valid = atom < a.natoms;
    const size_t at = valid ? atom : a.natoms - 1;
    const float T1 = a.params[at * 4 + 0];
    const float T2 = a.params[at * 4 + 1];
    const float B0 = a.params[at * 4 + 2];
    const float k1 = decay_k(T1), k2 = decay_k(T2);
    float2 *out = a.atoms + at * (size_t)a.T;
    const float inv_n = 1.0f / (float)a.n_iso;
    const double D = (double)a.spread / (double)a.n_iso;
    const double f0 = (double)B0 + (double)a.spread * (0.5 / (double)a.n_iso - 0.5);
    double nrm = 0.0;
    float sa0, ca0;
    sincosf((float)(180.0 * M_PI / 180.0), &sa0, &ca0);
    for (int p0 = 0; p0 < a.n_iso; p0 += G * NI) {
        float mx[NI], my[NI], mz[NI];
#pragma unroll
        for (int u = 0; u < NI; u++) {
            mx[u] = 0.f;
            my[u] = sa0;
            mz[u] = ca0;
        }
        const double rev_first = (f0 + (double)(p0 + l) * D) / 2000.0;   // per ms of TR
        const double rev_step = (double)G * D / 2000.0;
        const bool last = p0 + G * NI >= a.n_iso;
        float2 pend = make_float2(0.f, 0.f);
        for (int n = 0; n < a.T; n++) {
            const TrCoef c = a.coef[n];
            const float e1 = decay(c.tr, k1), e2 = decay(c.tr, k2);
            float sb, cb, sg, cg;
            sincos_rev(rev_first * (double)c.tr, &sb, &cb);
            sincos_rev(rev_step * (double)c.tr, &sg, &cg);
            float sx = 0.f, sy = 0.f;
#pragma unroll
            for (int u = 0; u < NI; u++) {
                if (p0 + l + G * u < a.n_iso) {
                    const float nx = c.c2p * mx[u] + c.s2p * c.ca * my[u] + c.s2p * c.sa * mz[u];
                    const float ny = -c.s2p * mx[u] + c.c2p * c.ca * my[u] + c.c2p * c.sa * mz[u];
                    const float nz = -c.sa * my[u] + c.ca * mz[u];
                    const float x = e2 * nx, y = e2 * ny;
                    mz[u] = e1 * nz + (1.f - e1);
                    const float x2 = cb * x + sb * y, y2 = cb * y - sb * x;
                    sx += x2;
                    sy += y2;
                    mx[u] = cb * x2 + sb * y2;
                    my[u] = cb * y2 - sb * x2;
                }
                if (u + 1 < NI) {
                    const float c2 = cb * cg - sb * sg;
                    sb = sb * cg + cb * sg;
                    cb = c2;
                }
            }
#pragma unroll
            for (int off = G / 2; off >= 1; off >>= 1) {
                sx += __shfl_xor(sx, off, G);
                sy += __shfl_xor(sy, off, G);
            }
            if ((n & (G - 1)) == l) pend = make_float2(sx, sy);
            if ((n & (G - 1)) == G - 1 || n + 1 == a.T) {
                const int t = (n & ~(G - 1)) + l;
                if (valid && t <= n) {
                    float2 v = pend;
                    if (p0 > 0) {
                        const float2 o = out[t];
                        v = make_float2(o.x + v.x, o.y + v.y);
                    }
                    if (last) {
                        v = make_float2(v.x * inv_n, v.y * inv_n);
                        nrm += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
                    }
                    out[t] = v;
                }
            }
        }
    }
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1) nrm += __shfl_xor(nrm, off, G);
    if (a.norm2 && valid && l == 0) a.norm2[atom] = nrm;
}

__global__ __launch_bounds__(kBlock) void normalize_kernel(float2 *atoms, const double *norm2,
                                                            size_t n, int T)
{
    const size_t row = blockIdx.x;
    if (row >= n) return;
    const double s = norm2[row];
    const float f = s > 0.0 ? (float)(1.0 / sqrt(s)) : 1.0f;
    float2 *p = atoms + row * (size_t)T;
    for (int t = threadIdx.x; t < T; t += kBlock) {
        const float2 v = p[t];
        p[t] = make_float2(v.x * f, v.y * f);
    }
}

// ------------------------------------------------------------------------
// dispatch
// ------------------------------------------------------------------------

// Compiled state counts per lane.
constexpr int kMs[] = {4, 8, 10, 12, 15, 16, 20, 24, 28, 30, 32};

template <int M>
int launch_m(const EpgLaunch &l, const EpgArgs &args, dim3 grid, hipStream_t s)
{
    const bool multi = args.L > 1;
    const bool phase = l.echo == kEpgSsfp;
    if (!l.generic && (l.echo == kEpgSsfp || l.echo == kEpgFisp)) {
        const bool trunc = args.istar >= 0;
#define A3 args.coef, args.params, args.atoms
#define MRF_FUSED(PH, HR)                                                                      \
    do {                                                                                       \
        if (multi) {                                                                           \
            if (trunc) epg_fused_trunc_kernel<M, true, PH, HR><<<grid, kBlock, 0, s>>>(args, A3);   \
            else epg_fused_kernel<M, true, PH, HR><<<grid, kBlock, 0, s>>>(args, A3);               \
        } else {                                                                               \
            if (trunc) epg_fused_trunc_kernel<M, false, PH, HR><<<grid, kBlock, 0, s>>>(args, A3);  \
            else epg_fused_kernel<M, false, PH, HR><<<grid, kBlock, 0, s>>>(args, A3);              \
        }                                                                                      \
    } while (0)
        if (phase) MRF_FUSED(true, false);
        else if (l.h_real) MRF_FUSED(false, true);
        else MRF_FUSED(false, false);
#undef MRF_FUSED
#undef A3
        return 0;
    }
    if (M != 8 && M != 32) return -1;   // reference-order path: two lane widths
#define MRF_GENERIC(PROG)                                                              \
    do {                                                                               \
        if (multi) epg_generic_kernel<M, true, PROG><<<grid, kBlock, 0, s>>>(args);    \
        else epg_generic_kernel<M, false, PROG><<<grid, kBlock, 0, s>>>(args);         \
    } while (0)
    switch (l.echo) {
    case kEpgSsfp: MRF_GENERIC(kEpgSsfp); break;
    case kEpgFisp: MRF_GENERIC(kEpgFisp); break;
    case kEpgSr: MRF_GENERIC(kEpgSr); break;
    case kEpgSe: MRF_GENERIC(kEpgSe); break;
    case kEpgTse: MRF_GENERIC(kEpgTse); break;
    default: return -1;
    }
#undef MRF_GENERIC
    return 0;
}

// Gauss-Newton trial cost, one voxel per lane: the fused epg_ssfp recursion
// at the voxel's trial (T1, T2, B0, B1) with its own RF angle alpha_n * B1
// (sincosf per TR, as the tangent passes form it), folded against the
// voxel's signal into s^H s and s^H x (float, summed in blocks of 32 TRs).
// MULTI: each voxel on a lane pair (j = 0, 1) holding M states each (the
// generator's multi-lane form; TRUNC when 2M > w), which doubles the waves of
// a launch that would otherwise leave most SIMDs with one latency-bound wave.
template <int M, bool MULTI, bool TRUNC>
__global__ __launch_bounds__(kBlock) void gn_trial_kernel(EpgArgs a, const float2 *__restrict__ sig,
                                                          size_t V, float *rec, int stride)
{
    const size_t gl = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t v = MULTI ? gl >> 1 : gl;
    const int j = MULTI ? (int)(gl & 1) : 0;
    if (v >= V) return;
    float *r = rec + v * (size_t)stride;
    const float T1 = r[kGnRecTrial], T2 = r[kGnRecTrial + 1], B0 = r[kGnRecTrial + 2],
                B1 = r[kGnRecTrial + 3];
    if (!(T1 > 0.f && T2 > 0.f && B1 > 0.f && isfinite(T1) && isfinite(T2) && isfinite(B0) &&
          isfinite(B1))) {
        if (j == 0) {
            r[kGnRecCt] = INFINITY;
            r[kGnRecHs] = r[kGnRecHsx] = r[kGnRecHsx + 1] = 0.f;
        }
        return;
    }
    FusedCtx x;
    x.j = j;
    x.k1 = decay_k(T1);
    x.k2 = decay_k(T2);
    x.b0r = (double)B0 * 1e-3;
    const TrCoef *__restrict__ cf = a.coef;
    FusedState<M> s;
    {
        EpgState<M> s0;
        epg_init(s0, a, j, T1, T2);
        const float2 z0 = make_float2(cf[0].zr, cf[0].zi);
#pragma unroll
        for (int i = 0; i < M; i++) {
            s.fp[i] = s0.fp[i];
            s.fm[i] = s0.fm[i];
            s.z[i] = s0.z[i];
        }
#pragma unroll
        for (int i = M; i < M + kUnroll; i++) s.fm[i] = make_float2(0.f, 0.f);
        s.fp[0] = cmulc(z0, s.fp[0]);
        s.fm[0] = cmul(z0, s.fm[0]);
    }
    auto coef_at = [&](int n) {
        TrCoef c = cf[n];
        float sa, ca;
        sincosf(c.pad[0] * B1, &sa, &ca);
        c.sa = sa;
        c.ca = ca;
        c.ca2 = 0.5f * (1.f + ca);
        c.sa2 = 0.5f * (1.f - ca);
        return c;
    };
    const float2 *xs = sig + v * (size_t)a.T;
    float S = 0.f, Sr = 0.f, Si = 0.f, bS = 0.f, bR = 0.f, bI = 0.f;
    auto fold = [&](float2 p, float2 q) {
        bS += p.x * p.x + p.y * p.y;
        bR += p.x * q.x + p.y * q.y;     // Re conj(p) q
        bI += p.x * q.y - p.y * q.x;     // Im conj(p) q
    };
    int n = 0;
    for (; n + kUnroll <= a.T; n += kUnroll) {
        const float2 v0 = fused_step<M, 0, MULTI, true, false, TRUNC>(s, coef_at(n), x, a);
        const float2 v1 = fused_step<M, 1, MULTI, true, false, TRUNC>(s, coef_at(n + 1), x, a);
        fm_copy_down<M, 2>(s);
        fold(v0, xs[n]);
        fold(v1, xs[n + 1]);
        if (((n + 2) & 31) == 0) {
            S += bS;
            Sr += bR;
            Si += bI;
            bS = bR = bI = 0.f;
        }
    }
    for (; n < a.T; n++) {
        const float2 v0 = fused_step<M, 0, MULTI, true, false, TRUNC>(s, coef_at(n), x, a);
        fm_copy_down<M, 1>(s);
        fold(v0, xs[n]);
    }
    if (j != 0) return;   // the sample is state 0, on lane 0 of the pair
    S += bS;
    Sr += bR;
    Si += bI;
    r[kGnRecCt] = r[kGnRecXx] - (Sr * Sr + Si * Si) / S;
    r[kGnRecHs] = S;
    r[kGnRecHsx] = Sr;
    r[kGnRecHsx + 1] = Si;
}

}  // namespace

int epg_form(const EpgLaunch &l, int *L_out, int *M_out)
{
    if (l.w < 2) return mrf_fail(MRF_E_ARG, "nstates must be >= 2 (got %d)", l.w);
    const bool fused = !l.generic && (l.echo == kEpgSsfp || l.echo == kEpgFisp);
    int L = 1, M = 0;
    if (l.lanes) {
        // requested lanes per atom: the smallest compiled M that covers w
        if (l.lanes != 1 && l.lanes != 2 && l.lanes != 4 && l.lanes != 8 && l.lanes != 16)
            return mrf_fail(MRF_E_ARG, "lanes per atom must be 1, 2, 4, 8 or 16 (got %d)", l.lanes);
        L = l.lanes;
        if (fused) {
            for (int m : kMs)
                if (m * L >= l.w) { M = m; break; }
        } else {
            M = (8 * L >= l.w) ? 8 : ((32 * L >= l.w) ? 32 : 0);
        }
        if (!M)
            return mrf_fail(MRF_E_UNSUPPORTED, "%d lanes per atom cannot hold %d states", L, l.w);
    } else if (fused) {
        // smallest L, then smallest compiled M with L*M >= w
        for (L = 1; L <= 16 && !M; L *= 2)
            for (int m : kMs)
                if (m * L >= l.w) { M = m; break; }
        L /= 2;
        // small dictionaries: one atom per lane leaves most SIMDs idle and
        // each lane a long per-TR chain; spread the atoms over more lanes
        // (fewer states each, at least 8) until the launch has 2^17 lanes
        while (M && L < 16 && l.natoms * (size_t)l.n_slices * (size_t)L < ((size_t)1 << 17)) {
            int m2 = 0;
            for (int m : kMs)
                if (m * 2 * L >= l.w) { m2 = m; break; }
            if (m2 < 8) break;
            L *= 2;
            M = m2;
        }
    } else {
        // reference-order path: M = 8 up to w = 128, else M = 32
        const int mw = (l.w <= 128) ? 8 : 32;
        for (L = 1; L <= 16 && mw * L < l.w; L *= 2) {}
        if (L <= 16) M = mw;
    }
    if (!M || L > 16) return mrf_fail(MRF_E_UNSUPPORTED, "nstates %d exceeds 512", l.w);
    *L_out = L;
    *M_out = M;
    return 0;
}

int launch_epg(const EpgLaunch &l, hipStream_t s)
{
    const bool fused = !l.generic && (l.echo == kEpgSsfp || l.echo == kEpgFisp);
    int L = 1, M = 0;
    if (int rc = epg_form(l, &L, &M)) return rc;

    EpgArgs args{};
    args.coef = l.coef;
    args.params = l.params;
    args.atoms = l.atoms;
    args.norm2 = l.norm2;
    args.natoms = l.natoms;
    args.T = l.T;
    args.L = L;
    args.lshift = 0;
    while ((1 << args.lshift) < L) args.lshift++;
    const int cap = L * M;
    args.istar = (cap > l.w) ? (l.w - 1) % M : -1;
    args.jstar = (l.w - 1) / M;
    // initial pulse angles as init_rf_pulse casts them (src/functions.cu:465-466)
    float a_deg = 0.f, p_deg = 0.f, TI = 0.f;
    switch (l.echo) {
    case kEpgSsfp: a_deg = 180.f; p_deg = 0.f; TI = 40.f; break;   // :965-966, :1242
    case kEpgSe:
    case kEpgTse: a_deg = 90.f; p_deg = 90.f; break;                // :764-765, :833-834
    default: break;                                                 // sr/fisp: (0, 0)
    }
    args.init_alpha = (float)((double)a_deg * M_PI / 180.0);
    args.init_phi = (float)((double)p_deg * M_PI / 180.0);
    args.TI = TI;
    args.normalize = (l.normalize && fused) ? 1 : 0;

    const size_t lanes = l.natoms * (size_t)L;
    const size_t blocks = (lanes + kBlock - 1) / kBlock;
    if (blocks > 0x7fffffff) return mrf_fail(MRF_E_ARG, "too many atoms");
    dim3 grid((unsigned)blocks, (unsigned)l.n_slices);
    int rc = -1;
    switch (M) {
    case 4: rc = launch_m<4>(l, args, grid, s); break;
    case 8: rc = launch_m<8>(l, args, grid, s); break;
    case 10: rc = launch_m<10>(l, args, grid, s); break;
    case 12: rc = launch_m<12>(l, args, grid, s); break;
    case 15: rc = launch_m<15>(l, args, grid, s); break;
    case 16: rc = launch_m<16>(l, args, grid, s); break;
    case 20: rc = launch_m<20>(l, args, grid, s); break;
    case 24: rc = launch_m<24>(l, args, grid, s); break;
    case 28: rc = launch_m<28>(l, args, grid, s); break;
    case 30: rc = launch_m<30>(l, args, grid, s); break;
    case 32: rc = launch_m<32>(l, args, grid, s); break;
    default: break;
    }
    if (rc) return mrf_fail(MRF_E_UNSUPPORTED, "no EPG kernel for w=%d", l.w);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mrf_fail(MRF_E_GPU, "EPG launch: %s", hipGetErrorString(e));
    if (l.normalize && !args.normalize)
        return launch_normalize(l.atoms, l.norm2, l.natoms * (size_t)l.n_slices, l.T, s);
    return 0;
}

int launch_roa(const RoaLaunch &l, hipStream_t s)
{
    RoaArgs a{l.coef, l.params, l.atoms, l.norm2, l.natoms, l.T, l.n_iso < 1 ? 1 : l.n_iso,
              l.iso_spread_hz};
    if (l.natoms == 0 || l.T <= 0) return 0;
    if (a.n_iso == 1) {
        const size_t blocks = (l.natoms + kBlock - 1) / kBlock;
        roa_kernel<<<dim3((unsigned)blocks), kBlock, 0, s>>>(a);
    } else {
        if (l.natoms > ((size_t)1 << 31) / kIsoLanes)
            return mrf_fail(MRF_E_ARG, "too many ROA atoms for the isochromat kernel");
        const size_t blocks = (l.natoms * kIsoLanes + kBlock - 1) / kBlock;
        const dim3 grid((unsigned)blocks);
        switch ((a.n_iso + kIsoLanes - 1) / kIsoLanes) {
        case 1: roa_iso_kernel<1><<<grid, kBlock, 0, s>>>(a); break;
        case 2: roa_iso_kernel<2><<<grid, kBlock, 0, s>>>(a); break;
        case 3: roa_iso_kernel<3><<<grid, kBlock, 0, s>>>(a); break;
        case 4: roa_iso_kernel<4><<<grid, kBlock, 0, s>>>(a); break;
        case 5: roa_iso_kernel<5><<<grid, kBlock, 0, s>>>(a); break;
        case 6: roa_iso_kernel<6><<<grid, kBlock, 0, s>>>(a); break;
        case 7: roa_iso_kernel<7><<<grid, kBlock, 0, s>>>(a); break;
        case 8: roa_iso_kernel<8><<<grid, kBlock, 0, s>>>(a); break;
        case 9: case 10: roa_iso_kernel<10><<<grid, kBlock, 0, s>>>(a); break;
        case 11: case 12: case 13: roa_iso_kernel<13><<<grid, kBlock, 0, s>>>(a); break;
        case 14: case 15: case 16: roa_iso_kernel<16><<<grid, kBlock, 0, s>>>(a); break;
        case 17: case 18: case 19: case 20: roa_iso_kernel<20><<<grid, kBlock, 0, s>>>(a); break;
        default: roa_iso_kernel<25><<<grid, kBlock, 0, s>>>(a); break;
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mrf_fail(MRF_E_GPU, "ROA launch: %s", hipGetErrorString(e));
    return 0;
}

int launch_normalize(float2 *atoms, const double *norm2, size_t n, int T, hipStream_t s)
{
    if (n == 0) return 0;
    if (n > 0x7fffffffULL) return mrf_fail(MRF_E_ARG, "too many atoms to normalise");
    normalize_kernel<<<dim3((unsigned)n), kBlock, 0, s>>>(atoms, norm2, n, T);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mrf_fail(MRF_E_GPU, "normalize launch: %s", hipGetErrorString(e));
    return 0;
}

int launch_gn_trial(const TrCoef *coef, int T, int w, const float2 *sig, size_t V, float *rec,
                    int stride, hipStream_t s)
{
    if (V == 0) return 0;
    EpgArgs a{};
    a.coef = coef;
    a.T = T;
    a.L = 1;
    a.lshift = 0;
    a.istar = -1;
    a.jstar = 0;
    a.init_alpha = (float)M_PI;   // init_rf_pulse(180 deg, 0 deg, TI = 40) (src/functions.cu:965-966)
    a.init_phi = 0.f;
    a.TI = 40.f;
    if (w >= 9 && w <= 32) {
        // lane pairs of M = 8, 12 or 16 states
        const int M = w <= 16 ? 8 : (w <= 24 ? 12 : 16);
        a.L = 2;
        a.lshift = 1;
        const bool trunc = 2 * M > w;
        a.istar = trunc ? (w - 1) % M : -1;
        a.jstar = (w - 1) / M;
        const dim3 grid((unsigned)((2 * V + kBlock - 1) / kBlock));
#define MRF_GN_TRIAL2(MM)                                                                    \
    if (M == MM) {                                                                           \
        if (trunc) gn_trial_kernel<MM, true, true><<<grid, kBlock, 0, s>>>(a, sig, V, rec, stride); \
        else gn_trial_kernel<MM, true, false><<<grid, kBlock, 0, s>>>(a, sig, V, rec, stride);     \
    }
        MRF_GN_TRIAL2(8)
        MRF_GN_TRIAL2(12)
        MRF_GN_TRIAL2(16)
#undef MRF_GN_TRIAL2
        return 0;
    }
    const dim3 grid((unsigned)((V + kBlock - 1) / kBlock));
    switch (w) {
    case 4: gn_trial_kernel<4, false, false><<<grid, kBlock, 0, s>>>(a, sig, V, rec, stride); break;
    case 8: gn_trial_kernel<8, false, false><<<grid, kBlock, 0, s>>>(a, sig, V, rec, stride); break;
    default: return 1;
    }
    return 0;
}

}  // namespace mrf
