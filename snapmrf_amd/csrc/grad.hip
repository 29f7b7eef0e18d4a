// grad.hip -- differentiation form of the EPG-SSFP program: Jacobians with
// respect to (T1, T2, B0, B1), reverse-mode adjoints with respect to those
// and to the per-timepoint flip angles and repetition times (checkpointed
// recomputation), per-voxel Gauss-Newton refinement, and the
// gradient of a Cramer-Rao-bound sequence-design loss.
//
// Model: the reference's epg_ssfp program (src/functions.cu:936-1019 with
// init_rf_pulse :456-481, apply_rf_pulse :483-528, decay_signal :530-583,
// shift_phase :585-607, dephase_gradients :619-668) as a continuous function
// of its inputs; the flip angles of an atom are alpha_n * B1, the CSV unit
// transform (functions.cu:1242-1261) is folded into the per-TR table.
//
// Layout: one group of SUB lanes per atom (SUB = 32 for w <= 32, two atoms
// per wave; else 64), lane k holds EPG state k (F+_k, F-_k, Z_k) and its ND
// tangents, so an atom's whole state sits in 6 (1 + ND) registers per lane;
// the dephasing shift is a cross-lane move inside the group (Jacobian,
// Gauss-Newton, CRB).  The adjoint puts an atom on 8 or 16 lanes with four
// states per lane (adj_kernel).  Per TR n:
//   RF(alpha_n B1, phi_n) -> relax(TE_n) -> phase(TE_n) -> s_n = F+_0 (lane 0)
//   -> phase(TR_n - TE_n) -> dephase -> relax(TR_n - TE_n)
// TR n runs in the frame of its RF phase, x~ = P_n^-1 x with
// P_n = diag(e^{i phi_n}, e^{-i phi_n}, 1): relaxation, precession and the
// gradient shift commute with P_n, so the RF there is the zero-phase matrix
// and the change to frame n+1, F+ e^{i(phi_n - phi_{n+1})} and F- times its
// conjugate, has the form of the TR - TE precession and is folded into its
// factor.  Readouts go back to true coordinates by e^{i phi_n}, the
// adjoint's cotangents by e^{-i phi_n}, and the initial state (the inversion
// at phase 0) enters frame 0 by e^{-/+ i phi_0}.
// Arithmetic is FP32; FP64 is used only for the 4x4 Gauss-Newton solve and
// the Fisher-matrix inverse.
#include "kernels.h"
#include "snapmrf.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>

namespace mrf {

namespace {

constexpr int kCkpt = 50;                 // TRs between checkpoints (CRB)
constexpr int kWaves = 4;                 // atoms per work-group
constexpr float kCB0 = (float)(2.0 * M_PI / 1000.0);
constexpr float kTI = 40.0f;              // inversion delay, functions.cu:1242

// complex float helpers
__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ float2 c_mul(float2 a, float2 b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ float2 c_conj(float2 a) { return {a.x, -a.y}; }
__device__ __forceinline__ float2 c_scale(float2 a, float s) { return {a.x * s, a.y * s}; }
__device__ __forceinline__ float2 c_i(float2 a) { return {-a.y, a.x}; }      // i * a
__device__ __forceinline__ float re_dot(float2 l, float2 v) { return l.x * v.x + l.y * v.y; }  // Re(conj(l) v)

// Lane groups of SUB lanes (32 or 64) each hold one atom; shifts,
// broadcasts and sums stay inside the group.  The shifts are whole-wave DPP
// moves (wave_shr:1, wave_shl:1) on the VALU: the lanes that receive a value
// from across a group boundary (k = 0 for the up shift, k = SUB - 1 for the
// down shift) are the ones every caller overwrites (F+_0 <- conj(F-_0),
// states k >= w - 1 zeroed, w <= SUB).
__device__ __forceinline__ float wave_shr1(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float wave_shl1(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                 0x130, 0xF, 0xF, true));
}
template <int SUB>
__device__ __forceinline__ float2 shfl_up1(float2 v)
{
    return {wave_shr1(v.x), wave_shr1(v.y)};
}
template <int SUB>
__device__ __forceinline__ float2 shfl_down1(float2 v)
{
    return {wave_shl1(v.x), wave_shl1(v.y)};
}
// lane 0 of the group (lanes 0 and 32 of the wave for SUB = 32)
template <int SUB>
__device__ __forceinline__ float bcast0(float v)
{
    const float a = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 0));
    if (SUB == 64) return a;
    const float b = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 32));
    return (threadIdx.x & 32) ? b : a;
}
template <int SUB>
__device__ __forceinline__ float2 bcast0(float2 v)
{
    return {bcast0<SUB>(v.x), bcast0<SUB>(v.y)};
}

// v + the value of the lane given by a DPP pattern
template <int CTRL>
__device__ __forceinline__ float dpp_add(float v)
{
    return v + __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
                                             0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}

// Sum over the SUB lanes of the group, on every lane: pairs within quads
// (quad_perm [1,0,3,2] and [2,3,0,1]), then 8-lane half rows (row_half_mirror)
// and 16-lane rows (row_mirror) by DPP on the VALU, then across rows.
template <int SUB>
__device__ __forceinline__ float group_sum(float v)
{
    v = dpp_add<0xB1>(v);
    v = dpp_add<0x4E>(v);
    v = dpp_add<0x141>(v);
    v = dpp_add<0x140>(v);
    if (SUB >= 32) v += __shfl_xor(v, 16, SUB);
    if (SUB >= 64) v += __shfl_xor(v, 32, SUB);
    return v;
}

// The zero-phase RF matrix M0(a) = [[ca2, sa2, -i sa], [sa2, ca2, i sa],
// [-i sa/2, i sa/2, ca]] (apply_rf_pulse, functions.cu:483-528, at phi = 0;
// ca2 = (1 + cos a)/2, sa2 = (1 - cos a)/2) or one of its a-derivatives,
// stored as hs = (ca2 + sa2)/2, hd = (ca2 - sa2)/2, s = sa, c = ca: with
// S = F+ + F-, D = F+ - F-,
//     F+' = hs S + W,  F-' = hs S - W,  W = hd D + s (Z.y, -Z.x),
//     Z'  = c Z + (s/2) (D.y, -D.x).
// hs is 1/2 for M0 and 0 for its derivatives (HS = false drops S).
struct C4 {
    float hs, hd, s, c;
};

template <bool HS>
__device__ __forceinline__ void rf3(const C4 &m, float2 fp, float2 fm, float2 z, float2 &op,
                                    float2 &om, float2 &oz)
{
    const float dx = fp.x - fm.x, dy = fp.y - fm.y;
    const float wx = m.hd * dx + m.s * z.y, wy = m.hd * dy - m.s * z.x;
    if (HS) {
        const float sx = fp.x + fm.x, sy = fp.y + fm.y;
        op = {m.hs * sx + wx, m.hs * sy + wy};
        om = {m.hs * sx - wx, m.hs * sy - wy};
    } else {
        op = {wx, wy};
        om = {-wx, -wy};
    }
    const float h = 0.5f * m.s;
    oz = {m.c * z.x + h * dy, m.c * z.y - h * dx};
}

// l = M^H l': F+ = hs S + W, F- = hs S - W with S = l+ + l-, D = l+ - l-,
// W = hd D + (s/2) (-lz.y, lz.x); Z = c lz + s (-D.y, D.x)
template <bool HS>
__device__ __forceinline__ void rf3_adj(const C4 &m, float2 lp, float2 lm, float2 lz, float2 &op,
                                        float2 &om, float2 &oz)
{
    const float dx = lp.x - lm.x, dy = lp.y - lm.y;
    const float h = 0.5f * m.s;
    const float wx = m.hd * dx - h * lz.y, wy = m.hd * dy + h * lz.x;
    if (HS) {
        const float sx = lp.x + lm.x, sy = lp.y + lm.y;
        op = {m.hs * sx + wx, m.hs * sy + wy};
        om = {m.hs * sx - wx, m.hs * sy - wy};
    } else {
        op = {wx, wy};
        om = {-wx, -wy};
    }
    oz = {m.c * lz.x - m.s * dy, m.c * lz.y + m.s * dx};
}

// M0(a), dM0/da, d2M0/da2 (the entries' derivatives: ca2' = -sa/2, sa2' = sa/2,
// sa' = ca, ca' = -sa)
__device__ __forceinline__ C4 rf_c(float sa, float ca) { return {0.5f, 0.5f * ca, sa, ca}; }
__device__ __forceinline__ C4 rf_d1(float sa, float ca) { return {0.f, -0.5f * sa, ca, -sa}; }
__device__ __forceinline__ C4 rf_d2(float sa, float ca) { return {0.f, -0.5f * ca, -sa, -ca}; }

struct Atom {
    float T1, T2, B0, B1;
    float iT1, iT2;
};

// Augmented state: index 0 is the signal state, 1 + d the tangent along
// theta_d, d = 0 (T1), 1 (T2), 2 (B0), 3 (B1).
template <int ND>
struct Aug {
    float2 fp[1 + ND], fm[1 + ND], z[1 + ND];
};

// Per-TR coefficients, identical on all lanes of the wave.
struct TrK {
    float sa, ca;            // sin/cos of alpha_n B1
    float2 ro;               // e^{i phi_n}: frame n -> true coordinates
    float al;
    float e1a, e2a, e1b, e2b;     // relaxation over TE and TR - TE
    float2 pa, pb;                // e^{i beta TE}, e^{i beta (TR - TE)} e^{i(phi_n - phi_{n+1})}
    float te, t2;
};

__device__ __forceinline__ TrK tr_coef(const DTr &q, const Atom &p)
{
    TrK c;
    c.al = q.al;
    sincosf(q.al * p.B1, &c.sa, &c.ca);
    c.ro = {q.cph, q.sph};
    c.te = q.te;
    c.t2 = q.t2;
    c.e1a = expf(-q.te * p.iT1);
    c.e2a = expf(-q.te * p.iT2);
    c.e1b = expf(-q.t2 * p.iT1);
    c.e2b = expf(-q.t2 * p.iT2);
    const float beta = kCB0 * p.B0;
    float s, co;
    sincosf(beta * q.te, &s, &co);
    c.pa = {co, s};
    sincosf(beta * q.t2, &s, &co);
    c.pb = c_mul({co, s}, {q.rr, q.ri});   // with the change to frame n+1
    return c;
}

// orders LDS writes and reads of one wave's lanes
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Atom-dependent per-TR coefficients, cached SUB TRs at a time: lane k of a
// group computes TR b * SUB + k (three sincosf, four expf) into LDS and the
// group reads each TR's entry back, so a lane does 1 / SUB of that work.  The
// two groups of a wave walk the same TRs, so a block changes for the whole
// wave at once; a group's entries are written and read by that wave only.
struct alignas(8) TrA {
    float sa, ca, e1a, e2a, e1b, e2b;
    float2 pa, pb;
};
static_assert(sizeof(TrA) == 40, "TrA layout");

constexpr int kCoefLds = kWaves * 64;   // TrA entries per work-group

template <int SUB>
struct CoefCache {
    TrA *blk;   // the group's SUB entries
    int cur;    // block held, -1: none
};

template <int SUB>
__device__ __forceinline__ CoefCache<SUB> coef_cache(TrA *lds)
{
    return {lds + (threadIdx.x >> 6) * 64 + ((threadIdx.x & 63) / SUB) * SUB, -1};
}

template <int SUB>
__device__ __forceinline__ TrK coef(CoefCache<SUB> &cc, const DTr *seq, int T, const Atom &p,
                                    int t, int k)
{
    const int b = t / SUB;
    if (b != cc.cur) {
        const int tt = b * SUB + k;
        if (tt < T) {
            const TrK c = tr_coef(seq[tt], p);
            TrA e;
            e.sa = c.sa;
            e.ca = c.ca;
            e.e1a = c.e1a;
            e.e2a = c.e2a;
            e.e1b = c.e1b;
            e.e2b = c.e2b;
            e.pa = c.pa;
            e.pb = c.pb;
            cc.blk[k] = e;
        }
        cc.cur = b;
        wave_sync();
    }
    const TrA e = cc.blk[t - b * SUB];
    const DTr q = seq[t];
    TrK c;
    c.al = q.al;
    c.sa = e.sa;
    c.ca = e.ca;
    c.ro = {q.cph, q.sph};
    c.te = q.te;
    c.t2 = q.t2;
    c.e1a = e.e1a;
    c.e2a = e.e2a;
    c.e1b = e.e1b;
    c.e2b = e.e2b;
    c.pa = e.pa;
    c.pb = e.pb;
    return c;
}

// Readouts (state 0 and its ND tangents, on lane 0) staged for SUB TRs in
// LDS, so that the group's lanes consume them one TR per lane once per block
// (coalesced stores, sums spread over the lanes) instead of lane 0 alone
// every TR.  kReadLds float2 per work-group (ND <= 4).
constexpr int kReadLds = kWaves * 64 * 5;

// the group's region holds SUB x 5 float2 whatever ND is, so passes of
// different ND in different waves never overlap
template <int SUB>
__device__ __forceinline__ float2 *read_stage(float2 *lds)
{
    return lds + ((threadIdx.x >> 6) * 64 + ((threadIdx.x & 63) / SUB) * SUB) * 5;
}

// --- forward operators on the augmented state (lane k) -------------------

template <int ND>
__device__ __forceinline__ void f_rf(Aug<ND> &y, const TrK &c)
{
    const C4 m = rf_c(c.sa, c.ca);
    float2 sp, sm, sz;
    if constexpr (ND > 0) {
        rf3<false>(rf_d1(c.sa, c.ca), y.fp[0], y.fm[0], y.z[0], sp, sm, sz);
    }
#pragma unroll
    for (int j = 0; j <= ND; j++) {
        float2 a, b, d;
        rf3<true>(m, y.fp[j], y.fm[j], y.z[j], a, b, d);
        y.fp[j] = a;
        y.fm[j] = b;
        y.z[j] = d;
    }
    if constexpr (ND > 0) {   // d/dB1: alpha M_a y
        y.fp[4] = c_add(y.fp[4], c_scale(sp, c.al));
        y.fm[4] = c_add(y.fm[4], c_scale(sm, c.al));
        y.z[4] = c_add(y.z[4], c_scale(sz, c.al));
    }
}

template <int ND>
__device__ __forceinline__ void f_relax(Aug<ND> &y, float e1, float e2, float t, const Atom &p,
                                        bool k0)
{
    if constexpr (ND > 0) {
        const float e1d = e1 * t * p.iT1 * p.iT1, e2d = e2 * t * p.iT2 * p.iT2;
#pragma unroll
        for (int j = 1; j <= ND; j++) {
            y.fp[j] = c_scale(y.fp[j], e2);
            y.fm[j] = c_scale(y.fm[j], e2);
            y.z[j] = c_scale(y.z[j], e1);
        }
        y.fp[2] = c_add(y.fp[2], c_scale(y.fp[0], e2d));
        y.fm[2] = c_add(y.fm[2], c_scale(y.fm[0], e2d));
        float2 zr = y.z[0];
        if (k0) zr.x -= 1.f;
        y.z[1] = c_add(y.z[1], c_scale(zr, e1d));
    }
    y.fp[0] = c_scale(y.fp[0], e2);
    y.fm[0] = c_scale(y.fm[0], e2);
    y.z[0] = c_scale(y.z[0], e1);
    if (k0) y.z[0].x += 1.f - e1;
}

template <int ND>
__device__ __forceinline__ void f_phase(Aug<ND> &y, float2 P, float t)
{
    const float2 Q = c_conj(P);
    if constexpr (ND > 0) {
#pragma unroll
        for (int j = 1; j <= ND; j++) {
            y.fp[j] = c_mul(P, y.fp[j]);
            y.fm[j] = c_mul(Q, y.fm[j]);
        }
        const float ct = kCB0 * t;
        y.fp[3] = c_add(y.fp[3], c_scale(c_i(c_mul(P, y.fp[0])), ct));
        y.fm[3] = c_add(y.fm[3], c_scale(c_i(c_mul(Q, y.fm[0])), -ct));
    }
    y.fp[0] = c_mul(P, y.fp[0]);
    y.fm[0] = c_mul(Q, y.fm[0]);
}

// dephase_gradients (functions.cu:619-668): F+_k <- F+_{k-1}, F-_k <- F-_{k+1},
// F+_0 <- conj(F-_0), F-_{w-1} <- 0; lanes >= w stay zero.
template <int ND, int SUB>
__device__ __forceinline__ void f_dephase(Aug<ND> &y, int k, int w)
{
#pragma unroll
    for (int j = 0; j <= ND; j++) {
        float2 m = shfl_down1<SUB>(y.fm[j]);
        float2 pl = shfl_up1<SUB>(y.fp[j]);
        if (k >= w - 1) m = {0.f, 0.f};
        if (k == 0) pl = c_conj(m);
        if (k >= w) pl = {0.f, 0.f};
        y.fp[j] = pl;
        y.fm[j] = m;
    }
}

template <int ND, int SUB>
__device__ __forceinline__ void f_tr(Aug<ND> &y, const TrK &c, const Atom &p, int k, int w,
                                     Aug<ND> *at_readout)
{
    f_rf(y, c);
    f_relax(y, c.e1a, c.e2a, c.te, p, k == 0);
    f_phase(y, c.pa, c.te);
    if (at_readout) *at_readout = y;
    f_phase(y, c.pb, c.t2);
    f_dephase<ND, SUB>(y, k, w);
    f_relax(y, c.e1b, c.e2b, c.t2, p, k == 0);
}

// init_rf_pulse(180 deg, 0 deg, TI) (functions.cu:456-481) on lane 0, with
// the tangents of its T1/T2 dependence, in the frame of TR 0 (ro0 =
// e^{i phi_0}: F+ times conj(ro0), F- times ro0).
template <int ND>
__device__ __forceinline__ void f_init(Aug<ND> &y, const Atom &p, int k, float2 ro0)
{
#pragma unroll
    for (int j = 0; j <= ND; j++) y.fp[j] = y.fm[j] = y.z[j] = {0.f, 0.f};
    if (k != 0) return;
    const float a0 = (float)(180.0 * M_PI / 180.0f);
    const float sa0 = sinf(a0), ca0 = cosf(a0);
    const float e1 = expf(-kTI * p.iT1), e2 = expf(-kTI * p.iT2);
    const float2 tmp = {0.f, e2 * sa0};
    y.fp[0] = c_mul(c_conj(ro0), c_conj(tmp));
    y.fm[0] = c_mul(ro0, tmp);
    y.z[0] = {e1 * ca0 + 1.f - e1, 0.f};
    if constexpr (ND > 0) {
        const float e1d = e1 * kTI * p.iT1 * p.iT1, e2d = e2 * kTI * p.iT2 * p.iT2;
        y.z[1] = {e1d * (ca0 - 1.f), 0.f};
        const float2 t2 = {0.f, e2d * sa0};
        y.fp[2] = c_mul(c_conj(ro0), c_conj(t2));
        y.fm[2] = c_mul(ro0, t2);
    }
}

__device__ __forceinline__ float2 frame0(const DTr *seq) { return {seq[0].cph, seq[0].sph}; }

// --- reverse operators over the tangent-augmented state (CRB) -------------
// l holds the adjoint of the operator's output on entry and of its input on
// exit; x is the operator's input.  ga / gt collect the lane's share of
// dL/d(alpha_n) and dL/dt (t the interval of the operator).

template <int ND>
__device__ __forceinline__ void r_relax(Aug<ND> &l, const Aug<ND> &x, float e1, float e2, float t,
                                        const Atom &p, bool k0, float *gt)
{
    const float i1 = p.iT1, i2 = p.iT2;
    const float e1d = e1 * t * i1 * i1, e2d = e2 * t * i2 * i2;
    float2 zr = x.z[0];
    if (k0) zr.x -= 1.f;
    if (gt) {
        const float e1t = -e1 * i1, e2t = -e2 * i2;
        float g = re_dot(l.fp[0], x.fp[0]) * e2t + re_dot(l.fm[0], x.fm[0]) * e2t +
                  re_dot(l.z[0], zr) * e1t;
        if constexpr (ND > 0) {
#pragma unroll
            for (int j = 1; j <= ND; j++)
                g += (re_dot(l.fp[j], x.fp[j]) + re_dot(l.fm[j], x.fm[j])) * e2t +
                     re_dot(l.z[j], x.z[j]) * e1t;
            const float e1dt = e1 * i1 * i1 * (1.f - t * i1), e2dt = e2 * i2 * i2 * (1.f - t * i2);
            g += (re_dot(l.fp[2], x.fp[0]) + re_dot(l.fm[2], x.fm[0])) * e2dt +
                 re_dot(l.z[1], zr) * e1dt;
        }
        *gt += g;
    }
    float2 lp0 = c_scale(l.fp[0], e2), lm0 = c_scale(l.fm[0], e2), lz0 = c_scale(l.z[0], e1);
    if constexpr (ND > 0) {
        lp0 = c_add(lp0, c_scale(l.fp[2], e2d));
        lm0 = c_add(lm0, c_scale(l.fm[2], e2d));
        lz0 = c_add(lz0, c_scale(l.z[1], e1d));
#pragma unroll
        for (int j = 1; j <= ND; j++) {
            l.fp[j] = c_scale(l.fp[j], e2);
            l.fm[j] = c_scale(l.fm[j], e2);
            l.z[j] = c_scale(l.z[j], e1);
        }
    }
    l.fp[0] = lp0;
    l.fm[0] = lm0;
    l.z[0] = lz0;
}

template <int ND, int SUB>
__device__ __forceinline__ void r_dephase(Aug<ND> &l, int k, int w)
{
#pragma unroll
    for (int j = 0; j <= ND; j++) {
        const float2 lp0 = bcast0<SUB>(l.fp[j]);
        float2 p = shfl_down1<SUB>(l.fp[j]);
        float2 m = shfl_up1<SUB>(l.fm[j]);
        if (k >= w - 1) p = {0.f, 0.f};
        if (k == 0 || k >= w) m = {0.f, 0.f};
        if (k == 1) m = c_add(m, c_conj(lp0));
        l.fp[j] = p;
        l.fm[j] = m;
    }
}

template <int ND>
__device__ __forceinline__ void r_phase(Aug<ND> &l, const Aug<ND> &x, float2 P, float t,
                                        float beta, float *gt)
{
    const float2 Q = c_conj(P);
    const float ct = kCB0 * t;
    const float2 Px0 = c_mul(P, x.fp[0]), Qx0 = c_mul(Q, x.fm[0]);
    if (gt) {
        // d/dt of P = i beta P, of Q = -i beta Q
        float g = re_dot(l.fp[0], c_scale(c_i(Px0), beta)) + re_dot(l.fm[0], c_scale(c_i(Qx0), -beta));
        if constexpr (ND > 0) {
#pragma unroll
            for (int j = 1; j <= ND; j++)
                g += re_dot(l.fp[j], c_scale(c_i(c_mul(P, x.fp[j])), beta)) +
                     re_dot(l.fm[j], c_scale(c_i(c_mul(Q, x.fm[j])), -beta));
            // d/dt of the B0 sources i c t P and -i c t Q
            const float2 a = c_scale(c_i(Px0), kCB0);          // i c P x
            const float2 bp = c_sub(a, c_scale(Px0, kCB0 * beta * t));   // i c P x (1 + i beta t)
            const float2 am = c_scale(c_i(Qx0), -kCB0);         // -i c Q x
            const float2 bm = c_sub(am, c_scale(Qx0, kCB0 * beta * t));  // -i c Q x (1 - i beta t)
            g += re_dot(l.fp[3], bp) + re_dot(l.fm[3], bm);
        }
        *gt += g;
    }
    float2 lp0 = c_mul(Q, l.fp[0]), lm0 = c_mul(P, l.fm[0]);
    if constexpr (ND > 0) {
        // conj(i c t P) = -i c t Q ; conj(-i c t Q) = i c t P
        lp0 = c_add(lp0, c_scale(c_i(c_mul(Q, l.fp[3])), -ct));
        lm0 = c_add(lm0, c_scale(c_i(c_mul(P, l.fm[3])), ct));
#pragma unroll
        for (int j = 1; j <= ND; j++) {
            l.fp[j] = c_mul(Q, l.fp[j]);
            l.fm[j] = c_mul(P, l.fm[j]);
        }
    }
    l.fp[0] = lp0;
    l.fm[0] = lm0;
}

template <int ND>
__device__ __forceinline__ void r_rf(Aug<ND> &l, const Aug<ND> &x, const TrK &c, float B1,
                                     float *ga)
{
    const C4 m = rf_c(c.sa, c.ca), m1 = rf_d1(c.sa, c.ca);
    float g_a = 0.f, g_al = 0.f;
#pragma unroll
    for (int j = 0; j <= ND; j++) {
        float2 a, b, d;
        rf3<false>(m1, x.fp[j], x.fm[j], x.z[j], a, b, d);
        g_a += re_dot(l.fp[j], a) + re_dot(l.fm[j], b) + re_dot(l.z[j], d);
        if (ND > 0 && j == 0) {
            // explicit alpha of the B1 source alpha M_a x0
            g_al += re_dot(l.fp[4], a) + re_dot(l.fm[4], b) + re_dot(l.z[4], d);
        }
    }
    float2 sp, sm, sz;
    if constexpr (ND > 0) {
        float2 a, b, d;
        rf3<false>(rf_d2(c.sa, c.ca), x.fp[0], x.fm[0], x.z[0], a, b, d);
        g_a += c.al * (re_dot(l.fp[4], a) + re_dot(l.fm[4], b) + re_dot(l.z[4], d));
        rf3_adj<false>(m1, l.fp[4], l.fm[4], l.z[4], sp, sm, sz);
    }
    *ga += B1 * g_a + g_al;
#pragma unroll
    for (int j = 0; j <= ND; j++) {
        float2 a, b, d;
        rf3_adj<true>(m, l.fp[j], l.fm[j], l.z[j], a, b, d);
        l.fp[j] = a;
        l.fm[j] = b;
        l.z[j] = d;
    }
    if constexpr (ND > 0) {
        l.fp[0] = c_add(l.fp[0], c_scale(sp, c.al));
        l.fm[0] = c_add(l.fm[0], c_scale(sm, c.al));
        l.z[0] = c_add(l.z[0], c_scale(sz, c.al));
    }
}

// --- state storage (checkpoints / segment buffer), lanes contiguous -------

template <int ND>
__device__ __forceinline__ void st_store(float *base, int w, int k, const Aug<ND> &y)
{
    if (k >= w) return;
#pragma unroll
    for (int j = 0; j <= ND; j++) {
        base[(6 * j + 0) * w + k] = y.fp[j].x;
        base[(6 * j + 1) * w + k] = y.fp[j].y;
        base[(6 * j + 2) * w + k] = y.fm[j].x;
        base[(6 * j + 3) * w + k] = y.fm[j].y;
        base[(6 * j + 4) * w + k] = y.z[j].x;
        base[(6 * j + 5) * w + k] = y.z[j].y;
    }
}

template <int ND>
__device__ __forceinline__ void st_load(const float *base, int w, int k, Aug<ND> &y)
{
#pragma unroll
    for (int j = 0; j <= ND; j++) {
        if (k < w) {
            y.fp[j] = {base[(6 * j + 0) * w + k], base[(6 * j + 1) * w + k]};
            y.fm[j] = {base[(6 * j + 2) * w + k], base[(6 * j + 3) * w + k]};
            y.z[j] = {base[(6 * j + 4) * w + k], base[(6 * j + 5) * w + k]};
        } else {
            y.fp[j] = y.fm[j] = y.z[j] = {0.f, 0.f};
        }
    }
}

__device__ __forceinline__ Atom load_atom(const float *theta, size_t a)
{
    Atom p;
    p.T1 = theta[4 * a];
    p.T2 = theta[4 * a + 1];
    p.B0 = theta[4 * a + 2];
    p.B1 = theta[4 * a + 3];
    p.iT1 = 1.f / p.T1;
    p.iT2 = 1.f / p.T2;
    return p;
}

// ---------------------------------------------------------------------------
// Lane-per-state passes (the CRB gradient and the Gauss-Newton primal pass):
// atom of the calling lane's group
template <int SUB>
__device__ __forceinline__ size_t group_atom()
{
    return ((size_t)blockIdx.x * kWaves + (threadIdx.x >> 6)) * (64 / SUB) +
           ((threadIdx.x & 63) / SUB);
}


// ---------------------------------------------------------------------------
// Reverse sweeps.  ND = 0: L = Re sum conj(v_n) s_n (adjoint).  ND = 4: the
// CRB loss sum_d wts_d [F^-1]_dd / scale_d^2, F = Re(J^H J) (unit noise), whose
// gradient needs the adjoint of the tangent recursion.
struct RevArgs {
    const DTr *seq;
    int T, w, nck;
    const float *theta;
    size_t n;
    const float2 *cot;         // ND = 0: [n][T]
    float *grad_theta;         // ND = 0: [n][4] or null
    float *grad_seq;           // [2][T] accumulated (FA raw units, TR) or null
    float fa_unit;             // d alpha / d FA_raw
    float wts[4];              // ND = 4
    double *loss;              // ND = 4: [n]
    float *scratch;            // per atom diff_scratch_floats(ND, T, w) floats
    int seq_lds;               // grad_seq summed per work-group in LDS
    float *seq_part;           // or per atom, [n][2][T], summed by seq_reduce_kernel
};

// grad_seq partial sums go through LDS up to this T (2T floats)
constexpr int kSeqLdsMaxT = 8192;

size_t rev_lds_bytes(RevArgs &A)
{
    A.seq_lds = A.grad_seq != nullptr && !A.seq_part && A.T <= kSeqLdsMaxT;
    return A.seq_lds ? (size_t)2 * A.T * sizeof(float) : 0;
}

// Fisher matrix from lane 0's running sums; returns false if singular.
__device__ bool inv4(const double *A, double *X)
{
    double m[4][8];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 8; c++) m[r][c] = c < 4 ? A[4 * r + c] : (c - 4 == r ? 1.0 : 0.0);
    for (int c = 0; c < 4; c++) {
        int p = c;
        for (int r = c + 1; r < 4; r++)
            if (fabs(m[r][c]) > fabs(m[p][c])) p = r;
        if (m[p][c] == 0.0) return false;
        if (p != c)
            for (int j = 0; j < 8; j++) {
                const double t = m[c][j];
                m[c][j] = m[p][j];
                m[p][j] = t;
            }
        const double inv = 1.0 / m[c][c];
        for (int j = 0; j < 8; j++) m[c][j] *= inv;
        for (int r = 0; r < 4; r++)
            if (r != c) {
                const double f = m[r][c];
                for (int j = 0; j < 8; j++) m[r][j] -= f * m[c][j];
            }
    }
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) X[4 * r + c] = m[r][c + 4];
    return true;
}

// One atom's sweeps of the CRB gradient (the lane group of SUB lanes, lane k
// = state k); gs_lds is the work-group's grad_seq partial sums when A.seq_lds.
template <int ND, int SUB>
__device__ __forceinline__ void rev_atom(const RevArgs &A, size_t a, float *gs_lds, TrA *coef_lds)
{
    static_assert(ND == 4, "the adjoint (ND = 0) is adj_kernel");
    constexpr int CKPT = kCkpt;
    const int k = threadIdx.x & (SUB - 1);
    const int w = A.w, T = A.T;
    const int C = 6 * (1 + ND) * w;
    float *ck = A.scratch + a * (size_t)(A.nck + kCkpt) * C;
    float *seg = ck + (size_t)A.nck * C;      // the segment's TR-start states
    const Atom p = load_atom(A.theta, a);
    CoefCache<SUB> cc = coef_cache<SUB>(coef_lds);

    // forward sweep: checkpoints every kCkpt TRs; Fisher sums (ND = 4, from
    // frame readouts: Re(J^H J) does not see the frame phases)
    const float2 ro0 = frame0(A.seq);
    Aug<ND> y;
    f_init(y, p, k, ro0);
    float F[10];
#pragma unroll
    for (int i = 0; i < 10; i++) F[i] = 0.f;
    for (int t = 0; t < T; t++) {
        if (t % CKPT == 0) st_store(ck + (size_t)(t / CKPT) * C, w, k, y);
        const TrK c = coef<SUB>(cc, A.seq, T, p, t, k);
        Aug<ND> r;
        f_tr<ND, SUB>(y, c, p, k, w, &r);
        if (k == 0) {
            int i = 0;
#pragma unroll
            for (int d = 0; d < 4; d++)
#pragma unroll
                for (int e = d; e < 4; e++) F[i++] += re_dot(r.fp[1 + d], r.fp[1 + e]);
        }
    }
    // loss and dL/dF, on lane 0 in FP64, broadcast as G
    float G[10];
#pragma unroll
    for (int i = 0; i < 10; i++) G[i] = 0.f;
    if (k == 0) {
        double Fd[16], Fi[16];
        int i = 0;
        for (int d = 0; d < 4; d++)
            for (int e = d; e < 4; e++, i++) Fd[4 * d + e] = Fd[4 * e + d] = F[i];
        double loss = 0.0;
        if (inv4(Fd, Fi)) {
            double Wd[4];
            for (int d = 0; d < 4; d++) {
                const double sc = d == 0 ? p.T1 : (d == 1 ? p.T2 : 1.0);
                Wd[d] = A.wts[d] / (sc * sc);
                loss += Wd[d] * Fi[5 * d];
            }
            // dL/dF = -F^-1 W F^-1
            i = 0;
            for (int d = 0; d < 4; d++)
                for (int e = d; e < 4; e++, i++) {
                    double g = 0.0;
                    for (int m = 0; m < 4; m++) g -= Fi[4 * d + m] * Wd[m] * Fi[4 * m + e];
                    G[i] = (float)g;
                }
        } else {
            loss = __builtin_nan("");
        }
        A.loss[a] = loss;
    }
#pragma unroll
    for (int i = 0; i < 10; i++) G[i] = bcast0<SUB>(G[i]);

    // reverse sweep, segment by segment
    Aug<ND> l;
#pragma unroll
    for (int j = 0; j <= ND; j++) l.fp[j] = l.fm[j] = l.z[j] = {0.f, 0.f};
    const float beta = kCB0 * p.B0;
    // the checkpoint of a segment is loaded while the segment after it (in
    // time) is reversed
    Aug<ND> cknext;
    st_load(ck + (size_t)(A.nck - 1) * C, w, k, cknext);
    for (int sgi = A.nck - 1; sgi >= 0; sgi--) {
        const int t0 = sgi * CKPT;
        const int t1 = min(T, t0 + CKPT);
        y = cknext;
        if (sgi > 0) st_load(ck + (size_t)(sgi - 1) * C, w, k, cknext);
        for (int t = t0; t < t1; t++) {
            st_store(seg + (size_t)(t - t0) * C, w, k, y);
            f_tr<ND, SUB>(y, coef<SUB>(cc, A.seq, T, p, t, k), p, k, w, nullptr);
        }
        // the TR-start state of the next (earlier) TR is loaded while this
        // TR is reversed
        Aug<ND> ynext;
        st_load(seg + (size_t)(t1 - 1 - t0) * C, w, k, ynext);
        for (int t = t1 - 1; t >= t0; t--) {
            const Aug<ND> yn = ynext;
            if (t > t0) st_load(seg + (size_t)(t - 1 - t0) * C, w, k, ynext);
            const TrK c = coef<SUB>(cc, A.seq, T, p, t, k);
            float ga = 0.f, gt = 0.f;
            // the TR's forward chain once: x1 = rf(yn), x2 = relax_a(x1),
            // x3 = phase_a(x2); the reverse steps take their inputs from it
            Aug<ND> x1 = yn;
            f_rf(x1, c);
            Aug<ND> x2 = x1;
            f_relax(x2, c.e1a, c.e2a, c.te, p, k == 0);
            Aug<ND> x3 = x2;
            f_phase(x3, c.pa, c.te);
            // relax(TR - TE): input = dephase(phase_b(x3))
            {
                Aug<ND> x = x3;
                f_phase(x, c.pb, c.t2);
                f_dephase<ND, SUB>(x, k, w);
                r_relax(l, x, c.e1b, c.e2b, c.t2, p, k == 0, &gt);
            }
            r_dephase<ND, SUB>(l, k, w);
            // phase(TR - TE) and the readout: input = x3
            r_phase(l, x3, c.pb, c.t2, beta, &gt);
            if (k == 0) {
                // dL/dJ_e = 2 sum_d G_de J_d
                float2 J[4];
#pragma unroll
                for (int d = 0; d < 4; d++) J[d] = x3.fp[1 + d];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    float2 cg = {0.f, 0.f};
#pragma unroll
                    for (int d = 0; d < 4; d++) {
                        const int i = d <= e ? (d * 7 - d * d) / 2 + e : (e * 7 - e * e) / 2 + d;
                        cg = c_add(cg, c_scale(J[d], 2.f * G[i]));
                    }
                    l.fp[1 + e] = c_add(l.fp[1 + e], cg);
                }
            }
            // phase(TE): input = x2
            r_phase(l, x2, c.pa, c.te, beta, nullptr);
            // relax(TE): input = x1
            r_relax(l, x1, c.e1a, c.e2a, c.te, p, k == 0, nullptr);
            r_rf(l, yn, c, p.B1, &ga);
            if (A.grad_seq) {
                ga = group_sum<SUB>(ga);
                gt = group_sum<SUB>(gt);
                if (k == 0) {
                    if (A.seq_part) {
                        A.seq_part[a * 2 * T + t] = ga * A.fa_unit;
                        A.seq_part[a * 2 * T + T + t] = gt;
                    } else if (A.seq_lds) {
                        atomicAdd(&gs_lds[t], ga * A.fa_unit);
                        atomicAdd(&gs_lds[T + t], gt);
                    } else {
                        atomicAdd(&A.grad_seq[t], ga * A.fa_unit);
                        atomicAdd(&A.grad_seq[T + t], gt);
                    }
                }
            }
        }
    }
}

// grad_seq of the atoms of a work-group is summed in LDS (2T floats of
// dynamic LDS when A.seq_lds): one LDS add per atom and TR, then one global
// atomic per entry and work-group instead of one per atom.
template <int ND, int SUB>
__global__ __launch_bounds__(64 * kWaves, 1) void rev_kernel(RevArgs A)
{
    extern __shared__ float gs_lds[];
    __shared__ TrA coef_lds[kCoefLds];
    if (A.seq_lds) {
        for (int i = threadIdx.x; i < 2 * A.T; i += blockDim.x) gs_lds[i] = 0.f;
        __syncthreads();
    }
    const size_t a = group_atom<SUB>();
    if (a < A.n)
        rev_atom<ND, SUB>(A, a, gs_lds, coef_lds);
    if (A.seq_lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < 2 * A.T; i += blockDim.x) {
            const float v = gs_lds[i];
            if (v != 0.f) atomicAdd(&A.grad_seq[i], v);
        }
    }
}

// grad_seq[i] += sum over atoms of part[a][i], i < T2: block (x, y) sums
// entries 256 x .. 256 x + 255 over atoms y, y + gridDim.y, ...
__global__ __launch_bounds__(256) void seq_reduce_kernel(const float *part, size_t n, int T2,
                                                         float *grad_seq)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= T2) return;
    float v = 0.f;
    for (size_t a = blockIdx.y; a < n; a += gridDim.y) v += part[a * T2 + i];
    atomicAdd(&grad_seq[i], v);
}

// ---------------------------------------------------------------------------
// Adjoint (ND = 0) in the lane-group form: an atom on L lanes (8 for w <= 32,
// 16 for w <= 64), lane j holding EPG states 4j .. 4j + 3, so the per-atom
// work of a TR (its coefficients, the readout, the sums over states) is
// shared by L lanes and the four states of a lane are four independent
// chains.  The forward sweep stores a checkpoint every kAdjSeg = 4 TRs (the
// atom's 4L states, lanes contiguous, 16-byte pieces); the reverse sweep
// recomputes a segment's TR-start states and reverses its TRs from them, the
// intra-TR intermediates recomputed from each TR-start state.  The operators
// and their adjoints are those of the lane-per-state form above (r_relax,
// r_dephase, r_phase, r_rf), rearranged:
//  - states >= w are zero at every TR start, and their adjoints at every
//    step, so the truncation of dephase_gradients reduces to F+_w <- 0
//    (folded into the relaxation that follows the shift) and, in reverse,
//    l-_{w-1} <- 0 before the shift (folded into its relaxation's scaling);
//    the values that cross a group boundary in the shifts are replaced by
//    selects, never by products, so the lanes of another atom cannot leak in;
//  - the forward step applies relax(TE), phase(TE) and phase(TR - TE) as one
//    complex factor e2a e^{i beta TE} e^{i beta (TR - TE)} (times the frame
//    change) on F+ and its conjugate on F-;
//  - the phase adjoints take the operator's output y: dL/dt of y = P x is
//    Re <l, i beta y> on F+ (minus on F-), so phase(TR - TE)'s term is read off
//    the shifted state that the relaxation adjoint already holds (the shift
//    maps (i y+, -i y-) onto itself) and phase(TE)'s off its output.
constexpr int kAdjM = 4;     // states per lane
constexpr int kAdjSeg = 4;   // TRs per checkpoint segment

struct AState {
    float2 fp[kAdjM], fm[kAdjM], z[kAdjM];
};

// the lane's place in its group and in the state range
struct AdjLane {
    int j;
    bool first, last;        // j == 0, j == L - 1
    bool atw[kAdjM];         // state 4j + i == w
    bool atw1[kAdjM];        // state 4j + i == w - 1
};

// per-TR coefficients of one atom (tr_coef), 64 bytes
struct alignas(16) AdjCoef {
    float sa, ca, e1a, e2a;
    float e1b, e2b, al, te;
    float2 pa, pb;
    float2 ro;
    float t2, pad;
};
static_assert(sizeof(AdjCoef) == 64, "AdjCoef layout");

__device__ __forceinline__ AdjCoef adj_coef_of(const DTr &q, const Atom &p)
{
    const TrK c = tr_coef(q, p);
    AdjCoef e;
    e.sa = c.sa;
    e.ca = c.ca;
    e.e1a = c.e1a;
    e.e2a = c.e2a;
    e.e1b = c.e1b;
    e.e2b = c.e2b;
    e.al = c.al;
    e.te = c.te;
    e.pa = c.pa;
    e.pb = c.pb;
    e.ro = c.ro;
    e.t2 = c.t2;
    e.pad = 0.f;
    return e;
}

// The coefficients of the group's atom, cached L TRs at a time in LDS: lane
// j computes TR b L + j, the group reads each TR's entry back (segments of
// kAdjSeg TRs never straddle a block).  All groups of a wave walk the same
// TRs, so a block changes for the whole wave at once.
template <int L>
struct AdjCache {
    AdjCoef *blk;   // the group's L entries
    int cur;        // block held, -1: none
};

template <int L>
__device__ __forceinline__ AdjCoef adj_coef(AdjCache<L> &cc, const DTr *seq, int T, const Atom &p,
                                            int t, int j)
{
    const int b = t / L;
    if (b != cc.cur) {
        wave_sync();
        const int tt = b * L + j;
        if (tt < T) cc.blk[j] = adj_coef_of(seq[tt], p);
        cc.cur = b;
        wave_sync();
    }
    return cc.blk[t - b * L];
}

// Lane i of a 16-lane DPP row receives lane i - 1 (row_shr:1) or i + 1
// (row_shl:1); the row's end lanes receive 0.  Groups of 8 or 16 lanes never
// straddle a row.
__device__ __forceinline__ float row_shr1(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                 0x111, 0xF, 0xF, true));
}
__device__ __forceinline__ float row_shl1(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                 0x101, 0xF, 0xF, true));
}
__device__ __forceinline__ float2 row_shr1(float2 v) { return {row_shr1(v.x), row_shr1(v.y)}; }
__device__ __forceinline__ float2 row_shl1(float2 v) { return {row_shl1(v.x), row_shl1(v.y)}; }

// sum over the L lanes of the group (8: quads, then half rows; 16: rows)
template <int L>
__device__ __forceinline__ float adj_gsum(float v)
{
    v = dpp_add<0xB1>(v);
    v = dpp_add<0x4E>(v);
    v = dpp_add<0x141>(v);
    if (L == 16) v = dpp_add<0x140>(v);
    return v;
}

__device__ __forceinline__ float re_idot(float2 l, float2 y) { return re_dot(l, c_i(y)); }

__device__ __forceinline__ void a_rf(AState &y, float sa, float ca)
{
    const C4 m = rf_c(sa, ca);
#pragma unroll
    for (int i = 0; i < kAdjM; i++) rf3<true>(m, y.fp[i], y.fm[i], y.z[i], y.fp[i], y.fm[i], y.z[i]);
}

// relax over an interval with F+ <- P F+, F- <- conj(P) F- (P = e2 times the
// precession factors), Z <- e1 Z + (1 - e1) on state 0
__device__ __forceinline__ void a_decay(AState &y, float2 P, float e1, const AdjLane &ln)
{
    const float2 Q = c_conj(P);
#pragma unroll
    for (int i = 0; i < kAdjM; i++) {
        y.fp[i] = c_mul(P, y.fp[i]);
        y.fm[i] = c_mul(Q, y.fm[i]);
        y.z[i] = c_scale(y.z[i], e1);
    }
    if (ln.first) y.z[0].x += 1.f - e1;
}

// dephase_gradients (functions.cu:619-668): F+_k <- F+_{k-1}, F-_k <- F-_{k+1},
// F+_0 <- conj(F-_0); F+_w is left for the caller (states >= w are zero).
__device__ __forceinline__ void a_shift(AState &y, const AdjLane &ln)
{
    const float2 nxt = row_shl1(y.fm[0]);            // F-_{4j+4}
    const float2 prv = row_shr1(y.fp[kAdjM - 1]);    // F+_{4j-1}
#pragma unroll
    for (int i = 0; i < kAdjM - 1; i++) y.fm[i] = y.fm[i + 1];
    y.fm[kAdjM - 1] = ln.last ? float2{0.f, 0.f} : nxt;
#pragma unroll
    for (int i = kAdjM - 1; i > 0; i--) y.fp[i] = y.fp[i - 1];
    y.fp[0] = ln.first ? c_conj(y.fm[0]) : prv;
}

// one TR from its TR-start state to the next
__device__ __forceinline__ void a_tr(AState &y, const AdjCoef &c, const AdjLane &ln)
{
    a_rf(y, c.sa, c.ca);
    a_decay(y, c_scale(c_mul(c.pa, c.pb), c.e2a), c.e1a, ln);
    a_shift(y, ln);
#pragma unroll
    for (int i = 0; i < kAdjM; i++) {
        y.fp[i] = c_scale(y.fp[i], ln.atw[i] ? 0.f : c.e2b);   // F+_w <- 0
        y.fm[i] = c_scale(y.fm[i], c.e2b);
        y.z[i] = c_scale(y.z[i], c.e1b);
    }
    if (ln.first) y.z[0].x += 1.f - c.e1b;
}

// init_rf_pulse(180 deg, 0 deg, TI) in the frame of TR 0 (f_init)
__device__ __forceinline__ void a_init(AState &y, const Atom &p, const AdjLane &ln, float2 ro0)
{
#pragma unroll
    for (int i = 0; i < kAdjM; i++) y.fp[i] = y.fm[i] = y.z[i] = {0.f, 0.f};
    if (!ln.first) return;
    const float a0 = (float)(180.0 * M_PI / 180.0f);
    const float sa0 = sinf(a0), ca0 = cosf(a0);
    const float e1 = expf(-kTI * p.iT1), e2 = expf(-kTI * p.iT2);
    const float2 tmp = {0.f, e2 * sa0};
    y.fp[0] = c_mul(c_conj(ro0), c_conj(tmp));
    y.fm[0] = c_mul(ro0, tmp);
    y.z[0] = {e1 * ca0 + 1.f - e1, 0.f};
}

// adjoint of relax (r_relax) over an interval t with input x: *gt (if
// given) receives dL/dt, *ub the derivative Re <l, (i x+, -i x-)> that a
// phase step whose output was shifted into x contributes (scaled by the
// relaxation, as the adjoint there carries it).  MASK: l-_{w-1} <- 0.
template <bool MASK>
__device__ __forceinline__ void a_rrelax(AState &l, const AState &x, float e1, float e2, float t,
                                         const Atom &p, const AdjLane &ln, float gth[4], float *gt,
                                         float *ub)
{
    const float i1 = p.iT1, i2 = p.iT2;
    float sf = 0.f, sz = 0.f, u = 0.f;
#pragma unroll
    for (int i = 0; i < kAdjM; i++) {
        sf += re_dot(l.fp[i], x.fp[i]) + re_dot(l.fm[i], x.fm[i]);
        sz += re_dot(l.z[i], x.z[i]);
        if (ub) u += re_idot(l.fp[i], x.fp[i]) - re_idot(l.fm[i], x.fm[i]);
    }
    if (ln.first) sz -= l.z[0].x;   // the source term: Z_0 - 1
    gth[0] += sz * (e1 * t * i1 * i1);
    gth[1] += sf * (e2 * t * i2 * i2);
    if (gt) *gt += sf * (-e2 * i2) + sz * (-e1 * i1);
    if (ub) *ub = u * e2;
#pragma unroll
    for (int i = 0; i < kAdjM; i++) {
        l.fp[i] = c_scale(l.fp[i], e2);
        l.fm[i] = c_scale(l.fm[i], (MASK && ln.atw1[i]) ? 0.f : e2);
        l.z[i] = c_scale(l.z[i], e1);
    }
}

// adjoint of the shift (r_dephase)
__device__ __forceinline__ void a_rshift(AState &l, const AdjLane &ln)
{
    const float2 nxt = row_shl1(l.fp[0]);            // l+_{4j+4}
    const float2 prv = row_shr1(l.fm[kAdjM - 1]);    // l-_{4j-1}
    const float2 lp0 = l.fp[0];                      // l+_0 on lane 0
#pragma unroll
    for (int i = 0; i < kAdjM - 1; i++) l.fp[i] = l.fp[i + 1];
    l.fp[kAdjM - 1] = ln.last ? float2{0.f, 0.f} : nxt;
#pragma unroll
    for (int i = kAdjM - 1; i > 0; i--) l.fm[i] = l.fm[i - 1];
    l.fm[0] = ln.first ? float2{0.f, 0.f} : prv;
    if (ln.first) l.fm[1] = c_add(l.fm[1], c_conj(lp0));
}

// adjoint of a phase step: l+ <- conj(P) l+, l- <- P l-
__device__ __forceinline__ void a_rphase(AState &l, float2 P)
{
    const float2 Q = c_conj(P);
#pragma unroll
    for (int i = 0; i < kAdjM; i++) {
        l.fp[i] = c_mul(Q, l.fp[i]);
        l.fm[i] = c_mul(P, l.fm[i]);
    }
}

// adjoint of the RF (r_rf): x the operator's input
__device__ __forceinline__ void a_rrf(AState &l, const AState &x, const AdjCoef &c, float B1,
                                      float *ga, float gth[4])
{
    const C4 m = rf_c(c.sa, c.ca), m1 = rf_d1(c.sa, c.ca);
    float g_a = 0.f;
#pragma unroll
    for (int i = 0; i < kAdjM; i++) {
        float2 a, b, d;
        rf3<false>(m1, x.fp[i], x.fm[i], x.z[i], a, b, d);
        g_a += re_dot(l.fp[i], a) + re_dot(l.fm[i], b) + re_dot(l.z[i], d);
        rf3_adj<true>(m, l.fp[i], l.fm[i], l.z[i], l.fp[i], l.fm[i], l.z[i]);
    }
    *ga += B1 * g_a;
    gth[3] += c.al * g_a;
}

// the reverse of TR n from its TR-start state yn; cot = dL/ds_n
__device__ __forceinline__ void a_rev_tr(AState &l, const AState &yn, const AdjCoef &c,
                                         const Atom &p, float beta, const AdjLane &ln, float2 cot,
                                         float &ga, float &gt, float gth[4])
{
    AState x3 = yn;   // x3 = phase_a(relax_a(rf(yn))): the readout state
    a_rf(x3, c.sa, c.ca);
    a_decay(x3, c_scale(c.pa, c.e2a), c.e1a, ln);
    float ub;
    {
        // relax(TR - TE) over x = dephase(phase_b(x3)); phase(TR - TE)'s
        // derivative read off x
        AState x = x3;
        const float2 Q = c_conj(c.pb);
#pragma unroll
        for (int i = 0; i < kAdjM; i++) {
            x.fp[i] = c_mul(c.pb, x.fp[i]);
            x.fm[i] = c_mul(Q, x.fm[i]);
        }
        a_shift(x, ln);
        a_rrelax<true>(l, x, c.e1b, c.e2b, c.t2, p, ln, gth, &gt, &ub);
    }
    a_rshift(l, ln);
    gt += beta * ub;
    gth[2] += kCB0 * c.t2 * ub;
    a_rphase(l, c.pb);
    // the readout s = ro F+_0: L = Re conj(v) s gives the cotangent conj(ro) v
    if (ln.first) l.fp[0] = c_add(l.fp[0], c_mul(c_conj(c.ro), cot));
    {
        // phase(TE), output x3
        float u = 0.f;
#pragma unroll
        for (int i = 0; i < kAdjM; i++) u += re_idot(l.fp[i], x3.fp[i]) - re_idot(l.fm[i], x3.fm[i]);
        gth[2] += kCB0 * c.te * u;
        a_rphase(l, c.pa);
    }
    AState x1 = yn;   // relax(TE) over x1 = rf(yn)
    a_rf(x1, c.sa, c.ca);
    a_rrelax<false>(l, x1, c.e1a, c.e2a, c.te, p, ln, gth, nullptr, nullptr);
    a_rrf(l, yn, c, p.B1, &ga, gth);
}

// a group's states as 6 pieces of 16 bytes per lane, lanes contiguous per
// piece (stride S pieces)
template <int S>
__device__ __forceinline__ void a_store(float4 *c, const AState &y)
{
    c[0 * S] = make_float4(y.fp[0].x, y.fp[0].y, y.fp[1].x, y.fp[1].y);
    c[1 * S] = make_float4(y.fp[2].x, y.fp[2].y, y.fp[3].x, y.fp[3].y);
    c[2 * S] = make_float4(y.fm[0].x, y.fm[0].y, y.fm[1].x, y.fm[1].y);
    c[3 * S] = make_float4(y.fm[2].x, y.fm[2].y, y.fm[3].x, y.fm[3].y);
    c[4 * S] = make_float4(y.z[0].x, y.z[0].y, y.z[1].x, y.z[1].y);
    c[5 * S] = make_float4(y.z[2].x, y.z[2].y, y.z[3].x, y.z[3].y);
}

template <int S>
__device__ __forceinline__ void a_load(const float4 *c, AState &y)
{
    float4 v;
    v = c[0 * S]; y.fp[0] = {v.x, v.y}; y.fp[1] = {v.z, v.w};
    v = c[1 * S]; y.fp[2] = {v.x, v.y}; y.fp[3] = {v.z, v.w};
    v = c[2 * S]; y.fm[0] = {v.x, v.y}; y.fm[1] = {v.z, v.w};
    v = c[3 * S]; y.fm[2] = {v.x, v.y}; y.fm[3] = {v.z, v.w};
    v = c[4 * S]; y.z[0] = {v.x, v.y}; y.z[1] = {v.z, v.w};
    v = c[5 * S]; y.z[2] = {v.x, v.y}; y.z[3] = {v.z, v.w};
}

// checkpoint floats per atom of adj_kernel<L>
size_t adj_ck_floats(int T, int L)
{
    return (size_t)((T + kAdjSeg - 1) / kAdjSeg) * 6 * 4 * L;
}

// One wave per work-group (64 / L atoms; small work-groups keep the last
// round of a launch short), two waves per SIMD (256 registers, 16 KB of LDS
// per wave): the segment's first and last TR-start states, the adjoint, the
// next segment's checkpoint and the intermediates of a TR in registers, the
// segment's middle TR-start states in LDS (the lane's own slots).
template <int L>
__global__ __launch_bounds__(64, 2) void adj_kernel(RevArgs A)
{
    static_assert(L == 8 || L == 16, "8 or 16 lanes per atom");
    static_assert(L % kAdjSeg == 0, "segments must not straddle a coefficient block");
    static_assert(kAdjSeg == 4, "the segment loop holds TR-start states 1, 2 in LDS");
    __shared__ AdjCoef coef_lds[64];
    __shared__ float4 seg_lds[2 * 6 * 64];
    const int lane = threadIdx.x;
    const size_t ag = (size_t)blockIdx.x * (64 / L) + lane / L;
    // lanes of a missing atom run on the last one and store nothing, so that
    // every lane of a row takes part in the shifts and sums
    const bool valid = ag < A.n;
    const size_t a = valid ? ag : A.n - 1;
    const int T = A.T, w = A.w;
    AdjLane ln;
    ln.j = lane & (L - 1);
    ln.first = ln.j == 0;
    ln.last = ln.j == L - 1;
#pragma unroll
    for (int i = 0; i < kAdjM; i++) {
        ln.atw[i] = kAdjM * ln.j + i == w;
        ln.atw1[i] = kAdjM * ln.j + i == w - 1;
    }
    const int j = ln.j;
    const Atom p = load_atom(A.theta, a);
    AdjCache<L> cc{coef_lds + (lane & ~(L - 1)), -1};
    const int nck = (T + kAdjSeg - 1) / kAdjSeg;
    float4 *ck = reinterpret_cast<float4 *>(A.scratch) + a * (size_t)nck * 6 * L + j;

    // forward sweep: a checkpoint every kAdjSeg TRs
    const float2 ro0 = frame0(A.seq);
    AState y;
    a_init(y, p, ln, ro0);
    for (int t = 0; t < T; t++) {
        if (t % kAdjSeg == 0 && valid) a_store<L>(ck + (size_t)(t / kAdjSeg) * 6 * L, y);
        a_tr(y, adj_coef<L>(cc, A.seq, T, p, t, j), ln);
    }

    // reverse sweep, segment by segment; the checkpoint and the cotangents
    // of the segment before (in time) are loaded while a segment is reversed
    AState l;
#pragma unroll
    for (int i = 0; i < kAdjM; i++) l.fp[i] = l.fm[i] = l.z[i] = {0.f, 0.f};
    float gth[4] = {0.f, 0.f, 0.f, 0.f};
    const float beta = kCB0 * p.B0;
    const float2 *cr = A.cot + a * (size_t)T;
    AState ckn;
    float2 cvn;
    auto seg_in = [&](int sg) {
        a_load<L>(ck + (size_t)sg * 6 * L, ckn);
        const int tt = sg * kAdjSeg + j;
        cvn = (j < kAdjSeg && tt < T) ? cr[tt] : float2{0.f, 0.f};
    };
    const int gbase = lane & ~(L - 1);
    float4 *slot = seg_lds + lane;   // slot s: TR-start state s + 1
    auto rev_one = [&](int t, int s, const AState &yn, float2 cv) {
        const AdjCoef c = adj_coef<L>(cc, A.seq, T, p, t, j);
        const float2 cot = {__shfl(cv.x, gbase + s), __shfl(cv.y, gbase + s)};
        float ga = 0.f, gt = 0.f;
        a_rev_tr(l, yn, c, p, beta, ln, cot, ga, gt, gth);
        if (A.seq_part) {
            ga = adj_gsum<L>(ga);
            gt = adj_gsum<L>(gt);
            if (ln.first && valid) {
                A.seq_part[a * 2 * T + t] = ga * A.fa_unit;
                A.seq_part[a * 2 * T + T + t] = gt;
            }
        }
    };
    seg_in(nck - 1);
    int sg = nck - 1;
    const int tail = T - sg * kAdjSeg;   // 1 .. kAdjSeg TRs in the last segment
    if (tail < kAdjSeg) {
        const int t0 = sg * kAdjSeg;
        const AState s0 = ckn;
        const float2 cv = cvn;
        if (sg > 0) seg_in(sg - 1);
        AState yy = s0;
#pragma unroll 1
        for (int s = 1; s < tail; s++) {
            a_tr(yy, adj_coef<L>(cc, A.seq, T, p, t0 + s - 1, j), ln);
            a_store<64>(slot + (s - 1) * 6 * 64, yy);
        }
#pragma unroll 1
        for (int s = tail - 1; s >= 1; s--) {
            AState yn;
            a_load<64>(slot + (s - 1) * 6 * 64, yn);
            rev_one(t0 + s, s, yn, cv);
        }
        rev_one(t0, 0, s0, cv);
        sg--;
    }
    for (; sg >= 0; sg--) {
        const int t0 = sg * kAdjSeg;
        const AState s0 = ckn;
        const float2 cv = cvn;
        if (sg > 0) seg_in(sg - 1);
        AState yy = s0;
        a_tr(yy, adj_coef<L>(cc, A.seq, T, p, t0, j), ln);
        a_store<64>(slot, yy);
        a_tr(yy, adj_coef<L>(cc, A.seq, T, p, t0 + 1, j), ln);
        a_store<64>(slot + 6 * 64, yy);
        a_tr(yy, adj_coef<L>(cc, A.seq, T, p, t0 + 2, j), ln);
        rev_one(t0 + 3, 3, yy, cv);
#pragma unroll 1
        for (int s = 2; s >= 1; s--) {
            AState yn;
            a_load<64>(slot + (s - 1) * 6 * 64, yn);
            rev_one(t0 + s, s, yn, cv);
        }
        rev_one(t0, 0, s0, cv);
    }
    if (A.grad_theta) {
        // init_rf_pulse's T1/T2 dependence (state 0, lane 0)
        if (ln.first) {
            const float a0 = (float)(180.0 * M_PI / 180.0f);
            const float sa0 = sinf(a0), ca0 = cosf(a0);
            const float e1 = expf(-kTI * p.iT1), e2 = expf(-kTI * p.iT2);
            const float e1d = e1 * kTI * p.iT1 * p.iT1, e2d = e2 * kTI * p.iT2 * p.iT2;
            gth[0] += l.z[0].x * e1d * (ca0 - 1.f);
            const float2 t2 = {0.f, e2d * sa0};
            gth[1] += re_dot(l.fp[0], c_mul(c_conj(ro0), c_conj(t2))) +
                      re_dot(l.fm[0], c_mul(ro0, t2));
        }
#pragma unroll
        for (int d = 0; d < 4; d++) gth[d] = adj_gsum<L>(gth[d]);
        if (ln.first && valid)
#pragma unroll
            for (int d = 0; d < 4; d++) A.grad_theta[a * 4 + d] = gth[d];
    }
}

// ---------------------------------------------------------------------------
// Forward mode in the lane-group layout (the Jacobian and the Gauss-Newton
// tangent passes): the adjoint's layout (L lanes per atom, lane j holding
// states 4j .. 4j + 3) with each state's four tangents (T1, T2, B0, B1), 120
// registers per lane.  The tangent rules are those of f_rf / f_relax /
// f_phase, rearranged like the adjoint's forward steps: relax(TE) and
// phase(TE) are one factor F = e2 e^{i beta TE} on F+ (conj(F) on F-), and
// so are relax(TR - TE) and phase(TR - TE) after the shift, with
// their T2 and B0 source terms folded in before it as y_T2 += (t / T2^2) y
// and y_B0 += i c t y (on F-: -i c t y), since d(e2)/dT2 = e2 t / T2^2 and
// d(e^{i beta t})/dB0 = i c t e^{i beta t}; the truncation is the adjoint's
// (states >= w stay zero, F+_w <- 0 in the relaxation after the shift).
constexpr int kJd = 5;   // the state and its four tangents

struct JState {
    float2 fp[kAdjM][kJd], fm[kAdjM][kJd], z[kAdjM][kJd];
};

// init_rf_pulse with the tangents of its T1/T2 dependence (f_init)
__device__ __forceinline__ void j_init(JState &y, const Atom &p, const AdjLane &ln, float2 ro0)
{
#pragma unroll
    for (int i = 0; i < kAdjM; i++)
#pragma unroll
        for (int d = 0; d < kJd; d++) y.fp[i][d] = y.fm[i][d] = y.z[i][d] = {0.f, 0.f};
    if (!ln.first) return;
    const float a0 = (float)(180.0 * M_PI / 180.0f);
    const float sa0 = sinf(a0), ca0 = cosf(a0);
    const float e1 = expf(-kTI * p.iT1), e2 = expf(-kTI * p.iT2);
    const float2 tmp = {0.f, e2 * sa0};
    y.fp[0][0] = c_mul(c_conj(ro0), c_conj(tmp));
    y.fm[0][0] = c_mul(ro0, tmp);
    y.z[0][0] = {e1 * ca0 + 1.f - e1, 0.f};
    const float e1d = e1 * kTI * p.iT1 * p.iT1, e2d = e2 * kTI * p.iT2 * p.iT2;
    y.z[0][1] = {e1d * (ca0 - 1.f), 0.f};
    const float2 t2 = {0.f, e2d * sa0};
    y.fp[0][2] = c_mul(c_conj(ro0), c_conj(t2));
    y.fm[0][2] = c_mul(ro0, t2);
}

// RF: every component through M(alpha); the B1 tangent also gets alpha M_a y
__device__ __forceinline__ void j_rf(JState &y, const AdjCoef &c)
{
    const C4 m = rf_c(c.sa, c.ca), m1 = rf_d1(c.sa, c.ca);
#pragma unroll
    for (int i = 0; i < kAdjM; i++) {
        float2 sp, sm, sz;
        rf3<false>(m1, y.fp[i][0], y.fm[i][0], y.z[i][0], sp, sm, sz);
#pragma unroll
        for (int d = 0; d < kJd; d++)
            rf3<true>(m, y.fp[i][d], y.fm[i][d], y.z[i][d], y.fp[i][d], y.fm[i][d], y.z[i][d]);
        y.fp[i][4] = c_add(y.fp[i][4], c_scale(sp, c.al));
        y.fm[i][4] = c_add(y.fm[i][4], c_scale(sm, c.al));
        y.z[i][4] = c_add(y.z[i][4], c_scale(sz, c.al));
    }
}

// B0 source of a precession over c t (before its factor)
__device__ __forceinline__ void j_b0_source(JState &y, float ct)
{
#pragma unroll
    for (int i = 0; i < kAdjM; i++) {
        y.fp[i][3] = c_add(y.fp[i][3], c_scale(c_i(y.fp[i][0]), ct));
        y.fm[i][3] = c_add(y.fm[i][3], c_scale(c_i(y.fm[i][0]), -ct));
    }
}

// relax over an interval t with the precession over it: F (e2 times the
// precession factor) on F+, conj(F) on F-, e1 on Z with the source on state
// 0; g2 = t / T2^2 and e1d = e1 t / T1^2 give the T2 and T1 source terms.
// MASK: F+_w <- 0.
template <bool MASK>
__device__ __forceinline__ void j_decay(JState &y, float2 F, float e1, float g2, float e1d,
                                        const AdjLane &ln)
{
#pragma unroll
    for (int i = 0; i < kAdjM; i++) {
        float2 zr = y.z[i][0];
        if (i == 0 && ln.first) zr.x -= 1.f;
        y.fp[i][2] = c_add(y.fp[i][2], c_scale(y.fp[i][0], g2));
        y.fm[i][2] = c_add(y.fm[i][2], c_scale(y.fm[i][0], g2));
        const float2 Fi = (MASK && ln.atw[i]) ? float2{0.f, 0.f} : F;
        const float2 Qi = c_conj(F);
#pragma unroll
        for (int d = 0; d < kJd; d++) {
            y.fp[i][d] = c_mul(Fi, y.fp[i][d]);
            y.fm[i][d] = c_mul(Qi, y.fm[i][d]);
        }
#pragma unroll
        for (int d = 0; d < kJd; d++) y.z[i][d] = c_scale(y.z[i][d], e1);
        y.z[i][1] = c_add(y.z[i][1], c_scale(zr, e1d));
    }
    if (ln.first) y.z[0][0].x += 1.f - e1;
}

// the shift of every component (a_shift)
__device__ __forceinline__ void j_shift(JState &y, const AdjLane &ln)
{
#pragma unroll
    for (int d = 0; d < kJd; d++) {
        const float2 nxt = row_shl1(y.fm[0][d]);
        const float2 prv = row_shr1(y.fp[kAdjM - 1][d]);
#pragma unroll
        for (int i = 0; i < kAdjM - 1; i++) y.fm[i][d] = y.fm[i + 1][d];
        y.fm[kAdjM - 1][d] = ln.last ? float2{0.f, 0.f} : nxt;
#pragma unroll
        for (int i = kAdjM - 1; i > 0; i--) y.fp[i][d] = y.fp[i - 1][d];
        y.fp[0][d] = ln.first ? c_conj(y.fm[0][d]) : prv;
    }
}

// one TR; the readouts (state 0 and its tangents, true coordinates) into rd
// on the group's first lane
__device__ __forceinline__ void j_tr(JState &y, const AdjCoef &c, const Atom &p,
                                     const AdjLane &ln, float2 *rd)
{
    const float i1s = p.iT1 * p.iT1, i2s = p.iT2 * p.iT2;
    j_rf(y, c);
    j_b0_source(y, kCB0 * c.te);
    j_decay<false>(y, c_scale(c.pa, c.e2a), c.e1a, c.te * i2s, c.e1a * c.te * i1s, ln);
    if (ln.first)
#pragma unroll
        for (int d = 0; d < kJd; d++) rd[d] = c_mul(c.ro, y.fp[0][d]);
    // phase(TR - TE) commutes with the shift (a uniform factor on F+, its
    // conjugate on F-, and conj(conj(P) F-) = P conj(F-)), so it joins the
    // relaxation after the shift, its B0 source with it
    j_shift(y, ln);
    j_b0_source(y, kCB0 * c.t2);
    j_decay<true>(y, c_scale(c.pb, c.e2b), c.e1b, c.t2 * i2s, c.e1b * c.t2 * i1s, ln);
}

// the lane's place in its group of L (all lanes of a wave, 64 / L atoms)
template <int L>
__device__ __forceinline__ AdjLane adj_lane(int lane, int w)
{
    AdjLane ln;
    ln.j = lane & (L - 1);
    ln.first = ln.j == 0;
    ln.last = ln.j == L - 1;
#pragma unroll
    for (int i = 0; i < kAdjM; i++) {
        ln.atw[i] = kAdjM * ln.j + i == w;
        ln.atw1[i] = kAdjM * ln.j + i == w - 1;
    }
    return ln;
}

// One atom's forward sweep with tangents; every L TRs (and at the end) the
// group's lanes take one TR each of the staged readouts: fold(t, e) with e
// the state and its four tangents at TR t.
template <int L, typename Fold>
__device__ __forceinline__ void j_sweep(const DTr *seq, int T, const Atom &p, const AdjLane &ln,
                                        AdjCoef *coef, float2 (*rd)[kJd], Fold fold)
{
    AdjCache<L> cc{coef, -1};
    JState y;
    j_init(y, p, ln, frame0(seq));
    for (int t = 0; t < T; t++) {
        const int jj = t % L;
        j_tr(y, adj_coef<L>(cc, seq, T, p, t, ln.j), p, ln, rd[jj]);
        if (jj == L - 1 || t == T - 1) {
            wave_sync();
            if (ln.j <= jj) fold(t - jj + ln.j, rd[ln.j]);
            wave_sync();
        }
    }
}

// Jacobian: s [n][T] and J [n][4][T] (either may be null); one wave per
// work-group, two per SIMD
template <int L>
__global__ __launch_bounds__(64, 2) void jacl_kernel(const DTr *seq, int T, int w,
                                                     const float *theta, size_t n, float2 *sig,
                                                     float2 *jac)
{
    __shared__ AdjCoef coef_lds[64];
    __shared__ float2 rd_lds[64 / L][L][kJd];
    const int lane = threadIdx.x;
    const size_t a = (size_t)blockIdx.x * (64 / L) + lane / L;
    if (a >= n) return;   // whole groups: their lanes only meet their own group
    const AdjLane ln = adj_lane<L>(lane, w);
    j_sweep<L>(seq, T, load_atom(theta, a), ln, coef_lds + (lane & ~(L - 1)), rd_lds[lane / L],
               [&](int t, const float2 *e) {
                   if (sig) sig[a * T + t] = e[0];
                   if (jac)
#pragma unroll
                       for (int d = 0; d < 4; d++) jac[(a * 4 + d) * T + t] = e[1 + d];
               });
}

// ---------------------------------------------------------------------------
// Gauss-Newton (Levenberg-Marquardt damped) refinement of (T1, T2, B0, B1)
// per voxel with the complex proton density eliminated (variable
// projection): cost c = |x|^2 - |s^H x|^2 / |s|^2.  A step is kept only if
// it lowers c, so the cost history is monotone.
struct GnState;
struct GnArgs {
    const DTr *seq;
    int T, w, iters;
    const float2 *sig;         // [V][T]
    size_t V;
    float *theta;              // [V][4] in/out
    float *pd;                 // [V] or null: |s^H x| / |s|
    float *cost;               // [V][iters + 1] or null
    GnState *st;               // [V] state between launches
    int *list;                 // [V] voxels due a tangent pass
    int *count;                // their number
};

// a group's sums of one pass: S = s^H s, Sx = s^H x, U_d = s^H ds_d,
// W_d = ds_d^H x, D_de = ds_d^H ds_e (d <= e)
struct GnSums {
    float S;
    float2 Sx, U[4], W[4], D[10];
};

template <int ND, int SUB>
__device__ GnSums gn_pass(const DTr *seq, int T, int w, const Atom &p, const float2 *x, int k,
                         TrA *coef_lds, float2 *read_lds)
{
    GnSums g{};
    CoefCache<SUB> cc = coef_cache<SUB>(coef_lds);   // p differs per pass
    float2 *rd = read_stage<SUB>(read_lds);
    Aug<ND> y;
    f_init(y, p, k, frame0(seq));
    for (int t = 0; t < T; t++) {
        Aug<ND> r;
        const TrK c = coef<SUB>(cc, seq, T, p, t, k);
        f_tr<ND, SUB>(y, c, p, k, w, &r);
        const int j = t % SUB;
        if (k == 0)   // readouts in true coordinates
#pragma unroll
            for (int d = 0; d <= ND; d++) rd[j * (1 + ND) + d] = c_mul(c.ro, r.fp[d]);
        if (j == SUB - 1 || t == T - 1) {
            // lane k folds TR t - j + k into its partial sums
            wave_sync();
            if (k <= j) {
                const float2 *e = rd + k * (1 + ND);
                const float2 s = e[0], xv = x[t - j + k];
                g.S += re_dot(s, s);
                g.Sx = c_add(g.Sx, c_mul(c_conj(s), xv));
                if constexpr (ND > 0) {
                    int i = 0;
#pragma unroll
                    for (int d = 0; d < 4; d++) {
                        const float2 ds = e[1 + d];
                        g.U[d] = c_add(g.U[d], c_mul(c_conj(s), ds));
                        g.W[d] = c_add(g.W[d], c_mul(c_conj(ds), xv));
#pragma unroll
                        for (int q = d; q < 4; q++, i++) g.D[i] = c_add(g.D[i], c_mul(c_conj(ds), e[1 + q]));
                    }
                }
            }
            wave_sync();
        }
    }
    // the group's partial sums
    g.S = group_sum<SUB>(g.S);
    g.Sx = {group_sum<SUB>(g.Sx.x), group_sum<SUB>(g.Sx.y)};
    if constexpr (ND > 0) {
#pragma unroll
        for (int d = 0; d < 4; d++) {
            g.U[d] = {group_sum<SUB>(g.U[d].x), group_sum<SUB>(g.U[d].y)};
            g.W[d] = {group_sum<SUB>(g.W[d].x), group_sum<SUB>(g.W[d].y)};
        }
#pragma unroll
        for (int i = 0; i < 10; i++) g.D[i] = {group_sum<SUB>(g.D[i].x), group_sum<SUB>(g.D[i].y)};
    }
    return g;
}

// gn_pass with tangents in the lane-group layout: the group's lanes fold
// one TR each of every L staged readouts
template <int L>
__device__ GnSums gnl_pass(const DTr *seq, int T, const Atom &p, const float2 *x, const AdjLane &ln,
                           AdjCoef *coef, float2 (*rd)[kJd])
{
    GnSums g{};
    j_sweep<L>(seq, T, p, ln, coef, rd, [&](int t, const float2 *e) {
        const float2 s = e[0], xv = x[t];
        g.S += re_dot(s, s);
        g.Sx = c_add(g.Sx, c_mul(c_conj(s), xv));
        int i = 0;
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const float2 ds = e[1 + d];
            g.U[d] = c_add(g.U[d], c_mul(c_conj(s), ds));
            g.W[d] = c_add(g.W[d], c_mul(c_conj(ds), xv));
#pragma unroll
            for (int q = d; q < 4; q++, i++) g.D[i] = c_add(g.D[i], c_mul(c_conj(ds), e[1 + q]));
        }
    });
    g.S = adj_gsum<L>(g.S);
    g.Sx = {adj_gsum<L>(g.Sx.x), adj_gsum<L>(g.Sx.y)};
#pragma unroll
    for (int d = 0; d < 4; d++) {
        g.U[d] = {adj_gsum<L>(g.U[d].x), adj_gsum<L>(g.U[d].y)};
        g.W[d] = {adj_gsum<L>(g.W[d].x), adj_gsum<L>(g.W[d].y)};
    }
#pragma unroll
    for (int i = 0; i < 10; i++) g.D[i] = {adj_gsum<L>(g.D[i].x), adj_gsum<L>(g.D[i].y)};
    return g;
}

__device__ __forceinline__ bool theta_ok(const float *th)
{
    return th[0] > 0.f && th[1] > 0.f && th[3] > 0.f && isfinite(th[0]) && isfinite(th[1]) &&
           isfinite(th[2]) && isfinite(th[3]);
}

__device__ __forceinline__ Atom make_atom(const float *th)
{
    Atom p;
    p.T1 = th[0];
    p.T2 = th[1];
    p.B0 = th[2];
    p.B1 = th[3];
    p.iT1 = 1.f / p.T1;
    p.iT2 = 1.f / p.T2;
    return p;
}

// Gauss-Newton state of one voxel between the launches of an iteration.
struct GnState {
    float th[4];       // current estimate
    float trial[4];    // this iteration's step
    float c, ct;       // cost at th and at trial
    float xx;          // |x|^2
    float S, Sx[2];    // s^H s and s^H x at th (the proton density)
    float hS, hSx[2];  // the same at trial
    double lam;        // Levenberg-Marquardt damping
    GnSums g;          // tangent-pass sums at th
};

static_assert(offsetof(GnState, trial) == kGnRecTrial * sizeof(float) &&
                  offsetof(GnState, ct) == kGnRecCt * sizeof(float) &&
                  offsetof(GnState, xx) == kGnRecXx * sizeof(float) &&
                  offsetof(GnState, hS) == kGnRecHs * sizeof(float) &&
                  offsetof(GnState, hSx) == kGnRecHsx * sizeof(float) &&
                  sizeof(GnState) % sizeof(float) == 0,
              "Gauss-Newton record layout");

// group's voxel: one group of L lanes per voxel (or per list entry), one
// wave per work-group
template <int L>
__global__ __launch_bounds__(64, 2) void gn_init_kernel(GnArgs A)
{
    __shared__ AdjCoef coef_lds[64];
    __shared__ float2 rd_lds[64 / L][L][kJd];
    const int lane = threadIdx.x;
    const size_t v = (size_t)blockIdx.x * (64 / L) + lane / L;
    if (v >= A.V) return;   // whole groups: their lanes only meet their own group
    const AdjLane ln = adj_lane<L>(lane, A.w);
    const int k = ln.j;
    const float2 *x = A.sig + v * A.T;
    float xx = 0.f;
    for (int t = k; t < A.T; t += L) xx += re_dot(x[t], x[t]);
    xx = adj_gsum<L>(xx);
    float th[4];
#pragma unroll
    for (int d = 0; d < 4; d++) th[d] = A.theta[v * 4 + d];
    // a start outside the domain stays where it is: the pass runs at a valid
    // point (the group keeps one control flow) and its sums are dropped, so
    // the solve finds a singular system and every trial is rejected
    const bool ok = theta_ok(th);
    float tp[4] = {1000.f, 100.f, 0.f, 1.f};
    if (ok)
#pragma unroll
        for (int d = 0; d < 4; d++) tp[d] = th[d];
    GnSums g = gnl_pass<L>(A.seq, A.T, make_atom(tp), x, ln, coef_lds + (lane & ~(L - 1)),
                           rd_lds[lane / L]);
    if (k == 0) {
        GnState &st = A.st[v];
        float c = xx - (g.Sx.x * g.Sx.x + g.Sx.y * g.Sx.y) / g.S;
        if (!ok) {
            const float2 z = make_float2(0.f, 0.f);
            g.S = 1.f;
            g.Sx = z;
            for (int d = 0; d < 4; d++) g.U[d] = g.W[d] = z;
            for (int i = 0; i < 10; i++) g.D[i] = z;
            c = NAN;
        }
#pragma unroll
        for (int d = 0; d < 4; d++) st.th[d] = th[d];
        st.c = c;
        st.xx = xx;
        st.S = ok ? g.S : 0.f;
        st.Sx[0] = g.Sx.x;
        st.Sx[1] = g.Sx.y;
        st.lam = 1e-3;
        st.g = g;
        if (A.cost) A.cost[v * (A.iters + 1)] = c;
    }
}

// the damped normal equations in FP64 (one thread per voxel); thread 0 also
// clears the accepted-voxel count of this iteration
__global__ void gn_solve_kernel(GnArgs A)
{
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v == 0) *A.count = 0;
    if (v >= A.V) return;
    GnState &st = A.st[v];
    const GnSums &g = st.g;
    const double S = g.S;
    const double rr = g.Sx.x / S, ri = g.Sx.y / S;     // rho = Sx / S
    const double r2 = rr * rr + ri * ri;
    double Am[16], b[4];
    int i = 0;
    for (int d = 0; d < 4; d++)
        for (int e = d; e < 4; e++, i++) {
            // Re[D_de - conj(U_d) U_e / S]
            const double cu = (double)g.U[d].x * g.U[e].x + (double)g.U[d].y * g.U[e].y;
            Am[4 * d + e] = Am[4 * e + d] = r2 * ((double)g.D[i].x - cu / S);
        }
    for (int d = 0; d < 4; d++) {
        // g_d = -Re[conj(rho) (W_d - rho conj(U_d))]
        const double wr = g.W[d].x - (rr * g.U[d].x + ri * g.U[d].y);
        const double wi = g.W[d].y - (ri * g.U[d].x - rr * g.U[d].y);
        b[d] = rr * wr + ri * wi;   // = -g_d
    }
    double Md[16], Mi[16];
    for (int q = 0; q < 16; q++) Md[q] = Am[q];
    for (int d = 0; d < 4; d++) Md[5 * d] = Am[5 * d] * (1.0 + st.lam);
    double dl[4] = {0, 0, 0, 0};
    if (inv4(Md, Mi))
        for (int d = 0; d < 4; d++)
            for (int e = 0; e < 4; e++) dl[d] += Mi[4 * d + e] * b[e];
    for (int d = 0; d < 4; d++) st.trial[d] = (float)(st.th[d] + dl[d]);
}

// the cost at the trial step: a primal pass (no tangents, so this kernel
// keeps the register budget of the primal state)
template <int SUB>
__global__ __launch_bounds__(64 * kWaves) void gn_eval_kernel(GnArgs A)
{
    const size_t v = group_atom<SUB>();
    if (v >= A.V) return;
    __shared__ TrA coef_lds[kCoefLds];
    __shared__ float2 read_lds[kReadLds];
    const int k = threadIdx.x & (SUB - 1);
    GnState &st = A.st[v];
    float trial[4];
#pragma unroll
    for (int d = 0; d < 4; d++) trial[d] = st.trial[d];
    float ct = INFINITY, hS = 0.f, hx = 0.f, hy = 0.f;
    if (theta_ok(trial)) {
        const GnSums h = gn_pass<0, SUB>(A.seq, A.T, A.w, make_atom(trial), A.sig + v * A.T, k,
                                         coef_lds, read_lds);
        ct = st.xx - (h.Sx.x * h.Sx.x + h.Sx.y * h.Sx.y) / h.S;
        hS = h.S;
        hx = h.Sx.x;
        hy = h.Sx.y;
    }
    if (k == 0) {
        st.ct = ct;
        st.hS = hS;
        st.hSx[0] = hx;
        st.hSx[1] = hy;
    }
}

// keep the step if it lowers the cost (one thread per voxel); voxels that
// moved and have iterations left are listed for a new tangent pass
__global__ void gn_accept_kernel(GnArgs A, int it)
{
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= A.V) return;
    GnState &st = A.st[v];
    if (st.ct < st.c) {
        for (int d = 0; d < 4; d++) st.th[d] = st.trial[d];
        st.c = st.ct;
        st.S = st.hS;
        st.Sx[0] = st.hSx[0];
        st.Sx[1] = st.hSx[1];
        st.lam = fmax(st.lam * 0.1, 1e-9);
        if (it + 1 < A.iters) A.list[atomicAdd(A.count, 1)] = (int)v;
    } else {
        st.lam = fmin(st.lam * 10.0, 1e12);
    }
    if (A.cost) A.cost[v * (A.iters + 1) + it + 1] = st.c;
}

// tangent pass at the new estimate of the listed voxels
template <int L>
__global__ __launch_bounds__(64, 2) void gn_tangent_kernel(GnArgs A)
{
    __shared__ AdjCoef coef_lds[64];
    __shared__ float2 rd_lds[64 / L][L][kJd];
    const int lane = threadIdx.x;
    const size_t p = (size_t)blockIdx.x * (64 / L) + lane / L;
    if (p >= (size_t)*A.count) return;
    const AdjLane ln = adj_lane<L>(lane, A.w);
    const size_t v = (size_t)A.list[p];
    GnState &st = A.st[v];
    float th[4];
#pragma unroll
    for (int d = 0; d < 4; d++) th[d] = st.th[d];
    const GnSums g = gnl_pass<L>(A.seq, A.T, make_atom(th), A.sig + v * A.T, ln,
                                 coef_lds + (lane & ~(L - 1)), rd_lds[lane / L]);
    if (ln.first) st.g = g;
}

__global__ void gn_final_kernel(GnArgs A)
{
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= A.V) return;
    const GnState &st = A.st[v];
    for (int d = 0; d < 4; d++) A.theta[v * 4 + d] = st.th[d];
    if (A.pd) A.pd[v] = sqrtf(st.Sx[0] * st.Sx[0] + st.Sx[1] * st.Sx[1]) / sqrtf(st.S);
}

}  // namespace

namespace {

// work-groups for n atoms at 64 / SUB atoms per wave
dim3 groups(size_t n, int sub)
{
    const size_t per = (size_t)kWaves * (64 / sub);
    return dim3((unsigned)((n + per - 1) / per));
}

}  // namespace

size_t diff_scratch_floats(int nd, int T, int w)
{
    // ND = 0: checkpoints, then the atom's grad_seq partials (2T)
    if (nd == 0) return adj_ck_floats(T, w <= 32 ? 8 : 16) + (size_t)2 * T;
    const int nck = (T + kCkpt - 1) / kCkpt;
    return (size_t)(nck + kCkpt) * 6 * (1 + nd) * w;
}

int launch_jacobian(const DTr *seq, int T, int w, const float *theta, size_t n, float2 *sig,
                    float2 *jac, hipStream_t s)
{
    if (n == 0) return 0;
    if (w <= 32)
        jacl_kernel<8><<<(unsigned)((n + 7) / 8), 64, 0, s>>>(seq, T, w, theta, n, sig, jac);
    else
        jacl_kernel<16><<<(unsigned)((n + 3) / 4), 64, 0, s>>>(seq, T, w, theta, n, sig, jac);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : mrf_fail(MRF_E_GPU, "jacobian launch: %s", hipGetErrorString(e));
}

int launch_adjoint(const DTr *seq, int T, int w, const float *theta, size_t n, const float2 *cot,
                   float *grad_theta, float *grad_seq, float fa_unit, float *scratch, hipStream_t s)
{
    if (n == 0) return 0;
    RevArgs A{};
    A.seq = seq;
    A.T = T;
    A.w = w;
    const int L = w <= 32 ? 8 : 16;
    A.nck = (T + kAdjSeg - 1) / kAdjSeg;
    A.theta = theta;
    A.n = n;
    A.cot = cot;
    A.grad_theta = grad_theta;
    A.grad_seq = grad_seq;
    A.fa_unit = fa_unit;
    A.scratch = scratch;
    // the checkpoints of all n atoms, then their grad_seq partials
    A.seq_part = grad_seq ? scratch + n * adj_ck_floats(T, L) : nullptr;
    const size_t per_block = 64 / L;
    const unsigned nb = (unsigned)((n + per_block - 1) / per_block);
    if (L == 8) adj_kernel<8><<<nb, 64, 0, s>>>(A);
    else adj_kernel<16><<<nb, 64, 0, s>>>(A);
    if (grad_seq) {
        const unsigned gy = (unsigned)(n < 64 ? n : 64);
        seq_reduce_kernel<<<dim3((unsigned)((2 * T + 255) / 256), gy), 256, 0, s>>>(
            A.seq_part, n, 2 * T, grad_seq);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : mrf_fail(MRF_E_GPU, "adjoint launch: %s", hipGetErrorString(e));
}

int launch_crb(const DTr *seq, int T, int w, const float *theta, size_t n, const float *wts,
               double *loss, float *grad_seq, float fa_unit, float *scratch, hipStream_t s)
{
    if (n == 0) return 0;
    RevArgs A{};
    A.seq = seq;
    A.T = T;
    A.w = w;
    A.nck = (T + kCkpt - 1) / kCkpt;
    A.theta = theta;
    A.n = n;
    A.grad_seq = grad_seq;
    A.fa_unit = fa_unit;
    for (int d = 0; d < 4; d++) A.wts[d] = wts[d];
    A.loss = loss;
    A.scratch = scratch;
    const size_t lds = rev_lds_bytes(A);
    if (w <= 32) rev_kernel<4, 32><<<groups(n, 32), 64 * kWaves, lds, s>>>(A);
    else rev_kernel<4, 64><<<groups(n, 64), 64 * kWaves, lds, s>>>(A);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : mrf_fail(MRF_E_GPU, "crb launch: %s", hipGetErrorString(e));
}

size_t gn_scratch_bytes(size_t V)
{
    return (V * sizeof(GnState) + 255) / 256 * 256 + (V + 1) * sizeof(int);
}

// Each iteration: FP64 solve per voxel, primal pass at the trial, accept,
// tangent pass for the voxels that moved (compacted, so a rejected voxel
// never waits in a wave for an accepted one).  The decisions are those of a
// single-kernel loop; the passes get their own register budgets.
int launch_gn(const DTr *seq, const TrCoef *coef, int T, int w, const float2 *sig, size_t V,
              float *theta, int iters, float *pd, float *cost, void *scratch, hipStream_t s)
{
    if (V == 0) return 0;
    GnArgs A{};
    A.seq = seq;
    A.T = T;
    A.w = w;
    A.iters = iters;
    A.sig = sig;
    A.V = V;
    A.theta = theta;
    A.pd = pd;
    A.cost = cost;
    A.st = reinterpret_cast<GnState *>(scratch);
    A.list = reinterpret_cast<int *>(reinterpret_cast<char *>(scratch) +
                                     (V * sizeof(GnState) + 255) / 256 * 256);
    A.count = A.list + V;
    const int sub = w <= 32 ? 32 : 64;
    const dim3 gg = groups(V, sub), gb(64 * kWaves);
    const dim3 vb((unsigned)((V + 255) / 256));
    // the tangent passes: L = 8 (16) lanes per voxel, one wave per work-group
    const dim3 lg((unsigned)(w <= 32 ? (V + 7) / 8 : (V + 3) / 4));
    if (w <= 32) gn_init_kernel<8><<<lg, 64, 0, s>>>(A);
    else gn_init_kernel<16><<<lg, 64, 0, s>>>(A);
    for (int it = 0; it < iters; it++) {
        gn_solve_kernel<<<vb, 256, 0, s>>>(A);
        // the trial cost: one voxel per lane where nstates allows, else
        // one voxel per lane group
        if (launch_gn_trial(coef, T, w, sig, V, reinterpret_cast<float *>(A.st),
                            (int)(sizeof(GnState) / sizeof(float)), s)) {
            if (sub == 32) gn_eval_kernel<32><<<gg, gb, 0, s>>>(A);
            else gn_eval_kernel<64><<<gg, gb, 0, s>>>(A);
        }
        gn_accept_kernel<<<vb, 256, 0, s>>>(A, it);
        if (it + 1 < iters) {
            if (w <= 32) gn_tangent_kernel<8><<<lg, 64, 0, s>>>(A);
            else gn_tangent_kernel<16><<<lg, 64, 0, s>>>(A);
        }
    }
    gn_final_kernel<<<vb, 256, 0, s>>>(A);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : mrf_fail(MRF_E_GPU, "gn launch: %s", hipGetErrorString(e));
}

}  // namespace mrf
