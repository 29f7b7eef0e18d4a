/*
 * oracle.h -- FP64 CPU restatement of snapMRF's dictionary-generation and
 * template-matching path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the FP64 reference the HIP
 * product path (snapmrf_amd/libsnapmrf.so) is compared against.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only to check or to time the
 * CPU baseline.  The product never links, loads or calls it.
 *
 * Every function states the reference file:line it restates (paths relative to
 * the snapMRF source tree).  Inputs that the reference holds in float32 (the
 * sequence table after its unit transform, T1/T2/B0/B1, the RF angle after the
 * degree->radian cast, the B0 phase angle passed to cexpf) are taken as the
 * same float32 values; all arithmetic after that is IEEE double.
 *
 * Parity pins (what each function is checked against, see tests/):
 *   orc_seq_prepare   tests/golden/atoms_epg_ssfp.ra via orc_epg (fixture)
 *   orc_epg  se/sr/tse/fisp   tests/golden/atoms_epg_{se,sr,tse1,tse2,fisp}.ra
 *                             + closed forms (e^-2TE/T2, e^-nTR/T2, CPMG 0.75/
 *                             0.9375/0.84375, sin(a)e^-TE/T2)
 *   orc_epg  ssfp             tests/golden/atoms_epg_ssfp.ra; invariants:
 *                             s(-B0) = -conj(s(B0)) for zero RF phase,
 *                             constant RF phase multiplies by e^{i phi},
 *                             w >= T truncation-free (w and w+5 agree),
 *                             B1 slice j == plain run with FA*B1_j
 *   orc_roa                   tests/golden/atoms_roa_ssfp.ra; B0 conjugate
 *                             invariant; n_iso=1 / zero spread == single
 *   orc_normalize             unit norm (closed form)
 *   orc_match_*               matching KAT (atoms_mrf.ra/params_mrf.ra with
 *                             glibc rand() picks); brute force; first-index ties
 */
#ifndef SNAPMRF_ORACLE_H
#define SNAPMRF_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* echo program ids (functions.cu:1238-1311 dispatch strings) */
enum { ORC_EPG_SSFP = 0, ORC_ROA_SSFP = 1, ORC_EPG_SR = 2, ORC_EPG_SE = 3,
       ORC_EPG_TSE = 4, ORC_EPG_FISP = 5 };

int orc_echo_id(const char *echo_type);

/* Unit transform of the CSV columns, float32-exact (functions.cu:1240-1261,
 * 1271-1291).  raw/out are [4][T] (FA | phase | TR | TE).  Only epg_ssfp and
 * roa_ssfp transform; other programs take raw radians/ms (functions.cu:1300-1309). */
void orc_seq_prepare(int echo, const float *raw, size_t T, float *out);

/* EPG dictionary (functions.cu:456-1019).  seq is the prepared [4][T] table.
 * params [A][4] = (T1, T2, B0, B1).  For SSFP all A = natoms*l_b1 atoms are
 * generated (B1 slice loop, functions.cu:958-1014); for SR/SE/TSE/FISP only the
 * first natoms.  out is [A][T] complex as interleaved doubles (re, im). */
int orc_epg(int echo, const float *seq, size_t T, const float *params,
            size_t A, size_t natoms, int w, double *out);

/* ROA single/multi-isochromat bSSFP (functions.cu:1022-1222).  n_iso
 * isochromats with off-resonance offsets spread uniformly over
 * [-spread/2, spread/2] Hz (midpoints) are averaged; n_iso = 1 is the
 * reference.  Only the first natoms atoms are generated (functions.cu:1293). */
int orc_roa(const float *seq, size_t T, const float *params, size_t natoms,
            int n_iso, double iso_spread_hz, double *out);

/* Per-atom L2 normalisation (functions.cu:670-685), FP64 sum. */
void orc_normalize(double *atoms, size_t T, size_t n);

/* Matching, FP64 scores (functions.cu:1317-1349 generate_maps semantics:
 * first index wins ties, PD = |<a,s>|).  atoms: [na][T] complex float32
 * interleaved; sig: [V][T] complex float32 interleaved.  Scores are computed
 * against atoms [a0, a0+na).  best_idx/best_pd: [V]; second_pd optional [V]
 * (runner-up score, for tie-margin analysis). */
void orc_match_slice(const float *atoms, size_t T, size_t na,
                     const float *sig, size_t V,
                     long long *best_idx, double *best_pd, double *second_pd);

/* ---- differentiable EPG-SSFP model (diff.c) ------------------------------
 * Continuous FP64 form of the epg_ssfp program (functions.cu:936-1019): raw
 * CSV columns [4][T] (float) with the unit transform applied in double, one
 * atom theta = (T1 ms, T2 ms, B0 Hz, B1), flip angles alpha_n * B1.  Signals
 * are the unnormalised F+_0 readouts, complex interleaved doubles. */

/* s [T] and (J != NULL) dS/dtheta [4][T] by forward-mode propagation. */
void orc_dm_jacobian(const float *raw, size_t T, const double *theta, int w,
                     double *s, double *J);

/* L = Re sum_n conj(v_n) s_n: dL/dtheta [4], dL/dFA_raw [T], dL/dTR_raw [T]
 * (TE held fixed) by reverse mode over stored TR-start states. */
void orc_dm_adjoint(const float *raw, size_t T, const double *theta, int w,
                    const double *v, double *g_theta, double *g_fa, double *g_tr);

/* Cramer-Rao-bound loss sum_i sum_d wts[d] [F_i^-1]_dd / scale_d^2 with
 * F_i = Re(J_i^H J_i), scale = (T1, T2, 1, 1); returns -1 for a singular F. */
int orc_dm_crb(const float *raw, size_t T, const double *tissues, size_t n, int w,
               const double *wts, double *loss);

/* 4x4 inverse, Gauss-Jordan with partial pivoting (-1 if singular). */
int orc_inv4(const double *A, double *X);

/* B1-map merge (functions.cu:1351-1384): slice k = argmin_k |b1map - B1_k|,
 * strict <, init 10 (float32 arithmetic as in the reference). */
int orc_b1_slice(float b1map_v, const float *slice_b1, int l_b1);

/* Synthetic / timing helpers */
double orc_wall(void);
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
