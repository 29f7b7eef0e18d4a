/*
 * snapmrf.h -- C ABI of the MI355X (gfx950) snapMRF dictionary-generation and
 * template-matching path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns 0
 * on success and a negative MRF_E_* code on failure (the message is kept in a
 * thread-local buffer read by mrf_last_error()).  Nothing here exits,
 * aborts or blocks on stdin, unlike the reference's gpuAssert/err paths
 * (src/functions.cu:37-74).  Device pointers are caller-owned.
 *
 * Reference interface each entry point replaces is cited as file:line of the
 * snapMRF source tree (src/...).
 */
#ifndef SNAPMRF_H
#define SNAPMRF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* complex64, layout-identical to CUDA/HIP float2 and std::complex<float>
 * (src/float2math.h) */
typedef struct { float x, y; } mrf_complex;

/* error codes */
#define MRF_OK 0
#define MRF_E_ARG (-1)       /* invalid argument / size */
#define MRF_E_IO (-2)        /* file open/read/write failure */
#define MRF_E_FORMAT (-3)    /* malformed CSV / RA content */
#define MRF_E_ECHO (-4)      /* unknown echo type (src/functions.cu:1310-1311) */
#define MRF_E_GPU (-5)       /* HIP runtime error or no device */
#define MRF_E_NOMEM (-6)     /* allocation failure */
#define MRF_E_UNSUPPORTED (-7)

/* Thread-local message of the last failing call on this thread. */
const char *mrf_last_error(void);

/* ------------------------------------------------------------------------ */
/* Host helpers (drop-in counterparts of the reference's host functions)      */
/* ------------------------------------------------------------------------ */

/* Number of '\n' in the file (src/functions.cu:100-120 csv_dim); the
 * sequence length is csv_dim - 1.  Returns (size_t)-1 on open failure. */
size_t mrf_csv_dim(const char *filepath);

/* Reads nreps rows of the 4-column rfpulse CSV into h_mrf[4][nreps] (FA |
 * phase | TR | TE), reading 100-byte fgets chunks and skipping the first two
 * chunks exactly as src/functions.cu:122-158 csv_read does.  Rows missing
 * because of that skip are zero-filled and reported in mrf_last_error()
 * (return value 1).  Returns MRF_E_FORMAT for a row with fewer than 4 fields. */
int mrf_csv_read(float *h_mrf, size_t nreps, const char *filepath);

/* Grid strings "a", "a:b" (step 1), "a:s:b", joined by '+'
 * (src/functions.cu:196-353 parse_range/parse_length/parse_params), float32
 * arithmetic, including the reference's dedup and end bookkeeping.
 * mrf_parse_length returns 0 on a malformed string. */
size_t mrf_parse_length(const char *str);
int mrf_parse_params(float *h_p, const char *str, size_t len);

/* logspace of the -N mode (src/functions.cu:269-304).  For min <= 0 the
 * reference writes one element past len; h_p must hold len + 1 floats. */
void mrf_logspace(float *h_p, float min, float max, size_t len);

/* Atoms with T1 > T2, times l_b0 * l_b1 (src/functions.cu:393-414). */
size_t mrf_compute_natoms(const float *h_t1, const float *h_t2, size_t l_t1,
                          size_t l_t2, size_t l_b0, size_t l_b1);

/* params[A][nparams] in the reference atom order: T1 fastest, then T2, B0,
 * B1 (src/functions.cu:355-391 trans_params). */
int mrf_trans_params(float *h_params, const float *h_t1, const float *h_t2,
                     const float *h_b0, const float *h_b1, size_t l_t1,
                     size_t l_t2, size_t l_b0, size_t l_b1, size_t NATOMS,
                     int nparams);

/* The -N hard-coded grid of src/mrf.cu:179-203 (l_b1=11, l_b0=21,
 * l_t1=l_t2=round(sqrt(N/231))).  Fills the lengths; if the four arrays are
 * non-NULL they must hold l_t1, l_t2, l_b0 + 1 and l_b1 floats. */
int mrf_grid_natoms_mode(size_t N, size_t *l_t1, size_t *l_t2, size_t *l_b0,
                         size_t *l_b1, float *h_t1, float *h_t2, float *h_b0,
                         float *h_b1);

/* RA files (src/ra.h:20-74, src/ra.cu:23-105).  Same bytes; headers are
 * validated and writes larger than 2 GiB are correct. */
typedef struct {
    uint64_t flags, eltype, elbyte, size, ndims;
    uint64_t *dims;
    uint8_t *data;
} mrf_ra_t;
int mrf_ra_read(mrf_ra_t *a, const char *path);
int mrf_ra_write(const mrf_ra_t *a, const char *path);
void mrf_ra_free(mrf_ra_t *a);

/* ------------------------------------------------------------------------ */
/* Dictionary generation                                                      */
/* ------------------------------------------------------------------------ */

/* Drop-in for MRF_dict (src/functions.cu:1224-1315, declared
 * src/functions.h:115-117).  d_atoms [NATOMS][nreps] complex64 (device, RA
 * dims (nreps, NATOMS)); d_params [NATOMS][nparams] (device); h_mrf host
 * [4][NREPS] csv columns of which the first nreps rows are used.
 * echo_type: epg_ssfp, roa_ssfp, epg_sr, epg_se, epg_tse1, epg_tse2,
 * epg_fisp.  Semantics follow the reference: csv unit transform and L2
 * normalisation for epg_ssfp/roa_ssfp only; roa_ssfp and the sr/se/tse/fisp
 * programs fill the first natoms atoms only.  Runs on the null stream and
 * returns after completion. */
int mrf_dict(mrf_complex *d_atoms, const float *d_params, const float *h_mrf,
             size_t nreps, size_t NREPS, size_t NATOMS, size_t natoms,
             int nparams, int nstates, const char *echo_type);

/* Options of the extended entry points.  Zero-initialise, then set. */
typedef struct {
    int normalize;        /* -1: reference default for the echo type, 0/1 force */
    int n_iso;            /* roa_ssfp isochromats per atom (<=1: reference) */
    float iso_spread_hz;  /* roa_ssfp off-resonance band of the isochromats:
                             isochromat q of n precesses at
                             B0 + spread*((q + 1/2)/n - 1/2) Hz; the signal
                             is their mean */
    int kernel_form;      /* 0: fused frame-form EPG kernel for epg_ssfp/epg_fisp;
                             1: the reference's unfused step order (cross-check) */
    int lanes_per_atom;   /* EPG lanes per atom, 1, 2, 4, 8 or 16; 0 = chosen from the
                             dictionary size.  The choice only changes FP32 rounding
                             (every form computes the same recursion); it is exposed so
                             that every lane form can be tested at any size */
    int reserved[6];
} mrf_dict_opts;

/* Extended dictionary generation on a caller stream (hipStream_t passed as
 * void*; NULL = null stream).  Asynchronous with respect to the host.
 * d_norm2 (optional, [NATOMS] double) receives each atom's squared L2 norm
 * before normalisation. */
int mrf_dict_ex(void *stream, const char *echo_type, const float *h_mrf,
                size_t nreps, size_t NREPS, const float *d_params,
                size_t NATOMS, size_t natoms, int nstates,
                const mrf_dict_opts *opts, mrf_complex *d_atoms,
                double *d_norm2);

/* ------------------------------------------------------------------------ */
/* Matching                                                                   */
/* ------------------------------------------------------------------------ */

/* Drop-in for MRF_match (src/functions.cu:1493-1584, declared
 * src/functions.h:132-134).  d_maps [nvoxels][nmaps] (device) receives
 * (T1, T2, B0, B1, PD) of the best atom per voxel; h_img host
 * [nvoxels][NREPS] complex64 of which nreps frames are matched (as in the
 * reference, the voxel stride is NREPS); d_atoms/d_params as produced by
 * mrf_dict.  in_b1map: path of a float32 RA B1 map with nvoxels entries, or
 * NULL for the max-PD merge.  Synchronous. */
int mrf_match(float *d_maps, const mrf_complex *h_img,
              const mrf_complex *d_atoms, const float *d_params, size_t NREPS,
              size_t nreps, size_t NATOMS, size_t natoms, size_t nvoxels,
              int nparams, int nmaps, const char *in_b1map);

/* Matching precisions of mrf_match_ex */
#define MRF_MATCH_FP32 0        /* complex64 dictionary, FP32 MFMA */
#define MRF_MATCH_BF16_TOP8 1   /* BF16 MFMA + top-8 + rescore of the candidates (FP64 sums) */
#define MRF_MATCH_SPLIT3 2      /* split-BF16 3 products (hi*hi + lo*hi + hi*lo) + top-8 + rescore */
#define MRF_MATCH_BF16_TOP16 3  /* BF16 MFMA + top-16 + rescore */
#define MRF_MATCH_SPLIT2 4      /* dictionary split (a_hi + a_lo) * s_hi + top-8 + rescore */
#define MRF_MATCH_BF16_TOP32 5  /* BF16 MFMA + TOP16's candidates over twice its atom chunks
                                   + top-32 + rescore */
#define MRF_MATCH_BF16_TOP64 6  /* BF16 MFMA + candidates over 32 atom chunks + top-64 +
                                   rescore: for dense grids (B0 steps of 2 Hz), where the FP32
                                   argmax can rank past 32 in BF16 */
/* Flag OR-ed into a low-precision mode: the candidates are rescored with FP32
 * sums instead of FP64 sums (FP64 makes the pick the exact argmax among the
 * candidates; FP32 keeps the rescore in the matcher's own precision). */
#define MRF_MATCH_RESCORE_FP32 0x100

/* Extended matching, all device-resident, on a caller stream.
 *   d_atoms   [NATOMS][T] complex64 (used by FP32 and for the rescore)
 *   d_lowp    [NATOMS][mrf_lowp_ld(T)] bf16x2 hi copy (all low-precision
 *             modes; from mrf_lowp_prepare), d_lowp_lo the lo residual
 *             (SPLIT2, SPLIT3)
 *   d_sig     [V][T] complex64 signals
 *   d_b1map   [V] float B1 map or NULL (max-PD merge over all slices)
 *   d_maps    [V][5] (T1, T2, B0, B1, PD); d_idx optional [V] int32 global
 *             atom index of the match.
 * B1-map semantics: each voxel is matched only against the slice k whose B1
 * minimises |b1map - B1_k| (strict <, first k on ties), which equals the
 * reference's per-slice argmax followed by merge_maps(b1map)
 * (src/functions.cu:1351-1384) because every atom of slice k carries B1_k. */
int mrf_match_ex(void *stream, int mode, const mrf_complex *d_atoms,
                 const void *d_lowp, const void *d_lowp_lo,
                 const float *d_params, size_t T, size_t NATOMS,
                 size_t natoms, const mrf_complex *d_sig, size_t V,
                 const float *d_b1map, float *d_maps, int32_t *d_idx);

/* FP32 matching (mode MRF_MATCH_FP32 of mrf_match_ex: same arguments, same
 * kernels, the same d_maps / d_idx) that allocates no device memory.  No
 * counterpart in the reference (MRF_match allocates its buffers per call,
 * src/functions.cu:1493-1584).  For callers that cannot allocate: code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc.  Its
 * scratch (voxel slices, per-voxel keys, voxel permutation, tile list) is a
 * static 20 MiB device array, one per GPU, held until the process ends:
 * calls with more voxels than fit (2^20 voxels over up to 1,024 B1 slices
 * always fit) fail with MRF_E_ARG before any launch.  Every call on a GPU
 * uses that GPU's one array, so calls that overlap on one GPU (several
 * streams or host threads) race: serialise them.  The scratch is cleared by
 * kernels on the stream at each call, so consecutive calls are independent. */
int mrf_match_ex_noalloc(void *stream, const mrf_complex *d_atoms,
                         const float *d_params, size_t T, size_t NATOMS,
                         size_t natoms, const mrf_complex *d_sig, size_t V,
                         const float *d_b1map, float *d_maps, int32_t *d_idx);

/* Dictionary generation streamed through the matcher: the atoms exist one
 * block at a time (block_atoms per B1 slice, every slice; 0 = blocks of at
 * most 1 GiB), each block is generated as mrf_dict_ex would generate it and
 * matched at once, and the running per-voxel best merges across blocks.
 * mode MRF_MATCH_FP32: d_maps / d_idx equal mrf_dict_ex followed by
 * mrf_match_ex (MRF_MATCH_FP32) on the same inputs, without the NATOMS x T
 * dictionary in memory.  Low-precision modes round block_atoms up to a
 * multiple of 128; each block is converted to BF16 and its candidates are
 * rescored against its complex64 rows.  Every candidate the whole-dictionary
 * call rescores is among the candidates of its block, so the pick scores at
 * least as high in FP64.  No counterpart in the reference, which
 * always materialises the dictionary (src/mrf.cu:286-321).  epg_ssfp with
 * any number of B1 slices; the other echo types need NATOMS == natoms (they
 * fill one slice). */
int mrf_dict_match_ex(void *stream, const char *echo_type, const float *h_mrf,
                      size_t nreps, size_t NREPS, const float *d_params,
                      size_t NATOMS, size_t natoms, int nstates,
                      const mrf_dict_opts *opts, int mode, const mrf_complex *d_sig,
                      size_t V, const float *d_b1map, size_t block_atoms,
                      float *d_maps, int32_t *d_idx);

/* Rank-K (SVD) compressed matching, the dictionary compression of McGivney
 * et al. (IEEE TMI 2014) as matlab/MRF_dict_B1.m:125-159 builds it; the
 * reference's CUDA path has no counterpart.
 *
 * mrf_svd_basis: left singular vectors of the (normalised) atoms with the
 * largest singular values, per B1 slice (per_slice = 1; NATOMS/natoms bases)
 * or for the whole dictionary (0), from the FP64 Gram matrix and FP64
 * eigensolver.  Rank: with energy_frac > 0, the number of singular values
 * whose cumulative energy fraction is <= energy_frac on the first basis (the
 * MATLAB rule), at least 1 and at most K_max; else K_max.  d_basis receives
 * [bases][K][T] complex64 (allocate bases * K_max * T), h_sv (optional,
 * host) [bases][K_max] singular values, *K_out the rank K.  Synchronous.
 *
 * mrf_svd_compress: d_catoms [NATOMS][K] = U^H * atoms, each slice with its
 * basis.
 *
 * mrf_svd_match_ex: compresses each signal with its B1 slice's basis (the
 * slice the B1 map selects, as mrf_match_ex) and runs the FP32 matcher over
 * the K compressed values; d_maps / d_idx as mrf_match_ex.  Per-slice bases
 * need d_b1map. */
int mrf_svd_basis(void *stream, const mrf_complex *d_atoms, size_t T, size_t NATOMS,
                  size_t natoms, int per_slice, double energy_frac, size_t K_max,
                  mrf_complex *d_basis, double *h_sv, size_t *K_out);
int mrf_svd_compress(void *stream, const mrf_complex *d_basis, size_t K, size_t T,
                     int per_slice, const mrf_complex *d_atoms, size_t NATOMS, size_t natoms,
                     mrf_complex *d_catoms);
int mrf_svd_match_ex(void *stream, const mrf_complex *d_basis, size_t K, size_t T,
                     int per_slice, const mrf_complex *d_catoms, const float *d_params,
                     size_t NATOMS, size_t natoms, const mrf_complex *d_sig, size_t V,
                     const float *d_b1map, float *d_maps, int32_t *d_idx);

/* Builds the BF16 copies of a complex64 dictionary: d_hi = bf16(a),
 * d_lo = bf16(a - hi) (optional), round to nearest even.  Each is
 * [NATOMS][mrf_lowp_ld(T)] words of two bf16 (re in the low half, im in the
 * high half), the pad past T zero.  No reference counterpart (the reference
 * matches in FP32 only, src/functions.cu:1386-1426). */
int mrf_lowp_prepare(void *stream, const mrf_complex *d_atoms, size_t T,
                     size_t NATOMS, void *d_hi, void *d_lo);

/* Row stride in 32-bit words of the mrf_lowp_prepare copies: T rounded up to
 * a multiple of 32. */
size_t mrf_lowp_ld(size_t T);

/* ------------------------------------------------------------------------ */
/* Differentiation form                                                       */
/* ------------------------------------------------------------------------ */
/* No reference counterpart: the reference generates forward atoms only
 * (src/functions.cu:936-1019).  These entry points differentiate the
 * epg_ssfp program as a continuous function: per atom theta = (T1 ms, T2 ms,
 * B0 Hz, B1) [n][4] (device), flip angles alpha_n * B1 with the CSV unit
 * transform of src/functions.cu:1242-1261, unnormalised readouts s_n = F+_0.
 * h_mrf is the host [4][NREPS] CSV table (FA | phase | TR | TE) of which
 * nreps rows are used; sequence gradients are with respect to the raw FA and
 * TR columns (TE held fixed).  2 <= nstates <= 64.  Asynchronous on the
 * caller stream (NULL = null stream). */

/* d_sig [n][nreps] (optional) and d_jac [n][4][nreps] = ds/dtheta (optional). */
int mrf_epg_jacobian(void *stream, const float *h_mrf, size_t nreps, size_t NREPS,
                     const float *d_theta, size_t n, int nstates, mrf_complex *d_sig,
                     mrf_complex *d_jac);

/* Reverse mode for L = sum_atoms Re sum_n conj(cot_n) s_n with d_cot
 * [n][nreps]: d_grad_theta [n][4] = dL/dtheta (optional) and d_grad_seq
 * [2][nreps] += (dL/dFA_raw, dL/dTR_raw) summed over atoms (optional,
 * accumulated: zero it first).  States are checkpointed every 4 TRs and
 * recomputed per segment.  Scratch: 4 * (96 L ceil(nreps / 4) + 2 nreps)
 * bytes per atom, L = 8 for nstates <= 32 and 16 above (about 200 KB at
 * nreps = 1000, nstates = 30), allocated per call on the stream; calls
 * needing more than 16 GiB (or MRF_DIFF_SCRATCH_BYTES when set in the
 * environment) run the atoms in equal batches. */
int mrf_epg_adjoint(void *stream, const float *h_mrf, size_t nreps, size_t NREPS,
                    const float *d_theta, size_t n, int nstates, const mrf_complex *d_cot,
                    float *d_grad_theta, float *d_grad_seq);

/* Per-voxel Gauss-Newton (Levenberg-Marquardt damped, variable projection of
 * the complex proton density) refinement of d_theta [V][4] (in: start, e.g.
 * the matched grid point; out: refined) against d_sig [V][nreps], iters
 * iterations; the cost |x|^2 - |s^H x|^2/|s|^2 never increases.  d_pd [V]
 * (optional) = |s^H x|/|s|; d_cost [V][iters + 1] (optional) cost history.
 * The 4x4 normal equations are solved in FP64.  A start outside the domain
 * (T1, T2, B1 > 0, all four finite) is returned unchanged with a NaN cost and
 * proton density.  Allocates a 240-byte state record per voxel on the stream. */
int mrf_gn_refine(void *stream, const float *h_mrf, size_t nreps, size_t NREPS,
                  const mrf_complex *d_sig, size_t V, int nstates, int iters, float *d_theta,
                  float *d_pd, float *d_cost);

/* Cramer-Rao-bound sequence-design loss over n tissues d_tissues [n][4]:
 * d_loss[i] = sum_d wts[d] [F_i^-1]_dd / scale_d^2, F_i = Re(J_i^H J_i) (unit
 * noise and proton density), scale = (T1, T2, 1, 1), inverted in FP64 (NaN
 * if singular); d_grad_seq [2][nreps] += d(sum_i loss_i)/d(FA_raw, TR_raw)
 * (accumulated: zero it first).  wts is host [4]. */
int mrf_crb_loss_grad(void *stream, const float *h_mrf, size_t nreps, size_t NREPS,
                      const float *d_tissues, size_t n, int nstates, const float *wts,
                      double *d_loss, float *d_grad_seq);

/* Device count visible to the library (0 when no GPU). */
int mrf_device_count(void);

#ifdef __cplusplus
}
#endif
#endif
