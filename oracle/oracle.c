/*
 * oracle.c -- FP64 CPU restatement of snapMRF's EPG / ROA dictionary and
 * matching logic.  TEST INFRASTRUCTURE ONLY (see oracle.h): the reference
 * the HIP path is compared against, and the CPU baseline that bench.py times
 * beside it.
 *
 * The step order of every program follows the reference host loops one to
 * one (file:line cited per function); each step is applied to the whole
 * state in IEEE double.  Float32 inputs of the reference (sequence table,
 * parameters, degree->radian casts, the B0 phase angle handed to cexpf) are
 * reproduced as float32 values before they enter double arithmetic.
 */
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex zd;

int orc_echo_id(const char *e)
{
    /* functions.cu:1238-1311 */
    if (!e) return -1;
    if (!strcmp(e, "epg_ssfp")) return ORC_EPG_SSFP;
    if (!strcmp(e, "roa_ssfp")) return ORC_ROA_SSFP;
    if (!strcmp(e, "epg_sr")) return ORC_EPG_SR;
    if (!strcmp(e, "epg_se")) return ORC_EPG_SE;
    if (!strcmp(e, "epg_tse1") || !strcmp(e, "epg_tse2")) return ORC_EPG_TSE;
    if (!strcmp(e, "epg_fisp")) return ORC_EPG_FISP;
    return -1;
}

void orc_seq_prepare(int echo, const float *raw, size_t T, float *out)
{
    memcpy(out, raw, 4 * T * sizeof(float));
    if (echo != ORC_EPG_SSFP && echo != ORC_ROA_SSFP) return;
    /* functions.cu:1242-1261: factors are computed in double and passed to
     * float kernel arguments; the products are float32. */
    const float fa_scale = (float)(60.0f * M_PI / 180.0f);
    const float ph_scale = (float)(M_PI / 180.0f);
    for (size_t i = 0; i < T; i++) {
        out[i] = out[i] * fa_scale;
        out[T + i] = out[T + i] * ph_scale;
        out[2 * T + i] = out[2 * T + i] + 16.0f;
        out[3 * T + i] = out[3 * T + i] + 3.5f;
    }
}

/* ---------------------------------------------------------------- EPG ---- */

typedef struct {
    int w;
    zd *fp, *fm, *z;   /* F+_k, F-_k, Z_k, k = 0..w-1 */
} epg_state;

static void st_alloc(epg_state *s, int w)
{
    s->w = w;
    s->fp = (zd *)calloc((size_t)w, sizeof(zd));
    s->fm = (zd *)calloc((size_t)w, sizeof(zd));
    s->z = (zd *)calloc((size_t)w, sizeof(zd));
}

static void st_free(epg_state *s)
{
    free(s->fp); free(s->fm); free(s->z);
}

static void st_zero(epg_state *s)
{
    memset(s->fp, 0, (size_t)s->w * sizeof(zd));
    memset(s->fm, 0, (size_t)s->w * sizeof(zd));
    memset(s->z, 0, (size_t)s->w * sizeof(zd));
}

/* init_rf_pulse, functions.cu:456-481.  alpha/phi in degrees, cast to float
 * radians as the kernel does. */
static void epg_init(epg_state *s, float alpha_deg, float phi_deg, float TI,
                     double T1, double T2)
{
    float a = (float)((double)alpha_deg * M_PI / 180.0f);
    float p = (float)((double)phi_deg * M_PI / 180.0f);
    double e1 = exp(-(double)TI / T1), e2 = exp(-(double)TI / T2);
    double sa = sin((double)a), ca = cos((double)a);
    double sph = sin((double)p), cph = cos((double)p);
    zd tmp = e2 * sa * (sph + I * cph);
    s->fp[0] = conj(tmp);
    s->fm[0] = tmp;
    s->z[0] = e1 * ca + 1.0 - e1;
}

/* fill_transition_matrix + apply_rf_pulse, functions.cu:483-528.  The cuBLAS
 * call computes C = T_m^T * w, i.e. M[r][c] = T_m[3r+c]. */
static void epg_rf(epg_state *s, float alpha, float phi)
{
    double sa = sin((double)alpha), ca = cos((double)alpha);
    double sa2 = (1.0 - ca) / 2.0, ca2 = (1.0 + ca) / 2.0;
    zd zz = cexp(I * (double)phi), zi = conj(zz);
    zd m00 = ca2, m01 = zz * zz * sa2, m02 = -I * zz * sa;
    zd m10 = sa2 * zi * zi, m11 = ca2, m12 = I * sa * zi;
    zd m20 = -I * 0.5 * sa * zi, m21 = I * 0.5 * zz * sa, m22 = ca;
    for (int k = 0; k < s->w; k++) {
        zd a = s->fp[k], b = s->fm[k], c = s->z[k];
        s->fp[k] = m00 * a + m01 * b + m02 * c;
        s->fm[k] = m10 * a + m11 * b + m12 * c;
        s->z[k] = m20 * a + m21 * b + m22 * c;
    }
}

/* compute_exp + decay_signal_stage1/2, functions.cu:530-583. */
static void epg_relax(epg_state *s, float t, double T1, double T2)
{
    double e1 = exp(-(double)t / T1), e2 = exp(-(double)t / T2);
    for (int k = 0; k < s->w; k++) {
        s->fp[k] *= e2;
        s->fm[k] *= e2;
        s->z[k] *= e1;
    }
    s->z[0] += 1.0 - e1;
}

/* shift_phase, functions.cu:585-607: the angle 2*M_PI*B0*t/1000.f is
 * evaluated in double and handed to cexpf as a float. */
static void epg_phase(epg_state *s, float t, float B0)
{
    float th = (float)(2.0 * M_PI * (double)B0 * (double)t / 1000.0f);
    zd e = cexp(I * (double)th);
    for (int k = 0; k < s->w; k++) {
        s->fp[k] *= e;
        s->fm[k] *= conj(e);
    }
}

/* dephase_gradients_stage1/2, functions.cu:619-668. */
static void epg_dephase(epg_state *s)
{
    int w = s->w;
    for (int k = w - 1; k >= 1; k--) s->fp[k] = s->fp[k - 1];
    for (int k = 0; k <= w - 2; k++) s->fm[k] = s->fm[k + 1];
    s->fp[0] = conj(s->fm[0]);
    s->fm[w - 1] = 0.0;
}

int orc_epg(int echo, const float *seq, size_t T, const float *params,
            size_t A, size_t natoms, int w, double *out)
{
    if (w < 1 || T == 0 || natoms == 0) return -1;
    const float *FA = seq, *PH = seq + T, *TR = seq + 2 * T, *TE = seq + 3 * T;

    if (echo == ORC_EPG_SSFP) {
        if (A % natoms) return -1;
        size_t l_b1 = A / natoms;
        float *fa = (float *)malloc(T * sizeof(float));
        memcpy(fa, FA, T * sizeof(float));
        const float TI = 40.0f;   /* functions.cu:1242 */
        for (size_t j = 0; j < l_b1; j++) {
            /* functions.cu:970-973: b1 of the slice's first atom, FA *= b1 */
            float b1 = params[(j * natoms) * 4 + 3];
            for (size_t i = 0; i < T; i++) fa[i] = fa[i] * b1;
#pragma omp parallel
            {
                epg_state s;
                st_alloc(&s, w);
#pragma omp for schedule(dynamic, 16)
                for (long long id = 0; id < (long long)natoms; id++) {
                    /* kernels index slice-0 parameters (functions.cu:462-463) */
                    double T1 = params[id * 4], T2 = params[id * 4 + 1];
                    float B0 = params[id * 4 + 2];
                    st_zero(&s);
                    epg_init(&s, 180.0f, 0.0f, TI, T1, T2);
                    double *o = out + 2 * ((j * natoms + (size_t)id) * T);
                    for (size_t i = 0; i < T; i++) {
                        epg_rf(&s, fa[i], PH[i]);
                        epg_relax(&s, TE[i], T1, T2);
                        epg_phase(&s, TE[i], B0);
                        o[2 * i] = creal(s.fp[0]);
                        o[2 * i + 1] = cimag(s.fp[0]);
                        epg_phase(&s, TR[i] - TE[i], B0);
                        epg_dephase(&s);
                        epg_relax(&s, TR[i] - TE[i], T1, T2);
                    }
                }
                st_free(&s);
            }
            /* functions.cu:1012-1013: FA *= 1/b1 (float round trip) */
            float inv = 1.0f / b1;
            for (size_t i = 0; i < T; i++) fa[i] = fa[i] * inv;
        }
        free(fa);
        orc_normalize(out, T, A);     /* functions.cu:1268-1269 */
        return 0;
    }
    if (echo != ORC_EPG_SR && echo != ORC_EPG_SE && echo != ORC_EPG_TSE &&
        echo != ORC_EPG_FISP)
        return -1;

#pragma omp parallel
    {
        epg_state s;
        st_alloc(&s, w);
#pragma omp for schedule(dynamic, 16)
        for (long long id = 0; id < (long long)natoms; id++) {
            double T1 = params[id * 4], T2 = params[id * 4 + 1];
            double *o = out + 2 * ((size_t)id * T);
            st_zero(&s);
            if (echo == ORC_EPG_SR) {
                /* epg_sr, functions.cu:687-739 */
                epg_init(&s, 0.0f, 0.0f, 0.0f, T1, T2);
                for (size_t i = 0; i < T; i++) {
                    epg_rf(&s, FA[i], PH[i]);
                    epg_relax(&s, TE[i], T1, T2);
                    o[2 * i] = creal(s.fp[0]);
                    o[2 * i + 1] = cimag(s.fp[0]);
                    epg_relax(&s, TR[i] - TE[i], T1, T2);
                }
            } else if (echo == ORC_EPG_SE || echo == ORC_EPG_TSE) {
                /* epg_se functions.cu:741-807; epg_tse :809-876 (same steps) */
                epg_init(&s, 90.0f, 90.0f, 0.0f, T1, T2);
                epg_dephase(&s);
                epg_relax(&s, TE[0], T1, T2);
                for (size_t i = 0; i < T; i++) {
                    epg_rf(&s, FA[i], PH[i]);
                    epg_dephase(&s);
                    epg_relax(&s, TE[i], T1, T2);
                    o[2 * i] = creal(s.fp[0]);
                    o[2 * i + 1] = cimag(s.fp[0]);
                    epg_dephase(&s);
                    epg_relax(&s, TR[i] - TE[i], T1, T2);
                }
            } else {
                /* epg_fisp, functions.cu:878-933 */
                epg_init(&s, 0.0f, 0.0f, 0.0f, T1, T2);
                for (size_t i = 0; i < T; i++) {
                    epg_rf(&s, FA[i], PH[i]);
                    epg_relax(&s, TE[i], T1, T2);
                    o[2 * i] = creal(s.fp[0]);
                    o[2 * i + 1] = cimag(s.fp[0]);
                    epg_dephase(&s);
                    epg_relax(&s, TR[i] - TE[i], T1, T2);
                }
            }
        }
        st_free(&s);
    }
    return 0;
}

/* ---------------------------------------------------------------- ROA ---- */

/* roa_ssfp, functions.cu:1155-1222, with kernels :1022-1153. */
int orc_roa(const float *seq, size_t T, const float *params, size_t natoms,
            int n_iso, double iso_spread_hz, double *out)
{
    if (n_iso < 1 || T == 0 || natoms == 0) return -1;
    const float *FA = seq, *PH = seq + T, *TR = seq + 2 * T;
#pragma omp parallel for schedule(dynamic, 16)
    for (long long id = 0; id < (long long)natoms; id++) {
        double T1 = params[id * 4], T2 = params[id * 4 + 1];
        float B0a = params[id * 4 + 2];
        double *o = out + 2 * ((size_t)id * T);
        for (size_t i = 0; i < T; i++) { o[2 * i] = 0.0; o[2 * i + 1] = 0.0; }
        for (int q = 0; q < n_iso; q++) {
            /* isochromat offset (new semantics; q-th midpoint of the band) */
            double off = (n_iso == 1) ? 0.0
                : iso_spread_hz * (((double)q + 0.5) / (double)n_iso - 0.5);
            /* the isochromat frequency in double (a float frequency would
             * turn its rounding into a phase drift over the TRs) */
            double B0 = (double)B0a + off;
            /* init_rf_pulse(180, 0, TI=0), functions.cu:1022-1047 */
            float a = (float)(180.0 * M_PI / 180.0f);
            double sa = sin((double)a), ca = cos((double)a);
            double mx = 0.0 * sa;            /* e2 * sin(2*0) * sa */
            double my = 1.0 * 1.0 * sa;      /* e2 * cos(2*0) * sa, e2 = 1 */
            double mz = 1.0 * ca + 1.0 - 1.0;
            for (size_t i = 0; i < T; i++) {
                /* fill_transition_matrix (:1049-1075): column-major A used
                 * by Sgemm(N,N) => A[r][c] = T_m[c*3+r] */
                double s_a = sin((double)FA[i]), c_a = cos((double)FA[i]);
                double sp = sin((double)PH[i]), cp = cos((double)PH[i]);
                double s2 = 2.0 * cp * sp, c2 = cp * cp - sp * sp;
                double nx = c2 * mx + s2 * c_a * my + s2 * s_a * mz;
                double ny = -s2 * mx + c2 * c_a * my + c2 * s_a * mz;
                double nz = -s_a * my + c_a * mz;
                /* compute_exp(type 4 = TR) + decay_signal (:1094-1105) */
                double e1 = exp(-(double)TR[i] / T1);
                double e2 = exp(-(double)TR[i] / T2);
                mx = e2 * nx; my = e2 * ny; mz = e1 * nz + (1.0 - e1);
                /* fill_offresoance_operator (:1107-1127): beta is a float */
                float beta = (float)(M_PI * B0 * (double)TR[i] / 1000.0f);
                double cb = cos((double)beta), sb = sin((double)beta);
                double x = cb * mx + sb * my, y = cb * my - sb * mx;
                mx = x; my = y;
                o[2 * i] += mx;
                o[2 * i + 1] += my;
                x = cb * mx + sb * my; y = cb * my - sb * mx;
                mx = x; my = y;
            }
        }
        if (n_iso > 1)
            for (size_t i = 0; i < 2 * T; i++) o[i] /= (double)n_iso;
    }
    orc_normalize(out, T, natoms);   /* functions.cu:1297-1298 */
    return 0;
}

void orc_normalize(double *atoms, size_t T, size_t n)
{
    /* normalize, functions.cu:670-685 (sum in double instead of float) */
#pragma omp parallel for schedule(static)
    for (long long a = 0; a < (long long)n; a++) {
        double *p = atoms + 2 * (size_t)a * T, s = 0.0;
        for (size_t i = 0; i < 2 * T; i++) s += p[i] * p[i];
        if (s > 0.0) {
            double inv = 1.0 / sqrt(s);
            for (size_t i = 0; i < 2 * T; i++) p[i] *= inv;
        }
    }
}

/* ----------------------------------------------------------- matching ---- */

void orc_match_slice(const float *atoms, size_t T, size_t na,
                     const float *sig, size_t V,
                     long long *best_idx, double *best_pd, double *second_pd)
{
    /* cublasCgemm(OP_C, OP_N) + generate_maps, functions.cu:1317-1349 and
     * :1464-1491: S = sum_t conj(a_t) s_t; strict '>' from (0, index 0). */
#pragma omp parallel for schedule(dynamic, 4)
    for (long long v = 0; v < (long long)V; v++) {
        const float *x = sig + 2 * (size_t)v * T;
        double best = 0.0, second = 0.0;
        long long bi = 0;
        for (size_t a = 0; a < na; a++) {
            const float *p = atoms + 2 * a * T;
            double re = 0.0, im = 0.0;
            for (size_t t = 0; t < T; t++) {
                double ar = p[2 * t], ai = p[2 * t + 1];
                double xr = x[2 * t], xi = x[2 * t + 1];
                re += ar * xr + ai * xi;
                im += ar * xi - ai * xr;
            }
            double sc = sqrt(re * re + im * im);
            if (sc > best) {
                second = best;
                best = sc;
                bi = (long long)a;
            } else if (sc > second) {
                second = sc;
            }
        }
        best_idx[v] = bi;
        best_pd[v] = best;
        if (second_pd) second_pd[v] = second;
    }
}

int orc_b1_slice(float b1map_v, const float *slice_b1, int l_b1)
{
    /* merge_maps(b1map), functions.cu:1351-1384 */
    int match = 0;
    float best = 10.0f;
    for (int i = 0; i < l_b1; i++) {
        float d = fabsf(b1map_v - slice_b1[i]);
        if (d < best) { best = d; match = i; }
    }
    return match;
}

double orc_wall(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
