/*
 * diff.c -- FP64 CPU restatement of the differentiable EPG-SSFP model:
 * signal, forward-mode Jacobian with respect to (T1, T2, B0, B1), reverse-mode
 * adjoint with respect to (T1, T2, B0, B1) and the per-timepoint flip angles
 * and repetition times, and the Cramer-Rao-bound design loss.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The model is the reference's epg_ssfp program (functions.cu:936-1019 with
 * init_rf_pulse :456-481, apply_rf_pulse :483-528, decay_signal :530-583,
 * shift_phase :585-607, dephase_gradients :619-668) written as a continuous
 * function of its inputs: every quantity is double, the CSV unit transform
 * (functions.cu:1242-1261) is applied in double, and each atom's flip angles
 * are alpha_n * B1 (the reference's per-slice FA *= b1).  The inversion
 * pulse is the reference's float32 pi (init_rf_pulse casts 180 deg to float
 * radians).
 *
 * Per TR n (state y = F+_k, F-_k, Z_k, k < w):
 *   RF(a = alpha_n B1, phi_n) -> relax(TE_n) -> phase(TE_n) -> s_n = F+_0
 *   -> phase(TR_n - TE_n) -> dephase -> relax(TR_n - TE_n)
 * relax(t): F *= E2, Z = E1 Z + (1 - E1) [k = 0]; E = exp(-t / T)
 * phase(t): F+ *= P, F- *= conj(P), P = exp(i 2 pi B0 t / 1000)
 * dephase: F+_k <- F+_{k-1}, F-_k <- F-_{k+1}, F+_0 <- conj(F-_0), F-_{w-1} <- 0
 */
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex zd;

enum { D_T1 = 0, D_T2 = 1, D_B0 = 2, D_B1 = 3 };

#define FA_UNIT (60.0 * M_PI / 180.0)
#define PH_UNIT (M_PI / 180.0)
#define TI_MS 40.0
#define C_B0 (2.0 * M_PI / 1000.0)

typedef struct {
    double a0, sa0, ca0;   /* inversion angle (float32 pi) */
    double *al, *ph, *tr, *te;
} dseq;

static void seq_load(const float *raw, size_t T, dseq *q)
{
    q->al = (double *)malloc(4 * T * sizeof(double));
    q->ph = q->al + T;
    q->tr = q->al + 2 * T;
    q->te = q->al + 3 * T;
    for (size_t n = 0; n < T; n++) {
        q->al[n] = (double)raw[n] * FA_UNIT;
        q->ph[n] = (double)raw[T + n] * PH_UNIT;
        q->tr[n] = (double)raw[2 * T + n] + 16.0;
        q->te[n] = (double)raw[3 * T + n] + 3.5;
    }
    q->a0 = (double)(float)(180.0 * M_PI / 180.0f);
    q->sa0 = sin(q->a0);
    q->ca0 = cos(q->a0);
}

/* RF matrix M(a, phi) (same entries as oracle.c epg_rf) and its a-derivative
 * of order `ord` (0, 1, 2). */
static void rf_mat(double a, double phi, int ord, zd M[3][3])
{
    double C, S;
    switch (ord) {
    case 0: C = cos(a); S = sin(a); break;
    case 1: C = -sin(a); S = cos(a); break;
    default: C = -cos(a); S = -sin(a); break;
    }
    const double d0 = ord == 0 ? 1.0 : 0.0;
    const double ca2 = (d0 + C) / 2.0, sa2 = (d0 - C) / 2.0;
    const zd zz = cexp(I * phi), zi = conj(zz);
    M[0][0] = ca2;            M[0][1] = zz * zz * sa2;   M[0][2] = -I * zz * S;
    M[1][0] = sa2 * zi * zi;  M[1][1] = ca2;             M[1][2] = I * S * zi;
    M[2][0] = -I * 0.5 * S * zi; M[2][1] = I * 0.5 * zz * S; M[2][2] = C;
}

/* state: y[c][k], c = 0 (F+), 1 (F-), 2 (Z) */
typedef struct {
    int w;
    zd *y[3];
} st3;

static void st3_alloc(st3 *s, int w)
{
    s->w = w;
    for (int c = 0; c < 3; c++) s->y[c] = (zd *)calloc((size_t)w, sizeof(zd));
}

static void st3_free(st3 *s)
{
    for (int c = 0; c < 3; c++) free(s->y[c]);
}

static void st3_copy(st3 *d, const st3 *s)
{
    for (int c = 0; c < 3; c++) memcpy(d->y[c], s->y[c], (size_t)s->w * sizeof(zd));
}

static void st3_zero(st3 *s)
{
    for (int c = 0; c < 3; c++) memset(s->y[c], 0, (size_t)s->w * sizeof(zd));
}

static void op_rf(st3 *s, zd M[3][3])
{
    for (int k = 0; k < s->w; k++) {
        zd a = s->y[0][k], b = s->y[1][k], c = s->y[2][k];
        for (int r = 0; r < 3; r++) s->y[r][k] = M[r][0] * a + M[r][1] * b + M[r][2] * c;
    }
}

/* out = M in (no recovery term), for tangent source terms */
static void op_rf_apply(st3 *out, zd M[3][3], const st3 *in, int accumulate)
{
    for (int k = 0; k < in->w; k++) {
        zd a = in->y[0][k], b = in->y[1][k], c = in->y[2][k];
        for (int r = 0; r < 3; r++) {
            zd v = M[r][0] * a + M[r][1] * b + M[r][2] * c;
            out->y[r][k] = accumulate ? out->y[r][k] + v : v;
        }
    }
}

static void op_relax(st3 *s, double e1, double e2)
{
    for (int k = 0; k < s->w; k++) {
        s->y[0][k] *= e2;
        s->y[1][k] *= e2;
        s->y[2][k] *= e1;
    }
    s->y[2][0] += 1.0 - e1;
}

/* relax without the recovery term (tangents and adjoints) */
static void op_scale(st3 *s, double e1, double e2)
{
    for (int k = 0; k < s->w; k++) {
        s->y[0][k] *= e2;
        s->y[1][k] *= e2;
        s->y[2][k] *= e1;
    }
}

static void op_phase(st3 *s, zd p)
{
    for (int k = 0; k < s->w; k++) {
        s->y[0][k] *= p;
        s->y[1][k] *= conj(p);
    }
}

static void op_dephase(st3 *s)
{
    const int w = s->w;
    for (int k = w - 1; k >= 1; k--) s->y[0][k] = s->y[0][k - 1];
    for (int k = 0; k <= w - 2; k++) s->y[1][k] = s->y[1][k + 1];
    s->y[0][0] = conj(s->y[1][0]);
    s->y[1][w - 1] = 0.0;
}

static void op_init(st3 *s, const dseq *q, double T1, double T2)
{
    st3_zero(s);
    const double e1 = exp(-TI_MS / T1), e2 = exp(-TI_MS / T2);
    const zd tmp = e2 * q->sa0 * I;   /* RF phase 0: sin 0 + i cos 0 */
    s->y[0][0] = conj(tmp);
    s->y[1][0] = tmp;
    s->y[2][0] = e1 * q->ca0 + 1.0 - e1;
}

/* ------------------------------------------------------------ forward ---- */

/* Signal and (J != NULL) forward-mode Jacobian: s [T] complex, J [4][T]
 * complex, interleaved doubles. */
void orc_dm_jacobian(const float *raw, size_t T, const double *theta, int w,
                     double *s_out, double *J_out)
{
    dseq q;
    seq_load(raw, T, &q);
    const double T1 = theta[0], T2 = theta[1], B0 = theta[2], B1 = theta[3];
    st3 y, d[4], tmp;
    st3_alloc(&y, w);
    st3_alloc(&tmp, w);
    for (int j = 0; j < 4; j++) st3_alloc(&d[j], w);
    op_init(&y, &q, T1, T2);
    {
        /* init tangents */
        const double e1 = exp(-TI_MS / T1), e2 = exp(-TI_MS / T2);
        const double e1d = e1 * TI_MS / (T1 * T1), e2d = e2 * TI_MS / (T2 * T2);
        d[D_T1].y[2][0] = e1d * (q.ca0 - 1.0);
        const zd t2 = e2d * q.sa0 * I;
        d[D_T2].y[0][0] = conj(t2);
        d[D_T2].y[1][0] = t2;
    }
    for (size_t n = 0; n < T; n++) {
        const double a = q.al[n] * B1;
        zd M[3][3], Ma[3][3];
        rf_mat(a, q.ph[n], 0, M);
        rf_mat(a, q.ph[n], 1, Ma);
        /* RF: d_B1 gets alpha M_a y (old y) */
        op_rf_apply(&tmp, Ma, &y, 0);
        for (int j = 0; j < 4; j++) op_rf(&d[j], M);
        for (int k = 0; k < w; k++)
            for (int c = 0; c < 3; c++) d[D_B1].y[c][k] += q.al[n] * tmp.y[c][k];
        op_rf(&y, M);
        /* two relax/phase intervals around the readout */
        for (int half = 0; half < 2; half++) {
            const double t = half == 0 ? q.te[n] : q.tr[n] - q.te[n];
            if (half == 0) {
                const double e1 = exp(-t / T1), e2 = exp(-t / T2);
                const double e1d = e1 * t / (T1 * T1), e2d = e2 * t / (T2 * T2);
                /* tangent of relax uses the pre-relax primal */
                for (int j = 0; j < 4; j++) op_scale(&d[j], e1, e2);
                for (int k = 0; k < w; k++) {
                    d[D_T2].y[0][k] += e2d * y.y[0][k];
                    d[D_T2].y[1][k] += e2d * y.y[1][k];
                    d[D_T1].y[2][k] += e1d * (y.y[2][k] - (k == 0 ? 1.0 : 0.0));
                }
                op_relax(&y, e1, e2);
                const zd p = cexp(I * C_B0 * B0 * t);
                for (int j = 0; j < 4; j++) op_phase(&d[j], p);
                for (int k = 0; k < w; k++) {
                    d[D_B0].y[0][k] += I * C_B0 * t * p * y.y[0][k];
                    d[D_B0].y[1][k] += -I * C_B0 * t * conj(p) * y.y[1][k];
                }
                op_phase(&y, p);
                s_out[2 * n] = creal(y.y[0][0]);
                s_out[2 * n + 1] = cimag(y.y[0][0]);
                if (J_out)
                    for (int j = 0; j < 4; j++) {
                        J_out[2 * (j * T + n)] = creal(d[j].y[0][0]);
                        J_out[2 * (j * T + n) + 1] = cimag(d[j].y[0][0]);
                    }
            } else {
                const zd p = cexp(I * C_B0 * B0 * t);
                for (int j = 0; j < 4; j++) op_phase(&d[j], p);
                for (int k = 0; k < w; k++) {
                    d[D_B0].y[0][k] += I * C_B0 * t * p * y.y[0][k];
                    d[D_B0].y[1][k] += -I * C_B0 * t * conj(p) * y.y[1][k];
                }
                op_phase(&y, p);
                op_dephase(&y);
                for (int j = 0; j < 4; j++) op_dephase(&d[j]);
                const double e1 = exp(-t / T1), e2 = exp(-t / T2);
                const double e1d = e1 * t / (T1 * T1), e2d = e2 * t / (T2 * T2);
                for (int j = 0; j < 4; j++) op_scale(&d[j], e1, e2);
                for (int k = 0; k < w; k++) {
                    d[D_T2].y[0][k] += e2d * y.y[0][k];
                    d[D_T2].y[1][k] += e2d * y.y[1][k];
                    d[D_T1].y[2][k] += e1d * (y.y[2][k] - (k == 0 ? 1.0 : 0.0));
                }
                op_relax(&y, e1, e2);
            }
        }
    }
    st3_free(&y);
    st3_free(&tmp);
    for (int j = 0; j < 4; j++) st3_free(&d[j]);
    free(q.al);
}

/* ------------------------------------------------------------ reverse ---- */

/* Reverse mode for L = Re sum_n conj(v_n) s_n: gradients with respect to
 * theta (g_theta[4]) and the raw CSV columns FA (g_fa[T]) and TR (g_tr[T]);
 * v is [T] complex interleaved.  Every TR-start state is kept. */
void orc_dm_adjoint(const float *raw, size_t T, const double *theta, int w,
                    const double *v, double *g_theta, double *g_fa, double *g_tr)
{
    dseq q;
    seq_load(raw, T, &q);
    const double T1 = theta[0], T2 = theta[1], B0 = theta[2], B1 = theta[3];
    st3 *ys = (st3 *)malloc((T + 1) * sizeof(st3));
    for (size_t n = 0; n <= T; n++) st3_alloc(&ys[n], w);
    op_init(&ys[0], &q, T1, T2);
    for (size_t n = 0; n < T; n++) {
        st3_copy(&ys[n + 1], &ys[n]);
        st3 *y = &ys[n + 1];
        zd M[3][3];
        rf_mat(q.al[n] * B1, q.ph[n], 0, M);
        op_rf(y, M);
        op_relax(y, exp(-q.te[n] / T1), exp(-q.te[n] / T2));
        op_phase(y, cexp(I * C_B0 * B0 * q.te[n]));
        const double t2 = q.tr[n] - q.te[n];
        op_phase(y, cexp(I * C_B0 * B0 * t2));
        op_dephase(y);
        op_relax(y, exp(-t2 / T1), exp(-t2 / T2));
    }
    double gT1 = 0, gT2 = 0, gB0 = 0, gB1 = 0;
    st3 l, ya, yb, yc, ye;
    st3_alloc(&l, w);
    st3_alloc(&ya, w);
    st3_alloc(&yb, w);
    st3_alloc(&yc, w);
    st3_alloc(&ye, w);
    for (size_t nn = T; nn-- > 0;) {
        const double te = q.te[nn], t2 = q.tr[nn] - q.te[nn];
        const double a = q.al[nn] * B1;
        zd M[3][3], Ma[3][3];
        rf_mat(a, q.ph[nn], 0, M);
        rf_mat(a, q.ph[nn], 1, Ma);
        /* recompute intermediates */
        st3_copy(&ya, &ys[nn]);
        op_rf(&ya, M);
        st3_copy(&yb, &ya);
        op_relax(&yb, exp(-te / T1), exp(-te / T2));
        const zd p1 = cexp(I * C_B0 * B0 * te), p2 = cexp(I * C_B0 * B0 * t2);
        st3_copy(&yc, &yb);
        op_phase(&yc, p1);
        st3_copy(&ye, &yc);
        op_phase(&ye, p2);
        op_dephase(&ye);
        /* relax(t2) */
        {
            const double e1 = exp(-t2 / T1), e2 = exp(-t2 / T2);
            double gt = 0.0;
            for (int k = 0; k < w; k++) {
                const zd zr = ye.y[2][k] - (k == 0 ? 1.0 : 0.0);
                gT1 += creal(conj(l.y[2][k]) * zr) * e1 * t2 / (T1 * T1);
                gT2 += creal(conj(l.y[0][k]) * ye.y[0][k] + conj(l.y[1][k]) * ye.y[1][k]) *
                       e2 * t2 / (T2 * T2);
                gt += creal(conj(l.y[2][k]) * zr) * (-e1 / T1) +
                      creal(conj(l.y[0][k]) * ye.y[0][k] + conj(l.y[1][k]) * ye.y[1][k]) * (-e2 / T2);
                l.y[0][k] *= e2;
                l.y[1][k] *= e2;
                l.y[2][k] *= e1;
            }
            g_tr[nn] = gt;
        }
        /* dephase^T */
        {
            const zd lp0 = l.y[0][0];
            for (int k = 0; k <= w - 2; k++) l.y[0][k] = l.y[0][k + 1];
            l.y[0][w - 1] = 0.0;
            for (int k = w - 1; k >= 1; k--) l.y[1][k] = l.y[1][k - 1];
            l.y[1][0] = 0.0;
            if (w > 1) l.y[1][1] += conj(lp0);
            /* w == 1: F+_0 <- conj(F-_0) with F-_0 <- 0 afterwards: no path */
        }
        /* phase(t2)^T, input yc */
        {
            double gt = 0.0, gb = 0.0;
            const double beta = C_B0 * B0;
            for (int k = 0; k < w; k++) {
                const zd fp = p2 * yc.y[0][k], fm = conj(p2) * yc.y[1][k];
                gt += creal(conj(l.y[0][k]) * (I * beta * fp)) + creal(conj(l.y[1][k]) * (-I * beta * fm));
                gb += creal(conj(l.y[0][k]) * (I * C_B0 * t2 * fp)) +
                      creal(conj(l.y[1][k]) * (-I * C_B0 * t2 * fm));
                l.y[0][k] *= conj(p2);
                l.y[1][k] *= p2;
            }
            g_tr[nn] += gt;
            gB0 += gb;
        }
        /* readout */
        l.y[0][0] += v[2 * nn] + I * v[2 * nn + 1];
        /* phase(te)^T, input yb */
        for (int k = 0; k < w; k++) {
            const zd fp = p1 * yb.y[0][k], fm = conj(p1) * yb.y[1][k];
            gB0 += creal(conj(l.y[0][k]) * (I * C_B0 * te * fp)) +
                   creal(conj(l.y[1][k]) * (-I * C_B0 * te * fm));
            l.y[0][k] *= conj(p1);
            l.y[1][k] *= p1;
        }
        /* relax(te)^T, input ya */
        {
            const double e1 = exp(-te / T1), e2 = exp(-te / T2);
            for (int k = 0; k < w; k++) {
                const zd zr = ya.y[2][k] - (k == 0 ? 1.0 : 0.0);
                gT1 += creal(conj(l.y[2][k]) * zr) * e1 * te / (T1 * T1);
                gT2 += creal(conj(l.y[0][k]) * ya.y[0][k] + conj(l.y[1][k]) * ya.y[1][k]) *
                       e2 * te / (T2 * T2);
                l.y[0][k] *= e2;
                l.y[1][k] *= e2;
                l.y[2][k] *= e1;
            }
        }
        /* RF^T, input ys[nn] */
        {
            double ga = 0.0;
            const st3 *y0 = &ys[nn];
            for (int k = 0; k < w; k++) {
                zd in[3] = {y0->y[0][k], y0->y[1][k], y0->y[2][k]};
                zd lo[3] = {l.y[0][k], l.y[1][k], l.y[2][k]};
                for (int r = 0; r < 3; r++) {
                    zd mv = Ma[r][0] * in[0] + Ma[r][1] * in[1] + Ma[r][2] * in[2];
                    ga += creal(conj(lo[r]) * mv);
                }
                for (int c = 0; c < 3; c++)
                    l.y[c][k] = conj(M[0][c]) * lo[0] + conj(M[1][c]) * lo[1] + conj(M[2][c]) * lo[2];
            }
            g_fa[nn] = ga * B1 * FA_UNIT;
            gB1 += ga * q.al[nn];
        }
    }
    /* init^T */
    {
        const double e1 = exp(-TI_MS / T1), e2 = exp(-TI_MS / T2);
        const double e1d = e1 * TI_MS / (T1 * T1), e2d = e2 * TI_MS / (T2 * T2);
        gT1 += creal(conj(l.y[2][0]) * (e1d * (q.ca0 - 1.0)));
        const zd t2 = e2d * q.sa0 * I;
        gT2 += creal(conj(l.y[0][0]) * conj(t2) + conj(l.y[1][0]) * t2);
    }
    g_theta[0] = gT1;
    g_theta[1] = gT2;
    g_theta[2] = gB0;
    g_theta[3] = gB1;
    st3_free(&l);
    st3_free(&ya);
    st3_free(&yb);
    st3_free(&yc);
    st3_free(&ye);
    for (size_t n = 0; n <= T; n++) st3_free(&ys[n]);
    free(ys);
    free(q.al);
}

/* ---------------------------------------------------------------- CRB ---- */

/* Inverse of a symmetric positive definite 4x4 by Gauss-Jordan with partial
 * pivoting; returns -1 if singular. */
int orc_inv4(const double *A, double *X)
{
    double m[4][8];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 8; c++) m[r][c] = c < 4 ? A[4 * r + c] : (c - 4 == r ? 1.0 : 0.0);
    for (int c = 0; c < 4; c++) {
        int p = c;
        for (int r = c + 1; r < 4; r++)
            if (fabs(m[r][c]) > fabs(m[p][c])) p = r;
        if (m[p][c] == 0.0) return -1;
        if (p != c)
            for (int j = 0; j < 8; j++) {
                double t = m[c][j];
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
    return 0;
}

/* CRB design loss: sum over tissues of sum_d wts[d] * [F^-1]_dd / scale_d^2,
 * F = Re(J^H J) (unit noise, unit proton density), scale_d = T1, T2 for
 * d = 0, 1 and 1 for B0, B1.  tissues [n][4]. */
int orc_dm_crb(const float *raw, size_t T, const double *tissues, size_t n, int w,
               const double *wts, double *loss)
{
    double *s = (double *)malloc(2 * T * sizeof(double));
    double *J = (double *)malloc(8 * T * sizeof(double));
    double L = 0.0;
    int rc = 0;
    for (size_t i = 0; i < n && !rc; i++) {
        const double *th = tissues + 4 * i;
        orc_dm_jacobian(raw, T, th, w, s, J);
        double F[16], Fi[16];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) {
                double acc = 0.0;
                for (size_t t = 0; t < T; t++)
                    acc += J[2 * (a * T + t)] * J[2 * (b * T + t)] +
                           J[2 * (a * T + t) + 1] * J[2 * (b * T + t) + 1];
                F[4 * a + b] = acc;
            }
        if (orc_inv4(F, Fi)) {
            rc = -1;
            break;
        }
        for (int d = 0; d < 4; d++) {
            const double sc = d < 2 ? th[d] : 1.0;
            L += wts[d] * Fi[5 * d] / (sc * sc);
        }
    }
    free(s);
    free(J);
    *loss = L;
    return rc;
}
