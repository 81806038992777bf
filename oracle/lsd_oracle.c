/*
 * lsd_oracle.c -- CPU FP64 oracle of the Leeds Spherical Dynamo per-timestep
 * path (Boussinesq, insulating inner core i_Nic = 1, no-slip / fixed
 * codensity / insulating boundaries, fixed dt).
 *
 * TEST INFRASTRUCTURE.  This file restates the reference algorithm in plain C
 * so that tests can check the HIP library against it; only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product library never links it.
 *
 * Each function cites the reference file:line it follows.  Floating-point
 * expressions keep the reference's association order where it is visible in
 * the Fortran source.  What pins each part is listed in tests/ (closed forms,
 * numpy/scipy/mpmath, the reference's own auxiliary routines built into
 * oracle/_ref, identities); nothing here is pinned by retyping it in a test.
 */
#include "lsd_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define KL 3
#define LDAB 10 /* 3*KL+1 */
#define MBW 7   /* 2*KL+1 */

static const double D_PI = 3.1415926535897931; /* parameters.F90_template:263 */

static char g_err[256];
static int g_threads = 1;

static void set_err(const char* m) { snprintf(g_err, sizeof g_err, "%s", m); }
const char* orc_last_error(void) { return g_err; }

void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int orc_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------ sizes */

/* Scripts/parameters.F90_template:250-264 (derived sizes) and :274-288
 * (par_precompute checks).  Th, Ph default to the 3/2 rule. */
int orc_sizes(const lsd_config* c, int* H, int* Th, int* Ph) {
    if (c->L <= (c->M - 1) * c->Mp) { set_err("set L > (M-1)*Mp"); return LSD_E_CONFIG; }
    if (c->M > 1 && (c->M % 2) == 1) { set_err("if(M>1) set M even"); return LSD_E_CONFIG; }
    if (c->N < 7 || c->N > 1024) { set_err("N must be in [7, 1024]"); return LSD_E_CONFIG; }
    int th = c->Th > 0 ? c->Th : c->L * 3 / 2;
    int ph = c->Ph > 0 ? c->Ph : (c->M * 3 / 2) * 2;
    if (ph < 2 * c->M) { set_err("Ph must be >= 2M"); return LSD_E_CONFIG; }
    if (ph % 2) { set_err("Ph must be even"); return LSD_E_CONFIG; }
    if (H) *H = c->L * c->M - c->Mp * (c->M - 1) * c->M / 2;
    if (Th) *Th = th;
    if (Ph) *Ph = ph;
    return LSD_OK;
}

/* ---------------------------------------------------- Gauss-Legendre nodes */

/* auxiliary/gauwts.f:6-75: Newton iteration on the Legendre recurrence. */
void orc_gauwts(double x1, double x2, double* gaux, double* gauw, int n) {
    const double EPS = 1.0e-13, PI = 3.14159265358979312;
    int m = (n + 1) / 2;
    double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
    for (int i = 1; i <= m; i++) {
        double z = cos(PI * ((double)i - 0.25) / ((double)n + 0.5));
        double z1, pp, p1, p2, p3;
        do {
            p1 = 1.0;
            p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / (double)j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
        } while (fabs(z - z1) > EPS);
        gaux[i - 1] = xm - xl * z;
        gaux[n - i] = xm + xl * z;
        gauw[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
        gauw[n - i] = gauw[i - 1];
    }
}

/* ------------------------------------------ Schmidt-normalised P_l^m(x) */

/* auxiliary/pmm.f:20-46 */
static double pmm(int m, double s) {
    if (m == 0) return 1.0;
    double p = sqrt(2.0);
    for (int i = 1; i <= m; i++) {
        double twoi = 2.0 * (double)i;
        p = p * s * sqrt((twoi - 1.0) / twoi);
    }
    return p;
}
/* auxiliary/dpmm.f:11-40 */
static double dpmm(int m, double c, double s) {
    if (m == 0) return 0.0;
    double p = sqrt(2.0) * (double)m * c / s;
    for (int i = 1; i <= m; i++) {
        double twoi = 2.0 * (double)i;
        p = p * s * sqrt((twoi - 1.0) / twoi);
    }
    return p;
}
/* auxiliary/pmm1.f:11-31 */
static double pmm1(int m, double x, double p0) { return x * p0 * sqrt(2.0 * m + 1.0); }
/* auxiliary/dpmm1.f:7-31 */
static double dpmm1(int m, double x, double s, double p0, double dp0) {
    return sqrt(2.0 * m + 1.0) * (x * dp0 - s * p0);
}
/* auxiliary/plm.f:14-54 */
static double plm(int l, int m, double x, double p1, double p2) {
    double rm = m, rl = l;
    double p = (2.0 * rl - 1.0) * x * p1;
    p = p - p2 * sqrt((rl + rm - 1.0) * (rl - rm - 1.0));
    return p / sqrt((rl + rm) * (rl - rm));
}
/* auxiliary/dplm.f:8-51 */
static double dplm(int l, int m, double x, double s, double p1, double dp1, double dp2) {
    double rm = m, rl = l;
    double p = (2.0 * rl - 1.0) * (x * dp1 - s * p1);
    p = p - dp2 * sqrt((rl + rm - 1.0) * (rl - rm - 1.0));
    return p / sqrt((rl + rm) * (rl - rm));
}

/* auxiliary/schnla.f:21-115 for one x; packed index l(l+1)/2+m (0-based). */
void orc_schnla(double* pa, double* dpa, double x, int lh) {
    double s = sqrt((1.0 + x) * (1.0 - x));
    int idx, i1, i2;
    for (int m = 0; m <= lh - 2; m++) {
        int l = m;
        idx = l * (l + 1) / 2 + m;
        pa[idx] = pmm(m, s);
        dpa[idx] = dpmm(m, x, s);
        i2 = idx;
        l = l + 1;
        idx = l * (l + 1) / 2 + m;
        pa[idx] = pmm1(m, x, pa[i2]);
        dpa[idx] = dpmm1(m, x, s, pa[i2], dpa[i2]);
        i1 = idx;
        for (l = m + 2; l <= lh; l++) {
            idx = l * (l + 1) / 2 + m;
            pa[idx] = plm(l, m, x, pa[i1], pa[i2]);
            dpa[idx] = dplm(l, m, x, s, pa[i1], dpa[i1], dpa[i2]);
            i2 = i1;
            i1 = idx;
        }
    }
    int m = lh - 1, l = m;
    if (m >= 0) {
        idx = l * (l + 1) / 2 + m;
        pa[idx] = pmm(m, s);
        dpa[idx] = dpmm(m, x, s);
        i2 = idx;
        l = lh;
        idx = l * (l + 1) / 2 + m;
        pa[idx] = pmm1(m, x, pa[i2]);
        dpa[idx] = dpmm1(m, x, s, pa[i2], dpa[i2]);
    }
    m = lh;
    l = lh;
    idx = l * (l + 1) / 2 + m;
    pa[idx] = pmm(m, s);
    dpa[idx] = dpmm(m, x, s);
}

/* ---------------------------------------------------------- harmonics */

/* Scripts/parallel.h_template:65-80 (_loop_setlm_begin), single rank. */
int orc_harmonics(int L, int M, int Mp, int* lo, int* mo) {
    int H = L * M - Mp * (M - 1) * M / 2;
    int Hc = (M + 1) / 2, Hr = H / Hc, nh = -1;
    for (int jj = 1; jj <= Hc; jj++)
        for (int ii = 1; ii <= Hr; ii++) {
            int m = jj - 1, l = m * Mp + ii - 1;
            if (l >= L) {
                m = 2 * Hc - jj;
                l = L - Hr + ii - 1;
            }
            nh++;
            lo[nh] = l;
            mo[nh] = m * Mp;
        }
    return nh + 1;
}

/* ---------------------------------------------------------------- FFT */

/* Small DFT of size R on (vr, vi) in place; sign -1 forward. */
static void small_dft(int R, double* vr, double* vi, int sign) {
    double tr[16], ti[16];
    if (R == 2) {
        double ar = vr[0] + vr[1], ai = vi[0] + vi[1];
        vr[1] = vr[0] - vr[1];
        vi[1] = vi[0] - vi[1];
        vr[0] = ar;
        vi[0] = ai;
        return;
    }
    for (int k = 0; k < R; k++) {
        double sr = 0.0, si = 0.0;
        for (int q = 0; q < R; q++) {
            double a = sign * 2.0 * D_PI * (double)((k * q) % R) / (double)R;
            double c = cos(a), s = sin(a);
            sr += vr[q] * c - vi[q] * s;
            si += vr[q] * s + vi[q] * c;
        }
        tr[k] = sr;
        ti[k] = si;
    }
    memcpy(vr, tr, R * sizeof(double));
    memcpy(vi, ti, R * sizeof(double));
}

/* Stockham autosort mixed-radix complex FFT (radices 4, 2, 3, 5, then any
 * remaining factor by a direct DFT).  Unnormalised, like FFTW. */
void orc_cfft(int n, double* re, double* im, int sign) {
    int fac[32], nf = 0, rem = n;
    while (rem % 4 == 0) { fac[nf++] = 4; rem /= 4; }
    while (rem % 2 == 0) { fac[nf++] = 2; rem /= 2; }
    while (rem % 3 == 0) { fac[nf++] = 3; rem /= 3; }
    while (rem % 5 == 0) { fac[nf++] = 5; rem /= 5; }
    if (rem > 1) fac[nf++] = rem;
    double* xr = re; double* xi = im;
    double* yr = (double*)malloc(sizeof(double) * n * 2);
    double* yi = yr + n;
    double* br[2] = {xr, yr}; double* bi[2] = {xi, yi};
    int cur = 0, Ns = 1;
    for (int f = 0; f < nf; f++) {
        int R = fac[f], nR = n / R;
        double* ir = br[cur]; double* ii = bi[cur];
        double* orr = br[1 - cur]; double* oi = bi[1 - cur];
        double vr[64], vi[64];
        double* gr = NULL; double* gi = NULL;
        if (R > 64) { gr = (double*)malloc(sizeof(double) * R * 2); gi = gr + R; }
        double* ur = R > 64 ? gr : vr; double* ui = R > 64 ? gi : vi;
        for (int j = 0; j < nR; j++) {
            int jm = j % Ns;
            for (int r = 0; r < R; r++) {
                ur[r] = ir[j + r * nR];
                ui[r] = ii[j + r * nR];
            }
            if (jm) {
                for (int r = 1; r < R; r++) {
                    double a = sign * 2.0 * D_PI * (double)(jm * r) / (double)(Ns * R);
                    double c = cos(a), s = sin(a);
                    double tr = ur[r] * c - ui[r] * s;
                    ui[r] = ur[r] * s + ui[r] * c;
                    ur[r] = tr;
                }
            }
            if (R == 4) {
                double a0r = ur[0] + ur[2], a0i = ui[0] + ui[2];
                double a1r = ur[0] - ur[2], a1i = ui[0] - ui[2];
                double a2r = ur[1] + ur[3], a2i = ui[1] + ui[3];
                double a3r = ur[1] - ur[3], a3i = ui[1] - ui[3];
                /* multiply a3 by sign*i */
                double b3r = -sign * a3i, b3i = sign * a3r;
                ur[0] = a0r + a2r; ui[0] = a0i + a2i;
                ur[2] = a0r - a2r; ui[2] = a0i - a2i;
                ur[1] = a1r + b3r; ui[1] = a1i + b3i;
                ur[3] = a1r - b3r; ui[3] = a1i - b3i;
            } else {
                small_dft(R, ur, ui, sign);
            }
            int idxD = (j / Ns) * Ns * R + jm;
            for (int r = 0; r < R; r++) {
                orr[idxD + r * Ns] = ur[r];
                oi[idxD + r * Ns] = ui[r];
            }
        }
        if (gr) free(gr);
        cur = 1 - cur;
        Ns *= R;
    }
    if (cur == 1) {
        memcpy(re, yr, sizeof(double) * n);
        memcpy(im, yi, sizeof(double) * n);
    }
    free(yr);
}

/* FFTW c2r of length n from X[0..nx-1] (zero beyond, Hermitian extension):
 * x_j = sum_k X_k exp(+2 pi i jk/n).  tmp holds 2n doubles. */
static void c2r(int n, int nx, const double* Xr, const double* Xi, double* x, double* tmp) {
    double* zr = tmp; double* zi = tmp + n;
    memset(tmp, 0, sizeof(double) * 2 * n);
    for (int k = 0; k < nx && k <= n / 2; k++) {
        zr[k] = Xr[k];
        zi[k] = Xi[k];
        if (k > 0 && k < n - k) {
            zr[n - k] = Xr[k];
            zi[n - k] = -Xi[k];
        }
    }
    orc_cfft(n, zr, zi, +1);
    for (int j = 0; j < n; j++) x[j] = zr[j];
}

/* FFTW r2c: X_k = sum_j x_j exp(-2 pi i jk/n), k < nx; Im X_0 = 0. */
static void r2c(int n, int nx, const double* x, double* Xr, double* Xi, double* tmp) {
    double* zr = tmp; double* zi = tmp + n;
    for (int j = 0; j < n; j++) { zr[j] = x[j]; zi[j] = 0.0; }
    orc_cfft(n, zr, zi, -1);
    for (int k = 0; k < nx; k++) { Xr[k] = zr[k]; Xi[k] = zi[k]; }
    Xi[0] = 0.0;
}

/* --------------------------------------------------------- banded LU */

#define AB(i, j) ab[((j) - 1) * LDAB + ((i) - 1)]

/* LAPACK dgbtf2 (the unblocked path dgbtrf takes for KL = 3), as called by
 * mes_lu_find (modules/meshs.F90:447-456).  ab(LDAB, n) with
 * A(i,j) = ab(KV+1+i-j, j), KV = KU+KL; ipiv 1-based.  Returns info. */
int orc_gbtrf(int n, double* ab, int* ipiv) {
    const int ku = KL, kl = KL, kv = ku + kl;
    int info = 0;
    for (int j = ku + 2; j <= (kv < n ? kv : n); j++)
        for (int i = kv - j + 2; i <= kl; i++) AB(i, j) = 0.0;
    int ju = 1;
    for (int j = 1; j <= n; j++) {
        if (j + kv <= n)
            for (int i = 1; i <= kl; i++) AB(i, j + kv) = 0.0;
        int km = kl < n - j ? kl : n - j;
        /* idamax over ab(kv+1 .. kv+1+km, j) */
        int jp = 1;
        double amax = fabs(AB(kv + 1, j));
        for (int i = 2; i <= km + 1; i++) {
            double a = fabs(AB(kv + i, j));
            if (a > amax) { amax = a; jp = i; }
        }
        ipiv[j - 1] = jp + j - 1;
        if (AB(kv + jp, j) != 0.0) {
            int t = j + ku + jp - 1;
            if (t > n) t = n;
            if (t > ju) ju = t;
            if (jp != 1) {
                /* dswap(ju-j+1, ab(kv+jp,j), ldab-1, ab(kv+1,j), ldab-1) */
                for (int c = 0; c <= ju - j; c++) {
                    double* a = &ab[(j - 1 + c) * LDAB + (kv + jp - 1 - c)];
                    double* b = &ab[(j - 1 + c) * LDAB + (kv - c)];
                    double tmp = *a; *a = *b; *b = tmp;
                }
            }
            if (km > 0) {
                double rp = 1.0 / AB(kv + 1, j);
                for (int i = 2; i <= km + 1; i++) AB(kv + i, j) *= rp;
                /* dger(km, ju-j, -1, ab(kv+2,j),1, ab(kv,j+1),ldab-1, ab(kv+1,j+1),ldab-1) */
                for (int c = 1; c <= ju - j; c++) {
                    double y = ab[(j - 1 + c) * LDAB + (kv - c)];
                    if (y != 0.0) {
                        double tmp = -1.0 * y;
                        for (int i = 1; i <= km; i++)
                            ab[(j - 1 + c) * LDAB + (kv - c + i)] += AB(kv + 1 + i, j) * tmp;
                    }
                }
            }
        } else if (info == 0) {
            info = j;
        }
    }
    return info;
}

/* LAPACK dgbtrs('N', n, KL, KU, 1, ...) as called by tim_invX
 * (modules/timestep.F90:566-588). */
void orc_gbtrs(int n, const double* ab, const int* ipiv, double* b) {
    const int kl = KL, kd = KL + KL + 1; /* kd = ku+kl+1 */
    for (int j = 1; j <= n - 1; j++) {
        int lm = kl < n - j ? kl : n - j;
        int l = ipiv[j - 1];
        if (l != j) { double t = b[l - 1]; b[l - 1] = b[j - 1]; b[j - 1] = t; }
        /* dger(lm, 1, -1, ab(kd+1,j), 1, b(j), ldb, b(j+1), ldb) */
        double y = b[j - 1];
        if (y != 0.0) {
            double tmp = -1.0 * y;
            for (int i = 1; i <= lm; i++) b[j - 1 + i] += AB(kd + i, j) * tmp;
        }
    }
    /* dtbsv('Upper','No transpose','Non-unit', n, kl+ku, ab, ldab, b, 1) */
    const int k = 2 * KL;
    for (int j = n; j >= 1; j--) {
        if (b[j - 1] != 0.0) {
            int l = k + 1 - j;
            b[j - 1] = b[j - 1] / AB(k + 1, j);
            double tmp = b[j - 1];
            int i0 = j - k > 1 ? j - k : 1;
            for (int i = j - 1; i >= i0; i--) b[i - 1] -= tmp * AB(l + i, j);
        }
    }
}
#undef AB

/* Dense LU inverse with LAPACK dgetf2 + dgetri (unblocked) semantics, as
 * called by mes_mat_invert (modules/meshs.F90:427-440).  a(lda, n). */
static int mat_invert(int n, double* a, int lda) {
#define A_(i, j) a[((j) - 1) * lda + ((i) - 1)]
    int ipiv[16];
    for (int j = 1; j <= n; j++) {
        int jp = j;
        double amax = fabs(A_(j, j));
        for (int i = j + 1; i <= n; i++)
            if (fabs(A_(i, j)) > amax) { amax = fabs(A_(i, j)); jp = i; }
        ipiv[j - 1] = jp;
        if (A_(jp, j) == 0.0) return j;
        if (jp != j)
            for (int c = 1; c <= n; c++) { double t = A_(j, c); A_(j, c) = A_(jp, c); A_(jp, c) = t; }
        if (j < n) {
            double rp = 1.0 / A_(j, j);
            for (int i = j + 1; i <= n; i++) A_(i, j) *= rp;
            for (int c = j + 1; c <= n; c++) {
                double y = A_(j, c);
                if (y != 0.0)
                    for (int i = j + 1; i <= n; i++) A_(i, c) += A_(i, j) * (-y);
            }
        }
    }
    /* dtrti2('Upper','Non-unit') */
    for (int j = 1; j <= n; j++) {
        A_(j, j) = 1.0 / A_(j, j);
        double ajj = -A_(j, j);
        /* dtrmv('Upper','No transpose','Non-unit', j-1, a, lda, a(1,j), 1) */
        for (int jj = 1; jj <= j - 1; jj++) {
            double x = A_(jj, j);
            if (x != 0.0) {
                for (int i = 1; i <= jj - 1; i++) A_(i, j) += x * A_(i, jj);
                A_(jj, j) = x * A_(jj, jj);
            }
        }
        for (int i = 1; i <= j - 1; i++) A_(i, j) *= ajj;
    }
    /* dgetri unblocked: solve inv(A)*L = inv(U) */
    double work[16];
    for (int j = n; j >= 1; j--) {
        for (int i = j + 1; i <= n; i++) { work[i - 1] = A_(i, j); A_(i, j) = 0.0; }
        if (j < n) {
            /* dgemv('N', n, n-j, -1, a(1,j+1), lda, work(j+1), 1, 1, a(1,j), 1) */
            for (int jj = j + 1; jj <= n; jj++) {
                double x = work[jj - 1];
                if (x != 0.0) {
                    double t = -1.0 * x;
                    for (int i = 1; i <= n; i++) A_(i, j) += t * A_(i, jj);
                }
            }
        }
    }
    for (int j = n - 1; j >= 1; j--) {
        int jp = ipiv[j - 1];
        if (jp != j)
            for (int i = 1; i <= n; i++) { double t = A_(i, j); A_(i, j) = A_(i, jp); A_(i, jp) = t; }
    }
    return 0;
#undef A_
}

/* libgfortran's integer power x**k with a runtime exponent (pow_r8_i4),
 * as used for D%r(:,j) = D%r(:,1)**j (modules/meshs.F90:262-265). */
static double ipow(double x, int k) {
    double pw = 1.0;
    unsigned u;
    if (k == 0) return 1.0;
    if (k < 0) { u = (unsigned)(-k); x = pw / x; }
    else u = (unsigned)k;
    for (;;) {
        if (u & 1) pw *= x;
        u >>= 1;
        if (u) x *= x; else break;
    }
    return pw;
}

/* ------------------------------------------------------------- the sim */

typedef struct { double *re, *im; } coll_t;
typedef struct { double ab[LDAB * 1]; } dummy_t;

struct orc_sim {
    lsd_config c;
    int L, M, Mp, N, Th, Ph, H, Th2, nslot;
    int *hl, *hm, *hs;
    int *slot_start, *slot_len, *slot_fft;
    double *gp, *gw, *sn, *sn1;
    double *tab[8]; /* P, P_, sin1mP, llP_, llsin1mP, gwP, gwllP_, gwllsin1mP; [h*Th2+j] */
    double *rp;     /* r(n,p) = rp[(p+3)*N + n], p = -3..3 */
    double *dr1, *dr2, *dr3, *radlap, *r1dr; /* mesh band [(j)*7 + (3+n-j)] 0-based */
    double *intr2dr;
    double *ll1, *ll11, *rtll1, *rtll11;
    /* LU operators [l*N*LDAB], pivots [l*N]; kinds 0 vT,1 vP,2 vG,3 C,4 X,5 bT,6 bP */
    double* lu[7];
    int* piv[7];
    double* ym[6]; /* Y operators [l*N*7]: 0 vT,1 vP,2 C,3 X,4 bT,5 bP */
    double *gre, *invg; /* gre[(l*2+k)*N+n], invg[l*4 + (i + 2*j)] (2x2 col-major) */
    coll_t f[LSD_NUM_COLL];
    coll_t hist[6];       /* saved N^n (module-private NTor/NPol/NC ...) */
    double* phys[LSD_NUM_PHYS];
    double dt;
    double t;
    int it, step;
    double dterr, lasterr, corr_dt;
    int adaptive;          /* d_timestep < 0: CFL / corrector controlled dt */
    double cfl_dt;         /* tim_cfl_dt of the last non_maxtstep */
    /* conducting inner core (i_Nic > 1) and the inner-core rotation */
    int Nic, NT;                      /* inner-core points; coupled length Nic + N - 1 */
    double* icrp;                     /* inner-core r(n,p) = icrp[(p+3)*Nic + n] */
    double *icdr1, *icdr2, *icdr3, *icradlap;
    double* iclu[2];                  /* coupled X of BT, BP [l*NT*LDAB] */
    int* icpiv[2];                    /* [l*NT] */
    double* icy;                      /* inner-core Y [l*Nic*MBW] */
    coll_t icf[2], icn[2], ichist[2]; /* icBT, icBP, their N, saved N: [h*Nic + n] */
    double* gwp0;                     /* leg_gwP0 [Th] */
    double omega, inertia;            /* rot_omega, rot_inertia */
    double vtorq, mtorq;              /* rot_velTorq, rot_magTorq of the last evaluation */
    double vtorq_s, mtorq_s;          /* velTorq, magTorq saved by rot_predictor/corrector */
};

/* d_courant = 0.8 is a default-real literal in the reference
 * (Scripts/parameters.F90_template:238), so its value is 0.8 rounded to
 * single precision. */
static const double COURANT = (double)0.8f;

static coll_t coll_new(int n) {
    coll_t c;
    c.re = (double*)calloc((size_t)n, sizeof(double));
    c.im = (double*)calloc((size_t)n, sizeof(double));
    return c;
}
static void coll_free(coll_t* c) { free(c->re); free(c->im); }

/* FD weights dr(1..3) of one radial domain of N points, r(n,p) =
 * rp[(p+3)*N + n], and if w != NULL its r^2 dr integration weights
 * (mes_rdom_init, modules/meshs.F90:292-336). */
static int rdom_weights(int N, const double* rp, double* drs[3], double* w) {
    const double* r1 = rp + 4 * N;
    const double* r2 = rp + 5 * N;
    for (int n = 1; n <= N; n++) {
        int l = n - KL > 1 ? n - KL : 1;
        int r = n + KL < N ? n + KL : N;
        int nn = r - l + 1;
        double A[MBW * MBW];
        for (int i = 1; i <= nn; i++) A[(i - 1)] = 1.0;
        for (int j = 2; j <= nn; j++)
            for (int i = 1; i <= nn; i++)
                A[(j - 1) * MBW + (i - 1)] =
                    A[(j - 2) * MBW + (i - 1)] * (r1[l + i - 2] - r1[n - 1]) / (double)(j - 1);
        if (mat_invert(nn, A, MBW)) { set_err("matrix inversion error 1"); return LSD_E_SINGULAR; }
        for (int i = 1; i <= KL; i++)
            for (int j = l; j <= r; j++)
                drs[i - 1][(j - 1) * MBW + (KL + n - j)] = A[(j - l) * MBW + i];
        if (!w) continue;
        /* integration weights r^2 dr (modules/meshs.F90:311-336) */
        double c = 1.0, e = 1.0;
        int np = n + 1 < N ? n + 1 : N, nm = n - 1 > 1 ? n - 1 : 1;
        for (int i = 1; i <= nn; i++) {
            c = c * (r1[np - 1] - r1[n - 1]) / (double)i;
            e = e * (r1[nm - 1] - r1[n - 1]) / (double)i;
            for (int j = l; j <= r; j++) {
                double wt = 0.5 * (c - e) * A[(j - l) * MBW + (i - 1)] * r2[n - 1];
                w[j - 1] += wt;
            }
        }
    }
    return LSD_OK;
}

/* radLap = dr2 + dr1 * 2/r (modules/meshs.F90:339-347). */
static void rdom_radlap(int N, const double* rp, const double* dr1, const double* dr2, double* radlap) {
    const double* rm1 = rp + 2 * N;
    for (int j = 1; j <= N; j++) {
        int l = j - KL > 1 ? j - KL : 1, r = j + KL < N ? j + KL : N;
        for (int n = l; n <= r; n++) {
            int q = (j - 1) * MBW + (KL + n - j);
            radlap[q] = dr2[q] + dr1[q] * 2.0 * rm1[n - 1];
        }
    }
}

/* mes_precompute / mes_rdom_init (modules/meshs.F90:116-219, 228-419),
 * Boussinesq: rho = Temp0 = 0, eta = 1 (program/io.F90:547-550), f-layer
 * off so radLap_flayer = radLap, GreLap = MagTorLap = MagPolLap = radLap. */
static int mesh_init(orc_sim* s) {
    int N = s->N;
    double ri = s->c.rratio / (1.0 - s->c.rratio);
    s->rp = (double*)calloc((size_t)7 * N, sizeof(double));
    double* r1 = s->rp + 4 * N;
    for (int n = 1; n <= N; n++)
        r1[n - 1] = 0.5 * (1.0 + cos(D_PI * (N - n) / (double)(N - 1))) + ri;
    for (int p = -3; p <= 3; p++) {
        if (p == 1) continue;
        for (int n = 0; n < N; n++) s->rp[(p + 3) * N + n] = ipow(r1[n], p);
    }
    s->dr1 = (double*)calloc((size_t)MBW * N, sizeof(double));
    s->dr2 = (double*)calloc((size_t)MBW * N, sizeof(double));
    s->dr3 = (double*)calloc((size_t)MBW * N, sizeof(double));
    s->radlap = (double*)calloc((size_t)MBW * N, sizeof(double));
    s->r1dr = (double*)calloc((size_t)MBW * N, sizeof(double));
    s->intr2dr = (double*)calloc((size_t)N, sizeof(double));
    double* drs[3] = {s->dr1, s->dr2, s->dr3};
    if (rdom_weights(N, s->rp, drs, s->intr2dr)) return LSD_E_SINGULAR;
    rdom_radlap(N, s->rp, s->dr1, s->dr2, s->radlap);
    /* r1dr = dr1 + 1/r on the diagonal (modules/meshs.F90:408-409) */
    const double* rm1 = s->rp + 2 * N;
    memcpy(s->r1dr, s->dr1, sizeof(double) * MBW * N);
    for (int j = 0; j < N; j++) s->r1dr[j * MBW + KL] += rm1[j];
    return LSD_OK;
}

/* The inner-core domain of mes_precompute (modules/meshs.F90:165-169):
 * Nic points r = r_i cos(pi (2 Nic - n) / (2 Nic - 1)), n = Nic+1 .. 2 Nic,
 * the outer half of a Chebyshev grid over the diameter, with its own FD
 * weights and radLap (mes_rdom_init). */
static int ic_mesh_init(orc_sim* s) {
    int Ni = s->Nic, N_ = 2 * Ni;
    double ri = s->c.rratio / (1.0 - s->c.rratio);
    s->icrp = (double*)calloc((size_t)7 * Ni, sizeof(double));
    double* r1 = s->icrp + 4 * Ni;
    for (int n = Ni + 1; n <= N_; n++) r1[n - Ni - 1] = ri * cos(D_PI * (N_ - n) / (double)(N_ - 1));
    for (int p = -3; p <= 3; p++) {
        if (p == 1) continue;
        for (int n = 0; n < Ni; n++) s->icrp[(p + 3) * Ni + n] = ipow(r1[n], p);
    }
    s->icdr1 = (double*)calloc((size_t)MBW * Ni, sizeof(double));
    s->icdr2 = (double*)calloc((size_t)MBW * Ni, sizeof(double));
    s->icdr3 = (double*)calloc((size_t)MBW * Ni, sizeof(double));
    s->icradlap = (double*)calloc((size_t)MBW * Ni, sizeof(double));
    double* drs[3] = {s->icdr1, s->icdr2, s->icdr3};
    if (rdom_weights(Ni, s->icrp, drs, NULL)) return LSD_E_SINGULAR;
    rdom_radlap(Ni, s->icrp, s->icdr1, s->icdr2, s->icradlap);
    return LSD_OK;
}

/* leg_precompute (modules/legendre.F90:63-128). */
static void leg_init(orc_sim* s) {
    int Th = s->Th, Th2 = s->Th2, H = s->H, L = s->L;
    s->gp = (double*)malloc(sizeof(double) * Th);
    s->gw = (double*)malloc(sizeof(double) * Th);
    s->sn = (double*)malloc(sizeof(double) * Th);
    s->sn1 = (double*)malloc(sizeof(double) * Th);
    orc_gauwts(-1.0, 1.0, s->gp, s->gw, Th);
    for (int j = 0; j < Th; j++) {
        double th = acos(s->gp[j]);
        s->sn[j] = sin(th);
        s->sn1[j] = 1.0 / s->sn[j];
    }
    for (int t = 0; t < 8; t++) s->tab[t] = (double*)calloc((size_t)H * Th2, sizeof(double));
    int np = L * (L + 1) / 2;
    double* pa = (double*)malloc(sizeof(double) * np);
    double* dpa = (double*)malloc(sizeof(double) * np);
    for (int j = 0; j < Th2; j++) {
        orc_schnla(pa, dpa, s->gp[j], L - 1);
        for (int nh = 0; nh < H; nh++) {
            int l = s->hl[nh], m_ = s->hm[nh];
            int k = l * (l + 1) / 2 + m_;
            size_t q = (size_t)nh * Th2 + j;
            s->tab[0][q] = pa[k];
            s->tab[1][q] = dpa[k];
            s->tab[2][q] = pa[k] * m_ * s->sn1[j];
            if (l == 0) {
                s->tab[3][q] = 0.0;
                s->tab[4][q] = 0.0;
            } else {
                double w = 1.0 / sqrt((double)(l * (l + 1)));
                s->tab[3][q] = dpa[k] * w;
                s->tab[4][q] = pa[k] * w * m_ * s->sn1[j];
            }
        }
    }
    for (int nh = 0; nh < H; nh++) {
        int l = s->hl[nh], m_ = s->hm[nh];
        for (int j = 0; j < Th2; j++) {
            size_t q = (size_t)nh * Th2 + j;
            double w = s->gw[j] * ((2.0 * l + 1.0) / 4.0);
            if (m_ == 0) w = 2.0 * w;
            s->tab[5][q] = s->tab[0][q] * w;
            if (l == 0) {
                s->tab[6][q] = 0.0;
                s->tab[7][q] = 0.0;
            } else {
                w = w / sqrt((double)(l * (l + 1)));
                s->tab[6][q] = s->tab[1][q] * w;
                s->tab[7][q] = s->tab[0][q] * w * m_ * s->sn1[j];
            }
        }
    }
    free(pa);
    free(dpa);
}

/* ------------------------------------------------------- matrices */

/* Band builders of modules/timestep.F90.  lu is ab(LDAB, N) column-major.
 * kind 0: tim_lumesh_flayer_X / tim_lumesh_X  diag (A+c1_) + c2_*(ll1*r^-2)
 * kind 1: tim_iclumesh_Tor_X/Pol_X (Nic=1)    diag (A+c1_) + (c2_*ll1)*r^-2 */
static void build_X(orc_sim* s, double c1, double c2, int l, int kind, double* ab) {
    int N = s->N;
    double c1_ = c1 / s->dt;
    double c2_ = c2 * s->c.implicit;
    if (s->c.implicit == 0.0) c2_ = 1.0;
    double ll1 = (double)(l * (l + 1));
    const double* rm2 = s->rp + 1 * N;
    memset(ab, 0, sizeof(double) * LDAB * N);
    for (int j = 0; j < N; j++)
        for (int q = 0; q < MBW; q++) ab[j * LDAB + KL + q] = -c2_ * s->radlap[j * MBW + q];
    for (int j = 0; j < N; j++) {
        double* d = &ab[j * LDAB + 2 * KL];
        if (kind == 0) *d = (*d + c1_) + c2_ * (ll1 * rm2[j]);
        else *d = (*d + c1_) + c2_ * ll1 * rm2[j];
    }
    /* zero first and last rows (timestep.F90:166-171) */
    for (int j = 1; j <= 1 + KL; j++) ab[(j - 1) * LDAB + (2 * KL + 1 - j)] = 0.0;
    for (int j = N - KL; j <= N; j++) ab[(j - 1) * LDAB + (2 * KL + N - j)] = 0.0;
}

/* tim_mesh_flayer_Y / tim_mesh_Y (modules/timestep.F90:421-473): (7, N). */
static void build_Y(orc_sim* s, double c1, double c2, int l, double* m) {
    int N = s->N;
    double c1_ = c1 / s->dt;
    double c2_ = c2 * (1.0 - s->c.implicit);
    const double* rm2 = s->rp + 1 * N;
    double ll = (double)(l * (l + 1));
    for (int j = 0; j < N; j++)
        for (int q = 0; q < MBW; q++) m[j * MBW + q] = c2_ * s->radlap[j * MBW + q];
    for (int j = 0; j < N; j++) {
        double* d = &m[j * MBW + KL];
        *d = (*d + c1_) - c2_ * (ll * rm2[j]);
    }
}

/* A(row, j) of a (LDAB, N) LU band, 1-based row/col of the full matrix. */
#define LUA(ab, i, j) (ab)[((j) - 1) * LDAB + (2 * KL + (i) - (j))]
#define MSH(mm, i, j) (mm)[((j) - 1) * MBW + (KL + (i) - (j))]

static int factor(double* ab, int* ipiv, int N, const char* who) {
    if (orc_gbtrf(N, ab, ipiv)) { set_err(who); return LSD_E_SINGULAR; }
    return LSD_OK;
}

/* tim_iclumesh_Tor_X / tim_iclumesh_Pol_X (modules/timestep.F90:218-344)
 * for i_Nic > 1 with the BC rows of mag_bc_Tor / mag_bc_Pol
 * (modules/magnetic.F90:137-282), Boussinesq without a diffusivity profile
 * (MagTorLap = MagPolLap = radLap): one band over the inner-core points
 * 1..Nic and the outer-core points Nic..NT (the inner-core boundary is point
 * Nic of both), row 1 regular at the innermost point, row Nic the
 * continuity of dB/dr across the inner-core boundary, row NT insulating. */
static void build_icX(orc_sim* s, int l, int pol, double* ab) {
    const int Ni = s->Nic, No = s->N, NT = s->NT;
    const double c1_ = 1.0 / s->dt;
    double c2_ = 1.0 * s->c.implicit;
    if (s->c.implicit == 0.0) c2_ = 1.0;
    const double ll1 = (double)(l * (l + 1));
    const double* icrm2 = s->icrp + 1 * Ni;
    const double* icrm1 = s->icrp + 2 * Ni;
    const double* rm2 = s->rp + 1 * No;
    const double* rm1 = s->rp + 2 * No;
    memset(ab, 0, sizeof(double) * LDAB * NT);
    for (int j = 0; j < Ni; j++)
        for (int q = 0; q < MBW; q++) ab[j * LDAB + KL + q] = -c2_ * s->icradlap[j * MBW + q];
    for (int j = 0; j < Ni; j++) {
        double* d = &ab[j * LDAB + 2 * KL];
        *d = (*d + c1_) + c2_ * ll1 * icrm2[j];
    }
    for (int j = 0; j < No; j++)
        for (int q = 0; q < MBW; q++)
            ab[(Ni - 1 + j) * LDAB + KL + q] = ab[(Ni - 1 + j) * LDAB + KL + q] - c2_ * s->radlap[j * MBW + q];
    for (int j = 0; j < No; j++) {
        double* d = &ab[(Ni - 1 + j) * LDAB + 2 * KL];
        *d = (*d + c1_) + c2_ * ll1 * rm2[j];
    }
    for (int j = 1; j <= 1 + KL; j++) LUA(ab, 1, j) = 0.0;
    for (int j = NT - KL; j <= NT; j++) LUA(ab, NT, j) = 0.0;
    for (int j = Ni - KL; j <= Ni + KL; j++) LUA(ab, Ni, j) = 0.0;
    if (!pol) {
        LUA(ab, 1, 1) = 1.0;   /* BTor = 0 at the innermost point */
        LUA(ab, NT, NT) = 1.0; /* insulating exterior */
    } else {
        /* (dr - l/r) BPol = 0 at the innermost point */
        for (int j = 1; j <= 1 + KL; j++) LUA(ab, 1, j) = MSH(s->icdr1, 1, j);
        LUA(ab, 1, 1) = LUA(ab, 1, 1) - icrm1[0] * l;
        /* (dr + (l+1)/r) BPol = 0 at r_o */
        for (int j = No - KL; j <= No; j++) LUA(ab, NT, j + Ni - 1) = MSH(s->dr1, No, j);
        LUA(ab, NT, NT) = LUA(ab, NT, NT) + rm1[No - 1] * (l + 1);
    }
    /* dB/dr continuous across the inner-core boundary */
    for (int j = 2; j <= 1 + KL; j++) LUA(ab, Ni, j + Ni - 1) = MSH(s->dr1, 1, j);
    for (int j = Ni - KL; j <= Ni - 1; j++) LUA(ab, Ni, j) = -MSH(s->icdr1, Ni, j);
    LUA(ab, Ni, Ni) = MSH(s->dr1, 1, 1) - MSH(s->icdr1, Ni, Ni);
}

/* tim_mesh_Y on the inner-core domain (modules/timestep.F90:421-449),
 * c1 = c2 = 1: (7, Nic). */
static void build_icY(orc_sim* s, int l, double* m) {
    const int Ni = s->Nic;
    const double c1_ = 1.0 / s->dt;
    const double c2_ = 1.0 * (1.0 - s->c.implicit);
    const double* icrm2 = s->icrp + 1 * Ni;
    const double ll = (double)(l * (l + 1));
    for (int j = 0; j < Ni; j++)
        for (int q = 0; q < MBW; q++) m[j * MBW + q] = c2_ * s->icradlap[j * MBW + q];
    for (int j = 0; j < Ni; j++) {
        double* d = &m[j * MBW + KL];
        *d = (*d + c1_) - c2_ * (ll * icrm2[j]);
    }
}

/* vel_matrices, cod_matrices, comp_matrices, mag_matrices
 * (modules/velocity.F90:63-170, codensity.F90:46-75, composition.F90:43-72,
 * magnetic.F90:72-133) with the BC rows of vel_bc_Tor/Pol/Gre
 * (velocity.F90:177-273), cod_bc_C (codensity.F90:85-119),
 * comp_bc_C (composition.F90:82-105), mag_bc_Tor/Pol (magnetic.F90:137-282). */
static int matrices(orc_sim* s) {
    int N = s->N, L = s->L;
    const lsd_config* c = &s->c;
    const double* rm1 = s->rp + 2 * N;
    for (int l = 0; l < L; l++) {
        double* ab;
        /* velocity toroidal: no-slip T = 0, stress-free (dr - 1/r) T = 0
         * (vel_bc_Tor: inner no-slip for i_vel_bc 1, 2; outer for 1, 3) */
        const int ns_in = c->vel_bc == 1 || c->vel_bc == 2, ns_out = c->vel_bc == 1 || c->vel_bc == 3;
        if (l >= 1) {
            ab = s->lu[0] + (size_t)l * N * LDAB;
            build_X(s, 1.0, c->Pm, l, 0, ab);
            if (ns_in) {
                LUA(ab, 1, 1) = 1.0;
            } else {
                for (int j = 1; j <= 1 + KL; j++) LUA(ab, 1, j) = MSH(s->dr1, 1, j);
                LUA(ab, 1, 1) = LUA(ab, 1, 1) - rm1[0];
            }
            if (ns_out) {
                LUA(ab, N, N) = 1.0;
            } else {
                for (int j = N - KL; j <= N; j++) LUA(ab, N, j) = MSH(s->dr1, N, j);
                LUA(ab, N, N) = LUA(ab, N, N) - rm1[N - 1];
            }
            if (factor(ab, s->piv[0] + l * N, N, "mes_lu_find")) return LSD_E_SINGULAR;
            build_Y(s, 1.0, c->Pm, l, s->ym[0] + (size_t)l * N * MBW);
            /* velocity poloidal: no-slip dP/dr = 0, stress-free d2P/dr2 = 0
             * (vel_bc_Pol, Boussinesq) */
            ab = s->lu[1] + (size_t)l * N * LDAB;
            build_X(s, 1.0, c->Pm, l, 0, ab);
            for (int j = 1; j <= 1 + KL; j++) LUA(ab, 1, j) = MSH(ns_in ? s->dr1 : s->dr2, 1, j);
            for (int j = N - KL; j <= N; j++) LUA(ab, N, j) = MSH(ns_out ? s->dr1 : s->dr2, N, j);
            if (factor(ab, s->piv[1] + l * N, N, "mes_lu_find")) return LSD_E_SINGULAR;
            build_Y(s, 1.0, c->Pm, l, s->ym[1] + (size_t)l * N * MBW);
            /* Green's function operator */
            ab = s->lu[2] + (size_t)l * N * LDAB;
            build_X(s, 0.0, 1.0 / c->implicit, l, 0, ab);
            LUA(ab, 1, 1) = 1.0;
            LUA(ab, N, N) = 1.0;
            if (factor(ab, s->piv[2] + l * N, N, "mes_lu_find")) return LSD_E_SINGULAR;
            /* magnetic toroidal */
            ab = s->lu[5] + (size_t)l * N * LDAB;
            build_X(s, 1.0, 1.0, l, 1, ab);
            LUA(ab, 1, 1) = 1.0;
            LUA(ab, N, N) = 1.0;
            if (factor(ab, s->piv[5] + l * N, N, "mes_iclu_find")) return LSD_E_SINGULAR;
            build_Y(s, 1.0, 1.0, l, s->ym[4] + (size_t)l * N * MBW);
            /* magnetic poloidal: (dr - l/r) at r_i, (dr + (l+1)/r) at r_o */
            ab = s->lu[6] + (size_t)l * N * LDAB;
            build_X(s, 1.0, 1.0, l, 1, ab);
            for (int j = 1; j <= 1 + KL; j++) LUA(ab, 1, j) = MSH(s->dr1, 1, j);
            LUA(ab, 1, 1) = LUA(ab, 1, 1) - rm1[0] * l;
            for (int j = N - KL; j <= N; j++) LUA(ab, N, j) = MSH(s->dr1, N, j);
            LUA(ab, N, N) = LUA(ab, N, N) + rm1[N - 1] * (l + 1);
            if (factor(ab, s->piv[6] + l * N, N, "mes_iclu_find")) return LSD_E_SINGULAR;
            build_Y(s, 1.0, 1.0, l, s->ym[5] + (size_t)l * N * MBW);
            if (s->Nic > 1) {
                /* conducting inner core: the coupled bands replace the two above */
                for (int k = 0; k < 2; k++) {
                    ab = s->iclu[k] + (size_t)l * s->NT * LDAB;
                    build_icX(s, l, k, ab);
                    if (factor(ab, s->icpiv[k] + l * s->NT, s->NT, "mes_iclu_find")) return LSD_E_SINGULAR;
                }
                build_icY(s, l, s->icy + (size_t)l * s->Nic * MBW);
            }
        }
        /* codensity (l = 0 included) */
        ab = s->lu[3] + (size_t)l * N * LDAB;
        build_X(s, c->Pr / c->Pm, 1.0, l, 0, ab);
        LUA(ab, 1, 1) = 1.0;
        LUA(ab, N, N) = 1.0;
        if (factor(ab, s->piv[3] + l * N, N, "mes_lu_find")) return LSD_E_SINGULAR;
        build_Y(s, c->Pr / c->Pm, 1.0, l, s->ym[2] + (size_t)l * N * MBW);
        /* composition */
        ab = s->lu[4] + (size_t)l * N * LDAB;
        build_X(s, c->Sc / c->Pm, 1.0, l, 0, ab);
        LUA(ab, 1, 1) = 1.0;
        LUA(ab, N, N) = 1.0;
        if (factor(ab, s->piv[4] + l * N, N, "mes_lu_find")) return LSD_E_SINGULAR;
        build_Y(s, c->Sc / c->Pm, 1.0, l, s->ym[3] + (size_t)l * N * MBW);
    }
    /* Green's functions and influence matrix (velocity.F90:134-168) */
    double* re = (double*)malloc(sizeof(double) * 2 * N);
    double* im = re + N;
    for (int l = 1; l < L; l++) {
        memset(re, 0, sizeof(double) * 2 * N);
        re[0] = 1.0;
        im[N - 1] = 1.0;
        const double* ab = s->lu[2] + (size_t)l * N * LDAB;
        const int* pv = s->piv[2] + l * N;
        orc_gbtrs(N, ab, pv, re);
        orc_gbtrs(N, ab, pv, im);
        re[0] = re[N - 1] = im[0] = im[N - 1] = 0.0;
        ab = s->lu[1] + (size_t)l * N * LDAB;
        pv = s->piv[1] + l * N;
        orc_gbtrs(N, ab, pv, re);
        orc_gbtrs(N, ab, pv, im);
        memcpy(s->gre + (size_t)(l * 2 + 0) * N, re, sizeof(double) * N);
        memcpy(s->gre + (size_t)(l * 2 + 1) * N, im, sizeof(double) * N);
        double* g = s->invg + l * 4;
        g[0] = re[0];      /* invG(1,1) */
        g[2] = im[0];      /* invG(1,2) */
        g[1] = re[N - 1];  /* invG(2,1) */
        g[3] = im[N - 1];  /* invG(2,2) */
        if (mat_invert(2, g, 2)) { free(re); set_err("matrix inversion error 1"); return LSD_E_SINGULAR; }
    }
    free(re);
    return LSD_OK;
}

/* ------------------------------------------------------------ create */

orc_sim* orc_create(const lsd_config* cfg) {
    int H, Th, Ph;
    if (orc_sizes(cfg, &H, &Th, &Ph)) return NULL;
    if (!cfg->vel) { set_err("vel must be time-stepped"); return NULL; }
    if (cfg->vel_bc < 1 || cfg->vel_bc > 4 || cfg->cod_bc != 1 || cfg->comp_bc != 1) {
        set_err("i_vel_bc must be 1..4, i_cod_bc = i_comp_bc = 1");
        return NULL;
    }
    if (cfg->Nic > 1 && cfg->vel_bc > 2) {
        set_err("a conducting inner core needs a no-slip inner boundary (i_vel_bc = 1 or 2)");
        return NULL;
    }
    if (cfg->Nic < 0 || (cfg->Nic > 1 && (!cfg->mag || cfg->Nic < 2 * KL + 2))) {
        set_err("i_Nic > 1 needs the magnetic field and at least 8 inner-core points");
        return NULL;
    }
    orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
    s->c = *cfg;
    s->L = cfg->L; s->M = cfg->M; s->Mp = cfg->Mp; s->N = cfg->N;
    s->Th = Th; s->Ph = Ph; s->H = H; s->Th2 = (Th + 1) / 2;
    /* d_timestep < 0 selects the adaptive step; its magnitude stands for the
     * dt io_load_state reads from the state file (program/io.F90:429) */
    s->adaptive = cfg->dt < 0.0;
    s->dt = fabs(cfg->dt);
    int N = s->N, L = s->L;
    s->hl = (int*)malloc(sizeof(int) * H);
    s->hm = (int*)malloc(sizeof(int) * H);
    s->hs = (int*)malloc(sizeof(int) * H);
    orc_harmonics(L, s->M, s->Mp, s->hl, s->hm);
    for (int h = 0; h < H; h++) s->hs[h] = (s->hl[h] + s->hm[h]) % 2;
    /* slots: runs of equal m (_loop_sumth_begin, parallel.h_template:88-96) */
    s->slot_start = (int*)malloc(sizeof(int) * (H + 1));
    s->slot_len = (int*)malloc(sizeof(int) * (H + 1));
    s->slot_fft = (int*)malloc(sizeof(int) * (H + 1));
    int ns = 0;
    for (int h = 0; h < H; h++) {
        if (h == 0 || s->hm[h] != s->hm[h - 1]) {
            s->slot_start[ns] = h;
            s->slot_len[ns] = 0;
            s->slot_fft[ns] = s->hm[h] / s->Mp; /* var_Th%m (variables.F90:179-182) */
            ns++;
        }
        s->slot_len[ns - 1]++;
    }
    s->nslot = ns;
    s->ll1 = (double*)calloc(L, sizeof(double));
    s->ll11 = (double*)calloc(L, sizeof(double));
    s->rtll1 = (double*)calloc(L, sizeof(double));
    s->rtll11 = (double*)calloc(L, sizeof(double));
    for (int l = 1; l < L; l++) { /* var_precompute (variables.F90:101-115) */
        s->ll1[l] = (double)(l * (l + 1));
        s->ll11[l] = 1.0 / s->ll1[l];
        s->rtll1[l] = sqrt(s->ll1[l]);
        s->rtll11[l] = 1.0 / s->rtll1[l];
    }
    s->Nic = cfg->Nic > 1 ? cfg->Nic : 1;
    s->NT = s->Nic + N - 1;
    if (mesh_init(s)) { orc_destroy(s); return NULL; }
    if (s->Nic > 1 && ic_mesh_init(s)) { orc_destroy(s); return NULL; }
    leg_init(s);
    /* leg_gwP0 (modules/legendre.F90:103-104) */
    s->gwp0 = (double*)calloc((size_t)Th, sizeof(double));
    for (int j = 0; j < s->Th2; j++) {
        s->gwp0[j] = s->tab[0][j] * s->gw[j] * 0.5;
        s->gwp0[Th - 1 - j] = s->gwp0[j];
    }
    /* rot_precompute (modules/rotation.F90:20-36) */
    s->omega = 0.0;
    s->inertia = (8.0 * D_PI / 15.0) * ipow(s->rp[4 * N], 5);
    for (int k = 0; k < 2; k++) {
        s->icf[k] = coll_new(H * s->Nic);
        s->icn[k] = coll_new(H * s->Nic);
        s->ichist[k] = coll_new(H * s->Nic);
        if (s->Nic > 1) {
            s->iclu[k] = (double*)calloc((size_t)L * s->NT * LDAB, sizeof(double));
            s->icpiv[k] = (int*)calloc((size_t)L * s->NT, sizeof(int));
        }
    }
    if (s->Nic > 1) s->icy = (double*)calloc((size_t)L * s->Nic * MBW, sizeof(double));
    for (int k = 0; k < 7; k++) {
        s->lu[k] = (double*)calloc((size_t)L * N * LDAB, sizeof(double));
        s->piv[k] = (int*)calloc((size_t)L * N, sizeof(int));
    }
    for (int k = 0; k < 6; k++) s->ym[k] = (double*)calloc((size_t)L * N * MBW, sizeof(double));
    s->gre = (double*)calloc((size_t)L * 2 * N, sizeof(double));
    s->invg = (double*)calloc((size_t)L * 4, sizeof(double));
    for (int k = 0; k < LSD_NUM_COLL; k++) s->f[k] = coll_new(H * N);
    for (int k = 0; k < 6; k++) s->hist[k] = coll_new(H * N);
    for (int k = 0; k < LSD_NUM_PHYS; k++)
        s->phys[k] = (double*)calloc((size_t)Th * Ph * N, sizeof(double));
    s->t = 0.0; s->it = 0; s->step = 0; s->dterr = 0.0; s->lasterr = 1e99;
    if (matrices(s)) { orc_destroy(s); return NULL; }
    return s;
}

void orc_destroy(orc_sim* s) {
    if (!s) return;
    free(s->hl); free(s->hm); free(s->hs);
    free(s->slot_start); free(s->slot_len); free(s->slot_fft);
    free(s->gp); free(s->gw); free(s->sn); free(s->sn1);
    for (int t = 0; t < 8; t++) free(s->tab[t]);
    free(s->rp); free(s->dr1); free(s->dr2); free(s->dr3); free(s->radlap); free(s->r1dr);
    free(s->intr2dr);
    free(s->ll1); free(s->ll11); free(s->rtll1); free(s->rtll11);
    for (int k = 0; k < 7; k++) { free(s->lu[k]); free(s->piv[k]); }
    for (int k = 0; k < 6; k++) free(s->ym[k]);
    free(s->gre); free(s->invg);
    for (int k = 0; k < LSD_NUM_COLL; k++) coll_free(&s->f[k]);
    for (int k = 0; k < 6; k++) coll_free(&s->hist[k]);
    for (int k = 0; k < LSD_NUM_PHYS; k++) free(s->phys[k]);
    free(s->icrp); free(s->icdr1); free(s->icdr2); free(s->icdr3); free(s->icradlap);
    for (int k = 0; k < 2; k++) {
        free(s->iclu[k]); free(s->icpiv[k]);
        coll_free(&s->icf[k]); coll_free(&s->icn[k]); coll_free(&s->ichist[k]);
    }
    free(s->icy); free(s->gwp0);
    free(s);
}

/* Inner-core fields icBT (0) and icBP (1), re/im[h*Nic + n]. */
int orc_set_ic(orc_sim* s, int k, const double* re, const double* im) {
    if (k < 0 || k > 1) { set_err("inner-core field must be 0 or 1"); return LSD_E_ARG; }
    memcpy(s->icf[k].re, re, sizeof(double) * s->H * s->Nic);
    memcpy(s->icf[k].im, im, sizeof(double) * s->H * s->Nic);
    return 0;
}
int orc_get_ic(orc_sim* s, int k, double* re, double* im) {
    if (k < 0 || k > 1) { set_err("inner-core field must be 0 or 1"); return LSD_E_ARG; }
    memcpy(re, s->icf[k].re, sizeof(double) * s->H * s->Nic);
    memcpy(im, s->icf[k].im, sizeof(double) * s->H * s->Nic);
    return 0;
}
/* out[0..2] = rot_omega, rot_velTorq, rot_magTorq. */
int orc_get_rot(orc_sim* s, double* out) {
    out[0] = s->omega; out[1] = s->vtorq; out[2] = s->mtorq;
    return 0;
}
int orc_set_rot(orc_sim* s, double omega) {
    s->omega = omega;
    return 0;
}
/* Inner-core radii r(1..Nic). */
int orc_get_icr(orc_sim* s, double* out) {
    memcpy(out, s->icrp + 4 * s->Nic, sizeof(double) * s->Nic);
    return 0;
}

int orc_set_dt(orc_sim* s, double dt) {
    s->dt = dt;
    return matrices(s);
}

int orc_get_table(orc_sim* s, int t, double* out) {
    size_t n = (size_t)s->H * s->Th2;
    if (t == LSD_T_GP) { memcpy(out, s->gp, sizeof(double) * s->Th); return 0; }
    if (t == LSD_T_GW) { memcpy(out, s->gw, sizeof(double) * s->Th); return 0; }
    if (t >= LSD_T_P && t <= LSD_T_GWLLSIN1MP) { memcpy(out, s->tab[t - 2], sizeof(double) * n); return 0; }
    set_err("unknown table");
    return LSD_E_ARG;
}

int orc_get_mesh(orc_sim* s, int w, double* out) {
    int N = s->N;
    switch (w) {
    case LSD_M_R: memcpy(out, s->rp + 4 * N, sizeof(double) * N); return 0;
    case LSD_M_DR1: memcpy(out, s->dr1, sizeof(double) * MBW * N); return 0;
    case LSD_M_DR2: memcpy(out, s->dr2, sizeof(double) * MBW * N); return 0;
    case LSD_M_RADLAP: memcpy(out, s->radlap, sizeof(double) * MBW * N); return 0;
    case LSD_M_R1DR: memcpy(out, s->r1dr, sizeof(double) * MBW * N); return 0;
    case LSD_M_INTR2DR: memcpy(out, s->intr2dr, sizeof(double) * N); return 0;
    }
    set_err("unknown mesh array");
    return LSD_E_ARG;
}

int orc_set_coll(orc_sim* s, int k, const double* re, const double* im) {
    if (k < 0 || k >= LSD_NUM_COLL) { set_err("bad field"); return LSD_E_ARG; }
    memcpy(s->f[k].re, re, sizeof(double) * s->H * s->N);
    memcpy(s->f[k].im, im, sizeof(double) * s->H * s->N);
    return 0;
}
int orc_get_coll(orc_sim* s, int k, double* re, double* im) {
    if (k < 0 || k >= LSD_NUM_COLL) { set_err("bad field"); return LSD_E_ARG; }
    memcpy(re, s->f[k].re, sizeof(double) * s->H * s->N);
    memcpy(im, s->f[k].im, sizeof(double) * s->H * s->N);
    return 0;
}
int orc_get_phys(orc_sim* s, int k, double* p) {
    if (k < 0 || k >= LSD_NUM_PHYS) { set_err("bad field"); return LSD_E_ARG; }
    memcpy(p, s->phys[k], sizeof(double) * s->Th * s->Ph * s->N);
    return 0;
}

/* ---------------------------------------------------- spectral ops */

/* var_coll_meshmult (modules/variables.F90:476-505): out = A.in along r. */
static void meshmult(const orc_sim* s, const double* A, const double* in, double* out) {
    int N = s->N;
    for (int n = 0; n < N; n++) out[n] = 0.0;
    for (int j = 1; j <= N; j++) {
        int l = j - KL > 1 ? j - KL : 1, r = j + KL < N ? j + KL : N;
        for (int n = l; n <= r; n++) out[n - 1] = out[n - 1] + MSH(A, n, j) * in[j - 1];
    }
}

/* var_coll_TorPol2qst_mag (modules/variables.F90:565-605). */
static void torpol2qst(const orc_sim* s, const coll_t* T, const coll_t* P,
                       coll_t* q, coll_t* sf, coll_t* t) {
    int N = s->N, H = s->H;
    const double* rm1 = s->rp + 2 * N;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int nh = 0; nh < H; nh++) {
        int l = s->hl[nh];
        size_t o = (size_t)nh * N;
        for (int n = 0; n < N; n++) {
            double a = s->ll1[l] * rm1[n];
            q->re[o + n] = P->re[o + n] * a;
            q->im[o + n] = P->im[o + n] * a;
        }
        meshmult(s, s->r1dr, P->re + o, sf->re + o);
        meshmult(s, s->r1dr, P->im + o, sf->im + o);
        for (int n = 0; n < N; n++) {
            sf->re[o + n] = sf->re[o + n] * s->rtll1[l];
            sf->im[o + n] = sf->im[o + n] * s->rtll1[l];
            t->re[o + n] = -T->re[o + n] * s->rtll1[l];
            t->im[o + n] = -T->im[o + n] * s->rtll1[l];
        }
    }
}

/* var_coll_qstcurl (modules/variables.F90:713-754), out of place. */
static void qstcurl(const orc_sim* s, const coll_t* q, const coll_t* sf, const coll_t* t,
                    coll_t* oq, coll_t* os, coll_t* ot) {
    int N = s->N, H = s->H;
    const double* rm1 = s->rp + 2 * N;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int nh = 0; nh < H; nh++) {
        int l = s->hl[nh];
        size_t o = (size_t)nh * N;
        meshmult(s, s->r1dr, sf->re + o, ot->re + o);
        meshmult(s, s->r1dr, sf->im + o, ot->im + o);
        for (int n = 0; n < N; n++) {
            double a = s->rtll1[l] * rm1[n];
            ot->re[o + n] = ot->re[o + n] - a * q->re[o + n];
            ot->im[o + n] = ot->im[o + n] - a * q->im[o + n];
        }
        meshmult(s, s->r1dr, t->re + o, os->re + o);
        meshmult(s, s->r1dr, t->im + o, os->im + o);
        for (int n = 0; n < N; n++) {
            os->re[o + n] = -os->re[o + n];
            os->im[o + n] = -os->im[o + n];
            double a = -s->rtll1[l] * rm1[n];
            oq->re[o + n] = t->re[o + n] * a;
            oq->im[o + n] = t->im[o + n] * a;
        }
    }
}

/* -------------------------------------------------------- transforms */

/* Fourier work array layout: Y[((n*nslot + slot)*Th + j)*2 + {re,im}]. */

/* Combine the parity sums of a synthesis into the Fourier row of one
 * m-slot: rows 1..Th2 = Y0+Y1, mirrored rows = Y0-Y1
 * (modules/transform.fftw3.F90:497-498). */
static void mirror_store(const orc_sim* s, const double* y0, const double* y1, double* F) {
    int Th = s->Th, Th2 = s->Th2, Th_ = s->Th / 2;
    for (int j = 0; j < Th2; j++) {
        F[2 * j] = y0[2 * j] + y1[2 * j];
        F[2 * j + 1] = y0[2 * j + 1] + y1[2 * j + 1];
    }
    for (int k = 0; k < Th_; k++) { /* row Th2+k <- Y(Th_-k) */
        int src = Th_ - 1 - k;
        F[2 * (Th2 + k)] = y0[2 * src] - y1[2 * src];
        F[2 * (Th2 + k) + 1] = y0[2 * src + 1] - y1[2 * src + 1];
    }
    (void)Th;
}

/* tra_s2p (modules/transform.fftw3.F90:280-364), serial branch: scatter
 * m-slot -> FFT index, zero m >= M, c2r over each theta row. */
static void s2p_shell(const orc_sim* s, const double* Yn, double* pn, double* Xr, double* Xi, double* x,
                      double* tmp) {
    int Th = s->Th, Ph = s->Ph, nx = Ph / 2 + 1;
    for (int j = 0; j < Th; j++) {
        for (int k = 0; k < nx; k++) { Xr[k] = 0.0; Xi[k] = 0.0; }
        for (int sl = 0; sl < s->nslot; sl++) {
            int mp = s->slot_fft[sl];
            Xr[mp] = Yn[((size_t)sl * Th + j) * 2];
            Xi[mp] = Yn[((size_t)sl * Th + j) * 2 + 1];
        }
        c2r(Ph, nx, Xr, Xi, x, tmp);
        for (int k = 0; k < Ph; k++) pn[(size_t)k * Th + j] = x[k];
    }
}

/* tra_p2s (modules/transform.fftw3.F90:379-466), serial branch: scale 1/Ph,
 * r2c, gather m-slots. */
static void p2s_shell(const orc_sim* s, const double* pn, double* Yn, double* Xr, double* Xi, double* x,
                      double* tmp) {
    int Th = s->Th, Ph = s->Ph, nx = Ph / 2 + 1;
    double scale = 1.0 / (double)Ph;
    for (int j = 0; j < Th; j++) {
        for (int k = 0; k < Ph; k++) x[k] = pn[(size_t)k * Th + j] * scale;
        r2c(Ph, nx, x, Xr, Xi, tmp);
        for (int sl = 0; sl < s->nslot; sl++) {
            int mp = s->slot_fft[sl];
            Yn[((size_t)sl * Th + j) * 2] = Xr[mp];
            Yn[((size_t)sl * Th + j) * 2 + 1] = Xi[mp];
        }
    }
}

/* Per-thread scratch for one shell. */
typedef struct { double *Y, *y0, *y1, *y2, *y3, *Xr, *Xi, *x, *tmp; } shell_ws;
static shell_ws ws_new(const orc_sim* s, int nfield) {
    shell_ws w;
    int Th = s->Th, Ph = s->Ph;
    w.Y = (double*)calloc((size_t)nfield * s->nslot * Th * 2, sizeof(double));
    w.y0 = (double*)calloc((size_t)s->Th2 * 2 * 4, sizeof(double));
    w.y1 = w.y0 + s->Th2 * 2;
    w.y2 = w.y1 + s->Th2 * 2;
    w.y3 = w.y2 + s->Th2 * 2;
    w.Xr = (double*)calloc((size_t)Ph + 2, sizeof(double));
    w.Xi = (double*)calloc((size_t)Ph + 2, sizeof(double));
    w.x = (double*)calloc((size_t)Ph, sizeof(double));
    w.tmp = (double*)calloc((size_t)Ph * 2, sizeof(double));
    return w;
}
static void ws_free(shell_ws* w) { free(w->Y); free(w->y0); free(w->Xr); free(w->Xi); free(w->x); free(w->tmp); }

/* Spectral value of harmonic h at shell n in a spec-layout pair. */
#define SRE(a, h, n) (a)[(size_t)(n) * H + (h)]

/* Scalar synthesis of one shell: Y(:,m,s) += c * T(:,nh)
 * (modules/transform.fftw3.F90:484-499).  par selects which parity bucket
 * (0: s, 1: 1-s).  rot multiplies c by i.  fac scales c first. */
static void synth_scalar(const orc_sim* s, const double* re, const double* im, int n, const double* tab,
                         int flip, int rot, double fac, int usefac, double* Yfield, shell_ws* w) {
    int H = s->H, Th2 = s->Th2, Th = s->Th;
    for (int sl = 0; sl < s->nslot; sl++) {
        double* y[2] = {w->y0, w->y1};
        memset(w->y0, 0, sizeof(double) * Th2 * 4);
        for (int h = s->slot_start[sl]; h < s->slot_start[sl] + s->slot_len[sl]; h++) {
            double cr = SRE(re, h, n), ci = SRE(im, h, n);
            if (rot) { double t = cr; cr = -ci; ci = t; }
            if (usefac) { cr = cr * fac; ci = ci * fac; }
            int b = flip ? 1 - s->hs[h] : s->hs[h];
            const double* P = tab + (size_t)h * Th2;
            double* Y = y[b];
            for (int j = 0; j < Th2; j++) {
                Y[2 * j] = Y[2 * j] + cr * P[j];
                Y[2 * j + 1] = Y[2 * j + 1] + ci * P[j];
            }
        }
        mirror_store(s, w->y0, w->y1, Yfield + (size_t)sl * Th * 2);
    }
}

/* tra_spec2phys over all shells (modules/transform.fftw3.F90:472-520). */
static void spec2phys(orc_sim* s, const double* sre, const double* sim, double* p) {
    int N = s->N, Th = s->Th, Ph = s->Ph;
#pragma omp parallel num_threads(g_threads)
    {
        shell_ws w = ws_new(s, 1);
#pragma omp for schedule(dynamic)
        for (int n = 0; n < N; n++) {
            synth_scalar(s, sre, sim, n, s->tab[0], 0, 0, 0.0, 0, w.Y, &w);
            s2p_shell(s, w.Y, p + (size_t)n * Th * Ph, w.Xr, w.Xi, w.x, w.tmp);
        }
        ws_free(&w);
    }
}

/* tra_grad (modules/transform.fftw3.F90:955-1000). */
static void tra_grad(orc_sim* s, const double* sre, const double* sim, double* pt, double* pp) {
    int N = s->N, Th = s->Th, Ph = s->Ph;
    const double* rm1 = s->rp + 2 * N;
#pragma omp parallel num_threads(g_threads)
    {
        shell_ws w = ws_new(s, 1);
#pragma omp for schedule(dynamic)
        for (int n = 0; n < N; n++) {
            synth_scalar(s, sre, sim, n, s->tab[1], 1, 0, rm1[n], 1, w.Y, &w);
            s2p_shell(s, w.Y, pt + (size_t)n * Th * Ph, w.Xr, w.Xi, w.x, w.tmp);
            synth_scalar(s, sre, sim, n, s->tab[2], 0, 1, rm1[n], 1, w.Y, &w);
            s2p_shell(s, w.Y, pp + (size_t)n * Th * Ph, w.Xr, w.Xi, w.x, w.tmp);
        }
        ws_free(&w);
    }
}

/* tra_qst2rtp (modules/transform.fftw3.F90:719-818). */
static void qst2rtp(orc_sim* s, const double* qre, const double* qim, const double* sre, const double* sim,
                    const double* tre, const double* tim, double* pr, double* pt, double* pp) {
    int N = s->N, Th = s->Th, Ph = s->Ph, H = s->H, Th2 = s->Th2;
#pragma omp parallel num_threads(g_threads)
    {
        shell_ws w = ws_new(s, 2);
#pragma omp for schedule(dynamic)
        for (int n = 0; n < N; n++) {
            synth_scalar(s, qre, qim, n, s->tab[0], 0, 0, 0.0, 0, w.Y, &w);
            s2p_shell(s, w.Y, pr + (size_t)n * Th * Ph, w.Xr, w.Xi, w.x, w.tmp);
            double* Y2 = w.Y;
            double* Y3 = w.Y + (size_t)s->nslot * Th * 2;
            for (int sl = 0; sl < s->nslot; sl++) {
                double* y2[2] = {w.y0, w.y1};
                double* y3[2] = {w.y2, w.y3};
                memset(w.y0, 0, sizeof(double) * Th2 * 8);
                for (int h = s->slot_start[sl]; h < s->slot_start[sl] + s->slot_len[sl]; h++) {
                    int sp = s->hs[h];
                    double c1r = SRE(sre, h, n), c1i = SRE(sim, h, n);
                    double c2r = SRE(tim, h, n), c2i = -SRE(tre, h, n);
                    double c3r = -SRE(sim, h, n), c3i = SRE(sre, h, n);
                    double c4r = SRE(tre, h, n), c4i = SRE(tim, h, n);
                    const double* lp = s->tab[3] + (size_t)h * Th2;
                    const double* ls = s->tab[4] + (size_t)h * Th2;
                    double* a = y2[1 - sp]; double* b = y2[sp];
                    double* c = y3[sp]; double* d = y3[1 - sp];
                    for (int j = 0; j < Th2; j++) {
                        a[2 * j] = a[2 * j] + c1r * lp[j];
                        a[2 * j + 1] = a[2 * j + 1] + c1i * lp[j];
                        b[2 * j] = b[2 * j] + c2r * ls[j];
                        b[2 * j + 1] = b[2 * j + 1] + c2i * ls[j];
                        c[2 * j] = c[2 * j] + c3r * ls[j];
                        c[2 * j + 1] = c[2 * j + 1] + c3i * ls[j];
                        d[2 * j] = d[2 * j] + c4r * lp[j];
                        d[2 * j + 1] = d[2 * j + 1] + c4i * lp[j];
                    }
                }
                mirror_store(s, w.y0, w.y1, Y2 + (size_t)sl * Th * 2);
                mirror_store(s, w.y2, w.y3, Y3 + (size_t)sl * Th * 2);
            }
            s2p_shell(s, Y2, pt + (size_t)n * Th * Ph, w.Xr, w.Xi, w.x, w.tmp);
            s2p_shell(s, Y3, pp + (size_t)n * Th * Ph, w.Xr, w.Xi, w.x, w.tmp);
        }
        ws_free(&w);
    }
}

/* Even/odd fold of one m-slot's Fourier column (modules/transform.fftw3.F90
 * :862-864): Y0 = north+south, Y1 = south-north, equator halved. */
static void fold(const orc_sim* s, const double* F, double* y0, double* y1) {
    int Th = s->Th, Th2 = s->Th2, Th_ = s->Th / 2 + 1;
    for (int j = 0; j < Th2; j++) {
        int m = Th - 1 - j; /* a3c_Y3(i_Th:Th_:-1) */
        y0[2 * j] = F[2 * j] + F[2 * m];
        y0[2 * j + 1] = F[2 * j + 1] + F[2 * m + 1];
        y1[2 * j] = F[2 * j] - F[2 * m];
        y1[2 * j + 1] = F[2 * j + 1] - F[2 * m + 1];
    }
    if (Th2 == Th_) {
        y0[2 * (Th2 - 1)] = y0[2 * (Th2 - 1)] * 0.5;
        y0[2 * (Th2 - 1) + 1] = y0[2 * (Th2 - 1) + 1] * 0.5;
    }
}

/* tra_phys2spec + tra_leg_trans_p2s_native (modules/transform.fftw3.F90
 * :527-602). */
static void phys2spec(orc_sim* s, const double* p, double* sre, double* sim) {
    int N = s->N, Th = s->Th, Ph = s->Ph, H = s->H, Th2 = s->Th2;
#pragma omp parallel num_threads(g_threads)
    {
        shell_ws w = ws_new(s, 1);
#pragma omp for schedule(dynamic)
        for (int n = 0; n < N; n++) {
            p2s_shell(s, p + (size_t)n * Th * Ph, w.Y, w.Xr, w.Xi, w.x, w.tmp);
            for (int sl = 0; sl < s->nslot; sl++) {
                double* y[2] = {w.y0, w.y1};
                fold(s, w.Y + (size_t)sl * Th * 2, w.y0, w.y1);
                for (int h = s->slot_start[sl]; h < s->slot_start[sl] + s->slot_len[sl]; h++) {
                    const double* g = s->tab[5] + (size_t)h * Th2;
                    const double* Y = y[s->hs[h]];
                    double ar = 0.0, ai = 0.0;
                    for (int j = 0; j < Th2; j++) { ar += Y[2 * j] * g[j]; ai += Y[2 * j + 1] * g[j]; }
                    SRE(sre, h, n) = ar;
                    SRE(sim, h, n) = s->hm[h] == 0 ? 0.0 : ai;
                }
            }
        }
        ws_free(&w);
    }
}

/* tra_rtp2qst (modules/transform.fftw3.F90:825-949).  Note the reference
 * zeroes s%Im (not q%Im) for m = 0 in the q loop; q%Im(m=0) is a sum of
 * exact zeros there because the DC term of an r2c transform is real. */
static void rtp2qst(orc_sim* s, const double* pr, const double* pt, const double* pp, double* qre,
                    double* qim, double* sre, double* sim, double* tre, double* tim) {
    int N = s->N, Th = s->Th, Ph = s->Ph, H = s->H, Th2 = s->Th2;
#pragma omp parallel num_threads(g_threads)
    {
        shell_ws w = ws_new(s, 3);
        double* Y1f = w.Y;
        double* Y2f = w.Y + (size_t)s->nslot * Th * 2;
        double* Y3f = w.Y + (size_t)2 * s->nslot * Th * 2;
        double* z = (double*)malloc(sizeof(double) * Th2 * 2 * 2);
#pragma omp for schedule(dynamic)
        for (int n = 0; n < N; n++) {
            p2s_shell(s, pr + (size_t)n * Th * Ph, Y1f, w.Xr, w.Xi, w.x, w.tmp);
            p2s_shell(s, pt + (size_t)n * Th * Ph, Y2f, w.Xr, w.Xi, w.x, w.tmp);
            p2s_shell(s, pp + (size_t)n * Th * Ph, Y3f, w.Xr, w.Xi, w.x, w.tmp);
            for (int sl = 0; sl < s->nslot; sl++) {
                double* y1[2] = {z, z + Th2 * 2};
                double* y2[2] = {w.y0, w.y1};
                double* y3[2] = {w.y2, w.y3};
                fold(s, Y1f + (size_t)sl * Th * 2, y1[0], y1[1]);
                fold(s, Y2f + (size_t)sl * Th * 2, y2[0], y2[1]);
                fold(s, Y3f + (size_t)sl * Th * 2, y3[0], y3[1]);
                for (int h = s->slot_start[sl]; h < s->slot_start[sl] + s->slot_len[sl]; h++) {
                    int sp = s->hs[h];
                    const double* gP = s->tab[5] + (size_t)h * Th2;
                    const double* gl = s->tab[6] + (size_t)h * Th2;
                    const double* gs = s->tab[7] + (size_t)h * Th2;
                    const double* Y = y1[sp];
                    double a, b;
                    a = 0.0; b = 0.0;
                    for (int j = 0; j < Th2; j++) { a += Y[2 * j] * gP[j]; b += Y[2 * j + 1] * gP[j]; }
                    SRE(qre, h, n) = a;
                    SRE(qim, h, n) = b;
                    const double* A2 = y2[1 - sp]; const double* B2 = y2[sp];
                    const double* A3 = y3[sp]; const double* B3 = y3[1 - sp];
                    double s_re = 0, s_im = 0, t_re = 0, t_im = 0, u;
                    for (int j = 0; j < Th2; j++) s_re += A2[2 * j] * gl[j];
                    for (int j = 0; j < Th2; j++) s_im += A2[2 * j + 1] * gl[j];
                    u = 0; for (int j = 0; j < Th2; j++) u += B2[2 * j + 1] * gs[j];
                    t_re = -u;
                    for (int j = 0; j < Th2; j++) t_im += B2[2 * j] * gs[j];
                    u = 0; for (int j = 0; j < Th2; j++) u += A3[2 * j + 1] * gs[j];
                    s_re = s_re + u;
                    u = 0; for (int j = 0; j < Th2; j++) u += A3[2 * j] * gs[j];
                    s_im = s_im - u;
                    if (s->hm[h] == 0) s_im = 0.0;
                    u = 0; for (int j = 0; j < Th2; j++) u += B3[2 * j] * gl[j];
                    t_re = t_re + u;
                    u = 0; for (int j = 0; j < Th2; j++) u += B3[2 * j + 1] * gl[j];
                    t_im = t_im + u;
                    if (s->hm[h] == 0) t_im = 0.0;
                    SRE(sre, h, n) = s_re;
                    SRE(sim, h, n) = s_im;
                    SRE(tre, h, n) = t_re;
                    SRE(tim, h, n) = t_im;
                }
            }
        }
        free(z);
        ws_free(&w);
    }
}

/* var_coll2spec / var_spec2coll, serial branch (modules/variables.F90
 * :246-268, 365-389): plain transposes. */
static void c2s(const orc_sim* s, const coll_t* c, double* re, double* im) {
    int H = s->H, N = s->N;
    for (int h = 0; h < H; h++)
        for (int n = 0; n < N; n++) {
            re[(size_t)n * H + h] = c->re[(size_t)h * N + n];
            im[(size_t)n * H + h] = c->im[(size_t)h * N + n];
        }
}
static void s2c(const orc_sim* s, const double* re, const double* im, coll_t* c) {
    int H = s->H, N = s->N;
    for (int h = 0; h < H; h++)
        for (int n = 0; n < N; n++) {
            c->re[(size_t)h * N + n] = re[(size_t)n * H + h];
            c->im[(size_t)h * N + n] = im[(size_t)n * H + h];
        }
}

/* Drop-in transform entry points (spec/phys in reference layouts). */
int orc_tra_spec2phys(orc_sim* s, const double* sre, const double* sim, double* p) {
    spec2phys(s, sre, sim, p);
    return 0;
}
int orc_tra_phys2spec(orc_sim* s, const double* p, double* sre, double* sim) {
    phys2spec(s, p, sre, sim);
    return 0;
}
int orc_tra_grad(orc_sim* s, const double* sre, const double* sim, double* t, double* p) {
    tra_grad(s, sre, sim, t, p);
    return 0;
}
int orc_tra_qst2rtp(orc_sim* s, const double* qre, const double* qim, const double* sre,
                    const double* sim, const double* tre, const double* tim, double* r, double* th,
                    double* ph) {
    qst2rtp(s, qre, qim, sre, sim, tre, tim, r, th, ph);
    return 0;
}
int orc_tra_rtp2qst(orc_sim* s, const double* r, const double* th, const double* ph, double* qre,
                    double* qim, double* sre, double* sim, double* tre, double* tim) {
    rtp2qst(s, r, th, ph, qre, qim, sre, sim, tre, tim);
    return 0;
}

/* -------------------------------------------------- transform drivers */

typedef struct { double *re, *im; } specbuf;
static specbuf spec_new(const orc_sim* s) {
    specbuf b;
    b.re = (double*)malloc(sizeof(double) * s->H * s->N);
    b.im = (double*)malloc(sizeof(double) * s->H * s->N);
    return b;
}
static void spec_free(specbuf* b) { free(b->re); free(b->im); }

/* vel_transform / mag_transform, Boussinesq branch
 * (modules/velocity.F90:310-350, magnetic.F90:336-360). */
static void vec_transform(orc_sim* s, const coll_t* T, const coll_t* P, int o) {
    int HN = s->H * s->N;
    coll_t q = coll_new(HN), sf = coll_new(HN), t = coll_new(HN);
    coll_t oq = coll_new(HN), os = coll_new(HN), ot = coll_new(HN);
    specbuf a = spec_new(s), b = spec_new(s), c = spec_new(s);
    torpol2qst(s, T, P, &q, &sf, &t);
    c2s(s, &q, a.re, a.im); c2s(s, &sf, b.re, b.im); c2s(s, &t, c.re, c.im);
    qst2rtp(s, a.re, a.im, b.re, b.im, c.re, c.im, s->phys[o], s->phys[o + 1], s->phys[o + 2]);
    qstcurl(s, &q, &sf, &t, &oq, &os, &ot);
    c2s(s, &oq, a.re, a.im); c2s(s, &os, b.re, b.im); c2s(s, &ot, c.re, c.im);
    qst2rtp(s, a.re, a.im, b.re, b.im, c.re, c.im, s->phys[o + 3], s->phys[o + 4], s->phys[o + 5]);
    coll_free(&q); coll_free(&sf); coll_free(&t);
    coll_free(&oq); coll_free(&os); coll_free(&ot);
    spec_free(&a); spec_free(&b); spec_free(&c);
}

/* cod_transform / comp_transform (modules/codensity.F90:195-213,
 * composition.F90:181-198). */
static void scal_transform(orc_sim* s, const coll_t* C, int o) {
    int HN = s->H * s->N, N = s->N;
    coll_t cC = coll_new(HN);
    for (int h = 0; h < s->H; h++) {
        meshmult(s, s->dr1, C->re + (size_t)h * N, cC.re + (size_t)h * N);
        meshmult(s, s->dr1, C->im + (size_t)h * N, cC.im + (size_t)h * N);
    }
    specbuf a = spec_new(s), b = spec_new(s);
    c2s(s, C, a.re, a.im);
    c2s(s, &cC, b.re, b.im);
    spec2phys(s, b.re, b.im, s->phys[o]);
    tra_grad(s, a.re, a.im, s->phys[o + 1], s->phys[o + 2]);
    coll_free(&cC);
    spec_free(&a); spec_free(&b);
}

/* vel_angmomclr, Boussinesq (modules/velocity.F90:475-513): with both
 * boundaries stress-free the rigid rotations omz, omx, omy are removed
 * from T of (l, m) = (1, 0) and (1, 1), each the ratio of the moments
 * int r^2 (r T) dr and int r^2 r^2 dr with the intr2dr weights. */
static void angmomclr(orc_sim* s) {
    if (s->c.vel_bc != 4) return;
    const int N = s->N;
    const double* r1 = s->rp + 4 * N;
    const double* r2 = s->rp + 5 * N;
    const double w = 8.0 / 3.0 * D_PI;
    coll_t* T = &s->f[LSD_VEL_TOR];
    for (int nh = 0; nh < s->H; nh++) {
        if (s->hl[nh] != 1 || s->hm[nh] > 1) continue;
        double* re = T->re + (size_t)nh * N;
        double* im = T->im + (size_t)nh * N;
        double d1 = 0.0, d2 = 0.0;
        for (int n = 0; n < N; n++) d1 = d1 + (re[n] * r1[n]) * s->intr2dr[n];
        for (int n = 0; n < N; n++) d2 = d2 + r2[n] * s->intr2dr[n];
        if (s->hm[nh] == 0) {
            const double omz = (w * d1) / (w * d2);
            for (int n = 0; n < N; n++) re[n] = re[n] - omz * r1[n];
        } else {
            double d3 = 0.0;
            for (int n = 0; n < N; n++) d3 = d3 + (im[n] * r1[n]) * s->intr2dr[n];
            const double omx = (2.0 * w * d1) / (2.0 * w * d2);
            const double omy = (-2.0 * w * d3) / (-2.0 * w * d2);
            for (int n = 0; n < N; n++) re[n] = re[n] - omx * r1[n];
            for (int n = 0; n < N; n++) im[n] = im[n] - omy * r1[n];
        }
    }
}

int orc_transform(orc_sim* s) {
    angmomclr(s);
    vec_transform(s, &s->f[LSD_VEL_TOR], &s->f[LSD_VEL_POL], LSD_P_VEL_R);
    if (s->c.cod) scal_transform(s, &s->f[LSD_COD], LSD_P_COD_GR);
    if (s->c.comp) scal_transform(s, &s->f[LSD_COMP], LSD_P_COMP_GR);
    if (s->c.mag) vec_transform(s, &s->f[LSD_MAG_TOR], &s->f[LSD_MAG_POL], LSD_P_MAG_R);
    return 0;
}

/* ---------------------------------------------------------- nonlinear */

/* tra_surfintegral (modules/transform.fftw3.F90:1039-1062) of one shell
 * p[k*Th + j]: (4 pi / Ph) sum_j gwP0(j) sum_k p(j, k). */
static double surfintegral(const orc_sim* s, const double* p) {
    const int Th = s->Th, Ph = s->Ph;
    double S_ = 0.0;
    for (int j = 0; j < Th; j++) {
        double Y = 0.0;
        for (int k = 0; k < Ph; k++) Y = Y + p[(size_t)k * Th + j];
        S_ = S_ + Y * s->gwp0[j];
    }
    return S_ * 4.0 * D_PI / (double)Ph;
}

/* non_rotation (modules/nonlinear.F90:479-537), no-slip: the viscous
 * torque r_i^3 Pm surf((-2/r_i u_phi - (curl u)_theta) sin) and the
 * magnetic torque r_i^3 (2 Pm / E) surf(B_r B_phi sin) on the inner core,
 * from the physical fields of the last transform at r_i. */
static void rotation_torques(orc_sim* s) {
    const int Th = s->Th, Ph = s->Ph, N = s->N;
    const size_t plane = (size_t)Th * Ph;
    const double ri = s->rp[4 * N], ri3 = s->rp[6 * N];
    const double vfac = ri3 * s->c.Pm;
    const double mfac = ri3 * 2 * s->c.Pm / s->c.E;
    double* p = (double*)calloc(plane, sizeof(double));
    const double S = -2.0 / ri;
    const double* up = s->phys[LSD_P_VEL_P];
    const double* wt = s->phys[LSD_P_VCURL_T];
    for (int k = 0; k < Ph; k++)
        for (int j = 0; j < Th; j++) {
            const size_t q = (size_t)k * Th + j;
            p[q] = (S * up[q] - wt[q]) * s->sn[j];
        }
    s->vtorq = s->c.vel_bc > 2 ? 0.0 : surfintegral(s, p) * vfac;
    if (s->Nic == 1) {
        s->mtorq = 0.0;
    } else {
        const double* br = s->phys[LSD_P_MAG_R];
        const double* bp = s->phys[LSD_P_MAG_P];
        for (int k = 0; k < Ph; k++)
            for (int j = 0; j < Th; j++) {
                const size_t q = (size_t)k * Th + j;
                p[q] = br[q] * bp[q] * s->sn[j];
            }
        s->mtorq = surfintegral(s, p) * mfac;
    }
    free(p);
}

/* The inner-core part of non_magnetic (modules/nonlinear.F90:450-471):
 * induction by the solid-body rotation of the inner core,
 * N = -rot_omega dB/dphi, and zero at (l, m) = (0, 0). */
static void ic_nonlinear(orc_sim* s) {
    const int Ni = s->Nic, H = s->H;
    for (int k = 0; k < 2; k++) {
        const coll_t* in = &s->icf[k];
        coll_t* out = &s->icn[k];
        for (int nh = 0; nh < H; nh++) {
            const double m_ = (double)s->hm[nh];
            for (int n = 0; n < Ni; n++) {
                const size_t q = (size_t)nh * Ni + n;
                const double a = in->re[q];
                const double dre = -in->im[q] * m_, dim = a * m_;
                out->re[q] = -(s->omega * dre);
                out->im[q] = -(s->omega * dim);
            }
        }
        for (int n = 0; n < Ni; n++) out->re[n] = out->im[n] = 0.0;
    }
}

/* non_velocity, non_codensity, non_composition, non_magnetic, Boussinesq
 * branches (modules/nonlinear.F90:28-199, 205-289, 294-357, 363-473). */
int orc_nonlinear(orc_sim* s) {
    if (s->c.rot) rotation_torques(s);
    int N = s->N, Th = s->Th, Ph = s->Ph, H = s->H, HN = H * N;
    size_t np = (size_t)Th * Ph * N;
    const lsd_config* c = &s->c;
    double** P = s->phys;
    double* r = (double*)malloc(sizeof(double) * np);
    double* th = (double*)malloc(sizeof(double) * np);
    double* ph = (double*)malloc(sizeof(double) * np);
    specbuf a = spec_new(s), b = spec_new(s), d = spec_new(s);
    coll_t cq = coll_new(HN), cs = coll_new(HN), ct = coll_new(HN);
    const double* r1 = s->rp + 4 * N;
    const double* rm1 = s->rp + 2 * N;

    /* --- velocity --- */
    double d_Cor = 2.0 * c->Pm / c->E;
    double Cmag = 2 * c->Pm / c->E;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int n = 0; n < N; n++)
        for (int k = 0; k < Ph; k++)
            for (int j = 0; j < Th; j++) {
                size_t q = ((size_t)n * Ph + k) * Th + j;
                double ur = P[0][q], ut = P[1][q], up = P[2][q];
                double wr = P[3][q], wt = P[4][q], wp = P[5][q];
                double br = P[6][q], bt = P[7][q], bp = P[8][q];
                double jr = P[9][q], jt = P[10][q], jp = P[11][q];
                double lsin = s->sn[j], lgp = s->gp[j];
                double vr = ut * wp - up * wt;
                double vt = up * wr - ur * wp;
                double vp = ur * wt - ut * wr;
                vr = vr + d_Cor * lsin * up;
                vt = vt + d_Cor * lgp * up;
                vp = vp - d_Cor * (lgp * ut + lsin * ur);
                vr = vr + Cmag * (jt * bp - jp * bt);
                vt = vt + Cmag * (jp * br - jr * bp);
                vp = vp + Cmag * (jr * bt - jt * br);
                r[q] = vr; th[q] = vt; ph[q] = vp;
            }
    rtp2qst(s, r, th, ph, a.re, a.im, b.re, b.im, d.re, d.im);
    s2c(s, a.re, a.im, &cq); s2c(s, b.re, b.im, &cs); s2c(s, d.re, d.im, &ct);
    /* buoyancy (nonlinear.F90:158-185) */
    for (int nh = 0; nh < H; nh++)
        for (int n = 0; n < N; n++) {
            size_t q = (size_t)nh * N + n;
            double d2 = c->Ra * r1[n] * (c->Pm * c->Pm) / c->Pr;
            cq.re[q] = cq.re[q] + d2 * s->f[LSD_COD].re[q];
            cq.im[q] = cq.im[q] + d2 * s->f[LSD_COD].im[q];
        }
    for (int nh = 0; nh < H; nh++)
        for (int n = 0; n < N; n++) {
            size_t q = (size_t)nh * N + n;
            double d2 = c->Ra_comp * r1[n] * (c->Pm * c->Pm) / c->Sc;
            cq.re[q] = cq.re[q] + d2 * s->f[LSD_COMP].re[q];
            cq.im[q] = cq.im[q] + d2 * s->f[LSD_COMP].im[q];
        }
    /* var_coll_qstllcurlr, var_coll_qstllcurlcurlr (variables.F90:762-805) */
    coll_t* NT = &s->f[LSD_N_VEL_TOR];
    coll_t* NP = &s->f[LSD_N_VEL_POL];
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int nh = 0; nh < H; nh++) {
        int l = s->hl[nh];
        size_t o = (size_t)nh * N;
        for (int n = 0; n < N; n++) {
            NT->re[o + n] = -ct.re[o + n] * s->rtll11[l];
            NT->im[o + n] = -ct.im[o + n] * s->rtll11[l];
        }
        meshmult(s, s->r1dr, cs.re + o, NP->re + o);
        meshmult(s, s->r1dr, cs.im + o, NP->im + o);
        for (int n = 0; n < N; n++) {
            double v = NP->re[o + n] * s->rtll11[l];
            NP->re[o + n] = cq.re[o + n] * rm1[n] - v;
            v = NP->im[o + n] * s->rtll11[l];
            NP->im[o + n] = cq.im[o + n] * rm1[n] - v;
        }
    }
    for (int n = 0; n < N; n++) { NT->re[n] = NT->im[n] = NP->re[n] = NP->im[n] = 0.0; }

    /* --- codensity / composition: -(Pr/Pm) u.grad C --- */
    for (int pass = 0; pass < 2; pass++) {
        int on = pass == 0 ? s->c.cod : s->c.comp;
        if (!on) continue;
        int g = pass == 0 ? LSD_P_COD_GR : LSD_P_COMP_GR;
        double d_qm1 = pass == 0 ? c->Pr / c->Pm : c->Sc / c->Pm;
        double src = pass == 0 ? c->source : c->source_comp;
        coll_t* NC = &s->f[pass == 0 ? LSD_N_COD : LSD_N_COMP];
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (size_t q = 0; q < np; q++)
            r[q] = -d_qm1 * P[0][q] * P[g][q] - d_qm1 * P[1][q] * P[g + 1][q] -
                   d_qm1 * P[2][q] * P[g + 2][q];
        phys2spec(s, r, a.re, a.im);
        s2c(s, a.re, a.im, NC);
        for (int n = 0; n < N; n++) NC->re[n] = NC->re[n] + src;
    }

    /* --- magnetic: E = u x B --- */
    if (s->c.mag) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (size_t q = 0; q < np; q++) {
            double ur = P[0][q], ut = P[1][q], up = P[2][q];
            double br = P[6][q], bt = P[7][q], bp = P[8][q];
            r[q] = ut * bp - up * bt;
            th[q] = up * br - ur * bp;
            ph[q] = ur * bt - ut * br;
        }
        rtp2qst(s, r, th, ph, a.re, a.im, b.re, b.im, d.re, d.im);
        s2c(s, a.re, a.im, &cq); s2c(s, b.re, b.im, &cs); s2c(s, d.re, d.im, &ct);
        coll_t* MP = &s->f[LSD_N_MAG_POL];
        coll_t* MT = &s->f[LSD_N_MAG_TOR];
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int nh = 0; nh < H; nh++) {
            int l = s->hl[nh];
            size_t o = (size_t)nh * N;
            for (int n = 0; n < N; n++) {
                MP->re[o + n] = -ct.re[o + n] * s->rtll11[l];
                MP->im[o + n] = -ct.im[o + n] * s->rtll11[l];
            }
            meshmult(s, s->r1dr, cs.re + o, MT->re + o);
            meshmult(s, s->r1dr, cs.im + o, MT->im + o);
            for (int n = 0; n < N; n++) {
                double v = MT->re[o + n] * s->rtll11[l];
                MT->re[o + n] = cq.re[o + n] * rm1[n] - v;
                v = MT->im[o + n] * s->rtll11[l];
                MT->im[o + n] = cq.im[o + n] * rm1[n] - v;
            }
        }
        for (int n = 0; n < N; n++) { MT->re[n] = MT->im[n] = MP->re[n] = MP->im[n] = 0.0; }
        ic_nonlinear(s);
    }
    free(r); free(th); free(ph);
    spec_free(&a); spec_free(&b); spec_free(&d);
    coll_free(&cq); coll_free(&cs); coll_free(&ct);
    return 0;
}

/* ------------------------------------------------ predictor / corrector */

/* tim_multY (modules/timestep.F90:640-670) for one harmonic, in place safe:
 * out = c + A.b. */
static void multY(const orc_sim* s, const double* A, const double* b, const double* cc, double* out) {
    int N = s->N;
    double tmp[1024];
    for (int n = 0; n < N; n++) tmp[n] = cc[n];
    for (int j = 1; j <= N; j++) {
        int nl = j - KL > 1 ? j - KL : 1, nr = j + KL < N ? j + KL : N;
        for (int n = nl; n <= nr; n++) tmp[n - 1] = tmp[n - 1] + MSH(A, n, j) * b[j - 1];
    }
    for (int n = 0; n < N; n++) out[n] = tmp[n];
}

/* Per-field correction norm sum for tim_addcorr (timestep.F90:715-737). */
static double addcorr(const orc_sim* s, const coll_t* ac, coll_t* a) {
    size_t HN = (size_t)s->H * s->N;
    double e = 0.0;
    for (size_t q = 0; q < HN; q++) {
        e = e + ac->re[q] * ac->re[q];
        e = e + ac->im[q] * ac->im[q];
        a->re[q] = a->re[q] + ac->re[q];
        a->im[q] = a->im[q] + ac->im[q];
    }
    return sqrt(e);
}

/* tim_addcorr on a field of n entries (the inner-core fields). */
static double addcorr_n(const coll_t* ac, coll_t* a, size_t n) {
    double e = 0.0;
    for (size_t q = 0; q < n; q++) {
        e = e + ac->re[q] * ac->re[q];
        e = e + ac->im[q] * ac->im[q];
        a->re[q] = a->re[q] + ac->re[q];
        a->im[q] = a->im[q] + ac->im[q];
    }
    return sqrt(e);
}

/* vel_gre_adjust (modules/velocity.F90:445-465) on one harmonic vector. */
static void gre_adjust(const orc_sim* s, int l, double* v) {
    int N = s->N;
    const double* g = s->invg + l * 4;
    const double* g1 = s->gre + (size_t)(l * 2) * N;
    const double* g2 = s->gre + (size_t)(l * 2 + 1) * N;
    double BCi = v[0], BCo = v[N - 1];
    double a = g[0] * BCi + g[2] * BCo;
    double b = g[1] * BCi + g[3] * BCo;
    for (int n = 0; n < N; n++) v[n] = v[n] - a * g1[n] - b * g2[n];
}

static const double* LUp(const orc_sim* s, int k, int l) { return s->lu[k] + (size_t)l * s->N * LDAB; }
static const int* PVp(const orc_sim* s, int k, int l) { return s->piv[k] + l * s->N; }
static const double* YMp(const orc_sim* s, int k, int l) { return s->ym[k] + (size_t)l * s->N * MBW; }

/* tim_multY on the inner-core domain for one harmonic: out = c + A.b. */
static void multY_ic(const orc_sim* s, const double* A, const double* b, const double* cc, double* out) {
    const int Ni = s->Nic;
    double tmp[1024];
    for (int n = 0; n < Ni; n++) tmp[n] = cc[n];
    for (int j = 1; j <= Ni; j++) {
        int nl = j - KL > 1 ? j - KL : 1, nr = j + KL < Ni ? j + KL : Ni;
        for (int n = nl; n <= nr; n++) tmp[n - 1] = tmp[n - 1] + MSH(A, n, j) * b[j - 1];
    }
    for (int n = 0; n < Ni; n++) out[n] = tmp[n];
}

/* mag_predictor for i_Nic > 1 (modules/magnetic.F90:370-411), one harmonic
 * and part of field k (0 BT, 1 BP): Y B + N on each domain, zero BC rows
 * (mag_setbc_Tor / mag_setbc_Pol with zero BC arrays, tim_zerobc on the
 * inner core), then tim_icinvX (modules/timestep.F90:594-634). */
static void ic_predict(orc_sim* s, int nh, int part, int k) {
    const int N = s->N, Ni = s->Nic, NT = s->NT, l = s->hl[nh];
    const size_t o = (size_t)nh * N, oi = (size_t)nh * Ni;
    double* x = part ? s->f[LSD_MAG_TOR + k].im + o : s->f[LSD_MAG_TOR + k].re + o;
    const double* nn = part ? s->f[LSD_N_MAG_TOR + k].im + o : s->f[LSD_N_MAG_TOR + k].re + o;
    double* xi = part ? s->icf[k].im + oi : s->icf[k].re + oi;
    const double* ni = part ? s->icn[k].im + oi : s->icn[k].re + oi;
    double v[1024], vi[1024], b[2048];
    if (l == 0) memcpy(v, nn, sizeof(double) * N);
    else multY(s, YMp(s, 4 + k, l), x, nn, v);
    if (l == 0) memcpy(vi, ni, sizeof(double) * Ni);
    else multY_ic(s, s->icy + (size_t)l * Ni * MBW, xi, ni, vi);
    vi[0] = 0.0; vi[Ni - 1] = 0.0;
    v[0] = 0.0; v[N - 1] = 0.0;
    if (l != 0) {
        for (int n = 0; n < Ni - 1; n++) b[n] = vi[n];
        for (int n = 0; n < N; n++) b[Ni - 1 + n] = v[n];
        orc_gbtrs(NT, s->iclu[k] + (size_t)l * NT * LDAB, s->icpiv[k] + l * NT, b);
        for (int n = 0; n < Ni; n++) vi[n] = b[n];
        for (int n = 0; n < N; n++) v[n] = b[Ni - 1 + n];
    }
    memcpy(x, v, sizeof(double) * N);
    memcpy(xi, vi, sizeof(double) * Ni);
}

/* vel_predictor, cod_predictor, comp_predictor, mag_predictor
 * (modules/velocity.F90:369-395, codensity.F90:224-236,
 * composition.F90:209-223, magnetic.F90:370-411). */
int orc_predictor(orc_sim* s) {
    int N = s->N, H = s->H;
    size_t HN = (size_t)H * N;
    /* rot_predictor (modules/rotation.F90:43-64) */
    if (s->c.rot) {
        s->omega = s->omega + (s->dt / s->inertia) * (s->vtorq + s->mtorq);
        s->vtorq_s = s->vtorq;
        s->mtorq_s = s->mtorq;
    }
    /* save N at time t */
    int nmap[6] = {LSD_N_VEL_TOR, LSD_N_VEL_POL, LSD_N_COD, LSD_N_COMP, LSD_N_MAG_TOR, LSD_N_MAG_POL};
    for (int k = 0; k < 6; k++) {
        memcpy(s->hist[k].re, s->f[nmap[k]].re, sizeof(double) * HN);
        memcpy(s->hist[k].im, s->f[nmap[k]].im, sizeof(double) * HN);
    }
    const int ic = s->Nic > 1 && s->c.mag;
    if (ic)
        for (int k = 0; k < 2; k++) {
            memcpy(s->ichist[k].re, s->icn[k].re, sizeof(double) * H * s->Nic);
            memcpy(s->ichist[k].im, s->icn[k].im, sizeof(double) * H * s->Nic);
        }
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 16)
    for (int nh = 0; nh < H; nh++) {
        int l = s->hl[nh], m_ = s->hm[nh];
        size_t o = (size_t)nh * N;
        double v[1024];
        for (int part = 0; part < 2; part++) {
            /* velocity toroidal */
            double* x = part ? s->f[LSD_VEL_TOR].im + o : s->f[LSD_VEL_TOR].re + o;
            const double* nn = part ? s->f[LSD_N_VEL_TOR].im + o : s->f[LSD_N_VEL_TOR].re + o;
            if (l == 0) memcpy(v, nn, sizeof(double) * N);
            else multY(s, YMp(s, 0, l), x, nn, v);
            v[0] = 0.0; v[N - 1] = 0.0; /* vel_setbc_Tor (velocity.F90:279-303) */
            if (l == 1 && m_ == 0 && part == 0 && s->c.vel_bc <= 2) v[0] = s->omega * s->rp[4 * N];
            if (l != 0) orc_gbtrs(N, LUp(s, 0, l), PVp(s, 0, l), v);
            memcpy(x, v, sizeof(double) * N);
            /* velocity poloidal */
            x = part ? s->f[LSD_VEL_POL].im + o : s->f[LSD_VEL_POL].re + o;
            nn = part ? s->f[LSD_N_VEL_POL].im + o : s->f[LSD_N_VEL_POL].re + o;
            memcpy(v, nn, sizeof(double) * N);
            v[0] = 0.0; v[N - 1] = 0.0;
            if (l != 0) {
                orc_gbtrs(N, LUp(s, 2, l), PVp(s, 2, l), v);
                multY(s, YMp(s, 1, l), x, v, v);
                v[0] = 0.0; v[N - 1] = 0.0;
                orc_gbtrs(N, LUp(s, 1, l), PVp(s, 1, l), v);
                gre_adjust(s, l, v);
            }
            memcpy(x, v, sizeof(double) * N);
            /* codensity / composition (l = 0 included) */
            for (int f = 0; f < 2; f++) {
                int on = f == 0 ? s->c.cod : s->c.comp;
                if (!on) continue;
                int fid = f == 0 ? LSD_COD : LSD_COMP, nid = f == 0 ? LSD_N_COD : LSD_N_COMP;
                x = part ? s->f[fid].im + o : s->f[fid].re + o;
                nn = part ? s->f[nid].im + o : s->f[nid].re + o;
                multY(s, YMp(s, 2 + f, l), x, nn, v);
                v[0] = 0.0; v[N - 1] = 0.0; /* cod_setbc, i_inhombc_C = 1 */
                if (nh == 0 && part == 0) { v[0] = 1.0; v[N - 1] = 0.0; }
                orc_gbtrs(N, LUp(s, 3 + f, l), PVp(s, 3 + f, l), v);
                memcpy(x, v, sizeof(double) * N);
            }
            if (ic) {
                for (int k = 0; k < 2; k++) ic_predict(s, nh, part, k);
            } else if (s->c.mag) {
                /* magnetic toroidal: mag_setbc_Tor with zero BC arrays */
                x = part ? s->f[LSD_MAG_TOR].im + o : s->f[LSD_MAG_TOR].re + o;
                nn = part ? s->f[LSD_N_MAG_TOR].im + o : s->f[LSD_N_MAG_TOR].re + o;
                if (l == 0) memcpy(v, nn, sizeof(double) * N);
                else multY(s, YMp(s, 4, l), x, nn, v);
                v[0] = 0.0; v[N - 1] = 0.0;
                if (l != 0) orc_gbtrs(N, LUp(s, 5, l), PVp(s, 5, l), v);
                memcpy(x, v, sizeof(double) * N);
                /* magnetic poloidal: mag_setbc_Pol with zero BC arrays */
                x = part ? s->f[LSD_MAG_POL].im + o : s->f[LSD_MAG_POL].re + o;
                nn = part ? s->f[LSD_N_MAG_POL].im + o : s->f[LSD_N_MAG_POL].re + o;
                if (l == 0) memcpy(v, nn, sizeof(double) * N);
                else multY(s, YMp(s, 5, l), x, nn, v);
                v[0] = 0.0; v[N - 1] = 0.0;
                if (l != 0) orc_gbtrs(N, LUp(s, 6, l), PVp(s, 6, l), v);
                memcpy(x, v, sizeof(double) * N);
            }
        }
    }
    return 0;
}

/* vel_corrector, cod_corrector, comp_corrector, mag_corrector
 * (modules/velocity.F90:413-439, codensity.F90:248-259,
 * composition.F90:235-246, magnetic.F90:422-461) with tim_nlincorr and
 * tim_addcorr (timestep.F90:715-737).  *dterr_out receives this pass's sum
 * of the per-field correction norms; orc_step adds it to tim_dterr. */
int orc_corrector(orc_sim* s, double* dterr_out) {
    int N = s->N, H = s->H;
    size_t HN = (size_t)H * N;
    double cimp = s->c.implicit;
    /* rot_corrector (modules/rotation.F90:70-88) */
    if (s->c.rot) {
        s->omega = s->omega + (s->dt * s->c.implicit / s->inertia) *
                                  ((s->vtorq + s->mtorq) - (s->vtorq_s + s->mtorq_s));
        s->vtorq_s = s->vtorq;
        s->mtorq_s = s->mtorq;
    }
    const int ic = s->Nic > 1 && s->c.mag;
    const int Ni = s->Nic, NT = s->NT;
    coll_t iccorr[2];
    for (int k = 0; k < 2; k++) iccorr[k] = coll_new(ic ? H * Ni : 1);
    /* correction fields: 0 vT, 1 vP, 2 C, 3 X, 4 bT, 5 bP */
    coll_t corr[6];
    for (int k = 0; k < 6; k++) corr[k] = coll_new((int)HN);
    double* icT = (double*)calloc((size_t)H * 2, sizeof(double));
    double* icP = (double*)calloc((size_t)H * 2, sizeof(double));
    int nmap[6] = {LSD_N_VEL_TOR, LSD_N_VEL_POL, LSD_N_COD, LSD_N_COMP, LSD_N_MAG_TOR, LSD_N_MAG_POL};
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 16)
    for (int nh = 0; nh < H; nh++) {
        int l = s->hl[nh];
        size_t o = (size_t)nh * N;
        for (int part = 0; part < 2; part++) {
            for (int k = 0; k < 6; k++) {
                if ((k == 2 && !s->c.cod) || (k == 3 && !s->c.comp) || (k >= 4 && !s->c.mag)) continue;
                const double* nn = part ? s->f[nmap[k]].im + o : s->f[nmap[k]].re + o;
                const double* np_ = part ? s->hist[k].im + o : s->hist[k].re + o;
                double* v = part ? corr[k].im + o : corr[k].re + o;
                for (int n = 0; n < N; n++) v[n] = cimp * (nn[n] - np_[n]);
                v[0] = 0.0; v[N - 1] = 0.0;
                /* vel_setbc_Tor with tim_it /= 0 (velocity.F90:292-302):
                 * T_10 at r_i gets rot_omega r_i - vel_Tor there. */
                if (k == 0 && part == 0 && l == 1 && s->hm[nh] == 0 && s->c.vel_bc <= 2)
                    v[0] = s->omega * s->rp[4 * N] - s->f[LSD_VEL_TOR].re[o];
                if (ic && k >= 4) {
                    /* tim_nlincorr on the inner core, tim_zerobc, tim_icinvX */
                    const size_t oi = (size_t)nh * Ni;
                    const coll_t* icn = &s->icn[k - 4];
                    const coll_t* ich = &s->ichist[k - 4];
                    double* vi = part ? iccorr[k - 4].im + oi : iccorr[k - 4].re + oi;
                    for (int n = 0; n < Ni; n++)
                        vi[n] = cimp * ((part ? icn->im : icn->re)[oi + n] - (part ? ich->im : ich->re)[oi + n]);
                    vi[0] = 0.0; vi[Ni - 1] = 0.0;
                    if (l != 0) {
                        double b[2048];
                        for (int n = 0; n < Ni - 1; n++) b[n] = vi[n];
                        for (int n = 0; n < N; n++) b[Ni - 1 + n] = v[n];
                        orc_gbtrs(NT, s->iclu[k - 4] + (size_t)l * NT * LDAB, s->icpiv[k - 4] + l * NT, b);
                        for (int n = 0; n < Ni; n++) vi[n] = b[n];
                        for (int n = 0; n < N; n++) v[n] = b[Ni - 1 + n];
                    }
                    continue;
                }
                switch (k) {
                case 0: if (l != 0) orc_gbtrs(N, LUp(s, 0, l), PVp(s, 0, l), v); break;
                case 1:
                    if (l != 0) {
                        orc_gbtrs(N, LUp(s, 2, l), PVp(s, 2, l), v);
                        v[0] = 0.0; v[N - 1] = 0.0;
                        orc_gbtrs(N, LUp(s, 1, l), PVp(s, 1, l), v);
                        gre_adjust(s, l, v);
                    }
                    break;
                case 2: orc_gbtrs(N, LUp(s, 3, l), PVp(s, 3, l), v); break;
                case 3: orc_gbtrs(N, LUp(s, 4, l), PVp(s, 4, l), v); break;
                case 4:
                    if (l != 0) orc_gbtrs(N, LUp(s, 5, l), PVp(s, 5, l), v);
                    icT[nh * 2 + part] = l != 0 ? v[0] : 0.0;
                    break;
                case 5:
                    if (l != 0) orc_gbtrs(N, LUp(s, 6, l), PVp(s, 6, l), v);
                    icP[nh * 2 + part] = l != 0 ? v[0] : 0.0;
                    break;
                }
            }
        }
    }
    /* N := N* (var_coll_copy after tim_nlincorr) */
    for (int k = 0; k < 6; k++) {
        memcpy(s->hist[k].re, s->f[nmap[k]].re, sizeof(double) * HN);
        memcpy(s->hist[k].im, s->f[nmap[k]].im, sizeof(double) * HN);
    }
    double dterr = 0.0;
    dterr += addcorr(s, &corr[0], &s->f[LSD_VEL_TOR]);
    dterr += addcorr(s, &corr[1], &s->f[LSD_VEL_POL]);
    if (s->c.cod) dterr += addcorr(s, &corr[2], &s->f[LSD_COD]);
    if (s->c.comp) dterr += addcorr(s, &corr[3], &s->f[LSD_COMP]);
    if (ic) {
        for (int k = 0; k < 2; k++) {
            memcpy(s->ichist[k].re, s->icn[k].re, sizeof(double) * H * Ni);
            memcpy(s->ichist[k].im, s->icn[k].im, sizeof(double) * H * Ni);
        }
        dterr += addcorr(s, &corr[4], &s->f[LSD_MAG_TOR]);
        dterr += addcorr_n(&iccorr[0], &s->icf[0], (size_t)H * Ni);
        dterr += addcorr(s, &corr[5], &s->f[LSD_MAG_POL]);
        dterr += addcorr_n(&iccorr[1], &s->icf[1], (size_t)H * Ni);
    } else if (s->c.mag) {
        double e;
        dterr += addcorr(s, &corr[4], &s->f[LSD_MAG_TOR]);
        e = 0.0;
        for (int nh = 0; nh < H; nh++) { e = e + icT[2 * nh] * icT[2 * nh]; e = e + icT[2 * nh + 1] * icT[2 * nh + 1]; }
        dterr += sqrt(e);
        dterr += addcorr(s, &corr[5], &s->f[LSD_MAG_POL]);
        e = 0.0;
        for (int nh = 0; nh < H; nh++) { e = e + icP[2 * nh] * icP[2 * nh]; e = e + icP[2 * nh + 1] * icP[2 * nh + 1]; }
        dterr += sqrt(e);
    }
    for (int k = 0; k < 6; k++) coll_free(&corr[k]);
    for (int k = 0; k < 2; k++) coll_free(&iccorr[k]);
    free(icT); free(icP);
    if (dterr_out) *dterr_out = dterr;
    return 0;
}

/* tim_check_cgce (modules/timestep.F90:744-830): control 2 iterates while
 * tim_dterr > d_dterr with tim_corr_dt = dt sqrt(d_dterr / tim_dterr),
 * control 1 while tim_dterr > 1 with tim_corr_dt = tim_dterr; returns 1 if
 * the run must stop. */
static int check_cgce(orc_sim* s) {
    int stop = 0;
    const int ctl1 = s->c.step_control == 1;
    const double crit = ctl1 ? 1.0 : s->c.dterr;
    if (s->it == 1) {
        s->corr_dt = ctl1 ? s->dterr : s->dt * sqrt(s->c.dterr / s->dterr);
        s->lasterr = 1e99;
    }
    if (s->dt < 1e-12 && s->step > 30) { stop = 1; s->it = 0; }
    else if (s->it > 10) { stop = 1; s->it = 0; }
    else if (s->dterr > s->lasterr) {
        if (ctl1) {
            if (s->dterr > 1.0) stop = 1;
        } else {
            if (s->dterr > 2.0 * s->c.dterr) stop = 1;
            if (s->dterr < 2.0 * s->c.dterr) s->corr_dt = s->dt / (1e1 * COURANT);
        }
        s->it = 0;
    } else if (s->dterr > crit) {
        s->lasterr = s->dterr;
        s->it = s->it + 1;
    } else {
        s->it = 0;
    }
    s->dterr = 0.0;
    return stop;
}

/* non_maxtstep (modules/nonlinear.F90:545-615): the CFL limit of the
 * physical velocity and magnetic fields of the last transform, shell by
 * shell: radially dr / max(Br^2 / sqrt(Br^2 E/Pm + d) + |ur|) with
 * d = ((E + E/Pm) / (2 dr))^2 and dr the smaller neighbour spacing, and
 * horizontally dr_h / max(Bh^2 / sqrt(Bh^2 E/Pm + d_h) + |uh|) with
 * dr_h = r / sqrt(L (L-1)) and d_h = (E / (2 dr_h))^2; then capped by E/Pm
 * and sqrt(E). */
static double maxtstep(const orc_sim* s) {
    const int N = s->N, Th = s->Th, Ph = s->Ph;
    const double E = s->c.E, Pm = s->c.Pm;
    const double* r = s->rp + 4 * N; /* r^1 */
    const size_t plane = (size_t)Th * Ph;
    const double* zero = NULL;
    double cfl = 1e99;
    for (int n = 0; n < N; n++) {
        const double* ur = s->phys[LSD_P_VEL_R] + n * plane;
        const double* ut = s->phys[LSD_P_VEL_T] + n * plane;
        const double* up = s->phys[LSD_P_VEL_P] + n * plane;
        const double* br = s->c.mag ? s->phys[LSD_P_MAG_R] + n * plane : zero;
        const double* bt = s->c.mag ? s->phys[LSD_P_MAG_T] + n * plane : zero;
        const double* bp = s->c.mag ? s->phys[LSD_P_MAG_P] + n * plane : zero;
        double dr;
        if (n == 0) dr = r[1] - r[0];
        else if (n == N - 1) dr = r[N - 1] - r[N - 2];
        else dr = fmin(r[n] - r[n - 1], r[n + 1] - r[n]);
        double d = (E + E / Pm) / (2.0 * dr);
        d = d * d;
        double mx = -1e300;
        for (size_t i = 0; i < plane; i++) {
            const double b = br ? br[i] : 0.0;
            const double p = b * b;
            const double v = p / sqrt(p * (E / Pm) + d) + fabs(ur[i]);
            if (v > mx) mx = v;
        }
        if (mx != 0.0) cfl = fmin(cfl, dr / mx);
        dr = r[n] / sqrt((double)(s->L * (s->L - 1)));
        d = E / (2.0 * dr);
        d = d * d;
        mx = -1e300;
        for (size_t i = 0; i < plane; i++) {
            const double b1 = bt ? bt[i] : 0.0, b2 = bp ? bp[i] : 0.0;
            const double p = b1 * b1 + b2 * b2;
            const double v = p / sqrt(p * E / Pm + d) + sqrt(ut[i] * ut[i] + up[i] * up[i]);
            if (v > mx) mx = v;
        }
        if (mx != 0.0) cfl = fmin(cfl, dr / mx);
    }
    cfl = fmin(cfl, E / Pm);
    cfl = fmin(cfl, sqrt(E));
    return cfl;
}

/* tim_new_tstep (modules/timestep.F90:837-870); returns 1 if dt changed
 * (tim_new_dt).  d_corr_upper = 0.02, d_corr_lower = 0.002
 * (Scripts/parameters.F90_template:242-243). */
static int new_tstep(orc_sim* s) {
    double dt = s->dt;
    if (s->c.step_control == 1) {
        if (s->corr_dt > 0.020) dt = 0.5 * dt;
        if (s->corr_dt < 0.00200) dt = 1.273974 * dt;
        if (s->step == 0) dt = fmin(dt / 1e3, 1e-8);
    } else {
        if (s->corr_dt > 0.0) dt = fmin(s->corr_dt, dt * 2.0 / COURANT);
        if (s->cfl_dt > 0.0) dt = fmin(s->cfl_dt, dt);
        if (s->step == 0) dt = fmin(dt / 1e3, 1e-10);
        dt = dt * COURANT;
    }
    const int changed = dt < s->dt * 0.97 || dt > s->dt * 1.03;
    if (changed) s->dt = dt;
    return changed;
}

int orc_get_dt(orc_sim* s, double* dt, double* cfl_dt, double* corr_dt) {
    if (dt) *dt = s->dt;
    if (cfl_dt) *cfl_dt = s->cfl_dt;
    if (corr_dt) *corr_dt = s->corr_dt;
    return 0;
}

/* program/main.F90:100-164 without I/O; with d_timestep < 0 the step is
 * chosen at the start of each step by non_maxtstep and tim_new_tstep and
 * the operators are rebuilt when it changes (main.F90:124-135). */
int orc_step(orc_sim* s, int nsteps, lsd_step_stats* st) {
    lsd_step_stats z;
    memset(&z, 0, sizeof z);
    for (int k = 0; k < nsteps; k++) {
        int its = 0;
        double last = 0.0;
        do {
            orc_transform(s);
            orc_nonlinear(s);
            z.evaluations++;
            if (s->it == 0) {
                if (s->adaptive) {
                    s->cfl_dt = maxtstep(s);
                    if (new_tstep(s)) {
                        int rc = matrices(s);
                        if (rc) return rc;
                    }
                }
                orc_predictor(s);
                s->it = 1;
            } else {
                double d = 0.0;
                orc_corrector(s, &d);
                s->dterr = s->dterr + d;
                last = s->dterr;
                its++;
                if (check_cgce(s)) { set_err("tim_check_cgce: run stopped"); return LSD_E_DIVERGED; }
            }
        } while (s->it != 0);
        s->t = s->t + s->dt;
        s->step++;
        z.steps++;
        z.last_its = its;
        if (its > z.max_its) z.max_its = its;
        z.last_dterr = last;
    }
    z.t = s->t;
    if (st) *st = z;
    return 0;
}

/* var_coll_norm (modules/variables.F90:903-957). */
int orc_coll_norm(orc_sim* s, int field, double* out) {
    int N = s->N;
    double* Fod = (double*)calloc(N * 4, sizeof(double));
    double *Fev = Fod + N, *Fm0le = Fod + 2 * N, *Fm0 = Fod + 3 * N;
    const coll_t* a = &s->f[field];
    for (int nh = 0; nh < s->H; nh++) {
        int l = s->hl[nh], m_ = s->hm[nh];
        double w = 4.0 * D_PI / (double)(2 * l + 1);
        if (m_ != 0) w = 4.0 * w;
        for (int n = 0; n < N; n++) {
            size_t q = (size_t)nh * N + n;
            double f = w * (a->re[q] * a->re[q] + a->im[q] * a->im[q]);
            if (m_ == 0) Fm0[n] += f;
            if (m_ == 0 && l % 2 == 0) Fm0le[n] += f;
            if ((l - m_) % 2 == 1) Fod[n] += f; else Fev[n] += f;
        }
    }
    for (int k = 0; k < 4; k++) {
        double e = 0.0;
        for (int n = 0; n < N; n++) e += Fod[k * N + n] * s->intr2dr[n];
        out[k] = e;
    }
    free(Fod);
    return 0;
}

/* Bulk energies and their l- and m-spectra, as io_write_vel_energy /
 * io_write_mag_energy assemble them (program/io.F90:1799-1884): the
 * field in qst form (var_coll_TorPol2qst, Boussinesq: compr1dr = r1dr,
 * modules/variables.F90:511-556), E = Efac * sum over q, s, t of the
 * var_coll_norm integrals (variables.F90:901-957), El/Em the
 * var_coll_normlm parts (variables.F90:1267-1313) times Efac.
 * Efac = 1/2 for the velocity, Pm/E for the magnetic field in the
 * Boussinesq scaling (program/io.F90:238-247).  which: 0 velocity, 1 magnetic. */
int orc_energy(orc_sim* s, int which, double* E, double* El, double* Em) {
    int N = s->N, H = s->H;
    const coll_t* T = &s->f[which ? LSD_MAG_TOR : LSD_VEL_TOR];
    const coll_t* P = &s->f[which ? LSD_MAG_POL : LSD_VEL_POL];
    double efac = which ? s->c.Pm / s->c.E : 0.5;
    coll_t q = coll_new(H * N), sf = coll_new(H * N), t = coll_new(H * N);
    torpol2qst(s, T, P, &q, &sf, &t);
    const coll_t* parts[3] = {&q, &sf, &t};
    double* F = (double*)calloc((size_t)N, sizeof(double));
    for (int l = 0; l < s->c.L; l++) El[l] = 0.0;
    for (int m = 0; m < s->c.M; m++) Em[m] = 0.0;
    double tot = 0.0;
    for (int k = 0; k < 3; k++) {
        const coll_t* a = parts[k];
        for (int n = 0; n < N; n++) F[n] = 0.0;
        for (int nh = 0; nh < H; nh++) {
            int l = s->hl[nh], m_ = s->hm[nh];
            double w = 4.0 * D_PI / (double)(2 * l + 1);
            if (m_ != 0) w = 4.0 * w;
            double e = 0.0;
            for (int n = 0; n < N; n++) {
                size_t o = (size_t)nh * N + n;
                double f = w * (a->re[o] * a->re[o] + a->im[o] * a->im[o]);
                F[n] += f;
                e += f * s->intr2dr[n];
            }
            El[l] += efac * e;
            Em[m_ / s->c.Mp] += efac * e;
        }
        double ek = 0.0;
        for (int n = 0; n < N; n++) ek += F[n] * s->intr2dr[n];
        tot += ek;
    }
    *E = efac * tot;
    free(F);
    coll_free(&q); coll_free(&sf); coll_free(&t);
    return 0;
}
