// FP64 set-up kernels: Gauss-Legendre quadrature, Schmidt-normalised
// P_l^m tables, radial finite-difference weights and the banded IMEX
// operators with their LU factors.  Everything runs on the device; the host
// only supplies index bookkeeping (harmonic order, packed-index map).
#include "lsd_internal.h"

// Set-up arithmetic keeps the reference's rounding steps: no fused
// multiply-adds in this file.
#pragma clang fp contract(off)

namespace lsd {

// ---------------------------------------------------------------- Gauss

// auxiliary/gauwts.f:6-75 -- Newton iteration on the three-term recurrence,
// one work-item per root pair (x1 = -1, x2 = 1, so xm = 0, xl = 1).
__global__ void k_gauss(int n, double* __restrict__ gp, double* __restrict__ gw,
                        double* __restrict__ sn) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int m = (n + 1) / 2;
    if (i > m) return;
    const double EPS = 1.0e-13, PI = 3.14159265358979312;
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
    double xlo = 0.0 - 1.0 * z, xhi = 0.0 + 1.0 * z;
    gp[i - 1] = xlo;
    gp[n - i] = xhi;
    double w = 2.0 * 1.0 / ((1.0 - z * z) * pp * pp);
    gw[i - 1] = w;
    gw[n - i] = w;
    sn[i - 1] = sin(acos(xlo));
    sn[n - i] = sin(acos(xhi));
}

// ------------------------------------------------------- Schmidt P_l^m

// Sectoral seed and its theta derivative (auxiliary/pmm.f:20-46,
// auxiliary/dpmm.f:11-40).
__device__ inline double pmm_d(int m, double s) {
    if (m == 0) return 1.0;
    double p = sqrt(2.0);
    for (int i = 1; i <= m; i++) {
        double twoi = 2.0 * (double)i;
        p = p * s * sqrt((twoi - 1.0) / twoi);
    }
    return p;
}
__device__ inline double dpmm_d(int m, double c, double s) {
    if (m == 0) return 0.0;
    double p = sqrt(2.0) * (double)m * c / s;
    for (int i = 1; i <= m; i++) {
        double twoi = 2.0 * (double)i;
        p = p * s * sqrt((twoi - 1.0) / twoi);
    }
    return p;
}

// One work-item per (theta point j < Th2, order index mi): the recursion in
// l of auxiliary/schnla.f:44-115 (pmm1.f, plm.f, dpmm1.f, dplm.f), written
// straight into the folded-order tables of modules/legendre.F90:84-125.
//   syn tables [t][h][j]:  P, P_, sin1mP, llP_, llsin1mP
//   ana tables [t][j][apos(h)]:  gwP, gwllP_, gwllsin1mP  (apos groups each
//   slot's even-k harmonics before its odd-k ones)
__global__ void k_schmidt(int L, int M, int Mp, int H, int Th2, const double* __restrict__ gp,
                          const double* __restrict__ gw, const double* __restrict__ sn,
                          const int* __restrict__ pidx, const int* __restrict__ apos,
                          double* __restrict__ syn, double* __restrict__ ana) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int mi = blockIdx.y;
    if (j >= Th2 || mi >= M) return;
    int m = mi * Mp;
    if (m > L - 1) return;
    double x = gp[j];
    double s = sqrt((1.0 + x) * (1.0 - x));
    double sin1 = 1.0 / sn[j];
    size_t HT = (size_t)H * Th2;
    double p2 = 0, dp2 = 0, p1 = 0, dp1 = 0;
    for (int l = m; l <= L - 1; l++) {
        double p, dp;
        if (l == m) {
            p = pmm_d(m, s);
            dp = dpmm_d(m, x, s);
        } else if (l == m + 1) {
            double rm = m;
            double sq = sqrt(2.0 * rm + 1.0);
            p = x * p1 * sq;
            dp = sq * (x * dp1 - s * p1);
        } else {
            double rm = m, rl = l;
            double a = sqrt((rl + rm - 1.0) * (rl - rm - 1.0));
            double b = sqrt((rl + rm) * (rl - rm));
            double t = (2.0 * rl - 1.0) * x * p1;
            t = t - p2 * a;
            p = t / b;
            double u = (2.0 * rl - 1.0) * (x * dp1 - s * p1);
            u = u - dp2 * a;
            dp = u / b;
        }
        p2 = p1;
        dp2 = dp1;
        p1 = p;
        dp1 = dp;
        int h = pidx[l * (l + 1) / 2 + m];
        if (h < 0) continue;
        size_t q = (size_t)h * Th2 + j;
        syn[0 * HT + q] = p;
        syn[1 * HT + q] = dp;
        syn[2 * HT + q] = p * (double)m * sin1;
        double llp = 0.0, llsm = 0.0;
        if (l != 0) {
            double w = 1.0 / sqrt((double)(l * (l + 1)));
            llp = dp * w;
            llsm = p * w * (double)m * sin1;
        }
        syn[3 * HT + q] = llp;
        syn[4 * HT + q] = llsm;
        // quadrature-weighted analysis tables (legendre.F90:108-125)
        double w = gw[j] * ((2.0 * l + 1.0) / 4.0);
        if (m == 0) w = 2.0 * w;
        size_t qa = (size_t)j * H + apos[h];
        ana[0 * HT + qa] = p * w;
        if (l == 0) {
            ana[1 * HT + qa] = 0.0;
            ana[2 * HT + qa] = 0.0;
        } else {
            w = w / sqrt((double)(l * (l + 1)));
            ana[1 * HT + qa] = dp * w;
            ana[2 * HT + qa] = p * w * (double)m * sin1;
        }
    }
}

int precompute_tables(Ctx& c) {
    int th = c.Th;
    hipLaunchKernelGGL(k_gauss, dim3((th + 127) / 128), dim3(64), 0, c.stream, th, c.d_gp, c.d_gw,
                       c.d_sin);
    LSD_HIP(hipGetLastError());
    dim3 grid((c.Th2 + 63) / 64, c.M);
    hipLaunchKernelGGL(k_schmidt, grid, dim3(64), 0, c.stream, c.L, c.M, c.Mp, c.H, c.Th2, c.d_gp,
                       c.d_gw, c.d_sin, c.d_pidx, c.d_apos, c.d_tab, c.d_atab);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

// ---------------------------------------------------------------- mesh

// libgfortran's x**k for a runtime integer exponent (pow_r8_i4), as used by
// D%r(:,j) = D%r(:,1)**j (modules/meshs.F90:262-265).
__device__ inline double ipow_d(double x, int k) {
    double pw = 1.0;
    unsigned u;
    if (k == 0) return 1.0;
    if (k < 0) {
        u = (unsigned)(-k);
        x = pw / x;
    } else {
        u = (unsigned)k;
    }
    for (;;) {
        if (u & 1) pw = pw * x;
        u >>= 1;
        if (u) x = x * x;
        else break;
    }
    return pw;
}

// Radial Chebyshev grid (modules/meshs.F90:135-143, d_alpha < 0 branch)
// and its powers r^p, p = -3..3, stored r[(p+3)*N + n].
__global__ void k_grid(int N, double rratio, double* __restrict__ r) {
    int n = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (n > N) return;
    const double PI = 3.1415926535897931;
    double ri = rratio / (1.0 - rratio);
    double r1 = 0.5 * (1.0 + cos(PI * (N - n) / (double)(N - 1))) + ri;
    for (int p = -3; p <= 3; p++) r[(p + 3) * N + n - 1] = (p == 1) ? r1 : ipow_d(r1, p);
}

// dgetf2 + dgetri (unblocked) on an n x n column-major matrix, lda = 7:
// the inverse used by mes_mat_invert (modules/meshs.F90:427-440).
__device__ int invert7(int n, double* a) {
#define A_(i, j) a[((j) - 1) * 7 + ((i) - 1)]
    int ipiv[7];
    for (int j = 1; j <= n; j++) {
        int jp = j;
        double amax = fabs(A_(j, j));
        for (int i = j + 1; i <= n; i++)
            if (fabs(A_(i, j)) > amax) {
                amax = fabs(A_(i, j));
                jp = i;
            }
        ipiv[j - 1] = jp;
        if (A_(jp, j) == 0.0) return j;
        if (jp != j)
            for (int cc = 1; cc <= n; cc++) {
                double t = A_(j, cc);
                A_(j, cc) = A_(jp, cc);
                A_(jp, cc) = t;
            }
        if (j < n) {
            double rp = 1.0 / A_(j, j);
            for (int i = j + 1; i <= n; i++) A_(i, j) = A_(i, j) * rp;
            for (int cc = j + 1; cc <= n; cc++) {
                double y = A_(j, cc);
                if (y != 0.0)
                    for (int i = j + 1; i <= n; i++) A_(i, cc) = A_(i, cc) + A_(i, j) * (-y);
            }
        }
    }
    for (int j = 1; j <= n; j++) {  // inverse of U (dtrti2)
        A_(j, j) = 1.0 / A_(j, j);
        double ajj = -A_(j, j);
        for (int jj = 1; jj <= j - 1; jj++) {
            double x = A_(jj, j);
            if (x != 0.0) {
                for (int i = 1; i <= jj - 1; i++) A_(i, j) = A_(i, j) + x * A_(i, jj);
                A_(jj, j) = x * A_(jj, jj);
            }
        }
        for (int i = 1; i <= j - 1; i++) A_(i, j) = A_(i, j) * ajj;
    }
    double work[7];
    for (int j = n; j >= 1; j--) {  // solve inv(A) L = inv(U)
        for (int i = j + 1; i <= n; i++) {
            work[i - 1] = A_(i, j);
            A_(i, j) = 0.0;
        }
        for (int jj = j + 1; jj <= n; jj++) {
            double x = work[jj - 1];
            if (x != 0.0) {
                double t = -1.0 * x;
                for (int i = 1; i <= n; i++) A_(i, j) = A_(i, j) + t * A_(i, jj);
            }
        }
    }
    for (int j = n - 1; j >= 1; j--) {
        int jp = ipiv[j - 1];
        if (jp != j)
            for (int i = 1; i <= n; i++) {
                double t = A_(i, j);
                A_(i, j) = A_(i, jp);
                A_(i, jp) = t;
            }
    }
    return 0;
#undef A_
}

// One work-item per radial point n: local Taylor matrix, its inverse, and
// the derivative weights dr(1), dr(2) (modules/meshs.F90:292-309).  The
// inverse is kept in ainv[n][7][7] for the integration weights.
__global__ void k_fd(int N, const double* __restrict__ r, double* __restrict__ dr1,
                     double* __restrict__ dr2, double* __restrict__ ainv, int* __restrict__ info) {
    int n = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (n > N) return;
    const double* r1 = r + 4 * N;
    int l = n - LSD_KL > 1 ? n - LSD_KL : 1;
    int rr = n + LSD_KL < N ? n + LSD_KL : N;
    int nn = rr - l + 1;
    double A[49];
    for (int i = 0; i < 49; i++) A[i] = 0.0;
    for (int i = 1; i <= nn; i++) A[i - 1] = 1.0;
    for (int j = 2; j <= nn; j++)
        for (int i = 1; i <= nn; i++)
            A[(j - 1) * 7 + (i - 1)] = A[(j - 2) * 7 + (i - 1)] * (r1[l + i - 2] - r1[n - 1]) / (double)(j - 1);
    if (invert7(nn, A)) *info = n;
    for (int q = 0; q < 49; q++) ainv[(size_t)(n - 1) * 49 + q] = A[q];
    // band column j holds M(KL+1+n-j, j) -> dr[(j-1)*7 + KL+n-j]
    for (int j = l; j <= rr; j++) {
        dr1[(j - 1) * 7 + (LSD_KL + n - j)] = A[(j - l) * 7 + 1];
        dr2[(j - 1) * 7 + (LSD_KL + n - j)] = A[(j - l) * 7 + 2];
    }
}

// radLap = dr2 + dr1*2/r, r1dr = dr1 + 1/r, and the r^2 dr integration
// weights (modules/meshs.F90:311-347, 408-409).  One work-item per column j.
__global__ void k_mesh_ops(int N, const double* __restrict__ r, const double* __restrict__ dr1,
                           const double* __restrict__ dr2, const double* __restrict__ ainv,
                           double* __restrict__ radlap, double* __restrict__ r1dr,
                           double* __restrict__ intr2dr) {
    int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (j > N) return;
    const double* rm1 = r + 2 * N;
    const double* r1 = r + 4 * N;
    const double* r2 = r + 5 * N;
    int l = j - LSD_KL > 1 ? j - LSD_KL : 1;
    int rr = j + LSD_KL < N ? j + LSD_KL : N;
    for (int q = 0; q < 7; q++) {
        radlap[(j - 1) * 7 + q] = 0.0;
        r1dr[(j - 1) * 7 + q] = dr1[(j - 1) * 7 + q];
    }
    for (int n = l; n <= rr; n++) {
        int q = (j - 1) * 7 + (LSD_KL + n - j);
        radlap[q] = dr2[q] + dr1[q] * 2.0 * rm1[n - 1];
    }
    r1dr[(j - 1) * 7 + LSD_KL] = r1dr[(j - 1) * 7 + LSD_KL] + rm1[j - 1];
    // intr2dr(j): contributions of every point n whose stencil covers j,
    // n ascending then i ascending, as the reference accumulates them.
    double acc = 0.0;
    for (int n = l; n <= rr; n++) {
        int ln = n - LSD_KL > 1 ? n - LSD_KL : 1;
        int rn = n + LSD_KL < N ? n + LSD_KL : N;
        int nn = rn - ln + 1;
        int np = n + 1 < N ? n + 1 : N, nm = n - 1 > 1 ? n - 1 : 1;
        double cc = 1.0, e = 1.0;
        const double* A = ainv + (size_t)(n - 1) * 49;
        for (int i = 1; i <= nn; i++) {
            cc = cc * (r1[np - 1] - r1[n - 1]) / (double)i;
            e = e * (r1[nm - 1] - r1[n - 1]) / (double)i;
            double w = 0.5 * (cc - e) * A[(j - ln) * 7 + (i - 1)] * r2[n - 1];
            acc = acc + w;
        }
    }
    intr2dr[j - 1] = acc;
}

// ll1, 1/ll1, sqrt(ll1), 1/sqrt(ll1) per degree (modules/variables.F90:101-115).
__global__ void k_ll(int L, double* __restrict__ ll) {
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L) return;
    double a = 0, b = 0, c = 0, d = 0;
    if (l > 0) {
        a = (double)(l * (l + 1));
        b = 1.0 / a;
        c = sqrt(a);
        d = 1.0 / c;
    }
    ll[0 * L + l] = a;
    ll[1 * L + l] = b;
    ll[2 * L + l] = c;
    ll[3 * L + l] = d;
}

int precompute_mesh(Ctx& c) {
    int N = c.N;
    hipLaunchKernelGGL(k_grid, dim3((N + 63) / 64), dim3(64), 0, c.stream, N, c.cfg.rratio, c.d_r);
    double* ainv = nullptr;
    int* info = nullptr;
    AsyncScratch s_ainv(c.stream), s_info(c.stream);
    LSD_HIP(s_ainv.alloc(&ainv, sizeof(double) * 49 * N));
    LSD_HIP(s_info.alloc(&info, sizeof(int)));
    LSD_HIP(hipMemsetAsync(info, 0, sizeof(int), c.stream));
    LSD_HIP(hipMemsetAsync(c.d_dr1, 0, sizeof(double) * 7 * N, c.stream));
    LSD_HIP(hipMemsetAsync(c.d_dr2, 0, sizeof(double) * 7 * N, c.stream));
    hipLaunchKernelGGL(k_fd, dim3((N + 63) / 64), dim3(64), 0, c.stream, N, c.d_r, c.d_dr1, c.d_dr2, ainv,
                       info);
    hipLaunchKernelGGL(k_mesh_ops, dim3((N + 63) / 64), dim3(64), 0, c.stream, N, c.d_r, c.d_dr1, c.d_dr2,
                       ainv, c.d_radlap, c.d_r1dr, c.d_intr2dr);
    hipLaunchKernelGGL(k_ll, dim3((c.L + 63) / 64), dim3(64), 0, c.stream, c.L, c.d_ll);
    LSD_HIP(hipGetLastError());
    int h_info = 0;
    LSD_HIP(hipMemcpyAsync(&h_info, info, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    LSD_HIP(hipStreamSynchronize(c.stream));
    if (h_info) {
        set_error("matrix inversion error 1");
        return LSD_E_SINGULAR;
    }
    return LSD_OK;
}

// ------------------------------------------------------------ operators

// LAPACK dgbtf2 (KL = KU = 3) on one band ab(10, n), column-major: the
// factorisation of mes_lu_find (modules/meshs.F90:447-456).
__device__ int gbtrf_d(int n, double* ab, int* ipiv) {
#define AB(i, j) ab[((j) - 1) * LSD_LDAB + ((i) - 1)]
    const int ku = LSD_KL, kl = LSD_KL, kv = ku + kl;
    int info = 0;
    for (int j = ku + 2; j <= (kv < n ? kv : n); j++)
        for (int i = kv - j + 2; i <= kl; i++) AB(i, j) = 0.0;
    int ju = 1;
    for (int j = 1; j <= n; j++) {
        if (j + kv <= n)
            for (int i = 1; i <= kl; i++) AB(i, j + kv) = 0.0;
        int km = kl < n - j ? kl : n - j;
        int jp = 1;
        double amax = fabs(AB(kv + 1, j));
        for (int i = 2; i <= km + 1; i++) {
            double a = fabs(AB(kv + i, j));
            if (a > amax) {
                amax = a;
                jp = i;
            }
        }
        ipiv[j - 1] = jp + j - 1;
        if (AB(kv + jp, j) != 0.0) {
            int t = j + ku + jp - 1;
            if (t > n) t = n;
            if (t > ju) ju = t;
            if (jp != 1) {
                for (int cc = 0; cc <= ju - j; cc++) {
                    double* a = &ab[(j - 1 + cc) * LSD_LDAB + (kv + jp - 1 - cc)];
                    double* b = &ab[(j - 1 + cc) * LSD_LDAB + (kv - cc)];
                    double tmp = *a;
                    *a = *b;
                    *b = tmp;
                }
            }
            if (km > 0) {
                double rp = 1.0 / AB(kv + 1, j);
                for (int i = 2; i <= km + 1; i++) AB(kv + i, j) = AB(kv + i, j) * rp;
                for (int cc = 1; cc <= ju - j; cc++) {
                    double y = ab[(j - 1 + cc) * LSD_LDAB + (kv - cc)];
                    if (y != 0.0) {
                        double tmp = -1.0 * y;
                        for (int i = 1; i <= km; i++) {
                            double* d = &ab[(j - 1 + cc) * LSD_LDAB + (kv - cc + i)];
                            *d = *d + AB(kv + 1 + i, j) * tmp;
                        }
                    }
                }
            }
        } else if (info == 0) {
            info = j;
        }
    }
    return info;
#undef AB
}

// dgbtrs('N', ...) with one right-hand side (modules/timestep.F90:566-588).
__device__ void gbtrs_d(int n, const double* __restrict__ ab, const int* __restrict__ ipiv, double* b) {
#define AB(i, j) ab[((j) - 1) * LSD_LDAB + ((i) - 1)]
    const int kl = LSD_KL, kd = 2 * LSD_KL + 1;
    for (int j = 1; j <= n - 1; j++) {
        int lm = kl < n - j ? kl : n - j;
        int l = ipiv[j - 1];
        if (l != j) {
            double t = b[l - 1];
            b[l - 1] = b[j - 1];
            b[j - 1] = t;
        }
        double y = b[j - 1];
        if (y != 0.0) {
            double tmp = -1.0 * y;
            for (int i = 1; i <= lm; i++) b[j - 1 + i] = b[j - 1 + i] + AB(kd + i, j) * tmp;
        }
    }
    const int k = 2 * LSD_KL;
    for (int j = n; j >= 1; j--) {
        if (b[j - 1] != 0.0) {
            int l = k + 1 - j;
            b[j - 1] = b[j - 1] / AB(k + 1, j);
            double tmp = b[j - 1];
            int i0 = j - k > 1 ? j - k : 1;
            for (int i = j - 1; i >= i0; i--) b[i - 1] = b[i - 1] - tmp * AB(l + i, j);
        }
    }
#undef AB
}

struct OpParams {
    double dt, implicit, Pm, Pr, Sc;
    int vel_bc;  // i_vel_bc: 1 no-slip, 2 stress-free outer, 3 stress-free inner, 4 both stress-free
};

// One work-item per (operator kind, degree l): X = c1/dt - c2 c radLap_l
// with boundary rows (modules/timestep.F90:138-211, 218-344;
// velocity.F90:177-273; codensity.F90:85-119; composition.F90:82-105;
// magnetic.F90:137-282), factorised in place, plus the explicit operator
// Y = c1/dt + c2 (1-c) radLap_l (timestep.F90:421-473).
__global__ void k_operators(int L, int N, OpParams op, const double* __restrict__ r,
                            const double* __restrict__ radlap, const double* __restrict__ dr1,
                            const double* __restrict__ dr2, double* __restrict__ lu, int* __restrict__ piv, double* __restrict__ ym,
                            int* __restrict__ info, double* __restrict__ xband) {
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    int kind = blockIdx.y;
    if (l >= L) return;
    bool need = (kind == 3 || kind == 4) ? true : (l >= 1);
    double* ab = lu + ((size_t)kind * L + l) * N * LSD_LDAB;
    int* pv = piv + ((size_t)kind * L + l) * N;
    if (!need) {
        info[kind * L + l] = 0;
        return;
    }
    double c1, c2;
    int diagform = 0;  // 0: (A+c1_)+c2_*(ll1*r^-2)   1: (A+c1_)+(c2_*ll1)*r^-2
    switch (kind) {
    case 0:
    case 1: c1 = 1.0; c2 = op.Pm; break;
    case 2: c1 = 0.0; c2 = 1.0 / op.implicit; break;
    case 3: c1 = op.Pr / op.Pm; c2 = 1.0; break;
    case 4: c1 = op.Sc / op.Pm; c2 = 1.0; break;
    default: c1 = 1.0; c2 = 1.0; diagform = 1; break;
    }
    double c1_ = c1 / op.dt;
    double c2_ = c2 * op.implicit;
    if (op.implicit == 0.0) c2_ = 1.0;
    double ll1 = (double)(l * (l + 1));
    const double* rm2 = r + 1 * N;
    const double* rm1 = r + 2 * N;
    for (int j = 0; j < N; j++) {
        for (int q = 0; q < LSD_KL; q++) ab[j * LSD_LDAB + q] = 0.0;
        for (int q = 0; q < 7; q++) ab[j * LSD_LDAB + LSD_KL + q] = -c2_ * radlap[j * 7 + q];
        double* d = &ab[j * LSD_LDAB + 2 * LSD_KL];
        if (diagform == 0) *d = (*d + c1_) + c2_ * (ll1 * rm2[j]);
        else *d = (*d + c1_) + c2_ * ll1 * rm2[j];
    }
#define LUA(i, j) ab[((j) - 1) * LSD_LDAB + (2 * LSD_KL + (i) - (j))]
#define DR1(i, j) dr1[((j) - 1) * 7 + (LSD_KL + (i) - (j))]
#define DR2(i, j) dr2[((j) - 1) * 7 + (LSD_KL + (i) - (j))]
    for (int j = 1; j <= 1 + LSD_KL; j++) LUA(1, j) = 0.0;
    for (int j = N - LSD_KL; j <= N; j++) LUA(N, j) = 0.0;
    // no-slip walls: inner for i_vel_bc 1, 2; outer for 1, 3
    const bool ns_in = op.vel_bc == 1 || op.vel_bc == 2, ns_out = op.vel_bc == 1 || op.vel_bc == 3;
    if (kind == 0) {  // vel_bc_Tor: no-slip T = 0, stress-free (dr - 1/r) T = 0
        if (ns_in) {
            LUA(1, 1) = 1.0;
        } else {
            for (int j = 1; j <= 1 + LSD_KL; j++) LUA(1, j) = DR1(1, j);
            LUA(1, 1) = LUA(1, 1) - rm1[0];
        }
        if (ns_out) {
            LUA(N, N) = 1.0;
        } else {
            for (int j = N - LSD_KL; j <= N; j++) LUA(N, j) = DR1(N, j);
            LUA(N, N) = LUA(N, N) - rm1[N - 1];
        }
    } else if (kind == 1) {  // vel_bc_Pol: no-slip dP/dr = 0, stress-free d2P/dr2 = 0
        for (int j = 1; j <= 1 + LSD_KL; j++) LUA(1, j) = ns_in ? DR1(1, j) : DR2(1, j);
        for (int j = N - LSD_KL; j <= N; j++) LUA(N, j) = ns_out ? DR1(N, j) : DR2(N, j);
    } else if (kind == 6) {  // mag_bc_Pol: (dr - l/r) at r_i, (dr + (l+1)/r) at r_o
        for (int j = 1; j <= 1 + LSD_KL; j++) LUA(1, j) = DR1(1, j);
        LUA(1, 1) = LUA(1, 1) - rm1[0] * l;
        for (int j = N - LSD_KL; j <= N; j++) LUA(N, j) = DR1(N, j);
        LUA(N, N) = LUA(N, N) + rm1[N - 1] * (l + 1);
    } else {  // Dirichlet rows
        LUA(1, 1) = 1.0;
        LUA(N, N) = 1.0;
    }
#undef LUA
#undef DR1
#undef DR2
    // keep X itself (7 bands per column) for residuals of the mixed forms
    double* xb = xband + ((size_t)kind * L + l) * N * 7;
    for (int j = 0; j < N; j++)
        for (int q = 0; q < 7; q++) xb[j * 7 + q] = ab[j * LSD_LDAB + LSD_KL + q];
    info[kind * L + l] = gbtrf_d(N, ab, pv);
    // explicit operator Y
    int yk = kind < 2 ? kind : (kind == 2 ? -1 : kind - 1);
    if (yk < 0) return;
    double* y = ym + ((size_t)yk * L + l) * N * 7;
    double c2y = c2 * (1.0 - op.implicit);
    for (int j = 0; j < N; j++) {
        for (int q = 0; q < 7; q++) y[j * 7 + q] = c2y * radlap[j * 7 + q];
        double* d = &y[j * 7 + LSD_KL];
        *d = (*d + c1_) - c2y * (ll1 * rm2[j]);
    }
}

// Green's functions and the 2x2 influence matrix of the poloidal velocity
// (modules/velocity.F90:134-168), one work-item per degree l >= 1.
__global__ void k_green(int L, int N, const double* __restrict__ lu, const int* __restrict__ piv,
                        double* __restrict__ gre, double* __restrict__ invg, int* __restrict__ info,
                        double* __restrict__ scratch) {
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < 1 || l >= L) return;
    double* re = scratch + (size_t)l * 2 * N;
    double* im = re + N;
    for (int n = 0; n < N; n++) {
        re[n] = 0.0;
        im[n] = 0.0;
    }
    re[0] = 1.0;
    im[N - 1] = 1.0;
    const double* abg = lu + ((size_t)2 * L + l) * N * LSD_LDAB;
    const int* pg = piv + ((size_t)2 * L + l) * N;
    gbtrs_d(N, abg, pg, re);
    gbtrs_d(N, abg, pg, im);
    re[0] = re[N - 1] = im[0] = im[N - 1] = 0.0;
    const double* abp = lu + ((size_t)1 * L + l) * N * LSD_LDAB;
    const int* pp = piv + ((size_t)1 * L + l) * N;
    gbtrs_d(N, abp, pp, re);
    gbtrs_d(N, abp, pp, im);
    for (int n = 0; n < N; n++) {
        gre[((size_t)l * 2 + 0) * N + n] = re[n];
        gre[((size_t)l * 2 + 1) * N + n] = im[n];
    }
    double g[14];
    g[0] = re[0];
    g[1] = re[N - 1];
    g[7] = im[0];
    g[8] = im[N - 1];
    if (invert7(2, g)) info[l] = 1;
    invg[l * 4 + 0] = g[0];
    invg[l * 4 + 1] = g[1];
    invg[l * 4 + 2] = g[7];
    invg[l * 4 + 3] = g[8];
}

__global__ void k_to_f32(size_t n, const double* __restrict__ a, float* __restrict__ b) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

// FP32 synthesis tables in the per-slot, theta-permuted layout of the FP32
// synthesis: entry (kind, row R, position q) holds table kind at harmonic
// row_h[R] and theta 64 (q / 64) + 16 (q % 4) + (q % 64) / 4, zero for
// padding rows and theta >= Th2.
__global__ void k_tab_syn32(int H, int Th2, int Th2p, int Hp, const int* __restrict__ row_h,
                            const double* __restrict__ tab, float* __restrict__ out) {
    const size_t n = (size_t)5 * Hp * Th2p;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int q = (int)(i % Th2p);
    const size_t kr = i / Th2p;
    const int R = (int)(kr % Hp), kind = (int)(kr / Hp);
    const int w = q & 63, th = (q - w) + 16 * (w & 3) + (w >> 2);
    const int h = row_h[R];
    out[i] = (h >= 0 && th < Th2) ? (float)tab[((size_t)kind * H + h) * Th2 + th] : 0.0f;
}

// FP32 copies of the synthesis and analysis tables, rounded from the FP64
// recursion, for the mixed forms.
int tables_to_f32(Ctx& c) {
    const size_t HT = (size_t)c.H * c.Th2;
    const size_t ns = (size_t)5 * c.Hp * c.Th2p;
    hipLaunchKernelGGL(k_tab_syn32, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, c.stream, c.H, c.Th2, c.Th2p,
                       c.Hp, c.d_row_h, c.d_tab, c.d_tab32);
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((3 * HT + 255) / 256)), dim3(256), 0, c.stream, 3 * HT, c.d_atab,
                       c.d_atab32);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

// FP32 analysis tables of the wide-load kernel: entry (kind, theta, q)
// holds the FP64 table kind at column acol[q] and theta, zero for padding
// columns and theta >= Th2.
__global__ void k_tab_ana4(int H, int Th2, int Th2a, int Hpa, const int* __restrict__ acol,
                           const double* __restrict__ atab, float* __restrict__ out) {
    const size_t n = (size_t)3 * Th2a * Hpa;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int q = (int)(i % Hpa);
    const size_t kt = i / Hpa;
    const int th = (int)(kt % Th2a), kind = (int)(kt / Th2a);
    const int col = acol[q];
    out[i] = (col >= 0 && th < Th2) ? (float)atab[((size_t)kind * Th2 + th) * H + col] : 0.0f;
}

int tables_ana4(Ctx& c) {
    const size_t n = (size_t)3 * c.Th2a * c.Hpa;
    hipLaunchKernelGGL(k_tab_ana4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, c.H, c.Th2, c.Th2a,
                       c.Hpa, c.d_acol, c.d_atab, c.d_atab4);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

// ------------------------------------------------------------ inner core

// Inner-core points r = r_i cos(pi (2 Nic - n) / (2 Nic - 1)),
// n = Nic+1 .. 2 Nic (modules/meshs.F90:165-169), the outer half of a
// Chebyshev grid over the diameter, with their powers r^p, p = -3..3.
__global__ void k_icgrid(int Ni, double rratio, double* __restrict__ r) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (k > Ni) return;
    const double PI = 3.1415926535897931;
    const double ri = rratio / (1.0 - rratio);
    const int N_ = 2 * Ni, n = k + Ni;
    const double r1 = ri * cos(PI * (N_ - n) / (double)(N_ - 1));
    for (int p = -3; p <= 3; p++) r[(p + 3) * Ni + k - 1] = (p == 1) ? r1 : ipow_d(r1, p);
}

// rot_precompute (modules/rotation.F90:20-36): the inertia (8 pi / 15)
// r_i^5; the torque factors of non_rotation (modules/nonlinear.F90:497-499)
// r_i^3 Pm and r_i^3 2 Pm / E; and per theta leg_gwP0 sin(theta)
// (modules/legendre.F90:103-104: P_0^0 gw / 2, mirrored).
__global__ void k_rot_setup(int N, int Th, int Th2, const double* __restrict__ r, const double* __restrict__ gw,
                            const double* __restrict__ sn, const double* __restrict__ p00, double Pm, double E,
                            double* __restrict__ rot, double* __restrict__ gwsin) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < Th) {
        const int jj = j < Th2 ? j : Th - 1 - j;
        gwsin[j] = (p00[jj] * gw[jj] * 0.5) * sn[j];
    }
    if (j == 0) {
        const double PI = 3.1415926535897931;
        const double ri = r[4 * N], ri3 = r[6 * N];
        rot[ROT_OMEGA] = 0.0;
        rot[ROT_VTORQ] = rot[ROT_MTORQ] = rot[ROT_VTORQ_S] = rot[ROT_MTORQ_S] = 0.0;
        rot[ROT_INERTIA] = (8.0 * PI / 15.0) * ipow_d(ri, 5);
        rot[ROT_RI] = ri;
        rot[ROT_VFAC] = ri3 * Pm;
        rot[ROT_MFAC] = ri3 * 2 * Pm / E;
    }
}

int precompute_innercore(Ctx& c) {
    LSD_HIP(hipMemsetAsync(c.d_rot, 0, sizeof(double) * ROT_COUNT, c.stream));
    hipLaunchKernelGGL(k_rot_setup, dim3((c.Th + 63) / 64), dim3(64), 0, c.stream, c.N, c.Th, c.Th2, c.d_r, c.d_gw,
                       c.d_sin, c.d_tab, c.cfg.Pm, c.cfg.E, c.d_rot, c.d_gwsin);
    LSD_HIP(hipGetLastError());
    if (c.Nic <= 1) return LSD_OK;
    const int Ni = c.Nic;
    hipLaunchKernelGGL(k_icgrid, dim3((Ni + 63) / 64), dim3(64), 0, c.stream, Ni, c.cfg.rratio, c.d_icr);
    double *ainv = nullptr, *r1dr = nullptr, *w = nullptr;
    int* info = nullptr;
    AsyncScratch s_ainv(c.stream), s_info(c.stream), s_r1dr(c.stream), s_w(c.stream);
    LSD_HIP(s_ainv.alloc(&ainv, sizeof(double) * 49 * Ni));
    LSD_HIP(s_info.alloc(&info, sizeof(int)));
    LSD_HIP(s_r1dr.alloc(&r1dr, sizeof(double) * 7 * Ni));
    LSD_HIP(s_w.alloc(&w, sizeof(double) * Ni));
    LSD_HIP(hipMemsetAsync(info, 0, sizeof(int), c.stream));
    LSD_HIP(hipMemsetAsync(c.d_icdr1, 0, sizeof(double) * 7 * Ni, c.stream));
    LSD_HIP(hipMemsetAsync(c.d_icdr2, 0, sizeof(double) * 7 * Ni, c.stream));
    hipLaunchKernelGGL(k_fd, dim3((Ni + 63) / 64), dim3(64), 0, c.stream, Ni, c.d_icr, c.d_icdr1, c.d_icdr2, ainv,
                       info);
    hipLaunchKernelGGL(k_mesh_ops, dim3((Ni + 63) / 64), dim3(64), 0, c.stream, Ni, c.d_icr, c.d_icdr1, c.d_icdr2,
                       ainv, c.d_icradlap, r1dr, w);
    LSD_HIP(hipGetLastError());
    int h_info = 0;
    LSD_HIP(hipMemcpyAsync(&h_info, info, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    LSD_HIP(hipStreamSynchronize(c.stream));
    if (h_info) {
        set_error("matrix inversion error 1");
        return LSD_E_SINGULAR;
    }
    return LSD_OK;
}

// One work-item per (field, degree l >= 1): tim_iclumesh_Tor_X /
// tim_iclumesh_Pol_X (modules/timestep.F90:218-344) for Nic > 1 with the BC
// rows of mag_bc_Tor / mag_bc_Pol (modules/magnetic.F90:137-282), Boussinesq
// without a diffusivity profile (MagTorLap = MagPolLap = radLap): one band
// over the inner-core points 1..Nic and the outer-core points Nic..NT (the
// inner-core boundary is point Nic of both); row 1 regular at the innermost
// point, row Nic the continuity of dB/dr across the boundary, row NT
// insulating; factorised in place.  Field 0 also writes the inner core's
// explicit operator Y = c1/dt + (1-c) (radLap - l(l+1)/r^2)
// (timestep.F90:421-449).
__global__ void k_icoperators(int L, int N, int Ni, OpParams op, const double* __restrict__ r,
                              const double* __restrict__ radlap, const double* __restrict__ dr1,
                              const double* __restrict__ icr, const double* __restrict__ icradlap,
                              const double* __restrict__ icdr1, double* __restrict__ iclu, int* __restrict__ icpiv,
                              double* __restrict__ icy, int* __restrict__ info) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int kind = blockIdx.y;
    if (l >= L) return;
    const int NT = Ni + N - 1;
    double* ab = iclu + ((size_t)kind * L + l) * NT * LSD_LDAB;
    int* pv = icpiv + ((size_t)kind * L + l) * NT;
    if (l == 0) {
        info[kind * L] = 0;
        return;
    }
    const double c1_ = 1.0 / op.dt;
    double c2_ = 1.0 * op.implicit;
    if (op.implicit == 0.0) c2_ = 1.0;
    const double ll1 = (double)(l * (l + 1));
    const double* icrm2 = icr + 1 * Ni;
    const double* icrm1 = icr + 2 * Ni;
    const double* rm2 = r + 1 * N;
    const double* rm1 = r + 2 * N;
    for (int q = 0; q < NT * LSD_LDAB; q++) ab[q] = 0.0;
    for (int j = 0; j < Ni; j++)
        for (int q = 0; q < 7; q++) ab[j * LSD_LDAB + LSD_KL + q] = -c2_ * icradlap[j * 7 + q];
    for (int j = 0; j < Ni; j++) {
        double* d = &ab[j * LSD_LDAB + 2 * LSD_KL];
        *d = (*d + c1_) + c2_ * ll1 * icrm2[j];
    }
    for (int j = 0; j < N; j++)
        for (int q = 0; q < 7; q++)
            ab[(Ni - 1 + j) * LSD_LDAB + LSD_KL + q] = ab[(Ni - 1 + j) * LSD_LDAB + LSD_KL + q] - c2_ * radlap[j * 7 + q];
    for (int j = 0; j < N; j++) {
        double* d = &ab[(Ni - 1 + j) * LSD_LDAB + 2 * LSD_KL];
        *d = (*d + c1_) + c2_ * ll1 * rm2[j];
    }
#define LUA(i, j) ab[((j) - 1) * LSD_LDAB + (2 * LSD_KL + (i) - (j))]
#define DR1(i, j) dr1[((j) - 1) * 7 + (LSD_KL + (i) - (j))]
#define IDR1(i, j) icdr1[((j) - 1) * 7 + (LSD_KL + (i) - (j))]
    for (int j = 1; j <= 1 + LSD_KL; j++) LUA(1, j) = 0.0;
    for (int j = NT - LSD_KL; j <= NT; j++) LUA(NT, j) = 0.0;
    for (int j = Ni - LSD_KL; j <= Ni + LSD_KL; j++) LUA(Ni, j) = 0.0;
    if (kind == 0) {
        LUA(1, 1) = 1.0;
        LUA(NT, NT) = 1.0;
    } else {
        for (int j = 1; j <= 1 + LSD_KL; j++) LUA(1, j) = IDR1(1, j);
        LUA(1, 1) = LUA(1, 1) - icrm1[0] * l;
        for (int j = N - LSD_KL; j <= N; j++) LUA(NT, j + Ni - 1) = DR1(N, j);
        LUA(NT, NT) = LUA(NT, NT) + rm1[N - 1] * (l + 1);
    }
    for (int j = 2; j <= 1 + LSD_KL; j++) LUA(Ni, j + Ni - 1) = DR1(1, j);
    for (int j = Ni - LSD_KL; j <= Ni - 1; j++) LUA(Ni, j) = -IDR1(Ni, j);
    LUA(Ni, Ni) = DR1(1, 1) - IDR1(Ni, Ni);
#undef LUA
#undef DR1
#undef IDR1
    info[kind * L + l] = gbtrf_d(NT, ab, pv);
    if (kind != 0) return;
    double* y = icy + (size_t)l * Ni * 7;
    const double c2y = 1.0 * (1.0 - op.implicit);
    for (int j = 0; j < Ni; j++) {
        for (int q = 0; q < 7; q++) y[j * 7 + q] = c2y * icradlap[j * 7 + q];
        double* d = &y[j * 7 + LSD_KL];
        *d = (*d + c1_) - c2y * (ll1 * icrm2[j]);
    }
}

int build_ic_operators(Ctx& c) {
    if (c.cfg.rot || c.Nic > 1)
        LSD_HIP(hipMemcpyAsync(c.d_rot + ROT_DT, &c.dt, sizeof(double), hipMemcpyHostToDevice, c.stream));
    if (!c.ic()) {
        LSD_HIP(hipStreamSynchronize(c.stream));
        return LSD_OK;
    }
    OpParams op{c.dt, c.cfg.implicit, c.cfg.Pm, c.cfg.Pr, c.cfg.Sc, c.cfg.vel_bc};
    int* info = nullptr;
    AsyncScratch s_info(c.stream);
    LSD_HIP(s_info.alloc(&info, sizeof(int) * 2 * c.L));
    hipLaunchKernelGGL(k_icoperators, dim3((c.L + 63) / 64, 2), dim3(64), 0, c.stream, c.L, c.N, c.Nic, op, c.d_r,
                       c.d_radlap, c.d_dr1, c.d_icr, c.d_icradlap, c.d_icdr1, c.d_iclu, c.d_icpiv, c.d_icy, info);
    LSD_HIP(hipGetLastError());
    std::vector<int> h(2 * c.L);
    LSD_HIP(hipMemcpyAsync(h.data(), info, sizeof(int) * 2 * c.L, hipMemcpyDeviceToHost, c.stream));
    LSD_HIP(hipStreamSynchronize(c.stream));
    for (int v : h)
        if (v) {
            set_error("mes_iclu_find");
            return LSD_E_SINGULAR;
        }
    return LSD_OK;
}

int build_operators(Ctx& c) {
    OpParams op{c.dt, c.cfg.implicit, c.cfg.Pm, c.cfg.Pr, c.cfg.Sc, c.cfg.vel_bc};
    int L = c.L, N = c.N;
    dim3 grid((L + 63) / 64, NLU);
    hipLaunchKernelGGL(k_operators, grid, dim3(64), 0, c.stream, L, N, op, c.d_r, c.d_radlap, c.d_dr1, c.d_dr2,
                       c.d_lu, c.d_piv, c.d_y, c.d_info, c.d_xband);
    LSD_HIP(hipGetLastError());
    double* scratch = nullptr;
    int* ginfo = nullptr;
    AsyncScratch s_scr(c.stream), s_ginfo(c.stream);
    LSD_HIP(s_scr.alloc(&scratch, sizeof(double) * 2 * N * L));
    LSD_HIP(s_ginfo.alloc(&ginfo, sizeof(int) * L));
    LSD_HIP(hipMemsetAsync(ginfo, 0, sizeof(int) * L, c.stream));
    hipLaunchKernelGGL(k_green, dim3((L + 63) / 64), dim3(64), 0, c.stream, L, N, c.d_lu, c.d_piv, c.d_gre,
                       c.d_invg, ginfo, scratch);
    LSD_HIP(hipGetLastError());
    int rc_ic = build_ic_operators(c);
    if (rc_ic) return rc_ic;
    std::vector<int> info(NLU * L), gi(L);
    LSD_HIP(hipMemcpyAsync(info.data(), c.d_info, sizeof(int) * NLU * L, hipMemcpyDeviceToHost, c.stream));
    LSD_HIP(hipMemcpyAsync(gi.data(), ginfo, sizeof(int) * L, hipMemcpyDeviceToHost, c.stream));
    LSD_HIP(hipStreamSynchronize(c.stream));
    for (int k = 0; k < NLU; k++)
        for (int l = 0; l < L; l++)
            if (info[k * L + l]) {
                set_error(k >= 5 ? "mes_iclu_find" : "mes_lu_find");
                return LSD_E_SINGULAR;
            }
    for (int l = 0; l < L; l++)
        if (gi[l]) {
            set_error("matrix inversion error 1");
            return LSD_E_SINGULAR;
        }
    return launch_ops_layout(c);
}

}  // namespace lsd
