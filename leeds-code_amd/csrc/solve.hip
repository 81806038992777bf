// Implicit part of the IMEX predictor-corrector: per-harmonic banded solves
// with the LU factors built by precompute.hip.
//
// Predictor (modules/velocity.F90:369-395, codensity.F90:224-236,
// composition.F90:209-223, magnetic.F90:370-411):  X f* = Y f + N  with the
// boundary rows of the *_setbc routines; the poloidal velocity goes through
// the Green's-function sequence XGre, Y, XPol and vel_gre_adjust.
// Corrector (velocity.F90:413-439, codensity.F90:248-259,
// composition.F90:235-246, magnetic.F90:422-461):  X d = c (N* - N), f += d,
// with the per-field correction norms of tim_addcorr (timestep.F90:715-737).
//
// One work-item per (field, harmonic); Re and Im share the factors and run
// together as one complex right-hand side.
#include "lsd_internal.h"

namespace lsd {

__device__ __forceinline__ double2 c2(double a, double b) { return make_double2(a, b); }

// dgbtrs('N') on a complex right-hand side (timestep.F90:566-588).
__device__ void gbtrs_c(int n, const double* __restrict__ ab, const int* __restrict__ ipiv, double2* b) {
#define AB(i, j) ab[((j) - 1) * LSD_LDAB + ((i) - 1)]
    const int kl = LSD_KL, kd = 2 * LSD_KL + 1;
    for (int j = 1; j <= n - 1; j++) {
        int lm = kl < n - j ? kl : n - j;
        int l = ipiv[j - 1];
        if (l != j) {
            double2 t = b[l - 1];
            b[l - 1] = b[j - 1];
            b[j - 1] = t;
        }
        double2 y = b[j - 1];
        for (int i = 1; i <= lm; i++) {
            double a = AB(kd + i, j);
            b[j - 1 + i] = c2(b[j - 1 + i].x - a * y.x, b[j - 1 + i].y - a * y.y);
        }
    }
    const int k = 2 * LSD_KL;
    for (int j = n; j >= 1; j--) {
        int l = k + 1 - j;
        double d = AB(k + 1, j);
        double2 v = b[j - 1];
        v = c2(v.x / d, v.y / d);
        b[j - 1] = v;
        int i0 = j - k > 1 ? j - k : 1;
        for (int i = j - 1; i >= i0; i--) {
            double a = AB(l + i, j);
            b[i - 1] = c2(b[i - 1].x - v.x * a, b[i - 1].y - v.y * a);
        }
    }
#undef AB
}

// out = c + Y.x, rows accumulated in ascending column order (tim_multY,
// timestep.F90:640-670).  out may alias c.
__device__ void multy_c(int N, const double* __restrict__ y, const double2* __restrict__ x, const double2* c,
                        double2* out) {
    for (int n = 0; n < N; n++) {
        int lo = n - LSD_KL > 0 ? n - LSD_KL : 0;
        int hi = n + LSD_KL < N - 1 ? n + LSD_KL : N - 1;
        double2 acc = c[n];
        for (int j = lo; j <= hi; j++) {
            double a = y[j * 7 + (LSD_KL + n - j)];
            acc.x = fma(a, x[j].x, acc.x);
            acc.y = fma(a, x[j].y, acc.y);
        }
        out[n] = acc;
    }
}

// vel_gre_adjust (velocity.F90:445-465).
__device__ void gre_adjust_c(int N, const double* __restrict__ g1, const double* __restrict__ g2,
                             const double* __restrict__ ig, double2* v) {
    double2 bi = v[0], bo = v[N - 1];
    double2 a = c2(ig[0] * bi.x + ig[2] * bo.x, ig[0] * bi.y + ig[2] * bo.y);
    double2 b = c2(ig[1] * bi.x + ig[3] * bo.x, ig[1] * bi.y + ig[3] * bo.y);
    for (int n = 0; n < N; n++)
        v[n] = c2(v[n].x - a.x * g1[n] - b.x * g2[n], v[n].y - a.y * g1[n] - b.y * g2[n]);
}

struct SolveArgs {
    int H, N, L;
    int on[6];  // field enabled: vT, vP, C, X, bT, bP
    double cimp;
    const int* hl;
    const int* hm;
    const double* lu;
    const int* piv;
    const double* y;
    const double* gre;
    const double* invg;
    double2* x[6];     // state
    double2* nn[6];    // current N
    double2* hist[6];  // saved N
    double2* work;     // [6][H][N]
    double* partial;   // [8][gridDim.x]
};

__device__ __forceinline__ const double* lu_of(const SolveArgs& a, int kind, int l) {
    return a.lu + ((size_t)kind * a.L + l) * a.N * LSD_LDAB;
}
__device__ __forceinline__ const int* piv_of(const SolveArgs& a, int kind, int l) {
    return a.piv + ((size_t)kind * a.L + l) * a.N;
}
__device__ __forceinline__ const double* y_of(const SolveArgs& a, int kind, int l) {
    return a.y + ((size_t)kind * a.L + l) * a.N * 7;
}

// field f: 0 vT, 1 vP, 2 C, 3 X, 4 bT, 5 bP -> LU kinds 0,(2,1),3,4,5,6 and Y kinds 0,1,2,3,4,5
__global__ void __launch_bounds__(64) k_predict(SolveArgs a) {
    const int f = blockIdx.y;
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (!a.on[f] || h >= a.H) return;
    const int N = a.N, l = a.hl[h];
    const size_t o = (size_t)h * N;
    double2* x = a.x[f] + o;
    const double2* nn = a.nn[f] + o;
    double2* hs = a.hist[f] + o;
    double2* w = a.work + (size_t)f * a.H * N + o;
    for (int n = 0; n < N; n++) hs[n] = nn[n];
    switch (f) {
    case 0:
    case 4: {
        int luk = f == 0 ? 0 : 5;
        if (l == 0) {
            for (int n = 0; n < N; n++) w[n] = nn[n];
        } else {
            multy_c(N, y_of(a, f, l), x, nn, w);
        }
        w[0] = w[N - 1] = c2(0.0, 0.0);
        if (l != 0) gbtrs_c(N, lu_of(a, luk, l), piv_of(a, luk, l), w);
        break;
    }
    case 1: {
        for (int n = 0; n < N; n++) w[n] = nn[n];
        w[0] = w[N - 1] = c2(0.0, 0.0);
        if (l != 0) {
            gbtrs_c(N, lu_of(a, 2, l), piv_of(a, 2, l), w);
            multy_c(N, y_of(a, 1, l), x, w, w);
            w[0] = w[N - 1] = c2(0.0, 0.0);
            gbtrs_c(N, lu_of(a, 1, l), piv_of(a, 1, l), w);
            gre_adjust_c(N, a.gre + (size_t)l * 2 * N, a.gre + ((size_t)l * 2 + 1) * N, a.invg + l * 4, w);
        }
        break;
    }
    case 5: {
        if (l == 0) {
            for (int n = 0; n < N; n++) w[n] = nn[n];
        } else {
            multy_c(N, y_of(a, 5, l), x, nn, w);
        }
        w[0] = w[N - 1] = c2(0.0, 0.0);
        if (l != 0) gbtrs_c(N, lu_of(a, 6, l), piv_of(a, 6, l), w);
        break;
    }
    default: {  // codensity / composition, l = 0 included, inhomogeneous BC on (0,0)
        multy_c(N, y_of(a, f, l), x, nn, w);
        w[0] = w[N - 1] = c2(0.0, 0.0);
        if (h == 0) w[0].x = 1.0;
        gbtrs_c(N, lu_of(a, f + 1, l), piv_of(a, f + 1, l), w);
        break;
    }
    }
    for (int n = 0; n < N; n++) x[n] = w[n];
}

// partial slots: 0 vT, 1 vP, 2 C, 3 X, 4 bT, 5 icT, 6 bP, 7 icP
__global__ void __launch_bounds__(64) k_correct(SolveArgs a) {
    __shared__ double red[2][64];
    const int f = blockIdx.y;
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0, eic = 0.0;
    if (a.on[f] && h < a.H) {
        const int N = a.N, l = a.hl[h];
        const size_t o = (size_t)h * N;
        double2* x = a.x[f] + o;
        const double2* nn = a.nn[f] + o;
        double2* hs = a.hist[f] + o;
        double2* w = a.work + (size_t)f * a.H * N + o;
        for (int n = 0; n < N; n++) {
            double2 d = c2(a.cimp * (nn[n].x - hs[n].x), a.cimp * (nn[n].y - hs[n].y));
            w[n] = d;
            hs[n] = nn[n];
        }
        w[0] = w[N - 1] = c2(0.0, 0.0);
        if (f == 0 && l == 1 && a.hm[h] == 0) w[0].x = 0.0 - x[0].x;  // vel_setbc_Tor, tim_it /= 0
        switch (f) {
        case 0: if (l != 0) gbtrs_c(N, lu_of(a, 0, l), piv_of(a, 0, l), w); break;
        case 1:
            if (l != 0) {
                gbtrs_c(N, lu_of(a, 2, l), piv_of(a, 2, l), w);
                w[0] = w[N - 1] = c2(0.0, 0.0);
                gbtrs_c(N, lu_of(a, 1, l), piv_of(a, 1, l), w);
                gre_adjust_c(N, a.gre + (size_t)l * 2 * N, a.gre + ((size_t)l * 2 + 1) * N, a.invg + l * 4, w);
            }
            break;
        case 2: gbtrs_c(N, lu_of(a, 3, l), piv_of(a, 3, l), w); break;
        case 3: gbtrs_c(N, lu_of(a, 4, l), piv_of(a, 4, l), w); break;
        case 4: if (l != 0) gbtrs_c(N, lu_of(a, 5, l), piv_of(a, 5, l), w); break;
        default: if (l != 0) gbtrs_c(N, lu_of(a, 6, l), piv_of(a, 6, l), w); break;
        }
        for (int n = 0; n < N; n++) {
            double2 d = w[n];
            e = e + d.x * d.x;
            e = e + d.y * d.y;
            x[n] = c2(x[n].x + d.x, x[n].y + d.y);
        }
        if (f >= 4 && l != 0) {
            eic = eic + w[0].x * w[0].x;
            eic = eic + w[0].y * w[0].y;
        }
    }
    red[0][threadIdx.x] = e;
    red[1][threadIdx.x] = eic;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0, si = 0.0;
        for (int t = 0; t < 64; t++) {
            s += red[0][t];
            si += red[1][t];
        }
        int slot = f < 4 ? f : (f == 4 ? 4 : 6);
        a.partial[(size_t)slot * gridDim.x + blockIdx.x] = s;
        if (f >= 4) a.partial[(size_t)(slot + 1) * gridDim.x + blockIdx.x] = si;
    }
}

static SolveArgs make_args(Ctx& c) {
    SolveArgs a;
    a.H = c.H;
    a.N = c.N;
    a.L = c.L;
    a.on[0] = 1;
    a.on[1] = 1;
    a.on[2] = c.cfg.cod;
    a.on[3] = c.cfg.comp;
    a.on[4] = c.cfg.mag;
    a.on[5] = c.cfg.mag;
    a.cimp = c.cfg.implicit;
    a.hl = c.d_hl;
    a.hm = c.d_hm;
    a.lu = c.d_lu;
    a.piv = c.d_piv;
    a.y = c.d_y;
    a.gre = c.d_gre;
    a.invg = c.d_invg;
    const int fx[6] = {LSD_VEL_TOR, LSD_VEL_POL, LSD_COD, LSD_COMP, LSD_MAG_TOR, LSD_MAG_POL};
    const int fn[6] = {LSD_N_VEL_TOR, LSD_N_VEL_POL, LSD_N_COD, LSD_N_COMP, LSD_N_MAG_TOR, LSD_N_MAG_POL};
    for (int k = 0; k < 6; k++) {
        a.x[k] = c.d_f[fx[k]];
        a.nn[k] = c.d_f[fn[k]];
        a.hist[k] = c.d_hist[k];
    }
    a.work = c.d_work;
    a.partial = c.d_partial;
    return a;
}

static int active_fields(const Ctx& c) {
    return 2 + (c.cfg.cod ? 1 : 0) + (c.cfg.comp ? 1 : 0) + (c.cfg.mag ? 2 : 0);
}

int launch_predictor(Ctx& c) {
    SolveArgs a = make_args(c);
    // per field: read state and N, write state and saved N
    ProfScope ps(c, LSD_K_SOLVE, 4.0 * 16.0 * c.coll_elems() * active_fields(c));
    dim3 grid((c.H + 63) / 64, 6);
    hipLaunchKernelGGL(k_predict, grid, dim3(64), 0, c.stream, a);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

// Runs the corrector and returns the sum of the per-field correction norms
// in the reference's order (vel Tor, vel Pol, cod, comp, mag Tor, inner-core
// Tor, mag Pol, inner-core Pol).
int launch_corrector(Ctx& c, double* dterr) {
    SolveArgs a = make_args(c);
    int nb = (c.H + 63) / 64;
    LSD_HIP(hipMemsetAsync(c.d_partial, 0, sizeof(double) * 8 * nb, c.stream));
    dim3 grid(nb, 6);
    {
        // per field: read state, N and saved N, write state and saved N
        ProfScope ps(c, LSD_K_SOLVE, 5.0 * 16.0 * c.coll_elems() * active_fields(c));
        hipLaunchKernelGGL(k_correct, grid, dim3(64), 0, c.stream, a);
    }
    LSD_HIP(hipGetLastError());
    LSD_HIP(hipMemcpyAsync(c.h_partial, c.d_partial, sizeof(double) * 8 * nb, hipMemcpyDeviceToHost, c.stream));
    LSD_HIP(hipStreamSynchronize(c.stream));
    double tot = 0.0;
    const int order[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    for (int q = 0; q < 8; q++) {
        int slot = order[q];
        bool on = slot == 0 || slot == 1 || (slot == 2 && c.cfg.cod) || (slot == 3 && c.cfg.comp) ||
                  (slot >= 4 && c.cfg.mag);
        if (!on) continue;
        double s = 0.0;
        for (int b = 0; b < nb; b++) s += c.h_partial[(size_t)slot * nb + b];
        tot += sqrt(s);
    }
    *dterr = tot;
    return LSD_OK;
}

}  // namespace lsd
