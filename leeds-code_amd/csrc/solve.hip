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
// Work decomposition: one 64-lane wave per block of hb consecutive
// harmonics of one field.  A block's coll columns are one contiguous run of
// hb*N double2, so the wave moves them between HBM and LDS with coalesced
// loads and stores, transposing to W[n][2*h+part] on the way.  Each lane
// then owns one real column (the Re or Im part of one harmonic: both share
// the factors) and runs dgbtrs on it with the active band held in a
// register window, so the serial chain never waits on memory.  The factors
// are kept transposed, [kind][n][row][l], so lanes with consecutive degrees
// read consecutive words.
//
// Mixed forms: each solve runs on FP32 factors (rounded from the FP64
// factorisation), then one refinement sweep computes the residual
// b - X x0 in FP64 against the unfactored FP64 band of X, solves for the
// correction in FP32 and adds it in FP64.
#include "lsd_internal.h"

namespace lsd {

struct SolveArgs {
    int H, N, L, hb, S;  // harmonics per block, LDS row stride (doubles)
    int on[6];           // field enabled: vT, vP, C, X, bT, bP
    double cimp;
    const int* hl;
    const int* hm;
    const double* lu;    // [NLU][N][LDAB][L]
    const int* piv;      // [NLU][N][L]
    const double* y;     // [NY][N][7][L]
    const double* gre;   // [2][N][L]
    const double* invg;  // [L][4]
    const float* lu32;   // [NLU][N][LDAB][L]  FP32 copy (mixed forms)
    const double* xb;    // [NLU][N][7][L]     unfactored X bands (mixed forms)
    double2* x[6];       // state
    double2* nn[6];      // current N
    double2* hist[6];    // saved N
    double* partial;     // [8][gridDim.x]
};

// Lane view of one real column in LDS and of its harmonic's operators.
template <typename T>
struct ColT {
    T* w;  // W + lane, stride S
    int S, N, L, l;
};
typedef ColT<double> Col;

// dgbtrs('N') on one column (timestep.F90:566-588 -> LAPACK dgbtrs with
// KL = 3, KU = 3, LDAB = 10): row interchanges and L^-1 forward, then the
// upper band solve (dtbsv) backward.  ab/ip point at this kind's factors
// offset by the lane's degree; entry AB(i,j) (1-based) is ab[((j-1)*10+i-1)*L].
template <typename T>
__device__ void gbtrs_col(const ColT<T>& c, const T* __restrict__ ab, const int* __restrict__ ip) {
    const int N = c.N, S = c.S, L = c.L;
    T* w = c.w;
    T r0 = w[0], r1 = w[S], r2 = w[2 * S], r3 = w[3 * S];
    for (int j = 0; j < N - 1; j++) {
        const int p = ip[(size_t)j * L] - 1 - j;  // 0..3 (pivot inside the band)
        T t = r0;
        if (p == 1) { r0 = r1; r1 = t; }
        else if (p == 2) { r0 = r2; r2 = t; }
        else if (p == 3) { r0 = r3; r3 = t; }
        const T yv = r0;
        w[j * S] = yv;
        const T* a = ab + ((size_t)j * LSD_LDAB + 2 * LSD_KL + 1) * L;
        r0 = r1 - a[0] * yv;
        r1 = r2 - a[L] * yv;
        r2 = r3 - a[2 * L] * yv;
        r3 = j + 4 < N ? w[(j + 4) * S] : T(0);
    }
    // registers beyond row N-1 above only ever held rows that are never stored
    T b[7];
    b[0] = r0;
    for (int d = 1; d < 7; d++) b[d] = w[(N - 1 - d) * S];
    for (int j = N - 1; j >= 0; j--) {
        const T* a = ab + (size_t)j * LSD_LDAB * L;
        const T v = b[0] / a[2 * LSD_KL * L];
        w[j * S] = v;
#pragma unroll
        for (int d = 1; d < 7; d++) b[d] = b[d] - v * a[(2 * LSD_KL - d) * L];
#pragma unroll
        for (int d = 0; d < 6; d++) b[d] = b[d + 1];
        b[6] = j - 7 >= 0 ? w[(j - 7) * S] : T(0);
    }
}

// Mixed forms: x = X^-1 b with FP32 factors and one FP64-residual sweep.
// c holds b on entry and x on exit; c32 is the lane's FP32 work column;
// xb is X in band storage, X(n,j) = xb[(j*7 + 3+n-j)*L].
__device__ void solve_refined(const Col& c, const ColT<float>& c32, const float* __restrict__ ab32,
                              const int* __restrict__ ip, const double* __restrict__ xb) {
    const int N = c.N, S = c.S, L = c.L;
    for (int n = 0; n < N; n++) c32.w[n * S] = (float)c.w[n * S];
    gbtrs_col<float>(c32, ab32, ip);
    // r = b - X x0 row by row; b <- x0, work column <- r
    double xw[7];
#pragma unroll
    for (int d = 0; d < 7; d++) xw[d] = (d - 3 >= 0 && d - 3 < N) ? (double)c32.w[(d - 3) * S] : 0.0;
    for (int n = 0; n < N; n++) {
        double acc = c.w[n * S];
#pragma unroll
        for (int d = 0; d < 7; d++) {
            const int j = n - 3 + d;
            if (j >= 0 && j < N) acc = acc - xb[((size_t)j * 7 + LSD_KL + n - j) * L] * xw[d];
        }
        c.w[n * S] = xw[3];
        c32.w[n * S] = (float)acc;
#pragma unroll
        for (int d = 0; d < 6; d++) xw[d] = xw[d + 1];
        xw[6] = n + 4 < N ? (double)c32.w[(n + 4) * S] : 0.0;
    }
    gbtrs_col<float>(c32, ab32, ip);
    for (int n = 0; n < N; n++) c.w[n * S] = c.w[n * S] + (double)c32.w[n * S];
}

// w <- w + Y.x in place (tim_multY, timestep.F90:640-670): row n sums its
// band in ascending column order; x is this lane's real part of the old
// state column, read as a stream.
__device__ void multy_col(const Col& c, const double* __restrict__ yk, const double* __restrict__ xs) {
    const int N = c.N, S = c.S, L = c.L;
    double xw[7];
#pragma unroll
    for (int d = 0; d < 7; d++) xw[d] = (d - 3 >= 0 && d - 3 < N) ? xs[2 * (d - 3)] : 0.0;
    for (int n = 0; n < N; n++) {
        double acc = c.w[n * S];
#pragma unroll
        for (int d = 0; d < 7; d++) {
            const int j = n - 3 + d;
            if (j >= 0 && j < N) acc = fma(yk[((size_t)j * 7 + LSD_KL + n - j) * L], xw[d], acc);
        }
        c.w[n * S] = acc;
#pragma unroll
        for (int d = 0; d < 6; d++) xw[d] = xw[d + 1];
        xw[6] = n + 4 < N ? xs[2 * (n + 4)] : 0.0;
    }
}

// vel_gre_adjust (velocity.F90:445-465) on one column.
__device__ void gre_adjust_col(const Col& c, const double* __restrict__ g, const double* __restrict__ ig) {
    const int N = c.N, S = c.S, L = c.L;
    const double bi = c.w[0], bo = c.w[(N - 1) * S];
    const double a = ig[0] * bi + ig[2] * bo;
    const double b = ig[1] * bi + ig[3] * bo;
    const double* g1 = g;
    const double* g2 = g + (size_t)N * L;
    for (int n = 0; n < N; n++) c.w[n * S] = c.w[n * S] - a * g1[(size_t)n * L] - b * g2[(size_t)n * L];
}

__device__ __forceinline__ const double* lu_of(const SolveArgs& a, int kind, int l) {
    return a.lu + (size_t)kind * a.N * LSD_LDAB * a.L + l;
}
__device__ __forceinline__ const int* piv_of(const SolveArgs& a, int kind, int l) {
    return a.piv + (size_t)kind * a.N * a.L + l;
}
__device__ __forceinline__ const double* y_of(const SolveArgs& a, int kind, int l) {
    return a.y + (size_t)kind * a.N * 7 * a.L + l;
}

template <bool MIX>
__device__ __forceinline__ void solve_kind(const SolveArgs& a, const Col& c, const ColT<float>& c32, int kind) {
    if constexpr (MIX) {
        const size_t off = (size_t)kind * a.N * LSD_LDAB * a.L + c.l;
        solve_refined(c, c32, a.lu32 + off, piv_of(a, kind, c.l), a.xb + (size_t)kind * a.N * 7 * a.L + c.l);
    } else {
        gbtrs_col<double>(c, lu_of(a, kind, c.l), piv_of(a, kind, c.l));
    }
}

// field f: 0 vT, 1 vP, 2 C, 3 X, 4 bT, 5 bP -> LU kinds 0,(2,1),3,4,5,6 and Y kinds 0,1,2,3,4,5
template <bool MIX>
__global__ void __launch_bounds__(64) k_predict(SolveArgs a) {
    extern __shared__ double W[];
    float* W32 = reinterpret_cast<float*>(W + (size_t)a.N * a.S);
    const int f = blockIdx.y;
    if (!a.on[f]) return;
    const int N = a.N, S = a.S, t = threadIdx.x;
    const int h0 = blockIdx.x * a.hb;
    const int nh = min(a.hb, a.H - h0);
    const size_t base = (size_t)h0 * N;
    // stage N into LDS, save it as the history term
    {
        const double2* src = a.nn[f] + base;
        double2* hs = a.hist[f] + base;
        for (int e = t; e < nh * N; e += 64) {
            const int hh = e / N, n = e - hh * N;
            const double2 v = src[e];
            hs[e] = v;
            W[n * S + 2 * hh] = v.x;
            W[n * S + 2 * hh + 1] = v.y;
        }
    }
    __syncthreads();
    const int hh = t >> 1, part = t & 1;
    if (hh < nh) {
        const int h = h0 + hh, l = a.hl[h];
        Col c{W + t, S, N, a.L, l};
        ColT<float> c32{W32 + t, S, N, a.L, l};
        const double* xs = reinterpret_cast<const double*>(a.x[f] + (size_t)h * N) + part;
        switch (f) {
        case 0:
        case 4:
        case 5: {
            const int luk = f == 0 ? 0 : f + 1;
            if (l != 0) multy_col(c, y_of(a, f, l), xs);
            c.w[0] = c.w[(N - 1) * S] = 0.0;
            if (l != 0) solve_kind<MIX>(a, c, c32, luk);
            break;
        }
        case 1: {
            c.w[0] = c.w[(N - 1) * S] = 0.0;
            if (l != 0) {
                solve_kind<MIX>(a, c, c32, 2);
                multy_col(c, y_of(a, 1, l), xs);
                c.w[0] = c.w[(N - 1) * S] = 0.0;
                solve_kind<MIX>(a, c, c32, 1);
                gre_adjust_col(c, a.gre + l, a.invg + l * 4);
            }
            break;
        }
        default: {  // codensity / composition, l = 0 included, inhomogeneous BC on (0,0)
            multy_col(c, y_of(a, f, l), xs);
            c.w[0] = c.w[(N - 1) * S] = 0.0;
            if (h == 0 && part == 0) c.w[0] = 1.0;
            solve_kind<MIX>(a, c, c32, f + 1);
            break;
        }
        }
    }
    __syncthreads();
    double2* dst = a.x[f] + base;
    for (int e = t; e < nh * N; e += 64) {
        const int hh2 = e / N, n = e - hh2 * N;
        dst[e] = make_double2(W[n * S + 2 * hh2], W[n * S + 2 * hh2 + 1]);
    }
}

// partial slots: 0 vT, 1 vP, 2 C, 3 X, 4 bT, 5 icT, 6 bP, 7 icP
template <bool MIX>
__global__ void __launch_bounds__(64) k_correct(SolveArgs a) {
    extern __shared__ double W[];
    float* W32 = reinterpret_cast<float*>(W + (size_t)a.N * a.S);
    __shared__ double red[2][64];
    const int f = blockIdx.y;
    const int N = a.N, S = a.S, t = threadIdx.x;
    double e2 = 0.0, eic = 0.0;
    if (a.on[f]) {
        const int h0 = blockIdx.x * a.hb;
        const int nh = min(a.hb, a.H - h0);
        const size_t base = (size_t)h0 * N;
        {
            const double2* src = a.nn[f] + base;
            double2* hs = a.hist[f] + base;
            for (int e = t; e < nh * N; e += 64) {
                const int hh = e / N, n = e - hh * N;
                const double2 v = src[e], o = hs[e];
                hs[e] = v;
                W[n * S + 2 * hh] = a.cimp * (v.x - o.x);
                W[n * S + 2 * hh + 1] = a.cimp * (v.y - o.y);
            }
        }
        __syncthreads();
        const int hh = t >> 1, part = t & 1;
        if (hh < nh) {
            const int h = h0 + hh, l = a.hl[h];
            Col c{W + t, S, N, a.L, l};
            ColT<float> c32{W32 + t, S, N, a.L, l};
            c.w[0] = c.w[(N - 1) * S] = 0.0;
            // vel_setbc_Tor for (1,0) once tim_it /= 0
            if (f == 0 && l == 1 && a.hm[h] == 0 && part == 0) c.w[0] = 0.0 - a.x[0][(size_t)h * N].x;
            switch (f) {
            case 0: if (l != 0) solve_kind<MIX>(a, c, c32, 0); break;
            case 1:
                if (l != 0) {
                    solve_kind<MIX>(a, c, c32, 2);
                    c.w[0] = c.w[(N - 1) * S] = 0.0;
                    solve_kind<MIX>(a, c, c32, 1);
                    gre_adjust_col(c, a.gre + l, a.invg + l * 4);
                }
                break;
            case 2: solve_kind<MIX>(a, c, c32, 3); break;
            case 3: solve_kind<MIX>(a, c, c32, 4); break;
            case 4: if (l != 0) solve_kind<MIX>(a, c, c32, 5); break;
            default: if (l != 0) solve_kind<MIX>(a, c, c32, 6); break;
            }
        }
        __syncthreads();
        double2* x = a.x[f] + base;
        for (int e = t; e < nh * N; e += 64) {
            const int hh2 = e / N, n = e - hh2 * N;
            const double dx = W[n * S + 2 * hh2], dy = W[n * S + 2 * hh2 + 1];
            const double2 o = x[e];
            x[e] = make_double2(o.x + dx, o.y + dy);
            e2 = e2 + dx * dx;
            e2 = e2 + dy * dy;
            if (f >= 4 && n == 0 && a.hl[h0 + hh2] != 0) {
                eic = eic + dx * dx;
                eic = eic + dy * dy;
            }
        }
    }
    red[0][t] = e2;
    red[1][t] = eic;
    __syncthreads();
    if (t == 0) {
        double s = 0.0, si = 0.0;
        for (int k = 0; k < 64; k++) {
            s += red[0][k];
            si += red[1][k];
        }
        const int slot = f < 4 ? f : (f == 4 ? 4 : 6);
        a.partial[(size_t)slot * gridDim.x + blockIdx.x] = s;
        if (f >= 4) a.partial[(size_t)(slot + 1) * gridDim.x + blockIdx.x] = si;
    }
}

// Factors, Y bands and Green's functions from the per-degree layouts the
// operator builders write to the degree-fastest layouts read above.
__global__ void k_ops_layout(int L, int N, const double* __restrict__ lu, const int* __restrict__ piv,
                             const double* __restrict__ y, const double* __restrict__ gre,
                             const double* __restrict__ xband, double* __restrict__ lut, int* __restrict__ pivt,
                             double* __restrict__ yt, double* __restrict__ gret, float* __restrict__ lut32,
                             double* __restrict__ xt) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t per = (size_t)N * L;
    if (idx >= per) return;
    const int n = (int)(idx / L), l = (int)(idx - (size_t)n * L);
    for (int k = 0; k < NLU; k++) {
        const double* s = lu + (((size_t)k * L + l) * N + n) * LSD_LDAB;
        for (int i = 0; i < LSD_LDAB; i++) {
            lut[(((size_t)k * N + n) * LSD_LDAB + i) * L + l] = s[i];
            lut32[(((size_t)k * N + n) * LSD_LDAB + i) * L + l] = (float)s[i];
        }
        pivt[((size_t)k * N + n) * L + l] = piv[((size_t)k * L + l) * N + n];
        const double* xs = xband + (((size_t)k * L + l) * N + n) * 7;
        for (int i = 0; i < 7; i++) xt[(((size_t)k * N + n) * 7 + i) * L + l] = xs[i];
    }
    for (int k = 0; k < NY; k++) {
        const double* s = y + (((size_t)k * L + l) * N + n) * 7;
        for (int i = 0; i < 7; i++) yt[(((size_t)k * N + n) * 7 + i) * L + l] = s[i];
    }
    for (int k = 0; k < 2; k++) gret[((size_t)k * N + n) * L + l] = gre[((size_t)l * 2 + k) * N + n];
}

int launch_ops_layout(Ctx& c) {
    const size_t per = (size_t)c.N * c.L;
    hipLaunchKernelGGL(k_ops_layout, dim3((unsigned)((per + 255) / 256)), dim3(256), 0, c.stream, c.L, c.N, c.d_lu,
                       c.d_piv, c.d_y, c.d_gre, c.d_xband, c.d_lut, c.d_pivt, c.d_yt, c.d_gret, c.d_lut32, c.d_xt);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

int solve_block_harmonics(int N, bool mixed) {
    // LDS of one block: N rows of (2 hb + 1) doubles (plus as many floats in
    // the mixed forms), kept within 63 KiB so the reduction array still fits
    // the 64 KiB a workgroup may hold
    int hb = (64512 / ((mixed ? 12 : 8) * N) - 1) / 2;
    return hb < 32 ? hb : 32;
}

static SolveArgs make_args(Ctx& c) {
    SolveArgs a;
    a.H = c.H;
    a.N = c.N;
    a.L = c.L;
    a.hb = c.solve_hb;
    a.S = 2 * c.solve_hb + 1;
    a.on[0] = 1;
    a.on[1] = 1;
    a.on[2] = c.cfg.cod;
    a.on[3] = c.cfg.comp;
    a.on[4] = c.cfg.mag;
    a.on[5] = c.cfg.mag;
    a.cimp = c.cfg.implicit;
    a.hl = c.d_hl;
    a.hm = c.d_hm;
    a.lu = c.d_lut;
    a.piv = c.d_pivt;
    a.y = c.d_yt;
    a.gre = c.d_gret;
    a.invg = c.d_invg;
    a.lu32 = c.d_lut32;
    a.xb = c.d_xt;
    const int fx[6] = {LSD_VEL_TOR, LSD_VEL_POL, LSD_COD, LSD_COMP, LSD_MAG_TOR, LSD_MAG_POL};
    const int fn[6] = {LSD_N_VEL_TOR, LSD_N_VEL_POL, LSD_N_COD, LSD_N_COMP, LSD_N_MAG_TOR, LSD_N_MAG_POL};
    for (int k = 0; k < 6; k++) {
        a.x[k] = c.d_f[fx[k]];
        a.nn[k] = c.d_f[fn[k]];
        a.hist[k] = c.d_hist[k];
    }
    a.partial = c.d_partial;
    return a;
}

static int active_fields(const Ctx& c) {
    return 2 + (c.cfg.cod ? 1 : 0) + (c.cfg.comp ? 1 : 0) + (c.cfg.mag ? 2 : 0);
}

static size_t solve_lds(const Ctx& c) {
    return (c.mixed ? 12 : 8) * (size_t)c.N * (2 * c.solve_hb + 1);
}

int launch_predictor(Ctx& c) {
    SolveArgs a = make_args(c);
    // per field: read state and N, write state and saved N
    ProfScope ps(c, LSD_K_SOLVE, 4.0 * 16.0 * c.coll_elems() * active_fields(c));
    dim3 grid(c.n_partial_blocks, 6);
    if (c.mixed) hipLaunchKernelGGL(k_predict<true>, grid, dim3(64), solve_lds(c), c.stream, a);
    else hipLaunchKernelGGL(k_predict<false>, grid, dim3(64), solve_lds(c), c.stream, a);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

// Queues the corrector and the copy of its norm partials to pinned host
// memory; every enabled field's slots are written by every block, so no
// clearing is needed (and none is recorded when the step is captured).
int launch_corrector_async(Ctx& c) {
    SolveArgs a = make_args(c);
    const int nb = c.n_partial_blocks;
    dim3 grid(nb, 6);
    {
        // per field: read state, N and saved N, write state and saved N
        ProfScope ps(c, LSD_K_SOLVE, 5.0 * 16.0 * c.coll_elems() * active_fields(c));
        if (c.mixed) hipLaunchKernelGGL(k_correct<true>, grid, dim3(64), solve_lds(c), c.stream, a);
        else hipLaunchKernelGGL(k_correct<false>, grid, dim3(64), solve_lds(c), c.stream, a);
    }
    LSD_HIP(hipGetLastError());
    LSD_HIP(hipMemcpyAsync(c.h_partial, c.d_partial, sizeof(double) * 8 * nb, hipMemcpyDeviceToHost, c.stream));
    return LSD_OK;
}

// After the stream has drained: the sum of the per-field correction norms in
// the reference's order (vel Tor, vel Pol, cod, comp, mag Tor, inner-core
// Tor, mag Pol, inner-core Pol), tim_addcorr (modules/timestep.F90:715-737).
double corrector_norms(const Ctx& c) {
    const int nb = c.n_partial_blocks;
    double tot = 0.0;
    for (int slot = 0; slot < 8; slot++) {
        bool on = slot == 0 || slot == 1 || (slot == 2 && c.cfg.cod) || (slot == 3 && c.cfg.comp) ||
                  (slot >= 4 && c.cfg.mag);
        if (!on) continue;
        double s = 0.0;
        for (int b = 0; b < nb; b++) s += c.h_partial[(size_t)slot * nb + b];
        tot += sqrt(s);
    }
    return tot;
}

int launch_corrector(Ctx& c, double* dterr) {
    int rc = launch_corrector_async(c);
    if (rc) return rc;
    LSD_HIP(hipStreamSynchronize(c.stream));
    *dterr = corrector_norms(c);
    return LSD_OK;
}

}  // namespace lsd
