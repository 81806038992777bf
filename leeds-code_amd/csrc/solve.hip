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
// Work decomposition: one 64-lane wave per block of up to hb harmonics
// that share one degree l (the harmonics are grouped by l on the host), so
// every lane of a block needs the same factors.  The block moves its coll
// columns (each a contiguous run of N double2) between HBM and LDS with
// coalesced loads and stores, transposing to W[n][2*h+part], and stages the
// operator data of (kind, l) -- LU factors and pivots, the Y band, the
// Green's functions -- into LDS once.  Each lane then owns one real column
// (the Re or Im part of one harmonic) and runs dgbtrs on it with the active
// band in a register window and the factors read from LDS at addresses that
// are the same for every lane.
//
// Mixed forms: each solve runs on FP32 factors (rounded from the FP64
// factorisation), then one refinement sweep computes the residual
// b - X x0 in FP64 against the unfactored FP64 band of X, solves for the
// correction in FP32 and adds it in FP64.
#include "solve_common.h"

namespace lsd {

#define RC_SOLVE(x)        \
    do {                   \
        int r_ = (x);      \
        if (r_) return r_; \
    } while (0)


// Mixed forms: x = X^-1 b with FP32 factors and one FP64-residual sweep.
// c holds b on entry and x on exit; c32 is the lane's FP32 work column;
// xb is X in band storage, X(n,j) = xb[j*7 + 3+n-j] (LDS).
__device__ void solve_refined(const Col& c, const ColT<float>& c32, const float* ab32, const int* ip,
                              const double* xb) {
    const int N = c.N, S = c.S;
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
            if (j >= 0 && j < N) acc = acc - xb[j * 7 + LSD_KL + n - j] * xw[d];
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

// vel_gre_adjust (velocity.F90:445-465) on one column; g = LDS [2][N].
__device__ void gre_adjust_col(const Col& c, const double* g, const double* __restrict__ ig) {
    const int N = c.N, S = c.S;
    const double bi = c.w[0], bo = c.w[(N - 1) * S];
    const double a = ig[0] * bi + ig[2] * bo;
    const double b = ig[1] * bi + ig[3] * bo;
    for (int n = 0; n < N; n++) c.w[n * S] = c.w[n * S] - a * g[n] - b * g[N + n];
}

template <bool MIX>
__device__ __forceinline__ void solve_staged(const Smem& m, const Col& c, const ColT<float>& c32) {
    if constexpr (MIX) solve_refined(c, c32, m.lu32, m.piv, m.xb);
    else gbtrs_col<double>(c, m.lu, m.piv);
}

// field f: 0 vT, 1 vP, 2 C, 3 X, 4 bT, 5 bP -> LU kinds 0,(2,1),3,4,5,6 and Y kinds 0,1,2,3,4,5
template <bool MIX>
__global__ void __launch_bounds__(64) k_predict(SolveArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int f = blockIdx.y;
    if (!a.on[f]) return;
    const int4 blk = a.blocks[blockIdx.x];
    const int l = blk.x, first = blk.y, nh = blk.z;
    const int N = a.N, S = a.S, t = threadIdx.x;
    const Smem m = carve(smem_raw, N, S, MIX);
    // stage N into LDS, save it as the history term
    {
        const double2* src = a.nn[f];
        double2* hs = a.hist[f];
        for (int e = t; e < nh * N; e += 64) {
            const int hh = e / N, n = e - hh * N;
            const size_t q = (size_t)a.order[first + hh] * N + n;
            const double2 v = src[q];
            hs[q] = v;
            m.W[n * S + 2 * hh] = v.x;
            m.W[n * S + 2 * hh + 1] = v.y;
        }
    }
    const int hh = t >> 1, part = t & 1;
    const bool act = hh < nh;
    const int h = act ? a.order[first + hh] : 0;
    Col c{m.W + t, S, N};
    ColT<float> c32{MIX ? m.W32 + t : nullptr, S, N};
    const double* xs = reinterpret_cast<const double*>(a.x[f] + (size_t)h * N) + part;
    switch (f) {
    case 0:
    case 4:
    case 5: {
        if (l != 0) {
            stage_y(a, m, f, l);
            stage_lu<MIX>(a, m, f == 0 ? 0 : f + 1, l);
        }
        __syncthreads();
        if (act) {
            if (l != 0) multy_col(c, m.y, xs);
            c.w[0] = c.w[(N - 1) * S] = 0.0;
            // vel_setbc_Tor: the rotating inner core's wall, T_10(r_i) = rot_omega r_i
            if (f == 0 && a.rot && l == 1 && a.hm[h] == 0 && part == 0) c.w[0] = a.rot[ROT_OMEGA] * a.rot[ROT_RI];
            if (l != 0) solve_staged<MIX>(m, c, c32);
        }
        break;
    }
    case 1: {
        if (l != 0) {
            stage_lu<MIX>(a, m, 2, l);
            stage_y(a, m, 1, l);
            stage_g(a, m, l);
        }
        __syncthreads();
        if (act) {
            c.w[0] = c.w[(N - 1) * S] = 0.0;
            if (l != 0) {
                solve_staged<MIX>(m, c, c32);
                multy_col(c, m.y, xs);
                c.w[0] = c.w[(N - 1) * S] = 0.0;
            }
        }
        if (l != 0) {
            __syncthreads();
            stage_lu<MIX>(a, m, 1, l);
            __syncthreads();
            if (act) {
                solve_staged<MIX>(m, c, c32);
                gre_adjust_col(c, m.g, a.invg + l * 4);
            }
        }
        break;
    }
    default: {  // codensity / composition, l = 0 included, inhomogeneous BC on (0,0)
        stage_y(a, m, f, l);
        stage_lu<MIX>(a, m, f + 1, l);
        __syncthreads();
        if (act) {
            multy_col(c, m.y, xs);
            c.w[0] = c.w[(N - 1) * S] = 0.0;
            if (h == 0 && part == 0) c.w[0] = 1.0;
            solve_staged<MIX>(m, c, c32);
        }
        break;
    }
    }
    __syncthreads();
    double2* dst = a.x[f];
    for (int e = t; e < nh * N; e += 64) {
        const int hh2 = e / N, n = e - hh2 * N;
        dst[(size_t)a.order[first + hh2] * N + n] = make_double2(m.W[n * S + 2 * hh2], m.W[n * S + 2 * hh2 + 1]);
    }
}

// partial slots: 0 vT, 1 vP, 2 C, 3 X, 4 bT, 5 icT, 6 bP, 7 icP
template <bool MIX>
__global__ void __launch_bounds__(64) k_correct(SolveArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[2][64];
    const int f = blockIdx.y;
    const int N = a.N, S = a.S, t = threadIdx.x;
    double e2 = 0.0, eic = 0.0;
    if (a.on[f]) {
        const int4 blk = a.blocks[blockIdx.x];
        const int l = blk.x, first = blk.y, nh = blk.z;
        const Smem m = carve(smem_raw, N, S, MIX);
        {
            const double2* src = a.nn[f];
            double2* hs = a.hist[f];
            for (int e = t; e < nh * N; e += 64) {
                const int hh = e / N, n = e - hh * N;
                const size_t q = (size_t)a.order[first + hh] * N + n;
                const double2 v = src[q], o = hs[q];
                hs[q] = v;
                m.W[n * S + 2 * hh] = a.cimp * (v.x - o.x);
                m.W[n * S + 2 * hh + 1] = a.cimp * (v.y - o.y);
            }
        }
        const int kind = f == 0 ? 0 : (f == 1 ? 2 : f + 1);
        const bool solve = l != 0 || f == 2 || f == 3;
        if (solve) {
            stage_lu<MIX>(a, m, kind, l);
            if (f == 1) stage_g(a, m, l);
        }
        __syncthreads();
        const int hh = t >> 1, part = t & 1;
        const bool act = hh < nh;
        const int h = act ? a.order[first + hh] : 0;
        Col c{m.W + t, S, N};
        ColT<float> c32{MIX ? m.W32 + t : nullptr, S, N};
        if (act) {
            c.w[0] = c.w[(N - 1) * S] = 0.0;
            // vel_setbc_Tor for (1,0) once tim_it /= 0
            if (a.t10 && f == 0 && l == 1 && a.hm[h] == 0 && part == 0)
                c.w[0] = (a.rot ? a.rot[ROT_OMEGA] * a.rot[ROT_RI] : 0.0) - a.x[0][(size_t)h * N].x;
            if (solve) solve_staged<MIX>(m, c, c32);
            if (f == 1 && solve) c.w[0] = c.w[(N - 1) * S] = 0.0;
        }
        if (f == 1 && solve) {
            __syncthreads();
            stage_lu<MIX>(a, m, 1, l);
            __syncthreads();
            if (act) {
                solve_staged<MIX>(m, c, c32);
                gre_adjust_col(c, m.g, a.invg + l * 4);
            }
        }
        __syncthreads();
        double2* x = a.x[f];
        for (int e = t; e < nh * N; e += 64) {
            const int hh2 = e / N, n = e - hh2 * N;
            const size_t q = (size_t)a.order[first + hh2] * N + n;
            const double dx = m.W[n * S + 2 * hh2], dy = m.W[n * S + 2 * hh2 + 1];
            const double2 o = x[q];
            x[q] = make_double2(o.x + dx, o.y + dy);
            e2 = e2 + dx * dx;
            e2 = e2 + dy * dy;
            if (f >= 4 && n == 0 && l != 0) {
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

// FP32 copy of the LU factors for the mixed forms.
__global__ void k_lu_f32(size_t n, const double* __restrict__ lu, float* __restrict__ lu32) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) lu32[i] = (float)lu[i];
}

// Pivot offsets ip(j)-1-j (0..3: dgbtrf pivots within KL rows) packed 2 bits
// per row, 16 rows per word: pk[(kind*L + l)*npk + j/16].
__global__ void k_pack_piv(size_t nw, int N, int npk, const int* __restrict__ piv, unsigned* __restrict__ pk) {
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    const size_t kl = w / npk;
    const int j0 = (int)(w - kl * npk) * 16;
    unsigned v = 0;
    for (int j = j0; j < j0 + 16 && j < N; j++) v |= (unsigned)((piv[kl * N + j] - 1 - j) & 3) << (2 * (j - j0));
    pk[w] = v;
}

int launch_ops_layout(Ctx& c) {
    const size_t n = (size_t)NLU * c.L * c.N * LSD_LDAB;
    hipLaunchKernelGGL(k_lu_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, n, c.d_lu, c.d_lut32);
    LSD_HIP(hipGetLastError());
    const int npk = (c.N + 15) / 16;
    const size_t nw = (size_t)NLU * c.L * npk;
    hipLaunchKernelGGL(k_pack_piv, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, c.stream, nw, c.N, npk, c.d_piv,
                       c.d_pivpk);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

// Harmonics per solve block: the LDS of a block (columns plus one
// operator) kept within 63 KiB, so the reduction array still fits the
// 64 KiB a workgroup may hold.  0 if N is too large.
int solve_block_harmonics(int N, bool mixed) {
    int hb = 32;
    while (hb > 0 && solve_smem_bytes(N, 2 * hb + 1, mixed) > 64512) hb--;
    return hb;
}

// FP32 factors with the FP64-residual sweep: the mixed forms, unless the
// LDS-column form with FP64 factors (variant 2) is selected
static bool fp32_factors(const Ctx& c) { return c.mixed && c.var[LSD_V_SOLVE] != 2; }

size_t solve_lds_bytes(const Ctx& c) { return solve_smem_bytes(c.N, 2 * c.solve_hb + 1, fp32_factors(c)); }

static SolveArgs make_args(Ctx& c) {
    SolveArgs a;
    a.H = c.H;
    a.N = c.N;
    a.L = c.L;
    a.S = 2 * c.solve_hb + 1;
    a.on[0] = 1;
    a.on[1] = 1;
    a.on[2] = c.cfg.cod;
    a.on[3] = c.cfg.comp;
    // with a conducting inner core the magnetic field has its own coupled solves
    a.on[4] = c.cfg.mag && !c.ic();
    a.on[5] = c.cfg.mag && !c.ic();
    a.cimp = c.cfg.implicit;
    a.hm = c.d_hm;
    a.rot = c.rot_wall() ? c.d_rot : nullptr;
    a.t10 = c.cfg.vel_bc <= 2;
    a.blocks = c.d_sblocks;
    a.order = c.d_sorder;
    a.lu = c.d_lu;
    a.piv = c.d_piv;
    a.y = c.d_y;
    a.gre = c.d_gre;
    a.invg = c.d_invg;
    a.lu32 = c.d_lut32;
    a.xb = c.d_xband;
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
    return 2 + (c.cfg.cod ? 1 : 0) + (c.cfg.comp ? 1 : 0) + (c.cfg.mag && !c.ic() ? 2 : 0);
}

static size_t solve_lds(const Ctx& c) { return solve_lds_bytes(c); }

// the two-pass kernels when they apply and are selected (LSD_V_SOLVE = 0)
static bool use_direct(const Ctx& c) { return c.solve_direct_ok && c.var[LSD_V_SOLVE] == 0; }

int launch_predictor(Ctx& c) {
    RC_SOLVE(launch_rot_update(c, true));
    if (use_direct(c)) {
        RC_SOLVE(launch_solve_direct(c, true));
        return c.ic() ? launch_ic_solve(c, true) : LSD_OK;
    }
    SolveArgs a = make_args(c);
    // per field: read state and N, write state and saved N
    ProfScope ps(c, LSD_K_SOLVE, 4.0 * 16.0 * c.coll_elems() * active_fields(c));
    dim3 grid(c.n_partial_blocks, 6);
    if (fp32_factors(c)) hipLaunchKernelGGL(k_predict<true>, grid, dim3(64), solve_lds(c), c.stream, a);
    else hipLaunchKernelGGL(k_predict<false>, grid, dim3(64), solve_lds(c), c.stream, a);
    LSD_HIP(hipGetLastError());
    return c.ic() ? launch_ic_solve(c, true) : LSD_OK;
}

// Queues the corrector and the copy of its norm partials to pinned host
// memory; every enabled field's slots are written by every block, so no
// clearing is needed (and none is recorded when the step is captured).
int launch_corrector_async(Ctx& c) {
    RC_SOLVE(launch_rot_update(c, false));
    SolveArgs a = make_args(c);
    const int nb = c.n_partial_blocks;
    dim3 grid(nb, 6);
    if (use_direct(c)) {
        int rc = launch_solve_direct(c, false);
        if (rc) return rc;
    } else {
        // per field: read state, N and saved N, write state and saved N
        ProfScope ps(c, LSD_K_SOLVE, 5.0 * 16.0 * c.coll_elems() * active_fields(c));
        if (fp32_factors(c)) {
            hipLaunchKernelGGL(k_correct<true>, grid, dim3(64), solve_lds(c), c.stream, a);
        } else {
            hipLaunchKernelGGL(k_correct<false>, grid, dim3(64), solve_lds(c), c.stream, a);
        }
    }
    LSD_HIP(hipGetLastError());
    LSD_HIP(hipMemcpyAsync(c.h_partial, c.d_partial, sizeof(double) * 8 * nb, hipMemcpyDeviceToHost, c.stream));
    return c.ic() ? launch_ic_solve(c, false) : LSD_OK;
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
        if (slot >= 4 && c.ic()) {
            // the coupled inner-core solves' norms: BT, icBT, BP, icBP
            for (int b = 0; b < c.n_icblk; b++) s += c.h_icpartial[(size_t)(slot - 4) * c.n_icblk + b];
        } else {
            for (int b = 0; b < nb; b++) s += c.h_partial[(size_t)slot * nb + b];
        }
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
