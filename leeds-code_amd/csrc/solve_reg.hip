// Banded predictor / corrector solves with the columns in registers, for
// the radial resolutions the library is built for (N fixed at compile
// time, a multiple of 8).  Same algorithm and work decomposition as
// solve.hip -- one wave per block of up to 32 harmonics of one degree l,
// lane = (harmonic, Re/Im part), the operator data of (kind, l) staged in
// LDS and read at lane-uniform addresses -- but each lane keeps its whole
// column in VGPRs, with every loop over the radius unrolled, so LDS holds
// only the operator and an 8-row transposition tile and many more waves
// fit on a CU.  Columns move between HBM and registers 8 rows at a time:
// the 128-byte runs of the coll layout are read and written coalesced and
// transposed through the tile.
//
// The mixed forms refine each FP32 solve with one FP64-residual sweep, as
// in solve.hip.
#include "solve_common.h"

namespace lsd {

constexpr int RB = 8;        // rows per transposition chunk

// Keeps the scheduler from hoisting the operator loads of many unrolled
// iterations ahead of their use (each would hold registers until then).
#define SWEEP_FENCE() __builtin_amdgcn_sched_barrier(0)

// Materialises a value in its VGPRs at this point of the program, so that
// its computation is not moved next to a much later use.
template <typename T>
__device__ __forceinline__ void pin(T& v) {
    asm volatile("" : "+v"(v));
}
constexpr int TILE_LD = 65;  // tile row stride (doubles), padded against bank conflicts

// dgbtrs('N') (timestep.F90:566-588 -> LAPACK dgbtrs, KL = KU = 3,
// LDAB = 10) on a register column; ab/ip are the LDS factors of this
// (kind, l), AB(i,j) (1-based) = ab[(j-1)*10 + i-1].
// The row interchange is a select on the pivot offset, which keeps the
// pivot load off the serial chain; the backward sweep multiplies by the
// staged reciprocal of U(j,j) (rd) instead of dividing, which keeps the
// long-latency division off it too.
template <int N, typename T>
__device__ __forceinline__ void gbtrs_reg(T (&b)[N], const T* ab, const T* rd, const int* ip) {
#pragma unroll
    for (int j = 0; j < N - 1; j++) {
        const int p = ip[j] - 1 - j;  // 0..3 (pivot inside the band)
        const T* a = ab + j * LSD_LDAB + 2 * LSD_KL + 1;
        const T bj = b[j];
        T c1 = T(0), c2 = T(0), c3 = T(0);
        if (j + 1 < N) c1 = b[j + 1];
        if (j + 2 < N) c2 = b[j + 2];
        if (j + 3 < N) c3 = b[j + 3];
        const T y = p == 0 ? bj : (p == 1 ? c1 : (p == 2 ? c2 : c3));
        b[j] = y;
        if (j + 1 < N) b[j + 1] = (p == 1 ? bj : c1) - a[0] * y;
        if (j + 2 < N) b[j + 2] = (p == 2 ? bj : c2) - a[1] * y;
        if (j + 3 < N) b[j + 3] = (p == 3 ? bj : c3) - a[2] * y;
        SWEEP_FENCE();
    }
#pragma unroll
    for (int j = N - 1; j >= 0; j--) {
        const T* a = ab + j * LSD_LDAB;
        const T v = b[j] * rd[j];
        b[j] = v;
#pragma unroll
        for (int d = 1; d <= 2 * LSD_KL; d++)
            if (j - d >= 0) b[j - d] = b[j - d] - v * a[2 * LSD_KL - d];
        SWEEP_FENCE();
    }
}

// Mixed forms: b <- X^-1 b with FP32 factors and one FP64-residual sweep
// against X in band storage, X(n,j) = xb[j*7 + 3+n-j].
template <int N>
__device__ __forceinline__ void solve_refined_reg(double (&b)[N], const float* ab32, const float* rd32,
                                                  const int* ip, const double* xb) {
    float w[N];
#pragma unroll
    for (int n = 0; n < N; n++) w[n] = (float)b[n];
    gbtrs_reg<N, float>(w, ab32, rd32, ip);
    // residual r = b - X x0, kept in b (row n reads b[n] only once)
#pragma unroll
    for (int n = 0; n < N; n++) {
        double acc = b[n];
#pragma unroll
        for (int j = n - LSD_KL; j <= n + LSD_KL; j++)
            if (j >= 0 && j < N) acc = acc - xb[j * 7 + LSD_KL + n - j] * (double)w[j];
        b[n] = acc;
        pin(b[n]);
    }
    // hand over row by row: FP32 residual out of b, x0 into b
    float r[N];
#pragma unroll
    for (int n = 0; n < N; n++) {
        r[n] = (float)b[n];
        b[n] = (double)w[n];
        pin(r[n]);
        pin(b[n]);
    }
    gbtrs_reg<N, float>(r, ab32, rd32, ip);
#pragma unroll
    for (int n = 0; n < N; n++) b[n] = b[n] + (double)r[n];
}

template <int N, bool MIX>
__device__ __forceinline__ void solve_reg(double (&b)[N], const Smem& m) {
    if constexpr (MIX) solve_refined_reg<N>(b, m.lu32, m.rd32, m.piv, m.xb);
    else gbtrs_reg<N, double>(b, m.lu, m.rd, m.piv);
}

// Reads the 8-row chunk starting at n0 of every harmonic column of the block
// (coll layout, 128-byte runs) and returns this lane's part of it in v;
// optionally copies the rows to `copy_to` (same layout) on the way.  All
// 64 lanes take part.
template <bool COPY>
__device__ __forceinline__ void load_chunk(const double2* __restrict__ src, double2* __restrict__ copy_to,
                                           const SolveArgs& a, int first, int nh, int n0, double* tile,
                                           double (&v)[RB]) {
    const int t = threadIdx.x;
    // keeps these loads in program order (a one-wave barrier is not a fence
    // for the compiler), so chunks are not all in flight at once
    asm volatile("" ::: "memory");
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int e = t + 64 * q;  // 32 harmonics x 8 rows
        const int hh = e >> 3, r = e & 7;
        if (hh < nh) {
            const size_t idx = (size_t)a.order[first + hh] * a.N + n0 + r;
            const double2 x = src[idx];
            if constexpr (COPY) copy_to[idx] = x;
            tile[r * TILE_LD + 2 * hh] = x.x;
            tile[r * TILE_LD + 2 * hh + 1] = x.y;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RB; r++) v[r] = tile[r * TILE_LD + t];
    __syncthreads();
}

// A one-wave workgroup's barrier does not stop the scheduler from sinking
// the per-chunk arithmetic below later chunks' loads (every chunk would then
// be live at once); fence each chunk.
#define CHUNK_FENCE() __builtin_amdgcn_sched_barrier(0)


// k_predict with the columns in registers (field f: 0 vT, 1 vP, 2 C, 3 X,
// 4 bT, 5 bP -> LU kinds 0,(2,1),3,4,5,6 and Y kinds 0,1,2,3,4,5).
// Register budgets: the FP64 corrector (column of N doubles plus the band
// window) fits two waves per SIMD; the predictor, with the multY window and
// the Green's-function sequence, and the mixed forms, which hold an FP32
// column as well, take one.
#define SOLVE_OCC(W) __attribute__((amdgpu_waves_per_eu(W, W)))

template <int N, bool MIX>
__global__ void __launch_bounds__(64) SOLVE_OCC(1) k_predict_reg(SolveArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int f = blockIdx.y;
    if (!a.on[f]) return;
    const int4 blk = a.blocks[blockIdx.x];
    const int l = blk.x, first = blk.y, nh = blk.z;
    const int t = threadIdx.x, hh = t >> 1;
    const bool act = hh < nh;
    const Smem m = carve(smem_raw, N, 0, MIX);
    double* tile = reinterpret_cast<double*>(smem_raw + align8(solve_smem_bytes(N, 0, MIX)));
    const int h = act ? a.order[first + hh] : 0;

    double b[N];
    // N -> registers, saved as the history term
#pragma unroll
    for (int c = 0; c < N / RB; c++) {
        double v[RB];
        load_chunk<true>(a.nn[f], a.hist[f], a, first, nh, c * RB, tile, v);
#pragma unroll
        for (int r = 0; r < RB; r++) {
            b[c * RB + r] = v[r];
            pin(b[c * RB + r]);
        }
        CHUNK_FENCE();
    }
    const bool has_y = f == 2 || f == 3 || l != 0;
    if (has_y) stage_y(a, m, f, l);
    if (f == 1) {
        if (l != 0) {
            stage_lu<MIX>(a, m, 2, l);
            stage_g(a, m, l);
        }
    } else if (has_y) {
        stage_lu<MIX>(a, m, f == 0 ? 0 : f + 1, l);
    }
    __syncthreads();
    if (f == 1 && l != 0) {
        // XGre first (velocity.F90:377-383)
        b[0] = b[N - 1] = 0.0;
        solve_reg<N, MIX>(b, m);
    }
    // b <- b + Y x (tim_multY), x streamed in chunks
    if (has_y) {
        double x[N];
        int done = 0;
#pragma unroll
        for (int c = 0; c < N / RB; c++) {
            double v[RB];
            load_chunk<false>(a.x[f], nullptr, a, first, nh, c * RB, tile, v);
#pragma unroll
            for (int r = 0; r < RB; r++) {
                x[c * RB + r] = v[r];
                pin(x[c * RB + r]);
            }
            const int upto = c == N / RB - 1 ? N : (c + 1) * RB - LSD_KL;
#pragma unroll
            for (int n = 0; n < N; n++) {
                if (n < done || n >= upto) continue;
                double acc = b[n];
#pragma unroll
                for (int j = n - LSD_KL; j <= n + LSD_KL; j++)
                    if (j >= 0 && j < N) acc = fma(m.y[j * 7 + LSD_KL + n - j], x[j], acc);
                b[n] = acc;
                pin(b[n]);
                SWEEP_FENCE();
            }
            done = upto;
            CHUNK_FENCE();
        }
    }
    b[0] = b[N - 1] = 0.0;
    if ((f == 2 || f == 3) && h == 0 && (t & 1) == 0) b[0] = 1.0;
    if (f == 1) {
        if (l != 0) {
            __syncthreads();
            stage_lu<MIX>(a, m, 1, l);
            __syncthreads();
            solve_reg<N, MIX>(b, m);
            // vel_gre_adjust (velocity.F90:445-465)
            const double* ig = a.invg + l * 4;
            const double bi = b[0], bo = b[N - 1];
            const double ca = ig[0] * bi + ig[2] * bo;
            const double cb = ig[1] * bi + ig[3] * bo;
#pragma unroll
            for (int n = 0; n < N; n++) b[n] = b[n] - ca * m.g[n] - cb * m.g[N + n];
        }
    } else if (has_y) {
        solve_reg<N, MIX>(b, m);
    }
    // registers -> x
    double2* dst = a.x[f];
#pragma unroll
    for (int c = 0; c < N / RB; c++) {
#pragma unroll
        for (int r = 0; r < RB; r++) tile[r * TILE_LD + t] = b[c * RB + r];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = t + 64 * q;
            const int h2 = e >> 3, r = e & 7;
            if (h2 < nh)
                dst[(size_t)a.order[first + h2] * N + c * RB + r] =
                    make_double2(tile[r * TILE_LD + 2 * h2], tile[r * TILE_LD + 2 * h2 + 1]);
        }
        __syncthreads();
        CHUNK_FENCE();
    }
}

// k_correct with the columns in registers; partial slots 0 vT, 1 vP, 2 C,
// 3 X, 4 bT, 5 icT, 6 bP, 7 icP.
template <int N, bool MIX>
__global__ void __launch_bounds__(64) SOLVE_OCC(MIX ? 1 : 2) k_correct_reg(SolveArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[2][64];
    const int f = blockIdx.y;
    const int t = threadIdx.x;
    double e2 = 0.0, eic = 0.0;
    if (a.on[f]) {
        const int4 blk = a.blocks[blockIdx.x];
        const int l = blk.x, first = blk.y, nh = blk.z;
        const int hh = t >> 1;
        const bool act = hh < nh;
        const Smem m = carve(smem_raw, N, 0, MIX);
        double* tile = reinterpret_cast<double*>(smem_raw + align8(solve_smem_bytes(N, 0, MIX)));
        const int h = act ? a.order[first + hh] : 0;
        double b[N];
        // d = c (N* - N), saved N <- N*
#pragma unroll
        for (int c = 0; c < N / RB; c++) {
            double vn[RB], vo[RB];
            load_chunk<false>(a.hist[f], nullptr, a, first, nh, c * RB, tile, vo);
            load_chunk<true>(a.nn[f], a.hist[f], a, first, nh, c * RB, tile, vn);
#pragma unroll
            for (int r = 0; r < RB; r++) {
                b[c * RB + r] = a.cimp * (vn[r] - vo[r]);
                pin(b[c * RB + r]);
            }
            CHUNK_FENCE();
        }
        const int kind = f == 0 ? 0 : (f == 1 ? 2 : f + 1);
        const bool solve = l != 0 || f == 2 || f == 3;
        if (solve) {
            stage_lu<MIX>(a, m, kind, l);
            if (f == 1) stage_g(a, m, l);
        }
        __syncthreads();
        b[0] = b[N - 1] = 0.0;
        // vel_setbc_Tor for (1,0) once tim_it /= 0
        if (act && f == 0 && l == 1 && a.hm[h] == 0 && (t & 1) == 0) b[0] = 0.0 - a.x[0][(size_t)h * N].x;
        if (solve) solve_reg<N, MIX>(b, m);
        if (f == 1 && solve) {
            b[0] = b[N - 1] = 0.0;
            __syncthreads();
            stage_lu<MIX>(a, m, 1, l);
            __syncthreads();
            solve_reg<N, MIX>(b, m);
            const double* ig = a.invg + l * 4;
            const double bi = b[0], bo = b[N - 1];
            const double ca = ig[0] * bi + ig[2] * bo;
            const double cb = ig[1] * bi + ig[3] * bo;
#pragma unroll
            for (int n = 0; n < N; n++) b[n] = b[n] - ca * m.g[n] - cb * m.g[N + n];
        }
        // x += d, with the correction norms of tim_addcorr
        double2* x = a.x[f];
#pragma unroll
        for (int c = 0; c < N / RB; c++) {
#pragma unroll
            for (int r = 0; r < RB; r++) tile[r * TILE_LD + t] = b[c * RB + r];
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int e = t + 64 * q;
                const int h2 = e >> 3, r = e & 7;
                if (h2 < nh) {
                    const size_t idx = (size_t)a.order[first + h2] * N + c * RB + r;
                    const double dx = tile[r * TILE_LD + 2 * h2], dy = tile[r * TILE_LD + 2 * h2 + 1];
                    const double2 o = x[idx];
                    x[idx] = make_double2(o.x + dx, o.y + dy);
                    e2 = e2 + dx * dx;
                    e2 = e2 + dy * dy;
                    if (f >= 4 && c == 0 && r == 0 && l != 0) {
                        eic = eic + dx * dx;
                        eic = eic + dy * dy;
                    }
                }
            }
            __syncthreads();
            CHUNK_FENCE();
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

bool solve_reg_supported(int N) { return N == 32 || N == 96 || N == 128; }

size_t solve_reg_lds_bytes(int N, bool mixed) {
    return align8(solve_smem_bytes(N, 0, mixed)) + sizeof(double) * RB * TILE_LD;
}

template <int N>
static void launch_n(bool corr, bool mixed, dim3 grid, size_t lds, hipStream_t st, const SolveArgs& a) {
    if (corr) {
        if (mixed) hipLaunchKernelGGL((k_correct_reg<N, true>), grid, dim3(64), lds, st, a);
        else hipLaunchKernelGGL((k_correct_reg<N, false>), grid, dim3(64), lds, st, a);
    } else {
        if (mixed) hipLaunchKernelGGL((k_predict_reg<N, true>), grid, dim3(64), lds, st, a);
        else hipLaunchKernelGGL((k_predict_reg<N, false>), grid, dim3(64), lds, st, a);
    }
}

int launch_solve_reg(Ctx& c, bool corr, const SolveArgs& a, dim3 grid) {
    const size_t lds = solve_reg_lds_bytes(c.N, c.mixed);
    switch (c.N) {
    case 32: launch_n<32>(corr, c.mixed, grid, lds, c.stream, a); break;
    case 96: launch_n<96>(corr, c.mixed, grid, lds, c.stream, a); break;
    case 128: launch_n<128>(corr, c.mixed, grid, lds, c.stream, a); break;
    default: set_error("no register solve kernel for this N"); return LSD_E_CONFIG;
    }
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

}  // namespace lsd
