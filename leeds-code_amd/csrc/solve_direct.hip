// Predictor / corrector solves for the radial resolutions the library is
// built for (N fixed at compile time).
//
// One 64-lane wave per block of up to 32 harmonics of one degree l (lane =
// harmonic, Re/Im part), so every lane needs the same factors, which are
// staged once in LDS and read at lane-uniform addresses.  Each lane loads a
// whole column straight into registers -- every row's load in flight at
// once -- runs dgbtrs on it, and writes the result back; no LDS
// transposition and no barrier between the loads and the sweeps, so the
// forward sweep starts on the first rows while the later ones are still
// arriving.
//  * Predictor, two passes: k_rhs_pred forms Y f + N with the boundary rows
//    of the *_setbc routines (one workgroup per block of harmonics of one
//    degree, with the Y band and the state columns staged in LDS) and saves
//    N; k_spred solves.  (Forming Y f in the solve kernel would hold the
//    state column beside the right-hand side's, past the 256 VGPRs of two
//    waves per SIMD at N = 96.)
//  * Corrector, one pass: k_scorr forms c (N* - N) in the register column
//    (N* loaded, the saved N streamed and replaced) and solves, so that
//    right-hand side never reaches HBM.
//
// The reference's predictor x* = X^-1 (Y x + N) and corrector
// x += X^-1 c (N* - N), with the poloidal velocity's XGre / XPol /
// vel_gre_adjust sequence (modules/velocity.F90:369-465), in FP64 for both
// the plain and the mixed forms: the sweeps are latency- and
// bandwidth-bound, and FP32 arithmetic issues at the FP64 rate here, so
// FP32 factors would only add the refinement sweep they need when X is
// poorly conditioned (large dt or fine radial grids).
#include <type_traits>

#include "solve_common.h"

namespace lsd {

struct DirArgs {
    int H, N, L;
    int on[6];           // field enabled: vT, vP, C, X, bT, bP
    double cimp;
    const int* hl;
    const int* hm;
    const int4* blocks;  // per block: degree l, first index into order, harmonic count
    const int* order;    // harmonics grouped by degree
    const double* lu;    // [NLU][L][N][LDAB]  FP64 factors
    const unsigned* pk;  // [NLU][L][npk] pivot offsets ip(j)-1-j, 2 bits per row, 16 rows per word
    int npk;
    const double* y;     // [NY][L][N][7]
    const double* gre;   // [L][2][N]
    const double* invg;  // [L][4]
    double2* x[6];       // state
    const double2* nn[6];  // current N
    double2* hist[6];      // saved N
    double2* rhs[6];     // [H][N] predictor right-hand sides
    double2* rhs2;       // poloidal velocity: the Y x part of the XPol right-hand side
    double* partial;     // [8][gridDim.x]
    const double* rot;   // inner-core rotation state (RotSlot), null without b_rot_tstep
    int t10;             // the inner wall is no-slip: T_10(r_i) row of vel_setbc_Tor
};

// v[f] for a field index that is the same across the workgroup, as scalar
// selects (indexing the kernel-argument array directly would copy it to
// scratch memory).
template <typename P>
__device__ __forceinline__ P pick(const P (&v)[6], int f) {
    P r = v[0];
#pragma unroll
    for (int k = 1; k < 6; k++)
        if (f == k) r = v[k];
    return r;
}

// Column access of a lane pair.  Lanes 2 hh and 2 hh + 1 hold the Re and Im
// columns of one harmonic, whose coll column stores (Re, Im) of each row as
// 16 bytes.  Each lane moves the 16 bytes of one row of a pair of rows (lane
// part: row n + part) and the pair exchanges halves, so each lane gets (or
// gives) its part of both rows with one request per row pair instead of one
// 8-byte request per row and lane.  The exchange is a DPP swap within the
// lane quad; every lane of the wave takes part in it.
__device__ __forceinline__ double swap_pair(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0xB1, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xB1, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// v0, v1: this lane's part of rows n, n + 1 (n even).
__device__ __forceinline__ void load_pair(const double2* col, int n, int part, double& v0, double& v1) {
    const double2 v = col[n + part];
    const double recv = swap_pair(part ? v.x : v.y);
    v0 = part ? recv : v.x;
    v1 = part ? v.y : recv;
}

__device__ __forceinline__ void store_pair(double2* col, int n, int part, double v0, double v1, bool act) {
    const double recv = swap_pair(part ? v0 : v1);
    if (act) col[n + part] = part ? make_double2(recv, v1) : make_double2(v0, recv);
}

// The same for a column of FP32 (Re, Im) rows (8 bytes per row): the saved N
// of the mixed forms, whose N comes from FP32 transforms.
__device__ __forceinline__ float swap_pair(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ void load_pair(const float2* col, int n, int part, double& v0, double& v1) {
    const float2 v = col[n + part];
    const float recv = swap_pair(part ? v.x : v.y);
    v0 = (double)(part ? recv : v.x);
    v1 = (double)(part ? v.y : recv);
}
__device__ __forceinline__ void store_pair(float2* col, int n, int part, double v0, double v1, bool act) {
    const float f0 = (float)v0, f1 = (float)v1;
    const float recv = swap_pair(part ? f0 : f1);
    if (act) col[n + part] = part ? make_float2(recv, f1) : make_float2(f0, recv);
}

// Storage of the saved N: FP64, or FP32 in the mixed forms.
template <bool H32>
using HistT = typename std::conditional<H32, float2, double2>::type;

// LU kind of the (main) solve of field f; the poloidal velocity's XGre solve.
__host__ __device__ inline int lu_kind(int f) { return f == 0 ? 0 : (f == 1 ? 2 : f + 1); }

// Row n of (band . x) accumulated onto acc in ascending column order
// (tim_multY); band stored per column, A(n,j) = band[j*7 + 3+n-j].
__device__ __forceinline__ double2 band_acc(double2 acc, const double* __restrict__ band,
                                            const double2* __restrict__ xc, int n, int N) {
    const int lo = n - LSD_KL > 0 ? n - LSD_KL : 0;
    const int hi = n + LSD_KL < N - 1 ? n + LSD_KL : N - 1;
    for (int j = lo; j <= hi; j++) {
        const double a = band[j * 7 + LSD_KL + n - j];
        const double2 xj = xc[j];
        acc.x = fma(a, xj.x, acc.x);
        acc.y = fma(a, xj.y, acc.y);
    }
    return acc;
}

// Predictor right-hand sides, one workgroup per solve block (harmonics of one
// degree l) and field: the Y band of (field, l) and the block's state
// columns are staged in LDS, so each point's 7-term band row reads LDS only.
constexpr int RHS_BLK = 1024;
constexpr int RHS_IT = (32 * 96 + RHS_BLK - 1) / RHS_BLK;  // points per thread at 32 harmonics, N <= 96

template <bool H32>
__global__ void __launch_bounds__(RHS_BLK) k_rhs_pred(DirArgs a, float inv_n) {
    extern __shared__ double2 xs_l[];  // [nh][N] state columns, then the band
    const int f = blockIdx.y;
    if (!a.on[f]) return;
    const int N = a.N;
    const int4 blk = a.blocks[blockIdx.x];
    const int l = blk.x, first = blk.y, nh = blk.z;
    double* band = reinterpret_cast<double*>(xs_l + (size_t)nh * N);
    const double2* x = pick(a.x, f);
    const double2* nn = pick(a.nn, f);
    HistT<H32>* hist = reinterpret_cast<HistT<H32>*>(pick(a.hist, f));
    double2* rhs = pick(a.rhs, f);
    const bool has_y = f == 2 || f == 3 || l != 0;
    const double* yb = a.y + ((size_t)f * a.L + l) * N * 7;
    for (int i = threadIdx.x; i < N * 7; i += RHS_BLK) band[i] = yb[i];
    const int ne = nh * N;
    // every point's state and N values are requested before any is used
    double2 nvs[RHS_IT];
#pragma unroll
    for (int k = 0; k < RHS_IT; k++) {
        const int e = threadIdx.x + RHS_BLK * k;
        if (e < ne) {
            const int hh = (int)(((float)e + 0.5f) * inv_n), n = e - hh * N;
            const size_t idx = (size_t)a.order[first + hh] * N + n;
            xs_l[e] = x[idx];
            nvs[k] = nn[idx];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < RHS_IT; k++) {
        const int e = threadIdx.x + RHS_BLK * k;
        if (e >= ne) break;
        const int hh = (int)(((float)e + 0.5f) * inv_n), n = e - hh * N;
        const int h = a.order[first + hh];
        const size_t idx = (size_t)h * N + n;
        const double keep = (n == 0 || n == N - 1) ? 0.0 : 1.0;
        const double2 nv = nvs[k];
        if constexpr (H32) hist[idx] = make_float2((float)nv.x, (float)nv.y);
        else hist[idx] = nv;
        const double2* xc = xs_l + hh * N;
        if (f == 1 && l != 0) {
            // XGre acts on N with zero boundary rows; the Y x part of the
            // XPol right-hand side (its boundary rows are zero too)
            rhs[idx] = make_double2(keep * nv.x, keep * nv.y);
            const double2 r = band_acc(make_double2(0.0, 0.0), band, xc, n, N);
            a.rhs2[idx] = make_double2(keep * r.x, keep * r.y);
        } else {
            double2 b = nv;
            if (has_y) b = band_acc(b, band, xc, n, N);
            b.x = keep * b.x;
            b.y = keep * b.y;
            if ((f == 2 || f == 3) && h == 0 && n == 0) b.x = 1.0;
            // vel_setbc_Tor: the rotating inner core's wall, T_10(r_i) = rot_omega r_i
            if (f == 0 && a.rot && l == 1 && n == 0 && a.hm[h] == 0) b.x = a.rot[ROT_OMEGA] * a.rot[ROT_RI];
            rhs[idx] = b;
        }
    }
}

// Ends one row of a sweep: the scheduler moves nothing across, so the
// factor reads issued for later rows stay where they are placed.
#define SWEEP_FENCE() __builtin_amdgcn_sched_barrier(0)
// Keeps the memory operations of one streamed batch from being merged with
// those of the next (each batch's loads would otherwise all be hoisted).
#define BATCH_FENCE() asm volatile("" ::: "memory")

// Materialises a value in its registers at this point of the program, so
// the arithmetic that produces it is not sunk below a later fence.
template <typename T>
__device__ __forceinline__ void pin(T& v) {
    asm volatile("" : "+v"(v));
}

constexpr int SWEEP_AHEAD = 2;  // rows of factors read from LDS ahead of the sweep

// dgbtrs('N') (timestep.F90:566-588 -> LAPACK dgbtrs, KL = KU = 3,
// LDAB = 10) on a register column; ab holds the factors of this (kind, l)
// in LDS, AB(i,j) (1-based) = ab[(j-1)*10 + i-1], rd[j] = 1 / U(j,j), and
// pk the pivot offsets ip(j)-1-j (0..3) packed 2 bits per row, read with
// scalar loads.  The row interchange is a select on the (scalar) offset,
// the backward sweep multiplies by the reciprocal of U(j,j) instead of
// dividing, and each row's factors are read SWEEP_AHEAD rows early, so no
// LDS or division latency sits on the serial chain.
template <int N, typename T>
__device__ __forceinline__ void gbtrs_reg(T (&b)[N], const T* ab, const T* rd, const unsigned* __restrict__ pk) {
    constexpr int NPK = (N + 15) / 16;
    unsigned pw[NPK];
#pragma unroll
    for (int k = 0; k < NPK; k++) pw[k] = pk[k];
    T fl[N][3];
#pragma unroll
    for (int j = 0; j < SWEEP_AHEAD; j++)
#pragma unroll
        for (int e = 0; e < 3; e++) fl[j][e] = ab[j * LSD_LDAB + 2 * LSD_KL + 1 + e];
#pragma unroll
    for (int j = 0; j < N - 1; j++) {
        if (j + SWEEP_AHEAD < N - 1) {
#pragma unroll
            for (int e = 0; e < 3; e++) fl[j + SWEEP_AHEAD][e] = ab[(j + SWEEP_AHEAD) * LSD_LDAB + 2 * LSD_KL + 1 + e];
        }
        const unsigned p = (pw[j >> 4] >> (2 * (j & 15))) & 3u;
        const T bj = b[j];
        T c1 = T(0), c2 = T(0), c3 = T(0);
        if (j + 1 < N) c1 = b[j + 1];
        if (j + 2 < N) c2 = b[j + 2];
        if (j + 3 < N) c3 = b[j + 3];
        const T y = p == 0 ? bj : (p == 1 ? c1 : (p == 2 ? c2 : c3));
        b[j] = y;
        if (j + 1 < N) b[j + 1] = (p == 1 ? bj : c1) - fl[j][0] * y;
        if (j + 2 < N) b[j + 2] = (p == 2 ? bj : c2) - fl[j][1] * y;
        if (j + 3 < N) b[j + 3] = (p == 3 ? bj : c3) - fl[j][2] * y;
        SWEEP_FENCE();
    }
    T fu[N][2 * LSD_KL + 1];  // [0] = 1/U(j,j), [d] = U(j-d, j)
#pragma unroll
    for (int j = N - 1; j > N - 1 - SWEEP_AHEAD; j--) {
        fu[j][0] = rd[j];
#pragma unroll
        for (int d = 1; d <= 2 * LSD_KL; d++) fu[j][d] = ab[j * LSD_LDAB + 2 * LSD_KL - d];
    }
#pragma unroll
    for (int j = N - 1; j >= 0; j--) {
        const int q = j - SWEEP_AHEAD;
        if (q >= 0) {
            fu[q][0] = rd[q];
#pragma unroll
            for (int d = 1; d <= 2 * LSD_KL; d++) fu[q][d] = ab[q * LSD_LDAB + 2 * LSD_KL - d];
        }
        const T v = b[j] * fu[j][0];
        b[j] = v;
#pragma unroll
        for (int d = 1; d <= 2 * LSD_KL; d++)
            if (j - d >= 0) b[j - d] = b[j - d] - v * fu[j][d];
        SWEEP_FENCE();
    }
}

// LDS of one block: the factors of one (kind, l) and the Green's functions.
template <int N>
struct DirSmem {
    double ab[N * LSD_LDAB];
    double rd[N];
    double g[2 * N];
};

// Copies the factors of (kind, l) into LDS, with the reciprocal diagonal.
// Callers synchronise before the data are read and before they are replaced.
template <int N>
__device__ __forceinline__ void stage_factors(const DirArgs& a, DirSmem<N>& s, int kind, int l) {
    const double* src = a.lu + ((size_t)kind * a.L + l) * N * LSD_LDAB;
    constexpr int NA = (N * LSD_LDAB + 63) / 64;
    double v[NA];
#pragma unroll
    for (int k = 0; k < NA; k++) {
        const int i = threadIdx.x + 64 * k;
        if (i < N * LSD_LDAB) v[k] = src[i];
    }
#pragma unroll
    for (int k = 0; k < NA; k++) {
        const int i = threadIdx.x + 64 * k;
        if (i < N * LSD_LDAB) {
            s.ab[i] = v[k];
            if (i % LSD_LDAB == 2 * LSD_KL) s.rd[i / LSD_LDAB] = 1.0 / v[k];
        }
    }
}

__device__ __forceinline__ const unsigned* pivots(const DirArgs& a, int kind, int l) {
    return a.pk + ((size_t)kind * a.L + l) * a.npk;
}

// Rows per batch when a second column (the state, or the poloidal
// velocity's second right-hand side) is streamed against the register
// column: 16 rows keep the column (2N VGPRs) and the batch within the 256
// VGPRs of two waves per SIMD at N = 96.
constexpr int DIR_BATCH = 16;

// Predictor solves: x* = X^-1 b for the right-hand sides of k_rhs_pred.  At
// most 256 VGPRs, so two waves share each SIMD.
template <int N>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) k_spred(DirArgs a) {
    static_assert(N % DIR_BATCH == 0, "N must be a multiple of the batch");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    DirSmem<N>& s = *reinterpret_cast<DirSmem<N>*>(smem_raw);
    constexpr int BT = DIR_BATCH;
    const int f = blockIdx.y;
    const int t = threadIdx.x;
    if (a.on[f]) {
        const int4 blk = a.blocks[blockIdx.x];
        const int l = blk.x, first = blk.y, nh = blk.z;
        const int hh = t >> 1, part = t & 1;
        const bool act = hh < nh;
        const int h = a.order[first + (act ? hh : 0)];
        const bool solve = f == 2 || f == 3 || l != 0;
        // this lane's part of the right-hand side, every row at once, issued
        // ahead of the factor staging so both are in flight together
        // (8-byte loads into the column: pair loads would hold 16 bytes per
        // row until every row has arrived)
        const double* src = reinterpret_cast<const double*>(pick(a.rhs, f) + (size_t)h * N) + part;
        double b[N];
#pragma unroll
        for (int n = 0; n < N; n++) b[n] = src[2 * n];
        if (solve) {
            stage_factors<N>(a, s, lu_kind(f), l);
            if (f == 1)
                for (int i = t; i < 2 * N; i += 64) s.g[i] = a.gre[(size_t)l * 2 * N + i];
        }
        __syncthreads();
        if (solve) gbtrs_reg<N, double>(b, s.ab, s.rd, pivots(a, lu_kind(f), l));
        double2* xcol = pick(a.x, f) + (size_t)h * N;
        if (f == 1 && solve) {
            // XPol right-hand side: XGre^-1 N + Y x inside, zero on the
            // boundary rows (where the second part is zero)
            const double2* r2col = a.rhs2 + (size_t)h * N;
#pragma unroll
            for (int c = 0; c < N / BT; c++) {
                double r2[BT];
#pragma unroll
                for (int r = 0; r < BT; r += 2) load_pair(r2col, c * BT + r, part, r2[r], r2[r + 1]);
#pragma unroll
                for (int r = 0; r < BT; r++) {
                    const int n = c * BT + r;
                    b[n] = (n == 0 || n == N - 1) ? r2[r] : b[n] + r2[r];
                    pin(b[n]);
                }
                BATCH_FENCE();
            }
            __syncthreads();
            stage_factors<N>(a, s, 1, l);
            __syncthreads();
            gbtrs_reg<N, double>(b, s.ab, s.rd, pivots(a, 1, l));
            // vel_gre_adjust (velocity.F90:445-465) on x*
            const double* ig = a.invg + l * 4;
            const double bi = b[0], bo = b[N - 1];
            const double ca = ig[0] * bi + ig[2] * bo;
            const double cb = ig[1] * bi + ig[3] * bo;
#pragma unroll
            for (int c = 0; c < N / BT; c++) {
#pragma unroll
                for (int r = 0; r < BT; r += 2) {
                    const int n = c * BT + r;
                    store_pair(xcol, n, part, b[n] - ca * s.g[n] - cb * s.g[N + n],
                               b[n + 1] - ca * s.g[n + 1] - cb * s.g[N + n + 1], act);
                }
                BATCH_FENCE();
            }
        } else {
#pragma unroll
            for (int n = 0; n < N; n += 2) store_pair(xcol, n, part, b[n], b[n + 1], act);
        }
    }
}

// Corrector: b = c (N* - N) formed in registers (N* in the column, the saved
// N streamed and replaced), x += X^-1 b, with the tim_addcorr norms; one
// wave per (solve block, field).
template <int N, bool H32>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) k_scorr(DirArgs a) {
    static_assert(N % DIR_BATCH == 0, "N must be a multiple of the batch");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[2][64];
    DirSmem<N>& s = *reinterpret_cast<DirSmem<N>*>(smem_raw);
    constexpr int BT = DIR_BATCH;
    const int f = blockIdx.y;
    const int t = threadIdx.x;
    double e2 = 0.0, eic = 0.0;
    if (a.on[f]) {
        const int4 blk = a.blocks[blockIdx.x];
        const int l = blk.x, first = blk.y, nh = blk.z;
        const int hh = t >> 1, part = t & 1;
        const bool act = hh < nh;
        const int h = a.order[first + (act ? hh : 0)];
        // X is solved except for the l = 0 modes of the velocity and the
        // magnetic field
        const bool solve = f == 2 || f == 3 || l != 0;
        double2* xcol = pick(a.x, f) + (size_t)h * N;
        const double2* ncol = pick(a.nn, f) + (size_t)h * N;
        HistT<H32>* hcol = reinterpret_cast<HistT<H32>*>(pick(a.hist, f)) + (size_t)h * N;
        // N*, every row at once, issued ahead of the factor staging (8-byte
        // loads into the column; the streamed columns below move by pairs)
        const double* nsrc = reinterpret_cast<const double*>(ncol) + part;
        double b[N];
#pragma unroll
        for (int n = 0; n < N; n++) b[n] = nsrc[2 * n];
        if (solve) {
            stage_factors<N>(a, s, lu_kind(f), l);
            if (f == 1)
                for (int i = t; i < 2 * N; i += 64) s.g[i] = a.gre[(size_t)l * 2 * N + i];
        }
        __syncthreads();
        // b = c (N* - N) inside, zero on the boundary rows; N* saved
        const double cimp = a.cimp;
#pragma unroll
        for (int c = 0; c < N / BT; c++) {
            double hv[BT];
#pragma unroll
            for (int r = 0; r < BT; r += 2) load_pair(hcol, c * BT + r, part, hv[r], hv[r + 1]);
#pragma unroll
            for (int r = 0; r < BT; r += 2) store_pair(hcol, c * BT + r, part, b[c * BT + r], b[c * BT + r + 1], act);
#pragma unroll
            for (int r = 0; r < BT; r++) {
                const int n = c * BT + r;
                const double keep = (n == 0 || n == N - 1) ? 0.0 : 1.0;
                b[n] = keep * (cimp * (b[n] - hv[r]));
                pin(b[n]);
            }
            BATCH_FENCE();
        }
        // vel_setbc_Tor for (1, 0) once tim_it /= 0
        if (a.t10 && f == 0 && l == 1 && a.hm[h] == 0 && part == 0)
            b[0] = (a.rot ? a.rot[ROT_OMEGA] * a.rot[ROT_RI] : 0.0) - xcol[0].x;
        if (solve) gbtrs_reg<N, double>(b, s.ab, s.rd, pivots(a, lu_kind(f), l));
        if (f == 1 && solve) {
            b[0] = 0.0;
            b[N - 1] = 0.0;
            __syncthreads();
            stage_factors<N>(a, s, 1, l);
            __syncthreads();
            gbtrs_reg<N, double>(b, s.ab, s.rd, pivots(a, 1, l));
            // vel_gre_adjust (velocity.F90:445-465) on the correction
            const double* ig = a.invg + l * 4;
            const double bi = b[0], bo = b[N - 1];
            const double ca = ig[0] * bi + ig[2] * bo;
            const double cb = ig[1] * bi + ig[3] * bo;
#pragma unroll
            for (int c = 0; c < N / BT; c++) {
                double xo[BT];
                double* xs = reinterpret_cast<double*>(xcol) + part;
#pragma unroll
                for (int r = 0; r < BT; r++) xo[r] = xs[2 * (c * BT + r)];
#pragma unroll
                for (int r = 0; r < BT; r++) {
                    const int n = c * BT + r;
                    const double d = b[n] - ca * s.g[n] - cb * s.g[N + n];
                    if (act) {
                        xs[2 * n] = xo[r] + d;
                        e2 = e2 + d * d;
                    }
                }
                BATCH_FENCE();
            }
        } else {
            // x += d, with the correction norms of tim_addcorr
#pragma unroll
            for (int c = 0; c < N / BT; c++) {
                double xo[BT];
                double* xs = reinterpret_cast<double*>(xcol) + part;
#pragma unroll
                for (int r = 0; r < BT; r++) xo[r] = xs[2 * (c * BT + r)];
#pragma unroll
                for (int r = 0; r < BT; r++) {
                    const int n = c * BT + r;
                    const double d = b[n];
                    xo[r] = xo[r] + d;
                    if (act) {
                        e2 = e2 + d * d;
                        if (f >= 4 && n == 0 && l != 0) eic = eic + d * d;
                    }
                }
#pragma unroll
                for (int r = 0; r < BT; r += 2) store_pair(xcol, c * BT + r, part, xo[r], xo[r + 1], act);
                BATCH_FENCE();
            }
        }
    }
    red[0][t] = e2;
    red[1][t] = eic;
    __syncthreads();
    if (t == 0) {
        double sum = 0.0, si = 0.0;
        for (int k = 0; k < 64; k++) {
            sum += red[0][k];
            si += red[1][k];
        }
        const int slot = f < 4 ? f : (f == 4 ? 4 : 6);
        a.partial[(size_t)slot * gridDim.x + blockIdx.x] = sum;
        if (f >= 4) a.partial[(size_t)(slot + 1) * gridDim.x + blockIdx.x] = si;
    }
}

bool solve_direct_supported(int N) { return N == 32 || N == 48 || N == 64 || N == 96; }

static DirArgs dir_args(Ctx& c) {
    DirArgs a;
    a.H = c.H;
    a.N = c.N;
    a.L = c.L;
    a.on[0] = 1;
    a.on[1] = 1;
    a.on[2] = c.cfg.cod;
    a.on[3] = c.cfg.comp;
    // with a conducting inner core the magnetic field has its own coupled solves
    a.on[4] = c.cfg.mag && !c.ic();
    a.on[5] = c.cfg.mag && !c.ic();
    a.cimp = c.cfg.implicit;
    a.rot = c.rot_wall() ? c.d_rot : nullptr;
    a.t10 = c.cfg.vel_bc <= 2;
    a.hl = c.d_hl;
    a.hm = c.d_hm;
    a.blocks = c.d_sblocks;
    a.order = c.d_sorder;
    a.lu = c.d_lu;
    a.pk = c.d_pivpk;
    a.npk = (c.N + 15) / 16;
    a.y = c.d_y;
    a.gre = c.d_gre;
    a.invg = c.d_invg;
    const int fx[6] = {LSD_VEL_TOR, LSD_VEL_POL, LSD_COD, LSD_COMP, LSD_MAG_TOR, LSD_MAG_POL};
    const int fn[6] = {LSD_N_VEL_TOR, LSD_N_VEL_POL, LSD_N_COD, LSD_N_COMP, LSD_N_MAG_TOR, LSD_N_MAG_POL};
    for (int k = 0; k < 6; k++) {
        a.x[k] = c.d_f[fx[k]];
        a.nn[k] = c.d_f[fn[k]];
        a.hist[k] = c.d_hist[k];
        a.rhs[k] = c.d_rhs[k];
    }
    a.rhs2 = c.d_rhs2;
    a.partial = c.d_partial;
    return a;
}

template <int N, bool PRED>
static void launch_sdir(Ctx& c, const DirArgs& a) {
    const dim3 grid(c.n_partial_blocks, 6);
    if (PRED) hipLaunchKernelGGL(k_spred<N>, grid, dim3(64), sizeof(DirSmem<N>), c.stream, a);
    else if (c.mixed) hipLaunchKernelGGL((k_scorr<N, true>), grid, dim3(64), sizeof(DirSmem<N>), c.stream, a);
    else hipLaunchKernelGGL((k_scorr<N, false>), grid, dim3(64), sizeof(DirSmem<N>), c.stream, a);
}

template <bool PRED>
static int launch_sdir_n(Ctx& c, const DirArgs& a) {
    switch (c.N) {
    case 32: launch_sdir<32, PRED>(c, a); break;
    case 48: launch_sdir<48, PRED>(c, a); break;
    case 64: launch_sdir<64, PRED>(c, a); break;
    case 96: launch_sdir<96, PRED>(c, a); break;
    default: set_error("no two-pass solve kernel for this N"); return LSD_E_CONFIG;
    }
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

// The predictor (pred: both passes) or the corrector on the context's
// stream.  Work (bytes) per field and coll element: the predictor reads the
// state and N, writes the saved N and the right-hand side, reads that and
// writes the state; the corrector reads N, the saved N and the state and
// writes the saved N and the state.
int launch_solve_direct(Ctx& c, bool pred) {
    const DirArgs a = dir_args(c);
    const int nf = 2 + (c.cfg.cod ? 1 : 0) + (c.cfg.comp ? 1 : 0) + (c.cfg.mag && !c.ic() ? 2 : 0);
    // bytes per element and field: predictor 64 + 32 (48 + 8 + 32 with the
    // FP32 saved N of the mixed forms), corrector 80 (64)
    const double per = pred ? (c.mixed ? 88.0 : 96.0) : (c.mixed ? 64.0 : 80.0);
    ProfScope ps(c, LSD_K_SOLVE, per * c.coll_elems() * nf);
    if (pred) {
        const size_t lds = sizeof(double2) * (size_t)c.solve_hb * c.N + sizeof(double) * 7 * (size_t)c.N;
        if (c.mixed)
            hipLaunchKernelGGL(k_rhs_pred<true>, dim3(c.n_partial_blocks, 6), dim3(RHS_BLK), lds, c.stream, a,
                               1.0f / (float)c.N);
        else
            hipLaunchKernelGGL(k_rhs_pred<false>, dim3(c.n_partial_blocks, 6), dim3(RHS_BLK), lds, c.stream, a,
                               1.0f / (float)c.N);
        LSD_HIP(hipGetLastError());
    }
    return pred ? launch_sdir_n<true>(c, a) : launch_sdir_n<false>(c, a);
}

}  // namespace lsd
