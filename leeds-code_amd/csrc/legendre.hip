// Gauss-Legendre synthesis and analysis as per-m contractions on the matrix
// cores (v_mfma_*_16x16x4: 16x4 by 4x16 per wave instruction).
//
// Synthesis (modules/transform.fftw3.F90:472-520, 719-818, 955-1000).  For
// an m-slot (harmonics h0..h0+K-1, l = m..L-1, equatorial parity of the
// k-th harmonic = k mod 2):
//     Y_p[c, j] = sum_{k: parity p} coef[h0+k, c] * T[h0+k, j]   j < Th2
//     F[j] = Y0 + Y1,  F[Th-1-j] = Y0 - Y1
// with columns c = 2n + (re, im).  The MFMA A operand is the coefficient
// block (16 columns x 4 harmonics, read straight from the coll layout), the B
// operand the table block (4 harmonics x 16 theta points, tables stored
// [h][j]); each lane ends with 4 consecutive columns (two complex values)
// of one theta point, so the epilogue writes 16 consecutive theta points
// per shell.  Parity-flipped tables (dtheta P) accumulate into the other
// bucket; the i and -i rotations of the vector/gradient terms swap the
// re/im column of a lane pair (one lane exchange).
//
// Analysis (modules/transform.fftw3.F90:527-602, 825-949).  The forward FFT
// leaves folded data G_p[j, c] (north+south, south-north).  A workgroup
// takes one slot and 32 columns, stages 16-theta chunks of G through LDS,
// and each wave contracts 16-harmonic tiles of one parity class against
// them: D[k, c] = sum_j Tw[j, k] G[j, c] with the quadrature-weighted tables
// stored [j][h] and each slot's even-k harmonics before its odd-k ones, so
// table reads and result writes are 128-byte runs.
#include <type_traits>

#include "lsd_internal.h"

namespace lsd {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef v4d acc;
    static __device__ __forceinline__ acc run(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <>
struct Mfma<float> {
    typedef v4f acc;
    static __device__ __forceinline__ acc run(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Block numbering shared by both contractions.  Work is grouped by table
// block (slot, theta chunk); the ncc column chunks of a group read the same
// table block, so they are given consecutive places in one XCD's dispatch
// sequence (workgroups go round-robin over the 8 XCDs) and share it through
// that XCD's L2.  Groups keep the longest-slot-first order within each XCD.
struct BlockId {
    int group, cc;
    bool valid;
};
__device__ __forceinline__ BlockId decode_block(int ngroups, int ncc) {
    const int b = blockIdx.x, xcd = b & 7, q = b >> 3;
    const int gl = q / ncc;
    BlockId r;
    r.cc = q - gl * ncc;
    r.group = gl * 8 + xcd;
    r.valid = r.group < ngroups;
    return r;
}
static unsigned grid_blocks(int ngroups, int ncc) { return (unsigned)(((ngroups + 7) / 8) * 8 * ncc); }

template <typename T>
__device__ __forceinline__ T partner(T v) {
    return __shfl_xor(v, 1);
}

// Up to LG_MAXJ transforms of one kind run as one launch (jobs): they read
// the same tables, so their column chunks are placed in the same table
// groups.
constexpr int LG_MAXJ = 6;

// v[j] for a job index that is the same across the workgroup, as scalar
// selects (indexing the kernel-argument array directly would copy it to
// scratch memory).
template <typename P>
__device__ __forceinline__ P pickj(const P (&v)[LG_MAXJ], int j) {
    P r = v[0];
#pragma unroll
    for (int k = 1; k < LG_MAXJ; k++)
        if (j == k) r = v[k];
    return r;
}

template <typename T>
struct SynArgs {
    int N, Th, Th2, M, H;
    int nth, ncc, ngroups;  // theta chunks, column chunks per job, nslot * nth
    int njob;
    const Slot* slots;
    const int* order;
    const T* tabA;   // [h][Th2]
    const T* tabB;   // second table (vector / gradient)
    const double* rinv;
    const T* in0[LG_MAXJ];  // coll [h][2N] (re, im interleaved)
    const T* in1[LG_MAXJ];
    T* out0[LG_MAXJ];       // fourier [n][M][Th] complex
    T* out1[LG_MAXJ];
};

// KIND 0: scalar (P);  1: vector (llP_ flipped, llsin1mP);  2: gradient
// (P_ flipped, sin1mP, scaled by 1/r).
template <typename T, int KIND>
__global__ void __launch_bounds__(64) k_syn(SynArgs<T> a) {
    typedef typename Mfma<T>::acc acc_t;
    constexpr int NOUT = KIND == 0 ? 1 : 2;
    constexpr int CB = KIND == 0 ? 2 : 1;
    const BlockId bid = decode_block(a.ngroups, a.ncc * a.njob);
    if (!bid.valid) return;
    const int sl = uniform(a.order[bid.group / a.nth]);
    const Slot s = a.slots[sl];
    const int lane = threadIdx.x, l16 = lane & 15, q4 = lane >> 4;
    const int j0 = (bid.group % a.nth) * 64;
    const int job = bid.cc / a.ncc;
    const int c0 = (bid.cc - job * a.ncc) * 16 * CB;
    const T* in0 = pickj(a.in0, job);
    const T* in1 = pickj(a.in1, job);
    const int C2 = 2 * a.N;
    acc_t Y[NOUT][2][CB][4];
#pragma unroll
    for (int o = 0; o < NOUT; o++)
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int cb = 0; cb < CB; cb++)
#pragma unroll
                for (int jb = 0; jb < 4; jb++) Y[o][p][cb][jb] = acc_t{0, 0, 0, 0};
    const bool odd_col = (l16 & 1);
    // Operands of one step of 8 harmonics (4 per parity): B = table values
    // at this lane's theta points, A = coefficients at this lane's columns.
    // The next step's operands are loaded while this step's products run.
    struct Ops {
        T ta[2][4], tb[2][4], x[2][CB], t[2][CB];
    };
    auto load = [&](Ops& o, int k0) {
#pragma unroll
        for (int p = 0; p < 2; p++) {
            const int k = k0 + 2 * q4 + p;
            const bool kv = k < s.len;
            const size_t hrow = (size_t)s.h0 + (kv ? k : 0);
#pragma unroll
            for (int jb = 0; jb < 4; jb++) {
                int j = j0 + 16 * jb + l16;
                int jc = j < a.Th2 ? j : a.Th2 - 1;
                o.ta[p][jb] = kv ? a.tabA[hrow * a.Th2 + jc] : T(0);
                o.tb[p][jb] = (KIND != 0 && kv) ? a.tabB[hrow * a.Th2 + jc] : T(0);
            }
#pragma unroll
            for (int cb = 0; cb < CB; cb++) {
                const int c = c0 + 16 * cb + l16;
                const bool cv = kv && c < C2;
                o.x[p][cb] = cv ? in0[hrow * C2 + c] : T(0);
                o.t[p][cb] = (KIND == 1 && cv) ? in1[hrow * C2 + c] : T(0);
            }
        }
    };
    auto step = [&](const Ops& o) {
#pragma unroll
        for (int p = 0; p < 2; p++) {
#pragma unroll
            for (int cb = 0; cb < CB; cb++) {
                const T x = o.x[p][cb];
                if (KIND == 0) {
#pragma unroll
                    for (int jb = 0; jb < 4; jb++) Y[0][p][cb][jb] = Mfma<T>::run(x, o.ta[p][jb], Y[0][p][cb][jb]);
                } else if (KIND == 2) {
                    // theta: c * dP (bucket 1-p);  phi: (i c) * mP/sin (bucket p)
                    T xr = partner(x);
                    xr = odd_col ? xr : -xr;
#pragma unroll
                    for (int jb = 0; jb < 4; jb++) {
                        Y[0][1 - p][cb][jb] = Mfma<T>::run(x, o.ta[p][jb], Y[0][1 - p][cb][jb]);
                        Y[NOUT - 1][p][cb][jb] = Mfma<T>::run(xr, o.tb[p][jb], Y[NOUT - 1][p][cb][jb]);
                    }
                } else {
                    const T t = o.t[p][cb];
                    // (-i t): re <- t_im, im <- -t_re ;  (i s): re <- -s_im, im <- s_re
                    T tr = partner(t), sr = partner(x);
                    tr = odd_col ? -tr : tr;
                    sr = odd_col ? sr : -sr;
#pragma unroll
                    for (int jb = 0; jb < 4; jb++) {
                        Y[0][1 - p][cb][jb] = Mfma<T>::run(x, o.ta[p][jb], Y[0][1 - p][cb][jb]);
                        Y[0][p][cb][jb] = Mfma<T>::run(tr, o.tb[p][jb], Y[0][p][cb][jb]);
                        Y[NOUT - 1][p][cb][jb] = Mfma<T>::run(sr, o.tb[p][jb], Y[NOUT - 1][p][cb][jb]);
                        Y[NOUT - 1][1 - p][cb][jb] = Mfma<T>::run(t, o.ta[p][jb], Y[NOUT - 1][1 - p][cb][jb]);
                    }
                }
            }
        }
    };
    Ops cur, nxt;
    load(cur, 0);
    for (int k0 = 0; k0 < s.len; k0 += 8) {
        if (k0 + 8 < s.len) load(nxt, k0 + 8);
        step(cur);
        cur = nxt;
    }
    // Output layout of v_mfma_*_16x16x4: the lane holds column j = l16 of
    // rows mrow(q4, r); rows are our columns c (re/im of a shell), columns
    // are theta points j0 + 16jb + l16.
    typedef typename std::conditional<sizeof(T) == 8, double2, float2>::type T2;
#pragma unroll
    for (int o = 0; o < NOUT; o++) {
        T2* F = reinterpret_cast<T2*>(o == 0 ? pickj(a.out0, job) : pickj(a.out1, job));
#pragma unroll
        for (int cb = 0; cb < CB; cb++) {
#pragma unroll
            for (int jb = 0; jb < 4; jb++) {
                const int j = j0 + 16 * jb + l16;
                T re0[2], im0[2], re1[2], im1[2];
                int shell[2];
                if constexpr (sizeof(T) == 8) {
                    // rows q4 + 4r: even q4 holds re, odd q4 the im of the same
                    // shells; pair with lane ^ 16, each lane writes two shells
                    const bool lo = !(q4 & 1);
#pragma unroll
                    for (int hp = 0; hp < 2; hp++) {
                        const int r0 = hp, r1 = 2 + hp;
                        // values this lane sends: the half its partner writes
                        const T s0 = lo ? Y[o][0][cb][jb][r1] : Y[o][0][cb][jb][r0];
                        const T s1 = lo ? Y[o][1][cb][jb][r1] : Y[o][1][cb][jb][r0];
                        const T g0 = __shfl_xor(s0, 16);
                        const T g1 = __shfl_xor(s1, 16);
                        const int r = lo ? r0 : r1;
                        re0[hp] = lo ? Y[o][0][cb][jb][r] : g0;
                        im0[hp] = lo ? g0 : Y[o][0][cb][jb][r];
                        re1[hp] = lo ? Y[o][1][cb][jb][r] : g1;
                        im1[hp] = lo ? g1 : Y[o][1][cb][jb][r];
                        shell[hp] = (q4 >> 1) + 2 * r;
                    }
                } else {
                    // rows 4q4 + r: the lane holds two whole complex values
#pragma unroll
                    for (int hp = 0; hp < 2; hp++) {
                        re0[hp] = Y[o][0][cb][jb][2 * hp];
                        im0[hp] = Y[o][0][cb][jb][2 * hp + 1];
                        re1[hp] = Y[o][1][cb][jb][2 * hp];
                        im1[hp] = Y[o][1][cb][jb][2 * hp + 1];
                        shell[hp] = 2 * q4 + hp;
                    }
                }
                if (j >= a.Th2) continue;
#pragma unroll
                for (int hp = 0; hp < 2; hp++) {
                    const int n = ((c0 + 16 * cb) >> 1) + shell[hp];
                    if (n >= a.N) continue;
                    const T f = KIND == 2 ? (T)a.rinv[n] : T(1);
                    T2* row = F + ((size_t)n * a.M + s.mfft) * a.Th;
                    T2 v;
                    v.x = f * (re0[hp] + re1[hp]);
                    v.y = f * (im0[hp] + im1[hp]);
                    row[j] = v;
                    if (a.Th - 1 - j >= a.Th2) {
                        v.x = f * (re0[hp] - re1[hp]);
                        v.y = f * (im0[hp] - im1[hp]);
                        row[a.Th - 1 - j] = v;
                    }
                }
            }
        }
    }
}

template <typename T>
static int syn_launch(Ctx& c, int kind, int njob, const T* tA, const T* tB, const void* const* in0,
                      const void* const* in1, void* const* out0, void* const* out1) {
    if (njob < 1 || njob > LG_MAXJ) {
        set_error("synthesis: 1..6 transforms per launch");
        return LSD_E_ARG;
    }
    SynArgs<T> a;
    a.N = c.N;
    a.Th = c.Th;
    a.Th2 = c.Th2;
    a.M = c.M;
    a.H = c.H;
    a.slots = c.d_slots;
    a.order = c.d_slot_order;
    a.tabA = tA;
    a.tabB = tB;
    a.rinv = c.d_r + 2 * c.N;
    for (int k = 0; k < LG_MAXJ; k++) {
        const int q = k < njob ? k : 0;
        a.in0[k] = reinterpret_cast<const T*>(in0[q]);
        a.in1[k] = in1 ? reinterpret_cast<const T*>(in1[q]) : nullptr;
        a.out0[k] = reinterpret_cast<T*>(out0[q]);
        a.out1[k] = out1 ? reinterpret_cast<T*>(out1[q]) : nullptr;
    }
    const int cb = kind == SYN_SCALAR ? 2 : 1;
    a.nth = (c.Th2 + 63) / 64;
    a.ncc = (2 * c.N + 16 * cb - 1) / (16 * cb);
    a.ngroups = c.nslot * a.nth;
    a.njob = njob;
    dim3 grid(grid_blocks(a.ngroups, a.ncc * njob));
    if (kind == SYN_SCALAR) hipLaunchKernelGGL((k_syn<T, 0>), grid, dim3(64), 0, c.stream, a);
    else if (kind == SYN_VECTOR) hipLaunchKernelGGL((k_syn<T, 1>), grid, dim3(64), 0, c.stream, a);
    else hipLaunchKernelGGL((k_syn<T, 2>), grid, dim3(64), 0, c.stream, a);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

// njob transforms of one kind in one launch:
// kind SYN_SCALAR: in0[k] -> out0[k] (table P)
// kind SYN_VECTOR: (in0[k] = s, in1[k] = t) -> (out0[k] = theta, out1[k] = phi)
// kind SYN_GRAD:   in0[k] -> (out0[k] = theta, out1[k] = phi)
// f32: FP32 tables, inputs and outputs (float2 spectral / Fourier layouts).
int launch_synthesis_group(Ctx& c, int kind, int njob, const void* const* in0, const void* const* in1,
                           void* const* out0, void* const* out1, bool f32) {
    size_t HT = (size_t)c.H * c.Th2;
    const double unit = 4.0 * (double)c.H * c.Th2 * c.N;  // flops per product unit
    ProfScope ps(c, LSD_K_SYNTH, unit * (kind == SYN_SCALAR ? 1 : kind == SYN_VECTOR ? 4 : 2) * njob);
    const size_t o0 = (kind == SYN_SCALAR ? 0 : kind == SYN_VECTOR ? 3 : 1) * HT;
    const size_t o1 = (kind == SYN_VECTOR ? 4 : 2) * HT;
    if (f32) return syn_launch<float>(c, kind, njob, c.d_tab32 + o0, c.d_tab32 + o1, in0, in1, out0, out1);
    return syn_launch<double>(c, kind, njob, c.d_tab + o0, c.d_tab + o1, in0, in1, out0, out1);
}

int launch_synthesis_p(Ctx& c, int kind, const void* in0, const void* in1, void* out0, void* out1, bool f32) {
    const void* i0[1] = {in0};
    const void* i1[1] = {in1};
    void* o0[1] = {out0};
    void* o1[1] = {out1};
    return launch_synthesis_group(c, kind, 1, i0, in1 ? i1 : nullptr, o0, out1 ? o1 : nullptr, f32);
}

int launch_synthesis(Ctx& c, int kind, const double2* in0, const double2* in1, double2* out0, double2* out1) {
    return launch_synthesis_p(c, kind, in0, in1, out0, out1, false);
}

// --------------------------------------------------------------- analysis

template <typename T>
struct AnaArgs {
    int N, Th2, M, H;
    int ncc, ngroups;  // column chunks per job, slots
    int njob;
    const Slot* slots;
    const int* order;
    const T* tP;   // gwP        [Th2][H] (parity-grouped h)
    const T* tL;   // gwllP_
    const T* tS;   // gwllsin1mP
    const T* g1[LG_MAXJ];  // folded [n][M][2][Th2] complex
    const T* g2[LG_MAXJ];
    const T* g3[LG_MAXJ];
    T* o1[LG_MAXJ];        // coll [h][2N]
    T* o2[LG_MAXJ];
    T* o3[LG_MAXJ];
};

constexpr int ANA_JC = 16;     // theta points per LDS chunk
constexpr int ANA_COLS = 32;   // columns per workgroup (16 shells)
constexpr int ANA_TPW = 4;     // max harmonic tiles per wave

// VEC: q, s, t from three folded fields (tra_rtp2qst); else one field
// (tra_phys2spec).  Workgroup: 4 waves, one slot, 32 columns.
template <typename T, bool VEC>
__global__ void __launch_bounds__(256) k_ana(AnaArgs<T> a) {
    typedef typename Mfma<T>::acc acc_t;
    constexpr int NF = VEC ? 3 : 1;
    constexpr int NO = VEC ? 3 : 1;
    // LDS: [field][parity][theta][column]
    __shared__ T G[NF][2][ANA_JC][ANA_COLS + 1];
    const BlockId bid = decode_block(a.ngroups, a.ncc * a.njob);
    if (!bid.valid) return;
    const int sl = uniform(a.order[bid.group]);
    const Slot s = a.slots[sl];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, q4 = lane >> 4;
    const int job = bid.cc / a.ncc;
    const int c0 = (bid.cc - job * a.ncc) * ANA_COLS;
    const T* gf[3] = {pickj(a.g1, job), pickj(a.g2, job), pickj(a.g3, job)};
    T* of[3] = {pickj(a.o1, job), pickj(a.o2, job), pickj(a.o3, job)};
    const int C2 = 2 * a.N;
    const int n_even = (s.len + 1) / 2, n_odd = s.len / 2;
    const int t_even = (n_even + 15) / 16, t_odd = (n_odd + 15) / 16;
    const int ntiles = t_even + t_odd;
    acc_t acc[ANA_TPW][NO][2];
#pragma unroll
    for (int t = 0; t < ANA_TPW; t++)
#pragma unroll
        for (int o = 0; o < NO; o++) acc[t][o][0] = acc[t][o][1] = acc_t{0, 0, 0, 0};
    const size_t gstride = (size_t)2 * a.Th2;  // complex values per (n, m)
    for (int jc0 = 0; jc0 < a.Th2; jc0 += ANA_JC) {
        __syncthreads();
        {   // stage: thread -> (shell in tile 0..15, theta in chunk 0..15)
            const int sh = tid >> 4, jj = tid & 15;
            const int n = (c0 >> 1) + sh;
            const int j = jc0 + jj;
            const bool ok = n < a.N && j < a.Th2;
#pragma unroll
            for (int f = 0; f < NF; f++) {
                const T* g = gf[f];
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    T re = T(0), im = T(0);
                    if (ok) {
                        size_t idx = (((size_t)n * a.M + s.mfft) * gstride + (size_t)p * a.Th2 + j) * 2;
                        re = g[idx];
                        im = g[idx + 1];
                    }
                    G[f][p][jj][2 * sh] = re;
                    G[f][p][jj][2 * sh + 1] = im;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int js = 0; js < ANA_JC / 4; js++) {
            const int jl = 4 * js + q4;  // chunk-local theta of this lane's K index
            const int j = jc0 + jl;
            const bool jv = j < a.Th2;
            const size_t trow = (size_t)(jv ? j : a.Th2 - 1) * a.H;
#pragma unroll
            for (int t = 0; t < ANA_TPW; t++) {
                const int tile = wave + 4 * t;
                if (tile >= ntiles) break;
                const int par = tile < t_even ? 0 : 1;
                const int kt = par ? tile - t_even : tile;  // tile within the class
                const int ncls = par ? n_odd : n_even;
                const int kc = kt * 16 + l16;               // class-local row of this lane
                const int ap = (kc < ncls) ? s.h0 + (par ? n_even : 0) + kc : s.h0;
                const T wP = jv ? a.tP[trow + ap] : T(0);
                const T wL = (VEC && jv) ? a.tL[trow + ap] : T(0);
                const T wS = (VEC && jv) ? a.tS[trow + ap] : T(0);
#pragma unroll
                for (int cb = 0; cb < 2; cb++) {
                    const int col = 16 * cb + l16;
                    const bool evn = !(l16 & 1);
                    acc[t][0][cb] = Mfma<T>::run(wP, G[0][par][jl][col], acc[t][0][cb]);
                    if constexpr (VEC) {
                        const T y2o = G[NF - 2][1 - par][jl][col];
                        const T y3s_p = G[NF - 1][par][jl][col ^ 1];
                        const T y2s_p = G[NF - 2][par][jl][col ^ 1];
                        const T y3o = G[NF - 1][1 - par][jl][col];
                        // s.re += Y2o.re gl + Y3s.im gs ; s.im += Y2o.im gl - Y3s.re gs
                        acc[t][NO - 2][cb] = Mfma<T>::run(wL, y2o, acc[t][NO - 2][cb]);
                        acc[t][NO - 2][cb] = Mfma<T>::run(wS, evn ? y3s_p : -y3s_p, acc[t][NO - 2][cb]);
                        // t.re = -Y2s.im gs + Y3o.re gl ; t.im = Y2s.re gs + Y3o.im gl
                        acc[t][NO - 1][cb] = Mfma<T>::run(wS, evn ? -y2s_p : y2s_p, acc[t][NO - 1][cb]);
                        acc[t][NO - 1][cb] = Mfma<T>::run(wL, y3o, acc[t][NO - 1][cb]);
                    }
                }
            }
        }
    }
    // the lane holds class-local rows 16*kt + mrow of columns c0 + 16cb + l16
    // (mrow = q4 + 4r for the FP64 instruction, 4q4 + r for FP32)
#pragma unroll
    for (int t = 0; t < ANA_TPW; t++) {
        const int tile = wave + 4 * t;
        if (tile >= ntiles) break;
        const int par = tile < t_even ? 0 : 1;
        const int kt = par ? tile - t_even : tile;
        const int ncls = par ? n_odd : n_even;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int kc = kt * 16 + (sizeof(T) == 8 ? q4 + 4 * r : 4 * q4 + r);
            if (kc >= ncls) continue;
            const size_t h = (size_t)s.h0 + 2 * kc + par;
#pragma unroll
            for (int cb = 0; cb < 2; cb++) {
                const int c = c0 + 16 * cb + l16;
                if (c >= C2) continue;
                const bool zero_im = (s.m == 0) && (c & 1);
                const size_t o = h * C2 + c;
                if constexpr (VEC) {
                    of[0][o] = acc[t][0][cb][r];  // the reference leaves q%Im(m=0) as summed
                    of[1][o] = zero_im ? T(0) : acc[t][1][cb][r];
                    of[2][o] = zero_im ? T(0) : acc[t][2][cb][r];
                } else {
                    of[0][o] = zero_im ? T(0) : acc[t][0][cb][r];
                }
            }
        }
    }
}

template <typename T>
static int ana_launch(Ctx& c, int kind, int njob, const T* tP, const T* tL, const T* tS, const void* const* g0,
                      const void* const* g1, const void* const* g2, void* const* o0, void* const* o1,
                      void* const* o2) {
    int maxlen = 0;
    for (auto& s : c.slots) maxlen = maxlen > s.len ? maxlen : s.len;
    int tiles = ((maxlen + 1) / 2 + 15) / 16 + (maxlen / 2 + 15) / 16;
    if (tiles > 4 * ANA_TPW) {
        set_error("analysis: L too large for the tile plan");
        return LSD_E_CONFIG;
    }
    if (njob < 1 || njob > LG_MAXJ) {
        set_error("analysis: 1..6 transforms per launch");
        return LSD_E_ARG;
    }
    AnaArgs<T> a;
    a.N = c.N;
    a.Th2 = c.Th2;
    a.M = c.M;
    a.H = c.H;
    a.slots = c.d_slots;
    a.order = c.d_slot_order;
    a.tP = tP;
    a.tL = tL;
    a.tS = tS;
    for (int k = 0; k < LG_MAXJ; k++) {
        const int q = k < njob ? k : 0;
        a.g1[k] = reinterpret_cast<const T*>(g0[q]);
        a.g2[k] = g1 ? reinterpret_cast<const T*>(g1[q]) : nullptr;
        a.g3[k] = g2 ? reinterpret_cast<const T*>(g2[q]) : nullptr;
        a.o1[k] = reinterpret_cast<T*>(o0[q]);
        a.o2[k] = o1 ? reinterpret_cast<T*>(o1[q]) : nullptr;
        a.o3[k] = o2 ? reinterpret_cast<T*>(o2[q]) : nullptr;
    }
    a.ncc = (2 * c.N + ANA_COLS - 1) / ANA_COLS;
    a.ngroups = c.nslot;
    a.njob = njob;
    dim3 grid(grid_blocks(a.ngroups, a.ncc * njob));
    if (kind == ANA_SCALAR) hipLaunchKernelGGL((k_ana<T, false>), grid, dim3(256), 0, c.stream, a);
    else hipLaunchKernelGGL((k_ana<T, true>), grid, dim3(256), 0, c.stream, a);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

// njob analyses of one kind in one launch: ANA_VECTOR (g0, g1, g2)[k] ->
// (o0, o1, o2)[k] = q, s, t; ANA_SCALAR g0[k] -> o0[k].
int launch_analysis_group(Ctx& c, int kind, int njob, const void* const* g0, const void* const* g1,
                          const void* const* g2, void* const* o0, void* const* o1, void* const* o2, bool f32) {
    size_t HT = (size_t)c.H * c.Th2;
    const double unit = 4.0 * (double)c.H * c.Th2 * c.N;
    ProfScope ps(c, LSD_K_ANAL, unit * (kind == ANA_SCALAR ? 1 : 5) * njob);
    if (f32)
        return ana_launch<float>(c, kind, njob, c.d_atab32, c.d_atab32 + HT, c.d_atab32 + 2 * HT, g0, g1, g2, o0,
                                 o1, o2);
    return ana_launch<double>(c, kind, njob, c.d_atab, c.d_atab + HT, c.d_atab + 2 * HT, g0, g1, g2, o0, o1, o2);
}

int launch_analysis_p(Ctx& c, int kind, const void* g0, const void* g1, const void* g2, void* o0, void* o1,
                      void* o2, bool f32) {
    const void* a0[1] = {g0};
    const void* a1[1] = {g1};
    const void* a2[1] = {g2};
    void* b0[1] = {o0};
    void* b1[1] = {o1};
    void* b2[1] = {o2};
    const bool v = kind == ANA_VECTOR;
    return launch_analysis_group(c, kind, 1, a0, v ? a1 : nullptr, v ? a2 : nullptr, b0, v ? b1 : nullptr,
                                 v ? b2 : nullptr, f32);
}

int launch_analysis(Ctx& c, int kind, const double2* g0, const double2* g1, const double2* g2, double2* o0,
                    double2* o1, double2* o2) {
    return launch_analysis_p(c, kind, g0, g1, g2, o0, o1, o2, false);
}

}  // namespace lsd
