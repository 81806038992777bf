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
// [h][j]).  Parity-flipped tables (dtheta P) accumulate into the other
// bucket; the i and -i rotations of the vector/gradient terms swap the
// re/im column of a lane pair (one lane exchange).  The Fourier output is
// stored [n][j][m] (m fastest), the order in which the physical-space pass
// reads a theta row, so a workgroup covers several consecutive m and
// writes each row's run of them at once.  The FP32 forms (k_syn4, below)
// load their operands 16 and 8 bytes per lane from theta-permuted tables.
//
// Analysis (modules/transform.fftw3.F90:527-602, 825-949).  The forward FFT
// leaves folded data G_p[j, c] (north+south, south-north), contracted as
// D[k, c] = sum_j Tw[j, k] G[j, c] with the quadrature-weighted tables.  In
// k_ana a workgroup takes one slot and 32 columns, stages 16-theta chunks of
// G through LDS, and each wave contracts 16-harmonic tiles of one parity
// class against them, with the tables stored [j][h] and each slot's even-k
// harmonics before its odd-k ones, so table reads and result writes are
// 128-byte runs.  The FP32 form k_ana4 (below, for even Th2) reads both
// operands 16 bytes per lane without LDS.
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
__device__ __forceinline__ BlockId decode_block_at(int b, int ngroups, int ncc) {
    const int xcd = b & 7, q = b >> 3;
    const int gl = q / ncc;
    BlockId r;
    r.cc = q - gl * ncc;
    r.group = gl * 8 + xcd;
    r.valid = r.group < ngroups;
    return r;
}
__device__ __forceinline__ BlockId decode_block(int ngroups, int ncc) {
    return decode_block_at(blockIdx.x, ngroups, ncc);
}
static unsigned grid_blocks(int ngroups, int ncc) { return (unsigned)(((ngroups + 7) / 8) * 8 * ncc); }

// The value of the neighbouring lane (lane ^ 1): a DPP quad permutation
// [1, 0, 3, 2], a register move rather than a trip through the LDS crossbar.
constexpr int DPP_SWAP_PAIRS = 0xB1;
__device__ __forceinline__ float partner(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), DPP_SWAP_PAIRS, 0xF, 0xF, false));
}
__device__ __forceinline__ double partner(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, DPP_SWAP_PAIRS, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), DPP_SWAP_PAIRS, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
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
    int N, Th, Th2, M, MP, H;
    int nth, ncc, nmg, njob;  // theta blocks, column blocks per job, m groups, jobs
    const Slot* slots;
    const int* mslot;         // [M] slot of each FFT index m
    const T* tabA;            // [h][Th2]
    const T* tabB;            // second table (vector / gradient)
    const double* rinv;
    const T* in0[LG_MAXJ];    // coll [h][2N] (re, im interleaved)
    const T* in1[LG_MAXJ];
    T* out0[LG_MAXJ];         // fourier [n][Th][MP] complex
    T* out1[LG_MAXJ];
    int out32;                // outputs rounded to FP32 (float2 Fourier layout)
};

constexpr int SY_TH = 16;     // theta points per tile (the MFMA's 16 columns)
constexpr int SY_COLS = 16;   // coefficient columns per tile (8 shells, re/im)
constexpr int SY_MG = 8;      // FFT indices m per workgroup
constexpr int SY_WAVES = 4;   // waves per workgroup (SY_MG / SY_WAVES m each)
constexpr int SY_SP = 2 * SY_MG + 1;  // LDS row of one (output, mirror, shell, theta): MG re/im pairs + pad
constexpr int SY_PD = 4;      // K steps per prefetch group

// Row (coefficient column) of accumulator element r held by lane group q4:
// v_mfma_f32_16x16x4f32 gives rows 4 q4 + r, the FP64 instruction q4 + 4 r.
template <typename T>
__device__ __forceinline__ int acc_row(int q4, int r) {
    return sizeof(T) == 8 ? q4 + 4 * r : 4 * q4 + r;
}

// Synthesis into the transposed Fourier layout [n][j][m].  A workgroup
// takes one tile of 16 theta points x 16 columns for SY_MG consecutive m:
// each wave contracts its m values one after another (KIND 0: scalar P;
// 1: vector (llP_ flipped, llsin1mP); 2: gradient (P_ flipped, sin1mP,
// scaled by 1/r)), with the operands of the next group of SY_PD K steps
// loaded while the current group's products run, and leaves its results in
// LDS; the workgroup then writes, for every (shell, theta row), the SY_MG
// values of consecutive m as one contiguous run.
template <typename T, int KIND>
__global__ void __launch_bounds__(64 * SY_WAVES) k_syn(SynArgs<T> a) {
    typedef typename Mfma<T>::acc acc_t;
    typedef typename std::conditional<sizeof(T) == 8, double2, float2>::type T2;
    constexpr int NOUT = KIND == 0 ? 1 : 2;
    __shared__ T St[NOUT * 2 * 8 * SY_TH * SY_SP];
    // groups (m group, job, column block), the theta blocks of one group
    // consecutive on one XCD: they read the same coefficient rows
    const BlockId bid = decode_block(a.nmg * a.njob * a.ncc, a.nth);
    if (!bid.valid) return;
    const int mg = bid.group / (a.njob * a.ncc);
    const int gjc = bid.group - mg * a.njob * a.ncc;
    const int job = gjc / a.ncc, cc = gjc - job * a.ncc, tb = bid.cc;
    const int j0 = tb * SY_TH, c0 = cc * SY_COLS, m0 = mg * SY_MG;
    const T* in0 = pickj(a.in0, job);
    const T* in1 = pickj(a.in1, job);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, l16 = lane & 15, q4 = lane >> 4;
    const int C2 = 2 * a.N;
    const bool odd_col = (l16 & 1);
    const int jc = j0 + l16 < a.Th2 ? j0 + l16 : a.Th2 - 1;
    const int c = c0 + l16;
    struct Ops {
        T ta[2], tb[2], x[2], t[2];
    };
#pragma unroll 1
    for (int mi = wave; mi < SY_MG; mi += SY_WAVES) {
        const int m = m0 + mi;
        acc_t Y[NOUT][2];
#pragma unroll
        for (int o = 0; o < NOUT; o++) Y[o][0] = Y[o][1] = acc_t{0, 0, 0, 0};
        if (m < a.M) {
            const Slot s = a.slots[a.mslot[m]];
            auto load = [&](Ops& op, int k0) {
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    const int k = k0 + 2 * q4 + p;
                    const bool kv = k < s.len;
                    const size_t hrow = (size_t)s.h0 + (kv ? k : 0);
                    op.ta[p] = kv ? a.tabA[hrow * a.Th2 + jc] : T(0);
                    op.tb[p] = (KIND != 0 && kv) ? a.tabB[hrow * a.Th2 + jc] : T(0);
                    const bool cv = kv && c < C2;
                    op.x[p] = cv ? in0[hrow * C2 + c] : T(0);
                    op.t[p] = (KIND == 1 && cv) ? in1[hrow * C2 + c] : T(0);
                }
            };
            auto step = [&](const Ops& op) {
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    const T x = op.x[p];
                    if (KIND == 0) {
                        Y[0][p] = Mfma<T>::run(x, op.ta[p], Y[0][p]);
                    } else if (KIND == 2) {
                        // theta: c * dP (bucket 1-p);  phi: (i c) * mP/sin (bucket p)
                        T xr = partner(x);
                        xr = odd_col ? xr : -xr;
                        Y[0][1 - p] = Mfma<T>::run(x, op.ta[p], Y[0][1 - p]);
                        Y[NOUT - 1][p] = Mfma<T>::run(xr, op.tb[p], Y[NOUT - 1][p]);
                    } else {
                        const T t = op.t[p];
                        // (-i t): re <- t_im, im <- -t_re ;  (i s): re <- -s_im, im <- s_re
                        T tr = partner(t), sr = partner(x);
                        tr = odd_col ? -tr : tr;
                        sr = odd_col ? sr : -sr;
                        Y[0][1 - p] = Mfma<T>::run(x, op.ta[p], Y[0][1 - p]);
                        Y[0][p] = Mfma<T>::run(tr, op.tb[p], Y[0][p]);
                        Y[NOUT - 1][p] = Mfma<T>::run(sr, op.tb[p], Y[NOUT - 1][p]);
                        Y[NOUT - 1][1 - p] = Mfma<T>::run(t, op.ta[p], Y[NOUT - 1][1 - p]);
                    }
                }
            };
            // groups of SY_PD steps (8 harmonics each), two groups in flight
            Ops ga[SY_PD], gb[SY_PD];
#pragma unroll
            for (int u = 0; u < SY_PD; u++) load(ga[u], 8 * u);
            for (int k0 = 0; k0 < s.len; k0 += 16 * SY_PD) {
                const int k1 = k0 + 8 * SY_PD;
                if (k1 < s.len) {
#pragma unroll
                    for (int u = 0; u < SY_PD; u++) load(gb[u], k1 + 8 * u);
                }
#pragma unroll
                for (int u = 0; u < SY_PD; u++)
                    if (k0 + 8 * u < s.len) step(ga[u]);
                if (k1 >= s.len) break;
                if (k1 + 8 * SY_PD < s.len) {
#pragma unroll
                    for (int u = 0; u < SY_PD; u++) load(ga[u], k1 + 8 * SY_PD + 8 * u);
                }
#pragma unroll
                for (int u = 0; u < SY_PD; u++)
                    if (k1 + 8 * u < s.len) step(gb[u]);
            }
        }
        // F[j] = Y0 + Y1, F[Th-1-j] = Y0 - Y1; lane: theta l16, rows acc_row
#pragma unroll
        for (int o = 0; o < NOUT; o++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = acc_row<T>(q4, r), sh = row >> 1, part = row & 1;
                const T y0 = Y[o][0][r], y1 = Y[o][1][r];
                T* d = St + (((o * 2 + 0) * 8 + sh) * SY_TH + l16) * SY_SP + 2 * mi + part;
                d[0] = y0 + y1;
                d[8 * SY_TH * SY_SP] = y0 - y1;
            }
    }
    __syncthreads();
    constexpr int NITEM = NOUT * 2 * 8 * SY_TH * SY_MG;
    for (int idx = tid; idx < NITEM; idx += 64 * SY_WAVES) {
        const int mi = idx % SY_MG;
        int rest = idx / SY_MG;
        const int th = rest % SY_TH;
        rest /= SY_TH;
        const int sh = rest % 8;
        rest /= 8;
        const int mir = rest & 1, o = rest >> 1;
        const int n = (c0 >> 1) + sh, j = j0 + th;
        if (n >= a.N || j >= a.Th2) continue;
        const int jrow = mir ? a.Th - 1 - j : j;
        if (mir && jrow < a.Th2) continue;
        const T* sp = St + (((o * 2 + mir) * 8 + sh) * SY_TH + th) * SY_SP + 2 * mi;
        const T f = KIND == 2 ? (T)a.rinv[n] : T(1);
        T2 v;
        v.x = f * sp[0];
        v.y = f * sp[1];
        T* out = o == 0 ? pickj(a.out0, job) : pickj(a.out1, job);
        const size_t at = ((size_t)n * a.Th + jrow) * a.MP + m0 + mi;
        if (a.out32) reinterpret_cast<float2*>(out)[at] = make_float2((float)v.x, (float)v.y);
        else reinterpret_cast<T2*>(out)[at] = v;
    }
}

template <typename T>
static int syn_launch(Ctx& c, int kind, int njob, const T* tA, const T* tB, const void* const* in0,
                      const void* const* in1, void* const* out0, void* const* out1, bool out32 = false) {
    if (njob < 1 || njob > LG_MAXJ) {
        set_error("synthesis: 1..6 transforms per launch");
        return LSD_E_ARG;
    }
    SynArgs<T> a;
    a.N = c.N;
    a.Th = c.Th;
    a.Th2 = c.Th2;
    a.M = c.M;
    a.MP = c.MP;
    a.H = c.H;
    a.slots = c.d_slots;
    a.mslot = c.d_mslot;
    a.tabA = tA;
    a.tabB = tB;
    a.rinv = c.d_r + 2 * c.N;
    a.out32 = out32 ? 1 : 0;
    for (int k = 0; k < LG_MAXJ; k++) {
        const int q = k < njob ? k : 0;
        a.in0[k] = reinterpret_cast<const T*>(in0[q]);
        a.in1[k] = in1 ? reinterpret_cast<const T*>(in1[q]) : nullptr;
        a.out0[k] = reinterpret_cast<T*>(out0[q]);
        a.out1[k] = out1 ? reinterpret_cast<T*>(out1[q]) : nullptr;
    }
    a.nth = (c.Th2 + SY_TH - 1) / SY_TH;
    a.ncc = (2 * c.N + SY_COLS - 1) / SY_COLS;
    a.nmg = (c.M + SY_MG - 1) / SY_MG;
    a.njob = njob;
    dim3 grid(grid_blocks(a.nmg * njob * a.ncc, a.nth));
    dim3 blk(64 * SY_WAVES);
    if (kind == SYN_SCALAR) hipLaunchKernelGGL((k_syn<T, 0>), grid, blk, 0, c.stream, a);
    else if (kind == SYN_VECTOR) hipLaunchKernelGGL((k_syn<T, 1>), grid, blk, 0, c.stream, a);
    else hipLaunchKernelGGL((k_syn<T, 2>), grid, blk, 0, c.stream, a);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

static int syn4_launch(Ctx& c, int kind, int njob, const float* tA, const float* tB, const void* const* in0,
                       const void* const* in1, void* const* out0, void* const* out1);

// njob transforms of one kind in one launch:
// kind SYN_SCALAR: in0[k] -> out0[k] (table P)
// kind SYN_VECTOR: (in0[k] = s, in1[k] = t) -> (out0[k] = theta, out1[k] = phi)
// kind SYN_GRAD:   in0[k] -> (out0[k] = theta, out1[k] = phi)
// f32: FP32 tables, inputs and outputs (float2 spectral / Fourier layouts),
// on the FP32 kernel below.  out32 (with f32 false): FP64 tables, inputs and
// sums, outputs rounded to FP32.
int launch_synthesis_group(Ctx& c, int kind, int njob, const void* const* in0, const void* const* in1,
                           void* const* out0, void* const* out1, bool f32, bool out32) {
    size_t HT = (size_t)c.H * c.Th2;
    const double unit = 4.0 * (double)c.H * c.Th2 * c.N;  // flops per product unit
    ProfScope ps(c, LSD_K_SYNTH, unit * (kind == SYN_SCALAR ? 1 : kind == SYN_VECTOR ? 4 : 2) * njob);
    const size_t o0 = (kind == SYN_SCALAR ? 0 : kind == SYN_VECTOR ? 3 : 1) * HT;
    const size_t o1 = (kind == SYN_VECTOR ? 4 : 2) * HT;
    if (f32) {
        if (njob < 1 || njob > LG_MAXJ) {
            set_error("synthesis: 1..6 transforms per launch");
            return LSD_E_ARG;
        }
        if (c.cfg.precision == LSD_MIXED_BF16X3) return synb_launch(c, kind, njob, in0, in1, out0, out1);
        const size_t HTp = (size_t)c.Hp * c.Th2p;
        const size_t p0 = (kind == SYN_SCALAR ? 0 : kind == SYN_VECTOR ? 3 : 1) * HTp;
        const size_t p1 = (kind == SYN_VECTOR ? 4 : 2) * HTp;
        return syn4_launch(c, kind, njob, c.d_tab32 + p0, c.d_tab32 + p1, in0, in1, out0, out1);
    }
    return syn_launch<double>(c, kind, njob, c.d_tab + o0, c.d_tab + o1, in0, in1, out0, out1, out32);
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

// ------------------------------------------------- FP32 synthesis (mixed forms)
//
// The same contractions with operands loaded 16 and 8 bytes per lane.  The
// FP32 tables are stored per slot (rows padded with zeros to a multiple of
// 8 harmonics) in rows of Th2p = Th2 rounded up to 64 theta points, each
// block of 64 permuted so that position 4 l + t holds theta 16 t + l: one
// 16-byte load per lane carries the B operands of four 16-theta MFMA tiles
// (theta 16 t + l16, t = 0..3).  A lane's 8-byte load of a coefficient row
// is the (re, im) pair of one shell, the A operands of the two column tiles
// (u = re, im), so the i rotations of the vector and gradient terms are a
// swap within the lane.  A wave computes 64 theta x 16 shells for one m; per
// two harmonics it runs 8 (scalar), 16 (gradient) or 32 (vector) MFMAs
// against two (three, four) loads.  The waves of a workgroup take
// consecutive m (4 scalar, 8 vector and gradient) and stage their results
// through LDS, so each (shell, theta row) is written as one run of 32 or 64
// bytes.
constexpr int S4_TH = 64;                  // theta points per wave tile
constexpr int S4_SH = 16;                  // shells per wave tile
constexpr int S4_MG = 16;                  // m per block group (one 128-byte line of a Fourier row)
// consecutive m per workgroup (one per wave), 4 or 8: a kernel variant
// (LSD_V_SYN_*); one hemisphere of one output is staged in LDS per pass
template <int MQ>
constexpr int S4_NMIR = 1;
// operand buffers (one step of 8 harmonics each), each refilled three steps
// ahead: at L = 170, N = 96 three run the scalar kind in 105 us against
// 114-117 us with four or six, and six spill the vector and gradient kinds
constexpr int S4_NB = 3;
// LDS staging row of one shell: 64 theta x (MQ + 1) float2, the m of one
// theta padded by one slot, so the 16 theta of a lane group's writes (a
// stride of 2 MQ + 2 banks) land in 16 distinct bank pairs; plus one slot
// between shells
template <int MQ>
constexpr int S4_MQP = MQ + 1;
template <int MQ>
constexpr int S4_SHS = S4_TH * S4_MQP<MQ> + 1;

struct Syn4Args {
    int N, Th, Th2, Th2p, M, MP, H;
    int nth, nsb, nmo, njob;  // theta tiles, shell blocks, groups of 4 m quads, jobs
    const Slot* slots;
    const int* mslot;
    const int* t0;            // [nslot] first table row of each slot
    const float* tabA;        // [Hp][Th2p]
    const float* tabB;
    const double* rinv;
    const float* in0[LG_MAXJ];  // FP32 coll [h][2N]
    const float* in1[LG_MAXJ];
    float* out0[LG_MAXJ];       // FP32 fourier [n][Th][MP] complex
    float* out1[LG_MAXJ];
};

struct S4Ops {
    float4 ta[2], tb[2];  // tables, harmonics k and k+1 of this lane group
    float2 x[2], t[2];    // coefficients (re, im) of the lane's shell
};

__device__ __forceinline__ v4f mfma4(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <int MQ>
constexpr int S4_STAGE = S4_NMIR<MQ> * S4_SH * S4_SHS<MQ>;

// One work item (block number b of the static numbering): groups (16 m,
// job, shell block); members (m quad, theta tile), the theta tiles of one
// m quad consecutive: they read the same coefficients.
template <int KIND, int MQ>
__device__ __forceinline__ void syn4_item(const Syn4Args& a, int b, float2* __restrict__ St) {
    constexpr int NOUT = KIND == 0 ? 1 : 2;
    constexpr int NB = S4_NB;
    constexpr int NMIR = S4_NMIR<MQ>, SHS = S4_SHS<MQ>;
    const BlockId bid = decode_block_at(b, a.nmo * a.njob * a.nsb, (S4_MG / MQ) * a.nth);
    if (!bid.valid) return;
    const int mo = bid.group / (a.njob * a.nsb);
    const int gjs = bid.group - mo * a.njob * a.nsb;
    const int job = gjs / a.nsb, sb = gjs - job * a.nsb;
    const int mq = mo * (S4_MG / MQ) + bid.cc / a.nth, tt = bid.cc % a.nth;
    const int m0 = mq * MQ;
    if (m0 >= a.M) return;
    const int jb = tt * S4_TH, n0 = sb * S4_SH;
    const float* in0 = pickj(a.in0, job);
    const float* in1 = pickj(a.in1, job);
    const int tid = threadIdx.x, wave = uniform(tid >> 6), lane = tid & 63, l16 = lane & 15, q4 = lane >> 4;
    const int C2 = 2 * a.N;
    const int ncl = min(n0 + l16, a.N - 1);  // rows of shells >= N are computed and not stored
    const int m = m0 + wave;
    v4f Y[NOUT][2][2][4];  // [output][parity bucket][re, im][theta tile]
#pragma unroll
    for (int o = 0; o < NOUT; o++)
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int u = 0; u < 2; u++)
#pragma unroll
                for (int t = 0; t < 4; t++) Y[o][p][u][t] = v4f{0, 0, 0, 0};
    if (m < a.M) {
        const int si = a.mslot[m];
        const Slot s = a.slots[si];
        const int rows = (s.len + 7) & ~7;
        const size_t tb0 = (size_t)a.t0[si] * a.Th2p + jb + 4 * l16;
        // rows are clamped rather than branched on, so no copy at a join
        // waits for the loads; steps past the slot's rows are never run
        auto load = [&](S4Ops& op, int k0) {
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const int k = min(k0 + 2 * q4 + p, rows - 1);
                op.ta[p] = *reinterpret_cast<const float4*>(a.tabA + tb0 + (size_t)k * a.Th2p);
                if (KIND != 0) op.tb[p] = *reinterpret_cast<const float4*>(a.tabB + tb0 + (size_t)k * a.Th2p);
                // padding rows (k >= len) have zero tables; their coefficient
                // row is any valid one
                const size_t h = (size_t)min(s.h0 + k, a.H - 1);
                op.x[p] = *reinterpret_cast<const float2*>(in0 + h * C2 + 2 * ncl);
                if (KIND == 1) op.t[p] = *reinterpret_cast<const float2*>(in1 + h * C2 + 2 * ncl);
            }
        };
        auto step = [&](const S4Ops& op) {
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const float ta[4] = {op.ta[p].x, op.ta[p].y, op.ta[p].z, op.ta[p].w};
                const float x[2] = {op.x[p].x, op.x[p].y};
                if constexpr (KIND == 0) {
#pragma unroll
                    for (int u = 0; u < 2; u++)
#pragma unroll
                        for (int t = 0; t < 4; t++) Y[0][p][u][t] = mfma4(x[u], ta[t], Y[0][p][u][t]);
                } else if constexpr (KIND == 2) {
                    // theta: c * dP (bucket 1-p);  phi: (i c) * mP/sin (bucket p)
                    const float tb[4] = {op.tb[p].x, op.tb[p].y, op.tb[p].z, op.tb[p].w};
                    const float xr[2] = {-x[1], x[0]};
#pragma unroll
                    for (int u = 0; u < 2; u++)
#pragma unroll
                        for (int t = 0; t < 4; t++) {
                            Y[0][1 - p][u][t] = mfma4(x[u], ta[t], Y[0][1 - p][u][t]);
                            Y[1][p][u][t] = mfma4(xr[u], tb[t], Y[1][p][u][t]);
                        }
                } else {
                    // (-i t): re <- t_im, im <- -t_re ;  (i s): re <- -s_im, im <- s_re
                    const float tb[4] = {op.tb[p].x, op.tb[p].y, op.tb[p].z, op.tb[p].w};
                    const float tv[2] = {op.t[p].x, op.t[p].y};
                    const float tr[2] = {tv[1], -tv[0]}, sr[2] = {-x[1], x[0]};
#pragma unroll
                    for (int u = 0; u < 2; u++)
#pragma unroll
                        for (int t = 0; t < 4; t++) {
                            Y[0][1 - p][u][t] = mfma4(x[u], ta[t], Y[0][1 - p][u][t]);
                            Y[0][p][u][t] = mfma4(tr[u], tb[t], Y[0][p][u][t]);
                            Y[1][p][u][t] = mfma4(sr[u], tb[t], Y[1][p][u][t]);
                            Y[1][1 - p][u][t] = mfma4(tv[u], ta[t], Y[1][1 - p][u][t]);
                        }
                }
            }
        };
        // a ring of NB one-step buffers (8 harmonics each): a buffer is
        // refilled, NB steps ahead, as soon as its step has been issued.  The
        // refill is unconditional (rows clamped), so no join makes the loads
        // wait early, and scheduling barriers keep the buffers' loads in ring
        // order, so each step waits for its own buffer only.
        S4Ops buf[NB];
#pragma unroll
        for (int u = 0; u < NB; u++) {
            load(buf[u], 8 * u);
            __builtin_amdgcn_sched_barrier(0);
        }
        for (int k0 = 0; k0 < s.len; k0 += 8 * NB) {
#pragma unroll
            for (int u = 0; u < NB; u++) {
                if (k0 + 8 * u < s.len) step(buf[u]);
                load(buf[u], k0 + 8 * u + 8 * NB);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
    // F[j] = Y0 + Y1 (north), F[Th-1-j] = Y0 - Y1 (south), NMIR hemispheres
    // of one output per pass, into LDS [hemisphere][shell][theta][m] (lane:
    // theta 16 t + l16, shells 4 q4 + r), then one run of MQ m per (shell,
    // theta row), written 16 bytes per lane by adjacent lanes (one request)
    constexpr int NPASS = NOUT * (2 / NMIR), NQ = MQ / 2;
#pragma unroll
    for (int pass = 0; pass < NPASS; pass++) {
        const int o = pass / (2 / NMIR), mir0 = NMIR == 2 ? 0 : pass % 2;
        if (pass) __syncthreads();  // the previous pass's reads are done
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int sh = 4 * q4 + r, th = 16 * t + l16;
                const float y0r = Y[o][0][0][t][r], y1r = Y[o][1][0][t][r];
                const float y0i = Y[o][0][1][t][r], y1i = Y[o][1][1][t][r];
#pragma unroll
                for (int mm = 0; mm < NMIR; mm++)
                    St[(mm * S4_SH + sh) * SHS + th * S4_MQP<MQ> + wave] =
                        (mir0 + mm) ? make_float2(y0r - y1r, y0i - y1i) : make_float2(y0r + y1r, y0i + y1i);
            }
        __syncthreads();
        float* out = o == 0 ? pickj(a.out0, job) : pickj(a.out1, job);
        for (int idx = tid; idx < NQ * NMIR * S4_SH * S4_TH; idx += 64 * MQ) {
            const int qq = idx % NQ, item = idx / NQ;
            const int th = item % S4_TH, rest = item / S4_TH;
            const int sh = rest % S4_SH, mm = rest / S4_SH, mir = mir0 + mm;
            const int n = n0 + sh, j = jb + th;
            if (n >= a.N || j >= a.Th2) continue;
            const int jrow = mir ? a.Th - 1 - j : j;
            if (mir && jrow < a.Th2) continue;
            const float2* sp = St + (mm * S4_SH + sh) * SHS + th * S4_MQP<MQ> + 2 * qq;
            float4 v = make_float4(sp[0].x, sp[0].y, sp[1].x, sp[1].y);
            if constexpr (KIND == 2) {
                const float f = (float)a.rinv[n];
                v.x *= f; v.y *= f; v.z *= f; v.w *= f;
            }
            reinterpret_cast<float4*>(out + 2 * (((size_t)n * a.Th + jrow) * a.MP + m0))[qq] = v;
        }
    }
}

template <int KIND, int MQ>
__global__ void __launch_bounds__(64 * MQ) __attribute__((amdgpu_waves_per_eu(2)))
k_syn4(Syn4Args a) {
    __shared__ float2 St[S4_STAGE<MQ>];
    syn4_item<KIND, MQ>(a, blockIdx.x, St);
}

// Persistent form: as many workgroups as the CUs hold at once, each pulling
// work items from its XCD's queue (wq[xcd] counts the items taken, item q of
// XCD x is block 8 q + x of the static numbering, so the items sharing a
// table block stay on one XCD and in the longest-slot-first order) until the
// queue is empty; uneven slot lengths then leave no tail of idle CUs.  The
// last workgroup to finish resets the queues (wq[8] counts finished
// workgroups) for the next launch.
template <int KIND, int MQ>
__global__ void __launch_bounds__(64 * MQ) __attribute__((amdgpu_waves_per_eu(2)))
k_syn4p(Syn4Args a, int* __restrict__ wq, int nq) {
    __shared__ float2 St[S4_STAGE<MQ>];
    __shared__ int s_q;
    const int xcd = blockIdx.x & 7;
    for (;;) {
        if (threadIdx.x == 0) s_q = atomicAdd(&wq[xcd], 1);
        __syncthreads();
        const int q = s_q;
        if (q >= nq) break;
        syn4_item<KIND, MQ>(a, 8 * q + xcd, St);
        __syncthreads();  // the item's staging reads are done before the next item's writes
    }
    if (threadIdx.x == 0 && atomicAdd(&wq[8], 1) == (int)gridDim.x - 1)
        for (int k = 0; k < 9; k++) atomicExch(&wq[k], 0);
}

// Persistent form that overlaps each item's epilogue with the next item's
// first operand loads: a workgroup takes its next item from the queue as
// soon as its contraction ends, issues that item's ring of loads, and only
// then stages and stores the finished item.  The item's steps are those of
// syn4_item, split into begin (decode, first loads), run (contraction) and
// end (epilogue).
struct S4Item {
    bool live;  // a valid block with m0 < M (uniform over the workgroup)
    bool run;   // this wave's m < M (uniform over the wave)
    int job, m0, jb, n0, ncl, h0, len, rows;
    size_t tb0;
    const float* in0;
    const float* in1;
};

template <int KIND>
__device__ __forceinline__ void syn4q_load(const Syn4Args& a, const S4Item& it, S4Ops& op, int k0, int q4) {
    const int C2 = 2 * a.N;
#pragma unroll
    for (int p = 0; p < 2; p++) {
        const int k = min(k0 + 2 * q4 + p, it.rows - 1);
        op.ta[p] = *reinterpret_cast<const float4*>(a.tabA + it.tb0 + (size_t)k * a.Th2p);
        if (KIND != 0) op.tb[p] = *reinterpret_cast<const float4*>(a.tabB + it.tb0 + (size_t)k * a.Th2p);
        const size_t h = (size_t)min(it.h0 + k, a.H - 1);
        op.x[p] = *reinterpret_cast<const float2*>(it.in0 + h * C2 + 2 * it.ncl);
        if (KIND == 1) op.t[p] = *reinterpret_cast<const float2*>(it.in1 + h * C2 + 2 * it.ncl);
    }
}

template <int KIND, int MQ>
__device__ __forceinline__ S4Item syn4q_begin(const Syn4Args& a, int b, S4Ops (&buf)[S4_NB]) {
    S4Item it;
    it.live = false;
    it.run = false;
    const BlockId bid = decode_block_at(b, a.nmo * a.njob * a.nsb, (S4_MG / MQ) * a.nth);
    if (!bid.valid) return it;
    const int mo = bid.group / (a.njob * a.nsb);
    const int gjs = bid.group - mo * a.njob * a.nsb;
    const int sb = gjs - (gjs / a.nsb) * a.nsb;
    it.job = gjs / a.nsb;
    const int mq = mo * (S4_MG / MQ) + bid.cc / a.nth, tt = bid.cc % a.nth;
    it.m0 = mq * MQ;
    if (it.m0 >= a.M) return it;
    it.live = true;
    it.jb = tt * S4_TH;
    it.n0 = sb * S4_SH;
    it.in0 = pickj(a.in0, it.job);
    it.in1 = pickj(a.in1, it.job);
    const int tid = threadIdx.x, wave = uniform(tid >> 6), lane = tid & 63, l16 = lane & 15, q4 = lane >> 4;
    it.ncl = min(it.n0 + l16, a.N - 1);
    const int m = it.m0 + wave;
    if (m < a.M) {
        const int si = a.mslot[m];
        const Slot s = a.slots[si];
        it.h0 = s.h0;
        it.len = s.len;
        it.rows = (s.len + 7) & ~7;
        it.tb0 = (size_t)a.t0[si] * a.Th2p + it.jb + 4 * l16;
        it.run = true;
#pragma unroll
        for (int u = 0; u < S4_NB; u++) {
            syn4q_load<KIND>(a, it, buf[u], 8 * u, q4);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    return it;
}

template <int KIND>
__device__ __forceinline__ void syn4q_step(const S4Ops& op, v4f (&Y)[KIND == 0 ? 1 : 2][2][2][4]) {
#pragma unroll
    for (int p = 0; p < 2; p++) {
        const float ta[4] = {op.ta[p].x, op.ta[p].y, op.ta[p].z, op.ta[p].w};
        const float x[2] = {op.x[p].x, op.x[p].y};
        if constexpr (KIND == 0) {
#pragma unroll
            for (int u = 0; u < 2; u++)
#pragma unroll
                for (int t = 0; t < 4; t++) Y[0][p][u][t] = mfma4(x[u], ta[t], Y[0][p][u][t]);
        } else if constexpr (KIND == 2) {
            const float tb[4] = {op.tb[p].x, op.tb[p].y, op.tb[p].z, op.tb[p].w};
            const float xr[2] = {-x[1], x[0]};
#pragma unroll
            for (int u = 0; u < 2; u++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    Y[0][1 - p][u][t] = mfma4(x[u], ta[t], Y[0][1 - p][u][t]);
                    Y[1][p][u][t] = mfma4(xr[u], tb[t], Y[1][p][u][t]);
                }
        } else {
            const float tb[4] = {op.tb[p].x, op.tb[p].y, op.tb[p].z, op.tb[p].w};
            const float tv[2] = {op.t[p].x, op.t[p].y};
            const float tr[2] = {tv[1], -tv[0]}, sr[2] = {-x[1], x[0]};
#pragma unroll
            for (int u = 0; u < 2; u++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    Y[0][1 - p][u][t] = mfma4(x[u], ta[t], Y[0][1 - p][u][t]);
                    Y[0][p][u][t] = mfma4(tr[u], tb[t], Y[0][p][u][t]);
                    Y[1][p][u][t] = mfma4(sr[u], tb[t], Y[1][p][u][t]);
                    Y[1][1 - p][u][t] = mfma4(tv[u], ta[t], Y[1][1 - p][u][t]);
                }
        }
    }
}

template <int KIND, int MQ>
__device__ __forceinline__ void syn4q_end(const Syn4Args& a, const S4Item& it,
                                          const v4f (&Y)[KIND == 0 ? 1 : 2][2][2][4], float2* __restrict__ St) {
    constexpr int NOUT = KIND == 0 ? 1 : 2;
    constexpr int NMIR = S4_NMIR<MQ>, SHS = S4_SHS<MQ>;
    constexpr int NPASS = NOUT * (2 / NMIR), NQ = MQ / 2;
    const int tid = threadIdx.x, wave = uniform(tid >> 6), lane = tid & 63, l16 = lane & 15, q4 = lane >> 4;
#pragma unroll
    for (int pass = 0; pass < NPASS; pass++) {
        const int o = pass / (2 / NMIR), mir0 = NMIR == 2 ? 0 : pass % 2;
        if (pass) __syncthreads();  // the previous pass's reads are done
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int sh = 4 * q4 + r, th = 16 * t + l16;
                const float y0r = Y[o][0][0][t][r], y1r = Y[o][1][0][t][r];
                const float y0i = Y[o][0][1][t][r], y1i = Y[o][1][1][t][r];
#pragma unroll
                for (int mm = 0; mm < NMIR; mm++)
                    St[(mm * S4_SH + sh) * SHS + th * S4_MQP<MQ> + wave] =
                        (mir0 + mm) ? make_float2(y0r - y1r, y0i - y1i) : make_float2(y0r + y1r, y0i + y1i);
            }
        __syncthreads();
        float* out = o == 0 ? pickj(a.out0, it.job) : pickj(a.out1, it.job);
        for (int idx = tid; idx < NQ * NMIR * S4_SH * S4_TH; idx += 64 * MQ) {
            const int qq = idx % NQ, item = idx / NQ;
            const int th = item % S4_TH, rest = item / S4_TH;
            const int sh = rest % S4_SH, mm = rest / S4_SH, mir = mir0 + mm;
            const int n = it.n0 + sh, j = it.jb + th;
            if (n >= a.N || j >= a.Th2) continue;
            const int jrow = mir ? a.Th - 1 - j : j;
            if (mir && jrow < a.Th2) continue;
            const float2* sp = St + (mm * S4_SH + sh) * SHS + th * S4_MQP<MQ> + 2 * qq;
            float4 v = make_float4(sp[0].x, sp[0].y, sp[1].x, sp[1].y);
            if constexpr (KIND == 2) {
                const float f = (float)a.rinv[n];
                v.x *= f; v.y *= f; v.z *= f; v.w *= f;
            }
            reinterpret_cast<float4*>(out + 2 * (((size_t)n * a.Th + jrow) * a.MP + it.m0))[qq] = v;
        }
    }
}

template <int KIND, int MQ>
__global__ void __launch_bounds__(64 * MQ) __attribute__((amdgpu_waves_per_eu(2)))
k_syn4q(Syn4Args a, int* __restrict__ wq, int nq) {
    constexpr int NOUT = KIND == 0 ? 1 : 2;
    __shared__ float2 St[S4_STAGE<MQ>];
    __shared__ int s_q;
    const int xcd = blockIdx.x & 7, q4 = (threadIdx.x & 63) >> 4;
    // s_q is read by every wave before thread 0 writes it again
    auto pull = [&]() {
        if (threadIdx.x == 0) s_q = atomicAdd(&wq[xcd], 1);
        __syncthreads();
        const int q = s_q;
        __syncthreads();
        return q;
    };
    S4Ops buf[S4_NB];
    int q = pull();
    S4Item it;
    it.live = false;
    it.run = false;
    if (q < nq) it = syn4q_begin<KIND, MQ>(a, 8 * q + xcd, buf);
    while (q < nq) {
        v4f Y[NOUT][2][2][4];
#pragma unroll
        for (int o = 0; o < NOUT; o++)
#pragma unroll
            for (int p = 0; p < 2; p++)
#pragma unroll
                for (int u = 0; u < 2; u++)
#pragma unroll
                    for (int t = 0; t < 4; t++) Y[o][p][u][t] = v4f{0, 0, 0, 0};
        if (it.run) {
            for (int k0 = 0; k0 < it.len; k0 += 8 * S4_NB) {
#pragma unroll
                for (int u = 0; u < S4_NB; u++) {
                    if (k0 + 8 * u < it.len) syn4q_step<KIND>(buf[u], Y);
                    syn4q_load<KIND>(a, it, buf[u], k0 + 8 * u + 8 * S4_NB, q4);
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
        // the next item's loads go out before this item's epilogue (the
        // queue's barriers also order the previous epilogue's LDS reads
        // before this one's writes)
        const int qn = pull();
        S4Item nx;
        nx.live = false;
        nx.run = false;
        if (qn < nq) nx = syn4q_begin<KIND, MQ>(a, 8 * qn + xcd, buf);
        if (it.live) syn4q_end<KIND, MQ>(a, it, Y, St);
        it = nx;
        q = qn;
    }
    if (threadIdx.x == 0 && atomicAdd(&wq[8], 1) == (int)gridDim.x - 1)
        for (int k = 0; k < 9; k++) atomicExch(&wq[k], 0);
}

static int syn4_launch(Ctx& c, int kind, int njob, const float* tA, const float* tB, const void* const* in0,
                       const void* const* in1, void* const* out0, void* const* out1) {
    Syn4Args a;
    a.N = c.N;
    a.Th = c.Th;
    a.Th2 = c.Th2;
    a.Th2p = c.Th2p;
    a.M = c.M;
    a.MP = c.MP;
    a.H = c.H;
    a.slots = c.d_slots;
    a.mslot = c.d_mslot;
    a.t0 = c.d_slot_t0;
    a.tabA = tA;
    a.tabB = tB;
    a.rinv = c.d_r + 2 * c.N;
    for (int k = 0; k < LG_MAXJ; k++) {
        const int q = k < njob ? k : 0;
        a.in0[k] = reinterpret_cast<const float*>(in0[q]);
        a.in1[k] = in1 ? reinterpret_cast<const float*>(in1[q]) : nullptr;
        a.out0[k] = reinterpret_cast<float*>(out0[q]);
        a.out1[k] = out1 ? reinterpret_cast<float*>(out1[q]) : nullptr;
    }
    a.nth = c.Th2p / S4_TH;
    a.nsb = (c.N + S4_SH - 1) / S4_SH;
    a.nmo = (c.M + S4_MG - 1) / S4_MG;
    a.njob = njob;
    auto grid = [&](int mq) { return dim3(grid_blocks(a.nmo * njob * a.nsb, (S4_MG / mq) * a.nth)); };
    const int v = c.var[LSD_V_SYN_SCALAR + kind];
    const int mq = (v >= 4 && v <= 6) ? 4 : 8;
    if (v == 6 || v == 10) {
        // persistent form with the next item's loads issued before each
        // epilogue
        const void* fn;
        if (kind == SYN_SCALAR) fn = mq == 4 ? (const void*)&k_syn4q<0, 4> : (const void*)&k_syn4q<0, 8>;
        else if (kind == SYN_VECTOR) fn = mq == 4 ? (const void*)&k_syn4q<1, 4> : (const void*)&k_syn4q<1, 8>;
        else fn = mq == 4 ? (const void*)&k_syn4q<2, 4> : (const void*)&k_syn4q<2, 8>;
        int& pg = c.syn_qgrid[kind][mq == 8 ? 1 : 0];
        if (pg == 0) {
            int per = 0;
            LSD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, 64 * mq, 0));
            pg = (std::max(per, 1) * c.ncu + 7) / 8 * 8;
        }
        const unsigned nb = grid(mq).x;
        const dim3 pgrid(std::min((unsigned)pg, nb));
        const int nq = (int)(nb / 8);
        if (kind == SYN_SCALAR) {
            if (mq == 4) hipLaunchKernelGGL((k_syn4q<0, 4>), pgrid, dim3(64 * 4), 0, c.stream, a, c.d_wq, nq);
            else hipLaunchKernelGGL((k_syn4q<0, 8>), pgrid, dim3(64 * 8), 0, c.stream, a, c.d_wq, nq);
        } else if (kind == SYN_VECTOR) {
            if (mq == 4) hipLaunchKernelGGL((k_syn4q<1, 4>), pgrid, dim3(64 * 4), 0, c.stream, a, c.d_wq, nq);
            else hipLaunchKernelGGL((k_syn4q<1, 8>), pgrid, dim3(64 * 8), 0, c.stream, a, c.d_wq, nq);
        } else {
            if (mq == 4) hipLaunchKernelGGL((k_syn4q<2, 4>), pgrid, dim3(64 * 4), 0, c.stream, a, c.d_wq, nq);
            else hipLaunchKernelGGL((k_syn4q<2, 8>), pgrid, dim3(64 * 8), 0, c.stream, a, c.d_wq, nq);
        }
        LSD_HIP(hipGetLastError());
        return LSD_OK;
    }
    if (v == 5 || v == 9) {
        // persistent form: the CUs' worth of workgroups (a multiple of 8,
        // so each XCD gets the same number) pulling the static blocks
        const void* fn;
        if (kind == SYN_SCALAR) fn = mq == 4 ? (const void*)&k_syn4p<0, 4> : (const void*)&k_syn4p<0, 8>;
        else if (kind == SYN_VECTOR) fn = mq == 4 ? (const void*)&k_syn4p<1, 4> : (const void*)&k_syn4p<1, 8>;
        else fn = mq == 4 ? (const void*)&k_syn4p<2, 4> : (const void*)&k_syn4p<2, 8>;
        int& pg = c.syn_pgrid[kind][mq == 8 ? 1 : 0];
        if (pg == 0) {
            int per = 0;
            LSD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, 64 * mq, 0));
            pg = (std::max(per, 1) * c.ncu + 7) / 8 * 8;
        }
        const unsigned nb = grid(mq).x;
        const dim3 pgrid(std::min((unsigned)pg, nb));
        const int nq = (int)(nb / 8);
        if (kind == SYN_SCALAR) {
            if (mq == 4) hipLaunchKernelGGL((k_syn4p<0, 4>), pgrid, dim3(64 * 4), 0, c.stream, a, c.d_wq, nq);
            else hipLaunchKernelGGL((k_syn4p<0, 8>), pgrid, dim3(64 * 8), 0, c.stream, a, c.d_wq, nq);
        } else if (kind == SYN_VECTOR) {
            if (mq == 4) hipLaunchKernelGGL((k_syn4p<1, 4>), pgrid, dim3(64 * 4), 0, c.stream, a, c.d_wq, nq);
            else hipLaunchKernelGGL((k_syn4p<1, 8>), pgrid, dim3(64 * 8), 0, c.stream, a, c.d_wq, nq);
        } else {
            if (mq == 4) hipLaunchKernelGGL((k_syn4p<2, 4>), pgrid, dim3(64 * 4), 0, c.stream, a, c.d_wq, nq);
            else hipLaunchKernelGGL((k_syn4p<2, 8>), pgrid, dim3(64 * 8), 0, c.stream, a, c.d_wq, nq);
        }
        LSD_HIP(hipGetLastError());
        return LSD_OK;
    }
    if (kind == SYN_SCALAR) {
        if (mq == 4) hipLaunchKernelGGL((k_syn4<0, 4>), grid(4), dim3(64 * 4), 0, c.stream, a);
        else hipLaunchKernelGGL((k_syn4<0, 8>), grid(8), dim3(64 * 8), 0, c.stream, a);
    } else if (kind == SYN_VECTOR) {
        if (mq == 4) hipLaunchKernelGGL((k_syn4<1, 4>), grid(4), dim3(64 * 4), 0, c.stream, a);
        else hipLaunchKernelGGL((k_syn4<1, 8>), grid(8), dim3(64 * 8), 0, c.stream, a);
    } else {
        if (mq == 4) hipLaunchKernelGGL((k_syn4<2, 4>), grid(4), dim3(64 * 4), 0, c.stream, a);
        else hipLaunchKernelGGL((k_syn4<2, 8>), grid(8), dim3(64 * 8), 0, c.stream, a);
    }
    LSD_HIP(hipGetLastError());
    return LSD_OK;
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
// FP32: two waves per SIMD (the compiler then fits the vector kind in 168
// registers, without spills); FP64 keeps the whole register file
template <typename T, bool VEC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 2 : 1)))
k_ana(AnaArgs<T> a) {
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
    // staging: thread -> (shell in tile 0..15, theta in chunk 0..15); the next
    // chunk's values are loaded into registers while the current chunk's
    // products run
    const int sh_s = tid >> 4, jj_s = tid & 15;
    const int n_s = (c0 >> 1) + sh_s;
    T gv[NF][2][2];
    auto gload = [&](int jc0) {
        const int j = jc0 + jj_s;
        const bool ok = n_s < a.N && j < a.Th2;
#pragma unroll
        for (int f = 0; f < NF; f++)
#pragma unroll
            for (int p = 0; p < 2; p++) {
                T re = T(0), im = T(0);
                if (ok) {
                    const size_t idx = (((size_t)n_s * a.M + s.mfft) * gstride + (size_t)p * a.Th2 + j) * 2;
                    re = gf[f][idx];
                    im = gf[f][idx + 1];
                }
                gv[f][p][0] = re;
                gv[f][p][1] = im;
            }
    };
    // table values of this lane for one chunk: [theta step][tile][P, L, S];
    // the scalar kind loads the next chunk's while the current one runs (the
    // vector kind has no registers to spare for it and loads them as it goes)
    T wc[ANA_JC / 4][VEC ? 1 : ANA_TPW][3], wn[ANA_JC / 4][VEC ? 1 : ANA_TPW][3];
    auto tload = [&](auto& w, int jc0) {
#pragma unroll
        for (int js = 0; js < ANA_JC / 4; js++) {
            const int j = jc0 + 4 * js + q4;
            const bool jv = j < a.Th2;
            const size_t trow = (size_t)(jv ? j : a.Th2 - 1) * a.H;
#pragma unroll
            for (int t = 0; t < (int)(sizeof(w[0]) / sizeof(w[0][0])); t++) {
                const int tile = wave + 4 * t;
                const bool tv = tile < ntiles && jv;
                const int par = tile < t_even ? 0 : 1;
                const int kt = par ? tile - t_even : tile;  // tile within the class
                const int ncls = par ? n_odd : n_even;
                const int kc = kt * 16 + l16;               // class-local row of this lane
                const int ap = (kc < ncls) ? s.h0 + (par ? n_even : 0) + kc : s.h0;
                w[js][t][0] = tv ? a.tP[trow + ap] : T(0);
                w[js][t][1] = (VEC && tv) ? a.tL[trow + ap] : T(0);
                w[js][t][2] = (VEC && tv) ? a.tS[trow + ap] : T(0);
            }
        }
    };
    gload(0);
    if constexpr (!VEC) tload(wc, 0);
    for (int jc0 = 0; jc0 < a.Th2; jc0 += ANA_JC) {
        __syncthreads();
#pragma unroll
        for (int f = 0; f < NF; f++)
#pragma unroll
            for (int p = 0; p < 2; p++) {
                G[f][p][jj_s][2 * sh_s] = gv[f][p][0];
                G[f][p][jj_s][2 * sh_s + 1] = gv[f][p][1];
            }
        __syncthreads();
        if (jc0 + ANA_JC < a.Th2) {
            gload(jc0 + ANA_JC);
            if constexpr (!VEC) tload(wn, jc0 + ANA_JC);
        }
#pragma unroll
        for (int js = 0; js < ANA_JC / 4; js++) {
            const int jl = 4 * js + q4;  // chunk-local theta of this lane's K index
#pragma unroll
            for (int t = 0; t < ANA_TPW; t++) {
                const int tile = wave + 4 * t;
                if (tile >= ntiles) break;
                const int par = tile < t_even ? 0 : 1;
                T wP, wL, wS;
                if constexpr (VEC) {
                    const int j = jc0 + jl;
                    const bool jv = j < a.Th2;
                    const size_t trow = (size_t)(jv ? j : a.Th2 - 1) * a.H;
                    const int kt = par ? tile - t_even : tile;  // tile within the class
                    const int ncls = par ? n_odd : n_even;
                    const int kc = kt * 16 + l16;               // class-local row of this lane
                    const int ap = (kc < ncls) ? s.h0 + (par ? n_even : 0) + kc : s.h0;
                    wP = jv ? a.tP[trow + ap] : T(0);
                    wL = jv ? a.tL[trow + ap] : T(0);
                    wS = jv ? a.tS[trow + ap] : T(0);
                } else {
                    wP = wc[js][t][0];
                    wL = wS = T(0);
                }
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
        if constexpr (!VEC) {
#pragma unroll
            for (int js = 0; js < ANA_JC / 4; js++)
#pragma unroll
                for (int t = 0; t < ANA_TPW; t++)
#pragma unroll
                    for (int e = 0; e < 3; e++) wc[js][t][e] = wn[js][t][e];
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

// ----------------------------------------------------- FP32 analysis (mixed forms)
//
// The same contractions with 16-byte loads and no LDS.  A wave takes one
// (slot, parity class, block of 64 harmonics) and 16 shells and contracts
// over all theta:
// * the A operand is the table: one 16-byte load per lane (tables stored per
//   slot and class in 64-harmonic blocks, position 4 l + t holding class
//   harmonic 16 t + l) carries four 16-harmonic tiles;
// * the B operand is the folded data: one 16-byte load is (re, im) of the
//   lane's shell at two consecutive theta, so K steps come in pairs, the
//   first over theta 2 q4 and the second over 2 q4 + 1 of each group of 8;
// * the i rotations are a swap of re and im within the lane.
// Operands sit in a ring of A4_NB buffers (one pair of K steps each),
// refilled A4_NB pairs ahead in load order; theta past Th2 meets zero table
// rows.  The lane writes (re, im) of its shell for each of its harmonics:
// 16 lanes write one 128-byte run of a coll row.
constexpr int A4_SH = 16;  // shells per wave
constexpr int A4_W = 2;    // waves (shell blocks) per workgroup
constexpr int A4_NB = 2;   // operand buffers (one or three measured 7-12 % slower)

struct Ana4Args {
    int N, Th2, Th2a, M, Hpa;
    int nsbp, maxhb, njob, ngroups;
    const Slot* slots;
    const int* order;
    const int* base;  // [2 nslot] first table column of (slot, class)
    const float* tP;  // [Th2a][Hpa]
    const float* tL;
    const float* tS;
    const float* g1[LG_MAXJ];  // folded [n][M][2][Th2] complex
    const float* g2[LG_MAXJ];
    const float* g3[LG_MAXJ];
    float* o1[LG_MAXJ];  // coll [h][2N]
    float* o2[LG_MAXJ];
    float* o3[LG_MAXJ];
};

struct A4Ops {
    float4 g[5];     // folded data: r(p), theta(1-p), phi(p), theta(p), phi(1-p); re, im at two theta
    float4 w[2][3];  // tables P, L, S at the two theta of each K step
};

// One work item (block b of the static numbering) of the FP32 analysis.
template <bool VEC>
__device__ __forceinline__ void ana4_item(const Ana4Args& a, int b) {
    constexpr int NO = VEC ? 3 : 1;
    constexpr int NG = VEC ? 5 : 1;
    // groups (slot, job), longest slots first; members (class, 64-harmonic
    // block, shell-block pair) on one XCD, reading the same folded rows
    const BlockId bid = decode_block_at(b, a.ngroups, 2 * a.maxhb * a.nsbp);
    if (!bid.valid) return;
    const int job = bid.group % a.njob;
    const int sl = uniform(a.order[bid.group / a.njob]);
    const Slot s = a.slots[sl];
    const int par = bid.cc / (a.maxhb * a.nsbp), rest = bid.cc - par * a.maxhb * a.nsbp;
    const int hb = rest / a.nsbp, sbp = rest - hb * a.nsbp;
    const int n_even = (s.len + 1) / 2;
    const int ncls = par ? s.len / 2 : n_even;
    if (64 * hb >= ncls) return;
    const int tid = threadIdx.x, wave = uniform(tid >> 6), lane = tid & 63, l16 = lane & 15, q4 = lane >> 4;
    const int n0 = (sbp * A4_W + wave) * A4_SH;
    if (n0 >= a.N) return;  // no barriers: a wave may leave alone
    const int ntile = min(4, (ncls - 64 * hb + 15) / 16);
    const int ncl = min(n0 + l16, a.N - 1);
    const float* gf[3] = {pickj(a.g1, job), pickj(a.g2, job), pickj(a.g3, job)};
    // folded rows of this lane's shell: (field, parity) -> complex row of Th2
    const size_t grow = ((size_t)ncl * a.M + s.mfft) * 2;
    const float* gr[NG];
    {
        const int fsel[5] = {0, 1, 2, 1, 2};
        const int psel[5] = {par, 1 - par, par, par, 1 - par};
#pragma unroll
        for (int g = 0; g < NG; g++) gr[g] = gf[fsel[g]] + (grow + psel[g]) * a.Th2 * 2;
    }
    const size_t tcol = (size_t)a.base[2 * sl + par] + 64 * hb + 4 * l16;
    auto load = [&](A4Ops& op, int tb) {
        tb = min(tb, a.Th2a - 8);
        const int th = min(tb + 2 * q4, a.Th2 - 2);  // theta pair of the folded loads
#pragma unroll
        for (int g = 0; g < NG; g++) op.g[g] = *reinterpret_cast<const float4*>(gr[g] + 2 * th);
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const size_t row = (size_t)(tb + 2 * q4 + k) * a.Hpa + tcol;
            op.w[k][0] = *reinterpret_cast<const float4*>(a.tP + row);
            if (VEC) {
                op.w[k][1] = *reinterpret_cast<const float4*>(a.tL + row);
                op.w[k][2] = *reinterpret_cast<const float4*>(a.tS + row);
            }
        }
    };
    v4f acc[4][NO][2];  // [harmonic tile][q, s, t][re, im]
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int o = 0; o < NO; o++) acc[t][o][0] = acc[t][o][1] = v4f{0, 0, 0, 0};
    auto step = [&](const A4Ops& op) {
#pragma unroll
        for (int k = 0; k < 2; k++) {
            float gv[NG][2];
#pragma unroll
            for (int g = 0; g < NG; g++) {
                gv[g][0] = k ? op.g[g].z : op.g[g].x;
                gv[g][1] = k ? op.g[g].w : op.g[g].y;
            }
            const float wp[4] = {op.w[k][0].x, op.w[k][0].y, op.w[k][0].z, op.w[k][0].w};
#pragma unroll
            for (int t = 0; t < 4; t++) {
                if (t >= ntile) break;
#pragma unroll
                for (int u = 0; u < 2; u++) acc[t][0][u] = mfma4(wp[t], gv[0][u], acc[t][0][u]);
                if constexpr (VEC) {
                    const float wl = t == 0 ? op.w[k][1].x : t == 1 ? op.w[k][1].y : t == 2 ? op.w[k][1].z : op.w[k][1].w;
                    const float ws = t == 0 ? op.w[k][2].x : t == 1 ? op.w[k][2].y : t == 2 ? op.w[k][2].z : op.w[k][2].w;
                    // s.re += L Y2o.re + S Y3s.im ; s.im += L Y2o.im - S Y3s.re
                    acc[t][1][0] = mfma4(wl, gv[1][0], acc[t][1][0]);
                    acc[t][1][0] = mfma4(ws, gv[2][1], acc[t][1][0]);
                    acc[t][1][1] = mfma4(wl, gv[1][1], acc[t][1][1]);
                    acc[t][1][1] = mfma4(ws, -gv[2][0], acc[t][1][1]);
                    // t.re = -Y2s.im S + Y3o.re L ; t.im = Y2s.re S + Y3o.im L
                    acc[t][2][0] = mfma4(ws, -gv[3][1], acc[t][2][0]);
                    acc[t][2][0] = mfma4(wl, gv[4][0], acc[t][2][0]);
                    acc[t][2][1] = mfma4(ws, gv[3][0], acc[t][2][1]);
                    acc[t][2][1] = mfma4(wl, gv[4][1], acc[t][2][1]);
                }
            }
        }
    };
    A4Ops buf[A4_NB];
#pragma unroll
    for (int u = 0; u < A4_NB; u++) {
        load(buf[u], 8 * u);
        __builtin_amdgcn_sched_barrier(0);
    }
    for (int tb = 0; tb < a.Th2a; tb += 8 * A4_NB) {
#pragma unroll
        for (int u = 0; u < A4_NB; u++) {
            if (tb + 8 * u < a.Th2a) step(buf[u]);
            load(buf[u], tb + 8 * u + 8 * A4_NB);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // the lane holds class harmonics 64 hb + 16 t + 4 q4 + r of shell n0 + l16
    const int n = n0 + l16;
    if (n >= a.N) return;
    const int C2 = 2 * a.N;
    float* of[3] = {pickj(a.o1, job), pickj(a.o2, job), pickj(a.o3, job)};
#pragma unroll
    for (int t = 0; t < 4; t++) {
        if (t >= ntile) break;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int kc = 64 * hb + 16 * t + 4 * q4 + r;
            if (kc >= ncls) continue;
            const size_t o = ((size_t)s.h0 + 2 * kc + par) * C2 + 2 * n;
            if constexpr (VEC) {
                // the reference leaves q%Im(m=0) as summed
                *reinterpret_cast<float2*>(of[0] + o) = make_float2(acc[t][0][0][r], acc[t][0][1][r]);
                *reinterpret_cast<float2*>(of[1] + o) = make_float2(acc[t][1][0][r], s.m == 0 ? 0.f : acc[t][1][1][r]);
                *reinterpret_cast<float2*>(of[2] + o) = make_float2(acc[t][2][0][r], s.m == 0 ? 0.f : acc[t][2][1][r]);
            } else {
                *reinterpret_cast<float2*>(of[0] + o) = make_float2(acc[t][0][0][r], s.m == 0 ? 0.f : acc[t][0][1][r]);
            }
        }
    }
}

template <bool VEC>
__global__ void __launch_bounds__(64 * A4_W) __attribute__((amdgpu_waves_per_eu(2)))
k_ana4(Ana4Args a) {
    ana4_item<VEC>(a, blockIdx.x);
}

// Persistent form: the CUs' worth of workgroups pulling items from per-XCD
// queues (item q of XCD x is block 8 q + x of the static numbering), as the
// persistent synthesis does; the item has no barriers, so every wave comes
// back to the loop.  The last workgroup to finish resets the queues.
template <bool VEC>
__global__ void __launch_bounds__(64 * A4_W) __attribute__((amdgpu_waves_per_eu(2)))
k_ana4p(Ana4Args a, int* __restrict__ wq, int nq) {
    __shared__ int s_q;
    const int xcd = blockIdx.x & 7;
    for (;;) {
        if (threadIdx.x == 0) s_q = atomicAdd(&wq[xcd], 1);
        __syncthreads();
        const int q = s_q;
        __syncthreads();  // every wave has read s_q before the next write
        if (q >= nq) break;
        ana4_item<VEC>(a, 8 * q + xcd);
    }
    if (threadIdx.x == 0 && atomicAdd(&wq[8], 1) == (int)gridDim.x - 1)
        for (int k = 0; k < 9; k++) atomicExch(&wq[k], 0);
}

static int ana4_launch(Ctx& c, int kind, int njob, const void* const* g0, const void* const* g1,
                       const void* const* g2, void* const* o0, void* const* o1, void* const* o2) {
    Ana4Args a;
    a.N = c.N;
    a.Th2 = c.Th2;
    a.Th2a = c.Th2a;
    a.M = c.M;
    a.Hpa = c.Hpa;
    a.slots = c.d_slots;
    a.order = c.d_slot_order;
    a.base = c.d_ana_base;
    const size_t HT = (size_t)c.Th2a * c.Hpa;
    a.tP = c.d_atab4;
    a.tL = c.d_atab4 + HT;
    a.tS = c.d_atab4 + 2 * HT;
    for (int k = 0; k < LG_MAXJ; k++) {
        const int q = k < njob ? k : 0;
        a.g1[k] = reinterpret_cast<const float*>(g0[q]);
        a.g2[k] = g1 ? reinterpret_cast<const float*>(g1[q]) : a.g1[k];
        a.g3[k] = g2 ? reinterpret_cast<const float*>(g2[q]) : a.g1[k];
        a.o1[k] = reinterpret_cast<float*>(o0[q]);
        a.o2[k] = o1 ? reinterpret_cast<float*>(o1[q]) : nullptr;
        a.o3[k] = o2 ? reinterpret_cast<float*>(o2[q]) : nullptr;
    }
    a.nsbp = ((c.N + A4_SH - 1) / A4_SH + A4_W - 1) / A4_W;
    a.maxhb = c.ana_maxhb;
    a.njob = njob;
    a.ngroups = c.nslot * njob;
    dim3 grid(grid_blocks(a.ngroups, 2 * a.maxhb * a.nsbp));
    if (c.var[LSD_V_ANA] == 2) {
        // persistent form on the analysis half of the queue counters
        int& pg = c.ana_pgrid[kind == ANA_SCALAR ? 0 : 1];
        if (pg == 0) {
            int per = 0;
            const void* fn = kind == ANA_SCALAR ? (const void*)&k_ana4p<false> : (const void*)&k_ana4p<true>;
            LSD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, 64 * A4_W, 0));
            pg = (std::max(per, 1) * c.ncu + 7) / 8 * 8;
        }
        const dim3 pgrid(std::min((unsigned)pg, grid.x));
        const int nq = (int)(grid.x / 8);
        if (kind == ANA_SCALAR) hipLaunchKernelGGL(k_ana4p<false>, pgrid, dim3(64 * A4_W), 0, c.stream, a, c.d_wq + 16, nq);
        else hipLaunchKernelGGL(k_ana4p<true>, pgrid, dim3(64 * A4_W), 0, c.stream, a, c.d_wq + 16, nq);
        LSD_HIP(hipGetLastError());
        return LSD_OK;
    }
    if (kind == ANA_SCALAR) hipLaunchKernelGGL(k_ana4<false>, grid, dim3(64 * A4_W), 0, c.stream, a);
    else hipLaunchKernelGGL(k_ana4<true>, grid, dim3(64 * A4_W), 0, c.stream, a);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
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
    if (f32 && c.cfg.precision == LSD_MIXED_BF16X3) return anab_launch(c, kind, njob, g0, g1, g2, o0, o1, o2);
    if (f32 && c.d_atab4 && (c.var[LSD_V_ANA] == 0 || c.var[LSD_V_ANA] == 2)) {
        if (njob < 1 || njob > LG_MAXJ) {
            set_error("analysis: 1..6 transforms per launch");
            return LSD_E_ARG;
        }
        return ana4_launch(c, kind, njob, g0, g1, g2, o0, o1, o2);
    }
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
