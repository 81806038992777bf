// Gauss-Legendre synthesis and analysis in BF16x3 split precision
// (LSD_MIXED_BF16X3): each FP32 operand x is split as x = hi + lo with
// hi = bf16(x), lo = bf16(x - hi), and a product a b is formed as
// hi_a hi_b + hi_a lo_b + lo_a hi_b on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation (the reference's sums: modules/transform.fftw3.F90:746-774
// for the synthesis, 860-910 for the analysis).  The dropped lo lo term and
// the rounding of lo leave about 16 significant bits per product.
//
// The contractions are those of the FP32 kernels (legendre.hip): per m-slot
// and equatorial parity class p, Y_p[shell, theta] = sum_k coef[k, shell]
// T[k, theta] (synthesis) and D[k, shell] = sum_theta Tw[theta, k]
// G[theta, shell] (analysis), with the same parity buckets, rotations and
// output layouts.  What differs is the shape of a step: the bf16
// instruction takes 32 values of K per operand row (lane l holds
// A[l % 16][8 (l / 16) .. + 7] and B[8 (l / 16) .. + 7][l % 16]), so
//  * the synthesis steps over 32 harmonics of one parity class at a time;
//    its tables are stored split and blocked so that one lane's operands
//    for the four 16-theta tiles of a step are one 128-byte run:
//    [table][class block][theta block of 64][lane 16][lane group 4][tile 4]
//    [hi, lo][8]; the coefficients (FP32, coll layout) are split on load;
//  * the analysis steps over 32 theta points at a time; its tables are
//    stored [table][class block of 64 harmonics][theta block of 32][tile 4]
//    [lane 16][lane group 4][hi, lo][8]; the folded data (FP32) are split on
//    load.
#include <algorithm>

#include "lsd_internal.h"

namespace lsd {

namespace {

typedef float v4f_b __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

constexpr int BX_MAXJ = 6;

template <typename P>
__device__ __forceinline__ P pick_job(const P (&v)[BX_MAXJ], int j) {
    P r = v[0];
#pragma unroll
    for (int k = 1; k < BX_MAXJ; k++)
        if (j == k) r = v[k];
    return r;
}

__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Workgroup b of a grid of ((ngroups + 7) / 8) * 8 * ncc: groups round-robin
// over the 8 XCDs, the ncc members of a group consecutive on one XCD.
struct BxId {
    int group, cc;
    bool valid;
};
__device__ __forceinline__ BxId bx_block(int ngroups, int ncc) {
    const int b = blockIdx.x, xcd = b & 7, q = b >> 3;
    const int gl = q / ncc;
    BxId r;
    r.cc = q - gl * ncc;
    r.group = gl * 8 + xcd;
    r.valid = r.group < ngroups;
    return r;
}
unsigned bx_grid(int ngroups, int ncc) { return (unsigned)(((ngroups + 7) / 8) * 8 * ncc); }

__device__ __forceinline__ v4f_b mfma_bf(const bf16x8& a, const bf16x8& b, v4f_b c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// hi lo split of eight FP32 values
struct Split8 {
    bf16x8 hi, lo;
};
__device__ __forceinline__ Split8 split8(const float (&x)[8]) {
    Split8 s;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const __bf16 h = (__bf16)x[e];
        s.hi[e] = h;
        s.lo[e] = (__bf16)(x[e] - (float)h);
    }
    return s;
}

__device__ __forceinline__ bf16x8 neg8(const bf16x8& v) {
    u32x4 u = __builtin_bit_cast(u32x4, v);
    u ^= (u32x4){0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
    return __builtin_bit_cast(bf16x8, u);
}

// acc += a b in BF16x3
__device__ __forceinline__ v4f_b mfma3(const Split8& a, const bf16x8& bh, const bf16x8& bl, v4f_b c) {
    c = mfma_bf(a.hi, bh, c);
    c = mfma_bf(a.hi, bl, c);
    return mfma_bf(a.lo, bh, c);
}

}  // namespace

// ---------------------------------------------------------------- tables

// Synthesis tables: entry (kind, block, theta block, l16, q4, t, part, e)
// holds table `kind` (FP64 [5][H][Th2]) at harmonic h0 + 2 (c0 + 8 q4 + e) + p
// of the block's slot and class, theta 64 tb + 16 t + l16; zero past the
// class or past Th2.  part 0 = hi, 1 = lo.
__global__ void k_tab_synb(int H, int Th2, int nthb, int nblk, const int4* __restrict__ blocks,
                           const double* __restrict__ tab, __bf16* __restrict__ out) {
    const size_t per_kind = (size_t)nblk * nthb * 4096;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 5 * per_kind) return;
    const int kind = (int)(i / per_kind);
    size_t r = i - kind * per_kind;
    const int e = (int)(r & 7), part = (int)((r >> 3) & 1), t = (int)((r >> 4) & 3), q4 = (int)((r >> 6) & 3),
              l16 = (int)((r >> 8) & 15);
    r >>= 12;
    const int tb = (int)(r % nthb), b = (int)(r / nthb);
    const int4 bk = blocks[b];  // h0 of the slot, class p, first class index c0, class length
    const int c = bk.z + 8 * q4 + e, th = 64 * tb + 16 * t + l16;
    double v = 0.0;
    if (c < bk.w && th < Th2) v = tab[((size_t)kind * H + bk.x + 2 * c + bk.y) * Th2 + th];
    const __bf16 hi = (__bf16)v;
    out[i] = part ? (__bf16)(v - (double)(float)hi) : hi;
}

// Analysis tables: entry (kind, class block, theta block, t, l16, q4, part, e)
// holds table `kind` (FP64 [3][Th2][H], each slot's class columns from
// base) at class harmonic 64 hb + 16 t + l16 and theta 32 jb + 8 q4 + e.
__global__ void k_tab_anab(int H, int Th2, int njb, int nblk, const int4* __restrict__ blocks,
                           const double* __restrict__ atab, __bf16* __restrict__ out) {
    const size_t per_kind = (size_t)nblk * njb * 4096;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * per_kind) return;
    const int kind = (int)(i / per_kind);
    size_t r = i - kind * per_kind;
    const int e = (int)(r & 7), part = (int)((r >> 3) & 1), q4 = (int)((r >> 4) & 3), l16 = (int)((r >> 6) & 15),
              t = (int)((r >> 10) & 3);
    r >>= 12;
    const int jb = (int)(r % njb), b = (int)(r / njb);
    const int4 bk = blocks[b];  // first analysis-table column of the class, class length, 64-block index
    const int kc = 64 * bk.z + 16 * t + l16, th = 32 * jb + 8 * q4 + e;
    double v = 0.0;
    if (kc < bk.y && th < Th2) v = atab[((size_t)kind * Th2 + th) * H + bk.x + kc];
    const __bf16 hi = (__bf16)v;
    out[i] = part ? (__bf16)(v - (double)(float)hi) : hi;
}

// Block lists and split tables of a context (mixed BF16X3 form).
int tables_bf16(Ctx& c) {
    std::vector<int4> sb, ab;            // synthesis blocks; analysis blocks
    std::vector<int> sbase(2 * c.nslot), abase(2 * c.nslot), scount(2 * c.nslot), acount(2 * c.nslot);
    for (int s = 0; s < c.nslot; s++) {
        const Slot& q = c.slots[s];
        const int n_even = (q.len + 1) / 2;
        for (int p = 0; p < 2; p++) {
            const int ncls = p ? q.len / 2 : n_even;
            sbase[2 * s + p] = (int)sb.size();
            scount[2 * s + p] = (ncls + 31) / 32;
            for (int c0 = 0; c0 < ncls; c0 += 32) sb.push_back(make_int4(q.h0, p, c0, ncls));
            abase[2 * s + p] = (int)ab.size();
            acount[2 * s + p] = (ncls + 63) / 64;
            for (int hb = 0; 64 * hb < ncls; hb++)
                ab.push_back(make_int4(q.h0 + (p ? n_even : 0), ncls, hb, 0));
        }
    }
    c.bx_nsb = (int)sb.size();
    c.bx_nab = (int)ab.size();
    c.bx_nthb = (c.Th2 + 63) / 64;
    c.bx_njb = (c.Th2 + 31) / 32;
    int maxsb = 0, maxab = 0;
    for (int k = 0; k < 2 * c.nslot; k++) {
        maxsb = std::max(maxsb, scount[k]);
        maxab = std::max(maxab, acount[k]);
    }
    c.bx_maxsb = maxsb;
    c.bx_maxab = maxab;
    const size_t ns = (size_t)5 * c.bx_nsb * c.bx_nthb * 4096, na = (size_t)3 * c.bx_nab * c.bx_njb * 4096;
    int4* d_sb = nullptr;
    int4* d_ab = nullptr;
    if (hipMalloc((void**)&c.d_tabb, sizeof(__bf16) * ns) != hipSuccess ||
        hipMalloc((void**)&c.d_atabb, sizeof(__bf16) * na) != hipSuccess ||
        hipMalloc((void**)&c.d_bx_sbase, sizeof(int) * 4 * c.nslot) != hipSuccess) {
        set_error("tables_bf16: hipMalloc failed");
        return LSD_E_HIP;
    }
    AsyncScratch s1(c.stream), s2(c.stream);
    LSD_HIP(s1.alloc(&d_sb, sizeof(int4) * std::max<size_t>(sb.size(), 1)));
    LSD_HIP(s2.alloc(&d_ab, sizeof(int4) * std::max<size_t>(ab.size(), 1)));
    std::vector<int> bases(4 * c.nslot);
    for (int k = 0; k < 2 * c.nslot; k++) {
        bases[k] = sbase[k];
        bases[2 * c.nslot + k] = abase[k];
    }
    LSD_HIP(hipMemcpyAsync(c.d_bx_sbase, bases.data(), sizeof(int) * 4 * c.nslot, hipMemcpyHostToDevice, c.stream));
    LSD_HIP(hipMemcpyAsync(d_sb, sb.data(), sizeof(int4) * sb.size(), hipMemcpyHostToDevice, c.stream));
    LSD_HIP(hipMemcpyAsync(d_ab, ab.data(), sizeof(int4) * ab.size(), hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_tab_synb, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, c.stream, c.H, c.Th2,
                       c.bx_nthb, c.bx_nsb, d_sb, c.d_tab, c.d_tabb);
    hipLaunchKernelGGL(k_tab_anab, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, c.stream, c.H, c.Th2,
                       c.bx_njb, c.bx_nab, d_ab, c.d_atab, c.d_atabb);
    LSD_HIP(hipGetLastError());
    LSD_HIP(hipStreamSynchronize(c.stream));
    return LSD_OK;
}

// ------------------------------------------------------------- synthesis

constexpr int BS_TH = 64;  // theta per wave (four 16-theta tiles)
constexpr int BS_SH = 16;  // shells per wave
constexpr int BS_MQ = 4;   // m per workgroup (one per wave), both hemispheres staged per pass
constexpr int BS_SHS = BS_TH * BS_MQ + 2;

struct SynbArgs {
    int N, Th, Th2, M, MP, H;
    int nth, nsb, nmo, njob, nthb;
    const Slot* slots;
    const int* mslot;
    const int* sbase;       // [2 nslot] first synthesis block of (slot, class)
    const __bf16* tabA;     // split tables, layout of k_tab_synb
    const __bf16* tabB;
    const double* rinv;
    const float* in0[BX_MAXJ];
    const float* in1[BX_MAXJ];
    float* out0[BX_MAXJ];
    float* out1[BX_MAXJ];
};

// One wave: 64 theta x 16 shells of one m; a workgroup BS_MQ consecutive m.
// KIND 0 scalar (P), 1 vector (llP_, llsin1mP; inputs s, t), 2 gradient
// (P_, sin1mP, scaled by 1/r).
template <int KIND>
__global__ void __launch_bounds__(64 * BS_MQ) k_synb(SynbArgs a) {
    constexpr int NOUT = KIND == 0 ? 1 : 2;
    __shared__ float2 St[2 * BS_SH * BS_SHS];
    const BxId bid = bx_block(a.nmo * a.njob * a.nsb, (16 / BS_MQ) * a.nth);
    if (!bid.valid) return;
    const int mo = bid.group / (a.njob * a.nsb);
    const int gjs = bid.group - mo * a.njob * a.nsb;
    const int job = gjs / a.nsb, sb = gjs - job * a.nsb;
    const int mq = mo * (16 / BS_MQ) + bid.cc / a.nth, tt = bid.cc % a.nth;
    const int m0 = mq * BS_MQ;
    if (m0 >= a.M) return;
    const int jb = tt * BS_TH, n0 = sb * BS_SH;
    const float* in0 = pick_job(a.in0, job);
    const float* in1 = pick_job(a.in1, job);
    const int tid = threadIdx.x, wave = wave_uniform(tid >> 6), lane = tid & 63, l16 = lane & 15, q4 = lane >> 4;
    const int C2 = 2 * a.N;
    const int ncl = min(n0 + l16, a.N - 1);
    const int m = m0 + wave;
    v4f_b Y[NOUT][2][2][4];  // [output][parity bucket][re, im][theta tile]
#pragma unroll
    for (int o = 0; o < NOUT; o++)
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int u = 0; u < 2; u++)
#pragma unroll
                for (int t = 0; t < 4; t++) Y[o][p][u][t] = v4f_b{0, 0, 0, 0};
    if (m < a.M) {
        const int si = a.mslot[m];
        const Slot s = a.slots[si];
        for (int p = 0; p < 2; p++) {
            const int ncls = p ? s.len / 2 : (s.len + 1) / 2;
            const int b0 = a.sbase[2 * si + p];
            for (int c0 = 0, b = b0; c0 < ncls; c0 += 32, b++) {
                // coefficients of this lane's shell, class harmonics c0 + 8 q4 + e
                float xr[8], xi[8], tr_[8], ti_[8];
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const int cc = min(c0 + 8 * q4 + e, ncls - 1);
                    const size_t h = (size_t)s.h0 + 2 * cc + p;
                    const float2 v = *reinterpret_cast<const float2*>(in0 + h * C2 + 2 * ncl);
                    xr[e] = v.x;
                    xi[e] = v.y;
                    if (KIND == 1) {
                        const float2 w = *reinterpret_cast<const float2*>(in1 + h * C2 + 2 * ncl);
                        tr_[e] = w.x;
                        ti_[e] = w.y;
                    }
                }
                // table operands: 4 tiles x (hi, lo) x 8 for table A (and B)
                const size_t tofs = ((((size_t)b * a.nthb + jb / 64) * 16 + l16) * 4 + q4) * 64;
                bf16x8 TA[4][2], TB[4][2];
#pragma unroll
                for (int t = 0; t < 4; t++)
#pragma unroll
                    for (int part = 0; part < 2; part++) {
                        TA[t][part] = *reinterpret_cast<const bf16x8*>(a.tabA + tofs + (t * 2 + part) * 8);
                        if (KIND != 0) TB[t][part] = *reinterpret_cast<const bf16x8*>(a.tabB + tofs + (t * 2 + part) * 8);
                    }
                const Split8 X[2] = {split8(xr), split8(xi)};
                if constexpr (KIND == 0) {
#pragma unroll
                    for (int u = 0; u < 2; u++)
#pragma unroll
                        for (int t = 0; t < 4; t++) Y[0][p][u][t] = mfma3(X[u], TA[t][0], TA[t][1], Y[0][p][u][t]);
                } else if constexpr (KIND == 2) {
                    // theta: c dP (bucket 1-p); phi: (i c) mP/sin (bucket p), (i c) = (-im, re)
                    const Split8 XR[2] = {Split8{neg8(X[1].hi), neg8(X[1].lo)}, X[0]};
#pragma unroll
                    for (int u = 0; u < 2; u++)
#pragma unroll
                        for (int t = 0; t < 4; t++) {
                            Y[0][1 - p][u][t] = mfma3(X[u], TA[t][0], TA[t][1], Y[0][1 - p][u][t]);
                            Y[1][p][u][t] = mfma3(XR[u], TB[t][0], TB[t][1], Y[1][p][u][t]);
                        }
                } else {
                    // (-i t) = (t_im, -t_re); (i s) = (-s_im, s_re)
                    const Split8 T[2] = {split8(tr_), split8(ti_)};
                    const Split8 TR[2] = {T[1], Split8{neg8(T[0].hi), neg8(T[0].lo)}};
                    const Split8 SR[2] = {Split8{neg8(X[1].hi), neg8(X[1].lo)}, X[0]};
#pragma unroll
                    for (int u = 0; u < 2; u++)
#pragma unroll
                        for (int t = 0; t < 4; t++) {
                            Y[0][1 - p][u][t] = mfma3(X[u], TA[t][0], TA[t][1], Y[0][1 - p][u][t]);
                            Y[0][p][u][t] = mfma3(TR[u], TB[t][0], TB[t][1], Y[0][p][u][t]);
                            Y[1][p][u][t] = mfma3(SR[u], TB[t][0], TB[t][1], Y[1][p][u][t]);
                            Y[1][1 - p][u][t] = mfma3(T[u], TA[t][0], TA[t][1], Y[1][1 - p][u][t]);
                        }
                }
            }
        }
    }
    // F[j] = Y0 + Y1 (north), F[Th-1-j] = Y0 - Y1 (south), both hemispheres
    // of one output per pass, staged [hemisphere][shell][theta][m], then one
    // run of BS_MQ m per (shell, theta row)
    constexpr int NQ = BS_MQ / 2;
#pragma unroll
    for (int o = 0; o < NOUT; o++) {
        if (o) __syncthreads();
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int sh = 4 * q4 + r, th = 16 * t + l16;
                const float y0r = Y[o][0][0][t][r], y1r = Y[o][1][0][t][r];
                const float y0i = Y[o][0][1][t][r], y1i = Y[o][1][1][t][r];
                St[(0 * BS_SH + sh) * BS_SHS + th * BS_MQ + wave] = make_float2(y0r + y1r, y0i + y1i);
                St[(1 * BS_SH + sh) * BS_SHS + th * BS_MQ + wave] = make_float2(y0r - y1r, y0i - y1i);
            }
        __syncthreads();
        float* out = o == 0 ? pick_job(a.out0, job) : pick_job(a.out1, job);
        for (int idx = tid; idx < NQ * 2 * BS_SH * BS_TH; idx += 64 * BS_MQ) {
            const int qq = idx % NQ, item = idx / NQ;
            const int th = item % BS_TH, rest = item / BS_TH;
            const int sh = rest % BS_SH, mir = rest / BS_SH;
            const int n = n0 + sh, j = jb + th;
            if (n >= a.N || j >= a.Th2) continue;
            const int jrow = mir ? a.Th - 1 - j : j;
            if (mir && jrow < a.Th2) continue;
            const float2* sp = St + (mir * BS_SH + sh) * BS_SHS + th * BS_MQ + 2 * qq;
            float4 v = make_float4(sp[0].x, sp[0].y, sp[1].x, sp[1].y);
            if constexpr (KIND == 2) {
                const float f = (float)a.rinv[n];
                v.x *= f;
                v.y *= f;
                v.z *= f;
                v.w *= f;
            }
            reinterpret_cast<float4*>(out + 2 * (((size_t)n * a.Th + jrow) * a.MP + m0))[qq] = v;
        }
    }
}

int synb_launch(Ctx& c, int kind, int njob, const void* const* in0, const void* const* in1, void* const* out0,
                void* const* out1) {
    if (njob < 1 || njob > BX_MAXJ) {
        set_error("synthesis: 1..6 transforms per launch");
        return LSD_E_ARG;
    }
    SynbArgs a;
    a.N = c.N;
    a.Th = c.Th;
    a.Th2 = c.Th2;
    a.M = c.M;
    a.MP = c.MP;
    a.H = c.H;
    a.slots = c.d_slots;
    a.mslot = c.d_mslot;
    a.sbase = c.d_bx_sbase;
    a.nthb = c.bx_nthb;
    const size_t per_kind = (size_t)c.bx_nsb * c.bx_nthb * 4096;
    const int ka = kind == SYN_SCALAR ? 0 : kind == SYN_VECTOR ? 3 : 1;
    const int kb = kind == SYN_VECTOR ? 4 : 2;
    a.tabA = c.d_tabb + ka * per_kind;
    a.tabB = c.d_tabb + kb * per_kind;
    a.rinv = c.d_r + 2 * c.N;
    for (int k = 0; k < BX_MAXJ; k++) {
        const int q = k < njob ? k : 0;
        a.in0[k] = reinterpret_cast<const float*>(in0[q]);
        a.in1[k] = in1 ? reinterpret_cast<const float*>(in1[q]) : nullptr;
        a.out0[k] = reinterpret_cast<float*>(out0[q]);
        a.out1[k] = out1 ? reinterpret_cast<float*>(out1[q]) : nullptr;
    }
    a.nth = c.bx_nthb;
    a.nsb = (c.N + BS_SH - 1) / BS_SH;
    a.nmo = (c.M + 15) / 16;
    a.njob = njob;
    const dim3 grid(bx_grid(a.nmo * njob * a.nsb, (16 / BS_MQ) * a.nth));
    if (kind == SYN_SCALAR) hipLaunchKernelGGL(k_synb<0>, grid, dim3(64 * BS_MQ), 0, c.stream, a);
    else if (kind == SYN_VECTOR) hipLaunchKernelGGL(k_synb<1>, grid, dim3(64 * BS_MQ), 0, c.stream, a);
    else hipLaunchKernelGGL(k_synb<2>, grid, dim3(64 * BS_MQ), 0, c.stream, a);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

// -------------------------------------------------------------- analysis

constexpr int BA_SH = 16;  // shells per wave
constexpr int BA_W = 2;    // waves (shell blocks) per workgroup

struct AnabArgs {
    int N, Th2, M, njb;
    int nsbp, maxab, njob, ngroups;
    const Slot* slots;
    const int* order;
    const int* abase;       // [2 nslot] first analysis block of (slot, class)
    const __bf16* tP;       // split tables, layout of k_tab_anab
    const __bf16* tL;
    const __bf16* tS;
    const float* g1[BX_MAXJ];  // folded [n][M][2][Th2] complex
    const float* g2[BX_MAXJ];
    const float* g3[BX_MAXJ];
    float* o1[BX_MAXJ];  // coll [h][2N]
    float* o2[BX_MAXJ];
    float* o3[BX_MAXJ];
};

// One wave: a 64-harmonic block of one (slot, class) and 16 shells, over all
// theta in steps of 32.  VEC: q, s, t from (r, theta, phi); else one field.
template <bool VEC>
__global__ void __launch_bounds__(64 * BA_W) k_anab(AnabArgs a) {
    constexpr int NO = VEC ? 3 : 1;
    constexpr int NG = VEC ? 5 : 1;
    const BxId bid = bx_block(a.ngroups, 2 * a.maxab * a.nsbp);
    if (!bid.valid) return;
    const int job = bid.group % a.njob;
    const int sl = wave_uniform(a.order[bid.group / a.njob]);
    const Slot s = a.slots[sl];
    const int par = bid.cc / (a.maxab * a.nsbp), rest = bid.cc - par * a.maxab * a.nsbp;
    const int hb = rest / a.nsbp, sbp = rest - hb * a.nsbp;
    const int n_even = (s.len + 1) / 2;
    const int ncls = par ? s.len / 2 : n_even;
    if (64 * hb >= ncls) return;
    const int tid = threadIdx.x, wave = wave_uniform(tid >> 6), lane = tid & 63, l16 = lane & 15, q4 = lane >> 4;
    const int n0 = (sbp * BA_W + wave) * BA_SH;
    if (n0 >= a.N) return;  // no barriers: a wave may leave alone
    const int ntile = min(4, (ncls - 64 * hb + 15) / 16);
    const int ncl = min(n0 + l16, a.N - 1);
    const float* gf[3] = {pick_job(a.g1, job), pick_job(a.g2, job), pick_job(a.g3, job)};
    const size_t grow = ((size_t)ncl * a.M + s.mfft) * 2;
    const float* gr[NG];
    {
        const int fsel[5] = {0, 1, 2, 1, 2};
        const int psel[5] = {par, 1 - par, par, par, 1 - par};
#pragma unroll
        for (int g = 0; g < NG; g++) gr[g] = gf[fsel[g]] + (grow + psel[g]) * a.Th2 * 2;
    }
    const int blk = a.abase[2 * sl + par] + hb;
    v4f_b acc[4][NO][2];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int o = 0; o < NO; o++) acc[t][o][0] = acc[t][o][1] = v4f_b{0, 0, 0, 0};
    for (int jb = 0; jb < a.njb; jb++) {
        // folded data of this lane's shell at theta 32 jb + 8 q4 + e, split
        Split8 G[NG][2];
#pragma unroll
        for (int g = 0; g < NG; g++) {
            float re[8], im[8];
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int th = 32 * jb + 8 * q4 + e;
                const float2 v = th < a.Th2 ? *reinterpret_cast<const float2*>(gr[g] + 2 * th) : make_float2(0.f, 0.f);
                re[e] = v.x;
                im[e] = v.y;
            }
            G[g][0] = split8(re);
            G[g][1] = split8(im);
        }
        const size_t tb = ((size_t)blk * a.njb + jb) * 4096;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            if (t >= ntile) break;
            const size_t o = tb + ((size_t)(t * 16 + l16) * 4 + q4) * 16;
            const bf16x8 ph = *reinterpret_cast<const bf16x8*>(a.tP + o);
            const bf16x8 pl = *reinterpret_cast<const bf16x8*>(a.tP + o + 8);
            const Split8 P = {ph, pl};
#pragma unroll
            for (int u = 0; u < 2; u++) acc[t][0][u] = mfma3(P, G[0][u].hi, G[0][u].lo, acc[t][0][u]);
            if constexpr (VEC) {
                const Split8 Lt = {*reinterpret_cast<const bf16x8*>(a.tL + o),
                                   *reinterpret_cast<const bf16x8*>(a.tL + o + 8)};
                const Split8 St_ = {*reinterpret_cast<const bf16x8*>(a.tS + o),
                                    *reinterpret_cast<const bf16x8*>(a.tS + o + 8)};
                // s.re += L Y2o.re + S Y3s.im ; s.im += L Y2o.im - S Y3s.re
                acc[t][1][0] = mfma3(Lt, G[1][0].hi, G[1][0].lo, acc[t][1][0]);
                acc[t][1][0] = mfma3(St_, G[2][1].hi, G[2][1].lo, acc[t][1][0]);
                acc[t][1][1] = mfma3(Lt, G[1][1].hi, G[1][1].lo, acc[t][1][1]);
                acc[t][1][1] = mfma3(St_, neg8(G[2][0].hi), neg8(G[2][0].lo), acc[t][1][1]);
                // t.re = -Y2s.im S + Y3o.re L ; t.im = Y2s.re S + Y3o.im L
                acc[t][2][0] = mfma3(St_, neg8(G[3][1].hi), neg8(G[3][1].lo), acc[t][2][0]);
                acc[t][2][0] = mfma3(Lt, G[4][0].hi, G[4][0].lo, acc[t][2][0]);
                acc[t][2][1] = mfma3(St_, G[3][0].hi, G[3][0].lo, acc[t][2][1]);
                acc[t][2][1] = mfma3(Lt, G[4][1].hi, G[4][1].lo, acc[t][2][1]);
            }
        }
    }
    // the lane holds class harmonics 64 hb + 16 t + 4 q4 + r of shell n0 + l16
    const int n = n0 + l16;
    if (n >= a.N) return;
    const int C2 = 2 * a.N;
    float* of[3] = {pick_job(a.o1, job), pick_job(a.o2, job), pick_job(a.o3, job)};
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

int anab_launch(Ctx& c, int kind, int njob, const void* const* g0, const void* const* g1, const void* const* g2,
                void* const* o0, void* const* o1, void* const* o2) {
    if (njob < 1 || njob > BX_MAXJ) {
        set_error("analysis: 1..6 transforms per launch");
        return LSD_E_ARG;
    }
    AnabArgs a;
    a.N = c.N;
    a.Th2 = c.Th2;
    a.M = c.M;
    a.njb = c.bx_njb;
    a.slots = c.d_slots;
    a.order = c.d_slot_order;
    a.abase = c.d_bx_sbase + 2 * c.nslot;
    const size_t per_kind = (size_t)c.bx_nab * c.bx_njb * 4096;
    a.tP = c.d_atabb;
    a.tL = c.d_atabb + per_kind;
    a.tS = c.d_atabb + 2 * per_kind;
    for (int k = 0; k < BX_MAXJ; k++) {
        const int q = k < njob ? k : 0;
        a.g1[k] = reinterpret_cast<const float*>(g0[q]);
        a.g2[k] = g1 ? reinterpret_cast<const float*>(g1[q]) : a.g1[k];
        a.g3[k] = g2 ? reinterpret_cast<const float*>(g2[q]) : a.g1[k];
        a.o1[k] = reinterpret_cast<float*>(o0[q]);
        a.o2[k] = o1 ? reinterpret_cast<float*>(o1[q]) : nullptr;
        a.o3[k] = o2 ? reinterpret_cast<float*>(o2[q]) : nullptr;
    }
    a.nsbp = ((c.N + BA_SH - 1) / BA_SH + BA_W - 1) / BA_W;
    a.maxab = c.bx_maxab;
    a.njob = njob;
    a.ngroups = c.nslot * njob;
    const dim3 grid(bx_grid(a.ngroups, 2 * a.maxab * a.nsbp));
    if (kind == ANA_SCALAR) hipLaunchKernelGGL(k_anab<false>, grid, dim3(64 * BA_W), 0, c.stream, a);
    else hipLaunchKernelGGL(k_anab<true>, grid, dim3(64 * BA_W), 0, c.stream, a);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

}  // namespace lsd
