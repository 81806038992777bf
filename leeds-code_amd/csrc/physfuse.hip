// Physical-space part of one nonlinear evaluation in a single pass:
// inverse phi FFTs of the 18 synthesised fields (tra_s2p,
// modules/transform.fftw3.F90:280-364), the pointwise Boussinesq products
// (modules/nonlinear.F90:71-146, 266-272, 335-341, 420-430), and the
// forward FFTs with the even/odd fold of tra_p2s / tra_rtp2qst
// (:379-466, 862-864) -- without the physical fields ever reaching HBM.
//
// A workgroup owns one shell and one theta row pair (j, Th-1-j).  The two
// real rows of a field are packed as one complex sequence (row j real, row
// Th-1-j imaginary), so the inverse transform, the products and the forward
// transform of both rows share every FFT, and the forward output separates
// directly into the folded north+south / north-south buckets the analysis
// consumes.  Inputs go through LDS in batches of up to six fields (u and
// curl u; B and curl B; grad C and grad xi), each loaded into registers one
// batch ahead; each thread keeps the values of its own phi point in
// registers between batches, and does one radix-8 butterfly (length 512) or
// one point (general lengths) per FFT stage.  The sequences are stored with
// one pad slot per 8 complex values, against bank conflicts of the stride-8
// Stockham writes.
//
// Fourier data are stored [f][n][j][m] (m padded to MP), so a row's values
// are contiguous.  Folded outputs are stored [o][n][m][parity][j]: a value
// is one piece of a 128-byte line shared by 16 (FP32) or 8 (FP64)
// consecutive row pairs.  Blocks are numbered so that the row pairs of one
// line are consecutive workgroups on the same XCD (workgroups go
// round-robin over the 8 XCDs), which makes them share each line through
// that XCD's L2.
#include "fft_lds.h"
#include "lsd_internal.h"

namespace lsd {

// Threads per workgroup: one per phi point, 512 for Ph <= 512 and 1024 up
// to Ph = 1024.
constexpr int PF_THREADS = 512;
constexpr int PF_THREADS_MAX = 1024;
constexpr int PF_EPT = 8;      // 8 sequences x 512 points = one radix-8 butterfly per thread
constexpr int PF_MAXSEQ = 8;

// Row pairs per group: as many as fill one 128-byte line of a Fourier or
// folded row with their complex values (16 in FP32, 8 in FP64).
template <typename R>
constexpr int PF_GROUP = 128 / (int)sizeof(typename Cx<R>::t);

template <typename R>
struct PhysArgs {
    FftPlan pl;
    int N, Th, Th2, M, MP, ngroups, gpr;  // gpr: groups of PF_GROUP<R> row pairs per shell
    float inv_m, inv_w, inv_2m;       // 1/M, 1/(Ph - 2M + 1), 1/(2M) for quot()
    R cor, cmag, qc, qx, scale;
    int cod, comp, mag;
    const double* gp;   // [Th] cos(theta)
    const double* sn;   // [Th] sin(theta)
    const typename Cx<R>::t* fourier;  // [18][N][Th][MP]
    typename Cx<R>::t* folded;         // [8][N][M][2][Th2]
    const typename Cx<R>::t* tw;
    // CFL partial maxima of non_maxtstep (null: not formed): per workgroup
    // cfl[(shell Th2 + row pair) 2 + k], with the per-shell constants
    // cfl_d[2 shell + k] and ep = E/Pm, eE = E, ePm = Pm
    double* cfl;
    const double* cfl_d;
    R ep, eE, ePm;
};

// i / d for 0 <= i < 2^21, d > 0, given inv = 1/d rounded to FP32: the true
// quotient q has q + 1/(2d) <= (i + 1/2)/d <= q + 1 - 1/(2d), and the
// computed product is within 2^-23 (i + 1/2)/d < 1/(2d) of (i + 1/2)/d, so
// its floor is q.
__device__ __forceinline__ int quot(int i, float inv) { return (int)(((float)i + 0.5f) * inv); }

// One batch of inputs -- nb fields of row pair (jp, jm): f0, f0+1, f0+2,
// then f1, f1+1, f1+2 -- held in registers from its loads to its LDS pack,
// so that the loads of the next batch are in flight while the current one
// is transformed (a barrier waits on LDS operations only).  Item i of
// thread t is value idx = t + i NT of the batch's nb M (field, m) values;
// since M <= Ph / 2 <= NT / 2, 6 M <= 3 NT and three items cover a batch.
constexpr int PF_ITEMS = 3;

template <typename V>
struct Batch {
    V A[PF_ITEMS], B[PF_ITEMS];  // rows jp and jm
};

template <int NT, typename V>
__device__ __forceinline__ void load_batch(Batch<V>& bt, const V* __restrict__ fourier, int f0, int f1, int nb,
                                           int sh, int N, int M, int MP, int Th, int jp, int jm, float inv_m) {
    // no branch holds the loads (lanes past the end re-read the last value;
    // an empty batch reads value 0 of f0), so the values are not copied
    // at a join, which would wait for them before the pack
    const int last = max(nb * M, 1) - 1;
#pragma unroll
    for (int i = 0; i < PF_ITEMS; i++) {
        const int idx = min((int)threadIdx.x + i * NT, last);
        const int b = quot(idx, inv_m), k = idx - b * M;
        const int fid = b < 3 ? f0 + b : f1 + b - 3;
        const V* X = fourier + ((size_t)fid * N + sh) * Th * MP + k;
        bt.A[i] = X[(size_t)jp * MP];
        bt.B[i] = X[(size_t)jm * MP];
    }
}

// Packs a loaded batch into Hermitian-extended complex sequences (row jp
// real part, row jm imaginary part); the DC imaginary parts are dropped as a
// c2r transform does.  Unless the first transform stage skips the zero
// band k = M .. Ph-M (PRUNED), it is written with zeros.  The caller
// synchronises before the transform.
template <int NT, bool PRUNED, typename V>
__device__ __forceinline__ void pack_batch(V* Z, const Batch<V>& bt, int nb, int M, int Ph, float inv_m,
                                           float inv_w) {
#pragma unroll
    for (int i = 0; i < PF_ITEMS; i++) {
        const int idx = (int)threadIdx.x + i * NT;
        if (i * NT < nb * M && idx < nb * M) {
            const int b = quot(idx, inv_m), k = idx - b * M;
            const V A = bt.A[i], B = bt.B[i];
            const int z = b * Ph;
            if (k == 0) {
                Z[lidx<true>(z)] = cx<V>(A.x, B.x);
            } else {
                Z[lidx<true>(z + k)] = cx<V>(A.x - B.y, A.y + B.x);
                Z[lidx<true>(z + Ph - k)] = cx<V>(A.x + B.y, B.x - A.y);
            }
        }
    }
    if constexpr (!PRUNED) {
        const int W = Ph - 2 * M + 1;  // zero band k = M .. Ph-M
        for (int idx = threadIdx.x; idx < nb * W; idx += NT) {
            const int b = quot(idx, inv_w), k = M + (idx - b * W);
            Z[lidx<true>(b * Ph + k)] = cx<V>(0, 0);
        }
    }
}

// Loads and packs a batch in one pass (the form without lookahead).
template <int NT, typename V>
__device__ void pack_direct(V* Z, const V* __restrict__ fourier, int f0, int f1, int nb, int sh, int N, int M,
                            int MP, int Th, int Ph, int jp, int jm, float inv_m, float inv_w) {
    for (int idx = threadIdx.x; idx < nb * M; idx += NT) {
        const int b = quot(idx, inv_m), k = idx - b * M;
        const int fid = b < 3 ? f0 + b : f1 + b - 3;
        const V* X = fourier + ((size_t)fid * N + sh) * Th * MP + k;
        const V A = X[(size_t)jp * MP], B = X[(size_t)jm * MP];
        const int z = b * Ph;
        if (k == 0) {
            Z[lidx<true>(z)] = cx<V>(A.x, B.x);
        } else {
            Z[lidx<true>(z + k)] = cx<V>(A.x - B.y, A.y + B.x);
            Z[lidx<true>(z + Ph - k)] = cx<V>(A.x + B.y, B.x - A.y);
        }
    }
    const int W = Ph - 2 * M + 1;
    for (int idx = threadIdx.x; idx < nb * W; idx += NT) {
        const int b = quot(idx, inv_w), k = M + (idx - b * W);
        Z[lidx<true>(b * Ph + k)] = cx<V>(0, 0);
    }
}

// Radix-8 stage of a length-512 transform (512 = 8^3) with every index a
// compile-time shift: NS = 1, 8, 64 is the Stockham sub-transform size.
// One butterfly per thread (nf <= 8 sequences x 64 butterflies).  With one
// pad slot per 8 values, lidx(x + 8c) = lidx(x) + 9c, so all eight reads
// and all eight writes of a butterfly are one base address plus
// compile-time offsets.  Twiddles come from the stage's own table
// tw[(r-1)*NS + jm] = exp(2 pi i jm r / (8 NS)) (see stage_twiddles), read
// by consecutive lanes at consecutive entries.
template <int NS, int SG, typename V>
__device__ __forceinline__ void stage512(V* Z, int nf, const V* tw) {
    typedef decltype(V().x) S;
    constexpr S sg = (S)SG;
    const int b = threadIdx.x;
    const bool on = b < nf * 64;
    const int p = b >> 6, j = b & 63, jm = j & (NS - 1);
    V v[8];
    if (on) {
        const int rb = lidx<true>(p * 512 + j);
#pragma unroll
        for (int r = 0; r < 8; r++) v[r] = Z[rb + 72 * r];
        if (NS > 1) {
            const V* t = tw + (NS == 8 ? 0 : 56) + jm;
#pragma unroll
            for (int r = 1; r < 8; r++) {
                V w = t[(r - 1) * NS];
                w.y *= sg;
                v[r] = cmulv(v[r], w);
            }
        }
        dft<8>(v, sg);
    }
    __syncthreads();
    if (on) {
        const int wb = lidx<true>(p * 512 + (j / NS) * NS * 8 + jm);
        constexpr int step = NS == 1 ? 1 : (NS == 8 ? 9 : 72);
#pragma unroll
        for (int r = 0; r < 8; r++) Z[wb + step * r] = v[r];
    }
    __syncthreads();
}

// The first stage of the inverse length-512 transform when 128 < M <= 192,
// without reading the zero band k = M .. 512-M: input k = j + 64 r is zero
// for r = 3, 4, for r = 2 when j >= M - 128 and for r = 5 when j <= 192 - M,
// so the band need not be written and the butterfly drops the terms of
// inputs 3 and 4.  Same results as stage512<1, 1> on a zero-filled band
// (up to the sign of exact zeros).
template <typename V>
__device__ __forceinline__ void stage512_inv_first(V* Z, int nf, int M) {
    typedef decltype(V().x) S;
    const int b = threadIdx.x;
    const bool on = b < nf * 64;
    const int p = b >> 6, j = b & 63;
    V v[8];
    if (on) {
        const int rb = lidx<true>(p * 512 + j);
        const V zero = cx<V>((S)0, (S)0);
        const V x0 = Z[rb], x1 = Z[rb + 72], x2r = Z[rb + 144], x5r = Z[rb + 360], x6 = Z[rb + 432],
                x7 = Z[rb + 504];
        const V x2 = j < M - 128 ? x2r : zero, x5 = j > 192 - M ? x5r : zero;
        // dft<8> with inputs 3 and 4 zero: even part (x0, x2, 0, x6), odd
        // part (x1, 0, x5, x7); w = exp(2 pi i / 8)
        const S c = (S)0.70710678118654752440;
        const V t2 = caddv(x2, x6), t3 = mul_si(csubv(x2, x6), (S)1);
        const V e0 = caddv(x0, t2), e2 = csubv(x0, t2), e1 = caddv(x0, t3), e3 = csubv(x0, t3);
        const V u0 = caddv(x1, x5), u1 = csubv(x1, x5), u3 = mul_si(cx<V>(-x7.x, -x7.y), (S)1);
        const V o0 = caddv(u0, x7), o2 = csubv(u0, x7), o1 = caddv(u1, u3), o3 = csubv(u1, u3);
        const V w1 = cx<V>(c * (o1.x - o1.y), c * (o1.y + o1.x));
        const V w2 = mul_si(o2, (S)1);
        const V w3 = cx<V>(c * (-o3.x - o3.y), c * (-o3.y + o3.x));
        v[0] = caddv(e0, o0);
        v[4] = csubv(e0, o0);
        v[1] = caddv(e1, w1);
        v[5] = csubv(e1, w1);
        v[2] = caddv(e2, w2);
        v[6] = csubv(e2, w2);
        v[3] = caddv(e3, w3);
        v[7] = csubv(e3, w3);
    }
    __syncthreads();
    if (on) {
        const int wb = lidx<true>(p * 512 + j * 8);
#pragma unroll
        for (int r = 0; r < 8; r++) Z[wb + r] = v[r];
    }
    __syncthreads();
}

// The last stage of the forward length-512 transform when M <= 192: outputs
// k = jm + 64 r for r = 3, 4 (k = 192 .. 319) are never read, since only
// k < M and k > 512 - M are gathered, so they are neither formed nor
// written.
template <typename V>
__device__ __forceinline__ void stage512_fwd_last(V* Z, int nf, const V* tw) {
    typedef decltype(V().x) S;
    constexpr S sg = (S)-1;
    const int b = threadIdx.x;
    const bool on = b < nf * 64;
    const int p = b >> 6, j = b & 63;
    V v[8];
    if (on) {
        const int rb = lidx<true>(p * 512 + j);
#pragma unroll
        for (int r = 0; r < 8; r++) v[r] = Z[rb + 72 * r];
        const V* t = tw + 56 + j;
#pragma unroll
        for (int r = 1; r < 8; r++) {
            V w = t[(r - 1) * 64];
            w.y *= sg;
            v[r] = cmulv(v[r], w);
        }
        const S c = (S)0.70710678118654752440;
        V e[4] = {v[0], v[2], v[4], v[6]};
        V o[4] = {v[1], v[3], v[5], v[7]};
        dft<4>(e, sg);
        dft<4>(o, sg);
        const V w1 = cx<V>(c * (o[1].x - sg * o[1].y), c * (o[1].y + sg * o[1].x));
        const V w2 = mul_si(o[2], sg);
        const V w3 = cx<V>(c * (-o[3].x - sg * o[3].y), c * (-o[3].y + sg * o[3].x));
        v[0] = caddv(e[0], o[0]);
        v[1] = caddv(e[1], w1);
        v[5] = csubv(e[1], w1);
        v[2] = caddv(e[2], w2);
        v[6] = csubv(e[2], w2);
        v[7] = csubv(e[3], w3);
    }
    __syncthreads();
    if (on) {
        const int wb = lidx<true>(p * 512 + j);
        Z[wb] = v[0];
        Z[wb + 72] = v[1];
        Z[wb + 144] = v[2];
        Z[wb + 360] = v[5];
        Z[wb + 432] = v[6];
        Z[wb + 504] = v[7];
    }
    __syncthreads();
}

// Twiddle tables of the two twiddled stages of a length-512 transform,
// from the full table exp(2 pi i k / 512): stage NS = 8 at [0, 56),
// entry (r-1)*8 + jm = w^(8 jm r); stage NS = 64 at [56, 504), entry
// 56 + (r-1)*64 + jm = w^(jm r).
template <typename V>
__device__ void stage_twiddles(V* twl, const V* __restrict__ tw) {
    for (int i = threadIdx.x; i < 504; i += PF_THREADS) {
        int k;
        if (i < 56) {
            const int r = i / 8 + 1, jm = i & 7;
            k = 8 * jm * r;
        } else {
            const int q = i - 56, r = q / 64 + 1, jm = q & 63;
            k = jm * r;
        }
        twl[i] = tw[k];
    }
}

// PRUNED (length 512 with 128 < M <= 192): the inverse transform skips the
// zero band of its input, the forward one the outputs that are not gathered.
template <int SG, bool R512, bool PRUNED, int NT, typename V>
__device__ __forceinline__ void fft_pf(V* Z, const FftPlan& pl, int nf, const V* tw, int M) {
    if constexpr (R512) {
        if constexpr (PRUNED && SG == 1) stage512_inv_first(Z, nf, M);
        else stage512<1, SG>(Z, nf, tw);
        stage512<8, SG>(Z, nf, tw);
        if constexpr (PRUNED && SG == -1) stage512_fwd_last(Z, nf, tw);
        else stage512<64, SG>(Z, nf, tw);
    } else {
        fft_lds<V, NT, PF_EPT, true>(Z, pl, nf, tw, (decltype(V().x))SG);
    }
}

// Length-512 transforms (radices 8, 8, 8) have their own kernel, which
// needs far fewer registers than the general-length path and a smaller
// twiddle area (504 entries), so more workgroups fit on a CU.
constexpr int TW512 = 504;

bool physfuse_r512(int Ph) {
    const FftPlan pl = make_fft_plan(Ph);
    return Ph == 512 && pl.nstage == 3 && pl.radix[0] == 8 && pl.radix[1] == 8 && pl.radix[2] == 8;
}

// LDS of one workgroup: the twiddles, then PF_MAXSEQ padded sequences
// (one pad slot per 8 values).
template <typename V>
size_t physfuse_lds(int Ph, bool r512) {
    return sizeof(V) * ((r512 ? TW512 : Ph) + (size_t)PF_MAXSEQ * Ph / 8 * 9);
}

// CFL: also form the CFL partial maxima (a separate instantiation, so the
// usual pass carries no registers for them).  PRUNED (length 512 with
// 128 < M <= 192): the transforms skip the zero band (fft_pf).
template <typename R, bool R512, int NT, bool CFL, bool PRUNED>
__global__ void __launch_bounds__(NT) k_physfuse(PhysArgs<R> a) {
    static_assert(!R512 || NT == PF_THREADS, "length-512 stages take one butterfly per thread of 512");
    static_assert(R512 || !PRUNED, "only the length-512 stages are pruned");
    constexpr int PP = 1;  // phi points per thread
    typedef typename Cx<R>::t V;
    extern __shared__ unsigned char lds_raw[];
    // twiddles first (read by every FFT stage), then the padded sequences
    V* twl = reinterpret_cast<V*>(lds_raw);
    V* Z = twl + (R512 ? TW512 : a.pl.n);
    constexpr int GS = PF_GROUP<R>;
    const int b = blockIdx.x, xcd = b & 7, q = b >> 3;
    const int G = (q / GS) * 8 + xcd, member = q % GS;
    if (G >= a.ngroups) return;
    const int sh = G / a.gpr, jg = G - sh * a.gpr;
    const int jp = jg * GS + member;
    if (jp >= a.Th2) return;
    const int jm = a.Th - 1 - jp;
    const int Ph = a.pl.n, N = a.N, M = a.M, Th = a.Th;
    const int t = threadIdx.x;
    // batches in order: u and curl u; B and curl B (if magnetic); grad C
    // and grad xi (as enabled).  Each is loaded one batch ahead.
    const int ng = 3 * (a.cod + a.comp);
    const int gf0 = a.cod ? LSD_P_COD_GR : LSD_P_COMP_GR;
    // (the FP64 form at 1024 threads has no registers to spare: it loads
    // each batch as it packs it)
    constexpr bool AHEAD = !(sizeof(R) == 8 && NT == PF_THREADS_MAX);
    Batch<V> inb;
    auto load_grad = [&]() {
        load_batch<NT>(inb, a.fourier, gf0, LSD_P_COMP_GR, ng, sh, N, M, a.MP, Th, jp, jm, a.inv_m);
    };
    auto pack = [&](int f0, int f1, int nb) {
        if constexpr (AHEAD) pack_batch<NT, PRUNED>(Z, inb, nb, M, Ph, a.inv_m, a.inv_w);
        else pack_direct<NT>(Z, a.fourier, f0, f1, nb, sh, N, M, a.MP, Th, Ph, jp, jm, a.inv_m, a.inv_w);
    };
    if constexpr (AHEAD)
        load_batch<NT>(inb, a.fourier, LSD_P_VEL_R, LSD_P_VCURL_R, 6, sh, N, M, a.MP, Th, jp, jm, a.inv_m);
    if constexpr (R512) stage_twiddles(twl, a.tw);
    else
        for (int i = t; i < Ph; i += NT) twl[i] = a.tw[i];
    const R sn = (R)a.sn[jp];
    const R gpr[2] = {(R)a.gp[jp], (R)a.gp[jm]};

    R u[PP][2][3], nu[PP][2][3], e[PP][2][3], nc[PP][2], nx[PP][2];
    // CFL maxima over this thread's points (radial, horizontal)
    R cm1 = (R)-1e30, cm2 = (R)-1e30;

    // u and curl u: advection and Coriolis
    {
        pack(LSD_P_VEL_R, LSD_P_VCURL_R, 6);
        __syncthreads();
        if constexpr (AHEAD) {
            // the next batch, chosen by uniform selects rather than branches
            const int f0 = a.mag ? LSD_P_MAG_R : (ng ? gf0 : LSD_P_VEL_R);
            const int f1 = a.mag ? LSD_P_MCURL_R : LSD_P_COMP_GR;
            const int nb = a.mag ? 6 : ng;
            load_batch<NT>(inb, a.fourier, f0, f1, nb, sh, N, M, a.MP, Th, jp, jm, a.inv_m);
        }
        fft_pf<1, R512, PRUNED, NT>(Z, a.pl, 6, twl, M);
#pragma unroll
        for (int p = 0; p < PP; p++) {
            const int x = t + p * NT;
            if (x >= Ph) continue;
            V z[6];
#pragma unroll
            for (int f = 0; f < 6; f++) z[f] = Z[lidx<true>(f * Ph + x)];
#pragma unroll
            for (int r = 0; r < 2; r++) {
                const R ur = r ? z[0].y : z[0].x, ut = r ? z[1].y : z[1].x, up = r ? z[2].y : z[2].x;
                const R wr = r ? z[3].y : z[3].x, wt = r ? z[4].y : z[4].x, wp = r ? z[5].y : z[5].x;
                R vr = ut * wp - up * wt;
                R vt = up * wr - ur * wp;
                R vp = ur * wt - ut * wr;
                const R lgp = gpr[r];
                vr = vr + a.cor * sn * up;
                vt = vt + a.cor * lgp * up;
                vp = vp - a.cor * (lgp * ut + sn * ur);
                u[p][r][0] = ur;
                u[p][r][1] = ut;
                u[p][r][2] = up;
                if (CFL && !a.mag) {
                    // no magnetic field: the Alfven terms vanish
                    cm1 = fmax(cm1, fabs(ur));
                    cm2 = fmax(cm2, sqrt(ut * ut + up * up));
                }
                nu[p][r][0] = vr;
                nu[p][r][1] = vt;
                nu[p][r][2] = vp;
            }
        }
        __syncthreads();
    }
    // B and curl B: Lorentz force and induction E = u x B
    if (a.mag) {
        pack(LSD_P_MAG_R, LSD_P_MCURL_R, 6);
        __syncthreads();
        if constexpr (AHEAD) {
            if (ng) load_grad();
        }
        fft_pf<1, R512, PRUNED, NT>(Z, a.pl, 6, twl, M);
#pragma unroll
        for (int p = 0; p < PP; p++) {
            const int x = t + p * NT;
            if (x >= Ph) continue;
            V z[6];
#pragma unroll
            for (int f = 0; f < 6; f++) z[f] = Z[lidx<true>(f * Ph + x)];
#pragma unroll
            for (int r = 0; r < 2; r++) {
                const R br = r ? z[0].y : z[0].x, bt = r ? z[1].y : z[1].x, bp = r ? z[2].y : z[2].x;
                const R jr = r ? z[3].y : z[3].x, jt = r ? z[4].y : z[4].x, jpv = r ? z[5].y : z[5].x;
                const R ur = u[p][r][0], ut = u[p][r][1], up = u[p][r][2];
                nu[p][r][0] = nu[p][r][0] + a.cmag * (jt * bp - jpv * bt);
                nu[p][r][1] = nu[p][r][1] + a.cmag * (jpv * br - jr * bp);
                nu[p][r][2] = nu[p][r][2] + a.cmag * (jr * bt - jt * br);
                e[p][r][0] = ut * bp - up * bt;
                e[p][r][1] = up * br - ur * bp;
                e[p][r][2] = ur * bt - ut * br;
                if (CFL) {
                    const R d1 = (R)a.cfl_d[2 * sh], d2 = (R)a.cfl_d[2 * sh + 1];
                    const R pr = br * br, ph = bt * bt + bp * bp;
                    cm1 = fmax(cm1, pr / sqrt(pr * a.ep + d1) + fabs(ur));
                    cm2 = fmax(cm2, ph / sqrt(ph * a.eE / a.ePm + d2) + sqrt(ut * ut + up * up));
                }
            }
        }
        __syncthreads();
    }
    if constexpr (CFL) {
        // workgroup maxima: across the wave, then across the waves
        __shared__ R cfl_red[2][NT / 64];
        for (int o = 32; o > 0; o >>= 1) {
            cm1 = fmax(cm1, __shfl_xor(cm1, o));
            cm2 = fmax(cm2, __shfl_xor(cm2, o));
        }
        if ((t & 63) == 0) {
            cfl_red[0][t >> 6] = cm1;
            cfl_red[1][t >> 6] = cm2;
        }
        __syncthreads();
        if (t == 0) {
            for (int w = 1; w < NT / 64; w++) {
                cm1 = fmax(cm1, cfl_red[0][w]);
                cm2 = fmax(cm2, cfl_red[1][w]);
            }
            a.cfl[((size_t)sh * a.Th2 + jp) * 2] = (double)cm1;
            a.cfl[((size_t)sh * a.Th2 + jp) * 2 + 1] = (double)cm2;
        }
    }
    // grad C and grad xi: advection of codensity and composition
    if (ng) {
        pack(gf0, LSD_P_COMP_GR, ng);
        __syncthreads();
        fft_pf<1, R512, PRUNED, NT>(Z, a.pl, ng, twl, M);
        const int oc = 0, ox = a.cod ? 3 : 0;
#pragma unroll
        for (int p = 0; p < PP; p++) {
            const int x = t + p * NT;
            if (x >= Ph) continue;
#pragma unroll
            for (int r = 0; r < 2; r++) {
                const R ur = u[p][r][0], ut = u[p][r][1], up = u[p][r][2];
                if (a.cod) {
                    const V g0 = Z[lidx<true>((oc + 0) * Ph + x)], g1 = Z[lidx<true>((oc + 1) * Ph + x)],
                            g2 = Z[lidx<true>((oc + 2) * Ph + x)];
                    nc[p][r] = -a.qc * ur * (r ? g0.y : g0.x) - a.qc * ut * (r ? g1.y : g1.x) -
                               a.qc * up * (r ? g2.y : g2.x);
                }
                if (a.comp) {
                    const V g0 = Z[lidx<true>((ox + 0) * Ph + x)], g1 = Z[lidx<true>((ox + 1) * Ph + x)],
                            g2 = Z[lidx<true>((ox + 2) * Ph + x)];
                    nx[p][r] = -a.qx * ur * (r ? g0.y : g0.x) - a.qx * ut * (r ? g1.y : g1.x) -
                               a.qx * up * (r ? g2.y : g2.x);
                }
            }
        }
        __syncthreads();
    }
    // outputs, scaled by 1/Ph before the forward transform as tra_p2s does
    // sequences: N_u r, theta, phi; then codensity, composition, E r, theta, phi as enabled
    const int ncx = a.cod + a.comp;
    const int nout = 3 + ncx + (a.mag ? 3 : 0);
    const R sc = a.scale;
#pragma unroll
    for (int p = 0; p < PP; p++) {
        const int x = t + p * NT;
        if (x >= Ph) continue;
        int s = 0;
#pragma unroll
        for (int c = 0; c < 3; c++) Z[lidx<true>((s++) * Ph + x)] = cx<V>(nu[p][0][c] * sc, nu[p][1][c] * sc);
        if (a.cod) Z[lidx<true>((s++) * Ph + x)] = cx<V>(nc[p][0] * sc, nc[p][1] * sc);
        if (a.comp) Z[lidx<true>((s++) * Ph + x)] = cx<V>(nx[p][0] * sc, nx[p][1] * sc);
        if (a.mag) {
#pragma unroll
            for (int c = 0; c < 3; c++) Z[lidx<true>((s++) * Ph + x)] = cx<V>(e[p][0][c] * sc, e[p][1][c] * sc);
        }
    }
    __syncthreads();
    fft_pf<-1, R512, PRUNED, NT>(Z, a.pl, nout, twl, M);
    // separate the two rows and fold: G0 = X_j + X_mirror, G1 = X_j - X_mirror
    const size_t fstride = (size_t)N * M * 2 * a.Th2;
    for (int idx = t; idx < nout * M * 2; idx += NT) {
        const int s = quot(idx, a.inv_2m), rem = idx - s * 2 * M;
        const int k = rem >> 1, par = rem & 1;
        const V zk = Z[lidx<true>(s * Ph + k)];
        V zc = Z[lidx<true>(s * Ph + (k ? Ph - k : 0))];
        zc.y = -zc.y;
        const V A = cx<V>((R)0.5 * (zk.x + zc.x), (R)0.5 * (zk.y + zc.y));
        const V D = csubv(zk, zc);
        const V B = cx<V>((R)0.5 * D.y, (R)-0.5 * D.x);
        const V o = par ? csubv(A, B) : caddv(A, B);
        const int oid = s < 3 ? NO_UR + s : (s < 3 + ncx ? ((s == 3 && a.cod) ? NO_C : NO_X) : NO_ER + (s - 3 - ncx));
        a.folded[oid * fstride + (((size_t)sh * M + k) * 2 + par) * a.Th2 + jp] = o;
    }
}

bool physfuse_supported(const Ctx& c) {
    const int nt = c.Ph <= PF_THREADS ? PF_THREADS : PF_THREADS_MAX;
    return c.Ph <= PF_THREADS_MAX && c.Th % 2 == 0 && 6 * c.M <= PF_ITEMS * nt;
}

template <typename R>
static int physfuse_launch(Ctx& c, const void* fourier, void* folded, const void* tw) {
    typedef typename Cx<R>::t V;
    const lsd_config& k = c.cfg;
    PhysArgs<R> a;
    a.pl = make_fft_plan(c.Ph);
    a.N = c.N;
    a.Th = c.Th;
    a.Th2 = c.Th2;
    a.M = c.M;
    a.MP = c.MP;
    a.gpr = (c.Th2 + PF_GROUP<R> - 1) / PF_GROUP<R>;
    a.inv_m = 1.0f / (float)c.M;
    a.inv_w = 1.0f / (float)(c.Ph - 2 * c.M + 1);
    a.inv_2m = 1.0f / (float)(2 * c.M);
    a.ngroups = c.N * a.gpr;
    a.cor = (R)(2.0 * k.Pm / k.E);
    a.cmag = (R)(2.0 * k.Pm / k.E);
    a.qc = (R)(k.Pr / k.Pm);
    a.qx = (R)(k.Sc / k.Pm);
    a.scale = (R)(1.0 / (double)c.Ph);
    a.cod = k.cod;
    a.comp = k.comp;
    a.mag = k.mag;
    a.gp = c.d_gp;
    a.sn = c.d_sin;
    a.fourier = reinterpret_cast<const V*>(fourier);
    a.folded = reinterpret_cast<V*>(folded);
    a.tw = reinterpret_cast<const V*>(tw);
    a.cfl = c.cfl_on ? c.d_cfl_part : nullptr;
    a.cfl_d = c.d_cfl_d;
    a.ep = (R)(k.E / k.Pm);
    a.eE = (R)k.E;
    a.ePm = (R)k.Pm;
    if (c.cfl_on) c.cfl_parts = c.Th2;
    const int nblk = ((a.ngroups + 7) / 8) * 8 * PF_GROUP<R>;
    const bool r512 = physfuse_r512(c.Ph);
    const size_t lds = physfuse_lds<V>(c.Ph, r512);
    const int nin = 6 + (k.mag ? 6 : 0) + (k.cod ? 3 : 0) + (k.comp ? 3 : 0);
    const int nout = 3 + (k.mag ? 3 : 0) + (k.cod ? 1 : 0) + (k.comp ? 1 : 0);
    // algorithmic bytes: Fourier fields in, folded fields out
    ProfScope ps(c, LSD_K_PHYS,
                 (double)sizeof(V) * c.N * c.M * ((double)nin * c.Th + (double)nout * 2 * c.Th2));
    const bool pruned = r512 && c.M > 128 && c.M <= 192;
    const dim3 g(nblk);
    if (c.cfl_on) {
        if (pruned) hipLaunchKernelGGL((k_physfuse<R, true, PF_THREADS, true, true>), g, dim3(PF_THREADS), lds, c.stream, a);
        else if (r512)
            hipLaunchKernelGGL((k_physfuse<R, true, PF_THREADS, true, false>), g, dim3(PF_THREADS), lds, c.stream, a);
        else if (c.Ph <= PF_THREADS)
            hipLaunchKernelGGL((k_physfuse<R, false, PF_THREADS, true, false>), g, dim3(PF_THREADS), lds, c.stream, a);
        else
            hipLaunchKernelGGL((k_physfuse<R, false, PF_THREADS_MAX, true, false>), g, dim3(PF_THREADS_MAX), lds,
                               c.stream, a);
    } else if (pruned) {
        hipLaunchKernelGGL((k_physfuse<R, true, PF_THREADS, false, true>), g, dim3(PF_THREADS), lds, c.stream, a);
    } else if (r512) {
        hipLaunchKernelGGL((k_physfuse<R, true, PF_THREADS, false, false>), g, dim3(PF_THREADS), lds, c.stream, a);
    } else if (c.Ph <= PF_THREADS) {
        hipLaunchKernelGGL((k_physfuse<R, false, PF_THREADS, false, false>), g, dim3(PF_THREADS), lds, c.stream, a);
    } else {
        hipLaunchKernelGGL((k_physfuse<R, false, PF_THREADS_MAX, false, false>), g, dim3(PF_THREADS_MAX), lds,
                           c.stream, a);
    }
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

int launch_physfuse(Ctx& c, bool f32) {
    if (!physfuse_supported(c)) {
        set_error("fused physical-space pass needs Ph <= 1024 and Th even");
        return LSD_E_CONFIG;
    }
    if (f32) return physfuse_launch<float>(c, c.d_fourier, c.d_folded, c.d_tw32);
    return physfuse_launch<double>(c, c.d_fourier, c.d_folded, c.d_tw);
}

}  // namespace lsd
