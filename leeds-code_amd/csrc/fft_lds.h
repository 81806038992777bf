// Stockham FFT over sequences held in LDS, radices 8, 4, 2, 3, for a
// 256-thread workgroup; templated on the real type (double or float).
// Z holds nf sequences of length n back to back; the transform is in place
// (each stage reads its butterfly inputs into registers, synchronises, and
// writes the outputs), unnormalised, with exp(sg * 2 pi i jk / n).
#pragma once

#include <hip/hip_runtime.h>

namespace lsd {

struct FftPlan {
    int n;        // transform length Ph
    int nstage;
    int radix[16];
};

constexpr int FFT_THREADS = 256;
constexpr int FFT_EPT = 16;  // complex elements per thread per stage (nf * n <= 4096)

template <typename R>
struct Cx;
template <>
struct Cx<double> {
    typedef double2 t;
};
template <>
struct Cx<float> {
    typedef float2 t;
};

template <typename V>
__device__ __forceinline__ V cx(decltype(V().x) a, decltype(V().x) b) {
    V v;
    v.x = a;
    v.y = b;
    return v;
}
template <typename V>
__device__ __forceinline__ V cmulv(V a, V b) {
    return cx<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename V>
__device__ __forceinline__ V caddv(V a, V b) {
    return cx<V>(a.x + b.x, a.y + b.y);
}
template <typename V>
__device__ __forceinline__ V csubv(V a, V b) {
    return cx<V>(a.x - b.x, a.y - b.y);
}
// multiply by sg * i
template <typename V>
__device__ __forceinline__ V mul_si(V a, decltype(V().x) sg) {
    return cx<V>(-sg * a.y, sg * a.x);
}

template <int RAD, typename V>
__device__ __forceinline__ void dft(V* v, decltype(V().x) sg) {
    typedef decltype(V().x) S;
    if (RAD == 2) {
        V a = v[0], b = v[1];
        v[0] = caddv(a, b);
        v[1] = csubv(a, b);
    } else if (RAD == 3) {
        const S h = (S)0.86602540378443864676;  // sqrt(3)/2
        V t = caddv(v[1], v[2]);
        V d = csubv(v[1], v[2]);
        V m = cx<V>(v[0].x - (S)0.5 * t.x, v[0].y - (S)0.5 * t.y);
        V e = mul_si(cx<V>(h * d.x, h * d.y), sg);
        v[0] = caddv(v[0], t);
        v[1] = caddv(m, e);
        v[2] = csubv(m, e);
    } else if (RAD == 4) {
        V t0 = caddv(v[0], v[2]), t1 = csubv(v[0], v[2]);
        V t2 = caddv(v[1], v[3]), t3 = mul_si(csubv(v[1], v[3]), sg);
        v[0] = caddv(t0, t2);
        v[2] = csubv(t0, t2);
        v[1] = caddv(t1, t3);
        v[3] = csubv(t1, t3);
    } else if (RAD == 8) {
        const S c = (S)0.70710678118654752440;
        V e[4] = {v[0], v[2], v[4], v[6]};
        V o[4] = {v[1], v[3], v[5], v[7]};
        dft<4>(e, sg);
        dft<4>(o, sg);
        // w^k, w = exp(sg*2*pi*i/8)
        V w1 = cx<V>(c * (o[1].x - sg * o[1].y), c * (o[1].y + sg * o[1].x));
        V w2 = mul_si(o[2], sg);
        V w3 = cx<V>(c * (-o[3].x - sg * o[3].y), c * (-o[3].y + sg * o[3].x));
        v[0] = caddv(e[0], o[0]);
        v[4] = csubv(e[0], o[0]);
        v[1] = caddv(e[1], w1);
        v[5] = csubv(e[1], w1);
        v[2] = caddv(e[2], w2);
        v[6] = csubv(e[2], w2);
        v[3] = caddv(e[3], w3);
        v[7] = csubv(e[3], w3);
    }
}

// Index of element i of the sequences in LDS: dense, or with one pad slot
// per 8 complex values, which spreads the stride-8 writes of the first
// Stockham stages over the banks.
template <bool PAD>
__device__ __forceinline__ int lidx(int i) {
    return PAD ? i + (i >> 3) : i;
}

// One Stockham stage over nf sequences of length n held in Z[lidx(p*n + i)]
// by a workgroup of THREADS threads, each taking up to EPT elements.
template <int RAD, int THREADS, int EPT, bool PAD, typename V>
__device__ void fft_stage(V* Z, int n, int nf, int Ns, const V* tw, decltype(V().x) sg) {
    constexpr int BPT = (EPT + RAD - 1) / RAD;
    const int nb = n / RAD;
    const int total = nf * nb;
    const int step = n / (Ns * RAD);
    V v[BPT][RAD];
#pragma unroll
    for (int q = 0; q < BPT; q++) {
        int b = threadIdx.x + q * THREADS;
        if (b < total) {
            int p = b / nb, j = b - p * nb;
            int jm = j % Ns;
            const int base = p * n + j;
#pragma unroll
            for (int r = 0; r < RAD; r++) v[q][r] = Z[lidx<PAD>(base + r * nb)];
            if (jm) {
#pragma unroll
                for (int r = 1; r < RAD; r++) {
                    V w = tw[jm * r * step];
                    w.y *= sg;
                    v[q][r] = cmulv(v[q][r], w);
                }
            }
            dft<RAD>(v[q], sg);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < BPT; q++) {
        int b = threadIdx.x + q * THREADS;
        if (b < total) {
            int p = b / nb, j = b - p * nb;
            int jm = j % Ns;
            const int base = p * n + (j / Ns) * Ns * RAD + jm;
#pragma unroll
            for (int r = 0; r < RAD; r++) Z[lidx<PAD>(base + r * Ns)] = v[q][r];
        }
    }
    __syncthreads();
}

// Requires nf * n <= EPT * THREADS; ends synchronised.
template <typename V, int THREADS = FFT_THREADS, int EPT = FFT_EPT, bool PAD = false>
__device__ void fft_lds(V* Z, const FftPlan& pl, int nf, const V* tw, decltype(V().x) sg) {
    int Ns = 1;
    for (int s = 0; s < pl.nstage; s++) {
        int R = pl.radix[s];
        switch (R) {
        case 8: fft_stage<8, THREADS, EPT, PAD>(Z, pl.n, nf, Ns, tw, sg); break;
        case 4: fft_stage<4, THREADS, EPT, PAD>(Z, pl.n, nf, Ns, tw, sg); break;
        case 2: fft_stage<2, THREADS, EPT, PAD>(Z, pl.n, nf, Ns, tw, sg); break;
        default: fft_stage<3, THREADS, EPT, PAD>(Z, pl.n, nf, Ns, tw, sg); break;
        }
        Ns *= R;
    }
}

FftPlan make_fft_plan(int n);

}  // namespace lsd
