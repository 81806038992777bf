// Shared declarations of the banded-solve kernels (solve.hip: columns in
// LDS, any N; solve_direct.hip: two passes, N fixed at compile time).
#pragma once

#include "lsd_internal.h"

namespace lsd {

struct SolveArgs {
    int H, N, L, S;      // LDS row stride S = 2 hb + 1
    int on[6];           // field enabled: vT, vP, C, X, bT, bP
    double cimp;
    const int* hm;
    const int4* blocks;  // per block: degree l, first index into order, harmonic count
    const int* order;    // harmonics grouped by degree
    const double* lu;    // [NLU][L][N][LDAB]  FP64 factors
    const float* lu32;   // [NLU][L][N][LDAB]  FP32 copy (mixed forms)
    const int* piv;      // [NLU][L][N]
    const double* xb;    // [NLU][L][N][7]     unfactored X bands (mixed forms)
    const double* y;     // [NY][L][N][7]
    const double* gre;   // [L][2][N]
    const double* invg;  // [L][4]
    double2* x[6];       // state
    double2* nn[6];      // current N
    double2* hist[6];    // saved N
    double* partial;     // [8][gridDim.x]
};

// LDS of one block.  Columns first, then the operator data of one (kind, l).
struct Smem {
    double* W;     // [N][S] columns
    float* W32;    // [N][S] FP32 work columns (mixed forms)
    double* lu;    // [N][LDAB] (plain)
    double* xb;    // [N][7] (mixed)
    double* y;     // [N][7]
    double* g;     // [2][N]
    double* rd;    // [N] 1 / U(j,j) (plain)
    float* lu32;   // [N][LDAB] (mixed)
    float* rd32;   // [N] 1 / U(j,j) of the FP32 factors (mixed)
    int* piv;      // [N]
};

__host__ __device__ inline size_t align8(size_t b) { return (b + 7) & ~(size_t)7; }

__host__ __device__ inline size_t solve_smem_bytes(int N, int S, bool mix) {
    size_t b = 8 * (size_t)N * S;
    if (mix) b += align8(4 * (size_t)N * S) + 56 * (size_t)N + 40 * (size_t)N;
    else b += 8 * LSD_LDAB * (size_t)N;
    return b + 56 * (size_t)N + 16 * (size_t)N + 8 * (size_t)N + (mix ? 4 * (size_t)N : 0) + 4 * (size_t)N;
}

__device__ inline Smem carve(unsigned char* p, int N, int S, bool mix) {
    Smem m{};
    m.W = reinterpret_cast<double*>(p);
    p += 8 * (size_t)N * S;
    if (mix) {
        m.W32 = reinterpret_cast<float*>(p);
        p += align8(4 * (size_t)N * S);
        m.xb = reinterpret_cast<double*>(p);
        p += 56 * (size_t)N;
    } else {
        m.lu = reinterpret_cast<double*>(p);
        p += 8 * LSD_LDAB * (size_t)N;
    }
    m.y = reinterpret_cast<double*>(p);
    p += 56 * (size_t)N;
    m.g = reinterpret_cast<double*>(p);
    p += 16 * (size_t)N;
    m.rd = reinterpret_cast<double*>(p);
    p += 8 * (size_t)N;
    if (mix) {
        m.lu32 = reinterpret_cast<float*>(p);
        p += 40 * (size_t)N;
        m.rd32 = reinterpret_cast<float*>(p);
        p += 4 * (size_t)N;
    }
    m.piv = reinterpret_cast<int*>(p);
    return m;
}

// Block-cooperative copies of one (kind, l) operator into LDS.  Callers
// synchronise before the data are read and before they are replaced.
template <bool MIX>
__device__ void stage_lu(const SolveArgs& a, const Smem& m, int kind, int l) {
    const int N = a.N;
    const size_t off = ((size_t)kind * a.L + l) * N;
    if constexpr (MIX) {
        for (int i = threadIdx.x; i < N * LSD_LDAB; i += 64) m.lu32[i] = a.lu32[off * LSD_LDAB + i];
        for (int i = threadIdx.x; i < N * 7; i += 64) m.xb[i] = a.xb[off * 7 + i];
        for (int i = threadIdx.x; i < N; i += 64) m.rd32[i] = 1.0f / a.lu32[(off + i) * LSD_LDAB + 2 * LSD_KL];
    } else {
        for (int i = threadIdx.x; i < N * LSD_LDAB; i += 64) m.lu[i] = a.lu[off * LSD_LDAB + i];
        for (int i = threadIdx.x; i < N; i += 64) m.rd[i] = 1.0 / a.lu[(off + i) * LSD_LDAB + 2 * LSD_KL];
    }
    for (int i = threadIdx.x; i < N; i += 64) m.piv[i] = a.piv[off + i];
}
__device__ inline void stage_y(const SolveArgs& a, const Smem& m, int yk, int l) {
    const size_t off = ((size_t)yk * a.L + l) * a.N * 7;
    for (int i = threadIdx.x; i < a.N * 7; i += 64) m.y[i] = a.y[off + i];
}
__device__ inline void stage_g(const SolveArgs& a, const Smem& m, int l) {
    const size_t off = (size_t)l * 2 * a.N;
    for (int i = threadIdx.x; i < 2 * a.N; i += 64) m.g[i] = a.gre[off + i];
}


}  // namespace lsd
