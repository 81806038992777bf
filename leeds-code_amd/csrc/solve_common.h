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
    const double* rot;   // inner-core rotation state (RotSlot), null without b_rot_tstep
    int t10;             // the inner wall is no-slip: T_10(r_i) row of vel_setbc_Tor
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


// Lane view of one real column in LDS.
template <typename T>
struct ColT {
    T* w;  // column base, stride S
    int S, N;
};
typedef ColT<double> Col;

// dgbtrs('N') on one column (timestep.F90:566-588 -> LAPACK dgbtrs with
// KL = 3, KU = 3, LDAB = 10): row interchanges and L^-1 forward, then the
// upper band solve (dtbsv) backward.  ab = LDS factors of this (kind, l),
// entry AB(i,j) (1-based) at ab[(j-1)*10 + i-1]; ip = LDS pivots.
template <typename T>
__device__ void gbtrs_col(const ColT<T>& c, const T* ab, const int* ip) {
    const int N = c.N, S = c.S;
    T* w = c.w;
    T r0 = w[0], r1 = w[S], r2 = w[2 * S], r3 = w[3 * S];
    for (int j = 0; j < N - 1; j++) {
        const int p = ip[j] - 1 - j;  // 0..3 (pivot inside the band)
        const T* a = ab + j * LSD_LDAB + 2 * LSD_KL + 1;
        const T a0 = a[0], a1 = a[1], a2 = a[2];
        T t = r0;
        if (p == 1) { r0 = r1; r1 = t; }
        else if (p == 2) { r0 = r2; r2 = t; }
        else if (p == 3) { r0 = r3; r3 = t; }
        const T yv = r0;
        w[j * S] = yv;
        r0 = r1 - a0 * yv;
        r1 = r2 - a1 * yv;
        r2 = r3 - a2 * yv;
        r3 = j + 4 < N ? w[(j + 4) * S] : T(0);
    }
    // registers beyond row N-1 above only ever held rows that are never stored
    T b[7];
    b[0] = r0;
    for (int d = 1; d < 7; d++) b[d] = w[(N - 1 - d) * S];
    for (int j = N - 1; j >= 0; j--) {
        const T* a = ab + j * LSD_LDAB;
        T u[7];
#pragma unroll
        for (int d = 0; d < 7; d++) u[d] = a[2 * LSD_KL - d];
        const T v = b[0] / u[0];
        w[j * S] = v;
#pragma unroll
        for (int d = 1; d < 7; d++) b[d] = b[d] - v * u[d];
#pragma unroll
        for (int d = 0; d < 6; d++) b[d] = b[d + 1];
        b[6] = j - 7 >= 0 ? w[(j - 7) * S] : T(0);
    }
}

// w <- w + Y.x in place (tim_multY, timestep.F90:640-670): row n sums its
// band in ascending column order; yk = LDS band, Y(n,j) = yk[j*7 + 3+n-j];
// xs is this lane's real part of the old state column, read as a stream.
__device__ inline void multy_col(const Col& c, const double* yk, const double* __restrict__ xs) {
    const int N = c.N, S = c.S;
    double xw[7];
#pragma unroll
    for (int d = 0; d < 7; d++) xw[d] = (d - 3 >= 0 && d - 3 < N) ? xs[2 * (d - 3)] : 0.0;
    for (int n = 0; n < N; n++) {
        double acc = c.w[n * S];
#pragma unroll
        for (int d = 0; d < 7; d++) {
            const int j = n - 3 + d;
            if (j >= 0 && j < N) acc = fma(yk[j * 7 + LSD_KL + n - j], xw[d], acc);
        }
        c.w[n * S] = acc;
#pragma unroll
        for (int d = 0; d < 6; d++) xw[d] = xw[d + 1];
        xw[6] = n + 4 < N ? xs[2 * (n + 4)] : 0.0;
    }
}

}  // namespace lsd
