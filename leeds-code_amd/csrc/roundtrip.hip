// Kernels of the synthesis -> analysis round trip (lsd_sht_roundtrip): the
// precision conversions of coll fields and the equatorial fold of Fourier
// rows.  At the Fourier level the phi FFTs of a round trip cancel (an
// unnormalised c2r followed by an r2c scaled by 1/Ph, tra_s2p / tra_p2s,
// modules/transform.fftw3.F90:280-466), so the fold of the synthesised
// Fourier data is exactly what the forward pass hands to the analysis.
#include "lsd_internal.h"

namespace lsd {

// coll double2 [n] <-> the transform precision
template <typename S, typename D>
__global__ void k_cast2(size_t n, const S* __restrict__ a, D* __restrict__ b) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = D{(decltype(D().x))a[i].x, (decltype(D().x))a[i].y};
}

// Fourier [n][j][MP] (Th rows) -> folded [n][m][p][j], j < Th2:
// p = 0: F[j] + F[Th-1-j], p = 1: F[j] - F[Th-1-j]
template <typename V>
__global__ void k_fold(int N, int Th, int Th2, int M, int MP, const V* __restrict__ F, V* __restrict__ G) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t tot = (size_t)N * M * 2 * Th2;
    if (i >= tot) return;
    const int j = (int)(i % Th2);
    size_t r = i / Th2;
    const int p = (int)(r % 2);
    r /= 2;
    const int m = (int)(r % M), n = (int)(r / M);
    const V a = F[((size_t)n * Th + j) * MP + m];
    const V b = F[((size_t)n * Th + (Th - 1 - j)) * MP + m];
    G[i] = p ? V{a.x - b.x, a.y - b.y} : V{a.x + b.x, a.y + b.y};
}

int launch_cast_coll(Ctx& c, const void* src, void* dst, bool to_f32, bool from_f32) {
    const size_t n = c.coll_elems();
    const dim3 g((unsigned)((n + 255) / 256)), b(256);
    if (!from_f32 && to_f32)
        hipLaunchKernelGGL((k_cast2<double2, float2>), g, b, 0, c.stream, n, (const double2*)src, (float2*)dst);
    else if (from_f32 && !to_f32)
        hipLaunchKernelGGL((k_cast2<float2, double2>), g, b, 0, c.stream, n, (const float2*)src, (double2*)dst);
    else
        hipLaunchKernelGGL((k_cast2<double2, double2>), g, b, 0, c.stream, n, (const double2*)src, (double2*)dst);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

int launch_fold(Ctx& c, const void* fourier, void* folded, bool f32) {
    const size_t tot = (size_t)c.N * c.M * 2 * c.Th2;
    const dim3 g((unsigned)((tot + 255) / 256)), b(256);
    if (f32)
        hipLaunchKernelGGL(k_fold<float2>, g, b, 0, c.stream, c.N, c.Th, c.Th2, c.M, c.MP, (const float2*)fourier,
                           (float2*)folded);
    else
        hipLaunchKernelGGL(k_fold<double2>, g, b, 0, c.stream, c.N, c.Th, c.Th2, c.M, c.MP, (const double2*)fourier,
                           (double2*)folded);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

}  // namespace lsd
