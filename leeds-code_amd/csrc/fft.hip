// Phi FFTs staged through LDS (modules/transform.fftw3.F90:280-466).
//
// Inverse (tra_s2p): each workgroup takes one (field, shell) and 2*NF theta
// rows; rows (j, j+1) are packed into one complex sequence
// Z = A + iB with Hermitian-extended spectra A, B (zero for m >= M, the
// imaginary parts of the DC terms ignored as FFTW's c2r does), transformed
// with an in-LDS Stockham FFT (radices 8, 4, 2, 3), and the real and
// imaginary parts are the two physical rows.
//
// Forward (tra_p2s + the even/odd fold of tra_rtp2qst, :862-864): rows j and
// Th-1-j (its mirror) are packed, scaled by 1/Ph as the reference does
// before r2c, transformed, separated, and written as
// G0 = X_j + X_mirror, G1 = X_j - X_mirror for m < M.
#include "lsd_internal.h"
#include "fft_lds.h"

namespace lsd {

// fourier [f][n][Th][MP] -> phys [f][n][Ph][Th]; grid (row blocks, N, fields).
__global__ void __launch_bounds__(FFT_THREADS) k_fft_c2r(FftPlan pl, int nf, int N, int Th, int M, int MP,
                                                         const double2* __restrict__ fourier,
                                                         double* __restrict__ phys, const double2* __restrict__ tw) {
    extern __shared__ double2 Z[];
    const int n = pl.n;
    const int f = blockIdx.z, sh = blockIdx.y;
    const int j0 = blockIdx.x * 2 * nf;
    const double2* X = fourier + ((size_t)f * N + sh) * (size_t)Th * MP;
    for (int i = threadIdx.x; i < nf * n; i += FFT_THREADS) Z[i] = make_double2(0.0, 0.0);
    __syncthreads();
    for (int t = threadIdx.x; t < M * nf; t += FFT_THREADS) {
        int k = t / nf, p = t - k * nf;
        int ja = j0 + 2 * p, jb = ja + 1;
        double2 a = ja < Th ? X[(size_t)ja * MP + k] : make_double2(0.0, 0.0);
        double2 b = jb < Th ? X[(size_t)jb * MP + k] : make_double2(0.0, 0.0);
        double2* z = Z + (size_t)p * n;
        if (k == 0) {
            z[0] = make_double2(a.x, b.x);
        } else {
            z[k] = make_double2(a.x - b.y, a.y + b.x);
            z[n - k] = make_double2(a.x + b.y, b.x - a.y);
        }
    }
    __syncthreads();
    fft_lds<double2>(Z, pl, nf, tw, 1.0);
    double* P = phys + ((size_t)f * N + sh) * (size_t)n * Th;
    const int rows = 2 * nf;
    for (int t = threadIdx.x; t < n * rows; t += FFT_THREADS) {
        int k = t / rows, r = t - k * rows;
        int j = j0 + r;
        if (j < Th) {
            double2 z = Z[(size_t)(r >> 1) * n + k];
            P[(size_t)k * Th + j] = (r & 1) ? z.y : z.x;
        }
    }
}

// phys [f][n][Ph][Th] -> folded [f][n][M][2][Th2]; grid (row blocks, N, fields).
__global__ void __launch_bounds__(FFT_THREADS) k_fft_r2c_fold(FftPlan pl, int nf, int N, int Th, int Th2, int M,
                                                              const double* __restrict__ phys,
                                                              double2* __restrict__ folded,
                                                              const double2* __restrict__ tw) {
    extern __shared__ double2 Z[];
    const int n = pl.n;
    const int f = blockIdx.z, sh = blockIdx.y;
    const int j0 = blockIdx.x * nf;
    const double scale = 1.0 / (double)n;
    const double* P = phys + ((size_t)f * N + sh) * (size_t)n * Th;
    for (int t = threadIdx.x; t < n * nf; t += FFT_THREADS) {
        int k = t / nf, p = t - k * nf;
        int j = j0 + p;
        double a = 0.0, b = 0.0;
        if (j < Th2) {
            a = P[(size_t)k * Th + j] * scale;
            b = P[(size_t)k * Th + (Th - 1 - j)] * scale;
        }
        Z[(size_t)p * n + k] = make_double2(a, b);
    }
    __syncthreads();
    fft_lds<double2>(Z, pl, nf, tw, -1.0);
    double2* G = folded + ((size_t)f * N + sh) * (size_t)M * 2 * Th2;
    for (int t = threadIdx.x; t < M * 2 * nf; t += FFT_THREADS) {
        int k = t / (2 * nf), rem = t - k * 2 * nf;
        int par = rem / nf, p = rem - par * nf;
        int j = j0 + p;
        if (j >= Th2) continue;
        double2 zk = Z[(size_t)p * n + k];
        double2 zc = Z[(size_t)p * n + (k ? n - k : 0)];
        zc.y = -zc.y;
        double2 A = make_double2(0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y));
        double2 D = csubv(zk, zc);
        double2 B = make_double2(0.5 * D.y, -0.5 * D.x);
        double2 out = par ? csubv(A, B) : caddv(A, B);
        G[((size_t)k * 2 + par) * Th2 + j] = out;
    }
}

FftPlan make_fft_plan(int n) {
    FftPlan p{};
    p.n = n;
    int rem = n;
    while (rem % 8 == 0) { p.radix[p.nstage++] = 8; rem /= 8; }
    while (rem % 4 == 0) { p.radix[p.nstage++] = 4; rem /= 4; }
    while (rem % 2 == 0) { p.radix[p.nstage++] = 2; rem /= 2; }
    while (rem % 3 == 0) { p.radix[p.nstage++] = 3; rem /= 3; }
    if (rem != 1) p.nstage = -1;
    return p;
}

// Twiddle table exp(+2 pi i k / Ph), k < Ph, built on the device.
__global__ void k_twiddle(int n, double2* __restrict__ tw, float2* __restrict__ tw32) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double s, c;
    sincospi(2.0 * (double)k / (double)n, &s, &c);
    tw[k] = make_double2(c, s);
    tw32[k] = make_float2((float)c, (float)s);
}

int fft_setup(Ctx& c) {
    FftPlan p = make_fft_plan(c.Ph);
    if (p.nstage < 0) {
        set_error("Ph must be a product of 2s and 3s");
        return LSD_E_CONFIG;
    }
    LSD_HIP(hipMalloc((void**)&c.d_tw, sizeof(double2) * c.Ph));
    LSD_HIP(hipMalloc((void**)&c.d_tw32, sizeof(float2) * c.Ph));
    hipLaunchKernelGGL(k_twiddle, dim3((c.Ph + 255) / 256), dim3(256), 0, c.stream, c.Ph, c.d_tw, c.d_tw32);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

static int nfft_per_block(const Ctx& c, int rows_needed) {
    int nf = (FFT_EPT * FFT_THREADS) / c.Ph;
    if (nf < 1) nf = 1;
    if (nf > rows_needed) nf = rows_needed;
    return nf;
}

int launch_fft_c2r(Ctx& c, const double2* fourier, double* phys, int nfield) {
    FftPlan p = make_fft_plan(c.Ph);
    int nf = nfft_per_block(c, (c.Th + 1) / 2);
    dim3 grid((c.Th + 2 * nf - 1) / (2 * nf), c.N, nfield);
    size_t lds = sizeof(double2) * nf * c.Ph;
    ProfScope ps(c, LSD_K_FFT_INV, (double)nfield * c.N * ((double)c.M * c.Th * 16 + (double)c.Ph * c.Th * 8));
    hipLaunchKernelGGL(k_fft_c2r, grid, dim3(FFT_THREADS), lds, c.stream, p, nf, c.N, c.Th, c.M, c.MP, fourier, phys,
                       c.d_tw);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

int launch_fft_r2c_fold(Ctx& c, const double* phys, double2* folded, int nfield) {
    FftPlan p = make_fft_plan(c.Ph);
    int nf = nfft_per_block(c, c.Th2);
    dim3 grid((c.Th2 + nf - 1) / nf, c.N, nfield);
    size_t lds = sizeof(double2) * nf * c.Ph;
    ProfScope ps(c, LSD_K_FFT_FWD, (double)nfield * c.N * ((double)c.Ph * c.Th * 8 + (double)c.M * 2 * c.Th2 * 16));
    hipLaunchKernelGGL(k_fft_r2c_fold, grid, dim3(FFT_THREADS), lds, c.stream, p, nf, c.N, c.Th, c.Th2, c.M, phys,
                       folded, c.d_tw);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

}  // namespace lsd
