// Small multilayer perceptron for the learned closure: training and
// inference on the GPU.  Three dense layers (n_in -> W -> W -> n_out) with
// GELU between them; BF16 weights and activations on the matrix cores
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, FP32 master weights
// updated by Adam.
//
// GEMM: C[M x N] = op(A)[M x K] op(B)[K x N] (+ bias, + C), operands read
// from FP32 memory with arbitrary strides (so transposes cost nothing),
// rounded to BF16 while staged through LDS.  Workgroup tile 64 x 64 by 4
// waves (32 x 32 each, 2 x 2 MFMA tiles), K in steps of 32.
// MFMA operand layout (16x16x32 bf16): lane l holds A[i = l%16][k = 8(l/16)
// .. 8(l/16)+7] and B[k = 8(l/16) ..][j = l%16]; the result lane holds
// C[4(l/16) + r][l%16], r = 0..3.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lsd.h"
#include "lsd_internal.h"

namespace lsd {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

struct GemmArgs {
    int M, N, K;
    const float* A;
    long sam, sak;  // A(i,k) = A[i*sam + k*sak]
    const float* B;
    long sbk, sbn;  // B(k,j) = B[k*sbk + j*sbn]
    float* C;
    long ldc;       // C(i,j) = C[i*ldc + j]
    const float* bias;  // [N] or null
    int accumulate;     // C += result
    int act;            // 0 none, 1 GELU (Z keeps the pre-activation)
    float* Z;           // [M][ldc] pre-activation when act = 1
};

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

constexpr int GT = 64;   // tile edge
constexpr int GK = 32;   // K step
constexpr int GKP = 40;  // padded LDS row (bf16) against bank conflicts

__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g) {
    __shared__ __bf16 As[GT][GKP];
    __shared__ __bf16 Bs[GT][GKP];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int l16 = lane & 15, q = lane >> 4;
    const int i0 = blockIdx.y * GT, j0 = blockIdx.x * GT;
    const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
    v4f acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += GK) {
        // stage: 64 x 32 of A and of B^T, 8 elements per thread each
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int idx = t + 256 * e;  // 0 .. 2047
            const int r = idx >> 5, kk = idx & 31;
            const int i = i0 + r, k = k0 + kk, j = j0 + r;
            const float av = (i < g.M && k < g.K) ? g.A[(long)i * g.sam + (long)k * g.sak] : 0.f;
            const float bv = (j < g.N && k < g.K) ? g.B[(long)k * g.sbk + (long)j * g.sbn] : 0.f;
            As[r][kk] = (__bf16)av;
            Bs[r][kk] = (__bf16)bv;
        }
        __syncthreads();
        bf16x8 af[2], bfr[2];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int v = 0; v < 8; v++) af[a][v] = As[wi + 16 * a + l16][8 * q + v];
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 8; v++) bfr[b][v] = Bs[wj + 16 * b + l16][8 * q + v];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = i0 + wi + 16 * a + 4 * q + r;
                const int j = j0 + wj + 16 * b + l16;
                if (i >= g.M || j >= g.N) continue;
                float v = acc[a][b][r];
                if (g.bias) v += g.bias[j];
                float* c = g.C + (long)i * g.ldc + j;
                if (g.accumulate) v += *c;
                if (g.act == 1) {
                    g.Z[(long)i * g.ldc + j] = v;
                    v = gelu(v);
                }
                *c = v;
            }
}

static int gemm(hipStream_t st, GemmArgs g) {
    dim3 grid((g.N + GT - 1) / GT, (g.M + GT - 1) / GT);
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, g);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

// dZ = dH * gelu'(Z), elementwise
__global__ void k_gelu_back(long n, const float* __restrict__ Z, const float* __restrict__ dH, float* __restrict__ dZ) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dZ[i] = dH[i] * gelu_grad(Z[i]);
}

// Column sums of dY [rows x cols] into db (bias gradients).
__global__ void k_colsum(int rows, int cols, const float* __restrict__ dY, float* __restrict__ db) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    float s = 0.f;
    for (int i = 0; i < rows; i++) s += dY[(long)i * cols + j];
    db[j] = s;
}

// dY = 2 (Y - T) / (rows * cols); per-block partial sums of squared error.
__global__ void k_mse(long n, const float* __restrict__ Y, const float* __restrict__ T, float scale,
                      float* __restrict__ dY, float* __restrict__ part) {
    __shared__ float red[256];
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float e = 0.f;
    if (i < n) {
        const float d = Y[i] - T[i];
        dY[i] = 2.f * d * scale;
        e = d * d;
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam on the FP32 master parameters.
__global__ void k_adam(long n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    p[i] = p[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
}

}  // namespace lsd

using namespace lsd;

// Parameter block (FP32, contiguous): W1[W][in] b1[W] W2[W][W] b2[W] W3[out][W] b3[out]
struct lsd_mlp {
    int nin, width, nout;
    long np;
    hipStream_t stream = nullptr;
    float* p = nullptr;   // parameters
    float* g = nullptr;   // gradients
    float* m1 = nullptr;  // Adam moments
    float* m2 = nullptr;
    long step = 0;
    // activations of the last forward pass (capacity rows)
    long cap = 0;
    float *X = nullptr, *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *Y = nullptr, *T = nullptr;
    float *dY = nullptr, *dH = nullptr, *dZ = nullptr, *part = nullptr;
    float* h_part = nullptr;
    long o_w1() const { return 0; }
    long o_b1() const { return (long)width * nin; }
    long o_w2() const { return o_b1() + width; }
    long o_b2() const { return o_w2() + (long)width * width; }
    long o_w3() const { return o_b2() + width; }
    long o_b3() const { return o_w3() + (long)nout * width; }
};

static void mlp_free_rows(lsd_mlp* m) {
    float* bufs[] = {m->X, m->Z1, m->H1, m->Z2, m->H2, m->Y, m->T, m->dY, m->dH, m->dZ, m->part};
    for (float* b : bufs)
        if (b) (void)hipFree(b);
    if (m->h_part) (void)hipHostFree(m->h_part);
    m->X = m->Z1 = m->H1 = m->Z2 = m->H2 = m->Y = m->T = m->dY = m->dH = m->dZ = m->part = nullptr;
    m->h_part = nullptr;
    m->cap = 0;
}

static int mlp_rows(lsd_mlp* m, long rows) {
    if (rows <= m->cap) return LSD_OK;
    mlp_free_rows(m);
    const long W = m->width;
    LSD_HIP(hipMalloc((void**)&m->X, sizeof(float) * rows * m->nin));
    LSD_HIP(hipMalloc((void**)&m->Z1, sizeof(float) * rows * W));
    LSD_HIP(hipMalloc((void**)&m->H1, sizeof(float) * rows * W));
    LSD_HIP(hipMalloc((void**)&m->Z2, sizeof(float) * rows * W));
    LSD_HIP(hipMalloc((void**)&m->H2, sizeof(float) * rows * W));
    LSD_HIP(hipMalloc((void**)&m->Y, sizeof(float) * rows * m->nout));
    LSD_HIP(hipMalloc((void**)&m->T, sizeof(float) * rows * m->nout));
    LSD_HIP(hipMalloc((void**)&m->dY, sizeof(float) * rows * m->nout));
    LSD_HIP(hipMalloc((void**)&m->dH, sizeof(float) * rows * W));
    LSD_HIP(hipMalloc((void**)&m->dZ, sizeof(float) * rows * W));
    const long nb = (rows * m->nout + 255) / 256;
    LSD_HIP(hipMalloc((void**)&m->part, sizeof(float) * nb));
    LSD_HIP(hipHostMalloc((void**)&m->h_part, sizeof(float) * nb, hipHostMallocDefault));
    m->cap = rows;
    return LSD_OK;
}

// forward over rows already in m->X
static int mlp_forward_dev(lsd_mlp* m, long rows) {
    const int W = m->width;
    GemmArgs g{};
    g.M = (int)rows;
    // layer 1: Z1 = X W1^T + b1, H1 = gelu(Z1)
    g.N = W; g.K = m->nin;
    g.A = m->X; g.sam = m->nin; g.sak = 1;
    g.B = m->p + m->o_w1(); g.sbk = 1; g.sbn = m->nin;
    g.C = m->H1; g.ldc = W; g.bias = m->p + m->o_b1(); g.act = 1; g.Z = m->Z1;
    int rc = gemm(m->stream, g);
    if (rc) return rc;
    // layer 2
    g.N = W; g.K = W;
    g.A = m->H1; g.sam = W; g.sak = 1;
    g.B = m->p + m->o_w2(); g.sbk = 1; g.sbn = W;
    g.C = m->H2; g.ldc = W; g.bias = m->p + m->o_b2(); g.act = 1; g.Z = m->Z2;
    if ((rc = gemm(m->stream, g))) return rc;
    // layer 3 (linear)
    g.N = m->nout; g.K = W;
    g.A = m->H2; g.sam = W; g.sak = 1;
    g.B = m->p + m->o_w3(); g.sbk = 1; g.sbn = W;
    g.C = m->Y; g.ldc = m->nout; g.bias = m->p + m->o_b3(); g.act = 0; g.Z = nullptr;
    return gemm(m->stream, g);
}

// loss and gradients for targets in m->T; returns the mean squared error
static int mlp_backward_dev(lsd_mlp* m, long rows, float* loss) {
    const int W = m->width;
    const long ny = rows * m->nout;
    const long nb = (ny + 255) / 256;
    hipLaunchKernelGGL(k_mse, dim3((unsigned)nb), dim3(256), 0, m->stream, ny, m->Y, m->T, 1.0f / (float)ny, m->dY,
                       m->part);
    GemmArgs g{};
    int rc;
    // dW3 = dY^T H2  [nout x W],  db3 = colsum dY
    g.M = m->nout; g.N = W; g.K = (int)rows;
    g.A = m->dY; g.sam = 1; g.sak = m->nout;
    g.B = m->H2; g.sbk = W; g.sbn = 1;
    g.C = m->g + m->o_w3(); g.ldc = W;
    if ((rc = gemm(m->stream, g))) return rc;
    hipLaunchKernelGGL(k_colsum, dim3((m->nout + 255) / 256), dim3(256), 0, m->stream, (int)rows, m->nout, m->dY,
                       m->g + m->o_b3());
    // dH2 = dY W3  [rows x W];  dZ2 = dH2 gelu'(Z2)
    g = GemmArgs{};
    g.M = (int)rows; g.N = W; g.K = m->nout;
    g.A = m->dY; g.sam = m->nout; g.sak = 1;
    g.B = m->p + m->o_w3(); g.sbk = W; g.sbn = 1;
    g.C = m->dH; g.ldc = W;
    if ((rc = gemm(m->stream, g))) return rc;
    const long nw = rows * W;
    hipLaunchKernelGGL(k_gelu_back, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, m->stream, nw, m->Z2, m->dH,
                       m->dZ);
    // dW2 = dZ2^T H1, db2
    g = GemmArgs{};
    g.M = W; g.N = W; g.K = (int)rows;
    g.A = m->dZ; g.sam = 1; g.sak = W;
    g.B = m->H1; g.sbk = W; g.sbn = 1;
    g.C = m->g + m->o_w2(); g.ldc = W;
    if ((rc = gemm(m->stream, g))) return rc;
    hipLaunchKernelGGL(k_colsum, dim3((W + 255) / 256), dim3(256), 0, m->stream, (int)rows, W, m->dZ,
                       m->g + m->o_b2());
    // dH1 = dZ2 W2, dZ1 = dH1 gelu'(Z1)
    g = GemmArgs{};
    g.M = (int)rows; g.N = W; g.K = W;
    g.A = m->dZ; g.sam = W; g.sak = 1;
    g.B = m->p + m->o_w2(); g.sbk = W; g.sbn = 1;
    g.C = m->dH; g.ldc = W;
    if ((rc = gemm(m->stream, g))) return rc;
    hipLaunchKernelGGL(k_gelu_back, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, m->stream, nw, m->Z1, m->dH,
                       m->dZ);
    // dW1 = dZ1^T X, db1
    g = GemmArgs{};
    g.M = W; g.N = m->nin; g.K = (int)rows;
    g.A = m->dZ; g.sam = 1; g.sak = W;
    g.B = m->X; g.sbk = m->nin; g.sbn = 1;
    g.C = m->g + m->o_w1(); g.ldc = m->nin;
    if ((rc = gemm(m->stream, g))) return rc;
    hipLaunchKernelGGL(k_colsum, dim3((W + 255) / 256), dim3(256), 0, m->stream, (int)rows, W, m->dZ,
                       m->g + m->o_b1());
    LSD_HIP(hipGetLastError());
    if (loss) {
        LSD_HIP(hipMemcpyAsync(m->h_part, m->part, sizeof(float) * nb, hipMemcpyDeviceToHost, m->stream));
        LSD_HIP(hipStreamSynchronize(m->stream));
        double s = 0.0;
        for (long b = 0; b < nb; b++) s += m->h_part[b];
        *loss = (float)(s / (double)ny);
    }
    return LSD_OK;
}

extern "C" {

int lsd_mlp_create(int n_in, int width, int n_out, unsigned long long seed, lsd_mlp** out) {
    if (!out || n_in < 1 || width < 1 || n_out < 1) {
        lsd::set_error("lsd_mlp_create: bad sizes");
        return LSD_E_ARG;
    }
    *out = nullptr;
    lsd_mlp* m = new lsd_mlp();
    m->nin = n_in;
    m->width = width;
    m->nout = n_out;
    m->np = m->o_b3() + n_out;
    auto fail = [&](void) {
        if (m->p) (void)hipFree(m->p);
        if (m->g) (void)hipFree(m->g);
        if (m->m1) (void)hipFree(m->m1);
        if (m->m2) (void)hipFree(m->m2);
        if (m->stream) (void)hipStreamDestroy(m->stream);
        delete m;
        return LSD_E_HIP;
    };
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void**)&m->p, sizeof(float) * m->np) != hipSuccess ||
        hipMalloc((void**)&m->g, sizeof(float) * m->np) != hipSuccess ||
        hipMalloc((void**)&m->m1, sizeof(float) * m->np) != hipSuccess ||
        hipMalloc((void**)&m->m2, sizeof(float) * m->np) != hipSuccess) {
        lsd::set_error("lsd_mlp_create: allocation failed");
        return fail();
    }
    // He-style initial weights from a 64-bit LCG (host), zero biases
    std::vector<float> h(m->np, 0.0f);
    unsigned long long s = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    auto uni = [&]() {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)(s >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](long off, long n, int fan_in) {
        const double a = std::sqrt(6.0 / fan_in);
        for (long i = 0; i < n; i++) h[off + i] = (float)((2.0 * uni() - 1.0) * a);
    };
    fill(m->o_w1(), (long)width * n_in, n_in);
    fill(m->o_w2(), (long)width * width, width);
    fill(m->o_w3(), (long)n_out * width, width);
    if (hipMemcpy(m->p, h.data(), sizeof(float) * m->np, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->m1, 0, sizeof(float) * m->np) != hipSuccess ||
        hipMemset(m->m2, 0, sizeof(float) * m->np) != hipSuccess) {
        lsd::set_error("lsd_mlp_create: initialisation failed");
        return fail();
    }
    *out = m;
    return LSD_OK;
}

void lsd_mlp_destroy(lsd_mlp* m) {
    if (!m) return;
    mlp_free_rows(m);
    (void)hipFree(m->p);
    (void)hipFree(m->g);
    (void)hipFree(m->m1);
    (void)hipFree(m->m2);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

long lsd_mlp_num_params(const lsd_mlp* m) { return m ? m->np : 0; }

int lsd_mlp_get_params(lsd_mlp* m, float* out) {
    if (!m || !out) { lsd::set_error("lsd_mlp_get_params: null"); return LSD_E_ARG; }
    LSD_HIP(hipMemcpy(out, m->p, sizeof(float) * m->np, hipMemcpyDeviceToHost));
    return LSD_OK;
}

int lsd_mlp_set_params(lsd_mlp* m, const float* in) {
    if (!m || !in) { lsd::set_error("lsd_mlp_set_params: null"); return LSD_E_ARG; }
    LSD_HIP(hipMemcpy(m->p, in, sizeof(float) * m->np, hipMemcpyHostToDevice));
    return LSD_OK;
}

int lsd_mlp_forward(lsd_mlp* m, long rows, const float* x, float* y) {
    if (!m || !x || !y || rows < 1) { lsd::set_error("lsd_mlp_forward: bad arguments"); return LSD_E_ARG; }
    int rc = mlp_rows(m, rows);
    if (rc) return rc;
    LSD_HIP(hipMemcpyAsync(m->X, x, sizeof(float) * rows * m->nin, hipMemcpyHostToDevice, m->stream));
    if ((rc = mlp_forward_dev(m, rows))) return rc;
    LSD_HIP(hipMemcpyAsync(y, m->Y, sizeof(float) * rows * m->nout, hipMemcpyDeviceToHost, m->stream));
    LSD_HIP(hipStreamSynchronize(m->stream));
    return LSD_OK;
}

int lsd_mlp_grad(lsd_mlp* m, long rows, const float* x, const float* target, float* loss, float* grad) {
    if (!m || !x || !target || rows < 1) { lsd::set_error("lsd_mlp_grad: bad arguments"); return LSD_E_ARG; }
    int rc = mlp_rows(m, rows);
    if (rc) return rc;
    LSD_HIP(hipMemcpyAsync(m->X, x, sizeof(float) * rows * m->nin, hipMemcpyHostToDevice, m->stream));
    LSD_HIP(hipMemcpyAsync(m->T, target, sizeof(float) * rows * m->nout, hipMemcpyHostToDevice, m->stream));
    if ((rc = mlp_forward_dev(m, rows))) return rc;
    if ((rc = mlp_backward_dev(m, rows, loss))) return rc;
    if (grad) LSD_HIP(hipMemcpyAsync(grad, m->g, sizeof(float) * m->np, hipMemcpyDeviceToHost, m->stream));
    LSD_HIP(hipStreamSynchronize(m->stream));
    return LSD_OK;
}

// Adam steps on minibatches drawn in order from (x, target) [rows x n],
// batch rows each (the last batch of a pass may be shorter); the data are
// copied to the device once per call.  loss[steps] receives the minibatch
// losses (may be null).
int lsd_mlp_train(lsd_mlp* m, long rows, const float* x, const float* target, long batch, int steps, float lr,
                  float* loss) {
    if (!m || !x || !target || rows < 1 || batch < 1 || steps < 0) {
        lsd::set_error("lsd_mlp_train: bad arguments");
        return LSD_E_ARG;
    }
    if (batch > rows) batch = rows;
    float *dx = nullptr, *dt = nullptr;
    LSD_HIP(hipMalloc((void**)&dx, sizeof(float) * rows * m->nin));
    LSD_HIP(hipMalloc((void**)&dt, sizeof(float) * rows * m->nout));
    int rc = mlp_rows(m, batch);
    if (!rc) {
        LSD_HIP(hipMemcpyAsync(dx, x, sizeof(float) * rows * m->nin, hipMemcpyHostToDevice, m->stream));
        LSD_HIP(hipMemcpyAsync(dt, target, sizeof(float) * rows * m->nout, hipMemcpyHostToDevice, m->stream));
        long pos = 0;
        for (int s = 0; s < steps && !rc; s++) {
            long nrow = std::min(batch, rows - pos);
            LSD_HIP(hipMemcpyAsync(m->X, dx + pos * m->nin, sizeof(float) * nrow * m->nin, hipMemcpyDeviceToDevice,
                                   m->stream));
            LSD_HIP(hipMemcpyAsync(m->T, dt + pos * m->nout, sizeof(float) * nrow * m->nout,
                                   hipMemcpyDeviceToDevice, m->stream));
            pos = (pos + nrow) % rows;
            if ((rc = mlp_forward_dev(m, nrow))) break;
            float l = 0.f;
            if ((rc = mlp_backward_dev(m, nrow, loss ? &l : nullptr))) break;
            if (loss) loss[s] = l;
            m->step++;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.f - powf(b1, (float)m->step), c2 = 1.f - powf(b2, (float)m->step);
            hipLaunchKernelGGL(k_adam, dim3((unsigned)((m->np + 255) / 256)), dim3(256), 0, m->stream, m->np, m->p,
                               m->g, m->m1, m->m2, lr, b1, b2, 1e-8f, c1, c2);
            LSD_HIP(hipGetLastError());
        }
        LSD_HIP(hipStreamSynchronize(m->stream));
    }
    (void)hipFree(dx);
    (void)hipFree(dt);
    return rc;
}

}  // extern "C"
