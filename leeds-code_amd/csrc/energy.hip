// Bulk energies and their l- and m-spectra (program/io.F90:1799-1884,
// modules/variables.F90:901-957, 1267-1313): the field in qst form,
// weights w = 4 pi / (2l+1) (times 4 for m /= 0), radial integrals with
// the int r^2 dr weights.  One work-item per harmonic integrates its q, s
// and t columns; the sums over harmonics run on the host in a fixed order.
#include <mutex>

#include "lsd_internal.h"

namespace lsd {

__global__ void k_energy_h(int H, int N, int L, const int* __restrict__ hl, const int* __restrict__ hm,
                           const double* __restrict__ r, const double* __restrict__ r1dr,
                           const double* __restrict__ ll, const double* __restrict__ intr2dr,
                           const double2* __restrict__ T, const double2* __restrict__ P, double* __restrict__ eh) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= H) return;
    const int l = hl[h];
    const double* rinv = r + 2 * N;
    const double ll1 = ll[l], rt = ll[2 * L + l];
    double w = 4.0 * 3.1415926535897931 / (double)(2 * l + 1);
    if (hm[h] != 0) w = 4.0 * w;
    const double2* Pc = P + (size_t)h * N;
    const double2* Tc = T + (size_t)h * N;
    double eq = 0.0, es = 0.0, et = 0.0;
    for (int n = 0; n < N; n++) {
        const double a = ll1 * rinv[n];
        const double qr = Pc[n].x * a, qi = Pc[n].y * a;
        // s = sqrt(l(l+1)) (1/r + d/dr) P, the 7-band stencil of r1dr
        const int lo = n - LSD_KL > 0 ? n - LSD_KL : 0;
        const int hi = n + LSD_KL < N - 1 ? n + LSD_KL : N - 1;
        double sr = 0.0, si = 0.0;
        for (int j = lo; j <= hi; j++) {
            const double b = r1dr[j * 7 + (LSD_KL + n - j)];
            sr = fma(b, Pc[j].x, sr);
            si = fma(b, Pc[j].y, si);
        }
        sr = sr * rt;
        si = si * rt;
        const double tr = -Tc[n].x * rt, ti = -Tc[n].y * rt;
        eq += w * (qr * qr + qi * qi) * intr2dr[n];
        es += w * (sr * sr + si * si) * intr2dr[n];
        et += w * (tr * tr + ti * ti) * intr2dr[n];
    }
    eh[h] = eq;
    eh[H + h] = es;
    eh[2 * H + h] = et;
}

// Sums of the per-harmonic parts in a fixed order: work-item 0 forms the
// total (the q, s and t parts each summed over h, then added), work-item
// 1 + l the degree-l part and 1 + L + m the order-m part, each over the
// parts and then h in harmonic order; all times efac.  out = [E, El[L],
// Em[M]].
__global__ void k_energy_sum(int H, int L, int M, int Mp, const int* __restrict__ hl, const int* __restrict__ hm,
                             double efac, const double* __restrict__ eh, double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > L + M) return;
    double acc = 0.0;
    if (t == 0) {
        for (int k = 0; k < 3; k++) {
            double ek = 0.0;
            for (int h = 0; h < H; h++) ek += eh[(size_t)k * H + h];
            acc += ek;
        }
        acc = efac * acc;
    } else if (t <= L) {
        const int l = t - 1;
        for (int k = 0; k < 3; k++)
            for (int h = 0; h < H; h++)
                if (hl[h] == l) acc += efac * eh[(size_t)k * H + h];
    } else {
        const int m = t - 1 - L;
        for (int k = 0; k < 3; k++)
            for (int h = 0; h < H; h++)
                if (hm[h] / Mp == m) acc += efac * eh[(size_t)k * H + h];
    }
    out[t] = acc;
}

size_t energy_scratch_doubles(const Ctx& c) { return (size_t)3 * c.H + 1 + c.L + c.M; }

int energy_run(Ctx& c, int which, double* scratch, double* E, double* El, double* Em) {
    const int H = c.H;
    double* d_eh = scratch;
    double* d_out = scratch + (size_t)3 * H;
    const double2* T = c.d_f[which ? LSD_MAG_TOR : LSD_VEL_TOR];
    const double2* P = c.d_f[which ? LSD_MAG_POL : LSD_VEL_POL];
    const double efac = which ? c.cfg.Pm / c.cfg.E : 0.5;
    hipLaunchKernelGGL(k_energy_h, dim3((H + 127) / 128), dim3(128), 0, c.stream, H, c.N, c.L, c.d_hl, c.d_hm, c.d_r,
                       c.d_r1dr, c.d_ll, c.d_intr2dr, T, P, d_eh);
    LSD_HIP(hipGetLastError());
    const int nt = 1 + c.L + c.M;
    hipLaunchKernelGGL(k_energy_sum, dim3((nt + 255) / 256), dim3(256), 0, c.stream, H, c.L, c.M, c.Mp, c.d_hl,
                       c.d_hm, efac, d_eh, d_out);
    LSD_HIP(hipGetLastError());
    LSD_HIP(hipMemcpyAsync(E, d_out, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (El) LSD_HIP(hipMemcpyAsync(El, d_out + 1, sizeof(double) * c.L, hipMemcpyDeviceToHost, c.stream));
    if (Em) LSD_HIP(hipMemcpyAsync(Em, d_out + 1 + c.L, sizeof(double) * c.M, hipMemcpyDeviceToHost, c.stream));
    LSD_HIP(hipStreamSynchronize(c.stream));
    return LSD_OK;
}

// Scratch of lsd_energy_noalloc: the per-harmonic parts [3][H] and the
// sums [1 + L + M], one copy per GPU for the life of the process.
constexpr int ENERGY_MAX_H = 65536;
constexpr int ENERGY_MAX_LM = 2048;
constexpr size_t ENERGY_SCRATCH = (size_t)3 * ENERGY_MAX_H + 1 + ENERGY_MAX_LM;
static __device__ __align__(256) double scratch[ENERGY_SCRATCH];

constexpr int ENERGY_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[ENERGY_MAX_DEVICES] = {};

// Device address of this device's copy of `scratch`; called once per device
// from lsd_create with that device current.
int energy_noalloc_init(int device) {
    if (device < 0 || device >= ENERGY_MAX_DEVICES) {
        set_error("lsd_create: device index beyond the static scratch table");
        return LSD_E_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return LSD_OK;
    double* ptr = nullptr;
    LSD_HIP(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = ptr;
    return LSD_OK;
}

int energy_noalloc(Ctx& c, int which, double* E, double* El, double* Em) {
    if (c.H > ENERGY_MAX_H || c.L + c.M > ENERGY_MAX_LM) {
        set_error("lsd_energy_noalloc: H = " + std::to_string(c.H) + ", L + M = " + std::to_string(c.L + c.M) +
                  " exceed the static scratch (H <= 65536, L + M <= 2048); use lsd_energy");
        return LSD_E_ARG;
    }
    double* ptr = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        ptr = (c.device >= 0 && c.device < ENERGY_MAX_DEVICES) ? g_scratch_ptr[c.device] : nullptr;
    }
    if (!ptr) {
        set_error("lsd_energy_noalloc: static scratch not set up on this device");
        return LSD_E_ARG;
    }
    return energy_run(c, which, ptr, E, El, Em);
}

int energy(Ctx& c, int which, double* E, double* El, double* Em) {
    double* scratch = nullptr;
    AsyncScratch s(c.stream);
    LSD_HIP(s.alloc(&scratch, sizeof(double) * energy_scratch_doubles(c)));
    return energy_run(c, which, scratch, E, El, Em);
}

}  // namespace lsd
