// Bulk energies and their l- and m-spectra (program/io.F90:1799-1884,
// modules/variables.F90:901-957, 1267-1313): the field in qst form,
// weights w = 4 pi / (2l+1) (times 4 for m /= 0), radial integrals with
// the int r^2 dr weights.  One work-item per harmonic integrates its q, s
// and t columns; the sums over harmonics run on the host in a fixed order.
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

int energy(Ctx& c, int which, double* E, double* El, double* Em) {
    const int H = c.H;
    double* d_eh = nullptr;
    LSD_HIP(hipMallocAsync((void**)&d_eh, sizeof(double) * 3 * H, c.stream));
    const double2* T = c.d_f[which ? LSD_MAG_TOR : LSD_VEL_TOR];
    const double2* P = c.d_f[which ? LSD_MAG_POL : LSD_VEL_POL];
    hipLaunchKernelGGL(k_energy_h, dim3((H + 127) / 128), dim3(128), 0, c.stream, H, c.N, c.L, c.d_hl, c.d_hm, c.d_r,
                       c.d_r1dr, c.d_ll, c.d_intr2dr, T, P, d_eh);
    LSD_HIP(hipGetLastError());
    std::vector<double> eh((size_t)3 * H);
    LSD_HIP(hipMemcpyAsync(eh.data(), d_eh, sizeof(double) * 3 * H, hipMemcpyDeviceToHost, c.stream));
    LSD_HIP(hipStreamSynchronize(c.stream));
    LSD_HIP(hipFree(d_eh));
    const double efac = which ? c.cfg.Pm / c.cfg.E : 0.5;
    double tot = 0.0;
    if (El)
        for (int l = 0; l < c.L; l++) El[l] = 0.0;
    if (Em)
        for (int m = 0; m < c.M; m++) Em[m] = 0.0;
    for (int k = 0; k < 3; k++) {
        double ek = 0.0;
        for (int h = 0; h < H; h++) {
            const double e = eh[(size_t)k * H + h];
            ek += e;
            if (El) El[c.hl[h]] += efac * e;
            if (Em) Em[c.hm[h] / c.Mp] += efac * e;
        }
        tot += ek;
    }
    *E = efac * tot;
    return LSD_OK;
}

}  // namespace lsd
