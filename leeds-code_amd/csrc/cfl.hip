// Adaptive timestep: the CFL limit non_maxtstep (modules/nonlinear.F90:545-615)
// and the step choice tim_new_tstep (modules/timestep.F90:837-870,
// i_time_step_control = 2), taken at the start of every step when
// d_timestep < 0 (program/main.F90:124-135).
//
// Per shell the limit needs two maxima over the shell's physical points:
//   radial      Br^2 / sqrt(Br^2 (E/Pm) + d_r) + |ur|,       d_r = ((E + E/Pm) / (2 dr))^2
//   horizontal  Bh^2 / sqrt(Bh^2 E / Pm + d_h) + |uh|,       d_h = (E / (2 dr_h))^2
// with dr the smaller neighbour spacing and dr_h = r / sqrt(L (L-1)).  The
// fused physical-space pass forms them for its row pair while the fields are
// in registers and writes one partial pair per workgroup; the unfused path
// (phase API, grids the fused pass does not take) reads the physical fields
// of the last transform.  A single workgroup then folds the partials into
// tim_cfl_dt on the device.
#include <algorithm>

#include "lsd_internal.h"

namespace lsd {

// d_courant = 0.8 is a default-real literal in the reference
// (Scripts/parameters.F90_template:238): 0.8 rounded to single precision.
const double COURANT = (double)0.8f;
// d_corr_upper, d_corr_lower (Scripts/parameters.F90_template:242-243)
const double CORR_UPPER = 0.020;
const double CORR_LOWER = 0.00200;

// Per-shell constants (d_r, d_h) of the two maxima, [N][2].
__global__ void k_cfl_consts(int N, int L, double E, double Pm, const double* __restrict__ r1,
                             double* __restrict__ cd) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double dr;
    if (n == 0) dr = r1[1] - r1[0];
    else if (n == N - 1) dr = r1[N - 1] - r1[N - 2];
    else dr = fmin(r1[n] - r1[n - 1], r1[n + 1] - r1[n]);
    double d = (E + E / Pm) / (2.0 * dr);
    cd[2 * n] = d * d;
    const double drh = r1[n] / sqrt((double)(L * (L - 1)));
    d = E / (2.0 * drh);
    cd[2 * n + 1] = d * d;
}

// Both maxima of one shell from the physical fields of the last transform
// (reference phys layout, FP64): one workgroup per shell.
__global__ void __launch_bounds__(256) k_cfl_phys(int Th, int Ph, double E, double Pm, int mag,
                                                  const double* __restrict__ ur, const double* __restrict__ ut,
                                                  const double* __restrict__ up, const double* __restrict__ br,
                                                  const double* __restrict__ bt, const double* __restrict__ bp,
                                                  const double* __restrict__ cd, double* __restrict__ part) {
    __shared__ double red[2][4];
    const int n = blockIdx.x;
    const size_t plane = (size_t)Th * Ph, off = (size_t)n * plane;
    const double d1 = cd[2 * n], d2 = cd[2 * n + 1], ep = E / Pm;
    double m1 = -1e300, m2 = -1e300;
    for (size_t i = threadIdx.x; i < plane; i += 256) {
        const double b0 = mag ? br[off + i] : 0.0, b1 = mag ? bt[off + i] : 0.0, b2 = mag ? bp[off + i] : 0.0;
        const double p = b0 * b0;
        m1 = fmax(m1, p / sqrt(p * ep + d1) + fabs(ur[off + i]));
        const double q = b1 * b1 + b2 * b2;
        const double a = ut[off + i], b = up[off + i];
        m2 = fmax(m2, q / sqrt(q * E / Pm + d2) + sqrt(a * a + b * b));
    }
    for (int o = 32; o > 0; o >>= 1) {
        m1 = fmax(m1, __shfl_xor(m1, o));
        m2 = fmax(m2, __shfl_xor(m2, o));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = m1;
        red[1][w] = m2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; k++) {
            m1 = fmax(m1, red[0][k]);
            m2 = fmax(m2, red[1][k]);
        }
        part[2 * n] = m1;
        part[2 * n + 1] = m2;
    }
}

// tim_cfl_dt from the partial maxima [N][parts][2]: per shell dr / max for
// each nonzero maximum, the least over shells, capped by E/Pm and sqrt(E).
__global__ void __launch_bounds__(256) k_cfl_final(int N, int parts, int L, double E, double Pm,
                                                   const double* __restrict__ r1,
                                                   const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double red[256];
    double best = 1e99;
    for (int n = threadIdx.x; n < N; n += 256) {
        double m1 = -1e300, m2 = -1e300;
        for (int q = 0; q < parts; q++) {
            m1 = fmax(m1, part[((size_t)n * parts + q) * 2]);
            m2 = fmax(m2, part[((size_t)n * parts + q) * 2 + 1]);
        }
        double dr;
        if (n == 0) dr = r1[1] - r1[0];
        else if (n == N - 1) dr = r1[N - 1] - r1[N - 2];
        else dr = fmin(r1[n] - r1[n - 1], r1[n + 1] - r1[n]);
        if (m1 != 0.0) best = fmin(best, dr / m1);
        const double drh = r1[n] / sqrt((double)(L * (L - 1)));
        if (m2 != 0.0) best = fmin(best, drh / m2);
    }
    red[threadIdx.x] = best;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double c = fmin(red[0], E / Pm);
        out[0] = fmin(c, sqrt(E));
    }
}

int cfl_setup(Ctx& c) {
    hipLaunchKernelGGL(k_cfl_consts, dim3((c.N + 63) / 64), dim3(64), 0, c.stream, c.N, c.L, c.cfg.E, c.cfg.Pm,
                       c.d_r + 4 * (size_t)c.N, c.d_cfl_d);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

int launch_cfl_phys(Ctx& c) {
    const lsd_config& k = c.cfg;
    const size_t pe = c.phys_elems();
    hipLaunchKernelGGL(k_cfl_phys, dim3(c.N), dim3(256), 0, c.stream, c.Th, c.Ph, k.E, k.Pm, k.mag,
                       c.d_phys + LSD_P_VEL_R * pe, c.d_phys + LSD_P_VEL_T * pe, c.d_phys + LSD_P_VEL_P * pe,
                       c.d_phys + LSD_P_MAG_R * pe, c.d_phys + LSD_P_MAG_T * pe, c.d_phys + LSD_P_MAG_P * pe,
                       c.d_cfl_d, c.d_cfl_part);
    LSD_HIP(hipGetLastError());
    c.cfl_parts = 1;
    return LSD_OK;
}

// Folds the partials of the last evaluation into tim_cfl_dt and queues its
// copy to the pinned host word (read after the stream is synchronised).
int launch_cfl_final(Ctx& c) {
    hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(256), 0, c.stream, c.N, c.cfl_parts, c.L, c.cfg.E, c.cfg.Pm,
                       c.d_r + 4 * (size_t)c.N, c.d_cfl_part, c.d_cfl);
    LSD_HIP(hipGetLastError());
    LSD_HIP(hipMemcpyAsync(c.h_cfl, c.d_cfl, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    return LSD_OK;
}

// tim_new_tstep: returns true if dt changed (tim_new_dt), in which case
// c.dt holds the new step.  Control 2 takes the least of the corrector's
// estimate (at most 2 dt / d_courant) and the CFL limit, times d_courant;
// control 1 halves dt when tim_corr_dt > d_corr_upper and grows it by
// 1.273974 when tim_corr_dt < d_corr_lower, without the CFL limit.  The
// first step of a run starts from min(dt / 1e3, 1e-10) (control 2) or
// min(dt / 1e3, 1e-8) (control 1).
bool new_tstep(Ctx& c) {
    double dt = c.dt;
    if (c.cfg.step_control == 1) {
        if (c.corr_dt > CORR_UPPER) dt = 0.5 * dt;
        if (c.corr_dt < CORR_LOWER) dt = 1.273974 * dt;
        if (c.step == 0) dt = std::min(dt / 1e3, 1e-8);
    } else {
        if (c.corr_dt > 0.0) dt = std::min(c.corr_dt, dt * 2.0 / COURANT);
        if (c.cfl_dt > 0.0) dt = std::min(c.cfl_dt, dt);
        if (c.step == 0) dt = std::min(dt / 1e3, 1e-10);
        dt = dt * COURANT;
    }
    const bool changed = dt < c.dt * 0.97 || dt > c.dt * 1.03;
    if (changed) c.dt = dt;
    return changed;
}

}  // namespace lsd
