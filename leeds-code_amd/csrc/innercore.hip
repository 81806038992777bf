// Conducting inner core (i_Nic > 1) and the inner-core rotation
// (b_rot_tstep): the coupled banded solves of the magnetic field over the
// inner and the outer core, the inner core's induction by its rotation, the
// torques on the inner core and the update of its rotation rate.
//
// The magnetic field of the inner core lives on its own radial points
// (modules/meshs.F90:165-169) as icBT, icBP [H][Nic]; the inner-core
// boundary is the last inner-core point and the first outer-core point.
// For l >= 1 the toroidal and the poloidal field each solve one band of
// NT = Nic + N - 1 rows (modules/timestep.F90:594-634, tim_icinvX): the
// inner-core rows 1..Nic-1 of the right-hand side, then the outer-core rows
// 1..N.  One 64-lane wave per block of harmonics of one degree, one lane
// per (harmonic, Re/Im) column in LDS, as the LDS-column solves of
// solve.hip; the factors of (field, l) are staged once per block.
#include "solve_common.h"

namespace lsd {

struct IcArgs {
    int N, Ni, NT, L, S;   // S: LDS row stride 2 hb + 1
    double cimp;
    const int4* blocks;    // degree, first index into order, count
    const int* order;
    const double* lu;      // [2][L][NT][LDAB]
    const int* piv;        // [2][L][NT]
    const double* yo;      // outer-core Y of BT, BP: [2][L][N][7]
    const double* yi;      // inner-core Y [L][Ni][7]
    double2* x[2];         // outer-core BT, BP [H][N]
    const double2* nn[2];  // their current N
    double2* hist[2];      // their saved N
    double2* xi[2];        // inner-core icBT, icBP [H][Nic]
    const double2* ni[2];
    double2* histi[2];
    double* partial;       // [4][gridDim.x]: BT, icBT, BP, icBP
};

size_t ic_smem_bytes(int N, int Ni, int S) {
    const int NT = Ni + N - 1;
    return 8 * (size_t)NT * S + 8 * (size_t)LSD_LDAB * NT + 56 * (size_t)N + 56 * (size_t)Ni + 4 * (size_t)NT;
}

int ic_block_harmonics(int N, int Ni) {
    int hb = 32;
    while (hb > 0 && ic_smem_bytes(N, Ni, 2 * hb + 1) > 64512) hb--;
    return hb;
}

// Predictor (PRED, mag_predictor, modules/magnetic.F90:370-411): Y B + N on
// each domain with N saved, zero boundary rows (mag_setbc_Tor and
// mag_setbc_Pol with zero boundary arrays on the outer core, tim_zerobc on
// the inner core), the coupled solve.  Corrector (mag_corrector,
// magnetic.F90:422-461): c (N* - N) on each domain with N* saved, the same
// boundary rows, the coupled solve, the correction added with the
// tim_addcorr norms of both domains.  l = 0 is not solved (tim_icinvX skips
// it).
template <bool PRED>
__global__ void __launch_bounds__(64) k_ic_solve(IcArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[2][64];
    const int k = blockIdx.y;
    const int4 blk = a.blocks[blockIdx.x];
    const int l = blk.x, first = blk.y, nh = blk.z;
    const int N = a.N, Ni = a.Ni, NT = a.NT, S = a.S, t = threadIdx.x, o = Ni - 1;
    double* W = reinterpret_cast<double*>(smem_raw);
    double* lu = W + (size_t)NT * S;
    double* yo = lu + (size_t)NT * LSD_LDAB;
    double* yi = yo + 7 * N;
    int* piv = reinterpret_cast<int*>(yi + 7 * Ni);
    if (l != 0) {
        const double* ls = a.lu + ((size_t)k * a.L + l) * NT * LSD_LDAB;
        for (int i = t; i < NT * LSD_LDAB; i += 64) lu[i] = ls[i];
        const int* ps = a.piv + ((size_t)k * a.L + l) * NT;
        for (int i = t; i < NT; i += 64) piv[i] = ps[i];
        if (PRED) {
            const double* ys = a.yo + ((size_t)k * a.L + l) * N * 7;
            for (int i = t; i < 7 * N; i += 64) yo[i] = ys[i];
            const double* yis = a.yi + (size_t)l * Ni * 7;
            for (int i = t; i < 7 * Ni; i += 64) yi[i] = yis[i];
        }
    }
    // outer core: rows o .. o+N-1 of the columns
    {
        const double2* src = a.nn[k];
        double2* hs = a.hist[k];
        for (int e = t; e < nh * N; e += 64) {
            const int hh = e / N, n = e - hh * N;
            const size_t q = (size_t)a.order[first + hh] * N + n;
            const double2 v = src[q];
            if (PRED) {
                W[(o + n) * S + 2 * hh] = v.x;
                W[(o + n) * S + 2 * hh + 1] = v.y;
            } else {
                const double2 p = hs[q];
                W[(o + n) * S + 2 * hh] = a.cimp * (v.x - p.x);
                W[(o + n) * S + 2 * hh + 1] = a.cimp * (v.y - p.y);
            }
            hs[q] = v;
        }
    }
    // inner core: rows 0 .. Ni-2 (row Ni-1, the boundary, is the outer
    // core's row 0 in the band), every row's N saved
    {
        const double2* src = a.ni[k];
        double2* hs = a.histi[k];
        for (int e = t; e < nh * Ni; e += 64) {
            const int hh = e / Ni, n = e - hh * Ni;
            const size_t q = (size_t)a.order[first + hh] * Ni + n;
            const double2 v = src[q];
            if (n < Ni - 1) {
                if (PRED) {
                    W[n * S + 2 * hh] = v.x;
                    W[n * S + 2 * hh + 1] = v.y;
                } else {
                    const double2 p = hs[q];
                    W[n * S + 2 * hh] = a.cimp * (v.x - p.x);
                    W[n * S + 2 * hh + 1] = a.cimp * (v.y - p.y);
                }
            }
            hs[q] = v;
        }
    }
    __syncthreads();
    const int hh = t >> 1, part = t & 1;
    if (hh < nh) {
        const int h = a.order[first + hh];
        if (PRED && l != 0) {
            const Col co{W + t + (size_t)o * S, S, N};
            multy_col(co, yo, reinterpret_cast<const double*>(a.x[k] + (size_t)h * N) + part);
            // inner-core rows 0 .. Ni-2: N + Y x in ascending column order
            const double* xs = reinterpret_cast<const double*>(a.xi[k] + (size_t)h * Ni) + part;
            for (int n = 0; n < Ni - 1; n++) {
                double acc = W[n * S + t];
#pragma unroll
                for (int d = 0; d < 7; d++) {
                    const int j = n - 3 + d;
                    if (j >= 0 && j < Ni) acc = fma(yi[j * 7 + LSD_KL + n - j], xs[2 * j], acc);
                }
                W[n * S + t] = acc;
            }
        }
        W[t] = 0.0;
        W[(size_t)o * S + t] = 0.0;
        W[(size_t)(o + N - 1) * S + t] = 0.0;
        if (l != 0) gbtrs_col<double>(Col{W + t, S, NT}, lu, piv);
    }
    __syncthreads();
    double e2 = 0.0, eic = 0.0;
    {
        double2* x = a.x[k];
        for (int e = t; e < nh * N; e += 64) {
            const int hh2 = e / N, n = e - hh2 * N;
            const size_t q = (size_t)a.order[first + hh2] * N + n;
            const double dx = W[(o + n) * S + 2 * hh2], dy = W[(o + n) * S + 2 * hh2 + 1];
            if (PRED) {
                x[q] = make_double2(dx, dy);
            } else {
                const double2 p = x[q];
                x[q] = make_double2(p.x + dx, p.y + dy);
                e2 = e2 + dx * dx;
                e2 = e2 + dy * dy;
            }
        }
        double2* xi = a.xi[k];
        for (int e = t; e < nh * Ni; e += 64) {
            const int hh2 = e / Ni, n = e - hh2 * Ni;
            const size_t q = (size_t)a.order[first + hh2] * Ni + n;
            const int row = n < Ni - 1 ? n : o;
            const double dx = W[row * S + 2 * hh2], dy = W[row * S + 2 * hh2 + 1];
            if (PRED) {
                xi[q] = make_double2(dx, dy);
            } else {
                const double2 p = xi[q];
                xi[q] = make_double2(p.x + dx, p.y + dy);
                eic = eic + dx * dx;
                eic = eic + dy * dy;
            }
        }
    }
    if (PRED) return;
    red[0][t] = e2;
    red[1][t] = eic;
    __syncthreads();
    if (t == 0) {
        double s = 0.0, si = 0.0;
        for (int q = 0; q < 64; q++) {
            s += red[0][q];
            si += red[1][q];
        }
        a.partial[(size_t)(2 * k) * gridDim.x + blockIdx.x] = s;
        a.partial[(size_t)(2 * k + 1) * gridDim.x + blockIdx.x] = si;
    }
}

int launch_ic_solve(Ctx& c, bool pred) {
    IcArgs a;
    a.N = c.N;
    a.Ni = c.Nic;
    a.NT = c.NT;
    a.L = c.L;
    a.S = 2 * c.ic_hb + 1;
    a.cimp = c.cfg.implicit;
    a.blocks = c.d_icblk;
    a.order = c.d_icord;
    a.lu = c.d_iclu;
    a.piv = c.d_icpiv;
    a.yo = c.d_y + (size_t)4 * c.L * c.N * 7;
    a.yi = c.d_icy;
    for (int k = 0; k < 2; k++) {
        a.x[k] = c.d_f[LSD_MAG_TOR + k];
        a.nn[k] = c.d_f[LSD_N_MAG_TOR + k];
        a.hist[k] = c.d_hist[4 + k];
        a.xi[k] = c.d_icf[k];
        a.ni[k] = c.d_icn[k];
        a.histi[k] = c.d_ichist[k];
    }
    a.partial = c.d_icpartial;
    const size_t lds = ic_smem_bytes(c.N, c.Nic, a.S);
    // per field and element: state, N and saved N read, state and saved N written
    ProfScope ps(c, LSD_K_SOLVE, 5.0 * 16.0 * 2.0 * (double)c.H * (c.N + c.Nic));
    if (pred) hipLaunchKernelGGL(k_ic_solve<true>, dim3(c.n_icblk, 2), dim3(64), lds, c.stream, a);
    else hipLaunchKernelGGL(k_ic_solve<false>, dim3(c.n_icblk, 2), dim3(64), lds, c.stream, a);
    LSD_HIP(hipGetLastError());
    if (!pred)
        LSD_HIP(hipMemcpyAsync(c.h_icpartial, c.d_icpartial, sizeof(double) * 4 * c.n_icblk, hipMemcpyDeviceToHost,
                               c.stream));
    return LSD_OK;
}

// The inner-core part of non_magnetic (modules/nonlinear.F90:450-471): the
// induction by the inner core's solid-body rotation, N = -rot_omega dB/dphi
// (var_coll_dph: (Re, Im) -> m (-Im, Re)), zero at (l, m) = (0, 0).
__global__ void k_ic_nonlinear(int HN, int Ni, const int* __restrict__ hm, const double* __restrict__ rot,
                               const double2* __restrict__ b0, const double2* __restrict__ b1,
                               double2* __restrict__ n0, double2* __restrict__ n1) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * HN) return;
    const int k = i >= HN, q = i - k * HN, h = q / Ni;
    const double om = rot[ROT_OMEGA];
    const double m_ = (double)hm[h];
    const double2 v = (k ? b1 : b0)[q];
    const double dre = -v.y * m_, dim = v.x * m_;
    double2 r = make_double2(-(om * dre), -(om * dim));
    if (h == 0) r = make_double2(0.0, 0.0);
    (k ? n1 : n0)[q] = r;
}

int launch_ic_nonlinear(Ctx& c) {
    if (!c.ic()) return LSD_OK;
    const int HN = c.H * c.Nic;
    hipLaunchKernelGGL(k_ic_nonlinear, dim3((2 * HN + 255) / 256), dim3(256), 0, c.stream, HN, c.Nic, c.d_hm,
                       c.d_rot, c.d_icf[0], c.d_icf[1], c.d_icn[0], c.d_icn[1]);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

// non_rotation (modules/nonlinear.F90:479-537), no-slip: the torques on the
// inner core from the synthesised fields at r_i, as tra_surfintegral's
// (4 pi / Ph) sum_theta gwP0 sum_phi (modules/transform.fftw3.F90:1039-1062)
// with the phi sums taken on the Fourier rows: sum_phi f = Ph Re F_0 for the
// viscous integrand (-2/r_i) u_phi - (curl u)_theta, and sum_phi Br Bphi =
// Ph (Re B_0 Re P_0 + 2 sum_{m>0} Re(B_m conj P_m)) for the magnetic one
// (the rows past m = M - 1 are zero).  Fourier rows [f][n][theta][MP], n = 0
// the inner boundary.  One workgroup.
template <typename V>
__global__ void __launch_bounds__(256) k_torques(const V* __restrict__ four, size_t per, int Th, int MP, int M,
                                                 int visc, int mag, const double* __restrict__ gwsin,
                                                 double* __restrict__ rot) {
    __shared__ double sv[256], sm[256];
    const double S = -2.0 / rot[ROT_RI];
    double v = 0.0, m = 0.0;
    for (int j = threadIdx.x; j < Th; j += 256) {
        const size_t row = (size_t)j * MP;
        const V up = four[(size_t)LSD_P_VEL_P * per + row];
        const V wt = four[(size_t)LSD_P_VCURL_T * per + row];
        v += gwsin[j] * (S * (double)up.x - (double)wt.x);
        if (mag) {
            const V* br = four + (size_t)LSD_P_MAG_R * per + row;
            const V* bp = four + (size_t)LSD_P_MAG_P * per + row;
            double f = (double)br[0].x * (double)bp[0].x;
            for (int k = 1; k < M; k++)
                f += 2.0 * ((double)br[k].x * (double)bp[k].x + (double)br[k].y * (double)bp[k].y);
            m += gwsin[j] * f;
        }
    }
    sv[threadIdx.x] = v;
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sv[threadIdx.x] += sv[threadIdx.x + w];
            sm[threadIdx.x] += sm[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double FOURPI = 4.0 * 3.1415926535897931;
        // a stress-free inner boundary transmits no viscous torque
        rot[ROT_VTORQ] = visc ? FOURPI * sv[0] * rot[ROT_VFAC] : 0.0;
        rot[ROT_MTORQ] = mag ? FOURPI * sm[0] * rot[ROT_MFAC] : 0.0;
    }
}

int launch_torques(Ctx& c, bool f32) {
    if (!c.cfg.rot) return LSD_OK;
    const int mag = c.ic() ? 1 : 0, visc = c.cfg.vel_bc <= 2 ? 1 : 0;
    if (f32)
        hipLaunchKernelGGL(k_torques<float2>, dim3(1), dim3(256), 0, c.stream,
                           reinterpret_cast<const float2*>(c.d_fourier), c.fourier_elems(), c.Th, c.MP, c.M, visc,
                           mag, c.d_gwsin, c.d_rot);
    else
        hipLaunchKernelGGL(k_torques<double2>, dim3(1), dim3(256), 0, c.stream, c.d_fourier, c.fourier_elems(), c.Th,
                           c.MP, c.M, visc, mag, c.d_gwsin, c.d_rot);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

// rot_predictor / rot_corrector (modules/rotation.F90:43-88): the rotation
// rate from the torques of the evaluation just made, the torques saved.
__global__ void k_rot_update(double* rot, int pred, double cimp) {
    if (threadIdx.x != 0) return;
    const double dt = rot[ROT_DT], I = rot[ROT_INERTIA];
    if (pred) {
        rot[ROT_OMEGA] = rot[ROT_OMEGA] + (dt / I) * (rot[ROT_VTORQ] + rot[ROT_MTORQ]);
    } else {
        rot[ROT_OMEGA] = rot[ROT_OMEGA] + (dt * cimp / I) * ((rot[ROT_VTORQ] + rot[ROT_MTORQ]) -
                                                             (rot[ROT_VTORQ_S] + rot[ROT_MTORQ_S]));
    }
    rot[ROT_VTORQ_S] = rot[ROT_VTORQ];
    rot[ROT_MTORQ_S] = rot[ROT_MTORQ];
}

int launch_rot_update(Ctx& c, bool pred) {
    if (!c.cfg.rot) return LSD_OK;
    hipLaunchKernelGGL(k_rot_update, dim3(1), dim3(64), 0, c.stream, c.d_rot, pred ? 1 : 0, c.cfg.implicit);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

// vel_angmomclr, Boussinesq (modules/velocity.F90:475-513): with both
// walls stress-free, the rigid rotations omz (from (l, m) = (1, 0)) and
// omx, omy (from (1, 1)) are removed from the toroidal velocity, each the
// ratio of int r^2 (r T) dr to int r^2 r^2 dr with the intr2dr weights.
// Thread 0 takes (1, 0), thread 1 (1, 1); sums in ascending radius.
__global__ void k_angmomclr(int N, int h10, int h11, const double* __restrict__ r,
                            const double* __restrict__ intr2dr, double2* __restrict__ uT) {
    const int t = threadIdx.x;
    const int h = t == 0 ? h10 : h11;
    if (t > 1 || h < 0) return;
    const double* r1 = r + 4 * N;
    const double* r2 = r + 5 * N;
    const double w = 8.0 / 3.0 * 3.1415926535897931;
    double2* T = uT + (size_t)h * N;
    double d1 = 0.0, d2 = 0.0, d3 = 0.0;
    for (int n = 0; n < N; n++) {
        d1 = d1 + (T[n].x * r1[n]) * intr2dr[n];
        d2 = d2 + r2[n] * intr2dr[n];
        d3 = d3 + (T[n].y * r1[n]) * intr2dr[n];
    }
    if (t == 0) {
        const double omz = (w * d1) / (w * d2);
        for (int n = 0; n < N; n++) T[n].x = T[n].x - omz * r1[n];
    } else {
        const double omx = (2.0 * w * d1) / (2.0 * w * d2);
        const double omy = (-2.0 * w * d3) / (-2.0 * w * d2);
        for (int n = 0; n < N; n++) T[n] = make_double2(T[n].x - omx * r1[n], T[n].y - omy * r1[n]);
    }
}

int launch_angmomclr(Ctx& c) {
    if (c.cfg.vel_bc != 4) return LSD_OK;
    hipLaunchKernelGGL(k_angmomclr, dim3(1), dim3(64), 0, c.stream, c.N, c.h10, c.h11, c.d_r, c.d_intr2dr,
                       c.d_f[LSD_VEL_TOR]);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

}  // namespace lsd
