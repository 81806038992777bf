// Gauss-Legendre synthesis and analysis as per-m contractions.
//
// Synthesis (modules/transform.fftw3.F90:472-520, 719-818, 955-1000): for an
// m-slot with K harmonics (l = m..L-1, equatorial parity s = (l+m) mod 2 =
// k mod 2 for the k-th harmonic of the slot),
//     Y_p[j, n] = sum_{k : parity p} T[h0+k, j] * c[h0+k, n]        j < Th2
//     F[j] = Y0 + Y1,  F[Th-1-j] = Y0 - Y1
// One wave per (slot, 64 theta points, S shells): lanes run over theta so
// table reads are coalesced; the coefficients of the wave's shells are
// wave-uniform and come through the scalar path.
//
// Analysis (modules/transform.fftw3.F90:527-602, 825-949): the folded
// Fourier data G_p[j, n] (north+south, south-north, from the forward FFT)
// are contracted with the quadrature-weighted tables,
//     out[h0+k, n] = sum_j Tw[j, h0+k] * G_{parity}[j, n]
// One wave per (slot, parity class, 64 harmonics, S shells): lanes run over
// harmonics of one parity class, the analysis tables are stored with each
// slot's even-k harmonics first, so reads are coalesced.
#include "lsd_internal.h"

namespace lsd {

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// -------------------------------------------------------------- synthesis

// SCALAR: one table (P), one input, one output, no parity flip.
template <int S>
__global__ void __launch_bounds__(64) k_synth_scalar(int N, int Th, int Th2, int M, const Slot* __restrict__ slots,
                                                     const int* __restrict__ order, const double* __restrict__ tab,
                                                     const double2* __restrict__ in, double2* __restrict__ out) {
    const int sl = uniform(order[blockIdx.x]);
    const Slot s = slots[sl];
    const int j = blockIdx.y * 64 + threadIdx.x;
    const int n0 = uniform(blockIdx.z * S);
    const int jj = j < Th2 ? j : Th2 - 1;
    double2 y0[S], y1[S];
#pragma unroll
    for (int i = 0; i < S; i++) y0[i] = y1[i] = make_double2(0.0, 0.0);
    const double* T = tab + (size_t)s.h0 * Th2 + jj;
    const double2* C = in + (size_t)s.h0 * N + n0;
    int k = 0;
    for (; k + 1 < s.len; k += 2) {
        double t0 = T[(size_t)k * Th2];
        double t1 = T[(size_t)(k + 1) * Th2];
#pragma unroll
        for (int i = 0; i < S; i++) {
            if (n0 + i < N) {
                double2 c0 = C[(size_t)k * N + i];
                double2 c1 = C[(size_t)(k + 1) * N + i];
                y0[i].x = fma(t0, c0.x, y0[i].x);
                y0[i].y = fma(t0, c0.y, y0[i].y);
                y1[i].x = fma(t1, c1.x, y1[i].x);
                y1[i].y = fma(t1, c1.y, y1[i].y);
            }
        }
    }
    if (k < s.len) {
        double t0 = T[(size_t)k * Th2];
#pragma unroll
        for (int i = 0; i < S; i++) {
            if (n0 + i < N) {
                double2 c0 = C[(size_t)k * N + i];
                y0[i].x = fma(t0, c0.x, y0[i].x);
                y0[i].y = fma(t0, c0.y, y0[i].y);
            }
        }
    }
    if (j >= Th2) return;
#pragma unroll
    for (int i = 0; i < S; i++) {
        if (n0 + i >= N) break;
        double2* F = out + (((size_t)(n0 + i) * M + s.mfft) * Th);
        F[j] = make_double2(y0[i].x + y1[i].x, y0[i].y + y1[i].y);
        if (Th - 1 - j >= Th2) F[Th - 1 - j] = make_double2(y0[i].x - y1[i].x, y0[i].y - y1[i].y);
    }
}

// VECTOR: tables llP_ (parity flipped) and llsin1mP, inputs s and t,
// outputs A_theta and A_phi:
//   A_th = sum s * llP_ + (-i t) * llsin1mP,  A_ph = sum (i s) * llsin1mP + t * llP_
// GRAD: tables P_ (flipped) and sin1mP, input c, outputs scaled by 1/r:
//   A_th = (1/r) sum c * P_,  A_ph = (1/r) sum (i c) * sin1mP
template <int S, bool GRAD>
__global__ void __launch_bounds__(64) k_synth_pair(int N, int Th, int Th2, int M, const Slot* __restrict__ slots,
                                                   const int* __restrict__ order, const double* __restrict__ tabA,
                                                   const double* __restrict__ tabB, const double2* __restrict__ in_s,
                                                   const double2* __restrict__ in_t, const double* __restrict__ rinv,
                                                   double2* __restrict__ out_th, double2* __restrict__ out_ph) {
    const int sl = uniform(order[blockIdx.x]);
    const Slot s = slots[sl];
    const int j = blockIdx.y * 64 + threadIdx.x;
    const int n0 = uniform(blockIdx.z * S);
    const int jj = j < Th2 ? j : Th2 - 1;
    // accumulators [output][parity bucket][shell]
    double2 a0[S], a1[S], b0[S], b1[S];
#pragma unroll
    for (int i = 0; i < S; i++) a0[i] = a1[i] = b0[i] = b1[i] = make_double2(0.0, 0.0);
    const double* TA = tabA + (size_t)s.h0 * Th2 + jj;
    const double* TB = tabB + (size_t)s.h0 * Th2 + jj;
    const double2* CS = in_s + (size_t)s.h0 * N + n0;
    const double2* CT = GRAD ? CS : in_t + (size_t)s.h0 * N + n0;
    for (int k = 0; k < s.len; k++) {
        const double ta = TA[(size_t)k * Th2];  // flipped parity table
        const double tb = TB[(size_t)k * Th2];
        const bool odd = (k & 1);
#pragma unroll
        for (int i = 0; i < S; i++) {
            if (n0 + i < N) {
                double2 cs = CS[(size_t)k * N + i];
                double2 ct = GRAD ? cs : CT[(size_t)k * N + i];
                // flip-parity contributions go to bucket (1 - k%2), the others to k%2
                if (GRAD) {
                    if (!odd) {
                        a1[i].x = fma(ta, cs.x, a1[i].x);
                        a1[i].y = fma(ta, cs.y, a1[i].y);
                        b0[i].x = fma(-tb, cs.y, b0[i].x);
                        b0[i].y = fma(tb, cs.x, b0[i].y);
                    } else {
                        a0[i].x = fma(ta, cs.x, a0[i].x);
                        a0[i].y = fma(ta, cs.y, a0[i].y);
                        b1[i].x = fma(-tb, cs.y, b1[i].x);
                        b1[i].y = fma(tb, cs.x, b1[i].y);
                    }
                } else {
                    if (!odd) {
                        a1[i].x = fma(ta, cs.x, a1[i].x);
                        a1[i].y = fma(ta, cs.y, a1[i].y);
                        b1[i].x = fma(ta, ct.x, b1[i].x);
                        b1[i].y = fma(ta, ct.y, b1[i].y);
                        a0[i].x = fma(tb, ct.y, a0[i].x);
                        a0[i].y = fma(-tb, ct.x, a0[i].y);
                        b0[i].x = fma(-tb, cs.y, b0[i].x);
                        b0[i].y = fma(tb, cs.x, b0[i].y);
                    } else {
                        a0[i].x = fma(ta, cs.x, a0[i].x);
                        a0[i].y = fma(ta, cs.y, a0[i].y);
                        b0[i].x = fma(ta, ct.x, b0[i].x);
                        b0[i].y = fma(ta, ct.y, b0[i].y);
                        a1[i].x = fma(tb, ct.y, a1[i].x);
                        a1[i].y = fma(-tb, ct.x, a1[i].y);
                        b1[i].x = fma(-tb, cs.y, b1[i].x);
                        b1[i].y = fma(tb, cs.x, b1[i].y);
                    }
                }
            }
        }
    }
    if (j >= Th2) return;
#pragma unroll
    for (int i = 0; i < S; i++) {
        const int n = n0 + i;
        if (n >= N) break;
        const double f = GRAD ? rinv[n] : 1.0;
        double2* Fa = out_th + (((size_t)n * M + s.mfft) * Th);
        double2* Fb = out_ph + (((size_t)n * M + s.mfft) * Th);
        Fa[j] = make_double2(f * (a0[i].x + a1[i].x), f * (a0[i].y + a1[i].y));
        Fb[j] = make_double2(f * (b0[i].x + b1[i].x), f * (b0[i].y + b1[i].y));
        if (Th - 1 - j >= Th2) {
            Fa[Th - 1 - j] = make_double2(f * (a0[i].x - a1[i].x), f * (a0[i].y - a1[i].y));
            Fb[Th - 1 - j] = make_double2(f * (b0[i].x - b1[i].x), f * (b0[i].y - b1[i].y));
        }
    }
}

constexpr int SYN_S = 8;

// kind SYN_SCALAR: in0 -> out0 (table P)
// kind SYN_VECTOR: (in0 = s, in1 = t) -> (out0 = theta, out1 = phi)
// kind SYN_GRAD:   in0 -> (out0 = theta, out1 = phi)
int launch_synthesis(Ctx& c, int kind, const double2* in0, const double2* in1, double2* out0, double2* out1) {
    size_t HT = (size_t)c.H * c.Th2;
    dim3 grid(c.nslot, (c.Th2 + 63) / 64, (c.N + SYN_S - 1) / SYN_S);
    // 4 flops per (harmonic, theta point, shell) and product unit
    const double unit = 4.0 * (double)c.H * c.Th2 * c.N;
    ProfScope ps(c, LSD_K_SYNTH, unit * (kind == SYN_SCALAR ? 1 : kind == SYN_VECTOR ? 4 : 2));
    const double* rinv = c.d_r + 2 * c.N;
    if (kind == SYN_SCALAR) {
        hipLaunchKernelGGL((k_synth_scalar<SYN_S>), grid, dim3(64), 0, c.stream, c.N, c.Th, c.Th2, c.M,
                           c.d_slots, c.d_slot_order, c.d_tab + 0 * HT, in0, out0);
    } else if (kind == SYN_VECTOR) {
        hipLaunchKernelGGL((k_synth_pair<SYN_S, false>), grid, dim3(64), 0, c.stream, c.N, c.Th, c.Th2, c.M,
                           c.d_slots, c.d_slot_order, c.d_tab + 3 * HT, c.d_tab + 4 * HT, in0, in1, rinv,
                           out0, out1);
    } else {
        hipLaunchKernelGGL((k_synth_pair<SYN_S, true>), grid, dim3(64), 0, c.stream, c.N, c.Th, c.Th2, c.M,
                           c.d_slots, c.d_slot_order, c.d_tab + 1 * HT, c.d_tab + 2 * HT, in0, in0, rinv,
                           out0, out1);
    }
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

// --------------------------------------------------------------- analysis

// Position of the k-th harmonic of a slot in the analysis tables: even k
// first, then odd k.
__host__ __device__ inline int ana_pos(int h0, int len, int k) {
    return h0 + ((k & 1) ? (len + 1) / 2 + (k >> 1) : (k >> 1));
}

// Folded layout: G[n][m][p][j], p = 0 north+south, p = 1 south-north.
template <int S, bool VEC>
__global__ void __launch_bounds__(64) k_anal(int N, int Th2, int M, int H, const Slot* __restrict__ slots,
                                             const int* __restrict__ order, const double* __restrict__ atab,
                                             const double2* __restrict__ g1, const double2* __restrict__ g2,
                                             const double2* __restrict__ g3, double2* __restrict__ o1,
                                             double2* __restrict__ o2, double2* __restrict__ o3) {
    const int sl = uniform(order[blockIdx.x]);
    const Slot s = slots[sl];
    const int par = uniform(blockIdx.y & 1);  // parity class = k mod 2
    const int lane_k = (blockIdx.y >> 1) * 64 + threadIdx.x;
    const int k = 2 * lane_k + par;
    const int n0 = uniform(blockIdx.z * S);
    const int ncls = par ? s.len / 2 : (s.len + 1) / 2;
    if (uniform((blockIdx.y >> 1) * 64) >= ncls) return;
    const bool active = lane_k < ncls;
    const int ap = active ? ana_pos(s.h0, s.len, k) : s.h0;
    size_t HT = (size_t)H * Th2;
    const double* TP = atab + 0 * HT + ap;   // gwP
    const double* TL = atab + 1 * HT + ap;   // gwllP_
    const double* TS = atab + 2 * HT + ap;   // gwllsin1mP
    // parity buckets: own = par, other = 1 - par
    double2 q[S], sv[S], tv[S];
#pragma unroll
    for (int i = 0; i < S; i++) q[i] = sv[i] = tv[i] = make_double2(0.0, 0.0);
    const size_t gstride = 2 * (size_t)Th2;  // per m
    for (int j = 0; j < Th2; j++) {
        const double tp = TP[(size_t)j * H];
        double tl = 0.0, ts = 0.0;
        if (VEC) {
            tl = TL[(size_t)j * H];
            ts = TS[(size_t)j * H];
        }
#pragma unroll
        for (int i = 0; i < S; i++) {
            const int n = n0 + i;
            if (n < N) {
                const size_t base = ((size_t)n * M + s.mfft) * gstride + j;
                const double2 y1 = g1[base + (size_t)par * Th2];
                q[i].x = fma(y1.x, tp, q[i].x);
                q[i].y = fma(y1.y, tp, q[i].y);
                if (VEC) {
                    const double2 y2o = g2[base + (size_t)(1 - par) * Th2];  // Y2(:,m,1-s)
                    const double2 y2s = g2[base + (size_t)par * Th2];        // Y2(:,m,s)
                    const double2 y3s = g3[base + (size_t)par * Th2];        // Y3(:,m,s)
                    const double2 y3o = g3[base + (size_t)(1 - par) * Th2];  // Y3(:,m,1-s)
                    sv[i].x = fma(y2o.x, tl, sv[i].x);
                    sv[i].x = fma(y3s.y, ts, sv[i].x);
                    sv[i].y = fma(y2o.y, tl, sv[i].y);
                    sv[i].y = fma(-y3s.x, ts, sv[i].y);
                    tv[i].x = fma(-y2s.y, ts, tv[i].x);
                    tv[i].x = fma(y3o.x, tl, tv[i].x);
                    tv[i].y = fma(y2s.x, ts, tv[i].y);
                    tv[i].y = fma(y3o.y, tl, tv[i].y);
                }
            }
        }
    }
    if (!active) return;
    const int h = s.h0 + k;
    const bool m0 = (s.m == 0);
#pragma unroll
    for (int i = 0; i < S; i++) {
        const int n = n0 + i;
        if (n >= N) break;
        size_t o = (size_t)h * N + n;
        if (VEC) {
            o1[o] = q[i];  // the reference leaves q%Im(m=0) as summed
            o2[o] = make_double2(sv[i].x, m0 ? 0.0 : sv[i].y);
            o3[o] = make_double2(tv[i].x, m0 ? 0.0 : tv[i].y);
        } else {
            o1[o] = make_double2(q[i].x, m0 ? 0.0 : q[i].y);
        }
    }
}

constexpr int ANA_S = 8;

int launch_analysis(Ctx& c, int kind, const double2* g0, const double2* g1, const double2* g2, double2* o0,
                    double2* o1, double2* o2) {
    int maxlen = 0;
    for (auto& s : c.slots) maxlen = maxlen > s.len ? maxlen : s.len;
    int ktiles = ((maxlen + 1) / 2 + 63) / 64;
    dim3 grid(c.nslot, 2 * ktiles, (c.N + ANA_S - 1) / ANA_S);
    const double unit = 4.0 * (double)c.H * c.Th2 * c.N;
    ProfScope ps(c, LSD_K_ANAL, unit * (kind == ANA_SCALAR ? 1 : 5));
    if (kind == ANA_SCALAR) {
        hipLaunchKernelGGL((k_anal<ANA_S, false>), grid, dim3(64), 0, c.stream, c.N, c.Th2, c.M, c.H, c.d_slots,
                           c.d_slot_order, c.d_atab, g0, g0, g0, o0, o0, o0);
    } else {
        hipLaunchKernelGGL((k_anal<ANA_S, true>), grid, dim3(64), 0, c.stream, c.N, c.Th2, c.M, c.H, c.d_slots,
                           c.d_slot_order, c.d_atab, g0, g1, g2, o0, o1, o2);
    }
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

}  // namespace lsd
