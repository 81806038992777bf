// Pointwise nonlinear products in physical space, Boussinesq branches of
// modules/nonlinear.F90: advection u x curl u, Coriolis 2Pm/E, Lorentz
// 2Pm/E (J x B) (:71-146), -(Pr/Pm) u.grad C (:266-272), -(Sc/Pm) u.grad xi
// (:335-341) and the induction E = u x B (:420-430).  One pass streams the
// 18 input fields once and writes 8 outputs; every access is coalesced
// along theta.
#include "lsd_internal.h"

namespace lsd {

struct NlParams {
    double cor;   // 2 Pm / E
    double cmag;  // 2 Pm / E
    double qc;    // Pr / Pm
    double qx;    // Sc / Pm
    int cod, comp, mag;
};

__global__ void k_nonlinear(size_t npts, int Th, NlParams p, const double* __restrict__ gp,
                            const double* __restrict__ sn, const double* __restrict__ in, double* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    int j = (int)(i % Th);
    const double lsin = sn[j], lgp = gp[j];
#define IN(f) in[(size_t)(f) * npts + i]
#define OUT(f) out[(size_t)(f) * npts + i]
    const double ur = IN(LSD_P_VEL_R), ut = IN(LSD_P_VEL_T), up = IN(LSD_P_VEL_P);
    const double wr = IN(LSD_P_VCURL_R), wt = IN(LSD_P_VCURL_T), wp = IN(LSD_P_VCURL_P);
    double vr = ut * wp - up * wt;
    double vt = up * wr - ur * wp;
    double vp = ur * wt - ut * wr;
    vr = vr + p.cor * lsin * up;
    vt = vt + p.cor * lgp * up;
    vp = vp - p.cor * (lgp * ut + lsin * ur);
    if (p.mag) {
        const double br = IN(LSD_P_MAG_R), bt = IN(LSD_P_MAG_T), bp = IN(LSD_P_MAG_P);
        const double jr = IN(LSD_P_MCURL_R), jt = IN(LSD_P_MCURL_T), jp = IN(LSD_P_MCURL_P);
        vr = vr + p.cmag * (jt * bp - jp * bt);
        vt = vt + p.cmag * (jp * br - jr * bp);
        vp = vp + p.cmag * (jr * bt - jt * br);
        OUT(NO_ER) = ut * bp - up * bt;
        OUT(NO_ET) = up * br - ur * bp;
        OUT(NO_EP) = ur * bt - ut * br;
    }
    OUT(NO_UR) = vr;
    OUT(NO_UT) = vt;
    OUT(NO_UP) = vp;
    if (p.cod)
        OUT(NO_C) = -p.qc * ur * IN(LSD_P_COD_GR) - p.qc * ut * IN(LSD_P_COD_GT) - p.qc * up * IN(LSD_P_COD_GP);
    if (p.comp)
        OUT(NO_X) = -p.qx * ur * IN(LSD_P_COMP_GR) - p.qx * ut * IN(LSD_P_COMP_GT) - p.qx * up * IN(LSD_P_COMP_GP);
#undef IN
#undef OUT
}

int launch_nonlinear(Ctx& c) {
    NlParams p;
    const lsd_config& k = c.cfg;
    p.cor = 2.0 * k.Pm / k.E;
    p.cmag = 2 * k.Pm / k.E;
    p.qc = k.Pr / k.Pm;
    p.qx = k.Sc / k.Pm;
    p.cod = k.cod;
    p.comp = k.comp;
    p.mag = k.mag;
    size_t np = c.phys_elems();
    int fin = 6 + (k.mag ? 6 : 0) + (k.cod ? 3 : 0) + (k.comp ? 3 : 0);
    int fout = 3 + (k.mag ? 3 : 0) + (k.cod ? 1 : 0) + (k.comp ? 1 : 0);
    ProfScope ps(c, LSD_K_NONLIN, (double)np * 8.0 * (fin + fout));
    hipLaunchKernelGGL(k_nonlinear, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c.stream, np, c.Th, p,
                       c.d_gp, c.d_sin, c.d_phys, c.d_nlphys);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

}  // namespace lsd
