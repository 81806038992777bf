// Spectral-side operators on coll fields (radius fastest): the 7-point
// radial stencils of var_coll_meshmult (modules/variables.F90:476-505),
// the toroidal/poloidal -> qst map, the qst curl, and the r/(l(l+1))
// r-hat.curl / r-hat.curl.curl projections of the nonlinear terms.
// One work-item per (harmonic, radial point); a workgroup holds whole
// columns, whose stencil inputs it stages in LDS.
#include "lsd_internal.h"

namespace lsd {

// Row n of A.x for a band A stored per column j: A(n,j) = band[j*7 + 3+n-j];
// contributions in ascending j, as var_coll_meshmult accumulates them.
__device__ __forceinline__ double2 bandrow(const double* __restrict__ band, const double2* __restrict__ x, int n,
                                           int N) {
    int lo = n - LSD_KL > 0 ? n - LSD_KL : 0;
    int hi = n + LSD_KL < N - 1 ? n + LSD_KL : N - 1;
    double2 acc = make_double2(0.0, 0.0);
    for (int j = lo; j <= hi; j++) {
        double a = band[j * 7 + (LSD_KL + n - j)];
        acc.x = fma(a, x[j].x, acc.x);
        acc.y = fma(a, x[j].y, acc.y);
    }
    return acc;
}

// Spectral inputs of the transforms are stored in the transform precision
// (double2, or float2 for the mixed forms); the stencils always run in FP64.
__device__ __forceinline__ void put(double2* p, double2 v) { *p = v; }
__device__ __forceinline__ void put(float2* p, double2 v) { *p = make_float2((float)v.x, (float)v.y); }
__device__ __forceinline__ double2 get(const double2* p) { return *p; }
__device__ __forceinline__ double2 get(const float2* p) {
    float2 v = *p;
    return make_double2((double)v.x, (double)v.y);
}

// bandrow over a field stored in either precision.
template <typename TI>
__device__ __forceinline__ double2 bandrow_t(const double* __restrict__ band, const TI* __restrict__ x, int n,
                                             int N) {
    int lo = n - LSD_KL > 0 ? n - LSD_KL : 0;
    int hi = n + LSD_KL < N - 1 ? n + LSD_KL : N - 1;
    double2 acc = make_double2(0.0, 0.0);
    for (int j = lo; j <= hi; j++) {
        double a = band[j * 7 + (LSD_KL + n - j)];
        double2 xj = get(x + j);
        acc.x = fma(a, xj.x, acc.x);
        acc.y = fma(a, xj.y, acc.y);
    }
    return acc;
}

// Column blocking: a workgroup takes hb = max(1, 512 / N) whole columns
// (one work-item per radial point) and stages in LDS the columns its
// stencils read, and the band, so each stencil reads LDS rather than
// re-fetching its neighbours through the caches, and an inner stencil is
// computed once per point instead of once per use.
__host__ __device__ inline int spec_hb(int N) { return N >= 512 ? 1 : 512 / N; }

// Copies the N x 7 band into LDS.
__device__ __forceinline__ void stage_band(double* dst, const double* __restrict__ band, int N) {
    for (int i = threadIdx.x; i < N * 7; i += blockDim.x) dst[i] = band[i];
}

// Row n of A.x with the band and the column in LDS; contributions in
// ascending j, as var_coll_meshmult accumulates them.
__device__ __forceinline__ double2 bandrow_l(const double* band, const double2* x, int n, int N) {
    const int lo = n - LSD_KL > 0 ? n - LSD_KL : 0;
    const int hi = n + LSD_KL < N - 1 ? n + LSD_KL : N - 1;
    double2 acc = make_double2(0.0, 0.0);
    for (int j = lo; j <= hi; j++) {
        const double a = band[j * 7 + (LSD_KL + n - j)];
        acc.x = fma(a, x[j].x, acc.x);
        acc.y = fma(a, x[j].y, acc.y);
    }
    return acc;
}

// var_coll_TorPol2qst_mag (variables.F90:565-605) followed by
// var_coll_qstcurl (variables.F90:713-754).
// Up to two fields per launch (blockIdx.y), e.g. u and B of one evaluation.
template <typename TO>
struct VecPrepIO {
    const double2* T[2];
    const double2* P[2];
    TO* o[2][6];  // q, s, t, curl q, curl s, curl t
};

template <typename TO>
__global__ void __launch_bounds__(1024) k_vec_prep(int H, int N, const int* __restrict__ hl,
                                                  const double* __restrict__ r, const double* __restrict__ r1dr,
                                                  const double* __restrict__ ll, VecPrepIO<TO> io, int L) {
    // the field's pointers by scalar selects (indexing the argument arrays
    // would copy them to scratch memory)
    const bool f1 = blockIdx.y != 0;
    const double2* __restrict__ T = f1 ? io.T[1] : io.T[0];
    const double2* __restrict__ P = f1 ? io.P[1] : io.P[0];
    TO* __restrict__ q = f1 ? io.o[1][0] : io.o[0][0];
    TO* __restrict__ s = f1 ? io.o[1][1] : io.o[0][1];
    TO* __restrict__ t = f1 ? io.o[1][2] : io.o[0][2];
    TO* __restrict__ oq = f1 ? io.o[1][3] : io.o[0][3];
    TO* __restrict__ os = f1 ? io.o[1][4] : io.o[0][4];
    TO* __restrict__ ot = f1 ? io.o[1][5] : io.o[0][5];
    extern __shared__ double2 sm[];
    const int hb = spec_hb(N), nc = hb * N;
    double2* Ps = sm;            // [hb][N]
    double2* Ts = sm + nc;       // [hb][N]  -T rt
    double2* Ss = sm + 2 * nc;   // [hb][N]  s = rt r1dr.P
    double* band = reinterpret_cast<double*>(sm + 3 * nc);
    const int c = threadIdx.x / N, n = threadIdx.x - c * N;
    const int h = blockIdx.x * hb + c;
    const bool act = threadIdx.x < nc && h < H;
    const size_t o = (size_t)h * N + n;
    const int l = act ? hl[h] : 0;
    const double rt = ll[2 * L + l];
    stage_band(band, r1dr, N);
    if (act) {
        Ps[threadIdx.x] = P[o];
        const double2 tv = T[o];
        Ts[threadIdx.x] = make_double2(-tv.x * rt, -tv.y * rt);
    }
    __syncthreads();
    const double2* Pc = Ps + c * N;
    if (act) {
        double2 sj = bandrow_l(band, Pc, n, N);
        Ss[threadIdx.x] = make_double2(sj.x * rt, sj.y * rt);
    }
    __syncthreads();
    if (!act) return;
    const double* rinv = r + 2 * N;
    const double ll1 = ll[l];
    const double a = ll1 * rinv[n];
    const double2 pn = Pc[n];
    const double2 qn = make_double2(pn.x * a, pn.y * a);
    const double2* Sc = Ss + c * N;
    const double2* Tc = Ts + c * N;
    const int lo = n - LSD_KL > 0 ? n - LSD_KL : 0;
    const int hi = n + LSD_KL < N - 1 ? n + LSD_KL : N - 1;
    double2 otn = make_double2(0.0, 0.0), osn = make_double2(0.0, 0.0);
    for (int j = lo; j <= hi; j++) {
        const double2 sj = Sc[j], tj = Tc[j];
        const double b = band[j * 7 + (LSD_KL + n - j)];
        otn.x = fma(b, sj.x, otn.x);
        otn.y = fma(b, sj.y, otn.y);
        osn.x = fma(b, tj.x, osn.x);
        osn.y = fma(b, tj.y, osn.y);
    }
    const double2 sn = Sc[n], tn = Tc[n];
    const double cc = rt * rinv[n];
    put(q + o, qn);
    put(s + o, sn);
    put(t + o, tn);
    put(ot + o, make_double2(otn.x - cc * qn.x, otn.y - cc * qn.y));
    put(os + o, make_double2(-osn.x, -osn.y));
    const double d = -rt * rinv[n];
    put(oq + o, make_double2(tn.x * d, tn.y * d));
}

// dC/dr by the first-derivative stencil (codensity.F90:207); Cc, when
// given, receives C itself in the transform precision.
// Up to two fields per launch (blockIdx.y): codensity and composition.
template <typename TO>
struct ScalPrepIO {
    const double2* C[2];
    TO* dC[2];
    TO* Cc[2];
};

template <typename TO>
__global__ void __launch_bounds__(1024) k_scal_prep(int H, int N, const double* __restrict__ dr1,
                                                   ScalPrepIO<TO> io) {
    const bool f1 = blockIdx.y != 0;
    const double2* __restrict__ C = f1 ? io.C[1] : io.C[0];
    TO* __restrict__ dC = f1 ? io.dC[1] : io.dC[0];
    TO* __restrict__ Cc = f1 ? io.Cc[1] : io.Cc[0];
    extern __shared__ double2 sm[];
    const int hb = spec_hb(N), nc = hb * N;
    double2* Cs = sm;
    double* band = reinterpret_cast<double*>(sm + nc);
    const int c = threadIdx.x / N, n = threadIdx.x - c * N;
    const int h = blockIdx.x * hb + c;
    const bool act = threadIdx.x < nc && h < H;
    const size_t o = (size_t)h * N + n;
    stage_band(band, dr1, N);
    if (act) Cs[threadIdx.x] = C[o];
    __syncthreads();
    if (!act) return;
    put(dC + o, bandrow_l(band, Cs + c * N, n, N));
    if (Cc) put(Cc + o, Cs[threadIdx.x]);
}

struct PostParams {
    double Ra, Ra_comp, Pm2, Pr, Sc;
    double src_c, src_x;  // d_source, d_source_comp
    int cod, comp, mag;
};

// Nonlinear terms in coll form (nonlinear.F90:148-197, 277-288, 341-348,
// 432-471): buoyancy on q, r/(l(l+1)) r.curl and r.curl.curl projections,
// l = 0 zeroed, sources on the l = 0 scalar modes.
template <typename TI>
__global__ void __launch_bounds__(1024) k_spec_post(int H, int N, int L, const int* __restrict__ hl,
                                                   const double* __restrict__ r, const double* __restrict__ r1dr,
                                                   const double* __restrict__ ll, PostParams pp,
                                                   const TI* __restrict__ so, const double2* __restrict__ Cst,
                                                   const double2* __restrict__ Xst, double2* __restrict__ NuT,
                                                   double2* __restrict__ NuP, double2* __restrict__ NC,
                                                   double2* __restrict__ NX, double2* __restrict__ NbT,
                                                   double2* __restrict__ NbP) {
    extern __shared__ double2 sm[];
    const int hb = spec_hb(N), nc = hb * N;
    double2* Us = sm;            // analysis s of u
    double2* Es = sm + nc;       // analysis s of E (magnetic)
    double* band = reinterpret_cast<double*>(sm + 2 * nc);
    const int c = threadIdx.x / N, n = threadIdx.x - c * N;
    const int h = blockIdx.x * hb + c;
    const bool act = threadIdx.x < nc && h < H;
    const size_t HN = (size_t)H * N;
    const size_t idx = (size_t)h * N + n;
    stage_band(band, r1dr, N);
    if (act) {
        Us[threadIdx.x] = get(so + SO_US * HN + idx);
        if (pp.mag) Es[threadIdx.x] = get(so + SO_ES * HN + idx);
    }
    __syncthreads();
    if (!act) return;
    const int l = hl[h];
    const double* rinv = r + 2 * N;
    const double rn = r[4 * N + n];
    const double rt11 = ll[3 * L + l];
    // buoyancy (nonlinear.F90:158-185)
    double2 cq = get(so + SO_UQ * HN + idx);
    double d2 = pp.Ra * rn * pp.Pm2 / pp.Pr;
    const double2 cs = Cst[idx], xs = Xst[idx];
    cq.x = cq.x + d2 * cs.x;
    cq.y = cq.y + d2 * cs.y;
    d2 = pp.Ra_comp * rn * pp.Pm2 / pp.Sc;
    cq.x = cq.x + d2 * xs.x;
    cq.y = cq.y + d2 * xs.y;
    const double2 utv = get(so + SO_UT * HN + idx);
    double2 nt = make_double2(-utv.x * rt11, -utv.y * rt11);
    double2 v = bandrow_l(band, Us + c * N, n, N);
    v.x = v.x * rt11;
    v.y = v.y * rt11;
    double2 np = make_double2(cq.x * rinv[n] - v.x, cq.y * rinv[n] - v.y);
    if (h == 0) nt = np = make_double2(0.0, 0.0);
    NuT[idx] = nt;
    NuP[idx] = np;
    if (pp.cod) {
        double2 cv = get(so + SO_C * HN + idx);
        if (h == 0) cv.x = cv.x + pp.src_c;
        NC[idx] = cv;
    }
    if (pp.comp) {
        double2 cv = get(so + SO_X * HN + idx);
        if (h == 0) cv.x = cv.x + pp.src_x;
        NX[idx] = cv;
    }
    if (pp.mag) {
        const double2 etv = get(so + SO_ET * HN + idx), eqv = get(so + SO_EQ * HN + idx);
        double2 bp = make_double2(-etv.x * rt11, -etv.y * rt11);
        double2 w = bandrow_l(band, Es + c * N, n, N);
        w.x = w.x * rt11;
        w.y = w.y * rt11;
        double2 bt = make_double2(eqv.x * rinv[n] - w.x, eqv.y * rinv[n] - w.y);
        if (h == 0) bp = bt = make_double2(0.0, 0.0);
        NbP[idx] = bp;
        NbT[idx] = bt;
    }
}

// spec re[n*H+h], im -> coll double2 [h][n] and back (the serial
// var_coll2spec / var_spec2coll transposes, variables.F90:246-268, 365-389).
__global__ void k_spec2coll(int H, int N, const double* __restrict__ re, const double* __restrict__ im,
                            double2* __restrict__ out) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)H * N) return;
    int h = (int)(idx / N), n = (int)(idx - (size_t)h * N);
    out[idx] = make_double2(re[(size_t)n * H + h], im[(size_t)n * H + h]);
}
__global__ void k_coll2spec(int H, int N, const double2* __restrict__ in, double* __restrict__ re,
                            double* __restrict__ im) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)H * N) return;
    int n = (int)(idx / H), h = (int)(idx - (size_t)n * H);
    double2 v = in[(size_t)h * N + n];
    re[idx] = v.x;
    im[idx] = v.y;
}

static dim3 grid1(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

// Column-blocked launch shape of the spectral kernels, and their LDS
// (ncols staged columns of double2 plus the band).
static dim3 spec_grid(const Ctx& c) { return dim3((unsigned)((c.H + spec_hb(c.N) - 1) / spec_hb(c.N))); }
static dim3 spec_block(const Ctx& c) { return dim3((unsigned)(spec_hb(c.N) * c.N)); }
static size_t spec_lds(const Ctx& c, int ncols) {
    return sizeof(double2) * (size_t)ncols * spec_hb(c.N) * c.N + sizeof(double) * 7 * (size_t)c.N;
}

template <typename TO>
static int vec_prep_t(Ctx& c, int nf, const double2* const* T, const double2* const* P, void* const* const* o) {
    ProfScope ps(c, LSD_K_SPEC, nf * (2.0 * 16.0 + 6.0 * sizeof(TO)) * c.coll_elems());
    VecPrepIO<TO> io;
    for (int f = 0; f < 2; f++) {
        const int g = f < nf ? f : 0;
        io.T[f] = T[g];
        io.P[f] = P[g];
        for (int k = 0; k < 6; k++) io.o[f][k] = reinterpret_cast<TO*>(o[g][k]);
    }
    dim3 grid = spec_grid(c);
    grid.y = nf;
    hipLaunchKernelGGL(k_vec_prep<TO>, grid, spec_block(c), spec_lds(c, 3), c.stream, c.H, c.N, c.d_hl, c.d_r,
                       c.d_r1dr, c.d_ll, io, c.L);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

// nf (1 or 2) fields (T[f], P[f]) -> o[f][0..5] = q, s, t, curl q, curl s,
// curl t in the layout of spec_in slots
int launch_vec_prep(Ctx& c, int nf, const double2* const* T, const double2* const* P, void* const* const* o,
                    bool f32) {
    if (nf < 1 || nf > 2) {
        set_error("vector preparation: 1 or 2 fields per launch");
        return LSD_E_ARG;
    }
    return f32 ? vec_prep_t<float2>(c, nf, T, P, o) : vec_prep_t<double2>(c, nf, T, P, o);
}

template <typename TO>
static int scal_prep_t(Ctx& c, int nf, const double2* const* C, void* const* dC, void* const* Cc) {
    ProfScope ps(c, LSD_K_SPEC, nf * (16.0 + sizeof(TO) * (Cc[0] ? 2.0 : 1.0)) * c.coll_elems());
    ScalPrepIO<TO> io;
    for (int f = 0; f < 2; f++) {
        const int g = f < nf ? f : 0;
        io.C[f] = C[g];
        io.dC[f] = reinterpret_cast<TO*>(dC[g]);
        io.Cc[f] = reinterpret_cast<TO*>(Cc[g]);
    }
    dim3 grid = spec_grid(c);
    grid.y = nf;
    hipLaunchKernelGGL(k_scal_prep<TO>, grid, spec_block(c), spec_lds(c, 1), c.stream, c.H, c.N, c.d_dr1, io);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

// nf (1 or 2) scalar fields C[f] -> dC[f] = dC/dr and, when Cc[f] is not
// null, Cc[f] = C in the transform precision
int launch_scal_prep(Ctx& c, int nf, const double2* const* C, void* const* dC, void* const* Cc, bool f32) {
    if (nf < 1 || nf > 2) {
        set_error("scalar preparation: 1 or 2 fields per launch");
        return LSD_E_ARG;
    }
    return f32 ? scal_prep_t<float2>(c, nf, C, dC, Cc) : scal_prep_t<double2>(c, nf, C, dC, Cc);
}

int launch_spec_post(Ctx& c, bool f32) {
    PostParams pp;
    const lsd_config& k = c.cfg;
    // reads: the analysis outputs and the C, xi state; writes the N fields
    const int nso = 3 + (k.cod ? 1 : 0) + (k.comp ? 1 : 0) + (k.mag ? 3 : 0);
    const int nwr = 2 + (k.cod ? 1 : 0) + (k.comp ? 1 : 0) + (k.mag ? 2 : 0);
    const double rb = f32 ? 8.0 : 16.0;
    ProfScope ps(c, LSD_K_SPEC, (nso * rb + 2 * 16.0 + nwr * 16.0) * c.coll_elems());
    pp.Ra = k.Ra;
    pp.Ra_comp = k.Ra_comp;
    pp.Pm2 = k.Pm * k.Pm;
    pp.Pr = k.Pr;
    pp.Sc = k.Sc;
    pp.src_c = k.source;
    pp.src_x = k.source_comp;
    pp.cod = k.cod;
    pp.comp = k.comp;
    pp.mag = k.mag;
    if (f32)
        hipLaunchKernelGGL(k_spec_post<float2>, spec_grid(c), spec_block(c), spec_lds(c, 2), c.stream, c.H, c.N, c.L,
                           c.d_hl, c.d_r, c.d_r1dr, c.d_ll, pp, reinterpret_cast<const float2*>(c.d_spec_out),
                           c.d_f[LSD_COD], c.d_f[LSD_COMP], c.d_f[LSD_N_VEL_TOR], c.d_f[LSD_N_VEL_POL],
                           c.d_f[LSD_N_COD], c.d_f[LSD_N_COMP], c.d_f[LSD_N_MAG_TOR], c.d_f[LSD_N_MAG_POL]);
    else
        hipLaunchKernelGGL(k_spec_post<double2>, spec_grid(c), spec_block(c), spec_lds(c, 2), c.stream, c.H, c.N,
                           c.L, c.d_hl, c.d_r, c.d_r1dr, c.d_ll, pp, c.d_spec_out, c.d_f[LSD_COD], c.d_f[LSD_COMP],
                           c.d_f[LSD_N_VEL_TOR], c.d_f[LSD_N_VEL_POL], c.d_f[LSD_N_COD], c.d_f[LSD_N_COMP],
                           c.d_f[LSD_N_MAG_TOR], c.d_f[LSD_N_MAG_POL]);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

int launch_layout_spec2coll(Ctx& c, const double* re, const double* im, double2* out) {
    hipLaunchKernelGGL(k_spec2coll, grid1(c.coll_elems()), dim3(256), 0, c.stream, c.H, c.N, re, im, out);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

int launch_layout_coll2spec(Ctx& c, const double2* in, double* re, double* im) {
    hipLaunchKernelGGL(k_coll2spec, grid1(c.coll_elems()), dim3(256), 0, c.stream, c.H, c.N, in, re, im);
    LSD_HIP(hipGetLastError());
    return LSD_OK;
}

}  // namespace lsd
