// Internal declarations of the MI355X per-timestep library.
//
// Device layouts (all resident in HBM for the life of a context):
//   coll     double2 [h][n]            complex, radius fastest (the reference's
//                                       coll type with Re/Im interleaved)
//   fourier  double2 [f][n][j][MP]     m = FFT index fastest, MP = M rounded
//                                       up to 16 (a theta row is one run)
//   folded   double2 [f][n][m][p][j]   j < Th2, p = equatorial parity bucket
//                                       (north+south, south-north)
//   phys     double  [f][n][k][j]      the reference's (Th, Ph, N) layout
//   tables   synthesis  double [h][j]  (j < Th2 fastest, like leg_P(j,nh));
//                       FP32 form per slot, theta-permuted (see d_tab32)
//            analysis   double [j][h]  (h fastest, so lanes over harmonics
//                                       read coalesced)
//   bands    mesh  double [n][7]       M(KL+1+n-j, j) stored per column j
//            lu    double [l][n][10]   LAPACK band storage + int pivots
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/lsd.h"

#define LSD_KL 3
#define LSD_LDAB 10
#define LSD_MBW 7

namespace lsd {

void set_error(const std::string& msg);

// Folded harmonic order (Scripts/parallel.h_template:65-80): l and m_ of
// each harmonic index.
void harmonics(int L, int M, int Mp, std::vector<int>& hl, std::vector<int>& hm);

#define LSD_HIP(call)                                                         \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            lsd::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return LSD_E_HIP;                                                 \
        }                                                                     \
    } while (0)

// Stream-ordered scratch of a set-up or diagnostic call, returned to the
// pool on every exit path (the call synchronises the stream before using
// what the kernels wrote, so the free follows all uses).
struct AsyncScratch {
    void* p = nullptr;
    hipStream_t s;
    explicit AsyncScratch(hipStream_t st) : s(st) {}
    AsyncScratch(const AsyncScratch&) = delete;
    AsyncScratch& operator=(const AsyncScratch&) = delete;
    ~AsyncScratch() {
        if (p) (void)hipFreeAsync(p, s);
    }
    template <class T>
    hipError_t alloc(T** out, size_t bytes) {
        hipError_t e = hipMallocAsync(&p, bytes, s);
        *out = static_cast<T*>(p);
        return e;
    }
};

// One m-slot: a run of harmonics with equal m (the reference's _loop_sumth).
struct Slot {
    int h0;    // first harmonic index
    int len;   // number of harmonics (l = m .. L-1)
    int m;     // actual order m_
    int mfft;  // FFT index m_/Mp
};

// Scalar syntheses of one evaluation at most: q of u, curl u, B, curl B,
// dC/dr, dxi/dr.
constexpr int LG_SCALARS = 6;

// Synthesis job kinds.
enum SynthKind { SYN_SCALAR = 0, SYN_VECTOR = 1, SYN_GRAD = 2 };

// Analysis job kinds.
enum AnalKind { ANA_SCALAR = 0, ANA_VECTOR = 1 };

// Banded operator kinds (LU): 0 vel tor, 1 vel pol, 2 vel green, 3 cod,
// 4 comp, 5 mag tor, 6 mag pol.  Y kinds: 0 vel tor, 1 vel pol, 2 cod,
// 3 comp, 4 mag tor, 5 mag pol.
constexpr int NLU = 7;
constexpr int NY = 6;

// Synthesis input (coll-layout spectral) buffer ids.
enum SpecIn {
    SI_UQ = 0, SI_US, SI_UT, SI_WQ, SI_WS, SI_WT,
    SI_BQ, SI_BS, SI_BT, SI_JQ, SI_JS, SI_JT,
    SI_DC, SI_DX,
    SI_CC, SI_XC,  // C and xi themselves in the transform precision (mixed forms)
    SI_COUNT
};

// Analysis output (coll-layout) buffer ids.
enum SpecOut { SO_UQ = 0, SO_US, SO_UT, SO_C, SO_X, SO_EQ, SO_ES, SO_ET, SO_COUNT };

// Physical output field ids of the nonlinear products (folded buffers).
enum NlOut { NO_UR = 0, NO_UT, NO_UP, NO_C, NO_X, NO_ER, NO_ET, NO_EP, NO_COUNT };

struct Ctx {
    lsd_config cfg;
    int device = 0;
    int L, M, Mp, N, Th, Ph, H, Th2, nslot;
    int MP;                        // Fourier row length: M rounded up to 16
    double dt;
    bool mixed = false;            // FP32 transforms and solves (LSD_MIXED_*)
    bool mag64 = false;            // B and curl B synthesised in FP64 (LSD_MIXED_F32_MAG64)
    hipStream_t stream = nullptr;

    std::vector<int> hl, hm;       // per harmonic l, m_
    std::vector<Slot> slots;       // in harmonic order
    std::vector<int> slot_order;   // slots sorted by decreasing length

    // device: index data
    int* d_hl = nullptr;           // [H]
    int* d_hm = nullptr;           // [H]
    int* d_pidx = nullptr;         // packed P_l^m index -> h, [(L)(L+1)/2], -1 if absent
    int* d_apos = nullptr;         // analysis-table position of each harmonic [H]
    Slot* d_slots = nullptr;       // [nslot]
    int* d_slot_order = nullptr;   // [nslot]
    int* d_mslot = nullptr;        // [M] slot of each FFT index

    // device: quadrature and tables
    double* d_gp = nullptr;        // [Th]
    double* d_gw = nullptr;        // [Th]
    double* d_sin = nullptr;       // [Th]
    double* d_tab = nullptr;       // 5 synthesis tables [5][H][Th2]
    double* d_atab = nullptr;      // 3 analysis tables [3][Th2][H]
    // FP32 synthesis tables (mixed forms) [5][Hp][Th2p]: each slot's rows
    // from row slot_t0[s], padded with zero rows to a multiple of 8; theta
    // rounded up to Th2p = 64-multiple and permuted within each block of 64
    // (position 4 l + t holds theta 16 t + l)
    float* d_tab32 = nullptr;
    int Th2p = 0, Hp = 0;
    int* d_slot_t0 = nullptr;      // [nslot] first table row of each slot
    int* d_row_h = nullptr;        // [Hp] harmonic of each table row, -1 for padding
    float* d_atab32 = nullptr;     // FP32 copies of the analysis tables
    // FP32 analysis tables of the wide-load kernel [3][Th2a][Hpa]: per (slot,
    // parity class) a run of whole 64-harmonic blocks from column
    // d_ana_base[2 slot + class], permuted within each block (position 4 l +
    // t holds class harmonic 16 t + l); theta zero-padded to Th2a, a multiple
    // of 8; zero where a block runs past the class
    float* d_atab4 = nullptr;
    int Th2a = 0, Hpa = 0, ana_maxhb = 0;
    int* d_ana_base = nullptr;     // [2 nslot]
    int* d_acol = nullptr;         // [Hpa] column of the FP64 analysis tables, -1 for padding
    // BF16x3 split tables (LSD_MIXED_BF16X3), layouts in legendre_bf16.hip
    __bf16* d_tabb = nullptr;      // synthesis [5][class blocks of 32][theta blocks of 64][4096]
    __bf16* d_atabb = nullptr;     // analysis [3][class blocks of 64][theta blocks of 32][4096]
    int* d_bx_sbase = nullptr;     // [2 nslot] first synthesis block, then [2 nslot] first analysis block
    int bx_nsb = 0, bx_nab = 0, bx_nthb = 0, bx_njb = 0, bx_maxsb = 0, bx_maxab = 0;
    // device: radial mesh
    double* d_r = nullptr;         // [7][N]  r^p, p = -3..3
    double* d_dr1 = nullptr;       // [N][7]
    double* d_dr2 = nullptr;
    double* d_radlap = nullptr;
    double* d_r1dr = nullptr;
    double* d_intr2dr = nullptr;   // [N]
    double* d_ll = nullptr;        // [4][L] ll1, ll11, rtll1, rtll11
    // device: operators
    double* d_lu = nullptr;        // [NLU][L][N][10]
    int* d_piv = nullptr;          // [NLU][L][N]
    unsigned* d_pivpk = nullptr;   // [NLU][L][(N+15)/16] pivot offsets, 2 bits per row
    double* d_y = nullptr;         // [NY][L][N][7]
    double* d_gre = nullptr;       // [L][2][N]
    double* d_invg = nullptr;      // [L][4]
    int* d_info = nullptr;         // [NLU][L] dgbtrf info
    float* d_lut32 = nullptr;      // [NLU][L][N][10]  FP32 factors (mixed forms)
    double* d_xband = nullptr;     // [NLU][L][N][7]   X before factorisation
    int4* d_sblocks = nullptr;     // solve blocks: degree, first index into d_sorder, count
    int* d_sorder = nullptr;       // [H] harmonics grouped by degree

    // device: fields
    double2* d_f[LSD_NUM_COLL] = {};   // state + current N
    double2* d_hist[6] = {};           // saved N^n
    double2* d_rhs[6] = {};            // solve right-hand sides [H][N] (two-pass solves)
    double2* d_rhs2 = nullptr;         // poloidal velocity's second right-hand side
    double2* d_spec_in = nullptr;      // [SI_COUNT][H][N]
    double2* d_spec_out = nullptr;     // [SO_COUNT][H][N]
    double2* d_fourier = nullptr;      // [18][N][M][Th]
    double2* d_folded = nullptr;       // [8][N][M][2][Th2]
    double* d_phys = nullptr;          // [18][N][Ph][Th]
    double* d_nlphys = nullptr;        // [8][N][Ph][Th]
    double2* d_tw = nullptr;           // FFT twiddles exp(2 pi i k/Ph) [Ph]
    float2* d_tw32 = nullptr;          // same, rounded to FP32
    double* d_partial = nullptr;       // corrector norm partials
    double* h_partial = nullptr;       // pinned copy
    int n_partial_blocks = 0;          // solve blocks per field
    int solve_hb = 0;                  // harmonics per solve block

    // kernel variants (LSD_V_*), see lsd_set_variant
    int var[LSD_V_COUNT] = {4, 8, 8, 0, 0};
    bool ana_wide_ok = false;          // the wide-load FP32 analysis applies (even Th2)
    bool solve_direct_ok = false;      // the two-pass solves apply (N in {32, 48, 64, 96})
    // solve blocks of the two forms: [0] two-pass (32 harmonics), [1] LDS-column
    // solve block sets: 0 two-pass, 1 LDS-column, 2 LDS-column with FP64
    // factors in the mixed forms
    int4* d_sblk[3] = {nullptr, nullptr, nullptr};
    int* d_sord[3] = {nullptr, nullptr, nullptr};
    int nblk_form[3] = {0, 0, 0};
    int hb_form[3] = {0, 0, 0};

    // captured step graphs: [evaluate, predictor, evaluate, corrector] and
    // [evaluate, corrector], each ending with the copy of the norm partials
    bool use_graph = true;
    hipGraphExec_t g_first = nullptr, g_more = nullptr;
    hipGraphExec_t g_eval = nullptr;       // adaptive step: predictor pass's evaluation + CFL
    int g_prof_class = -1;                 // class timed by events inside the graphs
    int cap = -1;                          // graph being captured (0 first, 1 more)
    std::vector<hipEvent_t> g_ev[2];       // (start, stop) pairs recorded by each graph
    double g_work[2] = {0.0, 0.0};         // work of the timed class per replay
    double prof_acc_ms = 0.0, prof_acc_work = 0.0;
    int prof_acc_launches = 0;

    // per-launch timing of one kernel class (lsd_profile_begin/end)
    int prof_class = -1;
    std::vector<hipEvent_t> prof_events;  // pool, pairs (start, stop)
    int prof_used = 0;
    double prof_work = 0.0;

    // host time-stepping state (modules/timestep.F90:12-20)
    double t = 0.0;
    int it = 0;
    int step = 0;
    double dterr = 0.0, lasterr = 1e99, corr_dt = 0.0;

    // adaptive step (d_timestep < 0): CFL partial maxima of the predictor
    // pass's evaluation, folded into tim_cfl_dt on the device
    bool adaptive = false;
    bool cfl_on = false;               // the next evaluation forms the CFL partials
    int cfl_parts = 1;                 // partial pairs per shell in d_cfl_part
    double cfl_dt = 0.0;               // tim_cfl_dt
    double* d_cfl_d = nullptr;         // [N][2] per-shell constants
    double* d_cfl_part = nullptr;      // [N][max(Th2, 1)][2]
    double* d_cfl = nullptr;           // [1] tim_cfl_dt
    double* h_cfl = nullptr;           // pinned copy

    // conducting inner core (cfg.Nic > 1) and the inner-core rotation
    // (cfg.rot), modules/magnetic.F90, modules/rotation.F90
    int Nic = 1, NT = 0;               // inner-core points; coupled band length Nic + N - 1
    double* d_icr = nullptr;           // [7][Nic] inner-core r^p, p = -3..3
    double* d_icdr1 = nullptr;         // [Nic][7]
    double* d_icdr2 = nullptr;         // [Nic][7]
    double* d_icradlap = nullptr;      // [Nic][7]
    double* d_iclu = nullptr;          // [2][L][NT][10] coupled factors of BT, BP
    int* d_icpiv = nullptr;            // [2][L][NT]
    double* d_icy = nullptr;           // [L][Nic][7] inner-core Y
    double2* d_icf[2] = {};            // icBT, icBP [H][Nic]
    double2* d_icn[2] = {};            // their N
    double2* d_ichist[2] = {};         // their saved N
    int4* d_icblk = nullptr;           // coupled-solve blocks (degree, first, count)
    int* d_icord = nullptr;            // [H] harmonics grouped by degree
    int n_icblk = 0, ic_hb = 0;
    double* d_icpartial = nullptr;     // [4][n_icblk] correction norms: BT, icBT, BP, icBP
    double* h_icpartial = nullptr;     // pinned copy
    double* d_rot = nullptr;           // [ROT_COUNT], see RotSlot
    double* d_gwsin = nullptr;         // [Th] leg_gwP0 sin(theta)
    int h10 = -1, h11 = -1;            // harmonics (l, m) = (1, 0), (1, 1) (-1 if absent)
    // persistent kernel forms: work-queue counters (zero between launches),
    // the device's CU count and the workgroups of each persistent launch
    int* d_wq = nullptr;               // [32]: synthesis queues [0, 9), analysis [16, 25)
    int ncu = 0;
    int syn_pgrid[3][2] = {};
    int syn_qgrid[3][2] = {};
    int ana_pgrid[2] = {};
    bool ic() const { return Nic > 1 && cfg.mag; }
    // the rotating inner core moves the wall: b_rot_tstep with a no-slip
    // inner boundary (vel_setbc_Tor, non_rotation)
    bool rot_wall() const { return cfg.rot && cfg.vel_bc <= 2; }

    size_t coll_elems() const { return (size_t)H * N; }
    size_t phys_elems() const { return (size_t)Th * Ph * N; }
    size_t fourier_elems() const { return (size_t)N * Th * MP; }
    size_t folded_elems() const { return (size_t)N * M * 2 * Th2; }
};

// Brackets one launch of a kernel class with events when that class is
// being timed; `work` is the launch's algorithmic flops or bytes.
struct ProfScope {
    Ctx& c;
    bool on;
    hipEvent_t stop = nullptr;
    ProfScope(Ctx& ctx, int cls, double work);
    ~ProfScope();
};

// precompute.hip
int precompute_tables(Ctx& c);
int tables_to_f32(Ctx& c);
int tables_ana4(Ctx& c);
int precompute_mesh(Ctx& c);

// Inner core and rotation (precompute.hip set-up, innercore.hip kernels).
// d_rot holds the rotation state and its constants.
enum RotSlot {
    ROT_OMEGA = 0, ROT_VTORQ, ROT_MTORQ, ROT_VTORQ_S, ROT_MTORQ_S,  // rot_omega, torques, saved torques
    ROT_INERTIA, ROT_RI, ROT_VFAC, ROT_MFAC, ROT_DT, ROT_COUNT
};
int precompute_innercore(Ctx& c);
int build_ic_operators(Ctx& c);
size_t ic_smem_bytes(int N, int Nic, int S);
int ic_block_harmonics(int N, int Nic);
int launch_ic_solve(Ctx& c, bool pred);
int launch_torques(Ctx& c, bool f32);
int launch_ic_nonlinear(Ctx& c);
int launch_rot_update(Ctx& c, bool pred);
int launch_angmomclr(Ctx& c);
int build_operators(Ctx& c);

// legendre.hip
int launch_synthesis(Ctx& c, int kind, const double2* in0, const double2* in1, double2* out0,
                     double2* out1);
int launch_analysis(Ctx& c, int kind, const double2* g0, const double2* g1, const double2* g2,
                     double2* o0, double2* o1, double2* o2);

int launch_synthesis_p(Ctx& c, int kind, const void* in0, const void* in1, void* out0, void* out1, bool f32);
int launch_synthesis_group(Ctx& c, int kind, int njob, const void* const* in0, const void* const* in1,
                           void* const* out0, void* const* out1, bool f32, bool out32 = false);
int launch_analysis_group(Ctx& c, int kind, int njob, const void* const* g0, const void* const* g1,
                          const void* const* g2, void* const* o0, void* const* o1, void* const* o2, bool f32);
int launch_analysis_p(Ctx& c, int kind, const void* g0, const void* g1, const void* g2, void* o0, void* o1,
                      void* o2, bool f32);

// physfuse.hip
bool physfuse_supported(const Ctx& c);
int launch_physfuse(Ctx& c, bool f32);

// fft.hip
int fft_setup(Ctx& c);
int launch_fft_c2r(Ctx& c, const double2* fourier, double* phys, int nfield);
int launch_fft_r2c_fold(Ctx& c, const double* phys, double2* folded, int nfield);

// spectral.hip
int launch_vec_prep(Ctx& c, int nf, const double2* const* T, const double2* const* P, void* const* const* out6,
                    bool f32);
int launch_scal_prep(Ctx& c, int nf, const double2* const* C, void* const* dC, void* const* Cc, bool f32);
int launch_spec_post(Ctx& c, bool f32);
int launch_layout_spec2coll(Ctx& c, const double* re, const double* im, double2* out);
int launch_layout_coll2spec(Ctx& c, const double2* in, double* re, double* im);

// nonlinear.hip
int launch_nonlinear(Ctx& c);

// energy.hip
int energy(Ctx& c, int which, double* E, double* El, double* Em);
size_t energy_scratch_doubles(const Ctx& c);
int energy_run(Ctx& c, int which, double* scratch, double* E, double* El, double* Em);
int energy_noalloc_init(int device);
int energy_noalloc(Ctx& c, int which, double* E, double* El, double* Em);

// state_io.hip: coll = [6][2][H][N] (uT, uP, C, Comp, BT, BP; Re, Im)
// The inner-core part of a state file: icBT, icBP on the inner-core radii
// and icOmega (program/io.F90:430-447, 1585-1614).
struct StateIc {
    int Nic = 1;                // the run's inner-core points
    std::vector<double> ric;    // the run's inner-core radii [Nic]
    std::vector<double> f;      // [2][2][H][Nic] icBT, icBP (Re, Im)
    bool loaded = false;        // read: the file held inner-core fields of more than one point
    double omega = 0.0;         // icOmega
};
int state_read(const char* path, const lsd_config& cfg, double* coll, double* t, double* dt, StateIc* ic = nullptr);
int state_write(const char* path, const lsd_config& cfg, const double* coll, double t, double dt,
                const StateIc* ic = nullptr);

// legendre_bf16.hip
int tables_bf16(Ctx& c);
int synb_launch(Ctx& c, int kind, int njob, const void* const* in0, const void* const* in1, void* const* out0,
                void* const* out1);
int anab_launch(Ctx& c, int kind, int njob, const void* const* g0, const void* const* g1, const void* const* g2,
                void* const* o0, void* const* o1, void* const* o2);

// roundtrip.hip
int launch_cast_coll(Ctx& c, const void* src, void* dst, bool to_f32, bool from_f32);
int launch_fold(Ctx& c, const void* fourier, void* folded, bool f32);

// cfl.hip
extern const double COURANT;
int cfl_setup(Ctx& c);
int launch_cfl_phys(Ctx& c);
int launch_cfl_final(Ctx& c);
bool new_tstep(Ctx& c);

// solve.hip
int solve_block_harmonics(int N, bool mixed);
size_t solve_lds_bytes(const Ctx& c);
int launch_ops_layout(Ctx& c);
int launch_predictor(Ctx& c);
int launch_corrector(Ctx& c, double* dterr);
int launch_corrector_async(Ctx& c);
double corrector_norms(const Ctx& c);

// solve_direct.hip
bool solve_direct_supported(int N);
int launch_solve_direct(Ctx& c, bool pred);

}  // namespace lsd
