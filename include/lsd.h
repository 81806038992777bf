/*
 * lsd.h -- C-ABI of the MI355X-native per-timestep path of the Leeds
 * Spherical Dynamo code (Boussinesq, insulating inner core, Nic = 1).
 *
 * Every entry point takes plain pointers and sizes, returns 0 on success and
 * a negative LSD_E* code on failure; lsd_last_error() then holds the message
 * (the reference stops the program with the same text, e.g.
 * modules/meshs.F90:454 'mes_lu_find', modules/timestep.F90:582 'tim_invX.1').
 *
 * Layouts (all FP64 on the host side, identical to the reference's types,
 * modules/variables.F90:65-86):
 *   coll  : re[h*N + n], im[h*N + n]        Fortran Re(N, 0:H-1)
 *   spec  : re[n*H + h], im[n*H + h]        Fortran Re(0:H-1, N)
 *   phys  : p[(n*Ph + k)*Th + j]            Fortran Re(Th, 0:Ph-1, N)
 * with harmonics h in the reference's folded order
 * (Scripts/parallel.h_template:65-80): column c = 0..M/2-1 holds m = c
 * (l = m..L-1) followed by m' = M-1-c (l = m'..L-1).
 * Theta points j ascend in x = cos(theta) from -1 (modules/legendre.F90:73).
 *
 * One context = one HIP device + one stream; all fields stay resident in
 * HBM between calls.  Contexts are not shared between host threads.
 */
#ifndef LSD_H
#define LSD_H

#ifdef __cplusplus
extern "C" {
#endif

#define LSD_OK            0
#define LSD_E_CONFIG     -1   /* parameter check failed (par_precompute)        */
#define LSD_E_HIP        -2   /* a HIP runtime call failed                        */
#define LSD_E_SINGULAR   -3   /* zero pivot in a banded LU (dgbtrf info /= 0)     */
#define LSD_E_ARG        -4   /* bad argument (null pointer, unknown id)          */
#define LSD_E_DIVERGED   -5   /* corrector loop gave up (tim_check_cgce)          */

/* Precision of the step.  PLAIN keeps every operation in FP64 and is the
 * drop-in; the mixed forms keep the state, the P_l^m recursion and the LU
 * factors in FP64.  MIXED_F32_MAG64 runs the Legendre sums of B and curl B
 * in FP64 (rounded to FP32 after the sum) and everything else as
 * MIXED_F32: the FP32 rounding of the Lorentz force's inputs is the largest
 * error term of the mixed form on large grids. */
#define LSD_PLAIN_F64     0
#define LSD_MIXED_F32     1   /* FP32 transforms + FP32 solves, FP64 refinement   */
#define LSD_MIXED_BF16X3  2   /* as F32, Legendre on bf16x3 split MFMA            */
#define LSD_MIXED_F32_MAG64 3 /* as F32, B and curl B synthesised in FP64        */

/* Run parameters; names follow Scripts/parameters.F90_template:19-264. */
typedef struct lsd_config {
    int L;          /* i_L  : l in [0, L)                                       */
    int M;          /* i_M  : m index in [0, M), must be even                  */
    int Mp;         /* i_Mp : azimuthal periodicity                            */
    int N;          /* i_N  : radial points                                     */
    int Th;         /* theta points; 0 selects the reference rule 3L/2          */
    int Ph;         /* phi points; 0 selects the reference rule 3M              */
    int vel, cod, comp, mag;        /* b_*_tstep; vel must be 1                  */
    int vel_bc, cod_bc, comp_bc;    /* i_vel_bc, i_cod_bc, i_comp_bc (1 only)    */
    int precision;                  /* LSD_PLAIN_F64 / LSD_MIXED_*               */
    double rratio;  /* d_rratio                                                 */
    double Ra;      /* d_Ra                                                     */
    double Ra_comp; /* d_Ra_comp                                                */
    double Pr;      /* d_Pr                                                     */
    double Sc;      /* d_Sc                                                     */
    double Pm;      /* d_Pm                                                     */
    double E;       /* d_E                                                      */
    double dt;      /* d_timestep (> 0: fixed step; < 0: adaptive, see lsd_get_dt) */
    double implicit;/* d_implicit                                               */
    double dterr;   /* d_dterr                                                  */
    double source;  /* d_source                                                 */
    double source_comp; /* d_source_comp                                        */
    int step_control;   /* i_time_step_control: 2 (or 0, the default) or 1     */
    int Nic;            /* i_Nic: inner-core radial points; 1 (or 0) = insulating
                           inner core, > 1 = conducting inner core            */
    int rot;            /* b_rot_tstep: time-step the inner-core rotation rate */
} lsd_config;

/* Coll-field identifiers. */
#define LSD_VEL_TOR   0
#define LSD_VEL_POL   1
#define LSD_COD       2
#define LSD_COMP      3
#define LSD_MAG_TOR   4
#define LSD_MAG_POL   5
#define LSD_N_VEL_TOR 6    /* nonlinear terms of the last evaluation           */
#define LSD_N_VEL_POL 7
#define LSD_N_COD     8
#define LSD_N_COMP    9
#define LSD_N_MAG_TOR 10
#define LSD_N_MAG_POL 11
#define LSD_NUM_COLL  12

/* Physical-field identifiers (main.F90:36-95 names). */
#define LSD_P_VEL_R    0
#define LSD_P_VEL_T    1
#define LSD_P_VEL_P    2
#define LSD_P_VCURL_R  3
#define LSD_P_VCURL_T  4
#define LSD_P_VCURL_P  5
#define LSD_P_MAG_R    6
#define LSD_P_MAG_T    7
#define LSD_P_MAG_P    8
#define LSD_P_MCURL_R  9
#define LSD_P_MCURL_T  10
#define LSD_P_MCURL_P  11
#define LSD_P_COD_GR   12
#define LSD_P_COD_GT   13
#define LSD_P_COD_GP   14
#define LSD_P_COMP_GR  15
#define LSD_P_COMP_GT  16
#define LSD_P_COMP_GP  17
#define LSD_NUM_PHYS   18

/* Legendre / quadrature tables (modules/legendre.F90:41-56). */
#define LSD_T_GP         0   /* leg_gp(Th)                                      */
#define LSD_T_GW         1   /* leg_gw(Th)                                      */
#define LSD_T_P          2   /* leg_P(Th2, 0:H-1), j fastest                    */
#define LSD_T_P_         3
#define LSD_T_SIN1MP     4
#define LSD_T_LLP_       5
#define LSD_T_LLSIN1MP   6
#define LSD_T_GWP        7
#define LSD_T_GWLLP_     8
#define LSD_T_GWLLSIN1MP 9

/* Radial mesh arrays (modules/meshs.F90:74-103); band matrices are returned
 * in the reference storage M(KL+1+n-j, j), i.e. out[(j-1)*7 + (3+n-j)]. */
#define LSD_M_R        0     /* r(N)                                            */
#define LSD_M_DR1      1     /* dr(1)%M (7, N)                                  */
#define LSD_M_DR2      2     /* dr(2)%M                                         */
#define LSD_M_RADLAP   3     /* radLap%M                                        */
#define LSD_M_R1DR     4     /* r1dr%M                                          */
#define LSD_M_INTR2DR  5     /* intr2dr(N)                                      */

typedef struct lsd_step_stats {
    int    steps;          /* completed timesteps                               */
    int    evaluations;    /* nonlinear evaluations performed                   */
    int    last_its;       /* corrector iterations of the last step             */
    int    max_its;        /* max corrector iterations over the call            */
    double last_dterr;     /* tim_dterr of the last accepted corrector          */
    double t;              /* tim_t after the call                              */
    int    status;         /* LSD_OK, or the code that stopped this run         */
} lsd_step_stats;

typedef struct lsd_ctx lsd_ctx;

/* Last error message of the calling thread ("" if none). */
const char* lsd_last_error(void);

/* Harmonic count H = L*M - Mp*(M-1)*M/2 (Scripts/parameters.F90_template:255)
 * and the resolved Th, Ph of a configuration; returns LSD_E_CONFIG on a
 * failed par_precompute check (Scripts/parameters.F90_template:274-288). */
int lsd_config_sizes(const lsd_config* cfg, int* H, int* Th, int* Ph);

/* Replaces program/main.F90:initialise (main.F90:220-400) minus I/O:
 * mes_precompute, leg_precompute (FP64 recursion on the GPU), var_precompute,
 * tra_precompute, tim_precompute, *_precompute and *_matrices at cfg->dt.
 * All fields start at zero. */
int lsd_create(const lsd_config* cfg, int device, lsd_ctx** out);
void lsd_destroy(lsd_ctx* ctx);

/* Copy a coll field in (reference layout, host memory).  Replaces the field
 * fill of io_load_state / io_load_coll (program/io.F90:349-483,1278-1355). */
int lsd_set_coll(lsd_ctx* ctx, int field, const double* re, const double* im);
int lsd_get_coll(lsd_ctx* ctx, int field, double* re, double* im);
/* The inner core's magnetic field with cfg->Nic > 1: which = 0 mag_icTor,
 * 1 mag_icPol, re/im [H][Nic] on the inner-core points
 * r = r_i cos(pi (2 Nic - n) / (2 Nic - 1)), n = Nic+1 .. 2 Nic
 * (modules/meshs.F90:165-169); its last point is the inner-core boundary,
 * the outer core's first.  Replaces the icBT/icBP fill of io_load_state
 * (program/io.F90:440-447).  LSD_E_ARG for an insulating inner core. */
int lsd_set_ic(lsd_ctx* ctx, int which, const double* re, const double* im);
int lsd_get_ic(lsd_ctx* ctx, int which, double* re, double* im);
/* Inner-core rotation (modules/rotation.F90): out[0..2] = rot_omega,
 * rot_velTorq and rot_magTorq of the last nonlinear evaluation; set sets
 * rot_omega (io_load_state reads it as icOmega, program/io.F90:430-431). */
int lsd_get_rotation(lsd_ctx* ctx, double* out);
int lsd_set_rotation(lsd_ctx* ctx, double omega);
/* State files: io_save_state / io_load_state (program/io.F90:1502-1621,
 * 349-483) in the netCDF classic format (CDF-1 written; CDF-1 and CDF-2
 * read), variables uT, uP, C, Comp, BT, BP, icBT, icBP as (ReIm, H, r)
 * in the m-major io_state_nh order with attributes L, M, Mp, the radial
 * grids r and icr, dt, icOmega and the global attribute t (these two
 * calls carry no inner core: zero icBT, icBP).  Reading maps each
 * (l, m) of the run to the file's index (modes the file lacks stay zero;
 * the reference's optional d_init_pert noise is not added), interpolates
 * radially when the file's grid differs (7-point, io_interp_wts), loads C
 * into the composition when the file has no Comp, and leaves B zero for a
 * run without a magnetic field.
 *   coll: [6][2][H][N], fields uT, uP, C, Comp, BT, BP, Re then Im, each
 *         in the coll layout.  No context or GPU needed. */
int lsd_read_state(const char* path, const lsd_config* cfg, double* coll, double* t, double* dt);
int lsd_write_state(const char* path, const lsd_config* cfg, const double* coll, double t, double dt);
/* The same on a context's state: load sets the six fields and tim_t, and
 * with cfg->dt < 0 takes tim_dt from the file (refactorising the
 * operators); save writes the current fields, tim_t and tim_dt.  With a
 * conducting inner core (cfg->Nic > 1) load takes icBT, icBP from the file
 * (rows 1..Nic of the file's inner core, regridded to this run's inner-core
 * points when they differ) or, when the file has none, extends the outer
 * core's field into the inner core as io_ic_extend does
 * (program/io.F90:440-447, 1438-1477); with either Nic > 1 or cfg->rot it
 * sets rot_omega from icOmega.  Save writes the inner-core fields, radii
 * and rot_omega. */
int lsd_load_state(lsd_ctx* ctx, const char* path);
int lsd_save_state(lsd_ctx* ctx, const char* path);

/* Physical field of the last lsd_transform (reference phys layout). */
int lsd_get_phys(lsd_ctx* ctx, int field, double* p);

/* Tables and mesh arrays as built on the device, for checking. */
int lsd_get_table(lsd_ctx* ctx, int table, double* out);
int lsd_get_mesh(lsd_ctx* ctx, int which, double* out);

/* The four phases of one pass of the main loop (program/main.F90:101-157):
 *  lsd_transform   vel_transform, cod_transform, comp_transform, mag_transform
 *  lsd_nonlinear   non_velocity, non_codensity, non_composition, non_magnetic
 *  lsd_predictor   *_predictor (tim_it == 0)
 *  lsd_corrector   *_corrector; *dterr receives tim_dterr (tim_addcorr sums)
 * The predictor opens a step (tim_it = 1) and each corrector advances the
 * context's tim_it by the tim_check_cgce rule of modules/timestep.F90:791-827
 * (the caller's own check still decides whether the run stops), so the
 * context knows when a step is in progress (see lsd_set_dt). */
int lsd_transform(lsd_ctx* ctx);
int lsd_nonlinear(lsd_ctx* ctx);
int lsd_predictor(lsd_ctx* ctx);
int lsd_corrector(lsd_ctx* ctx, double* dterr);

/* nsteps full timesteps of program/main.F90:100-164: one predictor and
 * correctors until tim_check_cgce (modules/timestep.F90:744-830, control
 * cfg->step_control) accepts; with cfg->dt < 0 a new dt is chosen at the
 * start of every step (see lsd_get_dt). */
int lsd_step(lsd_ctx* ctx, int nsteps, lsd_step_stats* stats);

/* Ensembles: nsteps timesteps of each of n independent contexts (any mix of
 * truncations and parameters), with every member's step graph launched on
 * its own stream before any is waited for, so the runs execute
 * concurrently.  Each member follows the same control as lsd_step and
 * gives the same results as stepping it alone.  A member that
 * tim_check_cgce stops (or whose launch fails) is left where the reference
 * would stop it and takes no further steps; the others carry on to nsteps.
 * Every member's stream is drained before the call returns.  stats[n] (may
 * be NULL) receives each member's counts and status; the return code is
 * that of the first failing member, and lsd_last_error() names every
 * failing member.  No counterpart in the reference (one run per program
 * there). */
int lsd_step_many(lsd_ctx** ctxs, int n, int nsteps, lsd_step_stats* stats);

/* Kinetic (which = 0) or magnetic (which = 1) energy of the current state
 * and its l- and m-spectra, as io_write_vel_energy / io_write_mag_energy
 * compute them (program/io.F90:1799-1884; var_coll_norm, var_coll_normlm,
 * modules/variables.F90:901-957, 1267-1313): E = Efac * sum over the q, s,
 * t parts of int |a|^2 dV, Efac = 1/2 (velocity) or Pm/E (magnetic,
 * program/io.F90:238-247).  El[L] and Em[M] (may be NULL) receive the
 * per-degree and per-order parts, times Efac. */
int lsd_energy(lsd_ctx* ctx, int which, double* E, double* El, double* Em);

/* lsd_energy without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The per-harmonic parts and sums live in one static device
 * array per GPU (1.6 MB, held until the process ends), so H <= 65536 and
 * L + M <= 2048 (LSD_E_ARG beyond).  Results equal lsd_energy's bit for
 * bit.  All contexts on one GPU share the array: calls that overlap on one
 * GPU (from several host threads) race. */
int lsd_energy_noalloc(lsd_ctx* ctx, int which, double* E, double* El, double* Em);

/* Whether lsd_step replays the timestep from captured HIP graphs (default 1)
 * or launches every kernel from the host (0).  No counterpart in the
 * reference; results are identical either way. */
int lsd_set_graph(lsd_ctx* ctx, int on);

/* Refactorise all banded operators for a new dt (vel_matrices, cod_matrices,
 * comp_matrices, mag_matrices; program/main.F90:126-135).  Only between
 * steps (tim_it == 0), as the reference rebuilds them; LSD_E_ARG inside a
 * step. */
int lsd_set_dt(lsd_ctx* ctx, double dt);

/* tim_dt, tim_cfl_dt and tim_corr_dt (modules/timestep.F90:13-18): the
 * current step, the last CFL limit of non_maxtstep and the corrector's
 * estimate of tim_check_cgce.  With cfg->dt < 0 (d_timestep < 0) lsd_step
 * chooses dt at the start of every step as program/main.F90:124-135 does
 * (non_maxtstep, tim_new_tstep, *_matrices on a change), starting from
 * |cfg->dt|, which stands for the dt the reference reads from the state
 * file (program/io.F90:429). */
int lsd_get_dt(lsd_ctx* ctx, double* dt, double* cfl_dt, double* corr_dt);

/* Drop-in transforms with host arrays in the reference spec/phys layouts
 * (modules/transform.fftw3.F90:472,527,719,825,955). */
int lsd_tra_spec2phys(lsd_ctx* ctx, const double* s_re, const double* s_im, double* p);
int lsd_tra_phys2spec(lsd_ctx* ctx, const double* p, double* s_re, double* s_im);
int lsd_tra_grad(lsd_ctx* ctx, const double* s_re, const double* s_im, double* t, double* p);
int lsd_tra_qst2rtp(lsd_ctx* ctx,
                    const double* q_re, const double* q_im,
                    const double* s_re, const double* s_im,
                    const double* t_re, const double* t_im,
                    double* r, double* th, double* ph);
int lsd_tra_rtp2qst(lsd_ctx* ctx, const double* r, const double* th, const double* ph,
                    double* q_re, double* q_im, double* s_re, double* s_im,
                    double* t_re, double* t_im);

/* Synthesis -> analysis round trip of the Legendre transforms in the
 * context's precision (FP64, FP32 or BF16x3), at the Fourier level, where the
 * phi FFTs of a round trip cancel: kind 0 takes one scalar field (tra_spec2phys
 * then tra_phys2spec), kind 1 the (q, s, t) of a vector field (tra_qst2rtp
 * then tra_rtp2qst).  in / out: [nfield][2][H][N], Re then Im, coll layout.
 * For measuring the transform kernels in isolation; no counterpart in the
 * reference. */
int lsd_sht_roundtrip(lsd_ctx* ctx, int kind, const double* in, double* out);

/* Wait for all work queued on the context's stream (get_time bracket,
 * modules/dyn_mpi.F90:81-92). */
int lsd_sync(lsd_ctx* ctx);

/* Kernel variants (no counterpart in the reference): launch shapes and
 * kernel forms that compute the same step, chosen per context.  Defaults
 * come from the tuning cache (lsd_tune) when one is found, else the values
 * given here.  Every variant gives the same results, bit for bit for the
 * synthesis shapes and to rounding for the analysis and solve forms.
 *  SYN_*     FP32 synthesis: m per workgroup, 4 or 8 (one wave per m, one
 *            hemisphere of one output staged in LDS per pass), or 5, 9:
 *            the same shapes as persistent kernels (the CUs' worth of
 *            workgroups pulling the work items from per-XCD queues), or
 *            6, 10: persistent, each workgroup issuing its next item's
 *            loads before the current item's epilogue;
 *            defaults 4, 8, 8
 *  ANA       FP32 analysis: 0 wide-load (even Th/2 only; default), 1 LDS-staged,
 *            2 the wide-load kernel in the persistent form
 *  SOLVE     banded solves: 0 two-pass register-column FP64 (N in {32, 48,
 *            64, 96}; default there), 1 LDS-column (the mixed forms: FP32
 *            factors with one FP64-residual refinement sweep), 2 LDS-column
 *            with FP64 factors in the mixed forms (their default for other
 *            N) */
#define LSD_V_SYN_SCALAR 0
#define LSD_V_SYN_VECTOR 1
#define LSD_V_SYN_GRAD   2
#define LSD_V_ANA        3
#define LSD_V_SOLVE      4
#define LSD_V_COUNT      5
int lsd_set_variant(lsd_ctx* ctx, int which, int value);
int lsd_get_variant(lsd_ctx* ctx, int which, int* value);

/* Autotuner: times every variant of this context's sizes on the GPU (the
 * synthesis and analysis launches on the current buffers, reps times each;
 * the solve forms over whole timesteps from a saved copy of the state,
 * which is restored afterwards), selects the fastest of each, and, if
 * cache_path is non-NULL, records the choice under this context's sizes
 * and precision in that cache file (created if absent).  lsd_create reads
 * the cache named by the environment variable LSD_TUNING_CACHE, else
 * tuning_cache.jsonl beside the library, when present.  times_us[LSD_V_COUNT
 * * 2] (may be NULL) receives the time of variant 0 and 1 of each group
 * (0 where a variant does not apply; for the synthesis the shapes 4 and 8,
 * after which the persistent form of the faster is timed and kept if it is
 * faster still). */
int lsd_tune(lsd_ctx* ctx, int reps, const char* cache_path, double* times_us);

/* Kernel classes for per-launch timing. */
#define LSD_K_SYNTH    0   /* Legendre synthesis                                */
#define LSD_K_ANAL     1   /* Legendre analysis                                 */
#define LSD_K_FFT_INV  2   /* phi FFT, Fourier -> physical                      */
#define LSD_K_FFT_FWD  3   /* phi FFT, physical -> folded Fourier               */
#define LSD_K_NONLIN   4   /* pointwise nonlinear products                      */
#define LSD_K_SPEC     5   /* radial stencils on spectral fields                */
#define LSD_K_SOLVE    6   /* banded predictor / corrector solves               */
#define LSD_K_PHYS     7   /* fused inverse FFT + products + forward FFT (step) */
#define LSD_K_COUNT    8

/* Timing hooks (no counterpart in the reference beyond get_time): between
 * begin and end, every launch of the given class is bracketed by HIP events
 * on the context's stream; end returns the summed kernel time, the launch
 * count and the algorithmic work of those launches (flops for SYNTH/ANAL,
 * HBM bytes for the other classes). */
int lsd_profile_begin(lsd_ctx* ctx, int kernel_class);
int lsd_profile_end(lsd_ctx* ctx, double* total_ms, int* launches, double* work);

/* ------------------------------------------------------------------------
 * Learned closure network (no counterpart in the reference): a 3-layer MLP
 * n_in -> width -> width -> n_out with GELU, BF16 weights and activations
 * on the matrix cores with FP32 accumulation, FP32 master weights trained
 * by Adam on a mean-squared-error loss.  Parameters, FP32 and contiguous:
 * W1[width][n_in] b1[width] W2[width][width] b2[width] W3[n_out][width]
 * b3[n_out].  Rows of x / y / target are contiguous, host memory.
 * ---------------------------------------------------------------------- */
typedef struct lsd_mlp lsd_mlp;
int  lsd_mlp_create(int n_in, int width, int n_out, unsigned long long seed, lsd_mlp** out);
void lsd_mlp_destroy(lsd_mlp* m);
long lsd_mlp_num_params(const lsd_mlp* m);
int  lsd_mlp_get_params(lsd_mlp* m, float* params);
int  lsd_mlp_set_params(lsd_mlp* m, const float* params);
/* y[rows][n_out] = MLP(x[rows][n_in]) */
int  lsd_mlp_forward(lsd_mlp* m, long rows, const float* x, float* y);
/* loss = mean (MLP(x) - target)^2 and its gradient w.r.t. the parameters
 * (grad may be NULL) */
int  lsd_mlp_grad(lsd_mlp* m, long rows, const float* x, const float* target, float* loss, float* grad);
/* `steps` Adam steps (lr, beta 0.9 / 0.999) on consecutive minibatches of
 * `batch` rows; loss[steps] (may be NULL) receives the minibatch losses */
int  lsd_mlp_train(lsd_mlp* m, long rows, const float* x, const float* target, long batch, int steps, float lr,
                   float* loss);

#ifdef __cplusplus
}
#endif
#endif /* LSD_H */
