/*
 * lsd_oracle.h -- CPU FP64 oracle of the Leeds Spherical Dynamo per-timestep
 * path.  TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, as the checker.  The product
 * library never links or calls it.
 *
 * Layouts are the reference's (see include/lsd.h): coll re[h*N+n],
 * spec re[n*H+h], phys p[(n*Ph+k)*Th+j].
 */
#ifndef LSD_ORACLE_H
#define LSD_ORACLE_H

#include "../include/lsd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

void orc_set_threads(int n);
int  orc_get_threads(void);
const char* orc_last_error(void);
int  orc_sizes(const lsd_config* cfg, int* H, int* Th, int* Ph);

/* auxiliary/gauwts.f, auxiliary/schnla.f */
void orc_gauwts(double x1, double x2, double* gaux, double* gauw, int n);
void orc_schnla(double* pa, double* dpa, double x, int lh);

/* Harmonic order of Scripts/parallel.h_template:65-80 (l, m_ per nh). */
int  orc_harmonics(int L, int M, int Mp, int* l, int* m);

/* Banded LU with partial pivoting and its solve, LAPACK dgbtf2/dgbtrs
 * semantics for KL = KU = 3, ab(10, n) column-major, 1-based ipiv. */
int  orc_gbtrf(int n, double* ab, int* ipiv);
void orc_gbtrs(int n, const double* ab, const int* ipiv, double* b);

/* Complex DFT (sign -1 forward, +1 backward, unnormalised). */
void orc_cfft(int n, double* re, double* im, int sign);

orc_sim* orc_create(const lsd_config* cfg);
void orc_destroy(orc_sim* s);
int  orc_get_table(orc_sim* s, int table, double* out);
int  orc_get_mesh(orc_sim* s, int which, double* out);
int  orc_set_coll(orc_sim* s, int field, const double* re, const double* im);
int  orc_get_coll(orc_sim* s, int field, double* re, double* im);
int  orc_get_phys(orc_sim* s, int field, double* p);
int  orc_transform(orc_sim* s);
int  orc_nonlinear(orc_sim* s);
int  orc_predictor(orc_sim* s);
int  orc_corrector(orc_sim* s, double* dterr);
int  orc_step(orc_sim* s, int nsteps, lsd_step_stats* st);
int  orc_set_dt(orc_sim* s, double dt);
/* tim_dt, tim_cfl_dt and tim_corr_dt (modules/timestep.F90:13-18). */
int  orc_get_dt(orc_sim* s, double* dt, double* cfl_dt, double* corr_dt);

int  orc_tra_spec2phys(orc_sim* s, const double* sre, const double* sim, double* p);
int  orc_tra_phys2spec(orc_sim* s, const double* p, double* sre, double* sim);
int  orc_tra_grad(orc_sim* s, const double* sre, const double* sim, double* t, double* p);
int  orc_tra_qst2rtp(orc_sim* s, const double* qre, const double* qim,
                     const double* sre, const double* sim,
                     const double* tre, const double* tim,
                     double* r, double* th, double* ph);
int  orc_tra_rtp2qst(orc_sim* s, const double* r, const double* th, const double* ph,
                     double* qre, double* qim, double* sre, double* sim,
                     double* tre, double* tim);

/* Conducting inner core (cfg.Nic > 1) and inner-core rotation (cfg.rot):
 * the inner-core fields icBT (0) and icBP (1) as re/im[h*Nic + n], the
 * rotation rate and the torques of the last evaluation
 * (out[0..2] = rot_omega, rot_velTorq, rot_magTorq), the inner-core radii. */
int  orc_set_ic(orc_sim* s, int k, const double* re, const double* im);
int  orc_get_ic(orc_sim* s, int k, double* re, double* im);
int  orc_get_rot(orc_sim* s, double* out);
int  orc_set_rot(orc_sim* s, double omega);
int  orc_get_icr(orc_sim* s, double* out);

/* Energy diagnostics, var_coll_norm (modules/variables.F90:903-957):
 * out[0..3] = Eodd, Eeven, Em0le, Em0 of one coll field. */
int  orc_coll_norm(orc_sim* s, int field, double* out);
int  orc_energy(orc_sim* s, int which, double* E, double* El, double* Em);

#ifdef __cplusplus
}
#endif
#endif
