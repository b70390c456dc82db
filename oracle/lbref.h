/* Test infrastructure, not product code.
 *
 * lbref: scalar C restatement of laboetie's lattice-Boltzmann + moment-
 * propagation path (maxlevesque/laboetie, Fortran).  It is the checker the
 * GPU tests compare the HIP library against, on machines where the reference
 * sources are not available; its own correctness is pinned against fixtures
 * written by the reference program (tests/golden) and against closed forms.
 *
 * Layouts follow the reference exactly (Fortran column-major, x fastest):
 *   scalar field  f[i + lx*(j + ly*k)]               (lx,ly,lz)
 *   population    n[node + N*l]                      (lx,ly,lz,19)   LB phase
 *   tracer pop.   nt[l + 19*node]                    (19,lx,ly,lz)   MP phase
 *   moment field  P[a + 3*node]                      (3,lx,ly,lz)
 * with N = lx*ly*lz, l = 0..18 standing for the reference's l = 1..19.
 * nature: 0 fluid, 1 solid (module_system.f90:32).
 */
#ifndef LBREF_H
#define LBREF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes shared with the HIP library's C-ABI */
#define LBREF_OK          0
#define LBREF_NEG_POP     1   /* equilibration.f90:248 */
#define LBREF_RESTPART    2   /* module_moment_propagation.f90:249,257 */
#define LBREF_BAD_TAU     3   /* module_collision.f90:39 */
#define LBREF_BAD_ARG     5

/* exit reasons of lbref_equilibrate */
#define LBREF_EXIT_CONVERGED 1  /* second convergence, equilibration.f90:373-374 */
#define LBREF_EXIT_CAP       2  /* step cap reached (extension; the reference has none) */
#define LBREF_EXIT_FLOOR     3  /* FP32-shift form: l2err stopped improving (lbshift_equilibrate) */

void lbref_tables(int32_t c[19 * 3], double a0[19], double a1[19], double a2[19], int32_t inv[19]);

double lbref_norm2(const double* v, int n);
void lbref_norm2_rows(const double* v, int n, int64_t count, double* out);

int lbref_geometry(int32_t label, int32_t lx, int32_t ly, int32_t lz,
                   const int32_t* solid_ijk, int64_t npts, int8_t* nature);
void lbref_interfacial(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, int8_t* interfacial);

int lbref_collide(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, double tau,
                  double* n, const double* rho, const double* jx, const double* jy, const double* jz,
                  const double* fx, const double* fy, const double* fz);

int lbref_equilibrate(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, double tau,
                      double rho0, const double fext[3], double target_error, int64_t max_steps,
                      double* n, double* rho, double* jx, double* jy, double* jz,
                      double* l2err_trace, int64_t trace_cap, int64_t* steps_done, int32_t* exit_reason);

void lbref_tracer_population(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature,
                             const double* rho, const double* jx, const double* jy, const double* jz,
                             const double fext[3], double* nt);

int lbref_mp_init(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, const int8_t* interfacial,
                  const double* nt, const double* rho, double Db, double ka, double kd,
                  double* P, double vacf0[3], int32_t* consider_adsorption);

int lbref_mp_step(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, const int8_t* interfacial,
                  const double* nt, const double* rho, double Db, double ka, double kd,
                  int32_t consider_adsorption,
                  const double* P_now, const double* Pa_now, double* P_next, double* Pa_next,
                  double vacf[3]);

int lbref_mp_converged(int64_t it, const double vacf[3], int32_t lx, int32_t ly, int32_t lz, double Db);

void lbref_deff(const double* vacf, int64_t nsteps, double deff[3]);

/* lbshift.c: the FP32-shift precision form (not in the reference), in the
 * per-node operation order of the library's FP32-shift kernels.  check_every:
 * stage 2 measures l2err = max|j(t) - j(t - check_every)| / check_every at the
 * steps switch + m check_every only (a power of two; 1: every step); the
 * trace holds NaN at the other steps. */
int lbshift_equilibrate(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, double tau, double rho0,
                        const double fext[3], double target_error, int64_t max_steps, int64_t floor_window,
                        int64_t check_every, double* n, double* rho, double* jx, double* jy, double* jz, double* l2err_trace,
                        int64_t trace_cap, int64_t* steps_done, int32_t* exit_reason);

/* The FP32-store form: the same FP32 deviation storage, collision and
 * j_old in FP64 (the library's LBX_PREC_FP32_STORE). */
int lbstore_equilibrate(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, double tau, double rho0,
                        const double fext[3], double target_error, int64_t max_steps, int64_t floor_window,
                        int64_t check_every, double* n, double* rho, double* jx, double* jy, double* jz, double* l2err_trace,
                        int64_t trace_cap, int64_t* steps_done, int32_t* exit_reason);

int lbshift_mp(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, const int8_t* interfacial,
               const double* rho, const double* jx, const double* jy, const double* jz, const double fext[3],
               double Db, double ka, double kd, int64_t nsteps, double* P, double* Pa, double* vacf);

#ifdef __cplusplus
}
#endif
#endif
