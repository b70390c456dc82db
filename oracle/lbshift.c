/* Test infrastructure, not product code: only tests/ may load it, and only as
 * the checker.
 *
 * Scalar C restatement of the FP32-shift precision form of the hot path (a
 * form the reference does not have): populations are stored as FP32
 * deviations g = n - a0 rho0 from the rest state, the collision runs in FP32
 * on those deviations, and the moment propagation stores its scattering
 * probabilities in FP32 while P, the rest probability, VACF and D_eff stay
 * FP64.  It mirrors the per-node operation order of the library's FP32-shift
 * kernels, so, built with -O2 -ffp-contract=off (IEEE single and double, no
 * FMA, no excess precision on x86-64), it reproduces their per-node results
 * bit for bit; global sums (VACF) differ only in summation order.
 *
 * The algorithm is the reference's (equilibration.f90:143-491,
 * module_collision.f90:69-108, module_moment_propagation.f90:30-289) with
 * these changes of representation:
 *   n' - a0 rho0 = (1-1/tau) g + (1/tau) (neq - a0 rho0) + (1-1/(2 tau)) G,
 *   neq - a0 rho0 = a0 drho + a1 (c.j) + a2 sum_ab j_a u_b (c_a c_b - cs^2 d_ab),
 *   drho = sum_l g_l, u = j / (rho0 + drho),
 * and s(r,l) rounded to FP32 once, every MP quantity built from the rounded
 * value (rem = 1 - sum s in FP64 keeps sum_r P conserved).
 *
 * Parity status: not a restatement of reference code, so nothing the
 * reference writes pins it directly.  It is pinned (tests/test_oracle_shift.py)
 * against lbref (itself pinned to the reference) within FP32 error bounds,
 * against the exact conservation of sum_r P by its propagator, against the
 * rest state being a fixed point, and, where the reference's own behaviour is
 * exact in any precision (a bulk lattice at rest gives D_eff = Db), against
 * that closed form.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "lbref.h"

static int32_t CV[19][3], INV[19];
static double A0[19], A1[19], A2[19];
static int ready = 0;

static void init(void) {
    if (ready) return;
    int32_t c[57];
    lbref_tables(c, A0, A1, A2, INV);
    for (int l = 0; l < 19; ++l)
        for (int a = 0; a < 3; ++a) CV[l][a] = c[3 * l + a];
    ready = 1;
}

static inline int32_t pbc(int32_t i, int32_t n) {
    if (i < 0) return i + n;
    if (i >= n) return i - n;
    return i;
}

/* the FP32-shift form: collision in FP32 */
#define REAL float
#define STORE_T float
#define FABS fabsf
#define FMAX fmaxf
#define SFX(x) lbshift_##x
/* FP32: a product added to a term, unfused; u = j / rho per component */
#define MADD(a, b, c) ((c) + (a) * (b))
#define SHIFT_VELOCITY(jx, jy, jz, full, ux, uy, uz) \
    do {                                            \
        ux = (jx) / (full);                         \
        uy = (jy) / (full);                         \
        uz = (jz) / (full);                         \
    } while (0)
#include "lbshift_eq.inc"
#undef REAL
#undef FABS
#undef FMAX
#undef SFX
#undef MADD
#undef SHIFT_VELOCITY

/* the FP32-store form: the same storage, collision in FP64 */
#define REAL double
#define FABS fabs
#define FMAX fmax
#define SFX(x) lbstore_##x
/* FP64 (the library's collide_shift<L, double>, madd): one fused multiply-add,
 * C99 fma (correctly rounded); u = j * (1 / rho) */
#define MADD(a, b, c) fma((a), (b), (c))
#define SHIFT_VELOCITY(jx, jy, jz, full, ux, uy, uz) \
    do {                                            \
        const double inv_ = 1.0 / (full);           \
        ux = (jx) * inv_;                           \
        uy = (jy) * inv_;                           \
        uz = (jz) * inv_;                           \
    } while (0)
#include "lbshift_eq.inc"
#undef REAL
#undef STORE_T
#undef FABS
#undef FMAX
#undef SFX
#undef MADD
#undef SHIFT_VELOCITY

/* Moment propagation in the FP32-shift form (probabilities stored in FP32),
 * per node in the library's gather order: P(r,t+1) = sum_l P(r+c_l,t)
 * s(r+c_l, inv l) + rem(r) P(r,t) (+ adsorption terms).  Runs nsteps steps;
 * vacf receives (nsteps+1) rows of 3 (row 0: VACF(0)); P, Pa ((3,lx,ly,lz))
 * the final fields.  Returns LBREF_RESTPART when a rest probability is below
 * epsilon (before any step, as the library and the reference do). */
int lbshift_mp(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, const int8_t* interfacial,
               const double* rho, const double* jx, const double* jy, const double* jz, const double fext[3],
               double Db, double ka, double kd, int64_t nsteps, double* P, double* Pa, double* vacf) {
    init();
    const double eps = DBL_EPSILON;
    if (ka < -eps || kd < -eps || Db <= eps) return LBREF_BAD_ARG;
    const double K = (fabs(kd) <= eps) ? 0.0 : ka / kd;
    const int ads = fabs(K) > eps;
    const double lambda = 4.0 * Db / (1.0 / 3.0);
    const int64_t N = (int64_t)lx * ly * lz;
    int64_t nf = 0, nfi = 0;
    for (int64_t r = 0; r < N; ++r) {
        nf += nature[r] == 0;
        nfi += nature[r] == 0 && interfacial[r];
    }
    const double bw = 1.0 / ((double)nf + K * (double)nfi);
    double* win = calloc((size_t)(18 * N), sizeof(double)); /* FP32-rounded, held as double */
    double* rem = calloc((size_t)N, sizeof(double));
    double* us = calloc((size_t)(3 * N), sizeof(double));   /* FP32-rounded */
    int64_t* nbr = malloc((size_t)(18 * N) * sizeof(int64_t));
    double* Pn = calloc((size_t)(3 * N), sizeof(double));
    double* Pan = calloc((size_t)(3 * N), sizeof(double));
    double* Pan2 = calloc((size_t)(3 * N), sizeof(double));
    double v0[3] = {0.0, 0.0, 0.0};
    int restpart = 0;
    memset(P, 0, (size_t)(3 * N) * sizeof(double));
    for (int32_t k = 0; k < lz; ++k)
        for (int32_t j = 0; j < ly; ++j)
            for (int32_t i = 0; i < lx; ++i) {
                const int64_t r = i + (int64_t)lx * (j + (int64_t)ly * k);
                if (nature[r] != 0) continue;
                const double rr = rho[r];
                const double vx = jx[r] + fext[0], vy = jy[r] + fext[1], vz = jz[r] + fext[2];
                double usx = 0.0, usy = 0.0, usz = 0.0, p1[3] = {0.0, 0.0, 0.0}, vp[3] = {0.0, 0.0, 0.0};
                double rv = 1.0;
                for (int l = 1; l < 19; ++l) {
                    const int32_t ip = pbc(i + CV[l][0], lx), jp = pbc(j + CV[l][1], ly), kp = pbc(k + CV[l][2], lz);
                    const int64_t q = ip + (int64_t)lx * (jp + (int64_t)ly * kp);
                    nbr[(l - 1) * N + r] = nature[q] == 0 ? q : -1;
                    if (nature[q] != 0) continue;
                    const int li = INV[l];
                    const double cx = CV[l][0], cy = CV[l][1], cz = CV[l][2];
                    const double nt = A0[l] * rr + A1[l] * ((cx * vx + cy * vy) + cz * vz);
                    const double s = (double)(float)(nt / rr - A0[l] + lambda * A0[l] * 0.5);
                    rv = rv - s;
                    if (CV[l][0] != 0) usx = usx + cx * s;
                    if (CV[l][1] != 0) usy = usy + cy * s;
                    if (CV[l][2] != 0) usz = usz + cz * s;
                    for (int a = 0; a < 3; ++a)
                        if (CV[l][a] != 0) vp[a] = vp[a] + bw * s;
                    const double rn = rho[q];
                    const double wx = jx[q] + fext[0], wy = jy[q] + fext[1], wz = jz[q] + fext[2];
                    const double ntn = A0[li] * rn + A1[li] * ((-cx * wx + -cy * wy) + -cz * wz);
                    const double sp = (double)(float)(ntn / rn - A0[li] + lambda * A0[li] * 0.5);
                    win[(l - 1) * N + r] = sp;
                    for (int a = 0; a < 3; ++a)
                        if (CV[l][a] != 0) p1[a] = p1[a] + (double)(-CV[l][a]) * sp * bw;
                }
                rem[r] = rv;
                const double rt = (ads && interfacial[r]) ? rv - ka : rv;
                if (rt < eps) restpart = 1;
                us[r] = (double)(float)usx;
                us[N + r] = (double)(float)usy;
                us[2 * N + r] = (double)(float)usz;
                for (int a = 0; a < 3; ++a) {
                    P[3 * r + a] = p1[a];
                    v0[a] = v0[a] + vp[a];
                }
            }
    for (int a = 0; a < 3; ++a) vacf[a] = v0[a];
    int status = restpart ? LBREF_RESTPART : LBREF_OK;
    for (int64_t it = 1; it <= nsteps && !restpart; ++it) {
        double v[3] = {0.0, 0.0, 0.0};
        for (int64_t r = 0; r < N; ++r) {
            if (nature[r] != 0) continue;
            double acc[3] = {0.0, 0.0, 0.0};
            for (int l = 1; l < 19; ++l) {
                const int64_t q = nbr[(l - 1) * N + r];
                if (q < 0) continue;
                const double w = win[(l - 1) * N + r];
                for (int a = 0; a < 3; ++a) acc[a] = acc[a] + P[3 * q + a] * w;
            }
            for (int a = 0; a < 3; ++a) v[a] = v[a] + P[3 * r + a] * us[a * N + r];
            if (!ads || !interfacial[r]) {
                for (int a = 0; a < 3; ++a) Pn[3 * r + a] = acc[a] + rem[r] * P[3 * r + a];
            } else {
                const double rt = rem[r] - ka, omk = 1.0 - kd;
                for (int a = 0; a < 3; ++a) {
                    const double pa = Pan[3 * r + a];
                    Pn[3 * r + a] = (acc[a] + rt * P[3 * r + a]) + pa * kd;
                    Pan2[3 * r + a] = pa * omk + P[3 * r + a] * ka;
                }
            }
        }
        memcpy(P, Pn, (size_t)(3 * N) * sizeof(double));
        if (ads) memcpy(Pan, Pan2, (size_t)(3 * N) * sizeof(double));
        for (int a = 0; a < 3; ++a) vacf[3 * it + a] = v[a];
    }
    if (Pa) memcpy(Pa, Pan, (size_t)(3 * N) * sizeof(double));
    free(win);
    free(rem);
    free(us);
    free(nbr);
    free(Pn);
    free(Pan);
    free(Pan2);
    return status;
}
