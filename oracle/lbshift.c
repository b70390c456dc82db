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
static double A0[19], A1[19];
static float W0F[19], W1F[19], W2F[19], KF[19][3]; /* FP32 weights, c_a^2 - cs^2 */
static int ready = 0;

static void init(void) {
    if (ready) return;
    int32_t c[57];
    double a2[19];
    lbref_tables(c, A0, A1, a2, INV);
    for (int l = 0; l < 19; ++l) {
        for (int a = 0; a < 3; ++a) {
            CV[l][a] = c[3 * l + a];
            KF[l][a] = (float)((double)(CV[l][a] * CV[l][a]) - 1.0 / 3.0);
        }
        W0F[l] = (float)A0[l];
        W1F[l] = (float)A1[l];
        W2F[l] = (float)a2[l];
    }
    ready = 1;
}

static inline int32_t pbc(int32_t i, int32_t n) {
    if (i < 0) return i + n;
    if (i >= n) return i - n;
    return i;
}

/* the FP32-shift collision of one population (library: collide_shift) */
static float collide_shift(int l, float g, float drho, float jx, float jy, float jz, float ux, float uy, float uz,
                           float fx, float fy, float fz, float A, float B, float C) {
    const int cx = CV[l][0], cy = CV[l][1], cz = CV[l][2];
    const float w0 = W0F[l], w1 = W1F[l], w2 = W2F[l];
    const float Kx = KF[l][0], Ky = KF[l][1], Kz = KF[l][2];
    float q = (jx * ux) * Kx;
    if (cx * cy != 0) q = q + (float)(cx * cy) * (jx * uy);
    if (cx * cz != 0) q = q + (float)(cx * cz) * (jx * uz);
    if (cy * cx != 0) q = q + (float)(cy * cx) * (jy * ux);
    q = q + (jy * uy) * Ky;
    if (cy * cz != 0) q = q + (float)(cy * cz) * (jy * uz);
    if (cz * cx != 0) q = q + (float)(cz * cx) * (jz * ux);
    if (cz * cy != 0) q = q + (float)(cz * cy) * (jz * uy);
    q = q + (jz * uz) * Kz;
    float dneq = w0 * drho;
    if (l != 0) {
        /* c.j with zero terms dropped, in x, y, z order */
        float d = 0.0f;
        int first = 1;
        const float jv[3] = {jx, jy, jz};
        const int cv[3] = {cx, cy, cz};
        for (int a = 0; a < 3; ++a)
            if (cv[a] != 0) {
                d = first ? (float)cv[a] * jv[a] : d + (float)cv[a] * jv[a];
                first = 0;
            }
        dneq = dneq + w1 * d;
    }
    dneq = dneq + w2 * q;
    float G = w1 * ((((float)cx - ux) * fx + ((float)cy - uy) * fy) + ((float)cz - uz) * fz);
    if (l != 0) {
        const float uv[3] = {ux, uy, uz}, fv[3] = {fx, fy, fz};
        const int cv[3] = {cx, cy, cz};
        float du = 0.0f, df = 0.0f;
        int first = 1;
        for (int a = 0; a < 3; ++a)
            if (cv[a] != 0) {
                du = first ? (float)cv[a] * uv[a] : du + (float)cv[a] * uv[a];
                df = first ? (float)cv[a] * fv[a] : df + (float)cv[a] * fv[a];
                first = 0;
            }
        G = G + ((2.0f * w2) * du) * df;
    }
    return (A * g + B * dneq) + C * G;
}

/* The equilibration loop in the FP32-shift form, from the rest state of
 * density rho0 (g = 0, j = 0).  Same stage logic as lbref_equilibrate; the
 * l2err of a step is max|j - j_old| of the FP32 fluxes.  floor_window > 0:
 * a stage also ends when l2err has set no new minimum for that many steps
 * after falling 100-fold below the stage's largest value (the FP32 rounding
 * floor; a stage-2 floor exits with LBREF_EXIT_FLOOR).  On return n (FP64,
 * (lx,ly,lz,19)) = a0 rho0 + g of the post-streaming state of the last step
 * (0 at solids), rho = rho0 + sum g and j = F/2 + sum g c, both summed in
 * FP64 from the FP32 deviations. */
int lbshift_equilibrate(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, double tau, double rho0,
                        const double fext[3], double target_error, int64_t max_steps, int64_t floor_window,
                        double* n, double* rho, double* jx, double* jy, double* jz, double* l2err_trace,
                        int64_t trace_cap, int64_t* steps_done, int32_t* exit_reason) {
    init();
    if (tau < 0.5) return LBREF_BAD_TAU;
    const int64_t N = (int64_t)lx * ly * lz;
    float* gc = calloc((size_t)(19 * N), sizeof(float)); /* post-collision */
    float* gs = calloc((size_t)(19 * N), sizeof(float)); /* post-streaming */
    float* j32 = calloc((size_t)(3 * N), sizeof(float));
    const float A = (float)(1.0 - 1.0 / tau), B = (float)(1.0 / tau), C = (float)(1.0 - 1.0 / (2.0 * tau));
    const float r0 = (float)rho0;
    int status = LBREF_OK;
    *exit_reason = 0;
    *steps_done = 0;
    /* opening collision of the rest state (drho = 0, j = 0, no force) */
    for (int64_t r = 0; r < N; ++r) {
        if (nature[r] != 0) continue;
        const float drho = (float)(rho0 - rho0), full = r0 + drho;
        for (int l = 0; l < 19; ++l)
            gc[r + N * l] = collide_shift(l, 0.0f, drho, 0.0f, 0.0f, 0.0f, 0.0f / full, 0.0f / full, 0.0f / full,
                                          0.0f, 0.0f, 0.0f, A, B, C);
    }
    int conv_without = 0;
    int64_t switch_step = INT64_MAX;
    int64_t check_after = 2;              /* stop tests apply to steps after this one */
    double rec_min = HUGE_VAL;            /* floor test: the stage's smallest l2err so far, */
    int64_t rec_step = check_after;       /* the step that set it */
    double rec_max = 0.0;                 /* and the stage's largest */
    double F[3] = {0.0, 0.0, 0.0}; /* force of the current step */
    int64_t t;
    for (t = 1;; ++t) {
        const float fmx = (float)(F[0] / 2.0), fmy = (float)(F[1] / 2.0), fmz = (float)(F[2] / 2.0);
        double l2 = 0.0;
        int neg = 0;
        /* pull-streaming with half-way bounce-back, moments, l2err */
        for (int32_t k = 0; k < lz; ++k)
            for (int32_t j = 0; j < ly; ++j)
                for (int32_t i = 0; i < lx; ++i) {
                    const int64_t r = i + (int64_t)lx * (j + (int64_t)ly * k);
                    if (nature[r] != 0) continue;
                    float g[19];
                    for (int l = 0; l < 19; ++l) {
                        const int32_t ip = pbc(i - CV[l][0], lx), jp = pbc(j - CV[l][1], ly),
                                      kp = pbc(k - CV[l][2], lz);
                        const int64_t rp = ip + (int64_t)lx * (jp + (int64_t)ly * kp);
                        g[l] = nature[rp] == 0 ? gc[rp + N * l] : gc[r + N * INV[l]];
                        gs[r + N * l] = g[l];
                        if (g[l] < -((float)A0[l] * r0)) neg = 1;
                    }
                    float drho = g[0];
                    for (int l = 1; l < 19; ++l) drho = drho + g[l];
                    float ax = fmx, ay = fmy, az = fmz;
                    for (int l = 0; l < 19; ++l) {
                        if (CV[l][0] != 0) ax = ax + (float)CV[l][0] * g[l];
                        if (CV[l][1] != 0) ay = ay + (float)CV[l][1] * g[l];
                        if (CV[l][2] != 0) az = az + (float)CV[l][2] * g[l];
                    }
                    const float dx = fabsf(ax - j32[r]), dy = fabsf(ay - j32[N + r]), dz = fabsf(az - j32[2 * N + r]);
                    const double m = (double)fmaxf(fmaxf(dx, dy), dz);
                    if (m > l2) l2 = m;
                    j32[r] = ax;
                    j32[N + r] = ay;
                    j32[2 * N + r] = az;
                }
        *steps_done = t;
        if (neg) {
            status = LBREF_NEG_POP;
            break;
        }
        if (l2err_trace && t <= trace_cap) l2err_trace[t - 1] = l2;
        int stop = 0, floor = 0;
        if (floor_window > 0 && t > check_after) {
            if (l2 < rec_min) {
                rec_min = l2;
                rec_step = t;
            }
            rec_max = fmax(rec_max, l2);
            floor = t - rec_step >= floor_window && rec_min <= 1e-2 * rec_max;
        }
        if (t > check_after && (l2 <= target_error || floor)) {
            if (!conv_without) {
                conv_without = 1;
                switch_step = t;
                check_after = t;
                rec_min = HUGE_VAL;
                rec_step = t;
                rec_max = 0.0;
                F[0] = fext[0];
                F[1] = fext[1];
                F[2] = fext[2];
            } else {
                *exit_reason = l2 <= target_error ? LBREF_EXIT_CONVERGED : LBREF_EXIT_FLOOR;
                stop = 1;
            }
        }
        if (!stop && max_steps > 0 && t >= max_steps) {
            *exit_reason = LBREF_EXIT_CAP;
            stop = 1;
        }
        if (stop) break;
        /* the collision that opens step t+1, with the force of step t+1 */
        const float fcx = (float)F[0], fcy = (float)F[1], fcz = (float)F[2];
        for (int64_t r = 0; r < N; ++r) {
            if (nature[r] != 0) continue;
            float drho = gs[r];
            for (int l = 1; l < 19; ++l) drho = drho + gs[r + N * l];
            const float ax = j32[r], ay = j32[N + r], az = j32[2 * N + r];
            const float full = r0 + drho;
            const float ux = ax / full, uy = ay / full, uz = az / full;
            for (int l = 0; l < 19; ++l)
                gc[r + N * l] = collide_shift(l, gs[r + N * l], drho, ax, ay, az, ux, uy, uz, fcx, fcy, fcz, A, B, C);
        }
    }
    /* FP64 output state of step t */
    const double fm[3] = {t > switch_step ? fext[0] / 2.0 : 0.0, t > switch_step ? fext[1] / 2.0 : 0.0,
                          t > switch_step ? fext[2] / 2.0 : 0.0};
    for (int64_t r = 0; r < N; ++r) {
        if (nature[r] != 0) {
            for (int l = 0; l < 19; ++l) n[r + N * l] = 0.0;
            rho[r] = 0.0;
            jx[r] = jy[r] = jz[r] = 0.0;
            continue;
        }
        double drho = 0.0, ax = fm[0], ay = fm[1], az = fm[2];
        for (int l = 0; l < 19; ++l) {
            const double d = (double)gs[r + N * l];
            n[r + N * l] = A0[l] * rho0 + d;
            drho = l == 0 ? d : drho + d;
            if (CV[l][0] != 0) ax = ax + (double)CV[l][0] * d;
            if (CV[l][1] != 0) ay = ay + (double)CV[l][1] * d;
            if (CV[l][2] != 0) az = az + (double)CV[l][2] * d;
        }
        rho[r] = rho0 + drho;
        jx[r] = ax;
        jy[r] = ay;
        jz[r] = az;
    }
    free(gc);
    free(gs);
    free(j32);
    return status;
}

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
