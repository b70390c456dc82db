/* Test infrastructure, not product code: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker.
 *
 * Scalar C restatement of laboetie's lattice-Boltzmann + moment-propagation
 * path (maxlevesque/laboetie).  Every function cites the reference lines it
 * restates and keeps the reference's operation order, so that, compiled with
 * -O2 -ffp-contract=off, it reproduces the reference program's IEEE results
 * bit for bit (checked against fixtures the reference writes, tests/golden).
 *
 * Parity status per function (what pins it):
 *   lbref_tables           reference dump of lbm%vel (ref_dump tables mode)
 *   lbref_norm2            bit-for-bit comparison with the compiler's NORM2 on
 *                          grids of geometry inputs (tests/test_oracle_pins.py)
 *   lbref_geometry         reference node%nature dumps (labels 0-3)
 *   lbref_interfacial      reference node%isInterfacial dumps
 *   lbref_collide          reference collide() on random inputs
 *   lbref_equilibrate      reference n/rho/j dumps + the full l2err.dat trace
 *   lbref_tracer_population reference tracer population dump
 *   lbref_mp_init/step     reference vacf.dat (every step, 1 OpenMP thread)
 *                          + probability conservation + bulk closed form
 *   lbref_deff             not in the reference (drop_tracers.f90:53-54 is
 *                          commented out): pinned by the bulk closed form D = Db
 *                          and by brute-force summation.
 */
#include "lbref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

/* D3Q19 velocity set, module_lbmodel.f90:66-86 (l = 1..19 -> 0..18) */
static const int32_t CV[19][3] = {
    {0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
    {1, 1, 0},  {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1}, {-1, 0, 1}, {1, 0, -1},
    {-1, 0, -1}, {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};

static double A0[19], A1[19], A2[19];
static int32_t INV[19];
static int tables_ready = 0;

/* module_lbmodel.f90:122-152 (weights; the parameters are folded in double)
 * and :156-162 (inverse velocity). */
static void init_tables(void) {
    if (tables_ready) return;
    const double csq = 1.0 / 3.0;
    const double a_00 = 1.0 / 3.0, a_01 = 1.0 / 18.0, a_02 = 1.0 / 36.0;
    const double a_10 = a_00 / csq, a_11 = a_01 / csq, a_12 = a_02 / csq;
    const double a_20 = a_00 / (2 * (csq * csq)), a_21 = a_01 / (2 * (csq * csq)),
                 a_22 = a_02 / (2 * (csq * csq));
    for (int l = 0; l < 19; ++l) {
        int cls = (l == 0) ? 0 : (l <= 6 ? 1 : 2);
        A0[l] = cls == 0 ? a_00 : (cls == 1 ? a_01 : a_02);
        A1[l] = cls == 0 ? a_10 : (cls == 1 ? a_11 : a_12);
        A2[l] = cls == 0 ? a_20 : (cls == 1 ? a_21 : a_22);
        for (int m = 0; m < 19; ++m)
            if (CV[m][0] == -CV[l][0] && CV[m][1] == -CV[l][1] && CV[m][2] == -CV[l][2]) INV[l] = m;
    }
    tables_ready = 1;
}

void lbref_tables(int32_t c[19 * 3], double a0[19], double a1[19], double a2[19], int32_t inv[19]) {
    init_tables();
    for (int l = 0; l < 19; ++l) {
        for (int a = 0; a < 3; ++a) c[3 * l + a] = CV[l][a];
        a0[l] = A0[l];
        a1[l] = A1[l];
        a2[l] = A2[l];
        inv[l] = INV[l];
    }
}

/* Periodic index, module_system.f90:99-112, 0-based: only a +-1 step wraps. */
static inline int32_t pbc(int32_t i, int32_t n) {
    if (i == -1) return n - 1;
    if (i == n) return 0;
    return i;
}

/* The Fortran NORM2 intrinsic as the flang runtime evaluates it: a scaled sum
 * of squares, max*sqrt(1 + sum((x/max)^2)) with the running rescale when a
 * larger element arrives.  Reproduced from black-box comparison with flang's
 * NORM2 (identical on the grids of tests/test_oracle_pins.py); the plain sqrt(sum x^2)
 * differs in the last bit often enough to flip BCC solid/fluid decisions. */
double lbref_norm2(const double* v, int n) {
    double mx = 0.0, s = 0.0;
    for (int i = 0; i < n; ++i) {
        double a = fabs(v[i]);
        if (mx == 0.0) {
            mx = a;
        } else if (a > mx) {
            double t = mx / a;
            double tsq = t * t;
            s = s * tsq;
            s = s + tsq;
            mx = a;
        } else {
            double t = a / mx;
            s = s + t * t;
        }
    }
    return mx * sqrt(1.0 + s);
}

/* supercell_definition.f90:48-90 with the builders of module_geometry.f90:
 * label -1 all fluid; 0 construct_custom (:380-426, solid_ijk 1-based triples);
 * 1 construct_slit (:158-166); 2 construct_cylinder (:253-277);
 * 3 construct_cc (:206-245). Returns LBREF_BAD_ARG on the reference's stops. */
int lbref_geometry(int32_t label, int32_t lx, int32_t ly, int32_t lz,
                   const int32_t* solid_ijk, int64_t npts, int8_t* nature) {
    if (lx <= 0 || ly <= 0 || lz <= 0) return LBREF_BAD_ARG;
    const int64_t N = (int64_t)lx * ly * lz;
    memset(nature, 0, (size_t)N);
    switch (label) {
    case -1:
        break;
    case 0:
        for (int64_t p = 0; p < npts; ++p) {
            int32_t i = solid_ijk[3 * p], j = solid_ijk[3 * p + 1], k = solid_ijk[3 * p + 2];
            if (i <= 0 || j <= 0 || k <= 0 || i > lx || j > ly || k > lz) return LBREF_BAD_ARG;
            nature[(i - 1) + (int64_t)lx * ((j - 1) + (int64_t)ly * (k - 1))] = 1;
        }
        break;
    case 1:
        for (int64_t p = 0; p < (int64_t)lx * ly; ++p) {
            nature[p] = 1;
            nature[p + (int64_t)lx * ly * (lz - 1)] = 1;
        }
        break;
    case 2: {
        if (lx != ly || lx < 3) return LBREF_BAD_ARG;
        const double ox = (double)(lx + 1) / 2.0, oy = (double)(ly + 1) / 2.0;
        const double radius = (double)(lx - 1) / 2.0;
        for (int32_t i = 1; i <= lx; ++i)
            for (int32_t j = 1; j <= ly; ++j) {
                double r[2] = {(double)i - ox, (double)j - oy};
                int8_t s = lbref_norm2(r, 2) >= radius ? 1 : 0;
                for (int32_t k = 1; k <= lz; ++k)
                    nature[(i - 1) + (int64_t)lx * ((j - 1) + (int64_t)ly * (k - 1))] = s;
            }
        break;
    }
    case 3: {
        if (lx != ly || lx != lz) return LBREF_BAD_ARG;
        const double rad = (double)(lx - 1) * sqrt(3.0) / 4.0;
        const double pts[9][3] = {{1, 1, 1},
                                  {(double)lx, 1, 1},
                                  {1, (double)ly, 1},
                                  {1, 1, (double)lz},
                                  {(double)lx, (double)ly, 1},
                                  {(double)lx, 1, (double)lz},
                                  {1, (double)ly, (double)lz},
                                  {(double)lx, (double)ly, (double)lz},
                                  {(double)(lx + 1) / 2.0, (double)(ly + 1) / 2.0, (double)(lz + 1) / 2.0}};
        for (int32_t i = 1; i <= lx; ++i)
            for (int32_t j = 1; j <= ly; ++j)
                for (int32_t k = 1; k <= lz; ++k) {
                    int8_t s = 0;
                    for (int q = 0; q < 9; ++q) {
                        double r[3] = {(double)i - pts[q][0], (double)j - pts[q][1], (double)k - pts[q][2]};
                        if (lbref_norm2(r, 3) <= rad) s = 1;
                    }
                    nature[(i - 1) + (int64_t)lx * ((j - 1) + (int64_t)ly * (k - 1))] = s;
                }
        break;
    }
    default:
        return LBREF_BAD_ARG;
    }
    int64_t nf = 0;
    for (int64_t p = 0; p < N; ++p) nf += nature[p] == 0;
    if (nf == 0) return LBREF_BAD_ARG; /* supercell_definition.f90:88-90 */
    return LBREF_OK;
}

/* detectInterfacialNodes, supercell_definition.f90:115-147: a node (fluid or
 * solid) is interfacial when any of its 18 D3Q19 neighbours differs in nature. */
void lbref_interfacial(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, int8_t* interfacial) {
    for (int32_t k = 0; k < lz; ++k)
        for (int32_t j = 0; j < ly; ++j)
            for (int32_t i = 0; i < lx; ++i) {
                int64_t r = i + (int64_t)lx * (j + (int64_t)ly * k);
                int8_t f = 0;
                for (int l = 1; l < 19; ++l) {
                    int32_t ip = pbc(i + CV[l][0], lx), jp = pbc(j + CV[l][1], ly), kp = pbc(k + CV[l][2], lz);
                    if (nature[r] != nature[ip + (int64_t)lx * (jp + (int64_t)ly * kp)]) {
                        f = 1;
                        break;
                    }
                }
                interfacial[r] = f;
            }
}

/* module_collision.f90:69-108, one fluid node; expression order as written
 * (Fortran evaluates a+b+c as (a+b)+c, a*b*c as (a*b)*c). */
static inline void collide_node(double* nl, int64_t stride, double tau, double rho, double jx, double jy,
                                double jz, double fx, double fy, double fz) {
    const double csq = 1.0 / 3.0;
    const double ux = jx / rho, uy = jy / rho, uz = jz / rho;
    for (int l = 0; l < 19; ++l) {
        const double cx = CV[l][0], cy = CV[l][1], cz = CV[l][2];
        const double neq = A0[l] * rho + A1[l] * (cx * jx + cy * jy + cz * jz) +
                           A2[l] * (jx * ux * (cx * cx - csq) + jx * uy * cx * cy + jx * uz * cx * cz +
                                    jy * ux * cy * cx + jy * uy * (cy * cy - csq) + jy * uz * cy * cz +
                                    jz * ux * cz * cx + jz * uy * cz * cy + jz * uz * (cz * cz - csq));
        double* p = nl + stride * l;
        *p = (1.0 - 1.0 / tau) * (*p) + (1.0 / tau) * neq +
             (1.0 - 1.0 / (2.0 * tau)) * (A1[l] * ((cx - ux) * fx + (cy - uy) * fy + (cz - uz) * fz) +
                                          2.0 * A2[l] * (cx * ux + cy * uy + cz * uz) * (cx * fx + cy * fy + cz * fz));
    }
}

/* collide(n, density, jx, jy, jz, f_ext_x, f_ext_y, f_ext_z),
 * module_collision.f90:15-130 (second-order branch; fluid nodes only). */
int lbref_collide(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, double tau,
                  double* n, const double* rho, const double* jx, const double* jy, const double* jz,
                  const double* fx, const double* fy, const double* fz) {
    init_tables();
    if (tau < 0.5) return LBREF_BAD_TAU; /* module_collision.f90:39 */
    const int64_t N = (int64_t)lx * ly * lz;
    for (int64_t r = 0; r < N; ++r)
        if (nature[r] == 0) collide_node(n + r, N, tau, rho[r], jx[r], jy[r], jz[r], fx[r], fy[r], fz[r]);
    return LBREF_OK;
}

/* The equilibration time loop, equilibration.f90:143-491 (without the
 * compensate_f_ext branch :388-487), starting from init_simu.f90:31-39:
 * n = rho0*a0 at fluid nodes, 0 at solids; density = rho0 (fluid), j = 0.
 * Per step: collide (:194), bounce-back link swap for odd l (:204-222),
 * push streaming (:227-243), negativity check (:248), density = SUM(n,4)
 * (:254), j = F/2 + sum_l n c (:293-300), l2err = max|j - j_old| (:339-344),
 * two-stage convergence and force switch (:346-386).  max_steps > 0 adds a
 * step cap (not a reference feature).  On return n, rho, j hold the state
 * after the last step taken (equilibration.f90:551-554). */
int lbref_equilibrate(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, double tau,
                      double rho0, const double fext[3], double target_error, int64_t max_steps,
                      double* n, double* rho, double* jx, double* jy, double* jz,
                      double* l2err_trace, int64_t trace_cap, int64_t* steps_done, int32_t* exit_reason) {
    init_tables();
    if (tau < 0.5) return LBREF_BAD_TAU;
    const int64_t N = (int64_t)lx * ly * lz;
    double* fx = calloc((size_t)N, sizeof(double));
    double* fy = calloc((size_t)N, sizeof(double));
    double* fz = calloc((size_t)N, sizeof(double));
    double* jox = malloc((size_t)N * sizeof(double));
    double* joy = malloc((size_t)N * sizeof(double));
    double* joz = malloc((size_t)N * sizeof(double));
    double* nold = malloc((size_t)N * sizeof(double));
    int32_t *il = malloc(19 * sizeof(int32_t) * (size_t)lx), *jl = malloc(19 * sizeof(int32_t) * (size_t)ly),
            *kl = malloc(19 * sizeof(int32_t) * (size_t)lz);
    int status = LBREF_OK;
    *exit_reason = 0;
    *steps_done = 0;
    for (int64_t r = 0; r < N; ++r) {
        rho[r] = nature[r] == 0 ? rho0 : 0.0;
        jx[r] = jy[r] = jz[r] = 0.0;
    }
    for (int l = 0; l < 19; ++l)
        for (int64_t r = 0; r < N; ++r) n[r + N * l] = rho[r] * A0[l];
    /* equilibration.f90:109-119 */
    for (int l = 0; l < 19; ++l) {
        for (int32_t i = 0; i < lx; ++i) il[l * lx + i] = pbc(i + CV[l][0], lx);
        for (int32_t j = 0; j < ly; ++j) jl[l * ly + j] = pbc(j + CV[l][1], ly);
        for (int32_t k = 0; k < lz; ++k) kl[l * lz + k] = pbc(k + CV[l][2], lz);
    }
    int conv_without = 0, conv_with = 0;
    for (int64_t t = 1;; ++t) {
        lbref_collide(lx, ly, lz, nature, tau, n, rho, jx, jy, jz, fx, fy, fz);
        /* bounce-back link swap, odd reference l (0-based even l) */
        for (int l = 0; l < 19; l += 2)
            for (int32_t k = 0; k < lz; ++k) {
                int32_t kp = kl[l * lz + k];
                for (int32_t j = 0; j < ly; ++j) {
                    int32_t jp = jl[l * ly + j];
                    for (int32_t i = 0; i < lx; ++i) {
                        int32_t ip = il[l * lx + i];
                        int64_t r = i + (int64_t)lx * (j + (int64_t)ly * k);
                        int64_t rp = ip + (int64_t)lx * (jp + (int64_t)ly * kp);
                        if (nature[r] != nature[rp]) {
                            double tmp = n[r + N * l];
                            n[r + N * l] = n[rp + N * INV[l]];
                            n[rp + N * INV[l]] = tmp;
                        }
                    }
                }
            }
        /* push streaming */
        for (int l = 0; l < 19; ++l) {
            memcpy(nold, n + N * l, (size_t)N * sizeof(double));
            for (int32_t k = 0; k < lz; ++k) {
                int32_t kp = kl[l * lz + k];
                for (int32_t j = 0; j < ly; ++j) {
                    int32_t jp = jl[l * ly + j];
                    for (int32_t i = 0; i < lx; ++i) {
                        int32_t ip = il[l * lx + i];
                        n[ip + (int64_t)lx * (jp + (int64_t)ly * kp) + N * l] =
                            nold[i + (int64_t)lx * (j + (int64_t)ly * k)];
                    }
                }
            }
        }
        *steps_done = t;
        for (int64_t p = 0; p < N * 19; ++p)
            if (n[p] < 0) {
                status = LBREF_NEG_POP;
                goto done;
            }
        for (int64_t r = 0; r < N; ++r) {
            double s = 0.0;
            for (int l = 0; l < 19; ++l) s = s + n[r + N * l];
            rho[r] = s;
        }
        memcpy(jox, jx, (size_t)N * sizeof(double));
        memcpy(joy, jy, (size_t)N * sizeof(double));
        memcpy(joz, jz, (size_t)N * sizeof(double));
        for (int64_t r = 0; r < N; ++r) {
            double ax = fx[r] / 2.0, ay = fy[r] / 2.0, az = fz[r] / 2.0;
            for (int l = 0; l < 19; ++l) {
                double v = n[r + N * l];
                ax = ax + v * (double)CV[l][0];
                ay = ay + v * (double)CV[l][1];
                az = az + v * (double)CV[l][2];
            }
            jx[r] = ax;
            jy[r] = ay;
            jz[r] = az;
        }
        double mx = 0.0, my = 0.0, mz = 0.0;
        for (int64_t r = 0; r < N; ++r) {
            double dx = fabs(jx[r] - jox[r]), dy = fabs(jy[r] - joy[r]), dz = fabs(jz[r] - joz[r]);
            if (dx > mx) mx = dx;
            if (dy > my) my = dy;
            if (dz > mz) mz = dz;
        }
        double l2err = mx;
        if (my > l2err) l2err = my;
        if (mz > l2err) l2err = mz;
        if (l2err_trace && t <= trace_cap) l2err_trace[t - 1] = l2err;

        int converged = (l2err <= target_error && t > 2);
        if (converged) {
            if (!conv_without)
                conv_without = 1;
            else
                conv_with = 1;
        }
        if (converged) {
            if (conv_without && conv_with) {
                *exit_reason = LBREF_EXIT_CONVERGED;
                break;
            } else if (conv_without && !conv_with) {
                for (int64_t r = 0; r < N; ++r)
                    if (nature[r] == 0) {
                        fx[r] = fext[0];
                        fy[r] = fext[1];
                        fz[r] = fext[2];
                    }
            }
        }
        if (max_steps > 0 && t >= max_steps) {
            *exit_reason = LBREF_EXIT_CAP;
            break;
        }
    }
done:
    free(fx); free(fy); free(fz); free(jox); free(joy); free(joz); free(nold);
    free(il); free(jl); free(kl);
    return status;
}

/* update_tracer_population, drop_tracers.f90:85-105 with tracer_z = 0 (the
 * only value module_moment_propagation.f90:71-82 accepts):
 * n(l,r) = a0*rho + a1*sum(c*(j + f_ext - rho*q*D*E)), E = 0 (elec_slope is
 * never set on this path); solids drop the f_ext term. */
void lbref_tracer_population(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature,
                             const double* rho, const double* jx, const double* jy, const double* jz,
                             const double fext[3], double* nt) {
    init_tables();
    const int64_t N = (int64_t)lx * ly * lz;
    const double qDE = 0.0 * 0.0; /* q*D*elec_slope with q = 0, E = 0 */
    for (int64_t r = 0; r < N; ++r) {
        for (int l = 0; l < 19; ++l) {
            double s;
            if (nature[r] == 0) {
                double vx = jx[r] + fext[0] - rho[r] * qDE;
                double vy = jy[r] + fext[1] - rho[r] * qDE;
                double vz = jz[r] + fext[2] - rho[r] * qDE;
                s = (((double)CV[l][0] * vx + (double)CV[l][1] * vy) + (double)CV[l][2] * vz);
            } else {
                s = (((double)CV[l][0] * jx[r] + (double)CV[l][1] * jy[r]) + (double)CV[l][2] * jz[r]);
            }
            nt[l + 19 * r] = A0[l] * rho[r] + A1[l] * s;
        }
    }
}

/* calc_scattprop, module_moment_propagation.f90:341-346 */
static inline double scattprop(double n, double rho, double w, double lambda, double fermi) {
    return n / rho - w + lambda * w * fermi;
}

/* calc_lambda, module_moment_propagation.f90:331-337 with kBT = 1/3
 * (module_system.f90:94) */
static inline double calc_lambda(double d) { return 4.0 * d / (1.0 / 3.0); }

static int mp_params(double Db, double ka, double kd, double* K, int32_t* ads) {
    const double eps = DBL_EPSILON;
    if (ka < -eps || kd < -eps) return LBREF_BAD_ARG;       /* :43-44 */
    if (Db <= eps) return LBREF_BAD_ARG;                     /* drop_tracers.f90:89 */
    *K = (fabs(kd) <= eps) ? 0.0 : ka / kd;                  /* :46-50 */
    *ads = fabs(*K) > eps;                                   /* :52-56 */
    return LBREF_OK;
}

/* moment_propagation::init, module_moment_propagation.f90:30-160 (neutral
 * tracer): Pstat, VACF(0) and P(r, t=1).  P is (3,lx,ly,lz), zeroed here.
 * DO CONCURRENT (i,j,k) runs i outermost, k innermost under flang. */
int lbref_mp_init(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, const int8_t* interfacial,
                  const double* nt, const double* rho, double Db, double ka, double kd,
                  double* P, double vacf0[3], int32_t* consider_adsorption) {
    init_tables();
    double K;
    int st = mp_params(Db, ka, kd, &K, consider_adsorption);
    if (st) return st;
    const double lambda = calc_lambda(Db);
    const int64_t N = (int64_t)lx * ly * lz;
    int64_t nf = 0, nfi = 0;
    for (int64_t r = 0; r < N; ++r) {
        nf += nature[r] == 0;
        nfi += nature[r] == 0 && interfacial[r];
    }
    const double Pstat = (double)nf + K * (double)nfi; /* :97 */
    memset(P, 0, (size_t)(3 * N) * sizeof(double));
    vacf0[0] = vacf0[1] = vacf0[2] = 0.0;
    const double fermi = 1.0 / (1.0 + 1.0);
    for (int32_t i = 0; i < lx; ++i)
        for (int32_t j = 0; j < ly; ++j)
            for (int32_t k = 0; k < lz; ++k) {
                int64_t r = i + (int64_t)lx * (j + (int64_t)ly * k);
                if (nature[r] != 0) continue;
                const double bw = 1.0 / Pstat;
                for (int l = 1; l < 19; ++l) {
                    int32_t ip = pbc(i + CV[l][0], lx), jp = pbc(j + CV[l][1], ly), kp = pbc(k + CV[l][2], lz);
                    int64_t rp = ip + (int64_t)lx * (jp + (int64_t)ly * kp);
                    if (nature[rp] == 1) continue;
                    const double exp_min_dphi = 1.0 / 1.0;
                    double s = scattprop(nt[l + 19 * r], rho[r], A0[l], lambda, fermi);
                    for (int a = 0; a < 3; ++a)
                        vacf0[a] = vacf0[a] + bw * s * (double)(CV[l][a] * CV[l][a]);
                    int li = INV[l];
                    double sp = scattprop(nt[li + 19 * rp], rho[rp], A0[li], lambda, 1.0 - fermi);
                    for (int a = 0; a < 3; ++a)
                        P[a + 3 * r] = P[a + 3 * r] + exp_min_dphi * sp * (double)CV[li][a] * bw;
                }
            }
    return LBREF_OK;
}

/* moment_propagation::propagate, module_moment_propagation.f90:201-257, one
 * step: P_next must hold zeros on entry (the roll at :262-267 zeroes it);
 * Pa_* may be NULL when adsorption is off.  VACF(it) is accumulated over
 * fluid nodes with k outermost, i innermost (1 OpenMP thread). */
int lbref_mp_step(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, const int8_t* interfacial,
                  const double* nt, const double* rho, double Db, double ka, double kd,
                  int32_t consider_adsorption,
                  const double* P_now, const double* Pa_now, double* P_next, double* Pa_next,
                  double vacf[3]) {
    init_tables();
    const double lambda = calc_lambda(Db);
    const double eps = DBL_EPSILON;
    const double fermi = 1.0 / (1.0 + 1.0);
    int error = 0;
    vacf[0] = vacf[1] = vacf[2] = 0.0;
    for (int32_t k = 0; k < lz; ++k)
        for (int32_t j = 0; j < ly; ++j)
            for (int32_t i = 0; i < lx; ++i) {
                int64_t r = i + (int64_t)lx * (j + (int64_t)ly * k);
                if (nature[r] != 0) continue;
                double us[3] = {0.0, 0.0, 0.0};
                double rem = 1.0;
                double loc[3] = {P_next[3 * r], P_next[3 * r + 1], P_next[3 * r + 2]};
                for (int l = 1; l < 19; ++l) {
                    int32_t ip = pbc(i + CV[l][0], lx), jp = pbc(j + CV[l][1], ly), kp = pbc(k + CV[l][2], lz);
                    int64_t rp = ip + (int64_t)lx * (jp + (int64_t)ly * kp);
                    if (nature[rp] != 0) continue;
                    double s = scattprop(nt[l + 19 * r], rho[r], A0[l], lambda, fermi);
                    rem = rem - s;
                    for (int a = 0; a < 3; ++a) us[a] = us[a] + s * (double)CV[l][a];
                    int li = INV[l];
                    double sp = scattprop(nt[li + 19 * rp], rho[rp], A0[li], lambda, 1.0 - fermi);
                    for (int a = 0; a < 3; ++a) loc[a] = loc[a] + P_now[a + 3 * rp] * sp;
                }
                for (int a = 0; a < 3; ++a) P_next[a + 3 * r] = loc[a];
                for (int a = 0; a < 3; ++a) vacf[a] = vacf[a] + P_now[a + 3 * r] * us[a];
                if ((!interfacial[r] && consider_adsorption) || !consider_adsorption) {
                    for (int a = 0; a < 3; ++a) P_next[a + 3 * r] = P_next[a + 3 * r] + rem * P_now[a + 3 * r];
                } else {
                    rem = rem - ka;
                    for (int a = 0; a < 3; ++a) {
                        P_next[a + 3 * r] = P_next[a + 3 * r] + rem * P_now[a + 3 * r] + Pa_now[a + 3 * r] * kd;
                        Pa_next[a + 3 * r] = Pa_now[a + 3 * r] * (1.0 - kd) + P_now[a + 3 * r] * ka;
                    }
                }
                if (rem < eps) error = 1;
            }
    return error ? LBREF_RESTPART : LBREF_OK;
}

/* Convergence test of module_moment_propagation.f90:284-288; 1.e-12 there is a
 * default-real (single precision) literal. */
int lbref_mp_converged(int64_t it, const double vacf[3], int32_t lx, int32_t ly, int32_t lz, double Db) {
    const double lim1 = 1.0 / (2.0 * (double)lx * (double)ly * (double)lz / Db);
    const double lim2 = (double)1.e-12f;
    if (it <= 2) return 0;
    for (int a = 0; a < 3; ++a)
        if (!(fabs(vacf[a]) < lim1) || !(fabs(vacf[a]) < lim2)) return 0;
    return 1;
}

/* Effective diffusion coefficient from the VACF trace (not computed by the
 * reference; drop_tracers.f90:53-54 leaves the integration commented out):
 * discrete Green-Kubo D_a = VACF_a(0)/2 + sum_{t=1..T} VACF_a(t), summed
 * with Neumaier compensation.  vacf holds (T+1) rows of 3. */
void lbref_deff(const double* vacf, int64_t nsteps, double deff[3]) {
    for (int a = 0; a < 3; ++a) {
        double s = 0.5 * vacf[a], c = 0.0;
        for (int64_t t = 1; t <= nsteps; ++t) {
            double x = vacf[3 * t + a];
            double u = s + x;
            if (fabs(s) >= fabs(x))
                c += (s - u) + x;
            else
                c += (x - u) + s;
            s = u;
        }
        deff[a] = s + c;
    }
}

/* lbref_norm2 over count vectors of n components (v[n*k + a]), for the
 * comparison with the compiler's NORM2 in the tests. */
void lbref_norm2_rows(const double* v, int n, int64_t count, double* out) {
    for (int64_t k = 0; k < count; ++k) out[k] = lbref_norm2(v + (int64_t)n * k, n);
}
