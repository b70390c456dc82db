/* lbx: MI355X (gfx950) lattice-Boltzmann + moment-propagation hot path of
 * laboetie (maxlevesque/laboetie), as a C-ABI shared library (liblbx.so).
 *
 * The reference has no FFI: its hot path is a set of Fortran subroutines on
 * module globals.  Each entry point below replaces one of them (cited), so a
 * Fortran shim with the reference's unchanged interfaces can forward to it
 * through iso_c_binding (see laboetie_amd/fortran).
 *
 * Conventions
 *  - Arrays are host arrays in the reference's Fortran column-major layouts,
 *    flattened (x fastest, N = lx*ly*lz):
 *      scalar field   f[i + lx*(j + ly*k)]            n(lx,ly,lz)
 *      population     n[node + N*l], l = 0..18         n(lx,ly,lz,19)
 *      moment field   P[a + 3*node], a = 0..2          P(3,lx,ly,lz)
 *    l = 0..18 stands for the reference's velocity index 1..19
 *    (module_lbmodel.f90:66-86).
 *  - nature: int8, 0 fluid, 1 solid (module_system.f90:32).
 *  - Every function returns an int status (LBX_OK = 0); the reference's fatal
 *    stops become status codes the caller maps back to the same messages.
 *  - A context owns its device memory and one HIP stream; it is not
 *    thread-safe (one caller at a time), like the reference's routines.
 *  - Copies in and out are synchronous unless stated; state between calls
 *    stays resident in HBM.
 */
#ifndef LBX_H
#define LBX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define LBX_OK            0
#define LBX_ERR_NEG_POP   1  /* "In equilibration, the population n(x,y,z,vel) < 0" equilibration.f90:248 */
#define LBX_ERR_RESTPART  2  /* "somewhere restpart is negative" module_moment_propagation.f90:249,257 */
#define LBX_ERR_TAU       3  /* "relaxation_time must be > 0.5" module_collision.f90:39 */
#define LBX_ERR_NOMEM     4  /* device allocation failed / HBM admission refused */
#define LBX_ERR_ARG       5  /* invalid argument (also the reference's input stops) */
#define LBX_ERR_DEVICE    6  /* HIP runtime error or no device */
#define LBX_ERR_STATE     7  /* call out of order (e.g. mp step before mp init) */
#define LBX_ERR_SIZE      8  /* input larger than a fixed-size staging area (lbx_collide_noalloc) */

/* precision forms */
#define LBX_PREC_FP64        0  /* populations and arithmetic in FP64 (parity form) */
#define LBX_PREC_FP32_SHIFT  1  /* populations stored as FP32 f_l - w_l*rho0, FP32 collision
                                   (constants rounded to conserve momentum), FP64
                                   moments/reductions/VACF/D_eff */
#define LBX_PREC_FP32_STORE  2  /* the same FP32 storage, collision and the stop test's flux in
                                   FP64 (half the FP64 form's population bytes at a fraction of
                                   the FP32-shift form's error) */

/* lbx_equilibrate exit reasons */
#define LBX_EXIT_CONVERGED 1    /* second convergence, equilibration.f90:373-374 */
#define LBX_EXIT_CAP       2    /* max_steps reached (extension: the reference has no cap) */
#define LBX_EXIT_FLOOR     3    /* FP32-storage forms only: l2err stopped improving above target_error
                                   (the form's rounding floor, lbx_set_floor_window) */

typedef struct lbx_ctx lbx_ctx;

/* Message of a status code (static string). */
const char* lbx_status_string(int status);

/* Library version, e.g. 100 for 1.0.0. */
int lbx_version(void);

/* D3Q19 tables exactly as module_lbmodel.f90:66-86,122-162 builds them:
 * c[3*l+a], a0 = w, a1 = w/cs^2, a2 = w/(2 cs^4), inv (0-based). */
int lbx_d3q19_tables(int32_t c[57], double a0[19], double a1[19], double a2[19], int32_t inv[19]);

/* ------------------------------------------------------------- geometry
 * Host-side builders of node%nature, replacing supercell_definition.f90:48-90
 * for geometryLabel -1 (no solid), 0 (custom list, module_geometry.f90:380-426:
 * solid_ijk holds npts 1-based (i,j,k) triples), 1 (slit, :158-166),
 * 2 (cylinder along z, :253-277), 3 (BCC of touching spheres, :206-245).
 * LBX_ERR_ARG on the reference's stops (shape checks, index out of range,
 * all nodes solid). */
int lbx_build_geometry(int32_t label, int32_t lx, int32_t ly, int32_t lz,
                       const int32_t* solid_ijk, int64_t npts, int8_t* nature);

/* detectInterfacialNodes, supercell_definition.f90:115-147. */
int lbx_detect_interfacial(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, int8_t* interfacial);

/* Seeded synthetic media (no counterpart in the reference, which reads such
 * media from geom.in files).  Random numbers: SplitMix64(seed).
 *
 * lbx_gaussian_field: i.i.d. N(0,1) per node (Box-Muller pairs, x fastest)
 * convolved periodically with exp(-d^2/(2 sigma^2)) along each axis
 * (|d| <= min(ceil(4 sigma), (n-1)/2)); field is (lx,ly,lz).
 * lbx_geometry_random: that field thresholded so that exactly
 * round(porosity*N) nodes are fluid (the lowest values; ties by index);
 * solid clusters are all kept.
 * lbx_geometry_disks: ndisks solid disks/spheres of the given radius at
 * seeded centres, uniform in [1, n+1) per axis (1 on an axis of length 1),
 * no two closer than 2*radius (periodic minimum image); a node is solid when
 * its per-axis minimum-image distance to a centre is <= radius, the rule of
 * tools/chromato_generator.f90:40-64.  centres (may be NULL) receives 3*ndisks
 * 1-based coordinates. */
int lbx_gaussian_field(int32_t lx, int32_t ly, int32_t lz, double sigma, uint64_t seed, double* field);
int lbx_geometry_random(int32_t lx, int32_t ly, int32_t lz, double sigma, double porosity, uint64_t seed,
                        int8_t* nature);
int lbx_geometry_disks(int32_t lx, int32_t ly, int32_t lz, int32_t ndisks, double radius, uint64_t seed,
                       int8_t* nature, double* centres);

/* ---------------------------------------------------------- input files
 * lb.in: the reference's parser (module_input.f90:561-633, the tag lookup of
 * input_dp and siblings :62-75): lines left-adjusted, cut to 100 characters,
 * '#' comments and blank lines dropped; a tag matches the first line that
 * starts with it followed by a blank, and its value is read list-directed
 * (blanks or commas, '/' ends; d exponents) from columns len(tag)+4 ..
 * len(tag)+50.  Each getter sets *found = 0 and leaves out untouched when no
 * line holds the tag (the caller's default applies), and returns LBX_ERR_ARG
 * on too few or malformed values.  lbx_lbin_dp / lbx_lbin_int read count
 * values (getinput%dp, %dp3, %int); lbx_lbin_log reads a logical (T/F,
 * .true./.false.), lbx_lbin_char the first token without quotes.
 * lbx_geom_in_parse: geom.in of the custom geometry (construct_custom,
 * module_geometry.f90:380-426): 1-based `i j k` solid nodes, one per line;
 * nature (lx,ly,lz, x fastest) becomes 1 there and 0 elsewhere; an index
 * out of range is LBX_ERR_ARG (the reference stops); *nsolid (may be NULL)
 * receives the number of distinct solid nodes. */
typedef struct lbx_lbin lbx_lbin;
int lbx_lbin_parse(const char* text, int64_t len, lbx_lbin** out);
int lbx_lbin_free(lbx_lbin* h);
int lbx_lbin_dp(const lbx_lbin* h, const char* tag, int32_t count, double* out, int32_t* found);
int lbx_lbin_int(const lbx_lbin* h, const char* tag, int32_t count, int64_t* out, int32_t* found);
int lbx_lbin_log(const lbx_lbin* h, const char* tag, int32_t* out, int32_t* found);
int lbx_lbin_char(const lbx_lbin* h, const char* tag, char* out, int64_t cap, int32_t* found);
int lbx_geom_in_parse(const char* text, int64_t len, int32_t lx, int32_t ly, int32_t lz, int8_t* nature,
                      int64_t* nsolid);

/* output/supercell.xsf of print_supercell_xsf (module_io.f90:86-160), byte
 * for byte as the reference program writes it: nature and interfacial
 * (lx,ly,lz, x fastest) to text in out (cap bytes, not NUL-terminated);
 * *len receives the text's size, and cap = 0 (out may be NULL) only asks
 * for it.  LBX_ERR_ARG when cap is too small or an axis has more than
 * 950000 nodes (flang writes the coordinates from 950000 on in E form). */
int lbx_supercell_xsf(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, const int8_t* interfacial,
                      char* out, int64_t cap, int64_t* len);

/* --------------------------------------------------------------- context */

/* HBM bytes a context of this shape needs (admission control for batching
 * and serving); with_mp > 0 adds the moment-propagation arrays of that many
 * tracers (lbx_mp_init_batch). */
int64_t lbx_footprint_bytes(int32_t lx, int32_t ly, int32_t lz, int32_t precision, int32_t with_mp);

/* Create a context on the current HIP device holding one lattice: nature
 * (N int8), relaxation time tau (must be >= 0.5, module_collision.f90:37-41).
 * Allocates everything the LB phase needs; the MP arrays come with
 * lbx_mp_init. */
int lbx_create(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, double tau,
               int32_t precision, lbx_ctx** out);
int lbx_destroy(lbx_ctx* ctx);

/* Populations n = rho0*w_l at fluid nodes, 0 at solids; density rho0 at
 * fluid, 0 at solid; j = 0 (init_simu.f90:31-39, equilibration.f90:59-80). */
int lbx_init_rest(lbx_ctx* ctx, double rho0);

/* Arbitrary state: n (post-streaming populations, (lx,ly,lz,19)), and the
 * density and flux that the next collision uses (collide's arguments).  In
 * the FP32-storage forms the rest density rho0 of the deviations becomes the
 * density of the first fluid node (lbx_init_rest sets it to its rho0). */
int lbx_set_state(lbx_ctx* ctx, const double* n, const double* rho,
                  const double* jx, const double* jy, const double* jz);

/* The equilibration time loop, equilibration.f90:143-491 (without the
 * compensate_f_ext branch): per step collide, bounce-back, streaming,
 * negativity check, density and flux, l2err = max|j - j_old|; stage 1 runs
 * with no force until l2err <= target_error (t > 2), then fext acts at every
 * fluid node; the run exits at the next convergence.  max_steps > 0 caps the
 * number of steps (extension).  l2err_trace (may be NULL) receives the first
 * trace_cap values of the per-step l2err (what l2err.dat holds).
 * FP32-storage forms (extension): stage 2 measures l2err only at every 64th
 * step after the force switch, as max|j(t) - j(t-64)| / 64 (the per-step
 * difference sits on the FP32 rounding of j); the stop tests apply there,
 * and the trace holds NaN at the other steps.
 * Returns LBX_ERR_NEG_POP with *steps_done = the failing step. */
int lbx_equilibrate(lbx_ctx* ctx, double target_error, const double fext[3], int64_t max_steps,
                    double* l2err_trace, int64_t trace_cap, int64_t* steps_done, int32_t* exit_reason);

/* The same loop in pieces, for callers that write outputs along the run
 * (the profiles every print_files_frequency steps and output/
 * total_mass_flux.dat of equilibration.f90:154-176,257-261, the stdout lines
 * of :149).  lbx_equilibrate_begin starts a run from the current state (the
 * opening collision of step 1); with flux_sums != 0 every step also sums j
 * over the fluid nodes (lbx_flux_sums).  lbx_equilibrate_advance runs it on
 * until step until_step (INT64_MAX: to its end) or its end, whichever comes
 * first: *exit_reason = 0 while the run goes on, else the reason as for
 * lbx_equilibrate, with the state extracted as lbx_equilibrate leaves it.
 * lbx_equilibrate is begin + advance to the end.  Between the calls of a
 * paused run, lbx_flow gives the density and flux after the last step taken,
 * lbx_l2err_trace and lbx_flux_sums the steps taken so far; lbx_get_state
 * and lbx_mp_init* need the run's end (LBX_ERR_STATE); a new begin,
 * lbx_init_rest or lbx_set_state ends or drops a paused run. */
int lbx_equilibrate_begin(lbx_ctx* ctx, double target_error, const double fext[3], int64_t max_steps,
                          int32_t flux_sums);
int lbx_equilibrate_advance(lbx_ctx* ctx, int64_t until_step, int64_t* steps_done, int32_t* exit_reason);

/* sum over the fluid nodes of j(t) = F(t)/2 + sum n c, 3 values per step, for
 * steps first_step .. first_step+count-1 of a run begun with flux_sums (the
 * value output/total_mass_flux.dat prints, in single precision, at step t+1);
 * *available (may be NULL) receives how many steps are held. */
int lbx_flux_sums(lbx_ctx* ctx, int64_t first_step, int64_t count, double* out, int64_t* available);

/* Density and flux (full box, 0 at solid nodes) after the last step taken:
 * of a paused run (without disturbing it), else what lbx_get_state gives.
 * LBX_ERR_STATE before the first step of a run. */
int lbx_flow(lbx_ctx* ctx, double* rho, double* jx, double* jy, double* jz);

/* The per-step l2err of the last lbx_equilibrate call, steps first_step ..
 * first_step+count-1 (1-based, as written to output/l2err.dat by
 * equilibration.f90:339-344); *available (may be NULL) receives how many
 * steps are held.  count = 0 only queries *available. */
int lbx_l2err_trace(lbx_ctx* ctx, int64_t first_step, int64_t count, double* out, int64_t* available);

/* State after the last step (equilibration.f90:551-555): n is the
 * post-streaming population, rho = SUM(n,4), j = F/2 + sum n c.  Any pointer
 * may be NULL. */
int lbx_get_state(lbx_ctx* ctx, double* n, double* rho, double* jx, double* jy, double* jz);

/* ------------------------------------------------- moment propagation
 * moment_propagation::init (module_moment_propagation.f90:30-160) after
 * update_tracer_population (drop_tracers.f90:85-105): rebuilds the tracer
 * population from the resident (rho, j) and fext, computes Pstat, VACF(0)
 * and P(r,1).  Tracers are neutral (tracer_z = 0, tracer_Ds = 0 are the only
 * values the reference accepts).  Db > eps; ka, kd >= 0; adsorption is on
 * when ka/kd > eps. */
int lbx_mp_init(lbx_ctx* ctx, const double fext[3], double Db, double ka, double kd, double vacf0[3]);

/* Same, from host density and flux fields (node%solventDensity/solventFlux)
 * instead of the resident flow. */
int lbx_mp_init_from_flow(lbx_ctx* ctx, const double* rho, const double* jx, const double* jy,
                          const double* jz, const double fext[3], double Db, double ka, double kd,
                          double vacf0[3]);

/* Several tracers on the same flow with the same Db and different ka, kd
 * (1 <= ntracers <= 4; ka, kd: ntracers values), propagated together: the
 * shared scattering probabilities are read once per node per step for all of
 * them.  Each tracer's P, Pa, VACF and D_eff are bit-identical to those of a
 * single-tracer lbx_mp_init with its ka, kd (module_moment_propagation.f90
 * :30-160 per tracer).  vacf0 (may be NULL) receives 3*ntracers values. */
int lbx_mp_init_batch(lbx_ctx* ctx, const double fext[3], double Db, int32_t ntracers, const double* ka,
                      const double* kd, double* vacf0);

/* Number of tracers of the current MP state (0 before lbx_mp_init*). */
int lbx_mp_tracers(lbx_ctx* ctx, int32_t* ntracers);

/* nsteps calls of moment_propagation::propagate (module_moment_propagation.f90
 * :164-289).  vacf_trace (may be NULL) receives VACF(it) for each step taken,
 * 3 per tracer per step (tracer-major within a step).  With
 * stop_on_convergence (single tracer only, else LBX_ERR_ARG), stops after the
 * first step whose VACF passes the test of :284-288 and sets *converged.
 * LBX_ERR_RESTPART when a rest probability is below epsilon (the reference
 * stops in the first propagate call). */
int lbx_mp_run(lbx_ctx* ctx, int64_t nsteps, int32_t stop_on_convergence, double* vacf_trace,
               int64_t* steps_done, int32_t* converged);

/* lbx_mp_run in two halves, so that the caller's host work (writing the
 * previous steps' vacf.dat lines, module_moment_propagation.f90:269) runs
 * while the GPU steps: lbx_mp_run_start launches 1 <= nsteps <= 512 steps
 * and returns; lbx_mp_run_wait waits for them and returns what lbx_mp_run
 * would have.  One begun chunk at a time; until it is waited for, the other
 * lbx_mp_* calls return LBX_ERR_STATE (lbx_mp_free and lbx_mp_init* drop it).
 * The same steps and bits as lbx_mp_run.  No counterpart in the reference. */
int lbx_mp_run_start(lbx_ctx* ctx, int64_t nsteps, int32_t stop_on_convergence);
int lbx_mp_run_wait(lbx_ctx* ctx, double* vacf_trace, int64_t* steps_done, int32_t* converged);

/* Current P(now) and, with adsorption, Pa(now), (3,lx,ly,lz), of tracer 0;
 * either may be NULL. */
int lbx_mp_get(lbx_ctx* ctx, double* P, double* Pa);

/* The same for tracer `tracer` (0-based) of a batch. */
int lbx_mp_get_tracer(lbx_ctx* ctx, int32_t tracer, double* P, double* Pa);

/* Effective diffusion coefficient of the steps run so far (not in the
 * reference: drop_tracers.f90:53-54 is commented out):
 * D_a = VACF_a(0)/2 + sum_{t>=1} VACF_a(t), FP64 with compensated summation. */
int lbx_mp_deff(lbx_ctx* ctx, double deff[3]);

/* D_eff of every tracer of a batch, 3*ntracers values. */
int lbx_mp_deff_batch(lbx_ctx* ctx, double* deff);

/* Release the MP arrays (deallocate_propagated_quantity,
 * module_moment_propagation.f90:350-353). */
int lbx_mp_free(lbx_ctx* ctx);

/* ------------------------------------------------------- stateless call
 * collide(n, density, jx, jy, jz, f_ext_x, f_ext_y, f_ext_z),
 * module_collision.f90:15-130: one BGK + Guo-force collision of every fluid
 * node of host arrays, on the device; n is updated in place. */
int lbx_collide(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, double tau, double* n,
                const double* density, const double* jx, const double* jy, const double* jz,
                const double* f_ext_x, const double* f_ext_y, const double* f_ext_z);

/* lbx_collide without any device allocation: no counterpart in the reference.
 * It is for callers that cannot allocate device memory, such as a loop that
 * forwards the reference's per-step collide() call and must not hipMalloc /
 * hipFree five fields every step.  Same arguments, same results bit for bit
 * as lbx_collide.  The fields are staged in a static device array of 64 MiB
 * that each GPU the library runs on holds until the process ends; lattices of
 * more than LBX_NOALLOC_MAX_NODES nodes (lx*ly*lz) return LBX_ERR_SIZE.  The
 * call runs on the null stream and returns when n holds the result.  All
 * calls on one GPU share its array: calls that overlap on one GPU race. */
#define LBX_NOALLOC_MAX_NODES 320000
int lbx_collide_noalloc(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, double tau, double* n,
                        const double* density, const double* jx, const double* jy, const double* jz,
                        const double* f_ext_x, const double* f_ext_y, const double* f_ext_z);

/* ------------------------------------------------------------- timing
 * Average device duration (ms) of the launches of one kernel family since
 * the last reset, measured with HIP events on the context's stream around
 * back-to-back launch runs.  which: 0 = LB step, 1 = MP step. */
#define LBX_KERNEL_LB 0
#define LBX_KERNEL_MP 1
int lbx_timing_enable(lbx_ctx* ctx, int32_t on);
int lbx_kernel_time(lbx_ctx* ctx, int32_t which, double* total_ms, int64_t* launches);

/* FP32-storage forms: a stage of lbx_equilibrate also ends when l2err has
 * not set a new minimum for `steps` consecutive steps after falling at least
 * 100-fold below the stage's largest l2err (its FP32 rounding floor;
 * a stage-1 floor switches the force on like a convergence, a stage-2 floor
 * exits with LBX_EXIT_FLOOR).  0 disables the test.  Default 1000 for
 * FP32-shift and FP32-store contexts; FP64 contexts keep the reference's
 * semantics and reject a non-zero window (LBX_ERR_ARG).  No counterpart in
 * the reference. */
int lbx_set_floor_window(lbx_ctx* ctx, int64_t steps);

/* FP32-storage forms: the steps between the l2err measurements of stage 2
 * of lbx_equilibrate (a power of two up to 4096; default 64).  l2err is then
 * max|j(t) - j(t-steps)| / steps at the steps switch + m*steps, and NaN in
 * the trace at the others; 1 measures every step as the FP64 form does.
 * FP64 contexts accept only 1 (LBX_ERR_ARG otherwise).  No counterpart in
 * the reference. */
int lbx_set_check_interval(lbx_ctx* ctx, int32_t steps);

/* Element addressing of the step kernels.  The population and moment arrays
 * are addressed through 32-bit buffer offsets while 19*Np*8 < 4 GiB (Np: the
 * fluid-node count rounded up to 64), and through 64-bit pointers above that;
 * the library picks the form from the size.  wide = 1 selects the 64-bit form
 * on a smaller lattice too, so the two forms can be checked against each other
 * (they give the same bits); wide = 0 returns to the size's choice.
 * lbx_addressing reports the form in use (1: 64-bit).  No counterpart in the
 * reference, whose arrays the Fortran runtime addresses. */
int lbx_set_addressing(lbx_ctx* ctx, int32_t wide);
int lbx_addressing(lbx_ctx* ctx);

/* How the moment-propagation step gets the incoming scattering probabilities
 * s(r+c_l, inv l) (module_moment_propagation.f90:218-229): LBX_MP_RECOMPUTED
 * keeps each node's density and flow (32 bytes) and recomputes them in the
 * step with the set-up's own arithmetic; LBX_MP_STORED keeps the 18
 * probabilities of each node (144 bytes in FP64, 72 in the FP32 forms);
 * LBX_MP_MIXED keeps the 10 of the directions with c_z != 0 (whose neighbours
 * lie a plane away) and recomputes the 8 others from the density and flow.
 * All give the same bits.  LBX_MP_AUTO (default): recomputed for two or more
 * FP64 tracers (their coefficients are shared, and 112 bytes less per node
 * pays), stored for one tracer and in the FP32 forms (where the recomputation
 * costs more than the bytes it saves).  Takes effect at the next
 * lbx_mp_init*.  No counterpart in the reference. */
#define LBX_MP_AUTO       (-1)
#define LBX_MP_STORED     0
#define LBX_MP_RECOMPUTED 1
#define LBX_MP_MIXED      2
int lbx_set_mp_coefficients(lbx_ctx* ctx, int32_t form);

/* Fluid-node count of the context (the MFLUPS denominator). */
int64_t lbx_fluid_nodes(lbx_ctx* ctx);

/* Measurement helper with no counterpart in the reference: read+write
 * bandwidth (GB/s, 1e9 B/s) of a plain streaming copy of `bytes` bytes,
 * iters back-to-back launches timed with HIP events.  It is the achievable
 * ceiling the roofline fractions are compared with, and the known-size
 * kernel that calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE counters. */
int lbx_copy_bandwidth(int64_t bytes, int32_t iters, double* gb_per_s);

/* Measurement helper with no counterpart in the reference: read+write
 * bandwidth (GB/s) of streaming `planes` (19: the populations; 22: with the
 * flux) FP64 planes of `nodes` values in and out, one node per lane, all
 * loads before the stores: the LB step's streams without its gathers, the
 * ceiling of the structure-of-arrays population layout (nodes in
 * [1024, 2^31 - 1]). */
int lbx_planes_bandwidth(int32_t planes, int64_t nodes, int32_t iters, double* gb_per_s);

/* Block until the context's stream is idle. */
int lbx_sync(lbx_ctx* ctx);

/* ----------------------------------------------------------- batch engine
 * Many independent lattices ("instances", each an lb.in-style solve: rest
 * start, two-stage equilibration, then moment propagation of one tracer)
 * resident in HBM at once, advanced together: one LB launch steps every
 * instance in its LB phase, one MP launch every instance in its MP phase,
 * over one packed fluid-node list with per-instance offsets (lbx_batch.hip).
 * Each instance's results are bit-identical to the single-lattice path with
 * the same input (LB fields, P; VACF with 256-node blocks).  No counterpart
 * in the reference (it solves one lattice per process); it serves batched
 * and continuously batched workloads.  A batch is used by one thread. */
typedef struct lbxb lbxb;

/* HBM bytes of a batch of this capacity (nodes of the packed list, full-box
 * map entries, instance slots). */
int64_t lbxb_footprint_bytes(int32_t precision, int64_t capacity_nodes, int64_t capacity_box,
                             int32_t max_instances);
int lbxb_create(int32_t precision, int64_t capacity_nodes, int64_t capacity_box, int32_t max_instances,
                lbxb** out);
int lbxb_destroy(lbxb* b);

/* Admit one instance (HBM admission: LBX_ERR_NOMEM when its fluid nodes,
 * full-box map or instance slot do not fit now).  It starts from rest at
 * rho0; equilibrates with the reference's two-stage loop to target_error
 * (max_lb_steps > 0 caps it), then runs mp_steps moment-propagation steps of
 * a tracer (Db, ka, kd); *id receives its slot. */
int lbxb_add(lbxb* b, int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, double tau, double rho0,
             const double fext[3], double target_error, int64_t max_lb_steps, double Db, double ka, double kd,
             int64_t mp_steps, int32_t* id);
/* Free an instance's slot and nodes (any phase). */
int lbxb_remove(lbxb* b, int32_t id);

/* One round: up to lb_steps LB launches and mp_steps MP launches
 * (each <= 1024), then phase changes (stage switch, end of LB -> MP set-up,
 * end of MP).  *active (may be NULL) = instances not yet done or failed. */
int lbxb_round(lbxb* b, int32_t lb_steps, int32_t mp_steps, int32_t* active);

/* Instance state: phase 1 LB, 2 LB stage switch pending, 3 LB done, 4 MP,
 * 5 done, 6 failed (status = the lbx status); LB steps taken and exit reason,
 * MP steps done. */
#define LBXB_PHASE_LB     1
#define LBXB_PHASE_DONE   5
#define LBXB_PHASE_FAILED 6
int lbxb_info(lbxb* b, int32_t id, int32_t* phase, int32_t* status, int64_t* lb_steps, int32_t* exit_reason,
              int64_t* mp_steps_done);
/* D_eff, the VACF trace (first vacf_rows rows of 3, row 0 = VACF(0)) and the
 * final flow (full box (lx,ly,lz), 0 at solids); any pointer may be NULL. */
int lbxb_result(lbxb* b, int32_t id, double deff[3], double* vacf, int64_t vacf_rows, double* rho, double* jx,
                double* jy, double* jz);

/* Block size (64, 128 or 256 nodes) of the LB and MP launches.  LB results
 * do not depend on it; VACF and D_eff sum per block, so they agree across
 * block sizes to rounding. */
int lbxb_set_block(lbxb* b, int32_t lb_block, int32_t mp_block);

/* lbx_set_check_interval for every instance of a batch (set before the
 * instances run; default 64 in the FP32-storage forms). */
int lbxb_set_check_interval(lbxb* b, int32_t steps);
/* lbx_set_floor_window for the instances added to the batch from now on
 * (default 1000 in the FP32-storage forms; FP64 batches accept only 0). */
int lbxb_set_floor_window(lbxb* b, int64_t steps);
/* How a round's steps are launched: one launch per step on the stream, the
 * same launches replayed from a captured HIP graph (default), or one
 * persistent cooperative launch per phase that loops over the steps with a
 * grid-wide barrier between them (LB block size for both phases).  All give
 * the same bits. */
#define LBXB_LAUNCH_STREAM     0
#define LBXB_LAUNCH_GRAPH      1
#define LBXB_LAUNCH_PERSISTENT 2
int lbxb_set_launch_mode(lbxb* b, int32_t mode);
int lbxb_launch_mode(lbxb* b);
/* Element addressing of the batch's step kernels, as lbx_set_addressing
 * (the capacity decides by itself; wide = 1 forces the 64-bit form). */
int lbxb_set_addressing(lbxb* b, int32_t wide);
int lbxb_addressing(lbxb* b);
/* Autotuner: on the current state, without changing it, times 64-, 128-
 * and 256-node blocks per phase with one launch per step and keeps the
 * fastest, then times those launches replayed from a captured graph and as
 * one persistent cooperative launch per phase, and keeps the launch mode
 * with the least LB + MP time.  times[LBXB_AUTOTUNE_TIMES] (may be NULL) =
 * mean ms per step of LB-64, LB-128, LB-256, MP-64, MP-128, MP-256,
 * LB-graph, MP-graph, LB-persistent, MP-persistent (0: phase not timed). */
#define LBXB_AUTOTUNE_TIMES 10
int lbxb_autotune(lbxb* b, int32_t launches, int32_t* lb_block, int32_t* mp_block, double* times);
/* Device time (HIP events around each phase of each round) and work done. */
int lbxb_stats(lbxb* b, double* lb_ms, int64_t* lb_box_updates, int64_t* lb_fluid_updates, double* mp_ms,
               int64_t* mp_box_updates, int64_t* mp_fluid_updates);
/* Largest free run of 256-node tiles (admission headroom). */
int64_t lbxb_free_tiles(lbxb* b);

#ifdef __cplusplus
}
#endif
#endif
