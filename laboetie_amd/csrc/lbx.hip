// liblbx: MI355X (gfx950) kernels and C-ABI of the laboetie hot path.
//
// Device layout of one lattice.  Only fluid nodes are stored (solid nodes
// hold n = 0 forever in the reference: they start at 0, collide() skips them
// and the bounce-back swap only moves fluid values into solid slots that the
// next streaming sends straight back), packed in the reference's node order
// (x fastest, then y, then z):
//   nodes[Nf]     full-box index r of each fluid node; bit 31 = interfacial
//   map[N]        full-box index -> packed index, kSolidSlot for solid nodes
//                 (the neighbour lookup; it also replaces the nature array)
//   f[2][19][Nfp] post-collision populations, ping-pong (step t reads
//                 f[(t-1)&1], writes f[t&1]); plane stride Nfp = Nf rounded
//                 up to 64
//   j[3][Nfp]     flux of the last completed step (the next step's j_old)
//   rho[Nfp]      density (written at set-up and when a run ends)
// Moment propagation adds win[18][Nfp] (incoming scattering probabilities),
// rem[Nfp], ustar[3][Nfp], P[2][3][Nfp] and, with adsorption, Pa[2][3][Nfp].
// Every kernel runs one thread per packed fluid node, 256-node tiles, so all
// lanes work and every fetched cache line holds fluid data.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <type_traits>
#include <vector>

#include "../../include/lbx.h"
#include "lbx_device.h"
#include "lbx_kernels.h"
#include "lbx_tables.h"

#pragma clang fp contract(off)

using namespace lbx;


// ====================================================================== host

// The host state of an equilibration between lbx_equilibrate_begin and the
// end of its run (the loop of equilibration.f90:143-491, in launch runs).
struct EqRun {
    bool active = false;    // a run is under way
    bool began = false;     // a run has been started and has ended
    bool flux_sums = false; // per-step sums of j
    double target = 0;
    double F[3] = {0, 0, 0};
    int64_t max_steps = 0;
    int stage = 1;
    int64_t switch_step = INT64_MAX;  // the stage-1 convergence step: F acts from the step after it
    int64_t t = 0;                    // last completed step
    int64_t check_after = 2;          // stop tests apply to steps > check_after (t > 2, equilibration.f90:346)
    int64_t chunk = 4;                // steps of the next launch run
    int status = LBX_OK;
    int32_t reason = 0;
};

struct lbx_ctx {
    int32_t lx = 0, ly = 0, lz = 0;
    FastDiv dlx{}, dly{};
    uint32_t N = 0, Nf = 0;
    uint64_t Np = 0;  // packed plane stride
    bool wide = false; // 19*Np*8 >= 2^32: 64-bit addressing of the population buffers
    bool wide_auto = false;  // what the size alone selects
    bool nt = false;  // LB population stores with the streaming hint (sets beyond the last-level cache)
    int32_t prec = LBX_PREC_FP64;
    double tau = 1.0, A = 0, B = 0, C = 0;
    float A32 = 0, B32 = 0;      // FP32-shift collision: A + B = 1 exactly (fp32_relaxation)
    int32_t check_every = 1;     // stage-2 steps between l2err measurements (kCheckEvery in FP32 storage)
    uint32_t* mp_tile_order = nullptr;  // MP: tiles in band-major order (NULL: packed order)
    int64_t nfi = 0;
    int64_t floor_window = 0;    // FP32-storage forms: l2err floor test (0: off)
    int device = 0;
    hipStream_t stream = nullptr;
    double* f[2] = {nullptr, nullptr};  // FP64 populations (ping-pong)
    float* g[2] = {nullptr, nullptr};   // FP32 storage: deviations g = n - a0 rho0 (ping-pong)
    float* j32 = nullptr;               // FP32-shift: the flux of the last step (j_old)
    double rho0 = 1.0;                  // FP32 storage: the rest density of the deviations
    double* j = nullptr;                // FP64 flux (FP32-shift: of the final state only)
    double* rho = nullptr;
    uint32_t* nodes = nullptr;
    uint32_t* map = nullptr;
    uint32_t ntiles = 0;
    double* part = nullptr;      // per-block VACF partials, two buffers of 3*kMaxTracers per tile
    unsigned long long* slots = nullptr;  // [3][kSlots] l2err slots
    Ctrl* ctrl = nullptr;
    Ctrl ctrl_h{};               // host copy that reset_ctrl uploads from (read by the stream's next copy)
    double* l2ring = nullptr;
    int pre = 0;                 // buffer holding the post-streaming state
    std::vector<uint32_t> nodes_h;
    // moment propagation
    bool mp_ready = false;
    double* win = nullptr;       // FP64 form
    double* ustar = nullptr;
    float* win32 = nullptr;      // FP32-storage forms
    float* ustar32 = nullptr;
    double* rv = nullptr;        // (rho, v) per node: the step recomputes the incoming probabilities from it
    int32_t mp_coef = LBX_MP_AUTO;        // where the next MP set-up's step takes its incoming probabilities
    int32_t mp_coef_on = LBX_MP_STORED;   // ... and the current MP state's (kCoef* in the kernels)
    double* rem = nullptr;
    double* P[2] = {nullptr, nullptr};
    double* Pa[2] = {nullptr, nullptr};
    double* vtrace = nullptr;
    int mp_now = 0;
    int64_t mp_it = 0;
    int32_t ntr = 0;             // tracers propagated together (1..kMaxTracers)
    double Db = 0, lambda = 0;
    TracerParams tp{};
    bool any_ads = false;
    bool mp_stopped = false;
    // a chunk of moment-propagation steps begun by lbx_mp_run_start and not yet waited for
    struct {
        bool active = false;
        int64_t S = 0;               // steps launched
        double* vt_h = nullptr;      // pinned: its VACF trace (3 kMaxTracers kMpChunk doubles)
        Ctrl* ctrl_h = nullptr;      // pinned: the control block after it
    } mpp;
    // timing
    bool timing = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double t_ms[2] = {0, 0};
    int64_t t_n[2] = {0, 0};
    std::vector<double> l2host;
    std::vector<double> l2trace;  // every l2err of the last lbx_equilibrate call
    // a running equilibration (lbx_equilibrate_begin .. the end of the run)
    EqRun eq;
    double* jpart = nullptr;          // flux sums: per-tile partial sums of j, [step % kChunk][tile][3]
    std::vector<double> jpart_h;
    std::vector<double> jsum_trace;   // sum of j over the fluid nodes, 3 per step
    double* jflow = nullptr;          // lbx_flow: the flux of the last step
};

// the forms that store populations as FP32 deviations (FP32-shift: FP32
// collision; FP32-store: FP64 collision), whose moment propagation keeps its
// probabilities in FP32
inline bool f32_storage(int32_t prec) { return prec == LBX_PREC_FP32_SHIFT || prec == LBX_PREC_FP32_STORE; }

namespace {

int hip_status(hipError_t e) {
    if (e == hipSuccess) return LBX_OK;
    return e == hipErrorOutOfMemory ? LBX_ERR_NOMEM : LBX_ERR_DEVICE;
}
int hip_status(int st) { return st; }  // an lbx status passes through

#define LBX_TRY(expr)                            \
    do {                                         \
        int _st = hip_status(expr);              \
        if (_st != LBX_OK) return _st;           \
    } while (0)

uint32_t nblocks(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

// f(std::integral_constant<int, T>) for the run-time tracer count T
template <typename F>
void with_tracers(int ntr, F&& f) {
    switch (ntr) {
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 3: f(std::integral_constant<int, 3>{}); break;
        default: f(std::integral_constant<int, 4>{}); break;
    }
}
static_assert(kMaxTracers == 4, "with_tracers dispatches 1..4");

template <typename T>
int dalloc(T** p, uint64_t count) {
    *p = nullptr;
    if (count == 0) return LBX_OK;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return e == hipErrorOutOfMemory ? LBX_ERR_NOMEM : LBX_ERR_DEVICE;
    }
    return LBX_OK;
}

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void free_mp(lbx_ctx* c) {
    if (c->mpp.active) {  // a begun chunk is dropped once its kernels are done with the buffers
        (void)hipStreamSynchronize(c->stream);
        c->mpp.active = false;
    }
    dfree(c->win);
    dfree(c->rv);
    dfree(c->rem);
    dfree(c->ustar);
    dfree(c->win32);
    dfree(c->ustar32);
    dfree(c->P[0]);
    dfree(c->P[1]);
    dfree(c->Pa[0]);
    dfree(c->Pa[1]);
    dfree(c->vtrace);
    dfree(c->mp_tile_order);
    c->mp_ready = false;
}

#ifndef LBX_MP_BAND
#define LBX_MP_BAND 32
#endif
// MP tile order: bands of kMpBandRows box rows, plane by plane within a band,
// so that the z-neighbour rows a tile gathers were read by tiles that ran just
// before it on the same XCD (0: the packed order).  Each tile still covers
// the same nodes and writes its VACF partial by tile, so the bits do not
// depend on the order.
constexpr int32_t kMpBandRows = LBX_MP_BAND;

int band_tile_order(lbx_ctx* c, int32_t band_rows, uint32_t** out) {
    if (band_rows <= 0 || c->ntiles < 2 || *out) return LBX_OK;
    std::vector<std::pair<uint64_t, uint32_t>> key(c->ntiles);
    const uint64_t lx = (uint64_t)c->lx, ly = (uint64_t)c->ly;
    for (uint32_t t = 0; t < c->ntiles; ++t) {
        const uint64_t r = c->nodes_h[std::min<uint64_t>((uint64_t)t * kBlock, c->Nf - 1)] & kIndexMask;
        const uint64_t y = (r / lx) % ly, z = r / (lx * ly);
        key[t] = {((y / (uint64_t)band_rows) * (uint64_t)c->lz + z) << 32 | t, t};
    }
    std::sort(key.begin(), key.end());
    std::vector<uint32_t> order(c->ntiles);
    for (uint32_t t = 0; t < c->ntiles; ++t) order[t] = key[t].second;
    LBX_TRY(dalloc(out, c->ntiles));
    LBX_TRY(hipMemcpyAsync(*out, order.data(), c->ntiles * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    return hip_status(hipStreamSynchronize(c->stream));  // order is a local
}

void free_all(lbx_ctx* c) {
    free_mp(c);
    if (c->mpp.vt_h) (void)hipHostFree(c->mpp.vt_h);
    if (c->mpp.ctrl_h) (void)hipHostFree(c->mpp.ctrl_h);
    c->mpp.vt_h = nullptr;
    c->mpp.ctrl_h = nullptr;
    dfree(c->f[0]);
    dfree(c->f[1]);
    dfree(c->g[0]);
    dfree(c->g[1]);
    dfree(c->j32);
    dfree(c->j);
    dfree(c->rho);
    dfree(c->nodes);
    dfree(c->map);
    dfree(c->part);
    dfree(c->slots);
    dfree(c->ctrl);
    dfree(c->l2ring);
    dfree(c->jpart);
    dfree(c->jflow);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e), e = nullptr;
    if (c->stream) (void)hipStreamDestroy(c->stream), c->stream = nullptr;
}

// A = 1-1/tau, B = 1/tau, C = 1-1/(2 tau), module_collision.f90:99-101
void relaxation_coeffs(double tau, double* A, double* B, double* C) {
    *A = 1.0 - 1.0 / tau;
    *B = 1.0 / tau;
    *C = 1.0 - 1.0 / (2.0 * tau);
}

// check_after: the step after which an LB stage's stop tests apply; the
// floor-test record starts there (no minimum yet).  The upload reads the
// context's ctrl_h, which stays untouched until the stream has synchronised
// (every caller synchronises before it resets the control block again).
int reset_ctrl(lbx_ctx* c, int64_t check_after = 2) {
    Ctrl& h = c->ctrl_h;
    h = Ctrl{};
    h.negpop = -1;
    for (int k = 0; k < 2; ++k) {
        h.l2min[k] = HUGE_VAL;
        h.l2min_step[k] = check_after;
    }
    return hip_status(hipMemcpyAsync(c->ctrl, &h, sizeof(Ctrl), hipMemcpyHostToDevice, c->stream));
}

int read_ctrl(lbx_ctx* c, Ctrl* h) {
    LBX_TRY(hipMemcpyAsync(h, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, c->stream));
    return hip_status(hipStreamSynchronize(c->stream));
}

// Kernel timing: events on the context's stream bracket a run of back-to-back
// launches (begin before the first, stop right after the last); the elapsed
// time is collected once the host has synchronised anyway.
void timer_begin(lbx_ctx* c) {
    if (c->timing) (void)hipEventRecord(c->ev[0], c->stream);
}
void timer_stop(lbx_ctx* c) {
    if (c->timing) (void)hipEventRecord(c->ev[1], c->stream);
}
void timer_end(lbx_ctx* c, int which, int64_t launches) {
    if (!c->timing || launches <= 0) return;
    (void)hipEventSynchronize(c->ev[1]);
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) {
        c->t_ms[which] += ms;
        c->t_n[which] += launches;
    }
}

template <typename E>
E* pop_buf(lbx_ctx* c, int b) {
    if constexpr (std::is_same<E, float>::value) return c->g[b]; else return c->f[b];
}

template <typename E, typename R = E>
StepArgs<E, R> step_args(lbx_ctx* c, int64_t t, const double fm[3], const double fc[3], int64_t check_after,
                         double target, int write_l2, int32_t check_every, int64_t check_base) {
    StepArgs<E, R> a;
    a.fin = pop_buf<E>(c, (t - 1) & 1);
    a.fout = pop_buf<E>(c, t & 1);
    if constexpr (std::is_same<R, float>::value) a.j = c->j32; else a.j = c->j;
    a.nodes = c->nodes;
    a.map = c->map;
    a.slots = c->slots;
    a.l2ring = c->l2ring;
    a.ctrl = c->ctrl;
    a.dlx = c->dlx;
    a.dly = c->dly;
    a.lz = c->lz;
    a.Nf = c->Nf;
    a.Np = c->Np;
    if constexpr (std::is_same<R, float>::value) {
        a.A = c->A32;
        a.B = c->B32;
    } else {
        a.A = (R)c->A;
        a.B = (R)c->B;
    }
    a.C = (R)c->C;
    a.fmx = (R)(fm[0] / 2.0);
    a.fmy = (R)(fm[1] / 2.0);
    a.fmz = (R)(fm[2] / 2.0);
    a.fcx = (R)fc[0];
    a.fcy = (R)fc[1];
    a.fcz = (R)fc[2];
    a.rho0 = (R)c->rho0;
    a.t = t;
    a.check_after = check_after;
    a.floor_window = c->floor_window;
    a.target = target;
    a.write_l2 = write_l2;
    a.have_prev = 0;
    a.jpart = c->eq.flux_sums ? c->jpart + (uint64_t)((t - 1) % kChunk) * 3 * c->ntiles : nullptr;
    a.check_every = check_every;
    a.check_base = check_base;
    a.check_inv = (R)(1.0 / check_every);
    return a;
}

// one LB step kernel of the context's precision and addressing
void launch_lb_step(lbx_ctx* c, int64_t t, const double fm[3], const double fc[3], int64_t check_after,
                    double target, int write_l2, int have_prev, int32_t check_every = 1, int64_t check_base = 0) {
    const uint32_t gt = c->ntiles;
    if (c->prec == LBX_PREC_FP32_SHIFT) {
        StepArgs<float> a = step_args<float>(c, t, fm, fc, check_after, target, write_l2, check_every, check_base);
        a.have_prev = have_prev;
        auto kern = c->wide ? (c->nt ? k_lb_step<true, float, float, true> : k_lb_step<true, float>)
                            : (c->nt ? k_lb_step<false, float, float, true> : k_lb_step<false, float>);
        hipLaunchKernelGGL(kern, dim3(gt), dim3(kBlock), 0, c->stream, a);
    } else if (c->prec == LBX_PREC_FP32_STORE) {
        StepArgs<float, double> a =
            step_args<float, double>(c, t, fm, fc, check_after, target, write_l2, check_every, check_base);
        a.have_prev = have_prev;
        auto kern = c->wide ? (c->nt ? k_lb_step<true, float, double, true> : k_lb_step<true, float, double>)
                            : (c->nt ? k_lb_step<false, float, double, true> : k_lb_step<false, float, double>);
        hipLaunchKernelGGL(kern, dim3(gt), dim3(kBlock), 0, c->stream, a);
    } else {
        StepArgs<double> a = step_args<double>(c, t, fm, fc, check_after, target, write_l2, check_every, check_base);
        a.have_prev = have_prev;
        auto kern = c->wide ? (c->nt ? k_lb_step<true, double, double, true> : k_lb_step<true, double>)
                            : (c->nt ? k_lb_step<false, double, double, true> : k_lb_step<false, double>);
        hipLaunchKernelGGL(kern, dim3(gt), dim3(kBlock), 0, c->stream, a);
    }
}

// packed <-> full-box copies on the host (not on the hot path)
int download_full(lbx_ctx* c, const double* dsrc, int planes, double* host_full) {
    std::vector<double> tmp((size_t)planes * c->Np);
    LBX_TRY(hipMemcpyAsync(tmp.data(), dsrc, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    LBX_TRY(hipStreamSynchronize(c->stream));
    std::memset(host_full, 0, (size_t)planes * c->N * sizeof(double));
    for (int p = 0; p < planes; ++p)
        for (uint32_t i = 0; i < c->Nf; ++i)
            host_full[(size_t)p * c->N + (c->nodes_h[i] & kIndexMask)] = tmp[(size_t)p * c->Np + i];
    return LBX_OK;
}

int upload_full(lbx_ctx* c, const double* const* host_planes, int planes, double* ddst) {
    std::vector<double> tmp((size_t)planes * c->Np, 0.0);
    for (int p = 0; p < planes; ++p)
        for (uint32_t i = 0; i < c->Nf; ++i) tmp[(size_t)p * c->Np + i] = host_planes[p][c->nodes_h[i] & kIndexMask];
    LBX_TRY(hipMemcpyAsync(ddst, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return hip_status(hipStreamSynchronize(c->stream));
}

// FP32-shift populations <-> full-box FP64 n = a0 rho0 + g (0 at solid nodes)
int download_shift(lbx_ctx* c, const float* gsrc, double* n_full) {
    std::vector<float> tmp((size_t)kQ * c->Np);
    LBX_TRY(hipMemcpyAsync(tmp.data(), gsrc, tmp.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    LBX_TRY(hipStreamSynchronize(c->stream));
    std::memset(n_full, 0, (size_t)kQ * c->N * sizeof(double));
    for (int l = 0; l < kQ; ++l) {
        const double rest = a0(l) * c->rho0;
        for (uint32_t i = 0; i < c->Nf; ++i)
            n_full[(size_t)l * c->N + (c->nodes_h[i] & kIndexMask)] = rest + (double)tmp[(size_t)l * c->Np + i];
    }
    return LBX_OK;
}

int upload_shift(lbx_ctx* c, const double* n_full, float* gdst) {
    std::vector<float> tmp((size_t)kQ * c->Np, 0.0f);
    for (int l = 0; l < kQ; ++l) {
        const double rest = a0(l) * c->rho0;
        for (uint32_t i = 0; i < c->Nf; ++i)
            tmp[(size_t)l * c->Np + i] = (float)(n_full[(size_t)l * c->N + (c->nodes_h[i] & kIndexMask)] - rest);
    }
    LBX_TRY(hipMemcpyAsync(gdst, tmp.data(), tmp.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    return hip_status(hipStreamSynchronize(c->stream));
}

template <typename W>
W* win_buf(lbx_ctx* c) {
    if constexpr (std::is_same<W, float>::value) return c->win32; else return c->win;
}
template <typename W>
W* ustar_buf(lbx_ctx* c) {
    if constexpr (std::is_same<W, float>::value) return c->ustar32; else return c->ustar;
}

// k_mp_prep + the VACF(0) reduction for the context's tracers
template <typename W>
void launch_mp_prep(lbx_ctx* c, const double fext[3], double eps) {
    MpPrepArgs<W> a;
    a.rho = c->rho;
    a.j = c->j;
    a.nodes = c->nodes;
    a.map = c->map;
    a.win = c->mp_coef_on == LBX_MP_RECOMPUTED ? nullptr : win_buf<W>(c);
    a.win_z = c->mp_coef_on == LBX_MP_MIXED ? 1 : 0;
    a.rv = c->mp_coef_on == LBX_MP_STORED ? nullptr : c->rv;
    a.rem = c->rem;
    a.ustar = ustar_buf<W>(c);
    a.P0 = c->P[0];
    a.partials = c->part;
    a.ctrl = c->ctrl;
    a.dlx = c->dlx;
    a.dly = c->dly;
    a.lz = c->lz;
    a.Nf = c->Nf;
    a.Np = c->Np;
    a.fx = fext[0];
    a.fy = fext[1];
    a.fz = fext[2];
    a.lambda = c->lambda;
    a.eps = eps;
    a.tile0 = 0;
    a.tp = c->tp;
    const uint32_t gt = c->ntiles;
    with_tracers(c->ntr, [&](auto TT) {
        constexpr int T = TT;
        hipLaunchKernelGGL((k_mp_prep<T, W>), dim3(gt), dim3(kBlock), 0, c->stream, a);
        hipLaunchKernelGGL(k_mp_reduce<3 * T>, dim3(1), dim3(kBlock), 0, c->stream, c->part, gt, nullptr, c->ctrl, 0,
                           (int64_t)0, 0, 0.0, 0.0);
    });
}

// one k_mp_step launch (and, with the convergence test, its VACF reduction)
template <typename W>
void launch_mp_step(lbx_ctx* c, int now, int64_t it, int64_t k, int32_t stop_on_convergence, double lim1,
                           double lim2) {
    const uint32_t gt = c->ntiles;
    const int nv = 3 * c->ntr;
    MpStepArgs<W> a;
    a.win = win_buf<W>(c);
    a.rem = c->rem;
    a.ustar = ustar_buf<W>(c);
    a.nodes = c->nodes;
    a.map = c->map;
    a.Pn = c->P[now];
    a.Px = c->P[now ^ 1];
    a.Pan = c->any_ads ? c->Pa[now] : nullptr;
    a.Pax = c->any_ads ? c->Pa[now ^ 1] : nullptr;
    // partial buffers alternate by step; the previous step's VACF (and its
    // convergence test) is reduced by block 0 of this step's kernel
    a.partials = c->part + (uint64_t)nv * gt * (it & 1);
    a.prev_partials = k > 0 ? c->part + (uint64_t)nv * gt * ((it - 1) & 1) : nullptr;
    a.prev_vtrace = c->vtrace + nv * (k - 1);
    a.check_conv = stop_on_convergence;
    a.lim1 = lim1;
    a.lim2 = lim2;
    a.win = c->mp_coef_on == LBX_MP_RECOMPUTED ? nullptr : win_buf<W>(c);
    a.rv = c->mp_coef_on == LBX_MP_STORED ? nullptr : c->rv;
    a.lambda = c->lambda;
    a.it = it;
    a.ctrl = c->ctrl;
    a.dlx = c->dlx;
    a.dly = c->dly;
    a.lz = c->lz;
    a.Nf = c->Nf;
    a.Np = c->Np;
    a.any_ads = c->any_ads;
    a.tp = c->tp;
    a.tile_order = c->mp_tile_order;
    with_tracers(c->ntr, [&](auto TT) {
        constexpr int T = TT;
        auto pick = [&](auto CFT) {
            constexpr int CF = decltype(CFT)::value;
            return c->wide ? k_mp_step<true, T, W, CF> : k_mp_step<false, T, W, CF>;
        };
        auto kern = c->mp_coef_on == LBX_MP_RECOMPUTED ? pick(std::integral_constant<int, kCoefRecomputed>{})
                    : c->mp_coef_on == LBX_MP_MIXED   ? pick(std::integral_constant<int, kCoefMixed>{})
                                                      : pick(std::integral_constant<int, kCoefStored>{});
        hipLaunchKernelGGL(kern, dim3(gt), dim3(kBlock), 0, c->stream, a);
    });
}

}  // namespace

// ------------------------------------------------------------------ C-ABI

extern "C" {

const char* lbx_status_string(int status) {
    switch (status) {
    case LBX_OK: return "ok";
    case LBX_ERR_NEG_POP: return "In equilibration, the population n(x,y,z,vel) < 0";
    case LBX_ERR_RESTPART: return "somewhere restpart is negative";
    case LBX_ERR_TAU: return "relaxation_time must be > 0.5";
    case LBX_ERR_NOMEM: return "device memory exhausted (HBM admission refused)";
    case LBX_ERR_ARG: return "invalid argument";
    case LBX_ERR_DEVICE: return "HIP runtime error";
    case LBX_ERR_STATE: return "call out of order";
    case LBX_ERR_SIZE: return "lattice larger than the no-allocation staging area (LBX_NOALLOC_MAX_NODES nodes)";
    default: return "unknown status";
    }
}

int lbx_version(void) { return 100; }

int lbx_d3q19_tables(int32_t c[57], double a0o[19], double a1o[19], double a2o[19], int32_t inv[19]) {
    for (int l = 0; l < kQ; ++l) {
        if (c) {
            c[3 * l] = kCX[l];
            c[3 * l + 1] = kCY[l];
            c[3 * l + 2] = kCZ[l];
        }
        if (a0o) a0o[l] = a0(l);
        if (a1o) a1o[l] = a1(l);
        if (a2o) a2o[l] = a2(l);
        if (inv) inv[l] = kInv[l];
    }
    return LBX_OK;
}

int64_t lbx_footprint_bytes(int32_t lx, int32_t ly, int32_t lz, int32_t precision, int32_t with_mp) {
    if (lx <= 0 || ly <= 0 || lz <= 0) return -1;
    // upper bound: every node fluid
    const uint64_t N = (uint64_t)lx * ly * lz, Np = (N + 63) / 64 * 64;
    const uint64_t pop = f32_storage(precision) ? 4 : 8;  // populations; also win and u* of moment propagation
    uint64_t b = 2 * kQ * Np * pop + (precision == LBX_PREC_FP32_SHIFT ? 3 * Np * 4 : 0) + 4 * Np * 8 + N * 8 +
                 nblocks(N) * 6 * kMaxTracers * 8 + sizeof(Ctrl) + kRing * 8;
    if (with_mp > 0) {  // with_mp tracers, each with adsorption (the upper bound)
        const uint64_t T = (uint64_t)std::min(with_mp, kMaxTracers);
        b += (18 + 3) * Np * pop + Np * 8 + 12 * T * Np * 8 + kMpChunk * 3 * T * 8;
    }
    return (int64_t)b;
}

int lbx_create(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, double tau, int32_t precision,
               lbx_ctx** out) {
    if (!out) return LBX_ERR_ARG;
    *out = nullptr;
    if (lx <= 0 || ly <= 0 || lz <= 0 || !nature) return LBX_ERR_ARG;
    // packed indices times 8 bytes must fit in 32 bits (the 32-bit offset loads)
    if ((uint64_t)lx * ly * lz >= (1ull << 29)) return LBX_ERR_ARG;
    if (precision != LBX_PREC_FP64 && !f32_storage(precision)) return LBX_ERR_ARG;
    if (!(tau >= 0.5)) return LBX_ERR_TAU;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return LBX_ERR_DEVICE;
    lbx_ctx* c = new (std::nothrow) lbx_ctx();
    if (!c) return LBX_ERR_NOMEM;
    c->lx = lx;
    c->ly = ly;
    c->lz = lz;
    c->N = (uint32_t)((uint64_t)lx * ly * lz);
    c->dlx.set((uint32_t)lx);
    c->dly.set((uint32_t)ly);
    c->prec = precision;
    c->floor_window = f32_storage(precision) ? 1000 : 0;
    c->check_every = f32_storage(precision) ? kCheckEvery : 1;
    c->tau = tau;
    relaxation_coeffs(tau, &c->A, &c->B, &c->C);
    fp32_relaxation(tau, &c->A32, &c->B32);
    (void)hipGetDevice(&c->device);
    std::vector<int8_t> itf(c->N);
    lbx_detect_interfacial(lx, ly, lz, nature, itf.data());
    std::vector<uint32_t> map(c->N, kSolidSlot);
    for (uint32_t r = 0; r < c->N; ++r)
        if (nature[r] == 0) {
            map[r] = (uint32_t)c->nodes_h.size();
            c->nodes_h.push_back(r | (itf[r] ? kNodeInterfacial : 0u));
            c->nfi += itf[r] != 0;
        }
    c->Nf = (uint32_t)c->nodes_h.size();
    if (c->Nf == 0) {
        delete c;
        return LBX_ERR_ARG;
    }
    c->Np = ((uint64_t)c->Nf + 63) / 64 * 64;
    // element indices of the population set (19*Np) are 32-bit
    if (kQ * c->Np >= (1ull << 32)) {
        delete c;
        return LBX_ERR_ARG;
    }
    c->wide = c->wide_auto = kQ * c->Np * 8 >= (1ull << 32);
    c->ntiles = nblocks(c->Nf);
    int st = LBX_OK;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) st = LBX_ERR_DEVICE;
    const bool shift = f32_storage(precision);
    // a step reads one population set and writes the other: once the two
    // together exceed the 256 MiB last-level cache, no written line survives
    // there to the next step
    c->nt = 2 * kQ * c->Np * (shift ? sizeof(float) : sizeof(double)) > (uint64_t)LBX_NT_MIN_BYTES;
    if (!st && !shift) st = dalloc(&c->f[0], kQ * c->Np);
    if (!st && !shift) st = dalloc(&c->f[1], kQ * c->Np);
    if (!st && shift) st = dalloc(&c->g[0], kQ * c->Np);
    if (!st && shift) st = dalloc(&c->g[1], kQ * c->Np);
    if (!st && precision == LBX_PREC_FP32_SHIFT) st = dalloc(&c->j32, 3 * c->Np);
    if (!st) st = dalloc(&c->j, 3 * c->Np);
    if (!st) st = dalloc(&c->rho, c->Np);
    if (!st) st = dalloc(&c->nodes, (uint64_t)c->Nf);
    if (!st) st = dalloc(&c->map, (uint64_t)c->N);
    if (!st) st = dalloc(&c->part, 6 * (uint64_t)kMaxTracers * c->ntiles);
    if (!st) st = dalloc(&c->slots, 3 * kSlots);
    if (!st) st = dalloc(&c->ctrl, 1);
    if (!st) st = dalloc(&c->l2ring, kRing);
    if (!st && (hipEventCreate(&c->ev[0]) != hipSuccess || hipEventCreate(&c->ev[1]) != hipSuccess))
        st = LBX_ERR_DEVICE;
    if (!st) st = hip_status(hipMemcpyAsync(c->nodes, c->nodes_h.data(), c->Nf * sizeof(uint32_t),
                                            hipMemcpyHostToDevice, c->stream));
    if (!st) st = hip_status(hipMemcpyAsync(c->map, map.data(), c->N * sizeof(uint32_t), hipMemcpyHostToDevice,
                                            c->stream));
    for (int b = 0; b < 2; ++b) {
        if (!st && !shift) st = hip_status(hipMemsetAsync(c->f[b], 0, kQ * c->Np * sizeof(double), c->stream));
        if (!st && shift) st = hip_status(hipMemsetAsync(c->g[b], 0, kQ * c->Np * sizeof(float), c->stream));
    }
    if (!st && c->j32) st = hip_status(hipMemsetAsync(c->j32, 0, 3 * c->Np * sizeof(float), c->stream));
    if (!st) st = hip_status(hipMemsetAsync(c->j, 0, 3 * c->Np * sizeof(double), c->stream));
    if (!st) st = hip_status(hipMemsetAsync(c->rho, 0, c->Np * sizeof(double), c->stream));
    if (!st) st = reset_ctrl(c);
    if (!st) st = hip_status(hipStreamSynchronize(c->stream));
    if (st) {
        free_all(c);
        delete c;
        return st;
    }
    c->l2host.resize(kRing);
    *out = c;
    return LBX_OK;
}

int lbx_destroy(lbx_ctx* c) {
    if (!c) return LBX_ERR_ARG;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
    return LBX_OK;
}

int lbx_init_rest(lbx_ctx* c, double rho0) {
    if (!c) return LBX_ERR_ARG;
    c->eq.active = false;  // a paused run is abandoned
    if (f32_storage(c->prec)) c->rho0 = rho0;  // the deviations follow the rest state
    hipLaunchKernelGGL(k_init_rest, dim3(c->ntiles), dim3(kBlock), 0, c->stream, c->f[0], c->g[0], c->j32, c->rho,
                       c->j, 0u, c->Nf, c->Np, rho0);
    c->pre = 0;
    LBX_TRY(hipGetLastError());
    return hip_status(hipStreamSynchronize(c->stream));
}

int lbx_set_state(lbx_ctx* c, const double* n, const double* rho, const double* jx, const double* jy,
                  const double* jz) {
    if (!c || !n || !rho || !jx || !jy || !jz) return LBX_ERR_ARG;
    c->eq.active = false;  // a paused run is abandoned
    if (f32_storage(c->prec)) {
        // the rest density of the deviations: the density of the first fluid
        // node (the uniform rest density when the state is a rest state)
        c->rho0 = rho[c->nodes_h[0] & kIndexMask];
        LBX_TRY(upload_shift(c, n, c->g[0]));
    } else {
        const double* planes[kQ];
        for (int l = 0; l < kQ; ++l) planes[l] = n + (size_t)l * c->N;
        LBX_TRY(upload_full(c, planes, kQ, c->f[0]));
    }
    LBX_TRY(upload_full(c, &rho, 1, c->rho));
    const double* jp[3] = {jx, jy, jz};
    LBX_TRY(upload_full(c, jp, 3, c->j));
    c->pre = 0;
    return LBX_OK;
}

int lbx_l2err_trace(lbx_ctx* c, int64_t first_step, int64_t count, double* out, int64_t* available) {
    if (!c || first_step < 1 || count < 0 || (count > 0 && !out)) return LBX_ERR_ARG;
    const int64_t n = (int64_t)c->l2trace.size();
    if (available) *available = n;
    for (int64_t k = 0; k < count; ++k) {
        const int64_t s = first_step + k;
        if (s > n) return LBX_ERR_ARG;
        out[k] = c->l2trace[s - 1];
    }
    return LBX_OK;
}

int lbx_get_state(lbx_ctx* c, double* n, double* rho, double* jx, double* jy, double* jz) {
    if (!c) return LBX_ERR_ARG;
    if (c->eq.active) return LBX_ERR_STATE;  // during a paused run: lbx_flow
    if (n && f32_storage(c->prec)) LBX_TRY(download_shift(c, c->g[c->pre], n));
    if (n && !f32_storage(c->prec)) LBX_TRY(download_full(c, c->f[c->pre], kQ, n));
    if (rho) LBX_TRY(download_full(c, c->rho, 1, rho));
    double* jp[3] = {jx, jy, jz};
    for (int a = 0; a < 3; ++a)
        if (jp[a]) LBX_TRY(download_full(c, c->j + a * c->Np, 1, jp[a]));
    return LBX_OK;
}

// The end of an equilibration run: the post-streaming state of its last step
// into the buffer lbx_get_state reads (f[t&1] holds the unused opening
// collision of step t+1), density alongside.
int end_run(lbx_ctx* c) {
    EqRun& q = c->eq;
    const double zero[3] = {0.0, 0.0, 0.0};
    const uint32_t gt = c->ntiles;
    const bool shift = f32_storage(c->prec);
    const int64_t t = q.t;
    if (t > 0) {
        // post-streaming state of step t into the other buffer (f[t&1] holds
        // the unused opening collision of step t+1), density alongside
        const int src = (t - 1) & 1, dst = t & 1;
        if (shift) {
            // FP64 density and flux of step t from its FP32 deviations; j(t)
            // carries F(t)/2 as in the step kernel
            const double* Ft = t > q.switch_step ? q.F : zero;
            hipLaunchKernelGGL(k_extract_shift, dim3(gt), dim3(kBlock), 0, c->stream, c->g[src], c->g[dst], c->rho,
                               c->j, c->nodes, c->map, c->dlx, c->dly, c->lz, 0u, c->Nf, c->Np, c->rho0,
                               Ft[0] / 2.0, Ft[1] / 2.0, Ft[2] / 2.0);
        } else {
            hipLaunchKernelGGL(k_extract, dim3(gt), dim3(kBlock), 0, c->stream, c->f[src], c->f[dst], c->rho,
                               c->nodes, c->map, c->dlx, c->dly, c->lz, 0u, c->Nf, c->Np);
        }
        LBX_TRY(hipGetLastError());
        c->pre = dst;
    } else {
        c->pre = 0;
    }
    q.active = false;
    q.began = true;
    return hip_status(hipStreamSynchronize(c->stream));
}

int lbx_equilibrate_begin(lbx_ctx* c, double target_error, const double fext[3], int64_t max_steps,
                          int32_t flux_sums) {
    if (!c || !fext || max_steps < 0) return LBX_ERR_ARG;
    if (c->eq.active) LBX_TRY(end_run(c));  // a paused run ends where it is
    const uint32_t gt = c->ntiles;
    const bool shift = f32_storage(c->prec);
    // opening collision of step 1 (stage 1: no force), into buffer 0
    if (c->pre != 0 && !shift)
        LBX_TRY(hipMemcpyAsync(c->f[0], c->f[c->pre], kQ * c->Np * sizeof(double), hipMemcpyDeviceToDevice,
                               c->stream));
    if (c->pre != 0 && shift)
        LBX_TRY(hipMemcpyAsync(c->g[0], c->g[c->pre], kQ * c->Np * sizeof(float), hipMemcpyDeviceToDevice,
                               c->stream));
    if (c->prec == LBX_PREC_FP32_SHIFT) {
        hipLaunchKernelGGL(k_collide_packed_shift<float>, dim3(gt), dim3(kBlock), 0, c->stream, c->g[0], c->rho,
                           c->j, c->j32, 0.0f, 0.0f, 0.0f, 0u, c->Nf, c->Np, c->A32, c->B32, (float)c->C,
                           c->rho0);
    } else if (c->prec == LBX_PREC_FP32_STORE) {  // j_old is j itself
        hipLaunchKernelGGL(k_collide_packed_shift<double>, dim3(gt), dim3(kBlock), 0, c->stream, c->g[0], c->rho,
                           c->j, c->j, 0.0, 0.0, 0.0, 0u, c->Nf, c->Np, c->A, c->B, c->C, c->rho0);
    } else {
        hipLaunchKernelGGL(k_collide_packed, dim3(gt), dim3(kBlock), 0, c->stream, c->f[0], c->rho, c->j, 0.0, 0.0,
                           0.0, 0u, c->Nf, c->Np, c->A, c->B, c->C);
    }
    LBX_TRY(hipGetLastError());
    LBX_TRY(reset_ctrl(c));
    c->l2trace.clear();
    c->jsum_trace.clear();
    LBX_TRY(hipMemsetAsync(c->l2ring, 0, kRing * sizeof(double), c->stream));
    LBX_TRY(hipMemsetAsync(c->slots, 0, 3 * kSlots * sizeof(unsigned long long), c->stream));
    if (flux_sums && !c->jpart) LBX_TRY(dalloc(&c->jpart, (uint64_t)kChunk * 3 * gt));
    if (flux_sums) c->jpart_h.resize((size_t)kChunk * 3 * gt);
    c->eq = EqRun{};
    c->eq.active = true;
    c->eq.target = target_error;
    for (int a = 0; a < 3; ++a) c->eq.F[a] = fext[a];
    c->eq.max_steps = max_steps;
    c->eq.flux_sums = flux_sums != 0;
    c->pre = -1;  // no extracted state until the run ends
    return hip_status(hipStreamSynchronize(c->stream));
}

int lbx_equilibrate_advance(lbx_ctx* c, int64_t until_step, int64_t* steps_done, int32_t* exit_reason) {
    if (!c) return LBX_ERR_ARG;
    EqRun& q = c->eq;
    if (steps_done) *steps_done = q.t;
    if (exit_reason) *exit_reason = q.reason;
    if (!q.active) return q.began ? q.status : LBX_ERR_STATE;
    const double zero[3] = {0.0, 0.0, 0.0};
    const uint32_t gt = c->ntiles;
    for (;;) {
        if (q.max_steps > 0 && q.t >= q.max_steps) {
            // a convergence at the capped step is still reported as such
            const double l2 = q.t > 0 ? c->l2trace[q.t - 1] : HUGE_VAL;
            q.reason = (q.stage == 2 && q.t > q.check_after && l2 <= q.target) ? LBX_EXIT_CONVERGED : LBX_EXIT_CAP;
            break;
        }
        if (q.t >= until_step) {  // paused: the run goes on at the next call
            if (steps_done) *steps_done = q.t;
            return LBX_OK;
        }
        int64_t S = std::min<int64_t>(q.chunk, until_step - q.t);
        q.chunk = std::min<int64_t>(2 * q.chunk, kChunk);  // grows to kChunk: the force switch normally comes at t = 3
        if (q.max_steps > 0) S = std::min<int64_t>(S, q.max_steps - q.t);
        const double* Fs = q.stage == 1 ? zero : q.F;
        // l2err of every step in stage 1; in stage 2 every check_every steps
        // from the switch (every step in FP64)
        const int32_t every = q.stage == 1 ? 1 : c->check_every;
        const int64_t base = q.stage == 1 ? 0 : q.switch_step;
        const int64_t t = q.t;
        timer_begin(c);
        for (int64_t k = 1; k <= S; ++k)
            launch_lb_step(c, t + k, Fs, Fs, q.check_after, q.target, 1, t + k > 1, every, base);
        timer_stop(c);
        hipLaunchKernelGGL(k_lb_finalize, dim3(1), dim3(kSlots), 0, c->stream, c->slots, c->l2ring, c->ctrl, t + S,
                           base, every);
        LBX_TRY(hipGetLastError());
        Ctrl h;
        LBX_TRY(hipMemcpyAsync(c->l2host.data(), c->l2ring, kRing * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
        if (q.flux_sums)
            LBX_TRY(hipMemcpyAsync(c->jpart_h.data(), c->jpart, (size_t)kChunk * 3 * gt * sizeof(double),
                                   hipMemcpyDeviceToHost, c->stream));
        LBX_TRY(read_ctrl(c, &h));
        int64_t last = t + S;  // steps actually executed in this chunk
        if (h.negpop >= 0) last = std::min<int64_t>(last, h.negpop);
        if (h.stop) last = std::min<int64_t>(last, h.exit_step);
        timer_end(c, LBX_KERNEL_LB, last - t);
        // l2err (and flux sums) of the steps that completed
        const int64_t last_l2 = (h.negpop >= 0 && h.negpop <= last) ? last - 1 : last;
        for (int64_t st = t + 1; st <= last_l2; ++st) {
            if ((int64_t)c->l2trace.size() < st) c->l2trace.resize(st);
            c->l2trace[st - 1] = c->l2host[st % kRing];
            if (q.flux_sums) {
                // sum over the tiles' partials in tile order
                const double* pp = c->jpart_h.data() + (size_t)((st - 1) % kChunk) * 3 * gt;
                double sum[3] = {pp[0], pp[1], pp[2]};
                for (uint32_t b = 1; b < gt; ++b)
                    for (int a = 0; a < 3; ++a) sum[a] = sum[a] + pp[3 * (size_t)b + a];
                c->jsum_trace.resize(3 * (size_t)st);
                for (int a = 0; a < 3; ++a) c->jsum_trace[3 * (size_t)(st - 1) + a] = sum[a];
            }
        }
        if (h.negpop >= 0) {
            q.status = LBX_ERR_NEG_POP;
            q.t = h.negpop;
            break;
        }
        if (h.stop) {
            const int64_t e = h.exit_step;
            if (q.stage == 1) {
                // the first convergence switches the force on at the end of step e
                // (equilibration.f90:377-386): redo the collision that opens e+1
                LBX_TRY(reset_ctrl(c, e));
                launch_lb_step(c, e, zero, q.F, INT64_MAX, q.target, 0, 0);
                LBX_TRY(hipGetLastError());
                q.stage = 2;
                q.switch_step = e;
                q.t = e;
                q.check_after = e;  // l2err(e) belongs to stage 1
                continue;
            }
            q.t = e;
            q.reason = h.floor_hit ? LBX_EXIT_FLOOR : LBX_EXIT_CONVERGED;
            break;
        }
        q.t += S;
    }
    LBX_TRY(end_run(c));
    if (steps_done) *steps_done = q.t;
    if (exit_reason) *exit_reason = q.reason;
    return q.status;
}

int lbx_equilibrate(lbx_ctx* c, double target_error, const double fext[3], int64_t max_steps,
                    double* l2err_trace, int64_t trace_cap, int64_t* steps_done, int32_t* exit_reason) {
    if (!c || !fext) return LBX_ERR_ARG;
    if (steps_done) *steps_done = 0;
    if (exit_reason) *exit_reason = 0;
    LBX_TRY(lbx_equilibrate_begin(c, target_error, fext, max_steps, 0));
    const int st = lbx_equilibrate_advance(c, INT64_MAX, steps_done, exit_reason);
    if (l2err_trace)
        for (int64_t s = 0; s < std::min<int64_t>(trace_cap, (int64_t)c->l2trace.size()); ++s)
            l2err_trace[s] = c->l2trace[s];
    return st;
}

int lbx_flux_sums(lbx_ctx* c, int64_t first_step, int64_t count, double* out, int64_t* available) {
    if (!c || first_step < 1 || count < 0 || (count > 0 && !out)) return LBX_ERR_ARG;
    const int64_t n = (int64_t)c->jsum_trace.size() / 3;
    if (available) *available = n;
    if (first_step + count - 1 > n) return LBX_ERR_ARG;
    for (int64_t k = 0; k < 3 * count; ++k) out[k] = c->jsum_trace[3 * (size_t)(first_step - 1) + k];
    return LBX_OK;
}

int lbx_flow(lbx_ctx* c, double* rho, double* jx, double* jy, double* jz) {
    if (!c) return LBX_ERR_ARG;
    if (!c->eq.active) return lbx_get_state(c, nullptr, rho, jx, jy, jz);
    const int64_t t = c->eq.t;
    if (t == 0) return LBX_ERR_STATE;  // no step taken yet: the caller holds the initial state
    if (!c->jflow) LBX_TRY(dalloc(&c->jflow, 4 * c->Np));
    const uint32_t gt = c->ntiles;
    const double zero[3] = {0.0, 0.0, 0.0};
    const double* Ft = t > c->eq.switch_step ? c->eq.F : zero;
    const int src = (t - 1) & 1;
    if (f32_storage(c->prec))
        hipLaunchKernelGGL(k_flow_shift, dim3(gt), dim3(kBlock), 0, c->stream, c->g[src], c->jflow, c->nodes,
                           c->map, c->dlx, c->dly, c->lz, c->Nf, c->Np, c->rho0, Ft[0] / 2.0, Ft[1] / 2.0,
                           Ft[2] / 2.0);
    else
        hipLaunchKernelGGL(k_flow, dim3(gt), dim3(kBlock), 0, c->stream, c->f[src], c->jflow, c->nodes,
                           c->map, c->dlx, c->dly, c->lz, c->Nf, c->Np, Ft[0] / 2.0, Ft[1] / 2.0, Ft[2] / 2.0);
    LBX_TRY(hipGetLastError());
    if (rho) LBX_TRY(download_full(c, c->jflow, 1, rho));
    double* jp[3] = {jx, jy, jz};
    for (int a = 0; a < 3; ++a)
        if (jp[a]) LBX_TRY(download_full(c, c->jflow + (a + 1) * c->Np, 1, jp[a]));
    return LBX_OK;
}

static int mp_setup(lbx_ctx* c, const double fext[3], double Db, int32_t ntr, const double* ka, const double* kd,
                    double* vacf0) {
    const double eps = DBL_EPSILON;
    if (ntr < 1 || ntr > kMaxTracers) return LBX_ERR_ARG;
    for (int t = 0; t < ntr; ++t)
        if (!(ka[t] >= -eps) || !(kd[t] >= -eps)) return LBX_ERR_ARG;  // module_moment_propagation.f90:43-44
    if (!(Db > eps)) return LBX_ERR_ARG;                                  // drop_tracers.f90:89
    free_mp(c);
    TracerParams tp{};
    bool any_ads = false;
    c->Db = Db;
    c->lambda = 4.0 * Db / kKbt;  // calc_lambda, :331-337
    for (int t = 0; t < ntr; ++t) {
        const double K = (std::fabs(kd[t]) <= eps) ? 0.0 : ka[t] / kd[t];  // :46-50
        tp.ads[t] = std::fabs(K) > eps;                                     // :52-56
        const double Pstat = (double)c->Nf + K * (double)c->nfi;           // :97
        tp.bw[t] = 1.0 / Pstat;
        tp.ka[t] = ka[t];
        tp.kd[t] = kd[t];
        tp.one_minus_kd[t] = 1.0 - kd[t];
        any_ads = any_ads || tp.ads[t];
    }
    const uint64_t np3 = 3 * (uint64_t)ntr * c->Np;
    int st = LBX_OK;
    const bool shift = f32_storage(c->prec);
    // the automatic form: the recomputation pays for several FP64 tracers (measured at 128^3 and
    // 256^3); for one tracer, and in the FP32 forms, the stored probabilities are as fast or faster
    c->mp_coef_on = c->mp_coef != LBX_MP_AUTO ? c->mp_coef
                                             : (!shift && ntr >= 2 ? LBX_MP_RECOMPUTED : LBX_MP_STORED);
    // stored probabilities: 18 per node, or the kZDirs with c_z != 0 in the mixed form
    const uint64_t nwin = c->mp_coef_on == LBX_MP_STORED ? 18 : (c->mp_coef_on == LBX_MP_MIXED ? kZDirs : 0);
    if (!st && !shift && nwin) st = dalloc(&c->win, nwin * c->Np);
    if (!st && !shift) st = dalloc(&c->ustar, 3 * c->Np);
    if (!st && shift && nwin) st = dalloc(&c->win32, nwin * c->Np);
    if (!st && shift) st = dalloc(&c->ustar32, 3 * c->Np);
    if (!st && c->mp_coef_on != LBX_MP_STORED) st = dalloc(&c->rv, 4 * c->Np);
    if (!st) st = dalloc(&c->rem, c->Np);
    if (!st) st = dalloc(&c->P[0], np3);
    if (!st) st = dalloc(&c->P[1], np3);
    if (!st) st = dalloc(&c->vtrace, 3 * (uint64_t)ntr * kMpChunk);
    // Pa of a tracer without adsorption is never read or written
    if (!st && any_ads) st = dalloc(&c->Pa[0], np3);
    if (!st && any_ads) st = dalloc(&c->Pa[1], np3);
    if (st) {
        free_mp(c);
        return st;
    }
    for (int b = 0; b < 2; ++b) {
        LBX_TRY(hipMemsetAsync(c->P[b], 0, np3 * sizeof(double), c->stream));
        if (any_ads) LBX_TRY(hipMemsetAsync(c->Pa[b], 0, np3 * sizeof(double), c->stream));
    }
    LBX_TRY(reset_ctrl(c));
    LBX_TRY(band_tile_order(c, kMpBandRows, &c->mp_tile_order));
    c->ntr = ntr;
    c->tp = tp;
    c->any_ads = any_ads;
    if (shift)
        launch_mp_prep<float>(c, fext, eps);
    else
        launch_mp_prep<double>(c, fext, eps);
    LBX_TRY(hipGetLastError());
    Ctrl h;
    LBX_TRY(read_ctrl(c, &h));
    if (vacf0)
        for (int k = 0; k < 3 * ntr; ++k) vacf0[k] = h.vacf0[k];
    c->mp_ready = true;
    c->mp_now = 0;
    c->mp_it = 0;
    c->mp_stopped = false;
    return LBX_OK;
}

int lbx_mp_init(lbx_ctx* c, const double fext[3], double Db, double ka, double kd, double vacf0[3]) {
    if (!c || !fext) return LBX_ERR_ARG;
    if (c->eq.active) return LBX_ERR_STATE;  // the flow of a paused run is not extracted
    return mp_setup(c, fext, Db, 1, &ka, &kd, vacf0);
}

int lbx_mp_init_batch(lbx_ctx* c, const double fext[3], double Db, int32_t ntracers, const double* ka,
                      const double* kd, double* vacf0) {
    if (!c || !fext || !ka || !kd) return LBX_ERR_ARG;
    if (c->eq.active) return LBX_ERR_STATE;  // the flow of a paused run is not extracted
    return mp_setup(c, fext, Db, ntracers, ka, kd, vacf0);
}

int lbx_mp_init_from_flow(lbx_ctx* c, const double* rho, const double* jx, const double* jy, const double* jz,
                          const double fext[3], double Db, double ka, double kd, double vacf0[3]) {
    if (!c || !rho || !jx || !jy || !jz || !fext) return LBX_ERR_ARG;
    LBX_TRY(upload_full(c, &rho, 1, c->rho));
    const double* jp[3] = {jx, jy, jz};
    LBX_TRY(upload_full(c, jp, 3, c->j));
    return mp_setup(c, fext, Db, 1, &ka, &kd, vacf0);
}

// The convergence test's limits (module_moment_propagation.f90:284): 1/(2 N/Db)
// and a default-real literal
static void mp_limits(const lbx_ctx* c, double* lim1, double* lim2) {
    *lim1 = 1.0 / (2.0 * (double)c->lx * (double)c->ly * (double)c->lz / c->Db);
    *lim2 = (double)1.e-12f;
}

// Launches S <= kMpChunk steps from the context's current MP step, the last
// step's VACF reduction (and convergence test), and the copies of the chunk's
// VACF trace (3T S doubles) and of the control block into vt_h, ctrl_h.
static int mp_chunk_launch(lbx_ctx* c, int64_t S, int32_t stop, double* vt_h, Ctrl* ctrl_h) {
    const int nv = 3 * c->ntr;
    double lim1, lim2;
    mp_limits(c, &lim1, &lim2);
    timer_begin(c);
    for (int64_t k = 0; k < S; ++k) {
        const int now = c->mp_now ^ (int)(k & 1);
        const int64_t it = c->mp_it + k + 1;
        if (f32_storage(c->prec))
            launch_mp_step<float>(c, now, it, k, stop, lim1, lim2);
        else
            launch_mp_step<double>(c, now, it, k, stop, lim1, lim2);
    }
    timer_stop(c);
    {  // the last step's VACF (and convergence test)
        const int64_t it = c->mp_it + S;
        with_tracers(c->ntr, [&](auto TT) {
            constexpr int T = TT;
            hipLaunchKernelGGL(k_mp_reduce<3 * T>, dim3(1), dim3(kBlock), 0, c->stream,
                               c->part + (uint64_t)nv * c->ntiles * (it & 1), c->ntiles, c->vtrace + nv * (S - 1),
                               c->ctrl, 1, it, stop, lim1, lim2);
        });
    }
    LBX_TRY(hipGetLastError());
    LBX_TRY(hipMemcpyAsync(vt_h, c->vtrace, nv * S * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return hip_status(hipMemcpyAsync(ctrl_h, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, c->stream));
}

// After a launched chunk of S steps has completed (h: the control block after
// it): the number of its steps that ran (all, or up to the convergence step),
// and the context's MP state moved past them.
static int64_t mp_chunk_done(lbx_ctx* c, int64_t S, const Ctrl& h) {
    int64_t ran = S;
    if (h.mp_stop) ran = std::min<int64_t>(S, h.mp_conv_step - c->mp_it);
    timer_end(c, LBX_KERNEL_MP, ran);
    c->mp_it += ran;
    if (ran & 1) c->mp_now ^= 1;
    if (h.mp_stop) c->mp_stopped = true;
    return ran;
}

int lbx_mp_run(lbx_ctx* c, int64_t nsteps, int32_t stop_on_convergence, double* vacf_trace, int64_t* steps_done,
               int32_t* converged) {
    if (!c || nsteps < 0) return LBX_ERR_ARG;
    if (steps_done) *steps_done = 0;
    if (converged) *converged = 0;
    if (!c->mp_ready || c->mpp.active) return LBX_ERR_STATE;
    if (stop_on_convergence && c->ntr != 1) return LBX_ERR_ARG;  // the test is per tracer
    Ctrl h;
    LBX_TRY(read_ctrl(c, &h));
    if (h.restpart) return LBX_ERR_RESTPART;  // the reference stops in the first propagate call
    if (c->mp_stopped) return LBX_OK;
    const int nv = 3 * c->ntr;
    int64_t done = 0;
    std::vector<double> vt((size_t)nv * kMpChunk);
    while (done < nsteps) {
        const int64_t S = std::min<int64_t>(kMpChunk, nsteps - done);
        LBX_TRY(mp_chunk_launch(c, S, stop_on_convergence, vt.data(), &h));
        LBX_TRY(hip_status(hipStreamSynchronize(c->stream)));
        const int64_t ran = mp_chunk_done(c, S, h);
        if (vacf_trace) std::memcpy(vacf_trace + nv * done, vt.data(), nv * ran * sizeof(double));
        done += ran;
        if (h.mp_stop) {
            if (converged) *converged = 1;
            break;
        }
    }
    if (steps_done) *steps_done = done;
    return LBX_OK;
}

int lbx_mp_run_start(lbx_ctx* c, int64_t nsteps, int32_t stop_on_convergence) {
    if (!c || nsteps < 1 || nsteps > kMpChunk) return LBX_ERR_ARG;
    if (!c->mp_ready || c->mpp.active) return LBX_ERR_STATE;
    if (stop_on_convergence && c->ntr != 1) return LBX_ERR_ARG;
    Ctrl h;
    LBX_TRY(read_ctrl(c, &h));
    if (h.restpart) return LBX_ERR_RESTPART;
    if (!c->mpp.vt_h)
        LBX_TRY(hip_status(hipHostMalloc((void**)&c->mpp.vt_h, 3 * kMaxTracers * kMpChunk * sizeof(double))));
    if (!c->mpp.ctrl_h) LBX_TRY(hip_status(hipHostMalloc((void**)&c->mpp.ctrl_h, sizeof(Ctrl))));
    c->mpp.S = c->mp_stopped ? 0 : nsteps;  // after the convergence step nothing more runs
    if (c->mpp.S) LBX_TRY(mp_chunk_launch(c, c->mpp.S, stop_on_convergence, c->mpp.vt_h, c->mpp.ctrl_h));
    c->mpp.active = true;
    return LBX_OK;
}

int lbx_mp_run_wait(lbx_ctx* c, double* vacf_trace, int64_t* steps_done, int32_t* converged) {
    if (!c) return LBX_ERR_ARG;
    if (steps_done) *steps_done = 0;
    if (converged) *converged = 0;
    if (!c->mpp.active) return LBX_ERR_STATE;
    c->mpp.active = false;
    if (!c->mpp.S) return LBX_OK;
    LBX_TRY(hip_status(hipStreamSynchronize(c->stream)));
    const Ctrl& h = *c->mpp.ctrl_h;
    const int64_t ran = mp_chunk_done(c, c->mpp.S, h);
    if (vacf_trace) std::memcpy(vacf_trace, c->mpp.vt_h, 3 * c->ntr * ran * sizeof(double));
    if (steps_done) *steps_done = ran;
    if (converged && h.mp_stop) *converged = 1;
    return LBX_OK;
}

int lbx_mp_tracers(lbx_ctx* c, int32_t* ntracers) {
    if (!c || !ntracers) return LBX_ERR_ARG;
    *ntracers = c->mp_ready ? c->ntr : 0;
    return LBX_OK;
}

int lbx_mp_get_tracer(lbx_ctx* c, int32_t tracer, double* P, double* Pa) {
    if (!c) return LBX_ERR_ARG;
    if (!c->mp_ready || c->mpp.active) return LBX_ERR_STATE;  // a begun chunk is waited for first
    if (tracer < 0 || tracer >= c->ntr) return LBX_ERR_ARG;
    const int K = 3 * c->ntr;
    std::vector<double> packed((size_t)K * c->Np);
    // the tracer's three components (comp_index layout) into (3,lx,ly,lz), 0 at solids
    auto fetch = [&](const double* src, double* dst) -> int {
        LBX_TRY(hipMemcpyAsync(packed.data(), src, packed.size() * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
        LBX_TRY(hipStreamSynchronize(c->stream));
        std::memset(dst, 0, 3 * (size_t)c->N * sizeof(double));
        for (uint32_t i = 0; i < c->Nf; ++i) {
            const size_t r = c->nodes_h[i] & kIndexMask;
            for (int a = 0; a < 3; ++a) dst[3 * r + a] = packed[comp_index(K, 3 * tracer + a, i, c->Np)];
        }
        return LBX_OK;
    };
    if (P) LBX_TRY(fetch(c->P[c->mp_now], P));
    if (Pa) {
        if (c->tp.ads[tracer])
            LBX_TRY(fetch(c->Pa[c->mp_now], Pa));
        else
            std::memset(Pa, 0, 3 * (size_t)c->N * sizeof(double));
    }
    return LBX_OK;
}

int lbx_mp_get(lbx_ctx* c, double* P, double* Pa) { return lbx_mp_get_tracer(c, 0, P, Pa); }

int lbx_mp_deff_batch(lbx_ctx* c, double* deff) {
    if (!c || !deff) return LBX_ERR_ARG;
    if (!c->mp_ready || c->mpp.active) return LBX_ERR_STATE;  // a begun chunk is waited for first
    Ctrl h;
    LBX_TRY(read_ctrl(c, &h));
    for (int a = 0; a < 3 * c->ntr; ++a) deff[a] = h.deff_s[a] + h.deff_c[a];
    return LBX_OK;
}

int lbx_mp_deff(lbx_ctx* c, double deff[3]) {
    if (!c || !deff) return LBX_ERR_ARG;
    if (!c->mp_ready || c->mpp.active) return LBX_ERR_STATE;  // a begun chunk is waited for first
    Ctrl h;
    LBX_TRY(read_ctrl(c, &h));
    for (int a = 0; a < 3; ++a) deff[a] = h.deff_s[a] + h.deff_c[a];
    return LBX_OK;
}

int lbx_mp_free(lbx_ctx* c) {
    if (!c) return LBX_ERR_ARG;
    (void)hipStreamSynchronize(c->stream);
    free_mp(c);
    return LBX_OK;
}

int lbx_collide(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, double tau, double* n,
                const double* density, const double* jx, const double* jy, const double* jz,
                const double* f_ext_x, const double* f_ext_y, const double* f_ext_z) {
    if (lx <= 0 || ly <= 0 || lz <= 0 || !nature || !n || !density || !jx || !jy || !jz || !f_ext_x ||
        !f_ext_y || !f_ext_z)
        return LBX_ERR_ARG;
    if (!(tau >= 0.5)) return LBX_ERR_TAU;
    const uint64_t N = (uint64_t)lx * ly * lz;
    if (N >= (1ull << 31)) return LBX_ERR_ARG;
    double *dn = nullptr, *dr = nullptr, *dj = nullptr, *dF = nullptr;
    int8_t* dnat = nullptr;
    int st = dalloc(&dn, kQ * N);
    if (!st) st = dalloc(&dr, N);
    if (!st) st = dalloc(&dj, 3 * N);
    if (!st) st = dalloc(&dF, 3 * N);
    if (!st) st = dalloc(&dnat, N);
    hipStream_t s = nullptr;
    if (!st) st = hip_status(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const size_t b = N * sizeof(double);
    if (!st) st = hip_status(hipMemcpyAsync(dn, n, kQ * b, hipMemcpyHostToDevice, s));
    if (!st) st = hip_status(hipMemcpyAsync(dr, density, b, hipMemcpyHostToDevice, s));
    if (!st) st = hip_status(hipMemcpyAsync(dj, jx, b, hipMemcpyHostToDevice, s));
    if (!st) st = hip_status(hipMemcpyAsync(dj + N, jy, b, hipMemcpyHostToDevice, s));
    if (!st) st = hip_status(hipMemcpyAsync(dj + 2 * N, jz, b, hipMemcpyHostToDevice, s));
    if (!st) st = hip_status(hipMemcpyAsync(dF, f_ext_x, b, hipMemcpyHostToDevice, s));
    if (!st) st = hip_status(hipMemcpyAsync(dF + N, f_ext_y, b, hipMemcpyHostToDevice, s));
    if (!st) st = hip_status(hipMemcpyAsync(dF + 2 * N, f_ext_z, b, hipMemcpyHostToDevice, s));
    if (!st) st = hip_status(hipMemcpyAsync(dnat, nature, N, hipMemcpyHostToDevice, s));
    if (!st) {
        double A, B, C;
        relaxation_coeffs(tau, &A, &B, &C);
        hipLaunchKernelGGL(k_collide_box, dim3(nblocks(N)), dim3(kBlock), 0, s, dn, dr, dj, dF, dnat, (uint32_t)N,
                           A, B, C);
        st = hip_status(hipGetLastError());
    }
    if (!st) st = hip_status(hipMemcpyAsync(n, dn, kQ * b, hipMemcpyDeviceToHost, s));
    if (!st) st = hip_status(hipStreamSynchronize(s));
    if (s) (void)hipStreamDestroy(s);
    dfree(dn);
    dfree(dr);
    dfree(dj);
    dfree(dF);
    dfree(dnat);
    return st;
}

}  // extern "C"

// ------------------------------------------- collide() without allocation
// lbx_collide_noalloc stages its fields in one file-scope device array of
// kNoAllocBytes instead of allocating them: each GPU the library runs on holds
// its own copy until the process ends.  In host code `scratch` names a
// host-side stand-in, never the device memory, so the array is reached only
// through the per-device address that hipGetSymbolAddress returns on the
// first call on each device (with that device current).  Every call on a GPU
// shares that GPU's array: calls that overlap on one GPU race.
namespace {

constexpr uint64_t kNoAllocBytes = 64ull << 20;
static_assert((uint64_t)LBX_NOALLOC_MAX_NODES * (kQ * 8 + 8 + 3 * 8 + 3 * 8 + 1) + 5 * 256 <= kNoAllocBytes,
              "LBX_NOALLOC_MAX_NODES must fit the staging array");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current device
int scratch_address(void** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return LBX_ERR_DEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return LBX_ERR_DEVICE;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return LBX_OK;
}

uint64_t up256(uint64_t b) { return (b + 255) / 256 * 256; }

}  // namespace

extern "C" {

int lbx_collide_noalloc(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, double tau, double* n,
                        const double* density, const double* jx, const double* jy, const double* jz,
                        const double* f_ext_x, const double* f_ext_y, const double* f_ext_z) {
    if (lx <= 0 || ly <= 0 || lz <= 0 || !nature || !n || !density || !jx || !jy || !jz || !f_ext_x ||
        !f_ext_y || !f_ext_z)
        return LBX_ERR_ARG;
    if (!(tau >= 0.5)) return LBX_ERR_TAU;
    const uint64_t N = (uint64_t)lx * ly * lz;
    if (N > LBX_NOALLOC_MAX_NODES) return LBX_ERR_SIZE;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return LBX_ERR_DEVICE;
    void* base = nullptr;
    LBX_TRY(scratch_address(&base));
    // the fields of collide() side by side in the array, each 256-byte aligned
    const size_t b = N * sizeof(double);
    char* p = (char*)base;
    double* dn = (double*)p;
    p += up256(kQ * b);
    double* dr = (double*)p;
    p += up256(b);
    double* dj = (double*)p;
    p += up256(3 * b);
    double* dF = (double*)p;
    p += up256(3 * b);
    int8_t* dnat = (int8_t*)p;
    // every byte the kernel reads is written by these copies first, so
    // nothing a previous call left in the array is seen
    hipStream_t s = nullptr;  // the null stream: no stream object to create
    LBX_TRY(hipMemcpyAsync(dn, n, kQ * b, hipMemcpyHostToDevice, s));
    LBX_TRY(hipMemcpyAsync(dr, density, b, hipMemcpyHostToDevice, s));
    LBX_TRY(hipMemcpyAsync(dj, jx, b, hipMemcpyHostToDevice, s));
    LBX_TRY(hipMemcpyAsync(dj + N, jy, b, hipMemcpyHostToDevice, s));
    LBX_TRY(hipMemcpyAsync(dj + 2 * N, jz, b, hipMemcpyHostToDevice, s));
    LBX_TRY(hipMemcpyAsync(dF, f_ext_x, b, hipMemcpyHostToDevice, s));
    LBX_TRY(hipMemcpyAsync(dF + N, f_ext_y, b, hipMemcpyHostToDevice, s));
    LBX_TRY(hipMemcpyAsync(dF + 2 * N, f_ext_z, b, hipMemcpyHostToDevice, s));
    LBX_TRY(hipMemcpyAsync(dnat, nature, N, hipMemcpyHostToDevice, s));
    double A, B, C;
    relaxation_coeffs(tau, &A, &B, &C);
    hipLaunchKernelGGL(k_collide_box, dim3(nblocks(N)), dim3(kBlock), 0, s, dn, dr, dj, dF, dnat, (uint32_t)N, A, B,
                       C);
    LBX_TRY(hipGetLastError());
    LBX_TRY(hipMemcpyAsync(n, dn, kQ * b, hipMemcpyDeviceToHost, s));
    return hip_status(hipStreamSynchronize(s));
}

int lbx_timing_enable(lbx_ctx* c, int32_t on) {
    if (!c) return LBX_ERR_ARG;
    c->timing = on != 0;
    c->t_ms[0] = c->t_ms[1] = 0.0;
    c->t_n[0] = c->t_n[1] = 0;
    return LBX_OK;
}

int lbx_kernel_time(lbx_ctx* c, int32_t which, double* total_ms, int64_t* launches) {
    if (!c || which < 0 || which > 1) return LBX_ERR_ARG;
    if (total_ms) *total_ms = c->t_ms[which];
    if (launches) *launches = c->t_n[which];
    return LBX_OK;
}

int lbx_set_floor_window(lbx_ctx* c, int64_t steps) {
    if (!c || steps < 0) return LBX_ERR_ARG;
    if (steps > 0 && !f32_storage(c->prec)) return LBX_ERR_ARG;
    c->floor_window = steps;
    return LBX_OK;
}

int lbx_set_check_interval(lbx_ctx* c, int32_t steps) {
    if (!c || steps < 1 || steps > 4096 || (steps & (steps - 1)) != 0) return LBX_ERR_ARG;
    if (steps != 1 && !f32_storage(c->prec)) return LBX_ERR_ARG;
    c->check_every = steps;
    return LBX_OK;
}

int lbx_set_addressing(lbx_ctx* c, int32_t wide) {
    if (!c || wide < 0 || wide > 1) return LBX_ERR_ARG;
    c->wide = c->wide_auto || wide == 1;
    return LBX_OK;
}

int lbx_addressing(lbx_ctx* c) { return c ? (c->wide ? 1 : 0) : -1; }

int64_t lbx_fluid_nodes(lbx_ctx* c) { return c ? (int64_t)c->Nf : -1; }

int lbx_planes_bandwidth(int32_t planes, int64_t nodes, int32_t iters, double* gb_per_s) {
    if ((planes != 19 && planes != 22) || nodes < 1024 || nodes > (int64_t)INT32_MAX || iters <= 0 || !gb_per_s)
        return LBX_ERR_ARG;
    const uint64_t n = (uint64_t)nodes, count = (uint64_t)planes * n;
    double *a = nullptr, *b = nullptr;
    int st = dalloc(&a, count);
    if (!st) st = dalloc(&b, count);
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!st) st = hip_status(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    if (!st) st = hip_status(hipEventCreate(&e0));
    if (!st) st = hip_status(hipEventCreate(&e1));
    if (!st) st = hip_status(hipMemsetAsync(a, 0, count * sizeof(double), s));
    if (!st) st = hip_status(hipMemsetAsync(b, 0, count * sizeof(double), s));
    const uint32_t grid = (uint32_t)((n + kBlock - 1) / kBlock);
    auto launch = [&]() {
        if (planes == 19)
            hipLaunchKernelGGL(k_planes<19>, dim3(grid), dim3(kBlock), 0, s, a, b, (uint32_t)n);
        else
            hipLaunchKernelGGL(k_planes<22>, dim3(grid), dim3(kBlock), 0, s, a, b, (uint32_t)n);
    };
    if (!st) {
        launch();  // warm-up
        (void)hipEventRecord(e0, s);
        for (int32_t k = 0; k < iters; ++k) launch();
        (void)hipEventRecord(e1, s);
        st = hip_status(hipEventSynchronize(e1));
    }
    if (!st) {
        float ms = 0.f;
        st = hip_status(hipEventElapsedTime(&ms, e0, e1));
        *gb_per_s = 2.0 * (double)(count * sizeof(double)) * iters / (ms * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    dfree(a);
    dfree(b);
    return st;
}

int lbx_set_mp_coefficients(lbx_ctx* c, int32_t form) {
    if (!c || form < LBX_MP_AUTO || form > LBX_MP_MIXED) return LBX_ERR_ARG;
    c->mp_coef = form;
    return LBX_OK;
}

int lbx_copy_bandwidth(int64_t bytes, int32_t iters, double* gb_per_s) {
    if (bytes < (1 << 20) || iters <= 0 || !gb_per_s) return LBX_ERR_ARG;
    const uint64_t n = (uint64_t)bytes / 16;
    double2 *a = nullptr, *b = nullptr;
    int st = dalloc(&a, n);
    if (!st) st = dalloc(&b, n);
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!st) st = hip_status(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    if (!st) st = hip_status(hipEventCreate(&e0));
    if (!st) st = hip_status(hipEventCreate(&e1));
    if (!st) st = hip_status(hipMemsetAsync(a, 0, n * sizeof(double2), s));
    if (!st) st = hip_status(hipMemsetAsync(b, 0, n * sizeof(double2), s));
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + kBlock - 1) / kBlock, 1u << 20);
    if (!st) {
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kBlock), 0, s, a, b, n);  // warm-up
        (void)hipEventRecord(e0, s);
        for (int32_t k = 0; k < iters; ++k) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kBlock), 0, s, a, b, n);
        (void)hipEventRecord(e1, s);
        st = hip_status(hipEventSynchronize(e1));
    }
    if (!st) {
        float ms = 0.f;
        st = hip_status(hipEventElapsedTime(&ms, e0, e1));
        *gb_per_s = 2.0 * (double)(n * sizeof(double2)) * iters / (ms * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    dfree(a);
    dfree(b);
    return st;
}

int lbx_sync(lbx_ctx* c) {
    if (!c) return LBX_ERR_ARG;
    return hip_status(hipStreamSynchronize(c->stream));
}

}  // extern "C"
