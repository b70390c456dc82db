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
#include <new>
#include <vector>

#include "../../include/lbx.h"
#include "lbx_device.h"
#include "lbx_tables.h"

#pragma clang fp contract(off)

using namespace lbx;

namespace {

constexpr int kBlock = 256;          // one tile = 256 consecutive packed nodes
constexpr int kSlots = 64;           // striped l2err slots per step
constexpr int kRing = 1024;          // l2err ring slots (>= 2 chunks)
constexpr int kChunk = 256;          // LB steps launched between host checks
constexpr int kMpChunk = 512;        // MP steps per host check
constexpr uint32_t kSolidSlot = 0xFFFFFFFFu;
constexpr uint32_t kIndexMask = 0x7FFFFFFFu;
constexpr uint32_t kNodeInterfacial = 0x80000000u;

// Device-side control block.  Every field has a single writer per launch
// (block 0 of a step kernel, or the one-block reduction kernel), and is read
// by later launches on the same stream, so plain stores suffice.
struct Ctrl {
    int32_t stop;                // LB: a step kernel saw the previous step converge
    int32_t restpart;            // MP: a rest probability < eps (set by mp prep)
    int64_t exit_step;           // LB: that converged step
    int64_t negpop;              // LB: first step with a negative population (-1: none)
    int32_t mp_stop;             // MP: VACF convergence reached
    int32_t pad;
    int64_t mp_conv_step;
    double vacf0[3];
    double deff_s[3], deff_c[3]; // Neumaier running sums of the VACF integral
};

__device__ __forceinline__ int32_t vload(const int32_t* p) { return *(const volatile int32_t*)p; }
__device__ __forceinline__ int64_t vload(const int64_t* p) { return *(const volatile int64_t*)p; }
__device__ __forceinline__ double vload(const double* p) { return *(const volatile double*)p; }

// packed indices of the 18 neighbours r + c_l (l = 1..18) of full-box node r
struct Neighbours {
    uint32_t q[kQ];
    __device__ __forceinline__ void load(const uint32_t* __restrict__ map, uint32_t r, const FastDiv& dlx,
                                         const FastDiv& dly, uint32_t lz) {
        Nbr nb;
        nb.init(r, dlx, dly, lz);
        static_for<1, kQ>([&](auto I) {
            constexpr int L = I;
            q[L] = ld(map, nb.at<kCX[L], kCY[L], kCZ[L]>());
        });
    }
};

// ---------------------------------------------------------------- LB kernels

struct StepArgs {
    const double* __restrict__ fin;
    double* __restrict__ fout;
    double* __restrict__ j;
    const uint32_t* __restrict__ nodes;
    const uint32_t* __restrict__ map;
    unsigned long long* slots;            // [3][kSlots] striped per-step max|dj| bit patterns
    double* l2ring;
    Ctrl* ctrl;
    FastDiv dlx, dly;
    uint32_t lz, Nf;
    uint64_t Np;
    double A, B, C;
    double fmx, fmy, fmz;   // F(t)/2, the force of step t in j(t) (equilibration.f90:293-295)
    double fcx, fcy, fcz;   // F(t+1), the force of the collision that opens step t+1
    int64_t t;
    int64_t check_after;    // stop check on l2err(t-1) only when t-1 > check_after
    double target;
    int32_t write_l2;       // 0 for a redo of a step whose l2err is already known
    int32_t have_prev;      // step t-1 ran in this run: reduce its l2err first
};

// A negative population among a block's nodes is reported as +inf in place of
// its max|dj| (max over bit patterns of non-negative doubles = max of values).
constexpr unsigned long long kNegMark = 0x7FF0000000000000ull;

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// max over the kSlots slots of one step (called by the first wave of a block)
__device__ __forceinline__ unsigned long long slots_max(const unsigned long long* set) {
    return wave_max_u64(*(const volatile unsigned long long*)&set[threadIdx.x & (kSlots - 1)]);
}

// Pull-streaming with half-way bounce-back (equilibration.f90:204-243): the
// population arriving at node i in direction L left r - c_L, or, when that
// node is solid, left i itself in direction inv(L).
// The two candidate sources differ only in the lane's element index, so one
// load per direction serves both.
template <int L, bool Wide>
__device__ __forceinline__ double pull(const View<Wide>& fin, uint32_t Np, uint32_t i, const Neighbours& nb) {
    if constexpr (L == 0) {
        return fin.ld(i);
    } else {
        const uint32_t q = nb.q[kInv[L]];  // r - c_L = r + c_inv(L)
        return fin.ld(q != kSolidSlot ? (uint32_t)L * Np + q : (uint32_t)kInv[L] * Np + i);
    }
}

// One time step, fused: pull-streaming with bounce-back of the post-collision
// populations of step t, the negativity check (equilibration.f90:248),
// density (:254) and flux (:293-300), the l2err contribution (:339-344), then
// the collision that opens step t+1 (module_collision.f90:87-108).
#ifndef LBX_LB_WAVES
#define LBX_LB_WAVES 1
#endif
template <bool Wide>
__global__ __launch_bounds__(kBlock, LBX_LB_WAVES) void k_lb_step(StepArgs a) {
    if (vload(&a.ctrl->stop) || vload(&a.ctrl->negpop) >= 0) return;
    if (a.have_prev) {
        // l2err of step t-1 (equilibration.f90:339-344) from its striped
        // slots, then the negativity (:248) and convergence (:346) verdicts
        __shared__ unsigned long long prev_bits;
        if (threadIdx.x < 64) {
            const unsigned long long v = slots_max(a.slots + ((a.t - 1) % 3) * kSlots);
            if (threadIdx.x == 0) prev_bits = v;
        }
        __syncthreads();
        const bool first = blockIdx.x == 0 && threadIdx.x == 0;
        if (prev_bits == kNegMark) {
            if (first) a.ctrl->negpop = a.t - 1;
            return;
        }
        const double prev = __longlong_as_double((long long)prev_bits);
        if (first) a.l2ring[(a.t - 1) % kRing] = prev;
        if (a.t - 1 > a.check_after && prev <= a.target) {
            if (first) {
                a.ctrl->exit_step = a.t - 1;
                a.ctrl->stop = 1;
            }
            return;
        }
    }
    // clear the slots of step t+1 (last used by step t-2, read by step t-1)
    if (a.write_l2 && blockIdx.x == 0 && threadIdx.x < kSlots) a.slots[((a.t + 1) % 3) * kSlots + threadIdx.x] = 0ull;
    const uint32_t i = xcd_tile() * kBlock + threadIdx.x;
    double l2 = 0.0;
    bool neg = false;
    if (i < a.Nf) {
        const uint32_t Np = (uint32_t)a.Np;
        const View<Wide> fin(a.fin, kQ * a.Np), fout(a.fout, kQ * a.Np), jv(a.j, 3 * a.Np);
        Neighbours nb;
        nb.load(a.map, ld(a.nodes, i) & kIndexMask, a.dlx, a.dly, a.lz);
        double jo[3] = {0.0, 0.0, 0.0};
        if (a.write_l2) {  // j_old, issued before the populations so its latency overlaps theirs
            jo[0] = jv.ld(i);
            jo[1] = jv.ld(Np + i);
            jo[2] = jv.ld(2 * Np + i);
        }
        double f[kQ];
        static_for<0, kQ>([&](auto I) { f[I] = pull<I, Wide>(fin, Np, i, nb); });
        static_for<0, kQ>([&](auto I) { neg = neg || (f[I] < 0.0); });
        double rho = f[0];
        static_for<1, kQ>([&](auto I) { rho = rho + f[I]; });
        double jx = a.fmx, jy = a.fmy, jz = a.fmz;
        static_for<0, kQ>([&](auto I) {
            constexpr int L = I;
            if constexpr (kCX[L] != 0) jx = jx + sgn<kCX[L]>(f[L]);
            if constexpr (kCY[L] != 0) jy = jy + sgn<kCY[L]>(f[L]);
            if constexpr (kCZ[L] != 0) jz = jz + sgn<kCZ[L]>(f[L]);
        });
        if (a.write_l2) {
            const double dx = fabs(jx - jo[0]), dy = fabs(jy - jo[1]), dz = fabs(jz - jo[2]);
            l2 = fmax(fmax(dx, dy), dz);
        }
        jv.st(i, jx);
        jv.st(Np + i, jy);
        jv.st(2 * Np + i, jz);
        const double ux = jx / rho, uy = jy / rho, uz = jz / rho;
        static_for<0, kQ>([&](auto I) {
            constexpr int L = I;
            fout.st((uint32_t)L * Np + i,
                    collide_pop<L>(f[L], rho, jx, jy, jz, ux, uy, uz, a.fcx, a.fcy, a.fcz, a.A, a.B, a.C));
        });
    }
    if (a.write_l2) {
        __shared__ double lds[kBlock / 64];
        l2 = wave_max(l2);
        const bool wneg = __ballot(neg) != 0ull;
        if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = wneg ? -1.0 : l2;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = lds[0];
            bool bneg = m < 0.0;
            for (int w = 1; w < kBlock / 64; ++w) {
                bneg = bneg || lds[w] < 0.0;
                m = fmax(m, lds[w]);
            }
            if (bneg || m > 0.0)
                atomicMax(&a.slots[(a.t % 3) * kSlots + (blockIdx.x & (kSlots - 1))],
                          bneg ? kNegMark : (unsigned long long)__double_as_longlong(m));
        }
    }
}

// l2err of the last step of a launch run, which no later step kernel reads.
__global__ __launch_bounds__(64) void k_lb_finalize(const unsigned long long* slots, double* l2ring, Ctrl* ctrl,
                                                    int64_t t) {
    if (vload(&ctrl->stop) || vload(&ctrl->negpop) >= 0) return;
    const unsigned long long v = slots_max(slots + (t % 3) * kSlots);
    if (threadIdx.x == 0) {
        if (v == kNegMark)
            ctrl->negpop = t;
        else
            l2ring[t % kRing] = __longlong_as_double((long long)v);
    }
}

// Streaming only: the post-streaming state f_s(t) of the step whose input is
// fin, and its density SUM(n,4) (equilibration.f90:254).
__global__ __launch_bounds__(kBlock) void k_extract(const double* __restrict__ fin, double* __restrict__ nout,
                                                    double* __restrict__ rho_out, const uint32_t* __restrict__ nodes,
                                                    const uint32_t* __restrict__ map, FastDiv dlx, FastDiv dly,
                                                    uint32_t lz, uint32_t Nf, uint64_t Np) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= Nf) return;
    Neighbours nb;
    nb.load(map, nodes[i] & kIndexMask, dlx, dly, lz);
    const View<true> fv(fin, kQ * Np);
    double rho = 0.0;
    static_for<0, kQ>([&](auto I) {
        constexpr int L = I;
        const double v = pull<L, true>(fv, (uint32_t)Np, i, nb);
        nout[(uint64_t)L * Np + i] = v;
        rho = (L == 0) ? v : rho + v;
    });
    rho_out[i] = rho;
}

// The collision on packed fields with a uniform force (the opening collision
// of a run), in place: module_collision.f90:87-108.
__global__ __launch_bounds__(kBlock) void k_collide_packed(double* f, const double* __restrict__ rho,
                                                           const double* __restrict__ j, double fx, double fy,
                                                           double fz, uint32_t Nf, uint64_t Np, double A, double B,
                                                           double C) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= Nf) return;
    const double d = rho[i], jx = j[i], jy = j[Np + i], jz = j[2 * Np + i];
    const double ux = jx / d, uy = jy / d, uz = jz / d;
    static_for<0, kQ>([&](auto I) {
        constexpr int L = I;
        const double v = f[(uint64_t)L * Np + i];
        f[(uint64_t)L * Np + i] = collide_pop<L>(v, d, jx, jy, jz, ux, uy, uz, fx, fy, fz, A, B, C);
    });
}

// collide() on whole full-box fields with per-node forces, in place
// (module_collision.f90:15-130, fluid nodes only): the stateless entry point.
__global__ __launch_bounds__(kBlock) void k_collide_box(double* f, const double* __restrict__ rho,
                                                        const double* __restrict__ j,
                                                        const double* __restrict__ F,
                                                        const int8_t* __restrict__ nature, uint32_t N, double A,
                                                        double B, double C) {
    const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= N || nature[r] != 0) return;
    const uint64_t n = N;
    const double d = rho[r], jx = j[r], jy = j[n + r], jz = j[2 * n + r];
    const double fx = F[r], fy = F[n + r], fz = F[2 * n + r];
    const double ux = jx / d, uy = jy / d, uz = jz / d;
    static_for<0, kQ>([&](auto I) {
        constexpr int L = I;
        const double v = f[(uint64_t)L * n + r];
        f[(uint64_t)L * n + r] = collide_pop<L>(v, d, jx, jy, jz, ux, uy, uz, fx, fy, fz, A, B, C);
    });
}

// Rest state, init_simu.f90:31-39: n = rho0*a0 at fluid nodes, j = 0.
__global__ __launch_bounds__(kBlock) void k_init_rest(double* __restrict__ f, double* __restrict__ rho,
                                                      double* __restrict__ j, uint32_t Nf, uint64_t Np, double rho0) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= Nf) return;
    static_for<0, kQ>([&](auto I) {
        constexpr int L = I;
        f[(uint64_t)L * Np + i] = rho0 * a0(L);
    });
    rho[i] = rho0;
    j[i] = 0.0;
    j[Np + i] = 0.0;
    j[2 * Np + i] = 0.0;
}

// ---------------------------------------------------------------- MP kernels

// Per-block triple sums in a fixed order (the block partial of a step kernel).
__device__ __forceinline__ void block_partial3(double v[3], double* part, double* lds) {
    block_sum3<kBlock>(v, lds);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 0] = v[0];
        part[3 * blockIdx.x + 1] = v[1];
        part[3 * blockIdx.x + 2] = v[2];
    }
}

// VACF of one moment-propagation step from the per-block partial sums, in a
// fixed order that does not depend on which kernel runs it (so every run gives
// the same bits), then the running D_eff sum and, optionally, the convergence
// test of module_moment_propagation.f90:284-288.  Called by all kBlock threads
// of one block.  mode 0: VACF(0) of mp init; mode 1: VACF(it) of a step.
__device__ __forceinline__ void mp_reduce_block(const double* part, uint32_t nparts, double* vtrace, Ctrl* ctrl,
                                                int32_t mode, int64_t it, int32_t check_conv, double lim1,
                                                double lim2, double* lds) {
    double s[3] = {0.0, 0.0, 0.0};
    for (uint32_t k = threadIdx.x; k < nparts; k += kBlock)
        for (int c = 0; c < 3; ++c) {
            const double x = *(const volatile double*)&part[3 * k + c];
            s[c] = (k == threadIdx.x) ? x : s[c] + x;
        }
    block_sum3<kBlock>(s, lds);
    if (threadIdx.x != 0) return;
    if (mode == 0) {
        for (int c = 0; c < 3; ++c) {
            ctrl->vacf0[c] = s[c];
            ctrl->deff_s[c] = 0.5 * s[c];
            ctrl->deff_c[c] = 0.0;
        }
        return;
    }
    for (int c = 0; c < 3; ++c) {
        vtrace[c] = s[c];
        // Neumaier update of D = VACF(0)/2 + sum VACF(t)
        const double d = ctrl->deff_s[c], x = s[c], u = d + x;
        ctrl->deff_c[c] += (fabs(d) >= fabs(x)) ? (d - u) + x : (x - u) + d;
        ctrl->deff_s[c] = u;
    }
    if (check_conv && it > 2) {
        bool conv = true;
        for (int c = 0; c < 3; ++c) conv = conv && fabs(s[c]) < lim1 && fabs(s[c]) < lim2;
        if (conv) {
            ctrl->mp_conv_step = it;
            ctrl->mp_stop = 1;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_mp_reduce(const double* __restrict__ part, uint32_t nparts,
                                                      double* vtrace, Ctrl* ctrl, int32_t mode, int64_t it,
                                                      int32_t check_conv, double lim1, double lim2) {
    __shared__ double lds[kBlock / 64 * 3];
    if (mode == 1 && vload(&ctrl->mp_stop)) return;
    mp_reduce_block(part, nparts, vtrace, ctrl, mode, it, check_conv, lim1, lim2, lds);
}

struct MpPrepArgs {
    const double* __restrict__ rho;
    const double* __restrict__ j;
    const uint32_t* __restrict__ nodes;
    const uint32_t* __restrict__ map;
    double* __restrict__ win;
    double* __restrict__ rem;
    double* __restrict__ ustar;
    double* __restrict__ P0;
    double* partials;
    Ctrl* ctrl;
    FastDiv dlx, dly;
    uint32_t lz, Nf;
    uint64_t Np;
    double fx, fy, fz;   // lb.in f_ext (drop_tracers.f90:92)
    double lambda, bw, ka, eps;
    int32_t ads;
};

// update_tracer_population (drop_tracers.f90:97-100) fused with
// moment_propagation::init (module_moment_propagation.f90:105-137) and the
// static part of propagate (:218-249): per fluid node, the scattering
// probabilities s(r,l) = n_l/rho - w_l + lambda w_l / 2 towards fluid
// neighbours, the rest probability rem (minus ka at adsorbing interfacial
// nodes), u* = sum s c, the incoming probabilities win(r,l) = s(r+c_l, inv l),
// P(r,1) and the VACF(0) contribution.
__global__ __launch_bounds__(kBlock) void k_mp_prep(MpPrepArgs a) {
    __shared__ double lds[kBlock / 64 * 3];
    const uint32_t i = xcd_tile() * kBlock + threadIdx.x;
    double vp[3] = {0.0, 0.0, 0.0};
    if (i < a.Nf) {
        const uint32_t e = a.nodes[i];
        Neighbours nb;
        nb.load(a.map, e & kIndexMask, a.dlx, a.dly, a.lz);
        const double rr = a.rho[i];
        const double vx = a.j[i] + a.fx, vy = a.j[a.Np + i] + a.fy, vz = a.j[2 * a.Np + i] + a.fz;
        double us[3] = {0.0, 0.0, 0.0};
        double p1[3] = {0.0, 0.0, 0.0};
        double remv = 1.0;
        static_for<1, kQ>([&](auto I) {
            constexpr int L = I;
            constexpr int Li = kInv[L];
            constexpr int cx = kCX[L], cy = kCY[L], cz = kCZ[L];
            const uint32_t q = nb.q[L];
            if (q != kSolidSlot) {
                const double nt = a0(L) * rr + a1(L) * dot3<cx, cy, cz>(vx, vy, vz);
                const double s = nt / rr - a0(L) + a.lambda * a0(L) * 0.5;
                remv = remv - s;
                if constexpr (cx != 0) us[0] = us[0] + sgn<cx>(s);
                if constexpr (cy != 0) us[1] = us[1] + sgn<cy>(s);
                if constexpr (cz != 0) us[2] = us[2] + sgn<cz>(s);
                if constexpr (cx != 0) vp[0] = vp[0] + a.bw * s;
                if constexpr (cy != 0) vp[1] = vp[1] + a.bw * s;
                if constexpr (cz != 0) vp[2] = vp[2] + a.bw * s;
                const double rn = a.rho[q];
                const double wx = a.j[q] + a.fx, wy = a.j[a.Np + q] + a.fy, wz = a.j[2 * a.Np + q] + a.fz;
                const double ntn = a0(Li) * rn + a1(Li) * dot3<-cx, -cy, -cz>(wx, wy, wz);
                const double sp = ntn / rn - a0(Li) + a.lambda * a0(Li) * 0.5;
                a.win[(uint64_t)(L - 1) * a.Np + i] = sp;
                if constexpr (cx != 0) p1[0] = p1[0] + sgn<-cx>(sp) * a.bw;
                if constexpr (cy != 0) p1[1] = p1[1] + sgn<-cy>(sp) * a.bw;
                if constexpr (cz != 0) p1[2] = p1[2] + sgn<-cz>(sp) * a.bw;
            } else {
                a.win[(uint64_t)(L - 1) * a.Np + i] = 0.0;
            }
        });
        if (a.ads && (e & kNodeInterfacial)) remv = remv - a.ka;
        if (remv < a.eps) a.ctrl->restpart = 1;
        a.rem[i] = remv;
        for (int c = 0; c < 3; ++c) {
            a.ustar[c * a.Np + i] = us[c];
            a.P0[c * a.Np + i] = p1[c];
        }
    }
    block_partial3(vp, a.partials, lds);
}

struct MpStepArgs {
    const double* __restrict__ win;
    const double* __restrict__ rem;
    const double* __restrict__ ustar;
    const uint32_t* __restrict__ nodes;
    const uint32_t* __restrict__ map;
    const double* __restrict__ Pn;
    double* __restrict__ Px;
    const double* __restrict__ Pan;
    double* __restrict__ Pax;
    double* partials;           // this step's per-block VACF partials
    const double* prev_partials; // the previous step's, reduced by block 0 (NULL: none)
    double* prev_vtrace;        // where VACF(it-1) goes
    int64_t it;
    Ctrl* ctrl;
    FastDiv dlx, dly;
    uint32_t lz, Nf;
    uint64_t Np;
    double ka, kd, one_minus_kd;
    int32_t ads;
};

// One moment-propagation step, module_moment_propagation.f90:207-253:
// P(r,t+1) = sum_l P(r+c_l,t) s(r+c_l, inv l) + rem(r) P(r,t) [+ kd Pa],
// Pa(r,t+1) = (1-kd) Pa(r,t) + ka P(r,t) at adsorbing interfacial nodes,
// VACF(t) = sum_r P(r,t) . u*(r).
template <bool Wide>
__global__ __launch_bounds__(kBlock) void k_mp_step(MpStepArgs a) {
    __shared__ double lds[kBlock / 64 * 3];
    if (vload(&a.ctrl->mp_stop)) return;
    if (a.prev_partials && blockIdx.x == 0)
        mp_reduce_block(a.prev_partials, gridDim.x, a.prev_vtrace, a.ctrl, 1, a.it - 1, 0, 0.0, 0.0, lds);
    const uint32_t i = xcd_tile() * kBlock + threadIdx.x;
    double vp[3] = {0.0, 0.0, 0.0};
    if (i < a.Nf) {
        const uint32_t Np = (uint32_t)a.Np;
        const View<Wide> win(a.win, 18 * a.Np), Pn(a.Pn, 3 * a.Np), Px(a.Px, 3 * a.Np), us(a.ustar, 3 * a.Np),
            rem(a.rem, a.Np);
        const uint32_t e = ld(a.nodes, i);
        Neighbours nb;
        nb.load(a.map, e & kIndexMask, a.dlx, a.dly, a.lz);
        double p[3];
        for (int c = 0; c < 3; ++c) {
            p[c] = Pn.ld(c * Np + i);
            vp[c] = us.ld(c * Np + i);
        }
        const double rm = rem.ld(i);
        double acc[3] = {0.0, 0.0, 0.0};
        static_for<1, kQ>([&](auto I) {
            constexpr int L = I;
            const uint32_t q = nb.q[L];
            if (q != kSolidSlot) {
                const double w = win.ld((uint32_t)(L - 1) * Np + i);
                acc[0] = acc[0] + Pn.ld(q) * w;
                acc[1] = acc[1] + Pn.ld(Np + q) * w;
                acc[2] = acc[2] + Pn.ld(2 * Np + q) * w;
            }
        });
        for (int c = 0; c < 3; ++c) vp[c] = p[c] * vp[c];
        if (!a.ads || !(e & kNodeInterfacial)) {
            for (int c = 0; c < 3; ++c) Px.st(c * Np + i, acc[c] + rm * p[c]);
        } else {
            const View<Wide> Pan(a.Pan, 3 * a.Np), Pax(a.Pax, 3 * a.Np);
            for (int c = 0; c < 3; ++c) {
                const double pa = Pan.ld(c * Np + i);
                Px.st(c * Np + i, (acc[c] + rm * p[c]) + pa * a.kd);
                Pax.st(c * Np + i, pa * a.one_minus_kd + p[c] * a.ka);
            }
        }
    }
    block_partial3(vp, a.partials, lds);
}

// Plain streaming copy (16 B per lane per iteration, grid-stride): the
// measured ceiling that the roofline fractions are read against, and the
// known-size kernel that calibrates the HBM traffic counters.
__global__ __launch_bounds__(kBlock) void k_copy(const double2* __restrict__ a, double2* __restrict__ b, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
        b[i] = a[i];
}

}  // namespace

// ====================================================================== host

struct lbx_ctx {
    int32_t lx = 0, ly = 0, lz = 0;
    FastDiv dlx{}, dly{};
    uint32_t N = 0, Nf = 0;
    uint64_t Np = 0;  // packed plane stride
    bool wide = false; // 19*Np*8 >= 2^32: 64-bit addressing of the population buffers
    int32_t prec = LBX_PREC_FP64;
    double tau = 1.0, A = 0, B = 0, C = 0;
    int64_t nfi = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    double* f[2] = {nullptr, nullptr};
    double* j = nullptr;
    double* rho = nullptr;
    uint32_t* nodes = nullptr;
    uint32_t* map = nullptr;
    uint32_t ntiles = 0;
    double* part = nullptr;      // per-block partials, two buffers of 3 per tile
    unsigned long long* slots = nullptr;  // [3][kSlots] l2err slots
    Ctrl* ctrl = nullptr;
    double* l2ring = nullptr;
    int pre = 0;                 // buffer holding the post-streaming state
    std::vector<uint32_t> nodes_h;
    // moment propagation
    bool mp_ready = false;
    double* win = nullptr;
    double* rem = nullptr;
    double* ustar = nullptr;
    double* P[2] = {nullptr, nullptr};
    double* Pa[2] = {nullptr, nullptr};
    double* vtrace = nullptr;
    int mp_now = 0;
    int64_t mp_it = 0;
    double Db = 0, ka = 0, kd = 0, lambda = 0, bw = 0;
    int32_t ads = 0;
    bool mp_stopped = false;
    // timing
    bool timing = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double t_ms[2] = {0, 0};
    int64_t t_n[2] = {0, 0};
    std::vector<double> l2host;
    std::vector<double> l2trace;  // every l2err of the last lbx_equilibrate call
};

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
    dfree(c->win);
    dfree(c->rem);
    dfree(c->ustar);
    dfree(c->P[0]);
    dfree(c->P[1]);
    dfree(c->Pa[0]);
    dfree(c->Pa[1]);
    dfree(c->vtrace);
    c->mp_ready = false;
}

void free_all(lbx_ctx* c) {
    free_mp(c);
    dfree(c->f[0]);
    dfree(c->f[1]);
    dfree(c->j);
    dfree(c->rho);
    dfree(c->nodes);
    dfree(c->map);
    dfree(c->part);
    dfree(c->slots);
    dfree(c->ctrl);
    dfree(c->l2ring);
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

int reset_ctrl(lbx_ctx* c) {
    Ctrl h{};
    h.negpop = -1;
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

StepArgs step_args(lbx_ctx* c, int64_t t, const double fm[3], const double fc[3], int64_t check_after,
                   double target, int write_l2) {
    StepArgs a;
    a.fin = c->f[(t - 1) & 1];
    a.fout = c->f[t & 1];
    a.j = c->j;
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
    a.A = c->A;
    a.B = c->B;
    a.C = c->C;
    a.fmx = fm[0] / 2.0;
    a.fmy = fm[1] / 2.0;
    a.fmz = fm[2] / 2.0;
    a.fcx = fc[0];
    a.fcy = fc[1];
    a.fcz = fc[2];
    a.t = t;
    a.check_after = check_after;
    a.target = target;
    a.write_l2 = write_l2;
    a.have_prev = 0;
    return a;
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
    const uint64_t pop = precision == LBX_PREC_FP32_SHIFT ? 4 : 8;
    uint64_t b = 2 * kQ * Np * pop + 4 * Np * 8 + N * 8 + nblocks(N) * 3 * 8 + sizeof(Ctrl) + kRing * 8;
    if (with_mp) b += (18 + 1 + 3 + 12) * Np * 8 + kMpChunk * 3 * 8;
    return (int64_t)b;
}

int lbx_create(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, double tau, int32_t precision,
               lbx_ctx** out) {
    if (!out) return LBX_ERR_ARG;
    *out = nullptr;
    if (lx <= 0 || ly <= 0 || lz <= 0 || !nature) return LBX_ERR_ARG;
    // packed indices times 8 bytes must fit in 32 bits (the 32-bit offset loads)
    if ((uint64_t)lx * ly * lz >= (1ull << 29)) return LBX_ERR_ARG;
    if (precision != LBX_PREC_FP64) return LBX_ERR_ARG;
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
    c->tau = tau;
    relaxation_coeffs(tau, &c->A, &c->B, &c->C);
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
    c->wide = kQ * c->Np * 8 >= (1ull << 32);
    c->ntiles = nblocks(c->Nf);
    int st = LBX_OK;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) st = LBX_ERR_DEVICE;
    if (!st) st = dalloc(&c->f[0], kQ * c->Np);
    if (!st) st = dalloc(&c->f[1], kQ * c->Np);
    if (!st) st = dalloc(&c->j, 3 * c->Np);
    if (!st) st = dalloc(&c->rho, c->Np);
    if (!st) st = dalloc(&c->nodes, (uint64_t)c->Nf);
    if (!st) st = dalloc(&c->map, (uint64_t)c->N);
    if (!st) st = dalloc(&c->part, 6 * (uint64_t)c->ntiles);
    if (!st) st = dalloc(&c->slots, 3 * kSlots);
    if (!st) st = dalloc(&c->ctrl, 1);
    if (!st) st = dalloc(&c->l2ring, kRing);
    if (!st && (hipEventCreate(&c->ev[0]) != hipSuccess || hipEventCreate(&c->ev[1]) != hipSuccess))
        st = LBX_ERR_DEVICE;
    if (!st) st = hip_status(hipMemcpyAsync(c->nodes, c->nodes_h.data(), c->Nf * sizeof(uint32_t),
                                            hipMemcpyHostToDevice, c->stream));
    if (!st) st = hip_status(hipMemcpyAsync(c->map, map.data(), c->N * sizeof(uint32_t), hipMemcpyHostToDevice,
                                            c->stream));
    if (!st) st = hip_status(hipMemsetAsync(c->f[0], 0, kQ * c->Np * sizeof(double), c->stream));
    if (!st) st = hip_status(hipMemsetAsync(c->f[1], 0, kQ * c->Np * sizeof(double), c->stream));
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
    hipLaunchKernelGGL(k_init_rest, dim3(c->ntiles), dim3(kBlock), 0, c->stream, c->f[0], c->rho, c->j, c->Nf,
                       c->Np, rho0);
    c->pre = 0;
    LBX_TRY(hipGetLastError());
    return hip_status(hipStreamSynchronize(c->stream));
}

int lbx_set_state(lbx_ctx* c, const double* n, const double* rho, const double* jx, const double* jy,
                  const double* jz) {
    if (!c || !n || !rho || !jx || !jy || !jz) return LBX_ERR_ARG;
    const double* planes[kQ];
    for (int l = 0; l < kQ; ++l) planes[l] = n + (size_t)l * c->N;
    LBX_TRY(upload_full(c, planes, kQ, c->f[0]));
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
    if (n) LBX_TRY(download_full(c, c->f[c->pre], kQ, n));
    if (rho) LBX_TRY(download_full(c, c->rho, 1, rho));
    double* jp[3] = {jx, jy, jz};
    for (int a = 0; a < 3; ++a)
        if (jp[a]) LBX_TRY(download_full(c, c->j + a * c->Np, 1, jp[a]));
    return LBX_OK;
}

int lbx_equilibrate(lbx_ctx* c, double target_error, const double fext[3], int64_t max_steps,
                    double* l2err_trace, int64_t trace_cap, int64_t* steps_done, int32_t* exit_reason) {
    if (!c || !fext) return LBX_ERR_ARG;
    if (steps_done) *steps_done = 0;
    if (exit_reason) *exit_reason = 0;
    const double zero[3] = {0.0, 0.0, 0.0};
    const double F[3] = {fext[0], fext[1], fext[2]};
    const uint32_t gt = c->ntiles;
    // opening collision of step 1 (stage 1: no force), into f[0]
    if (c->pre != 0)
        LBX_TRY(hipMemcpyAsync(c->f[0], c->f[c->pre], kQ * c->Np * sizeof(double), hipMemcpyDeviceToDevice,
                               c->stream));
    hipLaunchKernelGGL(k_collide_packed, dim3(gt), dim3(kBlock), 0, c->stream, c->f[0], c->rho, c->j, 0.0, 0.0, 0.0,
                       c->Nf, c->Np, c->A, c->B, c->C);
    LBX_TRY(hipGetLastError());
    LBX_TRY(reset_ctrl(c));
    c->l2trace.clear();
    LBX_TRY(hipMemsetAsync(c->l2ring, 0, kRing * sizeof(double), c->stream));
    LBX_TRY(hipMemsetAsync(c->slots, 0, 3 * kSlots * sizeof(unsigned long long), c->stream));

    int stage = 1;
    int64_t t = 0;            // last completed step
    int64_t check_after = 2;  // stop tests apply to steps > check_after (t > 2, equilibration.f90:346)
    int status = LBX_OK;
    int32_t reason = 0;
    int64_t chunk = 4;  // grows to kChunk: the force switch normally comes at t = 3
    for (;;) {
        int64_t S = chunk;
        chunk = std::min<int64_t>(2 * chunk, kChunk);
        if (max_steps > 0) S = std::min<int64_t>(S, max_steps - t);
        if (S <= 0) {
            reason = LBX_EXIT_CAP;
            break;
        }
        const double* Fs = stage == 1 ? zero : F;
        timer_begin(c);
        for (int64_t k = 1; k <= S; ++k) {
            StepArgs a = step_args(c, t + k, Fs, Fs, check_after, target_error, 1);
            a.have_prev = t + k > 1;
            hipLaunchKernelGGL(c->wide ? k_lb_step<true> : k_lb_step<false>, dim3(gt), dim3(kBlock), 0, c->stream, a);
        }
        timer_stop(c);
        hipLaunchKernelGGL(k_lb_finalize, dim3(1), dim3(kSlots), 0, c->stream, c->slots, c->l2ring, c->ctrl, t + S);
        LBX_TRY(hipGetLastError());
        Ctrl h;
        LBX_TRY(hipMemcpyAsync(c->l2host.data(), c->l2ring, kRing * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
        LBX_TRY(read_ctrl(c, &h));
        int64_t last = t + S;  // steps actually executed in this chunk
        if (h.negpop >= 0) last = std::min<int64_t>(last, h.negpop);
        if (h.stop) last = std::min<int64_t>(last, h.exit_step);
        timer_end(c, LBX_KERNEL_LB, last - t);
        // l2err of the steps that completed
        const int64_t last_l2 = (h.negpop >= 0 && h.negpop <= last) ? last - 1 : last;
        for (int64_t s = t + 1; s <= last_l2; ++s) {
            if ((int64_t)c->l2trace.size() < s) c->l2trace.resize(s);
            c->l2trace[s - 1] = c->l2host[s % kRing];
            if (l2err_trace && s <= trace_cap) l2err_trace[s - 1] = c->l2host[s % kRing];
        }
        if (h.negpop >= 0) {
            status = LBX_ERR_NEG_POP;
            t = h.negpop;
            break;
        }
        if (h.stop) {
            const int64_t e = h.exit_step;
            if (stage == 1) {
                // the first convergence switches the force on at the end of step e
                // (equilibration.f90:377-386): redo the collision that opens e+1
                StepArgs a = step_args(c, e, zero, F, INT64_MAX, target_error, 0);
                LBX_TRY(reset_ctrl(c));
                hipLaunchKernelGGL(c->wide ? k_lb_step<true> : k_lb_step<false>, dim3(gt), dim3(kBlock), 0, c->stream, a);
                LBX_TRY(hipGetLastError());
                stage = 2;
                t = e;
                check_after = e;  // l2err(e) belongs to stage 1
                continue;
            }
            t = e;
            reason = LBX_EXIT_CONVERGED;
            break;
        }
        t += S;
        if (max_steps > 0 && t >= max_steps) {
            // a convergence at the capped step is still reported as such
            const double l2 = c->l2host[t % kRing];
            reason = (stage == 2 && t > check_after && l2 <= target_error) ? LBX_EXIT_CONVERGED : LBX_EXIT_CAP;
            break;
        }
    }
    if (steps_done) *steps_done = t;
    if (exit_reason) *exit_reason = reason;
    if (t > 0) {
        // post-streaming state of step t into the other buffer (f[t&1] holds
        // the unused opening collision of step t+1), density alongside
        const int src = (t - 1) & 1, dst = t & 1;
        hipLaunchKernelGGL(k_extract, dim3(gt), dim3(kBlock), 0, c->stream, c->f[src], c->f[dst], c->rho, c->nodes,
                           c->map, c->dlx, c->dly, c->lz, c->Nf, c->Np);
        LBX_TRY(hipGetLastError());
        c->pre = dst;
    }
    LBX_TRY(hipStreamSynchronize(c->stream));
    return status;
}

static int mp_setup(lbx_ctx* c, const double fext[3], double Db, double ka, double kd, double vacf0[3]) {
    const double eps = DBL_EPSILON;
    if (ka < -eps || kd < -eps) return LBX_ERR_ARG;  // module_moment_propagation.f90:43-44
    if (Db <= eps) return LBX_ERR_ARG;                // drop_tracers.f90:89
    free_mp(c);
    int st = LBX_OK;
    const uint32_t gt = c->ntiles;
    if (!st) st = dalloc(&c->win, 18 * c->Np);
    if (!st) st = dalloc(&c->rem, c->Np);
    if (!st) st = dalloc(&c->ustar, 3 * c->Np);
    if (!st) st = dalloc(&c->P[0], 3 * c->Np);
    if (!st) st = dalloc(&c->P[1], 3 * c->Np);
    if (!st) st = dalloc(&c->vtrace, 3 * (uint64_t)kMpChunk);
    const double K = (std::fabs(kd) <= eps) ? 0.0 : ka / kd;  // :46-50
    c->ads = std::fabs(K) > eps;                                // :52-56
    if (!st && c->ads) st = dalloc(&c->Pa[0], 3 * c->Np);
    if (!st && c->ads) st = dalloc(&c->Pa[1], 3 * c->Np);
    if (st) {
        free_mp(c);
        return st;
    }
    for (int b = 0; b < 2; ++b) {
        LBX_TRY(hipMemsetAsync(c->P[b], 0, 3 * c->Np * sizeof(double), c->stream));
        if (c->ads) LBX_TRY(hipMemsetAsync(c->Pa[b], 0, 3 * c->Np * sizeof(double), c->stream));
    }
    LBX_TRY(reset_ctrl(c));
    c->Db = Db;
    c->ka = ka;
    c->kd = kd;
    c->lambda = 4.0 * Db / kKbt;                                // calc_lambda, :331-337
    const double Pstat = (double)c->Nf + K * (double)c->nfi;    // :97
    c->bw = 1.0 / Pstat;
    MpPrepArgs a;
    a.rho = c->rho;
    a.j = c->j;
    a.nodes = c->nodes;
    a.map = c->map;
    a.win = c->win;
    a.rem = c->rem;
    a.ustar = c->ustar;
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
    a.bw = c->bw;
    a.ka = ka;
    a.eps = eps;
    a.ads = c->ads;
    hipLaunchKernelGGL(k_mp_prep, dim3(gt), dim3(kBlock), 0, c->stream, a);
    hipLaunchKernelGGL(k_mp_reduce, dim3(1), dim3(kBlock), 0, c->stream, c->part, gt, nullptr, c->ctrl, 0,
                       (int64_t)0, 0, 0.0, 0.0);
    LBX_TRY(hipGetLastError());
    Ctrl h;
    LBX_TRY(read_ctrl(c, &h));
    if (vacf0)
        for (int k = 0; k < 3; ++k) vacf0[k] = h.vacf0[k];
    c->mp_ready = true;
    c->mp_now = 0;
    c->mp_it = 0;
    c->mp_stopped = false;
    return LBX_OK;
}

int lbx_mp_init(lbx_ctx* c, const double fext[3], double Db, double ka, double kd, double vacf0[3]) {
    if (!c || !fext) return LBX_ERR_ARG;
    return mp_setup(c, fext, Db, ka, kd, vacf0);
}

int lbx_mp_init_from_flow(lbx_ctx* c, const double* rho, const double* jx, const double* jy, const double* jz,
                          const double fext[3], double Db, double ka, double kd, double vacf0[3]) {
    if (!c || !rho || !jx || !jy || !jz || !fext) return LBX_ERR_ARG;
    LBX_TRY(upload_full(c, &rho, 1, c->rho));
    const double* jp[3] = {jx, jy, jz};
    LBX_TRY(upload_full(c, jp, 3, c->j));
    return mp_setup(c, fext, Db, ka, kd, vacf0);
}

int lbx_mp_run(lbx_ctx* c, int64_t nsteps, int32_t stop_on_convergence, double* vacf_trace, int64_t* steps_done,
               int32_t* converged) {
    if (!c || nsteps < 0) return LBX_ERR_ARG;
    if (steps_done) *steps_done = 0;
    if (converged) *converged = 0;
    if (!c->mp_ready) return LBX_ERR_STATE;
    Ctrl h;
    LBX_TRY(read_ctrl(c, &h));
    if (h.restpart) return LBX_ERR_RESTPART;  // the reference stops in the first propagate call
    if (c->mp_stopped) return LBX_OK;
    const uint32_t gt = c->ntiles;
    const double lim1 = 1.0 / (2.0 * (double)c->lx * (double)c->ly * (double)c->lz / c->Db);  // :284
    const double lim2 = (double)1.e-12f;  // a default-real literal in the reference
    int64_t done = 0;
    std::vector<double> vt(3 * kMpChunk);
    while (done < nsteps) {
        const int64_t S = std::min<int64_t>(kMpChunk, nsteps - done);
        timer_begin(c);
        for (int64_t k = 0; k < S; ++k) {
            const int now = c->mp_now ^ (int)(k & 1);
            MpStepArgs a;
            a.win = c->win;
            a.rem = c->rem;
            a.ustar = c->ustar;
            a.nodes = c->nodes;
            a.map = c->map;
            a.Pn = c->P[now];
            a.Px = c->P[now ^ 1];
            a.Pan = c->ads ? c->Pa[now] : nullptr;
            a.Pax = c->ads ? c->Pa[now ^ 1] : nullptr;
            const int64_t it = c->mp_it + k + 1;
            // partial buffers alternate by step; without a convergence test the
            // previous step's VACF is reduced by block 0 of this step's kernel
            a.partials = c->part + 3 * (uint64_t)gt * (it & 1);
            a.prev_partials = (!stop_on_convergence && k > 0) ? c->part + 3 * (uint64_t)gt * ((it - 1) & 1) : nullptr;
            a.prev_vtrace = c->vtrace + 3 * (k - 1);
            a.it = it;
            a.ctrl = c->ctrl;
            a.dlx = c->dlx;
            a.dly = c->dly;
            a.lz = c->lz;
            a.Nf = c->Nf;
            a.Np = c->Np;
            a.ka = c->ka;
            a.kd = c->kd;
            a.one_minus_kd = 1.0 - c->kd;
            a.ads = c->ads;
            hipLaunchKernelGGL(c->wide ? k_mp_step<true> : k_mp_step<false>, dim3(gt), dim3(kBlock), 0, c->stream, a);
            if (stop_on_convergence)
                hipLaunchKernelGGL(k_mp_reduce, dim3(1), dim3(kBlock), 0, c->stream, a.partials, gt,
                                   c->vtrace + 3 * k, c->ctrl, 1, it, 1, lim1, lim2);
        }
        timer_stop(c);
        if (!stop_on_convergence) {  // the last step's VACF
            const int64_t it = c->mp_it + S;
            hipLaunchKernelGGL(k_mp_reduce, dim3(1), dim3(kBlock), 0, c->stream,
                               c->part + 3 * (uint64_t)gt * (it & 1), gt, c->vtrace + 3 * (S - 1), c->ctrl, 1, it, 0,
                               lim1, lim2);
        }
        LBX_TRY(hipGetLastError());
        LBX_TRY(hipMemcpyAsync(vt.data(), c->vtrace, 3 * S * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        LBX_TRY(read_ctrl(c, &h));
        int64_t ran = S;
        if (h.mp_stop) ran = std::min<int64_t>(S, h.mp_conv_step - c->mp_it);
        timer_end(c, LBX_KERNEL_MP, ran);
        if (vacf_trace) std::memcpy(vacf_trace + 3 * done, vt.data(), 3 * ran * sizeof(double));
        c->mp_it += ran;
        if (ran & 1) c->mp_now ^= 1;
        done += ran;
        if (h.mp_stop) {
            c->mp_stopped = true;
            if (converged) *converged = 1;
            break;
        }
    }
    if (steps_done) *steps_done = done;
    return LBX_OK;
}

int lbx_mp_get(lbx_ctx* c, double* P, double* Pa) {
    if (!c) return LBX_ERR_ARG;
    if (!c->mp_ready) return LBX_ERR_STATE;
    std::vector<double> full(3 * (size_t)c->N);
    auto fetch = [&](const double* src, double* dst) -> int {
        LBX_TRY(download_full(c, src, 3, full.data()));
        for (uint32_t r = 0; r < c->N; ++r)
            for (int a = 0; a < 3; ++a) dst[3 * (size_t)r + a] = full[(size_t)a * c->N + r];
        return LBX_OK;
    };
    if (P) LBX_TRY(fetch(c->P[c->mp_now], P));
    if (Pa) {
        if (c->ads)
            LBX_TRY(fetch(c->Pa[c->mp_now], Pa));
        else
            std::memset(Pa, 0, 3 * (size_t)c->N * sizeof(double));
    }
    return LBX_OK;
}

int lbx_mp_deff(lbx_ctx* c, double deff[3]) {
    if (!c || !deff) return LBX_ERR_ARG;
    if (!c->mp_ready) return LBX_ERR_STATE;
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

int64_t lbx_fluid_nodes(lbx_ctx* c) { return c ? (int64_t)c->Nf : -1; }

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
