// Device code of liblbx shared by the single-lattice path (lbx.hip) and the
// multi-instance batch path (lbx_batch.hip): the control block, neighbour
// lookup, the LB and moment-propagation work of one block of nodes
// (lb_step_tile, mp_prep_tile, mp_step_tile, parameterised by the block size)
// and the small kernels both paths launch.  A lattice's nodes are a range of
// the packed fluid-node list; both paths run the same per-node arithmetic, so
// an instance gives the same bits alone or in a batch.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "lbx_device.h"
#include "lbx_tables.h"

#pragma clang fp contract(off)

namespace {

using namespace lbx;


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
    int32_t floor_hit;           // LB (FP32-shift): the stop at exit_step was the l2err floor test
    int32_t pad2;
    double l2min[2];             // LB floor test: running minimum of l2err of the stage up to step s,
    int64_t l2min_step[2];       // the step that set it, in slot s & 1,
    double l2max[2];             // and the running maximum
    double vacf0[12];            // 3 per tracer (at most 4)
    double deff_s[12], deff_c[12]; // Neumaier running sums of the VACF integrals
};

// A value another block or an earlier launch wrote (control fields, l2err
// slots, VACF partials): a relaxed device-scope atomic load, which reads past
// this CU's L1 like a volatile one but, unlike it, does not wait for each load
// before issuing the next (a volatile load is followed by a full wait).
template <typename T>
__device__ __forceinline__ T vload(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A control flag as thread 0 of the block read it, the same for every wave:
// another block may write the flag while this block's waves read it, and a
// block whose waves disagreed would split at its next barrier.
__device__ __forceinline__ bool block_flag(bool v) {
    __shared__ int32_t flag;
    __syncthreads();  // the previous use of `flag` by this block is over
    if (threadIdx.x == 0) flag = v ? 1 : 0;
    __syncthreads();
    return flag != 0;
}

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

// Populations are stored as E, the step's arithmetic, its flux and step
// constants are R: E = R = double for the FP64 form; E = float (deviations
// g = n - a0 rho0 from the rest state) with R = float for the FP32-shift form
// and R = double for the FP32-store form.
template <typename E, typename R = E>
struct StepArgs {
    const E* __restrict__ fin;
    E* __restrict__ fout;
    R* __restrict__ j;
    const uint32_t* __restrict__ nodes;
    const uint32_t* __restrict__ map;
    unsigned long long* slots;            // [3][kSlots] striped per-step max|dj| bit patterns
    double* l2ring;
    Ctrl* ctrl;
    FastDiv dlx, dly;
    uint32_t lz, Nf;
    uint64_t Np;
    R A, B, C;
    R fmx, fmy, fmz;   // F(t)/2, the force of step t in j(t) (equilibration.f90:293-295)
    R fcx, fcy, fcz;   // F(t+1), the force of the collision that opens step t+1
    R rho0;            // FP32 storage: the rest density of the deviations
    int64_t t;
    int64_t check_after;    // stop check on l2err(t-1) only when t-1 > check_after
    int64_t floor_window;   // > 0: also stop when l2err set no new minimum for this many steps
    double target;
    int32_t write_l2;       // 0 for a redo of a step whose l2err is already known
    int32_t have_prev;      // step t-1 ran in this run: reduce its l2err first
    // Steps whose l2err is measured (is_check): every step when check_every
    // is 1 (FP64, and stage 1 of the FP32 storage forms); else the steps
    // check_base + m check_every (check_every a power of two), where l2err is
    // max|j(t) - j(t - check_every)| / check_every and j is kept only there
    int32_t check_every = 1;
    int32_t pad_ = 0;
    int64_t check_base = 0;
    R check_inv = R(1);     // 1 / check_every
    // non-NULL: each block writes the sum of j(t) over its nodes (3 doubles,
    // in a fixed order) to jpart[3 * tile ...] (the per-step total mass flux)
    double* jpart = nullptr;
};

// A = 1 - 1/tau and B = 1/tau rounded to FP32 with A + B = 1 exactly, so that
// the FP32 relaxation keeps the momentum (rounded one by one they can miss 1
// by an ulp, a spurious force like the one a1_f32 removes).  1 - B is exact
// for B in [1/2, 2]; for smaller B, B is re-derived from the rounded A, which
// is exact for A in [1/2, 1].
inline void fp32_relaxation(double tau, float* A, float* B) {
    float b = (float)(1.0 / tau);
    const float a = 1.0f - b;
    if (!(b >= 0.5f)) b = 1.0f - a;
    *A = a;
    *B = b;
}

// The number of steps between the l2err measurements of the FP32 storage
// forms' stage 2 (a power of two)
constexpr int32_t kCheckEvery = 64;

// In the FP32 storage forms, j(t) - j(t-1) at a node is of the order of the
// rounding of j itself once the flow is nearly steady, so the per-step max|dj|
// settles on a rounding floor well above what the flow's decay would give and
// a stage would end early on the floor test.  Over check_every steps the decay
// adds up while the rounding does not, so the difference divided by
// check_every follows the decay down by that factor more; it also leaves j out
// of the other steps' memory traffic.
__host__ __device__ __forceinline__ bool is_check(int64_t s, int64_t base, int32_t every) {
    return every <= 1 || (s - base) % every == 0;
}

// A negative population among a block's nodes is reported as +inf in place of
// its max|dj| (max over bit patterns of non-negative doubles = max of values).
constexpr unsigned long long kNegMark = 0x7FF0000000000000ull;

// the floor test needs l2err this far below the stage's maximum
constexpr double kFloorDrop = 1e-2;

// maximum over the 64 lanes of a wave, returned to every lane (all 64 lanes
// active): DPP row steps, then the four row maxima
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
    const auto step = [&](auto ctrl) {
        constexpr int C = decltype(ctrl)::value;
        const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)v, C, 0xf, 0xf, false);
        const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(v >> 32), C, 0xf, 0xf, false);
        const unsigned long long o = ((unsigned long long)hi << 32) | lo;
        v = o > v ? o : v;
    };
    step(std::integral_constant<int, 0xB1>{});
    step(std::integral_constant<int, 0x4E>{});
    step(std::integral_constant<int, 0x141>{});
    step(std::integral_constant<int, 0x140>{});
    unsigned long long m = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, 16 * r);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), 16 * r);
        const unsigned long long o = ((unsigned long long)hi << 32) | lo;
        m = o > m ? o : m;
    }
    return m;
}

// max over the kSlots slots of one step (called by the first wave of a block)
__device__ __forceinline__ unsigned long long slots_max(const unsigned long long* set) {
    return wave_max_u64(vload(&set[threadIdx.x & (kSlots - 1)]));
}

// Per-block sums of NV values in a fixed order, written to out[0 .. NV-1]
// (value a by thread a).
template <int NV, int BLOCK>
__device__ __forceinline__ void block_partial(const double (&v)[NV], double* out, double* lds /* [BLOCK/64*NV] */) {
    static_assert(NV <= BLOCK, "one thread per value");
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double s[NV];
#pragma unroll
    for (int a = 0; a < NV; ++a) s[a] = v[a];
    wave_sums<NV>(s);
    if (lane == 0)
#pragma unroll
        for (int a = 0; a < NV; ++a) lds[w * NV + a] = s[a];
    __syncthreads();
    if (threadIdx.x < NV) {  // value a summed over the waves in wave order, by thread a
        const int a = threadIdx.x;
        double s = lds[a];
        for (int k = 1; k < BLOCK / 64; ++k) s = s + lds[k * NV + a];
        out[a] = s;
    }
    __syncthreads();
}

// Pull-streaming with half-way bounce-back (equilibration.f90:204-243): the
// population arriving at node i in direction L left r - c_L, or, when that
// node is solid, left i itself in direction inv(L).
// The two candidate sources differ only in the lane's element index, so one
// load per direction serves both.
template <int L, bool Wide, typename E>
__device__ __forceinline__ E pull(const View<Wide, E>& fin, uint32_t Np, uint32_t i, const Neighbours& nb) {
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
// E = float: the FP32-shift form; the l2err is still a max of FP64 values.
#ifndef LBX_MP_WAVES
#define LBX_MP_WAVES 1
#endif
#ifndef LBX_LB_WAVES
#define LBX_LB_WAVES 1
#endif
//
// lb_step_tile is one block's share: `tile` is the block's tile (BLOCK
// consecutive packed nodes, i = tile*BLOCK + lane, valid below a.Nf), `stripe`
// picks its l2err slot, and `first_block` is the one block of the lattice that
// writes the step's control fields and clears the next step's slots.
template <int BLOCK, bool Wide, typename E, typename R = E, bool NT = false>
__device__ __forceinline__ void lb_step_tile(const StepArgs<E, R>& a, uint32_t tile, uint32_t stripe,
                                             bool first_block) {
    constexpr bool kShift = std::is_same<E, float>::value;  // deviations stored in FP32
    const uint32_t i = tile * BLOCK + threadIdx.x;
    // Each wave's work is a chain of dependent memory round trips, so
    // independent requests go out together: the node word, the stop flags
    // and the previous step's l2err slots (read whether or not a flag is set).
    const uint32_t node = i < a.Nf ? ld(a.nodes, i) : 0u;
    __shared__ int32_t stop_flag;
    __shared__ unsigned long long prev_bits;
    __syncthreads();  // a previous step of this block (persistent launches) has read both
    if (threadIdx.x < 64) {
        const bool stop = ((int)(vload(&a.ctrl->stop) != 0) | (int)(vload(&a.ctrl->negpop) >= 0)) != 0;
        const unsigned long long v = a.have_prev ? slots_max(a.slots + ((a.t - 1) % 3) * kSlots) : 0ull;
        if (threadIdx.x == 0) {
            stop_flag = stop ? 1 : 0;
            prev_bits = v;
        }
    }
    __syncthreads();
    if (stop_flag) return;
    if (a.have_prev) {
        // l2err of step t-1 (equilibration.f90:339-344) from its striped
        // slots, then the negativity (:248) and convergence (:346) verdicts
        const bool first = first_block && threadIdx.x == 0;
        if (prev_bits == kNegMark) {
            if (first) a.ctrl->negpop = a.t - 1;
            return;
        }
        // a step without a measurement (is_check false) records NaN and
        // takes no stop decision
        const bool chk = is_check(a.t - 1, a.check_base, a.check_every);
        const double prev = chk ? __longlong_as_double((long long)prev_bits) : __builtin_nan("");
        if (first) a.l2ring[(a.t - 1) % kRing] = prev;
        bool floor = false;
        if (a.floor_window > 0 && a.t - 1 > a.check_after) {
            // record of the stage up to step t-1 from the one up to t-2 (every
            // block derives the same record; block 0 stores it for step t+1)
            const int s0 = (int)((a.t - 2) & 1), s1 = (int)((a.t - 1) & 1);
            double m = vload(&a.ctrl->l2min[s0]), mx = vload(&a.ctrl->l2max[s0]);
            int64_t ms = vload(&a.ctrl->l2min_step[s0]);
            if (chk) {
                if (prev < m) {
                    m = prev;
                    ms = a.t - 1;
                }
                mx = fmax(mx, prev);
            }
            if (first) {
                a.ctrl->l2min[s1] = m;
                a.ctrl->l2min_step[s1] = ms;
                a.ctrl->l2max[s1] = mx;
            }
            // a floor: no new minimum for floor_window steps after l2err has
            // fallen 100-fold below the stage's largest value (a plateau of a
            // still-accelerating flow is not one)
            floor = chk && (a.t - 1) - ms >= a.floor_window && m <= kFloorDrop * mx;
        }
        if (chk && a.t - 1 > a.check_after && (prev <= a.target || floor)) {
            if (first) {
                a.ctrl->exit_step = a.t - 1;
                a.ctrl->floor_hit = prev <= a.target ? 0 : 1;
                a.ctrl->stop = 1;
            }
            return;
        }
    }
    // clear the slots of step t+1 (last used by step t-2, read by step t-1)
    if (a.write_l2 && first_block && threadIdx.x < kSlots) a.slots[((a.t + 1) % 3) * kSlots + threadIdx.x] = 0ull;
    double l2 = 0.0;
    bool neg = false;
    double jsum[3] = {0.0, 0.0, 0.0};  // this node's j(t), for the flux sums
    if (i < a.Nf) {
        const uint32_t Np = (uint32_t)a.Np;
        const View<Wide, E> fin(a.fin, kQ * a.Np), fout(a.fout, kQ * a.Np);
        const View<Wide, R> jv(a.j, 3 * a.Np);
        Neighbours nb;
        nb.load(a.map, node & kIndexMask, a.dlx, a.dly, a.lz);
        // l2err of this step is measured: always in FP64; in the FP32 storage
        // forms only at the check steps, the only ones that read and keep j
        const bool chk = a.write_l2 && is_check(a.t, a.check_base, a.check_every);
        R jo[3] = {R(0), R(0), R(0)};
        if (chk) {  // j_old, issued before the populations so its latency overlaps theirs
            jo[0] = jv.ld(i);
            jo[1] = jv.ld(Np + i);
            jo[2] = jv.ld(2 * Np + i);
        }
        R f[kQ];
        static_for<0, kQ>([&](auto I) { f[I] = (R)pull<I, Wide, E>(fin, Np, i, nb); });
        if constexpr (kShift)  // n < 0  <=>  g < -a0 rho0
            static_for<0, kQ>([&](auto I) { neg = neg || (f[I] < -((R)a0(I) * a.rho0)); });
        else
            static_for<0, kQ>([&](auto I) { neg = neg || (f[I] < 0.0); });
        R rho = f[0];  // FP32 storage: drho = rho - rho0
        static_for<1, kQ>([&](auto I) { rho = rho + f[I]; });
        R jx = a.fmx, jy = a.fmy, jz = a.fmz;
        static_for<0, kQ>([&](auto I) {
            constexpr int L = I;
            if constexpr (kCX[L] != 0) jx = jx + sgn<kCX[L]>(f[L]);
            if constexpr (kCY[L] != 0) jy = jy + sgn<kCY[L]>(f[L]);
            if constexpr (kCZ[L] != 0) jz = jz + sgn<kCZ[L]>(f[L]);
        });
        if (chk) {
            R dx = fabs(jx - jo[0]), dy = fabs(jy - jo[1]), dz = fabs(jz - jo[2]);
            if constexpr (kShift) {  // exact: check_inv is a power of two
                dx = dx * a.check_inv;
                dy = dy * a.check_inv;
                dz = dz * a.check_inv;
            }
            l2 = (double)fmax(fmax(dx, dy), dz);
        }
        if (!kShift || chk) {  // FP64: j of every step is the run's output flux
            jv.st(i, jx);
            jv.st(Np + i, jy);
            jv.st(2 * Np + i, jz);
        }
        jsum[0] = (double)jx;
        jsum[1] = (double)jy;
        jsum[2] = (double)jz;
        if constexpr (kShift) {
            const R drho = rho, full = a.rho0 + drho;
            R ux, uy, uz;
            shift_velocity<R>(jx, jy, jz, full, ux, uy, uz);
            static_for<0, kQ>([&](auto I) {
                constexpr int L = I;
                const E v = (E)collide_shift<L, R>(f[L], drho, jx, jy, jz, ux, uy, uz, a.fcx, a.fcy, a.fcz, a.A,
                                                   a.B, a.C);
                if constexpr (NT)
                    fout.stnt((uint32_t)L * Np + i, v);
                else
                    fout.st((uint32_t)L * Np + i, v);
            });
        } else {
            const double ux = jx / rho, uy = jy / rho, uz = jz / rho;
            static_for<0, kQ>([&](auto I) {
                constexpr int L = I;
                const double v = collide_pop<L>(f[L], rho, jx, jy, jz, ux, uy, uz, a.fcx, a.fcy, a.fcz, a.A, a.B, a.C);
                if constexpr (NT)
                    fout.stnt((uint32_t)L * Np + i, v);
                else
                    fout.st((uint32_t)L * Np + i, v);
            });
        }
    }
    if (a.write_l2) {
        __shared__ double lds[BLOCK / 64];
        l2 = wave_max(l2);
        const bool wneg = __ballot(neg) != 0ull;
        if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = wneg ? -1.0 : l2;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = lds[0];
            bool bneg = m < 0.0;
            for (int w = 1; w < BLOCK / 64; ++w) {
                bneg = bneg || lds[w] < 0.0;
                m = fmax(m, lds[w]);
            }
            if (bneg || m > 0.0)
                atomicMax(&a.slots[(a.t % 3) * kSlots + (stripe & (kSlots - 1))],
                          bneg ? kNegMark : (unsigned long long)__double_as_longlong(m));
        }
        if (a.jpart) {
            __shared__ double jl[BLOCK / 64 * 3];
            block_partial<3, BLOCK>(jsum, a.jpart + 3 * (uint64_t)tile, jl);
        }
    }
}

// The FP32-store step (FP32 storage, FP64 arithmetic) fits five waves per
// SIMD without spilling; the compiler's own choice leaves it at four.
template <typename E, typename R>
constexpr int kLbMinWaves = (std::is_same<E, float>::value && std::is_same<R, double>::value) ? 5 : LBX_LB_WAVES;

// NT: the outgoing populations are written with the streaming hint, for a
// lattice whose two population sets outgrow the last-level cache, where the
// next step's reads find none of them there anyway.
#ifndef LBX_NT_MIN_BYTES  // population footprint (both sets) above which LB stores stream
#define LBX_NT_MIN_BYTES (256ull << 20)
#endif
template <bool Wide, typename E, typename R = E, bool NT = false>
__global__ __launch_bounds__(kBlock, (kLbMinWaves<E, R>)) void k_lb_step(StepArgs<E, R> a) {
    lb_step_tile<kBlock, Wide, E, R, NT>(a, xcd_tile(), blockIdx.x, blockIdx.x == 0);
}

// l2err of the last step of a launch run, which no later step kernel reads
// (called by one wave).
__device__ __forceinline__ void lb_finalize(const unsigned long long* slots, double* l2ring, Ctrl* ctrl, int64_t t,
                                            int64_t check_base, int32_t check_every) {
    if (vload(&ctrl->stop) || vload(&ctrl->negpop) >= 0) return;
    const unsigned long long v = slots_max(slots + (t % 3) * kSlots);
    if (threadIdx.x == 0) {
        if (v == kNegMark)
            ctrl->negpop = t;
        else
            l2ring[t % kRing] =
                is_check(t, check_base, check_every) ? __longlong_as_double((long long)v) : __builtin_nan("");
    }
}

__global__ __launch_bounds__(64) void k_lb_finalize(const unsigned long long* slots, double* l2ring, Ctrl* ctrl,
                                                    int64_t t, int64_t check_base, int32_t check_every) {
    lb_finalize(slots, l2ring, ctrl, t, check_base, check_every);
}

// Streaming only: the post-streaming state f_s(t) of the step whose input is
// fin, and its density SUM(n,4) (equilibration.f90:254).
__global__ __launch_bounds__(kBlock) void k_extract(const double* __restrict__ fin, double* __restrict__ nout,
                                                    double* __restrict__ rho_out, const uint32_t* __restrict__ nodes,
                                                    const uint32_t* __restrict__ map, FastDiv dlx, FastDiv dly,
                                                    uint32_t lz, uint32_t i0, uint32_t Nf, uint64_t Np) {
    const uint32_t i = i0 + blockIdx.x * kBlock + threadIdx.x;
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

// The same for the FP32-shift form: the post-streaming deviations g into
// gout, and the FP64 density rho0 + sum g and flux F/2 + sum g c, summed in
// FP64 from the FP32 deviations (the flow that moment propagation reads).
__global__ __launch_bounds__(kBlock) void k_extract_shift(const float* __restrict__ gin, float* __restrict__ gout,
                                                          double* __restrict__ rho_out, double* __restrict__ j_out,
                                                          const uint32_t* __restrict__ nodes,
                                                          const uint32_t* __restrict__ map, FastDiv dlx, FastDiv dly,
                                                          uint32_t lz, uint32_t i0, uint32_t Nf, uint64_t Np,
                                                          double rho0, double fmx, double fmy, double fmz) {
    const uint32_t i = i0 + blockIdx.x * kBlock + threadIdx.x;
    if (i >= Nf) return;
    Neighbours nb;
    nb.load(map, nodes[i] & kIndexMask, dlx, dly, lz);
    const View<true, float> gv(gin, kQ * Np);
    double drho = 0.0, jx = fmx, jy = fmy, jz = fmz;
    static_for<0, kQ>([&](auto I) {
        constexpr int L = I;
        const float v = pull<L, true>(gv, (uint32_t)Np, i, nb);
        gout[(uint64_t)L * Np + i] = v;
        const double d = v;
        drho = (L == 0) ? d : drho + d;
        if constexpr (kCX[L] != 0) jx = jx + sgn<kCX[L]>(d);
        if constexpr (kCY[L] != 0) jy = jy + sgn<kCY[L]>(d);
        if constexpr (kCZ[L] != 0) jz = jz + sgn<kCZ[L]>(d);
    });
    rho_out[i] = rho0 + drho;
    j_out[i] = jx;
    j_out[Np + i] = jy;
    j_out[2 * Np + i] = jz;
}

// The density SUM(n,4) (equilibration.f90:254) and flux F(t)/2 + sum n c
// (:293-300) of the post-streaming state of the step whose input is fin,
// without writing the populations (the run goes on from fin): out[0][Np] the
// density, out[1..3][Np] the flux.  The same sums in the same order as
// k_extract and the step kernel.
__global__ __launch_bounds__(kBlock) void k_flow(const double* __restrict__ fin, double* __restrict__ out,
                                                 const uint32_t* __restrict__ nodes, const uint32_t* __restrict__ map,
                                                 FastDiv dlx, FastDiv dly, uint32_t lz, uint32_t Nf, uint64_t Np,
                                                 double fmx, double fmy, double fmz) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= Nf) return;
    Neighbours nb;
    nb.load(map, nodes[i] & kIndexMask, dlx, dly, lz);
    const View<true> fv(fin, kQ * Np);
    double rho = 0.0, jx = fmx, jy = fmy, jz = fmz;
    static_for<0, kQ>([&](auto I) {
        constexpr int L = I;
        const double v = pull<L, true>(fv, (uint32_t)Np, i, nb);
        rho = (L == 0) ? v : rho + v;
        if constexpr (kCX[L] != 0) jx = jx + sgn<kCX[L]>(v);
        if constexpr (kCY[L] != 0) jy = jy + sgn<kCY[L]>(v);
        if constexpr (kCZ[L] != 0) jz = jz + sgn<kCZ[L]>(v);
    });
    out[i] = rho;
    out[Np + i] = jx;
    out[2 * Np + i] = jy;
    out[3 * Np + i] = jz;
}

// The same for the FP32 storage forms, as k_extract_shift sums them: FP64
// density rho0 + sum g and flux F(t)/2 + sum g c from the FP32 deviations.
__global__ __launch_bounds__(kBlock) void k_flow_shift(const float* __restrict__ gin, double* __restrict__ out,
                                                       const uint32_t* __restrict__ nodes,
                                                       const uint32_t* __restrict__ map, FastDiv dlx, FastDiv dly,
                                                       uint32_t lz, uint32_t Nf, uint64_t Np, double rho0,
                                                       double fmx, double fmy, double fmz) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= Nf) return;
    Neighbours nb;
    nb.load(map, nodes[i] & kIndexMask, dlx, dly, lz);
    const View<true, float> gv(gin, kQ * Np);
    double drho = 0.0, jx = fmx, jy = fmy, jz = fmz;
    static_for<0, kQ>([&](auto I) {
        constexpr int L = I;
        const double d = pull<L, true>(gv, (uint32_t)Np, i, nb);
        drho = (L == 0) ? d : drho + d;
        if constexpr (kCX[L] != 0) jx = jx + sgn<kCX[L]>(d);
        if constexpr (kCY[L] != 0) jy = jy + sgn<kCY[L]>(d);
        if constexpr (kCZ[L] != 0) jz = jz + sgn<kCZ[L]>(d);
    });
    out[i] = rho0 + drho;
    out[Np + i] = jx;
    out[2 * Np + i] = jy;
    out[3 * Np + i] = jz;
}

// The collision on packed fields with a uniform force (the opening collision
// of a run), in place: module_collision.f90:87-108.
__global__ __launch_bounds__(kBlock) void k_collide_packed(double* f, const double* __restrict__ rho,
                                                           const double* __restrict__ j, double fx, double fy,
                                                           double fz, uint32_t i0, uint32_t Nf, uint64_t Np,
                                                           double A, double B, double C) {
    const uint32_t i = i0 + blockIdx.x * kBlock + threadIdx.x;
    if (i >= Nf) return;
    const double d = rho[i], jx = j[i], jy = j[Np + i], jz = j[2 * Np + i];
    const double ux = jx / d, uy = jy / d, uz = jz / d;
    static_for<0, kQ>([&](auto I) {
        constexpr int L = I;
        const double v = f[(uint64_t)L * Np + i];
        f[(uint64_t)L * Np + i] = collide_pop<L>(v, d, jx, jy, jz, ux, uy, uz, fx, fy, fz, A, B, C);
    });
}

// The same for the FP32 storage forms: the FP64 density and flux of the state
// become the step's drho = rho - rho0 and j (type R: float for FP32-shift,
// double for FP32-store) of the collision and j_old of the first step
// (jold may be j itself in the FP32-store form).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_collide_packed_shift(float* g, const double* __restrict__ rho,
                                                                 const double* j, R* jold, R fx, R fy, R fz,
                                                                 uint32_t i0, uint32_t Nf, uint64_t Np, R A, R B,
                                                                 R C, double rho0) {
    const uint32_t i = i0 + blockIdx.x * kBlock + threadIdx.x;
    if (i >= Nf) return;
    const R drho = (R)(rho[i] - rho0), full = (R)rho0 + drho;
    const R jx = (R)j[i], jy = (R)j[Np + i], jz = (R)j[2 * Np + i];
    jold[i] = jx;  // j_old of the first step
    jold[Np + i] = jy;
    jold[2 * Np + i] = jz;
    R ux, uy, uz;
    shift_velocity<R>(jx, jy, jz, full, ux, uy, uz);
    static_for<0, kQ>([&](auto I) {
        constexpr int L = I;
        const R v = g[(uint64_t)L * Np + i];
        g[(uint64_t)L * Np + i] = (float)collide_shift<L, R>(v, drho, jx, jy, jz, ux, uy, uz, fx, fy, fz, A, B, C);
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

// Rest state, init_simu.f90:31-39: n = rho0*a0 at fluid nodes, j = 0; in the
// FP32-shift form (g != NULL) the deviations g = 0 and the FP32 flux j32 = 0.
__global__ __launch_bounds__(kBlock) void k_init_rest(double* __restrict__ f, float* __restrict__ g,
                                                      float* __restrict__ j32, double* __restrict__ rho,
                                                      double* __restrict__ j, uint32_t i0, uint32_t Nf, uint64_t Np,
                                                      double rho0) {
    const uint32_t i = i0 + blockIdx.x * kBlock + threadIdx.x;
    if (i >= Nf) return;
    static_for<0, kQ>([&](auto I) {
        constexpr int L = I;
        if (f) f[(uint64_t)L * Np + i] = rho0 * a0(L);
        if (g) g[(uint64_t)L * Np + i] = 0.0f;
    });
    rho[i] = rho0;
    for (int c = 0; c < 3; ++c) {
        j[c * Np + i] = 0.0;
        if (j32) j32[c * Np + i] = 0.0f;
    }
}

// ---------------------------------------------------------------- MP kernels

// packed nodes on either side of an MP block whose P the block stages in LDS
#ifndef LBX_MP_HALO
#define LBX_MP_HALO 128
#endif
constexpr int kMpHalo = LBX_MP_HALO;

// a block's LDS array with View's element / pair access (load_comps, store_comps)
struct LdsView {
    double* b;
    __device__ __forceinline__ double ld(uint32_t idx) const { return b[idx]; }
    __device__ __forceinline__ void st(uint32_t idx, double v) const { b[idx] = v; }
    __device__ __forceinline__ Vec2<double> ld2(uint32_t pidx) const { return ((const Vec2<double>*)b)[pidx]; }
    __device__ __forceinline__ void st2(uint32_t pidx, Vec2<double> v) const { ((Vec2<double>*)b)[pidx] = v; }
};
//
// Several tracers that share the flow and Db (so the scattering
// probabilities win, rem and u*) but differ in ka, kd can be propagated in one
// pass: the shared coefficients are read once per node for all of them.  The
// arithmetic of each tracer is exactly that of a single-tracer run.

constexpr int kMaxTracers = 4;

// VACF of one moment-propagation step (NV = 3 per tracer) from the per-block
// partial sums, in a fixed order that does not depend on which kernel runs it
// (so every run gives the same bits), then the running D_eff sums and,
// optionally, the convergence test of module_moment_propagation.f90:284-288
// (all NV components below both limits).  Called by all BLOCK threads of one
// block.  mode 0: VACF(0) of mp init; mode 1: VACF(it) of a step.
template <int NV, int BLOCK>
__device__ __forceinline__ void mp_reduce_block(const double* part, uint32_t nparts, double* vtrace, Ctrl* ctrl,
                                                int32_t mode, int64_t it, int32_t check_conv, double lim1,
                                                double lim2, double* lds /* [4*NV] */) {
    double s[NV];
#pragma unroll
    for (int a = 0; a < NV; ++a) s[a] = 0.0;
    for (uint32_t k = threadIdx.x; k < nparts; k += BLOCK)
#pragma unroll
        for (int a = 0; a < NV; ++a) {
            const double x = vload(&part[(uint64_t)NV * k + a]);
            s[a] = (k == threadIdx.x) ? x : s[a] + x;
        }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < NV; ++a) {
        const double t = wave_sum(s[a]);
        if (lane == 0) lds[w * NV + a] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int a = 0; a < NV; ++a) {
            double t = lds[a];
            for (int k = 1; k < BLOCK / 64; ++k) t = t + lds[k * NV + a];
            s[a] = t;
        }
        if (mode == 0) {
#pragma unroll
            for (int a = 0; a < NV; ++a) {
                ctrl->vacf0[a] = s[a];
                ctrl->deff_s[a] = 0.5 * s[a];
                ctrl->deff_c[a] = 0.0;
            }
        } else {
#pragma unroll
            for (int a = 0; a < NV; ++a) {
                vtrace[a] = s[a];
                // Neumaier update of D = VACF(0)/2 + sum VACF(t)
                const double d = ctrl->deff_s[a], x = s[a], u = d + x;
                ctrl->deff_c[a] += (fabs(d) >= fabs(x)) ? (d - u) + x : (x - u) + d;
                ctrl->deff_s[a] = u;
            }
            if (check_conv && it > 2) {
                bool conv = true;
#pragma unroll
                for (int a = 0; a < NV; ++a) conv = conv && fabs(s[a]) < lim1 && fabs(s[a]) < lim2;
                if (conv) {
                    ctrl->mp_conv_step = it;
                    ctrl->mp_stop = 1;
                }
            }
        }
    }
    __syncthreads();
}

template <int NV>
__global__ __launch_bounds__(kBlock) void k_mp_reduce(const double* __restrict__ part, uint32_t nparts,
                                                      double* vtrace, Ctrl* ctrl, int32_t mode, int64_t it,
                                                      int32_t check_conv, double lim1, double lim2) {
    __shared__ double lds[kBlock / 64 * NV];
    if (mode == 1 && vload(&ctrl->mp_stop)) return;
    mp_reduce_block<NV, kBlock>(part, nparts, vtrace, ctrl, mode, it, check_conv, lim1, lim2, lds);
}

struct TracerParams {
    double bw[kMaxTracers];    // 1/Pstat (module_moment_propagation.f90:97,108)
    double ka[kMaxTracers], kd[kMaxTracers], one_minus_kd[kMaxTracers];
    int32_t ads[kMaxTracers];  // adsorption on (K = ka/kd > eps, :46-56)
};

// W is the type the scattering probabilities win and u* are stored in:
// double, or float in the FP32-shift form.  There each probability is rounded
// to FP32 once, and every quantity built from it (the rest probability
// rem = 1 - sum s, which stays FP64, u*, P(r,1), VACF(0)) uses the rounded
// value, so the propagator still conserves sum_r P exactly up to FP64 rounding.
template <typename W>
__device__ __forceinline__ double as_coef(double v) {
    return (double)(W)v;
}

// The scattering probability s(r', l') = n_l'/rho - w_l' + lambda w_l' / 2
// (module_moment_propagation.f90:218-228,331-346) of a node r' with density
// rho and flow v = j + f, from its tracer population n_l' = a0 rho + a1 c.v
// (drop_tracers.f90:97-100), rounded as the coefficients are kept (W).  The
// set-up and the recomputing step kernel call this one function, so they get
// the same bits.
template <int L, typename W>
__device__ __forceinline__ double scatter_prob(double rho, double vx, double vy, double vz, double lambda) {
    constexpr int cx = kCX[L], cy = kCY[L], cz = kCZ[L];
    const double nt = a0(L) * rho + a1(L) * dot3<cx, cy, cz>(vx, vy, vz);
    return as_coef<W>(nt / rho - a0(L) + lambda * a0(L) * 0.5);
}

// Where the MP step takes the incoming probabilities s(r+c_l, inv l) from
// (the C-ABI's LBX_MP_STORED / _RECOMPUTED / _MIXED): the 18 stored per node,
// all 18 recomputed from the neighbours' (rho, v), or the 10 with c_z != 0
// stored (those neighbours are a plane away, outside the LDS window, where
// reading (rho, v) costs more than a stored value) and the 8 others
// recomputed from the window.
constexpr int kCoefStored = 0, kCoefRecomputed = 1, kCoefMixed = 2;
constexpr int kZDirs = 10;  // directions with c_z != 0: 5, 6, 11..18
// index of direction L (c_z != 0) among them, and of its pair (L, L+1 for odd L)
constexpr int zdir(int L) { return L <= 6 ? L - 5 : L - 9; }

template <typename W>
struct MpPrepArgs {
    const double* __restrict__ rho;
    const double* __restrict__ j;
    const uint32_t* __restrict__ nodes;
    const uint32_t* __restrict__ map;
    W* __restrict__ win = nullptr;       // incoming probabilities (NULL: not kept, the step recomputes them)
    int32_t win_z = 0;                   // win holds only the kZDirs directions with c_z != 0
    double* __restrict__ rv = nullptr;   // (rho, v) of each node, 4 doubles (NULL: not kept)
    double* __restrict__ rem;
    W* __restrict__ ustar;
    double* __restrict__ P0;   // [T][3][Np]
    double* partials;
    Ctrl* ctrl;
    FastDiv dlx, dly;
    uint32_t lz, Nf;
    uint64_t Np;
    double fx, fy, fz;   // lb.in f_ext (drop_tracers.f90:92)
    double lambda, eps;
    uint32_t tile0;        // first tile of the lattice's node range
    TracerParams tp;
};

// update_tracer_population (drop_tracers.f90:97-100) fused with
// moment_propagation::init (module_moment_propagation.f90:105-137) and the
// static part of propagate (:218-249): per fluid node, the scattering
// probabilities s(r,l) = n_l/rho - w_l + lambda w_l / 2 towards fluid
// neighbours, the rest probability rem, u* = sum s c, the incoming
// probabilities win(r,l) = s(r+c_l, inv l), and per tracer P(r,1) and the
// VACF(0) contribution.
//
// mp_prep_tile: one block's share; the block's nodes are tile*BLOCK + lane
// (valid below a.Nf) and its VACF(0) partial goes to a.partials[3T*tile ...].
template <int BLOCK, int T, typename W>
__device__ __forceinline__ void mp_prep_tile(const MpPrepArgs<W>& a, uint32_t tile) {
    __shared__ double lds[BLOCK / 64 * 3 * T];
    const uint32_t i = tile * BLOCK + threadIdx.x;
    double vp[3 * T];
#pragma unroll
    for (int k = 0; k < 3 * T; ++k) vp[k] = 0.0;
    if (i < a.Nf) {
        const uint32_t e = a.nodes[i];
        Neighbours nb;
        nb.load(a.map, e & kIndexMask, a.dlx, a.dly, a.lz);
        const double rr = a.rho[i];
        const double vx = a.j[i] + a.fx, vy = a.j[a.Np + i] + a.fy, vz = a.j[2 * a.Np + i] + a.fz;
        if (a.rv) {
            a.rv[4 * (uint64_t)i] = rr;
            a.rv[4 * (uint64_t)i + 1] = vx;
            a.rv[4 * (uint64_t)i + 2] = vy;
            a.rv[4 * (uint64_t)i + 3] = vz;
        }
        double us[3] = {0.0, 0.0, 0.0};
        double p1[3 * T];
#pragma unroll
        for (int k = 0; k < 3 * T; ++k) p1[k] = 0.0;
        double remv = 1.0;
        static_for<1, kQ>([&](auto I) {
            constexpr int L = I;
            constexpr int Li = kInv[L];
            constexpr int cx = kCX[L], cy = kCY[L], cz = kCZ[L];
            const uint32_t q = nb.q[L];
            if (q != kSolidSlot) {
                const double s = scatter_prob<L, W>(rr, vx, vy, vz, a.lambda);
                remv = remv - s;
                if constexpr (cx != 0) us[0] = us[0] + sgn<cx>(s);
                if constexpr (cy != 0) us[1] = us[1] + sgn<cy>(s);
                if constexpr (cz != 0) us[2] = us[2] + sgn<cz>(s);
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    if constexpr (cx != 0) vp[3 * t + 0] = vp[3 * t + 0] + a.tp.bw[t] * s;
                    if constexpr (cy != 0) vp[3 * t + 1] = vp[3 * t + 1] + a.tp.bw[t] * s;
                    if constexpr (cz != 0) vp[3 * t + 2] = vp[3 * t + 2] + a.tp.bw[t] * s;
                }
                const double rn = a.rho[q];
                const double wx = a.j[q] + a.fx, wy = a.j[a.Np + q] + a.fy, wz = a.j[2 * a.Np + q] + a.fz;
                const double sp = scatter_prob<Li, W>(rn, wx, wy, wz, a.lambda);
                if (a.win) {
                    if (!a.win_z)
                        a.win[comp_index(18, L - 1, i, a.Np)] = (W)sp;
                    else if constexpr (cz != 0)
                        a.win[comp_index(kZDirs, zdir(L), i, a.Np)] = (W)sp;
                }
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    if constexpr (cx != 0) p1[3 * t + 0] = p1[3 * t + 0] + sgn<-cx>(sp) * a.tp.bw[t];
                    if constexpr (cy != 0) p1[3 * t + 1] = p1[3 * t + 1] + sgn<-cy>(sp) * a.tp.bw[t];
                    if constexpr (cz != 0) p1[3 * t + 2] = p1[3 * t + 2] + sgn<-cz>(sp) * a.tp.bw[t];
                }
            } else if (a.win) {
                if (!a.win_z)
                    a.win[comp_index(18, L - 1, i, a.Np)] = W(0);
                else if constexpr (cz != 0)
                    a.win[comp_index(kZDirs, zdir(L), i, a.Np)] = W(0);
            }
        });
        a.rem[i] = remv;
        bool low = false;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const double rt = (a.tp.ads[t] && (e & kNodeInterfacial)) ? remv - a.tp.ka[t] : remv;
            low = low || rt < a.eps;
        }
        if (low) a.ctrl->restpart = 1;
        for (int c = 0; c < 3; ++c) a.ustar[comp_index(3, c, i, a.Np)] = (W)us[c];
#pragma unroll
        for (int k = 0; k < 3 * T; ++k) a.P0[comp_index(3 * T, k, i, a.Np)] = p1[k];
    }
    block_partial<3 * T, BLOCK>(vp, a.partials + 3 * T * (uint64_t)tile, lds);
}

template <int T, typename W>
__global__ __launch_bounds__(kBlock) void k_mp_prep(MpPrepArgs<W> a) {
    mp_prep_tile<kBlock, T, W>(a, a.tile0 + xcd_tile());
}

template <typename W>
struct MpStepArgs {
    const W* __restrict__ win;
    const double* __restrict__ rem;
    const W* __restrict__ ustar;
    const uint32_t* __restrict__ nodes;
    const uint32_t* __restrict__ map;
    const double* __restrict__ Pn;   // [T][3][Np]
    double* __restrict__ Px;
    const double* __restrict__ Pan;
    double* __restrict__ Pax;
    double* partials;                // this step's per-block VACF partials
    const double* prev_partials;     // the previous step's, reduced by block 0 (NULL: none)
    int32_t check_conv = 0;          // block 0 also applies the convergence test to VACF(it-1)
    double lim1 = 0.0, lim2 = 0.0;   // its limits (module_moment_propagation.f90:284)
    // RV steps: (rho, v) of each node, 4 doubles, and lambda, from which the
    // incoming probabilities are recomputed instead of read from win
    const double* rv = nullptr;
    double lambda = 0.0;
    double* prev_vtrace;             // where VACF(it-1) goes
    int64_t it;
    Ctrl* ctrl;
    FastDiv dlx, dly;
    uint32_t lz, Nf;
    uint64_t Np;
    int32_t any_ads;                 // some tracer of the batch adsorbs (Pan, Pax allocated)
    TracerParams tp;
    const uint32_t* tile_order = nullptr;  // tile of each logical tile (NULL: the same)
};

// One moment-propagation step of T tracers, module_moment_propagation.f90:207-253:
// P(r,t+1) = sum_l P(r+c_l,t) s(r+c_l, inv l) + rem(r) P(r,t) [+ kd Pa],
// Pa(r,t+1) = (1-kd) Pa(r,t) + ka P(r,t) at adsorbing interfacial nodes,
// VACF(t) = sum_r P(r,t) . u*(r).
//
// mp_step_tile: one block's share; nodes tile*BLOCK + lane (valid below a.Nf),
// VACF partial to a.partials[K*tile ...]; first_block also reduces the
// previous step's nparts partials (a.prev_partials) into VACF(it-1).
//
// XR: directions 11..14 may reuse the z-neighbours of the lanes beside (below).
// CF (kCoef*): recomputed probabilities come from the neighbours' (rho, v)
// (scatter_prob, the set-up's own arithmetic: the same bits): 32 bytes per
// node kept and staged in the LDS window beside P; stored ones from win (18 or,
// mixed, kZDirs per node, read in pairs).
template <int BLOCK, bool Wide, int T, typename W, int CF = kCoefStored, bool XR = true>
__device__ __forceinline__ void mp_step_tile(const MpStepArgs<W>& a, uint32_t tile, bool first_block,
                                             uint32_t nparts) {
    constexpr bool RV = CF != kCoefStored;  // (rho, v) staged in the window
    constexpr bool RVZ = CF == kCoefRecomputed;  // the directions with c_z != 0 recomputed too
    constexpr int K = 3 * T;  // moment components of the batch
    // The edge-lane records and the early gathers of the neighbours along 5
    // and 6 (below) serve the forms that store the z-directions'
    // probabilities.  With those recomputed (RVZ) the added loads and (rho, v)
    // work before the window barrier cost more than the records save
    // (measured), so RVZ leaves both out.
    constexpr bool XCH = !RVZ && XR;
    // the mixed form with three tracers: 4 fewer halo nodes a side keep the
    // records within a third of the CU's LDS (three blocks per CU)
    constexpr int HALO = (XCH && RV && T == 3) ? kMpHalo - 4 : kMpHalo;
    constexpr int WIN = BLOCK + 2 * HALO;
    constexpr int NW = BLOCK / 64;  // waves per block
    // Edge-lane records, [wave boundary][side][plane][RW]: P of a wave's first
    // (side 0) or last (side 1) lane's neighbour along 5 (plane 0) or 6
    // (plane 1), for the lane beside it in the adjacent wave.  The array
    // holds the VACF partial's wave sums after the step's last record read.
    constexpr int RW = K;  // doubles per record
    constexpr int XN = (XCH && NW > 1) ? (NW - 1) * 4 * RW : 0;
    constexpr int LN = NW * K > XN ? NW * K : XN;
    __shared__ double lds[LN];
    __shared__ double pwin[K * WIN];  // P(now) of packed nodes [w0, w0 + WIN), comp_index layout
    __shared__ double rvwin[RV ? 4 * WIN : 2];  // RV: (rho, v) of the same nodes
    const uint32_t i = tile * BLOCK + threadIdx.x;
    const bool valid = i < a.Nf;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    // Each wave's work is a chain of dependent memory round trips (stop flag,
    // node word, neighbour indices, neighbour values), so independent
    // requests are issued together: the node word with the stop flag, the
    // neighbour indices and the node's coefficients with the LDS window.
    const uint32_t e = valid ? ld(a.nodes, i) : 0u;
    if (block_flag(vload(&a.ctrl->mp_stop) != 0)) return;
    // VACF(it-1) from the previous step's partials and, with check_conv, the
    // convergence test: a stop set here ends the run at step it-1 (this
    // launch's other blocks may have started step it already; they write only
    // the other buffer of P, and every later launch returns at once)
    if (a.prev_partials && first_block) {
        mp_reduce_block<K, BLOCK>(a.prev_partials, nparts, a.prev_vtrace, a.ctrl, 1, a.it - 1, a.check_conv, a.lim1,
                                  a.lim2, lds);
        if constexpr (XCH && NW > 1) __syncthreads();  // its last read of lds before the records are written there
    }
    const uint32_t Np = (uint32_t)a.Np;
    const View<Wide> Pn(a.Pn, K * a.Np);
    const View<Wide, W> win(a.win, (CF == kCoefMixed ? kZDirs : 18) * a.Np), us(a.ustar, 3 * a.Np);
    const View<Wide> rem(a.rem, a.Np);
    Neighbours nb;
    W uw[3] = {W(0), W(0), W(0)};
    double rm = 0.0;
    if (valid) {
        nb.load(a.map, e & kIndexMask, a.dlx, a.dly, a.lz);
        load_comps<3>(us, Np, i, uw);
        rm = rem.ld(i);
    }
    // The block's nodes and HALO on either side: neighbours along x, and
    // along y where a row holds fewer fluid nodes than that, are read from LDS
    const int64_t w0 = (int64_t)tile * BLOCK - HALO;
    const LdsView pw{pwin};
    const View<Wide> rvv(a.rv, 4 * a.Np);
    // The node's neighbours along 5 (0,0,1) and 6 (0,0,-1), whose values the
    // directions 11..14 reuse (below): gathered here, with the window, where
    // they lie outside it or the lane is a wave's first or last.
    // A lattice one node wide along x (the 2-D instances) has no x-neighbour
    // in another lane: it skips all of this (the same for every lane of a
    // block, whose nodes belong to one lattice), and XR = false compiles it
    // out (a batch of such lattices only: the registers it holds cost them
    // occupancy).
    const bool xreuse = XR && a.dlx.d > 1u;
    double zc[2][K];         // their P (0 if solid)
    Vec2<double> zr[2][2];   // RVZ: their (rho, v), kept in the direction loop
    bool zearly[2] = {false, false};
    const bool edge_lane = XCH && NW > 1 && xreuse && (lane == 0u || lane == 63u);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
#pragma unroll
        for (int k = 0; k < K; ++k) zc[s][k] = 0.0;
        zr[s][0] = Vec2<double>{0.0, 0.0};
        zr[s][1] = Vec2<double>{0.0, 0.0};
        if (valid) {
            const uint32_t q = nb.q[5 + s];
            zearly[s] = XCH && xreuse && q != kSolidSlot && ((uint32_t)((int64_t)q - w0) >= (uint32_t)WIN || edge_lane);
            if (zearly[s]) load_comps<K>(Pn, Np, q, zc[s]);
        }
    }
    {
        // every slot's loads go out before the first LDS write waits on them
        constexpr int NIT = (WIN + BLOCK - 1) / BLOCK;
        double fv[NIT][K];
        Vec2<double> fr[NIT][2];
        bool in[NIT];
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int n = (int)threadIdx.x + it * BLOCK;
            const int64_t g = w0 + n;
            in[it] = n < WIN && g >= 0 && g < (int64_t)Np;
            if (in[it]) {
                load_comps<K>(Pn, Np, (uint32_t)g, fv[it]);
                if constexpr (RV) {
                    fr[it][0] = rvv.ld2(2 * (uint32_t)g);
                    fr[it][1] = rvv.ld2(2 * (uint32_t)g + 1);
                }
            }
        }
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int n = (int)threadIdx.x + it * BLOCK;
            if (in[it]) {
                store_comps<K>(pw, WIN, (uint32_t)n, fv[it]);
                if constexpr (RV) {
                    ((Vec2<double>*)rvwin)[2 * n] = fr[it][0];
                    ((Vec2<double>*)rvwin)[2 * n + 1] = fr[it][1];
                }
            }
        }
    }
    // records of the lanes at the block's inner wave boundaries: lane 0 of
    // wave w serves lane 63 of wave w-1 (directions 11, 13), lane 63 of wave w
    // serves lane 0 of wave w+1 (12, 14)
    if constexpr (XCH && NW > 1) {
        if (valid && xreuse && ((lane == 0u && wave > 0u) || (lane == 63u && wave + 1u < (uint32_t)NW))) {
            const uint32_t side = lane == 0u ? 0u : 1u, bnd = lane == 0u ? wave - 1u : wave;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                double* r = lds + ((bnd * 2u + side) * 2u + (uint32_t)s) * RW;
#pragma unroll
                for (int k = 0; k < K; ++k) r[k] = zc[s][k];
            }
        }
    }
    __syncthreads();
    double vp[K];
#pragma unroll
    for (int k = 0; k < K; ++k) vp[k] = 0.0;
    if (valid) {
        const View<Wide> Px(a.Px, K * a.Np);
        double p[K], acc[K];
        load_comps<K>(pw, WIN, (uint32_t)(HALO + threadIdx.x), p);
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = 0.0;
        // The node's x-neighbour in its row is packed node i+1 (i-1) whenever
        // that neighbour is fluid and not across the periodic wrap: held by the
        // next (previous) lane of the wave, or by the first (last) lane of the
        // next (previous) wave of the block.  Its neighbour one plane up (down)
        // is then the node's neighbour along (+-1, 0, +-1): directions 11..14
        // take P (and the probability from (rho, v)) of that lane's neighbour
        // along 5 or 6, moved across the wave or from its edge-lane record,
        // instead of gathering them again; the sum keeps its order and bits.
        const bool adj_r = nb.q[1] == i + 1u;
        const bool adj_l = i > 0u && nb.q[2] == i - 1u;
        Vec2<W> wp;  // the probabilities of directions L, L+1 (L odd) from one paired load
        static_for<1, kQ>([&](auto I) {
            constexpr int L = I;
            constexpr bool stored = CF == kCoefStored || (CF == kCoefMixed && kCZ[L] != 0);
            // 11, 12 from 5 of lane +-1; 13, 14 from 6 of lane +-1; 5, 6 kept
            constexpr int zsrc = (L >= 11 && L <= 14) ? (kCZ[L] > 0 ? 0 : 1) : -1;
            constexpr int zkeep = L == 5 ? 0 : (L == 6 ? 1 : -1);
            if constexpr (CF == kCoefStored && (L & 1)) wp = win.ld2((uint32_t)((L - 1) / 2) * Np + i);
            if constexpr (CF == kCoefMixed && stored && (L & 1)) wp = win.ld2((uint32_t)(zdir(L) / 2) * Np + i);
            double sh[K];          // moved across the wave while every valid lane is active
            Vec2<double> sr[2];
            if constexpr (zsrc >= 0) {
                if (xreuse) {
#pragma unroll
                    for (int k = 0; k < K; ++k) sh[k] = lane_shift<kCX[L]>(zc[zsrc][k]);
                    if constexpr (!stored)
                        for (int h = 0; h < 2; ++h)
                            sr[h] = Vec2<double>{lane_shift<kCX[L]>(zr[zsrc][h][0]),
                                                 lane_shift<kCX[L]>(zr[zsrc][h][1])};
                }
            }
            const uint32_t q = nb.q[L];
            if (q != kSolidSlot) {
                double pq[K];
                double w = 0.0;
                Vec2<double> r0, r1;  // (rho, vx), (vy, vz) of the neighbour (RV directions)
                int how = 0;          // 0: gather, 1: from the lane beside, 2: from an edge-lane record
                if constexpr (zsrc >= 0) {
                    constexpr bool right = kCX[L] > 0;
                    if (xreuse && (right ? adj_r : adj_l)) {  // else gathered (sh is not set)
                        if (!(right ? lane == 63u : lane == 0u))
                            how = 1;
                        else if (XCH && (right ? wave + 1u < (uint32_t)NW : wave > 0u))
                            how = 2;
                    }
                    if (how == 1) {
#pragma unroll
                        for (int k = 0; k < K; ++k) pq[k] = sh[k];
                        if constexpr (!stored) {
                            r0 = sr[0];
                            r1 = sr[1];
                        }
                    } else if (how == 2) {  // XCH: these directions' probabilities are stored
                        const double* r = lds + (((right ? wave : wave - 1u) * 2u + (right ? 0u : 1u)) * 2u +
                                                 (uint32_t)zsrc) * RW;
#pragma unroll
                        for (int k = 0; k < K; ++k) pq[k] = r[k];
                    }
                }
                if constexpr (zkeep >= 0) {
                    if (zearly[zkeep]) {
#pragma unroll
                        for (int k = 0; k < K; ++k) pq[k] = zc[zkeep][k];
                        if constexpr (!stored) {
                            r0 = zr[zkeep][0];
                            r1 = zr[zkeep][1];
                        }
                        how = 1;
                    }
                }
                if (how == 0) {
                    const uint32_t d = (uint32_t)((int64_t)q - w0);
                    if (d < (uint32_t)WIN) {
                        load_comps<K>(pw, WIN, d, pq);
                        if constexpr (!stored) {
                            r0 = ((const Vec2<double>*)rvwin)[2 * d];
                            r1 = ((const Vec2<double>*)rvwin)[2 * d + 1];
                        }
                    } else {
                        load_comps<K>(Pn, Np, q, pq);
                        if constexpr (!stored) {
                            r0 = rvv.ld2(2 * q);
                            r1 = rvv.ld2(2 * q + 1);
                        }
                    }
                }
                if constexpr (zkeep >= 0) {
#pragma unroll
                    for (int k = 0; k < K; ++k) zc[zkeep][k] = pq[k];
                    if constexpr (!stored) {
                        zr[zkeep][0] = r0;
                        zr[zkeep][1] = r1;
                    }
                }
                if constexpr (!stored)
                    w = scatter_prob<kInv[L], W>(r0[0], r0[1], r1[0], r1[1], a.lambda);
                else
                    w = (double)wp[(L - 1) & 1];
#pragma unroll
                for (int k = 0; k < K; ++k) acc[k] = acc[k] + pq[k] * w;
            }
        });
#pragma unroll
        for (int t = 0; t < T; ++t)
            for (int c = 0; c < 3; ++c) vp[3 * t + c] = p[3 * t + c] * (double)uw[c];
        double px[K];
        if (!(a.any_ads && (e & kNodeInterfacial))) {
#pragma unroll
            for (int k = 0; k < K; ++k) px[k] = acc[k] + rm * p[k];
            store_comps<K>(Px, Np, i, px);
        } else {
            const View<Wide> Pan(a.Pan, K * a.Np), Pax(a.Pax, K * a.Np);
            double pa[K], pax[K];
            load_comps<K>(Pan, Np, i, pa);
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const double rt = rm - a.tp.ka[t];
                for (int c = 0; c < 3; ++c) {
                    const int k = 3 * t + c;
                    if (a.tp.ads[t]) {
                        px[k] = (acc[k] + rt * p[k]) + pa[k] * a.tp.kd[t];
                        pax[k] = pa[k] * a.tp.one_minus_kd[t] + p[k] * a.tp.ka[t];
                    } else {  // a tracer without adsorption keeps Pa = 0
                        px[k] = acc[k] + rm * p[k];
                        pax[k] = pa[k];
                    }
                }
            }
            store_comps<K>(Px, Np, i, px);
            store_comps<K>(Pax, Np, i, pax);
        }
    }
    if constexpr (XCH && NW > 1)
        if (xreuse) __syncthreads();  // every wave's last edge-lane record read is over
    block_partial<K, BLOCK>(vp, a.partials + (uint64_t)K * tile, lds);
}

template <bool Wide, int T, typename W, int CF>
__global__ __launch_bounds__(kBlock, LBX_MP_WAVES) void k_mp_step(MpStepArgs<W> a) {
    mp_step_tile<kBlock, Wide, T, W, CF>(a, a.tile_order ? a.tile_order[xcd_tile()] : xcd_tile(), blockIdx.x == 0,
                                         gridDim.x);
}

// Plain streaming copy (16 B per lane per iteration, grid-stride): the
// measured ceiling that the roofline fractions are read against, and the
// known-size kernel that calibrates the HBM traffic counters.
__global__ __launch_bounds__(kBlock) void k_copy(const double2* __restrict__ a, double2* __restrict__ b, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
        b[i] = a[i];
}

// The LB step's streams without its gathers: S FP64 planes of n nodes read
// and S written, one node per lane, every load issued before the first store
// (the order of the step's pulls and writes).  Its rate is the ceiling of the
// structure-of-arrays layout itself, below the plain copy's.
template <int S>
__global__ __launch_bounds__(kBlock) void k_planes(const double* __restrict__ in, double* __restrict__ out, uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double v[S];
#pragma unroll
    for (int s = 0; s < S; ++s) v[s] = in[(uint64_t)s * n + i];
#pragma unroll
    for (int s = 0; s < S; ++s) out[(uint64_t)s * n + i] = v[s];
}

}  // namespace
