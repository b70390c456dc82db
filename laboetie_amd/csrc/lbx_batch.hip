// liblbx batch engine: many independent lattices ("instances") resident in
// HBM at once and advanced together, for batched and served solves.
//
// Layout.  All instances share one packed fluid-node list: instance b owns the
// node range [node_off, node_end) (node_off a multiple of 256, so every 128-
// or 256-node tile belongs to one instance) and a range of a shared full-box
// map array.  Populations, flux, density and the moment-propagation arrays
// are the single-lattice arrays over the whole list (plane stride = the node
// capacity).  tile_inst[] maps each 256-node tile to its instance, and a
// device table of BInst entries holds every instance's geometry, physics and
// the work of the current round.
//
// One LB launch advances every instance in the LB phase by one step, one MP
// launch every instance in the MP phase; the blocks of other instances leave
// at once.  The per-block work is lb_step_tile / mp_step_tile, the functions
// the single-lattice kernels run, on a per-instance view (its own step index,
// control block, l2err slots, VACF partials), so an instance's results are
// bit-identical to a single-lattice run of the same input (with 256-node
// blocks; 128-node blocks group the VACF sums differently).
//
// Phases (host-driven between rounds, as the single-lattice host loop does):
// LB stage 1 -> (redo of the switch step with the force on) -> LB stage 2 ->
// extract + MP prep -> MP steps -> done.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "../../include/lbx.h"
#include "lbx_device.h"
#include "lbx_kernels.h"
#include "lbx_tables.h"

#pragma clang fp contract(off)

using namespace lbx;

namespace {

constexpr int kTile = 256;                 // allocation unit (nodes) of an instance's range
constexpr uint32_t kNoInst = 0xFFFFFFFFu;
constexpr int kRoundMax = 1024;            // launches per phase per round (vtrace rows)
constexpr size_t kGraphCacheMax = 64;      // captured round graphs kept at once

struct BInst {
    // geometry (set when the instance is added)
    uint32_t node_off, node_end;  // packed range
    uint32_t tile_first, ntiles;  // 256-node tiles
    uint32_t map_off, lz;
    FastDiv dlx, dly;
    double A, B, C, rho0, target;
    // this round's LB work
    int32_t lb_on, lb_redo;       // takes part; the launch is the redo of step lb_t0 (force switch)
    int64_t lb_t0, lb_n;          // launch k runs step lb_t0 + k while k < lb_n
    int64_t check_after, floor_window;
    int64_t check_base;           // l2err measured at check_base + m check_every (StepArgs)
    int32_t check_every, pad2_;
    double check_inv;             // 1 / check_every
    double fm[3], fc[3];          // F(t)/2 and the collision force (the stage's)
    // this round's MP work
    int32_t mp_on, any_ads;
    int64_t mp_it0, mp_n;         // launch k runs MP step mp_it0 + k + 1 while k < mp_n
    int32_t mp_repeat, pad_;      // every launch recomputes step mp_it0 + 1 (timing)
    TracerParams tp;
};

template <typename E, typename R = E>
struct BatchView {
    const BInst* insts;
    const uint32_t* tile_inst;
    const uint32_t* act;          // the 256-node tiles of the instances in the phase, ascending
    E* f0;
    E* f1;
    R* j;                         // the step's flux (FP32-shift: j32)
    const uint32_t* nodes;
    const uint32_t* maps;
    unsigned long long* slots;    // [inst][3][kSlots]
    double* l2ring;               // [inst][kRing]
    Ctrl* ctrl;                   // [inst]
    uint64_t Np;
};

// A launch covers only the tiles of the instances in its phase: logical
// block l of the launch is BLOCK-node tile `active_tile(act, l)`, in the
// 256-node tile act[l / (256/BLOCK)] (instances in the other phase, or done,
// cost no blocks).
template <int BLOCK>
__device__ __forceinline__ uint32_t active_tile(const uint32_t* act, uint32_t l) {
    constexpr uint32_t per = kTile / BLOCK;
    return act[l / per] * per + l % per;
}

// LB launch k of a round on one BLOCK-node tile (a redo repeats step lb_t0);
// NT: population stores with the streaming hint
template <int BLOCK, bool Wide, typename E, typename R = E, bool NT = false>
__device__ __forceinline__ void kb_lb_tile(const BatchView<E, R>& v, uint32_t tile, int64_t k) {
    const uint32_t id = v.tile_inst[tile / (kTile / BLOCK)];
    if (id == kNoInst) return;
    const BInst& in = v.insts[id];
    if (!in.lb_on || k >= in.lb_n) return;
    const int64_t t = in.lb_t0 + (in.lb_redo ? 0 : k);
    StepArgs<E, R> a;
    a.fin = (t - 1) & 1 ? v.f1 : v.f0;
    a.fout = t & 1 ? v.f1 : v.f0;
    a.j = v.j;
    a.nodes = v.nodes;
    a.map = v.maps + in.map_off;
    a.slots = v.slots + (uint64_t)id * 3 * kSlots;
    a.l2ring = v.l2ring + (uint64_t)id * kRing;
    a.ctrl = v.ctrl + id;
    a.dlx = in.dlx;
    a.dly = in.dly;
    a.lz = in.lz;
    a.Nf = in.node_end;
    a.Np = v.Np;
    a.A = (R)in.A;
    a.B = (R)in.B;
    a.C = (R)in.C;
    a.fmx = (R)in.fm[0];
    a.fmy = (R)in.fm[1];
    a.fmz = (R)in.fm[2];
    a.fcx = (R)in.fc[0];
    a.fcy = (R)in.fc[1];
    a.fcz = (R)in.fc[2];
    a.rho0 = (R)in.rho0;
    a.t = t;
    a.check_after = in.check_after;
    a.floor_window = in.floor_window;
    a.target = in.target;
    a.write_l2 = in.lb_redo ? 0 : 1;
    a.have_prev = (!in.lb_redo && t > 1) ? 1 : 0;
    a.check_every = in.check_every;
    a.check_base = in.check_base;
    a.check_inv = (R)in.check_inv;
    const uint32_t first = in.tile_first * (kTile / BLOCK);
    lb_step_tile<BLOCK, Wide, E, R, NT>(a, tile, tile - first, tile == first);
}

template <int BLOCK, bool Wide, typename E, typename R = E, bool NT = false>
__global__ __launch_bounds__(BLOCK, (kLbMinWaves<E, R>)) void kb_lb_step(BatchView<E, R> v, int64_t k) {
    kb_lb_tile<BLOCK, Wide, E, R, NT>(v, active_tile<BLOCK>(v.act, xcd_tile()), k);
}

// l2err of each LB instance's last step of the round (one wave per instance slot)
// (called by the first wave of a block)
__device__ __forceinline__ void kb_lb_finalize_inst(const BInst* insts, unsigned long long* slots, double* l2ring,
                                                    Ctrl* ctrl, uint32_t id) {
    const BInst& in = insts[id];
    if (!in.lb_on || in.lb_redo || in.lb_n <= 0) return;
    lb_finalize(slots + (uint64_t)id * 3 * kSlots, l2ring + (uint64_t)id * kRing, ctrl + id, in.lb_t0 + in.lb_n - 1,
                in.check_base, in.check_every);
}

__global__ __launch_bounds__(64) void kb_lb_finalize(const BInst* insts, unsigned long long* slots, double* l2ring,
                                                     Ctrl* ctrl) {
    kb_lb_finalize_inst(insts, slots, l2ring, ctrl, blockIdx.x);
}

template <typename W>
struct BatchMpView {
    const BInst* insts;
    const uint32_t* tile_inst;
    const uint32_t* act;          // the 256-node tiles of the instances in the phase, ascending
    const W* win;
    const double* rem;
    const W* ustar;
    const uint32_t* nodes;
    const uint32_t* maps;
    double* P0;
    double* P1;
    double* Pa0;
    double* Pa1;
    double* part0;                // per-block VACF partials by step parity, indexed by BLOCK-node tile
    double* part1;
    double* vtrace;               // [k][inst][3]
    uint32_t vt_stride;           // instance slots per vtrace row
    Ctrl* ctrl;
    uint64_t Np;
};

// MP launch k of a round on one BLOCK-node tile (mp_repeat: every launch
// recomputes step mp_it0 + 1, for timing)
template <int BLOCK, bool Wide, typename W, bool XR>
__device__ __forceinline__ void kb_mp_tile(const BatchMpView<W>& v, uint32_t tile, int64_t k, int32_t reduce_prev) {
    const uint32_t id = v.tile_inst[tile / (kTile / BLOCK)];
    if (id == kNoInst) return;
    const BInst& in = v.insts[id];
    if (!in.mp_on || k >= in.mp_n) return;
    const int64_t it = in.mp_it0 + (in.mp_repeat ? 0 : k) + 1;  // P(it) -> P(it+1); P(m) lives in buffer (m-1)&1
    const uint32_t first = in.tile_first * (kTile / BLOCK);
    MpStepArgs<W> a;
    a.win = v.win;
    a.rem = v.rem;
    a.ustar = v.ustar;
    a.nodes = v.nodes;
    a.map = v.maps + in.map_off;
    a.Pn = (it - 1) & 1 ? v.P1 : v.P0;
    a.Px = it & 1 ? v.P1 : v.P0;
    a.Pan = (it - 1) & 1 ? v.Pa1 : v.Pa0;
    a.Pax = it & 1 ? v.Pa1 : v.Pa0;
    a.partials = it & 1 ? v.part1 : v.part0;
    a.prev_partials = (reduce_prev && !in.mp_repeat && k > 0)
                          ? ((it - 1) & 1 ? v.part1 : v.part0) + 3 * (uint64_t)first
                          : nullptr;
    a.prev_vtrace = v.vtrace + ((uint64_t)(k - 1) * v.vt_stride + id) * 3;
    a.it = it;
    a.ctrl = v.ctrl + id;
    a.dlx = in.dlx;
    a.dly = in.dly;
    a.lz = in.lz;
    a.Nf = in.node_end;
    a.Np = v.Np;
    a.any_ads = in.any_ads;
    a.tp = in.tp;
    mp_step_tile<BLOCK, Wide, 1, W, kCoefStored, XR>(a, tile, tile == first, in.ntiles * (kTile / BLOCK));
}

// XR: some live instance is more than one node wide along x (the MP step's
// reuse of the lanes' z-neighbours, compiled out for batches of 2-D instances)
template <int BLOCK, bool Wide, typename W, bool XR>
__global__ __launch_bounds__(BLOCK) void kb_mp_step(BatchMpView<W> v, int64_t k, int32_t reduce_prev) {
    kb_mp_tile<BLOCK, Wide, W, XR>(v, active_tile<BLOCK>(v.act, xcd_tile()), k, reduce_prev);
}

// VACF of an MP instance's last step of the round: partials of TILE-node
// tiles, reduced by the NT threads of one block
template <int TILE, int NT>
__device__ __forceinline__ void kb_mp_reduce_inst(const BInst* insts, const double* part0, const double* part1,
                                                  double* vtrace, uint32_t vt_stride, Ctrl* ctrl, uint32_t id) {
    __shared__ double lds[NT / 64 * 3];
    const BInst& in = insts[id];
    if (!in.mp_on || in.mp_n <= 0 || in.mp_repeat) return;
    const int64_t it = in.mp_it0 + in.mp_n;
    const uint32_t first = in.tile_first * (kTile / TILE);
    const double* part = (it & 1 ? part1 : part0) + 3 * (uint64_t)first;
    mp_reduce_block<3, NT>(part, in.ntiles * (kTile / TILE), vtrace + ((uint64_t)(in.mp_n - 1) * vt_stride + id) * 3,
                           ctrl + id, 1, it, 0, 0.0, 0.0, lds);
}

// (one block per instance slot)
template <int BLOCK>
__global__ __launch_bounds__(kBlock) void kb_mp_reduce_last(const BInst* insts, const double* part0,
                                                            const double* part1, double* vtrace, uint32_t vt_stride,
                                                            Ctrl* ctrl) {
    kb_mp_reduce_inst<BLOCK, kBlock>(insts, part0, part1, vtrace, vt_stride, ctrl, blockIdx.x);
}

// A whole round in one cooperative launch: S_lb LB steps, the LB closing
// reduction, S_mp MP steps and the MP closing reduction, every block looping
// over its share of the tiles with a grid-wide barrier between steps.  The
// per-tile work and its order of partial sums are those of the per-step
// launches, so the results are the same bits.  Every block runs every
// barrier: the loops have grid-uniform bounds and the per-tile functions
// only return from the tile, never from the kernel.
template <int BLOCK, bool Wide, typename E, typename R, typename W, bool XR>
__global__ __launch_bounds__(BLOCK) void kb_round(BatchView<E, R> lv, BatchMpView<W> mv, int64_t S_lb, int64_t S_mp,
                                                  uint32_t ntiles, uint32_t ninst, int32_t lb_redo_only) {
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    // the tiles of XCD x (blocks b = x mod 8) are one contiguous range
    const uint32_t nb = gridDim.x, b = blockIdx.x, x = b & 7u, local = b >> 3;
    const uint32_t nblk = (nb >> 3) + (x < (nb & 7u) ? 1u : 0u);
    const uint32_t per = (ntiles + 7u) / 8u;
    const uint32_t t0 = x * per, t1 = min(ntiles, t0 + per);
    for (int64_t k = 0; k < S_lb; ++k) {
        for (uint32_t l = t0 + local; l < t1; l += nblk) kb_lb_tile<BLOCK, Wide, E, R>(lv, active_tile<BLOCK>(lv.act, l), k);
        grid.sync();
    }
    if (S_lb > 0 && !lb_redo_only) {
        for (uint32_t id = b; id < ninst; id += nb)
            if (threadIdx.x < 64) kb_lb_finalize_inst(lv.insts, lv.slots, lv.l2ring, lv.ctrl, id);
        grid.sync();
    }
    for (int64_t k = 0; k < S_mp; ++k) {
        for (uint32_t l = t0 + local; l < t1; l += nblk)
            kb_mp_tile<BLOCK, Wide, W, XR>(mv, active_tile<BLOCK>(mv.act, l), k, 1);
        grid.sync();
    }
    if (S_mp > 0)
        for (uint32_t id = b; id < ninst; id += nb)
            kb_mp_reduce_inst<BLOCK, BLOCK>(mv.insts, mv.part0, mv.part1, mv.vtrace, mv.vt_stride, mv.ctrl, id);
}

// ------------------------------------------------------------------ host

int hip_status(hipError_t e) {
    if (e == hipSuccess) return LBX_OK;
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? LBX_ERR_NOMEM : LBX_ERR_DEVICE;
}
int hip_status(int st) { return st; }

#define LBX_TRY(expr)                            \
    do {                                         \
        int _st = hip_status(expr);              \
        if (_st != LBX_OK) return _st;           \
    } while (0)

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

// first-fit allocator of [0, cap) in units
struct RangeAlloc {
    std::map<uint64_t, uint64_t> free_;  // start -> length
    void init(uint64_t cap) {
        free_.clear();
        if (cap) free_[0] = cap;
    }
    bool take(uint64_t n, uint64_t* start) {
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                *start = it->first;
                const uint64_t rest = it->second - n, s = it->first + n;
                free_.erase(it);
                if (rest) free_[s] = rest;
                return true;
            }
        return false;
    }
    void give(uint64_t start, uint64_t n) {
        auto it = free_.emplace(start, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t largest() const {
        uint64_t m = 0;
        for (auto& kv : free_) m = std::max(m, kv.second);
        return m;
    }
};

enum Phase : int32_t { kFree = 0, kLb = 1, kRedo = 2, kLbDone = 3, kMp = 4, kDone = 5, kFailed = 6 };

struct HostInst {
    Phase phase = kFree;
    int32_t status = LBX_OK;
    int32_t lx = 0, ly = 0, lz = 0, stage = 1, exit_reason = 0;
    uint32_t N = 0, Nf = 0;
    uint64_t map_n = 0;
    double F[3] = {0, 0, 0};
    double Db = 0, ka = 0, kd = 0, lambda = 0;
    int64_t t = 0, max_lb = 0, switch_step = INT64_MAX, check_after = 2;
    int64_t mp_it = 0, mp_target = 0;
    std::vector<uint32_t> nodes_h;  // full-box index | interfacial bit
    std::vector<double> vacf;       // (mp_it+1) x 3
    double deff[3] = {0, 0, 0};
    int64_t nfi = 0;
};

}  // namespace

struct lbxb {
    int32_t prec = LBX_PREC_FP64;
    bool wide = false;       // 64-bit addressing of the population buffers
    uint32_t n_xrows = 0;    // live instances more than one node wide along x
    bool wide_auto = false;  // what the capacity alone selects
    uint64_t Np = 0, cap_box = 0;
    uint32_t cap_inst = 0, ntiles_cap = 0;
    int32_t lb_block = 256, mp_block = 256;
    bool graphs = true;
    bool persistent = false;  // a round's launches of each phase as one cooperative launch
    int32_t num_cu = 0;
    int32_t check_every = kCheckEvery;  // FP32 storage: stage-2 steps between l2err measurements
    int64_t floor_window = 1000;        // FP32 storage: l2err floor test of the instances added from now on
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    // device arrays
    double* f[2] = {nullptr, nullptr};
    float* g[2] = {nullptr, nullptr};
    float* j32 = nullptr;
    double* j = nullptr;
    double* rho = nullptr;
    uint32_t* nodes = nullptr;
    uint32_t* maps = nullptr;
    uint32_t* tile_inst = nullptr;
    BInst* insts = nullptr;
    unsigned long long* slots = nullptr;
    double* l2ring = nullptr;
    Ctrl* ctrl = nullptr;
    double* win = nullptr;
    float* win32 = nullptr;
    double* ustar = nullptr;
    float* ustar32 = nullptr;
    double* rem = nullptr;
    double* P[2] = {nullptr, nullptr};
    double* Pa[2] = {nullptr, nullptr};
    double* part[2] = {nullptr, nullptr};
    double* vtrace = nullptr;
    // host state
    std::vector<BInst> tab;
    std::vector<Ctrl> ctrl_h;  // per instance slot: what reset_inst_ctrl uploads (kept until the stream syncs)
    std::vector<HostInst> hi;
    std::vector<uint32_t> tile_inst_h;
    // the tiles of the instances with LB (0) or MP (1) work in the launches
    // about to run (built by upload_active; host copy kept until the stream syncs)
    uint32_t* act[2] = {nullptr, nullptr};
    std::vector<uint32_t> act_h[2];
    uint32_t n_act[2] = {0, 0};
    RangeAlloc node_alloc, map_alloc;
    // graphs of a round's launches, keyed by (phase, launches)
    struct Graph {
        hipGraphExec_t exec = nullptr;
        int32_t block = 0;
        uint32_t grid_tiles = 0;  // active 256-node tiles of the phase
        bool xr = false;  // the MP launches' kernel reuses the lanes' z-neighbours
    };
    std::map<std::pair<int, int64_t>, Graph> graph_cache;
    // statistics
    double lb_ms = 0, mp_ms = 0;
    int64_t lb_box_updates = 0, lb_fluid_updates = 0, mp_box_updates = 0, mp_fluid_updates = 0;
};

namespace {

void clear_graphs(lbxb* b) {
    for (auto& kv : b->graph_cache)
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    b->graph_cache.clear();
}

void free_all(lbxb* b) {
    clear_graphs(b);
    for (int k = 0; k < 2; ++k) {
        dfree(b->f[k]);
        dfree(b->g[k]);
        dfree(b->P[k]);
        dfree(b->Pa[k]);
        dfree(b->part[k]);
    }
    dfree(b->j32);
    dfree(b->j);
    dfree(b->rho);
    dfree(b->nodes);
    dfree(b->maps);
    dfree(b->tile_inst);
    dfree(b->act[0]);
    dfree(b->act[1]);
    dfree(b->insts);
    dfree(b->slots);
    dfree(b->l2ring);
    dfree(b->ctrl);
    dfree(b->win);
    dfree(b->win32);
    dfree(b->ustar);
    dfree(b->ustar32);
    dfree(b->rem);
    dfree(b->vtrace);
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e), e = nullptr;
    if (b->stream) (void)hipStreamDestroy(b->stream), b->stream = nullptr;
}

// populations stored as FP32 deviations (FP32-shift and FP32-store forms);
// their moment propagation keeps its probabilities in FP32
bool shift(const lbxb* b) { return b->prec == LBX_PREC_FP32_SHIFT || b->prec == LBX_PREC_FP32_STORE; }

// The upload reads the batch's ctrl_h[id], which no other call changes before
// the stream has synchronised (an instance is reset at most once between two
// synchronisations).
int reset_inst_ctrl(lbxb* b, uint32_t id, int64_t check_after) {
    Ctrl& h = b->ctrl_h[id];
    h = Ctrl{};
    h.negpop = -1;
    for (int k = 0; k < 2; ++k) {
        h.l2min[k] = HUGE_VAL;
        h.l2min_step[k] = check_after;
    }
    return hip_status(hipMemcpyAsync(b->ctrl + id, &h, sizeof(Ctrl), hipMemcpyHostToDevice, b->stream));
}

uint32_t grid_of(const lbxb* b, int block, int phase /*0 LB, 1 MP*/) { return b->n_act[phase] * (kTile / block); }

// The active tile lists of the work table as it stands (lb_on / mp_on), for
// the launches that follow; call after the table changes which instances take
// part and before those launches.
int upload_active(lbxb* b) {
    LBX_TRY(hipStreamSynchronize(b->stream));  // the previous lists' uploads have been read
    for (int ph = 0; ph < 2; ++ph) {
        std::vector<std::pair<uint32_t, uint32_t>> r;  // (tile_first, ntiles) of the phase's instances
        for (uint32_t id = 0; id < b->cap_inst; ++id) {
            const BInst& in = b->tab[id];
            if (ph == 0 ? in.lb_on : in.mp_on) r.emplace_back(in.tile_first, in.ntiles);
        }
        std::sort(r.begin(), r.end());
        std::vector<uint32_t>& a = b->act_h[ph];
        a.clear();
        for (const auto& x : r)
            for (uint32_t t = 0; t < x.second; ++t) a.push_back(x.first + t);
        b->n_act[ph] = (uint32_t)a.size();
        if (!a.empty())
            LBX_TRY(hipMemcpyAsync(b->act[ph], a.data(), a.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                                   b->stream));
    }
    return LBX_OK;
}

template <typename E, typename R = E>
BatchView<E, R> lb_view(lbxb* b) {
    BatchView<E, R> v;
    v.insts = b->insts;
    v.tile_inst = b->tile_inst;
    v.act = b->act[0];
    if constexpr (std::is_same<E, float>::value) {
        v.f0 = b->g[0];
        v.f1 = b->g[1];
        if constexpr (std::is_same<R, float>::value) v.j = b->j32; else v.j = b->j;
    } else {
        v.f0 = b->f[0];
        v.f1 = b->f[1];
        v.j = b->j;
    }
    v.nodes = b->nodes;
    v.maps = b->maps;
    v.slots = b->slots;
    v.l2ring = b->l2ring;
    v.ctrl = b->ctrl;
    v.Np = b->Np;
    return v;
}

template <typename W>
BatchMpView<W> mp_view(lbxb* b) {
    BatchMpView<W> v;
    v.insts = b->insts;
    v.tile_inst = b->tile_inst;
    v.act = b->act[1];
    if constexpr (std::is_same<W, float>::value) {
        v.win = b->win32;
        v.ustar = b->ustar32;
    } else {
        v.win = b->win;
        v.ustar = b->ustar;
    }
    v.rem = b->rem;
    v.nodes = b->nodes;
    v.maps = b->maps;
    v.P0 = b->P[0];
    v.P1 = b->P[1];
    v.Pa0 = b->Pa[0];
    v.Pa1 = b->Pa[1];
    v.part0 = b->part[0];
    v.part1 = b->part[1];
    v.vtrace = b->vtrace;
    v.vt_stride = b->cap_inst;
    v.ctrl = b->ctrl;
    v.Np = b->Np;
    return v;
}

// f(std::integral_constant<int, B>{}) for the block size B of a launch (64, 128 or 256 nodes)
template <typename F>
void with_block(int block, F&& f) {
    if (block == 64)
        f(std::integral_constant<int, 64>{});
    else if (block == 128)
        f(std::integral_constant<int, 128>{});
    else
        f(std::integral_constant<int, 256>{});
}

// the LB kernel of one launch: addressing form, and streaming stores when the
// active tiles' two population sets outgrow the last-level cache (a function
// of the active tile count, which a cached graph's validity already checks)
template <int BL, typename E, typename R = E>
auto lb_kernel(const lbxb* b) {
    const bool nt = 2ull * kQ * b->n_act[0] * kTile * sizeof(E) > (uint64_t)LBX_NT_MIN_BYTES;
    if (b->wide) return nt ? kb_lb_step<BL, true, E, R, true> : kb_lb_step<BL, true, E, R>;
    return nt ? kb_lb_step<BL, false, E, R, true> : kb_lb_step<BL, false, E, R>;
}

// launch LB step k of a round over the tiles of the instances with LB work
void launch_lb(lbxb* b, int64_t k, int block) {
    const uint32_t grid = grid_of(b, block, 0);
    if (!grid) return;
    with_block(block, [&](auto BT) {
        constexpr int BL = decltype(BT)::value;
        if (b->prec == LBX_PREC_FP32_SHIFT) {
            auto v = lb_view<float>(b);
            hipLaunchKernelGGL((lb_kernel<BL, float>(b)), dim3(grid), dim3(BL), 0, b->stream, v, k);
        } else if (b->prec == LBX_PREC_FP32_STORE) {
            auto v = lb_view<float, double>(b);
            hipLaunchKernelGGL((lb_kernel<BL, float, double>(b)), dim3(grid), dim3(BL), 0, b->stream, v, k);
        } else {
            auto v = lb_view<double>(b);
            hipLaunchKernelGGL((lb_kernel<BL, double>(b)), dim3(grid), dim3(BL), 0, b->stream, v, k);
        }
    });
}

void launch_mp(lbxb* b, int64_t k, int block, int reduce_prev) {
    const uint32_t grid = grid_of(b, block, 1);
    if (!grid) return;
    const bool xr = b->n_xrows > 0;
    with_block(block, [&](auto BT) {
        constexpr int BL = decltype(BT)::value;
        if (shift(b)) {
            auto v = mp_view<float>(b);
            auto kern = b->wide ? (xr ? kb_mp_step<BL, true, float, true> : kb_mp_step<BL, true, float, false>)
                                : (xr ? kb_mp_step<BL, false, float, true> : kb_mp_step<BL, false, float, false>);
            hipLaunchKernelGGL(kern, dim3(grid), dim3(BL), 0, b->stream, v, k, reduce_prev);
        } else {
            auto v = mp_view<double>(b);
            auto kern = b->wide ? (xr ? kb_mp_step<BL, true, double, true> : kb_mp_step<BL, true, double, false>)
                                : (xr ? kb_mp_step<BL, false, double, true> : kb_mp_step<BL, false, double, false>);
            hipLaunchKernelGGL(kern, dim3(grid), dim3(BL), 0, b->stream, v, k, reduce_prev);
        }
    });
}

void launch_lb_finalize(lbxb* b) {
    hipLaunchKernelGGL(kb_lb_finalize, dim3(b->cap_inst), dim3(64), 0, b->stream, b->insts, b->slots, b->l2ring,
                       b->ctrl);
}

void launch_mp_reduce_last(lbxb* b, int block) {
    with_block(block, [&](auto BT) {
        constexpr int BL = decltype(BT)::value;
        hipLaunchKernelGGL(kb_mp_reduce_last<BL>, dim3(b->cap_inst), dim3(kBlock), 0, b->stream, b->insts,
                           b->part[0], b->part[1], b->vtrace, b->cap_inst, b->ctrl);
    });
}

// the S launches of one phase of a round (+ its closing reduction), replayed
// from a captured graph when graphs are on
int run_phase(lbxb* b, int phase /*0 LB, 1 MP*/, int64_t S) {
    if (S <= 0) return LBX_OK;
    const int block = phase == 0 ? b->lb_block : b->mp_block;
    auto emit = [&]() {
        for (int64_t k = 0; k < S; ++k) {
            if (phase == 0)
                launch_lb(b, k, block);
            else
                launch_mp(b, k, block, 1);
        }
        if (phase == 0)
            launch_lb_finalize(b);
        else
            launch_mp_reduce_last(b, block);
    };
    if (!b->graphs) {
        emit();
        return hip_status(hipGetLastError());
    }
    auto key = std::make_pair(phase, S);
    auto it = b->graph_cache.find(key);
    const bool xr = b->n_xrows > 0;
    if (it != b->graph_cache.end() &&
        (it->second.block != block || it->second.grid_tiles != b->n_act[phase] || it->second.xr != xr)) {
        (void)hipGraphExecDestroy(it->second.exec);
        b->graph_cache.erase(it);
        it = b->graph_cache.end();
    }
    if (it == b->graph_cache.end()) {
        // one graph per (phase, round length): bounded by the callers' round
        // sizes, but callers that vary them freely would grow the cache; start
        // over past a few dozen entries
        if (b->graph_cache.size() >= kGraphCacheMax) clear_graphs(b);
        hipGraph_t graph = nullptr;
        LBX_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        emit();
        const hipError_t ce = hipStreamEndCapture(b->stream, &graph);
        if (ce != hipSuccess) return hip_status(ce);
        lbxb::Graph gr;
        const hipError_t ie = hipGraphInstantiate(&gr.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) return hip_status(ie);
        gr.block = block;
        gr.grid_tiles = b->n_act[phase];
        gr.xr = xr;
        it = b->graph_cache.emplace(key, gr).first;
    }
    return hip_status(hipGraphLaunch(it->second.exec, b->stream));
}

// One phase of a round as a single cooperative launch of kb_round (grid: as
// many blocks as can be resident at once, at most one per tile)
int run_phase_persistent(lbxb* b, int phase, int64_t S, int32_t redo_only) {
    if (S <= 0) return LBX_OK;
    const int block = b->lb_block;
    const uint32_t ntiles = grid_of(b, block, phase);
    if (!ntiles) return LBX_OK;
    const bool xr = b->n_xrows > 0;
    auto go = [&](auto etag, auto rtag, auto btag, auto wide) -> int {
        using E = decltype(etag);  // population storage
        using R = decltype(rtag);  // LB arithmetic
        using W = E;               // MP probabilities are stored like the populations
        constexpr int BL = decltype(btag)::value;
        constexpr bool WD = decltype(wide)::value;
        auto kern = xr ? kb_round<BL, WD, E, R, W, true> : kb_round<BL, WD, E, R, W, false>;
        int per_cu = 0;
        LBX_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, BL, 0));
        if (per_cu <= 0) return LBX_ERR_DEVICE;
        const uint32_t grid = std::min<uint32_t>(ntiles, (uint32_t)per_cu * (uint32_t)b->num_cu);
        BatchView<E, R> lv = lb_view<E, R>(b);
        BatchMpView<W> mv = mp_view<W>(b);
        int64_t S_lb = phase == 0 ? S : 0, S_mp = phase == 1 ? S : 0;
        uint32_t nt = ntiles, ni = b->cap_inst;
        int32_t ro = redo_only;
        void* args[] = {&lv, &mv, &S_lb, &S_mp, &nt, &ni, &ro};
        return hip_status(hipLaunchCooperativeKernel((const void*)kern, dim3(grid), dim3(BL), args, 0, b->stream));
    };
    using T_ = std::true_type;
    using F_ = std::false_type;
    int st = LBX_OK;
    with_block(block, [&](auto BT) {
        if (b->prec == LBX_PREC_FP32_SHIFT)
            st = b->wide ? go(float{}, float{}, BT, T_{}) : go(float{}, float{}, BT, F_{});
        else if (b->prec == LBX_PREC_FP32_STORE)
            st = b->wide ? go(float{}, double{}, BT, T_{}) : go(float{}, double{}, BT, F_{});
        else
            st = b->wide ? go(double{}, double{}, BT, T_{}) : go(double{}, double{}, BT, F_{});
    });
    return st;
}

template <typename T>
int download(lbxb* b, const T* src, uint64_t count, T* dst) {
    LBX_TRY(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, b->stream));
    return hip_status(hipStreamSynchronize(b->stream));
}

// extract + MP prep + VACF(0) of instance id, whose LB phase ended at step t
int mp_setup_inst(lbxb* b, uint32_t id) {
    HostInst& h = b->hi[id];
    BInst& in = b->tab[id];
    const int64_t t = h.t;
    const uint32_t gt = in.ntiles;
    const uint32_t* map = b->maps + in.map_off;
    if (t > 0) {
        const int src = (t - 1) & 1, dst = t & 1;
        if (shift(b)) {
            const double* Ft = t > h.switch_step ? h.F : nullptr;
            const double fm[3] = {Ft ? Ft[0] / 2.0 : 0.0, Ft ? Ft[1] / 2.0 : 0.0, Ft ? Ft[2] / 2.0 : 0.0};
            hipLaunchKernelGGL(k_extract_shift, dim3(gt), dim3(kBlock), 0, b->stream, b->g[src], b->g[dst], b->rho,
                               b->j, b->nodes, map, in.dlx, in.dly, in.lz, in.node_off, in.node_end, b->Np, in.rho0,
                               fm[0], fm[1], fm[2]);
        } else {
            hipLaunchKernelGGL(k_extract, dim3(gt), dim3(kBlock), 0, b->stream, b->f[src], b->f[dst], b->rho,
                               b->nodes, map, in.dlx, in.dly, in.lz, in.node_off, in.node_end, b->Np);
        }
    }
    const double eps = DBL_EPSILON;
    TracerParams tp{};
    const double K = (std::fabs(h.kd) <= eps) ? 0.0 : h.ka / h.kd;  // module_moment_propagation.f90:46-50
    tp.ads[0] = std::fabs(K) > eps;                                 // :52-56
    tp.bw[0] = 1.0 / ((double)h.Nf + K * (double)h.nfi);            // :97
    tp.ka[0] = h.ka;
    tp.kd[0] = h.kd;
    tp.one_minus_kd[0] = 1.0 - h.kd;
    in.tp = tp;
    in.any_ads = tp.ads[0];
    const uint64_t n = (uint64_t)gt * kTile;
    // Pa = 0 on the instance's range of both buffers (comp_index layout: pair plane + single plane)
    for (int k = 0; k < 2; ++k) {
        LBX_TRY(hipMemsetAsync(b->Pa[k] + 2 * (uint64_t)in.node_off, 0, 2 * n * sizeof(double), b->stream));
        LBX_TRY(hipMemsetAsync(b->Pa[k] + 2 * b->Np + in.node_off, 0, n * sizeof(double), b->stream));
    }
    LBX_TRY(reset_inst_ctrl(b, id, 2));
    auto prep = [&](auto wtag) {
        using W = decltype(wtag);
        MpPrepArgs<W> a;
        a.rho = b->rho;
        a.j = b->j;
        a.nodes = b->nodes;
        a.map = map;
        if constexpr (std::is_same<W, float>::value) {
            a.win = b->win32;
            a.ustar = b->ustar32;
        } else {
            a.win = b->win;
            a.ustar = b->ustar;
        }
        a.rem = b->rem;
        a.P0 = b->P[0];
        a.partials = b->part[0];
        a.ctrl = b->ctrl + id;
        a.dlx = in.dlx;
        a.dly = in.dly;
        a.lz = in.lz;
        a.Nf = in.node_end;
        a.Np = b->Np;
        a.fx = h.F[0];
        a.fy = h.F[1];
        a.fz = h.F[2];
        a.lambda = h.lambda;
        a.eps = eps;
        a.tile0 = in.tile_first;
        a.tp = tp;
        hipLaunchKernelGGL((k_mp_prep<1, W>), dim3(gt), dim3(kBlock), 0, b->stream, a);
    };
    if (shift(b))
        prep(float{});
    else
        prep(double{});
    hipLaunchKernelGGL(k_mp_reduce<3>, dim3(1), dim3(kBlock), 0, b->stream, b->part[0] + 3 * (uint64_t)in.tile_first,
                       gt, nullptr, b->ctrl + id, 0, (int64_t)0, 0, 0.0, 0.0);
    LBX_TRY(hipGetLastError());
    Ctrl c;
    LBX_TRY(download(b, b->ctrl + id, 1, &c));
    if (c.restpart) {
        h.phase = kFailed;
        h.status = LBX_ERR_RESTPART;
        return LBX_OK;
    }
    h.vacf.assign(c.vacf0, c.vacf0 + 3);
    h.mp_it = 0;
    h.phase = h.mp_target > 0 ? kMp : kDone;
    if (h.phase == kDone)
        for (int a_ = 0; a_ < 3; ++a_) h.deff[a_] = c.deff_s[a_] + c.deff_c[a_];
    return LBX_OK;
}

}  // namespace

// ------------------------------------------------------------------ C-ABI

extern "C" {

int64_t lbxb_footprint_bytes(int32_t precision, int64_t capacity_nodes, int64_t capacity_box, int32_t max_instances) {
    if (capacity_nodes <= 0 || capacity_box <= 0 || max_instances <= 0) return -1;
    const uint64_t Np = ((uint64_t)capacity_nodes + kTile - 1) / kTile * kTile;
    const uint64_t pop = (precision == LBX_PREC_FP32_SHIFT || precision == LBX_PREC_FP32_STORE) ? 4 : 8;
    uint64_t by = 2 * kQ * Np * pop + (precision == LBX_PREC_FP32_SHIFT ? 3 * Np * 4 : 0) + 4 * Np * 8 + Np * 4 +
                  (uint64_t)capacity_box * 4 + Np / kTile * 4;
    by += (uint64_t)max_instances * (sizeof(BInst) + 3 * kSlots * 8 + kRing * 8 + sizeof(Ctrl));
    by += (18 + 3) * Np * pop + Np * 8 + 4 * 3 * Np * 8 + 2 * 2 * (Np / 128) * 3 * 8;
    by += (uint64_t)kRoundMax * max_instances * 3 * 8;
    return (int64_t)by;
}

int lbxb_create(int32_t precision, int64_t capacity_nodes, int64_t capacity_box, int32_t max_instances,
                lbxb** out) {
    if (!out) return LBX_ERR_ARG;
    *out = nullptr;
    if (precision != LBX_PREC_FP64 && precision != LBX_PREC_FP32_SHIFT && precision != LBX_PREC_FP32_STORE)
        return LBX_ERR_ARG;
    if (capacity_nodes <= 0 || capacity_box <= 0 || max_instances <= 0) return LBX_ERR_ARG;
    const uint64_t Np = ((uint64_t)capacity_nodes + kTile - 1) / kTile * kTile;
    if (kQ * Np >= (1ull << 32) || (uint64_t)capacity_box >= (1ull << 32)) return LBX_ERR_ARG;
    lbxb* b = new (std::nothrow) lbxb();
    if (!b) return LBX_ERR_NOMEM;
    b->prec = precision;
    b->Np = Np;
    b->cap_box = (uint64_t)capacity_box;
    b->cap_inst = (uint32_t)max_instances;
    b->ntiles_cap = (uint32_t)(Np / kTile);
    // 32-bit byte offsets reach every array while the populations (the
    // largest: 19 per node, FP64 or FP32 by form) stay below 4 GiB
    b->wide = b->wide_auto =
        (uint64_t)kQ * Np * (precision == LBX_PREC_FP64 ? sizeof(double) : sizeof(float)) >= (1ull << 32);
    {
        int dev = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&b->num_cu, hipDeviceAttributeMultiprocessorCount, dev);
    }
    const bool sh = precision == LBX_PREC_FP32_SHIFT || precision == LBX_PREC_FP32_STORE;
    int st = LBX_OK;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) st = LBX_ERR_DEVICE;
    for (int k = 0; k < 2; ++k) {
        if (!st && !sh) st = dalloc(&b->f[k], kQ * Np);
        if (!st && sh) st = dalloc(&b->g[k], kQ * Np);
        if (!st) st = dalloc(&b->P[k], 3 * Np);
        if (!st) st = dalloc(&b->Pa[k], 3 * Np);
        if (!st) st = dalloc(&b->part[k], (Np / 64) * 3);  // one VACF partial per 64-node block
    }
    if (!st && precision == LBX_PREC_FP32_SHIFT) st = dalloc(&b->j32, 3 * Np);
    if (!st) st = dalloc(&b->j, 3 * Np);
    if (!st) st = dalloc(&b->rho, Np);
    if (!st) st = dalloc(&b->nodes, Np);
    if (!st) st = dalloc(&b->maps, (uint64_t)capacity_box);
    if (!st) st = dalloc(&b->tile_inst, (uint64_t)b->ntiles_cap);
    if (!st) st = dalloc(&b->act[0], (uint64_t)b->ntiles_cap);
    if (!st) st = dalloc(&b->act[1], (uint64_t)b->ntiles_cap);
    if (!st) st = dalloc(&b->insts, (uint64_t)max_instances);
    if (!st) st = dalloc(&b->slots, (uint64_t)max_instances * 3 * kSlots);
    if (!st) st = dalloc(&b->l2ring, (uint64_t)max_instances * kRing);
    if (!st) st = dalloc(&b->ctrl, (uint64_t)max_instances);
    if (!st && !sh) st = dalloc(&b->win, 18 * Np);
    if (!st && !sh) st = dalloc(&b->ustar, 3 * Np);
    if (!st && sh) st = dalloc(&b->win32, 18 * Np);
    if (!st && sh) st = dalloc(&b->ustar32, 3 * Np);
    if (!st) st = dalloc(&b->rem, Np);
    if (!st) st = dalloc(&b->vtrace, (uint64_t)kRoundMax * max_instances * 3);
    if (!st && (hipEventCreate(&b->ev[0]) != hipSuccess || hipEventCreate(&b->ev[1]) != hipSuccess))
        st = LBX_ERR_DEVICE;
    if (!st) {
        b->tab.assign(max_instances, BInst{});
        b->ctrl_h.assign(max_instances, Ctrl{});
        b->hi.assign(max_instances, HostInst{});
        b->tile_inst_h.assign(b->ntiles_cap, kNoInst);
        b->node_alloc.init(b->ntiles_cap);
        b->map_alloc.init(b->cap_box);
        st = hip_status(hipMemcpyAsync(b->tile_inst, b->tile_inst_h.data(), b->ntiles_cap * sizeof(uint32_t),
                                       hipMemcpyHostToDevice, b->stream));
    }
    if (!st) st = hip_status(hipMemcpyAsync(b->insts, b->tab.data(), b->tab.size() * sizeof(BInst),
                                            hipMemcpyHostToDevice, b->stream));
    if (!st) st = hip_status(hipStreamSynchronize(b->stream));
    if (st) {
        free_all(b);
        delete b;
        return st;
    }
    *out = b;
    return LBX_OK;
}

int lbxb_destroy(lbxb* b) {
    if (!b) return LBX_ERR_ARG;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    free_all(b);
    delete b;
    return LBX_OK;
}

int lbxb_add(lbxb* b, int32_t lx, int32_t ly, int32_t lz, const int8_t* nature, double tau, double rho0,
             const double fext[3], double target_error, int64_t max_lb_steps, double Db, double ka, double kd,
             int64_t mp_steps, int32_t* id_out) {
    if (!b || !nature || !fext || !id_out || lx <= 0 || ly <= 0 || lz <= 0 || mp_steps < 0 || max_lb_steps < 0)
        return LBX_ERR_ARG;
    *id_out = -1;
    if (!(tau >= 0.5)) return LBX_ERR_TAU;
    const double eps = DBL_EPSILON;
    if (ka < -eps || kd < -eps || !(Db > eps)) return LBX_ERR_ARG;
    const uint64_t N = (uint64_t)lx * ly * lz;
    if (N >= (1ull << 29)) return LBX_ERR_ARG;
    uint32_t id = 0;
    while (id < b->cap_inst && b->hi[id].phase != kFree) ++id;
    if (id == b->cap_inst) return LBX_ERR_NOMEM;
    HostInst h;
    std::vector<int8_t> itf(N);
    lbx_detect_interfacial(lx, ly, lz, nature, itf.data());
    for (uint32_t r = 0; r < N; ++r)
        if (nature[r] == 0) {
            h.nodes_h.push_back(r | (itf[r] ? kNodeInterfacial : 0u));
            h.nfi += itf[r] != 0;
        }
    h.Nf = (uint32_t)h.nodes_h.size();
    if (h.Nf == 0) return LBX_ERR_ARG;
    const uint64_t ntiles = (h.Nf + kTile - 1) / kTile;
    uint64_t tile0 = 0, map0 = 0;
    if (!b->node_alloc.take(ntiles, &tile0)) return LBX_ERR_NOMEM;  // HBM admission
    if (!b->map_alloc.take(N, &map0)) {
        b->node_alloc.give(tile0, ntiles);
        return LBX_ERR_NOMEM;
    }
    h.lx = lx;
    h.ly = ly;
    h.lz = lz;
    h.N = (uint32_t)N;
    h.map_n = N;
    for (int a = 0; a < 3; ++a) h.F[a] = fext[a];
    h.Db = Db;
    h.ka = ka;
    h.kd = kd;
    h.lambda = 4.0 * Db / kKbt;  // calc_lambda, module_moment_propagation.f90:331-337
    h.max_lb = max_lb_steps;
    h.mp_target = mp_steps;
    h.phase = kLb;
    BInst in{};
    in.node_off = (uint32_t)(tile0 * kTile);
    in.node_end = in.node_off + h.Nf;
    in.tile_first = (uint32_t)tile0;
    in.ntiles = (uint32_t)ntiles;
    in.map_off = (uint32_t)map0;
    in.lz = (uint32_t)lz;
    in.dlx.set((uint32_t)lx);
    in.dly.set((uint32_t)ly);
    in.A = 1.0 - 1.0 / tau;  // module_collision.f90:99-101
    in.B = 1.0 / tau;
    in.C = 1.0 - 1.0 / (2.0 * tau);
    if (b->prec == LBX_PREC_FP32_SHIFT) {  // FP32 collision: A + B = 1 exactly
        float a32, b32;
        fp32_relaxation(tau, &a32, &b32);
        in.A = a32;
        in.B = b32;
    }
    in.check_every = 1;
    in.check_base = 0;
    in.check_inv = 1.0;
    in.rho0 = rho0;
    in.target = target_error;
    in.floor_window = shift(b) ? b->floor_window : 0;
    // nodes (global packed indices of the instance's fluid nodes) and its map
    std::vector<uint32_t> gnodes(h.nodes_h);
    std::vector<uint32_t> map(N, kSolidSlot);
    for (uint32_t i = 0; i < h.Nf; ++i) map[h.nodes_h[i] & kIndexMask] = in.node_off + i;
    int st = hip_status(hipMemcpyAsync(b->nodes + in.node_off, gnodes.data(), h.Nf * sizeof(uint32_t),
                                       hipMemcpyHostToDevice, b->stream));
    if (!st) st = hip_status(hipMemcpyAsync(b->maps + map0, map.data(), N * sizeof(uint32_t), hipMemcpyHostToDevice,
                                            b->stream));
    // rest state and the opening collision of step 1 (no force)
    if (!st) {
        if (shift(b)) {
            hipLaunchKernelGGL(k_init_rest, dim3(ntiles), dim3(kBlock), 0, b->stream, nullptr, b->g[0], b->j32,
                               b->rho, b->j, in.node_off, in.node_end, b->Np, rho0);
            if (b->prec == LBX_PREC_FP32_SHIFT)
                hipLaunchKernelGGL(k_collide_packed_shift<float>, dim3(ntiles), dim3(kBlock), 0, b->stream, b->g[0],
                                   b->rho, b->j, b->j32, 0.0f, 0.0f, 0.0f, in.node_off, in.node_end, b->Np,
                                   (float)in.A, (float)in.B, (float)in.C, rho0);
            else  // FP32-store: j_old is j itself
                hipLaunchKernelGGL(k_collide_packed_shift<double>, dim3(ntiles), dim3(kBlock), 0, b->stream, b->g[0],
                                   b->rho, b->j, b->j, 0.0, 0.0, 0.0, in.node_off, in.node_end, b->Np, in.A, in.B,
                                   in.C, rho0);
        } else {
            hipLaunchKernelGGL(k_init_rest, dim3(ntiles), dim3(kBlock), 0, b->stream, b->f[0], nullptr, nullptr,
                               b->rho, b->j, in.node_off, in.node_end, b->Np, rho0);
            hipLaunchKernelGGL(k_collide_packed, dim3(ntiles), dim3(kBlock), 0, b->stream, b->f[0], b->rho, b->j,
                               0.0, 0.0, 0.0, in.node_off, in.node_end, b->Np, in.A, in.B, in.C);
        }
        st = hip_status(hipGetLastError());
    }
    if (!st) st = reset_inst_ctrl(b, id, 2);
    if (!st) st = hip_status(hipMemsetAsync(b->slots + (uint64_t)id * 3 * kSlots, 0,
                                            3 * kSlots * sizeof(unsigned long long), b->stream));
    if (!st) st = hip_status(hipMemsetAsync(b->l2ring + (uint64_t)id * kRing, 0, kRing * sizeof(double), b->stream));
    for (uint64_t t = 0; t < ntiles; ++t) b->tile_inst_h[tile0 + t] = id;
    if (!st) st = hip_status(hipMemcpyAsync(b->tile_inst + tile0, b->tile_inst_h.data() + tile0,
                                            ntiles * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream));
    if (!st) st = hip_status(hipStreamSynchronize(b->stream));
    if (st) {
        for (uint64_t t = 0; t < ntiles; ++t) b->tile_inst_h[tile0 + t] = kNoInst;
        b->node_alloc.give(tile0, ntiles);
        b->map_alloc.give(map0, N);
        return st;
    }
    b->tab[id] = in;
    b->hi[id] = std::move(h);
    if (lx > 1) ++b->n_xrows;
    *id_out = (int32_t)id;
    return LBX_OK;
}

int lbxb_remove(lbxb* b, int32_t id) {
    if (!b || id < 0 || (uint32_t)id >= b->cap_inst || b->hi[id].phase == kFree) return LBX_ERR_ARG;
    (void)hipStreamSynchronize(b->stream);
    BInst& in = b->tab[id];
    for (uint32_t t = 0; t < in.ntiles; ++t) b->tile_inst_h[in.tile_first + t] = kNoInst;
    LBX_TRY(hipMemcpyAsync(b->tile_inst + in.tile_first, b->tile_inst_h.data() + in.tile_first,
                           in.ntiles * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream));
    b->node_alloc.give(in.tile_first, in.ntiles);
    b->map_alloc.give(in.map_off, b->hi[id].map_n);
    if (b->hi[id].lx > 1) --b->n_xrows;
    b->hi[id] = HostInst{};
    in = BInst{};
    return hip_status(hipStreamSynchronize(b->stream));
}

// One round: redo launches of instances whose stage 1 ended, lb_steps LB
// launches, mp_steps MP launches, then the host bookkeeping (stage switches,
// ends of the LB phase -> MP set-up, ends of the MP phase).
int lbxb_round(lbxb* b, int32_t lb_steps, int32_t mp_steps, int32_t* active) {
    if (!b || lb_steps < 0 || mp_steps < 0 || lb_steps > kRoundMax || mp_steps > kRoundMax) return LBX_ERR_ARG;
    const uint32_t ni = b->cap_inst;
    const double zero[3] = {0.0, 0.0, 0.0};
    // 1. redo of the switch step (equilibration.f90:377-386), with the force on
    bool any_redo = false;
    for (uint32_t id = 0; id < ni; ++id) {
        BInst& in = b->tab[id];
        in.lb_on = in.mp_on = 0;
        HostInst& h = b->hi[id];
        if (h.phase != kRedo) continue;
        any_redo = true;
        LBX_TRY(reset_inst_ctrl(b, id, h.t));
        in.lb_on = 1;
        in.lb_redo = 1;
        in.lb_t0 = h.t;
        in.lb_n = 1;
        in.check_after = INT64_MAX;
        for (int a = 0; a < 3; ++a) {
            in.fm[a] = zero[a] / 2.0;
            in.fc[a] = h.F[a];
        }
        h.stage = 2;
        h.switch_step = h.t;
        h.check_after = h.t;
        h.phase = kLb;
    }
    if (any_redo) {
        LBX_TRY(hipMemcpyAsync(b->insts, b->tab.data(), ni * sizeof(BInst), hipMemcpyHostToDevice, b->stream));
        LBX_TRY(upload_active(b));
        launch_lb(b, 0, b->lb_block);
        LBX_TRY(hipGetLastError());
        LBX_TRY(hipStreamSynchronize(b->stream));  // the table upload has been read
        for (uint32_t id = 0; id < ni; ++id) b->tab[id].lb_on = 0, b->tab[id].lb_redo = 0;
    }
    // 2. this round's work table
    int64_t lb_box = 0, lb_fluid = 0, mp_box = 0, mp_fluid = 0;
    int64_t S_lb = 0, S_mp = 0;
    for (uint32_t id = 0; id < ni; ++id) {
        BInst& in = b->tab[id];
        HostInst& h = b->hi[id];
        if (h.phase == kLb && lb_steps > 0) {
            int64_t n = lb_steps;
            if (h.max_lb > 0) n = std::min<int64_t>(n, h.max_lb - h.t);
            if (n <= 0) continue;
            in.lb_on = 1;
            in.lb_redo = 0;
            in.lb_t0 = h.t + 1;
            in.lb_n = n;
            in.check_after = h.check_after;
            // l2err of every step in stage 1; in stage 2 of the FP32 storage
            // forms every kCheckEvery steps from the force switch
            in.check_every = (h.stage == 2 && shift(b)) ? b->check_every : 1;
            in.check_base = h.stage == 2 ? h.switch_step : 0;
            in.check_inv = 1.0 / in.check_every;
            for (int a = 0; a < 3; ++a) {
                const double F = h.stage == 1 ? 0.0 : h.F[a];
                in.fm[a] = F / 2.0;
                in.fc[a] = F;
            }
            S_lb = std::max(S_lb, n);
        } else if (h.phase == kMp && mp_steps > 0) {
            const int64_t n = std::min<int64_t>(mp_steps, h.mp_target - h.mp_it);
            if (n <= 0) continue;
            in.mp_on = 1;
            in.mp_it0 = h.mp_it;
            in.mp_n = n;
            S_mp = std::max(S_mp, n);
        }
    }
    LBX_TRY(hipMemcpyAsync(b->insts, b->tab.data(), ni * sizeof(BInst), hipMemcpyHostToDevice, b->stream));
    LBX_TRY(upload_active(b));
    if (S_lb > 0) {
        LBX_TRY(hipEventRecord(b->ev[0], b->stream));
        LBX_TRY(b->persistent ? run_phase_persistent(b, 0, S_lb, 0) : run_phase(b, 0, S_lb));
        LBX_TRY(hipEventRecord(b->ev[1], b->stream));
        LBX_TRY(hipEventSynchronize(b->ev[1]));
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, b->ev[0], b->ev[1]);
        b->lb_ms += ms;
    }
    if (S_mp > 0) {
        LBX_TRY(hipEventRecord(b->ev[0], b->stream));
        LBX_TRY(b->persistent ? run_phase_persistent(b, 1, S_mp, 0) : run_phase(b, 1, S_mp));
        LBX_TRY(hipEventRecord(b->ev[1], b->stream));
        LBX_TRY(hipEventSynchronize(b->ev[1]));
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, b->ev[0], b->ev[1]);
        b->mp_ms += ms;
    }
    // 3. bookkeeping
    std::vector<Ctrl> ctrl(ni);
    LBX_TRY(download(b, b->ctrl, ni, ctrl.data()));
    std::vector<double> vt;
    if (S_mp > 0) {
        vt.resize((size_t)S_mp * ni * 3);
        LBX_TRY(download(b, b->vtrace, vt.size(), vt.data()));
    }
    std::vector<uint32_t> to_mp;
    for (uint32_t id = 0; id < ni; ++id) {
        BInst& in = b->tab[id];
        HostInst& h = b->hi[id];
        const Ctrl& c = ctrl[id];
        if (in.lb_on) {
            int64_t last = in.lb_t0 + in.lb_n - 1;
            if (c.negpop >= 0) last = std::min<int64_t>(last, c.negpop);
            if (c.stop) last = std::min<int64_t>(last, c.exit_step);
            lb_box += (last - h.t) * (int64_t)h.N;
            lb_fluid += (last - h.t) * (int64_t)h.Nf;
            if (c.negpop >= 0) {
                h.phase = kFailed;
                h.status = LBX_ERR_NEG_POP;
                h.t = c.negpop;
            } else if (c.stop) {
                h.t = c.exit_step;
                if (h.stage == 1) {
                    h.phase = kRedo;
                } else {
                    h.exit_reason = c.floor_hit ? LBX_EXIT_FLOOR : LBX_EXIT_CONVERGED;
                    h.phase = kLbDone;
                }
            } else {
                h.t = last;
                if (h.max_lb > 0 && h.t >= h.max_lb) {
                    double l2 = 0.0;
                    LBX_TRY(download(b, b->l2ring + (uint64_t)id * kRing + (h.t % kRing), 1, &l2));
                    h.exit_reason = (h.stage == 2 && h.t > h.check_after && l2 <= in.target) ? LBX_EXIT_CONVERGED
                                                                                             : LBX_EXIT_CAP;
                    h.phase = kLbDone;
                }
            }
            if (h.phase == kLbDone) to_mp.push_back(id);
        } else if (in.mp_on) {
            mp_box += in.mp_n * (int64_t)h.N;
            mp_fluid += in.mp_n * (int64_t)h.Nf;
            for (int64_t k = 0; k < in.mp_n; ++k)
                for (int a = 0; a < 3; ++a) h.vacf.push_back(vt[((size_t)k * ni + id) * 3 + a]);
            h.mp_it += in.mp_n;
            if (h.mp_it >= h.mp_target) {
                h.phase = kDone;
                for (int a = 0; a < 3; ++a) h.deff[a] = c.deff_s[a] + c.deff_c[a];
            }
        }
        in.lb_on = in.mp_on = 0;
    }
    b->lb_box_updates += lb_box;
    b->lb_fluid_updates += lb_fluid;
    b->mp_box_updates += mp_box;
    b->mp_fluid_updates += mp_fluid;
    for (uint32_t id : to_mp) LBX_TRY(mp_setup_inst(b, id));
    int32_t n = 0;
    for (uint32_t id = 0; id < ni; ++id) n += b->hi[id].phase >= kLb && b->hi[id].phase <= kMp;
    if (active) *active = n;
    return hip_status(hipStreamSynchronize(b->stream));
}

int lbxb_info(lbxb* b, int32_t id, int32_t* phase, int32_t* status, int64_t* lb_steps, int32_t* exit_reason,
              int64_t* mp_steps_done) {
    if (!b || id < 0 || (uint32_t)id >= b->cap_inst) return LBX_ERR_ARG;
    const HostInst& h = b->hi[id];
    if (phase) *phase = h.phase;
    if (status) *status = h.status;
    if (lb_steps) *lb_steps = h.t;
    if (exit_reason) *exit_reason = h.exit_reason;
    if (mp_steps_done) *mp_steps_done = h.mp_it;
    return LBX_OK;
}

int lbxb_result(lbxb* b, int32_t id, double deff[3], double* vacf, int64_t vacf_rows, double* rho, double* jx,
                double* jy, double* jz) {
    if (!b || id < 0 || (uint32_t)id >= b->cap_inst || b->hi[id].phase == kFree) return LBX_ERR_ARG;
    const HostInst& h = b->hi[id];
    const BInst& in = b->tab[id];
    if (deff)
        for (int a = 0; a < 3; ++a) deff[a] = h.deff[a];
    if (vacf) {
        const int64_t rows = std::min<int64_t>(vacf_rows, (int64_t)h.vacf.size() / 3);
        std::memcpy(vacf, h.vacf.data(), (size_t)rows * 3 * sizeof(double));
    }
    // flow fields of the instance (valid once its LB phase has ended)
    double* dst[4] = {rho, jx, jy, jz};
    for (int k = 0; k < 4; ++k) {
        if (!dst[k]) continue;
        std::vector<double> tmp(h.Nf);
        const double* src = k == 0 ? b->rho + in.node_off : b->j + (uint64_t)(k - 1) * b->Np + in.node_off;
        LBX_TRY(download(b, src, h.Nf, tmp.data()));
        std::memset(dst[k], 0, (size_t)h.N * sizeof(double));
        for (uint32_t i = 0; i < h.Nf; ++i) dst[k][h.nodes_h[i] & kIndexMask] = tmp[i];
    }
    return LBX_OK;
}

int lbxb_set_block(lbxb* b, int32_t lb_block, int32_t mp_block) {
    auto ok = [](int32_t x) { return x == 64 || x == 128 || x == 256; };
    if (!b || !ok(lb_block) || !ok(mp_block)) return LBX_ERR_ARG;
    b->lb_block = lb_block;
    b->mp_block = mp_block;
    return LBX_OK;
}

int lbxb_set_check_interval(lbxb* b, int32_t steps) {
    if (!b || steps < 1 || steps > 4096 || (steps & (steps - 1)) != 0) return LBX_ERR_ARG;
    if (steps != 1 && !shift(b)) return LBX_ERR_ARG;
    b->check_every = steps;
    return LBX_OK;
}

int lbxb_set_floor_window(lbxb* b, int64_t steps) {
    if (!b || steps < 0) return LBX_ERR_ARG;
    if (steps > 0 && !shift(b)) return LBX_ERR_ARG;
    b->floor_window = steps;
    return LBX_OK;
}

int lbxb_set_launch_mode(lbxb* b, int32_t mode) {
    if (!b || mode < LBXB_LAUNCH_STREAM || mode > LBXB_LAUNCH_PERSISTENT) return LBX_ERR_ARG;
    b->graphs = mode == LBXB_LAUNCH_GRAPH;
    b->persistent = mode == LBXB_LAUNCH_PERSISTENT;
    if (!b->graphs) clear_graphs(b);
    return LBX_OK;
}

int lbxb_set_addressing(lbxb* b, int32_t wide) {
    if (!b || wide < 0 || wide > 1) return LBX_ERR_ARG;
    b->wide = b->wide_auto || wide == 1;
    clear_graphs(b);  // captured launches name the kernels of the previous form
    return LBX_OK;
}

int lbxb_addressing(lbxb* b) { return b ? (b->wide ? 1 : 0) : -1; }

int lbxb_launch_mode(lbxb* b) {
    if (!b) return -1;
    return b->persistent ? LBXB_LAUNCH_PERSISTENT : (b->graphs ? LBXB_LAUNCH_GRAPH : LBXB_LAUNCH_STREAM);
}

// The autotuner's search space: the block size of each phase (64, 128 or
// 256 nodes, i.e. 1, 2 or 4 waves per block and 4, 2 or 1 blocks per
// 256-node tile) and the launch mode of the rounds (one launch per step, the
// same launches replayed from a captured graph, or one cooperative
// persistent launch per phase).  Every point is timed on the current state
// without changing it: LB launches redo each LB instance's last completed step
// (same inputs, same outputs), MP launches recompute its last MP step without
// the VACF reduction.  The fastest block size per phase is kept, then the
// mode with the least LB + MP time at those sizes (the persistent launch
// runs both phases with the LB block size).  times[LBXB_AUTOTUNE_TIMES] (may be
// NULL): mean ms per step of LB-64, LB-128, LB-256, MP-64, MP-128, MP-256
// (per-step launches), LB-graph, MP-graph, LB-persistent, MP-persistent (0
// when a phase had no instance to time).
int lbxb_autotune(lbxb* b, int32_t launches, int32_t* lb_block, int32_t* mp_block, double* times) {
    if (!b || launches <= 0) return LBX_ERR_ARG;
    const uint32_t ni = b->cap_inst;
    constexpr int kSizes = 3;
    const int sizes[kSizes] = {64, 128, 256};
    double tm[2][kSizes] = {{0, 0, 0}, {0, 0, 0}};
    bool any[2] = {false, false};
    for (uint32_t id = 0; id < ni; ++id) {
        BInst& in = b->tab[id];
        const HostInst& h = b->hi[id];
        in.lb_on = in.mp_on = 0;
        if (h.phase == kLb && h.t >= 1) {
            in.lb_on = 1;
            in.lb_redo = 1;  // write_l2 = 0: recompute step t, slots and control untouched
            in.lb_t0 = h.t;
            in.lb_n = 1;
            in.check_after = INT64_MAX;
            const bool on = h.stage == 2 && h.t > h.switch_step;
            for (int a = 0; a < 3; ++a) {
                in.fm[a] = (on ? h.F[a] : 0.0) / 2.0;
                in.fc[a] = h.stage == 2 ? h.F[a] : 0.0;
            }
            any[0] = true;
        } else if (h.phase == kMp && h.mp_it >= 1) {
            in.mp_on = 1;
            in.mp_it0 = h.mp_it - 1;  // recompute step mp_it
            in.mp_n = 1;
            any[1] = true;
        }
    }
    LBX_TRY(hipMemcpyAsync(b->insts, b->tab.data(), ni * sizeof(BInst), hipMemcpyHostToDevice, b->stream));
    LBX_TRY(upload_active(b));
    auto launch = [&](int phase, int block) {
        if (phase == 0) launch_lb(b, 0, block); else launch_mp(b, 0, block, 0);
    };
    auto elapsed = [&](double* out) -> int {
        LBX_TRY(hipEventRecord(b->ev[1], b->stream));
        LBX_TRY(hipEventSynchronize(b->ev[1]));
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, b->ev[0], b->ev[1]);
        *out = ms / launches;
        return LBX_OK;
    };
    // per-step launches of every block size
    for (int s = 0; s < kSizes; ++s)
        for (int phase = 0; phase < 2; ++phase) {
            if (!any[phase]) continue;
            launch(phase, sizes[s]);  // warm
            LBX_TRY(hipEventRecord(b->ev[0], b->stream));
            for (int32_t k = 0; k < launches; ++k) launch(phase, sizes[s]);
            LBX_TRY(elapsed(&tm[phase][s]));
        }
    int best[2] = {b->lb_block, b->mp_block};
    for (int phase = 0; phase < 2; ++phase) {
        if (!any[phase]) continue;
        int k = 0;
        for (int s = 1; s < kSizes; ++s)
            if (tm[phase][s] < tm[phase][k]) k = s;
        best[phase] = sizes[k];
    }
    b->lb_block = best[0];
    b->mp_block = best[1];
    // the same launches at those sizes replayed from a captured graph
    double tg[2] = {0, 0};
    for (int phase = 0; phase < 2; ++phase) {
        if (!any[phase]) continue;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        LBX_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        for (int32_t k = 0; k < launches; ++k) launch(phase, best[phase]);
        const hipError_t ce = hipStreamEndCapture(b->stream, &graph);
        if (ce != hipSuccess) return hip_status(ce);
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) return hip_status(ie);
        int st = hip_status(hipGraphLaunch(exec, b->stream));  // warm
        if (!st) st = hip_status(hipEventRecord(b->ev[0], b->stream));
        if (!st) st = hip_status(hipGraphLaunch(exec, b->stream));
        if (!st) st = elapsed(&tg[phase]);
        (void)hipStreamSynchronize(b->stream);
        (void)hipGraphExecDestroy(exec);
        LBX_TRY(st);
    }
    // one cooperative launch of `launches` repeated steps per phase, with the
    // LB block size (kb_round runs both phases with one size)
    double tp[2] = {0, 0};
    for (uint32_t id = 0; id < ni; ++id) {
        BInst& in = b->tab[id];
        if (in.lb_on) in.lb_n = launches;  // lb_redo: every launch recomputes step t
        if (in.mp_on) in.mp_n = launches, in.mp_repeat = 1;
    }
    LBX_TRY(hipStreamSynchronize(b->stream));
    LBX_TRY(hipMemcpyAsync(b->insts, b->tab.data(), ni * sizeof(BInst), hipMemcpyHostToDevice, b->stream));
    for (int phase = 0; phase < 2; ++phase) {
        if (!any[phase]) continue;
        LBX_TRY(run_phase_persistent(b, phase, 1, 1));  // warm
        LBX_TRY(hipEventRecord(b->ev[0], b->stream));
        LBX_TRY(run_phase_persistent(b, phase, launches, 1));
        LBX_TRY(elapsed(&tp[phase]));
    }
    for (uint32_t id = 0; id < ni; ++id) {
        BInst& in = b->tab[id];
        in.lb_on = in.mp_on = in.lb_redo = in.mp_repeat = 0;
    }
    LBX_TRY(hipMemcpyAsync(b->insts, b->tab.data(), ni * sizeof(BInst), hipMemcpyHostToDevice, b->stream));
    // the launch mode with the least time over the timed phases
    if (any[0] || any[1]) {
        double t_stream = 0, t_graph = 0, t_pers = 0;
        for (int phase = 0; phase < 2; ++phase) {
            if (!any[phase]) continue;
            for (int s = 0; s < kSizes; ++s)
                if (sizes[s] == best[phase]) t_stream += tm[phase][s];
            t_graph += tg[phase];
            t_pers += tp[phase];
        }
        b->persistent = t_pers < t_graph && t_pers < t_stream;
        b->graphs = !b->persistent && t_graph <= t_stream;
    }
    clear_graphs(b);
    if (lb_block) *lb_block = b->lb_block;
    if (mp_block) *mp_block = b->mp_block;
    if (times) {
        for (int phase = 0; phase < 2; ++phase)
            for (int s = 0; s < kSizes; ++s) times[phase * kSizes + s] = tm[phase][s];
        times[6] = tg[0];
        times[7] = tg[1];
        times[8] = tp[0];
        times[9] = tp[1];
    }
    return hip_status(hipStreamSynchronize(b->stream));
}

int lbxb_stats(lbxb* b, double* lb_ms, int64_t* lb_box_updates, int64_t* lb_fluid_updates, double* mp_ms,
               int64_t* mp_box_updates, int64_t* mp_fluid_updates) {
    if (!b) return LBX_ERR_ARG;
    if (lb_ms) *lb_ms = b->lb_ms;
    if (lb_box_updates) *lb_box_updates = b->lb_box_updates;
    if (lb_fluid_updates) *lb_fluid_updates = b->lb_fluid_updates;
    if (mp_ms) *mp_ms = b->mp_ms;
    if (mp_box_updates) *mp_box_updates = b->mp_box_updates;
    if (mp_fluid_updates) *mp_fluid_updates = b->mp_fluid_updates;
    return LBX_OK;
}

int64_t lbxb_free_tiles(lbxb* b) {
    if (!b) return -1;
    return (int64_t)b->node_alloc.largest();
}

}  // extern "C"
