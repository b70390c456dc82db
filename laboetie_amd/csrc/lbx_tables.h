// D3Q19 tables shared by the host and device code of liblbx.
//
// Velocity order and weights as module_lbmodel.f90 builds them
// (init_velocities :66-86, init_weight_factors :122-152,
// determine_velocity_inverse :156-162); index l = 0..18 here stands for the
// reference's l = 1..19.  The weight expressions are the reference's, folded
// at compile time in IEEE double exactly as flang folds its parameters.
#pragma once

#if defined(__HIPCC__)
#define LBX_HD __host__ __device__
#else
#define LBX_HD
#endif

namespace lbx {

constexpr int kQ = 19;
constexpr int kCX[kQ] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int kCY[kQ] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int kCZ[kQ] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
constexpr int kInv[kQ] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};

constexpr double kCsq = 1.0 / 3.0;   // module_lbmodel.f90:67,127
constexpr double kW0 = 1.0 / 3.0, kW1 = 1.0 / 18.0, kW2 = 1.0 / 36.0;

constexpr LBX_HD int wclass(int l) { return l == 0 ? 0 : (l <= 6 ? 1 : 2); }
constexpr LBX_HD double a0(int l) { return wclass(l) == 0 ? kW0 : (wclass(l) == 1 ? kW1 : kW2); }
constexpr LBX_HD double a1(int l) { return a0(l) / kCsq; }
constexpr LBX_HD double a2(int l) { return a0(l) / (2 * (kCsq * kCsq)); }

// a1 for a collision in FP32 arithmetic.  Rounded one by one, fl(1/6) and
// fl(1/12) give sum_l a1_l c_lx^2 = 2 fl(1/6) + 8 fl(1/12) = 1 + 3.0e-8, so
// every step adds 3e-8 j of momentum: a spurious force of 3e-8 j per step,
// which relative to a body force F moves the steady flux by about
// 3e-8 j / F (1e-3 at j / F = 3e4).  Here the edges take fl(1/12) and the
// faces 1/2 - 4 fl(1/12), which FP32 holds exactly, so the sum is exactly 1.
constexpr LBX_HD float a1_f32(int l) {
    return wclass(l) == 2 ? (float)(kW2 / kCsq) : (wclass(l) == 1 ? 0.5f - 4.0f * (float)(kW2 / kCsq) : (float)a1(l));
}

// kBT = 1/3 (module_system.f90:94); lambda = 4 D / kBT (module_moment_propagation.f90:336)
constexpr double kKbt = 1.0 / 3.0;

// mask bits of the per-node neighbour word
constexpr unsigned kSolidBit = 1u << 31;        // the node itself is solid
constexpr unsigned kInterfacialBit = 1u << 30;  // node%isInterfacial
// bit l (1..18): the pull source r - c_l is solid

}  // namespace lbx
