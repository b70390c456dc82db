// Device-side building blocks of liblbx (gfx950): D3Q19 arithmetic in the
// reference's evaluation order, neighbour addressing, wave/block reductions.
//
// Bit-for-bit parity with the reference relies on three things:
//  * the translation unit is built with -ffp-contract=off (no FMA fusion),
//    like flang's x86-64 code;
//  * every sum and product is written in the reference's left-to-right
//    order (Fortran a+b+c is (a+b)+c, a*b*c is (a*b)*c);
//  * terms whose velocity factor is 0 are dropped and factors of +-1 become
//    a sign.  x*(+-1) is exact, and adding x*0 = +-0 to a sum leaves its value
//    unchanged (only the sign of an exact zero could differ, which no later
//    operation here observes: divisions are by the density, comparisons are
//    n < 0 and maxima of absolute values).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "lbx_tables.h"

#pragma clang fp contract(off)

namespace lbx {

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

template <int S, typename V>
__device__ __forceinline__ V sgn(V x) {
    static_assert(S == 1 || S == -1, "sign factor");
    if constexpr (S > 0) return x; else return -x;
}

// (a*x + b*y) + c*z for a, b, c in {-1, 0, 1}, zero terms dropped.
template <int A, int B, int C, typename V>
__device__ __forceinline__ V dot3(V x, V y, V z) {
    if constexpr (A != 0) {
        double s = sgn<A>(x);
        if constexpr (B != 0) s = s + sgn<B>(y);
        if constexpr (C != 0) s = s + sgn<C>(z);
        return s;
    } else if constexpr (B != 0) {
        double s = sgn<B>(y);
        if constexpr (C != 0) s = s + sgn<C>(z);
        return s;
    } else if constexpr (C != 0) {
        return sgn<C>(z);
    } else {
        return V(0);
    }
}

// One population of the BGK + Guo-force collision, module_collision.f90:90-104:
//   neq = a0 rho + a1 (c.j) + a2 sum_ab j_a u_b (c_a c_b - cs^2 delta_ab)
//   n'  = (1-1/tau) n + (1/tau) neq
//         + (1-1/(2 tau)) [a1 ((c-u).F) + 2 a2 (c.u)(c.F)]
// A = 1-1/tau, B = 1/tau, C = 1-1/(2 tau) are computed on the host with the
// reference's expressions.
template <int L>
__device__ __forceinline__ double collide_pop(double f, double rho, double jx, double jy, double jz,
                                              double ux, double uy, double uz, double fx, double fy,
                                              double fz, double A, double B, double C) {
    constexpr int cx = kCX[L], cy = kCY[L], cz = kCZ[L];
    constexpr double w0 = a0(L), w1 = a1(L), w2 = a2(L);
    constexpr double Kx = (double)(cx * cx) - kCsq;
    constexpr double Ky = (double)(cy * cy) - kCsq;
    constexpr double Kz = (double)(cz * cz) - kCsq;
    double q = (jx * ux) * Kx;
    if constexpr (cx * cy != 0) q = q + sgn<cx * cy>(jx * uy);
    if constexpr (cx * cz != 0) q = q + sgn<cx * cz>(jx * uz);
    if constexpr (cy * cx != 0) q = q + sgn<cy * cx>(jy * ux);
    q = q + (jy * uy) * Ky;
    if constexpr (cy * cz != 0) q = q + sgn<cy * cz>(jy * uz);
    if constexpr (cz * cx != 0) q = q + sgn<cz * cx>(jz * ux);
    if constexpr (cz * cy != 0) q = q + sgn<cz * cy>(jz * uy);
    q = q + (jz * uz) * Kz;
    double neq = w0 * rho;
    if constexpr (L != 0) neq = neq + w1 * dot3<cx, cy, cz>(jx, jy, jz);
    neq = neq + w2 * q;
    double g = w1 * ((((double)cx - ux) * fx + ((double)cy - uy) * fy) + ((double)cz - uz) * fz);
    if constexpr (L != 0) g = g + ((2.0 * w2) * dot3<cx, cy, cz>(ux, uy, uz)) * dot3<cx, cy, cz>(fx, fy, fz);
    return (A * f + B * neq) + C * g;
}

// The same collision for the FP32-shift form, where a population is stored as
// its deviation g = n - a0 rho0 from the rest state of density rho0:
//   g' = n' - a0 rho0 = (1-1/tau) g + (1/tau) (neq - a0 rho0) + (1-1/(2 tau)) G
// (because (1-1/tau) + 1/tau = 1), with neq - a0 rho0 = a0 drho + a1 (c.j) +
// a2 sum_ab j_a u_b (c_a c_b - cs^2 delta_ab) and drho = rho - rho0 = sum g.
// V = float (the FP32-shift form) runs all of it in FP32: the deviations
// stay small, so their rounding error is relative to the deviation, not to
// rho0.  V = double (the FP32-store form) runs the same operations in FP64 on
// the FP32-stored deviations.  In FP32 the constants are rounded so that the
// collision still conserves momentum: a1 from a1_f32, and A, B from
// fp32_relaxation (A + B = 1).  In FP64 each sum of a product and a term is
// one fused multiply-add (madd): fewer operations and registers, so the
// FP32-store step runs with five waves per SIMD; the restatement of the form
// fuses the same operations (C fma).
template <typename V>
__device__ __forceinline__ V madd(V a, V b, V c) {
    if constexpr (std::is_same<V, double>::value)
        return __builtin_fma(a, b, c);
    else
        return c + a * b;  // FP32: unfused, the form's own order
}

template <int L, typename V = float>
__device__ __forceinline__ V collide_shift(V g, V drho, V jx, V jy, V jz, V ux, V uy, V uz, V fx, V fy, V fz, V A,
                                           V B, V C) {
    constexpr int cx = kCX[L], cy = kCY[L], cz = kCZ[L];
    // FP32: a1 with an exact momentum sum (a1_f32); FP64: a1 as is
    constexpr V w0 = (V)a0(L), w1 = std::is_same<V, float>::value ? (V)a1_f32(L) : (V)a1(L), w2 = (V)a2(L);
    constexpr V Kx = (V)((double)(cx * cx) - kCsq);
    constexpr V Ky = (V)((double)(cy * cy) - kCsq);
    constexpr V Kz = (V)((double)(cz * cz) - kCsq);
    V q = (jx * ux) * Kx;
    if constexpr (cx * cy != 0) q = q + sgn<cx * cy>(jx * uy);
    if constexpr (cx * cz != 0) q = q + sgn<cx * cz>(jx * uz);
    if constexpr (cy * cx != 0) q = q + sgn<cy * cx>(jy * ux);
    q = madd<V>(jy * uy, Ky, q);
    if constexpr (cy * cz != 0) q = q + sgn<cy * cz>(jy * uz);
    if constexpr (cz * cx != 0) q = q + sgn<cz * cx>(jz * ux);
    if constexpr (cz * cy != 0) q = q + sgn<cz * cy>(jz * uy);
    q = madd<V>(jz * uz, Kz, q);
    V dneq = w0 * drho;
    if constexpr (L != 0) dneq = madd<V>(w1, dot3<cx, cy, cz>(jx, jy, jz), dneq);
    dneq = madd<V>(w2, q, dneq);
    V G = w1 * madd<V>((V)cz - uz, fz, madd<V>((V)cy - uy, fy, ((V)cx - ux) * fx));
    if constexpr (L != 0)
        G = madd<V>(((V)2 * w2) * dot3<cx, cy, cz>(ux, uy, uz), dot3<cx, cy, cz>(fx, fy, fz), G);
    return madd<V>(C, G, madd<V>(B, dneq, A * g));
}

// u = j / rho of the FP32 storage forms' collision: FP32-shift divides each
// component; FP32-store (R = double) multiplies by one reciprocal (the form's
// restatement does the same).
template <typename R>
__device__ __forceinline__ void shift_velocity(R jx, R jy, R jz, R full, R& ux, R& uy, R& uz) {
    if constexpr (std::is_same<R, double>::value) {
        const R inv = R(1) / full;
        ux = jx * inv;
        uy = jy * inv;
        uz = jz * inv;
    } else {
        ux = jx / full;
        uy = jy / full;
        uz = jz / full;
    }
}

// Division by a runtime-invariant divisor 1 <= d < 2^31 as a multiply-high
// and a shift: with s = ceil(log2 d) and m = ceil(2^(31+s) / d) (< 2^32),
// floor(n/d) = (n*m) >> (31+s) for every n < 2^31, because the rounding
// error of m times n stays below 2^(31+s).
struct FastDiv {
    uint32_t d, mul, shift;
    void set(uint32_t div) {
        d = div;
        uint32_t s = 0;
        while ((1ull << s) < div) ++s;
        shift = s;
        mul = (uint32_t)(((1ull << (31 + s)) + div - 1) / div);
    }
    __device__ __forceinline__ uint32_t div(uint32_t n) const {
        return d == 1 ? n : (__umulhi(n, mul) >> (shift - 1));
    }
};

// Load / store through a uniform base pointer and a 32-bit element index:
// the byte offset idx*8 fits in 32 bits (indices < 2^29), which lets the
// compiler use the SGPR-base + 32-bit VGPR-offset addressing form instead of
// a 64-bit address per lane.
template <typename T>
__device__ __forceinline__ T ld(const T* base, uint32_t idx) {
    return *(const T*)((const char*)base + (uint32_t)(idx * (uint32_t)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void st(T* base, uint32_t idx, T v) {
    *(T*)((char*)base + (uint32_t)(idx * (uint32_t)sizeof(T))) = v;
}

// One array (a population plane, a flux component) as a raw buffer resource:
// loads and stores take a 32-bit byte offset in a VGPR against a descriptor in
// SGPRs, so a lane needs one address register instead of a 64-bit pointer, and
// an access outside [0, bytes) reads 0 / writes nothing instead of faulting.
// 0x00020000 is the gfx9-family descriptor word 3 for raw 32-bit-dword data.
// E is the element type: double (b64 accesses) or float (b32).
constexpr int kBufWord3 = 0x00020000;
// the nt bit of a gfx940-family memory instruction's cache-policy field
constexpr int kNonTemporal = 2;
template <typename E>
using Vec2 = E __attribute__((ext_vector_type(2)));
template <typename E>
struct Plane {
    static_assert(std::is_same<E, double>::value || std::is_same<E, float>::value, "double or float elements");
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ Plane(const void* p, uint64_t bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0,
                                              (int)(uint32_t)(bytes < 0xFFFFFFFFull ? bytes : 0xFFFFFFFFull),
                                              kBufWord3)) {}
    __device__ __forceinline__ E ld(uint32_t idx) const {
        if constexpr (sizeof(E) == 8)
            return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, idx * 8u, 0, 0));
        else
            return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, idx * 4u, 0, 0));
    }
    // CP is the store's cache-policy field: 0 the default, kNonTemporal the
    // streaming hint (the line is not kept for reuse in the caches)
    template <int CP = 0>
    __device__ __forceinline__ void st(uint32_t idx, E v) const {
        if constexpr (sizeof(E) == 8)
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v),
                                                  r, idx * 8u, 0, CP);
        else
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, idx * 4u, 0, CP);
    }
    // elements 2*pidx and 2*pidx+1 in one access (b128 for double, b64 for float)
    __device__ __forceinline__ Vec2<E> ld2(uint32_t pidx) const {
        if constexpr (sizeof(E) == 8)
            return __builtin_bit_cast(Vec2<E>, __builtin_amdgcn_raw_buffer_load_b128(r, pidx * 16u, 0, 0));
        else
            return __builtin_bit_cast(Vec2<E>, __builtin_amdgcn_raw_buffer_load_b64(r, pidx * 8u, 0, 0));
    }
    __device__ __forceinline__ void st2(uint32_t pidx, Vec2<E> v) const {
        if constexpr (sizeof(E) == 8)
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, v),
                                                   r, pidx * 16u, 0, 0);
        else
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v),
                                                  r, pidx * 8u, 0, 0);
    }
};

// A set of planes of E addressed by a 32-bit element index.  Narrow: a
// buffer resource over the whole set (needs index*sizeof(E) < 2^32); Wide:
// plain 64-bit pointers (any size).
template <bool Wide, typename E = double>
struct View;
template <typename E>
struct View<false, E> {
    Plane<E> p;
    __device__ __forceinline__ View(const E* base, uint64_t n) : p(base, n * sizeof(E)) {}
    __device__ __forceinline__ E ld(uint32_t idx) const { return p.ld(idx); }
    __device__ __forceinline__ void st(uint32_t idx, E v) const { p.st(idx, v); }
    __device__ __forceinline__ void stnt(uint32_t idx, E v) const { p.template st<kNonTemporal>(idx, v); }
    __device__ __forceinline__ Vec2<E> ld2(uint32_t pidx) const { return p.ld2(pidx); }
    __device__ __forceinline__ void st2(uint32_t pidx, Vec2<E> v) const { p.st2(pidx, v); }
};
template <typename E>
struct View<true, E> {
    E* b;
    __device__ __forceinline__ View(const E* base, uint64_t) : b(const_cast<E*>(base)) {}
    __device__ __forceinline__ E ld(uint32_t idx) const { return b[(uint64_t)idx]; }
    __device__ __forceinline__ void st(uint32_t idx, E v) const { b[(uint64_t)idx] = v; }
    __device__ __forceinline__ void stnt(uint32_t idx, E v) const { __builtin_nontemporal_store(v, b + (uint64_t)idx); }
    __device__ __forceinline__ Vec2<E> ld2(uint32_t pidx) const { return ((const Vec2<E>*)b)[(uint64_t)pidx]; }
    __device__ __forceinline__ void st2(uint32_t pidx, Vec2<E> v) const { ((Vec2<E>*)b)[(uint64_t)pidx] = v; }
};

// K per-node values of Np nodes (a moment field, the scattering
// probabilities) laid out so that one 16-byte (double) or 8-byte (float)
// access reads two of them: values 2p and 2p+1 (p < K/2) side by side in pair
// plane p, an odd last value alone in a plane after the pairs.  A node's K
// values then take ceil(K/2) loads instead of K.
__host__ __device__ constexpr uint64_t comp_index(int K, int k, uint64_t i, uint64_t Np) {
    return k < (K & ~1) ? ((uint64_t)(k >> 1) * Np + i) * 2 + (uint64_t)(k & 1) : (uint64_t)(K & ~1) * Np + i;
}

template <int K, typename V, typename E>
__device__ __forceinline__ void load_comps(const V& v, uint32_t Np, uint32_t i, E (&out)[K]) {
    static_for<0, K / 2>([&](auto P) {
        const Vec2<E> x = v.ld2((uint32_t)P * Np + i);
        out[2 * P] = x[0];
        out[2 * P + 1] = x[1];
    });
    if constexpr (K & 1) out[K - 1] = v.ld((uint32_t)(K - 1) * Np + i);
}

template <int K, typename V, typename E>
__device__ __forceinline__ void store_comps(const V& v, uint32_t Np, uint32_t i, const E (&in)[K]) {
    static_for<0, K / 2>([&](auto P) {
        Vec2<E> x;
        x[0] = in[2 * P];
        x[1] = in[2 * P + 1];
        v.st2((uint32_t)P * Np + i, x);
    });
    if constexpr (K & 1) v.st((uint32_t)(K - 1) * Np + i, in[K - 1]);
}

// Periodic neighbour offsets of a node: X[1+d] is the x index of x+d, and Y,
// Z the row / plane offsets of y+d, z+d (module_system.f90:99-112: only a
// +-1 step wraps; an axis of length 1 wraps onto itself).
struct Nbr {
    uint32_t X[3], Y[3], Z[3];
    __device__ __forceinline__ void init(uint32_t r, const FastDiv& dlx, const FastDiv& dly, uint32_t lz) {
        const uint32_t lx = dlx.d, ly = dly.d;
        const uint32_t q = dlx.div(r);
        const uint32_t x = r - q * lx;
        const uint32_t z = dly.div(q);
        const uint32_t y = q - z * ly;
        const uint32_t lxy = lx * ly;
        X[0] = x == 0 ? lx - 1 : x - 1;
        X[1] = x;
        X[2] = x + 1 == lx ? 0 : x + 1;
        Y[0] = (y == 0 ? ly - 1 : y - 1) * lx;
        Y[1] = y * lx;
        Y[2] = (y + 1 == ly ? 0 : y + 1) * lx;
        Z[0] = (z == 0 ? lz - 1 : z - 1) * lxy;
        Z[1] = z * lxy;
        Z[2] = (z + 1 == lz ? 0 : z + 1) * lxy;
    }
    // index of r + d
    template <int dx, int dy, int dz>
    __device__ __forceinline__ uint32_t at() const {
        return X[1 + dx] + Y[1 + dy] + Z[1 + dz];
    }
};

// Workgroups are dispatched round-robin over the 8 XCDs; give consecutive
// logical tiles to the same XCD so that neighbouring tiles share its L2.
__device__ __forceinline__ uint32_t xcd_tile() {
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t x = b & 7u, i = b >> 3, per = nb >> 3, rem = nb & 7u;
    return x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}


// v of the lane that DPP control CTRL selects within the lane's row of 16
// (all 64 lanes active)
template <int CTRL>
__device__ __forceinline__ double dpp_row(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xf, 0xf, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ double read_lane(double v, int lane) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over the 64 lanes of a wave in a fixed order, returned to every lane
// (all 64 lanes active).  Within each row of 16 lanes, register moves only:
// lane pairs (quad_perm 1,0,3,2), pairs of pairs (quad_perm 2,3,0,1), half
// rows (row_half_mirror) and rows (row_mirror); each step adds two partial
// sums in either order, so every lane of a row ends with the same bits.  The
// four row sums are then read from lanes 0, 16, 32, 48 and added as
// (r0 + r1) + (r2 + r3).
__device__ __forceinline__ double wave_sum(double v) {
    v = v + dpp_row<0xB1>(v);
    v = v + dpp_row<0x4E>(v);
    v = v + dpp_row<0x141>(v);
    v = v + dpp_row<0x140>(v);
    return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// wave_sum of NV values at once, level by level, so that the NV independent
// chains fill each other's DPP wait states; the same additions per value.
template <int NV>
__device__ __forceinline__ void wave_sums(double (&v)[NV]) {
#pragma unroll
    for (int a = 0; a < NV; ++a) v[a] = v[a] + dpp_row<0xB1>(v[a]);
#pragma unroll
    for (int a = 0; a < NV; ++a) v[a] = v[a] + dpp_row<0x4E>(v[a]);
#pragma unroll
    for (int a = 0; a < NV; ++a) v[a] = v[a] + dpp_row<0x141>(v[a]);
#pragma unroll
    for (int a = 0; a < NV; ++a) v[a] = v[a] + dpp_row<0x140>(v[a]);
#pragma unroll
    for (int a = 0; a < NV; ++a)
        v[a] = (read_lane(v[a], 0) + read_lane(v[a], 16)) + (read_lane(v[a], 32) + read_lane(v[a], 48));
}

// Maximum over the 64 lanes of a wave (fmax: a NaN lane is passed over),
// returned to every lane (all 64 lanes active); the same row steps as wave_sum.
__device__ __forceinline__ double wave_max(double v) {
    v = fmax(v, dpp_row<0xB1>(v));
    v = fmax(v, dpp_row<0x4E>(v));
    v = fmax(v, dpp_row<0x141>(v));
    v = fmax(v, dpp_row<0x140>(v));
    return fmax(fmax(read_lane(v, 0), read_lane(v, 16)), fmax(read_lane(v, 32), read_lane(v, 48)));
}

// Lane k of the wave receives lane k+1's value (S = +1, DPP wave_shl:1) or
// lane k-1's (S = -1, wave_shr:1); lane 63 (S = +1) and lane 0 (S = -1), and
// a lane whose source lane is inactive, receive 0.  Register moves only.
template <int S>
__device__ __forceinline__ double lane_shift(double v) {
    static_assert(S == 1 || S == -1, "one lane up or down");
    constexpr int ctrl = S > 0 ? 0x130 : 0x138;
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, ctrl, 0xf, 0xf, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), ctrl, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Deterministic block sum of 3 values; the result is valid in thread 0.
template <int BLOCK>
__device__ __forceinline__ void block_sum3(double v[3], double* lds /* [BLOCK/64*3] */) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int a = 0; a < 3; ++a) {
        double s = wave_sum(v[a]);
        if (lane == 0) lds[w * 3 + a] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int a = 0; a < 3; ++a) {
            double s = lds[a];
            for (int k = 1; k < BLOCK / 64; ++k) s = s + lds[k * 3 + a];
            v[a] = s;
        }
    }
    __syncthreads();
}

}  // namespace lbx
