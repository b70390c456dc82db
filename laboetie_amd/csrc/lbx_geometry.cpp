// Host-side geometry builders of the lbx C-ABI (lbx_build_geometry,
// lbx_detect_interfacial).  One-off, init-time work: plain C++.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/lbx.h"
#include "lbx_tables.h"

namespace lbx {

// The Fortran NORM2 intrinsic as the flang runtime evaluates it (a scaled sum
// of squares, max*sqrt(1 + sum((x/max)^2)) with a running rescale whenever a
// larger element arrives).  Identified by black-box comparison with flang's
// NORM2 over 1.08e9 geometry inputs; a plain sqrt(sum x^2) differs in the last
// bit often enough to flip solid/fluid decisions of construct_cc.
double fortran_norm2(const double* v, int n) {
    double mx = 0.0, s = 0.0;
    for (int i = 0; i < n; ++i) {
        double a = std::fabs(v[i]);
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
    return mx * std::sqrt(1.0 + s);
}

static inline int64_t idx(int32_t i, int32_t j, int32_t k, int32_t lx, int32_t ly) {
    return (int64_t)i + (int64_t)lx * ((int64_t)j + (int64_t)ly * k);
}

// 0-based periodic step (module_system.f90:99-112: only a +-1 step wraps)
static inline int32_t wrap(int32_t i, int32_t n) { return i < 0 ? n - 1 : (i >= n ? 0 : i); }

}  // namespace lbx

extern "C" int lbx_build_geometry(int32_t label, int32_t lx, int32_t ly, int32_t lz,
                                  const int32_t* solid_ijk, int64_t npts, int8_t* nature) {
    using namespace lbx;
    if (lx <= 0 || ly <= 0 || lz <= 0 || !nature) return LBX_ERR_ARG;
    const int64_t N = (int64_t)lx * ly * lz;
    std::memset(nature, 0, (size_t)N);
    switch (label) {
    case -1:  // no solid node
        break;
    case 0:   // construct_custom, module_geometry.f90:380-426
        if (npts > 0 && !solid_ijk) return LBX_ERR_ARG;
        for (int64_t p = 0; p < npts; ++p) {
            int32_t i = solid_ijk[3 * p], j = solid_ijk[3 * p + 1], k = solid_ijk[3 * p + 2];
            if (i <= 0 || j <= 0 || k <= 0 || i > lx || j > ly || k > lz) return LBX_ERR_ARG;
            nature[idx(i - 1, j - 1, k - 1, lx, ly)] = 1;
        }
        break;
    case 1:   // construct_slit, :158-166: the planes k = 1 and k = lz
        for (int32_t j = 0; j < ly; ++j)
            for (int32_t i = 0; i < lx; ++i) {
                nature[idx(i, j, 0, lx, ly)] = 1;
                nature[idx(i, j, lz - 1, lx, ly)] = 1;
            }
        break;
    case 2: {  // construct_cylinder, :253-277: solid where |r - centre| >= (lx-1)/2
        if (lx != ly || lx < 3) return LBX_ERR_ARG;
        const double ox = (double)(lx + 1) / 2.0, oy = (double)(ly + 1) / 2.0;
        const double radius = (double)(lx - 1) / 2.0;
        for (int32_t i = 1; i <= lx; ++i)
            for (int32_t j = 1; j <= ly; ++j) {
                const double r[2] = {(double)i - ox, (double)j - oy};
                const int8_t s = fortran_norm2(r, 2) >= radius ? 1 : 0;
                for (int32_t k = 0; k < lz; ++k) nature[idx(i - 1, j - 1, k, lx, ly)] = s;
            }
        break;
    }
    case 3: {  // construct_cc, :206-245: spheres at the 8 corners and the centre
        if (lx != ly || lx != lz) return LBX_ERR_ARG;
        const double rad = (double)(lx - 1) * std::sqrt(3.0) / 4.0;
        const double c[9][3] = {{1, 1, 1},
                                {(double)lx, 1, 1},
                                {1, (double)ly, 1},
                                {1, 1, (double)lz},
                                {(double)lx, (double)ly, 1},
                                {(double)lx, 1, (double)lz},
                                {1, (double)ly, (double)lz},
                                {(double)lx, (double)ly, (double)lz},
                                {(double)(lx + 1) / 2.0, (double)(ly + 1) / 2.0, (double)(lz + 1) / 2.0}};
        for (int32_t k = 1; k <= lz; ++k)
            for (int32_t j = 1; j <= ly; ++j)
                for (int32_t i = 1; i <= lx; ++i) {
                    int8_t s = 0;
                    for (int q = 0; q < 9 && !s; ++q) {
                        const double r[3] = {(double)i - c[q][0], (double)j - c[q][1], (double)k - c[q][2]};
                        if (fortran_norm2(r, 3) <= rad) s = 1;
                    }
                    nature[idx(i - 1, j - 1, k - 1, lx, ly)] = s;
                }
        break;
    }
    default:
        return LBX_ERR_ARG;  // supercell_definition.f90:77-78
    }
    for (int64_t p = 0; p < N; ++p)
        if (nature[p] == 0) return LBX_OK;
    return LBX_ERR_ARG;  // "All nodes are solid", supercell_definition.f90:88-90
}

extern "C" int lbx_detect_interfacial(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature,
                                      int8_t* interfacial) {
    using namespace lbx;
    if (lx <= 0 || ly <= 0 || lz <= 0 || !nature || !interfacial) return LBX_ERR_ARG;
    // row by row: the 9 rows (y+dy, z+dz) a node's neighbours lie in, wrapped
    // once per row, and the x-neighbours wrapped per node (an axis of length 1
    // wraps onto itself); a node is interfacial if any of the 18 differs
    for (int32_t k = 0; k < lz; ++k)
        for (int32_t j = 0; j < ly; ++j) {
            const int8_t* rows[3][3];
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz)
                    rows[dy + 1][dz + 1] = nature + idx(0, wrap(j + dy, ly), wrap(k + dz, lz), lx, ly);
            const int8_t* c = rows[1][1];
            int8_t* out = interfacial + idx(0, j, k, lx, ly);
            for (int32_t i = 0; i < lx; ++i) {
                const int32_t xs[3] = {wrap(i - 1, lx), i, wrap(i + 1, lx)};
                const int8_t v = c[i];
                int f = 0;
                for (int l = 1; l < 19; ++l) f |= rows[kCY[l] + 1][kCZ[l] + 1][xs[kCX[l] + 1]] != v;
                out[i] = (int8_t)f;
            }
        }
    return LBX_OK;
}

// ------------------------------------------------------------ generators
// Seeded synthetic media for the workloads the reference runs on files it
// does not generate itself.  Both draw from SplitMix64 (Steele, Lea, Flood,
// "Fast splittable pseudorandom number generators", OOPSLA 2014) so that a
// seed gives the same medium on every machine with an IEEE libm.
namespace lbx {

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform in [0, 1) with 53 random bits
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

}  // namespace lbx

#include <algorithm>
#include <numeric>
#include <thread>
#include <vector>

namespace {

// out[i] = sum_{d=-R..R} w[d+R] * in[(i+d) mod n] along one axis of a
// (lx,ly,lz) x-fastest field, d in increasing order; planes split over threads.
void conv_axis(const std::vector<double>& in, std::vector<double>& out, int32_t lx, int32_t ly, int32_t lz,
               int axis, const std::vector<double>& w, int R) {
    const int32_t n[3] = {lx, ly, lz};
    const int64_t stride[3] = {1, (int64_t)lx, (int64_t)lx * ly};
    const int32_t len = n[axis];
    auto work = [&](int32_t k0, int32_t k1) {
        for (int32_t k = k0; k < k1; ++k)
            for (int32_t j = 0; j < ly; ++j)
                for (int32_t i = 0; i < lx; ++i) {
                    const int32_t c[3] = {i, j, k};
                    const int64_t base = (int64_t)i + (int64_t)lx * ((int64_t)j + (int64_t)ly * k) -
                                         (int64_t)c[axis] * stride[axis];
                    double s = 0.0;
                    for (int d = -R; d <= R; ++d) {
                        int32_t p = (c[axis] + d) % len;
                        if (p < 0) p += len;
                        s = s + w[d + R] * in[base + (int64_t)p * stride[axis]];
                    }
                    out[base + (int64_t)c[axis] * stride[axis]] = s;
                }
    };
    const int32_t nt = (int32_t)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    const int32_t per = (lz + nt - 1) / nt;
    for (int32_t t = 0; t < nt; ++t) {
        const int32_t k0 = t * per, k1 = std::min(lz, k0 + per);
        if (k0 < k1) pool.emplace_back(work, k0, k1);
    }
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" {

int lbx_gaussian_field(int32_t lx, int32_t ly, int32_t lz, double sigma, uint64_t seed, double* field) {
    if (lx <= 0 || ly <= 0 || lz <= 0 || !field || !(sigma > 0.0)) return LBX_ERR_ARG;
    const int64_t N = (int64_t)lx * ly * lz;
    // i.i.d. N(0,1) per node, x fastest, by Box-Muller pairs
    lbx::SplitMix64 rng{seed};
    std::vector<double> a((size_t)N), b((size_t)N);
    for (int64_t r = 0; r < N; r += 2) {
        const double u1 = 1.0 - rng.uniform();  // (0, 1]
        const double u2 = rng.uniform();
        const double m = std::sqrt(-2.0 * std::log(u1));
        a[(size_t)r] = m * std::cos(2.0 * M_PI * u2);
        if (r + 1 < N) a[(size_t)r + 1] = m * std::sin(2.0 * M_PI * u2);
    }
    // separable periodic convolution with exp(-d^2 / (2 sigma^2)), |d| <= R,
    // R = min(ceil(4 sigma), (n-1)/2) per axis so that no image is counted twice
    const int32_t n[3] = {lx, ly, lz};
    for (int axis = 0; axis < 3; ++axis) {
        const int R = (int)std::min<double>(std::ceil(4.0 * sigma), (double)((n[axis] - 1) / 2));
        std::vector<double> w((size_t)(2 * R + 1));
        for (int d = -R; d <= R; ++d) w[(size_t)(d + R)] = std::exp(-(double)d * d / (2.0 * sigma * sigma));
        conv_axis(a, b, lx, ly, lz, axis, w, R);
        a.swap(b);
    }
    std::memcpy(field, a.data(), (size_t)N * sizeof(double));
    return LBX_OK;
}

int lbx_geometry_random(int32_t lx, int32_t ly, int32_t lz, double sigma, double porosity, uint64_t seed,
                        int8_t* nature) {
    if (!nature || !(porosity > 0.0) || !(porosity <= 1.0)) return LBX_ERR_ARG;
    const int64_t N = (int64_t)lx * ly * lz;
    if (lx <= 0 || ly <= 0 || lz <= 0) return LBX_ERR_ARG;
    std::vector<double> f((size_t)N);
    int st = lbx_gaussian_field(lx, ly, lz, sigma, seed, f.data());
    if (st) return st;
    // the round(porosity N) lowest values are fluid (ties: lower index first);
    // every solid cluster is kept
    const int64_t nfluid = std::llround(porosity * (double)N);
    if (nfluid <= 0) return LBX_ERR_ARG;
    // (value, index) pairs in lexicographic order: the nfluid-th smallest is the cut
    std::vector<std::pair<double, int64_t>> v((size_t)N);
    for (int64_t r = 0; r < N; ++r) v[(size_t)r] = {f[(size_t)r], r};
    std::nth_element(v.begin(), v.begin() + (nfluid - 1), v.end());
    const std::pair<double, int64_t> cut = v[(size_t)(nfluid - 1)];
    for (int64_t r = 0; r < N; ++r) nature[r] = std::make_pair(f[(size_t)r], r) <= cut ? 0 : 1;
    return LBX_OK;
}

int lbx_geometry_disks(int32_t lx, int32_t ly, int32_t lz, int32_t ndisks, double radius, uint64_t seed,
                       int8_t* nature, double* centres) {
    if (lx <= 0 || ly <= 0 || lz <= 0 || !nature || ndisks < 0 || !(radius > 0.0)) return LBX_ERR_ARG;
    const int32_t n[3] = {lx, ly, lz};
    lbx::SplitMix64 rng{seed};
    std::vector<double> c((size_t)3 * ndisks);
    // per-axis periodic minimum-image distance, tools/chromato_generator.f90:53-61
    auto axis_d = [&](double p, double q, int a) {
        double d = std::fabs(p - q);
        return std::min(d, (double)n[a] - d);
    };
    const double dia = 2.0 * radius;
    for (int32_t k = 0; k < ndisks; ++k) {
        for (int tries = 0;; ++tries) {
            if (tries > 100000) return LBX_ERR_ARG;  // no room for another disk
            // centre uniform in [1, n+1) per axis (a flat axis of length 1 at 1),
            // chromato_generator.f90:20-24
            for (int a = 0; a < 3; ++a) c[(size_t)(3 * k + a)] = n[a] == 1 ? 1.0 : rng.uniform() * n[a] + 1.0;
            bool ok = true;
            for (int32_t m = 0; m < k && ok; ++m) {
                const double dx = axis_d(c[(size_t)3 * k], c[(size_t)3 * m], 0);
                const double dy = axis_d(c[(size_t)3 * k + 1], c[(size_t)3 * m + 1], 1);
                const double dz = axis_d(c[(size_t)3 * k + 2], c[(size_t)3 * m + 2], 2);
                ok = std::sqrt(dx * dx + dy * dy + dz * dz) > dia;  // disks neither overlap nor touch
            }
            if (ok) break;
        }
    }
    // solid where the distance to a centre is <= radius, chromato_generator.f90:40-64
    // (1-based node coordinates)
    for (int32_t k = 0; k < lz; ++k)
        for (int32_t j = 0; j < ly; ++j)
            for (int32_t i = 0; i < lx; ++i) {
                int8_t v = 0;
                for (int32_t m = 0; m < ndisks && !v; ++m) {
                    const double dx = axis_d((double)(i + 1), c[(size_t)3 * m], 0);
                    const double dy = axis_d((double)(j + 1), c[(size_t)3 * m + 1], 1);
                    const double dz = axis_d((double)(k + 1), c[(size_t)3 * m + 2], 2);
                    v = std::sqrt(dx * dx + dy * dy + dz * dz) <= radius;
                }
                nature[(int64_t)i + (int64_t)lx * ((int64_t)j + (int64_t)ly * k)] = v;
            }
    if (centres) std::memcpy(centres, c.data(), c.size() * sizeof(double));
    return LBX_OK;
}

}  // extern "C"
