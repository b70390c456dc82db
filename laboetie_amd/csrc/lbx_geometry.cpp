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
        return LBX_ERR_ARG;  // supercell_definition.f90:156-157
    }
    for (int64_t p = 0; p < N; ++p)
        if (nature[p] == 0) return LBX_OK;
    return LBX_ERR_ARG;  // "All nodes are solid", supercell_definition.f90:167-169
}

extern "C" int lbx_detect_interfacial(int32_t lx, int32_t ly, int32_t lz, const int8_t* nature,
                                      int8_t* interfacial) {
    using namespace lbx;
    if (lx <= 0 || ly <= 0 || lz <= 0 || !nature || !interfacial) return LBX_ERR_ARG;
    for (int32_t k = 0; k < lz; ++k)
        for (int32_t j = 0; j < ly; ++j)
            for (int32_t i = 0; i < lx; ++i) {
                const int64_t r = idx(i, j, k, lx, ly);
                int8_t f = 0;
                for (int l = 1; l < 19 && !f; ++l) {
                    const int64_t rp = idx(wrap(i + kCX[l], lx), wrap(j + kCY[l], ly), wrap(k + kCZ[l], lz), lx, ly);
                    if (nature[rp] != nature[r]) f = 1;
                }
                interfacial[r] = f;
            }
    return LBX_OK;
}
