"""Test infrastructure, not product code: numpy restatement of the library's
seeded medium generators (lbx_gaussian_field, lbx_geometry_random,
lbx_geometry_disks).  Only tests/ import it, as the checker.

These generators have no counterpart in the reference, which reads such media
from geom.in files; the disk inclusion rule is the reference's
(tools/chromato_generator.f90:40-64, per-axis periodic minimum image,
d <= radius).  Pinned by: exact fluid counts, the disk rule re-derived node by
node, the separable filter's variance and correlation closed forms
(tests/test_geometry_gen.py).
"""
import math

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64_uniform(seed, count, start=0):
    """The first `count` uniforms in [0,1) of SplitMix64(seed) after `start` draws."""
    with np.errstate(over="ignore"):
        k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def gaussian_field(lx, ly, lz, sigma, seed):
    N = lx * ly * lz
    npairs = (N + 1) // 2
    u = splitmix64_uniform(seed, 2 * npairs)
    u1 = 1.0 - u[0::2]
    u2 = u[1::2]
    m = np.sqrt(-2.0 * np.log(u1))
    a = np.empty(2 * npairs)
    a[0::2] = m * np.cos(2.0 * np.pi * u2)
    a[1::2] = m * np.sin(2.0 * np.pi * u2)
    f = a[:N].reshape(lz, ly, lx)  # [k][j][i], x fastest
    for axis, n in ((2, lx), (1, ly), (0, lz)):
        R = int(min(math.ceil(4.0 * sigma), (n - 1) // 2))
        s = np.zeros_like(f)
        for d in range(-R, R + 1):
            w = math.exp(-float(d) * d / (2.0 * sigma * sigma))
            s = s + w * np.roll(f, -d, axis=axis)
        f = s
    return f.reshape(-1)


def geometry_random(lx, ly, lz, sigma, porosity, seed):
    f = gaussian_field(lx, ly, lz, sigma, seed)
    N = f.size
    nfluid = int(round(porosity * N))
    order = np.lexsort((np.arange(N), f))
    nat = np.ones(N, np.int8)
    nat[order[:nfluid]] = 0
    return nat


def geometry_disks(lx, ly, lz, ndisks, radius, seed):
    n = (lx, ly, lz)
    draws = 0
    cs = []

    def axis_d(p, q, a):
        d = abs(p - q)
        return min(d, n[a] - d)

    for _ in range(ndisks):
        while True:
            c = []
            for a in range(3):
                if n[a] == 1:
                    c.append(1.0)
                else:
                    c.append(float(splitmix64_uniform(seed, 1, draws)[0]) * n[a] + 1.0)
                    draws += 1
            if all(math.sqrt(sum(axis_d(c[a], m[a], a) ** 2 for a in range(3))) > 2.0 * radius for m in cs):
                break
        cs.append(c)
    nat = np.zeros(lx * ly * lz, np.int8)
    for k in range(lz):
        for j in range(ly):
            for i in range(lx):
                p = (i + 1.0, j + 1.0, k + 1.0)
                for c in cs:
                    if math.sqrt(sum(axis_d(p[a], c[a], a) ** 2 for a in range(3))) <= radius:
                        nat[i + lx * (j + ly * k)] = 1
                        break
    return nat, np.array(cs).reshape(-1, 3)
