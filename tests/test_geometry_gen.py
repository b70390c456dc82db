"""Seeded medium generators of the C-ABI (host code, no GPU): Gaussian random
porous media (lbx_geometry_random) and chromatography disks
(lbx_geometry_disks), against their numpy restatement (oracle/geomgen.py)
and closed forms."""
import math

import numpy as np
import pytest

import geomgen
from laboetie_amd import lbx


def test_gaussian_field_matches_restatement():
    dims, sigma, seed = (24, 20, 18), 3.0, 7
    f = lbx.gaussian_field(*dims, sigma, seed)
    g = geomgen.gaussian_field(*dims, sigma, seed)
    assert np.abs(f - g).max() <= 1e-12 * np.abs(g).max()


@pytest.mark.parametrize("dims,sigma,phi,seed", [((24, 20, 18), 3.0, 0.6, 7), ((31, 17, 9), 2.0, 0.4, 11),
                                                 ((16, 16, 1), 1.5, 0.8, 3)])
def test_random_medium_matches_restatement_and_porosity(dims, sigma, phi, seed):
    nat = lbx.geometry_random(*dims, sigma, phi, seed)
    N = int(np.prod(dims))
    assert int((nat == 0).sum()) == round(phi * N)
    assert np.array_equal(nat, geomgen.geometry_random(*dims, sigma, phi, seed))
    assert not np.array_equal(nat, lbx.geometry_random(*dims, sigma, phi, seed + 1))


def test_filtered_noise_variance_and_correlation():
    """Unit white noise through the separable filter w(d) = exp(-d^2/2s^2):
    variance (sum w^2)^3 and, along each axis, correlation
    sum_d w(d) w(d+h) / sum_d w(d)^2 at lag h."""
    L, s = 96, 2.0
    f = lbx.gaussian_field(L, L, L, s, 5).reshape(L, L, L)
    R = math.ceil(4 * s)
    w = np.exp(-np.arange(-R, R + 1) ** 2 / (2 * s * s))
    var = (w ** 2).sum() ** 3
    assert abs(f.var() / var - 1.0) < 0.03
    h = 2
    rho_th = (w[:-h] * w[h:]).sum() / (w ** 2).sum()
    for axis in range(3):
        c = (f * np.roll(f, -h, axis=axis)).mean() / f.var()
        assert abs(c - rho_th) < 0.03


@pytest.mark.parametrize("dims,nd,seed", [((1, 50, 50), 8, 3), ((1, 100, 100), 5, 9), ((1, 50, 50), 1, 1)])
def test_disks_match_restatement_and_reference_rule(dims, nd, seed):
    nat, c = lbx.geometry_disks(*dims, nd, 5.0, seed)
    nat2, c2 = geomgen.geometry_disks(*dims, nd, 5.0, seed)
    assert np.array_equal(c, c2) and np.array_equal(nat, nat2)
    lx, ly, lz = dims
    assert np.all(c[:, 0] == 1.0)  # flat axis
    # no two disks closer than a diameter (minimum image)
    for a in range(nd):
        for b in range(a + 1, nd):
            d = [min(abs(c[a, k] - c[b, k]), dims[k] - abs(c[a, k] - c[b, k])) for k in range(3)]
            assert math.sqrt(sum(x * x for x in d)) > 10.0
    # every solid node lies within a radius of a centre and vice versa
    k, j, i = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
    inside = np.zeros((lz, ly, lx), bool)
    for m in range(nd):
        dy = np.abs(j + 1 - c[m, 1])
        dy = np.minimum(dy, ly - dy)
        dz = np.abs(k + 1 - c[m, 2])
        dz = np.minimum(dz, lz - dz)
        inside |= np.sqrt(dy * dy + dz * dz) <= 5.0
    assert np.array_equal(nat.reshape(lz, ly, lx) == 1, inside)


def test_generator_argument_errors():
    with pytest.raises(lbx.LbxError):
        lbx.geometry_random(8, 8, 8, 2.0, 0.0, 1)
    with pytest.raises(lbx.LbxError):
        lbx.geometry_disks(1, 20, 20, 50, 5.0, 1)  # no room
