"""lbx_collide_noalloc, the collide() entry point that allocates no device
memory, against lbx_collide on the same inputs: bit-identical results on two
calls in a row, the second on a smaller lattice (so it runs on what the first
left in the shared staging array), a lattice of exactly the largest size, the
reference's own collide() fixture, and the refusal of a lattice one node
larger."""
import sys

import numpy as np
import pytest

from golden_cases import load
from laboetie_amd import lbx

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_collide_inputs(dims, seed, solid_fraction=0.3):
    """A post-streaming state near rest with a random flow and per-node force."""
    rng = np.random.default_rng(seed)
    N = int(np.prod(dims))
    nature = (rng.random(N) < solid_fraction).astype(np.int8)
    _, a0, _, _, _ = lbx.d3q19_tables()
    rho = 1.0 + 0.05 * rng.standard_normal(N)
    n = (a0[:, None] * rho[None, :] * (1.0 + 0.1 * rng.standard_normal((19, N)))).ravel()
    j = 0.01 * rng.standard_normal((3, N))
    f = 1e-5 * rng.standard_normal((3, N))
    return nature, n, rho, j, f


@pytest.mark.parametrize("tau", [1.0, 0.8])
def test_noalloc_equals_collide_on_two_calls(tau):
    for dims, seed in (((41, 30, 23), 11), ((7, 5, 9), 12)):  # the second call is smaller
        say("collide_noalloc", dims, "tau", tau)
        nature, n, rho, j, f = random_collide_inputs(dims, seed)
        want = lbx.collide(dims, nature, tau, n, rho, j, f)
        got = lbx.collide(dims, nature, tau, n, rho, j, f, noalloc=True)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
        assert not np.array_equal(got, n)  # the fluid nodes did collide


def test_noalloc_at_the_largest_size_and_refused_beyond():
    dims = (80, 80, 50)
    assert np.prod(dims) == lbx.NOALLOC_MAX_NODES
    say("collide_noalloc largest", dims)
    nature, n, rho, j, f = random_collide_inputs(dims, 13)
    want = lbx.collide(dims, nature, 1.0, n, rho, j, f)
    got = lbx.collide(dims, nature, 1.0, n, rho, j, f, noalloc=True)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    # one node more is refused before any device work
    big = (lbx.NOALLOC_MAX_NODES + 1, 1, 1)
    nature, n, rho, j, f = random_collide_inputs(big, 14)
    with pytest.raises(lbx.LbxError) as e:
        lbx.collide(big, nature, 1.0, n, rho, j, f, noalloc=True)
    assert e.value.status == lbx.ERR_SIZE


def test_noalloc_matches_reference_fixture():
    g = load("collide_random")
    say("collide_noalloc collide_random", tuple(g["dims"]))
    n = lbx.collide(tuple(g["dims"]), g["nature"], float(g["tau"]), g["n_in"], g["rho"], g["j"], g["f"], noalloc=True)
    assert np.array_equal(n, g["n_out"])
