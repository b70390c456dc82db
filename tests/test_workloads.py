"""Host-side parts of the batch and serving drivers (no GPU): the seeded
workload generators and the HBM footprints admission relies on."""
import math

import numpy as np

from laboetie_amd import lbx, serving


def test_chromatography_batch_is_seeded_and_in_spec():
    a = serving.chromatography_requests(12, seed=7)
    b = serving.chromatography_requests(12, seed=7)
    for r, s in zip(a, b):
        assert np.array_equal(r["nature"], s["nature"]) and r["ka"] == s["ka"]
    for k, r in enumerate(a):
        L = 50 if k % 2 == 0 else 100
        assert r["dims"] == (1, L, L)
        assert r["fext"] == (0.0, 1e-5, 0.0) and r["Db"] == 0.01 and r["kd"] == lbx.KD_DEFAULT
        assert -3.0 <= r["dF"] <= 0.0 and math.isclose(r["ka"], r["kd"] * math.exp(-r["dF"]))
        nd = int(r["kind"][len("disks"):])
        assert 1 <= nd <= 8
        assert 0 < int((r["nature"] == 1).sum()) < L * L


def test_serving_requests_in_spec():
    rng = np.random.default_rng(3)
    kinds = set()
    for _ in range(24):
        r = serving.serving_request(rng)
        N = r["dims"][0]
        assert r["dims"] == (N, N, N) and 32 <= N <= 192
        F = max(abs(v) for v in r["fext"])
        assert 1e-7 <= F <= 1e-5
        assert r["target"] == 1e-9 and r["mp_steps"] == 2000
        assert r["nature"].size == N ** 3 and (r["nature"] == 0).any()
        kinds.add(r["kind"])
    assert kinds == {"slit", "cylinder", "bcc", "random"}


def test_footprints_grow_with_size_and_capacity():
    f = [lbx.footprint_bytes(n, n, n) for n in (16, 32, 64)]
    assert f[0] < f[1] < f[2]
    assert lbx.footprint_bytes(32, 32, 32, lbx.PREC_FP32_SHIFT) < lbx.footprint_bytes(32, 32, 32)
    assert lbx.footprint_bytes(32, 32, 32, with_mp=3) > lbx.footprint_bytes(32, 32, 32, with_mp=1)
    b = [lbx.Batch.footprint_bytes(n, 2 * n, 64) for n in (1 << 16, 1 << 20)]
    assert b[0] < b[1]
    reqs = serving.chromatography_requests(4, seed=1)
    nodes, box = serving.capacity_for(reqs)
    assert nodes >= sum(int((r["nature"] == 0).sum()) for r in reqs)
    assert box >= sum(r["nature"].size for r in reqs)


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU; --gpus 2 exits at once with a message (no
    device is touched, so this runs on the CPU)."""
    import os
    import subprocess
    import sys

    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, os.path.join(repo, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""


def test_geom_in_writer_round_trips():
    """bench.py's geom.in writer (the reference's label-0 input) read back by
    the reference-grammar reader gives the same solid nodes."""
    import os
    import sys

    from laboetie_amd.lbin import read_geom_in

    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench

    dims = (5, 7, 3)
    nat = (np.random.default_rng(1).random(105) < 0.3).astype(np.int8)
    pts = read_geom_in(bench.geom_in_text(dims, nat))
    back = np.zeros(105, np.int8)
    back[(pts[:, 0] - 1) + 5 * ((pts[:, 1] - 1) + 7 * (pts[:, 2] - 1))] = 1
    assert np.array_equal(back, nat)
