"""bench.py's command line and the batched workload's case set (no device)."""
import math
import os
import subprocess
import sys

import numpy as np

from hystrath_amd import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_cylinder_batch_spans_the_stated_ranges():
    """64 cases: Mach in [5, 25], Kn = mean free path / cylinder diameter in
    [0.01, 1] (hard-sphere mean free path of argon at d_ref), 50 to 300 cells
    a side, 10k to 2.1M freestream parcels; a quarter cell per step at the
    freestream speed; the same set for the same seed, another for another."""
    cs = cases.cylinder_batch(64)
    assert len(cs) == 64
    d_ref = 4.17e-10
    parcels = []
    for c in cs:
        p = c.batch_params
        assert 5.0 <= p["mach"] <= 25.0
        assert 50 <= p["n_side"] <= 300 and c.ncells == p["n_side"] ** 2
        lam = 1.0 / (math.sqrt(2.0) * math.pi * d_ref ** 2 * c.n_inf)
        assert 0.01 * (1 - 1e-9) <= lam / 0.2 <= 1.0 * (1 + 1e-9)
        assert abs(lam / 0.2 - p["kn"]) <= 1e-5 * max(p["kn"], 1.0) + 5e-6
        dx = 1.0 / p["n_side"]
        assert abs(c.U_inf * c.cfg.dt / dx - 0.25) < 1e-12
        n_est = c.n_inf * c.ncells * dx * dx * dx / c.cfg.fn
        assert abs(n_est / (p["n_side"] ** 2 * p["ppc"]) - 1) < 1e-2
        assert c.cfg.capacity >= 2.5 * n_est
        parcels.append(n_est)
    assert 1e4 <= min(parcels) and max(parcels) <= 2.1e6
    assert 2.0e7 < sum(parcels) < 4.5e7
    again = cases.cylinder_batch(64)
    assert [c.batch_params for c in again] == [c.batch_params for c in cs]
    other = cases.cylinder_batch(64, seed=1)
    assert [c.batch_params for c in other] != [c.batch_params for c in cs]
    assert len({round(c.batch_params["mach"], 3) for c in cs}) == 64
    assert np.std([c.batch_params["n_side"] for c in cs]) > 40
