"""bench.py's host-side pieces without a GPU: the one-GPU refusal and the
serving stream's job generator (grids of the drawn B-spline counts, inside
the range generate_grid accepts)."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode != 0
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["n_gpus"] == 2 and "one GPU" in out["error"]


def test_serving_jobs_have_the_drawn_sizes():
    sys.path.insert(0, ROOT)
    import bench
    import bs2e
    terms = bench.branch_terms(120, 8, 3)
    assert [tm.l for tm in terms] == [0, 1, 2] and all(tm.n_config > 0 for tm in terms)
    from bs2e.workloads import serving_jobs
    jobs = [(j.k, j.t, j.k_GL, j.max_k) for j in serving_jobs(64, 20261202)]
    again = [(j.k, j.t, j.k_GL, j.max_k) for j in serving_jobs(64, 20261202)]
    assert len(jobs) == 64
    for (k, t, k_gl, max_k), (k2, t2, _, _) in zip(jobs, again):
        assert k == k2 and np.array_equal(t, t2)            # deterministic for a seed
        assert k in (6, 7, 8, 9) and k_gl == k + 6 and max_k % 2 == 0
        n_b = len(t) - k - 2
        assert 40 <= n_b <= 400
        assert t[0] == 0.0 and np.all(np.diff(t) >= 0.0)
    from bs2e.workloads import largest_serving_job, solve_h_max
    h = solve_h_max(8, 3, 2, 100.0, 150)
    assert h is not None and len(bs2e.generate_grid(8, 3, 2, h, 100.0)) - 8 - 2 == 150
    big = largest_serving_job()
    assert (big.k, len(big.t) - big.k - 2, big.k_GL, big.max_k) == (9, 400, 15, 16)
