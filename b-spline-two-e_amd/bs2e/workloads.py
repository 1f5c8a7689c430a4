"""Named workloads of the benchmark and parity suite.

Knots come from generate_grid(k, m=3, Z=2, h_max, r_max) (grid_tools.f90:6-55);
h_max is the middle of the plateau of values that gives exactly the named
number of B-splines n_b.  A workload with perturb_seed set moves every interior
distinct breakpoint by a seeded multiplicative factor (see perturb_breakpoints).
"""
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class Workload:
    name: str
    k: int          # B-spline order
    n_b: int        # retained B-splines
    r_max: float
    h_max: float
    k_GL: int
    max_k: int
    lmax: int
    m: int = 3
    Z: int = 2
    description: str = ""
    perturb_seed: int | None = None   # seed of the breakpoint perturbation (None: the bare grid)
    perturb_rel: float = 0.01         # largest relative move of a breakpoint


WORKLOADS = {
    "tiny": Workload("tiny", 4, 36, 8.0, 1.0, 6, 2, 1, description="order 4, 36 B-splines, r<=8, GL 6, k<=2"),
    "t0": Workload("t0", 8, 46, 15.0, 1.5, 14, 4, 2,
                   description="the reference test's basis: order 8, r<=15, h_max 1.5, GL 14, k<=4"),
    "c1": Workload("c1", 7, 60, 50.0, 2.3384, 10, 4, 2,
                   description="He-like Z=2, order 7, 60 B-splines, r<=50, GL 10, l<=2, k<=4"),
    "c2": Workload("c2", 8, 200, 200.0, 1.2046, 12, 10, 5,
                   description="order 8, 200 B-splines, r<=200, GL 12, lmax 5, k<=10"),
    "c3": Workload("c3", 8, 300, 300.0, 1.1252, 12, 16, 8,
                   description="order 8, 300 B-splines, r<=300, GL 12, lmax 8, k<=16"),
    "c4": Workload("c4", 9, 400, 400.0, 1.0933, 15, 16, 8,
                   description="order 9, 400 B-splines, r<=400, GL 15, k<=16 (largest serving job)"),
    "c5": Workload("c5", 9, 600, 1000.0, 1.7802, 14, 24, 12,
                   description="order 9, 600 B-splines, r<=1000, GL 14, lmax 12, k<=24"),
    "c5p": Workload("c5p", 9, 600, 1000.0, 1.7802, 14, 24, 12, perturb_seed=42,
                    description="order 9, 600 B-splines, r<=1000, GL 14, lmax 12, k<=24, interior breakpoints "
                                "moved by a seeded (42) factor in [0.99, 1.01]"),
}


def perturb_breakpoints(t, k, seed, rel=0.01):
    """Knot vector t with every interior distinct breakpoint b_i (all but the
    first and the last) replaced by b_i * (1 + u), u drawn uniformly from
    [-rel, rel) by numpy's default_rng(seed), left to right; a draw that is not
    strictly above the previous (moved) breakpoint and strictly below the next
    (unmoved) one is drawn again.  The end knots keep their multiplicity k, so
    the spline count and the support rule are those of t."""
    t = np.asarray(t, dtype=np.float64)
    bp = np.unique(t)
    if len(bp) < 3:
        return t.copy()
    rng = np.random.default_rng(seed)
    out = bp.copy()
    for i in range(1, len(bp) - 1):
        while True:
            v = bp[i] * (1.0 + rng.uniform(-rel, rel))
            if out[i - 1] < v < bp[i + 1]:
                out[i] = v
                break
    head = np.full(int(np.count_nonzero(t == bp[0])), out[0])
    tail = np.full(int(np.count_nonzero(t == bp[-1])), out[-1])
    return np.concatenate([head, out[1:-1], tail])


def knots(w):
    """Knot vector of workload w (a Workload or its name)."""
    import bs2e
    if isinstance(w, str):
        w = WORKLOADS[w]
    t = bs2e.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    if w.perturb_seed is not None:
        t = perturb_breakpoints(t, w.k, w.perturb_seed, w.perturb_rel)
    return t


def solve_h_max(k, m, Z, r_max, n_b):
    """h_max whose grid (grid_tools.f90:6-55 rule) gives exactly n_b B-splines, or None.
    generate_grid refuses grids that overrun the reference's knot array (a
    linear part finer than 2^-m, or a geometric ramp longer than r_max); the
    search stays inside the h_max range it accepts."""
    import bs2e

    def nb(h):
        try:
            return len(bs2e.generate_grid(k, m, Z, h, r_max)) - k - 2
        except bs2e.Bs2eError:
            return None

    lo, hi = r_max / 4000.0, r_max
    while nb(lo) is None and lo < hi:
        lo *= 1.25
    while nb(hi) is None and hi > lo:
        hi /= 1.25
    n_lo, n_hi = nb(lo), nb(hi)
    if n_lo is None or n_hi is None or n_lo < n_b or n_hi > n_b:
        return None
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        v = nb(mid)
        if v is None:
            return None
        if v == n_b:
            return mid
        if v > n_b:
            lo = mid
        else:
            hi = mid
    return None


@dataclass(frozen=True)
class ServingJob:
    """One basis_setup request of the serving stream."""
    k: int
    t: np.ndarray
    k_GL: int
    max_k: int
    lmax: int


def serving_jobs(n, seed=7):
    """The serving stream's jobs: B-spline count uniform in [40, 400], order in
    {6, 7, 8, 9}, r_max uniform in [50, 500], lmax uniform in [2, 8],
    k <= 2 lmax, k_GL = order + 6 (input_tools.f90:833 default), m = 3, Z = 2,
    h_max solved for the drawn count (draws without such a grid are drawn
    again)."""
    import bs2e
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(100 * n):
        if len(jobs) == n:
            break
        n_b = int(rng.integers(40, 401))
        k = int(rng.choice([6, 7, 8, 9]))
        r_max = float(rng.uniform(50.0, 500.0))
        lmax = int(rng.integers(2, 9))
        h = solve_h_max(k, 3, 2, r_max, n_b)
        if h is None:
            continue
        jobs.append(ServingJob(k, bs2e.generate_grid(k, 3, 2, h, r_max), k + 6, 2 * lmax, lmax))
    if len(jobs) < n:
        raise RuntimeError(f"serving_jobs: only {len(jobs)} of {n} jobs have a grid with the drawn size")
    return jobs


def largest_serving_job():
    """The largest job the serving stream can draw: order 9, 400 B-splines,
    lmax 8 (k <= 16), GL 15 (the c4 workload's basis)."""
    w = WORKLOADS["c4"]
    return ServingJob(w.k, knots(w), w.k_GL, w.max_k, w.lmax)
