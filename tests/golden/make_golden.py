"""Writes tests/golden/ref_fixtures.npz from the compiled reference.

The reference modules without external dependencies (grid_tools, bspline_tools,
orbital_tools) are compiled from /root/reference by oracle/Makefile into
oracle/_ref/libbs2e_ref.so; this script runs them on fixed inputs and stores
inputs and outputs, so that the pins hold where /root/reference is absent.

    make -C oracle ref && python tests/golden/make_golden.py

It also writes tests/golden/workload_knots.npz: the knot vector of every named
workload (bs2e.workloads.knots, the seeded breakpoint perturbation included),
so that the workloads' bases are pinned as data.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "b-spline-two-e_amd"))

from oracle import oracle as O  # noqa: E402
from bs2e.workloads import WORKLOADS  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_fixtures.npz")
KNOTS_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "workload_knots.npz")


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libbs2e_ref.so missing: run `make -C oracle ref` where /root/reference exists")
    rng = np.random.default_rng(20261202)
    fx = {}
    for name in ("tiny", "t0", "c1", "c2", "c3", "c5"):
        w = WORKLOADS[name]
        fx[f"knots_{name}"] = O.ref_generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    # the h_max values recorded beside the benchmark configs
    for name, h in (("c1", 2.2476), ("c2", 1.2007), ("c3", 1.123), ("c5", 1.7786)):
        w = WORKLOADS[name]
        fx[f"knots_recorded_hmax_{name}"] = O.ref_generate_grid(w.k, w.m, w.Z, h, w.r_max)
    # B-spline values (BVALUE_D through eval_d / d_eval_d) at random points of
    # random intervals, for every local spline, value and 1st/2nd derivative.
    for name in ("tiny", "c1"):
        w = WORKLOADS[name]
        t = fx[f"knots_{name}"]
        n, n_b, n_bp, _ = O.ref_spline_info(t, w.k)
        niv = n_bp - 1
        iv = rng.integers(1, niv + 1, size=400)
        ls = rng.integers(0, w.k, size=400)
        s = iv + ls
        bp = t[w.k - 1:w.k - 1 + niv + 1]
        u = rng.random(400)
        x = bp[iv - 1] + u * (bp[iv] - bp[iv - 1])
        fx[f"bv_{name}_s"] = s
        fx[f"bv_{name}_iv"] = iv
        fx[f"bv_{name}_x"] = x
        for d in (0, 1, 2):
            fx[f"bv_{name}_d{d}"] = O.ref_bvalue_batch(t, w.k, s, x, iv, d)
    # count_configs for the C1 basis, z-polarised symmetries L = 0..2, with and
    # without the r_2_max / r_all_l truncations.
    w = WORKLOADS["c1"]
    t = fx["knots_c1"]
    _, n_b, _, _ = O.ref_spline_info(t, w.k)
    trunc = (O.ref_find_max_n_b(t, w.k, 15.0), O.ref_find_max_n_b(t, w.k, 25.0))
    fx["c1_find_max_n_b"] = np.array(trunc)
    for L in range(3):
        fx[f"cfg_c1_L{L}"] = O.ref_count_configs(L, L % 2, 2, n_b, w.k, n_b, n_b, 2)
        fx[f"cfg_c1_trunc_L{L}"] = O.ref_count_configs(L, L % 2, 3, n_b, w.k, trunc[0], trunc[1], 1)
    np.savez_compressed(OUT, **fx)
    print("wrote", OUT, sum(v.nbytes for v in fx.values()), "bytes")
    from bs2e.workloads import knots
    kn = {name: knots(w) for name, w in WORKLOADS.items()}
    np.savez_compressed(KNOTS_OUT, **kn)
    print("wrote", KNOTS_OUT, sum(v.nbytes for v in kn.values()), "bytes")


if __name__ == "__main__":
    main()
