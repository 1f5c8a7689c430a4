"""One-electron matrices on the device (bs2e_setup_S, bs2e_setup_H_one_particle)
against the oracle's restatement of mat_els.f90:47-118 / :294-346 (itself
bit-exact in its B-spline values and derivatives against the compiled
reference), and the hydrogenic levels -Z^2/(2 n^2) from the device matrices.

Bar: every element within 1e-12 of the matrix's largest element (the device
evaluates B-splines and their second derivatives by the Cox-de Boor triangle
and the derivative recurrence rather than BVALUE_D's coefficient
differencing: equal to rounding).
"""
import numpy as np
import pytest
import scipy.linalg as sl

pytestmark = pytest.mark.gpu

import bs2e  # noqa: E402
from bs2e.workloads import WORKLOADS  # noqa: E402

TOL = 1e-12


def normwise(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("name,lmax", [("tiny", 1), ("c1", 2), ("c2", 5)])
@pytest.mark.parametrize("cap", [(2, 10.0, 0.0), (2, 30.0, 0.02 + 0.005j), (3, 20.0, 1e-3)])
def test_one_electron_matches_oracle(oracle, name, lmax, cap):
    w = WORKLOADS[name]
    t = bs2e.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    spl = bs2e.b_spline(w.k, t)
    print(f"one-electron {name} lmax={lmax} cap={cap}", flush=True)
    S = bs2e.setup_S(spl, w.k_GL)
    assert normwise(S, oracle.setup_S(w.k, t, w.k_GL)) <= TOL
    H = bs2e.setup_H_one_particle(2, cap, range(lmax + 1), spl, w.k_GL)
    for l in range(lmax + 1):
        ref = oracle.setup_H1p(w.k, t, w.k_GL, 2, l, cap_order=cap[0], cap_r0=cap[1], cap_eta=cap[2])
        assert normwise(H[l], ref) <= TOL, l
    one = bs2e.setup_H_one_particle(2, cap, lmax, spl, w.k_GL)
    assert np.array_equal(one, H[lmax])


def test_hydrogenic_levels_from_device_matrices():
    w = WORKLOADS["c1"]
    t = bs2e.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    spl = bs2e.b_spline(w.k, t)
    S = bs2e.setup_S(spl, w.k_GL).real
    for l in (0, 1, 2):
        H = bs2e.setup_H_one_particle(2, (2, 100.0, 0.0), l, spl, w.k_GL).real
        E = sl.eigh(H, S, eigvals_only=True)
        for i, n in enumerate(range(l + 1, l + 3)):   # levels well inside the r <= 50 box
            assert abs(E[i] + 4.0 / (2 * n * n)) < 1e-9, (l, n, E[i])
