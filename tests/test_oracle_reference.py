"""Pins of the oracle against the reference itself.

The reference's dependency-free Fortran modules (grid_tools, bspline_tools,
orbital_tools) are compiled from its own sources into oracle/_ref; their
outputs on fixed inputs are stored in tests/golden/ref_fixtures.npz by
tests/golden/make_golden.py.  These tests compare the oracle bit for bit
with those fixtures (and with the live compiled reference where it exists).
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FX = np.load(os.path.join(HERE, "golden", "ref_fixtures.npz"))

from bs2e.workloads import WORKLOADS  # noqa: E402


@pytest.mark.parametrize("name", ["tiny", "t0", "c1", "c2", "c3", "c5"])
def test_generate_grid_bit_exact(oracle, name):
    """grid_tools.f90:6-55"""
    w = WORKLOADS[name]
    t = oracle.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    assert np.array_equal(t, FX[f"knots_{name}"])
    n, n_b, _ = oracle.dims(w.k, t)
    assert n_b == w.n_b


@pytest.mark.parametrize("name,k", [("tiny", 4), ("c1", 7)])
@pytest.mark.parametrize("deriv", [0, 1, 2])
def test_bvalue_bit_exact(oracle, name, k, deriv):
    """bspline_tools.f90:151-224 (values and the derivative branch)."""
    t = FX[f"knots_{name}"]
    s, x, iv = FX[f"bv_{name}_s"], FX[f"bv_{name}_x"], FX[f"bv_{name}_iv"]
    mine = np.array([oracle.bvalue_unit(t, k, int(a), float(b), int(c), deriv) for a, b, c in zip(s, x, iv)])
    assert np.array_equal(mine, FX[f"bv_{name}_d{deriv}"])


@pytest.mark.parametrize("L", [0, 1, 2])
def test_count_configs_bit_exact(oracle, L):
    """orbital_tools.f90:119-216, with and without the r_2_max / r_all_l truncations."""
    t = FX["knots_c1"]
    _, n_b, _ = oracle.dims(7, t)
    assert np.array_equal(oracle.count_configs(L, L % 2, 2, n_b, 7, n_b, n_b, 2), FX[f"cfg_c1_L{L}"])
    n2, nall = (int(v) for v in FX["c1_find_max_n_b"])
    assert np.array_equal(oracle.count_configs(L, L % 2, 3, n_b, 7, n2, nall, 1), FX[f"cfg_c1_trunc_L{L}"])


def test_c1_block_size_is_5311(oracle):
    assert FX["cfg_c1_L0"].shape[0] == 5311


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libbs2e_ref.so")),
                    reason="compiled reference (oracle/_ref) not built here")
def test_live_reference_random_points(oracle):
    """Fresh random bases and points against the compiled reference."""
    rng = np.random.default_rng(11)
    for _ in range(6):
        k = int(rng.integers(3, 11))
        h_max = float(rng.uniform(0.6, 2.5))
        r_max = float(rng.uniform(10, 60))
        t = oracle.ref_generate_grid(k, 3, 2, h_max, r_max)
        assert np.array_equal(t, oracle.generate_grid(k, 3, 2, h_max, r_max))
        n, n_b, n_bp, sup = oracle.ref_spline_info(t, k)
        assert (n, n_b) == oracle.dims(k, t)[:2]
        niv = n_bp - 1
        ivs = np.arange(1, niv + 1)[:, None]
        ss = np.arange(1, n + 1)[None, :]
        assert np.array_equal(sup, (ivs <= ss) & (ss <= ivs + k - 1))   # bspline_tools.f90:42-56
        iv = rng.integers(1, niv + 1, 200)
        s = iv + rng.integers(0, k, 200)
        bp = t[k - 1:k - 1 + niv + 1]
        x = bp[iv - 1] + rng.random(200) * (bp[iv] - bp[iv - 1])
        for d in (0, 2):
            ref = oracle.ref_bvalue_batch(t, k, s, x, iv, d)
            mine = np.array([oracle.bvalue_unit(t, k, int(a), float(b), int(c), d) for a, b, c in zip(s, x, iv)])
            assert np.array_equal(mine, ref)
