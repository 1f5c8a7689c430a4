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


@pytest.mark.parametrize("name,max_k,r0_of", [("tiny", 2, "mid"), ("t0", 4, "knot"), ("c1", 4, "mid"),
                                               ("c2", 10, "mid"), ("c1", 8, "far"), ("c1", 4, "near")])
def test_r_k_int_matches_oracle(oracle, name, max_k, r0_of):
    """RKH_hamiltonian.f90:91-159 setup_r_k_int on the device against the
    oracle, per multipole k within 1e-12 of that k's largest element; r_0
    inside the grid (between or on breakpoints), beyond it and below its
    first non-zero knot."""
    w = WORKLOADS[name]
    t = bs2e.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    spl = bs2e.b_spline(w.k, t)
    r0 = {"mid": 0.37 * t[-1], "knot": t[w.k + 9], "far": 3.0 * t[-1], "near": 1e-3 * t[w.k]}[r0_of]
    print(f"r_k_int {name} max_k={max_k} r_0={r0}", flush=True)
    got = bs2e.setup_r_k_int(spl, w.k_GL, max_k, r0)
    ref = oracle.setup_r_k_int(w.k, t, w.k_GL, max_k, r0)
    assert got.shape == ref.shape == (max_k + 1, spl.n_b, spl.n_b)
    for k in range(max_k + 1):
        assert normwise(got[k], ref[k]) <= TOL, k
    i, j = np.indices((spl.n_b, spl.n_b))
    assert np.all(got[:, np.abs(i - j) >= w.k] == 0)


def test_r_k_int_rejects_bad_input():
    w = WORKLOADS["tiny"]
    spl = bs2e.b_spline(w.k, bs2e.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max))
    for max_k, r0 in ((-1, 1.0), (2, 0.0), (2, float("nan")), (2, -3.0)):
        with pytest.raises(bs2e.Bs2eError):
            bs2e.setup_r_k_int(spl, w.k_GL, max_k, r0)


def test_fortran_rkh_multipoles(oracle, tmp_path):
    """The amdflang program over the drop-in module bs2e_rkh (init_g_l,
    setup_r_k_int with the module variables r_0 and max_k) writes the same
    multipoles as the oracle."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(bs2e.__file__), "..", "lib", "bs2e_rkh_multipoles")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is missing: build with `python b-spline-two-e_amd/build.py` (needs amdflang)")
    w = WORKLOADS["c1"]
    r0 = 7.25
    out = tmp_path / "rk.bin"
    cmd = [exe, str(w.k), str(w.m), str(w.Z), repr(w.h_max), repr(w.r_max), str(w.k_GL), "6", repr(r0), str(out)]
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    raw = out.read_bytes()
    n_b, max_k = np.frombuffer(raw[:16], np.int64)
    got = np.frombuffer(raw[16:], np.float64).reshape(max_k + 1, n_b, n_b, order="F")
    t = oracle.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    ref = oracle.setup_r_k_int(w.k, t, w.k_GL, 6, r0)
    assert max_k == 6 and got.shape == ref.shape
    for k in range(7):
        assert normwise(got[k], ref[k]) <= TOL, k
