"""Oracle pieces used by the full-size GPU tests and the CPU baseline.

* rk_rows / slater_diag(pair_mask) / construct_block_rows reproduce rows of the
  full oracle bit for bit (the GPU tests compare sampled rows of full-size
  tensors and blocks with them);
* rk_dok (the reference's hash-map schedule, timed by bench.py) builds the
  same tensor as rk_dense;
* the CSR entry predicate of construct_block is checked against a brute-force
  evaluation of hamiltonian.f90:157-205 (fill) and :376-413 (count_nnz) over
  every (row, column) pair, with the angular factor of wigner_tools.f90:126-138
  evaluated by sympy, structural zeros included;
* the workloads' knot vectors are pinned as fixtures.
"""
import functools
import os

import numpy as np
import pytest

import bs2e
from bs2e.workloads import WORKLOADS, knots, perturb_breakpoints

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _basis(oracle, name):
    w = WORKLOADS[name]
    t = oracle.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    return w, t, oracle.dims(w.k, t)[1]


@pytest.mark.parametrize("name", ["tiny", "c1"])
def test_rk_rows_and_masked_cells_match_dense(oracle, name):
    w, t, n_b = _basis(oracle, name)
    off = oracle.slater_off_diag(w.k, t, w.k_GL, w.max_k)
    dg = oracle.slater_diag(w.k, t, w.k_GL, w.max_k)
    R = oracle.rk_dense(w.k, n_b, w.max_k, off, dg)
    npairs = R.shape[0]
    rng = np.random.default_rng(5)
    rows = np.unique(np.concatenate([[0, npairs - 1], rng.integers(0, npairs, 12)]))
    mask = np.zeros(npairs, bool)
    mask[rows] = True
    dm = oracle.slater_diag(w.k, t, w.k_GL, w.max_k, pair_mask=mask)
    for key in ("iv", "i", "ip", "j", "jp"):
        assert np.array_equal(dm[key], dg[key])
    pi = np.array([oracle.pair_index(w.k, n_b, a, c) for a, c in zip(dg["i"], dg["ip"])])
    pj = np.array([oracle.pair_index(w.k, n_b, a, c) for a, c in zip(dg["j"], dg["jp"])])
    touched = mask[pi] | mask[pj]
    assert np.array_equal(dm["rdk"][touched], dg["rdk"][touched])
    assert not np.any(dm["rdk"][~touched])
    Rr = oracle.rk_rows(w.k, n_b, w.max_k, off, dm, rows)
    assert np.array_equal(Rr, R[rows])


def test_dok_schedule_matches_dense(oracle):
    w, t, n_b = _basis(oracle, "tiny")
    off = oracle.slater_off_diag(w.k, t, w.k_GL, w.max_k)
    dg = oracle.slater_diag(w.k, t, w.k_GL, w.max_k)
    R = oracle.rk_dense(w.k, n_b, w.max_k, off, dg)
    n, Rd = oracle.rk_dok(w.k, n_b, w.max_k, off, dg, want_dense=True)
    assert np.array_equal(Rd, R)
    iv = off["iv"]
    same = sum(int(np.count_nonzero(iv == v)) ** 2 for v in np.unique(iv))
    assert n == off["nnz"] ** 2 - same + 2 * dg["nnz"]
    # a leading fraction touches only the rows of its sparse_4d entries
    n_lead = off["nnz"] // 3
    n2 = oracle.rk_dok(w.k, n_b, w.max_k, off, dg, n_lead=n_lead)
    assert 2 * dg["nnz"] < n2 < n


@pytest.mark.parametrize("L,full", [(0, True), (1, False), (2, True)])
def test_block_rows_match_full_block(oracle, L, full):
    w, t, n_b = _basis(oracle, "c1")
    off = oracle.slater_off_diag(w.k, t, w.k_GL, w.max_k)
    dg = oracle.slater_diag(w.k, t, w.k_GL, w.max_k)
    R = oracle.rk_dense(w.k, n_b, w.max_k, off, dg)
    S = oracle.setup_S(w.k, t, w.k_GL)
    Hl = [oracle.setup_H1p(w.k, t, w.k_GL, w.Z, l, cap_order=2, cap_r0=40.0, cap_eta=0.0) for l in range(3)]
    cf = oracle.count_configs(L, L % 2, 2, n_b, w.k, n_b, n_b, 2)
    ref = oracle.construct_block(w.k, n_b, L, w.max_k, full, cf, R, Hl, S)
    n = len(cf)
    for lo, hi in [(0, 7), (n // 2, n // 2 + 40), (n - 5, n)]:
        rows = oracle.block_tensor_rows(w.k, n_b, cf, lo, hi)
        rmap = np.full(R.shape[0], -1, np.int64)
        rmap[rows] = np.arange(len(rows))
        Rr = oracle.rk_rows(w.k, n_b, w.max_k, off, dg, rows)
        got = oracle.construct_block_rows(w.k, n_b, L, w.max_k, full, cf, Rr, rmap, lo, hi, Hl, S)
        for which in ("H", "S"):
            ip, ix, d = ref[which]
            a, b = ip[lo] - 1, ip[hi] - 1
            gip, gix, gd = got[which]
            assert np.array_equal(gip - 1, ip[lo:hi + 1] - ip[lo])
            assert np.array_equal(gix, ix[a:b])
            assert np.array_equal(gd, d[a:b])
    with pytest.raises(ValueError):
        rmap = np.full(R.shape[0], -1, np.int64)
        oracle.construct_block_rows(w.k, n_b, L, w.max_k, full, cf, R[:1], rmap, 0, 3, Hl, S)


# --------------------------------------------------------------------------
# brute-force CSR predicate
# --------------------------------------------------------------------------
@functools.lru_cache(maxsize=None)
def _ang_sympy(k, la, lb, lc, ld, L):
    """wigner_tools.f90:126-138 ang_k_LS(k, confs, L), confs(1)%l = (la, lb),
    confs(2)%l = (lc, ld), from sympy's exact 3j/6j."""
    from sympy import sqrt
    from sympy.physics.wigner import wigner_3j, wigner_6j
    if (la + k + lc) % 2 or (lb + k + ld) % 2:
        return 0.0

    def cred(kk, a, b):   # :107-112 C_red_mat
        return (-1) ** a * sqrt((2 * a + 1) * (2 * b + 1)) * wigner_3j(a, kk, b, 0, 0, 0)

    v = (-1) ** (lb + lc + L) * wigner_6j(la, lb, L, ld, lc, k) * cred(k, la, lc) * cred(k, lb, ld)
    return float(v)


def _brute_force_pattern(k_s, max_k, L, cf, full):
    """Per (i, j): H entry, S entry and the count_nnz predicates, evaluated
    pair by pair from the definitions (no shared code with the oracle)."""
    n = len(cf)
    rowsH, rowsS = [], []
    cntH = cntS = 0
    zeros = 0                              # entries with neither a two-electron nor a one-electron term
    for i in range(n):
        na, nb_, la, lb = (int(x) for x in cf[i, :4])
        hs, ss = [], []
        for j in range(0 if full else i, n):
            nc, nd, lc, ld = (int(x) for x in cf[j, :4])
            support = abs(na - nc) < k_s and abs(nb_ - nd) < k_s
            support_ex = abs(na - nd) < k_s and abs(nb_ - nc) < k_s
            ang = [_ang_sympy(k, la, lb, lc, ld, L) for k in range(max_k + 1)]
            ang_ex = [_ang_sympy(k, la, lb, ld, lc, L) for k in range(max_k + 1)]
            l_eq = la == lc and lb == ld
            l_eq_ex = la == ld and lb == lc
            if support:                    # count_nnz, :385-396
                cntS += l_eq
                cntH += any(abs(a) > 5e-15 for a in ang)
            elif support_ex:               # :397-412
                cntS += l_eq_ex
                cntH += any(abs(a) > 5e-15 for a in ang_ex)
            if not (support or support_ex):
                continue
            r12 = any((abs(a) > 5e-15 and support) or (abs(b) > 5e-15 and support_ex) for a, b in zip(ang, ang_ex))
            if (support and l_eq) or (support_ex and l_eq_ex):
                ss.append(j + 1)
            if r12 or l_eq or l_eq_ex:
                hs.append(j + 1)
                zeros += not r12 and not ((support and l_eq) or (support_ex and l_eq_ex))
        rowsH.append(hs)
        rowsS.append(ss)
    return rowsH, rowsS, (cntH, cntS), zeros


def _csr_rows(ip, ix):
    return [list(ix[ip[r] - 1:ip[r + 1] - 1]) for r in range(len(ip) - 1)]


# The last four have max_k < 2 max_l: entries kept with no term (structural
# zeros: l_eq with exchange support only, no allowed multipole) and rows where
# the reference's count_nnz counts fewer entries than the fill writes.
@pytest.mark.parametrize("k,r_max,h_max,max_l,max_k,L,full,odd", [
    (3, 3.0, 1.0, 2, 4, 0, True, False),
    (3, 4.0, 0.8, 2, 2, 1, True, False),
    (4, 4.0, 0.7, 3, 3, 2, False, False),
    (5, 4.0, 0.8, 2, 4, 3, True, False),
    (2, 3.0, 0.7, 3, 0, 0, False, False),
    (3, 3.0, 1.0, 1, 0, 1, True, True),
    (3, 3.0, 1.0, 2, 1, 2, True, True),
    (4, 3.0, 1.0, 2, 0, 1, False, True),
    (4, 3.0, 1.0, 3, 1, 2, True, True),
])
def test_block_pattern_brute_force(oracle, k, r_max, h_max, max_l, max_k, L, full, odd):
    t = oracle.generate_grid(k, 1, 2, h_max, r_max)
    n_b = oracle.dims(k, t)[1]
    cf = oracle.count_configs(L, L % 2, max_l, n_b, k, n_b, n_b, max_l)
    assert len(cf) > 0
    npairs = oracle.n_pairs(k, n_b)
    R = np.ones((npairs, npairs, max_k + 1))           # the pattern does not read values
    eye = np.eye(n_b, dtype=np.complex128)
    got = oracle.construct_block(k, n_b, L, max_k, full, cf, R, [eye] * (max_l + 1), eye)
    rowsH, rowsS, cnt, zeros = _brute_force_pattern(k, max_k, L, cf, full)
    assert _csr_rows(*got["H"][:2]) == rowsH
    assert _csr_rows(*got["S"][:2]) == rowsS
    assert got["ref_count"] == cnt
    assert got["nnz"][0] == sum(len(r) for r in rowsH)
    if odd:
        assert zeros > 0 and cnt[0] < got["nnz"][0]


# --------------------------------------------------------------------------
# workload knots
# --------------------------------------------------------------------------
def test_workload_knots_fixture():
    fx = np.load(os.path.join(GOLDEN, "workload_knots.npz"))
    for name, w in WORKLOADS.items():
        t = knots(w)
        assert np.array_equal(t, fx[name]), name
        assert bs2e.b_spline(w.k, t).n_b == w.n_b, name


def test_c5p_perturbation_rule():
    w = WORKLOADS["c5p"]
    t0 = bs2e.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    t = knots(w)
    assert len(t) == len(t0)
    assert np.array_equal(t[:w.k], t0[:w.k]) and np.array_equal(t[-w.k:], t0[-w.k:])   # end multiplicities
    b0, b = np.unique(t0), np.unique(t)
    assert len(b) == len(b0) and np.all(np.diff(b) > 0)
    rel = b[1:-1] / b0[1:-1] - 1.0
    assert np.all(np.abs(rel) <= 0.01) and np.abs(rel).max() > 0.009
    assert np.array_equal(perturb_breakpoints(t0, w.k, 42), t)       # deterministic
    assert not np.array_equal(perturb_breakpoints(t0, w.k, 43), t)


def test_recorded_h_max_values_give_one_more_spline():
    """The h_max values recorded beside the configs (C1 2.2476, C2 1.2007,
    C3 1.123, C5 1.7786) give n_b + 1 B-splines under the reference's grid
    rule (the compiled reference agrees, ref_fixtures.npz); the workloads use
    h_max values that give the stated n_b."""
    fx = np.load(os.path.join(GOLDEN, "ref_fixtures.npz"))
    for name, h in (("c1", 2.2476), ("c2", 1.2007), ("c3", 1.123), ("c5", 1.7786)):
        w = WORKLOADS[name]
        t = bs2e.generate_grid(w.k, w.m, w.Z, h, w.r_max)
        assert np.array_equal(t, fx[f"knots_recorded_hmax_{name}"])
        assert bs2e.b_spline(w.k, t).n_b == w.n_b + 1
