"""GPU parity of the dipole operators against the oracle's restatement of
mat_els.f90:120-170 / :348-390 (radial dipole matrices) and dipole.f90:8-150
with mat_els.f90:717-828 (dipole blocks between two symmetries).

Bar: index_ptr and indices bit-exact; data within 1e-12 of the block's
largest entry (the 6j symbols come from two independent evaluations, and
direct and exchange terms can cancel, so an entry-wise relative bound would
be ill-conditioned); radial matrices within 1e-12 of their largest entry
(the device evaluates the B-splines by the recurrence, the oracle by de
Boor's bvalue).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import bs2e  # noqa: E402
from bs2e import capi  # noqa: E402
from bs2e.workloads import WORKLOADS  # noqa: E402

TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    return bs2e.Context(0)


def _basis(oracle, name):
    w = WORKLOADS[name]
    t = oracle.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    return w, t, oracle.dims(w.k, t)[1]


def _close(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() <= TOL * max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", ["tiny", "t0", "c1"])
def test_radial_dipole_matches_oracle(oracle, ctx, name):
    w, t, n_b = _basis(oracle, name)
    bsp = bs2e.b_spline(w.k, t)
    print("radial dipole", name, flush=True)
    rl = bs2e.setup_radial_dip(bsp, w.k_GL, "l", ctx=ctx)
    assert rl.gauge == "l" and rl.r_mat.shape == (n_b, n_b)
    assert _close(rl.r_mat, oracle.setup_radial_dip(w.k, t, w.k_GL, "l"))
    rv = bs2e.setup_radial_dip(bsp, w.k_GL, "v", ctx=ctx)
    dr, ri = oracle.setup_radial_dip(w.k, t, w.k_GL, "v")
    assert _close(rv.dr_mat, dr) and _close(rv.r_inv_mat, ri)
    assert np.abs(rv.dr_mat.real).max() == 0 and np.abs(rl.r_mat.imag).max() == 0


def test_radial_dipole_rejects_bad_gauge(ctx):
    bsp = bs2e.b_spline(4, bs2e.generate_grid(4, 3, 2, 1.0, 8.0))
    with pytest.raises(bs2e.Bs2eError):
        bs2e.setup_radial_dip(bsp, 6, "x", ctx=ctx)
    with pytest.raises(bs2e.Bs2eError):
        capi.check(capi.lib().bs2e_setup_radial_dip(ctx.handle, 4, 3, None, 6, 0, None, None), ctx.handle)


CASES = [
    # name, (L1, M1, pi1), (L2, M2, pi2), q
    ("t0", (0, 0, 0), (1, 0, 1), 0),
    ("t0", (1, 1, 1), (2, 0, 0), 1),
    ("t0", (1, -1, 1), (2, 0, 0), -1),
    ("t0", (2, 0, 0), (1, 1, 1), -1),
    ("t0", (1, 0, 1), (1, 1, 0), -1),
    ("t0", (2, 1, 0), (3, 1, 1), 0),
    ("c1", (0, 0, 0), (1, 0, 1), 0),
    ("c1", (1, 0, 1), (2, 0, 0), 0),
]


@pytest.mark.parametrize("name,s1,s2,q", CASES)
@pytest.mark.parametrize("gauge", ["l", "v"])
def test_dip_block_matches_oracle(oracle, ctx, name, s1, s2, q, gauge):
    w, t, n_b = _basis(oracle, name)
    max_l = 2
    cf1 = oracle.count_configs(s1[0], s1[2], max_l, n_b, w.k, n_b, n_b, max_l)
    cf2 = oracle.count_configs(s2[0], s2[2], max_l, n_b, w.k, n_b, n_b, max_l)
    S = oracle.setup_S(w.k, t, w.k_GL)
    rad_o = oracle.setup_radial_dip(w.k, t, w.k_GL, gauge)
    rad = bs2e.radial_dipole("l", r_mat=rad_o) if gauge == "l" else \
        bs2e.radial_dipole("v", dr_mat=rad_o[0], r_inv_mat=rad_o[1])
    syms = [bs2e.sym(s1[0], s1[1], bool(s1[2]), cf1), bs2e.sym(s2[0], s2[1], bool(s2[2]), cf2)]
    print("dip block", name, s1, s2, q, gauge, len(cf1), len(cf2), flush=True)
    blk = bs2e.DipBlock(syms, q, bs2e.b_spline(w.k, t), ctx=ctx)
    got = blk.fill(S, rad)
    ms = blk.kernel_ms()
    blk.release()
    ip, ix, d = oracle.construct_dip_block(w.k, n_b, s1, cf1, s2, cf2, q, 1, gauge, S, rad_o)
    assert got.shape == (len(cf1), len(cf2))
    assert len(ix) > 0 and got.nnz == len(ix)
    assert np.array_equal(got.index_ptr, ip) and np.array_equal(got.indices, ix)
    assert _close(got.data, d)
    assert ms[0] > 0 and ms[1] > 0


@pytest.mark.parametrize("seed,k,k_gl,gauge", [(11, 3, 6, "l"), (12, 5, 8, "v"), (13, 10, 13, "l"), (14, 12, 15, "v")])
def test_dip_block_random_small_bases(oracle, ctx, seed, k, k_gl, gauge):
    """Small bases of other orders with drawn boxes: the closed-form row
    counts and the patterns bit-exact, values within 1e-12 of the block scale,
    for a drawn symmetry pair and component."""
    rng = np.random.default_rng(seed)
    t = bs2e.generate_grid(k, 3, 2, float(rng.uniform(0.6, 1.6)), float(rng.uniform(6.0, 14.0)))
    n_b = oracle.dims(k, t)[1]
    L1 = int(rng.integers(0, 3))
    pi1 = L1 % 2
    M1 = int(rng.integers(-L1, L1 + 1))
    q = int(rng.integers(-1, 2))
    L2, M2 = L1 + 1, M1 - q           # the 3j symbol (L1 1 L2; -M1 q M2) needs -M1 + q + M2 = 0
    s1, s2 = (L1, M1, pi1), (L2, M2, 1 - pi1)
    max_l = 2
    cf1 = oracle.count_configs(L1, pi1, max_l, n_b, k, n_b, n_b, max_l)
    cf2 = oracle.count_configs(L2, 1 - pi1, max_l, n_b, k, n_b, n_b, max_l)
    print("dip random", seed, k, n_b, s1, s2, q, gauge, len(cf1), len(cf2), flush=True)
    S = oracle.setup_S(k, t, k_gl)
    rad_o = oracle.setup_radial_dip(k, t, k_gl, gauge)
    rad = bs2e.radial_dipole("l", r_mat=rad_o) if gauge == "l" else \
        bs2e.radial_dipole("v", dr_mat=rad_o[0], r_inv_mat=rad_o[1])
    syms = [bs2e.sym(L1, M1, bool(pi1), cf1), bs2e.sym(L2, M2, not bool(pi1), cf2)]
    blk = bs2e.DipBlock(syms, q, bs2e.b_spline(k, t), ctx=ctx)
    got = blk.fill(S, rad)
    blk.release()
    ip, ix, d = oracle.construct_dip_block(k, n_b, s1, cf1, s2, cf2, q, 1, gauge, S, rad_o)
    assert got.shape == (len(cf1), len(cf2)) and got.nnz == len(ix)
    assert np.array_equal(got.index_ptr, ip) and np.array_equal(got.indices, ix)
    assert len(ix) > 0 and _close(got.data, d)


@pytest.mark.parametrize("k,gauge", [(2, "l"), (3, "v"), (4, "l"), (4, "v")])
def test_dip_block_tiny_bases(oracle, ctx, k, gauge):
    """The smallest bases the grid allows (orders 2-4, 7-9 B-splines), where
    every row and column touches an end of the basis: L = 0 -> 1 for
    q = -1, 0, 1 and L = 1 -> 2 for q = 0, patterns bit-exact, values within
    1e-12 of the block scale."""
    t = bs2e.generate_grid(k, 1, 2, 1.0, 3.0)
    n_b = oracle.dims(k, t)[1]
    k_gl, max_l = k + 3, 2
    S = oracle.setup_S(k, t, k_gl)
    rad_o = oracle.setup_radial_dip(k, t, k_gl, gauge)
    rad = bs2e.radial_dipole("l", r_mat=rad_o) if gauge == "l" else \
        bs2e.radial_dipole("v", dr_mat=rad_o[0], r_inv_mat=rad_o[1])
    for L1, M1, q in ((0, 0, -1), (0, 0, 0), (0, 0, 1), (1, 1, 0)):
        pi1, L2, M2 = L1 % 2, L1 + 1, M1 - q
        s1, s2 = (L1, M1, pi1), (L2, M2, 1 - pi1)
        cf1 = oracle.count_configs(L1, pi1, max_l, n_b, k, n_b, n_b, max_l)
        cf2 = oracle.count_configs(L2, 1 - pi1, max_l, n_b, k, n_b, n_b, max_l)
        print("dip tiny", k, n_b, s1, s2, q, gauge, len(cf1), len(cf2), flush=True)
        syms = [bs2e.sym(L1, M1, bool(pi1), cf1), bs2e.sym(L2, M2, not bool(pi1), cf2)]
        blk = bs2e.DipBlock(syms, q, bs2e.b_spline(k, t), ctx=ctx)
        got = blk.fill(S, rad)
        blk.release()
        ip, ix, d = oracle.construct_dip_block(k, n_b, s1, cf1, s2, cf2, q, 1, gauge, S, rad_o)
        assert got.shape == (len(cf1), len(cf2)) and got.nnz == len(ix)
        assert np.array_equal(got.index_ptr, ip) and np.array_equal(got.indices, ix)
        assert len(ix) > 0 and _close(got.data, d)


def test_dip_block_end_to_end_and_empty_cases(oracle, ctx):
    """Radial matrices and overlap from the device feed the device block; the
    reference's empty blocks (no parity change, vanishing 3j, compute false)
    have no entries and index_ptr all 1."""
    w, t, n_b = _basis(oracle, "t0")
    bsp = bs2e.b_spline(w.k, t)
    cf0 = oracle.count_configs(0, 0, 2, n_b, w.k, n_b, n_b, 2)
    cf1 = oracle.count_configs(1, 1, 2, n_b, w.k, n_b, n_b, 2)
    cf1e = oracle.count_configs(1, 0, 2, n_b, w.k, n_b, n_b, 2)
    S = bs2e.setup_S(bsp, w.k_GL, ctx=ctx)
    for g in ("l", "v"):
        rad = bs2e.setup_radial_dip(bsp, w.k_GL, g, ctx=ctx)
        got = bs2e.construct_dip_block_tensor([bs2e.sym(0, 0, False, cf0), bs2e.sym(1, 0, True, cf1)], 0, bsp, S,
                                              rad, ctx=ctx)
        rad_o = oracle.setup_radial_dip(w.k, t, w.k_GL, g)
        ip, ix, d = oracle.construct_dip_block(w.k, n_b, (0, 0, 0), cf0, (1, 0, 1), cf1, 0, 1, g,
                                               oracle.setup_S(w.k, t, w.k_GL), rad_o)
        assert np.array_equal(got.index_ptr, ip) and np.array_equal(got.indices, ix)
        assert _close(got.data, d)
    rad = bs2e.setup_radial_dip(bsp, w.k_GL, "l", ctx=ctx)
    empty = [
        ([bs2e.sym(0, 0, False, cf0), bs2e.sym(1, 0, False, cf1e)], 0, True),    # no parity change
        ([bs2e.sym(1, 0, True, cf1), bs2e.sym(1, 0, False, cf1e)], 0, True),     # 3j(1 1 1; 0 0 0) = 0
        ([bs2e.sym(0, 0, False, cf0), bs2e.sym(1, 0, True, cf1)], 1, True),      # -M1 + q + M2 != 0
        ([bs2e.sym(0, 0, False, cf0), bs2e.sym(1, 0, True, cf1)], 0, False),     # compute = .false.
    ]
    for syms, q, compute in empty:
        got = bs2e.construct_dip_block_tensor(syms, q, bsp, S, rad, compute=compute, ctx=ctx)
        assert got.nnz == 0 and np.all(got.index_ptr == 1) and got.shape == (syms[0].n_config, syms[1].n_config)


def test_dip_block_rejects_bad_input(oracle, ctx):
    w, t, n_b = _basis(oracle, "tiny")
    bsp = bs2e.b_spline(w.k, t)
    cf0 = oracle.count_configs(0, 0, 1, n_b, w.k, n_b, n_b, 1)
    cf1 = oracle.count_configs(1, 1, 1, n_b, w.k, n_b, n_b, 1)
    with pytest.raises(bs2e.Bs2eError):   # |M| > L
        bs2e.DipBlock([bs2e.sym(0, 1, False, cf0), bs2e.sym(1, 0, True, cf1)], 0, bsp, ctx=ctx)
    with pytest.raises(bs2e.Bs2eError):   # q outside -1..1
        bs2e.DipBlock([bs2e.sym(0, 0, False, cf0), bs2e.sym(1, 0, True, cf1)], 2, bsp, ctx=ctx)
    bad = cf1.copy()
    bad[0, 0] = n_b + 1                   # radial index outside the basis
    with pytest.raises(bs2e.Bs2eError):
        bs2e.DipBlock([bs2e.sym(0, 0, False, cf0), bs2e.sym(1, 0, True, bad)], 0, bsp, ctx=ctx)
    shuffled = cf1[::-1].copy()           # not in count_configs order
    with pytest.raises(bs2e.Bs2eError):
        bs2e.DipBlock([bs2e.sym(0, 0, False, cf0), bs2e.sym(1, 0, True, shuffled)], 0, bsp, ctx=ctx)
    blk = bs2e.DipBlock([bs2e.sym(0, 0, False, cf0), bs2e.sym(1, 0, True, cf1)], 0, bsp, ctx=ctx)
    with pytest.raises(bs2e.Bs2eError):   # wrong matrix size
        blk.fill(np.eye(3), bs2e.radial_dipole("l", r_mat=np.eye(3)))
    blk.release()
