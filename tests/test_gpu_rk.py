"""GPU parity of the R^k path (libbs2e.so through its C ABI) against the oracle.

Bar: index arrays and key sets bit-exact; values within 1e-12 relative,
element by element (every primitive and every R^k value is >= 0: B-splines
are non-negative, GL weights positive, the kernel positive, so the relative
error is well conditioned).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import bs2e  # noqa: E402
from bs2e import capi  # noqa: E402
from bs2e.workloads import WORKLOADS  # noqa: E402

TOL = 1e-12


def max_rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape
    d = np.abs(a - b)
    s = np.abs(b)
    nz = s > 0
    rel = float((d[nz] / s[nz]).max()) if nz.any() else 0.0
    zero_part = float(np.abs(a[~nz]).max()) if (~nz).any() else 0.0
    return rel, zero_part


@pytest.fixture(scope="module")
def ctx():
    return bs2e.Context(0)


def knots(w):
    from bs2e.workloads import knots as workload_knots
    return workload_knots(w)


def gpu_containers(w, t, ctx):
    spl = bs2e.b_spline(w.k, t)
    r_k = bs2e.sparse_4d().init(w.max_k, spl)
    r_m_k = bs2e.sparse_4d().init(w.max_k, spl)
    r_d_k = bs2e.sparse_6d().init(w.max_k, spl)
    bs2e.setup_Slater_integrals(spl, w.max_k, w.k_GL, r_k, r_m_k, r_d_k, ctx=ctx)
    return spl, r_k, r_m_k, r_d_k


@pytest.mark.parametrize("name", ["tiny", "t0", "c1", "c2"])
def test_setup_slater_integrals_matches_oracle(name, ctx, oracle):
    w = WORKLOADS[name]
    t = knots(w)
    _, r_k, r_m_k, r_d_k = gpu_containers(w, t, ctx)
    off = oracle.slater_off_diag(w.k, t, w.k_GL, w.max_k)
    dg = oracle.slater_diag(w.k, t, w.k_GL, w.max_k)
    assert r_k.nnz == off["nnz"] and r_d_k.nnz == dg["nnz"]
    for a, b in ((r_k.iv, off["iv"]), (r_k.i, off["i"]), (r_k.j, off["j"]), (r_m_k.iv, off["iv"])):
        assert np.array_equal(a, b)
    for key in ("iv", "i", "ip", "j", "jp"):
        mine = {"ip": r_d_k.i_p, "jp": r_d_k.j_p}.get(key, getattr(r_d_k, key, None))
        assert np.array_equal(mine, dg[key]), key
    for got, ref in ((r_k.data, off["rk"]), (r_m_k.data, off["rmk"]), (r_d_k.data, dg["rdk"])):
        rel, zp = max_rel(got, ref)
        assert rel <= TOL and zp == 0.0, (name, rel, zp)


@pytest.mark.parametrize("name", ["tiny", "t0", "c1", "c2"])
def test_rk_tensor_matches_oracle(name, ctx, oracle):
    w = WORKLOADS[name]
    t = knots(w)
    rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx)
    R = rk.to_host()
    off = oracle.slater_off_diag(w.k, t, w.k_GL, w.max_k)
    dg = oracle.slater_diag(w.k, t, w.k_GL, w.max_k)
    n_b = oracle.dims(w.k, t)[1]
    Rref, touched = oracle.rk_dense(w.k, n_b, w.max_k, off, dg, want_touched=True)
    assert touched.all()          # every key of count_nnz_R_k is set by the reference
    assert R.shape == Rref.shape == (oracle.n_pairs(w.k, n_b),) * 2 + (w.max_k + 1,)
    rel, zp = max_rel(R, Rref)
    assert rel <= TOL and zp == 0.0, (name, rel, zp)
    assert (R > 0).all()
    rk.release()


@pytest.mark.parametrize("name", ["tiny", "c1"])
def test_compute_R_K_map_from_containers(name, ctx, oracle):
    """compute_R_K_map drop-in: the tensor is built from the caller's containers."""
    w = WORKLOADS[name]
    t = knots(w)
    spl, r_k, r_m_k, r_d_k = gpu_containers(w, t, ctx)
    R = bs2e.Nd_DOK()
    bs2e.compute_R_K_map(r_d_k, r_k, r_m_k, spl, w.max_k, R, ctx=ctx)
    got = R.tensor.to_host()
    off = oracle.slater_off_diag(w.k, t, w.k_GL, w.max_k)
    dg = oracle.slater_diag(w.k, t, w.k_GL, w.max_k)
    Rref = oracle.rk_dense(w.k, spl.n_b, w.max_k, off, dg)
    rel, zp = max_rel(got, Rref)
    assert rel <= TOL and zp == 0.0
    # and from the oracle's own containers (values the caller holds, any source)
    r_k.data[:] = off["rk"]
    r_m_k.data[:] = off["rmk"]
    r_d_k.data[:] = dg["rdk"]
    R2 = bs2e.compute_R_K_map(r_d_k, r_k, r_m_k, spl, w.max_k, bs2e.Nd_DOK(), ctx=ctx)
    rel, zp = max_rel(R2.tensor.to_host(), Rref)
    assert rel <= TOL and zp == 0.0


def test_compute_R_K_map_rejects_foreign_indices(ctx):
    w = WORKLOADS["tiny"]
    t = knots(w)
    spl, r_k, r_m_k, r_d_k = gpu_containers(w, t, ctx)
    r_k.j[3] = r_k.i[3] + w.k + 2      # outside the band |i-j| < k
    with pytest.raises(bs2e.Bs2eError) as e:
        bs2e.compute_R_K_map(r_d_k, r_k, r_m_k, spl, w.max_k, bs2e.Nd_DOK(), ctx=ctx)
    assert e.value.code == capi.BS2E_ERANGE


def test_get_val_and_pair_index(ctx, oracle):
    w = WORKLOADS["c1"]
    t = knots(w)
    rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx)
    full = rk.to_host()
    pairs = oracle.pair_table(w.k, rk.n_b)
    for p, (a, c) in enumerate(pairs):
        assert rk.pair_index(a, c) == p
    rng = np.random.default_rng(3)
    for _ in range(200):
        p, q = rng.integers(0, rk.n_pairs, size=2)
        (a, c), (b, d) = pairs[p], pairs[q]
        np.testing.assert_array_equal(rk.get_val([a, b, c, d]), full[p, q])
    with pytest.raises(bs2e.Bs2eError) as e:
        rk.get_val([1, 1, 1 + w.k, 1])
    assert e.value.code == capi.BS2E_ERANGE
    # symmetric integrals: R(a,b,c,d) = R(b,a,d,c) (exact math, ~1 ulp here:
    # elements (p, q) and (q, p) of overlapping windows are rounded apart)
    perm = np.array([rk.pair_index(c, a) for a, c in pairs])
    assert max_rel(full, full.transpose(1, 0, 2))[0] <= 1e-14
    # R(a,b,c,d) = R(c,b,a,d) = R(a,d,c,b) bit for bit: the window moments of
    # (a, c) and (c, a) are the same sums of one unordered local pair per
    # interval (the block fill takes factors through this symmetry)
    assert np.array_equal(full, full[perm])
    assert np.array_equal(full, full[:, perm])


def test_stage_times_and_repeat_determinism(ctx):
    w = WORKLOADS["c1"]
    rk = bs2e.RkTensor.build(w.k, knots(w), w.k_GL, w.max_k, ctx=ctx, compute=False)
    rk.set_profiling(True)
    rk.compute()
    ms = rk.stage_times()
    assert (ms > 0).all() and ms.sum() < 1000.0
    a = rk.to_host()
    rk.compute()
    b = rk.to_host()
    assert np.array_equal(a, b)       # bit-identical across builds


def test_invalid_arguments(ctx):
    w = WORKLOADS["tiny"]
    t = knots(w)
    with pytest.raises(bs2e.Bs2eError) as e:
        bs2e.RkTensor.build(21, t, w.k_GL, w.max_k, ctx=ctx)      # order > 20 (BVALUE_D limit)
    assert e.value.code == capi.BS2E_EINVAL
    with pytest.raises(bs2e.Bs2eError):
        bs2e.RkTensor.build(w.k, t[::-1].copy(), w.k_GL, w.max_k, ctx=ctx)
    h = C.c_void_p()
    assert capi.lib().bs2e_rk_create(ctx.handle, w.k, len(t), capi.ptr(t), w.k_GL, w.max_k, 99, C.byref(h)) \
        == capi.BS2E_EINVAL


def test_release_order_does_not_matter(oracle):
    """A context destroyed before its tensor, and a tensor released before its
    block, stay usable until the last dependent handle goes; afterwards new
    builds work normally."""
    w = WORKLOADS["tiny"]
    t = knots(w)
    n_b = oracle.dims(w.k, t)[1]
    c = bs2e.Context(0)
    rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=c)
    term = bs2e.sym(l=0, pi=False)
    term.configs = bs2e.count_configs(term, 1, n_b, w.k, n_b, n_b, 1)
    blk = bs2e.Block(rk, term)
    c.close()
    rk.release()
    S = oracle.setup_S(w.k, t, w.k_GL)
    Hl = [oracle.setup_H1p(w.k, t, w.k_GL, 2, l, cap_eta=0.0) for l in range(2)]
    H_sp, S_sp = blk.fill(Hl, S)
    assert H_sp.nnz > 0 and S_sp.nnz > 0
    blk.release()
    c2 = bs2e.Context(0)
    rk2 = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=c2)
    ref = oracle.rk_dense(w.k, n_b, w.max_k, oracle.slater_off_diag(w.k, t, w.k_GL, w.max_k),
                          oracle.slater_diag(w.k, t, w.k_GL, w.max_k))
    rel, _ = max_rel(rk2.to_host(), ref)
    assert rel <= TOL
    rk2.release()
    c2.close()


@pytest.mark.parametrize("name", ["c3", "c5p"])
def test_full_size_tensor_properties(name, ctx, oracle):
    """The two largest single-GPU configurations (c3: 2.69 GB; c5p: 20.5 GB,
    the order-9, 600-spline basis with its interior breakpoints moved by a
    seeded factor in [0.99, 1.01], one interval 1e4 times narrower than its
    neighbours): 42 sampled rows (the first and the last pair included)
    element by element against the oracle's rows (the same accumulation,
    restricted to the diagonal cells that land in them) within 1e-12; and,
    over the whole tensor, size-independent properties: every banded
    integral is positive (non-negative B-splines, positive weights and
    kernel); rows of the pairs (a, c) and (c, a) are equal bit for bit
    (R(a,b,c,d) = R(c,b,a,d)), as are the columns of (b, d) and (d, b)
    (R(a,b,c,d) = R(a,d,c,b)); and contracting the whole tensor with the
    hydrogenic orbitals of Z = 2 (from the oracle's one-electron matrices)
    gives the textbook Slater integrals F^0(1s,1s) = 5Z/8 ... F^2(2p,2p) =
    45Z/512 to 1e-9."""
    import scipy.linalg as sl
    w = WORKLOADS[name]
    t = knots(w)
    print(f"full-size properties {name}", flush=True)
    rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx)
    R = rk.to_host()
    n_pairs, K = rk.n_pairs, rk.K
    rk.release()
    n_b = oracle.dims(w.k, t)[1]
    pairs = oracle.pair_table(w.k, n_b)
    assert R.shape == (len(pairs), len(pairs), w.max_k + 1) == (n_pairs, n_pairs, K)
    rng = np.random.default_rng(11)
    sample = np.unique(np.concatenate([[0, n_pairs - 1], rng.choice(n_pairs, size=40, replace=False)]))
    mask = np.zeros(n_pairs, bool)
    mask[sample] = True
    ref = oracle.rk_rows(w.k, n_b, w.max_k, oracle.slater_off_diag(w.k, t, w.k_GL, w.max_k),
                         oracle.slater_diag(w.k, t, w.k_GL, w.max_k, pair_mask=mask), sample)
    rel, _ = max_rel(R[sample], ref)
    print(f"  {len(sample)} sampled rows vs oracle: max rel err {rel:.3e}", flush=True)
    assert rel <= TOL
    del ref
    assert bool((R > 0).all())
    index = {(int(a), int(c)): i for i, (a, c) in enumerate(pairs)}
    swap = np.array([index[(int(c), int(a))] for a, c in pairs])
    rows = rng.choice(n_pairs, size=64, replace=False)
    assert np.array_equal(R[rows], R[swap[rows]])
    assert np.array_equal(R[rows][:, swap], R[rows])
    Z = 2
    S = oracle.setup_S(w.k, t, w.k_GL).real
    orb = {}
    for l in (0, 1):
        E, V = sl.eigh(oracle.setup_H1p(w.k, t, w.k_GL, Z, l, cap_eta=0.0).real, S)
        for nprinc in (1, 2):
            i = nprinc - 1 - l
            if i >= 0:
                assert abs(E[i] + Z * Z / (2.0 * nprinc ** 2)) < 1e-9
                orb[(nprinc, l)] = V[:, i]
    flat = R.reshape(n_pairs, n_pairs * K)
    s1, s2, p2 = orb[(1, 0)], orb[(2, 0)], orb[(2, 1)]
    names = {id(s1): "1s", id(s2): "2s", id(p2): "2p"}
    rows_of = {}   # one pass over the tensor per electron-1 pair density

    def slater(k, a, b, c, d):
        key = (names[id(a)], names[id(c)])
        if key not in rows_of:
            A = a[pairs[:, 0] - 1] * c[pairs[:, 1] - 1]
            rows_of[key] = (A @ flat).reshape(n_pairs, K)
        B = b[pairs[:, 0] - 1] * d[pairs[:, 1] - 1]
        return float(rows_of[key][:, k] @ B)

    cases = [(slater(0, s1, s1, s1, s1), 5 * Z / 8), (slater(0, s1, s2, s1, s2), 17 * Z / 81),
             (slater(0, s1, s2, s2, s1), 16 * Z / 729), (slater(0, s2, s2, s2, s2), 77 * Z / 512),
             (slater(0, s1, p2, s1, p2), 59 * Z / 243), (slater(1, s1, p2, p2, s1), 112 * Z / 2187),
             (slater(0, p2, p2, p2, p2), 93 * Z / 512), (slater(2, p2, p2, p2, p2), 45 * Z / 512)]
    for got, exact in cases:
        assert abs(got - exact) < 1e-9 * exact, (got, exact)


# orders 2..12 and GL orders on both sides of 16: the matrix-core diagonal
# cells (order <= 9, k_GL <= 16) and the vector-ALU kernel (the rest)
# orders 13..16 take more than 64 KiB of LDS per workgroup (up to 160 KiB on gfx950)
SWEEP = [(2, 4, 0), (3, 5, 2), (5, 11, 4), (6, 17, 3), (9, 15, 6), (9, 18, 2), (10, 16, 5), (11, 17, 4),
         (12, 18, 3), (14, 20, 3), (16, 22, 2)]


@pytest.mark.parametrize("k,k_gl,max_k", SWEEP)
def test_order_and_gl_sweep_matches_oracle(k, k_gl, max_k, ctx, oracle):
    """Small bases (r <= 6, about 20-40 B-splines; r <= 4 from order 13) of
    every order the primitive kernels take, element by element against the
    oracle: primitives (P, Q, D containers) and the tensor."""
    t = bs2e.generate_grid(k, 3, 2, 0.5, 6.0 if k <= 12 else 4.0)
    n_b = oracle.dims(k, t)[1]
    print(f"sweep k={k} k_GL={k_gl} max_k={max_k} n_b={n_b}", flush=True)
    spl = bs2e.b_spline(k, t)
    r_k = bs2e.sparse_4d().init(max_k, spl)
    r_m_k = bs2e.sparse_4d().init(max_k, spl)
    r_d_k = bs2e.sparse_6d().init(max_k, spl)
    bs2e.setup_Slater_integrals(spl, max_k, k_gl, r_k, r_m_k, r_d_k, ctx=ctx)
    off = oracle.slater_off_diag(k, t, k_gl, max_k)
    dg = oracle.slater_diag(k, t, k_gl, max_k)
    for got, ref in ((r_k.data, off["rk"]), (r_m_k.data, off["rmk"]), (r_d_k.data, dg["rdk"])):
        rel, zp = max_rel(got, ref)
        assert rel <= TOL and zp == 0.0, (k, k_gl, rel, zp)
    rk = bs2e.RkTensor.build(k, t, k_gl, max_k, ctx=ctx)
    rel, zp = max_rel(rk.to_host(), oracle.rk_dense(k, n_b, max_k, off, dg))
    assert rel <= TOL and zp == 0.0, (k, k_gl, rel, zp)
    rk.release()


@pytest.mark.parametrize("k,max_k", [(2, 0), (2, 3), (3, 0), (3, 2), (4, 1), (4, 4)])
def test_smallest_bases_match_oracle(k, max_k, ctx, oracle):
    """The smallest bases the grid allows (orders 2-4, 7-9 B-splines, a few
    intervals), with one multipole (odd rows of the rank-one fill's 16-byte
    stores) and a few: every pair window touches an end of the basis.  The
    tensor element by element against the oracle, and its pair symmetries
    bit for bit."""
    t = bs2e.generate_grid(k, 1, 2, 1.0, 3.0)
    n_b = oracle.dims(k, t)[1]
    k_gl = k + 2
    print(f"smallest k={k} max_k={max_k} n_b={n_b}", flush=True)
    off = oracle.slater_off_diag(k, t, k_gl, max_k)
    dg = oracle.slater_diag(k, t, k_gl, max_k)
    rk = bs2e.RkTensor.build(k, t, k_gl, max_k, ctx=ctx)
    full = rk.to_host()
    perm = np.array([rk.pair_index(c, a) for a, c in oracle.pair_table(k, n_b)])
    rk.release()
    rel, zp = max_rel(full, oracle.rk_dense(k, n_b, max_k, off, dg))
    assert rel <= TOL and zp == 0.0, (k, max_k, rel, zp)
    assert np.array_equal(full, full[perm]) and np.array_equal(full, full[:, perm])
