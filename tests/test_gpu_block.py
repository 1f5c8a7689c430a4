"""GPU parity of the symmetry-block assembly (construct_block_tensor on the
device tensor) against the oracle's restatement of hamiltonian.f90:106-283.

Bar: index_ptr and indices bit-exact; S data bit-exact (the same products of
the same inputs in the same order, no fused multiply-adds); H data within
1e-12 of the block's largest entry (the R^k values under it differ from the
oracle's by ~1e-15 relative, and direct and exchange terms can cancel, so an
entry-wise relative bound would be ill-conditioned).
"""
import ctypes as C

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


@pytest.fixture(scope="module")
def c1(ctx, oracle):
    w = WORKLOADS["c1"]
    t = bs2e.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    n_b = oracle.dims(w.k, t)[1]
    rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx)
    off = oracle.slater_off_diag(w.k, t, w.k_GL, w.max_k)
    dg = oracle.slater_diag(w.k, t, w.k_GL, w.max_k)
    Rref = oracle.rk_dense(w.k, n_b, w.max_k, off, dg)
    S = oracle.setup_S(w.k, t, w.k_GL)
    Hl = [oracle.setup_H1p(w.k, t, w.k_GL, 2, l, cap_eta=0.0) for l in range(3)]
    return w, t, n_b, rk, Rref, S, Hl


def compare(got_H, got_S, ref):
    ipH, ixH, dH = ref["H"]
    ipS, ixS, dS = ref["S"]
    assert np.array_equal(got_H.index_ptr, ipH)
    assert np.array_equal(got_H.indices, ixH)
    assert np.array_equal(got_S.index_ptr, ipS)
    assert np.array_equal(got_S.indices, ixS)
    assert np.array_equal(got_S.data, dS)
    err = np.abs(got_H.data - dH).max() if dH.size else 0.0
    assert err <= TOL * np.abs(dH).max(), err


@pytest.mark.parametrize("L", [0, 1, 2])
@pytest.mark.parametrize("full", [True, False])
def test_block_matches_oracle(c1, oracle, L, full):
    w, t, n_b, rk, Rref, S, Hl = c1
    term = bs2e.sym(l=L, pi=bool(L % 2))
    term.configs = bs2e.count_configs(term, 2, n_b, w.k, n_b, n_b, 2)
    print(f"block c1 L={L} full={full} n_config={term.n_config}", flush=True)
    blk = bs2e.Block(rk, term, full)
    ref = oracle.construct_block(w.k, n_b, L, w.max_k, full, term.configs, Rref, Hl, S)
    assert (blk.nnz_H, blk.nnz_S) == ref["nnz"]
    assert (blk.ref_nnz_H, blk.ref_nnz_S) == ref["ref_count"]
    H_sp, S_sp = blk.fill(Hl, S)
    blk.release()
    compare(H_sp, S_sp, ref)


@pytest.mark.parametrize("L", [0, 1, 2])
def test_block_on_device_tensor(c1, oracle, L):
    """Fed the device's own tensor, the oracle's block assembly and the
    device's agree to 1e-14 of the block's largest entry: what remains are
    the angular factors (two independent 3j/6j evaluations, ~1 ulp) and, for
    rows with |n1 - n2| < k, the exchange term read as the tensor element
    (q, p) through R(a,b,d,c) = R(b,a,c,d).  The fill forms rank-one elements
    from the factor tables, with the p factors of a far-only row taken
    through the pair symmetry (a, c) <-> (c, a) (pinned bit for bit in
    test_gpu_rk)."""
    w, t, n_b, rk, Rref, S, Hl = c1
    term = bs2e.sym(l=L, pi=bool(L % 2))
    term.configs = bs2e.count_configs(term, 2, n_b, w.k, n_b, n_b, 2)
    print(f"block on device tensor c1 L={L} n_config={term.n_config}", flush=True)
    ref = oracle.construct_block(w.k, n_b, L, w.max_k, True, term.configs, rk.to_host(), Hl, S)
    blk = bs2e.Block(rk, term, True)
    H_sp, S_sp = blk.fill(Hl, S)
    blk.release()
    assert np.array_equal(H_sp.index_ptr, ref["H"][0]) and np.array_equal(H_sp.indices, ref["H"][1])
    assert np.array_equal(S_sp.data, ref["S"][2])
    err = np.abs(H_sp.data - ref["H"][2]).max()
    assert err <= 1e-14 * np.abs(ref["H"][2]).max(), err


@pytest.mark.parametrize("L", [0, 1])
def test_block_order_16_many_multipoles(ctx, oracle, L):
    """Order 16 with 34 multipoles (an even count): the fill's per-wavefront
    LDS (coefficients and the factor runs of 31 pairs) passes 64 KiB per
    workgroup, and rows with |n1 - n2| >= 16 take the staged path."""
    k, k_gl, max_k, max_l = 16, 20, 33, 1
    t = bs2e.generate_grid(k, 3, 2, 0.5, 4.0)
    n_b = oracle.dims(k, t)[1]
    print(f"block order 16 K=34 L={L} n_b={n_b}", flush=True)
    rk = bs2e.RkTensor.build(k, t, k_gl, max_k, ctx=ctx)
    Rref = oracle.rk_dense(k, n_b, max_k, oracle.slater_off_diag(k, t, k_gl, max_k),
                           oracle.slater_diag(k, t, k_gl, max_k))
    S = oracle.setup_S(k, t, k_gl)
    Hl = [oracle.setup_H1p(k, t, k_gl, 2, l, cap_eta=0.0) for l in range(max_l + 1)]
    term = bs2e.sym(l=L, pi=bool(L % 2))
    term.configs = bs2e.count_configs(term, max_l, n_b, k, n_b, n_b, max_l)
    assert (np.abs(term.configs[:, 0] - term.configs[:, 1]) >= k).any()
    blk = bs2e.Block(rk, term, True)
    H_sp, S_sp = blk.fill(Hl, S)
    blk.release()
    rk.release()
    compare(H_sp, S_sp, oracle.construct_block(k, n_b, L, max_k, True, term.configs, Rref, Hl, S))


@pytest.mark.parametrize("k,k_gl,max_k,L,full", [(9, 12, 2, 0, True), (12, 14, 3, 1, False), (7, 10, 2, 1, True)])
def test_block_separated_rows(ctx, oracle, k, k_gl, max_k, L, full):
    """Rows whose electrons sit at least 2k - 1 splines apart take the
    matrix-core fill: one 16 x 16 tile per window side for order 7, one and
    the 17th row and column as dot products for order 9, two for order 12
    (2k - 1 > 17).  A basis long enough to hold such rows;
    patterns bit-exact and values against the oracle, with the rows of both
    kinds (separated and not) present."""
    t = bs2e.generate_grid(k, 3, 2, 0.8, 40.0)
    n_b = oracle.dims(k, t)[1]
    max_l = 1
    print(f"block separated rows k={k} n_b={n_b} max_k={max_k} L={L} full={full}", flush=True)
    rk = bs2e.RkTensor.build(k, t, k_gl, max_k, ctx=ctx)
    Rref = oracle.rk_dense(k, n_b, max_k, oracle.slater_off_diag(k, t, k_gl, max_k),
                           oracle.slater_diag(k, t, k_gl, max_k))
    S = oracle.setup_S(k, t, k_gl)
    Hl = [oracle.setup_H1p(k, t, k_gl, 2, l, cap_eta=0.0) for l in range(max_l + 1)]
    term = bs2e.sym(l=L, pi=bool(L % 2))
    term.configs = bs2e.count_configs(term, max_l, n_b, k, n_b, n_b, max_l)
    gap = np.abs(term.configs[:, 0] - term.configs[:, 1])
    assert (gap >= 2 * k - 1).sum() > 100 and (gap < 2 * k - 1).any()
    ref = oracle.construct_block(k, n_b, L, max_k, full, term.configs, Rref, Hl, S)
    blk = bs2e.Block(rk, term, full)
    assert (blk.nnz_H, blk.nnz_S) == ref["nnz"] and (blk.ref_nnz_H, blk.ref_nnz_S) == ref["ref_count"]
    H_sp, S_sp = blk.fill(Hl, S)
    blk.release()
    rk.release()
    compare(H_sp, S_sp, ref)


@pytest.mark.parametrize("k,max_k", [(5, 11), (6, 12), (8, 15), (7, 17)])
def test_block_tiled_rows_multipole_counts(ctx, oracle, k, max_k):
    """The rows with |n1 - n2| < k are filled in tiles that hold 12 or 18
    multipoles per cell: K = max_k + 1 = 12 and 18 fill them exactly, 13 and
    16 leave 5 and 2 slots past the last multipole (their loads repeat the
    last one and meet zero coefficients).  Patterns bit-exact and values
    against the oracle, both triangles and both parities."""
    t = bs2e.generate_grid(k, 3, 2, 1.0, 14.0)
    n_b = oracle.dims(k, t)[1]
    max_l = 2
    print(f"block tiled rows k={k} n_b={n_b} max_k={max_k}", flush=True)
    rk = bs2e.RkTensor.build(k, t, k + 4, max_k, ctx=ctx)
    Rref = oracle.rk_dense(k, n_b, max_k, oracle.slater_off_diag(k, t, k + 4, max_k),
                           oracle.slater_diag(k, t, k + 4, max_k))
    S = oracle.setup_S(k, t, k + 4)
    Hl = [oracle.setup_H1p(k, t, k + 4, 2, l, cap_eta=0.0) for l in range(max_l + 1)]
    for L, pi, full in ((1, True, True), (2, False, False), (1, False, True)):
        term = bs2e.sym(l=L, pi=pi)
        term.configs = bs2e.count_configs(term, max_l, n_b, k, n_b, n_b, max_l)
        if term.n_config == 0:
            continue
        assert (np.abs(term.configs[:, 0] - term.configs[:, 1]) < k).any()
        ref = oracle.construct_block(k, n_b, L, max_k, full, term.configs, Rref, Hl, S)
        blk = bs2e.Block(rk, term, full)
        H_sp, S_sp = blk.fill(Hl, S)
        blk.release()
        compare(H_sp, S_sp, ref)
    rk.release()


RANDOM_BASES = [   # seed, order, k_GL, max_l, L, full
    (1, 2, 5, 1, 0, True), (2, 3, 6, 2, 1, False), (3, 4, 7, 3, 2, True), (4, 5, 8, 2, 3, False),
    (5, 6, 9, 1, 1, True), (6, 9, 12, 2, 0, False), (7, 11, 14, 1, 2, True), (8, 12, 15, 2, 1, False),
]


@pytest.mark.parametrize("seed,k,k_gl,max_l,L,full", RANDOM_BASES)
def test_block_random_small_bases(ctx, oracle, seed, k, k_gl, max_l, L, full):
    """Small bases of many orders with drawn box sizes, multipole counts and
    symmetries: patterns (so the closed-form counts, row by row) bit-exact and
    values against the oracle, through both parities."""
    rng = np.random.default_rng(seed)
    h_max, r_max = float(rng.uniform(0.6, 1.6)), float(rng.uniform(6.0, 14.0))
    max_k = int(rng.integers(1, 2 * max_l + 2))
    t = bs2e.generate_grid(k, 3, 2, h_max, r_max)
    n_b = oracle.dims(k, t)[1]
    print(f"block random seed={seed} k={k} n_b={n_b} max_k={max_k} L={L} full={full}", flush=True)
    rk = bs2e.RkTensor.build(k, t, k_gl, max_k, ctx=ctx)
    Rref = oracle.rk_dense(k, n_b, max_k, oracle.slater_off_diag(k, t, k_gl, max_k),
                           oracle.slater_diag(k, t, k_gl, max_k))
    S = oracle.setup_S(k, t, k_gl)
    Hl = [oracle.setup_H1p(k, t, k_gl, 2, l, cap_eta=0.0) for l in range(max_l + 1)]
    for pi in (bool(L % 2), not bool(L % 2)):
        term = bs2e.sym(l=L, pi=pi)
        term.configs = bs2e.count_configs(term, max_l, n_b, k, n_b, n_b, max_l)
        if term.n_config == 0:
            continue
        ref = oracle.construct_block(k, n_b, L, max_k, full, term.configs, Rref, Hl, S)
        blk = bs2e.Block(rk, term, full)
        assert (blk.nnz_H, blk.nnz_S) == ref["nnz"] and (blk.ref_nnz_H, blk.ref_nnz_S) == ref["ref_count"]
        H_sp, S_sp = blk.fill(Hl, S)
        blk.release()
        compare(H_sp, S_sp, ref)
    rk.release()


@pytest.mark.parametrize("k,max_k", [(2, 0), (3, 1), (4, 0), (4, 3)])
def test_block_tiny_bases(ctx, oracle, k, max_k):
    """Bases of 7-9 B-splines (every row and column touches a boundary of
    the basis) and one or a few multipoles (K = 1 included): patterns and
    values against the oracle for L = 0 and 1, full and upper triangle."""
    t = bs2e.generate_grid(k, 1, 2, 1.0, 3.0)
    n_b = oracle.dims(k, t)[1]
    k_gl, max_l = k + 3, 1
    rk = bs2e.RkTensor.build(k, t, k_gl, max_k, ctx=ctx)
    Rref = oracle.rk_dense(k, n_b, max_k, oracle.slater_off_diag(k, t, k_gl, max_k),
                           oracle.slater_diag(k, t, k_gl, max_k))
    S = oracle.setup_S(k, t, k_gl)
    Hl = [oracle.setup_H1p(k, t, k_gl, 2, l, cap_eta=0.0) for l in range(max_l + 1)]
    print(f"block tiny k={k} n_b={n_b} max_k={max_k}", flush=True)
    for L, full in ((0, True), (1, False)):
        term = bs2e.sym(l=L, pi=bool(L % 2))
        term.configs = bs2e.count_configs(term, max_l, n_b, k, n_b, n_b, max_l)
        ref = oracle.construct_block(k, n_b, L, max_k, full, term.configs, Rref, Hl, S)
        blk = bs2e.Block(rk, term, full)
        assert (blk.nnz_H, blk.nnz_S) == ref["nnz"] and (blk.ref_nnz_H, blk.ref_nnz_S) == ref["ref_count"]
        H_sp, S_sp = blk.fill(Hl, S)
        blk.release()
        compare(H_sp, S_sp, ref)
    rk.release()


def test_block_truncated_configs_and_cap(c1, oracle):
    """Truncated configuration lists (max_n_b, n_all_l, l_2_max) and a complex
    absorbing potential in the one-electron blocks."""
    w, t, n_b, rk, Rref, S, _ = c1
    Hc = [oracle.setup_H1p(w.k, t, w.k_GL, 2, l, cap_order=2, cap_r0=30.0, cap_eta=0.02 + 0.005j) for l in range(4)]
    term = bs2e.sym(l=1, pi=True)
    term.configs = bs2e.count_configs(term, 3, n_b, w.k, 40, 30, 1)
    print(f"block c1 truncated n_config={term.n_config}", flush=True)
    H_sp, S_sp = bs2e.construct_block_tensor(Hc, S, bs2e.b_spline(w.k, t), term, w.max_k, rk, full=True)
    ref = oracle.construct_block(w.k, n_b, 1, w.max_k, True, term.configs, Rref, Hc, S)
    compare(H_sp, S_sp, ref)
    assert np.abs(H_sp.data.imag).max() > 0.0


def test_helium_ground_state_on_device_blocks(c1):
    """The device tensor and block give He 1s^2 1S in the spd limit
    (-2.902767, Carroll, Silverstone and Metzger 1979)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    w, t, n_b, rk, _, S, Hl = c1
    term = bs2e.sym(l=0, pi=False)
    term.configs = bs2e.count_configs(term, 2, n_b, w.k, n_b, n_b, 2)
    H_sp, S_sp = bs2e.construct_block_tensor(Hl, S, bs2e.b_spline(w.k, t), term, w.max_k, rk)
    n = term.n_config
    H = sp.csr_matrix((H_sp.data.real, H_sp.indices - 1, H_sp.index_ptr - 1), shape=(n, n))
    M = sp.csr_matrix((S_sp.data.real, S_sp.indices - 1, S_sp.index_ptr - 1), shape=(n, n))
    E = sla.eigsh(H.tocsc(), k=1, M=M.tocsc(), sigma=-3.0, which="LM", return_eigenvectors=False)
    assert abs(E[0] + 2.902767) < 2e-6


def test_block_errors(c1, ctx):
    w, t, n_b, rk, _, S, Hl = c1
    term = bs2e.sym(l=0, pi=False)
    cf = bs2e.count_configs(term, 2, n_b, w.k, n_b, n_b, 2)
    # channels must be contiguous runs and n2 consecutive (count_configs order)
    bad = cf.copy()
    bad[[0, -1]] = bad[[-1, 0]]
    term.configs = bad
    with pytest.raises(capi.Bs2eError) as e:
        bs2e.Block(rk, term)
    assert e.value.code == capi.BS2E_EINVAL
    term.configs = cf
    blk = bs2e.Block(rk, term)
    with pytest.raises(capi.Bs2eError):
        blk.fill(Hl[:2], S)        # l = 2 channel without its one-electron block
    blk.release()
    # a tensor that has not been computed
    fresh = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx, compute=False)
    with pytest.raises(capi.Bs2eError) as e:
        bs2e.Block(fresh, term)
    assert e.value.code == capi.BS2E_ESTATE
    fresh.release()


@pytest.mark.parametrize("full", [True, False])
def test_blockset_equals_single_blocks_and_oracle(c1, oracle, full):
    """The symmetries L = 0..2 (and a truncated L = 3 list) in one count and
    one fill launch: every symmetry's CSR arrays equal the single-block path
    bit for bit and the oracle to the block bar; sizes report per symmetry."""
    w, t, n_b, rk, Rref, S, Hl = c1
    terms = []
    for L in range(3):
        term = bs2e.sym(l=L, pi=bool(L % 2))
        term.configs = bs2e.count_configs(term, 2, n_b, w.k, n_b, n_b, 2)
        terms.append(term)
    t3 = bs2e.sym(l=3, pi=True)
    t3.configs = bs2e.count_configs(t3, 2, n_b, w.k, 40, 30, 1)
    terms.append(t3)
    print("blockset c1 L=0..3 full", full, [t.n_config for t in terms], flush=True)
    bset = bs2e.BlockSet(rk, terms, full)
    bset.fill(Hl, S)
    ms = bset.kernel_ms()
    assert ms[0] > 0 and ms[1] > 0
    for i, term in enumerate(terms):
        H_sp, S_sp = bset.get(i)
        blk = bs2e.Block(rk, term, full)
        assert (blk.nnz_H, blk.nnz_S, blk.ref_nnz_H, blk.ref_nnz_S) == tuple(int(x) for x in bset.sizes[i, :4])
        H1, S1 = blk.fill(Hl, S)
        blk.release()
        for a, b in ((H_sp, H1), (S_sp, S1)):
            assert np.array_equal(a.index_ptr, b.index_ptr) and np.array_equal(a.indices, b.indices)
            assert np.array_equal(a.data, b.data)
        if i in (0, 3):
            ref = oracle.construct_block(w.k, n_b, term.l, w.max_k, full, term.configs, Rref, Hl, S)
            compare(H_sp, S_sp, ref)
    with pytest.raises(capi.Bs2eError) as e:
        bset.get(len(terms))
    assert e.value.code == capi.BS2E_ERANGE
    one = np.zeros(1, np.int64)
    cd = np.zeros(2)
    I, D = C.POINTER(C.c_int64), C.POINTER(C.c_double)
    args = [one.ctypes.data_as(I), one.ctypes.data_as(I), cd.ctypes.data_as(D)] * 2
    for bad in (0, len(terms) + 1):   # the C ABI counts symmetries from 1
        assert capi.lib().bs2e_blockset_get(bset.handle, bad, *args) == capi.BS2E_ERANGE
    bset.release()


def test_blockset_uneven_channels_t0(oracle, ctx):
    """The reference test's basis (order 8, r <= 15) with l <= 3 and every
    symmetry L = 0..4 of both parities in one launch: ten blocks of very
    different channel counts, each equal to the oracle's."""
    w = WORKLOADS["t0"]
    t = bs2e.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    n_b = oracle.dims(w.k, t)[1]
    max_l = 3
    rk = bs2e.RkTensor.build(w.k, t, w.k_GL, 2 * max_l, ctx=ctx)
    Rref = oracle.rk_dense(w.k, n_b, 2 * max_l, oracle.slater_off_diag(w.k, t, w.k_GL, 2 * max_l),
                           oracle.slater_diag(w.k, t, w.k_GL, 2 * max_l))
    S = oracle.setup_S(w.k, t, w.k_GL)
    Hl = [oracle.setup_H1p(w.k, t, w.k_GL, 2, l, cap_eta=0.0) for l in range(max_l + 1)]
    terms = []
    for L in range(5):
        for pi in (False, True):
            term = bs2e.sym(l=L, pi=pi)
            term.configs = bs2e.count_configs(term, max_l, n_b, w.k, n_b, n_b, max_l)
            if term.n_config:
                terms.append(term)
    print("blockset t0 terms", [(t_.l, t_.pi, t_.n_config) for t_ in terms], flush=True)
    blocks = bs2e.construct_blocks_tensor(Hl, S, bs2e.b_spline(w.k, t), terms, 2 * max_l, rk)
    for term, (H_sp, S_sp) in zip(terms, blocks):
        ref = oracle.construct_block(w.k, n_b, term.l, 2 * max_l, True, term.configs, Rref, Hl, S)
        compare(H_sp, S_sp, ref)
    rk.release()


def _dense(csr, n):
    import scipy.sparse as sp
    return sp.csr_matrix((csr.data, csr.indices - 1, csr.index_ptr - 1), shape=(n, n)).toarray()


@pytest.mark.parametrize("L,pi", [(0, False), (1, True), (2, False)])
def test_eigenvalues_device_blocks_match_oracle_blocks(ctx, oracle, L, pi):
    """He (Z = 2) on the c1 basis (order 7, 60 B-splines, r <= 50, GL 10,
    l1, l2 <= 2, k <= 4, no absorbing potential): the lowest six eigenvalues
    of H c = E S c for the block built entirely on the device (device tensor,
    device one-electron matrices, device channel sum) and for the block built
    entirely by the oracle (its tensor, its one-electron matrices, its
    construct_block), by the same dense solver on the CPU, agree within
    1e-9 Ha.  L = 0 holds 1s^2 1S, L = 1 odd 1s2p 1P, L = 2 even 1s3d 1D
    (the reference's diag solves these blocks, diagonalization.f90:86-118)."""
    import scipy.linalg as sl
    w = WORKLOADS["c1"]
    t = bs2e.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    spl = bs2e.b_spline(w.k, t)
    n_b = spl.n_b
    term = bs2e.sym(l=L, pi=pi)
    term.configs = bs2e.count_configs(term, 2, n_b, w.k, n_b, n_b, 2)
    n = term.n_config
    print(f"eigenvalue gate c1 L={L} n_config={n}", flush=True)
    rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx)
    S_d = bs2e.setup_S(spl, w.k_GL, ctx=ctx)
    H_d = bs2e.setup_H_one_particle(w.Z, (2, 0.8 * w.r_max, 0.0), range(3), spl, w.k_GL, ctx=ctx)
    H_sp, S_sp = bs2e.construct_block_tensor(list(H_d), S_d, spl, term, w.max_k, rk)
    rk.release()
    Rref = oracle.rk_dense(w.k, n_b, w.max_k, oracle.slater_off_diag(w.k, t, w.k_GL, w.max_k),
                           oracle.slater_diag(w.k, t, w.k_GL, w.max_k))
    S_o = oracle.setup_S(w.k, t, w.k_GL)
    H_o = [oracle.setup_H1p(w.k, t, w.k_GL, w.Z, l, cap_order=2, cap_r0=0.8 * w.r_max, cap_eta=0.0)
           for l in range(3)]
    ref = oracle.construct_block(w.k, n_b, L, w.max_k, True, term.configs, Rref, H_o, S_o)
    E = []
    for Hc, Sc in ((H_sp, S_sp), (bs2e.CSR_matrix((n, n), *ref["H"]), bs2e.CSR_matrix((n, n), *ref["S"]))):
        Hm, Sm = _dense(Hc, n), _dense(Sc, n)
        assert np.abs(Hm.imag).max() == 0 and np.abs(Sm.imag).max() == 0
        E.append(sl.eigh(Hm.real, Sm.real, subset_by_index=[0, 5], eigvals_only=True))
    print(f"  device {E[0]}\n  oracle {E[1]}  max |dE| {np.abs(E[0] - E[1]).max():.3e}", flush=True)
    assert np.abs(E[0] - E[1]).max() <= 1e-9
    lowest = {0: -2.9028, 1: -2.1238, 2: -2.0556}[L]   # spd-limit He levels, Ha
    assert abs(E[0][0] - lowest) < 2e-3


def test_c3_blockset_sampled_rows(ctx, oracle):
    """The c3 basis (order 8, 300 B-splines, r <= 300, GL 12, k <= 16) with
    every symmetry L = 0..4 (natural parity) of l1, l2 <= 4 in one count and
    one fill launch (1.86 M configuration rows, 3.5e9 H entries, every
    symmetry's channels of uneven size): per symmetry, windows of rows at the
    start, the middle and the end against the oracle's rows of the same block
    (its own tensor rows, the same one-electron matrices): patterns and S bit
    for bit, H within 1e-12 of the window's largest entry; the square part of
    each window is symmetric (full = true), and the per-row counts sum to the
    sizes the set reports."""
    w = WORKLOADS["c3"]
    t = bs2e.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    n_b = oracle.dims(w.k, t)[1]
    max_l, W = 4, 24
    rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx)
    S = oracle.setup_S(w.k, t, w.k_GL)
    Hl = [oracle.setup_H1p(w.k, t, w.k_GL, w.Z, l, cap_order=2, cap_r0=0.8 * w.r_max, cap_eta=0.0)
          for l in range(max_l + 1)]
    terms = []
    for L in range(5):
        term = bs2e.sym(l=L, pi=bool(L % 2))
        term.configs = bs2e.count_configs(term, max_l, n_b, w.k, n_b, n_b, max_l)
        terms.append(term)
    print("c3 blockset rows", [t_.n_config for t_ in terms], flush=True)
    bset = bs2e.BlockSet(rk, terms, True)
    bset.fill(Hl, S)
    windows = []
    for i, term in enumerate(terms):
        n = term.n_config
        for lo in (0, n // 2 - W // 2, n - W):
            windows.append((i, lo))
    got = {}
    for i, lo in windows:
        got[(i, lo)] = bset.get_rows(i, lo, W)
        H_sp, S_sp = got[(i, lo)]
        assert H_sp.index_ptr[0] == 1 and H_sp.nnz == H_sp.index_ptr[-1] - 1
    # every row's counts: the windows' offsets agree with the set's sizes
    for i, term in enumerate(terms):
        ipH, ipS = bset.row_offsets(i, 0, term.n_config)
        assert (ipH[-1] - 1, ipS[-1] - 1) == tuple(int(x) for x in bset.sizes[i, :2])
        lo = term.n_config // 2 - W // 2
        a, b = bset.row_offsets(i, lo, W)
        assert np.array_equal(a - 1, ipH[lo:lo + W + 1] - ipH[lo]) and np.array_equal(b - 1, ipS[lo:lo + W + 1] - ipS[lo])
    bset.release()
    rk.release()
    # the oracle's rows: tensor rows pair(n_a, .) of every window's configurations
    need = np.unique(np.concatenate([oracle.block_tensor_rows(w.k, n_b, terms[i].configs, lo, lo + W)
                                     for i, lo in windows]))
    npairs = oracle.n_pairs(w.k, n_b)
    mask = np.zeros(npairs, bool)
    mask[need] = True
    Rr = oracle.rk_rows(w.k, n_b, w.max_k, oracle.slater_off_diag(w.k, t, w.k_GL, w.max_k),
                        oracle.slater_diag(w.k, t, w.k_GL, w.max_k, pair_mask=mask), need)
    rmap = np.full(npairs, -1, np.int64)
    rmap[need] = np.arange(len(need))
    worst = 0.0
    for i, lo in windows:
        term = terms[i]
        ref = oracle.construct_block_rows(w.k, n_b, term.l, w.max_k, True, term.configs, Rr, rmap, lo, lo + W, Hl, S)
        H_sp, S_sp = got[(i, lo)]
        compare(H_sp, S_sp, ref)
        dH = ref["H"][2]
        worst = max(worst, float(np.abs(H_sp.data - dH).max() / np.abs(dH).max()))
        # the window's square part (columns lo+1..lo+W) is symmetric
        Hd = np.zeros((W, W), np.complex128)
        for r in range(W):
            a, b = H_sp.index_ptr[r] - 1, H_sp.index_ptr[r + 1] - 1
            cols, vals = H_sp.indices[a:b], H_sp.data[a:b]
            sel = (cols > lo) & (cols <= lo + W)
            Hd[r, cols[sel] - lo - 1] = vals[sel]
        assert np.abs(Hd - Hd.T).max() <= TOL * np.abs(dH).max()
    print(f"  {len(windows)} windows of {W} rows: H max err / scale {worst:.3e}", flush=True)


def _global_window_rows(bset, lo, n):
    """(H_sp, S_sp) pieces of global rows lo .. lo+n-1, one per symmetry the
    window touches: [(symmetry, first row within it, H_sp, S_sp)]."""
    r0 = bset.row0
    out = []
    for i in range(len(bset.terms)):
        a, b = max(lo, int(r0[i])), min(lo + n, int(r0[i + 1]))
        if a < b:
            out.append((i, a - int(r0[i])) + tuple(bset.get_rows(i, a - int(r0[i]), b - a)))
    return out


@pytest.mark.parametrize("full", [True, False])
def test_row_windows_equal_whole_fill(c1, full):
    """Windows of global rows (one row, windows across symmetry boundaries,
    the whole set, drawn sizes) filled one at a time into outputs that hold
    only the window give the whole set's fill bit for bit; refill() fills the
    window again; reading outside the window, or a symmetry it does not hold,
    is refused, as is a graph of a window."""
    w, t, n_b, rk, Rref, S, Hl = c1
    terms = []
    for L in range(3):
        term = bs2e.sym(l=L, pi=bool(L % 2))
        term.configs = bs2e.count_configs(term, 2, n_b, w.k, n_b, n_b, 2)
        terms.append(term)
    bset = bs2e.BlockSet(rk, terms, full)
    N = int(bset.row0[-1])
    with pytest.raises(bs2e.Bs2eError):   # no one-electron matrices on the device yet
        bset.fill_rows(0, 10)
    bset.fill(Hl, S)
    whole = [bset.get(i) for i in range(len(terms))]
    cH, cS = bset.row_counts()
    assert int(cH.sum()) == int(bset.sizes[:, 0].sum()) and int(cS.sum()) == int(bset.sizes[:, 1].sum())
    rng = np.random.default_rng(11)
    r0 = bset.row0
    wins = [(0, 1), (N - 1, 1), (int(r0[1]) - 5, 10), (int(r0[2]) - 100, 300), (0, N)]
    wins += [(int(a), int(rng.integers(1, N - a + 1))) for a in rng.integers(0, N, 4)]
    for k, (lo, n) in enumerate(wins):
        print(f"window fill c1 full={full} rows {lo}..{lo + n - 1}", flush=True)
        if k == 1:
            bset.fill_rows(lo, n, Hl, S)   # with host matrices
        else:
            bset.fill_rows(lo, n)          # the last fill's device copy
        for rep in range(2):
            for i, a, H_sp, S_sp in _global_window_rows(bset, lo, n):
                Hw, Sw = whole[i]
                m = H_sp.index_ptr.size - 1
                pa, pb = Hw.index_ptr[a] - 1, Hw.index_ptr[a + m] - 1
                qa, qb = Sw.index_ptr[a] - 1, Sw.index_ptr[a + m] - 1
                assert np.array_equal(H_sp.index_ptr - 1, Hw.index_ptr[a:a + m + 1] - 1 - pa)
                assert np.array_equal(S_sp.index_ptr - 1, Sw.index_ptr[a:a + m + 1] - 1 - qa)
                assert np.array_equal(H_sp.indices, Hw.indices[pa:pb])
                assert np.array_equal(H_sp.data, Hw.data[pa:pb])
                assert np.array_equal(S_sp.indices, Sw.indices[qa:qb])
                assert np.array_equal(S_sp.data, Sw.data[qa:qb])
            bset.refill()
    lo, n = int(r0[1]) - 5, 10   # a window across the boundary of symmetries 0 and 1
    bset.fill_rows(lo, n)
    with pytest.raises(bs2e.Bs2eError):
        bset.get_rows(0, 0, 3)              # rows outside the window
    with pytest.raises(bs2e.Bs2eError):
        bset.get(1)                         # a symmetry the window holds only in part
    with pytest.raises(bs2e.Bs2eError):
        bset.set_graph(True)                # graphs cover the whole set
    with pytest.raises(bs2e.Bs2eError):
        bset.fill_rows(N - 2, 3)            # rows outside the set
    bset.fill(Hl, S)                        # the whole set again
    for i in range(len(terms)):
        H_sp, S_sp = bset.get(i)
        assert np.array_equal(H_sp.data, whole[i][0].data) and np.array_equal(S_sp.indices, whole[i][1].indices)
    bset.release()


def test_c3_stated_size_streamed_in_windows(ctx, oracle):
    """The c3 channel batch at its stated size: the c3 basis with lmax 8 and
    every symmetry L = 0..4 (natural parity), all (l1, l2) channels: 4.47 M
    configuration rows whose CSR blocks (about 5e11 bytes) exceed the device's
    memory.  The set is counted in one launch and streamed through windows of
    at most 24 GB of entries; from every window two drawn runs of 24 rows are
    read back, and rows at the start, middle and end of every symmetry are
    filled as windows of their own.  All of them against the oracle's rows of
    the same blocks: patterns and S bit for bit, H within 1e-12 of the run's
    largest entry; the windows' counts sum to the set's sizes."""
    w = WORKLOADS["c3"]
    t = bs2e.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    n_b = oracle.dims(w.k, t)[1]
    max_l, W = 8, 24
    rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx)
    S = oracle.setup_S(w.k, t, w.k_GL)
    Hl = [oracle.setup_H1p(w.k, t, w.k_GL, w.Z, l, cap_order=2, cap_r0=0.8 * w.r_max, cap_eta=0.0)
          for l in range(max_l + 1)]
    terms = []
    for L in range(5):
        term = bs2e.sym(l=L, pi=bool(L % 2))
        term.configs = bs2e.count_configs(term, max_l, n_b, w.k, n_b, n_b, max_l)
        terms.append(term)
    bset = bs2e.BlockSet(rk, terms, True)
    total = 24 * int(bset.sizes[:, :2].sum())
    wins = bset.windows(24 << 30)
    print(f"c3 stated size: rows {[t_.n_config for t_ in terms]}, {total / 1e9:.1f} GB of CSR, "
          f"{len(wins)} windows", flush=True)
    assert total > 288e9   # larger than the device
    assert sum(n for _, n in wins) == int(bset.row0[-1])
    rng = np.random.default_rng(5)
    got = {}
    r0 = bset.row0
    for k, (lo, n) in enumerate(bset.stream(24 << 30, Hl, S)):
        print(f"  window {k}: rows {lo}..{lo + n - 1}", flush=True)
        assert (lo, n) == wins[k]
        for g in rng.integers(lo, lo + n - W, 2):
            i = int(np.searchsorted(r0, g, side="right") - 1)
            a = int(g - r0[i])
            if a + W <= terms[i].n_config:
                got[(i, a)] = bset.get_rows(i, a, W)
    for i, term in enumerate(terms):
        n = term.n_config
        for a in (0, n // 2 - W // 2, n - W):
            bset.fill_rows(int(r0[i]) + a, W)
            got[(i, a)] = bset.get_rows(i, a, W)
    bset.release()
    rk.release()
    print(f"  {len(got)} runs of {W} rows read back; oracle rows", flush=True)
    need = np.unique(np.concatenate([oracle.block_tensor_rows(w.k, n_b, terms[i].configs, a, a + W)
                                     for i, a in got]))
    npairs = oracle.n_pairs(w.k, n_b)
    mask = np.zeros(npairs, bool)
    mask[need] = True
    Rr = oracle.rk_rows(w.k, n_b, w.max_k, oracle.slater_off_diag(w.k, t, w.k_GL, w.max_k),
                        oracle.slater_diag(w.k, t, w.k_GL, w.max_k, pair_mask=mask), need)
    rmap = np.full(npairs, -1, np.int64)
    rmap[need] = np.arange(len(need))
    worst = 0.0
    for (i, a), (H_sp, S_sp) in got.items():
        term = terms[i]
        ref = oracle.construct_block_rows(w.k, n_b, term.l, w.max_k, True, term.configs, Rr, rmap, a, a + W, Hl, S)
        compare(H_sp, S_sp, ref)
        dH = ref["H"][2]
        worst = max(worst, float(np.abs(H_sp.data - dH).max() / np.abs(dH).max()))
    print(f"  H max err / scale {worst:.3e}", flush=True)


def test_branch_graph_replays_match_eager(ctx, oracle):
    """The whole branch of a set (tensor build, count, device row offsets,
    layouts, fill over L = 0..2) captured as one hipGraph: every replay, with
    eager builds and synchronisations between them, gives the eager tensor and
    every symmetry's CSR arrays bit for bit; refill() from the device copy of
    the one-electron matrices equals the first fill; a new fill drops the
    graph."""
    w = WORKLOADS["c1"]
    t = bs2e.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    n_b = oracle.dims(w.k, t)[1]
    S = oracle.setup_S(w.k, t, w.k_GL)
    Hl = [oracle.setup_H1p(w.k, t, w.k_GL, 2, l, cap_eta=1e-3) for l in range(3)]
    rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx)
    terms = []
    for L in range(3):
        term = bs2e.sym(l=L, pi=bool(L % 2))
        term.configs = bs2e.count_configs(term, 2, n_b, w.k, n_b, n_b, 2)
        terms.append(term)
    bset = bs2e.BlockSet(rk, terms, False)
    bset.fill(Hl, S)
    R0 = rk.to_host()
    ref = [bset.get(i) for i in range(len(terms))]

    def same():
        for i, (H0, S0) in enumerate(ref):
            H1, S1 = bset.get(i)
            for a, b in ((H0, H1), (S0, S1)):
                assert np.array_equal(a.index_ptr, b.index_ptr) and np.array_equal(a.indices, b.indices)
                assert np.array_equal(a.data, b.data)

    bset.refill()
    same()
    bset.set_graph(True)
    other = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx, compute=False)
    for rep in range(3):
        print(f"branch graph replay {rep}", flush=True)
        bset.replay()
        bset.synchronize()
        assert np.array_equal(rk.to_host(), R0), rep
        same()
        other.compute()               # eager work between replays
        other.synchronize()
    bset.fill(Hl, S)                  # a new fill drops the graph
    same()
    with pytest.raises(capi.Bs2eError) as e:
        bset.replay()
    assert e.value.code == capi.BS2E_ESTATE
    other.release()
    bset.release()
    rk.release()
