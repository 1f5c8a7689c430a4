"""Pins of the oracle by what the mathematics fixes.

* Gauss-Legendre: numpy's leggauss and exactness on polynomials.
* B-splines: scipy's BSpline, partition of unity, non-negativity.
* P^k: exact for polynomial integrands (degree 2(k-1)+k <= 2 k_GL - 1), so
  it must equal the analytic integral of the piecewise polynomial.
* Q^k, D^k: an independent numpy/scipy evaluation of the same rule (scipy
  spline values, vectorised sums in a different order); the triangle rule's
  diagonal cells summed both ways must equal an adaptive 2-D integral.
* R^k: 8-fold symmetry, positivity, every key of count_nnz_R_k set, and the
  textbook hydrogenic Slater integrals of Z=2 orbitals.
* Angular algebra: sympy's exact 3j/6j.
* One-electron Hamiltonian: hydrogenic levels -Z^2 / (2 n^2).
* Symmetry blocks (construct_block_tensor): the helium ground state in the
  s-, sp- and spd-partial-wave limits and the 1s2p 1P / 1s3d 1D levels, the
  symmetry of H and S, and the upper-triangle (full = .false.) layout.
"""
import numpy as np
import pytest
import scipy.linalg as sl
from scipy.interpolate import BSpline

from bs2e.workloads import WORKLOADS


def knots(oracle, name):
    w = WORKLOADS[name]
    return w, oracle.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)


# ------------------------------------------------------------------ GL
@pytest.mark.parametrize("N", [1, 2, 3, 6, 10, 12, 14, 16, 24, 33])
def test_gauss_legendre(oracle, N):
    from sympy.integrals.quadrature import gauss_legendre as sym_gl
    x, w = oracle.gauss_legendre(N)
    xs, ws = sym_gl(N, 40)                         # 40-digit nodes and weights
    ex = np.array([float(v) for v in xs])
    ew = np.array([float(v) for v in ws])
    o = np.argsort(ex)
    ex, ew = ex[o], ew[o]
    assert np.all(np.abs(x - ex) <= np.spacing(np.abs(ex) + 1e-300))   # within 1 ulp
    assert np.all(np.abs(w - ew) <= np.spacing(ew))
    xr, wr = np.polynomial.legendre.leggauss(N)     # numpy's weights are only good to ~1e-13
    assert np.allclose(x, xr, rtol=0, atol=4e-16) and np.allclose(w, wr, rtol=1e-12, atol=0)
    for deg in range(2 * N):                        # exact up to degree 2N-1
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        assert abs(np.dot(w, x ** deg) - exact) < 2e-15 * max(1, N)


# ------------------------------------------------------------------ splines
def scipy_basis(t, k, s, x):
    """B_s(x) (1-based s) from scipy: order k <-> degree k-1."""
    c = np.zeros(len(t) - k)
    c[s - 1] = 1.0
    return BSpline(t, c, k - 1, extrapolate=False)(x)


@pytest.mark.parametrize("name", ["tiny", "c1"])
def test_bvalue_vs_scipy_and_partition_of_unity(oracle, name):
    w, t = knots(oracle, name)
    n, n_b, niv = oracle.dims(w.k, t)
    bp = t[w.k - 1:w.k - 1 + niv + 1]
    rng = np.random.default_rng(1)
    for iv in range(1, niv + 1, max(1, niv // 15)):
        for x in bp[iv - 1] + rng.random(3) * (bp[iv] - bp[iv - 1]):
            vals = np.array([oracle.bvalue_unit(t, w.k, iv + ls, x, iv) for ls in range(w.k)])
            ref = np.array([scipy_basis(t, w.k, iv + ls, x) for ls in range(w.k)])
            assert np.allclose(vals, ref, rtol=1e-13, atol=1e-15)
            assert (vals >= 0).all()
            assert abs(vals.sum() - 1.0) < 1e-14


# ------------------------------------------------------------------ primitives
def spline_table(t, k, x):
    """values[s-1, i] = B_s(x_i) for every spline s."""
    n = len(t) - k
    return np.array([scipy_basis(t, k, s, x) for s in range(1, n + 1)])


def test_P_is_the_exact_integral(oracle):
    """P^k[v;a,c] = int_v B_a B_c r^k dr exactly when the rule is exact."""
    w, t = knots(oracle, "tiny")
    k_gl, max_k = 8, 4                                  # 2(k-1)+max_k = 10 <= 2*8-1
    off = oracle.slater_off_diag(w.k, t, k_gl, max_k)
    bp = t[w.k - 1:]
    # exact integral of the piecewise polynomial with a 40-point rule per interval
    xg, wg = np.polynomial.legendre.leggauss(40)
    for e in range(0, off["nnz"], 7):
        iv, a, c = off["iv"][e], off["i"][e], off["j"][e]
        lo, hi = bp[iv - 1], bp[iv]
        r = 0.5 * (hi - lo) * xg + 0.5 * (hi + lo)
        ba, bc = scipy_basis(t, w.k, a + 1, r), scipy_basis(t, w.k, c + 1, r)
        for k in range(max_k + 1):
            exact = 0.5 * (hi - lo) * np.sum(wg * ba * bc * r ** k)   # itself rounded to ~1e-15
            assert abs(off["rk"][e, k] - exact) <= 1e-13 * abs(exact) + 1e-300


@pytest.mark.parametrize("name", ["tiny", "c1"])
def test_Q_and_D_independent_restatement(oracle, name):
    """Same rule, independent code: scipy spline values, vectorised numpy."""
    w, t = knots(oracle, name)
    k_gl, max_k = w.k_GL, w.max_k
    off = oracle.slater_off_diag(w.k, t, k_gl, max_k)
    dg = oracle.slater_diag(w.k, t, k_gl, max_k)
    xq, wq = np.polynomial.legendre.leggauss(k_gl)
    bp = t[w.k - 1:]
    # off-diagonal Q
    for e in range(0, off["nnz"], 3):
        iv, a, c = off["iv"][e], off["i"][e], off["j"][e]
        lo, hi = bp[iv - 1], bp[iv]
        h = 0.5 * (hi - lo)
        r = h * xq + 0.5 * (hi + lo)
        bb = wq * scipy_basis(t, w.k, a + 1, r) * scipy_basis(t, w.k, c + 1, r)
        q = h * np.array([np.sum(bb / r ** (k + 1)) for k in range(max_k + 1)])
        assert np.allclose(off["rmk"][e], q, rtol=2e-14, atol=0)
    # diagonal cells (mat_els.f90:460-483): outer pair (i,i') at r, inner (j,j') on [bp_v, r]
    step = max(1, dg["nnz"] // 400)
    for e in range(0, dg["nnz"], step):
        iv = dg["iv"][e]
        lo, hi = bp[iv - 1], bp[iv]
        h = 0.5 * (hi - lo)
        r = h * xq + 0.5 * (hi + lo)
        bo = wq * scipy_basis(t, w.k, dg["i"][e] + 1, r) * scipy_basis(t, w.k, dg["ip"][e] + 1, r)
        s = 0.5 * (r - lo)
        rho = s[:, None] * xq[None, :] + 0.5 * (r + lo)[:, None]
        bi = wq[None, :] * scipy_basis(t, w.k, dg["j"][e] + 1, rho) * scipy_basis(t, w.k, dg["jp"][e] + 1, rho)
        for k in range(max_k + 1):
            inner = s * np.sum(bi * rho ** k, axis=1)
            ref = h * np.sum(bo * inner / r ** (k + 1))
            assert abs(dg["rdk"][e, k] - ref) <= 1e-13 * abs(ref)


def test_diagonal_cell_triangles_sum_to_the_cell_integral(oracle):
    """D[v;p,q] + D[v;q,p] = int int_{cell^2} B_p(r1) B_q(r2) r<^k / r>^(k+1)."""
    from scipy.integrate import dblquad
    w, t = knots(oracle, "tiny")
    dg = oracle.slater_diag(w.k, t, 12, 2)
    key = {}
    for e in range(dg["nnz"]):
        key[(dg["iv"][e], dg["i"][e], dg["ip"][e], dg["j"][e], dg["jp"][e])] = e
    bp = t[w.k - 1:]
    rng = np.random.default_rng(5)
    picks = rng.choice(dg["nnz"], 6, replace=False)
    for e in picks:
        iv, i, ip, j, jp = (dg[n][e] for n in ("iv", "i", "ip", "j", "jp"))
        swapped = key[(iv, j, jp, i, ip)]
        lo, hi = bp[iv - 1], bp[iv]
        for k in range(3):
            f = lambda r2, r1: (scipy_basis(t, w.k, i + 1, r1) * scipy_basis(t, w.k, ip + 1, r1) *  # noqa: E731
                                scipy_basis(t, w.k, j + 1, r2) * scipy_basis(t, w.k, jp + 1, r2) *
                                min(r1, r2) ** k / max(r1, r2) ** (k + 1))
            # split along the diagonal so each half is smooth
            a1, _ = dblquad(f, lo, hi, lambda r1: lo, lambda r1: r1, epsabs=0, epsrel=1e-11)
            a2, _ = dblquad(f, lo, hi, lambda r1: r1, lambda r1: hi, epsabs=0, epsrel=1e-11)
            both = dg["rdk"][e, k] + dg["rdk"][swapped, k]
            assert abs(both - (a1 + a2)) <= 1e-9 * abs(a1 + a2), (e, k)


# ------------------------------------------------------------------ tensor
@pytest.fixture(scope="module")
def c1_tensor(oracle):
    w, t = knots(oracle, "c1")
    n, n_b, _ = oracle.dims(w.k, t)
    off = oracle.slater_off_diag(w.k, t, w.k_GL, w.max_k)
    dg = oracle.slater_diag(w.k, t, w.k_GL, w.max_k)
    R, touched = oracle.rk_dense(w.k, n_b, w.max_k, off, dg, want_touched=True)
    return w, t, n_b, R, touched


def test_tensor_keys_symmetry_positivity(oracle, c1_tensor):
    w, t, n_b, R, touched = c1_tensor
    assert touched.all() and R.shape[0] == oracle.n_pairs(w.k, n_b)
    assert R.shape[0] == n_b * (2 * w.k - 1) - w.k * (w.k - 1)
    assert (R > 0).all()
    pairs = oracle.pair_table(w.k, n_b)
    idx = {tuple(p): i for i, p in enumerate(pairs)}
    swap = np.array([idx[(c, a)] for a, c in pairs])
    scale = np.abs(R).max()
    assert np.abs(R - R.transpose(1, 0, 2)).max() <= 1e-15 * scale * 8      # R(a,b,c,d) = R(b,a,d,c)
    assert np.abs(R - R[swap]).max() <= 1e-15 * scale * 8                   # R(a,b,c,d) = R(c,b,a,d)
    assert np.abs(R - R[:, swap]).max() <= 1e-15 * scale * 8                # R(a,b,c,d) = R(a,d,c,b)


def test_hydrogenic_slater_integrals(oracle, c1_tensor):
    """Textbook hydrogenic Slater integrals for Z=2 (CAP off)."""
    w, t, n_b, R, _ = c1_tensor
    Z = 2
    S = oracle.setup_S(w.k, t, w.k_GL).real
    orb = {}
    for l in (0, 1):
        H = oracle.setup_H1p(w.k, t, w.k_GL, Z, l, cap_eta=0.0).real
        E, V = sl.eigh(H, S)
        for nprinc in (1, 2):
            i = nprinc - 1 - l
            if i >= 0:
                assert abs(E[i] + Z * Z / (2.0 * nprinc ** 2)) < 1e-9
                orb[(nprinc, l)] = V[:, i]
    pairs = oracle.pair_table(w.k, n_b)

    def slater(k, a, b, c, d):
        A = np.array([a[x - 1] * c[y - 1] for x, y in pairs])
        B = np.array([b[x - 1] * d[y - 1] for x, y in pairs])
        return A @ R[:, :, k] @ B

    s1, s2, p2 = orb[(1, 0)], orb[(2, 0)], orb[(2, 1)]
    cases = [(slater(0, s1, s1, s1, s1), 5 * Z / 8), (slater(0, s1, s2, s1, s2), 17 * Z / 81),
             (slater(0, s1, s2, s2, s1), 16 * Z / 729), (slater(0, s2, s2, s2, s2), 77 * Z / 512),
             (slater(0, s1, p2, s1, p2), 59 * Z / 243), (slater(1, s1, p2, p2, s1), 112 * Z / 2187),
             (slater(0, p2, p2, p2, p2), 93 * Z / 512), (slater(2, p2, p2, p2, p2), 45 * Z / 512)]
    for got, exact in cases:
        assert abs(got - exact) < 1e-9 * exact, (got, exact)


# ------------------------------------------------------------------ angular
def test_wigner_vs_sympy(oracle):
    from sympy.physics.wigner import wigner_3j, wigner_6j
    rng = np.random.default_rng(2)
    for _ in range(300):
        a, b = rng.integers(0, 13, 2)
        c = int(rng.integers(abs(a - b), a + b + 1))
        ref = float(wigner_3j(int(a), int(b), c, 0, 0, 0))
        assert abs(oracle.three_j(a, b, c, 0, 0, 0) - ref) < 1e-14
    for _ in range(300):
        j = [int(x) for x in rng.integers(0, 13, 6)]
        ref = float(wigner_6j(*j))
        assert abs(oracle.six_j(*j) - ref) < 1e-13
    # large arguments used by lmax = 12, k <= 24
    for args in [(12, 12, 4, 12, 12, 24), (11, 12, 3, 10, 9, 20), (12, 7, 5, 9, 10, 17)]:
        assert abs(oracle.six_j(*args) - float(wigner_6j(*args))) < 1e-13


def test_ang_k_LS_closed_form(oracle):
    """ang_k_LS for (s s) -> (s s), L=0, k=0 is 1; parity forbids odd k between s and p."""
    assert abs(oracle.ang_k_LS(0, (0, 0), (0, 0), 0) - 1.0) < 1e-15
    assert oracle.ang_k_LS(1, (0, 0), (0, 0), 0) == 0.0
    # direct Coulomb between p^2 1S and p^2 1S: k=0 -> 1, k=2 -> 2/5 * sqrt? check with sympy
    from sympy import sqrt, Rational  # noqa: F401
    from sympy.physics.wigner import wigner_3j, wigner_6j
    for k in (0, 2):
        la, lb, lc, ld, L = 1, 1, 1, 1, 0
        ref = ((-1) ** (lb + lc + L) * float(wigner_6j(la, lb, L, ld, lc, k)) *
               (-1) ** la * np.sqrt((2 * la + 1) * (2 * lc + 1)) * float(wigner_3j(la, k, lc, 0, 0, 0)) *
               (-1) ** lb * np.sqrt((2 * lb + 1) * (2 * ld + 1)) * float(wigner_3j(lb, k, ld, 0, 0, 0)))
        assert abs(oracle.ang_k_LS(k, (la, lb), (lc, ld), L) - ref) < 1e-14


# ------------------------------------------------------------------ blocks
def _csr(b, key, n):
    import scipy.sparse as sp
    ip, ix, d = b[key]
    return sp.csr_matrix((d, ix - 1, ip - 1), shape=(n, n))


def _lowest(H, S, m=1):
    import scipy.sparse.linalg as sla
    E = sla.eigsh(H.real.tocsc(), k=m, M=S.real.tocsc(), sigma=-3.0, which="LM", return_eigenvectors=False)
    return np.sort(E)


@pytest.fixture(scope="module")
def c1_one_electron(oracle, c1_tensor):
    w, t, n_b, R, _ = c1_tensor
    S = oracle.setup_S(w.k, t, w.k_GL)
    Hl = [oracle.setup_H1p(w.k, t, w.k_GL, 2, l, cap_eta=0.0) for l in range(3)]
    return S, Hl


@pytest.mark.parametrize("max_l,exact,tol", [(0, -2.8790287673, 2e-8), (1, -2.900516, 2e-6), (2, -2.902767, 2e-6)])
def test_helium_partial_wave_limits(oracle, c1_tensor, c1_one_electron, max_l, exact, tol):
    """He 1s^2 1S in the l <= max_l partial-wave limits (Carroll, Silverstone
    and Metzger, J. Chem. Phys. 71, 4142 (1979): -2.8790287673, -2.900516,
    -2.902767).  A wrong exchange sign, a dropped direct or exchange term or a
    transposed index moves these by more than 1e-3."""
    w, t, n_b, R, _ = c1_tensor
    S, Hl = c1_one_electron
    cf = oracle.count_configs(0, 0, max_l, n_b, w.k, n_b, n_b, max_l)
    b = oracle.construct_block(w.k, n_b, 0, w.max_k, True, cf, R, Hl, S)
    H, SS = _csr(b, "H", len(cf)), _csr(b, "S", len(cf))
    assert abs(_lowest(H, SS)[0] - exact) < tol


@pytest.mark.parametrize("L,pi,exact", [(1, 1, -2.1238431), (2, 0, -2.0556207)])
def test_helium_singly_excited(oracle, c1_tensor, c1_one_electron, L, pi, exact):
    """1s2p 1P^o and 1s3d 1D^e of helium (Drake's non-relativistic values);
    the l <= 2 basis leaves a few 1e-4 Ha."""
    w, t, n_b, R, _ = c1_tensor
    S, Hl = c1_one_electron
    cf = oracle.count_configs(L, pi, 2, n_b, w.k, n_b, n_b, 2)
    b = oracle.construct_block(w.k, n_b, L, w.max_k, True, cf, R, Hl, S)
    E = _lowest(_csr(b, "H", len(cf)), _csr(b, "S", len(cf)))[0]
    assert abs(E - exact) < 5e-4


@pytest.mark.parametrize("L", [0, 1])
def test_block_symmetry_and_upper_triangle(oracle, c1_tensor, c1_one_electron, L):
    """H and S are real symmetric without a CAP; full=.false. keeps exactly
    the j >= i part of the full block; count_nnz never exceeds the fill."""
    w, t, n_b, R, _ = c1_tensor
    S, Hl = c1_one_electron
    cf = oracle.count_configs(L, L % 2, 2, n_b, w.k, n_b, n_b, 2)
    n = len(cf)
    full = oracle.construct_block(w.k, n_b, L, w.max_k, True, cf, R, Hl, S)
    up = oracle.construct_block(w.k, n_b, L, w.max_k, False, cf, R, Hl, S)
    for key in ("H", "S"):
        A = _csr(full, key, n)
        assert abs(A - A.T).max() <= 1e-14 * abs(A).max()
        U = _csr(up, key, n)
        import scipy.sparse as sp
        assert (sp.triu(A) != U).nnz == 0
        ip, ix, _ = up[key]
        rows = np.repeat(np.arange(1, n + 1), np.diff(ip))
        assert (ix >= rows).all()
    assert full["ref_count"][0] <= full["nnz"][0] and full["ref_count"][1] <= full["nnz"][1]


# ------------------------------------------------------------------ dipoles
def test_radial_dipole_hydrogenic(oracle):
    """Radial dipole matrices (mat_els.f90:120-170, :348-390) between the
    hydrogenic 1s and 2p orbitals of Z = 2 in the c1 basis: the length form
    gives the closed form <1s|r|2p> = 2^7 sqrt(6) / (3^5 Z); the velocity
    form, -(dr + r_inv) for l = 0 -> l' = 1 (mat_els.f90:796-808), equals
    i (E_1s - E_2p) times the length form's reduced element r C_red(1, 0, 1)
    (commutator [H, r] = -i p; C_red(1, 0, 1) = -1); r_mat and r_inv_mat are
    symmetric and dr_mat antisymmetric (the splines vanish at both ends)."""
    import scipy.linalg as sl
    from bs2e.workloads import WORKLOADS
    w = WORKLOADS["c1"]
    t = oracle.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    S = oracle.setup_S(w.k, t, w.k_GL).real
    rl = oracle.setup_radial_dip(w.k, t, w.k_GL, "l")
    dr, ri = oracle.setup_radial_dip(w.k, t, w.k_GL, "v")
    assert abs(rl - rl.T).max() <= 1e-14 * abs(rl).max() and abs(rl.imag).max() == 0
    assert abs(ri - ri.T).max() <= 1e-14 * abs(ri).max() and abs(ri.real).max() == 0
    assert abs(dr + dr.T).max() <= 1e-13 * abs(dr).max() and abs(dr.real).max() == 0
    orb, E = {}, {}
    for l in (0, 1):
        e, V = sl.eigh(oracle.setup_H1p(w.k, t, w.k_GL, 2, l, cap_eta=0.0).real, S)
        orb[l], E[l] = V[:, 0], e[0]
    assert abs(E[0] + 2.0) < 1e-10 and abs(E[1] + 0.5) < 1e-10
    ln = orb[0] @ rl.real @ orb[1]
    exact = 2**7 * np.sqrt(6) / 3**5 / 2
    assert abs(abs(ln) - exact) < 1e-11 * exact
    assert abs(oracle.C_red(1, 0, 1) + 1.0) < 1e-15
    vel = orb[0] @ (-(dr + ri)) @ orb[1]
    assert abs(vel - 1j * (E[0] - E[1]) * (ln * oracle.C_red(1, 0, 1))) < 1e-11 * abs(ln)


@pytest.mark.parametrize("L1,p1,L2,p2", [(0, 0, 1, 1), (1, 1, 2, 0), (1, 1, 1, 0), (2, 0, 3, 1)])
def test_dipole_block_adjoint(oracle, L1, p1, L2, p2):
    """d_q^dagger = (-1)^q d_-q: the block from syms (1, 2) with component q
    is (-1)^q times the transpose of the block from (2, 1) with -q in the
    length gauge (real) and -(-1)^q times it in the velocity gauge (imaginary:
    p is Hermitian); the patterns are transposes.  The two blocks reach the
    6j symbols, exchange terms and support tests from opposite sides, so a
    swapped argument, a wrong exchange sign or a transposed radial index
    breaks the relation."""
    import scipy.sparse as sp
    from bs2e.workloads import WORKLOADS
    w = WORKLOADS["tiny"]
    t = oracle.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    _, n_b, _ = oracle.dims(w.k, t)
    S = oracle.setup_S(w.k, t, w.k_GL)
    rad = {"l": oracle.setup_radial_dip(w.k, t, w.k_GL, "l"), "v": oracle.setup_radial_dip(w.k, t, w.k_GL, "v")}
    cf1 = oracle.count_configs(L1, p1, 2, n_b, w.k, n_b, n_b, 2)
    cf2 = oracle.count_configs(L2, p2, 2, n_b, w.k, n_b, n_b, 2)
    csr = lambda b, n, m: sp.csr_matrix((b[2], b[1] - 1, b[0] - 1), shape=(n, m))   # noqa: E731
    seen = 0
    for q in (-1, 0, 1):
        M2 = 0
        M1 = q + M2
        if abs(M1) > L1:
            continue
        for g, sgn in (("l", (-1) ** q), ("v", -(-1) ** q)):
            A = csr(oracle.construct_dip_block(w.k, n_b, (L1, M1, p1), cf1, (L2, M2, p2), cf2, q, 1, g, S, rad[g]),
                    len(cf1), len(cf2))
            B = csr(oracle.construct_dip_block(w.k, n_b, (L2, M2, p2), cf2, (L1, M1, p1), cf1, -q, 1, g, S, rad[g]),
                    len(cf2), len(cf1))
            assert A.nnz == B.nnz
            if A.nnz == 0:   # 3j(1 1 1; 0 0 0) = 0
                continue
            pa = sp.csr_matrix((np.ones(A.nnz), A.indices, A.indptr), shape=A.shape)
            pb = sp.csr_matrix((np.ones(B.nnz), B.indices, B.indptr), shape=B.shape)
            assert (pa != pb.T).nnz == 0
            assert abs(A - sgn * B.T).max() <= 1e-13 * abs(A).max()
            seen += 1
    assert seen >= 2


def test_helium_oscillator_strength(oracle, c1_tensor, c1_one_electron):
    """He 1s^2 1S -> 1s2p 1P^o oscillator strength from the dipole block
    (dipole.f90:8-47) between the lowest states of the L = 0 and L = 1 blocks,
    as the reference's test does (test_mat_els.f90:464-490: 2 dE |<0|z|1>|^2
    in the length gauge, 2 |<0|z|1>|^2 / dE in the velocity gauge).  The l <= 2
    basis gives 0.2747 / 0.2742 against the converged 0.27616 (Drake); the two
    gauges agree to 0.2 %."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    w, t, n_b, R, _ = c1_tensor
    S, Hl = c1_one_electron
    st = {}
    for L, pi in ((0, 0), (1, 1)):
        cf = oracle.count_configs(L, pi, 2, n_b, w.k, n_b, n_b, 2)
        b = oracle.construct_block(w.k, n_b, L, w.max_k, True, cf, R, Hl, S)
        H, M = _csr(b, "H", len(cf)).real.tocsc(), _csr(b, "S", len(cf)).real.tocsc()
        e, v = sla.eigsh(H, k=1, M=M, sigma=-3.0, which="LM")
        st[L] = (cf, e[0], v[:, 0] / np.sqrt(v[:, 0] @ (M @ v[:, 0])))
    (cf0, E0, c0), (cf1, E1, c1) = st[0], st[1]
    dE = E1 - E0
    f = {}
    for g in ("l", "v"):
        rad = oracle.setup_radial_dip(w.k, t, w.k_GL, g)
        ip, ix, d = oracle.construct_dip_block(w.k, n_b, (0, 0, 0), cf0, (1, 0, 1), cf1, 0, 1, g, S, rad)
        D = sp.csr_matrix((d, ix - 1, ip - 1), shape=(len(cf0), len(cf1)))
        m = c0 @ (D @ c1)
        f[g] = 2 * dE * abs(m) ** 2 if g == "l" else 2 * abs(m) ** 2 / dE
    assert abs(f["l"] - 0.27616) < 0.01 * 0.27616 and abs(f["v"] - 0.27616) < 0.01 * 0.27616
    assert abs(f["l"] - f["v"]) < 0.002 * f["l"]


# ------------------------------------------------------------------ RKH multipoles
def test_r_k_int_limits_and_closed_form(oracle):
    """RKH_hamiltonian.f90:91-159 setup_r_k_int.  With r_0 beyond the last
    knot every point has r < r_0, so r_k_int(k) is the interval sum of the
    P^k moments over r_0^(k+1); with r_0 below the first non-zero knot it is
    the Q^k moments times r_0^k (both from the independently pinned
    setup_Slater_off_diag restatement).  With r_0 on a breakpoint, the
    hydrogenic 1s orbital of Z = 2 gives the closed form
    <1s| 1/max(r, r_0) |1s> = [1 - e^-x (1 + x + x^2/2)] / r_0 + Z e^-x (1 + x),
    x = 2 Z r_0."""
    import scipy.linalg as sl
    from bs2e.workloads import WORKLOADS
    w = WORKLOADS["c1"]
    t = oracle.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    _, n_b, _ = oracle.dims(w.k, t)
    K = w.max_k + 1
    off = oracle.slater_off_diag(w.k, t, w.k_GL, w.max_k)
    P = np.zeros((K, n_b, n_b))
    Q = np.zeros((K, n_b, n_b))
    for e in range(off["nnz"]):
        P[:, off["i"][e] - 1, off["j"][e] - 1] += off["rk"][e]
        Q[:, off["i"][e] - 1, off["j"][e] - 1] += off["rmk"][e]
    r0 = 2 * t[-1]
    A = oracle.setup_r_k_int(w.k, t, w.k_GL, w.max_k, r0)
    for k in range(K):
        assert np.abs(A[k] - P[k] / r0 ** (k + 1)).max() <= 1e-14 * np.abs(P[k]).max() / r0 ** (k + 1)
    r0 = t[w.k] * 1e-3
    A = oracle.setup_r_k_int(w.k, t, w.k_GL, w.max_k, r0)
    for k in range(K):
        assert np.abs(A[k] - Q[k] * r0 ** k).max() <= 1e-14 * np.abs(Q[k]).max() * r0 ** k
    # banded: zero outside |i - j| < k
    i, j = np.indices((n_b, n_b))
    assert np.all(A[:, np.abs(i - j) >= w.k] == 0)
    S = oracle.setup_S(w.k, t, w.k_GL).real
    _, V = sl.eigh(oracle.setup_H1p(w.k, t, w.k_GL, 2, 0, cap_eta=0.0).real, S)
    c = V[:, 0]
    Z = 2.0
    for r0 in (t[w.k + 5], t[w.k + 12]):   # breakpoints: the kink of r_k sits on an interval edge
        x = 2 * Z * r0
        exact = (1 - np.exp(-x) * (1 + x + x * x / 2)) / r0 + Z * np.exp(-x) * (1 + x)
        got = c @ oracle.setup_r_k_int(w.k, t, w.k_GL, 0, r0)[0] @ c
        assert abs(got - exact) < 1e-9 * exact
