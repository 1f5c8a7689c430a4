"""ctypes wrapper of the oracle (bs2e_oracle.c) and of the compiled-reference
checker (oracle/_ref/libbs2e_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product (b-spline-two-e_amd/)
never imports it.

Each wrapper names the reference routine it restates; the C source carries
the file:line citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libbs2e_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libbs2e_ref.so")

i64 = C.c_int64
dbl = C.c_double
P = C.POINTER

_lib = None
_ref = None


def build(ref: bool = True) -> None:
    """Compile the oracle (and the compiled-reference checker when the
    reference sources are present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True,
                       stderr=subprocess.DEVNULL)


def _ptr(a, ct):
    return a.ctypes.data_as(P(ct))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(ref=False)
        L = C.CDLL(LIB_PATH)
        L.or_generate_grid.restype = i64
        L.or_generate_grid.argtypes = [i64, i64, i64, dbl, dbl, P(dbl), i64]
        L.or_bvalue_unit.restype = dbl
        L.or_bvalue_unit.argtypes = [P(dbl), i64, i64, i64, dbl, i64, i64]
        L.or_bvalue.restype = dbl
        L.or_bvalue.argtypes = [P(dbl), i64, i64, P(dbl), dbl, i64, i64]
        L.or_gauss_legendre.restype = None
        L.or_gauss_legendre.argtypes = [i64, P(dbl), P(dbl)]
        L.or_spline_dims.restype = None
        L.or_spline_dims.argtypes = [i64, i64, P(i64), P(i64), P(i64)]
        L.or_count_nnz_4d.restype = i64
        L.or_count_nnz_4d.argtypes = [i64, i64]
        L.or_count_nnz_6d.restype = i64
        L.or_count_nnz_6d.argtypes = [i64, i64]
        L.or_count_pairs.restype = i64
        L.or_count_pairs.argtypes = [i64, i64]
        L.or_pair_index.restype = i64
        L.or_pair_index.argtypes = [i64, i64, i64, i64]
        L.or_slater_off_diag.restype = None
        L.or_slater_off_diag.argtypes = [i64, i64, P(dbl), i64, i64, i64, P(dbl), P(dbl), P(i64), P(i64),
                                         P(i64), i64, i64]
        L.or_slater_diag.restype = None
        L.or_slater_diag.argtypes = [i64, i64, P(dbl), i64, i64, i64, P(dbl), P(i64), P(i64), P(i64),
                                     P(i64), P(i64), i64, i64]
        L.or_slater_diag_masked.restype = None
        L.or_slater_diag_masked.argtypes = [i64, i64, P(dbl), i64, i64, i64, P(dbl), P(i64), P(i64), P(i64),
                                            P(i64), P(i64), i64, P(C.c_ubyte)]
        L.or_slater_diag_sample.restype = i64
        L.or_slater_diag_sample.argtypes = [i64, i64, P(dbl), i64, i64, i64, i64, P(dbl)]
        L.or_rk_rows.restype = None
        L.or_rk_rows.argtypes = [i64, i64, i64, i64, P(dbl), P(dbl), P(i64), P(i64), P(i64), i64, P(dbl),
                                 P(i64), P(i64), P(i64), P(i64), i64, P(i64), P(dbl)]
        L.or_rk_dok.restype = i64
        L.or_rk_dok.argtypes = [i64, i64, i64, i64, P(dbl), P(dbl), P(i64), P(i64), P(i64), i64, P(dbl),
                                P(i64), P(i64), P(i64), P(i64), i64, P(dbl)]
        L.or_construct_block_rows.restype = i64
        L.or_construct_block_rows.argtypes = [i64, i64, i64, i64, i64, i64, P(i64), P(dbl), P(i64), i64, i64,
                                              P(dbl), P(dbl), P(i64), P(i64), P(dbl), P(i64), P(i64), P(dbl), i64,
                                              i64, P(i64), P(i64)]
        L.or_rk_dense.restype = None
        L.or_rk_dense.argtypes = [i64, i64, i64, i64, P(dbl), P(dbl), P(i64), P(i64), P(i64), i64, P(dbl),
                                  P(i64), P(i64), P(i64), P(i64), P(dbl), P(C.c_ubyte)]
        L.or_three_j.restype = dbl
        L.or_three_j.argtypes = [i64] * 6
        L.or_six_j.restype = dbl
        L.or_six_j.argtypes = [i64] * 6
        L.or_C_red.restype = dbl
        L.or_C_red.argtypes = [i64] * 3
        L.or_ang_k_LS.restype = dbl
        L.or_ang_k_LS.argtypes = [i64] * 6
        L.or_count_configs.restype = i64
        L.or_count_configs.argtypes = [i64] * 8 + [P(i64), i64]
        L.or_setup_S.restype = None
        L.or_setup_S.argtypes = [i64, i64, P(dbl), i64, P(C.c_double)]
        L.or_setup_H1p.restype = None
        L.or_setup_H1p.argtypes = [i64, i64, P(dbl), i64, i64, i64, i64, dbl, dbl, dbl, P(C.c_double)]
        L.or_construct_block.restype = i64
        L.or_construct_block.argtypes = [i64, i64, i64, i64, i64, i64, P(i64), P(dbl), P(dbl), P(dbl),
                                         P(i64), P(i64), P(dbl), P(i64), P(i64), P(dbl), i64, i64, P(i64),
                                         P(i64)]
        L.or_setup_radial_dip.restype = None
        L.or_setup_radial_dip.argtypes = [i64, i64, P(dbl), i64, i64, P(dbl), P(dbl), P(dbl)]
        L.or_setup_r_k_int.restype = None
        L.or_setup_r_k_int.argtypes = [i64, i64, P(dbl), i64, i64, dbl, P(dbl)]
        L.or_construct_dip_block.restype = i64
        L.or_construct_dip_block.argtypes = [i64, i64, i64, i64, i64, i64, P(i64), i64, i64, i64, i64, P(i64), i64,
                                             i64, i64, P(dbl), P(dbl), P(dbl), P(dbl), P(i64), P(i64), P(dbl), i64]
        _lib = L
    return _lib


# --------------------------------------------------------------------------
# grid / splines / quadrature
# --------------------------------------------------------------------------
def generate_grid(k, m, Z, h_max, r_max):
    """grid_tools.f90:6-55"""
    L = lib()
    n = L.or_generate_grid(k, m, Z, h_max, r_max, None, 0)
    if n < 0:
        raise ValueError("the grid overruns the reference's 2k + m + ceiling(r_max / 2^-m) knots (grid_tools.f90:20)")
    out = np.zeros(n)
    L.or_generate_grid(k, m, Z, h_max, r_max, _ptr(out, dbl), n)
    return out


def dims(k, t):
    """bspline_tools.f90:27-40: (n, n_b, n_iv)"""
    n, nb, niv = i64(), i64(), i64()
    lib().or_spline_dims(k, len(t), C.byref(n), C.byref(nb), C.byref(niv))
    return n.value, nb.value, niv.value


def bvalue_unit(t, k, s, x, iv, i_deriv=0):
    """bspline_tools.f90:151-224 with unit coefficients e_s (s 1-based)."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    n = len(t) - k
    return lib().or_bvalue_unit(_ptr(t, dbl), n, k, s, float(x), i_deriv, iv)


def gauss_legendre(N):
    """quad_tools.f90:14-27 (stdlib gauss_legendre on [-1,1])"""
    x = np.zeros(N)
    w = np.zeros(N)
    lib().or_gauss_legendre(N, _ptr(x, dbl), _ptr(w, dbl))
    return x, w


def count_nnz4(k, t):
    return lib().or_count_nnz_4d(k, len(t))


def count_nnz6(k, t):
    return lib().or_count_nnz_6d(k, len(t))


def n_pairs(k, n_b):
    return lib().or_count_pairs(k, n_b)


def pair_index(k, n_b, a, c):
    return lib().or_pair_index(k, n_b, a, c)


# --------------------------------------------------------------------------
# primitives and tensor
# --------------------------------------------------------------------------
def slater_off_diag(k, t, k_gl, max_k, faithful=False, nthreads=0):
    """mat_els.f90:184-228 + :392-439 -> dict(rk, rmk (nnz4, K) Fortran order, iv, i, j)"""
    t = np.ascontiguousarray(t, dtype=np.float64)
    nnz = count_nnz4(k, t)
    K = max_k + 1
    rk = np.zeros(nnz * K)
    rmk = np.zeros(nnz * K)
    iv = np.zeros(nnz, np.int64)
    ii = np.zeros(nnz, np.int64)
    jj = np.zeros(nnz, np.int64)
    lib().or_slater_off_diag(k, len(t), _ptr(t, dbl), k_gl, max_k, nnz, _ptr(rk, dbl), _ptr(rmk, dbl),
                             _ptr(iv, i64), _ptr(ii, i64), _ptr(jj, i64), int(faithful), nthreads)
    return dict(nnz=nnz, rk=rk.reshape(K, nnz).T, rmk=rmk.reshape(K, nnz).T, iv=iv, i=ii, j=jj)


def slater_diag(k, t, k_gl, max_k, faithful=False, nthreads=0, pair_mask=None):
    """mat_els.f90:230-292 + :441-491 -> dict(rdk (nnz6, K), iv, i, ip, j, jp).
    pair_mask (bool per 0-based banded pair): compute only the entries whose
    pair (i, i') or (j, j') is marked; the others stay zero (index arrays are
    complete; each computed entry has the bits of the full call)."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    nnz = count_nnz6(k, t)
    K = max_k + 1
    rdk = np.zeros(nnz * K)
    arrs = [np.zeros(nnz, np.int64) for _ in range(5)]
    if pair_mask is None:
        lib().or_slater_diag(k, len(t), _ptr(t, dbl), k_gl, max_k, nnz, _ptr(rdk, dbl),
                             *[_ptr(a, i64) for a in arrs], int(faithful), nthreads)
    else:
        mask = np.ascontiguousarray(pair_mask, dtype=np.uint8)
        lib().or_slater_diag_masked(k, len(t), _ptr(t, dbl), k_gl, max_k, nnz, _ptr(rdk, dbl),
                                    *[_ptr(a, i64) for a in arrs], nthreads, _ptr(mask, C.c_ubyte))
    iv, ii, ip, jj, jp = arrs
    return dict(nnz=nnz, rdk=rdk.reshape(K, nnz).T, iv=iv, i=ii, ip=ip, j=jj, jp=jp)


def slater_diag_sample(k, t, k_gl, k_sel, jp_hi, faithful=True):
    """Timing sample of the reference's diagonal-cell loop (mat_els.f90:247-289):
    multipole k_sel only, outer columns j_b_p <= jp_hi, one thread.  Returns the
    number of sparse_6d entries computed."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    sink = dbl(0.0)
    return lib().or_slater_diag_sample(k, len(t), _ptr(t, dbl), k_gl, k_sel, jp_hi, int(faithful), C.byref(sink))


def rk_dense(k, n_b, max_k, off, diag, want_touched=False):
    """sparse_array_tools.f90:452-493 (DOK -> dense banded R[p, q, k])."""
    K = max_k + 1
    npairs = n_pairs(k, n_b)
    R = np.zeros(npairs * npairs * K)
    touched = np.zeros(npairs * npairs, np.uint8) if want_touched else None
    rk = np.ascontiguousarray(off["rk"].T)
    rmk = np.ascontiguousarray(off["rmk"].T)
    rdk = np.ascontiguousarray(diag["rdk"].T)
    lib().or_rk_dense(k, n_b, max_k, off["nnz"], _ptr(rk, dbl), _ptr(rmk, dbl), _ptr(off["iv"], i64),
                      _ptr(off["i"], i64), _ptr(off["j"], i64), diag["nnz"], _ptr(rdk, dbl),
                      _ptr(diag["i"], i64), _ptr(diag["ip"], i64), _ptr(diag["j"], i64),
                      _ptr(diag["jp"], i64), _ptr(R, dbl),
                      _ptr(touched, C.c_ubyte) if want_touched else None)
    R = R.reshape(npairs, npairs, K)
    return (R, touched.reshape(npairs, npairs)) if want_touched else R


def rk_rows(k, n_b, max_k, off, diag, rows):
    """Rows `rows` (0-based pair indices) of rk_dense, bit for bit: shape
    (len(rows), n_pairs, K).  Only the diagonal cells that land in these rows
    are read, so `diag` may come from slater_diag(pair_mask=...) with the rows
    marked."""
    K = max_k + 1
    npairs = n_pairs(k, n_b)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    assert len(np.unique(rows)) == len(rows) and rows.min() >= 0 and rows.max() < npairs
    out = np.zeros(len(rows) * npairs * K)
    rk = np.ascontiguousarray(off["rk"].T)
    rmk = np.ascontiguousarray(off["rmk"].T)
    rdk = np.ascontiguousarray(diag["rdk"].T)
    lib().or_rk_rows(k, n_b, max_k, off["nnz"], _ptr(rk, dbl), _ptr(rmk, dbl), _ptr(off["iv"], i64),
                     _ptr(off["i"], i64), _ptr(off["j"], i64), diag["nnz"], _ptr(rdk, dbl),
                     _ptr(diag["i"], i64), _ptr(diag["ip"], i64), _ptr(diag["j"], i64),
                     _ptr(diag["jp"], i64), len(rows), _ptr(rows, i64), _ptr(out, dbl))
    return out.reshape(len(rows), npairs, K)


def rk_dok(k, n_b, max_k, off, diag, n_lead=None, want_dense=False):
    """compute_R_K_map (sparse_array_tools.f90:452-493) into a restated Nd_DOK
    hash map, the reference's serial schedule, for the products of the first
    n_lead sparse_4d entries (all when None) and every diagonal cell.  Returns
    the number of set_val calls (and the dense tensor of the map's contents
    with want_dense)."""
    K = max_k + 1
    npairs = n_pairs(k, n_b)
    rk = np.ascontiguousarray(off["rk"].T)
    rmk = np.ascontiguousarray(off["rmk"].T)
    rdk = np.ascontiguousarray(diag["rdk"].T)
    R = np.zeros(npairs * npairs * K) if want_dense else None
    n = lib().or_rk_dok(k, n_b, max_k, off["nnz"], _ptr(rk, dbl), _ptr(rmk, dbl), _ptr(off["iv"], i64),
                        _ptr(off["i"], i64), _ptr(off["j"], i64), diag["nnz"], _ptr(rdk, dbl),
                        _ptr(diag["i"], i64), _ptr(diag["ip"], i64), _ptr(diag["j"], i64), _ptr(diag["jp"], i64),
                        off["nnz"] if n_lead is None else int(n_lead), _ptr(R, dbl) if want_dense else None)
    if n < 0:
        raise MemoryError("rk_dok: host allocation failed")
    return (n, R.reshape(npairs, npairs, K)) if want_dense else n


def pair_table(k, n_b):
    """(a, c) of every banded pair in the loop order of count_nnz_R_k."""
    out = []
    for c in range(1, n_b + 1):
        for a in range(max(1, c - k + 1), min(n_b, c + k - 1) + 1):
            out.append((a, c))
    return np.array(out, dtype=np.int64)


# --------------------------------------------------------------------------
# angular algebra, configurations, one-electron matrices, blocks
# --------------------------------------------------------------------------
def three_j(*a):
    return lib().or_three_j(*[int(x) for x in a])


def six_j(*a):
    return lib().or_six_j(*[int(x) for x in a])


def C_red(k, a, b):
    return lib().or_C_red(k, a, b)


def ang_k_LS(k, l1, l2, L):
    """wigner_tools.f90:126-138; l1 = confs(1)%l, l2 = confs(2)%l."""
    return lib().or_ang_k_LS(k, l1[0], l1[1], l2[0], l2[1], L)


def count_configs(L, term_pi, max_l_1p, n_b, k_spline, max_n_b, n_all_l, l_2_max):
    """orbital_tools.f90:119-216 -> int64 array (n_config, 5): n1, n2, l1, l2, eqv"""
    Lb = lib()
    args = [L, int(term_pi), max_l_1p, n_b, k_spline, max_n_b, n_all_l, l_2_max]
    n = Lb.or_count_configs(*args, None, 0)
    out = np.zeros(5 * max(n, 1), np.int64)
    Lb.or_count_configs(*args, _ptr(out, i64), n)
    return out[:5 * n].reshape(n, 5)


def symmetries(max_L, z_pol=True, two_el=True):
    """orbital_tools.f90:245-343 init_basis ordering: list of (L, M, pi)."""
    syms = [(0, 0, 0)]
    if z_pol:
        for l in range(1, max_L + 1):
            syms.append((l, 0, l % 2))
    else:
        for l in range(1, max_L + 1):
            for p in (0, 1):
                for m in range(-l, l + 1):
                    if abs(m) % 2 != p:
                        continue
                    syms.append((l, m, p))
    return syms


def setup_S(k, t, k_gl):
    """mat_els.f90:85-118 -> complex (n_b, n_b)"""
    t = np.ascontiguousarray(t, dtype=np.float64)
    _, n_b, _ = dims(k, t)
    S = np.zeros(2 * n_b * n_b)
    lib().or_setup_S(k, len(t), _ptr(t, dbl), k_gl, _ptr(S, dbl))
    return S.view(np.complex128).reshape(n_b, n_b, order="F")


def setup_H1p(k, t, k_gl, Z, l, cap_order=2, cap_r0=10.0, cap_eta=0.0):
    """mat_els.f90:47-83 -> complex (n_b, n_b)"""
    t = np.ascontiguousarray(t, dtype=np.float64)
    _, n_b, _ = dims(k, t)
    H = np.zeros(2 * n_b * n_b)
    lib().or_setup_H1p(k, len(t), _ptr(t, dbl), k_gl, Z, l, cap_order, cap_r0, complex(cap_eta).real,
                       complex(cap_eta).imag, _ptr(H, dbl))
    return H.view(np.complex128).reshape(n_b, n_b, order="F")


def construct_block(k, n_b, L, max_k, full, configs, R, Hl, S):
    """hamiltonian.f90:106-283 (+ count_nnz :348-416).

    Returns dict(H=(index_ptr, indices, data), S=(...), nnz=(nH, nS),
    ref_count=(cH, cS)); CSR arrays are 1-based like the reference."""
    Lb = lib()
    cf = np.ascontiguousarray(configs, dtype=np.int64)
    n = cf.shape[0]
    Rf = np.ascontiguousarray(R, dtype=np.float64)
    Hf = np.ascontiguousarray(np.stack([np.asfortranarray(h).ravel(order="F") for h in Hl]), dtype=np.complex128)
    Sf = np.ascontiguousarray(np.asfortranarray(S).ravel(order="F"), dtype=np.complex128)
    nnz = np.zeros(2, np.int64)
    ref = np.zeros(2, np.int64)
    common = [k, n_b, L, max_k, int(full), n, _ptr(cf, i64), _ptr(Rf, dbl), Hf.ctypes.data_as(P(dbl)),
              Sf.ctypes.data_as(P(dbl))]
    rc = Lb.or_construct_block(*common, None, None, None, None, None, None, 0, 0, _ptr(nnz, i64), _ptr(ref, i64))
    assert rc == 0
    ipH = np.zeros(n + 1, np.int64)
    ixH = np.zeros(max(nnz[0], 1), np.int64)
    dH = np.zeros(max(nnz[0], 1), np.complex128)
    ipS = np.zeros(n + 1, np.int64)
    ixS = np.zeros(max(nnz[1], 1), np.int64)
    dS = np.zeros(max(nnz[1], 1), np.complex128)
    rc = Lb.or_construct_block(*common, _ptr(ipH, i64), _ptr(ixH, i64), dH.ctypes.data_as(P(dbl)),
                               _ptr(ipS, i64), _ptr(ixS, i64), dS.ctypes.data_as(P(dbl)), int(nnz[0]),
                               int(nnz[1]), _ptr(nnz, i64), _ptr(ref, i64))
    assert rc == 0, rc
    return dict(H=(ipH, ixH[:nnz[0]], dH[:nnz[0]]), S=(ipS, ixS[:nnz[1]], dS[:nnz[1]]),
                nnz=tuple(int(x) for x in nnz), ref_count=tuple(int(x) for x in ref))


def construct_block_rows(k, n_b, L, max_k, full, configs, R_rows, rmap, row_lo, row_hi, Hl, S):
    """Rows [row_lo, row_hi) of construct_block (hamiltonian.f90:106-283), the
    same loop restricted to them, with a tensor that holds only some rows:
    full-tensor row p is row rmap[p] of R_rows (rmap[p] = -1: absent).
    Returns dict(H=(index_ptr relative to row_lo, indices, data), S=(...),
    nnz=(nH, nS))."""
    Lb = lib()
    cf = np.ascontiguousarray(configs, dtype=np.int64)
    n = cf.shape[0]
    Rf = np.ascontiguousarray(R_rows, dtype=np.float64)
    rm = np.ascontiguousarray(rmap, dtype=np.int64)
    Hf = np.ascontiguousarray(np.stack([np.asfortranarray(h).ravel(order="F") for h in Hl]), dtype=np.complex128)
    Sf = np.ascontiguousarray(np.asfortranarray(S).ravel(order="F"), dtype=np.complex128)
    nnz = np.zeros(2, np.int64)
    ref = np.zeros(2, np.int64)
    common = [k, n_b, L, max_k, int(full), n, _ptr(cf, i64), _ptr(Rf, dbl), _ptr(rm, i64), row_lo, row_hi,
              Hf.ctypes.data_as(P(dbl)), Sf.ctypes.data_as(P(dbl))]
    rc = Lb.or_construct_block_rows(*common, None, None, None, None, None, None, 0, 0, _ptr(nnz, i64),
                                    _ptr(ref, i64))
    assert rc == 0, rc
    m = row_hi - row_lo
    ipH = np.zeros(m + 1, np.int64)
    ixH = np.zeros(max(nnz[0], 1), np.int64)
    dH = np.zeros(max(nnz[0], 1), np.complex128)
    ipS = np.zeros(m + 1, np.int64)
    ixS = np.zeros(max(nnz[1], 1), np.int64)
    dS = np.zeros(max(nnz[1], 1), np.complex128)
    rc = Lb.or_construct_block_rows(*common, _ptr(ipH, i64), _ptr(ixH, i64), dH.ctypes.data_as(P(dbl)),
                                    _ptr(ipS, i64), _ptr(ixS, i64), dS.ctypes.data_as(P(dbl)), int(nnz[0]),
                                    int(nnz[1]), _ptr(nnz, i64), _ptr(ref, i64))
    if rc == -3:   # the count pass reads no values; the fill pass finds the missing row
        raise ValueError("construct_block_rows: a tensor row the rows need is absent from R_rows")
    assert rc == 0, rc
    return dict(H=(ipH, ixH[:nnz[0]], dH[:nnz[0]]), S=(ipS, ixS[:nnz[1]], dS[:nnz[1]]),
                nnz=tuple(int(x) for x in nnz))


def block_tensor_rows(k, n_b, configs, row_lo, row_hi):
    """0-based pair indices of the tensor rows that rows [row_lo, row_hi) of a
    block read: pair(n_a, c) for every c within the band of n_a (direct and
    exchange values of config (n_a, n_b) are R[pair(n_a, .)][...])."""
    cf = np.asarray(configs)
    rows = set()
    for na in np.unique(cf[row_lo:row_hi, 0]):
        for c in range(max(1, na - k + 1), min(n_b, na + k - 1) + 1):
            rows.add(pair_index(k, n_b, int(na), c))
    return np.array(sorted(rows), dtype=np.int64)


# --------------------------------------------------------------------------
# compiled reference (oracle/_ref): only where /root/reference was present
# --------------------------------------------------------------------------
def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        R.ref_generate_grid.restype = i64
        R.ref_generate_grid.argtypes = [i64, i64, i64, dbl, dbl, P(dbl), i64]
        R.ref_bvalue_unit.restype = dbl
        R.ref_bvalue_unit.argtypes = [P(dbl), i64, i64, i64, dbl, i64, i64]
        R.ref_bvalue_batch.restype = None
        R.ref_bvalue_batch.argtypes = [P(dbl), i64, i64, i64, P(i64), P(dbl), P(i64), i64, P(dbl)]
        R.ref_spline_info.restype = None
        R.ref_spline_info.argtypes = [P(dbl), i64, i64, P(i64), P(i64), P(i64), P(i64), i64]
        R.ref_find_max_n_b.restype = i64
        R.ref_find_max_n_b.argtypes = [P(dbl), i64, i64, dbl]
        R.ref_count_configs.restype = i64
        R.ref_count_configs.argtypes = [i64] * 8 + [P(i64), i64]
        _ref = R
    return _ref


def ref_generate_grid(k, m, Z, h_max, r_max):
    cap = 2 * k + m + int(np.ceil(r_max * 2 ** m)) + 64
    out = np.zeros(cap)
    n = ref().ref_generate_grid(k, m, Z, h_max, r_max, _ptr(out, dbl), cap)
    assert n <= cap
    return out[:n].copy()


def ref_bvalue_batch(t, k, s_list, x_list, iv_list, i_deriv=0):
    t = np.ascontiguousarray(t, dtype=np.float64)
    s = np.ascontiguousarray(s_list, dtype=np.int64)
    x = np.ascontiguousarray(x_list, dtype=np.float64)
    iv = np.ascontiguousarray(iv_list, dtype=np.int64)
    out = np.zeros(len(s))
    ref().ref_bvalue_batch(_ptr(t, dbl), len(t), k, len(s), _ptr(s, i64), _ptr(x, dbl), _ptr(iv, i64), i_deriv,
                           _ptr(out, dbl))
    return out


def ref_spline_info(t, k):
    t = np.ascontiguousarray(t, dtype=np.float64)
    n, nb, nbp = i64(), i64(), i64()
    cap = len(t) * len(t)
    sup = np.zeros(cap, np.int64)
    ref().ref_spline_info(_ptr(t, dbl), len(t), k, C.byref(n), C.byref(nb), C.byref(nbp), _ptr(sup, i64), cap)
    niv = nbp.value - 1
    return n.value, nb.value, nbp.value, sup[:niv * n.value].reshape(n.value, niv).T.astype(bool)


def ref_find_max_n_b(t, k, x):
    t = np.ascontiguousarray(t, dtype=np.float64)
    return ref().ref_find_max_n_b(_ptr(t, dbl), len(t), k, float(x))


def ref_count_configs(L, term_pi, max_l_1p, n_b, k_spline, max_n_b, n_all_l, l_2_max):
    cap = (max_l_1p + 1) ** 2 * n_b * n_b + 1
    out = np.zeros(5 * cap, np.int64)
    n = ref().ref_count_configs(L, int(term_pi), max_l_1p, n_b, k_spline, max_n_b, n_all_l, l_2_max,
                                _ptr(out, i64), cap)
    return out[:5 * n].reshape(n, 5)


def setup_radial_dip(k, t, k_gl, gauge="l"):
    """mat_els.f90:123-170: gauge 'l' -> r_mat; 'v' -> (dr_mat, r_inv_mat); complex (n_b, n_b)."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    _, n_b, _ = dims(k, t)
    g = 0 if gauge == "l" else 1
    mats = [np.zeros(n_b * n_b, np.complex128) for _ in range(3)]
    P_ = lambda a: a.ctypes.data_as(P(dbl))   # noqa: E731
    lib().or_setup_radial_dip(k, len(t), _ptr(t, dbl), k_gl, g, P_(mats[0]), P_(mats[1]), P_(mats[2]))
    out = [m.reshape(n_b, n_b, order="F") for m in mats]
    return out[0] if g == 0 else (out[1], out[2])


def construct_dip_block(k, n_b, sym1, cf1, sym2, cf2, q, compute, gauge, S, rad):
    """dipole.f90:8-47 (+ init_dip_block :95-150): sym = (L, M, pi); rad = r_mat
    (gauge 'l') or (dr_mat, r_inv_mat) ('v').  Returns (index_ptr, indices, data)."""
    Lb = lib()
    c1 = np.ascontiguousarray(cf1, dtype=np.int64).reshape(-1, 5)
    c2 = np.ascontiguousarray(cf2, dtype=np.int64).reshape(-1, 5)
    g = 0 if gauge == "l" else 1
    F = lambda m: np.ascontiguousarray(np.asfortranarray(m).ravel(order="F"), dtype=np.complex128)   # noqa: E731
    Sf = F(S)
    if g == 0:
        r, dr, ri = F(rad), None, None
    else:
        r, dr, ri = None, F(rad[0]), F(rad[1])
    cp = lambda a: a.ctypes.data_as(P(dbl)) if a is not None else None   # noqa: E731
    args = [k, n_b, sym1[0], sym1[1], int(sym1[2]), len(c1), _ptr(c1, i64), sym2[0], sym2[1], int(sym2[2]), len(c2),
            _ptr(c2, i64), q, int(compute), g, cp(Sf), cp(r), cp(dr), cp(ri)]
    nnz = Lb.or_construct_dip_block(*args, None, None, None, 0)
    ip = np.ones(len(c1) + 1, np.int64)
    ix = np.zeros(max(nnz, 1), np.int64)
    d = np.zeros(max(nnz, 1), np.complex128)
    n2 = Lb.or_construct_dip_block(*args, _ptr(ip, i64), _ptr(ix, i64), d.ctypes.data_as(P(dbl)), max(nnz, 1))
    assert n2 == nnz
    return ip, ix[:nnz], d[:nnz]


def setup_r_k_int(k, t, k_gl, max_k, r_0):
    """RKH_hamiltonian.f90:91-159: r_k_int(0:max_k, n_b, n_b) as an array
    [k, i, j] (Fortran order: k fastest)."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    _, n_b, _ = dims(k, t)
    out = np.zeros((max_k + 1) * n_b * n_b, np.float64)
    lib().or_setup_r_k_int(k, len(t), _ptr(t, dbl), k_gl, max_k, float(r_0), _ptr(out, dbl))
    return out.reshape(max_k + 1, n_b, n_b, order="F")
