"""The Fortran drop-in (b-spline-two-e_amd/fortran, amdflang) over the C ABI:
bs2e_basis_setup runs the reference's basis_setup call sequence
(main_basis_setup.f90:75-118) on the device and its outputs are checked
against the oracle with the same bars as the Python-side parity tests.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from bs2e.workloads import WORKLOADS  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "b-spline-two-e_amd", "lib", "bs2e_basis_setup")
TOL = 1e-12


class Stream:
    def __init__(self, raw):
        self.raw = raw
        self.pos = 0

    def take(self, n, dtype):
        a = np.frombuffer(self.raw, dtype=dtype, count=n, offset=self.pos)
        self.pos += a.nbytes
        return a

    def i(self, n=None):
        return int(self.take(1, np.int64)[0]) if n is None else self.take(n, np.int64)


def max_rel(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


@pytest.mark.parametrize("L,pi,full,hs_on_device", [(0, 0, 1, False), (1, 1, 0, False), (0, 0, 1, True)])
def test_fortran_basis_setup_matches_oracle(oracle, tmp_path, L, pi, full, hs_on_device):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is missing: build with `python b-spline-two-e_amd/build.py` (needs amdflang)")
    w = WORKLOADS["c1"]
    max_l, stride = 2, 7
    t = oracle.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    _, n_b, _ = oracle.dims(w.k, t)
    S = oracle.setup_S(w.k, t, w.k_GL)
    Hl = [oracle.setup_H1p(w.k, t, w.k_GL, 2, l, cap_eta=0.0) for l in range(max_l + 1)]
    hs = tmp_path / "hs.bin"
    with open(hs, "wb") as f:
        for M in [S] + Hl:
            f.write(np.asfortranarray(M, dtype=np.complex128).tobytes(order="F"))
    out = tmp_path / "out.bin"
    cmd = [EXE, str(w.k), str(w.m), str(w.Z), repr(w.h_max), repr(w.r_max), str(w.k_GL), str(w.max_k), str(max_l),
           str(L), str(pi), str(full), str(stride), "device" if hs_on_device else str(hs), str(out)]
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    s = Stream(out.read_bytes())
    K = w.max_k + 1

    knots = s.take(s.i(), np.float64)
    assert np.array_equal(knots, t)

    off = oracle.slater_off_diag(w.k, t, w.k_GL, w.max_k)
    nnz4 = s.i()
    assert nnz4 == off["nnz"] and s.i() == K
    rk = s.take(nnz4 * K, np.float64).reshape(K, nnz4).T
    rmk = s.take(nnz4 * K, np.float64).reshape(K, nnz4).T
    for key in ("iv", "i", "j"):
        assert np.array_equal(s.i(nnz4), off[key]), key
    assert max_rel(rk, off["rk"]) <= TOL and max_rel(rmk, off["rmk"]) <= TOL

    dg = oracle.slater_diag(w.k, t, w.k_GL, w.max_k)
    nnz6 = s.i()
    assert nnz6 == dg["nnz"]
    rdk = s.take(nnz6 * K, np.float64).reshape(K, nnz6).T
    for key in ("iv", "i", "ip", "j", "jp"):
        assert np.array_equal(s.i(nnz6), dg[key]), key
    assert max_rel(rdk, dg["rdk"]) <= TOL

    R = oracle.rk_dense(w.k, n_b, w.max_k, off, dg)
    n_keys = s.i()
    keys = s.i(4 * n_keys).reshape(n_keys, 4)
    vals = s.take(K * n_keys, np.float64).reshape(n_keys, K)
    vals_direct = s.take(K * n_keys, np.float64).reshape(n_keys, K)
    pidx = {tuple(p): i for i, p in enumerate(oracle.pair_table(w.k, n_b))}
    ref = np.array([R[pidx[(a, c)], pidx[(b, d)]] for a, b, c, d in keys])
    assert n_keys > 1000
    assert max_rel(vals, ref) <= TOL
    assert max_rel(vals_direct, ref) <= TOL

    n_config = s.i()
    cf = s.i(5 * n_config).reshape(n_config, 5)
    assert np.array_equal(cf, oracle.count_configs(L, pi, max_l, n_b, w.k, n_b, n_b, max_l))
    blk = oracle.construct_block(w.k, n_b, L, w.max_k, bool(full), cf, R, Hl, S)
    for name in ("H", "S"):
        nnz = s.i()
        ip = s.i(n_config + 1)
        ix = s.i(nnz)
        d = s.take(2 * nnz, np.float64).view(np.complex128)
        rip, rix, rd = blk[name]
        assert np.array_equal(ip, rip) and np.array_equal(ix, rix), name
        assert np.abs(d - rd).max() <= TOL * np.abs(rd).max(), name   # one-electron inputs from the device too
    assert s.pos == len(s.raw)


@pytest.mark.parametrize("gauge,full", [("l", 1), ("v", 0)])
def test_fortran_basis_files_match_reference_format(oracle, tmp_path, gauge, full):
    """The driver stores H_diag.dat, S_diag.dat, D_-1.dat, D_0.dat, D_1.dat,
    basis.dat, basis_info.txt and splines.dat for every symmetry L = 0..2 (as
    main_basis_setup.f90:122-161 does); their contents equal the oracle's
    blocks, and the Python writer of the same format reproduces the
    amdflang-written files byte for byte."""
    from bs2e import fortran_io as fio
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is missing")
    w = WORKLOADS["t0"]
    max_l, max_L = 2, 2
    t = oracle.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    _, n_b, _ = oracle.dims(w.k, t)
    res = str(tmp_path) + "/"
    cmd = [EXE, str(w.k), str(w.m), str(w.Z), repr(w.h_max), repr(w.r_max), str(w.k_GL), str(w.max_k), str(max_l),
           "0", "0", str(full), "50", "device", str(tmp_path / "out.bin"), str(max_L), res, gauge]
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=180)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    k, knots = fio.load_splines(res)
    assert k == w.k and np.array_equal(knots, t)
    max_l_1p, mL, two_el, syms = fio.load_basis(res)
    assert (max_l_1p, mL, two_el, len(syms)) == (max_l, max_L, True, max_L + 1)
    H = fio.load_block_diag(res + "H_diag.dat")
    S = fio.load_block_diag(res + "S_diag.dat")
    off = oracle.slater_off_diag(w.k, t, w.k_GL, w.max_k)
    dg = oracle.slater_diag(w.k, t, w.k_GL, w.max_k)
    R = oracle.rk_dense(w.k, n_b, w.max_k, off, dg)
    Sm = oracle.setup_S(w.k, t, w.k_GL)
    Hl = [oracle.setup_H1p(w.k, t, w.k_GL, w.Z, l, cap_eta=0.0) for l in range(max_l + 1)]
    for L in range(max_L + 1):
        s = syms[L]
        assert (s.l, s.m, s.pi) == (L, 0, bool(L % 2))
        cf = oracle.count_configs(L, L % 2, max_l, n_b, w.k, n_b, n_b, max_l)
        assert np.array_equal(s.configs, cf)
        ref = oracle.construct_block(w.k, n_b, L, w.max_k, bool(full), cf, R, Hl, Sm)
        for got, key in ((H[L], "H"), (S[L], "S")):
            rip, rix, rd = ref[key]
            assert got.shape == (len(cf), len(cf))
            assert np.array_equal(got.index_ptr, rip) and np.array_equal(got.indices, rix), (L, key)
            assert np.abs(got.data - rd).max() <= TOL * np.abs(rd).max(), (L, key)
    # dipoles: every (i, j) pair for q = -1, 0, 1 (z-polarised: only q = 0
    # reaches M = 0 -> M = 0); upper triangle only when full = 0
    rad = oracle.setup_radial_dip(w.k, t, w.k_GL, gauge)
    D = {q: fio.load_block_CS(res + f"D_{q}.dat") for q in (-1, 0, 1)}
    n_nonempty = 0
    for q in (-1, 0, 1):
        assert len(D[q]) == max_L + 1 and all(len(r) == max_L + 1 for r in D[q])
        for i in range(max_L + 1):
            for j in range(max_L + 1):
                got = D[q][i][j]
                s1, s2 = syms[i], syms[j]
                assert got.shape == (s1.n_config, s2.n_config)
                ip, ix, d = oracle.construct_dip_block(w.k, n_b, (s1.l, s1.m, s1.pi), s1.configs,
                                                       (s2.l, s2.m, s2.pi), s2.configs, q, bool(full) or i <= j,
                                                       gauge, Sm, rad)
                assert got.nnz == len(ix), (q, i, j)
                if got.nnz:
                    n_nonempty += 1
                    assert np.array_equal(got.index_ptr, ip) and np.array_equal(got.indices, ix), (q, i, j)
                    assert np.abs(got.data - d).max() <= TOL * np.abs(d).max(), (q, i, j)
    assert n_nonempty == (4 if full else 2)
    # basis_input.dat: the run's parameters as the namelist the other programs read
    group, nml = fio.read_namelist(res + "basis_input.dat")
    assert group == "basis_input"
    expect = dict(k=w.k, m=w.m, z=w.Z, k_gl=w.k_GL, max_k=w.max_k, max_l=max_L, max_l_1p=max_l, max_l2=max_l,
                  gauge=gauge, full=bool(full), z_pol=True, two_el=True, cap_order=2, cap_eta=0j)
    for key, v in expect.items():
        assert nml[key] == v, (key, nml[key], v)
    assert abs(nml["h_max"] - w.h_max) < 1e-12 and abs(nml["r_max"] - w.r_max) < 1e-12
    assert nml["basis_output_dir"].strip() == res.strip()[:64].strip()
    # the Python writer reproduces the files byte for byte
    py = tmp_path / "py"
    py.mkdir()
    fio.store_block_diag(py / "H_diag.dat", H)
    fio.store_block_diag(py / "S_diag.dat", S)
    for q in (-1, 0, 1):
        fio.store_block_CS(py / f"D_{q}.dat", D[q])
    fio.store_basis(str(py), max_l_1p, mL, two_el, syms)
    fio.store_splines(str(py), k, knots)
    for name in ("H_diag.dat", "S_diag.dat", "D_-1.dat", "D_0.dat", "D_1.dat", "basis.dat", "basis_info.txt",
                 "splines.dat"):
        assert (py / name).read_bytes() == (tmp_path / name).read_bytes(), name
