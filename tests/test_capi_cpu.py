"""The C ABI library without a GPU: it loads, exports every symbol that
include/bs2e.h declares, and its host-side entry points agree with the
reference's pinned outputs."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from bs2e import capi
from bs2e.workloads import WORKLOADS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bs2e.h")
FX = np.load(os.path.join(ROOT, "tests", "golden", "ref_fixtures.npz"))


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bs2e_[A-Za-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    # and the Python binding covers all of them
    assert set(names) == {s[0] for s in capi.SIGNATURES}


@pytest.mark.parametrize("name", ["tiny", "t0", "c1", "c2", "c3", "c5"])
def test_generate_grid_bit_exact(name):
    import bs2e
    w = WORKLOADS[name]
    assert np.array_equal(bs2e.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max), FX[f"knots_{name}"])


def test_generate_grid_capacity_boundary(oracle):
    """The reference fills 2k + m + ceiling(r_max / 2^-m) knots
    (grid_tools.f90:20); a small r_max leaves the geometric ramp no room and the
    reference writes past that array.  The library refuses exactly the inputs
    the oracle's bounded restatement refuses, and agrees bit for bit on the
    rest."""
    import bs2e
    n_ok = n_bad = 0
    for k in (2, 4, 8, 12, 16):
        for m in (0, 3, 5):
            for Z in (1, 2, 5):
                for h_max in (0.25, 1.0):
                    for r_max in (0.05, 0.3, 1.0, 5.0, 40.0):
                        try:
                            ref = oracle.generate_grid(k, m, Z, h_max, r_max)
                        except ValueError:
                            ref = None
                        if ref is None:
                            n_bad += 1
                            with pytest.raises(bs2e.Bs2eError, match="2k \\+ m"):
                                bs2e.generate_grid(k, m, Z, h_max, r_max)
                        else:
                            n_ok += 1
                            assert np.array_equal(bs2e.generate_grid(k, m, Z, h_max, r_max), ref), (k, m, Z, h_max,
                                                                                                    r_max)
    assert n_ok >= 50 and n_bad >= 50   # m = 0 never leaves 0 and always overruns
    with pytest.raises(bs2e.Bs2eError):
        bs2e.generate_grid(4, 3, 2, 0.5, 0.3)


@pytest.mark.parametrize("name", ["tiny", "t0", "c1", "c2"])
def test_count_nnz(oracle, name):
    import bs2e
    w = WORKLOADS[name]
    t = FX[f"knots_{name}"]
    nnz4, nnz6, npairs = bs2e.count_nnz(w.k, t)
    assert nnz4 == oracle.count_nnz4(w.k, t)
    assert nnz6 == oracle.count_nnz6(w.k, t)
    assert npairs == oracle.n_pairs(w.k, w.n_b)


def test_angular_matches_oracle(oracle):
    import bs2e
    rng = np.random.default_rng(4)
    for _ in range(400):
        l1 = [int(x) for x in rng.integers(0, 9, 2)]
        l2 = [int(x) for x in rng.integers(0, 9, 2)]
        L = int(rng.integers(0, 9))
        k = int(rng.integers(0, 17))
        assert abs(bs2e.ang_k_LS(k, l1, l2, L) - oracle.ang_k_LS(k, l1, l2, L)) < 1e-14
    for _ in range(200):
        j = [int(x) for x in rng.integers(0, 25, 6)]
        assert abs(bs2e.six_j(*j) - oracle.six_j(*j)) < 1e-13
        a, b = j[0], j[1]
        c = int(rng.integers(abs(a - b), a + b + 1))
        assert abs(bs2e.three_j(a, b, c, 0, 0, 0) - oracle.three_j(a, b, c, 0, 0, 0)) < 1e-14


@pytest.mark.parametrize("L", [0, 1, 2])
def test_count_configs_bit_exact(oracle, L):
    """orbital_tools.f90:119-216 through the C ABI (host code): the compiled
    reference's configurations, with and without the truncations."""
    import bs2e
    t = FX["knots_c1"]
    _, n_b, _ = oracle.dims(7, t)
    term = bs2e.sym(l=L, pi=bool(L % 2))
    assert np.array_equal(bs2e.count_configs(term, 2, n_b, 7, n_b, n_b, 2), FX[f"cfg_c1_L{L}"])
    n2, nall = (int(v) for v in FX["c1_find_max_n_b"])
    assert np.array_equal(bs2e.count_configs(term, 3, n_b, 7, n2, nall, 1), FX[f"cfg_c1_trunc_L{L}"])
    # odd parity, and a cap that is too small
    odd = bs2e.sym(l=L, pi=not bool(L % 2))
    assert np.array_equal(bs2e.count_configs(odd, 2, n_b, 7, n_b, n_b, 2),
                          oracle.count_configs(L, 1 - L % 2, 2, n_b, 7, n_b, n_b, 2))
    buf = np.zeros(5 * 3, np.int64)
    assert capi.lib().bs2e_count_configs(L, L % 2, 2, n_b, 7, n_b, n_b, 2, capi.ptr(buf, C.c_int64), 3) \
        == capi.BS2E_EINVAL


def test_error_paths_without_device():
    L = capi.lib()
    h = C.c_void_p()
    rc = L.bs2e_create(C.byref(h), 0)
    if rc == capi.BS2E_OK:       # a device is present: nothing to check here
        L.bs2e_destroy(h)
        pytest.skip("a HIP device is visible")
    assert rc == capi.BS2E_EDEVICE
    assert b"no HIP device" in L.bs2e_last_error(None)
    assert L.bs2e_generate_grid(0, 3, 2, 1.0, 10.0, None, 0) == capi.BS2E_EINVAL
    t = np.ascontiguousarray(FX["knots_tiny"])
    a = C.c_int64()
    assert L.bs2e_count_nnz(21, len(t), capi.ptr(t), C.byref(a), None, None) == capi.BS2E_EINVAL
    assert b"order" in L.bs2e_last_error(None)
    with pytest.raises(capi.Bs2eError):
        import bs2e
        bs2e.Context(0)


def test_fortran_records_round_trip(tmp_path, monkeypatch):
    """Unformatted sequential records, including records split into
    subrecords (forced here by a small subrecord limit)."""
    from bs2e import CSR_matrix
    from bs2e import fortran_io as fio
    rng = np.random.default_rng(3)
    blocks = []
    for n in (3, 5, 0, 4):
        nnz = n * 2
        ip = np.concatenate([[1], 1 + np.sort(rng.integers(0, nnz + 1, n - 1)), [nnz + 1]]) if n else np.ones(1, int)
        if n:
            ip[-1] = nnz + 1
        blocks.append(CSR_matrix((n, n), np.asarray(ip, np.int64), rng.integers(1, n + 1, nnz).astype(np.int64) if n
                                 else np.zeros(0, np.int64), rng.normal(size=nnz) + 1j * rng.normal(size=nnz)))
    for limit in (fio.MAX_SUB, 24):
        monkeypatch.setattr(fio, "MAX_SUB", limit)
        p = tmp_path / f"b{limit}.dat"
        fio.store_block_diag(p, blocks)
        got = fio.load_block_diag(p)
        assert len(got) == len(blocks)
        for a, b in zip(got, blocks):
            assert a.shape == b.shape and a.nnz == b.nnz
            if b.nnz:
                assert np.array_equal(a.index_ptr, b.index_ptr) and np.array_equal(a.indices, b.indices)
                assert np.array_equal(a.data, b.data)
    raw = (tmp_path / "b24.dat").read_bytes()
    # first record 'CSR': 3 bytes in one subrecord; a 5x2 complex data record (160 bytes) is split
    assert raw[:4] == (3).to_bytes(4, "little") and raw[7:11] == (3).to_bytes(4, "little")
    assert (-24).to_bytes(4, "little", signed=True) in raw


def test_block_CS_round_trip(tmp_path):
    """block_tools.f90:386-456: a 3 x 3 matrix of CSR blocks (j outer, i
    inner), empty blocks carrying only shape and nnz; the stored shape is the
    sum over the diagonal blocks (compute_shape, :127-140, reads blocks(i,i)
    for both dimensions, so the block matrix is square)."""
    from bs2e import CSR_matrix
    from bs2e import fortran_io as fio
    rng = np.random.default_rng(5)
    rows, cols = (3, 4, 2), (3, 5, 2)
    blocks = []
    for i, n in enumerate(rows):
        row = []
        for j, m in enumerate(cols):
            if (i + j) % 2 == 0:
                row.append(CSR_matrix((n, m), np.ones(n + 1, np.int64), np.zeros(0, np.int64),
                                      np.zeros(0, np.complex128)))
                continue
            cnt = rng.integers(0, m + 1, n)
            ip = np.concatenate([[1], 1 + np.cumsum(cnt)]).astype(np.int64)
            ix = np.concatenate([np.sort(rng.choice(m, c, replace=False)) + 1 for c in cnt]).astype(np.int64)
            d = rng.normal(size=ix.size) + 1j * rng.normal(size=ix.size)
            row.append(CSR_matrix((n, m), ip, ix, d))
        blocks.append(row)
    p = tmp_path / "D_0.dat"
    fio.store_block_CS(p, blocks)
    got = fio.load_block_CS(p)
    for i in range(3):
        for j in range(3):
            a, b = got[i][j], blocks[i][j]
            assert a.shape == b.shape and a.nnz == b.nnz
            assert np.array_equal(a.index_ptr, b.index_ptr) and np.array_equal(a.indices, b.indices)
            assert np.array_equal(a.data, b.data)
    raw = p.read_bytes()
    # header records: 'CSR', block_shape (3, 3), shape (3+4+2, 3+5+2)
    assert raw[:11] == (3).to_bytes(4, "little") + b"CSR" + (3).to_bytes(4, "little")
    hdr = np.frombuffer(raw[15:31], "<i8")
    assert hdr.tolist() == [3, 3] and np.frombuffer(raw[39:55], "<i8").tolist() == [9, 10]


def test_read_namelist_formats(tmp_path):
    """basis_input.dat as gfortran and as flang write namelists (input_tools.f90:765-773)."""
    from bs2e import fortran_io as fio
    gfortran = ("&BASIS_INPUT\n BASIS_OUTPUT_DIR=\"/data/run one/                                     \",\n K=8,\n M=3,\n"
                " Z=2,\n H_MAX=  1.5000000000000000     ,\n R_MAX=  15.000000000000000     ,\n"
                " CAP_ETA=(1.0000000000000000E-003,0.0000000000000000),\n GAUGE=\"v\",\n Z_POL=T,\n FULL=F,\n"
                " TWO_EL=T,\n /\n")
    flang = (" &BASIS_INPUT BASIS_OUTPUT_DIR=\"/data/run\n one/\", K=8, M=3, Z=2, H_MAX=1.5, R_MAX=15.,\n"
             " CAP_ETA=(1.E-3,0.), GAUGE='v', Z_POL=.TRUE., FULL=.FALSE., TWO_EL=.TRUE./\n")
    for name, text in (("g", gfortran), ("f", flang)):
        p = tmp_path / f"{name}.dat"
        p.write_text(text)
        group, v = fio.read_namelist(p)
        assert group == "basis_input"
        assert v["basis_output_dir"].strip() == "/data/run one/"
        assert (v["k"], v["m"], v["z"]) == (8, 3, 2)
        assert v["h_max"] == 1.5 and v["r_max"] == 15.0 and v["cap_eta"] == 1e-3 + 0j
        assert v["gauge"] == "v" and v["z_pol"] is True and v["full"] is False and v["two_el"] is True


def test_row_windows_plan():
    """Windows cover every row once, in order, each within the byte budget
    unless a single row alone exceeds it."""
    from bs2e import row_windows
    rng = np.random.default_rng(2)
    cH = rng.integers(0, 5000, 10_000)
    cS = rng.integers(0, 800, 10_000)
    per = 24 * (cH + cS)
    for budget in (1, 24 * 3000, 10 << 20, int(per.sum()), int(per.sum()) * 2):
        wins = row_windows(cH, cS, budget)
        lo = 0
        for a, n in wins:
            assert a == lo and n >= 1
            assert n == 1 or per[a:a + n].sum() <= budget
            if a + n < len(per):   # a window stops only where the next row would not fit
                assert per[a:a + n + 1].sum() > budget
            lo = a + n
        assert lo == len(per)
    assert row_windows([], [], 100) == []
