"""The reference's output files (Fortran unformatted sequential), read and
written from Python, so that blocks built here can be handed to its `diag`
and its files read back:

    store_block_diag / load_block_diag   block_tools.f90:458-524  (H_diag.dat, S_diag.dat)
    store_block_CS / load_block_CS       block_tools.f90:386-456  (D_-1.dat, D_0.dat, D_1.dat)
    store_CS / load_CS                    sparse_array_tools.f90:695-742
    store_splines / load_splines          bspline_tools.f90:375-400 (splines.dat)
    store_basis / load_basis              orbital_tools.f90:364-397 (basis.dat, basis_info.txt)
    read_namelist                         input_tools.f90:765-773 (basis_input.dat, a namelist)

Records follow the gfortran convention (which amdflang's runtime writes as
well): a 4-byte little-endian length before and after each record; a record
longer than 2**31 - 9 bytes is split into subrecords whose leading length is
negative for every subrecord but the last and whose trailing length is
negative for every subrecord but the first.  Integers and logicals are 8
bytes (the reference builds with -fdefault-integer-8); logical .true. is 1.
"""
from __future__ import annotations

import os
import struct

import numpy as np

from . import CSR_matrix, sym

MAX_SUB = 2**31 - 9


def _write_record(f, payload: bytes):
    n = len(payload)
    if n <= MAX_SUB:
        f.write(struct.pack("<i", n))
        f.write(payload)
        f.write(struct.pack("<i", n))
        return
    pos, first = 0, True
    while pos < n:
        m = min(MAX_SUB, n - pos)
        last = pos + m == n
        f.write(struct.pack("<i", m if last else -m))
        f.write(payload[pos:pos + m])
        f.write(struct.pack("<i", m if first else -m))
        pos += m
        first = False


def _read_record(f) -> bytes:
    parts = []
    while True:
        head = f.read(4)
        if len(head) < 4:
            raise EOFError("unexpected end of file")
        n = struct.unpack("<i", head)[0]
        parts.append(f.read(abs(n)))
        f.read(4)
        if n >= 0:
            return b"".join(parts)


def _i8(*v) -> bytes:
    return np.asarray(v, dtype="<i8").tobytes()


class _Writer:
    def __init__(self, path):
        self.f = open(path, "wb")

    def rec(self, payload: bytes):
        _write_record(self.f, payload)

    def close(self):
        self.f.close()


class _Reader:
    def __init__(self, path):
        self.f = open(path, "rb")

    def raw(self) -> bytes:
        return _read_record(self.f)

    def ints(self):
        return np.frombuffer(self.raw(), dtype="<i8").astype(np.int64)

    def int(self):
        return int(self.ints()[0])

    def close(self):
        self.f.close()


# --------------------------------------------------------------------- CSR
def _put_csr_arrays(w: _Writer, m: CSR_matrix):
    w.rec(np.ascontiguousarray(m.index_ptr, dtype="<i8").tobytes())
    w.rec(np.ascontiguousarray(m.indices, dtype="<i8").tobytes())
    w.rec(np.ascontiguousarray(m.data, dtype="<c16").tobytes())


def store_CS(path, m: CSR_matrix, type_="CSR"):
    """sparse_array_tools.f90:695-708 store_CS"""
    w = _Writer(path)
    w.rec(type_.encode())
    w.rec(_i8(*m.shape))
    w.rec(_i8(m.nnz))
    _put_csr_arrays(w, m)
    w.close()


def load_CS(path) -> CSR_matrix:
    """sparse_array_tools.f90:724-742 load_CS (CSR: index_ptr of shape(1)+1)"""
    r = _Reader(path)
    kind = r.raw().decode()
    if kind != "CSR":
        raise ValueError(f"{path}: a {kind} matrix, not CSR")
    shape = tuple(int(x) for x in r.ints())
    r.int()
    ip = r.ints()
    ix = r.ints()
    d = np.frombuffer(r.raw(), dtype="<c16").copy()
    r.close()
    return CSR_matrix(shape, ip, ix, d)


def store_block_diag(path, blocks: list[CSR_matrix]):
    """block_tools.f90:458-485 CS_block_diag_store: 'CSR', block_shape,
    shape (the sum of the block shapes), then per block its shape, nnz and,
    when nnz > 0, index_ptr, indices and data."""
    n = len(blocks)
    shape = np.sum([b.shape for b in blocks], axis=0) if n else np.zeros(2, np.int64)
    w = _Writer(path)
    w.rec(b"CSR")
    w.rec(_i8(n, n))
    w.rec(_i8(*shape))
    for b in blocks:
        w.rec(_i8(*b.shape))
        w.rec(_i8(b.nnz))
        if b.nnz > 0:
            _put_csr_arrays(w, b)
    w.close()


def load_block_diag(path) -> list[CSR_matrix]:
    """block_tools.f90:487-524 CS_block_diag_load"""
    r = _Reader(path)
    kind = r.raw().decode()
    if kind != "CSR":
        raise ValueError(f"{path}: blocks must be CSR, found {kind}")
    block_shape = r.ints()
    r.ints()
    out = []
    for _ in range(int(block_shape[0])):
        shape = tuple(int(x) for x in r.ints())
        nnz = r.int()
        if nnz > 0:
            ip = r.ints()
            ix = r.ints()
            d = np.frombuffer(r.raw(), dtype="<c16").copy()
        else:
            ip = np.ones(shape[0] + 1, np.int64)
            ix = np.zeros(0, np.int64)
            d = np.zeros(0, np.complex128)
        out.append(CSR_matrix(shape, ip, ix, d))
    r.close()
    return out


def store_block_CS(path, blocks: list[list[CSR_matrix]]):
    """block_tools.f90:386-415 CS_block_store: 'CSR', block_shape, shape (the
    row counts of the diagonal blocks, then their column counts), then per
    block, column-block j outer and row-block i inner, its shape, nnz and,
    when nnz > 0, index_ptr, indices and data.  blocks[i][j]."""
    n1 = len(blocks)
    n2 = len(blocks[0]) if n1 else 0
    shape = [sum(blocks[i][i].shape[0] for i in range(n1)), sum(blocks[i][i].shape[1] for i in range(n2))]
    w = _Writer(path)
    w.rec(b"CSR")
    w.rec(_i8(n1, n2))
    w.rec(_i8(*shape))
    for j in range(n2):
        for i in range(n1):
            b = blocks[i][j]
            w.rec(_i8(*b.shape))
            w.rec(_i8(b.nnz))
            if b.nnz > 0:
                _put_csr_arrays(w, b)
    w.close()


def load_block_CS(path) -> list[list[CSR_matrix]]:
    """block_tools.f90:417-456 CS_block_load; returns blocks[i][j] (an empty
    block has index_ptr all 1)."""
    r = _Reader(path)
    kind = r.raw().decode()
    if kind != "CSR":
        raise ValueError(f"{path}: only CSR blocks are implemented, found {kind}")
    n1, n2 = (int(x) for x in r.ints())
    r.ints()
    out = [[None] * n2 for _ in range(n1)]
    for j in range(n2):
        for i in range(n1):
            shape = tuple(int(x) for x in r.ints())
            nnz = r.int()
            if nnz > 0:
                ip = r.ints()
                ix = r.ints()
                d = np.frombuffer(r.raw(), dtype="<c16").copy()
            else:
                ip = np.ones(shape[0] + 1, np.int64)
                ix = np.zeros(0, np.int64)
                d = np.zeros(0, np.complex128)
            out[i][j] = CSR_matrix(shape, ip, ix, d)
    r.close()
    return out


# --------------------------------------------------------------------- basis
def store_splines(directory, k, knots):
    """bspline_tools.f90:375-387: splines.dat = k, size(knots), knots"""
    w = _Writer(os.path.join(directory, "splines.dat"))
    w.rec(_i8(k))
    w.rec(_i8(len(knots)))
    w.rec(np.ascontiguousarray(knots, dtype="<f8").tobytes())
    w.close()


def load_splines(directory):
    r = _Reader(os.path.join(directory, "splines.dat"))
    k = r.int()
    n = r.int()
    t = np.frombuffer(r.raw(), dtype="<f8").copy()
    r.close()
    assert len(t) == n
    return k, t


def store_basis(directory, max_l_1p, max_L, two_el, syms: list[sym]):
    """orbital_tools.f90:364-397 store_basis (after compute_basis_size:
    n_states = sum of n_config, sym_ptr(i) = 1 + configurations before
    symmetry i): basis.dat and basis_info.txt."""
    n_sym = len(syms)
    n_conf = [s.n_config for s in syms]
    ptr = np.concatenate([[1], 1 + np.cumsum(n_conf)]).astype(np.int64)
    w = _Writer(os.path.join(directory, "basis.dat"))
    w.rec(_i8(max_l_1p))
    w.rec(_i8(max_L))
    w.rec(_i8(int(bool(two_el))))
    w.rec(_i8(n_sym))
    w.rec(_i8(int(sum(n_conf))))
    w.rec(ptr.astype("<i8").tobytes())
    for s in syms:
        w.rec(_i8(s.l))
        w.rec(_i8(s.m))
        w.rec(_i8(int(bool(s.pi))))
        w.rec(_i8(s.n_config))
        cf = np.zeros((s.n_config, 5), np.int64) if s.configs is None else np.asarray(s.configs, np.int64)
        w.rec(np.ascontiguousarray(cf, dtype="<i8").tobytes())   # n(2), l(2), eqv per config
    w.close()
    with open(os.path.join(directory, "basis_info.txt"), "w") as f:
        f.write(f"{'# Sym':>5}{'L':>5}{'M':>5}{'Pi':>5}{'states':>9}\n")
        for i, s in enumerate(syms, 1):
            f.write(f"{i:5d}{s.l:5d}{s.m:5d}{'T' if s.pi else 'F':>5}{s.n_config:9d}\n")


def load_basis(directory):
    """(max_l_1p, max_L, two_el, [sym])"""
    r = _Reader(os.path.join(directory, "basis.dat"))
    max_l_1p = r.int()
    max_L = r.int()
    two_el = bool(r.int())
    n_sym = r.int()
    r.int()
    r.ints()
    syms = []
    for _ in range(n_sym):
        l, m, pi, n = r.int(), r.int(), bool(r.int()), r.int()
        cf = r.ints().reshape(n, 5)
        syms.append(sym(l=l, m=m, pi=pi, configs=cf))
    r.close()
    return max_l_1p, max_L, two_el, syms


# --------------------------------------------------------------------- namelists
def read_namelist(path) -> tuple[str, dict]:
    """A Fortran namelist file such as basis_input.dat (written by
    `write(unit, nml=basis_input)`, input_tools.f90:765-773): the group name
    and its variables (lower-case names; integers, reals, complex numbers,
    logicals and strings converted)."""
    import re
    text = open(path).read()
    m = re.search(r"&(\w+)(.*?)(?:^\s*/|/\s*$)", text, re.S | re.M)
    if not m:
        raise ValueError(f"{path}: no namelist group")
    group, body = m.group(1).lower(), m.group(2)
    item = re.compile(r"(\w+)\s*=\s*('(?:[^']|'')*'|\"(?:[^\"]|\"\")*\"|\([^)]*\)|[^,\n]+)")

    def value(v):
        v = v.strip()
        if v[0] in "'\"":   # a character constant may continue over records: the record ends add nothing
            return v[1:-1].replace("\r", "").replace("\n", "").replace(v[0] * 2, v[0])
        if v.startswith("("):
            re_, im_ = (float(x.replace("d", "e").replace("D", "e")) for x in v[1:-1].split(","))
            return complex(re_, im_)
        u = v.upper().strip(".")
        if u in ("T", "TRUE"):
            return True
        if u in ("F", "FALSE"):
            return False
        try:
            return int(v)
        except ValueError:
            pass
        try:
            return float(v.replace("d", "e").replace("D", "e"))
        except ValueError:
            return v   # a string written without delimiters

    return group, {k.lower(): value(v) for k, v in item.findall(body)}
