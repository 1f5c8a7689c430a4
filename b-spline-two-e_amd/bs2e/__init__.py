"""bs2e -- host-side mirror of the reference's two-electron R^k interfaces,
running on the MI355X HIP library (libbs2e.so) through its C ABI.

Names, argument meaning and error behaviour follow edvinolo/b-spline-two-e:

    generate_grid(k, m, Z, h_max, r_max)                 grid_tools.f90:6-55
    b_spline(k, t)                                       bspline_tools.f90:4-56
    sparse_4d / sparse_6d  .init(max_k, b_splines)       sparse_array_tools.f90:9-31, 159-178
    setup_Slater_integrals(b_splines, max_k, k_GL,
                           r_k, r_m_k, r_d_k)            mat_els.f90:172-182
    Nd_DOK, compute_R_K_map(r_d_k, r_k, r_m_k,
                            b_splines, max_k, R)         sparse_array_tools.f90:452-555
    ang_k_LS(k, l1, l2, L)                               wigner_tools.f90:126-138
    sym, count_configs(term, max_l_1p, n_b, k_spline,
                       max_n_b, n_all_l, l_2_max)        orbital_tools.f90:10-19, 119-216
    CSR_matrix, construct_block_tensor(H, S, b_splines,
                       term, max_k, R_k, full)           hamiltonian.f90:106-283, 348-416

Everything numeric runs on the GPU; there is no CPU fallback.  A missing
library or device raises Bs2eError.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import capi
from .capi import Bs2eError, check, ptr

__all__ = ["Bs2eError", "Context", "generate_grid", "b_spline", "sparse_4d", "sparse_6d", "setup_Slater_integrals",
           "Nd_DOK", "RkTensor", "compute_R_K_map", "ang_k_LS", "three_j", "six_j", "count_nnz", "sym",
           "count_configs", "CSR_matrix", "Block", "construct_block_tensor", "BlockSet", "construct_blocks_tensor",
           "Server", "footprint", "setup_S", "setup_H_one_particle", "radial_dipole", "setup_radial_dip", "DipBlock",
           "construct_dip_block_tensor", "setup_r_k_int"]


class Context:
    """One device and one stream (bs2e_create)."""

    def __init__(self, device: int = 0):
        L = capi.lib()
        h = C.c_void_p()
        check(L.bs2e_create(C.byref(h), device))
        self.handle = h
        self.device = device

    def store_ceiling(self, dev_ptr: int, n_bytes: int, reps: int = 5) -> float:
        """GB/s of the library's plain store kernel over n_bytes of device memory at dev_ptr (bs2e_store_ceiling)."""
        out = C.c_double()
        check(capi.lib().bs2e_store_ceiling(self.handle, C.c_void_p(dev_ptr), int(n_bytes), int(reps), C.byref(out)),
              self.handle)
        return out.value

    def set_stream(self, stream_ptr: int | None):
        check(capi.lib().bs2e_set_stream(self.handle, C.c_void_p(stream_ptr) if stream_ptr else None), self.handle)

    def close(self):
        if self.handle:
            capi.lib().bs2e_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def generate_grid(k, m, Z, h_max, r_max) -> np.ndarray:
    """grid_tools.f90:6-55: the knot vector (divided by Z)."""
    L = capi.lib()
    n = check(L.bs2e_generate_grid(k, m, Z, float(h_max), float(r_max), None, 0))
    out = np.zeros(n)
    check(L.bs2e_generate_grid(k, m, Z, float(h_max), float(r_max), ptr(out), n))
    return out


def count_nnz(k, t):
    """(nnz4, nnz6, n_pairs): count_nnz_4d / count_nnz_6d / sqrt(count_nnz_R_k)."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
    check(capi.lib().bs2e_count_nnz(k, len(t), ptr(t), C.byref(a), C.byref(b), C.byref(c)))
    return a.value, b.value, c.value


class b_spline:
    """bspline_tools.f90:4-56: order k, knots t; n = size(t)-k, n_b = n-2,
    breakpoints = t(k:size(t)-k+1), support(iv, s) <=> iv <= s <= iv+k-1."""

    def __init__(self, k, t):
        self.k = int(k)
        self.knots = np.ascontiguousarray(t, dtype=np.float64)
        self.n = len(self.knots) - self.k
        self.n_b = self.n - 2
        self.breakpoints = self.knots[self.k - 1:len(self.knots) - self.k + 1].copy()
        niv = len(self.breakpoints) - 1
        iv = np.arange(1, niv + 1)[:, None]
        s = np.arange(1, self.n + 1)[None, :]
        self.support = (iv <= s) & (s <= iv + self.k - 1)   # (n_iv, n), 1-based semantics

    init = __init__


@dataclass
class sparse_4d:
    """sparse_array_tools.f90:9-18 (data(nnz, 0:max_k) column-major)."""
    nnz: int = 0
    data: np.ndarray = field(default_factory=lambda: np.zeros((0, 0), order="F"))
    iv: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    i: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    j: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))

    def init(self, max_k, b_splines):
        """init_4d (:159-167): count_nnz_4d, zeroed data and index arrays."""
        self.nnz = count_nnz(b_splines.k, b_splines.knots)[0]
        self.data = np.zeros((self.nnz, max_k + 1), order="F")
        self.iv = np.zeros(self.nnz, np.int64)
        self.i = np.zeros(self.nnz, np.int64)
        self.j = np.zeros(self.nnz, np.int64)
        return self


@dataclass
class sparse_6d:
    """sparse_array_tools.f90:20-31."""
    nnz: int = 0
    data: np.ndarray = field(default_factory=lambda: np.zeros((0, 0), order="F"))
    iv: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    i: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    j: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    i_p: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    j_p: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))

    def init(self, max_k, b_splines):
        """init_6d (:169-178)."""
        self.nnz = count_nnz(b_splines.k, b_splines.knots)[1]
        self.data = np.zeros((self.nnz, max_k + 1), order="F")
        for name in ("iv", "i", "j", "i_p", "j_p"):
            setattr(self, name, np.zeros(self.nnz, np.int64))
        return self


def setup_Slater_integrals(b_splines, max_k, k_GL, r_k, r_m_k, r_d_k, ctx: Context | None = None):
    """mat_els.f90:172-182 on the GPU: fills r_k, r_m_k (sparse_4d) and r_d_k
    (sparse_6d), which the caller has init'ed, in the reference's order."""
    ctx = ctx or default_context()
    t = b_splines.knots
    for c in (r_k, r_m_k, r_d_k):
        if not (c.data.flags.f_contiguous and c.data.shape[1] == max_k + 1):
            raise Bs2eError(capi.BS2E_EINVAL, "containers must be init'ed with the same max_k")
    check(capi.lib().bs2e_slater_primitives(
        ctx.handle, b_splines.k, len(t), ptr(t), k_GL, max_k,
        r_k.nnz, ptr(r_k.data), ptr(r_m_k.data), ptr(r_k.iv, C.c_int64), ptr(r_k.i, C.c_int64),
        ptr(r_k.j, C.c_int64),
        r_d_k.nnz, ptr(r_d_k.data), ptr(r_d_k.iv, C.c_int64), ptr(r_d_k.i, C.c_int64), ptr(r_d_k.i_p, C.c_int64),
        ptr(r_d_k.j, C.c_int64), ptr(r_d_k.j_p, C.c_int64)), ctx.handle)
    r_m_k.iv[:] = r_k.iv
    r_m_k.i[:] = r_k.i
    r_m_k.j[:] = r_k.j


class _ThreadErrors:
    """Stands in for a context: errors are read from the calling thread's last message."""
    handle = None


class RkTensor:
    """Device-resident banded R^k tensor R[pair(a,c), pair(b,d), k] of one basis."""

    def __init__(self, handle, ctx, owned=True):
        self.handle = handle
        self.ctx = ctx
        self.owned = owned
        info = np.zeros(11, np.int64)
        check(capi.lib().bs2e_rk_info(handle, ptr(info, C.c_int64), 11), ctx.handle)
        (self.n_b, self.n_pairs, self.K, self.nnz4, self.nnz6, self.n_iv, self.k, self.k_GL, self.U,
         self.n_knots, self.near_elems) = (int(x) for x in info)
        self.max_k = self.K - 1

    @classmethod
    def build(cls, k, t, k_GL, max_k, ctx: Context | None = None, compute=True, precision="fp64"):
        """Plans + allocates, then (optionally) enqueues the full device build.
        precision: "fp64" (plain) or "mixed" (the off-diagonal moments on the
        FP32 matrix cores from exact integer slices, BS2E_PREC_MIXED)."""
        ctx = ctx or default_context()
        t = np.ascontiguousarray(t, dtype=np.float64)
        h = C.c_void_p()
        prec = {"fp64": capi.PREC_FP64, "mixed": capi.PREC_MIXED}[precision]
        check(capi.lib().bs2e_rk_create(ctx.handle, k, len(t), ptr(t), k_GL, max_k, prec, C.byref(h)), ctx.handle)
        obj = cls(h, ctx)
        if compute:
            obj.compute()
        return obj

    def compute(self):
        check(capi.lib().bs2e_rk_compute(self.handle), self.ctx.handle)

    def synchronize(self):
        check(capi.lib().bs2e_rk_synchronize(self.handle), self.ctx.handle)

    def device_ptr(self):
        p = C.c_void_p()
        n = C.c_int64()
        check(capi.lib().bs2e_rk_device_ptr(self.handle, C.byref(p), C.byref(n)), self.ctx.handle)
        return p.value, n.value

    def pair_index(self, a, c):
        return capi.lib().bs2e_rk_pair_index(self.handle, a, c)

    def to_host(self) -> np.ndarray:
        out = np.empty(self.n_pairs * self.n_pairs * self.K)
        check(capi.lib().bs2e_rk_copy_to_host(self.handle, ptr(out), out.size), self.ctx.handle)
        return out.reshape(self.n_pairs, self.n_pairs, self.K)

    def mirror(self):
        check(capi.lib().bs2e_rk_copy_to_host(self.handle, None, 0), self.ctx.handle)

    def rows(self, rows) -> np.ndarray:
        """Tensor rows (0-based pair indices), shape (len(rows), n_pairs, K)."""
        rows = np.asarray(rows, dtype=np.int64)
        out = np.empty((len(rows), self.n_pairs, self.K))
        for i, r in enumerate(rows):
            check(capi.lib().bs2e_rk_copy_rows(self.handle, int(r), 1, ptr(out[i])), self.ctx.handle)
        return out

    def get_val(self, indices) -> np.ndarray:
        idx = np.ascontiguousarray(indices, dtype=np.int64)
        vals = np.zeros(self.K)
        check(capi.lib().bs2e_rk_get_val(self.handle, ptr(idx, C.c_int64), ptr(vals), self.K), self.ctx.handle)
        return vals

    def set_profiling(self, on=True):
        check(capi.lib().bs2e_rk_set_profiling(self.handle, int(on)), self.ctx.handle)

    def stage_times(self):
        """Device timeline averaged over the builds computed since profiling was
        switched on (or since the previous call), ms since each build's start:
        [moments + windows done, rank-one fill done, diagonal cells done,
        near elements done, end]."""
        ms = np.zeros(5)
        check(capi.lib().bs2e_rk_stage_times(self.handle, ptr(ms), 5), self.ctx.handle)
        return ms

    def autotune(self, force=False):
        """Time the fill shapes and diagonal-cell chunks on this basis; keep the fastest."""
        check(capi.lib().bs2e_rk_autotune(self.handle, int(force)), self.ctx.handle)

    def tuning(self):
        """(rank-one fill shape, multipoles per diagonal-cell workgroup) in use."""
        out = np.zeros(2, np.int64)
        check(capi.lib().bs2e_rk_tuning(self.handle, ptr(out, C.c_int64), 2), self.ctx.handle)
        return int(out[0]), int(out[1])

    def set_tuning(self, shape, kc):
        """Pins the rank-one fill shape and the diagonal-cell multipole chunk (bs2e_rk_set_tuning)."""
        check(capi.lib().bs2e_rk_set_tuning(self.handle, int(shape), int(kc)), self.ctx.handle)

    def set_graph(self, on=True):
        """Capture the build into a hipGraph at the next compute and replay it afterwards."""
        check(capi.lib().bs2e_rk_set_graph(self.handle, int(on)), self.ctx.handle)

    def release(self):
        if self.handle:
            if self.owned:
                capi.lib().bs2e_rk_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def footprint(k, t, k_GL, max_k) -> int:
    """Device bytes of one build (tables + tensor) of this basis."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    b = C.c_int64()
    check(capi.lib().bs2e_footprint(k, len(t), ptr(t), k_GL, max_k, C.byref(b)))
    return b.value


class Server:
    """A build service on one device (bs2e_server_*): jobs are admitted while a
    stream pair is free and their footprint fits the HBM budget; admitted
    builds overlap on the device."""

    QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3

    def __init__(self, hbm_bytes, slots=4, device=0):
        h = C.c_void_p()
        check(capi.lib().bs2e_server_create(device, int(hbm_bytes), slots, C.byref(h)))
        self.handle = h

    def _check(self, rc):
        if rc < 0:
            msg = capi.lib().bs2e_server_last_error(self.handle)
            raise Bs2eError(int(rc), msg.decode() if msg else "")
        return rc

    def submit(self, k, t, k_GL, max_k) -> int:
        t = np.ascontiguousarray(t, dtype=np.float64)
        return self._check(capi.lib().bs2e_server_submit(self.handle, k, len(t), ptr(t), k_GL, max_k))

    def submit_job(self, k, t, k_GL, max_k, precision="fp64", terms=(), full=True, Z=2, max_l_1p=None,
                   CAP_c=(2, 10.0, 0.0)) -> int:
        """A job of the two-electron branch: the tensor, the one-electron
        matrices (hydrogenic Z, CAP_c = (order, r_0, eta)) and the H/S blocks
        of `terms` (sym objects with their configs).  Returns the ticket."""
        t = np.ascontiguousarray(t, dtype=np.float64)
        # each symmetry's configurations by pointer (the service copies them)
        cfs = [np.ascontiguousarray(tm.configs, dtype=np.int64) for tm in terms]
        L = np.array([int(tm.l) for tm in terms], np.int64)
        n_config = np.array([c.shape[0] for c in cfs], np.int64)
        if max_l_1p is None:
            max_l_1p = max((int(c[:, 2:4].max()) for c in cfs if len(c)), default=0)
        order, r0, eta = CAP_c
        eta = complex(eta)
        I = C.c_int64
        per_sym = (C.POINTER(I) * max(len(cfs), 1))(*[ptr(c, I) for c in cfs])
        job = capi.Job(k=k, n_knots=len(t), knots=ptr(t), k_gl=k_GL, max_k=max_k,
                       precision={"fp64": capi.PREC_FP64, "mixed": capi.PREC_MIXED}[precision],
                       n_sym=len(cfs), L=ptr(L, I), n_config=ptr(n_config, I), configs=None,
                       full=int(bool(full)), Z=int(Z), max_l_1p=int(max_l_1p), cap_order=int(order), cap_r0=float(r0),
                       cap_eta_re=eta.real, cap_eta_im=eta.imag, sym_configs=per_sym)
        return self._check(capi.lib().bs2e_server_submit_job(self.handle, C.byref(job)))

    def blocks(self, ticket, terms):
        """The finished job's blocks (owned by the service until the job is
        retired), read as a BlockSet of `terms` (the job's symmetries)."""
        h = C.c_void_p()
        self._check(capi.lib().bs2e_server_blocks(self.handle, ticket, C.byref(h)))
        return BlockSet.wrap(h, terms)

    def cache_info(self):
        """(entries, device bytes, admissions served from the cache, admissions that computed their tables)"""
        v = np.zeros(4, np.int64)
        self._check(capi.lib().bs2e_server_cache(self.handle, ptr(v, C.c_int64), 4))
        return tuple(int(x) for x in v)

    def progress(self) -> int:
        return self._check(capi.lib().bs2e_server_progress(self.handle))

    def status(self, ticket):
        """(state, ms from submission to admission, ms from submission to completion)"""
        st = C.c_int64()
        tm = np.zeros(2)
        self._check(capi.lib().bs2e_server_status(self.handle, ticket, C.byref(st), ptr(tm)))
        return st.value, float(tm[0]), float(tm[1])

    def wait(self, ticket):
        self._check(capi.lib().bs2e_server_wait(self.handle, ticket))

    def tensor(self, ticket) -> RkTensor:
        h = C.c_void_p()
        self._check(capi.lib().bs2e_server_tensor(self.handle, ticket, C.byref(h)))
        return RkTensor(h, _ThreadErrors(), owned=False)

    def last_error(self) -> str:
        msg = capi.lib().bs2e_server_last_error(self.handle)
        return msg.decode() if msg else ""

    def finished(self, cap=64):
        """Progresses the service; [(ticket, latency ms)] of the jobs done since the previous call (latency -1: the
        job failed, last_error() says why)."""
        tk = np.zeros(cap, np.int64)
        lat = np.zeros(cap)
        n = self._check(capi.lib().bs2e_server_finished(self.handle, ptr(tk, C.c_int64), ptr(lat), cap))
        return [(int(tk[i]), float(lat[i])) for i in range(n)]

    def retire(self, ticket):
        self._check(capi.lib().bs2e_server_retire(self.handle, ticket))

    def memory(self):
        """(budget bytes, bytes held by admitted or finished jobs)"""
        m = np.zeros(2, np.int64)
        self._check(capi.lib().bs2e_server_memory(self.handle, ptr(m, C.c_int64)))
        return int(m[0]), int(m[1])

    def close(self):
        if self.handle:
            capi.lib().bs2e_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Nd_DOK:
    """sparse_array_tools.f90:49-61 replacement: the banded R^k tensor held on
    the device with arithmetic key -> offset instead of a hash map.  get_val
    keeps the reference's signature (indices(4) -> values(0:max_k))."""

    def __init__(self):
        self.tensor: RkTensor | None = None

    def get_val(self, indices, values=None):
        if self.tensor is None:
            raise Bs2eError(capi.BS2E_ESTATE, "Nd_DOK not built (call compute_R_K_map first)")
        v = self.tensor.get_val(indices)
        if values is not None:
            values[:] = v
        return v


def compute_R_K_map(r_d_k, r_k, r_m_k, b_splines, max_k, R: Nd_DOK, ctx: Context | None = None):
    """sparse_array_tools.f90:452-493: builds R from the given containers on the GPU."""
    ctx = ctx or default_context()
    t = b_splines.knots
    rk = np.asfortranarray(r_k.data)
    rmk = np.asfortranarray(r_m_k.data)
    rdk = np.asfortranarray(r_d_k.data)
    h = C.c_void_p()
    I = C.c_int64
    check(capi.lib().bs2e_rk_from_primitives(
        ctx.handle, b_splines.k, len(t), ptr(t), max_k,
        r_k.nnz, ptr(rk), ptr(rmk), ptr(r_k.iv, I), ptr(r_k.i, I), ptr(r_k.j, I),
        r_d_k.nnz, ptr(rdk), ptr(r_d_k.iv, I), ptr(r_d_k.i, I), ptr(r_d_k.i_p, I), ptr(r_d_k.j, I),
        ptr(r_d_k.j_p, I), C.byref(h)), ctx.handle)
    R.tensor = RkTensor(h, ctx)
    R.tensor.compute()
    R.tensor.mirror()
    return R


def setup_S(b_splines, k_GL, ctx: Context | None = None) -> np.ndarray:
    """mat_els.f90:85-118: the overlap matrix (n_b x n_b complex)."""
    ctx = ctx or default_context()
    t = b_splines.knots
    n_b = b_splines.n_b
    S = np.zeros(n_b * n_b, np.complex128)
    check(capi.lib().bs2e_setup_S(ctx.handle, b_splines.k, len(t), ptr(t), k_GL,
                                  S.ctypes.data_as(C.POINTER(C.c_double))), ctx.handle)
    return S.reshape(n_b, n_b, order="F")


def setup_H_one_particle(Z, CAP_c, l, b_splines, k_GL, ctx: Context | None = None) -> np.ndarray:
    """mat_els.f90:47-83 with potentials.f90 hydrogenic(Z) and CAP_c = (order, r_0, eta)
    (CAP_tools.f90); l an int or a range of l's (then an (n_l, n_b, n_b) array)."""
    ctx = ctx or default_context()
    t = b_splines.knots
    n_b = b_splines.n_b
    ls = [l] if np.isscalar(l) else list(l)
    l_min, l_max = min(ls), max(ls)
    order, r0, eta = CAP_c
    eta = complex(eta)
    H = np.zeros((l_max - l_min + 1) * n_b * n_b, np.complex128)
    check(capi.lib().bs2e_setup_H_one_particle(ctx.handle, b_splines.k, len(t), ptr(t), k_GL, int(Z), l_min, l_max,
                                               int(order), float(r0), eta.real, eta.imag,
                                               H.ctypes.data_as(C.POINTER(C.c_double))), ctx.handle)
    H = H.reshape(l_max - l_min + 1, n_b, n_b).transpose(0, 2, 1)   # each block column-major
    out = np.stack([H[x - l_min] for x in ls])
    return out[0] if np.isscalar(l) else out


def ang_k_LS(k, l1, l2, L):
    """wigner_tools.f90:126-138 with l1 = confs(1)%l, l2 = confs(2)%l."""
    return capi.lib().bs2e_ang_k_LS(k, l1[0], l1[1], l2[0], l2[1], L)


def three_j(*a):
    return capi.lib().bs2e_three_j(*[int(x) for x in a])


def six_j(*a):
    return capi.lib().bs2e_six_j(*[int(x) for x in a])


@dataclass
class sym:
    """orbital_tools.f90:10-19: a two-electron symmetry (total L, M, parity pi)
    and its configurations, one row (n1, n2, l1, l2, eqv) per config%."""
    l: int
    m: int = 0
    pi: bool = False
    configs: np.ndarray | None = None

    @property
    def n_config(self) -> int:
        return 0 if self.configs is None else int(self.configs.shape[0])


def count_configs(term: sym, max_l_1p, n_b, k_spline, max_n_b, n_all_l, l_2_max, eigs=None) -> np.ndarray:
    """orbital_tools.f90:119-216: the configurations of `term` as an (n, 5)
    int64 array (n1, n2, l1, l2, eqv).  `eigs` is accepted for signature
    parity; the reference computes an energy from it that no filter uses."""
    L = capi.lib()
    args = [int(term.l), int(bool(term.pi)), max_l_1p, n_b, k_spline, max_n_b, n_all_l, l_2_max]
    n = check(L.bs2e_count_configs(*args, None, 0))
    out = np.zeros((max(n, 1), 5), np.int64)
    check(L.bs2e_count_configs(*args, ptr(out, C.c_int64), n))
    return out[:n]


@dataclass
class CSR_matrix:
    """sparse_array_tools.f90:127-135: 1-based CSR (index_ptr, indices) with
    double complex data."""
    shape: tuple
    index_ptr: np.ndarray
    indices: np.ndarray
    data: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.indices.size)

    def get_dense(self) -> np.ndarray:
        out = np.zeros(self.shape, np.complex128)
        for r in range(self.shape[0]):
            lo, hi = self.index_ptr[r] - 1, self.index_ptr[r + 1] - 1
            out[r, self.indices[lo:hi] - 1] += self.data[lo:hi]
        return out


class Block:
    """One symmetry block on the device: configurations, CSR pattern counts
    (hamiltonian.f90:348-416) and the fill of H_sp / S_sp."""

    def __init__(self, tensor: RkTensor, term: sym, full=True):
        cf = np.ascontiguousarray(term.configs, dtype=np.int64)
        self.tensor = tensor
        self.ctx = tensor.ctx
        self.n_config = int(cf.shape[0])
        h = C.c_void_p()
        check(capi.lib().bs2e_block_create(tensor.handle, int(term.l), self.n_config, ptr(cf, C.c_int64),
                                           int(bool(full)), C.byref(h)), self.ctx.handle)
        self.handle = h
        sz = np.zeros(5, np.int64)
        check(capi.lib().bs2e_block_sizes(h, ptr(sz, C.c_int64), 5), self.ctx.handle)
        self.nnz_H, self.nnz_S, self.ref_nnz_H, self.ref_nnz_S = (int(x) for x in sz[:4])

    def fill(self, H, S):
        """H: one-electron blocks per l (n_b x n_b each, l = 0 first); S: overlap."""
        n_b = self.tensor.n_b
        Hl = np.ascontiguousarray(np.stack([np.asarray(h, np.complex128).ravel(order="F") for h in H]))
        Sf = np.ascontiguousarray(np.asarray(S, np.complex128).ravel(order="F"))
        if Hl.shape[1] != n_b * n_b or Sf.size != n_b * n_b:
            raise Bs2eError(capi.BS2E_EINVAL, f"one-electron matrices must be {n_b} x {n_b}")
        n = self.n_config
        ipH = np.zeros(n + 1, np.int64)
        ipS = np.zeros(n + 1, np.int64)
        ixH = np.zeros(max(self.nnz_H, 1), np.int64)
        ixS = np.zeros(max(self.nnz_S, 1), np.int64)
        dH = np.zeros(max(self.nnz_H, 1), np.complex128)
        dS = np.zeros(max(self.nnz_S, 1), np.complex128)
        I = C.c_int64
        D = C.POINTER(C.c_double)
        check(capi.lib().bs2e_block_fill(self.handle, len(H) - 1, Hl.ctypes.data_as(D), Sf.ctypes.data_as(D),
                                         ptr(ipH, I), ptr(ixH, I), dH.ctypes.data_as(D), ptr(ipS, I), ptr(ixS, I),
                                         dS.ctypes.data_as(D)), self.ctx.handle)
        H_sp = CSR_matrix((n, n), ipH, ixH[:self.nnz_H], dH[:self.nnz_H])
        S_sp = CSR_matrix((n, n), ipS, ixS[:self.nnz_S], dS[:self.nnz_S])
        return H_sp, S_sp

    def kernel_ms(self):
        """(count kernel, fill kernel) device ms, HIP events."""
        ms = np.zeros(2)
        check(capi.lib().bs2e_block_times(self.handle, ptr(ms), 2), self.ctx.handle)
        return float(ms[0]), float(ms[1])

    def release(self):
        if self.handle:
            capi.lib().bs2e_block_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def _one_electron_arrays(H, S, n_b):
    Hl = np.ascontiguousarray(np.stack([np.asarray(h, np.complex128).ravel(order="F") for h in H]))
    Sf = np.ascontiguousarray(np.asarray(S, np.complex128).ravel(order="F"))
    if Hl.shape[1] != n_b * n_b or Sf.size != n_b * n_b:
        raise Bs2eError(capi.BS2E_EINVAL, f"one-electron matrices must be {n_b} x {n_b}")
    return Hl, Sf


def row_windows(cnt_H, cnt_S, max_bytes):
    """Consecutive windows (row_lo, n_rows) covering rows 0 .. len(cnt_H)-1,
    each as long as it stays within max_bytes of CSR entries (24 bytes per H
    or S entry: column index and complex value); a row that alone needs more
    is a window of its own."""
    per = 24 * (np.asarray(cnt_H, np.int64) + np.asarray(cnt_S, np.int64))
    out, lo, acc = [], 0, 0
    for r, b in enumerate(per.tolist()):
        if r > lo and acc + b > max_bytes:
            out.append((lo, r - lo))
            lo, acc = r, 0
        acc += b
    if len(per) > lo:
        out.append((lo, len(per) - lo))
    return out


class BlockSet:
    """The H/S blocks of several symmetries against one device tensor, counted
    in one launch and filled in one launch over the configuration rows of all
    of them (main_basis_setup.f90:104-115 runs construct_block_tensor per
    symmetry)."""

    owned = True

    @classmethod
    def wrap(cls, handle, terms):
        """A set the library owns (a service job's blocks): read-only access."""
        obj = cls.__new__(cls)
        obj.owned = False
        obj.tensor = None
        obj.ctx = _ThreadErrors()
        obj.terms = list(terms)
        obj.n_config = [int(np.asarray(t.configs).shape[0]) for t in obj.terms]
        obj.handle = handle
        sz = np.zeros(5 * len(obj.terms), np.int64)
        check(capi.lib().bs2e_blockset_sizes(handle, ptr(sz, C.c_int64), sz.size))
        obj.sizes = sz.reshape(-1, 5)
        return obj

    def __init__(self, tensor: RkTensor, terms, full=True):
        self.tensor = tensor
        self.ctx = tensor.ctx
        self.terms = list(terms)
        cfs = [np.ascontiguousarray(t.configs, dtype=np.int64) for t in self.terms]
        self.n_config = [int(c.shape[0]) for c in cfs]
        Ls = np.array([int(t.l) for t in self.terms], np.int64)
        ns = np.array(self.n_config, np.int64)
        allc = np.ascontiguousarray(np.concatenate(cfs, axis=0)) if cfs else np.zeros((0, 5), np.int64)
        h = C.c_void_p()
        I = C.c_int64
        check(capi.lib().bs2e_blockset_create(tensor.handle, len(self.terms), ptr(Ls, I), ptr(ns, I), ptr(allc, I),
                                              int(bool(full)), C.byref(h)), self.ctx.handle)
        self.handle = h
        sz = np.zeros(5 * len(self.terms), np.int64)
        check(capi.lib().bs2e_blockset_sizes(h, ptr(sz, I), sz.size), self.ctx.handle)
        self.sizes = sz.reshape(-1, 5)

    def fill(self, H, S):
        """One launch over every row; results stay on the device (get())."""
        Hl, Sf = _one_electron_arrays(H, S, self.tensor.n_b)
        D = C.POINTER(C.c_double)
        check(capi.lib().bs2e_blockset_fill(self.handle, len(H) - 1, Hl.ctypes.data_as(D), Sf.ctypes.data_as(D)),
              self.ctx.handle)

    def get(self, i):
        """(H_sp, S_sp) of symmetry i (0-based position in `terms`)."""
        if not 0 <= i < len(self.terms):
            raise Bs2eError(capi.BS2E_ERANGE, f"symmetry index {i} outside 0..{len(self.terms) - 1}")
        n = self.n_config[i]
        nH, nS = int(self.sizes[i, 0]), int(self.sizes[i, 1])
        ipH = np.zeros(n + 1, np.int64)
        ipS = np.zeros(n + 1, np.int64)
        ixH = np.zeros(max(nH, 1), np.int64)
        ixS = np.zeros(max(nS, 1), np.int64)
        dH = np.zeros(max(nH, 1), np.complex128)
        dS = np.zeros(max(nS, 1), np.complex128)
        I = C.c_int64
        D = C.POINTER(C.c_double)
        check(capi.lib().bs2e_blockset_get(self.handle, i + 1, ptr(ipH, I), ptr(ixH, I), dH.ctypes.data_as(D),
                                           ptr(ipS, I), ptr(ixS, I), dS.ctypes.data_as(D)), self.ctx.handle)
        return (CSR_matrix((n, n), ipH, ixH[:nH], dH[:nH]), CSR_matrix((n, n), ipS, ixS[:nS], dS[:nS]))

    @property
    def row0(self):
        """First global row of each symmetry (and the total row count last)."""
        return np.concatenate([[0], np.cumsum(self.n_config)]).astype(np.int64)

    def row_counts(self, row_lo=0, n_rows=None):
        """H and S entries of global rows row_lo .. row_lo+n_rows-1 (the count pass)."""
        if n_rows is None:
            n_rows = int(self.row0[-1]) - row_lo
        I = C.c_int64
        cH = np.zeros(max(n_rows, 1), np.int64)
        cS = np.zeros(max(n_rows, 1), np.int64)
        check(capi.lib().bs2e_blockset_row_counts(self.handle, row_lo, n_rows, ptr(cH, I), ptr(cS, I)),
              self.ctx.handle)
        return cH[:n_rows], cS[:n_rows]

    def windows(self, max_bytes):
        """Consecutive windows (row_lo, n_rows) of global rows covering the
        set, each holding at most max_bytes of CSR entries (row_windows)."""
        return row_windows(*self.row_counts(), max_bytes)

    def reserve(self, nnz_H, nnz_S):
        """Output buffers with room for nnz_H and nnz_S entries (allocated once for a stream of windows)."""
        check(capi.lib().bs2e_blockset_reserve(self.handle, int(nnz_H), int(nnz_S)), self.ctx.handle)

    def stream(self, max_bytes, H=None, S=None):
        """Fills the set window by window (windows(max_bytes)), the outputs
        reserved once for the largest window, yielding (row_lo, n_rows) after
        each window's fill; H and S as fill_rows (the first window's fill
        takes them)."""
        wins = self.windows(max_bytes)
        cH, cS = self.row_counts()
        pH = np.concatenate([[0], np.cumsum(cH)])
        pS = np.concatenate([[0], np.cumsum(cS)])
        self.reserve(max((int(pH[a + n] - pH[a]) for a, n in wins), default=0),
                     max((int(pS[a + n] - pS[a]) for a, n in wins), default=0))
        for k, (lo, n) in enumerate(wins):
            if k == 0 and H is not None:
                self.fill_rows(lo, n, H, S)
            else:
                self.fill_rows(lo, n)
            yield lo, n

    def fill_rows(self, row_lo, n_rows, H=None, S=None):
        """Fills global rows row_lo .. row_lo+n_rows-1 only (the outputs hold
        that window); H and S None reuse the last fill's one-electron matrices."""
        D = C.POINTER(C.c_double)
        if H is None and S is None:
            check(capi.lib().bs2e_blockset_fill_rows(self.handle, 0, None, None, int(row_lo), int(n_rows)),
                  self.ctx.handle)
            return
        Hl, Sf = _one_electron_arrays(H, S, self.tensor.n_b)
        check(capi.lib().bs2e_blockset_fill_rows(self.handle, len(H) - 1, Hl.ctypes.data_as(D),
                                                 Sf.ctypes.data_as(D), int(row_lo), int(n_rows)), self.ctx.handle)

    def refill(self):
        """The fill again (all rows, or the last window) from the device copy of the last fill's one-electron
        matrices (waits)."""
        check(capi.lib().bs2e_blockset_refill(self.handle), self.ctx.handle)

    def set_graph(self, on=True):
        """Captures the whole branch (tensor build, count, row offsets, fill) as one hipGraph."""
        check(capi.lib().bs2e_blockset_set_graph(self.handle, int(bool(on))), self.ctx.handle)

    def replay(self):
        """One launch of the captured branch (asynchronous; synchronize() waits)."""
        check(capi.lib().bs2e_blockset_replay(self.handle), self.ctx.handle)

    def synchronize(self):
        check(capi.lib().bs2e_blockset_synchronize(self.handle), self.ctx.handle)

    def row_offsets(self, i, row_lo, n_rows):
        """index_ptr of H and S for rows row_lo .. row_lo+n_rows-1 of symmetry
        i, relative to the window (1-based); nothing else is copied."""
        if not 0 <= i < len(self.terms):
            raise Bs2eError(capi.BS2E_ERANGE, f"symmetry index {i} outside 0..{len(self.terms) - 1}")
        I = C.c_int64
        ipH = np.zeros(n_rows + 1, np.int64)
        ipS = np.zeros(n_rows + 1, np.int64)
        check(capi.lib().bs2e_blockset_get_rows(self.handle, i + 1, row_lo, n_rows, ptr(ipH, I), None, None,
                                                ptr(ipS, I), None, None), self.ctx.handle)
        return ipH, ipS

    def get_rows(self, i, row_lo, n_rows):
        """Rows row_lo .. row_lo+n_rows-1 of symmetry i (0-based position in
        `terms`): (H_sp, S_sp) of shape (n_rows, n_config) with index_ptr
        relative to the window (1-based)."""
        if not 0 <= i < len(self.terms):
            raise Bs2eError(capi.BS2E_ERANGE, f"symmetry index {i} outside 0..{len(self.terms) - 1}")
        I = C.c_int64
        D = C.POINTER(C.c_double)
        ipH = np.zeros(n_rows + 1, np.int64)
        ipS = np.zeros(n_rows + 1, np.int64)
        lib = capi.lib()
        check(lib.bs2e_blockset_get_rows(self.handle, i + 1, row_lo, n_rows, ptr(ipH, I), None, None, ptr(ipS, I),
                                         None, None), self.ctx.handle)
        nH, nS = int(ipH[-1] - 1), int(ipS[-1] - 1)
        ixH = np.zeros(max(nH, 1), np.int64)
        ixS = np.zeros(max(nS, 1), np.int64)
        dH = np.zeros(max(nH, 1), np.complex128)
        dS = np.zeros(max(nS, 1), np.complex128)
        check(lib.bs2e_blockset_get_rows(self.handle, i + 1, row_lo, n_rows, ptr(ipH, I), ptr(ixH, I),
                                         dH.ctypes.data_as(D), ptr(ipS, I), ptr(ixS, I), dS.ctypes.data_as(D)),
              self.ctx.handle)
        n = self.n_config[i]
        return (CSR_matrix((n_rows, n), ipH, ixH[:nH], dH[:nH]), CSR_matrix((n_rows, n), ipS, ixS[:nS], dS[:nS]))

    def kernel_ms(self):
        """(count launch, fill launch) device ms, HIP events."""
        ms = np.zeros(2)
        check(capi.lib().bs2e_blockset_times(self.handle, ptr(ms), 2), self.ctx.handle)
        return float(ms[0]), float(ms[1])

    def release(self):
        if self.handle and self.owned:
            capi.lib().bs2e_blockset_release(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def construct_blocks_tensor(H, S, b_splines, terms, max_k, R_k, full=True):
    """construct_block_tensor (hamiltonian.f90:106-283) for every symmetry of
    `terms` in one count and one fill launch: a list of (H_sp, S_sp)."""
    tensor = R_k.tensor if isinstance(R_k, Nd_DOK) else R_k
    if tensor is None:
        raise Bs2eError(capi.BS2E_ESTATE, "R_k holds no tensor (call compute_R_K_map first)")
    if tensor.K != max_k + 1 or tensor.k != b_splines.k:
        raise Bs2eError(capi.BS2E_EINVAL, "max_k / b_splines do not match the tensor")
    bset = BlockSet(tensor, terms, full)
    try:
        bset.fill(H, S)
        return [bset.get(i) for i in range(len(bset.terms))]
    finally:
        bset.release()


def construct_block_tensor(H, S, b_splines, term: sym, max_k, R_k, full=True):
    """hamiltonian.f90:106-283: returns (H_sp, S_sp) for `term` from the
    device tensor held by R_k (an Nd_DOK from compute_R_K_map, or an RkTensor).
    H: one-electron blocks H(l)%data for l = 0.. (n_b x n_b complex each)."""
    tensor = R_k.tensor if isinstance(R_k, Nd_DOK) else R_k
    if tensor is None:
        raise Bs2eError(capi.BS2E_ESTATE, "R_k holds no tensor (call compute_R_K_map first)")
    if tensor.K != max_k + 1 or tensor.k != b_splines.k:
        raise Bs2eError(capi.BS2E_EINVAL, "max_k / b_splines do not match the tensor")
    blk = Block(tensor, term, full)
    try:
        return blk.fill(H, S)
    finally:
        blk.release()


def setup_r_k_int(b_splines, k_GL, max_k, r_0, ctx: Context | None = None) -> np.ndarray:
    """RKH_hamiltonian.f90:91-159 setup_r_k_int on the device: the array
    r_k_int[k, i, j] (k = 0..max_k; real) for the multipole expansion about
    r_0 (adiabatic_potentials.f90:24-34 r_k)."""
    ctx = ctx or default_context()
    t = b_splines.knots
    n_b = b_splines.n_b
    out = np.zeros((max_k + 1) * n_b * n_b, np.float64)
    check(capi.lib().bs2e_setup_r_k_int(ctx.handle, b_splines.k, len(t), ptr(t), k_GL, int(max_k), float(r_0),
                                        ptr(out)), ctx.handle)
    return out.reshape(max_k + 1, n_b, n_b, order="F")


# --------------------------------------------------------------------- dipoles
@dataclass
class radial_dipole:
    """mat_els.f90:14-19: the radial dipole matrices of one gauge ('l': r_mat;
    'v': dr_mat and r_inv_mat), n_b x n_b complex."""
    gauge: str = "l"
    r_mat: np.ndarray | None = None
    dr_mat: np.ndarray | None = None
    r_inv_mat: np.ndarray | None = None


def _gauge_code(gauge) -> int:
    if gauge not in ("l", "v"):
        raise Bs2eError(capi.BS2E_EINVAL, f"Unrecognized gauge option for radial dipole: {gauge!r}")
    return 0 if gauge == "l" else 1


def setup_radial_dip(b_splines, k_GL, gauge="l", ctx: Context | None = None) -> radial_dipole:
    """mat_els.f90:120-170 setup_radial_dip(b_splines, k_GL, radial_dip, gauge)
    on the device (:348-390 compute_radial_dip_len / _vel)."""
    ctx = ctx or default_context()
    g = _gauge_code(gauge)
    t = b_splines.knots
    n_b = b_splines.n_b
    m1 = np.zeros(n_b * n_b, np.complex128)
    m2 = np.zeros(n_b * n_b if g else 1, np.complex128)
    D = C.POINTER(C.c_double)
    check(capi.lib().bs2e_setup_radial_dip(ctx.handle, b_splines.k, len(t), ptr(t), k_GL, g, m1.ctypes.data_as(D),
                                           m2.ctypes.data_as(D)), ctx.handle)
    m1 = m1.reshape(n_b, n_b, order="F")
    if g == 0:
        return radial_dipole("l", r_mat=m1)
    return radial_dipole("v", dr_mat=m1, r_inv_mat=m2.reshape(n_b, n_b, order="F"))


class DipBlock:
    """One dipole block between two symmetries on the device (dipole.f90:95-150
    init_dip_block counts; fill evaluates dip_mat_neq for a gauge)."""

    def __init__(self, syms, q, b_splines, compute=True, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        s1, s2 = syms
        cf1 = np.ascontiguousarray(s1.configs, dtype=np.int64)
        cf2 = np.ascontiguousarray(s2.configs, dtype=np.int64)
        self.shape = (int(cf1.shape[0]), int(cf2.shape[0]))
        self.n_b = int(b_splines.n_b)
        h = C.c_void_p()
        I = C.c_int64
        check(capi.lib().bs2e_dip_block_create(
            self.ctx.handle, int(b_splines.k), self.n_b, int(s1.l), int(s1.m), int(bool(s1.pi)), self.shape[0],
            ptr(cf1, I), int(s2.l), int(s2.m), int(bool(s2.pi)), self.shape[1], ptr(cf2, I), int(q),
            int(bool(compute)), C.byref(h)), self.ctx.handle)
        self.handle = h
        sz = np.zeros(3, np.int64)
        check(capi.lib().bs2e_dip_block_sizes(h, ptr(sz, I), 3), self.ctx.handle)
        self.nnz = int(sz[0])

    def fill(self, S, radial_dip: radial_dipole) -> CSR_matrix:
        n_b = self.n_b
        g = _gauge_code(radial_dip.gauge)
        mats = [S, radial_dip.r_mat] if g == 0 else [S, radial_dip.dr_mat, radial_dip.r_inv_mat]
        flat = []
        for M in mats:
            M = np.asarray(M, np.complex128)
            if M.shape != (n_b, n_b):
                raise Bs2eError(capi.BS2E_EINVAL, f"S and the radial dipole matrices must be {n_b} x {n_b}")
            flat.append(np.ascontiguousarray(M.ravel(order="F")))
        if g == 0:
            flat.append(flat[1])
        n1 = self.shape[0]
        ip = np.zeros(n1 + 1, np.int64)
        ix = np.zeros(max(self.nnz, 1), np.int64)
        d = np.zeros(max(self.nnz, 1), np.complex128)
        D = C.POINTER(C.c_double)
        I = C.c_int64
        check(capi.lib().bs2e_dip_block_fill(self.handle, g, flat[0].ctypes.data_as(D), flat[1].ctypes.data_as(D),
                                             flat[2].ctypes.data_as(D), ptr(ip, I), ptr(ix, I), d.ctypes.data_as(D)),
              self.ctx.handle)
        return CSR_matrix(self.shape, ip, ix[:self.nnz], d[:self.nnz])

    def kernel_ms(self):
        """(count kernel, fill kernel) device ms, HIP events."""
        ms = np.zeros(2)
        check(capi.lib().bs2e_dip_block_times(self.handle, ptr(ms), 2), self.ctx.handle)
        return float(ms[0]), float(ms[1])

    def release(self):
        if self.handle:
            capi.lib().bs2e_dip_block_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def construct_dip_block_tensor(syms, q, b_splines, S, radial_dip: radial_dipole, compute=True,
                               ctx: Context | None = None) -> CSR_matrix:
    """dipole.f90:8-47 construct_dip_block_tensor(syms, q, b_splines, S,
    radial_dip, dip_block, compute): the CSR dipole block between the
    configurations of syms[0] (rows) and syms[1] (columns) for spherical
    component q, on the device."""
    blk = DipBlock(syms, q, b_splines, compute, ctx)
    try:
        return blk.fill(S, radial_dip)
    finally:
        blk.release()
