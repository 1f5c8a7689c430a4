#!/usr/bin/env python3
"""bench.py -- R^k integrals/s of the device R^k build on one MI355X.

One step = one complete build of the banded R^k tensor of a basis from its
knot vector: de Boor tables, pair moments, diagonal cells and the tensor
(the reference's setup_Slater_integrals + compute_R_K_map), with the knots
and every work buffer already resident in HBM.  Integrals per step =
n_pairs^2 * (max_k+1), the key count of the reference's Nd_DOK times its
value length.

    python bench.py [--steps K] [--warmup W] [--workload c5p]

The default workload is the largest single-GPU configuration: order 9, 600
B-splines, r <= 1000, GL 14, k <= 24 (20.5 GB of FP64 tensor), with its
interior breakpoints moved by a seeded (42) factor in [0.99, 1.01].

Prints one JSON line.  The dominant kernel (the rank-one tensor fill) is
timed with HIP events on the stream it runs on and priced against HBM (8
bytes stored per integral, 8 TB/s peak).  cpu_baseline times the oracle (the
repository's C restatement of the reference path) on the host, on the
reference's schedule, over bounded samples of the same workload, and
extrapolates to the whole build.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "b-spline-two-e_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "R^k integrals/sec (mixed vs FP64 GPU) at <=1e-10 rel err; p99 build latency"
HBM_PEAK_GBS = 8000.0
# HBM bytes per k3_far launch, FETCH_SIZE + WRITE_SIZE from rocprofv3 --pmc,
# each scaled by what the counter reads for a device copy of known size
# (profiles/rk_build_c2.md, profiles/rk_build_c5p.md), per (workload, fill
# shape); absent for pairs not measured that way.
FAR_TRAFFIC_BYTES = {("c2", 0): 739.26e6, ("c2", 1): 739.82e6, ("c2", 2): 740.34e6, ("c2", 3): 739.93e6,
                     ("c2", 4): 739.86e6, ("c2", 5): 739.71e6, ("c2", 6): 739.04e6,
                     ("c5p", 0): 20260.12e6, ("c5p", 1): 20263.29e6, ("c5p", 2): 20273.51e6,
                     ("c5p", 3): 20267.81e6, ("c5p", 4): 20262.35e6, ("c5p", 5): 20271.02e6,
                     ("c5p", 6): 20259.39e6}


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def host_threads():
    """Threads for the host baseline: the CPUs this process may run on,
    capped at 16 (the GPU host's share per GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def sample_rows(n_pairs, n, seed=3):
    """n distinct tensor rows: the first, the last and n-2 drawn ones."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return np.unique(np.concatenate([[0, n_pairs - 1], rng.choice(n_pairs, size=n - 2, replace=False)]))


def oracle_rows(w, t, rows, nthreads):
    """The oracle's rows `rows` of the tensor (its diagonal cells computed only
    where they land in these rows): bit for bit its full tensor's rows."""
    import numpy as np

    from oracle import oracle as O
    n_b = O.dims(w.k, t)[1]
    mask = np.zeros(O.n_pairs(w.k, n_b), bool)
    mask[rows] = True
    off = O.slater_off_diag(w.k, t, w.k_GL, w.max_k, nthreads=nthreads)
    dg = O.slater_diag(w.k, t, w.k_GL, w.max_k, nthreads=nthreads, pair_mask=mask)
    return O.rk_rows(w.k, n_b, w.max_k, off, dg, rows)


def error_stats(got, ref):
    """Per multipole k: the largest element-wise relative error and the
    normwise error max|got - ref| / max|ref|; over all elements the p50 /
    p99 / max of the element-wise relative error (every R^k value is > 0)."""
    import numpy as np
    K = ref.shape[-1]
    g = got.reshape(-1, K)
    r = ref.reshape(-1, K)
    d = np.abs(g - r)
    rel = d / r
    return {"per_k_max_rel": [float(x) for x in rel.max(axis=0)],
            "per_k_normwise": [float(x) for x in d.max(axis=0) / np.abs(r).max(axis=0)],
            "elementwise_rel": {"p50": float(np.percentile(rel, 50)), "p99": float(np.percentile(rel, 99)),
                                "max": float(rel.max())},
            "elements": int(rel.size)}


def cpu_baseline_reference_schedule(w, t, nthreads, diag_target_s=4.0, dok_target_s=3.0):
    """The whole build on the reference's schedule, timed on the host over
    bounded samples and extrapolated:
      * off-diagonal moments (mat_els.f90:184-228): in full, every B-spline
        value evaluated inside the loops, OpenMP over the multipoles k;
      * diagonal cells (mat_els.f90:230-292): one multipole (one thread's
        unit of the reference's !$omp parallel do over k) on the leading
        outer columns j_b_p, B-spline values evaluated inside the loops;
        extrapolated by entry count to all sparse_6d entries, then to the K
        multipoles over the threads (ceil(K / threads) rounds);
      * R^k assembly (sparse_array_tools.f90:452-493): serial set_val calls
        into a restated Nd_DOK hash map, for the products of a leading
        fraction of the sparse_4d entries (extrapolated by set_val count) and
        for every diagonal cell (in full)."""
    import numpy as np

    from oracle import oracle as O
    K = w.max_k + 1
    n_b = O.dims(w.k, t)[1]
    npairs = O.n_pairs(w.k, n_b)
    n_int = npairs * npairs * K
    # off-diagonal moments, in full
    t0 = time.perf_counter()
    off = O.slater_off_diag(w.k, t, w.k_GL, w.max_k, faithful=True, nthreads=nthreads)
    t_off = time.perf_counter() - t0
    # diagonal cells: a calibration run, then a sample of about diag_target_s
    nnz6 = O.count_nnz6(w.k, t)
    k_sel = w.max_k // 2
    jp = max(w.k, n_b // 200)
    t0 = time.perf_counter()
    done = O.slater_diag_sample(w.k, t, w.k_GL, k_sel, jp, True)
    dt = time.perf_counter() - t0
    per_entry = dt / max(done, 1)
    jp = min(n_b, max(jp, int(jp * diag_target_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    done = O.slater_diag_sample(w.k, t, w.k_GL, k_sel, jp, True)
    t_diag_sample = time.perf_counter() - t0
    per_entry = t_diag_sample / max(done, 1)
    t_diag = per_entry * nnz6 * math.ceil(K / nthreads)
    # assembly: the diagonal cells' set_val calls need only the index arrays
    dg_idx = O.slater_diag(w.k, t, w.k_GL, w.max_k, nthreads=nthreads, pair_mask=np.zeros(npairs, bool))
    empty = dict(nnz=0, rdk=np.zeros((0, K)), iv=np.zeros(0, np.int64), i=np.zeros(0, np.int64),
                 ip=np.zeros(0, np.int64), j=np.zeros(0, np.int64), jp=np.zeros(0, np.int64))
    iv = off["iv"]
    total_products = off["nnz"] ** 2 - sum(int(np.count_nonzero(iv == v)) ** 2 for v in np.unique(iv))
    n_lead = max(8, off["nnz"] // 1000)
    t0 = time.perf_counter()
    calls = O.rk_dok(w.k, n_b, w.max_k, off, empty, n_lead=n_lead)
    dt = time.perf_counter() - t0
    n_lead = min(off["nnz"], max(n_lead, int(n_lead * dok_target_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    calls = O.rk_dok(w.k, n_b, w.max_k, off, empty, n_lead=n_lead)
    t_prod_sample = time.perf_counter() - t0
    t_prod = t_prod_sample * total_products / max(calls, 1)
    t0 = time.perf_counter()
    O.rk_dok(w.k, n_b, w.max_k, off, dg_idx, n_lead=0)
    t_dcells = time.perf_counter() - t0
    total = t_off + t_diag + t_prod + t_dcells
    sampled = t_off + t_diag_sample + t_prod_sample + t_dcells
    return {"value": n_int / total, "unit": "integrals/s", "cores": nthreads, "host_cpus": os.cpu_count(),
            "kind": "port", "extrapolated": True, "seconds_estimated": total, "seconds_sampled": sampled,
            "sample": (f"{w.name} build on the reference's schedule by the oracle (B-spline values evaluated "
                       f"inside the loops, OpenMP over the {K} multipoles on {nthreads} threads, serial hash-map "
                       f"assembly): off-diagonal moments in full ({t_off:.2f} s); diagonal cells of multipole "
                       f"{k_sel} for outer columns <= {jp} ({done} of {nnz6} entries, {t_diag_sample:.2f} s, one "
                       f"thread) extrapolated to {t_diag:.1f} s for all entries and multipoles; the products of "
                       f"the first {n_lead} of {off['nnz']} sparse_4d entries ({calls} of {total_products} set_val "
                       f"calls, {t_prod_sample:.2f} s) extrapolated to {t_prod:.1f} s; the {2 * dg_idx['nnz']} "
                       f"diagonal-cell set_val calls in full ({t_dcells:.2f} s); estimated whole build "
                       f"{total:.1f} s"),
            "stages_s": {"off_diagonal_moments": t_off, "diagonal_cells": t_diag, "assembly_products": t_prod,
                         "assembly_diagonal_cells": t_dcells}}


def cpu_block_baseline(w, t, R, gpu_block, block_max_l=2):
    """The oracle's count and fill of the L = 0 block (l1, l2 <= block_max_l)
    of the basis, on one thread (the reference builds each symmetry on one
    thread), and the device's block compared with it."""
    import numpy as np

    from oracle import oracle as O
    n_b = O.dims(w.k, t)[1]
    S = O.setup_S(w.k, t, w.k_GL)
    Hl = [O.setup_H1p(w.k, t, w.k_GL, w.Z, l, cap_order=2, cap_r0=0.8 * w.r_max, cap_eta=0.0)
          for l in range(block_max_l + 1)]
    cf = O.count_configs(0, 0, block_max_l, n_b, w.k, n_b, n_b, block_max_l)
    tb = time.perf_counter()
    blk = O.construct_block(w.k, n_b, 0, w.max_k, True, cf, R, Hl, S)
    db = time.perf_counter() - tb
    out = dict(value=blk["nnz"][0] / db, unit="H entries/s", cores=1, kind="port",
               sample=(f"{w.name} basis, L=0 block, l1,l2 <= {block_max_l} ({len(cf)} configurations, "
                       f"{blk['nnz'][0]} H entries): the oracle's count and fill passes; {db:.2f} s"),
               seconds=db)
    if gpu_block is not None:   # the whole block against the oracle's (oracle tensor and one-electron matrices)
        H, Sb = gpu_block
        pattern = bool(np.array_equal(H.index_ptr, blk["H"][0]) and np.array_equal(H.indices, blk["H"][1])
                       and np.array_equal(Sb.index_ptr, blk["S"][0]) and np.array_equal(Sb.indices, blk["S"][1]))
        out["parity_vs_oracle"] = dict(
            pattern_bit_exact=pattern,
            H_max_abs_err_over_max_abs=float(np.abs(H.data - blk["H"][2]).max() / np.abs(blk["H"][2]).max())
            if pattern else None,
            S_max_abs_err_over_max_abs=float(np.abs(Sb.data - blk["S"][2]).max() / np.abs(blk["S"][2]).max())
            if pattern else None)
    return out


def measure_build(rk, stream, steps, warmup, latency_samples):
    """Times `steps` builds (after `warmup`): the rate from the host clock
    around the synchronised loop, the per-stage device timeline from the
    library's HIP events (a second loop), the replay of a captured hipGraph of
    the build (a third loop) and the latency of single builds."""
    import numpy as np
    import torch
    n_int = rk.n_pairs * rk.n_pairs * rk.K
    for _ in range(warmup):
        rk.compute()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        rk.compute()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    rk.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        rk.compute()
    torch.cuda.synchronize()
    wall_prof = time.perf_counter() - t0
    st = rk.stage_times()        # ms since start: windows, far end, diag end, near end, total
    rk.set_profiling(False)
    rk.set_graph(True)
    rk.compute()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        rk.compute()
    torch.cuda.synchronize()
    wall_graph = time.perf_counter() - t0
    rk.set_graph(False)
    lat = []
    for _ in range(latency_samples):
        torch.cuda.synchronize()
        a = time.perf_counter()
        rk.compute()
        rk.synchronize()
        lat.append((time.perf_counter() - a) * 1e3)
    lat = np.array(lat) if lat else np.zeros(1)
    return {"value": n_int * steps / wall, "ms_per_step": wall * 1e3 / steps,
            "event_ms_per_step": ev0.elapsed_time(ev1) / steps, "profiled_ms_per_step": wall_prof * 1e3 / steps,
            "graph_ms_per_step": wall_graph * 1e3 / steps,
            "timeline_ms": {"moments_and_windows_done": float(st[0]), "rank_one_fill_done": float(st[1]),
                            "diag_cells_done_side_stream": float(st[2]),
                            "near_ranges_done_side_stream": float(st[3]), "build_done": float(st[4])},
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                           "samples": int(latency_samples)},
            "stages": st}


def store_ceiling(ctx, stream, n_elems):
    """(library, torch) GB/s of a store-only fill of a buffer of n_elems
    doubles on the build's stream: the library's plain store kernel (one
    16-byte store per thread, the faster of with and without the
    non-temporal hint; bs2e_store_ceiling) and torch's own fill (zero_)."""
    import torch
    with torch.cuda.stream(stream):
        buf = torch.empty(n_elems + 2, dtype=torch.float64, device="cuda")
        off = (-buf.data_ptr() % 16) // 8          # a 16-byte aligned start
        lib = ctx.store_ceiling(buf.data_ptr() + 8 * off, 8 * (n_elems // 2) * 2, 5)
        buf.zero_()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(5):
            buf.zero_()
        e1.record(stream)
    torch.cuda.synchronize()
    tfill = 8.0 * (n_elems + 2) * 5 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del buf
    torch.cuda.empty_cache()
    return lib, tfill


def branch_terms(n_b, k, lmax, max_L=2, l2=1):
    """The symmetries of one basis_setup request of the serving stream: total
    L = 0..max_L with natural parity (z_pol), one-electron l <= lmax, the
    second electron's n2 <= n_b // 4 (an r_2_max truncation) and its l <= l2
    for n1 > n_b // 2 (r_all_l; orbital_tools.f90:144-193)."""
    import bs2e
    terms = []
    for L in range(max_L + 1):
        term = bs2e.sym(l=L, pi=bool(L % 2))
        term.configs = bs2e.count_configs(term, lmax, n_b, k, n_b // 4, n_b // 2, min(l2, lmax))
        if term.n_config:
            terms.append(term)
    return terms


def run_stream(srv, jobs, arrivals, submit):
    """Submits job i at arrivals[i] (s after the start) through submit(srv,
    job), hands each finished ticket to its job's `done` hook, retires it;
    returns (latency ms per job, wall s)."""
    n = len(jobs)
    lat = [None] * n
    tickets = {}
    nxt = 0
    t0 = time.perf_counter()
    while nxt < n or tickets:
        now = time.perf_counter() - t0
        while nxt < n and arrivals[nxt] <= now:
            tickets[submit(srv, jobs[nxt])] = nxt
            nxt += 1
        for tk, ms in srv.finished():
            i = tickets.pop(tk)
            if ms < 0:
                raise RuntimeError(f"serving job {i} failed: {srv.last_error()}")
            lat[i] = ms
            srv.retire(tk)
    return lat, time.perf_counter() - t0


def serving_benchmark(n_jobs, seed=20261202, slots=4, budget_gib=160, load=0.8, precision="mixed", fixed_rate=60.0,
                      pause_s=0.0):
    """Continuous batching of a seeded stream of basis_setup requests through
    the build service: each job is the two-electron branch (R^k tensor in
    `precision`, one-electron matrices, the H/S blocks of L = 0..2); closed-
    loop capacity, then Poisson arrivals at `load` x capacity with p50/p99
    latency from arrival to the job's H/S CSR written, and the same arrivals
    rescaled to a fixed rate of `fixed_rate` jobs/s (comparable between
    builds of different capacity).  pause_s: an idle pause before each pass,
    so that a kernel trace shows the passes apart."""
    import numpy as np

    import bs2e
    from bs2e.workloads import serving_jobs
    jobs = []
    integrals, rows = [], []
    for j in serving_jobs(n_jobs, 7):
        n_b = len(j.t) - j.k - 2
        terms = branch_terms(n_b, j.k, j.lmax)
        jobs.append((j, terms))
        _, _, n_pairs = bs2e.count_nnz(j.k, j.t)
        integrals.append(n_pairs * n_pairs * (j.max_k + 1))
        rows.append(sum(tm.n_config for tm in terms))

    def submit(srv, job):
        j, terms = job
        return srv.submit_job(j.k, j.t, j.k_GL, j.max_k, precision=precision, terms=terms, full=False, Z=2,
                              max_l_1p=j.lmax, CAP_c=(2, 0.8 * j.t[-1], 1e-3))

    srv = bs2e.Server(budget_gib << 30, slots=slots)
    run_stream(srv, jobs[:4], [0.0] * 4, submit)                    # warm-up
    time.sleep(pause_s)
    _, wall_cap = run_stream(srv, jobs, [0.0] * n_jobs, submit)      # all at once: capacity
    cap = n_jobs / wall_cap
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / (load * cap), n_jobs))
    time.sleep(pause_s)
    lat, wall = run_stream(srv, jobs, arrivals, submit)
    time.sleep(pause_s)
    lat_fixed, _ = run_stream(srv, jobs, arrivals * (load * cap / fixed_rate), submit)
    cache = srv.cache_info()
    srv.close()
    lat = np.array(lat)
    lat_fixed = np.array(lat_fixed)
    return {"unit_of_work": "the two-electron branch of one basis_setup request: R^k tensor, S and H(l), H/S CSR "
                            "of L = 0..2 (l <= lmax, n2 <= n_b/4, l2 <= 1 beyond n_b/2, upper triangles)",
            "precision": precision, "jobs": n_jobs, "slots": slots, "budget_GiB": budget_gib, "job_seed": 7,
            "arrival_seed": seed, "integrals_total": int(sum(integrals)), "block_rows_total": int(sum(rows)),
            "capacity_jobs_per_s": cap, "capacity_integrals_per_s": sum(integrals) / wall_cap,
            "offered_load": load, "stream_integrals_per_s": sum(integrals) / wall,
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                           "mean": float(lat.mean()), "max": float(lat.max()),
                           "from": "submission", "to": "H/S CSR of every symmetry written (device event)"},
            "latency_ms_fixed_rate": {"jobs_per_s": fixed_rate, "p50": float(np.percentile(lat_fixed, 50)),
                                      "p99": float(np.percentile(lat_fixed, 99)), "mean": float(lat_fixed.mean())},
            "one_electron_cache": {"entries": cache[0], "bytes": cache[1], "hits": cache[2], "misses": cache[3]}}


def blocks_benchmark(rk, w, t, max_l=2, L=0, keep=None):
    """Channel summation into the CSR H/S block of one symmetry (L, even
    parity, l1, l2 <= max_l) from the built tensor; kernel times from the
    library's HIP events, wall time including the host transfers."""
    import bs2e
    spl = bs2e.b_spline(w.k, t)
    S = bs2e.setup_S(spl, w.k_GL, ctx=rk.ctx)
    Hl = bs2e.setup_H_one_particle(w.Z, (2, 0.8 * w.r_max, 0.0), range(max_l + 1), spl, w.k_GL, ctx=rk.ctx)
    term = bs2e.sym(l=L, pi=bool(L % 2))
    term.configs = bs2e.count_configs(term, max_l, spl.n_b, w.k, spl.n_b, spl.n_b, max_l)
    bs2e.Block(rk, term).release()                 # warm-up
    t0 = time.perf_counter()
    blk = bs2e.Block(rk, term)
    H_sp, S_sp = blk.fill(list(Hl), S)
    wall = time.perf_counter() - t0
    ms_count, ms_fill = blk.kernel_ms()
    blk.release()
    if keep is not None:
        keep.extend([H_sp, S_sp])
    written = 24.0 * (H_sp.nnz + S_sp.nnz)          # 1-based column index + complex value per entry
    return {"workload": f"{w.name} basis, L={L} block, l1,l2 <= {max_l}", "n_config": term.n_config,
            "nnz_H": H_sp.nnz, "nnz_S": S_sp.nnz, "count_kernel_ms": ms_count, "fill_kernel_ms": ms_fill,
            "H_entries_per_s": H_sp.nnz / (ms_fill * 1e-3), "fill_GBs": written / (ms_fill * 1e-3) / 1e9,
            "wall_ms_with_transfers": wall * 1e3}


def branch_benchmark(ctx, name="c2", max_l=2, max_L=2, steps=50, warmup=5):
    """The two-electron branch of one basis, repeated: the tensor build and the
    H/S blocks of L = 0..max_L (natural parity, l1, l2 <= max_l, upper
    triangles), eager (the build, then the fill from the device copy of the
    one-electron matrices, waited for) against one launch of the captured
    graph of the whole branch (waited for); ms per branch."""
    import torch

    import bs2e
    from bs2e.workloads import WORKLOADS, knots
    w = WORKLOADS[name]
    t = knots(w)
    spl = bs2e.b_spline(w.k, t)
    rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx)
    S = bs2e.setup_S(spl, w.k_GL, ctx=ctx)
    Hl = bs2e.setup_H_one_particle(w.Z, (2, 0.8 * w.r_max, 1e-3), range(max_l + 1), spl, w.k_GL, ctx=ctx)
    terms = []
    for L in range(max_L + 1):
        term = bs2e.sym(l=L, pi=bool(L % 2))
        term.configs = bs2e.count_configs(term, max_l, spl.n_b, w.k, spl.n_b, spl.n_b, max_l)
        terms.append(term)
    bset = bs2e.BlockSet(rk, terms, False)
    bset.fill(list(Hl), S)
    for _ in range(warmup):
        rk.compute()
        bset.refill()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        rk.compute()
        bset.refill()
    eager = (time.perf_counter() - t0) * 1e3 / steps
    bset.set_graph(True)
    for _ in range(warmup):
        bset.replay()
        bset.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        bset.replay()
        bset.synchronize()
    graph = (time.perf_counter() - t0) * 1e3 / steps
    nH = int(bset.sizes[:, 0].sum())
    bset.release()
    rk.release()
    return {"workload": f"{w.name} basis: tensor (k <= {w.max_k}) + H/S blocks of L = 0..{max_L}, l1,l2 <= {max_l}, "
                        f"upper triangles ({sum(t_.n_config for t_ in terms)} rows, {nH} H entries)",
            "eager_ms": eager, "graph_ms": graph, "steps": steps,
            "eager": "build enqueued, then layouts + fill launched and waited for",
            "graph": "one hipGraph launch: build, count, device row offsets, layouts, fill; waited for"}


def blockset_benchmark(ctx, name="c3", max_l=4, max_L=4, refills=3):
    """Channel summation of every symmetry L = 0..max_L (natural parity) of
    the named basis with l1, l2 <= max_l in one count and one fill launch
    (uneven channel counts per symmetry), from the basis's own tensor (k up to
    its max_k); kernel times from the library's HIP events.  The CSR arrays
    stay on the device (nothing is copied back)."""
    import bs2e
    from bs2e.workloads import WORKLOADS
    w = WORKLOADS[name]
    t = bs2e.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    spl = bs2e.b_spline(w.k, t)
    rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx)
    S = bs2e.setup_S(spl, w.k_GL, ctx=ctx)
    Hl = bs2e.setup_H_one_particle(w.Z, (2, 0.8 * w.r_max, 0.0), range(max_l + 1), spl, w.k_GL, ctx=ctx)
    terms = []
    for L in range(max_L + 1):
        term = bs2e.sym(l=L, pi=bool(L % 2))
        term.configs = bs2e.count_configs(term, max_l, spl.n_b, w.k, spl.n_b, spl.n_b, max_l)
        terms.append(term)
    out = {"workload": f"{w.name} basis (order {w.k}, {spl.n_b} B-splines, k <= {w.max_k}), L = 0..{max_L}, "
                       f"l1,l2 <= {max_l}, all channels, one count + one fill launch",
           "rows": int(sum(t_.n_config for t_ in terms))}
    try:
        t0 = time.perf_counter()
        bset = bs2e.BlockSet(rk, terms)
        bset.fill(list(Hl), S)
        wall = time.perf_counter() - t0
    except bs2e.Bs2eError as e:
        rk.release()
        out["error"] = str(e)
        return out
    ms_count, ms_first = bset.kernel_ms()
    warm = []
    for _ in range(refills):   # the fill again into the same outputs: the steady state of repeated fills
        bset.refill()
        warm.append(bset.kernel_ms()[1])
    ms_fill = sorted(warm)[len(warm) // 2] if warm else ms_first
    nH, nS = int(bset.sizes[:, 0].sum()), int(bset.sizes[:, 1].sum())
    bset.release()
    rk.release()
    out.update({"channels_per_symmetry": [int(len({(int(a), int(b)) for a, b in t_.configs[:, 2:4]})) for t_ in terms],
                "nnz_H": nH, "nnz_S": nS, "count_kernel_ms": ms_count, "fill_kernel_ms": ms_fill,
                "fill_kernel_ms_runs": warm, "first_fill_kernel_ms": ms_first,
                "fill_kernel_ms_is": f"median of {len(warm)} refills into the same outputs (HIP events around the "
                                     "layout and fill launches); first_fill_kernel_ms is the first fill into fresh "
                                     "outputs",
                "H_entries_per_s": nH / (ms_fill * 1e-3), "fill_GBs": 24.0 * (nH + nS) / (ms_fill * 1e-3) / 1e9,
                "wall_ms_create_and_fill": wall * 1e3})
    return out


def blockset_stream_benchmark(ctx, name="c3", max_l=8, max_L=4, window_gib=64):
    """The named basis's channel batch at its stated size: every symmetry
    L = 0..max_L (natural parity) with l1, l2 <= max_l, all channels, counted
    in one launch and filled window by window (bs2e_blockset_fill_rows) into
    outputs of at most window_gib GiB, since the whole set exceeds the
    device's memory; each window's fill is timed by the library's HIP events
    (layouts and the two fill launches) and the window's entries stay on the
    device until the next window overwrites them (a consumer would write them
    out in between)."""
    import bs2e
    from bs2e.workloads import WORKLOADS
    w = WORKLOADS[name]
    t = bs2e.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    spl = bs2e.b_spline(w.k, t)
    rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx)
    S = bs2e.setup_S(spl, w.k_GL, ctx=ctx)
    Hl = bs2e.setup_H_one_particle(w.Z, (2, 0.8 * w.r_max, 0.0), range(max_l + 1), spl, w.k_GL, ctx=ctx)
    terms = []
    for L in range(max_L + 1):
        term = bs2e.sym(l=L, pi=bool(L % 2))
        term.configs = bs2e.count_configs(term, max_l, spl.n_b, w.k, spl.n_b, spl.n_b, max_l)
        terms.append(term)
    out = {"workload": f"{w.name} basis (order {w.k}, {spl.n_b} B-splines, k <= {w.max_k}), L = 0..{max_L}, "
                       f"l1,l2 <= {max_l} (the stated channel batch), all channels, streamed in row windows of "
                       f"<= {window_gib} GiB",
           "rows": int(sum(t_.n_config for t_ in terms))}
    try:
        t0 = time.perf_counter()
        bset = bs2e.BlockSet(rk, terms)
        t_count = time.perf_counter() - t0
        ms = []
        t0 = time.perf_counter()
        for _ in bset.stream(window_gib << 30, list(Hl), S):   # outputs reserved once for the largest window
            ms.append(bset.kernel_ms()[1])
        wall = time.perf_counter() - t0
    except bs2e.Bs2eError as e:
        rk.release()
        out["error"] = str(e)
        return out
    nH, nS = int(bset.sizes[:, 0].sum()), int(bset.sizes[:, 1].sum())
    count_ms = bset.kernel_ms()[0]
    bset.release()
    rk.release()
    fill = float(sum(ms))
    out.update({"channels_per_symmetry": [int(len({(int(a), int(b)) for a, b in t_.configs[:, 2:4]})) for t_ in terms],
                "nnz_H": nH, "nnz_S": nS, "csr_GB": 24.0 * (nH + nS) / 1e9, "windows": len(ms),
                "count_kernel_ms": count_ms, "fill_kernel_ms": fill, "window_fill_ms": ms,
                "H_entries_per_s": nH / (fill * 1e-3), "fill_GBs": 24.0 * (nH + nS) / (fill * 1e-3) / 1e9,
                "wall_ms_create": t_count * 1e3, "wall_ms_windows": wall * 1e3,
                "wall_ms_windows_is": "host clock over the reservation of the outputs and every window's fill, "
                                      "each waited for"})
    return out


def dipoles_benchmark(ctx, w, t, max_l=2, L1=0, gauge="l"):
    """Dipole block between symmetries (L1, even) and (L1+1, odd), q = 0
    (dipole.f90:8-47): radial dipole matrices, pattern count and fill on the
    device; kernel times from the library's HIP events, wall time including
    the host transfers."""
    import bs2e
    spl = bs2e.b_spline(w.k, t)
    S = bs2e.setup_S(spl, w.k_GL, ctx=ctx)
    rad = bs2e.setup_radial_dip(spl, w.k_GL, gauge, ctx=ctx)
    syms = []
    for L in (L1, L1 + 1):
        s = bs2e.sym(l=L, pi=bool(L % 2))
        s.configs = bs2e.count_configs(s, max_l, spl.n_b, w.k, spl.n_b, spl.n_b, max_l)
        syms.append(s)
    bs2e.DipBlock(syms, 0, spl, ctx=ctx).release()   # warm-up
    t0 = time.perf_counter()
    blk = bs2e.DipBlock(syms, 0, spl, ctx=ctx)
    D = blk.fill(S, rad)
    wall = time.perf_counter() - t0
    ms_count, ms_fill = blk.kernel_ms()
    blk.release()
    written = 24.0 * D.nnz                            # 1-based column index + complex value per entry
    return {"workload": f"{w.name} basis, L={L1} -> L={L1 + 1} dipole block, q=0, gauge {gauge}, l1,l2 <= {max_l}",
            "shape": list(D.shape), "nnz": D.nnz, "count_kernel_ms": ms_count, "fill_kernel_ms": ms_fill,
            "entries_per_s": D.nnz / (ms_fill * 1e-3), "fill_GBs": written / (ms_fill * 1e-3) / 1e9,
            "wall_ms_with_transfers": wall * 1e3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c5p")
    ap.add_argument("--secondary", default="c2",
                    help="a second, smaller workload timed the same way (its blocks and dipoles too); none skips it")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sample-rows", type=int, default=48,
                    help="tensor rows of the headline workload compared with the oracle's rows")
    ap.add_argument("--latency-samples", type=int, default=50,
                    help="synchronised single builds timed for p50/p99 latency")
    ap.add_argument("--blocks", type=int, default=1, help="also time the L=0 symmetry block (0 skips it)")
    ap.add_argument("--blockset", type=int, default=1,
                    help="also time all symmetries L=0..4 (l <= 4) of the c3 basis in one launch (0 skips it)")
    ap.add_argument("--blockset-stream", type=int, default=1,
                    help="also stream the stated c3 channel batch (l <= 8, L=0..4) through row windows (0 skips it)")
    ap.add_argument("--dipoles", type=int, default=1, help="also time the L=0 -> L=1 dipole block (0 skips it)")
    ap.add_argument("--branch", type=int, default=1,
                    help="also time the c2 branch (tensor + L=0..2 blocks) eager against its captured graph (0 skips)")
    ap.add_argument("--serve-jobs", type=int, default=64,
                    help="jobs of the serving stream (differing bases through the build service); 0 skips it")
    ap.add_argument("--serve-pause-ms", type=float, default=0.0,
                    help="idle pause before each serving pass (kernel traces show the passes apart)")
    ap.add_argument("--far-shape", type=int, default=-1,
                    help="pin the rank-one fill shape (bs2e_rk_set_tuning) instead of autotuning (-1: autotune)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the run, copy a buffer of this size 3 times on the build stream "
                         "(reference kernel for HBM counter scaling under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU: the R^k build of one basis runs on one device "
                                   "(multi-GPU sharding is a design note, not a run)", "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    import bs2e
    from bs2e.workloads import WORKLOADS, knots

    w = WORKLOADS[args.workload]
    nthreads = host_threads()
    torch.cuda.set_device(0)
    stream = torch.cuda.Stream(device=0)     # default priority; the build and its events share it
    ctx = bs2e.Context(0)
    ctx.set_stream(stream.cuda_stream)
    t = knots(w)
    ref_rows = None
    runs = {}
    for precision in ("fp64", "mixed"):
        rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx, compute=False, precision=precision)
        n_int = rk.n_pairs * rk.n_pairs * rk.K
        log(f"bench {w.name} {precision}: n_b={rk.n_b} n_pairs={rk.n_pairs} K={rk.K} integrals/step={n_int}")
        if args.far_shape >= 0:
            rk.autotune()
            rk.set_tuning(args.far_shape, rk.tuning()[1])   # the pinned shape, the tuned multipole chunk
        else:
            rk.autotune()                 # fill shape and multipole chunk timed on this basis
        tuning = rk.tuning()
        m = measure_build(rk, stream, args.steps, args.warmup, args.latency_samples)
        log(f"{w.name} {precision}: {m['ms_per_step']:.4f} ms/build, {m['value']:.3e} integrals/s")
        st = m.pop("stages")
        t_far = float(st[1] - st[0]) * 1e-3
        far_elems = n_int - rk.near_elems
        traffic_alg = 8.0 * far_elems            # algorithmic bytes of one k3_far launch
        achieved = traffic_alg / t_far / 1e9
        m["tuning"] = {"far_fill_shape": tuning[0], "diag_multipole_chunk": tuning[1]}
        m["tensor_stage_GBs"] = 8.0 * n_int / (float(st[4] - st[0]) * 1e-3) / 1e9
        m["far"] = dict(achieved=achieved, t_far=t_far, far_elems=far_elems, traffic_alg=traffic_alg,
                        traffic=FAR_TRAFFIC_BYTES.get((w.name, int(tuning[0]))))
        if not args.no_cpu_baseline:
            if ref_rows is None:
                rows = sample_rows(rk.n_pairs, args.sample_rows)
                log(f"{w.name}: {len(rows)} sampled rows from the oracle")
                ref_rows = oracle_rows(w, t, rows, nthreads)
            m["error_vs_oracle"] = dict(error_stats(rk.rows(rows), ref_rows), rows=len(rows),
                                        sample=f"{len(rows)} tensor rows (first, last, {len(rows) - 2} drawn, seed 3)")
        rk.release()
        runs[precision] = m
    ceiling, torch_fill = store_ceiling(ctx, stream, n_int)
    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * (1 << 20) // 8
        with torch.cuda.stream(stream):
            src = torch.ones(n, dtype=torch.float64, device="cuda")
            dst = torch.empty_like(src)
            for _ in range(3):
                dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    mx, fp = runs["mixed"], runs["fp64"]
    mixed_ok = "error_vs_oracle" not in mx or max(mx["error_vs_oracle"]["per_k_normwise"]) <= 1e-10
    head = mx if mixed_ok else fp
    far = head.pop("far")
    fp_far = fp.pop("far") if head is mx else far
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "integrals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-mfma+f64" if head is mx else "f64",
        "data": "synthetic",
        "config": {"workload": f"{w.name}: {w.description}", "n_b": rk.n_b, "n_pairs": rk.n_pairs, "K": rk.K,
                   "k_GL": w.k_GL, "integrals_per_step": n_int,
                   "precision": ("mixed: off-diagonal moments on FP32 MFMA from exact 10-bit slices, FP64 de Boor, "
                                 "diagonal cells, window sums and tensor products") if head is mx else "plain FP64",
                   "parallelism": "single GPU"},
        "timeline_ms": head["timeline_ms"],
        "tensor_stage_GBs": head["tensor_stage_GBs"],
        "event_ms_per_step": head["event_ms_per_step"],
        "profiled_ms_per_step": head["profiled_ms_per_step"],
        "graph_ms_per_step": head["graph_ms_per_step"],
        "tuning": head["tuning"],
        "latency_ms": head["latency_ms"],
        "roofline": {"bound": "hbm", "kernel": "k3_far (rank-one tensor fill)", "achieved": far["achieved"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": far["achieved"] / HBM_PEAK_GBS,
                     "traffic": far["traffic"], "algorithmic_bytes": far["traffic_alg"], "elements": far["far_elems"],
                     "avg_launch_ms": far["t_far"] * 1e3, "store_ceiling_GBs": ceiling,
                     "frac_of_store_ceiling": far["achieved"] / ceiling,
                     "store_ceiling_is": "the library's plain store kernel (one 16-byte store per thread, the faster of "
                                         "with and without the non-temporal hint) over a tensor-sized buffer",
                     "torch_fill_GBs": torch_fill},
        "mixed_over_fp64": mx["value"] / fp["value"],
        "fp64": dict(fp, roofline_frac=fp_far["achieved"] / HBM_PEAK_GBS),
    }
    if "error_vs_oracle" in head:
        out["error_vs_oracle"] = head["error_vs_oracle"]
    if head is fp:
        out["mixed"] = dict(mx, note="per-k normwise error above 1e-10: not the headline")
    del ref_rows

    gpu_block = None
    if args.secondary != "none":
        w2 = WORKLOADS[args.secondary]
        t2 = knots(w2)
        sec = {"workload": f"{w2.name}: {w2.description}"}
        tensors = {}
        for precision in ("fp64", "mixed"):
            rk2 = bs2e.RkTensor.build(w2.k, t2, w2.k_GL, w2.max_k, ctx=ctx, compute=False, precision=precision)
            rk2.autotune()
            m2 = measure_build(rk2, stream, args.steps, args.warmup, 0)
            m2.pop("stages")
            m2.pop("latency_ms")
            sec["integrals_per_step"] = rk2.n_pairs * rk2.n_pairs * rk2.K
            sec[precision] = m2
            log(f"{w2.name} {precision}: {m2['ms_per_step']:.4f} ms/build")
            if not args.no_cpu_baseline:
                tensors[precision] = rk2.to_host()
            if precision == "fp64":
                if args.blocks:
                    gpu_block = [] if not args.no_cpu_baseline else None
                    out["blocks"] = blocks_benchmark(rk2, w2, t2, keep=gpu_block)
                if args.dipoles:
                    out["dipoles"] = dipoles_benchmark(ctx, w2, t2)
            rk2.release()
        sec["mixed_over_fp64"] = sec["mixed"]["value"] / sec["fp64"]["value"]
        out["secondary"] = sec
        if tensors:
            from oracle import oracle as O
            n_b2 = O.dims(w2.k, t2)[1]
            Rref = O.rk_dense(w2.k, n_b2, w2.max_k, O.slater_off_diag(w2.k, t2, w2.k_GL, w2.max_k, nthreads=nthreads),
                              O.slater_diag(w2.k, t2, w2.k_GL, w2.max_k, nthreads=nthreads))
            for precision, R2 in tensors.items():
                sec[precision]["error_vs_oracle"] = dict(error_stats(R2, Rref), rows="all")
            if args.blocks:
                log(f"{w2.name}: L=0 block on the oracle")
                out["blocks"]["cpu_baseline"] = cpu_block_baseline(w2, t2, Rref, tuple(gpu_block))
                out["blocks"]["speedup_vs_cpu"] = out["blocks"]["H_entries_per_s"] / out["blocks"]["cpu_baseline"]["value"]
            del Rref, tensors
    if args.branch:
        log("branch graph")
        out["branch"] = branch_benchmark(ctx)
    if args.blockset:
        log("c3 block set")
        out["blockset"] = blockset_benchmark(ctx)
    if args.blockset_stream:
        log("c3 stated channel batch in row windows")
        out["blockset_stream"] = blockset_stream_benchmark(ctx)
    if args.serve_jobs > 0:
        log("serving stream")
        out["serving"] = serving_benchmark(args.serve_jobs, pause_s=args.serve_pause_ms * 1e-3)
    if not args.no_cpu_baseline:
        log(f"{w.name}: host baseline on the reference's schedule ({nthreads} threads)")
        cb = cpu_baseline_reference_schedule(w, t, nthreads)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = out["value"] / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
