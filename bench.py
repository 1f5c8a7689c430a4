#!/usr/bin/env python3
"""bench.py -- R^k integrals/s of the device R^k build on one MI355X.

One step = one complete build of the banded R^k tensor of a basis from its
knot vector: de Boor tables, pair moments, diagonal cells and the tensor
(the reference's setup_Slater_integrals + compute_R_K_map), with the knots
and every work buffer already resident in HBM.  Integrals per step =
n_pairs^2 * (max_k+1), the key count of the reference's Nd_DOK times its
value length.

    python bench.py [--steps K] [--warmup W] [--workload c2]

Prints one JSON line.  The dominant kernel (the tensor fill) is timed with
HIP events on the stream it runs on and priced against HBM (8 bytes stored
per integral, 8 TB/s peak).  cpu_baseline times the oracle (the repository's
C restatement of the reference path) on the host for the same workload.
"""
from __future__ import annotations

import argparse
import json
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
# each scaled by what the counter reads for a 1 GiB device copy
# (profiles/rk_build_c2.md), per (workload, fill shape); None for pairs not
# measured that way.
FAR_TRAFFIC_BYTES = {("c2", 0): 739.35e6, ("c2", 1): 739.59e6, ("c2", 2): 739.99e6, ("c2", 3): 739.74e6,
                     ("c2", 4): 739.52e6, ("c2", 5): 739.51e6, ("c2", 6): 739.05e6}


def cpu_baseline(w, t, R_gpu, nthreads, block_max_l=None, gpu_block=None):
    """Oracle (C restatement, OpenMP over k like the reference) on the host:
    primitives + dense tensor assembly for the same basis; with block_max_l,
    also the L = 0 block of bench's `blocks` line (count and fill, one
    thread: the reference builds each symmetry on one thread), compared with
    the device's block (gpu_block = (H, S) CSR) when given."""
    import numpy as np

    from oracle import oracle as O
    K = w.max_k + 1
    n_b = O.dims(w.k, t)[1]
    t0 = time.perf_counter()
    off = O.slater_off_diag(w.k, t, w.k_GL, w.max_k, nthreads=nthreads)
    dg = O.slater_diag(w.k, t, w.k_GL, w.max_k, nthreads=nthreads)
    R = O.rk_dense(w.k, n_b, w.max_k, off, dg)
    dt = time.perf_counter() - t0
    n_int = R.size
    rel = None
    if R_gpu is not None:   # largest relative error per multipole k (every R^k value is positive)
        rel = [float(x) for x in (np.abs(R_gpu - R) / R).reshape(-1, K).max(axis=0)]
    blocks = None
    if block_max_l is not None:
        S = O.setup_S(w.k, t, w.k_GL)
        Hl = [O.setup_H1p(w.k, t, w.k_GL, w.Z, l, cap_order=2, cap_r0=0.8 * w.r_max, cap_eta=0.0)
              for l in range(block_max_l + 1)]
        cf = O.count_configs(0, 0, block_max_l, n_b, w.k, n_b, n_b, block_max_l)
        tb = time.perf_counter()
        blk = O.construct_block(w.k, n_b, 0, w.max_k, True, cf, R, Hl, S)
        db = time.perf_counter() - tb
        blocks = dict(value=blk["nnz"][0] / db, unit="H entries/s", cores=1, kind="port",
                      sample=(f"{w.name} basis, L=0 block, l1,l2 <= {block_max_l} ({len(cf)} configurations, "
                              f"{blk['nnz'][0]} H entries): the oracle's count and fill passes; {db:.2f} s"),
                      seconds=db)
        if gpu_block is not None:   # the whole block against the oracle's (oracle tensor and one-electron matrices)
            H, Sb = gpu_block
            pattern = bool(np.array_equal(H.index_ptr, blk["H"][0]) and np.array_equal(H.indices, blk["H"][1])
                           and np.array_equal(Sb.index_ptr, blk["S"][0]) and np.array_equal(Sb.indices, blk["S"][1]))
            blocks["parity_vs_oracle"] = dict(
                pattern_bit_exact=pattern,
                H_max_abs_err_over_max_abs=float(np.abs(H.data - blk["H"][2]).max() / np.abs(blk["H"][2]).max())
                if pattern else None,
                S_max_abs_err_over_max_abs=float(np.abs(Sb.data - blk["S"][2]).max() / np.abs(blk["S"][2]).max())
                if pattern else None)
    return dict(value=n_int / dt, unit="integrals/s", cores=int(min(nthreads, K)), kind="port",
                sample=(f"full {w.name} build (primitives + tensor, {n_int} integrals) by the oracle: "
                        f"OpenMP over the {K} multipoles as the reference's !$omp parallel do, "
                        f"serial assembly; {dt:.2f} s"),
                seconds=dt), rel, blocks


def solve_h_max(k, m, Z, r_max, n_b):
    """h_max whose grid (grid_tools.f90:6-55 rule) gives exactly n_b B-splines, or None.
    generate_grid refuses grids that overrun the reference's knot array (a
    linear part finer than 2^-m, or a geometric ramp longer than r_max); the
    search stays inside the h_max range it accepts."""
    import bs2e

    def nb(h):
        try:
            return len(bs2e.generate_grid(k, m, Z, h, r_max)) - k - 2
        except bs2e.Bs2eError:
            return None

    lo, hi = r_max / 4000.0, r_max
    while nb(lo) is None and lo < hi:
        lo *= 1.25
    while nb(hi) is None and hi > lo:
        hi /= 1.25
    n_lo, n_hi = nb(lo), nb(hi)
    if n_lo is None or n_hi is None or n_lo < n_b or n_hi > n_b:
        return None
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        v = nb(mid)
        if v is None:
            return None
        if v == n_b:
            return mid
        if v > n_b:
            lo = mid
        else:
            hi = mid
    return None


def serving_jobs(n, seed):
    """Serving stream jobs: B-spline count uniform in [40,400], order in {6,7,8,9},
    r_max uniform in [50,500], lmax uniform in [2,8], k <= 2 lmax, k_GL = order+6."""
    import numpy as np

    import bs2e
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(100 * n):
        if len(jobs) == n:
            break
        n_b = int(rng.integers(40, 401))
        k = int(rng.choice([6, 7, 8, 9]))
        r_max = float(rng.uniform(50.0, 500.0))
        lmax = int(rng.integers(2, 9))
        h = solve_h_max(k, 3, 2, r_max, n_b)
        if h is None:
            continue
        jobs.append((k, bs2e.generate_grid(k, 3, 2, h, r_max), k + 6, 2 * lmax))
    if len(jobs) < n:
        raise RuntimeError(f"serving_jobs: only {len(jobs)} of {n} jobs have a grid with the drawn size")
    return jobs


def run_stream(srv, jobs, arrivals):
    """Submits job i at arrivals[i] (s after the start), retires finished jobs;
    returns (latency ms per job, wall s)."""
    n = len(jobs)
    lat = [None] * n
    tickets = {}
    nxt = 0
    t0 = time.perf_counter()
    while nxt < n or tickets:
        now = time.perf_counter() - t0
        while nxt < n and arrivals[nxt] <= now:
            tickets[srv.submit(*jobs[nxt])] = nxt
            nxt += 1
        for tk, ms in srv.finished():
            lat[tickets.pop(tk)] = ms
            srv.retire(tk)
    return lat, time.perf_counter() - t0


def serving_benchmark(n_jobs, seed=20261202, slots=4, budget_gib=64, load=0.8):
    """Continuous batching of a seeded stream of differing builds through the
    build service: closed-loop capacity, then Poisson arrivals at `load` x
    capacity with p50/p99 latency (submission to completion)."""
    import numpy as np

    import bs2e
    jobs = serving_jobs(n_jobs, 7)
    integrals = []
    for k, t, kgl, mk in jobs:
        _, _, n_pairs = bs2e.count_nnz(k, t)
        integrals.append(n_pairs * n_pairs * (mk + 1))
    srv = bs2e.Server(budget_gib << 30, slots=slots)
    run_stream(srv, jobs[:8], [0.0] * 8)                     # warm-up
    _, wall_cap = run_stream(srv, jobs, [0.0] * n_jobs)       # all at once: capacity
    cap = n_jobs / wall_cap
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / (load * cap), n_jobs))
    lat, wall = run_stream(srv, jobs, arrivals)
    srv.close()
    lat = np.array(lat)
    return {"jobs": n_jobs, "slots": slots, "budget_GiB": budget_gib, "job_seed": 7, "arrival_seed": seed,
            "integrals_total": int(sum(integrals)),
            "capacity_jobs_per_s": cap, "capacity_integrals_per_s": sum(integrals) / wall_cap,
            "offered_load": load, "stream_integrals_per_s": sum(integrals) / wall,
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                           "mean": float(lat.mean()), "max": float(lat.max())}}


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


def blockset_benchmark(ctx, name="c3", max_l=4, max_L=4):
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
    ms_count, ms_fill = bset.kernel_ms()
    nH, nS = int(bset.sizes[:, 0].sum()), int(bset.sizes[:, 1].sum())
    bset.release()
    rk.release()
    out.update({"channels_per_symmetry": [int(len({(int(a), int(b)) for a, b in t_.configs[:, 2:4]})) for t_ in terms],
                "nnz_H": nH, "nnz_S": nS, "count_kernel_ms": ms_count, "fill_kernel_ms": ms_fill,
                "H_entries_per_s": nH / (ms_fill * 1e-3), "fill_GBs": 24.0 * (nH + nS) / (ms_fill * 1e-3) / 1e9,
                "wall_ms_create_and_fill": wall * 1e3})
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
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--latency-samples", type=int, default=50,
                    help="synchronised single builds timed for p50/p99 latency")
    ap.add_argument("--blocks", type=int, default=1, help="also time the L=0 symmetry block (0 skips it)")
    ap.add_argument("--blockset", type=int, default=1,
                    help="also time all symmetries L=0..4 (l <= 4) of the c3 basis in one launch (0 skips it)")
    ap.add_argument("--dipoles", type=int, default=1, help="also time the L=0 -> L=1 dipole block (0 skips it)")
    ap.add_argument("--serve-jobs", type=int, default=256,
                    help="jobs of the serving stream (differing bases through the build service); 0 skips it")
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
    from bs2e.workloads import WORKLOADS

    w = WORKLOADS[args.workload]
    torch.cuda.set_device(0)
    stream = torch.cuda.Stream(device=0)     # default priority; the build and its events share it
    ctx = bs2e.Context(0)
    ctx.set_stream(stream.cuda_stream)
    t = bs2e.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)
    rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx, compute=False)
    n_int = rk.n_pairs * rk.n_pairs * rk.K
    print(f"# bench {w.name}: n_b={rk.n_b} n_pairs={rk.n_pairs} K={rk.K} integrals/step={n_int}", file=sys.stderr,
          flush=True)

    rk.autotune()                 # fill shape and multipole chunk timed on this basis
    tuning = rk.tuning()
    for _ in range(args.warmup):
        rk.compute()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        rk.compute()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = n_int * args.steps / wall

    # The same K steps again with the library's HIP events recorded on both
    # build streams around every kernel stage: per-stage device times averaged
    # over the steps (and the profiled rate, to show the events cost nothing).
    rk.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rk.compute()
    torch.cuda.synchronize()
    wall_prof = time.perf_counter() - t0
    st = rk.stage_times()        # ms since start: windows, far end, diag end, near end, total
    rk.set_profiling(False)
    # the same K steps replayed from a captured hipGraph of the build
    rk.set_graph(True)
    rk.compute()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rk.compute()
    torch.cuda.synchronize()
    wall_graph = time.perf_counter() - t0
    rk.set_graph(False)
    t_far = float(st[1] - st[0]) * 1e-3
    far_elems = n_int - rk.near_elems
    traffic_alg = 8.0 * far_elems            # algorithmic bytes of one k3_far launch
    achieved = traffic_alg / t_far / 1e9
    # build latency: one build submitted to an idle device, host clock to completion
    lat = []
    for _ in range(args.latency_samples):
        torch.cuda.synchronize()
        a = time.perf_counter()
        rk.compute()
        rk.synchronize()
        lat.append((time.perf_counter() - a) * 1e3)
    lat = np.array(lat) if lat else np.zeros(1)
    # the practical store ceiling on this box: a store-only fill of a buffer of
    # the tensor's size on the same stream
    with torch.cuda.stream(stream):
        buf = torch.empty(n_int, dtype=torch.float64, device="cuda")
        buf.zero_()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(5):
            buf.zero_()
        e1.record(stream)
    torch.cuda.synchronize()
    store_ceiling = 8.0 * n_int * 5 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del buf
    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * (1 << 20) // 8
        with torch.cuda.stream(stream):
            src = torch.ones(n, dtype=torch.float64, device="cuda")
            dst = torch.empty_like(src)
            for _ in range(3):
                dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    R_gpu = rk.to_host() if not args.no_cpu_baseline else None

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "integrals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{w.name}: {w.description}", "n_b": rk.n_b, "n_pairs": rk.n_pairs, "K": rk.K,
                   "k_GL": w.k_GL, "integrals_per_step": n_int, "precision": "plain FP64",
                   "parallelism": "single GPU"},
        "timeline_ms": {"moments_and_windows_done": float(st[0]), "rank_one_fill_done": float(st[1]),
                        "diag_cells_done_side_stream": float(st[2]), "near_ranges_done_side_stream": float(st[3]),
                        "build_done": float(st[4])},
        "tensor_stage_GBs": 8.0 * n_int / (float(st[4] - st[0]) * 1e-3) / 1e9,
        "event_ms_per_step": ev0.elapsed_time(ev1) / args.steps,
        "profiled_ms_per_step": wall_prof * 1e3 / args.steps,
        "graph_ms_per_step": wall_graph * 1e3 / args.steps,
        "tuning": {"far_fill_shape": tuning[0], "diag_multipole_chunk": tuning[1]},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "samples": int(args.latency_samples)},
        "roofline": {"bound": "hbm", "kernel": "k3_far (rank-one tensor fill)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": FAR_TRAFFIC_BYTES.get((w.name, int(tuning[0]))),
                     "algorithmic_bytes": traffic_alg, "elements": far_elems,
                     "avg_launch_ms": t_far * 1e3, "store_ceiling_GBs": store_ceiling,
                     "frac_of_store_ceiling": achieved / store_ceiling},
    }
    if args.blocks:
        gpu_block = [] if not args.no_cpu_baseline else None
        out["blocks"] = blocks_benchmark(rk, w, t, keep=gpu_block)
    if args.dipoles:
        out["dipoles"] = dipoles_benchmark(rk.ctx, w, t)
    if args.blockset:
        out["blockset"] = blockset_benchmark(rk.ctx)
    if args.serve_jobs > 0:
        out["serving"] = serving_benchmark(args.serve_jobs)
    if not args.no_cpu_baseline:
        nthreads = min(16, os.cpu_count() or 1)
        cb, rel, cblocks = cpu_baseline(w, t, R_gpu, nthreads, block_max_l=2 if args.blocks else None,
                                        gpu_block=tuple(gpu_block) if args.blocks else None)
        out["cpu_baseline"] = cb
        if cblocks is not None:
            out["blocks"]["cpu_baseline"] = cblocks
            out["blocks"]["speedup_vs_cpu"] = out["blocks"]["H_entries_per_s"] / cblocks["value"]
        out["max_rel_err_vs_oracle"] = max(rel)
        out["max_rel_err_vs_oracle_per_k"] = rel
        out["speedup_vs_cpu"] = value / cb["value"]
    rk.release()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
