"""Graph replay of a build and the build service (bs2e_server_*).

* A captured hipGraph of the two-stream build gives the eager build's tensor
  bit for bit, on every replay, with other eager builds and synchronisations
  between the replays.
* Jobs of differing bases admitted concurrently by the service give, each,
  the tensor the same basis gives when built alone (bit for bit: the same
  kernels, no atomics, nothing shared between jobs), and the service never
  holds more device memory than its budget.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import bs2e  # noqa: E402
from bs2e import capi  # noqa: E402
from bs2e.workloads import WORKLOADS  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    return bs2e.Context(0)


def knots_of(name):
    w = WORKLOADS[name]
    return w, bs2e.generate_grid(w.k, w.m, w.Z, w.h_max, w.r_max)


def test_graph_replays_match_eager(ctx):
    w, t = knots_of("c1")
    eager = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx)
    ref = eager.to_host()
    g = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx, compute=False)
    g.set_graph(True)
    w0, t0 = knots_of("t0")
    other = bs2e.RkTensor.build(w0.k, t0, w0.k_GL, w0.max_k, ctx=ctx, compute=False)
    for rep in range(4):
        print(f"graph replay {rep}", flush=True)
        g.compute()
        assert np.array_equal(g.to_host(), ref), rep
        other.compute()               # eager work between replays
        other.synchronize()
    with pytest.raises(capi.Bs2eError) as e:
        g.set_profiling(True)
    assert e.value.code == capi.BS2E_ESTATE
    g.set_graph(False)
    g.compute()
    assert np.array_equal(g.to_host(), ref)
    for x in (g, other, eager):
        x.release()


def jobs():
    out = []
    for name in ("tiny", "t0", "c1"):
        w, t = knots_of(name)
        out.append((w.k, t, w.k_GL, w.max_k))
    rng = np.random.default_rng(5)
    for _ in range(7):
        k = int(rng.integers(5, 10))
        t = bs2e.generate_grid(k, 3, 2, float(rng.uniform(1.0, 3.0)), float(rng.uniform(20.0, 90.0)))
        out.append((k, t, k + 4, int(rng.integers(2, 9))))
    return out


def test_service_results_equal_alone(ctx):
    js = jobs()
    alone = []
    for k, t, kgl, mk in js:
        r = bs2e.RkTensor.build(k, t, kgl, mk, ctx=ctx)
        alone.append(r.to_host())
        r.release()
    fp = [bs2e.footprint(k, t, kgl, mk) for k, t, kgl, mk in js]
    budget = int(2.5 * max(fp))
    srv = bs2e.Server(budget, slots=3)
    tickets = []
    for n, (k, t, kgl, mk) in enumerate(js):
        tickets.append(srv.submit(k, t, kgl, mk))
        cap, used = srv.memory()
        assert cap >= budget and used <= cap
        if n == 2:
            # three slots and room for them: the first three jobs were admitted
            # at submission, none waited for another to finish
            states = [srv.status(x)[0] for x in tickets]
            assert all(st in (bs2e.Server.RUNNING, bs2e.Server.DONE) for st in states), states
    print(f"service: {len(js)} jobs, budget {budget / 1e6:.1f} MB", flush=True)
    for i, tk in enumerate(tickets):
        srv.wait(tk)
        st, t_admit, t_done = srv.status(tk)
        assert st == bs2e.Server.DONE and 0.0 <= t_admit <= t_done
        got = srv.tensor(tk).to_host()
        assert np.array_equal(got, alone[i]), i
        srv.retire(tk)
    assert srv.memory()[1] == 0
    # a job larger than the whole budget is refused at submission
    small = bs2e.Server(max(fp) // 4, slots=1)
    k, t, kgl, mk = js[int(np.argmax(fp))]
    with pytest.raises(capi.Bs2eError) as e:
        small.submit(k, t, kgl, mk)
    assert e.value.code == capi.BS2E_ENOMEM
    small.close()
    srv.close()


def test_service_waits_for_memory(ctx):
    """With room for one job at a time, later jobs start only after earlier
    ones are retired, and wait() reports a job that cannot be admitted."""
    js = jobs()[:3][::-1]          # c1, t0, tiny: the first takes nearly the whole budget
    fp = [bs2e.footprint(k, t, kgl, mk) for k, t, kgl, mk in js]
    assert fp[0] == max(fp) and min(fp) > 4096
    srv = bs2e.Server(max(fp) + 4096, slots=2)
    tk = [srv.submit(*j) for j in js]
    assert srv.status(tk[0])[0] in (bs2e.Server.RUNNING, bs2e.Server.DONE)
    assert srv.status(tk[1])[0] == bs2e.Server.QUEUED
    srv.wait(tk[0])
    done = srv.finished()
    assert [t for t, _ in done] == [tk[0]] and done[0][1] > 0.0
    assert srv.finished() == []
    with pytest.raises(capi.Bs2eError) as e:      # job 1 waits for job 0's memory, which is not retired
        srv.wait(tk[1])
    assert e.value.code == capi.BS2E_ESTATE
    srv.retire(tk[0])
    for x in tk[1:]:
        srv.wait(x)
        srv.retire(x)
    assert srv.memory()[1] == 0
    srv.close()


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_autotune_keeps_values(ctx, name):
    """Every rank-one fill shape and every diagonal-cell multipole chunk,
    pinned one at a time, computes the same tensor, bit for bit; the
    autotuner picks one of them, and a second basis of the same shape reuses
    the choice."""
    w, t = knots_of(name)
    ref_rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx)
    ref = ref_rk.to_host()
    ref_rk.release()
    rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx, compute=False)
    shapes, chunks = [], []
    for shape in range(16):
        try:
            rk.set_tuning(shape, 1)
        except bs2e.Bs2eError:
            continue
        rk.compute()
        assert np.array_equal(rk.to_host(), ref), shape
        shapes.append(shape)
    for kc in range(1, 17):
        try:
            rk.set_tuning(shapes[0], kc)
        except bs2e.Bs2eError:
            continue
        rk.compute()
        assert np.array_equal(rk.to_host(), ref), kc
        chunks.append(kc)
    print(f"{name}: fill shapes {shapes}, multipole chunks {chunks} give the same tensor", flush=True)
    assert len(shapes) >= 3 and chunks and chunks == list(range(1, len(chunks) + 1))
    with pytest.raises(bs2e.Bs2eError):
        rk.set_tuning(shapes[0], chunks[-1] + 1)
    rk.autotune(force=True)
    shape, kc = rk.tuning()
    print(f"{name}: autotuned fill shape {shape}, multipole chunk {kc}", flush=True)
    assert shape in shapes and kc in chunks
    rk.compute()
    assert np.array_equal(rk.to_host(), ref)
    again = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx, compute=False)
    again.autotune()
    assert again.tuning() == (shape, kc)
    again.compute()
    assert np.array_equal(again.to_host(), ref)
    rk.release()
    again.release()


def test_service_serving_stream_jobs_equal_alone(ctx, oracle):
    """The serving stream's own job distribution (seed 7: B-spline counts in
    [40, 400], orders 6-9, r_max in [50, 500], k <= 2 lmax <= 16, GL order + 6)
    plus its largest possible job (order 9, 400 B-splines, k <= 16, 6.2 GB of
    tensor), all submitted at once to a service with four slots: each job's
    tensor equals the same basis built alone, bit for bit; and the largest
    job's tensor, contracted with the hydrogenic Z = 2 orbitals, gives the
    textbook Slater integrals F^0(1s,1s) = 5Z/8, G^1(1s,2p) = 112Z/2187 and
    F^2(2p,2p) = 45Z/512 to 1e-9."""
    import scipy.linalg as sl
    from bs2e.workloads import largest_serving_job, serving_jobs
    js = [(j.k, j.t, j.k_GL, j.max_k) for j in serving_jobs(9, 7)]
    big = largest_serving_job()
    js.insert(3, (big.k, big.t, big.k_GL, big.max_k))
    print("serving jobs (order, n_b, k_GL, max_k):", [(k, len(t) - k - 2, g, m) for k, t, g, m in js], flush=True)
    alone = []
    for k, t, kgl, mk in js:
        r = bs2e.RkTensor.build(k, t, kgl, mk, ctx=ctx)
        alone.append(r.to_host())
        r.release()
    srv = bs2e.Server(64 << 30, slots=4)
    tickets = [srv.submit(*j) for j in js]
    for i, tk in enumerate(tickets):
        srv.wait(tk)
        assert np.array_equal(srv.tensor(tk).to_host(), alone[i]), i
        srv.retire(tk)
    assert srv.memory()[1] == 0
    srv.close()
    # Slater integrals of the largest job's tensor
    k, t, kgl, mk = js[3]
    R = alone[3]
    n_b = oracle.dims(k, t)[1]
    pairs = oracle.pair_table(k, n_b)
    n_pairs, K = R.shape[0], R.shape[2]
    S = oracle.setup_S(k, t, kgl).real
    E0, V0 = sl.eigh(oracle.setup_H1p(k, t, kgl, 2, 0, cap_eta=0.0).real, S)
    E1, V1 = sl.eigh(oracle.setup_H1p(k, t, kgl, 2, 1, cap_eta=0.0).real, S)
    assert abs(E0[0] + 2.0) < 1e-9 and abs(E1[0] + 0.5) < 1e-9   # 1s, 2p of Z = 2: -Z^2 / (2 n^2)
    s1, p2 = V0[:, 0], V1[:, 0]
    flat = R.reshape(n_pairs, n_pairs * K)

    def slater(kk, a, b, c, d):
        A = a[pairs[:, 0] - 1] * c[pairs[:, 1] - 1]
        B = b[pairs[:, 0] - 1] * d[pairs[:, 1] - 1]
        return float((A @ flat).reshape(n_pairs, K)[:, kk] @ B)

    Z = 2
    for got, exact in ((slater(0, s1, s1, s1, s1), 5 * Z / 8), (slater(1, s1, p2, p2, s1), 112 * Z / 2187),
                       (slater(2, p2, p2, p2, p2), 45 * Z / 512)):
        assert abs(got - exact) < 1e-9 * exact, (got, exact)


def branch_terms(n_b, k, lmax, max_L=2, l2=1):
    """The symmetries of a basis_setup request: L = 0..max_L with natural
    parity (z_pol), l1 <= lmax, the second electron's n2 <= n_b // 3 (an
    r_2_max truncation) and its l <= l2 beyond n1 = n_b // 2 (r_all_l)."""
    terms = []
    for L in range(max_L + 1):
        term = bs2e.sym(l=L, pi=bool(L % 2))
        term.configs = bs2e.count_configs(term, lmax, n_b, k, n_b // 3, n_b // 2, min(l2, lmax))
        if term.n_config:
            terms.append(term)
    return terms


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_service_whole_branch_equals_alone(ctx, precision):
    """Jobs of the serving distribution as whole two-electron branches (tensor,
    one-electron matrices with an absorbing potential, the H/S blocks of
    L = 0, 1, 2), all submitted at once to a service with three slots: every
    job's tensor and every symmetry's CSR arrays equal the same job built
    alone through the single-build calls, bit for bit.  A second job on the
    first job's knots (fewer multipoles and l's) takes its one-electron tables
    from the service's cache."""
    from bs2e.workloads import serving_jobs
    picked = [j for j in serving_jobs(40, 7) if len(j.t) - j.k - 2 <= 160][:4]
    assert len(picked) == 4
    first = picked[0]
    specs = [(j.k, j.t, j.k_GL, j.max_k, j.lmax) for j in picked]
    specs.append((first.k, first.t, first.k_GL, max(0, first.max_k - 2), max(1, first.lmax - 1)))
    cap = (2, 0.0, 1e-3)
    srv = bs2e.Server(32 << 30, slots=3)
    tickets, terms_of = [], []
    for k, t, kgl, mk, lmax in specs:
        n_b = len(t) - k - 2
        terms = branch_terms(n_b, k, lmax)
        terms_of.append(terms)
        c = (cap[0], 0.8 * t[-1], cap[2])
        tickets.append(srv.submit_job(k, t, kgl, mk, precision=precision, terms=terms, full=False, Z=2,
                                      max_l_1p=lmax, CAP_c=c))
    print("whole-branch jobs (order, n_b, max_k, lmax, rows):",
          [(k, len(t) - k - 2, mk, lm, sum(tm.n_config for tm in tr)) for (k, t, _, mk, lm), tr in
           zip(specs, terms_of)], flush=True)
    got = []
    for tk, terms in zip(tickets, terms_of):
        srv.wait(tk)
        st, t_admit, t_done = srv.status(tk)
        assert st == bs2e.Server.DONE and 0.0 <= t_admit <= t_done
        bset = srv.blocks(tk, terms)
        got.append((srv.tensor(tk).to_host(), [bset.get(i) for i in range(len(terms))]))
    entries, nbytes, hits, misses = srv.cache_info()
    assert hits >= 1 and misses >= 1 and entries == misses and nbytes > 0
    for tk in tickets:
        srv.retire(tk)
    assert srv.memory()[1] == nbytes    # only the cache stays
    srv.close()
    for (k, t, kgl, mk, lmax), terms, (R, blocks) in zip(specs, terms_of, got):
        spl = bs2e.b_spline(k, t)
        rk = bs2e.RkTensor.build(k, t, kgl, mk, ctx=ctx, precision=precision)
        assert np.array_equal(rk.to_host(), R)
        S = bs2e.setup_S(spl, kgl, ctx=ctx)
        Hl = bs2e.setup_H_one_particle(2, (cap[0], 0.8 * t[-1], cap[2]), range(lmax + 1), spl, kgl, ctx=ctx)
        alone = bs2e.BlockSet(rk, terms, False)
        alone.fill(list(Hl), S)
        for i, (H_sp, S_sp) in enumerate(blocks):
            H1, S1 = alone.get(i)
            for a, b in ((H_sp, H1), (S_sp, S1)):
                assert np.array_equal(a.index_ptr, b.index_ptr) and np.array_equal(a.indices, b.indices)
                assert np.array_equal(a.data, b.data)
        alone.release()
        rk.release()


def _job_blocks(srv, tk, terms):
    bset = srv.blocks(tk, terms)
    return [bset.get(i) for i in range(len(terms))]


def _same_blocks(a, b):
    for (ha, sa), (hb, sb) in zip(a, b):
        for x, y in ((ha, hb), (sa, sb)):
            assert np.array_equal(x.index_ptr, y.index_ptr) and np.array_equal(x.indices, y.indices)
            assert np.array_equal(x.data, y.data)


def test_service_requeued_job_admitted_again(ctx):
    """A job whose tensor and tables fit but whose blocks' outputs do not (the
    first job holds them) goes back to the head of the queue once counted;
    admitted again after the first job is retired, it rebuilds its set from
    the host tables prepared at submission and gives the first job's H/S bit
    for bit (both are the same request)."""
    w, t = knots_of("c1")
    n_b = len(t) - w.k - 2
    terms = branch_terms(n_b, w.k, 2)
    cap = (2, 0.8 * t[-1], 1e-3)

    def submit(s):
        return s.submit_job(w.k, t, w.k_GL, w.max_k, precision="fp64", terms=terms, full=False, Z=2, max_l_1p=2,
                            CAP_c=cap)

    probe = bs2e.Server(16 << 30, slots=1)
    tk = submit(probe)
    probe.wait(tk)
    used = probe.memory()[1]               # tensor, tables, outputs and the one-electron cache entry
    probe.retire(tk)
    cache = probe.memory()[1]
    probe.close()
    fp = bs2e.footprint(w.k, t, w.k_GL, w.max_k)
    rest = used - cache - fp               # the job's block tables and outputs
    assert rest > 0
    srv = bs2e.Server(used + fp + rest // 2, slots=2)
    a = submit(srv)
    srv.wait(a)
    b = submit(srv)                        # admitted, counted, then refused its outputs: queued again
    seen = [srv.status(b)[0]]
    for _ in range(100000):
        srv.progress()
        st = srv.status(b)[0]
        if st != seen[-1]:
            seen.append(st)
        if st == bs2e.Server.QUEUED:
            break
    assert seen == [bs2e.Server.RUNNING, bs2e.Server.QUEUED], seen
    assert srv.memory()[1] == used         # the requeued job holds nothing
    first = _job_blocks(srv, a, terms)
    assert [x for x, _ in srv.finished()] == [a]
    srv.retire(a)
    srv.wait(b)
    assert srv.status(b)[0] == bs2e.Server.DONE
    _same_blocks(_job_blocks(srv, b, terms), first)
    srv.retire(b)
    srv.close()


def test_service_hands_out_failed_job(ctx):
    """A job whose configurations break count_configs order fails at
    admission: finished() hands it out with latency -1, its status is
    FAILED and the service's last error says why; a good job after it runs."""
    w, t = knots_of("c1")
    n_b = len(t) - w.k - 2
    terms = branch_terms(n_b, w.k, 2)
    bad = bs2e.sym(l=terms[0].l, pi=terms[0].pi)
    bad.configs = terms[0].configs[::-1].copy()   # n1 descending within each channel
    srv = bs2e.Server(16 << 30, slots=1)
    tk = srv.submit_job(w.k, t, w.k_GL, w.max_k, terms=[bad], full=False, Z=2, max_l_1p=2)
    done = []
    for _ in range(20000):
        done += srv.finished()
        if done:
            break
    assert done == [(tk, -1.0)]
    assert srv.status(tk)[0] == bs2e.Server.FAILED
    assert "count_configs order" in srv.last_error()
    good = srv.submit_job(w.k, t, w.k_GL, w.max_k, terms=terms, full=False, Z=2, max_l_1p=2)
    srv.wait(good)
    assert srv.status(good)[0] == bs2e.Server.DONE
    srv.retire(tk)
    srv.retire(good)
    srv.close()
