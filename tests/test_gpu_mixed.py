"""The mixed-precision build (BS2E_PREC_MIXED): the off-diagonal moments on the
FP32 matrix cores from exact integer slices, everything else FP64.

Bar (against the FP64 oracle, the reference's arithmetic): per multipole k,
the normwise error max|R - R_ref| / max|R_ref| <= 1e-10, and every element
within 1e-10 relative; the index pattern is the plain path's (the same dense
banded layout).  The builds are deterministic (bit-identical repeats), and a
basis the mixed moments cannot take (k_GL > 16) is refused.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import bs2e  # noqa: E402
from bs2e import capi  # noqa: E402
from bs2e.workloads import WORKLOADS, knots  # noqa: E402

TOL = 1e-10


@pytest.fixture(scope="module")
def ctx():
    return bs2e.Context(0)


def check(got, ref, label):
    K = ref.shape[-1]
    g, r = got.reshape(-1, K), ref.reshape(-1, K)
    d = np.abs(g - r)
    normwise = d.max(axis=0) / np.abs(r).max(axis=0)
    rel = d / r
    print(f"{label}: per-k normwise max {normwise.max():.3e}, element-wise rel p50 {np.percentile(rel, 50):.3e} "
          f"p99 {np.percentile(rel, 99):.3e} max {rel.max():.3e}", flush=True)
    assert normwise.max() <= TOL, normwise
    assert rel.max() <= TOL


@pytest.mark.parametrize("name", ["tiny", "t0", "c1", "c2"])
def test_mixed_tensor_matches_fp64_oracle(name, ctx, oracle):
    w = WORKLOADS[name]
    t = knots(w)
    n_b = oracle.dims(w.k, t)[1]
    rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx, precision="mixed")
    R = rk.to_host()
    rk.compute()
    assert np.array_equal(rk.to_host(), R)            # deterministic
    rk.release()
    ref = oracle.rk_dense(w.k, n_b, w.max_k, oracle.slater_off_diag(w.k, t, w.k_GL, w.max_k),
                          oracle.slater_diag(w.k, t, w.k_GL, w.max_k))
    assert R.shape == ref.shape
    check(R, ref, f"mixed {name}")


@pytest.mark.parametrize("k,k_gl,max_k", [(2, 4, 0), (3, 5, 2), (6, 16, 3), (9, 15, 6), (12, 16, 3), (16, 16, 2)])
def test_mixed_order_sweep(k, k_gl, max_k, ctx, oracle):
    t = bs2e.generate_grid(k, 3, 2, 0.5, 6.0 if k <= 12 else 4.0)
    n_b = oracle.dims(k, t)[1]
    rk = bs2e.RkTensor.build(k, t, k_gl, max_k, ctx=ctx, precision="mixed")
    R = rk.to_host()
    rk.release()
    ref = oracle.rk_dense(k, n_b, max_k, oracle.slater_off_diag(k, t, k_gl, max_k),
                          oracle.slater_diag(k, t, k_gl, max_k))
    check(R, ref, f"mixed order {k} GL {k_gl} max_k {max_k}")


def test_mixed_c5p_sampled_rows(ctx, oracle):
    """The largest configuration (order 9, 600 B-splines, k <= 24, perturbed
    breakpoints): 34 rows against the oracle's rows, and the whole tensor
    against the plain FP64 device build (same layout, every element within
    1e-10 relative)."""
    w = WORKLOADS["c5p"]
    t = knots(w)
    n_b = oracle.dims(w.k, t)[1]
    rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx, precision="mixed")
    n_pairs = rk.n_pairs
    rng = np.random.default_rng(17)
    rows = np.unique(np.concatenate([[0, n_pairs - 1], rng.choice(n_pairs, size=32, replace=False)]))
    got = rk.rows(rows)
    mask = np.zeros(n_pairs, bool)
    mask[rows] = True
    ref = oracle.rk_rows(w.k, n_b, w.max_k, oracle.slater_off_diag(w.k, t, w.k_GL, w.max_k),
                         oracle.slater_diag(w.k, t, w.k_GL, w.max_k, pair_mask=mask), rows)
    check(got, ref, "mixed c5p rows")
    plain = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx)
    worst = 0.0
    for lo in range(0, n_pairs, 1024):   # row blocks of the two device tensors
        hi = min(n_pairs, lo + 1024)
        a = rk.rows(np.arange(lo, hi))
        b = plain.rows(np.arange(lo, hi))
        worst = max(worst, float((np.abs(a - b) / b).max()))
    print(f"mixed vs plain device tensor c5p: max rel {worst:.3e}", flush=True)
    assert worst <= TOL
    plain.release()
    rk.release()


def test_mixed_refuses_long_gl_rules(ctx):
    w = WORKLOADS["c1"]
    t = knots(w)
    with pytest.raises(capi.Bs2eError) as e:
        bs2e.RkTensor.build(w.k, t, 17, w.max_k, ctx=ctx, precision="mixed")
    assert e.value.code == capi.BS2E_EINVAL
