"""bs2e_rk_build_static: the tensor build into the caller's device buffer with
its work tables in the library's static scratch (no device allocation).

Checked through the C ABI against the allocating path (bs2e_rk_create +
bs2e_rk_compute) on two calls in a row, the second with a smaller basis,
bit for bit (the same kernels on the same inputs), in both precision modes;
the output buffer is filled with a NaN pattern before each call, so every
element must be written; a basis whose tables exceed the scratch and a wrong
output size are refused with BS2E_EINVAL and a message.  Device memory for
the outputs comes from the HIP runtime the library itself runs on.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import bs2e  # noqa: E402
from bs2e import capi  # noqa: E402
from bs2e.workloads import WORKLOADS, knots  # noqa: E402

vp, sz = C.c_void_p, C.c_size_t


def _hip():
    capi.lib()                                  # loads libbs2e.so and, with it, its HIP runtime
    h = C.CDLL("libamdhip64.so.7")              # the runtime already in the process
    h.hipMalloc.argtypes = [C.POINTER(vp), sz]
    h.hipFree.argtypes = [vp]
    h.hipMemset.argtypes = [vp, C.c_int, sz]
    h.hipMemcpy.argtypes = [vp, vp, sz, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(vp)]
    h.hipStreamSynchronize.argtypes = [vp]
    h.hipStreamDestroy.argtypes = [vp]
    h.hipSetDevice.argtypes = [C.c_int]
    for f in (h.hipMalloc, h.hipFree, h.hipMemset, h.hipMemcpy, h.hipStreamCreate, h.hipStreamSynchronize,
              h.hipStreamDestroy, h.hipSetDevice):
        f.restype = C.c_int
    return h


def _dims(w, t):
    n_b = len(t) - w.k - 2
    n_pairs = n_b * (2 * w.k - 1) - w.k * (w.k - 1)
    return n_pairs, n_pairs * n_pairs * (w.max_k + 1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_static_build_equals_allocating_build(precision):
    lib = capi.lib()
    prec = {"fp64": capi.PREC_FP64, "mixed": capi.PREC_MIXED}[precision]
    hip = _hip()
    assert hip.hipSetDevice(0) == 0
    stream = vp()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    ctx = bs2e.Context(0)
    P = C.POINTER(C.c_double)
    n_max = _dims(WORKLOADS["c2"], knots(WORKLOADS["c2"]))[1]
    dev = vp()
    assert hip.hipMalloc(C.byref(dev), 8 * n_max) == 0
    try:
        for name in ("c2", "c1"):          # the second call is smaller
            w = WORKLOADS[name]
            t = np.ascontiguousarray(knots(w), dtype=np.float64)
            n_pairs, n = _dims(w, t)
            print(f"static build {name} {precision}: {n} integrals", flush=True)
            assert hip.hipMemset(dev, 0xFF, 8 * n_max) == 0          # NaN in every element
            rc = lib.bs2e_rk_build_static(w.k, len(t), t.ctypes.data_as(P), w.k_GL, w.max_k, prec, dev, n, stream)
            assert rc == 0, lib.bs2e_last_error(None)
            assert hip.hipStreamSynchronize(stream) == 0
            got = np.empty(n)
            assert hip.hipMemcpy(got.ctypes.data_as(vp), dev, 8 * n, 2) == 0   # device to host
            rk = bs2e.RkTensor.build(w.k, t, w.k_GL, w.max_k, ctx=ctx, precision=precision)
            ref = rk.to_host().ravel()
            rk.release()
            assert np.array_equal(got, ref), name
        # refused: tables beyond the scratch (c5p: about 330 MB of work tables), a wrong output size
        w = WORKLOADS["c5p"]
        t = np.ascontiguousarray(knots(w), dtype=np.float64)
        rc = lib.bs2e_rk_build_static(w.k, len(t), t.ctypes.data_as(P), w.k_GL, w.max_k, prec, dev,
                                      _dims(w, t)[1], stream)
        assert rc == capi.BS2E_EINVAL and b"static scratch" in lib.bs2e_last_error(None)
        w = WORKLOADS["c1"]
        t = np.ascontiguousarray(knots(w), dtype=np.float64)
        rc = lib.bs2e_rk_build_static(w.k, len(t), t.ctypes.data_as(P), w.k_GL, w.max_k, prec, dev, 8, stream)
        assert rc == capi.BS2E_EINVAL and b"n_elems" in lib.bs2e_last_error(None)
    finally:
        hip.hipFree(dev)
        hip.hipStreamDestroy(stream)
        ctx.close()
