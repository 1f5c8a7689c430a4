// capi_onebody.hip -- C ABI of the one-electron matrices (overlap and
// hydrogenic one-particle Hamiltonian with a complex absorbing potential).
#include <cmath>
#include <cstring>
#include <vector>

#include "capi_internal.h"

using namespace bs2e;

static int64_t one_body(bs2e_ctx* ctx, const char* who, int64_t k, int64_t n_knots, const double* knots, int64_t k_gl,
                        int64_t Z, int64_t l_min, int64_t l_max, int64_t cap_order, double cap_r0, double eta_re,
                        double eta_im, double* S_out, double* H_out)
{
    if (!ctx) return fail(nullptr, BS2E_EINVAL, std::string(who) + ": ctx is NULL");
    if (!knots || (!S_out && !H_out)) return fail(ctx, BS2E_EINVAL, std::string(who) + ": NULL argument");
    if (l_min < 0 || l_max < l_min || l_max > 1000 || cap_order < 0 || cap_order > 64)
        return fail(ctx, BS2E_EINVAL, std::string(who) + ": need 0 <= l_min <= l_max and 0 <= CAP order <= 64");
    Plan p;
    std::string msg = make_plan(k, n_knots, knots, k_gl, 0, p);
    if (!msg.empty()) return fail(ctx, BS2E_EINVAL, std::string(who) + ": " + msg);
    if (p.k_s > 16) return fail(ctx, BS2E_EINVAL, std::string(who) + ": spline orders up to 16 are supported");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const int n_l = H_out ? (int)(l_max - l_min + 1) : 1;
    const size_t nn = (size_t)p.n_b * p.n_b;
    const size_t tab = (size_t)p.n_iv * p.k_gl * p.k_s;
    DevTables d;
    double *B = nullptr, *B2 = nullptr, *rw = nullptr, *dS = nullptr, *dH = nullptr;
    auto cleanup = [&]() {
        dev_free(d.t);
        dev_free(d.bp);
        dev_free(d.x);
        dev_free(d.w);
        dev_free(B);
        dev_free(B2);
        dev_free(rw);
        dev_free(dS);
        dev_free(dH);
    };
    auto run = [&]() -> int64_t {
        HIP_TRY(ctx, dev_alloc(&d.t, p.n_knots));
        HIP_TRY(ctx, dev_alloc(&d.bp, p.n_iv + 1));
        HIP_TRY(ctx, dev_alloc(&d.x, p.k_gl));
        HIP_TRY(ctx, dev_alloc(&d.w, p.k_gl));
        HIP_TRY(ctx, dev_alloc(&B, tab));
        HIP_TRY(ctx, dev_alloc(&B2, tab));
        HIP_TRY(ctx, dev_alloc(&rw, 2 * (size_t)p.n_iv * p.k_gl));
        if (S_out) HIP_TRY(ctx, dev_alloc(&dS, 2 * nn));
        if (H_out) HIP_TRY(ctx, dev_alloc(&dH, 2 * nn * n_l));
        hipStream_t s = ctx->stream;
        HIP_TRY(ctx, hipMemcpyAsync(d.t, p.t.data(), 8 * p.n_knots, hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(d.bp, p.bp.data(), 8 * (p.n_iv + 1), hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(d.x, p.x.data(), 8 * p.k_gl, hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(d.w, p.w.data(), 8 * p.k_gl, hipMemcpyHostToDevice, s));
        clear_stale_error();
        HIP_TRY(ctx, launch_one_body(p, d, B, B2, rw, (int)l_min, n_l, (int)Z, (int)cap_order, cap_r0, eta_re, eta_im,
                                     dS, dH, s));
        if (S_out) HIP_TRY(ctx, hipMemcpyAsync(S_out, dS, 16 * nn, hipMemcpyDeviceToHost, s));
        if (H_out) HIP_TRY(ctx, hipMemcpyAsync(H_out, dH, 16 * nn * n_l, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        return BS2E_OK;
    };
    const int64_t rc = run();
    cleanup();
    return rc;
}

static int64_t rk_int(bs2e_ctx* ctx, int64_t k, int64_t n_knots, const double* knots, int64_t k_gl, int64_t max_k,
                      double r0, double* out)
{
    const char* who = "bs2e_setup_r_k_int";
    if (!ctx) return fail(nullptr, BS2E_EINVAL, std::string(who) + ": ctx is NULL");
    if (!knots || !out) return fail(ctx, BS2E_EINVAL, std::string(who) + ": NULL argument");
    if (max_k < 0 || max_k > 200 || !(r0 > 0.0) || !std::isfinite(r0))
        return fail(ctx, BS2E_EINVAL, std::string(who) + ": need 0 <= max_k <= 200 and a finite r_0 > 0");
    Plan p;
    std::string msg = make_plan(k, n_knots, knots, k_gl, 0, p);
    if (!msg.empty()) return fail(ctx, BS2E_EINVAL, std::string(who) + ": " + msg);
    if (p.k_s > 16) return fail(ctx, BS2E_EINVAL, std::string(who) + ": spline orders up to 16 are supported");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t n_out = (size_t)(max_k + 1) * p.n_b * p.n_b;
    const size_t tab = (size_t)p.n_iv * p.k_gl * p.k_s;
    DevTables d;
    double *B = nullptr, *B2 = nullptr, *rw = nullptr, *dout = nullptr;
    auto cleanup = [&]() {
        dev_free(d.t);
        dev_free(d.bp);
        dev_free(d.x);
        dev_free(d.w);
        dev_free(B);
        dev_free(B2);
        dev_free(rw);
        dev_free(dout);
    };
    auto run = [&]() -> int64_t {
        HIP_TRY(ctx, dev_alloc(&d.t, p.n_knots));
        HIP_TRY(ctx, dev_alloc(&d.bp, p.n_iv + 1));
        HIP_TRY(ctx, dev_alloc(&d.x, p.k_gl));
        HIP_TRY(ctx, dev_alloc(&d.w, p.k_gl));
        HIP_TRY(ctx, dev_alloc(&B, tab));
        HIP_TRY(ctx, dev_alloc(&B2, tab));
        HIP_TRY(ctx, dev_alloc(&rw, 2 * (size_t)p.n_iv * p.k_gl));
        HIP_TRY(ctx, dev_alloc(&dout, n_out));
        hipStream_t s = ctx->stream;
        HIP_TRY(ctx, hipMemcpyAsync(d.t, p.t.data(), 8 * p.n_knots, hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(d.bp, p.bp.data(), 8 * (p.n_iv + 1), hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(d.x, p.x.data(), 8 * p.k_gl, hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(d.w, p.w.data(), 8 * p.k_gl, hipMemcpyHostToDevice, s));
        clear_stale_error();
        HIP_TRY(ctx, launch_rk_int(p, d, B, B2, rw, (int)max_k, r0, dout, s));
        HIP_TRY(ctx, hipMemcpyAsync(out, dout, 8 * n_out, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        return BS2E_OK;
    };
    const int64_t rc = run();
    cleanup();
    return rc;
}

extern "C" {

int64_t bs2e_setup_r_k_int(bs2e_ctx* ctx, int64_t k, int64_t n_knots, const double* knots, int64_t k_gl,
                           int64_t max_k, double r_0, double* r_k_int)
{
    return rk_int(ctx, k, n_knots, knots, k_gl, max_k, r_0, r_k_int);
}

int64_t bs2e_setup_S(bs2e_ctx* ctx, int64_t k, int64_t n_knots, const double* knots, int64_t k_gl, double* S)
{
    return one_body(ctx, "bs2e_setup_S", k, n_knots, knots, k_gl, 0, 0, 0, 0, 0.0, 0.0, 0.0, S, nullptr);
}

int64_t bs2e_setup_H_one_particle(bs2e_ctx* ctx, int64_t k, int64_t n_knots, const double* knots, int64_t k_gl,
                                  int64_t Z, int64_t l_min, int64_t l_max, int64_t cap_order, double cap_r0,
                                  double cap_eta_re, double cap_eta_im, double* H)
{
    return one_body(ctx, "bs2e_setup_H_one_particle", k, n_knots, knots, k_gl, Z, l_min, l_max, cap_order, cap_r0,
                    cap_eta_re, cap_eta_im, nullptr, H);
}

}  // extern "C"
