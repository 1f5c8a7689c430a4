// capi_dipole.hip -- C ABI of the dipole operators: radial dipole matrices
// (length and velocity gauge) and the CSR dipole blocks between two
// symmetries, counted and filled on the device.
#include <cmath>
#include <cstring>
#include <new>

#include "capi_internal.h"

using namespace bs2e;

static double neg1(int64_t e) { return (std::llabs(e) & 1) ? -1.0 : 1.0; }

struct bs2e_dip {
    bs2e_ctx* ctx = nullptr;
    int64_t ks = 0, n_b = 0, n1 = 0, n2 = 0, L1 = 0, L2 = 0, max_l = 0;
    bool empty = true;                  // parity, angular factor or compute = 0: no entries
    double ang = 0.0;
    float ms_count = 0.f, ms_fill = 0.f;
    std::vector<BlockChan> chan1, chan2;
    std::vector<BlockSeg> seg1, seg2;
    std::vector<int32_t> cfg_chan1, cfg_chan2;
    std::vector<DipCoef> coef;
    std::vector<int64_t> cnt, ptr;
    int64_t nnz = 0;
    int64_t* d_cfg1 = nullptr;
    int32_t* d_cfg_chan1 = nullptr;
    BlockChan* d_chan2 = nullptr;
    BlockSeg* d_seg2 = nullptr;
    DipCoef* d_coef = nullptr;
    int64_t* d_cnt = nullptr;
    int64_t* d_ptr = nullptr;
};

static void dip_free(bs2e_dip* b)
{
    dev_free(b->d_cfg1);
    dev_free(b->d_cfg_chan1);
    dev_free(b->d_chan2);
    dev_free(b->d_seg2);
    dev_free(b->d_coef);
    dev_free(b->d_cnt);
    dev_free(b->d_ptr);
}

// mat_els.f90:739-770 dip_red_mat and :811-828 ang_dip_red, per (row channel
// (l_a, l_b), column channel (l_c, l_d)): the four terms of dip_mat_neq
// (:717-737) -- direct electron 1 and 2, then the same with the column's
// electrons exchanged (confs_ex(2) = (n_d, n_c; l_d, l_c)).
static DipCoef dip_coef(int64_t L1, int64_t L2, int64_t la, int64_t lb, int64_t lc, int64_t ld)
{
    DipCoef c;
    std::memset(&c, 0, sizeof(c));
    const double sq = std::sqrt((double)((2 * L1 + 1) * (2 * L2 + 1)));
    // term: (condition, sign exponent, 6j arguments j4 j5 j6, l and l' of the 1p element)
    struct T {
        bool cond;
        int64_t e, j4, j5, j6, l, lp;
    };
    const T t[4] = {
        {lb == ld, la + lb + L2 + 1, lc, lb, la, la, lc},   // direct, electron 1
        {la == lc, la + ld + L1 + 1, ld, la, lb, lb, ld},   // direct, electron 2
        {lb == lc, la + lb + L2 + 1, ld, lb, la, la, ld},   // exchange, electron 1
        {la == ld, la + lc + L1 + 1, lc, la, lb, lb, lc},   // exchange, electron 2
    };
    for (int i = 0; i < 4; ++i) {
        c.cond[i] = t[i].cond ? 1 : 0;
        if (!t[i].cond) continue;
        c.pf[i] = neg1(t[i].e) * sq * wigner_6j(L1, 1, L2, t[i].j4, t[i].j5, t[i].j6);
        c.cred[i] = c_red(1, t[i].l, t[i].lp);
        const int64_t l = t[i].l, lp = t[i].lp;
        if (std::llabs(l - lp) != 1) {
            c.vkind[i] = 0;
        } else if (l > lp) {
            c.vkind[i] = 1;
            c.vsq[i] = std::sqrt((double)l);
            c.vm[i] = (double)(lp + 1);
        } else {
            c.vkind[i] = 2;
            c.vsq[i] = std::sqrt((double)lp);
            c.vm[i] = (double)lp;
        }
    }
    double adr = 0.0;   // ang_dip_red: the direct terms only
    if (lb == ld) adr = adr + std::fabs(wigner_6j(L1, 1, L2, lc, lb, la) * c_red(1, la, lc));
    if (la == lc) adr = adr + std::fabs(wigner_6j(L1, 1, L2, ld, la, lb) * c_red(1, lb, ld));
    c.nz = adr > 5e-16 ? 1 : 0;
    c.sgn_ex = neg1(L2 + lc + ld);
    return c;
}

static DipArgs dip_args(const bs2e_dip* b)
{
    DipArgs a{};
    a.n1 = (int)b->n1;
    a.n_chan2 = (int)b->chan2.size();
    a.ks = (int)b->ks;
    a.n_b = (int)b->n_b;
    a.ang = b->ang;
    a.cfg1 = b->d_cfg1;
    a.cfg_chan1 = b->d_cfg_chan1;
    a.chan2 = b->d_chan2;
    a.seg2 = b->d_seg2;
    a.coef = b->d_coef;
    return a;
}

extern "C" {

int64_t bs2e_setup_radial_dip(bs2e_ctx* ctx, int64_t k, int64_t n_knots, const double* knots, int64_t k_gl,
                              int64_t gauge, double* m1, double* m2)
{
    if (!ctx) return fail(nullptr, BS2E_EINVAL, "bs2e_setup_radial_dip: ctx is NULL");
    if (!knots || !m1 || (gauge == 1 && !m2)) return fail(ctx, BS2E_EINVAL, "bs2e_setup_radial_dip: NULL argument");
    if (gauge != 0 && gauge != 1)
        return fail(ctx, BS2E_EINVAL, "bs2e_setup_radial_dip: gauge must be 0 (length) or 1 (velocity)");
    Plan p;
    std::string msg = make_plan(k, n_knots, knots, k_gl, 0, p);
    if (!msg.empty()) return fail(ctx, BS2E_EINVAL, "bs2e_setup_radial_dip: " + msg);
    if (p.k_s > 16) return fail(ctx, BS2E_EINVAL, "bs2e_setup_radial_dip: spline orders up to 16 are supported");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t nn = (size_t)p.n_b * p.n_b;
    const size_t tab = (size_t)p.n_iv * p.k_gl * p.k_s;
    DevTables d;
    double *B = nullptr, *B1 = nullptr, *rw = nullptr, *dM1 = nullptr, *dM2 = nullptr;
    auto cleanup = [&]() {
        dev_free(d.t);
        dev_free(d.bp);
        dev_free(d.x);
        dev_free(d.w);
        dev_free(B);
        dev_free(B1);
        dev_free(rw);
        dev_free(dM1);
        dev_free(dM2);
    };
    auto run = [&]() -> int64_t {
        HIP_TRY(ctx, dev_alloc(&d.t, p.n_knots));
        HIP_TRY(ctx, dev_alloc(&d.bp, p.n_iv + 1));
        HIP_TRY(ctx, dev_alloc(&d.x, p.k_gl));
        HIP_TRY(ctx, dev_alloc(&d.w, p.k_gl));
        HIP_TRY(ctx, dev_alloc(&B, tab));
        HIP_TRY(ctx, dev_alloc(&B1, tab));
        HIP_TRY(ctx, dev_alloc(&rw, 2 * (size_t)p.n_iv * p.k_gl));
        HIP_TRY(ctx, dev_alloc(&dM1, 2 * nn));
        if (gauge == 1) HIP_TRY(ctx, dev_alloc(&dM2, 2 * nn));
        hipStream_t s = ctx->stream;
        HIP_TRY(ctx, hipMemcpyAsync(d.t, p.t.data(), 8 * p.n_knots, hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(d.bp, p.bp.data(), 8 * (p.n_iv + 1), hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(d.x, p.x.data(), 8 * p.k_gl, hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(d.w, p.w.data(), 8 * p.k_gl, hipMemcpyHostToDevice, s));
        clear_stale_error();
        HIP_TRY(ctx, launch_radial_dipole(p, d, B, B1, rw, (int)gauge, dM1, dM2, s));
        HIP_TRY(ctx, hipMemcpyAsync(m1, dM1, 16 * nn, hipMemcpyDeviceToHost, s));
        if (gauge == 1) HIP_TRY(ctx, hipMemcpyAsync(m2, dM2, 16 * nn, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        return BS2E_OK;
    };
    const int64_t rc = run();
    cleanup();
    return rc;
}

int64_t bs2e_dip_block_create(bs2e_ctx* ctx, int64_t k, int64_t n_b, int64_t L1, int64_t M1, int64_t pi1, int64_t n1,
                              const int64_t* cf1, int64_t L2, int64_t M2, int64_t pi2, int64_t n2, const int64_t* cf2,
                              int64_t q, int64_t compute, bs2e_dip** out)
{
    if (!ctx) return fail(nullptr, BS2E_EINVAL, "bs2e_dip_block_create: ctx is NULL");
    if (!out || (n1 > 0 && !cf1) || (n2 > 0 && !cf2))
        return fail(ctx, BS2E_EINVAL, "bs2e_dip_block_create: NULL argument");
    *out = nullptr;
    if (k < 2 || k > 16 || n_b < 1 || n_b >= ((int64_t)1 << 30))
        return fail(ctx, BS2E_EINVAL, "bs2e_dip_block_create: need 2 <= k <= 16 and n_b >= 1");
    if (L1 < 0 || L2 < 0 || L1 > 1000 || L2 > 1000 || q < -1 || q > 1 || std::llabs(M1) > L1 || std::llabs(M2) > L2)
        return fail(ctx, BS2E_EINVAL, "bs2e_dip_block_create: need L >= 0, |M| <= L and q in {-1, 0, 1}");
    if (n1 < 1 || n2 < 1 || n1 >= ((int64_t)1 << 31) || n2 >= ((int64_t)1 << 31))
        return fail(ctx, BS2E_EINVAL, "bs2e_dip_block_create: need 1 <= n_config < 2^31 on both sides");
    bs2e_dip* b = new (std::nothrow) bs2e_dip();
    if (!b) return fail(ctx, BS2E_ENOMEM, "bs2e_dip_block_create: host allocation failed");
    b->ctx = ctx;
    b->ks = k;
    b->n_b = n_b;
    b->n1 = n1;
    b->n2 = n2;
    b->L1 = L1;
    b->L2 = L2;
    int64_t ml1 = 0, ml2 = 0;
    std::string bad = channel_tables(cf1, n1, n_b, b->chan1, b->seg1, b->cfg_chan1, ml1);
    if (bad.empty()) bad = channel_tables(cf2, n2, n_b, b->chan2, b->seg2, b->cfg_chan2, ml2);
    if (!bad.empty()) {
        delete b;
        return fail(ctx, BS2E_EINVAL, "bs2e_dip_block_create: " + bad + " (count_configs order required)");
    }
    b->max_l = std::max(ml1, ml2);
    // dipole.f90:22-31: parity change, Wigner-Eckart factor, compute flag
    const bool parity_allowed = (pi1 != 0) != (pi2 != 0);
    b->ang = neg1(L1 - M1) * wigner_3j(L1, 1, L2, -M1, q, M2);
    b->empty = std::fabs(b->ang) < 5e-16 || !parity_allowed || !compute;
    b->cnt.assign((size_t)n1, 0);
    b->ptr.assign((size_t)n1, 0);
    if (b->empty) {
        ctx_hold(ctx);
        *out = b;
        return BS2E_OK;
    }
    const size_t nc1 = b->chan1.size(), nc2 = b->chan2.size();
    b->coef.resize(nc1 * nc2);
    for (size_t i = 0; i < nc1; ++i)
        for (size_t j = 0; j < nc2; ++j)
            b->coef[i * nc2 + j] = dip_coef(L1, L2, b->chan1[i].l1, b->chan1[i].l2, b->chan2[j].l1, b->chan2[j].l2);
    auto run = [&]() -> int64_t {
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        HIP_TRY(ctx, dev_alloc(&b->d_cfg1, (size_t)5 * n1));
        HIP_TRY(ctx, dev_alloc(&b->d_cfg_chan1, (size_t)n1));
        HIP_TRY(ctx, dev_alloc(&b->d_chan2, nc2));
        HIP_TRY(ctx, dev_alloc(&b->d_seg2, b->seg2.size()));
        HIP_TRY(ctx, dev_alloc(&b->d_coef, b->coef.size()));
        HIP_TRY(ctx, dev_alloc(&b->d_cnt, (size_t)n1));
        HIP_TRY(ctx, dev_alloc(&b->d_ptr, (size_t)n1));
        hipStream_t s = ctx->stream;
        HIP_TRY(ctx, hipMemcpyAsync(b->d_cfg1, cf1, sizeof(int64_t) * 5 * n1, hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(b->d_cfg_chan1, b->cfg_chan1.data(), sizeof(int32_t) * n1, hipMemcpyHostToDevice,
                                    s));
        HIP_TRY(ctx, hipMemcpyAsync(b->d_chan2, b->chan2.data(), sizeof(BlockChan) * nc2, hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(b->d_seg2, b->seg2.data(), sizeof(BlockSeg) * b->seg2.size(),
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(b->d_coef, b->coef.data(), sizeof(DipCoef) * b->coef.size(),
                                    hipMemcpyHostToDevice, s));
        clear_stale_error();
        DipArgs a = dip_args(b);
        a.cnt = b->d_cnt;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIP_TRY(ctx, hipEventCreate(&e0));
        HIP_TRY(ctx, hipEventCreate(&e1));
        hipError_t le = hipEventRecord(e0, s);
        if (le == hipSuccess) le = launch_dip_count(a, s);
        if (le == hipSuccess) le = hipEventRecord(e1, s);
        if (le == hipSuccess)
            le = hipMemcpyAsync(b->cnt.data(), b->d_cnt, sizeof(int64_t) * n1, hipMemcpyDeviceToHost, s);
        if (le == hipSuccess) le = hipStreamSynchronize(s);
        if (le == hipSuccess) le = hipEventElapsedTime(&b->ms_count, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIP_TRY(ctx, le);
        for (int64_t i = 0; i < n1; ++i) {
            b->ptr[i] = b->nnz;
            b->nnz += b->cnt[i];
        }
        HIP_TRY(ctx, hipMemcpyAsync(b->d_ptr, b->ptr.data(), sizeof(int64_t) * n1, hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        return BS2E_OK;
    };
    const int64_t rc = run();
    if (rc) {
        dip_free(b);
        delete b;
        return rc;
    }
    ctx_hold(ctx);
    *out = b;
    return BS2E_OK;
}

int64_t bs2e_dip_block_sizes(const bs2e_dip* b, int64_t* sizes, int64_t n)
{
    if (!b || !sizes) return fail(nullptr, BS2E_EINVAL, "bs2e_dip_block_sizes: NULL argument");
    const int64_t v[3] = {b->nnz, b->n1, b->n2};
    for (int64_t i = 0; i < n && i < 3; ++i) sizes[i] = v[i];
    return BS2E_OK;
}

int64_t bs2e_dip_block_times(const bs2e_dip* b, double* ms, int64_t n)
{
    if (!b || !ms) return fail(nullptr, BS2E_EINVAL, "bs2e_dip_block_times: NULL argument");
    const double v[2] = {b->ms_count, b->ms_fill};
    for (int64_t i = 0; i < n && i < 2; ++i) ms[i] = v[i];
    return BS2E_OK;
}

int64_t bs2e_dip_block_fill(bs2e_dip* b, int64_t gauge, const double* S, const double* rad1, const double* rad2,
                            int64_t* ip, int64_t* ix, double* data)
{
    if (!b || !ip || (b->nnz > 0 && (!S || !rad1 || !ix || !data)) || (b->nnz > 0 && gauge == 1 && !rad2))
        return fail(b ? b->ctx : nullptr, BS2E_EINVAL, "bs2e_dip_block_fill: NULL argument");
    bs2e_ctx* ctx = b->ctx;
    if (gauge != 0 && gauge != 1)
        return fail(ctx, BS2E_EINVAL, "bs2e_dip_block_fill: gauge must be 0 (length) or 1 (velocity)");
    ip[0] = 1;
    for (int64_t i = 0; i < b->n1; ++i) ip[i + 1] = ip[i] + b->cnt[i];
    if (b->nnz == 0) return BS2E_OK;
    const size_t nn = (size_t)b->n_b * b->n_b;
    double *dS = nullptr, *dR1 = nullptr, *dR2 = nullptr, *dd = nullptr, *dT = nullptr;
    int64_t* dix = nullptr;
    auto cleanup = [&]() {
        dev_free(dT);
        dev_free(dS);
        dev_free(dR1);
        dev_free(dR2);
        dev_free(dd);
        dev_free(dix);
    };
    auto run = [&]() -> int64_t {
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        HIP_TRY(ctx, dev_alloc(&dS, 2 * nn));
        HIP_TRY(ctx, dev_alloc(&dR1, 2 * nn));
        if (gauge == 1) HIP_TRY(ctx, dev_alloc(&dR2, 2 * nn));
        HIP_TRY(ctx, dev_alloc(&dT, 2 * nn * (gauge == 1 ? 3 : 2)));   // row-major S, rad1 (, rad2)
        HIP_TRY(ctx, dev_alloc(&dd, 2 * (size_t)b->nnz));
        HIP_TRY(ctx, dev_alloc(&dix, (size_t)b->nnz));
        hipStream_t s = ctx->stream;
        HIP_TRY(ctx, hipMemcpyAsync(dS, S, 16 * nn, hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(dR1, rad1, 16 * nn, hipMemcpyHostToDevice, s));
        if (gauge == 1) HIP_TRY(ctx, hipMemcpyAsync(dR2, rad2, 16 * nn, hipMemcpyHostToDevice, s));
        clear_stale_error();
        DipArgs a = dip_args(b);
        a.gauge = (int)gauge;
        a.S = dT;
        a.rad1 = dT + 2 * nn;
        a.rad2 = gauge == 1 ? dT + 4 * nn : a.rad1;
        a.ptr = b->d_ptr;
        a.lim = b->nnz;
        a.ix = dix;
        a.data = dd;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIP_TRY(ctx, hipEventCreate(&e0));
        HIP_TRY(ctx, hipEventCreate(&e1));
        hipError_t le = hipEventRecord(e0, s);
        if (le == hipSuccess)
            le = launch_dip_prepare(dS, dR1, gauge == 1 ? dR2 : nullptr, (int)b->n_b, dT, dT + 2 * nn,
                                    gauge == 1 ? dT + 4 * nn : nullptr, s);
        if (le == hipSuccess) le = launch_dip_fill(a, s);
        if (le == hipSuccess) le = hipEventRecord(e1, s);
        if (le == hipSuccess) le = hipEventSynchronize(e1);
        if (le == hipSuccess) le = hipEventElapsedTime(&b->ms_fill, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIP_TRY(ctx, le);
        HIP_TRY(ctx, hipMemcpyAsync(ix, dix, sizeof(int64_t) * b->nnz, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(data, dd, 16 * (size_t)b->nnz, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        return BS2E_OK;
    };
    const int64_t rc = run();
    cleanup();
    return rc;
}

int64_t bs2e_dip_block_release(bs2e_dip* b)
{
    if (!b) return BS2E_OK;
    bs2e_ctx* ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    dip_free(b);
    delete b;
    ctx_drop(ctx);
    return BS2E_OK;
}

}  // extern "C"
