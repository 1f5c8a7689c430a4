// capi.hip -- the extern "C" boundary (include/bs2e.h) over the HIP kernels.
#include <algorithm>
#include <cstdio>
#include <functional>
#include <map>
#include <tuple>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "capi_internal.h"

using namespace bs2e;

static thread_local std::string g_err;

int64_t bs2e::fail(bs2e_ctx* ctx, int64_t code, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    g_err = msg;
    return code;
}

extern "C" {

int64_t bs2e_create(bs2e_ctx** out, int64_t device)
{
    if (!out) return fail(nullptr, BS2E_EINVAL, "bs2e_create: out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(nullptr, BS2E_EDEVICE, "bs2e_create: no HIP device available");
    if (device < 0 || device >= n) return fail(nullptr, BS2E_EINVAL, "bs2e_create: device index out of range");
    bs2e_ctx* c = new (std::nothrow) bs2e_ctx();
    if (!c) return fail(nullptr, BS2E_ENOMEM, "bs2e_create: host allocation failed");
    c->device = (int)device;
    e = hipSetDevice(c->device);
    // both streams at the default priority
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e != hipSuccess) {
        if (c->own) (void)hipStreamDestroy(c->own);
        delete c;
        return fail(nullptr, BS2E_EDEVICE, std::string("bs2e_create: ") + hipGetErrorString(e));
    }
    c->stream = c->own;
    *out = c;
    return BS2E_OK;
}

static void ctx_free_now(bs2e_ctx* ctx)
{
    (void)hipSetDevice(ctx->device);
    for (hipStream_t s : {ctx->own, ctx->side})
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    if (ctx->pin) (void)hipHostFree(ctx->pin);
    if (ctx->up) (void)hipHostFree(ctx->up);
    if (ctx->up_ev) (void)hipEventDestroy(ctx->up_ev);
    delete ctx;
}

int64_t bs2e_destroy(bs2e_ctx* ctx)
{
    if (!ctx) return BS2E_OK;
    if (ctx->refs > 0) {      // tensors still use its streams: freed with the last of them
        ctx->closing = true;
        return BS2E_OK;
    }
    ctx_free_now(ctx);
    return BS2E_OK;
}

const char* bs2e_last_error(const bs2e_ctx* ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int64_t bs2e_set_stream(bs2e_ctx* ctx, void* hip_stream)
{
    if (!ctx) return fail(nullptr, BS2E_EINVAL, "bs2e_set_stream: ctx is NULL");
    ctx->stream = hip_stream ? (hipStream_t)hip_stream : ctx->own;
    return BS2E_OK;
}

int64_t bs2e_generate_grid(int64_t k, int64_t m, int64_t Z, double h_max, double r_max, double* knots, int64_t cap)
{
    if (k < 1 || m < 0 || m > 60 || Z < 1 || !(h_max > 0.0) || !(r_max > 0.0))
        return fail(nullptr, BS2E_EINVAL, "bs2e_generate_grid: need k >= 1, 0 <= m <= 60, Z >= 1, h_max > 0, r_max > 0");
    std::vector<double> g;
    int64_t n = generate_grid(k, m, Z, h_max, r_max, g);
    if (n < 0)
        return fail(nullptr, BS2E_EINVAL,
                    "bs2e_generate_grid: the grid does not fit the reference's 2k + m + ceiling(r_max / 2^-m) knots "
                    "(grid_tools.f90:20; r_max too small for the geometric ramp to Z h_max)");
    if (knots) {
        if (cap < n) return fail(nullptr, BS2E_EINVAL, "bs2e_generate_grid: cap smaller than the knot count");
        std::memcpy(knots, g.data(), sizeof(double) * (size_t)n);
    }
    return n;
}

int64_t bs2e_count_nnz(int64_t k, int64_t n_knots, const double* knots, int64_t* nnz4, int64_t* nnz6,
                       int64_t* n_pairs)
{
    Plan p;
    std::string msg = make_plan(k, n_knots, knots, 1, 0, p);
    if (!msg.empty()) return fail(nullptr, BS2E_EINVAL, "bs2e_count_nnz: " + msg);
    if (nnz4) *nnz4 = p.nnz4;
    if (nnz6) *nnz6 = p.nnz6;
    if (n_pairs) *n_pairs = p.n_pairs;
    return BS2E_OK;
}

double bs2e_ang_k_LS(int64_t k, int64_t l1a, int64_t l1b, int64_t l2a, int64_t l2b, int64_t L)
{
    return ang_k_LS(k, l1a, l1b, l2a, l2b, L);
}

double bs2e_three_j(int64_t ja, int64_t jb, int64_t jc, int64_t ma, int64_t mb, int64_t mc)
{
    return wigner_3j(ja, jb, jc, ma, mb, mc);
}

double bs2e_six_j(int64_t j1, int64_t j2, int64_t j3, int64_t j4, int64_t j5, int64_t j6)
{
    return wigner_6j(j1, j2, j3, j4, j5, j6);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// device arena
// ---------------------------------------------------------------------------
hipError_t bs2e::Arena::init(size_t bytes)
{
    bytes = (bytes + 255) / 256 * 256;
    hipError_t e = hipMalloc((void**)&base_, bytes);
    if (e != hipSuccess) {
        base_ = nullptr;
        return e;
    }
    cap_ = bytes;
    used_ = 0;
    free_.clear();
    taken_.clear();
    free_[0] = bytes;
    return hipSuccess;
}

void bs2e::Arena::destroy()
{
    if (base_) (void)hipFree(base_);
    base_ = nullptr;
    cap_ = used_ = 0;
    free_.clear();
    taken_.clear();
}

void* bs2e::Arena::alloc(size_t bytes)
{
    std::lock_guard<std::mutex> lk(mu_);
    bytes = (bytes + 255) / 256 * 256;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < bytes) continue;
        const size_t off = it->first, rest = it->second - bytes;
        free_.erase(it);
        if (rest) free_[off + bytes] = rest;
        taken_[off] = bytes;
        used_ += bytes;
        return base_ + off;
    }
    return nullptr;
}

void bs2e::Arena::release(void* p)
{
    std::lock_guard<std::mutex> lk(mu_);
    const size_t off = (size_t)((char*)p - base_);
    auto t = taken_.find(off);
    if (t == taken_.end()) return;
    size_t start = off, size = t->second;
    used_ -= size;
    taken_.erase(t);
    auto next = free_.lower_bound(start);
    if (next != free_.end() && next->first == start + size) {   // merge with the following range
        size += next->second;
        next = free_.erase(next);
    }
    if (next != free_.begin()) {                                 // and with the preceding one
        auto prev = std::prev(next);
        if (prev->first + prev->second == start) {
            start = prev->first;
            size += prev->second;
            free_.erase(prev);
        }
    }
    free_[start] = size;
}

size_t bs2e::Arena::largest_free() const
{
    std::lock_guard<std::mutex> lk(mu_);
    size_t m = 0;
    for (const auto& f : free_) m = std::max(m, f.second);
    return m;
}

// ---------------------------------------------------------------------------
// tensor handle
// ---------------------------------------------------------------------------
static void free_tables(bs2e_rk* rk)
{
    DevTables& d = rk->d;
    ctx_free(rk->ctx, rk->small);   // knots, breakpoints, GL rule, pairs, near rows, c_first
    d.t = d.bp = d.x = d.w = nullptr;
    d.pairs = nullptr;
    d.rows = d.c_first = nullptr;
    ctx_free(rk->ctx, d.Pl);
    ctx_free(rk->ctx, d.Ql);
    ctx_free(rk->ctx, d.Pw);
    ctx_free(rk->ctx, d.Qw);
    ctx_free(rk->ctx, d.PP);
    ctx_free(rk->ctx, d.QS);
    ctx_free(rk->ctx, d.Ptot);
    ctx_free(rk->ctx, d.Qtot);
    ctx_free(rk->ctx, d.D);
    ctx_free(rk->ctx, rk->R);
    for (auto& set : rk->evpool)
        for (hipEvent_t e : set)
            if (e) (void)hipEventDestroy(e);
    rk->evpool.clear();
    rk->evused = 0;
    for (hipEvent_t* e : {&rk->windows, &rk->join})
        if (*e) {
            (void)hipEventDestroy(*e);
            *e = nullptr;
        }
    if (rk->graph_exec) (void)hipGraphExecDestroy(rk->graph_exec);
    if (rk->graph) (void)hipGraphDestroy(rk->graph);
    rk->graph_exec = nullptr;
    rk->graph = nullptr;
}

static int64_t create_events(bs2e_ctx* ctx, bs2e_rk* rk)
{
    for (hipEvent_t* e : {&rk->windows, &rk->join})
        HIP_TRY(ctx, hipEventCreateWithFlags(e, hipEventDisableTiming));
    return BS2E_OK;
}

// The small tables of a build, packed into one device block (256-byte aligned
// pieces): knots, breakpoints, GL nodes and weights, pairs, near rows, c_first.
namespace {
struct SmallLayout {
    size_t t, bp, x, w, pairs, rows, c_first, total;
};
SmallLayout small_layout(const Plan& p, bool with_tensor)
{
    auto r = [](size_t bytes) { return (std::max<size_t>(bytes, 1) + 255) / 256 * 256; };
    SmallLayout L{};
    size_t o = 0;
    L.t = o, o += r(8 * p.n_knots);
    L.bp = o, o += r(8 * (p.n_iv + 1));
    L.x = o, o += r(8 * p.k_gl);
    L.w = o, o += r(8 * p.k_gl);
    L.pairs = o, o += r(sizeof(PairInfo) * p.n_pairs);
    L.rows = o, o += with_tensor ? r(4 * p.rows.size()) : 0;
    L.c_first = o, o += r(4 * p.c_first.size());
    L.total = o;
    return L;
}
}  // namespace

// The allocations of alloc_tables(with_tensor = true), rounded as the arena rounds.
size_t bs2e::rk_footprint(const Plan& p)
{
    auto r = [](size_t n, size_t elem) { return ((n ? n : 1) * elem + 255) / 256 * 256; };
    const size_t slots = (size_t)p.n_pairs * p.k_s * p.K;
    size_t b = small_layout(p, true).total;
    b += 2 * r((size_t)p.n_iv * p.U * p.K, 8) + 4 * r(slots, 8) + 2 * r((size_t)p.n_pairs * p.K, 8);
    b += r((size_t)p.n_iv * p.K * p.U * p.U, 8);
    b += r((size_t)p.n_pairs * p.n_pairs * p.K, 8);
    return b;
}

// Allocates the build's device tables (in the context's arena when it has
// one) and enqueues the upload of the small ones on the context's stream as
// one copy from the handle's packed host image (which lives as long as the
// handle); the kernels that read them follow on the same stream.
static int64_t alloc_tables(bs2e_ctx* ctx, bs2e_rk* rk, bool with_tensor)
{
    if (with_tensor && rk->plan.rows.empty()) build_near_rows(rk->plan);
    const Plan& p = rk->plan;
    DevTables& d = rk->d;
    const size_t slots = (size_t)p.n_pairs * p.k_s * p.K;
    const SmallLayout L = small_layout(p, with_tensor);
    HIP_TRY(ctx, ctx_alloc(ctx, &rk->small, L.total));
    rk->small_host.assign(L.total, 0);
    char* h = rk->small_host.data();
    std::memcpy(h + L.t, p.t.data(), 8 * p.n_knots);
    std::memcpy(h + L.bp, p.bp.data(), 8 * (p.n_iv + 1));
    std::memcpy(h + L.x, p.x.data(), 8 * p.k_gl);
    std::memcpy(h + L.w, p.w.data(), 8 * p.k_gl);
    std::memcpy(h + L.pairs, p.pairs.data(), sizeof(PairInfo) * p.n_pairs);
    if (with_tensor) std::memcpy(h + L.rows, p.rows.data(), 4 * p.rows.size());
    std::memcpy(h + L.c_first, p.c_first.data(), 4 * p.c_first.size());
    d.t = (double*)(rk->small + L.t);
    d.bp = (double*)(rk->small + L.bp);
    d.x = (double*)(rk->small + L.x);
    d.w = (double*)(rk->small + L.w);
    d.pairs = (PairInfo*)(rk->small + L.pairs);
    d.rows = with_tensor ? (int32_t*)(rk->small + L.rows) : nullptr;
    d.c_first = (int32_t*)(rk->small + L.c_first);
    HIP_TRY(ctx, ctx_alloc(ctx, &d.Pl, (size_t)p.n_iv * p.U * p.K));
    HIP_TRY(ctx, ctx_alloc(ctx, &d.Ql, (size_t)p.n_iv * p.U * p.K));
    HIP_TRY(ctx, ctx_alloc(ctx, &d.Pw, slots));
    HIP_TRY(ctx, ctx_alloc(ctx, &d.Qw, slots));
    HIP_TRY(ctx, ctx_alloc(ctx, &d.PP, slots));
    HIP_TRY(ctx, ctx_alloc(ctx, &d.QS, slots));
    HIP_TRY(ctx, ctx_alloc(ctx, &d.Ptot, (size_t)p.n_pairs * p.K));
    HIP_TRY(ctx, ctx_alloc(ctx, &d.Qtot, (size_t)p.n_pairs * p.K));
    HIP_TRY(ctx, ctx_alloc(ctx, &d.D, (size_t)p.n_iv * p.K * p.U * p.U));
    if (with_tensor) {
        rk->R_elems = (size_t)p.n_pairs * p.n_pairs * p.K;
        HIP_TRY(ctx, ctx_alloc(ctx, &rk->R, rk->R_elems));
    }
    // up through the context's pinned staging (a copy from pageable memory can
    // hold the host until the stream reaches it), reused once the previous
    // upload from it has run; it grows by doubling.  On a stream being
    // captured the copy reads the tensor's own host image (a graph node keeps
    // its source).
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIP_TRY(ctx, hipStreamIsCapturing(ctx->stream, &cap));
    if (cap != hipStreamCaptureStatusNone) {
        HIP_TRY(ctx, hipMemcpyAsync(rk->small, h, L.total, hipMemcpyHostToDevice, ctx->stream));
        return BS2E_OK;
    }
    std::lock_guard<std::mutex> up_lock(ctx->up_mu);
    if (ctx->up_ev) HIP_TRY(ctx, hipEventSynchronize(ctx->up_ev));
    else HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->up_ev, hipEventDisableTiming));
    if (ctx->up_bytes < L.total) {
        const size_t grow = std::max(L.total, std::max<size_t>(2 * ctx->up_bytes, (size_t)1 << 20));
        if (ctx->up) (void)hipHostFree(ctx->up);
        ctx->up = nullptr;
        ctx->up_bytes = 0;
        HIP_TRY(ctx, hipHostMalloc(&ctx->up, grow, hipHostMallocDefault));
        ctx->up_bytes = grow;
    }
    std::memcpy(ctx->up, h, L.total);
    HIP_TRY(ctx, hipMemcpyAsync(rk->small, ctx->up, L.total, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->up_ev, ctx->stream));
    return BS2E_OK;
}

int64_t bs2e::check_plan_limits(bs2e_ctx* ctx, const Plan& p, bool needs_primitives)
{
    if (needs_primitives &&
        (p.k_s > 16 || inner_kc(p.k_s, p.k_gl, p.K) == 0 || outer_lds_bytes(p.k_s, p.k_gl, p.K) > kLdsPerWorkgroup))
        return fail(ctx, BS2E_EINVAL,
                    "the per-interval primitive kernels support spline orders <= 16 whose tables fit 160 KiB of LDS "
                    "(about k_GL^2 * k + k_GL * k^2 doubles; e.g. k = 16 with k_GL <= 26)");
    if (p.n_pairs * p.K >= ((int64_t)1 << 31)) return fail(ctx, BS2E_EINVAL, "tensor row too long (n_pairs*K >= 2^31)");
    if (p.n_iv * p.U * p.U * p.K >= ((int64_t)1 << 31))
        return fail(ctx, BS2E_EINVAL, "diagonal-cell table too large (n_iv*U^2*K >= 2^31)");
    if (near_lds_bytes(p.k_s, p.K) > kLdsPerWorkgroup)
        return fail(ctx, BS2E_EINVAL,
                    "spline order x multipole count too large for the near kernel's LDS "
                    "((2k+2) x (max_k+1) doubles must fit 160 KiB)");
    return BS2E_OK;
}

// Primitives on one stream, in order (for callers that want the containers).
static int64_t run_primitives_serial(bs2e_ctx* ctx, bs2e_rk* rk)
{
    clear_stale_error();
    hipStream_t s = ctx->stream;
    HIP_TRY(ctx, launch_outer_moments(rk->plan, rk->d, s));
    HIP_TRY(ctx, launch_diag_cells(rk->plan, rk->d, s, false));
    HIP_TRY(ctx, launch_pair_windows(rk->plan, rk->d, s));
    return BS2E_OK;
}

// The full build.  Main stream: local moments -> pair windows -> rank-one
// fill.  Side stream (forked once the pair windows exist, so that those two
// short critical-path kernels run alone): diagonal cells -> near elements,
// concurrently with the rank-one fill.  The main stream waits for the side
// stream at the end, so the build is one unit of work on the caller's stream.
// Profiling events, all relative to ev[0]: ev[1] windows done, ev[2] far
// fill done, ev[3] diagonal cells done, ev[4] near elements done, ev[5] end.
static constexpr size_t kMaxProfiledBuilds = 1 << 16;

static int64_t run_build(bs2e_ctx* ctx, bs2e_rk* rk)
{
    clear_stale_error();
    hipStream_t m = ctx->stream, s = ctx->side;
    const bool prof = rk->profiling && rk->evused < kMaxProfiledBuilds;
    hipEvent_t* ev = nullptr;
    if (prof) {
        if (rk->evused == rk->evpool.size()) {
            std::array<hipEvent_t, 6> set{};
            for (auto& e : set) {
                hipError_t err = hipEventCreate(&e);
                if (err != hipSuccess) {
                    for (hipEvent_t x : set)
                        if (x) (void)hipEventDestroy(x);
                    return fail(ctx, BS2E_EDEVICE, std::string("hipEventCreate: ") + hipGetErrorString(err));
                }
            }
            rk->evpool.push_back(set);
        }
        ev = rk->evpool[rk->evused++].data();
    }
    if (prof) HIP_TRY(ctx, hipEventRecord(ev[0], m));
    if (!rk->from_primitives) {
        if (rk->precision == BS2E_PREC_MIXED)
            HIP_TRY(ctx, launch_outer_moments_mixed(rk->plan, rk->d, m));
        else
            HIP_TRY(ctx, launch_outer_moments(rk->plan, rk->d, m));
        HIP_TRY(ctx, launch_pair_windows(rk->plan, rk->d, m));
    }
    if (prof) HIP_TRY(ctx, hipEventRecord(ev[1], m));
    HIP_TRY(ctx, hipEventRecord(rk->windows, m));
    HIP_TRY(ctx, hipStreamWaitEvent(s, rk->windows, 0));
    if (!rk->from_primitives) HIP_TRY(ctx, launch_diag_cells(rk->plan, rk->d, s, true));
    if (prof) HIP_TRY(ctx, hipEventRecord(ev[3], s));
    HIP_TRY(ctx, launch_tensor_near(rk->plan, rk->d, rk->R, s));
    if (prof) HIP_TRY(ctx, hipEventRecord(ev[4], s));
    HIP_TRY(ctx, hipEventRecord(rk->join, s));
    HIP_TRY(ctx, launch_tensor_far(rk->plan, rk->d, rk->R, m));
    if (prof) HIP_TRY(ctx, hipEventRecord(ev[2], m));
    HIP_TRY(ctx, hipStreamWaitEvent(m, rk->join, 0));
    if (prof) HIP_TRY(ctx, hipEventRecord(ev[5], m));
    return BS2E_OK;
}

int64_t bs2e::rk_enqueue_build(bs2e_ctx* ctx, bs2e_rk* rk) { return run_build(ctx, rk); }

extern "C" {

int64_t bs2e_slater_primitives(bs2e_ctx* ctx, int64_t k, int64_t n_knots, const double* knots, int64_t k_gl,
                               int64_t max_k, int64_t nnz4, double* rk, double* rmk, int64_t* iv4, int64_t* i4,
                               int64_t* j4, int64_t nnz6, double* rdk, int64_t* iv6, int64_t* i6, int64_t* ip6,
                               int64_t* j6, int64_t* jp6)
{
    if (!ctx) return fail(nullptr, BS2E_EINVAL, "bs2e_slater_primitives: ctx is NULL");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    bs2e_rk h;
    h.ctx = ctx;
    std::string msg = make_plan(k, n_knots, knots, k_gl, max_k, h.plan);
    if (!msg.empty()) return fail(ctx, BS2E_EINVAL, "bs2e_slater_primitives: " + msg);
    const Plan& p = h.plan;
    if (nnz4 != p.nnz4 || nnz6 != p.nnz6)
        return fail(ctx, BS2E_EINVAL, "bs2e_slater_primitives: nnz4/nnz6 differ from count_nnz_4d/6d for this basis");
    if (!rk || !rmk || !iv4 || !i4 || !j4 || !rdk || !iv6 || !i6 || !ip6 || !j6 || !jp6)
        return fail(ctx, BS2E_EINVAL, "bs2e_slater_primitives: NULL output array");
    int64_t rc = check_plan_limits(ctx, p, true);
    if (rc) return rc;
    build_diag_entries(h.plan);
    struct Guard {
        bs2e_rk* h;
        std::vector<void*> extra;
        ~Guard()
        {
            free_tables(h);
            for (void* e : extra)
                if (e) (void)hipFree(e);
        }
    } guard{&h, {}};
    rc = alloc_tables(ctx, &h, false);
    if (rc) return rc;
    const int64_t K = p.K;
    double *d_rk = nullptr, *d_rmk = nullptr, *d_rdk = nullptr;
    int64_t *d_i4 = nullptr, *d_i6 = nullptr, *d_off4 = nullptr;
    DiagEntry* d_diag = nullptr;
    HIP_TRY(ctx, dev_alloc(&d_rk, (size_t)nnz4 * K));
    guard.extra.push_back(d_rk);
    HIP_TRY(ctx, dev_alloc(&d_rmk, (size_t)nnz4 * K));
    guard.extra.push_back(d_rmk);
    HIP_TRY(ctx, dev_alloc(&d_i4, (size_t)nnz4 * 3));
    guard.extra.push_back(d_i4);
    HIP_TRY(ctx, dev_alloc(&d_off4, (size_t)p.n_pairs));
    guard.extra.push_back(d_off4);
    HIP_TRY(ctx, dev_alloc(&d_rdk, (size_t)nnz6 * K));
    guard.extra.push_back(d_rdk);
    HIP_TRY(ctx, dev_alloc(&d_i6, (size_t)nnz6 * 5));
    guard.extra.push_back(d_i6);
    HIP_TRY(ctx, dev_alloc(&d_diag, p.diag.size()));
    guard.extra.push_back(d_diag);
    hipStream_t s = ctx->stream;
    HIP_TRY(ctx, hipMemcpyAsync(d_off4, p.off4.data(), sizeof(int64_t) * p.n_pairs, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(d_diag, p.diag.data(), sizeof(DiagEntry) * p.diag.size(), hipMemcpyHostToDevice, s));
    rc = run_primitives_serial(ctx, &h);
    if (rc) return rc;
    HIP_TRY(ctx, launch_scatter_4d(p, h.d, d_off4, d_rk, d_rmk, d_i4, d_i4 + nnz4, d_i4 + 2 * nnz4, s));
    HIP_TRY(ctx, launch_scatter_6d(p, h.d, d_diag, (int64_t)p.diag.size(), d_rdk, d_i6, d_i6 + nnz6,
                                   d_i6 + 2 * nnz6, d_i6 + 3 * nnz6, d_i6 + 4 * nnz6, s));
    HIP_TRY(ctx, hipMemcpyAsync(rk, d_rk, sizeof(double) * nnz4 * K, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(rmk, d_rmk, sizeof(double) * nnz4 * K, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(iv4, d_i4, sizeof(int64_t) * nnz4, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(i4, d_i4 + nnz4, sizeof(int64_t) * nnz4, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(j4, d_i4 + 2 * nnz4, sizeof(int64_t) * nnz4, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(rdk, d_rdk, sizeof(double) * nnz6 * K, hipMemcpyDeviceToHost, s));
    int64_t* outs6[5] = {iv6, i6, ip6, j6, jp6};
    for (int a = 0; a < 5; ++a)
        HIP_TRY(ctx, hipMemcpyAsync(outs6[a], d_i6 + a * nnz6, sizeof(int64_t) * nnz6, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    return BS2E_OK;
}

}  // extern "C"

// A tensor handle for an existing plan (whose near rows are built): checks
// the limits, allocates (in the context's arena when it has one), uploads.
int64_t bs2e::rk_from_plan(bs2e_ctx* ctx, Plan&& plan, bs2e_rk** out, int64_t precision)
{
    *out = nullptr;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    bs2e_rk* rk = new (std::nothrow) bs2e_rk();
    if (!rk) return fail(ctx, BS2E_ENOMEM, "bs2e_rk_create: host allocation failed");
    rk_attach(ctx, rk);
    rk->precision = precision;
    rk->plan = std::move(plan);
    int64_t rc = check_plan_limits(ctx, rk->plan, true);
    if (!rc) rc = alloc_tables(ctx, rk, true);
    if (!rc) rc = create_events(ctx, rk);
    if (rc) {
        free_tables(rk);
        rk_delete(rk);
        return rc;
    }
    *out = rk;
    return BS2E_OK;
}

extern "C" {

int64_t bs2e_rk_create(bs2e_ctx* ctx, int64_t k, int64_t n_knots, const double* knots, int64_t k_gl, int64_t max_k,
                       int64_t precision, bs2e_rk** out)
{
    if (!ctx) return fail(nullptr, BS2E_EINVAL, "bs2e_rk_create: ctx is NULL");
    if (!out) return fail(ctx, BS2E_EINVAL, "bs2e_rk_create: out is NULL");
    *out = nullptr;
    if (precision != BS2E_PREC_FP64 && precision != BS2E_PREC_MIXED)
        return fail(ctx, BS2E_EINVAL, "bs2e_rk_create: unknown precision mode");
    Plan plan;
    std::string msg = make_plan(k, n_knots, knots, k_gl, max_k, plan);
    if (!msg.empty()) return fail(ctx, BS2E_EINVAL, "bs2e_rk_create: " + msg);
    if (precision == BS2E_PREC_MIXED && !mixed_supported(plan.k_s, plan.k_gl, plan.K))
        return fail(ctx, BS2E_EINVAL, "bs2e_rk_create: the mixed-precision moments take k_GL <= 16");
    build_near_rows(plan);
    return rk_from_plan(ctx, std::move(plan), out, precision);
}

int64_t bs2e_rk_from_primitives(bs2e_ctx* ctx, int64_t k, int64_t n_knots, const double* knots, int64_t max_k,
                                int64_t nnz4, const double* rk4, const double* rmk, const int64_t* iv4,
                                const int64_t* i4, const int64_t* j4, int64_t nnz6, const double* rdk,
                                const int64_t* iv6, const int64_t* i6, const int64_t* ip6, const int64_t* j6,
                                const int64_t* jp6, bs2e_rk** out)
{
    if (!ctx) return fail(nullptr, BS2E_EINVAL, "bs2e_rk_from_primitives: ctx is NULL");
    if (!out) return fail(ctx, BS2E_EINVAL, "bs2e_rk_from_primitives: out is NULL");
    *out = nullptr;
    if (!rk4 || !rmk || !iv4 || !i4 || !j4 || !rdk || !iv6 || !i6 || !ip6 || !j6 || !jp6)
        return fail(ctx, BS2E_EINVAL, "bs2e_rk_from_primitives: NULL container array");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    bs2e_rk* rk = new (std::nothrow) bs2e_rk();
    if (!rk) return fail(ctx, BS2E_ENOMEM, "bs2e_rk_from_primitives: host allocation failed");
    rk_attach(ctx, rk);
    rk->from_primitives = true;
    // compute_R_K_map has no quadrature argument; the GL order only sizes work
    // tables that this path never fills, so the smallest rule is planned.
    std::string msg = make_plan(k, n_knots, knots, 1, max_k, rk->plan);
    int64_t rc = msg.empty() ? check_plan_limits(ctx, rk->plan, false)
                             : fail(ctx, BS2E_EINVAL, "bs2e_rk_from_primitives: " + msg);
    if (!rc && (nnz4 != rk->plan.nnz4 || nnz6 != rk->plan.nnz6))
        rc = fail(ctx, BS2E_EINVAL, "bs2e_rk_from_primitives: nnz4/nnz6 differ from count_nnz_4d/6d for this basis");
    if (!rc) rc = alloc_tables(ctx, rk, true);
    if (!rc) rc = create_events(ctx, rk);
    if (rc) {
        free_tables(rk);
        rk_delete(rk);
        return rc;
    }
    const Plan& p = rk->plan;
    const int64_t K = p.K;
    std::vector<void*> tmp;
    auto cleanup = [&]() {
        for (void* e : tmp)
            if (e) (void)hipFree(e);
    };
    auto bail = [&](int64_t code) {
        cleanup();
        free_tables(rk);
        rk_delete(rk);
        return code;
    };
#define STEP(expr)                                                                                           \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess)                                                                                \
            return bail(fail(ctx, e_ == hipErrorOutOfMemory ? BS2E_ENOMEM : BS2E_EDEVICE,                    \
                             std::string(#expr) + ": " + hipGetErrorString(e_)));                            \
    } while (0)
    double *d_rk = nullptr, *d_rmk = nullptr, *d_rdk = nullptr;
    int64_t *d_i4 = nullptr, *d_i6 = nullptr;
    int32_t* d_cf = nullptr;
    int* d_bad = nullptr;
    STEP(dev_alloc(&d_rk, (size_t)nnz4 * K));
    tmp.push_back(d_rk);
    STEP(dev_alloc(&d_rmk, (size_t)nnz4 * K));
    tmp.push_back(d_rmk);
    STEP(dev_alloc(&d_i4, (size_t)nnz4 * 3));
    tmp.push_back(d_i4);
    STEP(dev_alloc(&d_rdk, (size_t)nnz6 * K));
    tmp.push_back(d_rdk);
    STEP(dev_alloc(&d_i6, (size_t)nnz6 * 5));
    tmp.push_back(d_i6);
    STEP(dev_alloc(&d_cf, p.c_first.size()));
    tmp.push_back(d_cf);
    STEP(dev_alloc(&d_bad, 1));
    tmp.push_back(d_bad);
    hipStream_t s = ctx->stream;
    STEP(hipMemcpyAsync(d_rk, rk4, sizeof(double) * nnz4 * K, hipMemcpyHostToDevice, s));
    STEP(hipMemcpyAsync(d_rmk, rmk, sizeof(double) * nnz4 * K, hipMemcpyHostToDevice, s));
    STEP(hipMemcpyAsync(d_i4, iv4, sizeof(int64_t) * nnz4, hipMemcpyHostToDevice, s));
    STEP(hipMemcpyAsync(d_i4 + nnz4, i4, sizeof(int64_t) * nnz4, hipMemcpyHostToDevice, s));
    STEP(hipMemcpyAsync(d_i4 + 2 * nnz4, j4, sizeof(int64_t) * nnz4, hipMemcpyHostToDevice, s));
    STEP(hipMemcpyAsync(d_rdk, rdk, sizeof(double) * nnz6 * K, hipMemcpyHostToDevice, s));
    const int64_t* in6[5] = {iv6, i6, ip6, j6, jp6};
    for (int a = 0; a < 5; ++a)
        STEP(hipMemcpyAsync(d_i6 + a * nnz6, in6[a], sizeof(int64_t) * nnz6, hipMemcpyHostToDevice, s));
    STEP(hipMemcpyAsync(d_cf, p.c_first.data(), sizeof(int32_t) * p.c_first.size(), hipMemcpyHostToDevice, s));
    const int zero = 0;
    STEP(hipMemcpyAsync(d_bad, &zero, sizeof(int), hipMemcpyHostToDevice, s));
    const size_t slots = (size_t)p.n_pairs * p.k_s * p.K;
    clear_stale_error();
    STEP(launch_fill(rk->d.Pw, (int64_t)slots, 0.0, s));
    STEP(launch_fill(rk->d.Qw, (int64_t)slots, 0.0, s));
    STEP(launch_fill(rk->d.D, (int64_t)((size_t)p.n_iv * p.K * p.U * p.U), 0.0, s));
    STEP(launch_gather_4d(p, rk->d, d_cf, d_rk, d_rmk, d_i4, d_i4 + nnz4, d_i4 + 2 * nnz4, d_bad, s));
    STEP(launch_gather_6d(p, rk->d, d_rdk, d_i6, d_i6 + nnz6, d_i6 + 2 * nnz6, d_i6 + 3 * nnz6, d_i6 + 4 * nnz6,
                          d_bad, s));
    STEP(launch_window_sums(p, rk->d, s));
    STEP(launch_symmetrize(p, rk->d, s));
    int bad = 0;
    STEP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s));
    STEP(hipStreamSynchronize(s));
#undef STEP
    cleanup();
    if (bad != 0) {
        free_tables(rk);
        rk_delete(rk);
        return fail(ctx, BS2E_ERANGE,
                    "bs2e_rk_from_primitives: " + std::to_string(bad) +
                        " container entries have indices outside the banded, supported set of this basis");
    }
    *out = rk;
    return BS2E_OK;
}

int64_t bs2e_rk_compute(bs2e_rk* rk)
{
    if (!rk) return fail(nullptr, BS2E_EINVAL, "bs2e_rk_compute: rk is NULL");
    bs2e_ctx* ctx = rk->ctx;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    rk->mirror_valid = false;
    if (rk->use_graph) {
        if (!rk->graph_exec) {
            // capture the two-stream build once; its kernels and event
            // dependencies become the graph's nodes (no memset nodes)
            HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            int64_t rc = run_build(ctx, rk);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(ctx->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            HIP_TRY(ctx, e);
            rk->graph = g;
            HIP_TRY(ctx, hipGraphInstantiateWithFlags(&rk->graph_exec, rk->graph, 0));
        }
        HIP_TRY(ctx, hipGraphLaunch(rk->graph_exec, ctx->stream));
    } else {
        int64_t rc = run_build(ctx, rk);
        if (rc) return rc;
    }
    rk->computed = true;
    return BS2E_OK;
}

// Autotuner: per (order, GL order, multipoles, pairs, intervals) the fastest
// rank-one fill shape and diagonal-cell multipole chunk, timed on the
// handle's own tables with HIP events, kept for the process.
namespace {
struct TuneKey {
    int64_t ks, kgl, K, n_pairs, n_iv;
    bool operator<(const TuneKey& o) const
    {
        return std::tie(ks, kgl, K, n_pairs, n_iv) < std::tie(o.ks, o.kgl, o.K, o.n_pairs, o.n_iv);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, std::pair<int, int>> g_tuned;
}  // namespace

static int64_t time_ms(bs2e_ctx* ctx, const std::function<hipError_t()>& launch, float* ms)
{
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(ctx, hipEventCreate(&a));
    HIP_TRY(ctx, hipEventCreate(&b));
    hipError_t e = launch();   // warm-up
    for (int r = 0; r < 8 && e == hipSuccess; ++r) {
        if (r == 0) e = hipEventRecord(a, ctx->stream);
        if (e == hipSuccess) e = launch();
    }
    if (e == hipSuccess) e = hipEventRecord(b, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    HIP_TRY(ctx, e);
    return BS2E_OK;
}

int64_t bs2e_rk_autotune(bs2e_rk* rk, int64_t force)
{
    if (!rk) return fail(nullptr, BS2E_EINVAL, "bs2e_rk_autotune: rk is NULL");
    bs2e_ctx* ctx = rk->ctx;
    if (rk->from_primitives) return fail(ctx, BS2E_ESTATE, "bs2e_rk_autotune: tensors built from containers are not tuned");
    if (rk->use_graph) return fail(ctx, BS2E_ESTATE, "bs2e_rk_autotune: switch the graph off first");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    Plan& p = rk->plan;
    const TuneKey key{p.k_s, p.k_gl, p.K, p.n_pairs, p.n_iv};
    if (!force) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            p.far_variant = it->second.first;
            p.inner_kc = it->second.second;
            return BS2E_OK;
        }
    }
    const bool prof = rk->profiling;
    rk->profiling = false;
    // Each candidate is timed over whole two-stream builds: the fill shares
    // the chip with the diagonal cells and the near ranges, and the fastest
    // shape on its own is not always the fastest beside them.
    int64_t rc = BS2E_OK;
    auto build = [&]() -> hipError_t {
        const int64_t r = run_build(ctx, rk);
        return r ? hipErrorLaunchFailure : hipSuccess;
    };
    float best = 1e30f;
    int best_far = far_variant_default(p);
    for (int v = 0; v < far_variant_count() && !rc; ++v) {
        if (!far_variant_valid(p, v)) continue;
        p.far_variant = v;
        float ms = 0.f;
        rc = time_ms(ctx, build, &ms);
        if (!rc && ms < best) {
            best = ms;
            best_far = v;
        }
    }
    p.far_variant = best_far;
    best = 1e30f;
    int best_kc = 0;
    const int kmax = inner_kc(p.k_s, p.k_gl, p.K);
    std::vector<int> kcs;
    for (int kc : {1, 2, 4, 8, kmax})
        if (kc >= 1 && kc <= kmax && std::find(kcs.begin(), kcs.end(), kc) == kcs.end()) kcs.push_back(kc);
    for (int kc : kcs) {
        if (rc) continue;
        p.inner_kc = kc;
        float ms = 0.f;
        rc = time_ms(ctx, build, &ms);
        if (!rc && ms < best) {
            best = ms;
            best_kc = kc;
        }
    }
    p.inner_kc = best_kc;
    rk->profiling = prof;
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[key] = {p.far_variant, p.inner_kc};
    }
    rk->mirror_valid = false;   // the timing launches rewrote R and D with the same values
    return BS2E_OK;
}

int64_t bs2e_rk_tuning(const bs2e_rk* rk, int64_t* out, int64_t n)
{
    if (!rk || !out) return fail(nullptr, BS2E_EINVAL, "bs2e_rk_tuning: NULL argument");
    const Plan& p = rk->plan;
    const int64_t v[2] = {far_variant_valid(p, p.far_variant) ? p.far_variant : far_variant_default(p),
                          (p.inner_kc >= 1) ? p.inner_kc : inner_kc(p.k_s, p.k_gl, p.K)};
    for (int64_t i = 0; i < n && i < 2; ++i) out[i] = v[i];
    return BS2E_OK;
}

int64_t bs2e_store_ceiling(bs2e_ctx* ctx, void* dst, int64_t n_bytes, int64_t reps, double* gbs)
{
    if (!ctx || !dst || !gbs) return fail(ctx, BS2E_EINVAL, "bs2e_store_ceiling: NULL argument");
    if (n_bytes < 16 || n_bytes % 16 != 0 || (uintptr_t)dst % 16 != 0 || n_bytes / 16 >= ((int64_t)1 << 38))
        return fail(ctx, BS2E_EINVAL, "bs2e_store_ceiling: need a 16-byte aligned buffer of a positive multiple of 16 bytes");
    if (reps < 1 || reps > 1000) return fail(ctx, BS2E_EINVAL, "bs2e_store_ceiling: need 1 <= reps <= 1000");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(ctx, hipEventCreate(&a));
    HIP_TRY(ctx, hipEventCreate(&b));
    double best = 0.0;
    hipError_t e = hipSuccess;
    for (int nt = 0; nt < 2 && e == hipSuccess; ++nt) {
        e = launch_store_fill(dst, n_bytes / 16, nt != 0, st);   // warm-up
        if (e == hipSuccess) e = hipEventRecord(a, st);
        for (int64_t r = 0; r < reps && e == hipSuccess; ++r) e = launch_store_fill(dst, n_bytes / 16, nt != 0, st);
        if (e == hipSuccess) e = hipEventRecord(b, st);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        if (e == hipSuccess && ms > 0.f) best = std::max(best, (double)n_bytes * reps / (ms * 1e-3) / 1e9);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    HIP_TRY(ctx, e);
    *gbs = best;
    return BS2E_OK;
}

int64_t bs2e_rk_set_tuning(bs2e_rk* rk, int64_t shape, int64_t kc)
{
    if (!rk) return fail(nullptr, BS2E_EINVAL, "bs2e_rk_set_tuning: rk is NULL");
    bs2e_ctx* ctx = rk->ctx;
    if (rk->use_graph) return fail(ctx, BS2E_ESTATE, "bs2e_rk_set_tuning: switch the graph off first");
    Plan& p = rk->plan;
    if (shape < 0 || shape >= far_variant_count() || !far_variant_valid(p, (int)shape))
        return fail(ctx, BS2E_EINVAL, "bs2e_rk_set_tuning: no such fill shape for this basis");
    const int kmax = inner_kc(p.k_s, p.k_gl, p.K);
    if (kc < 1 || kc > std::max(kmax, 1))
        return fail(ctx, BS2E_EINVAL, "bs2e_rk_set_tuning: multipole chunk outside 1..the largest that fits");
    p.far_variant = (int)shape;
    p.inner_kc = (int)kc;
    return BS2E_OK;
}

int64_t bs2e_rk_set_graph(bs2e_rk* rk, int64_t on)
{
    if (!rk) return fail(nullptr, BS2E_EINVAL, "bs2e_rk_set_graph: rk is NULL");
    if (on && rk->profiling)
        return fail(rk->ctx, BS2E_ESTATE, "bs2e_rk_set_graph: switch profiling off first (events are not captured)");
    if (!on || rk->use_graph != (on != 0)) {
        if (rk->graph_exec) (void)hipGraphExecDestroy(rk->graph_exec);
        if (rk->graph) (void)hipGraphDestroy(rk->graph);
        rk->graph_exec = nullptr;
        rk->graph = nullptr;
    }
    rk->use_graph = on != 0;
    return BS2E_OK;
}

int64_t bs2e_rk_synchronize(bs2e_rk* rk)
{
    if (!rk) return fail(nullptr, BS2E_EINVAL, "bs2e_rk_synchronize: rk is NULL");
    HIP_TRY(rk->ctx, hipStreamSynchronize(rk->ctx->stream));
    return BS2E_OK;
}

int64_t bs2e_rk_info(const bs2e_rk* rk, int64_t* info, int64_t n)
{
    if (!rk || !info) return fail(nullptr, BS2E_EINVAL, "bs2e_rk_info: NULL argument");
    const Plan& p = rk->plan;
    if (p.near_elems < 0) const_cast<Plan&>(p).near_elems = count_near_elements(p);   // counted once, on demand
    const int64_t v[11] = {p.n_b, p.n_pairs, p.K, p.nnz4, p.nnz6, p.n_iv, p.k_s, p.k_gl, p.U, p.n_knots, p.near_elems};
    for (int64_t i = 0; i < n && i < 11; ++i) info[i] = v[i];
    return BS2E_OK;
}

int64_t bs2e_rk_device_ptr(const bs2e_rk* rk, void** ptr, int64_t* n_elems)
{
    if (!rk) return fail(nullptr, BS2E_EINVAL, "bs2e_rk_device_ptr: rk is NULL");
    if (ptr) *ptr = rk->R;
    if (n_elems) *n_elems = (int64_t)rk->R_elems;
    return BS2E_OK;
}

int64_t bs2e_rk_pair_index(const bs2e_rk* rk, int64_t a, int64_t c)
{
    if (!rk) return fail(nullptr, BS2E_EINVAL, "bs2e_rk_pair_index: rk is NULL");
    const Plan& p = rk->plan;
    if (a < 1 || c < 1 || a > p.n_b || c > p.n_b || std::llabs(a - c) >= p.k_s) return BS2E_ERANGE;
    return p.c_first[c] + (a - std::max<int64_t>(1, c - p.k_s + 1));
}

int64_t bs2e_rk_copy_to_host(bs2e_rk* rk, double* dst, int64_t n_elems)
{
    if (!rk) return fail(nullptr, BS2E_EINVAL, "bs2e_rk_copy_to_host: rk is NULL");
    bs2e_ctx* ctx = rk->ctx;
    if (!rk->computed) return fail(ctx, BS2E_ESTATE, "bs2e_rk_copy_to_host: tensor not computed");
    if (dst && n_elems != (int64_t)rk->R_elems)
        return fail(ctx, BS2E_EINVAL, "bs2e_rk_copy_to_host: n_elems differs from the tensor size");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    try {
        rk->mirror.resize(rk->R_elems);
    } catch (...) {
        return fail(ctx, BS2E_ENOMEM, "bs2e_rk_copy_to_host: host mirror allocation failed");
    }
    HIP_TRY(ctx, hipMemcpy(rk->mirror.data(), rk->R, sizeof(double) * rk->R_elems, hipMemcpyDeviceToHost));
    rk->mirror_valid = true;
    if (dst) std::memcpy(dst, rk->mirror.data(), sizeof(double) * rk->R_elems);
    return BS2E_OK;
}

int64_t bs2e_rk_copy_rows(const bs2e_rk* rk, int64_t row_lo, int64_t n_rows, double* dst)
{
    if (!rk || !dst) return fail(nullptr, BS2E_EINVAL, "bs2e_rk_copy_rows: NULL argument");
    bs2e_ctx* ctx = rk->ctx;
    const Plan& p = rk->plan;
    if (row_lo < 0 || n_rows < 0 || row_lo + n_rows > p.n_pairs)
        return fail(ctx, BS2E_ERANGE, "bs2e_rk_copy_rows: rows outside 0..n_pairs-1");
    if (!rk->computed) return fail(ctx, BS2E_ESTATE, "bs2e_rk_copy_rows: tensor not computed");
    const size_t row = (size_t)p.n_pairs * p.K;
    if (rk->mirror_valid) {
        std::memcpy(dst, rk->mirror.data() + (size_t)row_lo * row, sizeof(double) * row * (size_t)n_rows);
        return BS2E_OK;
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipMemcpy(dst, rk->R + (size_t)row_lo * row, sizeof(double) * row * (size_t)n_rows,
                           hipMemcpyDeviceToHost));
    return BS2E_OK;
}

int64_t bs2e_rk_get_val(const bs2e_rk* rk, const int64_t* idx, double* values, int64_t n_values)
{
    if (!rk || !idx || !values) return fail(nullptr, BS2E_EINVAL, "bs2e_rk_get_val: NULL argument");
    const Plan& p = rk->plan;
    if (n_values < p.K) return fail(rk->ctx, BS2E_EINVAL, "bs2e_rk_get_val: values holds fewer than max_k+1 entries");
    const int64_t pa = bs2e_rk_pair_index(rk, idx[0], idx[2]);
    const int64_t pb = bs2e_rk_pair_index(rk, idx[1], idx[3]);
    if (pa < 0 || pb < 0) {
        char buf[160];
        std::snprintf(buf, sizeof buf, "bs2e_rk_get_val: key [%lld,%lld,%lld,%lld] is outside the banded key set",
                      (long long)idx[0], (long long)idx[1], (long long)idx[2], (long long)idx[3]);
        return fail(rk->ctx, BS2E_ERANGE, buf);
    }
    const size_t off = ((size_t)pa * p.n_pairs + pb) * p.K;
    if (rk->mirror_valid) {
        std::memcpy(values, rk->mirror.data() + off, sizeof(double) * p.K);
        return BS2E_OK;
    }
    if (!rk->computed) return fail(rk->ctx, BS2E_ESTATE, "bs2e_rk_get_val: tensor not computed");
    hipError_t e = hipSetDevice(rk->ctx->device);
    if (e == hipSuccess) e = hipStreamSynchronize(rk->ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(values, rk->R + off, sizeof(double) * p.K, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(rk->ctx, BS2E_EDEVICE, std::string("bs2e_rk_get_val: ") + hipGetErrorString(e));
    return BS2E_OK;
}

int64_t bs2e_rk_set_profiling(bs2e_rk* rk, int64_t on)
{
    if (!rk) return fail(nullptr, BS2E_EINVAL, "bs2e_rk_set_profiling: rk is NULL");
    if (on && rk->use_graph)
        return fail(rk->ctx, BS2E_ESTATE, "bs2e_rk_set_profiling: the build is a captured graph; switch it off first");
    rk->profiling = on != 0;
    rk->evused = 0;
    return BS2E_OK;
}

int64_t bs2e_rk_stage_times(bs2e_rk* rk, double* ms, int64_t n)
{
    if (!rk || !ms) return fail(nullptr, BS2E_EINVAL, "bs2e_rk_stage_times: NULL argument");
    if (rk->evused == 0) return fail(rk->ctx, BS2E_ESTATE, "bs2e_rk_stage_times: no build was profiled");
    const size_t nb = rk->evused;
    HIP_TRY(rk->ctx, hipEventSynchronize(rk->evpool[nb - 1][5]));
    double sum[5] = {0, 0, 0, 0, 0};
    for (size_t b = 0; b < nb; ++b)
        for (int i = 0; i < 5; ++i) {
            float t = 0.f;
            HIP_TRY(rk->ctx, hipEventElapsedTime(&t, rk->evpool[b][0], rk->evpool[b][i + 1]));
            sum[i] += t;
        }
    for (int64_t i = 0; i < n && i < 5; ++i) ms[i] = sum[i] / (double)nb;
    rk->evused = 0;
    return BS2E_OK;
}

int64_t bs2e_rk_release(bs2e_rk* rk)
{
    if (!rk) return BS2E_OK;
    if (rk->blocks > 0) {     // blocks still read the tensor: freed with the last of them
        rk->closing = true;
        return BS2E_OK;
    }
    return rk_release_now(rk);
}

}  // extern "C"

void bs2e::rk_attach(bs2e_ctx* ctx, bs2e_rk* rk)
{
    rk->ctx = ctx;
    ctx->refs += 1;
}

void bs2e::ctx_hold(bs2e_ctx* ctx) { ctx->refs += 1; }

void bs2e::ctx_drop(bs2e_ctx* ctx)
{
    if (ctx && --ctx->refs == 0 && ctx->closing) ctx_free_now(ctx);
}

void bs2e::rk_delete(bs2e_rk* rk)
{
    bs2e_ctx* ctx = rk->ctx;
    delete rk;
    if (ctx && --ctx->refs == 0 && ctx->closing) ctx_free_now(ctx);
}

int64_t bs2e::rk_release_now(bs2e_rk* rk)
{
    (void)hipSetDevice(rk->ctx->device);
    // A build service's contexts (arena set) carry other jobs' work; the
    // service releases a tensor only after its build has completed.
    if (!rk->ctx->arena) {
        (void)hipStreamSynchronize(rk->ctx->stream);
        (void)hipStreamSynchronize(rk->ctx->side);
    }
    free_tables(rk);
    rk_delete(rk);
    return BS2E_OK;
}
