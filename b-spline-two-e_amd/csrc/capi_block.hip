// capi_block.hip -- C ABI of the symmetry blocks: configuration enumeration
// and the angular-channel summation of the device R^k tensor into CSR H/S.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "capi_internal.h"

using namespace bs2e;

namespace bs2e {

// orbital_tools.f90:46-72 consistent: no equivalent configuration for odd L,
// triangle |l1-l2| <= L <= l1+l2, parity of l1+l2 equal to the term's.
static bool consistent(int64_t l1, int64_t l2, int64_t L, bool odd, bool eqv)
{
    if (eqv && (L & 1)) return false;
    if (!(std::llabs(l1 - l2) <= L && L <= l1 + l2)) return false;
    return (((l1 + l2) & 1) != 0) == odd;
}

// orbital_tools.f90:119-216 count_configs: l1 = 0..max_l_1p, l2 = 0..l1,
// n1 from min(l1+1, k-1) to n_b (skipping n1 > n_all_l when l2 > l_2_max),
// n2 from min(l2+1, k-1) to min(n1, max_n_b) (l1 = l2) or min(n_b, max_n_b).
int64_t count_configs(int64_t L, int64_t term_pi, int64_t max_l_1p, int64_t n_b, int64_t k_spline, int64_t max_n_b,
                      int64_t n_all_l, int64_t l_2_max, int64_t* out, int64_t cap)
{
    int64_t n = 0;
    for (int64_t l1 = 0; l1 <= max_l_1p; ++l1)
        for (int64_t l2 = 0; l2 <= l1; ++l2)
            for (int64_t n1 = std::min(l1 + 1, k_spline - 1); n1 <= n_b; ++n1) {
                if (n1 > n_all_l && l2 > l_2_max) continue;
                const int64_t n2hi = (l1 == l2) ? std::min(n1, max_n_b) : std::min(n_b, max_n_b);
                for (int64_t n2 = std::min(l2 + 1, k_spline - 1); n2 <= n2hi; ++n2) {
                    const bool eqv = (l1 == l2) && (n1 == n2);
                    if (!consistent(l1, l2, L, term_pi != 0, eqv)) continue;
                    if (out && n < cap) {
                        int64_t* o = out + 5 * n;
                        o[0] = n1;
                        o[1] = n2;
                        o[2] = l1;
                        o[3] = l2;
                        o[4] = eqv ? 1 : 0;
                    }
                    ++n;
                }
            }
    return n;
}

}  // namespace bs2e

// Channel runs of a configuration list in count_configs order (each (l1,l2)
// contiguous, n1 ascending within it, n2 consecutive for each n1): per
// channel its first configuration and count; per (channel, n1) the offset of
// its first configuration within the channel and its n2 range; per
// configuration its channel.  Returns an error message or "".
std::string bs2e::channel_tables(const int64_t* configs, int64_t n_config, int64_t nb, std::vector<BlockChan>& chan,
                                 std::vector<BlockSeg>& seg, std::vector<int32_t>& cfg_chan, int64_t& max_l)
{
    chan.clear();
    seg.clear();
    cfg_chan.assign((size_t)n_config, 0);
    max_l = 0;
    int64_t prev_n1 = -1, prev_n2 = -1;
    for (int64_t i = 0; i < n_config; ++i) {
        const int64_t* c = configs + 5 * i;
        const int64_t n1 = c[0], n2 = c[1], l1 = c[2], l2 = c[3];
        if (n1 < 1 || n1 > nb || n2 < 1 || n2 > nb || l1 < 0 || l2 < 0 || l1 > 63 || l2 > 63)
            return "configuration " + std::to_string(i + 1) + " has indices outside the basis";
        max_l = std::max({max_l, l1, l2});
        if (chan.empty() || chan.back().l1 != l1 || chan.back().l2 != l2) {
            for (const BlockChan& ch : chan)
                if (ch.l1 == l1 && ch.l2 == l2) return "channel (l1,l2) appears in two separate runs";
            chan.push_back(BlockChan{(int32_t)l1, (int32_t)l2, (int32_t)i, 0});
            seg.resize(chan.size() * (size_t)(nb + 1), BlockSeg{0, 1, 0, 0});
            prev_n1 = -1;
            prev_n2 = -1;
        }
        BlockChan& ch = chan.back();
        BlockSeg& sg = seg[(chan.size() - 1) * (size_t)(nb + 1) + n1];
        if (n1 != prev_n1) {
            if (n1 < prev_n1 || sg.hi >= sg.lo) return "configurations are not ordered by n1 within a channel";
            sg.off = (int32_t)(i - ch.start);
            sg.lo = (int32_t)n2;
            sg.hi = (int32_t)n2;
        } else {
            if (n2 != prev_n2 + 1) return "n2 values of a channel's n1 are not consecutive";
            sg.hi = (int32_t)n2;
        }
        ch.n += 1;
        cfg_chan[i] = (int32_t)(chan.size() - 1);
        prev_n1 = n1;
        prev_n2 = n2;
    }
    return std::string();
}

struct bs2e_block {
    bs2e_rk* rk = nullptr;
    int64_t L = 0, n_config = 0, full = 1, max_l = 0;
    float ms_count = 0.f, ms_fill = 0.f;   // device time of the count and fill kernels (HIP events)
    std::vector<BlockChan> chan;
    std::vector<BlockSeg> seg;
    std::vector<int32_t> cfg_chan, flags;
    std::vector<double> coef;
    std::vector<int64_t> cntH, cntS, ptrH, ptrS;
    int64_t nnzH = 0, nnzS = 0, refH = 0, refS = 0;
    // device
    int64_t* d_cfg = nullptr;
    int32_t* d_cfg_chan = nullptr;
    BlockChan* d_chan = nullptr;
    BlockSeg* d_seg = nullptr;
    double* d_coef = nullptr;
    int32_t* d_flags = nullptr;
    int64_t* d_cnt = nullptr;   // 4 x n_config
    int64_t* d_ptr = nullptr;   // 2 x n_config
};

static void block_free(bs2e_block* b)
{
    dev_free(b->d_cfg);
    dev_free(b->d_cfg_chan);
    dev_free(b->d_chan);
    dev_free(b->d_seg);
    dev_free(b->d_coef);
    dev_free(b->d_flags);
    dev_free(b->d_cnt);
    dev_free(b->d_ptr);
}

static BlockArgs block_args(const bs2e_block* b)
{
    const Plan& p = b->rk->plan;
    BlockArgs a{};
    a.n_config = (int)b->n_config;
    a.n_chan = (int)b->chan.size();
    a.full = (int)b->full;
    a.L = (int)b->L;
    a.ks = (int)p.k_s;
    a.n_b = (int)p.n_b;
    a.n_pairs = (int)p.n_pairs;
    a.K = (int)p.K;
    a.cfg = b->d_cfg;
    a.cfg_chan = b->d_cfg_chan;
    a.chan = b->d_chan;
    a.seg = b->d_seg;
    a.coef = b->d_coef;
    a.flags = b->d_flags;
    a.c_first = b->rk->d.c_first;
    a.R = b->rk->R;
    return a;
}

extern "C" {

int64_t bs2e_count_configs(int64_t L, int64_t term_pi, int64_t max_l_1p, int64_t n_b, int64_t k_spline,
                           int64_t max_n_b, int64_t n_all_l, int64_t l_2_max, int64_t* configs, int64_t cap)
{
    if (L < 0 || max_l_1p < 0 || n_b < 1 || k_spline < 2 || (configs && cap < 0))
        return fail(nullptr, BS2E_EINVAL, "bs2e_count_configs: invalid argument");
    const int64_t n = count_configs(L, term_pi, max_l_1p, n_b, k_spline, max_n_b, n_all_l, l_2_max, configs, cap);
    if (configs && n > cap) return fail(nullptr, BS2E_EINVAL, "bs2e_count_configs: cap smaller than the count");
    return n;
}

int64_t bs2e_block_create(bs2e_rk* rk, int64_t L, int64_t n_config, const int64_t* configs, int64_t full,
                          bs2e_block** out)
{
    if (!rk || !out || (n_config > 0 && !configs))
        return fail(rk ? rk->ctx : nullptr, BS2E_EINVAL, "bs2e_block_create: NULL argument");
    bs2e_ctx* ctx = rk->ctx;
    *out = nullptr;
    if (!rk->computed) return fail(ctx, BS2E_ESTATE, "bs2e_block_create: the R^k tensor has not been computed");
    if (L < 0 || n_config < 1 || n_config >= ((int64_t)1 << 31))
        return fail(ctx, BS2E_EINVAL, "bs2e_block_create: need L >= 0 and 1 <= n_config < 2^31");
    const Plan& p = rk->plan;
    bs2e_block* b = new (std::nothrow) bs2e_block();
    if (!b) return fail(ctx, BS2E_ENOMEM, "bs2e_block_create: host allocation failed");
    b->rk = rk;
    b->L = L;
    b->n_config = n_config;
    b->full = full ? 1 : 0;
    // channel tables; the configurations must be in count_configs order
    const int64_t nb = p.n_b;
    std::string bad = channel_tables(configs, n_config, nb, b->chan, b->seg, b->cfg_chan, b->max_l);
    if (!bad.empty()) {
        delete b;
        return fail(ctx, BS2E_EINVAL, "bs2e_block_create: " + bad + " (count_configs order required)");
    }
    // angular coefficients per (row channel, column channel): wigner_tools.f90:126-138
    const int64_t nch = (int64_t)b->chan.size(), K = p.K;
    b->coef.assign((size_t)(nch * nch * 2 * K), 0.0);
    b->flags.assign((size_t)(nch * nch), 0);
    for (int64_t ci = 0; ci < nch; ++ci)
        for (int64_t cj = 0; cj < nch; ++cj) {
            const int64_t la = b->chan[ci].l1, lb = b->chan[ci].l2, lc = b->chan[cj].l1, ld = b->chan[cj].l2;
            const double sgn = ((lc + ld + L) & 1) ? -1.0 : 1.0;
            double* cd = b->coef.data() + (ci * nch + cj) * 2 * K;
            double* ce = cd + K;
            int32_t fl = 0;
            for (int64_t k = 0; k < K; ++k) {
                const double ang = ang_k_LS(k, la, lb, lc, ld, L);
                const double ang_ex = ang_k_LS(k, la, lb, ld, lc, L);
                if (std::fabs(ang) >= 5e-16) cd[k] = ang;              // r_12_tens, mat_els.f90:566-570
                if (std::fabs(ang_ex) >= 5e-16) ce[k] = sgn * ang_ex;
                if (std::fabs(ang) > 5e-15) fl |= 1;                    // hamiltonian.f90:171-178
                if (std::fabs(ang_ex) > 5e-15) fl |= 2;
            }
            b->flags[ci * nch + cj] = fl;
        }
    int64_t rc = BS2E_OK;
    auto run = [&]() -> int64_t {
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        HIP_TRY(ctx, dev_alloc(&b->d_cfg, (size_t)5 * n_config));
        HIP_TRY(ctx, dev_alloc(&b->d_cfg_chan, (size_t)n_config));
        HIP_TRY(ctx, dev_alloc(&b->d_chan, b->chan.size()));
        HIP_TRY(ctx, dev_alloc(&b->d_seg, b->seg.size()));
        HIP_TRY(ctx, dev_alloc(&b->d_coef, b->coef.size()));
        HIP_TRY(ctx, dev_alloc(&b->d_flags, b->flags.size()));
        HIP_TRY(ctx, dev_alloc(&b->d_cnt, (size_t)4 * n_config));
        HIP_TRY(ctx, dev_alloc(&b->d_ptr, (size_t)2 * n_config));
        hipStream_t s = ctx->stream;
        HIP_TRY(ctx, hipMemcpyAsync(b->d_cfg, configs, sizeof(int64_t) * 5 * n_config, hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(b->d_cfg_chan, b->cfg_chan.data(), sizeof(int32_t) * n_config,
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(b->d_chan, b->chan.data(), sizeof(BlockChan) * b->chan.size(),
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(b->d_seg, b->seg.data(), sizeof(BlockSeg) * b->seg.size(), hipMemcpyHostToDevice,
                                    s));
        HIP_TRY(ctx, hipMemcpyAsync(b->d_coef, b->coef.data(), sizeof(double) * b->coef.size(),
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(b->d_flags, b->flags.data(), sizeof(int32_t) * b->flags.size(),
                                    hipMemcpyHostToDevice, s));
        clear_stale_error();
        BlockArgs a = block_args(b);
        a.cntH = b->d_cnt;
        a.cntS = b->d_cnt + n_config;
        a.refH = b->d_cnt + 2 * n_config;
        a.refS = b->d_cnt + 3 * n_config;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIP_TRY(ctx, hipEventCreate(&e0));
        HIP_TRY(ctx, hipEventCreate(&e1));
        hipError_t le = hipEventRecord(e0, s);
        if (le == hipSuccess) le = launch_block_count(a, s);
        if (le == hipSuccess) le = hipEventRecord(e1, s);
        std::vector<int64_t> cnt((size_t)4 * n_config);
        if (le == hipSuccess)
            le = hipMemcpyAsync(cnt.data(), b->d_cnt, sizeof(int64_t) * 4 * n_config, hipMemcpyDeviceToHost, s);
        if (le == hipSuccess) le = hipStreamSynchronize(s);
        if (le == hipSuccess) le = hipEventElapsedTime(&b->ms_count, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIP_TRY(ctx, le);
        b->cntH.assign(cnt.begin(), cnt.begin() + n_config);
        b->cntS.assign(cnt.begin() + n_config, cnt.begin() + 2 * n_config);
        b->ptrH.resize(n_config);
        b->ptrS.resize(n_config);
        for (int64_t i = 0; i < n_config; ++i) {
            b->ptrH[i] = b->nnzH;
            b->ptrS[i] = b->nnzS;
            b->nnzH += b->cntH[i];
            b->nnzS += b->cntS[i];
            b->refH += cnt[2 * n_config + i];
            b->refS += cnt[3 * n_config + i];
        }
        HIP_TRY(ctx, hipMemcpyAsync(b->d_ptr, b->ptrH.data(), sizeof(int64_t) * n_config, hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(b->d_ptr + n_config, b->ptrS.data(), sizeof(int64_t) * n_config,
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        return BS2E_OK;
    };
    rc = run();
    if (rc) {
        block_free(b);
        delete b;
        return rc;
    }
    rk->blocks += 1;
    *out = b;
    return BS2E_OK;
}

int64_t bs2e_block_times(const bs2e_block* b, double* ms, int64_t n)
{
    if (!b || !ms) return fail(nullptr, BS2E_EINVAL, "bs2e_block_times: NULL argument");
    const double v[2] = {b->ms_count, b->ms_fill};
    for (int64_t i = 0; i < n && i < 2; ++i) ms[i] = v[i];
    return BS2E_OK;
}

int64_t bs2e_block_sizes(const bs2e_block* b, int64_t* sizes, int64_t n)
{
    if (!b || !sizes) return fail(nullptr, BS2E_EINVAL, "bs2e_block_sizes: NULL argument");
    const int64_t v[5] = {b->nnzH, b->nnzS, b->refH, b->refS, b->n_config};
    for (int64_t i = 0; i < n && i < 5; ++i) sizes[i] = v[i];
    return BS2E_OK;
}

int64_t bs2e_block_fill(bs2e_block* b, int64_t max_l_1p, const double* Hl, const double* S, int64_t* ip_H,
                        int64_t* ix_H, double* d_H, int64_t* ip_S, int64_t* ix_S, double* d_S)
{
    if (!b || !Hl || !S || !ip_H || !ix_H || !d_H || !ip_S || !ix_S || !d_S)
        return fail(b ? b->rk->ctx : nullptr, BS2E_EINVAL, "bs2e_block_fill: NULL argument");
    bs2e_ctx* ctx = b->rk->ctx;
    if (max_l_1p < b->max_l)
        return fail(ctx, BS2E_EINVAL, "bs2e_block_fill: one-electron matrices do not cover every l of the block");
    const Plan& p = b->rk->plan;
    const size_t nn = (size_t)p.n_b * p.n_b;
    double *dHl = nullptr, *dS = nullptr, *ddH = nullptr, *ddS = nullptr;
    int64_t *dixH = nullptr, *dixS = nullptr;
    auto cleanup = [&]() {
        dev_free(dHl);
        dev_free(dS);
        dev_free(ddH);
        dev_free(ddS);
        dev_free(dixH);
        dev_free(dixS);
    };
    auto run = [&]() -> int64_t {
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        HIP_TRY(ctx, dev_alloc(&dHl, 2 * nn * (size_t)(max_l_1p + 1)));
        HIP_TRY(ctx, dev_alloc(&dS, 2 * nn));
        HIP_TRY(ctx, dev_alloc(&ddH, 2 * (size_t)b->nnzH));
        HIP_TRY(ctx, dev_alloc(&ddS, 2 * (size_t)b->nnzS));
        HIP_TRY(ctx, dev_alloc(&dixH, (size_t)b->nnzH));
        HIP_TRY(ctx, dev_alloc(&dixS, (size_t)b->nnzS));
        hipStream_t s = ctx->stream;
        HIP_TRY(ctx, hipMemcpyAsync(dHl, Hl, sizeof(double) * 2 * nn * (max_l_1p + 1), hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(dS, S, sizeof(double) * 2 * nn, hipMemcpyHostToDevice, s));
        clear_stale_error();
        BlockArgs a = block_args(b);
        a.Hl = dHl;
        a.S = dS;
        a.ptrH = b->d_ptr;
        a.ptrS = b->d_ptr + b->n_config;
        a.ixH = dixH;
        a.dH = ddH;
        a.ixS = dixS;
        a.dS = ddS;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIP_TRY(ctx, hipEventCreate(&e0));
        HIP_TRY(ctx, hipEventCreate(&e1));
        hipError_t le = hipEventRecord(e0, s);
        if (le == hipSuccess) le = launch_block_fill(a, s);
        if (le == hipSuccess) le = hipEventRecord(e1, s);
        if (le == hipSuccess) le = hipEventSynchronize(e1);
        if (le == hipSuccess) le = hipEventElapsedTime(&b->ms_fill, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIP_TRY(ctx, le);
        HIP_TRY(ctx, hipMemcpyAsync(ix_H, dixH, sizeof(int64_t) * b->nnzH, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(d_H, ddH, sizeof(double) * 2 * b->nnzH, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(ix_S, dixS, sizeof(int64_t) * b->nnzS, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(d_S, ddS, sizeof(double) * 2 * b->nnzS, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        return BS2E_OK;
    };
    const int64_t rc = run();
    cleanup();
    if (rc) return rc;
    // CSR row pointers, 1-based as the reference's (hamiltonian.f90:142-143, 203-204)
    ip_H[0] = 1;
    ip_S[0] = 1;
    for (int64_t i = 0; i < b->n_config; ++i) {
        ip_H[i + 1] = ip_H[i] + b->cntH[i];
        ip_S[i + 1] = ip_S[i] + b->cntS[i];
    }
    return BS2E_OK;
}

int64_t bs2e_block_release(bs2e_block* b)
{
    if (!b) return BS2E_OK;
    bs2e_rk* rk = b->rk;
    (void)hipSetDevice(rk->ctx->device);
    block_free(b);
    delete b;
    if (--rk->blocks == 0 && rk->closing) return rk_release_now(rk);   // the tensor's release was deferred
    return BS2E_OK;
}

}  // extern "C"
