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
    if (nb >= (1 << 15)) return "the block fill takes fewer than 2^15 B-splines";   // n2 ranges packed in 16 bits
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

// The blocks of one or more symmetries against one computed tensor.  Their
// configuration rows are laid end to end (symmetry s owns global rows
// [row0[s], row0[s+1])); one count launch and one fill launch cover them
// all, each wavefront locating its block in the descriptor table.  The fill
// keeps the CSR arrays of every symmetry on the device, concatenated in
// symmetry order; bs2e_blockset_get copies one symmetry's arrays out.  A
// window fill (bs2e_blockset_fill_rows) holds only the entries of a range of
// global rows, so that a set larger than the device's memory streams through
// it window by window.
struct bs2e_blockset {
    bs2e_rk* rk = nullptr;
    int64_t n_sym = 0, full = 1, max_l = 0, n_rows = 0, n_l = 0;   // n_l: one-electron matrices of the last fill
    float ms_count = 0.f, ms_fill = 0.f;   // device time of the count and fill kernels (HIP events)
    std::vector<int64_t> L, row0;
    std::vector<int64_t> chan_off, seg_off, coef_off, flag_off, n_chan;
    std::vector<BlockChan> chan;
    std::vector<BlockSeg> seg;
    std::vector<int32_t> flags;
    std::vector<double> coef;
    std::vector<int64_t> nnzH, nnzS, refH, refS;   // per symmetry
    std::vector<int64_t> offH, offS;               // first entry of each symmetry in the concatenated output
    std::vector<int64_t> prefH, prefS;             // n_rows + 1: first entry of each global row in the whole set
                                                   // (row g holds prefH[g + 1] - prefH[g] H entries)
    int64_t totH = 0, totS = 0;
    // the rows the outputs hold: global rows [win_lo, win_lo + win_n), whose
    // entries are [winH, winH + nwinH) and [winS, winS + nwinS) of the whole
    // set; capH, capS: entries the output buffers have room for
    int64_t win_lo = 0, win_n = 0, winH = 0, winS = 0, nwinH = 0, nwinS = 0, capH = 0, capS = 0;
    std::vector<BlockArgs> tab;                    // host copy of the descriptor table
    bool filled = false;
    // device
    int32_t* d_cfg = nullptr;   // 2 per configuration: n1 | n2 << 16, l1 | l2 << 16
    int32_t* d_cfg_chan = nullptr;
    BlockChan* d_chan = nullptr;
    BlockSeg* d_seg = nullptr;
    double* d_coef = nullptr;
    int32_t* d_flags = nullptr;
    int64_t* d_cnt = nullptr;   // 4 x n_rows: H, S, reference H, reference S
    int64_t* d_ptr = nullptr;   // 2 x n_rows: H, S (0-based offsets into the concatenated output)
    BlockArgs* d_tab = nullptr;
    int64_t* d_row0 = nullptr;
    double *d_Hl = nullptr, *d_S = nullptr;   // the caller's column-major one-electron matrices
    double* d_T = nullptr;                     // the fill's layouts: Hl, S row-major; Ptot, Qtot [K][n_pairs]
    int64_t *d_ixH = nullptr, *d_ixS = nullptr;
    double *d_dH = nullptr, *d_dS = nullptr;
    // the whole branch as one captured graph (bs2e_blockset_set_graph)
    BlockArgs* d_tabg = nullptr;   // count and fill descriptors in one table
    int64_t* d_scan = nullptr;     // the row-offset scan's tile totals
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // the fill's rows in three ascending parts: the staged rows (the
    // matrix-core fill) whose windows cannot overlap (|n1 - n2| >= 2k - 1)
    // up to n_far, the other staged rows (k <= |n1 - n2| < 2k - 1) up to
    // n_sep, then the rest
    std::vector<int32_t> rowlist;
    int64_t n_far = 0, n_sep = 0;
    int32_t* d_rowlist = nullptr;
    // a count pass enqueued without waiting (the build service's admission):
    // its prefix sums' scratch, timing events and the event behind the
    // counts' copy to the context's pinned staging; set_count_finish reads them
    int64_t* d_count_tmp = nullptr;
    hipEvent_t count_ev[3] = {nullptr, nullptr, nullptr};
    bool count_pending = false;
};

// A single symmetry's block is a set of one.
struct bs2e_block : bs2e_blockset {};

static void set_drop_graph(bs2e_blockset* b)
{
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    ctx_free(b->rk->ctx, b->d_tabg);
    ctx_free(b->rk->ctx, b->d_scan);
}

// A set's device memory comes from its tensor's context: the context's arena
// when it has one (a build service's slot), else hipMalloc.
static void set_drop_outputs(bs2e_blockset* b)
{
    set_drop_graph(b);   // its nodes address these buffers
    bs2e_ctx* c = b->rk->ctx;
    ctx_free(c, b->d_Hl);
    ctx_free(c, b->d_S);
    ctx_free(c, b->d_T);
    ctx_free(c, b->d_ixH);
    ctx_free(c, b->d_dH);
    ctx_free(c, b->d_ixS);
    ctx_free(c, b->d_dS);
    b->capH = b->capS = 0;
    b->filled = false;
}

static void set_free(bs2e_blockset* b)
{
    bs2e_ctx* c = b->rk->ctx;
    set_drop_outputs(b);
    ctx_free(c, b->d_cfg);
    ctx_free(c, b->d_cfg_chan);
    ctx_free(c, b->d_chan);
    ctx_free(c, b->d_seg);
    ctx_free(c, b->d_coef);
    ctx_free(c, b->d_flags);
    ctx_free(c, b->d_cnt);
    ctx_free(c, b->d_ptr);
    ctx_free(c, b->d_tab);
    ctx_free(c, b->d_row0);
    ctx_free(c, b->d_rowlist);
    if (b->count_pending) (void)hipStreamSynchronize(c->stream);   // the pass still uses its scratch
    ctx_free(c, b->d_count_tmp);
    for (hipEvent_t& e : b->count_ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
}

// The fill launches over the rows of the window: the window's rows of each
// part of the row list are one contiguous run of it (the parts ascend).
static hipError_t set_fill_launch(bs2e_blockset* b, const BlockArgs* tab, hipStream_t st)
{
    const Plan& p = b->rk->plan;
    const int32_t lo = (int32_t)b->win_lo, hi = (int32_t)(b->win_lo + b->win_n);
    const auto l0 = b->rowlist.begin(), lf = l0 + b->n_far, ls = l0 + b->n_sep, le = b->rowlist.end();
    const int64_t f0 = std::lower_bound(l0, lf, lo) - l0, f1 = std::lower_bound(l0, lf, hi) - l0;
    const int64_t s0 = std::lower_bound(lf, ls, lo) - l0, s1 = std::lower_bound(lf, ls, hi) - l0;
    const int64_t g0 = std::lower_bound(ls, le, lo) - l0, g1 = std::lower_bound(ls, le, hi) - l0;
    return launch_block_fill(tab, b->d_row0, (int)b->n_sym, b->d_rowlist + f0, f1 - f0, b->d_rowlist + s0, s1 - s0,
                             b->d_rowlist + g0, g1 - g0, (int)p.k_s, (int)p.K, st);
}

// The outputs hold global rows [lo, lo + n).
static void set_window(bs2e_blockset* b, int64_t lo, int64_t n)
{
    b->win_lo = lo;
    b->win_n = n;
    b->winH = b->prefH[lo];
    b->winS = b->prefS[lo];
    b->nwinH = b->prefH[lo + n] - b->winH;
    b->nwinS = b->prefS[lo + n] - b->winS;
}

static bool window_is_whole(const bs2e_blockset* b) { return b->win_lo == 0 && b->win_n == b->n_rows; }

// descriptor of symmetry s for the count (fill = false) or fill pass
static BlockArgs set_args(const bs2e_blockset* b, int64_t s, bool fill)
{
    const Plan& p = b->rk->plan;
    const int64_t r0 = b->row0[s], N = b->n_rows;
    BlockArgs a{};
    a.n_config = (int)(b->row0[s + 1] - r0);
    a.n_chan = (int)b->n_chan[s];
    a.full = (int)b->full;
    a.L = (int)b->L[s];
    a.ks = (int)p.k_s;
    a.n_b = (int)p.n_b;
    a.n_pairs = (int)p.n_pairs;
    a.K = (int)p.K;
    a.n_iv = (int)p.n_iv;
    a.cfg = b->d_cfg + 2 * r0;
    a.cfg_chan = b->d_cfg_chan + r0;
    a.chan = b->d_chan + b->chan_off[s];
    a.seg = b->d_seg + b->seg_off[s];
    a.coef = b->d_coef + b->coef_off[s];
    a.flags = b->d_flags + b->flag_off[s];
    a.c_first = b->rk->d.c_first;
    a.R = b->rk->R;
    if (!fill) {
        a.cntH = b->d_cnt + r0;
        a.cntS = b->d_cnt + N + r0;
        a.refH = b->d_cnt + 2 * N + r0;
        a.refS = b->d_cnt + 3 * N + r0;
    } else {
        const size_t nn = (size_t)p.n_b * p.n_b, nf = (size_t)p.n_pairs * p.K;
        a.Hl = b->d_T;
        a.S = b->d_T + 2 * nn * (size_t)b->n_l;
        a.PtotT = a.S + 2 * nn;
        a.QtotT = a.PtotT + nf;
        a.ptrH = b->d_ptr + r0;
        a.ptrS = b->d_ptr + N + r0;
        a.winH = b->winH;
        a.winS = b->winS;
        a.limH = b->nwinH;
        a.limS = b->nwinS;
        a.ixH = b->d_ixH;
        a.dH = b->d_dH;
        a.ixS = b->d_ixS;
        a.dS = b->d_dS;
    }
    return a;
}

// Host tables, device upload and the count launch of a set whose rk, n_sym,
// full, L and row0 are given.  Returns BS2E_OK or the error code (set on ctx).
void bs2e::set_host_tables(int64_t nb, int64_t K, int64_t k_s, int64_t n_sym, const int64_t* L, const int64_t* row0,
                           const int64_t* configs, SetHost& h)
{
    const int64_t N = row0[n_sym];
    h = SetHost();
    if (N >= ((int64_t)1 << 31)) {
        h.err = "a set holds fewer than 2^31 configurations";
        return;
    }
    h.chan_off.assign(n_sym, 0);
    h.seg_off.assign(n_sym, 0);
    h.coef_off.assign(n_sym, 0);
    h.flag_off.assign(n_sym, 0);
    h.n_chan.assign(n_sym, 0);
    h.cfg_chan.assign((size_t)N, 0);
    for (int64_t s = 0; s < n_sym; ++s) {
        std::vector<BlockChan> ch;
        std::vector<BlockSeg> sg;
        std::vector<int32_t> cc;
        int64_t ml = 0;
        const int64_t r0 = row0[s], n = row0[s + 1] - r0;
        std::string bad = channel_tables(configs + 5 * r0, n, nb, ch, sg, cc, ml);
        if (!bad.empty()) {
            h.err = "symmetry " + std::to_string(s + 1) + ": " + bad + " (count_configs order required)";
            return;
        }
        h.max_l = std::max(h.max_l, ml);
        h.chan_off[s] = (int64_t)h.chan.size();
        h.seg_off[s] = (int64_t)h.seg.size();
        h.coef_off[s] = (int64_t)h.coef.size();
        h.flag_off[s] = (int64_t)h.flags.size();
        h.n_chan[s] = (int64_t)ch.size();
        std::copy(cc.begin(), cc.end(), h.cfg_chan.begin() + r0);
        h.chan.insert(h.chan.end(), ch.begin(), ch.end());
        h.seg.insert(h.seg.end(), sg.begin(), sg.end());
        // angular coefficients per (row channel, column channel): wigner_tools.f90:126-138
        const int64_t nch = (int64_t)ch.size(), Ls = L[s];
        const size_t c0 = h.coef.size(), f0 = h.flags.size();
        h.coef.resize(c0 + (size_t)(nch * nch * 2 * K), 0.0);
        h.flags.resize(f0 + (size_t)(nch * nch), 0);
        for (int64_t ci = 0; ci < nch; ++ci)
            for (int64_t cj = 0; cj < nch; ++cj) {
                const int64_t la = ch[ci].l1, lb = ch[ci].l2, lc = ch[cj].l1, ld = ch[cj].l2;
                const double sgn = ((lc + ld + Ls) & 1) ? -1.0 : 1.0;
                double* cd = h.coef.data() + c0 + (ci * nch + cj) * 2 * K;
                double* ce = cd + K;
                int32_t fl = 0;
                for (int64_t k = 0; k < K; ++k) {
                    const double ang = ang_k_LS(k, la, lb, lc, ld, Ls);
                    const double ang_ex = ang_k_LS(k, la, lb, ld, lc, Ls);
                    if (std::fabs(ang) >= 5e-16) cd[k] = ang;              // r_12_tens, mat_els.f90:566-570
                    if (std::fabs(ang_ex) >= 5e-16) ce[k] = sgn * ang_ex;
                    if (std::fabs(ang) > 5e-15) fl |= 1;                    // hamiltonian.f90:171-178
                    if (std::fabs(ang_ex) > 5e-15) fl |= 2;
                }
                h.flags[f0 + ci * nch + cj] = fl;
            }
    }
    // the configurations as the kernels read them (n1 | n2 << 16, l1 | l2 << 16:
    // channel_tables checked n < 2^15 and l < 64), and the row list in its
    // three parts
    h.pk.resize((size_t)2 * N);
    const int64_t kf = 2 * k_s - 1;
    for (int64_t g = 0; g < N; ++g) {
        const int64_t* c = configs + 5 * g;
        h.pk[2 * g] = (int32_t)(c[0] | (c[1] << 16));
        h.pk[2 * g + 1] = (int32_t)(c[2] | (c[3] << 16));
        const int64_t gap = std::llabs(c[0] - c[1]);
        h.n_far += gap >= kf;
        h.n_sep += gap >= k_s;
    }
    h.rowlist.resize((size_t)N);
    for (int64_t g = 0, a = 0, m = h.n_far, o = h.n_sep; g < N; ++g) {
        const int32_t w = h.pk[2 * g], gap = std::abs((w & 0xffff) - (w >> 16));
        h.rowlist[(size_t)(gap >= kf ? a++ : gap >= k_s ? m++ : o++)] = (int32_t)g;
    }
}

static int64_t set_create(bs2e_blockset* b, const int64_t* configs, const char* who, bool async,
                          const SetHost* prepared = nullptr)
{
    bs2e_ctx* ctx = b->rk->ctx;
    const Plan& p = b->rk->plan;
    const int64_t N = b->n_rows;
    // prepared tables are copied where the set keeps them and read in place
    // for the uploads, so that they serve again if this set is released
    // before its fill (the build service admits such a job anew)
    SetHost own;
    if (!prepared) set_host_tables(p.n_b, p.K, p.k_s, b->n_sym, b->L.data(), b->row0.data(), configs, own);
    const SetHost& h = prepared ? *prepared : own;
    if (!h.err.empty()) return fail(ctx, BS2E_EINVAL, std::string(who) + ": " + h.err);
    b->max_l = h.max_l;
    b->n_far = h.n_far;
    b->n_sep = h.n_sep;
    if (prepared) {
        b->chan_off = h.chan_off, b->seg_off = h.seg_off, b->coef_off = h.coef_off, b->flag_off = h.flag_off;
        b->n_chan = h.n_chan, b->chan = h.chan, b->seg = h.seg, b->coef = h.coef, b->flags = h.flags;
        b->rowlist = h.rowlist;
    } else {
        b->chan_off = std::move(own.chan_off), b->seg_off = std::move(own.seg_off);
        b->coef_off = std::move(own.coef_off), b->flag_off = std::move(own.flag_off);
        b->n_chan = std::move(own.n_chan), b->chan = std::move(own.chan), b->seg = std::move(own.seg);
        b->coef = std::move(own.coef), b->flags = std::move(own.flags);
        b->rowlist = std::move(own.rowlist);
    }
    const std::vector<int32_t>& pk = h.pk;
    const std::vector<int32_t>& cfg_chan = h.cfg_chan;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_rowlist, (size_t)N));
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_cfg, (size_t)2 * N));
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_cfg_chan, (size_t)N));
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_chan, b->chan.size()));
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_seg, b->seg.size()));
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_coef, std::max<size_t>(b->coef.size(), 1)));
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_flags, std::max<size_t>(b->flags.size(), 1)));
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_cnt, (size_t)4 * N));
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_ptr, (size_t)2 * N));
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_tab, (size_t)b->n_sym));
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_row0, (size_t)b->n_sym + 1));
    hipStream_t st = ctx->stream;
    // The context's pinned staging carries the set's tables up (a copy from
    // pageable memory can hold the host until the stream reaches it) and the
    // counts back down; stream order keeps the two uses apart.  It grows by
    // doubling (pinning memory takes milliseconds).
    b->tab.resize((size_t)b->n_sym);
    for (int64_t s = 0; s < b->n_sym; ++s) b->tab[s] = set_args(b, s, false);
    struct Up {
        void* dst;
        const void* src;
        size_t bytes;
    };
    const Up ups[] = {{b->d_cfg, pk.data(), sizeof(int32_t) * 2 * N},
                      {b->d_cfg_chan, cfg_chan.data(), sizeof(int32_t) * N},
                      {b->d_rowlist, b->rowlist.data(), sizeof(int32_t) * N},
                      {b->d_chan, b->chan.data(), sizeof(BlockChan) * b->chan.size()},
                      {b->d_seg, b->seg.data(), sizeof(BlockSeg) * b->seg.size()},
                      {b->d_coef, b->coef.data(), sizeof(double) * b->coef.size()},
                      {b->d_flags, b->flags.data(), sizeof(int32_t) * b->flags.size()},
                      {b->d_row0, b->row0.data(), sizeof(int64_t) * (b->n_sym + 1)},
                      {b->d_tab, b->tab.data(), sizeof(BlockArgs) * b->n_sym}};
    size_t up_bytes = 0;
    for (const Up& u : ups) up_bytes += (u.bytes + 255) / 256 * 256;
    std::unique_lock<std::mutex> pin_lock(ctx->pin_mu, std::defer_lock);
    if (!async) pin_lock.lock();
    const size_t need = std::max(sizeof(int64_t) * 4 * (size_t)N, up_bytes);
    if (ctx->pin_bytes < need) {
        const size_t grow = std::max(need, std::max<size_t>(2 * ctx->pin_bytes, (size_t)16 << 20));
        if (ctx->pin) (void)hipHostFree(ctx->pin);
        ctx->pin = nullptr;
        ctx->pin_bytes = 0;
        HIP_TRY(ctx, hipHostMalloc(&ctx->pin, grow, hipHostMallocDefault));
        ctx->pin_bytes = grow;
    }
    size_t off = 0;
    for (const Up& u : ups) {
        if (u.bytes == 0) continue;
        char* stage = (char*)ctx->pin + off;
        std::memcpy(stage, u.src, u.bytes);
        HIP_TRY(ctx, hipMemcpyAsync(u.dst, stage, u.bytes, hipMemcpyHostToDevice, st));
        off += (u.bytes + 255) / 256 * 256;
    }
    clear_stale_error();
    // the count pass, the rows' CSR offsets by a prefix sum on the device, and
    // the four per-row counts back to the host through the context's pinned
    // staging (the host keeps the offsets for the row windows and sizes).
    // async: enqueued behind an event and left running (the build service
    // owns its contexts and reads the counts with set_count_finish).
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_count_tmp, (size_t)row_scan_tmp_elems(N)));
    for (hipEvent_t& e : b->count_ev)
        if (!e) HIP_TRY(ctx, hipEventCreate(&e));
    hipError_t le = hipEventRecord(b->count_ev[0], st);
    if (le == hipSuccess) le = launch_block_count(b->d_tab, b->d_row0, (int)b->n_sym, N, st);
    if (le == hipSuccess) le = hipEventRecord(b->count_ev[1], st);
    if (le == hipSuccess) le = launch_row_scan(b->d_cnt, b->d_ptr, N, b->d_count_tmp, st);
    if (le == hipSuccess)
        le = hipMemcpyAsync(ctx->pin, b->d_cnt, sizeof(int64_t) * 4 * (size_t)N, hipMemcpyDeviceToHost, st);
    if (le == hipSuccess) le = hipEventRecord(b->count_ev[2], st);
    HIP_TRY(ctx, le);
    b->count_pending = true;
    return async ? BS2E_OK : set_count_finish(b);
}

// Waits for the count pass (set_create) and takes the rows' offsets, the
// symmetries' sizes and the reference's counts from the counts it copied to
// the context's pinned staging.
int64_t bs2e::set_count_finish(bs2e_blockset* b)
{
    bs2e_ctx* ctx = b->rk->ctx;
    if (!b->count_pending) return BS2E_OK;
    HIP_TRY(ctx, hipEventSynchronize(b->count_ev[2]));
    b->count_pending = false;
    HIP_TRY(ctx, hipEventElapsedTime(&b->ms_count, b->count_ev[0], b->count_ev[1]));
    ctx_free(ctx, b->d_count_tmp);
    const int64_t N = b->n_rows;
    const int64_t* cnt = (const int64_t*)ctx->pin;
    b->prefH.resize((size_t)N + 1);
    b->prefS.resize((size_t)N + 1);
    b->nnzH.assign(b->n_sym, 0);
    b->nnzS.assign(b->n_sym, 0);
    b->refH.assign(b->n_sym, 0);
    b->refS.assign(b->n_sym, 0);
    b->offH.assign(b->n_sym, 0);
    b->offS.assign(b->n_sym, 0);
    b->totH = b->totS = 0;
    int64_t* ph = b->prefH.data();
    int64_t* ps = b->prefS.data();
    for (int64_t s = 0; s < b->n_sym; ++s) {
        b->offH[s] = b->totH;
        b->offS[s] = b->totS;
        int64_t rh = 0, rs = 0;
        for (int64_t g = b->row0[s]; g < b->row0[s + 1]; ++g) {
            ph[g] = b->totH;
            ps[g] = b->totS;
            b->totH += cnt[g];
            b->totS += cnt[N + g];
            rh += cnt[2 * N + g];
            rs += cnt[3 * N + g];
        }
        b->refH[s] = rh;
        b->refS[s] = rs;
        b->nnzH[s] = b->totH - b->offH[s];
        b->nnzS[s] = b->totS - b->offS[s];
    }
    ph[N] = b->totH;
    ps[N] = b->totS;
    return BS2E_OK;
}

// Allocates the outputs and the fill's operand layouts, and enqueues the
// layouts and the fill launch over every row of the set on the context's
// stream, from the column-major one-electron matrices dHl (n_l of them) and
// dS already on the device.  wait: blocks until the fill is done and records
// its device time.  Results stay on the device.
int64_t bs2e::set_fill_alloc(bs2e_blockset* b, int64_t n_l, const char* who)
{
    bs2e_ctx* ctx = b->rk->ctx;
    if (n_l - 1 < b->max_l)
        return fail(ctx, BS2E_EINVAL, std::string(who) + ": one-electron matrices do not cover every l of the blocks");
    const Plan& p = b->rk->plan;
    const size_t nn = (size_t)p.n_b * p.n_b;
    const size_t nf = (size_t)p.n_pairs * p.K;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    set_drop_graph(b);   // its nodes address the buffers reallocated here
    ctx_free(ctx, b->d_T);
    ctx_free(ctx, b->d_ixH);
    ctx_free(ctx, b->d_dH);
    ctx_free(ctx, b->d_ixS);
    ctx_free(ctx, b->d_dS);
    b->filled = false;
    b->n_l = n_l;
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_T, 2 * nn * (size_t)(b->n_l + 1) + 2 * nf));
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_ixH, (size_t)std::max<int64_t>(b->totH, 1)));
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_dH, 2 * (size_t)std::max<int64_t>(b->totH, 1)));
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_ixS, (size_t)std::max<int64_t>(b->totS, 1)));
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_dS, 2 * (size_t)std::max<int64_t>(b->totS, 1)));
    b->capH = b->totH;
    b->capS = b->totS;
    set_window(b, 0, b->n_rows);
    return BS2E_OK;
}

int64_t bs2e::set_fill_enqueue(bs2e_blockset* b, const double* dHl, const double* dS, bool wait, const char* who)
{
    (void)who;
    bs2e_ctx* ctx = b->rk->ctx;
    const Plan& p = b->rk->plan;
    const size_t nn = (size_t)p.n_b * p.n_b;
    const size_t nf = (size_t)p.n_pairs * p.K;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    for (int64_t s = 0; s < b->n_sym; ++s) b->tab[s] = set_args(b, s, true);
    HIP_TRY(ctx, hipMemcpyAsync(b->d_tab, b->tab.data(), sizeof(BlockArgs) * b->n_sym, hipMemcpyHostToDevice, st));
    clear_stale_error();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (wait) {
        HIP_TRY(ctx, hipEventCreate(&e0));
        HIP_TRY(ctx, hipEventCreate(&e1));
    }
    hipError_t le = wait ? hipEventRecord(e0, st) : hipSuccess;
    if (le == hipSuccess)
        le = launch_block_prepare(dHl, dS, (int)b->n_l, (int)p.n_b, b->rk->d.Ptot, b->rk->d.Qtot, (int)p.n_pairs,
                                  (int)p.K, b->d_T, b->d_T + 2 * nn * b->n_l, b->d_T + 2 * nn * (b->n_l + 1),
                                  b->d_T + 2 * nn * (b->n_l + 1) + nf, st);
    if (le == hipSuccess)
        le = set_fill_launch(b, b->d_tab, st);
    if (wait) {
        if (le == hipSuccess) le = hipEventRecord(e1, st);
        if (le == hipSuccess) le = hipEventSynchronize(e1);
        if (le == hipSuccess) le = hipEventElapsedTime(&b->ms_fill, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    HIP_TRY(ctx, le);
    b->filled = true;   // stream-ordered: readers of the outputs on this stream see the fill
    return BS2E_OK;
}

int64_t bs2e::set_fill_device(bs2e_blockset* b, int64_t n_l, const double* dHl, const double* dS, bool wait,
                              const char* who)
{
    const int64_t rc = set_fill_alloc(b, n_l, who);
    return rc ? rc : set_fill_enqueue(b, dHl, dS, wait, who);
}

static size_t arena_round(size_t bytes) { return (std::max<size_t>(bytes, 1) + 255) / 256 * 256; }

size_t bs2e::blockset_fill_bytes(const bs2e_blockset* b, int64_t n_l)
{
    const Plan& p = b->rk->plan;
    const size_t nn = (size_t)p.n_b * p.n_b, nf = (size_t)p.n_pairs * p.K;
    const size_t H = (size_t)std::max<int64_t>(b->totH, 1), S = (size_t)std::max<int64_t>(b->totS, 1);
    return arena_round(8 * (2 * nn * (size_t)(n_l + 1) + 2 * nf)) + arena_round(8 * H) + arena_round(16 * H) +
           arena_round(8 * S) + arena_round(16 * S);
}

size_t bs2e::blockset_table_bytes(const bs2e_blockset* b)
{
    const size_t N = (size_t)b->n_rows;
    return arena_round(8 * N) + arena_round(4 * N) + arena_round(4 * N) +
           arena_round(sizeof(BlockChan) * b->chan.size()) + arena_round(sizeof(BlockSeg) * b->seg.size()) +
           arena_round(8 * b->coef.size()) +
           arena_round(4 * b->flags.size()) + arena_round(32 * N) + arena_round(16 * N) +
           arena_round(sizeof(BlockArgs) * (size_t)b->n_sym) + arena_round(8 * ((size_t)b->n_sym + 1));
}

// Uploads the caller's one-electron matrices, then fills as set_fill_device
// and waits.
static int64_t set_fill(bs2e_blockset* b, int64_t max_l_1p, const double* Hl, const double* S, const char* who)
{
    bs2e_ctx* ctx = b->rk->ctx;
    if (!Hl || !S) return fail(ctx, BS2E_EINVAL, std::string(who) + ": NULL argument");
    if (max_l_1p < b->max_l)
        return fail(ctx, BS2E_EINVAL, std::string(who) + ": one-electron matrices do not cover every l of the blocks");
    const Plan& p = b->rk->plan;
    const size_t nn = (size_t)p.n_b * p.n_b;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    set_drop_outputs(b);
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_Hl, 2 * nn * (size_t)(max_l_1p + 1)));
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_S, 2 * nn));
    hipStream_t st = ctx->stream;
    HIP_TRY(ctx, hipMemcpyAsync(b->d_Hl, Hl, sizeof(double) * 2 * nn * (max_l_1p + 1), hipMemcpyHostToDevice, st));
    HIP_TRY(ctx, hipMemcpyAsync(b->d_S, S, sizeof(double) * 2 * nn, hipMemcpyHostToDevice, st));
    return set_fill_device(b, max_l_1p + 1, b->d_Hl, b->d_S, true, who);
}

// Copies symmetry s's CSR arrays to the host (index_ptr 1-based as the
// reference's, hamiltonian.f90:142-143, 203-204).
static int64_t set_get(const bs2e_blockset* b, int64_t s, int64_t* ip_H, int64_t* ix_H, double* d_H, int64_t* ip_S,
                       int64_t* ix_S, double* d_S, const char* who)
{
    bs2e_ctx* ctx = b->rk->ctx;
    if (!ip_H || !ix_H || !d_H || !ip_S || !ix_S || !d_S)
        return fail(ctx, BS2E_EINVAL, std::string(who) + ": NULL argument");
    if (!b->filled) return fail(ctx, BS2E_ESTATE, std::string(who) + ": the blocks have not been filled");
    if (b->row0[s] < b->win_lo || b->row0[s + 1] > b->win_lo + b->win_n)
        return fail(ctx, BS2E_ESTATE, std::string(who) + ": the filled row window does not hold the whole symmetry");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int64_t nH = b->nnzH[s], nS = b->nnzS[s];
    const int64_t aH = b->offH[s] - b->winH, aS = b->offS[s] - b->winS;   // within the outputs
    if (nH > 0) {
        HIP_TRY(ctx, hipMemcpyAsync(ix_H, b->d_ixH + aH, sizeof(int64_t) * nH, hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipMemcpyAsync(d_H, b->d_dH + 2 * aH, sizeof(double) * 2 * nH, hipMemcpyDeviceToHost, st));
    }
    if (nS > 0) {
        HIP_TRY(ctx, hipMemcpyAsync(ix_S, b->d_ixS + aS, sizeof(int64_t) * nS, hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipMemcpyAsync(d_S, b->d_dS + 2 * aS, sizeof(double) * 2 * nS, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(ctx, hipStreamSynchronize(st));
    ip_H[0] = 1;
    ip_S[0] = 1;
    const int64_t r0 = b->row0[s], n = b->row0[s + 1] - r0;
    for (int64_t i = 0; i < n; ++i) {
        ip_H[i + 1] = 1 + b->prefH[r0 + i + 1] - b->prefH[r0];
        ip_S[i + 1] = 1 + b->prefS[r0 + i + 1] - b->prefS[r0];
    }
    return BS2E_OK;
}

// After a set's device memory is freed and the object deleted: drops its
// hold on the tensor.
static int64_t rk_unblock(bs2e_rk* rk)
{
    if (--rk->blocks == 0 && rk->closing) return rk_release_now(rk);   // the tensor's release was deferred
    return BS2E_OK;
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
    bs2e_block* b = new (std::nothrow) bs2e_block();
    if (!b) return fail(ctx, BS2E_ENOMEM, "bs2e_block_create: host allocation failed");
    b->rk = rk;
    b->n_sym = 1;
    b->full = full ? 1 : 0;
    b->L = {L};
    b->row0 = {0, n_config};
    b->n_rows = n_config;
    const int64_t rc = set_create(b, configs, "bs2e_block_create", false);
    if (rc) {
        set_free(b);
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
    const int64_t v[5] = {b->nnzH[0], b->nnzS[0], b->refH[0], b->refS[0], b->n_rows};
    for (int64_t i = 0; i < n && i < 5; ++i) sizes[i] = v[i];
    return BS2E_OK;
}

int64_t bs2e_block_fill(bs2e_block* b, int64_t max_l_1p, const double* Hl, const double* S, int64_t* ip_H,
                        int64_t* ix_H, double* d_H, int64_t* ip_S, int64_t* ix_S, double* d_S)
{
    if (!b || !Hl || !S || !ip_H || !ix_H || !d_H || !ip_S || !ix_S || !d_S)
        return fail(b ? b->rk->ctx : nullptr, BS2E_EINVAL, "bs2e_block_fill: NULL argument");
    int64_t rc = set_fill(b, max_l_1p, Hl, S, "bs2e_block_fill");
    if (rc == BS2E_OK) rc = set_get(b, 0, ip_H, ix_H, d_H, ip_S, ix_S, d_S, "bs2e_block_fill");
    set_drop_outputs(b);   // a single block hands its arrays over and keeps no device copy
    return rc;
}

int64_t bs2e_block_release(bs2e_block* b)
{
    if (!b) return BS2E_OK;
    bs2e_rk* rk = b->rk;
    (void)hipSetDevice(rk->ctx->device);
    set_free(b);
    delete b;
    return rk_unblock(rk);
}

int64_t bs2e_blockset_create(bs2e_rk* rk, int64_t n_sym, const int64_t* L, const int64_t* n_config,
                             const int64_t* configs, int64_t full, bs2e_blockset** out)
{
    if (!rk || !out) return fail(rk ? rk->ctx : nullptr, BS2E_EINVAL, "bs2e_blockset_create: NULL argument");
    *out = nullptr;
    if (!rk->computed)
        return fail(rk->ctx, BS2E_ESTATE, "bs2e_blockset_create: the R^k tensor has not been computed");
    return blockset_new(rk, n_sym, L, n_config, configs, full, out);
}

}  // extern "C"

// bs2e_blockset_create without the computed-tensor check: the count pass
// reads only the plan's tables, so a build service counts a job's blocks
// before its tensor is built (the fill, enqueued after the build on the same
// stream, reads the tensor).
int64_t bs2e::blockset_new(bs2e_rk* rk, int64_t n_sym, const int64_t* L, const int64_t* n_config,
                           const int64_t* configs, int64_t full, bs2e_blockset** out, bool async,
                           const SetHost* prepared)
{
    if (!rk || !out || !L || !n_config || (!configs && !prepared))
        return fail(rk ? rk->ctx : nullptr, BS2E_EINVAL, "bs2e_blockset_create: NULL argument");
    bs2e_ctx* ctx = rk->ctx;
    *out = nullptr;
    if (n_sym < 1 || n_sym > 100000) return fail(ctx, BS2E_EINVAL, "bs2e_blockset_create: need 1 <= n_sym <= 100000");
    int64_t total = 0;
    for (int64_t s = 0; s < n_sym; ++s) {
        if (L[s] < 0 || n_config[s] < 1 || n_config[s] >= ((int64_t)1 << 31))
            return fail(ctx, BS2E_EINVAL, "bs2e_blockset_create: need L >= 0 and 1 <= n_config < 2^31 per symmetry");
        total += n_config[s];
    }
    if (total >= ((int64_t)1 << 40)) return fail(ctx, BS2E_EINVAL, "bs2e_blockset_create: too many rows");
    bs2e_blockset* b = new (std::nothrow) bs2e_blockset();
    if (!b) return fail(ctx, BS2E_ENOMEM, "bs2e_blockset_create: host allocation failed");
    b->rk = rk;
    b->n_sym = n_sym;
    b->full = full ? 1 : 0;
    b->L.assign(L, L + n_sym);
    b->row0.assign((size_t)n_sym + 1, 0);
    for (int64_t s = 0; s < n_sym; ++s) b->row0[s + 1] = b->row0[s] + n_config[s];
    b->n_rows = total;
    const int64_t rc = set_create(b, configs, "bs2e_blockset_create", async, prepared);
    if (rc) {
        set_free(b);
        delete b;
        return rc;
    }
    rk->blocks += 1;
    *out = b;
    return BS2E_OK;
}

hipEvent_t bs2e::set_counted_event(const bs2e_blockset* b) { return b->count_ev[2]; }

extern "C" {

int64_t bs2e_blockset_sizes(const bs2e_blockset* b, int64_t* sizes, int64_t n)
{
    if (!b || !sizes) return fail(nullptr, BS2E_EINVAL, "bs2e_blockset_sizes: NULL argument");
    for (int64_t s = 0; s < b->n_sym; ++s) {
        const int64_t v[5] = {b->nnzH[s], b->nnzS[s], b->refH[s], b->refS[s], b->row0[s + 1] - b->row0[s]};
        for (int64_t i = 0; i < 5 && 5 * s + i < n; ++i) sizes[5 * s + i] = v[i];
    }
    return BS2E_OK;
}

int64_t bs2e_blockset_fill(bs2e_blockset* b, int64_t max_l_1p, const double* Hl, const double* S)
{
    if (!b) return fail(nullptr, BS2E_EINVAL, "bs2e_blockset_fill: NULL argument");
    return set_fill(b, max_l_1p, Hl, S, "bs2e_blockset_fill");
}

int64_t bs2e_blockset_get(const bs2e_blockset* b, int64_t sym, int64_t* ip_H, int64_t* ix_H, double* d_H,
                          int64_t* ip_S, int64_t* ix_S, double* d_S)
{
    if (!b) return fail(nullptr, BS2E_EINVAL, "bs2e_blockset_get: NULL argument");
    if (sym < 1 || sym > b->n_sym)
        return fail(b->rk->ctx, BS2E_ERANGE, "bs2e_blockset_get: symmetry index outside 1..n_sym");
    return set_get(b, sym - 1, ip_H, ix_H, d_H, ip_S, ix_S, d_S, "bs2e_blockset_get");
}

int64_t bs2e_blockset_get_rows(const bs2e_blockset* b, int64_t sym, int64_t row_lo, int64_t n_rows, int64_t* ip_H,
                               int64_t* ix_H, double* d_H, int64_t* ip_S, int64_t* ix_S, double* d_S)
{
    if (!b || !ip_H || !ip_S) return fail(nullptr, BS2E_EINVAL, "bs2e_blockset_get_rows: NULL argument");
    bs2e_ctx* ctx = b->rk->ctx;
    if (sym < 1 || sym > b->n_sym)
        return fail(ctx, BS2E_ERANGE, "bs2e_blockset_get_rows: symmetry index outside 1..n_sym");
    const int64_t s = sym - 1, r0 = b->row0[s], n = b->row0[s + 1] - r0;
    if (row_lo < 0 || n_rows < 0 || row_lo + n_rows > n)
        return fail(ctx, BS2E_ERANGE, "bs2e_blockset_get_rows: rows outside the symmetry's configurations");
    ip_H[0] = 1;
    ip_S[0] = 1;
    const int64_t g0 = r0 + row_lo;
    for (int64_t i = 0; i < n_rows; ++i) {
        ip_H[i + 1] = 1 + b->prefH[g0 + i + 1] - b->prefH[g0];
        ip_S[i + 1] = 1 + b->prefS[g0 + i + 1] - b->prefS[g0];
    }
    if (!ix_H && !d_H && !ix_S && !d_S) return BS2E_OK;   // sizes only
    if (!ix_H || !d_H || !ix_S || !d_S) return fail(ctx, BS2E_EINVAL, "bs2e_blockset_get_rows: NULL argument");
    if (!b->filled) return fail(ctx, BS2E_ESTATE, "bs2e_blockset_get_rows: the blocks have not been filled");
    const int64_t g = r0 + row_lo;   // first global row asked for
    if (g < b->win_lo || g + n_rows > b->win_lo + b->win_n)
        return fail(ctx, BS2E_ESTATE, "bs2e_blockset_get_rows: rows outside the filled row window");
    const int64_t aH = b->prefH[g] - b->winH, aS = b->prefS[g] - b->winS;   // first entry within the outputs
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int64_t nH = ip_H[n_rows] - 1, nS = ip_S[n_rows] - 1;
    if (nH > 0) {
        HIP_TRY(ctx, hipMemcpyAsync(ix_H, b->d_ixH + aH, sizeof(int64_t) * nH, hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipMemcpyAsync(d_H, b->d_dH + 2 * aH, sizeof(double) * 2 * nH, hipMemcpyDeviceToHost, st));
    }
    if (nS > 0) {
        HIP_TRY(ctx, hipMemcpyAsync(ix_S, b->d_ixS + aS, sizeof(int64_t) * nS, hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipMemcpyAsync(d_S, b->d_dS + 2 * aS, sizeof(double) * 2 * nS, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(ctx, hipStreamSynchronize(st));
    return BS2E_OK;
}

int64_t bs2e_blockset_row_counts(const bs2e_blockset* b, int64_t row_lo, int64_t n_rows, int64_t* cnt_H,
                                 int64_t* cnt_S)
{
    if (!b || !cnt_H || !cnt_S) return fail(b ? b->rk->ctx : nullptr, BS2E_EINVAL, "bs2e_blockset_row_counts: NULL argument");
    if (row_lo < 0 || n_rows < 0 || row_lo + n_rows > b->n_rows)
        return fail(b->rk->ctx, BS2E_ERANGE, "bs2e_blockset_row_counts: rows outside the set");
    for (int64_t r = 0; r < n_rows; ++r) {
        cnt_H[r] = b->prefH[row_lo + r + 1] - b->prefH[row_lo + r];
        cnt_S[r] = b->prefS[row_lo + r + 1] - b->prefS[row_lo + r];
    }
    return BS2E_OK;
}

// Output buffers with room for at least nH and nS entries: reallocated (their
// contents dropped) only when smaller.
static int64_t set_reserve(bs2e_blockset* b, int64_t nH, int64_t nS)
{
    bs2e_ctx* ctx = b->rk->ctx;
    if (nH > b->capH || !b->d_ixH) {
        set_drop_graph(b);
        b->filled = false;
        ctx_free(ctx, b->d_ixH);
        ctx_free(ctx, b->d_dH);
        b->capH = 0;
        HIP_TRY(ctx, ctx_alloc(ctx, &b->d_ixH, (size_t)std::max<int64_t>(nH, 1)));
        HIP_TRY(ctx, ctx_alloc(ctx, &b->d_dH, 2 * (size_t)std::max<int64_t>(nH, 1)));
        b->capH = std::max<int64_t>(nH, 1);
    }
    if (nS > b->capS || !b->d_ixS) {
        set_drop_graph(b);
        b->filled = false;
        ctx_free(ctx, b->d_ixS);
        ctx_free(ctx, b->d_dS);
        b->capS = 0;
        HIP_TRY(ctx, ctx_alloc(ctx, &b->d_ixS, (size_t)std::max<int64_t>(nS, 1)));
        HIP_TRY(ctx, ctx_alloc(ctx, &b->d_dS, 2 * (size_t)std::max<int64_t>(nS, 1)));
        b->capS = std::max<int64_t>(nS, 1);
    }
    return BS2E_OK;
}

int64_t bs2e_blockset_reserve(bs2e_blockset* b, int64_t nnz_H, int64_t nnz_S)
{
    if (!b) return fail(nullptr, BS2E_EINVAL, "bs2e_blockset_reserve: NULL argument");
    if (nnz_H < 0 || nnz_S < 0) return fail(b->rk->ctx, BS2E_EINVAL, "bs2e_blockset_reserve: negative size");
    HIP_TRY(b->rk->ctx, hipSetDevice(b->rk->ctx->device));
    return set_reserve(b, nnz_H, nnz_S);
}

int64_t bs2e_blockset_fill_rows(bs2e_blockset* b, int64_t max_l_1p, const double* Hl, const double* S, int64_t row_lo,
                                int64_t n_rows)
{
    const char* who = "bs2e_blockset_fill_rows";
    if (!b) return fail(nullptr, BS2E_EINVAL, std::string(who) + ": NULL argument");
    bs2e_ctx* ctx = b->rk->ctx;
    if (row_lo < 0 || n_rows < 1 || row_lo + n_rows > b->n_rows)
        return fail(ctx, BS2E_ERANGE, std::string(who) + ": rows outside the set");
    if ((Hl == nullptr) != (S == nullptr))
        return fail(ctx, BS2E_EINVAL, std::string(who) + ": give both Hl and S, or neither");
    if (!Hl && (!b->d_Hl || !b->d_S))
        return fail(ctx, BS2E_ESTATE, std::string(who) + ": no one-electron matrices on the device yet");
    if (Hl && max_l_1p < b->max_l)
        return fail(ctx, BS2E_EINVAL, std::string(who) + ": one-electron matrices do not cover every l of the blocks");
    const Plan& p = b->rk->plan;
    const size_t nn = (size_t)p.n_b * p.n_b, nf = (size_t)p.n_pairs * p.K;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    set_drop_graph(b);   // its fill addresses the whole set's outputs
    b->filled = false;
    hipStream_t st = ctx->stream;
    if (Hl) {
        ctx_free(ctx, b->d_Hl);
        ctx_free(ctx, b->d_S);
        ctx_free(ctx, b->d_T);
        b->n_l = max_l_1p + 1;
        HIP_TRY(ctx, ctx_alloc(ctx, &b->d_Hl, 2 * nn * (size_t)b->n_l));
        HIP_TRY(ctx, ctx_alloc(ctx, &b->d_S, 2 * nn));
        HIP_TRY(ctx, hipMemcpyAsync(b->d_Hl, Hl, sizeof(double) * 2 * nn * b->n_l, hipMemcpyHostToDevice, st));
        HIP_TRY(ctx, hipMemcpyAsync(b->d_S, S, sizeof(double) * 2 * nn, hipMemcpyHostToDevice, st));
    }
    if (!b->d_T) HIP_TRY(ctx, ctx_alloc(ctx, &b->d_T, 2 * nn * (size_t)(b->n_l + 1) + 2 * nf));
    // the outputs grow to the window (kept when they already hold it)
    const int64_t rc = set_reserve(b, b->prefH[row_lo + n_rows] - b->prefH[row_lo],
                                   b->prefS[row_lo + n_rows] - b->prefS[row_lo]);
    if (rc) return rc;
    set_window(b, row_lo, n_rows);
    return set_fill_enqueue(b, b->d_Hl, b->d_S, true, who);
}

int64_t bs2e_blockset_refill(bs2e_blockset* b)
{
    if (!b) return fail(nullptr, BS2E_EINVAL, "bs2e_blockset_refill: NULL argument");
    if (!b->filled || !b->d_Hl || !b->d_S)
        return fail(b->rk->ctx, BS2E_ESTATE, "bs2e_blockset_refill: fill the set from host matrices first");
    return set_fill_enqueue(b, b->d_Hl, b->d_S, true, "bs2e_blockset_refill");
}

int64_t bs2e_blockset_set_graph(bs2e_blockset* b, int64_t on)
{
    if (!b) return fail(nullptr, BS2E_EINVAL, "bs2e_blockset_set_graph: NULL argument");
    bs2e_rk* rk = b->rk;
    bs2e_ctx* ctx = rk->ctx;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    set_drop_graph(b);
    if (!on) return BS2E_OK;
    if (!b->filled || !b->d_Hl || !b->d_S)
        return fail(ctx, BS2E_ESTATE, "bs2e_blockset_set_graph: fill the set from host matrices first");
    if (!window_is_whole(b))
        return fail(ctx, BS2E_ESTATE, "bs2e_blockset_set_graph: the last fill was a row window; fill the whole set");
    if (rk->use_graph || rk->profiling || rk->from_primitives)
        return fail(ctx, BS2E_ESTATE,
                    "bs2e_blockset_set_graph: the tensor must build from its knots, without its own graph or profiling");
    const Plan& p = rk->plan;
    const int64_t N = b->n_rows;
    std::vector<BlockArgs> tg((size_t)b->n_sym);
    for (int64_t s = 0; s < b->n_sym; ++s) {
        const BlockArgs c = set_args(b, s, false);
        tg[s] = set_args(b, s, true);
        tg[s].cntH = c.cntH;
        tg[s].cntS = c.cntS;
        tg[s].refH = c.refH;
        tg[s].refS = c.refS;
    }
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_tabg, (size_t)b->n_sym));
    HIP_TRY(ctx, ctx_alloc(ctx, &b->d_scan, (size_t)row_scan_tmp_elems(N)));
    HIP_TRY(ctx, hipMemcpy(b->d_tabg, tg.data(), sizeof(BlockArgs) * tg.size(), hipMemcpyHostToDevice));
    const size_t nn = (size_t)p.n_b * p.n_b, nf = (size_t)p.n_pairs * p.K;
    hipStream_t st = ctx->stream;
    clear_stale_error();
    HIP_TRY(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int64_t rc = rk_enqueue_build(ctx, rk);
    hipError_t e = rc ? hipErrorLaunchFailure : hipSuccess;
    if (e == hipSuccess) e = launch_block_count(b->d_tabg, b->d_row0, (int)b->n_sym, N, st);
    if (e == hipSuccess) e = launch_row_scan(b->d_cnt, b->d_ptr, N, b->d_scan, st);
    if (e == hipSuccess)
        e = launch_block_prepare(b->d_Hl, b->d_S, (int)b->n_l, (int)p.n_b, rk->d.Ptot, rk->d.Qtot, (int)p.n_pairs,
                                 (int)p.K, b->d_T, b->d_T + 2 * nn * b->n_l, b->d_T + 2 * nn * (b->n_l + 1),
                                 b->d_T + 2 * nn * (b->n_l + 1) + nf, st);
    if (e == hipSuccess) e = set_fill_launch(b, b->d_tabg, st);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &g);
    if (e == hipSuccess) e = ec;
    if (e == hipSuccess) {
        b->graph = g;
        e = hipGraphInstantiateWithFlags(&b->graph_exec, b->graph, 0);
    } else if (g) {
        (void)hipGraphDestroy(g);
    }
    if (e != hipSuccess) {
        set_drop_graph(b);
        return rc ? rc : fail(ctx, BS2E_EDEVICE, std::string("bs2e_blockset_set_graph: ") + hipGetErrorString(e));
    }
    return BS2E_OK;
}

int64_t bs2e_blockset_replay(bs2e_blockset* b)
{
    if (!b) return fail(nullptr, BS2E_EINVAL, "bs2e_blockset_replay: NULL argument");
    bs2e_ctx* ctx = b->rk->ctx;
    if (!b->graph_exec) return fail(ctx, BS2E_ESTATE, "bs2e_blockset_replay: no graph (bs2e_blockset_set_graph)");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipGraphLaunch(b->graph_exec, ctx->stream));
    b->rk->computed = true;
    b->rk->mirror_valid = false;
    b->filled = true;
    return BS2E_OK;
}

int64_t bs2e_blockset_synchronize(bs2e_blockset* b)
{
    if (!b) return fail(nullptr, BS2E_EINVAL, "bs2e_blockset_synchronize: NULL argument");
    HIP_TRY(b->rk->ctx, hipStreamSynchronize(b->rk->ctx->stream));
    return BS2E_OK;
}

int64_t bs2e_blockset_times(const bs2e_blockset* b, double* ms, int64_t n)
{
    if (!b || !ms) return fail(nullptr, BS2E_EINVAL, "bs2e_blockset_times: NULL argument");
    const double v[2] = {b->ms_count, b->ms_fill};
    for (int64_t i = 0; i < n && i < 2; ++i) ms[i] = v[i];
    return BS2E_OK;
}

int64_t bs2e_blockset_release(bs2e_blockset* b)
{
    if (!b) return BS2E_OK;
    bs2e_rk* rk = b->rk;
    (void)hipSetDevice(rk->ctx->device);
    set_free(b);
    delete b;
    return rk_unblock(rk);
}

}  // extern "C"
