// capi_internal.h -- handle types and helpers shared by the C ABI sources.
#pragma once

#include <array>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bs2e.h"
#include "bs2e_internal.h"

namespace bs2e {

// First-fit sub-allocator over one device allocation: the memory of a build
// service, so that admitting and retiring jobs never calls hipMalloc/hipFree
// (hipFree waits for the whole device).  Host-side bookkeeping only.
class Arena {
public:
    hipError_t init(size_t bytes);
    void destroy();
    void* alloc(size_t bytes);   // 256-byte aligned; nullptr when no free range fits
    void release(void* p);
    size_t capacity() const { return cap_; }
    size_t in_use() const { return used_; }
    size_t largest_free() const;

private:
    char* base_ = nullptr;
    size_t cap_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;    // offset -> size
    std::map<size_t, size_t> taken_;   // offset -> size
    mutable std::mutex mu_;
};

}  // namespace bs2e

struct bs2e_ctx {
    int device = 0;
    int refs = 0;                   // live tensor handles created on this context
    bool closing = false;           // bs2e_destroy was called; freed with the last tensor
    bs2e::Arena* arena = nullptr;   // when set, the tensors of this context live in it
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;   // the build's stream (own, or the caller's)
    hipStream_t side = nullptr;     // diagonal cells + near ranges, concurrent with the rank-one fill
    std::string err;
    // pinned host staging for per-row tables: a block set's tables on their
    // way up and its counts on their way back (grown by doubling, freed with
    // the context; pin_mu held while it is in use)
    void* pin = nullptr;
    size_t pin_bytes = 0;
    std::mutex pin_mu;
    // pinned host staging for a tensor's small tables on their way up, reused
    // once up_ev (recorded behind the last upload from it) has passed
    void* up = nullptr;
    size_t up_bytes = 0;
    hipEvent_t up_ev = nullptr;
    std::mutex up_mu;
};

struct bs2e_rk {
    bs2e_ctx* ctx = nullptr;
    bs2e::Plan plan;
    bs2e::DevTables d;
    char* small = nullptr;          // one device block holding the small tables (knots .. c_first)
    std::vector<char> small_host;   // their packed host image, uploaded with one copy
    double* R = nullptr;
    size_t R_elems = 0;
    int64_t precision = BS2E_PREC_FP64;
    std::vector<double> mirror;
    bool mirror_valid = false;
    bool profiling = false;
    // profiling: one set of 6 timing events per build since profiling was
    // switched on (the pool is reused; see run_build for their meaning)
    std::vector<std::array<hipEvent_t, 6>> evpool;
    size_t evused = 0;
    hipEvent_t windows = nullptr, join = nullptr;                                  // stream ordering
    bool use_graph = false;         // compute launches a captured hipGraph of the build
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool computed = false;
    int blocks = 0;                 // live block handles reading this tensor
    bool closing = false;           // bs2e_rk_release was called; freed with the last block
    bool from_primitives = false;   // tables gathered from caller containers: compute runs the fill only
};

namespace bs2e {

int64_t fail(bs2e_ctx* ctx, int64_t code, const std::string& msg);

// Kernel launches report hipGetLastError(); an unrelated earlier call of the
// process that failed without being checked would otherwise be reported by
// the next launch, so a launch sequence starts by dropping it.
inline void clear_stale_error() { (void)hipGetLastError(); }

// Handle lifetimes: a tensor keeps its context, a block keeps its tensor; a
// destroy/release of the owner while dependents live is deferred to the
// release of the last dependent (so that no release order frees memory a
// live handle still reads).
// A handle other than a tensor that uses the context's streams holds it
// alive the same way (bs2e_destroy defers while any is held).
void ctx_hold(bs2e_ctx* ctx);
void ctx_drop(bs2e_ctx* ctx);
void rk_attach(bs2e_ctx* ctx, bs2e_rk* rk);
void rk_delete(bs2e_rk* rk);      // deletes the handle object, then drops its context reference
int64_t rk_release_now(bs2e_rk* rk);

template <class T>
inline hipError_t dev_alloc(T** p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void**)p, n * sizeof(T));
}

template <class T>
inline void dev_free(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Allocation of a context's tensors: its arena when it has one.
template <class T>
inline hipError_t ctx_alloc(bs2e_ctx* c, T** p, size_t n)
{
    if (c && c->arena) {
        *p = (T*)c->arena->alloc((n ? n : 1) * sizeof(T));
        return *p ? hipSuccess : hipErrorOutOfMemory;
    }
    return dev_alloc(p, n);
}

template <class T>
inline void ctx_free(bs2e_ctx* c, T*& p)
{
    if (p && c && c->arena) c->arena->release(p);
    else if (p) (void)hipFree(p);
    p = nullptr;
}

// Device bytes a tensor build of this plan allocates (tables and tensor,
// with the arena's alignment); the plan's near rows must be built.
size_t rk_footprint(const Plan& p);

// Channel runs of a configuration list in count_configs order (capi_block.hip).
std::string channel_tables(const int64_t* configs, int64_t n_config, int64_t nb, std::vector<BlockChan>& chan,
                           std::vector<BlockSeg>& seg, std::vector<int32_t>& cfg_chan, int64_t& max_l);

// Enqueues the two-stream tensor build of rk on ctx's streams (also under
// stream capture: events and kernels only).
int64_t rk_enqueue_build(bs2e_ctx* ctx, bs2e_rk* rk);

// The kernels' size limits for a plan (BS2E_EINVAL with a message when one is
// exceeded); needs_primitives: the plan's primitives are computed on the device.
int64_t check_plan_limits(bs2e_ctx* ctx, const Plan& p, bool needs_primitives);

// The host tables of a block set, built without the device: per symmetry the
// channel runs and the angular coefficients and flags of every pair of
// channels, the configurations packed as the kernels read them and the row
// list in its three parts.  The build service builds them on a worker thread
// while the job waits for a slot.
struct SetHost {
    std::string err;   // empty, or why the configurations were refused
    int64_t max_l = 0, n_far = 0, n_sep = 0;
    std::vector<int64_t> chan_off, seg_off, coef_off, flag_off, n_chan;
    std::vector<BlockChan> chan;
    std::vector<BlockSeg> seg;
    std::vector<int32_t> cfg_chan, flags, pk, rowlist;
    std::vector<double> coef;
};
// Fills h (or h.err) for the n_sym symmetries whose configurations are the
// rows [row0[s], row0[s + 1]) of configs, against a basis of n_b splines of
// order k_s and K multipoles.
void set_host_tables(int64_t n_b, int64_t K, int64_t k_s, int64_t n_sym, const int64_t* L, const int64_t* row0,
                     const int64_t* configs, SetHost& h);

// Blocks of several symmetries against a planned tensor (capi_block.hip):
// bs2e_blockset_create without the computed-tensor check (the count pass
// needs only the plan), and the fill from one-electron matrices already on the
// device (column-major, n_l of H then S), enqueued on the tensor context's
// stream (wait: block until done).  prepared: the set's host tables, built
// by set_host_tables from the same configurations (read, not changed; configs
// is then not read).
int64_t blockset_new(bs2e_rk* rk, int64_t n_sym, const int64_t* L, const int64_t* n_config, const int64_t* configs,
                     int64_t full, bs2e_blockset** out, bool async = false, const SetHost* prepared = nullptr);
// With async, blockset_new leaves the count pass running; set_count_finish
// waits for it and takes the sizes (set_counted_event: recorded behind it).
int64_t set_count_finish(bs2e_blockset* b);
hipEvent_t set_counted_event(const bs2e_blockset* b);
int64_t set_fill_device(bs2e_blockset* b, int64_t n_l, const double* dHl, const double* dS, bool wait,
                        const char* who);
// set_fill_device in two steps: the allocations (nothing enqueued), then the
// layouts and the fill launch.
int64_t set_fill_alloc(bs2e_blockset* b, int64_t n_l, const char* who);
int64_t set_fill_enqueue(bs2e_blockset* b, const double* dHl, const double* dS, bool wait, const char* who);
// Device bytes set_fill_device allocates for n_l one-electron matrices (with
// the arena's rounding), and the set's own tables (allocated at creation).
size_t blockset_fill_bytes(const bs2e_blockset* b, int64_t n_l);
size_t blockset_table_bytes(const bs2e_blockset* b);

// bs2e_rk_create for an existing plan with its near rows built.
int64_t rk_from_plan(bs2e_ctx* ctx, Plan&& plan, bs2e_rk** out, int64_t precision = BS2E_PREC_FP64);

}  // namespace bs2e

#define HIP_TRY(ctx, expr)                                                                                   \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess)                                                                                \
            return ::bs2e::fail((ctx), e_ == hipErrorOutOfMemory ? BS2E_ENOMEM : BS2E_EDEVICE,               \
                                std::string(#expr) + ": " + hipGetErrorString(e_));                          \
    } while (0)
