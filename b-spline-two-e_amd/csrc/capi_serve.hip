// capi_serve.hip -- a build service on one device: jobs (bases of differing
// orders, knot counts and multipole counts) are admitted as soon as a stream
// pair is free and their device footprint fits the service's HBM budget, and
// the admitted builds run concurrently.  No counterpart in the reference,
// which builds one basis per process (main_basis_setup.f90:75-86).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>

#include "capi_internal.h"

using namespace bs2e;

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

size_t arena_round(size_t bytes) { return (std::max<size_t>(bytes, 1) + 255) / 256 * 256; }

// The one-electron tables of one basis on the device: spline values and
// second derivatives at the GL points, S and H(l) for l = 0 .. n_l-1
// (column-major complex, as the fill reads them).  Jobs on the same knots, GL
// rule, charge and absorbing potential share one entry (a convergence scan
// resubmits a basis with other multipole counts and symmetries); an entry
// with more l's serves a job that needs fewer.
struct OneBody {
    int64_t k = 0, k_gl = 0, Z = 0, cap_order = 0, n_l = 0;
    double cap_r0 = 0, eta_re = 0, eta_im = 0;
    std::vector<double> knots;
    double *B = nullptr, *B2 = nullptr, *rw = nullptr, *S = nullptr, *H = nullptr;   // the service's arena
    hipEvent_t ready = nullptr;   // recorded after S and H are written
    hipStream_t made_on = nullptr;
    int refs = 0;                 // jobs admitted with it and not retired
    uint64_t last = 0;            // admission tick of its last use
    size_t bytes = 0;
    const void* maker = nullptr;  // the job it was computed for, until a job is admitted with it
};

size_t onebody_bytes(const Plan& p, int64_t n_l)
{
    const size_t tab = (size_t)p.n_iv * p.k_gl * p.k_s, nn = (size_t)p.n_b * p.n_b;
    return 2 * arena_round(8 * tab) + arena_round(16 * (size_t)p.n_iv * p.k_gl) + arena_round(16 * nn) +
           arena_round(16 * nn * (size_t)n_l);
}

// A job's block-set host tables, built on a worker thread from the job's
// configurations (copied at submission) while the job waits for a slot.
struct Prep {
    std::vector<int64_t> configs;   // freed once the tables are built
    std::vector<int64_t> L, row0;
    int64_t n_b = 0, K = 0, k_s = 0;
    SetHost h;
    std::mutex mu;
    std::condition_variable cv;
    bool done = false;

    void build()
    {
        try {
            set_host_tables(n_b, K, k_s, (int64_t)L.size(), L.data(), row0.data(), configs.data(), h);
        } catch (...) {   // (std::bad_alloc)
            h = SetHost();
            h.err = "host allocation failed";
        }
        std::vector<int64_t>().swap(configs);
        {
            std::lock_guard<std::mutex> lk(mu);
            done = true;
        }
        cv.notify_all();
    }
    void wait()
    {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [this] { return done; });
    }
};

// Host threads that build the jobs' tables, started with the first job; a
// task they cannot take runs on the caller's thread.
class Workers {
public:
    void post(std::function<void()> f)
    {
        try {
            std::lock_guard<std::mutex> lk(mu_);
            if (th_.empty()) {
                const unsigned n = std::max(1u, std::min(4u, std::thread::hardware_concurrency()));
                for (unsigned i = 0; i < n; ++i) th_.emplace_back([this] { run(); });
            }
            q_.push_back(std::move(f));
        } catch (...) {   // no thread or no queue entry: run it here
            if (f) f();
            return;
        }
        cv_.notify_one();
    }
    ~Workers()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (std::thread& t : th_) t.join();
    }

private:
    void run()
    {
        for (;;) {
            std::function<void()> f;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [this] { return stop_ || !q_.empty(); });
                if (q_.empty()) return;   // stopping, nothing left
                f = std::move(q_.front());
                q_.pop_front();
            }
            f();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::function<void()>> q_;
    std::vector<std::thread> th_;
    bool stop_ = false;
};

struct Job {
    Plan plan;                     // planned at submission, copied into the tensor at admission
    int64_t precision = BS2E_PREC_FP64;
    size_t bytes = 0;              // device footprint: tensor + tables (+ blocks, once counted)
    bool sized = false;            // bytes includes the blocks' exact sizes
    int state = 0;                 // 0 queued, 1 running, 2 done, 3 failed
    int phase = 0;                 // while running: 1 blocks being counted, 2 fill enqueued (done event recorded)
    int slot = -1;
    double t_submit = 0, t_admit = 0, t_done = 0;
    bs2e_rk* rk = nullptr;
    hipEvent_t done = nullptr;
    // the symmetry blocks of the job (n_sym = 0: the tensor only)
    int64_t n_sym = 0, full = 1, Z = 0, max_l_1p = 0, cap_order = 0;
    double cap_r0 = 0, eta_re = 0, eta_im = 0;
    std::vector<int64_t> L, n_config;
    std::shared_ptr<Prep> prep;    // the blocks' host tables (being) built
    bs2e_blockset* set = nullptr;
    OneBody* ob = nullptr;
};

}  // namespace

struct bs2e_server {
    int device = 0;
    Arena arena;
    std::vector<bs2e_ctx*> slots;
    std::vector<int64_t> slot_job;   // ticket running on each slot, or -1
    std::deque<int64_t> queue;
    std::map<int64_t, Job> jobs;
    std::deque<int64_t> finished;    // done since the last bs2e_server_finished
    std::list<OneBody> cache;        // one-electron tables by basis
    int64_t cache_hits = 0, cache_misses = 0;
    uint64_t tick = 0;
    int64_t next = 0;
    bs2e_ctx* err = nullptr;         // error messages of the service
    Workers workers;                 // (last: its threads end first)
};

static void onebody_free(bs2e_server* s, OneBody& o)
{
    for (double** p : {&o.B, &o.B2, &o.rw, &o.S, &o.H})
        if (*p) {
            s->arena.release(*p);
            *p = nullptr;
        }
    if (o.ready) (void)hipEventDestroy(o.ready);
    o.ready = nullptr;
}

// Frees the cache entries no job holds, least recently used first, until
// `need` bytes fit the largest free range; returns whether any were freed.
// `keep` (the entry of the job being admitted, if any) is spared.
static bool cache_evict(bs2e_server* s, size_t need, const OneBody* keep = nullptr)
{
    bool any = false;
    while (s->arena.largest_free() < need) {
        auto victim = s->cache.end();
        for (auto it = s->cache.begin(); it != s->cache.end(); ++it)
            if (it->refs == 0 && &*it != keep && (victim == s->cache.end() || it->last < victim->last)) victim = it;
        if (victim == s->cache.end()) break;
        onebody_free(s, *victim);
        s->cache.erase(victim);
        any = true;
    }
    return any;
}

static OneBody* cache_find(bs2e_server* s, const Job& j)
{
    const Plan& p = j.plan;
    for (OneBody& o : s->cache)
        if (o.k == p.k_s && o.k_gl == p.k_gl && o.Z == j.Z && o.cap_order == j.cap_order && o.cap_r0 == j.cap_r0 &&
            o.eta_re == j.eta_re && o.eta_im == j.eta_im && o.n_l >= j.max_l_1p + 1 && o.knots == p.t)
            return &o;
    return nullptr;
}

// A new entry for job j, computed on slot context c's stream from the job's
// tensor tables (knots, breakpoints, GL rule), which outlive the kernels: the
// job is retired only after its done event, recorded behind them.
static int64_t cache_make(bs2e_server* s, Job& j, bs2e_ctx* c, OneBody** out)
{
    const Plan& p = j.plan;
    s->cache.emplace_back();
    OneBody& o = s->cache.back();
    o.k = p.k_s, o.k_gl = p.k_gl, o.Z = j.Z, o.cap_order = j.cap_order, o.n_l = j.max_l_1p + 1;
    o.cap_r0 = j.cap_r0, o.eta_re = j.eta_re, o.eta_im = j.eta_im;
    o.knots = p.t;
    o.bytes = onebody_bytes(p, o.n_l);
    const size_t tab = (size_t)p.n_iv * p.k_gl * p.k_s, nn = (size_t)p.n_b * p.n_b;
    auto grab = [&](double** d, size_t n) { return (*d = (double*)s->arena.alloc(8 * n)) != nullptr; };
    if (!grab(&o.B, tab) || !grab(&o.B2, tab) || !grab(&o.rw, 2 * (size_t)p.n_iv * p.k_gl) || !grab(&o.S, 2 * nn) ||
        !grab(&o.H, 2 * nn * (size_t)o.n_l)) {
        onebody_free(s, o);
        s->cache.pop_back();
        return BS2E_ENOMEM;
    }
    hipError_t e = hipEventCreateWithFlags(&o.ready, hipEventDisableTiming);
    if (e == hipSuccess)
        e = launch_one_body(j.rk->plan, j.rk->d, o.B, o.B2, o.rw, 0, (int)o.n_l, (int)j.Z, (int)j.cap_order, j.cap_r0,
                            j.eta_re, j.eta_im, o.S, o.H, c->stream);
    if (e == hipSuccess) e = hipEventRecord(o.ready, c->stream);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(c->stream);
        onebody_free(s, o);
        s->cache.pop_back();
        return fail(s->err, BS2E_EDEVICE, std::string("bs2e_server: one-electron tables: ") + hipGetErrorString(e));
    }
    o.made_on = c->stream;
    o.maker = &j;
    *out = &o;
    return BS2E_OK;
}

// Frees what an admitted (or half-admitted) job holds on the device.
static void job_release(Job& j)
{
    if (j.set) {
        (void)bs2e_blockset_release(j.set);
        j.set = nullptr;
    }
    if (j.rk) {
        (void)bs2e_rk_release(j.rk);
        j.rk = nullptr;
    }
    if (j.ob) {
        j.ob->refs -= 1;
        j.ob = nullptr;
    }
}

static int64_t job_fail(bs2e_server* s, Job& j, int64_t code)
{
    (void)s;
    j.state = 3;
    job_release(j);
    return code;
}

// Admits job j on slot `slot`, phase one: the tensor (and its tables) in the
// arena and its build enqueued, the blocks' tables uploaded and their count
// pass enqueued behind it without waiting (the counts come back to the slot
// context's pinned staging while the build runs).  A job without blocks is
// done with its build.  Returns BS2E_ENOMEM (nothing held) when the budget
// lacks room now.
static int64_t admit(bs2e_server* s, Job& j, int slot)
{
    bs2e_ctx* c = s->slots[slot];
    Plan keep = j.plan;
    int64_t rc = rk_from_plan(c, std::move(keep), &j.rk, j.precision);
    if (rc) {
        j.rk = nullptr;
        return rc;
    }
    if (j.n_sym > 0) {
        j.prep->wait();
        rc = blockset_new(j.rk, j.n_sym, j.L.data(), j.n_config.data(), nullptr, j.full, &j.set, true, &j.prep->h);
        if (rc) {
            job_release(j);
            return rc;
        }
    }
    rc = bs2e_rk_compute(j.rk);
    if (!rc && !j.set && hipEventRecord(j.done, c->stream) != hipSuccess) rc = BS2E_EDEVICE;
    if (rc) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamSynchronize(c->side);
        job_release(j);
        return rc;
    }
    j.phase = j.set ? 1 : 2;
    return BS2E_OK;
}

// Phase two, once the counts are back: the blocks' exact sizes, the
// one-electron tables found in the cache or computed, the block outputs
// allocated (least recently used cache entries evicted if they are in the
// way), then the fill enqueued behind the build and the done event recorded.
// Returns BS2E_ENOMEM, with nothing held and j.bytes exact, when the outputs
// do not fit now.
static int64_t admit_fill(bs2e_server* s, Job& j)
{
    bs2e_ctx* c = s->slots[j.slot];
    int64_t rc = set_count_finish(j.set);
    OneBody* ob = rc ? nullptr : cache_find(s, j);
    if (!rc && !ob) {
        rc = cache_make(s, j, c, &ob);
        if (rc == BS2E_ENOMEM && cache_evict(s, onebody_bytes(j.plan, j.max_l_1p + 1))) rc = cache_make(s, j, c, &ob);
    }
    if (!rc) {
        j.ob = ob;
        ob->refs += 1;
        ob->last = ++s->tick;
        rc = set_fill_alloc(j.set, j.max_l_1p + 1, "bs2e_server");
        if (rc == BS2E_ENOMEM && cache_evict(s, blockset_fill_bytes(j.set, j.max_l_1p + 1), ob))
            rc = set_fill_alloc(j.set, j.max_l_1p + 1, "bs2e_server");
    }
    if (!rc) {
        if (j.ob->made_on != c->stream && hipStreamWaitEvent(c->stream, j.ob->ready, 0) != hipSuccess)
            rc = fail(s->err, BS2E_EDEVICE, "bs2e_server: hipStreamWaitEvent failed");
        if (!rc) rc = set_fill_enqueue(j.set, j.ob->H, j.ob->S, false, "bs2e_server");
        if (!rc && hipEventRecord(j.done, c->stream) != hipSuccess) rc = BS2E_EDEVICE;
    }
    if (j.set) {
        j.bytes = rk_footprint(j.plan) + blockset_table_bytes(j.set) + blockset_fill_bytes(j.set, j.max_l_1p + 1);
        j.sized = true;
    }
    if (rc) {
        (void)hipStreamSynchronize(c->stream);   // the build and a new entry's kernels read the job's tables
        (void)hipStreamSynchronize(c->side);
        job_release(j);
        return rc;
    }
    if (j.ob->maker == &j) s->cache_misses += 1;   // counted once per admitted job
    else s->cache_hits += 1;
    j.ob->maker = nullptr;
    j.phase = 2;
    return BS2E_OK;
}

// The event that marks a running job's next step: its counts (phase one) or
// its end.
static hipEvent_t pending_event(const Job& j) { return j.phase == 1 ? set_counted_event(j.set) : j.done; }

// Takes running jobs on (counted blocks to their fill, finished builds out of
// their slots) and admits queued jobs (FIFO; a job that does not fit yet is
// passed over by smaller ones behind it).
static int64_t progress(bs2e_server* s)
{
    (void)hipSetDevice(s->device);
    for (size_t i = 0; i < s->slots.size(); ++i) {
        const int64_t t = s->slot_job[i];
        if (t < 0) continue;
        Job& j = s->jobs[t];
        if (j.phase == 1) {
            const hipError_t e = hipEventQuery(set_counted_event(j.set));
            if (e == hipErrorNotReady) continue;
            if (e != hipSuccess)
                return fail(s->err, BS2E_EDEVICE, std::string("bs2e_server: count failed: ") + hipGetErrorString(e));
            const int64_t rc = admit_fill(s, j);
            if (rc) {
                s->slot_job[i] = -1;
                j.slot = -1;
                j.phase = 0;
                if (rc == BS2E_ENOMEM && j.bytes <= s->arena.capacity()) {   // room later: first in line again
                    j.state = 0;
                    s->queue.push_front(t);
                } else {
                    if (rc == BS2E_ENOMEM)
                        fail(s->err, BS2E_ENOMEM, "bs2e_server: the job's blocks and tensor exceed the service's budget");
                    else
                        fail(s->err, rc, "bs2e_server: job " + std::to_string(t) + ": " + bs2e_last_error(s->slots[i]));
                    job_fail(s, j, rc);
                    s->finished.push_back(t);   // handed out as failed
                }
                continue;
            }
        }
        const hipError_t e = hipEventQuery(j.done);
        if (e == hipErrorNotReady) continue;
        if (e != hipSuccess)
            return fail(s->err, BS2E_EDEVICE, std::string("bs2e_server: build failed: ") + hipGetErrorString(e));
        j.state = 2;
        j.phase = 0;
        j.t_done = now_ms();
        s->slot_job[i] = -1;
        s->finished.push_back(t);
    }
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        int slot = -1;
        for (size_t i = 0; i < s->slots.size(); ++i)
            if (s->slot_job[i] < 0) {
                slot = (int)i;
                break;
            }
        if (slot < 0) break;
        Job& j = s->jobs[*it];
        if (j.bytes > s->arena.largest_free() && !cache_evict(s, j.bytes, cache_find(s, j))) {
            ++it;
            continue;
        }
        int64_t rc = admit(s, j, slot);
        if (rc == BS2E_ENOMEM && cache_evict(s, j.bytes, cache_find(s, j))) rc = admit(s, j, slot);
        if (rc == BS2E_ENOMEM && j.bytes <= s->arena.capacity()) {   // room later: keep it queued
            ++it;
            continue;
        }
        if (rc == BS2E_ENOMEM)
            fail(s->err, BS2E_ENOMEM, "bs2e_server: the job's blocks and tensor exceed the service's budget");
        else if (rc)
            fail(s->err, rc, "bs2e_server: job " + std::to_string(*it) + ": " + bs2e_last_error(s->slots[slot]));
        if (rc) {
            job_fail(s, j, rc);
            s->finished.push_back(*it);   // handed out as failed
            it = s->queue.erase(it);
            continue;
        }
        j.state = 1;
        j.slot = slot;
        j.t_admit = now_ms();
        s->slot_job[slot] = *it;
        it = s->queue.erase(it);
    }
    int64_t pending = (int64_t)s->queue.size();
    for (int64_t t : s->slot_job) pending += t >= 0;
    return pending;
}

extern "C" {

int64_t bs2e_server_create(int64_t device, int64_t hbm_bytes, int64_t slots, bs2e_server** out)
{
    if (!out || hbm_bytes <= 0 || slots < 1 || slots > 16)
        return fail(nullptr, BS2E_EINVAL, "bs2e_server_create: need out, hbm_bytes > 0 and 1 <= slots <= 16");
    *out = nullptr;
    bs2e_server* s = new (std::nothrow) bs2e_server();
    if (!s) return fail(nullptr, BS2E_ENOMEM, "bs2e_server_create: host allocation failed");
    s->device = (int)device;
    int64_t rc = bs2e_create(&s->err, device);
    if (rc) {
        delete s;
        return rc;
    }
    (void)hipSetDevice(s->device);
    hipError_t e = s->arena.init((size_t)hbm_bytes);
    if (e != hipSuccess) {
        bs2e_destroy(s->err);
        delete s;
        return fail(nullptr, e == hipErrorOutOfMemory ? BS2E_ENOMEM : BS2E_EDEVICE,
                    std::string("bs2e_server_create: ") + hipGetErrorString(e));
    }
    for (int64_t i = 0; i < slots; ++i) {
        bs2e_ctx* c = nullptr;
        rc = bs2e_create(&c, device);
        if (rc) break;
        c->arena = &s->arena;
        s->slots.push_back(c);
        s->slot_job.push_back(-1);
    }
    if (rc) {
        for (bs2e_ctx* c : s->slots) bs2e_destroy(c);
        s->arena.destroy();
        bs2e_destroy(s->err);
        delete s;
        return rc;
    }
    *out = s;
    return BS2E_OK;
}

int64_t bs2e_server_destroy(bs2e_server* s)
{
    if (!s) return BS2E_OK;
    (void)hipSetDevice(s->device);
    for (bs2e_ctx* c : s->slots) {
        (void)hipStreamSynchronize(c->own);
        (void)hipStreamSynchronize(c->side);
    }
    for (auto& kv : s->jobs) {
        job_release(kv.second);
        if (kv.second.done) (void)hipEventDestroy(kv.second.done);
    }
    for (OneBody& o : s->cache) onebody_free(s, o);
    s->cache.clear();
    for (bs2e_ctx* c : s->slots) bs2e_destroy(c);
    s->arena.destroy();
    bs2e_destroy(s->err);
    delete s;
    return BS2E_OK;
}

const char* bs2e_server_last_error(const bs2e_server* s) { return bs2e_last_error(s ? s->err : nullptr); }

int64_t bs2e_footprint(int64_t k, int64_t n_knots, const double* knots, int64_t k_gl, int64_t max_k, int64_t* bytes)
{
    if (!bytes) return fail(nullptr, BS2E_EINVAL, "bs2e_footprint: bytes is NULL");
    Plan p;
    std::string msg = make_plan(k, n_knots, knots, k_gl, max_k, p);
    if (!msg.empty()) return fail(nullptr, BS2E_EINVAL, "bs2e_footprint: " + msg);
    build_near_rows(p);
    *bytes = (int64_t)rk_footprint(p);
    return BS2E_OK;
}

int64_t bs2e_server_submit_job(bs2e_server* s, const bs2e_job* job)
{
    if (!s || !job || !job->knots || job->n_knots < 1)
        return fail(s ? s->err : nullptr, BS2E_EINVAL, "bs2e_server_submit_job: bad argument");
    if (job->precision != BS2E_PREC_FP64 && job->precision != BS2E_PREC_MIXED)
        return fail(s->err, BS2E_EINVAL, "bs2e_server_submit_job: unknown precision mode");
    Job j;
    j.t_submit = now_ms();
    std::string msg = make_plan(job->k, job->n_knots, job->knots, job->k_gl, job->max_k, j.plan);
    if (!msg.empty()) return fail(s->err, BS2E_EINVAL, "bs2e_server_submit_job: " + msg);
    if (job->precision == BS2E_PREC_MIXED && !mixed_supported(j.plan.k_s, j.plan.k_gl, j.plan.K))
        return fail(s->err, BS2E_EINVAL, "bs2e_server_submit_job: the mixed-precision moments take k_GL <= 16");
    j.precision = job->precision;
    if (job->n_sym < 0) return fail(s->err, BS2E_EINVAL, "bs2e_server_submit_job: n_sym < 0");
    if (job->n_sym > 0) {
        // (sym_configs is read only without configs: callers built against
        // the structure without that member always pass configs)
        const int64_t* const* per_sym = job->configs ? nullptr : job->sym_configs;
        if (!job->L || !job->n_config || (!job->configs && !per_sym))
            return fail(s->err, BS2E_EINVAL, "bs2e_server_submit_job: NULL symmetry arrays");
        if (job->max_l_1p < 0 || job->cap_order < 0 || job->cap_order > 64)
            return fail(s->err, BS2E_EINVAL, "bs2e_server_submit_job: need max_l_1p >= 0, 0 <= cap_order <= 64");
        int64_t rows = 0;
        for (int64_t i = 0; i < job->n_sym; ++i) {
            if (job->L[i] < 0 || job->n_config[i] < 1 || (per_sym && !per_sym[i]))
                return fail(s->err, BS2E_EINVAL, "bs2e_server_submit_job: need L >= 0, n_config >= 1");
            rows += job->n_config[i];
        }
        std::shared_ptr<Prep> pr;
        try {
            pr = std::make_shared<Prep>();
            pr->configs.resize((size_t)(5 * rows));
            pr->L.assign(job->L, job->L + job->n_sym);
            pr->row0.assign((size_t)job->n_sym + 1, 0);
        } catch (...) {
            return fail(s->err, BS2E_ENOMEM, "bs2e_server_submit_job: host allocation failed");
        }
        for (int64_t i = 0; i < job->n_sym; ++i) pr->row0[i + 1] = pr->row0[i] + job->n_config[i];
        // the copy, symmetry by symmetry, with the l bound checked on the way
        int64_t* dst = pr->configs.data();
        for (int64_t i = 0; i < job->n_sym; ++i) {
            const int64_t n5 = 5 * job->n_config[i];
            const int64_t* src = per_sym ? per_sym[i] : job->configs + 5 * pr->row0[i];
            for (int64_t e = 0; e < n5; e += 5) {
                if (src[e + 2] > job->max_l_1p || src[e + 3] > job->max_l_1p)
                    return fail(s->err, BS2E_EINVAL, "bs2e_server_submit_job: a configuration's l exceeds max_l_1p");
                dst[e] = src[e], dst[e + 1] = src[e + 1], dst[e + 2] = src[e + 2], dst[e + 3] = src[e + 3];
                dst[e + 4] = src[e + 4];
            }
            dst += n5;
        }
        pr->n_b = j.plan.n_b, pr->K = j.plan.K, pr->k_s = j.plan.k_s;
        j.n_sym = job->n_sym;
        j.full = job->full ? 1 : 0;
        j.Z = job->Z, j.max_l_1p = job->max_l_1p, j.cap_order = job->cap_order;
        j.cap_r0 = job->cap_r0, j.eta_re = job->cap_eta_re, j.eta_im = job->cap_eta_im;
        j.L.assign(job->L, job->L + job->n_sym);
        j.n_config.assign(job->n_config, job->n_config + job->n_sym);
        j.prep = std::move(pr);
    }
    build_near_rows(j.plan);
    j.bytes = rk_footprint(j.plan);
    if (j.bytes > s->arena.capacity())
        return fail(s->err, BS2E_ENOMEM, "bs2e_server_submit_job: the job's device footprint exceeds the budget");
    (void)hipSetDevice(s->device);
    if (hipEventCreateWithFlags(&j.done, hipEventDisableTiming) != hipSuccess)
        return fail(s->err, BS2E_EDEVICE, "bs2e_server_submit_job: hipEventCreate failed");
    const int64_t t = s->next++;
    if (j.prep) {   // the blocks' host tables built on a worker thread meanwhile
        std::shared_ptr<Prep> pr = j.prep;
        s->workers.post([pr] { pr->build(); });
    }
    s->jobs[t] = std::move(j);
    s->queue.push_back(t);
    const int64_t rc = progress(s);
    return rc < 0 ? rc : t;
}

int64_t bs2e_server_submit(bs2e_server* s, int64_t k, int64_t n_knots, const double* knots, int64_t k_gl,
                           int64_t max_k)
{
    bs2e_job job{};
    job.k = k, job.n_knots = n_knots, job.knots = knots, job.k_gl = k_gl, job.max_k = max_k;
    job.precision = BS2E_PREC_FP64;
    return bs2e_server_submit_job(s, &job);
}

int64_t bs2e_server_progress(bs2e_server* s)
{
    if (!s) return fail(nullptr, BS2E_EINVAL, "bs2e_server_progress: NULL");
    return progress(s);
}

int64_t bs2e_server_status(bs2e_server* s, int64_t ticket, int64_t* state, double* times)
{
    if (!s) return fail(nullptr, BS2E_EINVAL, "bs2e_server_status: NULL");
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return fail(s->err, BS2E_EINVAL, "bs2e_server_status: unknown ticket");
    const Job& j = it->second;
    if (state) *state = j.state;
    if (times) {
        times[0] = j.state >= 1 ? j.t_admit - j.t_submit : -1.0;
        times[1] = j.state == 2 ? j.t_done - j.t_submit : -1.0;
    }
    return BS2E_OK;
}

int64_t bs2e_server_wait(bs2e_server* s, int64_t ticket)
{
    if (!s) return fail(nullptr, BS2E_EINVAL, "bs2e_server_wait: NULL");
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return fail(s->err, BS2E_EINVAL, "bs2e_server_wait: unknown ticket");
    for (;;) {
        const int64_t rc = progress(s);
        if (rc < 0) return rc;
        Job& j = it->second;
        if (j.state == 2) return BS2E_OK;
        if (j.state == 3) return fail(s->err, BS2E_EDEVICE, "bs2e_server_wait: the job failed to start");
        hipEvent_t e = nullptr;
        if (j.state == 1) {
            e = pending_event(j);
        } else {
            for (int64_t t : s->slot_job)
                if (t >= 0) {
                    e = pending_event(s->jobs[t]);
                    break;
                }
            if (!e)
                return fail(s->err, BS2E_ESTATE,
                            "bs2e_server_wait: the job cannot be admitted while finished jobs hold the budget "
                            "(retire them)");
        }
        HIP_TRY(s->err, hipEventSynchronize(e));
    }
}

int64_t bs2e_server_tensor(bs2e_server* s, int64_t ticket, bs2e_rk** rk)
{
    if (!s || !rk) return fail(nullptr, BS2E_EINVAL, "bs2e_server_tensor: NULL");
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end() || it->second.state != 2)
        return fail(s->err, BS2E_ESTATE, "bs2e_server_tensor: the job is unknown or not finished");
    *rk = it->second.rk;
    return BS2E_OK;
}

int64_t bs2e_server_blocks(bs2e_server* s, int64_t ticket, bs2e_blockset** set)
{
    if (!s || !set) return fail(nullptr, BS2E_EINVAL, "bs2e_server_blocks: NULL");
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end() || it->second.state != 2)
        return fail(s->err, BS2E_ESTATE, "bs2e_server_blocks: the job is unknown or not finished");
    if (!it->second.set) return fail(s->err, BS2E_ESTATE, "bs2e_server_blocks: the job has no blocks");
    *set = it->second.set;
    return BS2E_OK;
}

int64_t bs2e_server_cache(const bs2e_server* s, int64_t* info, int64_t n)
{
    if (!s || !info) return fail(nullptr, BS2E_EINVAL, "bs2e_server_cache: NULL");
    size_t bytes = 0;
    for (const OneBody& o : s->cache) bytes += o.bytes;
    const int64_t v[4] = {(int64_t)s->cache.size(), (int64_t)bytes, s->cache_hits, s->cache_misses};
    for (int64_t i = 0; i < n && i < 4; ++i) info[i] = v[i];
    return BS2E_OK;
}

int64_t bs2e_server_retire(bs2e_server* s, int64_t ticket)
{
    if (!s) return fail(nullptr, BS2E_EINVAL, "bs2e_server_retire: NULL");
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return fail(s->err, BS2E_EINVAL, "bs2e_server_retire: unknown ticket");
    Job& j = it->second;
    if (j.state == 0 || j.state == 1)
        return fail(s->err, BS2E_ESTATE, "bs2e_server_retire: the job is still queued or running");
    job_release(j);
    if (j.done) (void)hipEventDestroy(j.done);
    s->jobs.erase(it);
    const int64_t rc = progress(s);
    return rc < 0 ? rc : BS2E_OK;
}

int64_t bs2e_server_finished(bs2e_server* s, int64_t* tickets, double* latency_ms, int64_t cap)
{
    if (!s || (cap > 0 && !tickets)) return fail(nullptr, BS2E_EINVAL, "bs2e_server_finished: NULL");
    const int64_t rc = progress(s);
    if (rc < 0) return rc;
    int64_t n = 0;
    while (n < cap && !s->finished.empty()) {
        const int64_t t = s->finished.front();
        s->finished.pop_front();
        auto it = s->jobs.find(t);
        if (it == s->jobs.end()) continue;   // retired already
        tickets[n] = t;
        if (latency_ms) latency_ms[n] = it->second.state == 3 ? -1.0 : it->second.t_done - it->second.t_submit;
        ++n;
    }
    return n;
}

int64_t bs2e_server_memory(const bs2e_server* s, int64_t* mem)
{
    if (!s || !mem) return fail(nullptr, BS2E_EINVAL, "bs2e_server_memory: NULL");
    mem[0] = (int64_t)s->arena.capacity();
    mem[1] = (int64_t)s->arena.in_use();
    return BS2E_OK;
}

}  // extern "C"
