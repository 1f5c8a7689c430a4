// capi_serve.hip -- a build service on one device: jobs (bases of differing
// orders, knot counts and multipole counts) are admitted as soon as a stream
// pair is free and their device footprint fits the service's HBM budget, and
// the admitted builds run concurrently.  No counterpart in the reference,
// which builds one basis per process (main_basis_setup.f90:75-86).
#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <new>

#include "capi_internal.h"

using namespace bs2e;

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Job {
    Plan plan;                     // planned at submission, handed to the tensor at admission
    size_t bytes = 0;
    int state = 0;                 // 0 queued, 1 running, 2 done, 3 failed
    int slot = -1;
    double t_submit = 0, t_admit = 0, t_done = 0;
    bs2e_rk* rk = nullptr;
    hipEvent_t done = nullptr;
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
    int64_t next = 0;
    bs2e_ctx* err = nullptr;         // error messages of the service
};

static int64_t job_fail(bs2e_server* s, Job& j, int64_t code)
{
    j.state = 3;
    if (j.rk) {
        (void)bs2e_rk_release(j.rk);
        j.rk = nullptr;
    }
    return code;
}

// Retires finished builds and admits queued jobs (FIFO; a job that does not
// fit yet is passed over by smaller ones behind it).
static int64_t progress(bs2e_server* s)
{
    (void)hipSetDevice(s->device);
    for (size_t i = 0; i < s->slots.size(); ++i) {
        const int64_t t = s->slot_job[i];
        if (t < 0) continue;
        Job& j = s->jobs[t];
        const hipError_t e = hipEventQuery(j.done);
        if (e == hipErrorNotReady) continue;
        if (e != hipSuccess)
            return fail(s->err, BS2E_EDEVICE, std::string("bs2e_server: build failed: ") + hipGetErrorString(e));
        j.state = 2;
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
        if (j.bytes > s->arena.largest_free()) {
            ++it;
            continue;
        }
        bs2e_ctx* c = s->slots[slot];
        Plan keep = j.plan;        // rk_from_plan consumes its plan; keep one for a retry
        int64_t rc = rk_from_plan(c, std::move(keep), &j.rk);
        if (rc == BS2E_ENOMEM) {   // fragmentation: keep it queued
            j.rk = nullptr;
            ++it;
            continue;
        }
        j.plan = Plan();
        if (rc) {
            job_fail(s, j, rc);
            it = s->queue.erase(it);
            continue;
        }
        rc = bs2e_rk_compute(j.rk);
        if (!rc && hipEventRecord(j.done, c->stream) != hipSuccess) rc = BS2E_EDEVICE;
        if (rc) {
            job_fail(s, j, rc);
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
        if (kv.second.rk) (void)bs2e_rk_release(kv.second.rk);
        if (kv.second.done) (void)hipEventDestroy(kv.second.done);
    }
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

int64_t bs2e_server_submit(bs2e_server* s, int64_t k, int64_t n_knots, const double* knots, int64_t k_gl,
                           int64_t max_k)
{
    if (!s || !knots || n_knots < 1) return fail(s ? s->err : nullptr, BS2E_EINVAL, "bs2e_server_submit: bad argument");
    Job j;
    j.t_submit = now_ms();
    std::string msg = make_plan(k, n_knots, knots, k_gl, max_k, j.plan);
    if (!msg.empty()) return fail(s->err, BS2E_EINVAL, "bs2e_server_submit: " + msg);
    build_near_rows(j.plan);
    j.bytes = rk_footprint(j.plan);
    if (j.bytes > s->arena.capacity())
        return fail(s->err, BS2E_ENOMEM, "bs2e_server_submit: the job's device footprint exceeds the service's budget");
    (void)hipSetDevice(s->device);
    if (hipEventCreateWithFlags(&j.done, hipEventDisableTiming) != hipSuccess)
        return fail(s->err, BS2E_EDEVICE, "bs2e_server_submit: hipEventCreate failed");
    const int64_t t = s->next++;
    s->jobs[t] = std::move(j);
    s->queue.push_back(t);
    const int64_t rc = progress(s);
    return rc < 0 ? rc : t;
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
            e = j.done;
        } else {
            for (int64_t t : s->slot_job)
                if (t >= 0) {
                    e = s->jobs[t].done;
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

int64_t bs2e_server_retire(bs2e_server* s, int64_t ticket)
{
    if (!s) return fail(nullptr, BS2E_EINVAL, "bs2e_server_retire: NULL");
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return fail(s->err, BS2E_EINVAL, "bs2e_server_retire: unknown ticket");
    Job& j = it->second;
    if (j.state == 0 || j.state == 1)
        return fail(s->err, BS2E_ESTATE, "bs2e_server_retire: the job is still queued or running");
    if (j.rk) (void)bs2e_rk_release(j.rk);
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
        if (latency_ms) latency_ms[n] = it->second.t_done - it->second.t_submit;
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
