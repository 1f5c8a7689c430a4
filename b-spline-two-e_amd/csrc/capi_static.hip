// capi_static.hip -- the tensor build into a caller's buffer without any
// device allocation: the build's work tables live in one file-scope device
// array (each GPU that loads the library holds its own copy until the
// process ends).
#include <algorithm>
#include <cstring>
#include <mutex>

#include "capi_internal.h"

using namespace bs2e;

static constexpr size_t kScratchBytes = (size_t)128 << 20;   // 128 MiB

// Host code must not use this array's name as an address: on the host it is
// a stand-in.  The device address comes from hipGetSymbolAddress, per device.
static __device__ __align__(256) char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};   // the array's device address, per device

// Device address of `scratch` on the current device, looked up once per device.
hipError_t scratch_address(int device, char** out)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = (char*)g_scratch_ptr[device];
    return hipSuccess;
}

struct ScratchLayout {
    size_t t, bp, x, w, pairs, rows, c_first, small_end;
    size_t Pl, Ql, Pw, Qw, PP, QS, Ptot, Qtot, D, total;
};

ScratchLayout scratch_layout(const Plan& p)
{
    auto r = [](size_t bytes) { return (std::max<size_t>(bytes, 1) + 255) / 256 * 256; };
    ScratchLayout L{};
    size_t o = 0;
    L.t = o, o += r(8 * p.n_knots);
    L.bp = o, o += r(8 * (p.n_iv + 1));
    L.x = o, o += r(8 * p.k_gl);
    L.w = o, o += r(8 * p.k_gl);
    L.pairs = o, o += r(sizeof(PairInfo) * p.n_pairs);
    L.rows = o, o += r(4 * p.rows.size());
    L.c_first = o, o += r(4 * p.c_first.size());
    L.small_end = o;
    const size_t local = 8 * (size_t)p.n_iv * p.U * p.K, slots = 8 * (size_t)p.n_pairs * p.k_s * p.K;
    L.Pl = o, o += r(local);
    L.Ql = o, o += r(local);
    L.Pw = o, o += r(slots);
    L.Qw = o, o += r(slots);
    L.PP = o, o += r(slots);
    L.QS = o, o += r(slots);
    L.Ptot = o, o += r(8 * (size_t)p.n_pairs * p.K);
    L.Qtot = o, o += r(8 * (size_t)p.n_pairs * p.K);
    L.D = o, o += r(8 * (size_t)p.n_iv * p.U * p.U * p.K);
    L.total = o;
    return L;
}

}  // namespace

extern "C" {

int64_t bs2e_rk_build_static(int64_t k, int64_t n_knots, const double* knots, int64_t k_gl, int64_t max_k,
                             int64_t precision, double* R, int64_t n_elems, void* hip_stream)
{
    if (!knots || !R) return fail(nullptr, BS2E_EINVAL, "bs2e_rk_build_static: NULL argument");
    if (precision != BS2E_PREC_FP64 && precision != BS2E_PREC_MIXED)
        return fail(nullptr, BS2E_EINVAL, "bs2e_rk_build_static: unknown precision mode");
    Plan p;
    std::string msg = make_plan(k, n_knots, knots, k_gl, max_k, p);
    if (!msg.empty()) return fail(nullptr, BS2E_EINVAL, "bs2e_rk_build_static: " + msg);
    int64_t rc = check_plan_limits(nullptr, p, true);
    if (rc) return rc;
    if (precision == BS2E_PREC_MIXED && !mixed_supported(p.k_s, p.k_gl, p.K))
        return fail(nullptr, BS2E_EINVAL, "bs2e_rk_build_static: the mixed-precision moments take k_GL <= 16");
    if (n_elems != p.n_pairs * p.n_pairs * p.K)
        return fail(nullptr, BS2E_EINVAL, "bs2e_rk_build_static: n_elems differs from n_pairs^2 * (max_k+1)");
    build_near_rows(p);
    const ScratchLayout L = scratch_layout(p);
    if (L.total > kScratchBytes)
        return fail(nullptr, BS2E_EINVAL,
                    "bs2e_rk_build_static: the build's work tables (" + std::to_string(L.total) +
                        " bytes) exceed the static scratch of " + std::to_string(kScratchBytes) +
                        " bytes; use bs2e_rk_create / bs2e_rk_compute");
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    char* base = nullptr;
    if (e == hipSuccess) e = scratch_address(device, &base);
    if (e != hipSuccess)
        return fail(nullptr, BS2E_EDEVICE, std::string("bs2e_rk_build_static: ") + hipGetErrorString(e));
    // the small tables, packed as they lie in the scratch
    std::vector<char> host(L.small_end, 0);
    std::memcpy(host.data() + L.t, p.t.data(), 8 * p.n_knots);
    std::memcpy(host.data() + L.bp, p.bp.data(), 8 * (p.n_iv + 1));
    std::memcpy(host.data() + L.x, p.x.data(), 8 * p.k_gl);
    std::memcpy(host.data() + L.w, p.w.data(), 8 * p.k_gl);
    std::memcpy(host.data() + L.pairs, p.pairs.data(), sizeof(PairInfo) * p.n_pairs);
    std::memcpy(host.data() + L.rows, p.rows.data(), 4 * p.rows.size());
    std::memcpy(host.data() + L.c_first, p.c_first.data(), 4 * p.c_first.size());
    DevTables d;
    d.t = (double*)(base + L.t);
    d.bp = (double*)(base + L.bp);
    d.x = (double*)(base + L.x);
    d.w = (double*)(base + L.w);
    d.pairs = (PairInfo*)(base + L.pairs);
    d.rows = (int32_t*)(base + L.rows);
    d.c_first = (int32_t*)(base + L.c_first);
    d.Pl = (double*)(base + L.Pl);
    d.Ql = (double*)(base + L.Ql);
    d.Pw = (double*)(base + L.Pw);
    d.Qw = (double*)(base + L.Qw);
    d.PP = (double*)(base + L.PP);
    d.QS = (double*)(base + L.QS);
    d.Ptot = (double*)(base + L.Ptot);
    d.Qtot = (double*)(base + L.Qtot);
    d.D = (double*)(base + L.D);
    hipStream_t s = (hipStream_t)hip_stream;
    clear_stale_error();
    // Ordered after the stream's earlier work (an earlier call may still read
    // the scratch); the wait lets the host image go before the call returns.
    e = hipMemcpyAsync(base, host.data(), L.small_end, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    // one stream: moments, pair windows, diagonal cells (symmetrised), near
    // elements, rank-one fill; the same kernels and values as bs2e_rk_compute
    if (e == hipSuccess)
        e = precision == BS2E_PREC_MIXED ? launch_outer_moments_mixed(p, d, s) : launch_outer_moments(p, d, s);
    if (e == hipSuccess) e = launch_pair_windows(p, d, s);
    if (e == hipSuccess) e = launch_diag_cells(p, d, s, true);
    if (e == hipSuccess) e = launch_tensor_near(p, d, R, s);
    if (e == hipSuccess) e = launch_tensor_far(p, d, R, s);
    if (e != hipSuccess)
        return fail(nullptr, BS2E_EDEVICE, std::string("bs2e_rk_build_static: ") + hipGetErrorString(e));
    return BS2E_OK;
}

}  // extern "C"
