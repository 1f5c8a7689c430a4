// capi_internal.h -- handle types and helpers shared by the C ABI sources.
#pragma once

#include <array>
#include <string>
#include <vector>

#include "../../include/bs2e.h"
#include "bs2e_internal.h"

struct bs2e_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;   // the build's stream (own, or the caller's)
    hipStream_t side = nullptr;     // diagonal cells + near ranges, concurrent with the rank-one fill
    std::string err;
};

struct bs2e_rk {
    bs2e_ctx* ctx = nullptr;
    bs2e::Plan plan;
    bs2e::DevTables d;
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
    bool computed = false;
    bool from_primitives = false;   // tables gathered from caller containers: compute runs the fill only
};

namespace bs2e {

int64_t fail(bs2e_ctx* ctx, int64_t code, const std::string& msg);

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

}  // namespace bs2e

#define HIP_TRY(ctx, expr)                                                                                   \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess)                                                                                \
            return ::bs2e::fail((ctx), e_ == hipErrorOutOfMemory ? BS2E_ENOMEM : BS2E_EDEVICE,               \
                                std::string(#expr) + ": " + hipGetErrorString(e_));                          \
    } while (0)
