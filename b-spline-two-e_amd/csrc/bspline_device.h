// bspline_device.h -- B-spline values and derivatives on the device.
//
// Cox-de Boor triangle (de Boor, A Practical Guide to Splines, BSPLVB) for the
// k splines non-zero on an interval, from the differences d_p(j) =
// t(i+j) - x and d_m(j) = x - t(i+1-j), i = k-1+iv, that BVALUE_D forms
// (bspline_tools.f90:205-211); second derivatives by applying the
// derivative recurrence d/dx B_{s,m} = (m-1) [B_{s,m-1} / (t(s+m-1) - t(s))
// - B_{s+1,m-1} / (t(s+m) - t(s+1))] twice to the order k-2 values (the
// quantity BVALUE_D's derivative branch returns by differencing the
// coefficients, bspline_tools.f90:196-202).
#pragma once

#include <hip/hip_runtime.h>

namespace bs2e {

// Local pair u of an interval's k(k+1)/2 unordered spline pairs (row-major
// upper triangle) -> its splines ta <= tc.
__device__ __forceinline__ void local_pair(int u, int ks, int& ta, int& tc)
{
    int lo = 0, base = 0;
    while (base + (ks - lo) <= u) {
        base += ks - lo;
        ++lo;
    }
    ta = lo;
    tc = lo + (u - base);
}

__device__ __forceinline__ double ipow_d(double x, int m)
{
    // x**m for integer m >= 0 by square-and-multiply (how Fortran lowers an
    // integer power).
    unsigned n = (unsigned)m;
    double y = (n & 1u) ? x : 1.0;
    while (n >>= 1) {
        x = x * x;
        if (n & 1u) y = y * x;
    }
    return y;
}

// All KS splines non-zero on interval iv at x: out[ls] = B_{iv+ls}(x).
// t is stored 0-based (t(n) == t[n-1]); i1 = k-1+iv is the left knot index.
template <int KS>
__device__ __forceinline__ void splines_at(const double* __restrict__ t, int i1, double x, double* out, int stride)
{
    double dr[KS], dl[KS], b[KS + 1];
#pragma unroll
    for (int j = 1; j <= KS - 1; ++j) {
        dr[j] = t[i1 + j - 1] - x;   // d_p(j)
        dl[j] = x - t[i1 - j];       // d_m(j)
    }
    b[1] = 1.0;
#pragma unroll
    for (int j = 1; j <= KS - 1; ++j) {
        double saved = 0.0;
#pragma unroll
        for (int r = 1; r <= j; ++r) {
            const double term = b[r] / (dr[r] + dl[j + 1 - r]);
            b[r] = saved + dr[r] * term;
            saved = dl[j + 1 - r] * term;
        }
        b[j + 1] = saved;
    }
#pragma unroll
    for (int ls = 0; ls < KS; ++ls) out[ls * stride] = b[ls + 1];
}

// d1[ls] = B'_{s}(x) for the KS splines s = i1-KS+1+ls non-zero on the
// interval with left knot index i1: the recurrence once, on order KS-1 values.
template <int KS>
__device__ __forceinline__ void splines_d1_at(const double* __restrict__ t, int i1, double x, double* d1)
{
    if constexpr (KS < 2) {
        d1[0] = 0.0;
    } else {
        double low[KS - 1];   // order KS-1: splines i1-KS+2 .. i1
        splines_at<KS - 1>(t, i1, x, low, 1);
        const int s0 = i1 - KS + 1;
        auto bm1 = [&](int s) {
            const int ls = s - (i1 - KS + 2);
            return (ls >= 0 && ls < KS - 1) ? low[ls] : 0.0;
        };
        auto coef = [&](int s) {   // (KS-1) / (t(s+KS-1) - t(s))
            const double den = t[s + KS - 2] - t[s - 1];
            return den > 0.0 ? (double)(KS - 1) / den : 0.0;
        };
#pragma unroll
        for (int ls = 0; ls < KS; ++ls) {
            const int s = s0 + ls;
            d1[ls] = coef(s) * bm1(s) - coef(s + 1) * bm1(s + 1);
        }
    }
}

// d2[ls] = B''_{s}(x) for the KS splines s = i1-KS+1+ls non-zero on the
// interval with left knot index i1 (t stored 0-based, t(n) == t[n-1]); a term
// whose knot difference vanishes is zero.
template <int KS>
__device__ __forceinline__ void splines_d2_at(const double* __restrict__ t, int i1, double x, double* d2)
{
    if constexpr (KS < 3) {
#pragma unroll
        for (int ls = 0; ls < KS; ++ls) d2[ls] = 0.0;
    } else {
        double low[KS - 2];   // order KS-2: splines i1-KS+3 .. i1
        splines_at<KS - 2>(t, i1, x, low, 1);
        const int s0 = i1 - KS + 1;
        auto bm2 = [&](int s) {
            const int ls = s - (i1 - KS + 3);
            return (ls >= 0 && ls < KS - 2) ? low[ls] : 0.0;
        };
        auto coef = [&](int s, int m) {   // (m-1) / (t(s+m-1) - t(s))
            const double den = t[s + m - 2] - t[s - 1];
            return den > 0.0 ? (double)(m - 1) / den : 0.0;
        };
        double d1[KS + 1];   // first derivatives of order KS-1, splines s0 .. s0+KS
#pragma unroll
        for (int j = 0; j <= KS; ++j) {
            const int s = s0 + j;
            d1[j] = coef(s, KS - 1) * bm2(s) - coef(s + 1, KS - 1) * bm2(s + 1);
        }
#pragma unroll
        for (int ls = 0; ls < KS; ++ls) {
            const int s = s0 + ls;
            d2[ls] = coef(s, KS) * d1[ls] - coef(s + 1, KS) * d1[ls + 1];
        }
    }
}

}  // namespace bs2e
