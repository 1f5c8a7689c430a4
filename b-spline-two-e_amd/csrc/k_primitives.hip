// k_primitives.hip -- per-interval primitives of the R^k build.
//
// B-spline values: the k splines that are non-zero on interval v are
// evaluated together at a point by the Cox-de Boor triangle (de Boor,
// A Practical Guide to Splines, BSPLVB), from the same differences
// d_p(j) = t(i+j) - x and d_m(j) = x - t(i+1-j), i = k-1+iv, that BVALUE_D
// forms (bspline_tools.f90:205-211); its per-spline recurrence (:213-220)
// gives the same values to rounding.
//
// kp_outer  (one workgroup per interval): spline values at the k_GL outer GL
//   points and the local pair moments of the U = k(k+1)/2 unordered local
//   spline pairs u = (a,c)
//     Pl[v][u][k] = h_v sum_q w_q B_a B_c r_q^k,
//     Ql[v][u][k] = h_v sum_q w_q B_a B_c / r_q^(k+1)       (mat_els.f90:413-428)
// kp_inner  (one workgroup per interval and chunk of kc multipoles): spline
//   values at the k_GL^2 inner points of the triangle rule and the diagonal
//   cells of mat_els.f90:460-483 as the contraction over the outer points
//     D[v][uo][ui][k] = h_v sum_q (w_q B_a B_c)(r_q) * J'[k][q][ui],
//     J'[k][q][u]     = s_q / r_q^(k+1) * sum_p w_p rho_qp^k B_b B_d(rho_qp),
//   with s_q = (r_q - bp_v)/2 and rho_qp = s_q x_p + (r_q + bp_v)/2; each
//   thread accumulates kc multipoles in registers, staging in LDS.
// kw_pair_windows: per banded pair p and multipole k, the window moments
//   Pw/Qw[p][j][k] (j-th interval of p's support window), their prefix (P)
//   and suffix (Q) sums and totals, gathered from Pl/Ql.
#include <algorithm>

#include "bs2e_internal.h"
#include "bspline_device.h"

namespace bs2e {

constexpr int kPThreads = 256;

hipError_t allow_lds(const void* fn, size_t bytes)
{
    if (bytes <= 64 * 1024) return hipSuccess;
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

size_t outer_lds_bytes(int64_t ks, int64_t kgl, int64_t K)
{
    return sizeof(double) * (size_t)(kgl * ks + kgl + 2 * kgl * K);
}

// Orders up to kMfOrder (U <= 45 local pairs, three 16-wide tiles) with at
// most 16 GL points contract the diagonal cells on the FP64 matrix cores
// (kp_inner_mf); other bases use the vector-ALU kernel kp_inner.
constexpr int kMfOrder = 9;
constexpr int kMfMaxKc = 16;  // multipoles per matrix-core workgroup (register accumulators)
constexpr int kMfSteps = 4;   // 4-point steps of the contraction: the matrix-core path takes k_GL <= 16

size_t inner_lds_bytes(int64_t ks, int64_t kgl, int64_t kc)
{
    const int64_t U = ks * (ks + 1) / 2;
    if (ks <= kMfOrder && kgl <= 4 * kMfSteps)   // sBo, sBi, sR, sWB, sX0, sXr, sJ[kc]
        return sizeof(double) * (size_t)(kgl * ks + kgl * kgl * ks + kgl + kgl * U + 2 * kgl * kgl + kc * kgl * U);
    return sizeof(double) *
           (size_t)(kgl * ks + kgl * kgl * ks + kgl + kgl * kgl + kgl * U + kc * kgl * kgl + kc * kgl * U);
}

int inner_kc(int64_t ks, int64_t kgl, int64_t K)
{
    if (ks <= kMfOrder && kgl <= 4 * kMfSteps) {   // all multipoles of an interval in one workgroup when they fit
        for (int64_t kc = std::min<int64_t>(K, kMfMaxKc); kc >= 1; --kc)
            if (inner_lds_bytes(ks, kgl, kc) <= kLdsPerWorkgroup) return (int)kc;
        return 0;
    }
    for (int kc : {4, 2, 1})
        if (kc <= K && inner_lds_bytes(ks, kgl, kc) <= kLdsPerWorkgroup) return kc;
    return inner_lds_bytes(ks, kgl, 1) <= kLdsPerWorkgroup ? 1 : 0;
}

template <int KS>
__global__ __launch_bounds__(kPThreads) void kp_outer(int K, int kgl, const double* __restrict__ t,
                                                      const double* __restrict__ bp, const double* __restrict__ x,
                                                      const double* __restrict__ w, double* __restrict__ Pl,
                                                      double* __restrict__ Ql)
{
    constexpr int U = KS * (KS + 1) / 2;
    extern __shared__ double lds[];
    double* sB = lds;                 // [kgl][KS]
    double* sR = sB + kgl * KS;       // [kgl]
    double* sPow = sR + kgl;          // [kgl][K]  r^k
    double* sInv = sPow + kgl * K;    // [kgl][K]  1 / r^(k+1)
    const int v = blockIdx.x;
    const int tid = threadIdx.x;
    const double lim1 = bp[v], lim2 = bp[v + 1];
    const double h = 0.5 * (lim2 - lim1);
    const double c = 0.5 * (lim2 + lim1);
    const int i1 = KS - 1 + (v + 1);
    for (int q = tid; q < kgl; q += kPThreads) {
        const double r = h * x[q] + c;
        sR[q] = r;
        splines_at<KS>(t, i1, r, sB + q * KS, 1);
    }
    __syncthreads();
    for (int e = tid; e < kgl * K; e += kPThreads) {
        const int q = e / K, k = e - q * K;
        const double r = sR[q];
        sPow[e] = ipow_d(r, k);
        sInv[e] = 1.0 / ipow_d(r, k + 1);
    }
    __syncthreads();
    for (int e = tid; e < U * K; e += kPThreads) {
        const int u = e / K, k = e - u * K;
        int ta, tc;
        local_pair(u, KS, ta, tc);
        double accp = 0.0, accq = 0.0;
        for (int q = 0; q < kgl; ++q) {
            const double bb = w[q] * sB[q * KS + ta] * sB[q * KS + tc];
            accp = accp + bb * sPow[q * K + k];
            accq = accq + bb * sInv[q * K + k];
        }
        Pl[((size_t)v * U + u) * K + k] = h * accp;
        Ql[((size_t)v * U + u) * K + k] = h * accq;
    }
}

// SYM: write Dsym[v][u1][u2][k] = D[v][u1][u2][k] + D[v][u2][u1][k] (what the
// tensor consumes) for u1 <= u2 only, the symmetric matrix's upper triangle,
// instead of D (what the sparse_6d container holds).
template <int KS, bool SYM>
__global__ __launch_bounds__(kPThreads) void kp_inner(int K, int kgl, int kc, int kchunks,
                                                      const double* __restrict__ t, const double* __restrict__ bp,
                                                      const double* __restrict__ x, const double* __restrict__ w,
                                                      double* __restrict__ D)
{
    constexpr int U = KS * (KS + 1) / 2;
    constexpr int KCMAX = 4;
    extern __shared__ double lds[];
    double* sBo = lds;                       // [kgl][KS]       outer points
    double* sBi = sBo + kgl * KS;            // [kgl][kgl][KS]  inner points
    double* sR = sBi + kgl * kgl * KS;       // [kgl]
    double* sRho = sR + kgl;                 // [kgl][kgl]
    double* sWB = sRho + kgl * kgl;          // [kgl][U]        w_q B_a B_c at r_q
    double* sX = sWB + kgl * U;              // [kc][kgl][kgl]  s_q w_p rho^k / r_q^(k+1)
    double* sJ = sX + kc * kgl * kgl;        // [kc][kgl][U]

    const int v = blockIdx.x / kchunks;
    const int k0 = (blockIdx.x - v * kchunks) * kc;
    const int nk = (K - k0) < kc ? (K - k0) : kc;
    const int tid = threadIdx.x;
    const double lim1 = bp[v], lim2 = bp[v + 1];
    const double h = 0.5 * (lim2 - lim1);
    const double c = 0.5 * (lim2 + lim1);
    const int i1 = KS - 1 + (v + 1);

    for (int e = tid; e < kgl * (1 + kgl); e += kPThreads) {
        if (e < kgl) {
            const double r = h * x[e] + c;
            sR[e] = r;
            splines_at<KS>(t, i1, r, sBo + e * KS, 1);
        } else {
            const int q = (e - kgl) / kgl, p = (e - kgl) - q * kgl;
            const double rq = h * x[q] + c;
            const double rho = 0.5 * (rq - lim1) * x[p] + 0.5 * (rq + lim1);
            sRho[q * kgl + p] = rho;
            splines_at<KS>(t, i1, rho, sBi + (q * kgl + p) * KS, 1);
        }
    }
    __syncthreads();
    for (int e = tid; e < kgl * U; e += kPThreads) {
        const int q = e / U, u = e - q * U;
        int ta, tc;
        local_pair(u, KS, ta, tc);
        sWB[e] = w[q] * sBo[q * KS + ta] * sBo[q * KS + tc];
    }
    for (int e = tid; e < nk * kgl * kgl; e += kPThreads) {
        const int kk = e / (kgl * kgl), qp = e - kk * kgl * kgl;
        const int q = qp / kgl, p = qp - q * kgl;
        const int k = k0 + kk;
        const double rq = sR[q];
        sX[e] = 0.5 * (rq - lim1) * w[p] * ipow_d(sRho[qp], k) / ipow_d(rq, k + 1);
    }
    __syncthreads();
    // J'[kk][q][u] = sum_p X[kk][q][p] B_b(rho_qp) B_d(rho_qp)
    for (int e = tid; e < kgl * U; e += kPThreads) {
        const int q = e / U, u = e - q * U;
        int ta, tc;
        local_pair(u, KS, ta, tc);
        double acc[KCMAX] = {0.0, 0.0, 0.0, 0.0};
        const double* bq = sBi + q * kgl * KS;
        for (int p = 0; p < kgl; ++p) {
            const double bb = bq[p * KS + ta] * bq[p * KS + tc];
#pragma unroll
            for (int kk = 0; kk < KCMAX; ++kk)
                if (kk < nk) acc[kk] += sX[(kk * kgl + q) * kgl + p] * bb;
        }
#pragma unroll
        for (int kk = 0; kk < KCMAX; ++kk)
            if (kk < nk) sJ[(kk * kgl + q) * U + u] = acc[kk];
    }
    __syncthreads();
    // D[v][uo][ui][k0 .. k0+nk) = h * sum_q WB[q][uo] J'[kk][q][ui]
    for (int e = tid; e < U * U; e += kPThreads) {
        const int uo = e / U, ui = e - uo * U;
        if (SYM && uo > ui) continue;   // the lower triangle mirrors the upper
        double acc[KCMAX] = {0.0, 0.0, 0.0, 0.0};
        for (int q = 0; q < kgl; ++q) {
            const double a = sWB[q * U + uo];
            const double b = SYM ? sWB[q * U + ui] : 0.0;
#pragma unroll
            for (int kk = 0; kk < KCMAX; ++kk)
                if (kk < nk) {
                    acc[kk] += a * sJ[(kk * kgl + q) * U + ui];
                    if (SYM) acc[kk] += b * sJ[(kk * kgl + q) * U + uo];
                }
        }
        double* dst = D + (((size_t)v * U + uo) * U + ui) * K + k0;
#pragma unroll
        for (int kk = 0; kk < KCMAX; ++kk)
            if (kk < nk) dst[kk] = h * acc[kk];
    }
}

// Matrix-core variant of kp_inner for orders <= kMfOrder.  Per workgroup:
// interval v and multipoles k0 .. k0+nk-1 (the chunks of one interval sit on
// one XCD, so the K-contiguous records of D are completed in one L2).
//   stage 1: splines at the outer and inner points; X_k0[q][p] =
//            s_q w_p rho^k0 / r_q^(k0+1) (from X_0 by k0 multiplications with
//            the ratio rho / r_q, so that every chunking gives the same bits);
//   stage 2: WB[q][u] = w_q B_a B_c (r_q);
//   stage 3: J'[kk][q][u] = sum_p X_k[q][p] B_b B_d (rho_qp), the powers for
//            k > k0 by repeated multiplication with the ratio;
//   stage 4: per 16 x 16 tile of (uo, ui) and multipole, on v_mfma_f64_16x16x4:
//            D = WB^T J' (SYM: Dsym = WB^T J' + J'^T WB, the two products in
//            separate accumulators, upper-triangle tiles and elements only),
//            zero-padded to whole tiles and to 16 GL points; two multipoles
//            per pass.
template <int KS, bool SYM>
__global__ __launch_bounds__(kPThreads) void kp_inner_mf(int K, int kgl, int kc, int kchunks,
                                                         const double* __restrict__ t, const double* __restrict__ bp,
                                                         const double* __restrict__ x, const double* __restrict__ w,
                                                         double* __restrict__ D)
{
    constexpr int U = KS * (KS + 1) / 2;
    constexpr int TU = (U + 15) / 16;
    typedef double d4_t __attribute__((ext_vector_type(4)));
    extern __shared__ double lds[];
    double* sBo = lds;                       // [kgl][KS]       outer points
    double* sBi = sBo + kgl * KS;            // [kgl][kgl][KS]  inner points
    double* sR = sBi + kgl * kgl * KS;       // [kgl]
    double* sWB = sR + kgl;                  // [kgl][U]        w_q B_a B_c at r_q
    double* sX0 = sWB + kgl * U;             // [kgl][kgl]      X_k0
    double* sXr = sX0 + kgl * kgl;           // [kgl][kgl]      rho / r_q
    double* sJ = sXr + kgl * kgl;            // [kc][kgl][U]

    const int64_t lg = xcd_block(blockIdx.x, gridDim.x);
    const int v = (int)(lg / kchunks);
    const int k0 = (int)(lg - (int64_t)v * kchunks) * kc;
    const int nk = (K - k0) < kc ? (K - k0) : kc;
    const int tid = threadIdx.x;
    const double lim1 = bp[v], lim2 = bp[v + 1];
    const double h = 0.5 * (lim2 - lim1);
    const double c = 0.5 * (lim2 + lim1);
    const int i1 = KS - 1 + (v + 1);

    for (int e = tid; e < kgl * (1 + kgl); e += kPThreads) {
        if (e < kgl) {
            const double r = h * x[e] + c;
            sR[e] = r;
            splines_at<KS>(t, i1, r, sBo + e * KS, 1);
        } else {
            const int q = (e - kgl) / kgl, pp = (e - kgl) - q * kgl;
            const double rq = h * x[q] + c;
            const double rho = 0.5 * (rq - lim1) * x[pp] + 0.5 * (rq + lim1);
            splines_at<KS>(t, i1, rho, sBi + (q * kgl + pp) * KS, 1);
            // X_k = s_q w_p rho^k / r_q^(k+1) = X_0 (rho / r_q)^k, the same
            // multiplication sequence from k = 0 for every chunking
            const double ratio = rho / rq;
            double xk = 0.5 * (rq - lim1) * w[pp] / rq;
            for (int j = 0; j < k0; ++j) xk *= ratio;
            sX0[q * kgl + pp] = xk;
            sXr[q * kgl + pp] = ratio;
        }
    }
    __syncthreads();
    for (int e = tid; e < kgl * U; e += kPThreads) {
        const int q = e / U, u = e - q * U;
        int ta, tc;
        local_pair(u, KS, ta, tc);
        sWB[e] = w[q] * sBo[q * KS + ta] * sBo[q * KS + tc];
    }
    for (int e = tid; e < kgl * U; e += kPThreads) {
        const int q = e / U, u = e - q * U;
        int ta, tc;
        local_pair(u, KS, ta, tc);
        double acc[kMfMaxKc];
#pragma unroll
        for (int kk = 0; kk < kMfMaxKc; ++kk) acc[kk] = 0.0;
        const double* bq = sBi + q * kgl * KS;
        for (int pp = 0; pp < kgl; ++pp) {
            const double bb = bq[pp * KS + ta] * bq[pp * KS + tc];
            double xk = sX0[q * kgl + pp];
            const double ratio = sXr[q * kgl + pp];
#pragma unroll
            for (int kk = 0; kk < kMfMaxKc; ++kk)
                if (kk < nk) {
                    acc[kk] += xk * bb;
                    xk *= ratio;
                }
        }
#pragma unroll
        for (int kk = 0; kk < kMfMaxKc; ++kk)
            if (kk < nk) sJ[(kk * kgl + q) * U + u] = acc[kk];
    }
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    // Operands are read from clamped addresses and zeroed outside the U x kgl
    // range (no lane-predicated loads), so the loads of a pass issue back to
    // back ahead of its matrix instructions.
    constexpr int NT = SYM ? TU * (TU + 1) / 2 : TU * TU;
    for (int tile = wave; tile < NT; tile += kPThreads / 64) {
        int ta, tb;
        if (SYM) {   // upper-triangle tiles, row by row
            ta = 0;
            int r = tile;
            while (r >= TU - ta) {
                r -= TU - ta;
                ++ta;
            }
            tb = ta + r;
        } else {
            ta = tile / TU;
            tb = tile - ta * TU;
        }
        const int uo_a = ta * 16 + li, ui_b = tb * 16 + li;   // operand row / column of this lane
        const bool va = uo_a < U, vb = ui_b < U;
        const int ua = va ? uo_a : U - 1, ub = vb ? ui_b : U - 1;
        int qo[kMfSteps];
        bool qv[kMfSteps];
#pragma unroll
        for (int st = 0; st < kMfSteps; ++st) {
            const int q = st * 4 + lk;
            qv[st] = q < kgl;
            qo[st] = (qv[st] ? q : kgl - 1) * U;
        }
        // WB operands do not depend on the multipole: loaded once per tile
        double a1[kMfSteps], b2[kMfSteps];
#pragma unroll
        for (int st = 0; st < kMfSteps; ++st) {
            const double x1 = sWB[qo[st] + ua], x2 = sWB[qo[st] + ub];
            a1[st] = (qv[st] && va) ? x1 : 0.0;
            b2[st] = (qv[st] && vb) ? x2 : 0.0;
        }
        for (int kk = 0; kk < nk; kk += 2) {
            const bool two = kk + 1 < nk;   // two multipoles per pass: four independent chains
            const double* J0 = sJ + (size_t)kk * kgl * U;
            const double* J1 = J0 + (two ? kgl * U : 0);
            double b10[kMfSteps], b11[kMfSteps], a20[kMfSteps], a21[kMfSteps];
#pragma unroll
            for (int st = 0; st < kMfSteps; ++st) {
                const double y0 = J0[qo[st] + ub], y1 = J1[qo[st] + ub];
                b10[st] = (qv[st] && vb) ? y0 : 0.0;
                b11[st] = (qv[st] && vb) ? y1 : 0.0;
                if (SYM) {
                    const double z0 = J0[qo[st] + ua], z1 = J1[qo[st] + ua];
                    a20[st] = (qv[st] && va) ? z0 : 0.0;
                    a21[st] = (qv[st] && va) ? z1 : 0.0;
                }
            }
            d4_t c0 = {0.0, 0.0, 0.0, 0.0}, c1 = c0, e0 = c0, e1 = c0;
#pragma unroll
            for (int st = 0; st < kMfSteps; ++st) {
                c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[st], b10[st], c0, 0, 0, 0);
                c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[st], b11[st], c1, 0, 0, 0);
                if (SYM) {
                    e0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a20[st], b2[st], e0, 0, 0, 0);
                    e1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a21[st], b2[st], e1, 0, 0, 0);
                }
            }
            // result row (lane >> 4) + 4 r, column lane & 15 of the tile
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int uo = ta * 16 + lk + 4 * r, ui = tb * 16 + li;
                if (uo < U && ui < U && (!SYM || uo <= ui)) {
                    double* dst = D + (((size_t)v * U + uo) * U + ui) * K + k0 + kk;
                    dst[0] = h * (SYM ? c0[r] + e0[r] : c0[r]);
                    if (two) dst[1] = h * (SYM ? c1[r] + e1[r] : c1[r]);
                }
            }
        }
    }
}

// Per (pair, k): window moments from the local tables, prefix/suffix sums.
template <int KSMAX>
__global__ __launch_bounds__(256) void kw_pair_windows(int n_pairs, int K, int ks, int U,
                                                       const PairInfo* __restrict__ pairs,
                                                       const double* __restrict__ Pl, const double* __restrict__ Ql,
                                                       double* __restrict__ Pw, double* __restrict__ Qw,
                                                       double* __restrict__ PP, double* __restrict__ QS,
                                                       double* __restrict__ Ptot, double* __restrict__ Qtot)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t)n_pairs * K) return;
    const int p = (int)(gid / K), k = (int)(gid - (int64_t)p * K);
    const PairInfo pi = pairs[p];
    double pw[KSMAX], qw[KSMAX];
#pragma unroll
    for (int j = 0; j < KSMAX; ++j) {
        pw[j] = 0.0;
        qw[j] = 0.0;
        if (j < pi.m) {
            const int v = pi.v0 - 1 + j;                         // 0-based interval
            const int u = sym_local(pi.a - v, pi.c - v, ks);     // local splines s - iv
            pw[j] = Pl[((size_t)v * U + u) * K + k];
            qw[j] = Ql[((size_t)v * U + u) * K + k];
        }
    }
    const size_t base = (size_t)p * ks * K + k;
    double run = 0.0;
#pragma unroll
    for (int j = 0; j < KSMAX; ++j) {
        if (j < ks) {
            run += pw[j];
            Pw[base + (size_t)j * K] = pw[j];
            Qw[base + (size_t)j * K] = qw[j];
            PP[base + (size_t)j * K] = (j < pi.m) ? run : 0.0;
        }
    }
    Ptot[(size_t)p * K + k] = run;
    run = 0.0;
#pragma unroll
    for (int j = KSMAX - 1; j >= 0; --j) {
        if (j < ks) {
            run += qw[j];
            QS[base + (size_t)j * K] = (j < pi.m) ? run : 0.0;
        }
    }
    Qtot[(size_t)p * K + k] = run;
}

// Window prefix/suffix sums from Pw/Qw already in place (the moments were
// gathered from caller-provided containers).
__global__ __launch_bounds__(256) void kw_window_sums(int n_pairs, int K, int ks, const PairInfo* __restrict__ pairs,
                                                      const double* __restrict__ Pw, const double* __restrict__ Qw,
                                                      double* __restrict__ PP, double* __restrict__ QS,
                                                      double* __restrict__ Ptot, double* __restrict__ Qtot)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t)n_pairs * K) return;
    const int p = (int)(gid / K), k = (int)(gid - (int64_t)p * K);
    const int m = pairs[p].m;
    const size_t base = (size_t)p * ks * K + k;
    double run = 0.0;
    for (int j = 0; j < ks; ++j) {
        if (j < m) run += Pw[base + (size_t)j * K];
        PP[base + (size_t)j * K] = (j < m) ? run : 0.0;
    }
    Ptot[(size_t)p * K + k] = run;
    run = 0.0;
    for (int j = ks - 1; j >= 0; --j) {
        if (j < m) run += Qw[base + (size_t)j * K];
        QS[base + (size_t)j * K] = (j < m) ? run : 0.0;
    }
    Qtot[(size_t)p * K + k] = run;
}

static inline unsigned blocks_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

template <int KS>
static void launch_outer(const Plan& p, const DevTables& d, hipStream_t s)
{
    const size_t lds = outer_lds_bytes(KS, p.k_gl, p.K);
    if (allow_lds((const void*)kp_outer<KS>, lds) != hipSuccess) return;   // the error stays for hipGetLastError
    hipLaunchKernelGGL(kp_outer<KS>, dim3((unsigned)p.n_iv), dim3(kPThreads), lds, s, (int)p.K, (int)p.k_gl, d.t, d.bp,
                       d.x, d.w, d.Pl, d.Ql);
}

template <int KS, bool SYM>
static void launch_inner(const Plan& p, const DevTables& d, hipStream_t s)
{
    const int kmax = inner_kc(KS, p.k_gl, p.K);
    const int kc = (p.inner_kc >= 1 && p.inner_kc <= kmax) ? p.inner_kc : kmax;
    const int kchunks = (int)((p.K + kc - 1) / kc);
    const size_t lds = inner_lds_bytes(KS, p.k_gl, kc);
    if (KS <= kMfOrder && p.k_gl <= 4 * kMfSteps) {
        const auto fn = kp_inner_mf<KS <= kMfOrder ? KS : 2, SYM>;
        if (allow_lds((const void*)fn, lds) != hipSuccess) return;   // the error stays for hipGetLastError
        hipLaunchKernelGGL(fn, dim3((unsigned)(p.n_iv * kchunks)), dim3(kPThreads), lds, s, (int)p.K, (int)p.k_gl, kc,
                           kchunks, d.t, d.bp, d.x, d.w, d.D);
    } else {
        const auto fn = kp_inner<KS, SYM>;
        if (allow_lds((const void*)fn, lds) != hipSuccess) return;
        hipLaunchKernelGGL(fn, dim3((unsigned)(p.n_iv * kchunks)), dim3(kPThreads), lds, s, (int)p.K, (int)p.k_gl, kc,
                           kchunks, d.t, d.bp, d.x, d.w, d.D);
    }
}

template <int KS>
static void launch_inner_sym(const Plan& p, const DevTables& d, hipStream_t s)
{
    launch_inner<KS, true>(p, d, s);
}

template <int KS>
static void launch_inner_plain(const Plan& p, const DevTables& d, hipStream_t s)
{
    launch_inner<KS, false>(p, d, s);
}

#define BS2E_ORDER_SWITCH(FN)                                                                                \
    switch (p.k_s) {                                                                                         \
        case 2: FN<2>(p, d, s); break;                                                                       \
        case 3: FN<3>(p, d, s); break;                                                                       \
        case 4: FN<4>(p, d, s); break;                                                                       \
        case 5: FN<5>(p, d, s); break;                                                                       \
        case 6: FN<6>(p, d, s); break;                                                                       \
        case 7: FN<7>(p, d, s); break;                                                                       \
        case 8: FN<8>(p, d, s); break;                                                                       \
        case 9: FN<9>(p, d, s); break;                                                                       \
        case 10: FN<10>(p, d, s); break;                                                                     \
        case 11: FN<11>(p, d, s); break;                                                                     \
        case 12: FN<12>(p, d, s); break;                                                                     \
        case 13: FN<13>(p, d, s); break;                                                                     \
        case 14: FN<14>(p, d, s); break;                                                                     \
        case 15: FN<15>(p, d, s); break;                                                                     \
        case 16: FN<16>(p, d, s); break;                                                                     \
        default: return hipErrorInvalidValue;                                                                \
    }

hipError_t launch_outer_moments(const Plan& p, const DevTables& d, hipStream_t s)
{
    BS2E_ORDER_SWITCH(launch_outer)
    return hipGetLastError();
}

hipError_t launch_diag_cells(const Plan& p, const DevTables& d, hipStream_t s, bool symmetrised)
{
    if (symmetrised) {
        BS2E_ORDER_SWITCH(launch_inner_sym)
    } else {
        BS2E_ORDER_SWITCH(launch_inner_plain)
    }
    return hipGetLastError();
}

template <int KSMAX>
static void launch_kw(const Plan& p, const DevTables& d, hipStream_t s)
{
    const int64_t n = p.n_pairs * p.K;
    hipLaunchKernelGGL(kw_pair_windows<KSMAX>, dim3(blocks_for(n, 256)), dim3(256), 0, s, (int)p.n_pairs, (int)p.K,
                       (int)p.k_s, (int)p.U, d.pairs, d.Pl, d.Ql, d.Pw, d.Qw, d.PP, d.QS, d.Ptot, d.Qtot);
}

hipError_t launch_pair_windows(const Plan& p, const DevTables& d, hipStream_t s)
{
    if (p.k_s <= 6) launch_kw<6>(p, d, s);
    else if (p.k_s <= 8) launch_kw<8>(p, d, s);
    else if (p.k_s <= 10) launch_kw<10>(p, d, s);
    else if (p.k_s <= 12) launch_kw<12>(p, d, s);
    else launch_kw<16>(p, d, s);
    return hipGetLastError();
}

hipError_t launch_window_sums(const Plan& p, const DevTables& d, hipStream_t s)
{
    const int64_t n = p.n_pairs * p.K;
    hipLaunchKernelGGL(kw_window_sums, dim3(blocks_for(n, 256)), dim3(256), 0, s, (int)p.n_pairs, (int)p.K,
                       (int)p.k_s, d.pairs, d.Pw, d.Qw, d.PP, d.QS, d.Ptot, d.Qtot);
    return hipGetLastError();
}

}  // namespace bs2e
