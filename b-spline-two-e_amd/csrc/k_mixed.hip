// k_mixed.hip -- the mixed-precision local moments: the pair densities
// A[u][q] = w_q B_a(r_q) B_c(r_q) and the radial moments r_q^k, r_q^-(k+1)
// contracted on the FP32 matrix cores (mat_els.f90:413-428:
//   P^k[v;a,c] = h_v sum_q w_q B_a B_c r_q^k,  Q^k = h_v sum_q w_q B_a B_c / r_q^(k+1)).
//
// Each operand is split in FP64 into NS integer slices of BETA bits with a
// power-of-two scale per row of A (per local pair) and per column of the
// moments (per multipole): x = 2^e (d_1 2^-BETA + d_2 2^-2BETA + ...), d_s
// integers, |d_1| <= 2^BETA, |d_s| <= 2^(BETA-1) for s > 1 (round to nearest),
// the remainder below 2^(e - NS BETA - 1).  Every slice product d_i g_j is an
// integer of at most 2 BETA bits and a 16-point sum of them stays below 2^24,
// so v_mfma_f32_16x16x4f32 accumulates the products of one slice pair without
// rounding.  The slice pairs of one level i + j share an accumulator (the
// levels 2 and 3 stay exact; rounding in the deeper levels is below
// 2^-24 of terms already 2^-3BETA down), and the levels are recombined in
// FP64 with the scales: the error-free transformation of a matrix product
// into products of narrow slices of Ozaki, Ogita, Oishi and Rump ("Error-free
// transformations of matrix multiplication by using fast routines of matrix
// multiplication and its applications", Numer. Algorithms 59, 2012).  With
// NS = 5 and BETA = 10 the moments carry about 2^-51 of the largest term of
// their row and column (with NS = 4, 2^-41, the element-wise error of the
// tensor's smallest elements reached 1.4e-10 at the order-8, 200-spline basis).
// A moment's error relative to its own value grows with the spread of the
// moment over the interval's GL points (the ratio of the column's scale to
// the points where the pair density sits): the intervals next to r = 0, where
// r^k spreads over more than 2^kMxRange, keep FP64 (with k <= 24 on an
// order-9 grid, the first three intervals).  The B-spline values (de Boor in FP64), the moments'
// scales and the recombination stay FP64, as do every later stage of the
// build (window sums, diagonal cells, the tensor's products).
#include "bs2e_internal.h"
#include "bspline_device.h"

namespace bs2e {

constexpr int kMxThreads = 256;
constexpr int kMxSlices = 5;     // NS
constexpr int kMxBeta = 10;      // bits per slice
constexpr int kMxQ = 16;         // GL points per contraction (k_GL <= 16, zero-padded)
constexpr int kMxRange = 11;     // log2 of the largest spread of r^K over an interval's GL points taken on the matrix cores

static constexpr int mx_tiles(int n) { return (n + 15) / 16; }

size_t mixed_lds_bytes(int64_t ks, int64_t K)
{
    const int64_t U = ks * (ks + 1) / 2;
    const int64_t UP = 16 * ((U + 15) / 16), KP = 16 * ((K + 15) / 16);
    return sizeof(double) * (size_t)(kMxQ * ks + kMxQ) + sizeof(float) * (size_t)(kMxSlices * kMxQ * (UP + 2 * KP)) +
           sizeof(int) * (size_t)(UP + 2 * KP);
}

bool mixed_supported(int64_t ks, int64_t kgl, int64_t K)
{
    return ks <= 16 && kgl <= kMxQ && K <= kMaxK && mixed_lds_bytes(ks, K) <= kLdsPerWorkgroup;
}

// One workgroup per interval.  LDS: spline values and points (FP64), then the
// slices of the pair densities sA[s][q][UP], of the moments r^k sP[s][q][KP]
// and r^-(k+1) sQ[s][q][KP] (FP32), and the row / column exponents.
template <int KS>
__global__ __launch_bounds__(kMxThreads) void kp_outer_mx(int K, int kgl, const double* __restrict__ t,
                                                          const double* __restrict__ bp,
                                                          const double* __restrict__ x,
                                                          const double* __restrict__ w, double* __restrict__ Pl,
                                                          double* __restrict__ Ql)
{
    constexpr int U = KS * (KS + 1) / 2;
    constexpr int TU = mx_tiles(U), UP = 16 * TU;
    typedef float f4_t __attribute__((ext_vector_type(4)));
    extern __shared__ double lds[];
    const int TK = mx_tiles(K), KP = 16 * TK;
    double* sB = lds;                                    // [kMxQ][KS]
    double* sR = sB + kMxQ * KS;                         // [kMxQ]
    float* sA = (float*)(sR + kMxQ);                     // [NS][kMxQ][UP]
    float* sP = sA + kMxSlices * kMxQ * UP;              // [NS][kMxQ][KP]
    float* sQ = sP + kMxSlices * kMxQ * KP;              // [NS][kMxQ][KP]
    int* eA = (int*)(sQ + kMxSlices * kMxQ * KP);        // [UP]
    int* eP = eA + UP;                                   // [KP]
    int* eQ = eP + KP;                                   // [KP]
    const int v = blockIdx.x;
    const int tid = threadIdx.x;
    const double lim1 = bp[v], lim2 = bp[v + 1];
    const double h = 0.5 * (lim2 - lim1);
    const double c = 0.5 * (lim2 + lim1);
    const int i1 = KS - 1 + (v + 1);
    for (int q = tid; q < kgl; q += kMxThreads) {
        const double r = h * x[q] + c;
        sR[q] = r;
        splines_at<KS>(t, i1, r, sB + q * KS, 1);
    }
    __syncthreads();
    // Near r = 0 the moments r^k, r^-(k+1) span (r_max / r_min)^K over the
    // interval's GL points (about 1e54 on the first interval of an order-9
    // grid with k <= 24): beyond 2^11, one scale per column leaves the small
    // points too few bits, and the interval's moments are summed in FP64 as
    // the plain path sums them (the same operations, so the same bits).
    if ((double)K * log2(sR[kgl - 1] / sR[0]) > (double)kMxRange) {
        double* sPow = (double*)sA;          // [kgl][K]  r^k        (the slice area is free here)
        double* sInv = sPow + kgl * K;       // [kgl][K]  1 / r^(k+1)
        for (int e = tid; e < kgl * K; e += kMxThreads) {
            const int q = e / K, k = e - q * K;
            const double r = sR[q];
            sPow[e] = ipow_d(r, k);
            sInv[e] = 1.0 / ipow_d(r, k + 1);
        }
        __syncthreads();
        for (int e = tid; e < U * K; e += kMxThreads) {
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
        return;
    }
    // slices: sixteen threads per pair-density row or moment column, one per
    // GL point (zero beyond k_GL); the column's scale is the exponent of its
    // largest value, found across the sixteen lanes
    for (int e = tid; e < kMxQ * (UP + 2 * KP); e += kMxThreads) {
        const int col = e >> 4, q = e & (kMxQ - 1);
        double val = 0.0;
        float* dst;
        int* ex;
        int stride;
        if (col < UP) {
            const int u = col;
            if (u < U && q < kgl) {
                int ta, tc;
                local_pair(u, KS, ta, tc);
                val = w[q] * sB[q * KS + ta] * sB[q * KS + tc];
            }
            dst = sA + q * UP + u;
            ex = eA + u;
            stride = kMxQ * UP;
        } else {
            const bool inv = col >= UP + KP;
            const int k = col - UP - (inv ? KP : 0);
            if (k < K && q < kgl) val = inv ? 1.0 / ipow_d(sR[q], k + 1) : ipow_d(sR[q], k);
            dst = (inv ? sQ : sP) + q * KP + k;
            ex = (inv ? eQ : eP) + k;
            stride = kMxQ * KP;
        }
        double m = val;
#pragma unroll
        for (int o = 1; o < kMxQ; o <<= 1) m = fmax(m, __shfl_xor(m, o, kMxQ));
        int ee = 0;
        if (m > 0.0) (void)frexp(m, &ee);   // m < 2^ee
        if (q == 0) *ex = ee;
        double y = ldexp(val, -ee);          // in [0, 1)
#pragma unroll
        for (int sl = 0; sl < kMxSlices; ++sl) {
            y = ldexp(y, kMxBeta);
            const double d = rint(y);
            y -= d;   // exact: the integer part and the remainder of a double
            dst[sl * stride] = (float)d;
        }
    }
    __syncthreads();
    // [A | r^k] and [A | r^-(k+1)]: 16 x 16 output tiles, one wave per tile
    const int wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int NT = 2 * TU * TK;
    for (int tile = wave; tile < NT; tile += kMxThreads / 64) {
        const bool inv = tile >= TU * TK;
        const int tt = tile - (inv ? TU * TK : 0);
        const int ta = tt / TK, tb = tt - ta * TK;
        const float* sM = inv ? sQ : sP;
        f4_t acc[kMxSlices];   // level l = i + j (2 .. NS + 1) in acc[l - 2]
#pragma unroll
        for (int l = 0; l < kMxSlices; ++l) acc[l] = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int i = 0; i < kMxSlices; ++i)
#pragma unroll
            for (int j = 0; i + j < kMxSlices; ++j)
#pragma unroll
                for (int st = 0; st < kMxQ / 4; ++st) {
                    const int q = 4 * st + lk;
                    const float a = sA[(i * kMxQ + q) * UP + ta * 16 + li];   // A[row li][inner lk]
                    const float b = sM[(j * kMxQ + q) * KP + tb * 16 + li];   // B[inner lk][column li]
                    acc[i + j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[i + j], 0, 0, 0);
                }
        // result rows 4 (lane >> 4) + r, column lane & 15 of the tile
        const int k = tb * 16 + li;
        const int* ex = inv ? eQ : eP;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int u = ta * 16 + 4 * lk + r;
            if (u < U && k < K) {
                double s = 0.0;
#pragma unroll
                for (int l = kMxSlices - 1; l >= 0; --l) s += ldexp((double)acc[l][r], -(l + 2) * kMxBeta);
                const double val = h * ldexp(s, eA[u] + ex[k]);
                (inv ? Ql : Pl)[((size_t)v * U + u) * K + k] = val;
            }
        }
    }
}

template <int KS>
static void launch_outer_mx(const Plan& p, const DevTables& d, hipStream_t s)
{
    const size_t lds = mixed_lds_bytes(KS, p.K);
    if (allow_lds((const void*)kp_outer_mx<KS>, lds) != hipSuccess) return;   // the error stays for hipGetLastError
    hipLaunchKernelGGL(kp_outer_mx<KS>, dim3((unsigned)p.n_iv), dim3(kMxThreads), lds, s, (int)p.K, (int)p.k_gl, d.t,
                       d.bp, d.x, d.w, d.Pl, d.Ql);
}

hipError_t launch_outer_moments_mixed(const Plan& p, const DevTables& d, hipStream_t s)
{
    if (!mixed_supported(p.k_s, p.k_gl, p.K)) return hipErrorInvalidValue;
    switch (p.k_s) {
        case 2: launch_outer_mx<2>(p, d, s); break;
        case 3: launch_outer_mx<3>(p, d, s); break;
        case 4: launch_outer_mx<4>(p, d, s); break;
        case 5: launch_outer_mx<5>(p, d, s); break;
        case 6: launch_outer_mx<6>(p, d, s); break;
        case 7: launch_outer_mx<7>(p, d, s); break;
        case 8: launch_outer_mx<8>(p, d, s); break;
        case 9: launch_outer_mx<9>(p, d, s); break;
        case 10: launch_outer_mx<10>(p, d, s); break;
        case 11: launch_outer_mx<11>(p, d, s); break;
        case 12: launch_outer_mx<12>(p, d, s); break;
        case 13: launch_outer_mx<13>(p, d, s); break;
        case 14: launch_outer_mx<14>(p, d, s); break;
        case 15: launch_outer_mx<15>(p, d, s); break;
        case 16: launch_outer_mx<16>(p, d, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace bs2e
