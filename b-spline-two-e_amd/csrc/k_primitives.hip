// k_primitives.hip -- per-interval primitives of the R^k build.
//
//   K1 k1_bspline_tables : de Boor values of the k local splines of every
//                          interval at the outer GL points and at the inner
//                          points of the triangle rule (bspline_tools.f90:151-224,
//                          mat_els.f90:413-419, :462-475).
//   K2 k2_pair_moments   : P^k[v;a,c] = h_v sum_q w_q B_a B_c r^k and
//                          Q^k = h_v sum_q w_q B_a B_c / r^(k+1) over each banded
//                          pair's support window (mat_els.f90:413-428), with the
//                          window prefix/suffix sums the tensor kernel contracts.
//   K4 k4_diag_cells     : D^k[v; outer (i,i'), inner (j,j')] of mat_els.f90:460-483,
//                          computed as the per-interval contraction
//                          D = h_v * O^T J over the k_GL outer points, with
//                          O[q][u] = w_q B_i B_i' / r_q^(k+1) and
//                          J[q][u] = s_q sum_p w_p B_j B_j' rho_qp^k, staged in LDS.
#include "bs2e_internal.h"

namespace bs2e {

__device__ inline double ipow_d(double x, int m)
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

// ---------------------------------------------------------------------------
// K1: one thread per (interval, point); point < k_gl is an outer GL point,
// otherwise inner point (q, p) of the triangle rule on [bp_v, r_q].
// BVALUE_D with iv known uses left knot i = k-1+iv (bspline_tools.f90:175); the
// unit coefficient e_s selects local spline s-iv.  Each local spline runs the
// reference's convex-combination recurrence (:213-220).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k1_bspline_tables(int n_iv, int ks, int kgl, const double* __restrict__ t,
                                                         const double* __restrict__ bp, const double* __restrict__ x,
                                                         double* __restrict__ Bo, double* __restrict__ Bi,
                                                         double* __restrict__ ro)
{
    const int pts = kgl * (1 + kgl);
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t)n_iv * pts) return;
    const int v = (int)(gid / pts);          // 0-based interval
    const int pt = (int)(gid - (int64_t)v * pts);
    const double lim1 = bp[v], lim2 = bp[v + 1];
    const double scale = 0.5 * (lim2 - lim1);
    const double translate = 0.5 * (lim2 + lim1);
    double r;
    double* out;
    if (pt < kgl) {
        r = scale * x[pt] + translate;
        ro[(size_t)v * kgl + pt] = r;
        out = Bo + ((size_t)v * kgl + pt) * ks;
    } else {
        const int q = (pt - kgl) / kgl, p = (pt - kgl) - q * kgl;
        const double rq = scale * x[q] + translate;
        const double scale_j = 0.5 * (rq - lim1);
        const double translate_j = 0.5 * (rq + lim1);
        r = scale_j * x[p] + translate_j;
        out = Bi + (((size_t)v * kgl + q) * kgl + p) * ks;
    }
    // d_p(j) = t(i+j) - x, d_m(j) = x - t(i+1-j), i = k-1+iv, iv = v+1 (1-based);
    // t is stored 0-based so t(n) == t[n-1].
    const int i1 = ks - 1 + (v + 1);
    double dp[kMaxOrder + 1], dm[kMaxOrder + 1];
    for (int j = 1; j <= ks - 1; ++j) {
        dp[j] = t[i1 + j - 1] - r;
        dm[j] = r - t[i1 + 1 - j - 1];
    }
    for (int ls = 0; ls < ks; ++ls) {
        double a[kMaxOrder + 2];
        for (int j = 1; j <= ks; ++j) a[j] = (j == ls + 1) ? 1.0 : 0.0;
        for (int j = 1; j <= ks - 1; ++j) {
            const int kmj = ks - j;
            int ii = kmj;
            for (int jj = 1; jj <= kmj; ++jj) {
                a[jj] = (a[jj + 1] * dm[ii] + a[jj] * dp[jj]) / (dm[ii] + dp[jj]);
                --ii;
            }
        }
        out[ls] = a[1];
    }
}

// ---------------------------------------------------------------------------
// K2: one thread per (pair, k).  Writes the window moments Pw/Qw[p][j][k]
// (zero-padded to k_s slots), the prefix sums PP[p][j][k] = sum_{i<=j} Pw,
// the suffix sums QS[p][j][k] = sum_{i>=j} Qw and the totals.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k2_pair_moments(int n_pairs, int K, int ks, int kgl,
                                                       const PairInfo* __restrict__ pairs,
                                                       const double* __restrict__ bp, const double* __restrict__ w,
                                                       const double* __restrict__ Bo, const double* __restrict__ ro,
                                                       double* __restrict__ Pw, double* __restrict__ Qw,
                                                       double* __restrict__ PP, double* __restrict__ QS,
                                                       double* __restrict__ Ptot, double* __restrict__ Qtot)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t)n_pairs * K) return;
    const int p = (int)(gid / K), k = (int)(gid - (int64_t)p * K);
    const PairInfo pi = pairs[p];
    double pw[kMaxOrder], qw[kMaxOrder];
    for (int j = 0; j < ks; ++j) {
        pw[j] = 0.0;
        qw[j] = 0.0;
    }
    for (int j = 0; j < pi.m; ++j) {
        const int v = pi.v0 - 1 + j;                 // 0-based interval
        const int ta = pi.a - v, tc = pi.c - v;      // local spline index s - iv
        const double scale = 0.5 * (bp[v + 1] - bp[v]);
        const double* b = Bo + (size_t)v * kgl * ks;
        const double* rr = ro + (size_t)v * kgl;
        double accp = 0.0, accq = 0.0;
        for (int q = 0; q < kgl; ++q) {
            const double bb = w[q] * b[q * ks + ta] * b[q * ks + tc];
            const double r = rr[q];
            accp = accp + bb * ipow_d(r, k);
            accq = accq + bb / ipow_d(r, k + 1);
        }
        pw[j] = scale * accp;
        qw[j] = scale * accq;
    }
    const size_t base = (size_t)p * ks * K + k;
    double run = 0.0;
    for (int j = 0; j < ks; ++j) {
        run += pw[j];
        Pw[base + (size_t)j * K] = pw[j];
        Qw[base + (size_t)j * K] = qw[j];
        PP[base + (size_t)j * K] = (j < pi.m) ? run : 0.0;
    }
    Ptot[(size_t)p * K + k] = run;
    run = 0.0;
    for (int j = ks - 1; j >= 0; --j) {
        if (j < pi.m) run += qw[j];
        QS[base + (size_t)j * K] = (j < pi.m) ? run : 0.0;
    }
    Qtot[(size_t)p * K + k] = run;
}

// ---------------------------------------------------------------------------
// K4: one workgroup per (interval v, multipole k).
//   rk_tab[q][p] = w_p rho_qp^k
//   J[q][u]  = s_q * sum_p rk_tab[q][p] B_tb(rho_qp) B_td(rho_qp)   (inner pair u)
//   O[q][u]  = w_q B_ta(r_q) B_tc(r_q) / r_q^(k+1)                    (outer pair u)
//   D[v][k][uo][ui] = h_v * sum_q O[q][uo] J[q][ui]
// Local pairs are the U = k(k+1)/2 unordered pairs of the interval's k splines
// (B_a B_c is symmetric in (a, c)).
// ---------------------------------------------------------------------------
constexpr int kK4Threads = 256;

__global__ __launch_bounds__(kK4Threads) void k4_diag_cells(int n_iv, int K, int ks, int kgl, int U,
                                                            const double* __restrict__ bp,
                                                            const double* __restrict__ x, const double* __restrict__ w,
                                                            const double* __restrict__ Bo,
                                                            const double* __restrict__ Bi,
                                                            const double* __restrict__ ro, double* __restrict__ D)
{
    extern __shared__ double lds[];
    const int v = blockIdx.x / K;
    const int k = blockIdx.x - v * K;
    const int tid = threadIdx.x;
    double* sO = lds;                          // [kgl][U]
    double* sJ = sO + kgl * U;                 // [kgl][U]
    double* sRK = sJ + kgl * U;                // [kgl][kgl]
    int* sTa = (int*)(sRK + kgl * kgl);        // [U]
    int* sTc = sTa + U;                        // [U]

    const double lim1 = bp[v], lim2 = bp[v + 1];
    const double scale_i = 0.5 * (lim2 - lim1);
    for (int u = tid; u < U; u += kK4Threads) {
        // invert the symmetric local index (row-major upper triangle)
        int lo = 0, base = 0;
        while (base + (ks - lo) <= u) {
            base += ks - lo;
            ++lo;
        }
        sTa[u] = lo;
        sTc[u] = lo + (u - base);
    }
    const double* rr = ro + (size_t)v * kgl;
    for (int e = tid; e < kgl * kgl; e += kK4Threads) {
        const int q = e / kgl, p = e - q * kgl;
        const double rq = rr[q];
        const double scale_j = 0.5 * (rq - lim1);
        const double translate_j = 0.5 * (rq + lim1);
        const double rho = scale_j * x[p] + translate_j;
        sRK[e] = w[p] * ipow_d(rho, k);
    }
    __syncthreads();
    const double* bo = Bo + (size_t)v * kgl * ks;
    const double* bi = Bi + (size_t)v * kgl * kgl * ks;
    for (int e = tid; e < kgl * U; e += kK4Threads) {
        const int q = e / U, u = e - q * U;
        const int ta = sTa[u], tc = sTc[u];
        const double rq = rr[q];
        sO[e] = w[q] * bo[q * ks + ta] * bo[q * ks + tc] / ipow_d(rq, k + 1);
        const double scale_j = 0.5 * (rq - lim1);
        const double* bq = bi + (size_t)q * kgl * ks;
        double acc = 0.0;
        for (int p = 0; p < kgl; ++p) acc = acc + sRK[q * kgl + p] * bq[p * ks + ta] * bq[p * ks + tc];
        sJ[e] = scale_j * acc;
    }
    __syncthreads();
    double* Dv = D + ((size_t)v * K + k) * U * U;
    for (int e = tid; e < U * U; e += kK4Threads) {
        const int uo = e / U, ui = e - uo * U;
        double acc = 0.0;
        for (int q = 0; q < kgl; ++q) acc = acc + sO[q * U + uo] * sJ[q * U + ui];
        Dv[e] = scale_i * acc;
    }
}

static inline unsigned blocks_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

hipError_t launch_bspline_tables(const Plan& p, const DevTables& d, hipStream_t s)
{
    const int64_t n = p.n_iv * p.k_gl * (1 + p.k_gl);
    hipLaunchKernelGGL(k1_bspline_tables, dim3(blocks_for(n, 256)), dim3(256), 0, s, (int)p.n_iv, (int)p.k_s,
                       (int)p.k_gl, d.t, d.bp, d.x, d.Bo, d.Bi, d.ro);
    return hipGetLastError();
}

hipError_t launch_pair_moments(const Plan& p, const DevTables& d, hipStream_t s)
{
    const int64_t n = p.n_pairs * p.K;
    hipLaunchKernelGGL(k2_pair_moments, dim3(blocks_for(n, 256)), dim3(256), 0, s, (int)p.n_pairs, (int)p.K,
                       (int)p.k_s, (int)p.k_gl, d.pairs, d.bp, d.w, d.Bo, d.ro, d.Pw, d.Qw, d.PP, d.QS, d.Ptot,
                       d.Qtot);
    return hipGetLastError();
}

size_t diag_cells_lds_bytes(const Plan& p)
{
    return sizeof(double) * (size_t)(2 * p.k_gl * p.U + p.k_gl * p.k_gl) + sizeof(int) * (size_t)(2 * p.U);
}

hipError_t launch_diag_cells(const Plan& p, const DevTables& d, hipStream_t s)
{
    const size_t lds = diag_cells_lds_bytes(p);
    hipLaunchKernelGGL(k4_diag_cells, dim3((unsigned)(p.n_iv * p.K)), dim3(kK4Threads), lds, s, (int)p.n_iv,
                       (int)p.K, (int)p.k_s, (int)p.k_gl, (int)p.U, d.bp, d.x, d.w, d.Bo, d.Bi, d.ro, d.D);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// K2b: window prefix/suffix sums and totals from Pw/Qw already in place (used
// when the moments are gathered from caller-provided containers).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k2b_window_sums(int n_pairs, int K, int ks, const PairInfo* __restrict__ pairs,
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

hipError_t launch_window_sums(const Plan& p, const DevTables& d, hipStream_t s)
{
    const int64_t n = p.n_pairs * p.K;
    hipLaunchKernelGGL(k2b_window_sums, dim3(blocks_for(n, 256)), dim3(256), 0, s, (int)p.n_pairs, (int)p.K,
                       (int)p.k_s, d.pairs, d.Pw, d.Qw, d.PP, d.QS, d.Ptot, d.Qtot);
    return hipGetLastError();
}

}  // namespace bs2e
