// k_tensor.hip -- the banded R^k tensor (sparse_array_tools.f90:452-493).
//
// For banded pairs p = (a,c) (electron 1) and q = (b,d) (electron 2) with
// support windows W_p, W_q (intervals), the reference accumulates
//   R^k(a,b,c,d) = sum_{v<w} P^k[v;p] Q^k[w;q] + sum_{v>w} Q^k[v;p] P^k[w;q]
//                + sum_{v in W_p and W_q} ( D^k[v;p,q] + D^k[v;q,p] )
// over the nnz4^2 off-diagonal products and the nnz6 diagonal cells
// (sparse_array_tools.f90:472-488).  Because the windows are contiguous:
//   * W_p entirely left of W_q :  R = Ptot[p] * Qtot[q]            (rank one)
//   * W_p entirely right of W_q:  R = Qtot[p] * Ptot[q]            (rank one)
//   * in general                :  R = sum_{w in W_q} Qw_q(w) PP_p(w-1)
//                                     + Pw_q(w) QS_p(w+1) + sum_v Dsym[v;p,q]
// with PP_p / QS_p the prefix / suffix sums of P / Q over W_p and
// Dsym[v] = D[v] + D[v]^T the symmetrised diagonal cell.
//
// Two kernels write disjoint elements, each element exactly once:
//   * k3_far : a few rows x one column chunk of rank-one products (the pairs'
//     windows do not overlap), streamed with non-temporal stores.  Bound by
//     the FP64 store of R: 8 bytes per integral.
//   * k3_near: the elements whose pairs share a support interval (a few
//     percent of each row).  Rows p = (a,c) and (c,a) are identical
//     (R(a,b,c,d) = R(c,b,a,d)), so one workgroup evaluates row (a,c), a <= c,
//     and stores it into both rows.  With W_p = [lo, hi], for column q:
//       R = Qtot_p * sum_{w in W_q, w < lo} P_q(w)
//         + Ptot_p * sum_{w in W_q, w > hi} Q_q(w)
//         + sum_{w in W_p and W_q} Q_q(w) PP_p(w-1) + P_q(w) QS_p(w+1) + Dsym[w; p, q]
//     (the general formula with PP_p = 0 left of W_p and Ptot_p right of it,
//     QS_p = Qtot_p left of W_p and 0 right of it); the first two sums are
//     q's own prefix / suffix sums.  It runs on a second stream, concurrently
//     with k3_far.
#include "bs2e_internal.h"

namespace bs2e {

constexpr int kThreads = 256;

// LDS of a near workgroup (row p, window length m):
//   PPx[i][k], i in [0, m]: 0, PP[0], ..., PP[m-1] (= Ptot)
//   QSx[i][k], i in [0, m]: QS[0] (= Qtot), ..., QS[m-1], 0
// so that PP_p(w-1) = PPx[w - lo] and QS_p(w+1) = QSx[w - lo + 1] for w in
// W_p; the local pair index of p on each interval of W_p; and the row's
// column segments (one run of consecutive pairs per column spline d).
size_t near_lds_bytes(int64_t ks, int64_t K)
{
    return sizeof(double) * (size_t)(2 * (ks + 1) * K) + sizeof(int) * (size_t)(kMaxOrder + 2 * kNearMaxSeg + 4);
}

__device__ __forceinline__ int pair_at(const int32_t* __restrict__ c_first, int ks, int a, int c)
{
    const int lo = (c - ks + 1) < 1 ? 1 : (c - ks + 1);
    return c_first[c] + (a - lo);
}

__global__ __launch_bounds__(kThreads) void k3_near(int n_pairs, int n_b, int K, int ks, int U,
                                                    const PairInfo* __restrict__ pairs,
                                                    const int32_t* __restrict__ c_first,
                                                    const int32_t* __restrict__ rows,
                                                    const double* __restrict__ Pw, const double* __restrict__ Qw,
                                                    const double* __restrict__ PP, const double* __restrict__ QS,
                                                    const double* __restrict__ Dsym, double* __restrict__ R)
{
    extern __shared__ double lds[];
    double* sPPx = lds;                                    // [(ks+1)][K]
    double* sQSx = lds + (ks + 1) * K;                     // [(ks+1)][K]
    int* sUp = (int*)(lds + 2 * (ks + 1) * K);             // [kMaxOrder]
    int* sQ0 = sUp + kMaxOrder;                            // [kNearMaxSeg]: first pair of each segment
    int* sJ0 = sQ0 + kNearMaxSeg;                          // [kNearMaxSeg + 1]: segment starts in row order
    int* sNseg = sJ0 + kNearMaxSeg + 1;
    const int tid = threadIdx.x;
    const int p = rows[xcd_block(blockIdx.x, gridDim.x)];   // neighbouring rows gather the same Dsym cells
    const PairInfo pp = pairs[p];
    const int lo = pp.v0, m = pp.m, hi = pp.v0 + pp.m - 1;
    const int pt = (pp.a == pp.c) ? -1 : pair_at(c_first, ks, pp.c, pp.a);   // the identical row (c, a)
    for (int e = tid; e < (ks + 1) * K; e += kThreads) {
        const int i = e / K, k = e - i * K;
        const size_t pb = (size_t)p * ks * K + k;
        sPPx[e] = (i == 0) ? 0.0 : PP[pb + (size_t)(i - 1) * K];
        sQSx[e] = (i >= m) ? 0.0 : QS[pb + (size_t)i * K];
    }
    if (tid < m) {
        const int vz = lo - 1 + tid;
        sUp[tid] = sym_local(pp.a - vz, pp.c - vz, ks);
    }
    if (tid == 0) {
        // column pairs (b, d) share an interval with W_p iff
        // max(b,d) + 2 - k <= hi and min(b,d) + 1 >= lo
        const int dlo = max(1, lo - 1), dhi = min(n_b, hi + ks - 2);
        int ns = 0, J = 0;
        for (int d = dlo; d <= dhi; ++d) {
            const int b0 = max(1, d - ks + 1);
            const int blo = max(b0, lo - 1), bhi = min(min(d + ks - 1, n_b), hi + ks - 2);
            if (blo > bhi) continue;
            sQ0[ns] = c_first[d] + (blo - b0);
            sJ0[ns] = J;
            J += bhi - blo + 1;
            ++ns;
        }
        sJ0[ns] = J;
        *sNseg = ns;
    }
    __syncthreads();
    const int ns = *sNseg;
    const int total = sJ0[ns] * K;
    const int64_t rowlen = (int64_t)n_pairs * K;
    double* Rp = R + (size_t)p * rowlen;
    double* Rt = R + (size_t)(pt < 0 ? p : pt) * rowlen;
    const unsigned UK = (unsigned)(U * K);   // Dsym has n_iv*U*U*K < 2^31 entries (checked on the host)
    for (int e = tid; e < total; e += kThreads) {
        const int j = e / K, k = e - j * K;
        int s = 0;                       // last segment starting at or before j
        for (int step = kNearMaxSeg / 2; step > 0; step >>= 1)
            if (s + step < ns && sJ0[s + step] <= j) s += step;
        const int q = sQ0[s] + (j - sJ0[s]);
        const PairInfo pq = pairs[q];
        const int v0q = pq.v0, endq = pq.v0 + pq.m - 1;
        const size_t qb = (size_t)q * ks * K + k;
        double val = 0.0;
        if (v0q < lo) val = sQSx[k] * PP[qb + (size_t)(lo - 1 - v0q) * K];
        if (endq > hi) val += sPPx[m * K + k] * QS[qb + (size_t)(hi + 1 - v0q) * K];
        const int wlo = max(lo, v0q), whi = min(hi, endq);
        double acc = 0.0, dacc = 0.0;
#pragma unroll 2
        for (int w = wlo; w <= whi; ++w) {
            const int jq = w - v0q, jp = w - lo, vz = w - 1;
            acc += Qw[qb + (size_t)jq * K] * sPPx[jp * K + k] + Pw[qb + (size_t)jq * K] * sQSx[(jp + 1) * K + k];
            const unsigned uq = (unsigned)sym_local(pq.a - vz, pq.c - vz, ks), up = (unsigned)sUp[jp];
            const unsigned ulo = up < uq ? up : uq, uhi = up < uq ? uq : up;   // Dsym is read in its upper triangle
            dacc += Dsym[(unsigned)vz * U * UK + ulo * UK + uhi * (unsigned)K + (unsigned)k];
        }
        val += acc + dacc;
        const size_t o = (size_t)q * K + k;
        Rp[o] = val;
        if (pt >= 0) Rt[o] = val;
    }
}

// Far workgroups: ROWS rows x one column chunk of rank-one products; the
// elements whose pairs' windows overlap are left to k3_near.  V = 2: each
// thread stores two consecutive elements with one 16-byte store (rows of even
// length, so every row start is 16-byte aligned); V = 1: one element.
typedef double double2_t __attribute__((ext_vector_type(2)));

template <int V, int ROWS, int ITERS>
__global__ __launch_bounds__(kThreads) void k3_far(int n_pairs, int K, int far_cols, int x_cols,
                                                   const PairInfo* __restrict__ pairs,
                                                   const double* __restrict__ Ptot,
                                                   const double* __restrict__ Qtot, double* __restrict__ R)
{
    __shared__ double sA[ROWS * kMaxK];   // Ptot of the rows
    __shared__ double sB[ROWS * kMaxK];   // Qtot of the rows
    __shared__ int sLo[ROWS], sHi[ROWS];
    const int64_t rowlen = (int64_t)n_pairs * K;
    const int tid = threadIdx.x;
    // XCD x owns the x-th eighth of the column chunks (x_cols of them) and
    // walks it row group by row group: its L2 keeps that slice of the column
    // factors while the stores stream past, and the XCDs' stores in flight
    // spread over eight regions of a row.  Measured in one process against
    // the row-major order and against chunks x, x + 8, ... per XCD: c5p
    // 3.45-3.48 ms per build against 3.71-3.76 (k3_far<2,2,2>), c2, c3, c4
    // within 2 %.
    const int x = (int)(blockIdx.x % kXcds), j = (int)(blockIdx.x / kXcds);
    const int ry = j / x_cols, cx = x * x_cols + (j - ry * x_cols);
    if (cx >= far_cols) return;   // (the last XCD's slice can be short; uniform over the workgroup)
    const int p0 = ry * ROWS;
    const int nrows = (n_pairs - p0) < ROWS ? (n_pairs - p0) : ROWS;
    for (int e = tid; e < nrows * K; e += kThreads) {
        const int r = e / K, k = e - r * K;
        sA[r * kMaxK + k] = Ptot[(size_t)(p0 + r) * K + k];
        sB[r * kMaxK + k] = Qtot[(size_t)(p0 + r) * K + k];
    }
    if (tid < nrows) {
        const PairInfo pi = pairs[p0 + tid];
        sLo[tid] = pi.v0;
        sHi[tid] = pi.v0 + pi.m - 1;   // last interval of the row's window
    }
    __syncthreads();
    double* Rrow = R + (size_t)p0 * rowlen;
    const unsigned uK = (unsigned)K, urow = (unsigned)rowlen;   // rowlen = n_pairs * K < 2^31 (checked on the host)
    for (int i = 0; i < ITERS; ++i) {
        const unsigned e = (unsigned)V * ((unsigned)cx * kThreads * ITERS + (unsigned)i * kThreads + (unsigned)tid);
        if (e >= urow) break;
        const int q0 = (int)(e / uK), k0 = (int)(e - (unsigned)q0 * uK);
        const PairInfo a0 = pairs[q0];
        const int v1a = a0.v0 + a0.m - 1;
        if (V == 1) {
            const double pt = Ptot[e], qt = Qtot[e];
            for (int r = 0; r < nrows; ++r) {
                const bool left = sHi[r] < a0.v0, right = sLo[r] > v1a;
                if (!left && !right) continue;   // windows overlap: written by k3_near
                const double val = left ? sA[r * kMaxK + k0] * qt : sB[r * kMaxK + k0] * pt;
                __builtin_nontemporal_store(val, Rrow + (size_t)r * rowlen + e);
            }
        } else {
            const int q1 = (k0 + 1 == K) ? q0 + 1 : q0, k1 = (k0 + 1 == K) ? 0 : k0 + 1;
            const double2_t pt = *(const double2_t*)(Ptot + e), qt = *(const double2_t*)(Qtot + e);
            const PairInfo a1 = pairs[q1];
            const int v1b = a1.v0 + a1.m - 1;
            for (int r = 0; r < nrows; ++r) {
                const bool l0 = sHi[r] < a0.v0, f0 = l0 || sLo[r] > v1a;
                const bool l1 = sHi[r] < a1.v0, f1 = l1 || sLo[r] > v1b;
                double2_t val;
                val.x = l0 ? sA[r * kMaxK + k0] * qt.x : sB[r * kMaxK + k0] * pt.x;
                val.y = l1 ? sA[r * kMaxK + k1] * qt.y : sB[r * kMaxK + k1] * pt.y;
                double* dst = Rrow + (size_t)r * rowlen + e;
                if (f0 && f1) {
                    __builtin_nontemporal_store(val, (double2_t*)dst);
                } else {
                    if (f0) __builtin_nontemporal_store(val.x, dst);
                    if (f1) __builtin_nontemporal_store(val.y, dst + 1);
                }
            }
        }
    }
}

// The store ceiling the rank-one fill is measured against: zeros over n16
// 16-byte elements, one store per thread (NT: with the non-temporal hint),
// the write pattern that reached the highest rate among those timed (more
// stores per thread, grid-stride loops and 64 contiguous bytes per thread
// were slower).
template <bool NT>
__global__ __launch_bounds__(kThreads) void k_store_fill(double2_t* __restrict__ dst, int64_t n16)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n16) return;
    const double2_t z = {0.0, 0.0};
    if (NT) __builtin_nontemporal_store(z, dst + i);
    else dst[i] = z;
}

hipError_t launch_store_fill(void* dst, int64_t n16, bool nt, hipStream_t s)
{
    const unsigned grid = (unsigned)((n16 + kThreads - 1) / kThreads);
    if (nt) hipLaunchKernelGGL((k_store_fill<true>), dim3(grid), dim3(kThreads), 0, s, (double2_t*)dst, n16);
    else hipLaunchKernelGGL((k_store_fill<false>), dim3(grid), dim3(kThreads), 0, s, (double2_t*)dst, n16);
    return hipGetLastError();
}

// Dsym[v][u1][u2][k] = D[v][u1][u2][k] + D[v][u2][u1][k], in place (one
// thread per unordered local pair pair and multipole).
__global__ __launch_bounds__(256) void k_symmetrize(int n_iv, int U, int K, double* __restrict__ D)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per_v = (int64_t)U * U * K;
    if (gid >= (int64_t)n_iv * per_v) return;
    const int v = (int)(gid / per_v);
    const int64_t rem = gid - (int64_t)v * per_v;
    const int u1 = (int)(rem / ((int64_t)U * K));
    const int u2 = (int)((rem / K) % U);
    const int k = (int)(rem % K);
    if (u2 < u1) return;
    double* a = D + (((size_t)v * U + u1) * U + u2) * K + k;
    double* b = D + (((size_t)v * U + u2) * U + u1) * K + k;
    const double s = *a + *b;
    *a = s;
    *b = s;
}

// Host: the near kernel's rows (pairs (a, c) with a <= c).
void build_near_rows(Plan& p)
{
    p.rows.clear();
    for (int64_t P = 0; P < p.n_pairs; ++P)
        if (p.pairs[P].a <= p.pairs[P].c) p.rows.push_back((int32_t)P);
}

// Host: the number of tensor elements whose pairs' windows overlap.
int64_t count_near_elements(const Plan& p)
{
    int64_t n = 0;
    const int64_t ks = p.k_s, nb = p.n_b;
    for (int64_t P = 0; P < p.n_pairs; ++P) {
        const PairInfo& pp = p.pairs[P];
        const int64_t v1p = pp.v0 + pp.m - 1;
        const int64_t cmin = std::max<int64_t>(1, pp.v0 - 1);
        const int64_t cmax = std::min<int64_t>(nb, v1p + ks - 2);
        for (int64_t Q = p.c_first[cmin]; Q < p.c_first[cmax + 1]; ++Q) {
            const PairInfo& pq = p.pairs[Q];
            if (pq.v0 <= v1p && pp.v0 <= pq.v0 + pq.m - 1) n += p.K;
        }
    }
    return n;
}

hipError_t launch_tensor_near(const Plan& p, const DevTables& d, double* R, hipStream_t s)
{
    if (p.rows.empty()) return hipSuccess;
    hipLaunchKernelGGL(k3_near, dim3((unsigned)p.rows.size()), dim3(kThreads), near_lds_bytes(p.k_s, p.K), s,
                       (int)p.n_pairs, (int)p.n_b, (int)p.K, (int)p.k_s, (int)p.U, d.pairs, d.c_first, d.rows, d.Pw,
                       d.Qw, d.PP, d.QS, d.D, R);
    return hipGetLastError();
}

// Shapes of the rank-one fill the autotuner chooses from: elements per
// store (V), rows and element groups per workgroup.
struct FarShape {
    int V, rows, iters;
};
static constexpr FarShape kFarShapes[] = {{2, 2, 1}, {2, 2, 2}, {2, 4, 4}, {2, 2, 8}, {2, 1, 4}, {1, 4, 8}, {1, 2, 4}};

int far_variant_count() { return (int)(sizeof(kFarShapes) / sizeof(kFarShapes[0])); }

bool far_variant_valid(const Plan& p, int v)
{
    return v >= 0 && v < far_variant_count() && (kFarShapes[v].V == 1 || (p.n_pairs * p.K) % 2 == 0);
}

// 2 rows x 512 elements, one 16-byte store per thread and row: a short
// workgroup life keeps more stores in flight than longer loops (5.7 TB/s
// against 5.5 for 2 x 1024 at 0.76 GB); rows of odd length fall back to
// 8-byte stores.
int far_variant_default(const Plan& p) { return (p.n_pairs * p.K) % 2 == 0 ? 0 : 5; }

template <int V, int ROWS, int ITERS>
static void launch_far_shape(const Plan& p, const DevTables& d, double* R, hipStream_t s)
{
    const int64_t rowlen = p.n_pairs * p.K;
    const int64_t per = (int64_t)V * kThreads * ITERS;
    const int far_cols = (int)((rowlen + per - 1) / per);
    const int far_rows = (int)((p.n_pairs + ROWS - 1) / ROWS);
    const int x_cols = (far_cols + kXcds - 1) / kXcds;   // column chunks per XCD
    hipLaunchKernelGGL((k3_far<V, ROWS, ITERS>), dim3((unsigned)(kXcds * x_cols * far_rows)), dim3(kThreads), 0, s,
                       (int)p.n_pairs, (int)p.K, far_cols, x_cols, d.pairs, d.Ptot, d.Qtot, R);
}

hipError_t launch_tensor_far(const Plan& p, const DevTables& d, double* R, hipStream_t s)
{
    const int v = far_variant_valid(p, p.far_variant) ? p.far_variant : far_variant_default(p);
    switch (v) {
    case 0: launch_far_shape<2, 2, 1>(p, d, R, s); break;
    case 1: launch_far_shape<2, 2, 2>(p, d, R, s); break;
    case 2: launch_far_shape<2, 4, 4>(p, d, R, s); break;
    case 3: launch_far_shape<2, 2, 8>(p, d, R, s); break;
    case 4: launch_far_shape<2, 1, 4>(p, d, R, s); break;
    case 5: launch_far_shape<1, 4, 8>(p, d, R, s); break;
    default: launch_far_shape<1, 2, 4>(p, d, R, s); break;
    }
    return hipGetLastError();
}

hipError_t launch_symmetrize(const Plan& p, const DevTables& d, hipStream_t s)
{
    const int64_t n = p.n_iv * p.U * p.U * p.K;
    hipLaunchKernelGGL(k_symmetrize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (int)p.n_iv, (int)p.U,
                       (int)p.K, d.D);
    return hipGetLastError();
}

}  // namespace bs2e
