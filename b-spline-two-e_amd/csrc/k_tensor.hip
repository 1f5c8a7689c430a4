// k_tensor.hip -- K3: the banded R^k tensor (sparse_array_tools.f90:452-493).
//
// For banded pairs p = (a,c) (electron 1) and q = (b,d) (electron 2) with
// support windows W_p, W_q (intervals), the reference accumulates
//   R^k(a,b,c,d) = sum_{v<w} P^k[v;p] Q^k[w;q] + sum_{v>w} Q^k[v;p] P^k[w;q]
//                + sum_{v in W_p and W_q} ( D^k[v;p,q] + D^k[v;q,p] )
// over the nnz4^2 off-diagonal products and the nnz6 diagonal cells
// (sparse_array_tools.f90:472-488).  Because the windows are contiguous the
// first two sums collapse:
//   * W_p entirely left of W_q :  R = Ptot[p] * Qtot[q]            (rank one)
//   * W_p entirely right of W_q:  R = Qtot[p] * Ptot[q]            (rank one)
//   * otherwise (windows meet)  :  R = sum_{v in W_p} Pw[p](v) QS_q(v+1)
//                                     + Qw[p](v) PP_q(v-1) + diagonal cells
// with QS_q / PP_q the suffix / prefix sums of Q / P over W_q.  The kernel is
// bound by the FP64 store of R (8 bytes per integral): each workgroup streams
// ROWS rows of the tensor over one column chunk, reading the per-column totals
// once for all its rows and writing with non-temporal stores.
#include "bs2e_internal.h"

namespace bs2e {

constexpr int kK3Threads = 256;
constexpr int kK3Rows = 8;        // tensor rows per workgroup
constexpr int kK3Iters = 8;       // elements per thread per row

__device__ __noinline__ double near_element(int p, int q, int k, int K, int ks, int U,
                                            const PairInfo* __restrict__ pairs, const double* __restrict__ Pw,
                                            const double* __restrict__ Qw, const double* __restrict__ PP,
                                            const double* __restrict__ QS, double ptot_q, double qtot_q,
                                            const double* __restrict__ D)
{
    const PairInfo pp = pairs[p], pq = pairs[q];
    const int v1p = pp.v0 + pp.m - 1, v1q = pq.v0 + pq.m - 1;
    const size_t pbase = (size_t)p * ks * K + k;
    const size_t qbase = (size_t)q * ks * K + k;
    double acc = 0.0;
    for (int j = 0; j < pp.m; ++j) {
        const int v = pp.v0 + j;
        const int xs = v + 1;   // Q of q summed over w >= v+1
        const double qs = (xs <= pq.v0) ? qtot_q : ((xs > v1q) ? 0.0 : QS[qbase + (size_t)(xs - pq.v0) * K]);
        const int xp = v - 1;   // P of q summed over w <= v-1
        const double pv = (xp < pq.v0) ? 0.0 : ((xp >= v1q) ? ptot_q : PP[qbase + (size_t)(xp - pq.v0) * K]);
        acc += Pw[pbase + (size_t)j * K] * qs + Qw[pbase + (size_t)j * K] * pv;
    }
    const int lo = pp.v0 > pq.v0 ? pp.v0 : pq.v0;
    const int hi = v1p < v1q ? v1p : v1q;
    for (int v = lo; v <= hi; ++v) {
        const int vz = v - 1;
        const int up = sym_local(pp.a - vz, pp.c - vz, ks);
        const int uq = sym_local(pq.a - vz, pq.c - vz, ks);
        const double* Dv = D + ((size_t)vz * K + k) * (size_t)U * U;
        acc += Dv[up * U + uq] + Dv[uq * U + up];
    }
    return acc;
}

__global__ __launch_bounds__(kK3Threads) void k3_tensor_fill(int n_pairs, int K, int ks, int U,
                                                             const PairInfo* __restrict__ pairs,
                                                             const double* __restrict__ Pw,
                                                             const double* __restrict__ Qw,
                                                             const double* __restrict__ PP,
                                                             const double* __restrict__ QS,
                                                             const double* __restrict__ Ptot,
                                                             const double* __restrict__ Qtot,
                                                             const double* __restrict__ D, double* __restrict__ R)
{
    __shared__ double sP[kK3Rows * kMaxK];
    __shared__ double sQ[kK3Rows * kMaxK];
    __shared__ int sV0[kK3Rows], sV1[kK3Rows];

    const int64_t rowlen = (int64_t)n_pairs * K;
    const int p0 = blockIdx.y * kK3Rows;
    const int nrows = (n_pairs - p0) < kK3Rows ? (n_pairs - p0) : kK3Rows;
    const int tid = threadIdx.x;
    for (int e = tid; e < nrows * K; e += kK3Threads) {
        const int r = e / K, k = e - r * K;
        sP[r * kMaxK + k] = Ptot[(size_t)(p0 + r) * K + k];
        sQ[r * kMaxK + k] = Qtot[(size_t)(p0 + r) * K + k];
    }
    if (tid < nrows) {
        const PairInfo pi = pairs[p0 + tid];
        sV0[tid] = pi.v0;
        sV1[tid] = pi.v0 + pi.m - 1;
    }
    __syncthreads();

    const int64_t chunk = (int64_t)kK3Threads * kK3Iters;
    int64_t e = (int64_t)blockIdx.x * chunk + tid;
    int q = (int)(e / K);
    int k = (int)(e - (int64_t)q * K);
    const int dq = kK3Threads / K, dk = kK3Threads - (kK3Threads / K) * K;
    double* Rrow = R + (size_t)p0 * rowlen;
    for (int it = 0; it < kK3Iters; ++it) {
        if (e >= rowlen) break;
        const double pt = Ptot[e];
        const double qt = Qtot[e];
        const PairInfo pq = pairs[q];
        const int v0q = pq.v0, v1q = pq.v0 + pq.m - 1;
        for (int r = 0; r < nrows; ++r) {
            double val;
            if (sV1[r] < v0q) {
                val = sP[r * kMaxK + k] * qt;
            } else if (sV0[r] > v1q) {
                val = sQ[r * kMaxK + k] * pt;
            } else {
                val = near_element(p0 + r, q, k, K, ks, U, pairs, Pw, Qw, PP, QS, pt, qt, D);
            }
            __builtin_nontemporal_store(val, Rrow + (size_t)r * rowlen + e);
        }
        e += kK3Threads;
        q += dq;
        k += dk;
        if (k >= K) {
            k -= K;
            q += 1;
        }
    }
}

hipError_t launch_tensor_fill(const Plan& p, const DevTables& d, double* R, hipStream_t s)
{
    const int64_t rowlen = p.n_pairs * p.K;
    const int64_t chunk = (int64_t)kK3Threads * kK3Iters;
    dim3 grid((unsigned)((rowlen + chunk - 1) / chunk), (unsigned)((p.n_pairs + kK3Rows - 1) / kK3Rows));
    hipLaunchKernelGGL(k3_tensor_fill, grid, dim3(kK3Threads), 0, s, (int)p.n_pairs, (int)p.K, (int)p.k_s,
                       (int)p.U, d.pairs, d.Pw, d.Qw, d.PP, d.QS, d.Ptot, d.Qtot, d.D, R);
    return hipGetLastError();
}

}  // namespace bs2e
