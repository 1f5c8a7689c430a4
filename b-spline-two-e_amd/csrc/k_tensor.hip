// k_tensor.hip -- K3: the banded R^k tensor (sparse_array_tools.f90:452-493).
//
// For banded pairs p = (a,c) (electron 1) and q = (b,d) (electron 2) with
// support windows W_p, W_q (intervals), the reference accumulates
//   R^k(a,b,c,d) = sum_{v<w} P^k[v;p] Q^k[w;q] + sum_{v>w} Q^k[v;p] P^k[w;q]
//                + sum_{v in W_p and W_q} ( D^k[v;p,q] + D^k[v;q,p] )
// over the nnz4^2 off-diagonal products and the nnz6 diagonal cells
// (sparse_array_tools.f90:472-488).  Because the windows are contiguous:
//   * W_p entirely left of W_q :  R = Ptot[p] * Qtot[q]            (rank one)
//   * W_p entirely right of W_q:  R = Qtot[p] * Ptot[q]            (rank one)
//   * windows overlap           :  R = sum_{w in W_q} Qw_q(w) PP_p(w-1)
//                                     + Pw_q(w) QS_p(w+1) + diagonal cells
// with PP_p / QS_p the prefix / suffix sums of P / Q over W_p.
//
// For every row p the q whose windows overlap W_p form the index range
// [qlo(p), qhi(p)] (the "near" range, a few percent of the row).  Two kernels
// write disjoint elements, so every element is written exactly once:
//   * k3_far: kRows rows x one column chunk of rank-one products, streamed
//     with non-temporal stores, skipping each row's near range.  It is bound
//     by the FP64 store of R: 8 bytes per integral.
//   * k3_near: one chunk of one row's near range; the row's prefix/suffix
//     sums sit in LDS, each lane loads its column pair's whole window at once
//     (unrolled, independent loads) and adds the diagonal cells of the
//     common intervals.  It runs on a second stream, concurrently with k3_far.
#include "bs2e_internal.h"

namespace bs2e {

constexpr int kThreads = 256;
constexpr int kRows = 4;          // far workgroup: tensor rows
constexpr int kFarIters = 8;      // far workgroup: elements per thread per row
constexpr int kNearIters = 4;     // near workgroup: elements per thread

// Near element (p, q, k) in the symmetric form; KSMAX >= k_s bounds the
// unrolled window loops (lanes predicate on the actual window length).
template <int KSMAX>
__device__ __forceinline__ double near_value(int q, int k, int K, int ks, int U, const PairInfo& pp,
                                             const PairInfo* __restrict__ pairs, const double* __restrict__ Pw,
                                             const double* __restrict__ Qw, const double* __restrict__ sPP,
                                             const double* __restrict__ sQS, double ptot_p, double qtot_p,
                                             const double* __restrict__ D)
{
    const PairInfo pq = pairs[q];
    const size_t qbase = (size_t)q * ks * K + k;
    double pw[KSMAX], qw[KSMAX];
#pragma unroll
    for (int j = 0; j < KSMAX; ++j) {
        const bool in = j < pq.m;
        pw[j] = in ? Pw[qbase + (size_t)j * K] : 0.0;
        qw[j] = in ? Qw[qbase + (size_t)j * K] : 0.0;
    }
    const int v1p = pp.v0 + pp.m - 1, v1q = pq.v0 + pq.m - 1;
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < KSMAX; ++j) {
        if (j < pq.m) {
            const int w = pq.v0 + j;
            const int xp = w - 1;   // P of p summed over v <= w-1
            const double pp_v = (xp < pp.v0) ? 0.0 : ((xp >= v1p) ? ptot_p : sPP[(xp - pp.v0) * K + k]);
            const int xs = w + 1;   // Q of p summed over v >= w+1
            const double qs_v = (xs <= pp.v0) ? qtot_p : ((xs > v1p) ? 0.0 : sQS[(xs - pp.v0) * K + k]);
            acc += qw[j] * pp_v + pw[j] * qs_v;
        }
    }
    const int lo = pp.v0 > pq.v0 ? pp.v0 : pq.v0;
    const int hi = v1p < v1q ? v1p : v1q;
    double dacc = 0.0;
#pragma unroll
    for (int j = 0; j < KSMAX; ++j) {
        const int v = lo + j;
        if (v <= hi) {
            const int vz = v - 1;
            const int up = sym_local(pp.a - vz, pp.c - vz, ks);
            const int uq = sym_local(pq.a - vz, pq.c - vz, ks);
            const double* Dv = D + (size_t)vz * U * U * K + k;     // D[v][uo][ui][k]
            dacc += Dv[(size_t)(up * U + uq) * K] + Dv[(size_t)(uq * U + up) * K];
        }
    }
    return acc + dacc;
}

// Near workgroups: one chunk of one row's near range.
template <int KSMAX>
__global__ __launch_bounds__(kThreads) void k3_near(int n_pairs, int K, int ks, int U,
                                                    const PairInfo* __restrict__ pairs,
                                                    const NearItem* __restrict__ items,
                                                    const double* __restrict__ Pw, const double* __restrict__ Qw,
                                                    const double* __restrict__ PP, const double* __restrict__ QS,
                                                    const double* __restrict__ Ptot, const double* __restrict__ Qtot,
                                                    const double* __restrict__ D, double* __restrict__ R)
{
    __shared__ double sA[kMaxOrder * kMaxK];   // PP of the row
    __shared__ double sB[kMaxOrder * kMaxK];   // QS of the row
    __shared__ double sT[2 * kMaxK];           // Ptot, Qtot of the row
    const int64_t rowlen = (int64_t)n_pairs * K;
    const int tid = threadIdx.x;
    const NearItem it = items[blockIdx.x];
    const int p = it.p;
    const PairInfo pp = pairs[p];
    const size_t pbase = (size_t)p * ks * K;
    for (int e = tid; e < pp.m * K; e += kThreads) {
        sA[e] = PP[pbase + e];
        sB[e] = QS[pbase + e];
    }
    for (int e = tid; e < K; e += kThreads) {
        sT[e] = Ptot[(size_t)p * K + e];
        sT[kMaxK + e] = Qtot[(size_t)p * K + e];
    }
    __syncthreads();
    double* Rrow = R + (size_t)p * rowlen;
    int64_t e = it.e0 + tid;
    int q = (int)(e / K);
    int k = (int)(e - (int64_t)q * K);
    const int dq = kThreads / K, dk = kThreads - (kThreads / K) * K;
#pragma unroll 1
    for (int i = 0; i < kNearIters; ++i) {
        if (e >= it.e1) break;
        const double val = near_value<KSMAX>(q, k, K, ks, U, pp, pairs, Pw, Qw, sA, sB, sT[k], sT[kMaxK + k], D);
        __builtin_nontemporal_store(val, Rrow + e);
        e += kThreads;
        q += dq;
        k += dk;
        if (k >= K) {
            k -= K;
            q += 1;
        }
    }
}

// Far workgroups: kRows rows x one column chunk of rank-one products; the
// elements of each row's near range are left to the near workgroups.
__global__ __launch_bounds__(kThreads) void k3_far(int n_pairs, int K, int far_cols, const PairInfo* __restrict__ pairs,
                                                   const NearRange* __restrict__ near,
                                                   const double* __restrict__ Ptot,
                                                   const double* __restrict__ Qtot, double* __restrict__ R)
{
    __shared__ double sA[kRows * kMaxK];   // Ptot of the rows
    __shared__ double sB[kRows * kMaxK];   // Qtot of the rows
    __shared__ int sI[4 * kRows];
    const int64_t rowlen = (int64_t)n_pairs * K;
    const int tid = threadIdx.x;
    const int ry = (int)blockIdx.x / far_cols, cx = (int)blockIdx.x - ry * far_cols;
    const int p0 = ry * kRows;
    const int nrows = (n_pairs - p0) < kRows ? (n_pairs - p0) : kRows;
    for (int e = tid; e < nrows * K; e += kThreads) {
        const int r = e / K, k = e - r * K;
        sA[r * kMaxK + k] = Ptot[(size_t)(p0 + r) * K + k];
        sB[r * kMaxK + k] = Qtot[(size_t)(p0 + r) * K + k];
    }
    if (tid < nrows) {
        const PairInfo pi = pairs[p0 + tid];
        const NearRange nr = near[p0 + tid];
        sI[4 * tid + 0] = pi.v0 + pi.m - 1;   // last interval of the row's window
        sI[4 * tid + 1] = nr.qlo;
        sI[4 * tid + 2] = nr.qhi;
    }
    __syncthreads();
    int64_t e = (int64_t)cx * kThreads * kFarIters + tid;
    int q = (int)(e / K);
    int k = (int)(e - (int64_t)q * K);
    const int dq = kThreads / K, dk = kThreads - (kThreads / K) * K;
    double* Rrow = R + (size_t)p0 * rowlen;
    for (int i = 0; i < kFarIters; ++i) {
        if (e >= rowlen) break;
        const double pt = Ptot[e];
        const double qt = Qtot[e];
        const int v0q = pairs[q].v0;
        for (int r = 0; r < nrows; ++r) {
            if (q >= sI[4 * r + 1] && q <= sI[4 * r + 2]) continue;   // near range: written by a near workgroup
            const double val = (sI[4 * r + 0] < v0q) ? sA[r * kMaxK + k] * qt : sB[r * kMaxK + k] * pt;
            __builtin_nontemporal_store(val, Rrow + (size_t)r * rowlen + e);
        }
        e += kThreads;
        q += dq;
        k += dk;
        if (k >= K) {
            k -= K;
            q += 1;
        }
    }
}

// Host: near ranges of every row and the near work items (chunks of
// kThreads*kNearIters elements of each row's near range).
void build_near_ranges(Plan& p)
{
    p.near.assign((size_t)p.n_pairs, NearRange{0, -1});
    p.items.clear();
    const int64_t ks = p.k_s, nb = p.n_b;
    const int64_t chunk = (int64_t)kThreads * kNearIters;
    for (int64_t P = 0; P < p.n_pairs; ++P) {
        const PairInfo& pp = p.pairs[P];
        const int64_t v1p = pp.v0 + pp.m - 1;
        const int64_t cmin = std::max<int64_t>(1, pp.v0 - 1);
        const int64_t cmax = std::min<int64_t>(nb, v1p + ks - 2);
        int32_t lo = -1, hi = -2;
        for (int64_t Q = p.c_first[cmin]; Q < p.c_first[cmax + 1]; ++Q) {
            const PairInfo& pq = p.pairs[Q];
            if (pq.v0 <= v1p && pp.v0 <= pq.v0 + pq.m - 1) {
                if (lo < 0) lo = (int32_t)Q;
                hi = (int32_t)Q;
            }
        }
        p.near[P] = NearRange{lo, hi};
        const int64_t e0 = (int64_t)lo * p.K, e1 = (int64_t)(hi + 1) * p.K;
        for (int64_t e = e0; e < e1; e += chunk)
            p.items.push_back(NearItem{(int32_t)P, (int32_t)e, (int32_t)std::min(e1, e + chunk), 0});
    }
}

template <int KSMAX>
static void launch_near(const Plan& p, const DevTables& d, double* R, hipStream_t s)
{
    if (p.items.empty()) return;
    hipLaunchKernelGGL(k3_near<KSMAX>, dim3((unsigned)p.items.size()), dim3(kThreads), 0, s, (int)p.n_pairs, (int)p.K,
                       (int)p.k_s, (int)p.U, d.pairs, d.items, d.Pw, d.Qw, d.PP, d.QS, d.Ptot, d.Qtot, d.D, R);
}

hipError_t launch_tensor_near(const Plan& p, const DevTables& d, double* R, hipStream_t s)
{
    if (p.k_s <= 6) launch_near<6>(p, d, R, s);
    else if (p.k_s <= 8) launch_near<8>(p, d, R, s);
    else if (p.k_s <= 10) launch_near<10>(p, d, R, s);
    else if (p.k_s <= 12) launch_near<12>(p, d, R, s);
    else if (p.k_s <= 16) launch_near<16>(p, d, R, s);
    else launch_near<20>(p, d, R, s);
    return hipGetLastError();
}

hipError_t launch_tensor_far(const Plan& p, const DevTables& d, double* R, hipStream_t s)
{
    const int64_t rowlen = p.n_pairs * p.K;
    const int far_cols = (int)((rowlen + kThreads * kFarIters - 1) / (kThreads * kFarIters));
    const int far_rows = (int)((p.n_pairs + kRows - 1) / kRows);
    hipLaunchKernelGGL(k3_far, dim3((unsigned)(far_cols * far_rows)), dim3(kThreads), 0, s, (int)p.n_pairs, (int)p.K,
                       far_cols, d.pairs, d.near, d.Ptot, d.Qtot, R);
    return hipGetLastError();
}

}  // namespace bs2e
