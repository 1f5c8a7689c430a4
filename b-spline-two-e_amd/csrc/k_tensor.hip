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
// Rows are grouped in blocks of kNearRows consecutive pairs.  For each block
// the column pairs whose windows overlap the window of any of its rows form
// the index range [qlo, qhi] (the block's "near" range, a few percent of the
// row).  Two kernels write disjoint elements, each element exactly once:
//   * k3_far : kRows rows x one column chunk of rank-one products, streamed
//     with non-temporal stores, skipping the block's near range.  Bound by
//     the FP64 store of R: 8 bytes per integral.
//   * k3_near: one block x one chunk of its near range, general formula.
//     Each lane loads its column pair's window once into registers and reuses
//     it for all rows of the block, whose prefix/suffix sums sit in LDS; the
//     diagonal cells are one load per common interval.  It runs on a second
//     stream, concurrently with k3_far.
#include "bs2e_internal.h"

namespace bs2e {

constexpr int kThreads = 256;
constexpr int kRows = 4;          // far workgroup: tensor rows (divides kNearRows)
constexpr int kFarIters = 8;      // far workgroup: elements per thread per row
constexpr int kNearRows = 8;      // near block: tensor rows
constexpr int kNearIters = 2;     // near workgroup: elements per thread

static_assert(kNearRows % kRows == 0, "far row groups must not straddle near blocks");

// LDS of a near workgroup, per row r of the block:
//   PPx[r][i][k], i in [0, m]: 0, PP[0], ..., PP[m-1] (= Ptot)
//   QSx[r][i][k], i in [0, m]: QS[0] (= Qtot), ..., QS[m-1], 0
// so that, with i clamped to [0, m], PP_p(x) = PPx[x - v0p + 1] and
// QS_p(x) = QSx[x - v0p] for every interval x; and the local pair index of
// the row's pair on each interval of its window.
size_t near_lds_bytes(int64_t ks, int64_t K)
{
    return sizeof(double) * (size_t)(kNearRows * 2 * (ks + 1) * K) + sizeof(int) * (size_t)(kNearRows * kMaxOrder) +
           sizeof(PairInfo) * kNearRows;
}

template <int KSMAX>
__global__ __launch_bounds__(kThreads) void k3_near(int n_pairs, int K, int ks, int U,
                                                    const PairInfo* __restrict__ pairs,
                                                    const NearItem* __restrict__ items,
                                                    const double* __restrict__ Pw, const double* __restrict__ Qw,
                                                    const double* __restrict__ PP, const double* __restrict__ QS,
                                                    const double* __restrict__ Ptot, const double* __restrict__ Qtot,
                                                    const double* __restrict__ Dsym, double* __restrict__ R)
{
    extern __shared__ double lds[];
    const int rowdata = 2 * (ks + 1) * K;
    int* sUp = (int*)(lds + kNearRows * rowdata);              // [kNearRows][kMaxOrder]
    PairInfo* sPair = (PairInfo*)(sUp + kNearRows * kMaxOrder);
    const int64_t rowlen = (int64_t)n_pairs * K;
    const int tid = threadIdx.x;
    const NearItem it = items[blockIdx.x];
    const int p0 = it.p;
    const int nrows = (n_pairs - p0) < kNearRows ? (n_pairs - p0) : kNearRows;
    for (int e = tid; e < nrows * rowdata; e += kThreads) {
        const int r = e / rowdata, o = e - r * rowdata;
        const int half = (ks + 1) * K;
        const int which = o >= half, oi = o - which * half;
        const int i = oi / K, k = oi - i * K;
        const PairInfo pr = pairs[p0 + r];
        const size_t pb = (size_t)(p0 + r) * ks * K + k;
        double val;
        if (!which) val = (i == 0) ? 0.0 : PP[pb + (size_t)(i - 1) * K];
        else val = (i >= pr.m) ? 0.0 : QS[pb + (size_t)i * K];
        lds[e] = val;
    }
    for (int e = tid; e < nrows * kMaxOrder; e += kThreads) {
        const int r = e / kMaxOrder, j = e - r * kMaxOrder;
        const PairInfo pr = pairs[p0 + r];
        const int vz = pr.v0 - 1 + j;
        sUp[e] = (j < pr.m) ? sym_local(pr.a - vz, pr.c - vz, ks) : 0;
    }
    if (tid < nrows) sPair[tid] = pairs[p0 + tid];
    __syncthreads();

    const unsigned UK = (unsigned)(U * K);   // Dsym has n_iv*U*U*K < 2^31 entries (checked on the host)
    int64_t e = it.e0 + tid;
    int q = (int)(e / K);
    int k = (int)(e - (int64_t)q * K);
    const int dq = kThreads / K, dk = kThreads - (kThreads / K) * K;
#pragma unroll 1
    for (int i = 0; i < kNearIters; ++i) {
        if (e >= it.e1) break;
        const PairInfo pq = pairs[q];
        const size_t qbase = (size_t)q * ks * K + k;
        double pw[KSMAX], qw[KSMAX];
        unsigned dbase[KSMAX];   // Dsym offset of (interval v0q+j, pair q) without the row's pair term
#pragma unroll
        for (int j = 0; j < KSMAX; ++j) {
            const bool in = j < pq.m;
            pw[j] = in ? Pw[qbase + (size_t)j * K] : 0.0;
            qw[j] = in ? Qw[qbase + (size_t)j * K] : 0.0;
            const int vz = pq.v0 - 1 + j;
            dbase[j] = in ? ((unsigned)vz * U * UK + (unsigned)sym_local(pq.a - vz, pq.c - vz, ks) * K + k) : 0u;
        }
#pragma unroll 1
        for (int r = 0; r < nrows; ++r) {
            const PairInfo pp = sPair[r];
            const double* sPPx = lds + r * rowdata;
            const double* sQSx = sPPx + (ks + 1) * K;
            const int* up = sUp + r * kMaxOrder;
            const int o = pq.v0 - pp.v0;   // q's window start relative to p's
            double acc = 0.0, dacc = 0.0;
#pragma unroll
            for (int j = 0; j < KSMAX; ++j) {
                // sum_{v<w} P_p(v) Q_q(w) + sum_{v>w} Q_p(v) P_q(w) at w = v0q + j
                const int ip = min(max(o + j, 0), pp.m);        // PP_p(w-1)
                const int is = min(max(o + j + 1, 0), pp.m);    // QS_p(w+1)
                acc += qw[j] * sPPx[ip * K + k] + pw[j] * sQSx[is * K + k];
                // diagonal cell of interval w if it lies in both windows
                if (j < pq.m && (unsigned)(o + j) < (unsigned)pp.m) dacc += Dsym[dbase[j] + (unsigned)up[o + j] * UK];
            }
            __builtin_nontemporal_store(acc + dacc, R + (size_t)(p0 + r) * rowlen + e);
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

// Far workgroups: kRows rows x one column chunk of rank-one products; the
// elements of each row's (block's) near range are left to k3_near.
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
            if (q >= sI[4 * r + 1] && q <= sI[4 * r + 2]) continue;   // near range: written by k3_near
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

// Host: per-row window overlap ranges, merged per block of kNearRows rows, and
// the near work items (chunks of kThreads*kNearIters elements of a block's
// near range).
void build_near_ranges(Plan& p)
{
    p.near.assign((size_t)p.n_pairs, NearRange{0, -1});
    p.items.clear();
    const int64_t ks = p.k_s, nb = p.n_b;
    const int64_t chunk = (int64_t)kThreads * kNearIters;
    for (int64_t B0 = 0; B0 < p.n_pairs; B0 += kNearRows) {
        const int64_t B1 = std::min<int64_t>(p.n_pairs, B0 + kNearRows);
        int32_t lo = INT32_MAX, hi = -1;
        for (int64_t P = B0; P < B1; ++P) {
            const PairInfo& pp = p.pairs[P];
            const int64_t v1p = pp.v0 + pp.m - 1;
            // a column pair (b, d) can only share an interval with W_p if
            // d lies in [v0p - 1, v1p + k - 2]
            const int64_t cmin = std::max<int64_t>(1, pp.v0 - 1);
            const int64_t cmax = std::min<int64_t>(nb, v1p + ks - 2);
            for (int64_t Q = p.c_first[cmin]; Q < p.c_first[cmax + 1]; ++Q) {
                const PairInfo& pq = p.pairs[Q];
                if (pq.v0 <= v1p && pp.v0 <= pq.v0 + pq.m - 1) {
                    lo = std::min(lo, (int32_t)Q);
                    hi = std::max(hi, (int32_t)Q);
                }
            }
        }
        for (int64_t P = B0; P < B1; ++P) p.near[P] = NearRange{lo, hi};
        const int64_t e0 = (int64_t)lo * p.K, e1 = (int64_t)(hi + 1) * p.K;
        for (int64_t e = e0; e < e1; e += chunk)
            p.items.push_back(NearItem{(int32_t)B0, (int32_t)e, (int32_t)std::min(e1, e + chunk), 0});
    }
}

template <int KSMAX>
static void launch_near(const Plan& p, const DevTables& d, double* R, hipStream_t s)
{
    if (p.items.empty()) return;
    hipLaunchKernelGGL(k3_near<KSMAX>, dim3((unsigned)p.items.size()), dim3(kThreads), near_lds_bytes(p.k_s, p.K), s,
                       (int)p.n_pairs, (int)p.K, (int)p.k_s, (int)p.U, d.pairs, d.items, d.Pw, d.Qw, d.PP, d.QS,
                       d.Ptot, d.Qtot, d.D, R);
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

hipError_t launch_symmetrize(const Plan& p, const DevTables& d, hipStream_t s)
{
    const int64_t n = p.n_iv * p.U * p.U * p.K;
    hipLaunchKernelGGL(k_symmetrize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (int)p.n_iv, (int)p.U,
                       (int)p.K, d.D);
    return hipGetLastError();
}

}  // namespace bs2e
