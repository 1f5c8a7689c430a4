// k_scatter.hip -- writes the device primitives in the reference's container
// order, for callers of setup_Slater_integrals (mat_els.f90:172-182):
//   sparse_4d (r_k, r_m_k): data(ptr, 0:max_k) column-major, ptr over
//     j_b, i_b (|i_b-j_b| < k), i_r ascending (mat_els.f90:203-223);
//     index arrays iv, i = i_b, j = j_b (:430-438);
//   sparse_6d (r_d_k): ptr over j_b_p, j_b, i_b_p, i_b, i_r (:251-289);
//     index arrays iv, i, i_p, j, j_p (:484-490).
#include "bs2e_internal.h"

namespace bs2e {

__global__ __launch_bounds__(256) void k_scatter_4d(int n_pairs, int K, int ks, int64_t nnz4,
                                                    const PairInfo* __restrict__ pairs,
                                                    const int64_t* __restrict__ off4, const double* __restrict__ Pw,
                                                    const double* __restrict__ Qw, double* __restrict__ rk,
                                                    double* __restrict__ rmk, int64_t* __restrict__ iv,
                                                    int64_t* __restrict__ ii, int64_t* __restrict__ jj)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t)n_pairs * ks) return;
    const int p = (int)(gid / ks), j = (int)(gid - (int64_t)p * ks);
    const PairInfo pi = pairs[p];
    if (j >= pi.m) return;
    const int64_t ptr = off4[p] + j;
    const size_t src = ((size_t)p * ks + j) * K;
    for (int k = 0; k < K; ++k) {
        rk[(size_t)k * nnz4 + ptr] = Pw[src + k];
        rmk[(size_t)k * nnz4 + ptr] = Qw[src + k];
    }
    iv[ptr] = pi.v0 + j;
    ii[ptr] = pi.a;
    jj[ptr] = pi.c;
}

__global__ __launch_bounds__(256) void k_scatter_6d(int64_t n_diag, int K, int ks, int U, int64_t nnz6,
                                                    const PairInfo* __restrict__ pairs,
                                                    const DiagEntry* __restrict__ diag, const double* __restrict__ D,
                                                    double* __restrict__ rdk, int64_t* __restrict__ iv,
                                                    int64_t* __restrict__ i, int64_t* __restrict__ ip,
                                                    int64_t* __restrict__ j, int64_t* __restrict__ jp)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n_diag * ks) return;
    const int64_t e = gid / ks;
    const int s = (int)(gid - e * ks);
    const DiagEntry de = diag[e];
    if (s >= de.len) return;
    const int64_t ptr = de.ptr + s;
    const int v = de.v0 + s, vz = v - 1;
    const PairInfo pI = pairs[de.I], pJ = pairs[de.J];
    const int uo = sym_local(pI.a - vz, pI.c - vz, ks);   // outer pair (i, i_p)
    const int ui = sym_local(pJ.a - vz, pJ.c - vz, ks);   // inner pair (j, j_p)
    const double* Dv = D + (((size_t)vz * U + uo) * U + ui) * K;   // D[v][uo][ui][k]
    for (int k = 0; k < K; ++k) rdk[(size_t)k * nnz6 + ptr] = Dv[k];
    iv[ptr] = v;
    i[ptr] = pI.a;
    ip[ptr] = pI.c;
    j[ptr] = pJ.a;
    jp[ptr] = pJ.c;
}

hipError_t launch_scatter_4d(const Plan& p, const DevTables& d, const int64_t* d_off4, double* rk, double* rmk,
                             int64_t* iv, int64_t* i, int64_t* j, hipStream_t s)
{
    const int64_t n = p.n_pairs * p.k_s;
    hipLaunchKernelGGL(k_scatter_4d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (int)p.n_pairs, (int)p.K,
                       (int)p.k_s, p.nnz4, d.pairs, d_off4, d.Pw, d.Qw, rk, rmk, iv, i, j);
    return hipGetLastError();
}

hipError_t launch_scatter_6d(const Plan& p, const DevTables& d, const DiagEntry* d_diag, int64_t n_diag, double* rdk,
                             int64_t* iv, int64_t* i, int64_t* ip, int64_t* j, int64_t* jp, hipStream_t s)
{
    const int64_t n = n_diag * p.k_s;
    hipLaunchKernelGGL(k_scatter_6d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n_diag, (int)p.K,
                       (int)p.k_s, (int)p.U, p.nnz6, d.pairs, d_diag, d.D, rdk, iv, i, ip, j, jp);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Gathers for compute_R_K_map(r_d_k, r_k, r_m_k, ...) (sparse_array_tools.f90:
// 452-493) when the primitives come from the caller: each container entry is
// placed by its own index arrays (iv, i, j) / (iv, i, i_p, j, j_p), so the
// tensor is built from exactly the values the caller holds.  Entries whose
// indices are outside the banded, supported set are counted in *bad.
// Diagonal cells are stored once per unordered local pair: the entry with
// i <= i_p and j <= j_p is taken (its mirror images carry the same integral).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fill(double* __restrict__ p, int64_t n, double v)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < n) p[gid] = v;
}

__global__ __launch_bounds__(256) void k_gather_4d(int64_t nnz4, int K, int ks, int n_b,
                                                   const int32_t* __restrict__ c_first,
                                                   const PairInfo* __restrict__ pairs, const double* __restrict__ rk,
                                                   const double* __restrict__ rmk, const int64_t* __restrict__ iv,
                                                   const int64_t* __restrict__ ii, const int64_t* __restrict__ jj,
                                                   double* __restrict__ Pw, double* __restrict__ Qw,
                                                   int* __restrict__ bad)
{
    const int64_t ptr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ptr >= nnz4) return;
    const int64_t a = ii[ptr], c = jj[ptr], v = iv[ptr];
    if (a < 1 || a > n_b || c < 1 || c > n_b || (a > c ? a - c : c - a) >= ks) {
        atomicAdd(bad, 1);
        return;
    }
    const int lo = (int)(c - ks + 1 < 1 ? 1 : c - ks + 1);
    const int p = c_first[c] + (int)(a - lo);
    const PairInfo pi = pairs[p];
    const int64_t slot = v - pi.v0;
    if (slot < 0 || slot >= pi.m) {
        atomicAdd(bad, 1);
        return;
    }
    const size_t dst = ((size_t)p * ks + slot) * K;
    for (int k = 0; k < K; ++k) {
        Pw[dst + k] = rk[(size_t)k * nnz4 + ptr];
        Qw[dst + k] = rmk[(size_t)k * nnz4 + ptr];
    }
}

__global__ __launch_bounds__(256) void k_gather_6d(int64_t nnz6, int K, int ks, int U, int n_b, int n_iv,
                                                   const double* __restrict__ rdk, const int64_t* __restrict__ iv,
                                                   const int64_t* __restrict__ i, const int64_t* __restrict__ ip,
                                                   const int64_t* __restrict__ j, const int64_t* __restrict__ jp,
                                                   double* __restrict__ D, int* __restrict__ bad)
{
    const int64_t ptr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ptr >= nnz6) return;
    const int64_t v = iv[ptr];
    const int64_t b[4] = {i[ptr], ip[ptr], j[ptr], jp[ptr]};
    bool ok = v >= 1 && v <= n_iv;
    for (int e = 0; e < 4; ++e) {
        const int64_t s = b[e] + 1;   // spline index of basis function b
        ok = ok && b[e] >= 1 && b[e] <= n_b && v <= s && s <= v + ks - 1;
    }
    if (!ok) {
        atomicAdd(bad, 1);
        return;
    }
    if (b[0] > b[1] || b[2] > b[3]) return;
    const int vz = (int)v - 1;
    const int uo = sym_local((int)b[0] - vz, (int)b[1] - vz, ks);
    const int ui = sym_local((int)b[2] - vz, (int)b[3] - vz, ks);
    double* Dv = D + (((size_t)vz * U + uo) * U + ui) * K;         // D[v][uo][ui][k]
    for (int k = 0; k < K; ++k) Dv[k] = rdk[(size_t)k * nnz6 + ptr];
}

hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_gather_4d(const Plan& p, const DevTables& d, const int32_t* d_c_first, const double* rk,
                            const double* rmk, const int64_t* iv, const int64_t* i, const int64_t* j, int* bad,
                            hipStream_t s)
{
    if (p.nnz4 == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_4d, dim3((unsigned)((p.nnz4 + 255) / 256)), dim3(256), 0, s, p.nnz4, (int)p.K,
                       (int)p.k_s, (int)p.n_b, d_c_first, d.pairs, rk, rmk, iv, i, j, d.Pw, d.Qw, bad);
    return hipGetLastError();
}

hipError_t launch_gather_6d(const Plan& p, const DevTables& d, const double* rdk, const int64_t* iv, const int64_t* i,
                            const int64_t* ip, const int64_t* j, const int64_t* jp, int* bad, hipStream_t s)
{
    if (p.nnz6 == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_6d, dim3((unsigned)((p.nnz6 + 255) / 256)), dim3(256), 0, s, p.nnz6, (int)p.K,
                       (int)p.k_s, (int)p.U, (int)p.n_b, (int)p.n_iv, rdk, iv, i, ip, j, jp, d.D, bad);
    return hipGetLastError();
}

}  // namespace bs2e
