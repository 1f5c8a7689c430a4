// k_onebody.hip -- one-electron matrices on the device: the overlap
// (mat_els.f90:85-118 setup_S, :330-346 compute_S) and the one-particle
// Hamiltonian of a hydrogenic potential with a complex absorbing potential
// (mat_els.f90:47-83 setup_H_one_particle, :294-321 compute_H,
// potentials.f90 V_hydrogenic, CAP_tools.f90:24-34 V_CAP).
//
//   ko_tables: one workgroup per interval, one thread per GL point: the
//     point r = h x_q + c and weight h w_q of quad_tools.f90:29-52 and the
//     values and second derivatives of the k splines non-zero there.
//   ko_assemble: one thread per matrix element (i_b, j_b) (and per l): the
//     reference's running sum over the intervals holding both splines, in
//     ascending order, and over the GL points, in order:
//       S(i,j) += w B_i B_j
//       H(i,j) += w (-0.5 B_i B_j'' + (V_l(r) + V_CAP(r)) B_i B_j)
//     with V_l(r) = 0.5 l(l+1)/r^2 - Z/r and V_CAP(r) = -i eta (r - r_0)^order
//     for r >= r_0 (0 below).  Elements with |i_b - j_b| >= k are zero.
//
//   ko_rk_int: one thread per (k, i_b, j_b) of the RKH one-electron
//     multipoles (RKH_hamiltonian.f90:91-159 setup_r_k_int with
//     adiabatic_potentials.f90:24-34 r_k): per interval holding both splines,
//     in ascending order, the GL sum sum_q w r_k(r_q) B_i B_j, added to the
//     running total; r_k(r) = r_0^k / r^(k+1) for r > r_0 and
//     r^k / r_0^(k+1) otherwise, integer powers by squaring.
#include "bs2e_internal.h"
#include "bspline_device.h"

namespace bs2e {

template <int KS>
__global__ __launch_bounds__(64) void ko_tables(int kgl, const double* __restrict__ t, const double* __restrict__ bp,
                                                const double* __restrict__ x, const double* __restrict__ w,
                                                double* __restrict__ B, double* __restrict__ B2,
                                                double* __restrict__ rw)
{
    const int v = blockIdx.x;
    const double lim1 = bp[v], lim2 = bp[v + 1];
    const double h = (lim2 - lim1) / 2.0, c = (lim1 + lim2) / 2.0;
    const int i1 = KS - 1 + (v + 1);
    for (int q = threadIdx.x; q < kgl; q += blockDim.x) {
        const double r = h * x[q] + c;
        const size_t o = ((size_t)v * kgl + q) * KS;
        splines_at<KS>(t, i1, r, B + o, 1);
        double d2[KS];
        splines_d2_at<KS>(t, i1, r, d2);
#pragma unroll
        for (int ls = 0; ls < KS; ++ls) B2[o + ls] = d2[ls];
        rw[2 * ((size_t)v * kgl + q)] = r;
        rw[2 * ((size_t)v * kgl + q) + 1] = h * w[q];
    }
}

struct OneBodyArgs {
    int n_b, n_iv, ks, kgl, n_l, l_min, Z, cap_order;
    double cap_r0, eta_re, eta_im;
    const double *B, *B2, *rw;
    double *S, *H;   // complex, column-major n_b x n_b (H: n_l of them); S may be null
};

__global__ __launch_bounds__(256) void ko_assemble(OneBodyArgs A)
{
    const int64_t nn = (int64_t)A.n_b * A.n_b;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nn * A.n_l) return;
    const int li = (int)(e / nn);
    const int64_t ij = e - (int64_t)li * nn;
    const int jb = (int)(ij / A.n_b) + 1, ib = (int)(ij - (int64_t)(jb - 1) * A.n_b) + 1;
    const int ks = A.ks, kgl = A.kgl;
    double s = 0.0, hr = 0.0, hi = 0.0;
    const int d = ib > jb ? ib - jb : jb - ib;
    if (d < ks) {
        const int l = A.l_min + li;
        const double cl = 0.5 * l * (l + 1);
        const int mx = ib > jb ? ib : jb, mn = ib < jb ? ib : jb;
        const int lo = (mx + 2 - ks) > 1 ? (mx + 2 - ks) : 1;
        const int hiv = (mn + 1) < A.n_iv ? (mn + 1) : A.n_iv;
        for (int iv = lo; iv <= hiv; ++iv) {
            const int ta = ib + 1 - iv, tc = jb + 1 - iv;
            for (int q = 0; q < kgl; ++q) {
                const size_t o = ((size_t)(iv - 1) * kgl + q);
                const double bi = A.B[o * ks + ta], bj = A.B[o * ks + tc], d2j = A.B2[o * ks + tc];
                const double r = A.rw[2 * o], wq = A.rw[2 * o + 1];
                s += wq * bi * bj;
                // (V + V_CAP) B_i B_j, V real, V_CAP = (0,-1) eta (r - r_0)^order
                double vre = cl / (r * r) - (double)A.Z / r, vim = 0.0;
                if (r >= A.cap_r0) {
                    double pw = 1.0, base = r - A.cap_r0;
                    for (int n = A.cap_order; n > 0; n >>= 1) {   // integer power by squaring
                        if (n & 1) pw *= base;
                        base *= base;
                    }
                    vre += A.eta_im * pw;    // (0,-1) (eta_re, eta_im) = (eta_im, -eta_re)
                    vim = -A.eta_re * pw;
                }
                hr += wq * (-0.5 * bi * d2j + vre * bi * bj);
                hi += wq * (vim * bi * bj);
            }
        }
    }
    if (A.S && li == 0) {
        A.S[2 * ij] = s;
        A.S[2 * ij + 1] = 0.0;
    }
    if (A.H) {
        A.H[2 * e] = hr;
        A.H[2 * e + 1] = hi;
    }
}

__global__ __launch_bounds__(256) void ko_rk_int(int n_b, int n_iv, int ks, int kgl, int K, double r0,
                                                 const double* __restrict__ B, const double* __restrict__ rw,
                                                 double* __restrict__ out)
{
    const int64_t n = (int64_t)K * n_b * n_b;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int k = (int)(e % K);
    const int64_t ij = e / K;
    const int jb = (int)(ij / n_b) + 1, ib = (int)(ij - (int64_t)(jb - 1) * n_b) + 1;
    double acc = 0.0;
    const int d = ib > jb ? ib - jb : jb - ib;
    if (d < ks) {
        const int mx = ib > jb ? ib : jb, mn = ib < jb ? ib : jb;
        const int lo = (mx + 2 - ks) > 1 ? (mx + 2 - ks) : 1;
        const int hiv = (mn + 1) < n_iv ? (mn + 1) : n_iv;
        const double r0k = ipow_d(r0, k), r0k1 = ipow_d(r0, k + 1);
        for (int iv = lo; iv <= hiv; ++iv) {
            const int ta = ib + 1 - iv, tc = jb + 1 - iv;
            double res = 0.0;
            for (int q = 0; q < kgl; ++q) {
#pragma clang fp contract(off)
                const size_t o = (size_t)(iv - 1) * kgl + q;
                const double r = rw[2 * o], wq = rw[2 * o + 1];
                const double rk = r > r0 ? r0k / ipow_d(r, k + 1) : ipow_d(r, k) / r0k1;
                res = res + ((wq * rk) * B[o * ks + ta]) * B[o * ks + tc];   // each product rounded: no FMA
            }
            acc = acc + res;
        }
    }
    out[e] = acc;
}

template <int KS>
static void launch_tables(const Plan& p, const DevTables& d, double* B, double* B2, double* rw, hipStream_t s)
{
    hipLaunchKernelGGL(ko_tables<KS>, dim3((unsigned)p.n_iv), dim3(64), 0, s, (int)p.k_gl, d.t, d.bp, d.x, d.w, B, B2,
                       rw);
}

hipError_t launch_one_body(const Plan& p, const DevTables& d, double* B, double* B2, double* rw, int l_min, int n_l,
                           int Z, int cap_order, double cap_r0, double eta_re, double eta_im, double* S, double* H,
                           hipStream_t s)
{
    switch (p.k_s) {
        case 2: launch_tables<2>(p, d, B, B2, rw, s); break;
        case 3: launch_tables<3>(p, d, B, B2, rw, s); break;
        case 4: launch_tables<4>(p, d, B, B2, rw, s); break;
        case 5: launch_tables<5>(p, d, B, B2, rw, s); break;
        case 6: launch_tables<6>(p, d, B, B2, rw, s); break;
        case 7: launch_tables<7>(p, d, B, B2, rw, s); break;
        case 8: launch_tables<8>(p, d, B, B2, rw, s); break;
        case 9: launch_tables<9>(p, d, B, B2, rw, s); break;
        case 10: launch_tables<10>(p, d, B, B2, rw, s); break;
        case 11: launch_tables<11>(p, d, B, B2, rw, s); break;
        case 12: launch_tables<12>(p, d, B, B2, rw, s); break;
        case 13: launch_tables<13>(p, d, B, B2, rw, s); break;
        case 14: launch_tables<14>(p, d, B, B2, rw, s); break;
        case 15: launch_tables<15>(p, d, B, B2, rw, s); break;
        case 16: launch_tables<16>(p, d, B, B2, rw, s); break;
        default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    OneBodyArgs A{(int)p.n_b, (int)p.n_iv, (int)p.k_s, (int)p.k_gl, n_l, l_min, Z, cap_order, cap_r0, eta_re, eta_im,
                  B, B2, rw, S, H};
    const int64_t n = (int64_t)p.n_b * p.n_b * n_l;
    hipLaunchKernelGGL(ko_assemble, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_rk_int(const Plan& p, const DevTables& d, double* B, double* B2, double* rw, int max_k, double r0,
                         double* out, hipStream_t s)
{
    switch (p.k_s) {
        case 2: launch_tables<2>(p, d, B, B2, rw, s); break;
        case 3: launch_tables<3>(p, d, B, B2, rw, s); break;
        case 4: launch_tables<4>(p, d, B, B2, rw, s); break;
        case 5: launch_tables<5>(p, d, B, B2, rw, s); break;
        case 6: launch_tables<6>(p, d, B, B2, rw, s); break;
        case 7: launch_tables<7>(p, d, B, B2, rw, s); break;
        case 8: launch_tables<8>(p, d, B, B2, rw, s); break;
        case 9: launch_tables<9>(p, d, B, B2, rw, s); break;
        case 10: launch_tables<10>(p, d, B, B2, rw, s); break;
        case 11: launch_tables<11>(p, d, B, B2, rw, s); break;
        case 12: launch_tables<12>(p, d, B, B2, rw, s); break;
        case 13: launch_tables<13>(p, d, B, B2, rw, s); break;
        case 14: launch_tables<14>(p, d, B, B2, rw, s); break;
        case 15: launch_tables<15>(p, d, B, B2, rw, s); break;
        case 16: launch_tables<16>(p, d, B, B2, rw, s); break;
        default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t n = (int64_t)(max_k + 1) * p.n_b * p.n_b;
    hipLaunchKernelGGL(ko_rk_int, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (int)p.n_b, (int)p.n_iv,
                       (int)p.k_s, (int)p.k_gl, max_k + 1, r0, B, rw, out);
    return hipGetLastError();
}

}  // namespace bs2e
