// k_dipole.hip -- dipole operators on the device.
//
// Radial dipole matrices (mat_els.f90:123-170 setup_radial_dip, :348-390
// compute_radial_dip_len / _vel), one thread per matrix element running the
// reference's sum over the intervals holding both splines and their GL points:
//   length:   r_mat(i,j)     += w r B_i B_j
//   velocity: dr_mat(i,j)    += (0,-1) w B_i B_j'
//             r_inv_mat(i,j) += (0,-1) w B_i B_j / r
//
// Dipole blocks between two symmetries (dipole.f90:8-47 construct_dip_block
// _tensor, :95-150 init_dip_block, mat_els.f90:717-828 dip_mat_neq,
// dip_red_mat, dip_red_1p_len/_vel, ang_dip_red): one wavefront per row
// configuration; a column (same walk as the Hamiltonian blocks) carries an
// entry when the radial pairs share support (direct or exchange) and the
// channel pair's ang_dip_red exceeds 5e-16.  Its value is
//   ang ( red + (-1)^(L2 + l_c + l_d) red_ex ),
//   red = pf1 d1(n_a -> n_c) S(n_b, n_d) [l_b == l_d] + pf2 d1(n_b -> n_d) S(n_a, n_c) [l_a == l_c]
// (red_ex the same with (n_c, l_c) and (n_d, l_d) exchanged), d1 the one-
// electron reduced element of the gauge and pf the channel pair's
// (-1)^(..) sqrt((2L1+1)(2L2+1)) 6j(..) factors, all per channel pair from
// the host; complex products without fused multiply-adds, in the reference's
// order.
#include "bs2e_internal.h"
#include "bspline_device.h"
#include "channel_walk.h"

// Every product and sum below is rounded on its own: no fused multiply-adds.
#pragma clang fp contract(off)

namespace bs2e {

constexpr int kDThreads = 256;
constexpr int kDWave = 64;

// ---------------------------------------------------------------- radial
template <int KS>
__global__ __launch_bounds__(64) void kr_tables(int kgl, const double* __restrict__ t, const double* __restrict__ bp,
                                                const double* __restrict__ x, const double* __restrict__ w,
                                                double* __restrict__ B, double* __restrict__ B1,
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
        double d1[KS];
        splines_d1_at<KS>(t, i1, r, d1);
#pragma unroll
        for (int ls = 0; ls < KS; ++ls) B1[o + ls] = d1[ls];
        rw[2 * ((size_t)v * kgl + q)] = r;
        rw[2 * ((size_t)v * kgl + q) + 1] = h * w[q];
    }
}

__global__ __launch_bounds__(256) void kr_assemble(int n_b, int n_iv, int ks, int kgl, int gauge,
                                                   const double* __restrict__ B, const double* __restrict__ B1,
                                                   const double* __restrict__ rw, double* __restrict__ M1,
                                                   double* __restrict__ M2)
{
    const int64_t nn = (int64_t)n_b * n_b;
    const int64_t ij = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= nn) return;
    const int jb = (int)(ij / n_b) + 1, ib = (int)(ij - (int64_t)(jb - 1) * n_b) + 1;
    double a = 0.0, bsum = 0.0;
    const int d = ib > jb ? ib - jb : jb - ib;
    if (d < ks) {
        const int mx = ib > jb ? ib : jb, mn = ib < jb ? ib : jb;
        const int lo = (mx + 2 - ks) > 1 ? (mx + 2 - ks) : 1;
        const int hiv = (mn + 1) < n_iv ? (mn + 1) : n_iv;
        for (int iv = lo; iv <= hiv; ++iv) {
            const int ta = ib + 1 - iv, tc = jb + 1 - iv;
            for (int q = 0; q < kgl; ++q) {
                const size_t o = (size_t)(iv - 1) * kgl + q;
                const double bi = B[o * ks + ta], bj = B[o * ks + tc];
                const double r = rw[2 * o], wq = rw[2 * o + 1];
                if (gauge == 0) {
                    a += ((wq * r) * bi) * bj;
                } else {
                    const double mw = -wq;   // (0,-1) w: imaginary part
                    a += (mw * bi) * B1[o * ks + tc];
                    bsum += ((mw * bi) * bj) / r;
                }
            }
        }
    }
    if (gauge == 0) {
        M1[2 * ij] = a;
        M1[2 * ij + 1] = 0.0;
    } else {
        M1[2 * ij] = 0.0;
        M1[2 * ij + 1] = a;
        M2[2 * ij] = 0.0;
        M2[2 * ij + 1] = bsum;
    }
}

template <int KS>
static void launch_kr_tables(const Plan& p, const DevTables& d, double* B, double* B1, double* rw, hipStream_t s)
{
    hipLaunchKernelGGL(kr_tables<KS>, dim3((unsigned)p.n_iv), dim3(64), 0, s, (int)p.k_gl, d.t, d.bp, d.x, d.w, B, B1,
                       rw);
}

hipError_t launch_radial_dipole(const Plan& p, const DevTables& d, double* B, double* B1, double* rw, int gauge,
                                double* M1, double* M2, hipStream_t s)
{
    switch (p.k_s) {
        case 2: launch_kr_tables<2>(p, d, B, B1, rw, s); break;
        case 3: launch_kr_tables<3>(p, d, B, B1, rw, s); break;
        case 4: launch_kr_tables<4>(p, d, B, B1, rw, s); break;
        case 5: launch_kr_tables<5>(p, d, B, B1, rw, s); break;
        case 6: launch_kr_tables<6>(p, d, B, B1, rw, s); break;
        case 7: launch_kr_tables<7>(p, d, B, B1, rw, s); break;
        case 8: launch_kr_tables<8>(p, d, B, B1, rw, s); break;
        case 9: launch_kr_tables<9>(p, d, B, B1, rw, s); break;
        case 10: launch_kr_tables<10>(p, d, B, B1, rw, s); break;
        case 11: launch_kr_tables<11>(p, d, B, B1, rw, s); break;
        case 12: launch_kr_tables<12>(p, d, B, B1, rw, s); break;
        case 13: launch_kr_tables<13>(p, d, B, B1, rw, s); break;
        case 14: launch_kr_tables<14>(p, d, B, B1, rw, s); break;
        case 15: launch_kr_tables<15>(p, d, B, B1, rw, s); break;
        case 16: launch_kr_tables<16>(p, d, B, B1, rw, s); break;
        default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t n = p.n_b * p.n_b;
    hipLaunchKernelGGL(kr_assemble, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (int)p.n_b, (int)p.n_iv,
                       (int)p.k_s, (int)p.k_gl, gauge, B, B1, rw, M1, M2);
    return hipGetLastError();
}

// ---------------------------------------------------------------- blocks
struct DCplx {
    double re, im;
};
__device__ __forceinline__ DCplx dmul(DCplx a, DCplx b)
{
    return DCplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ DCplx dadd(DCplx a, DCplx b) { return DCplx{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ DCplx dsub(DCplx a, DCplx b) { return DCplx{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ DCplx dscale(double s, DCplx a) { return DCplx{s * a.re, s * a.im}; }
// element (r, c), 1-based, of a complex row-major n_b x n_b matrix: a
// wavefront's consecutive n_d are consecutive columns
__device__ __forceinline__ DCplx dmat(const double* M, int n_b, int r, int c)
{
    const size_t o = 2 * ((size_t)(r - 1) * n_b + (c - 1));
    return DCplx{M[o], M[o + 1]};
}

// one-electron reduced dipole element of term t between radial indices (n, n')
__device__ __forceinline__ DCplx red1p(const DipArgs& A, const DipCoef& c, int t, int n, int np)
{
    if (A.gauge == 0) return dscale(c.cred[t], dmat(A.rad1, A.n_b, n, np));
    if (c.vkind[t] == 0) return DCplx{0.0, 0.0};
    const DCplx dr = dmat(A.rad1, A.n_b, n, np), ri = dmat(A.rad2, A.n_b, n, np);
    if (c.vkind[t] == 1) return dscale(c.vsq[t], dsub(dr, dscale(c.vm[t], ri)));
    const DCplx v = dscale(c.vsq[t], dadd(dr, dscale(c.vm[t], ri)));
    return DCplx{-v.re, -v.im};
}

// The fill: entries go to the row's range from the count pass; stores beyond
// the outputs (lim entries) are dropped.
__global__ __launch_bounds__(kDThreads) void kd_block(DipArgs A)
{
    const int lane = threadIdx.x & (kDWave - 1);
    const int i = (int)blockIdx.x * (kDThreads / kDWave) + __builtin_amdgcn_readfirstlane((int)threadIdx.x / kDWave);
    if (i >= A.n1) return;   // whole wave exits together
    const int ks = A.ks;
    const int64_t* ci = A.cfg1 + 5 * (size_t)i;
    const int n_a = __builtin_amdgcn_readfirstlane((int)ci[0]), n_bb = __builtin_amdgcn_readfirstlane((int)ci[1]);
    const int ch_i = __builtin_amdgcn_readfirstlane(A.cfg_chan1[i]);
    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int64_t n = 0;
    const int64_t base = A.ptr[i];
    for (int ch = 0; ch < A.n_chan2; ++ch) {
        const DipCoef& c = A.coef[(size_t)ch_i * A.n_chan2 + ch];
        if (!c.nz) continue;
        const BlockChan cj = A.chan2[ch];
        const BlockSeg* sg = A.seg2 + (size_t)ch * (A.n_b + 1);
        walk_channel<false>(n_a, n_bb, ks, A.n_b, sg, lane, nullptr,
                            [&](bool active, int n_c, int n_d, int col, bool in_dir_c,
                                                          bool in_ex_c) {
            const bool support = in_dir_c && iabs_d(n_bb - n_d) < ks;
            const bool support_ex = in_ex_c && iabs_d(n_a - n_d) < ks;
            const bool entry = active && (support || support_ex);
            const unsigned long long m = __ballot(entry);
            if (entry) {
                const DCplx zero{0.0, 0.0};
                // direct: electron 1 n_a -> n_c, electron 2 n_b -> n_d
                const DCplx r1 = c.cond[0] ? dmul(dscale(c.pf[0], red1p(A, c, 0, n_a, n_c)), dmat(A.S, A.n_b, n_bb, n_d))
                                           : zero;
                const DCplx r2 = c.cond[1] ? dmul(dscale(c.pf[1], red1p(A, c, 1, n_bb, n_d)), dmat(A.S, A.n_b, n_a, n_c))
                                           : zero;
                // exchange: the column's electrons swapped
                const DCplx x1 = c.cond[2] ? dmul(dscale(c.pf[2], red1p(A, c, 2, n_a, n_d)), dmat(A.S, A.n_b, n_bb, n_c))
                                           : zero;
                const DCplx x2 = c.cond[3] ? dmul(dscale(c.pf[3], red1p(A, c, 3, n_bb, n_c)), dmat(A.S, A.n_b, n_a, n_d))
                                           : zero;
                const DCplx red = dadd(r1, r2), redx = dscale(c.sgn_ex, dadd(x1, x2));
                const DCplx v = dscale(A.ang, dadd(red, redx));
                const int64_t p = base + n + __popcll(m & lt_mask);
                if (p < A.lim) {
                    A.ix[p] = cj.start + col + 1;
                    A.data[2 * p] = v.re;
                    A.data[2 * p + 1] = v.im;
                }
            }
            n += __popcll(m);
        });
    }
}

// Row counts of the count pass, in closed form: within one column channel
// whose channel pair carries entries, candidate n_c (one per lane) owns the
// direct run D of n_d (within k of n_b, when |n_c - n_a| < k) and the
// exchange run E (within k of n_a, when |n_c - n_b| < k) inside its n2 range;
// the entries are D u E.
__global__ __launch_bounds__(kDThreads) void kd_count(DipArgs A)
{
    const int lane = threadIdx.x & (kDWave - 1);
    const int i = (int)blockIdx.x * (kDThreads / kDWave) + __builtin_amdgcn_readfirstlane((int)threadIdx.x / kDWave);
    if (i >= A.n1) return;   // whole wave exits together
    const int ks = A.ks, n_b = A.n_b;
    const int64_t* ci = A.cfg1 + 5 * (size_t)i;
    const int n_a = __builtin_amdgcn_readfirstlane((int)ci[0]), n_bb = __builtin_amdgcn_readfirstlane((int)ci[1]);
    const int ch_i = __builtin_amdgcn_readfirstlane(A.cfg_chan1[i]);
    const CandRanges cr = candidate_ranges(n_a, n_bb, ks, n_b);
    const bool cand = lane < cr.n1 + cr.n2;
    const int n_c = lane < cr.n1 ? cr.r1lo + lane : cr.r2lo + (lane - cr.n1);
    const bool in_dir_c = cand && iabs_d(n_c - n_a) < ks, in_ex_c = cand && iabs_d(n_c - n_bb) < ks;
    int n = 0;
    for (int ch = 0; ch < A.n_chan2; ++ch) {
        if (!A.coef[(size_t)ch_i * A.n_chan2 + ch].nz || !cand) continue;
        const BlockSeg sg = A.seg2[(size_t)ch * (n_b + 1) + n_c];
        if (sg.hi < sg.lo) continue;
        const int dlo = in_dir_c ? (sg.lo > n_bb - ks + 1 ? sg.lo : n_bb - ks + 1) : 1;
        const int dhi = in_dir_c ? (sg.hi < n_bb + ks - 1 ? sg.hi : n_bb + ks - 1) : 0;
        const int elo = in_ex_c ? (sg.lo > n_a - ks + 1 ? sg.lo : n_a - ks + 1) : 1;
        const int ehi = in_ex_c ? (sg.hi < n_a + ks - 1 ? sg.hi : n_a + ks - 1) : 0;
        const int ov_lo = dlo > elo ? dlo : elo, ov_hi = dhi < ehi ? dhi : ehi;
        n += (dhi >= dlo ? dhi - dlo + 1 : 0) + (ehi >= elo ? ehi - elo + 1 : 0) - (ov_hi >= ov_lo ? ov_hi - ov_lo + 1 : 0);
    }
    for (int off = kDWave / 2; off > 0; off >>= 1) n += __shfl_xor(n, off, kDWave);
    if (lane == 0) A.cnt[i] = n;
}

hipError_t launch_dip_count(const DipArgs& a, hipStream_t s)
{
    const int64_t threads = (int64_t)a.n1 * kDWave;
    hipLaunchKernelGGL(kd_count, dim3((unsigned)((threads + kDThreads - 1) / kDThreads)), dim3(kDThreads), 0, s,
                       a);
    return hipGetLastError();
}

// dst = the transpose of the complex n_b x n_b matrix src (column-major to
// row-major), one thread per element
__global__ __launch_bounds__(256) void kd_transpose(const double* __restrict__ src, double* __restrict__ dst, int n_b)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)n_b * n_b) return;
    const int64_t r = e / n_b, c = e - r * n_b;   // destination (r, c)
    dst[2 * e] = src[2 * (c * n_b + r)];
    dst[2 * e + 1] = src[2 * (c * n_b + r) + 1];
}

hipError_t launch_dip_prepare(const double* S, const double* R1, const double* R2, int n_b, double* ST, double* R1T,
                              double* R2T, hipStream_t s)
{
    const unsigned g = (unsigned)(((int64_t)n_b * n_b + 255) / 256);
    hipLaunchKernelGGL(kd_transpose, dim3(g), dim3(256), 0, s, S, ST, n_b);
    hipLaunchKernelGGL(kd_transpose, dim3(g), dim3(256), 0, s, R1, R1T, n_b);
    if (R2) hipLaunchKernelGGL(kd_transpose, dim3(g), dim3(256), 0, s, R2, R2T, n_b);
    return hipGetLastError();
}

hipError_t launch_dip_fill(const DipArgs& a, hipStream_t s)
{
    const int64_t threads = (int64_t)a.n1 * kDWave;
    hipLaunchKernelGGL(kd_block, dim3((unsigned)((threads + kDThreads - 1) / kDThreads)), dim3(kDThreads), 0, s,
                       a);
    return hipGetLastError();
}

}  // namespace bs2e
