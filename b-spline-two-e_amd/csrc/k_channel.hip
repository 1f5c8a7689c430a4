// k_channel.hip -- angular-channel summation of R^k into the CSR blocks of one
// symmetry (L, parity): hamiltonian.f90:106-283 construct_block_tensor with
// mat_els.f90:552-571 r_12_tens, :608-633 c_mat_neq_tens, :664-678 S_mat_neq,
// :697-715 H_1p_neq, and the sizes of :348-416 count_nnz.
//
// Configurations (n1, n2, l1, l2) of a block come in the reference's order
// (orbital_tools.f90:144-193): channel (l1, l2), then n1, then n2.  Row i =
// (n_a, n_b, l_a, l_b), column j = (n_c, n_d, l_c, l_d).  An entry can only
// exist where the radial pairs share support:
//   support    = |n_a-n_c| < k  and |n_b-n_d| < k
//   support_ex = |n_a-n_d| < k  and |n_b-n_c| < k,
// so one wavefront per row visits, channel by channel, only the columns with
// n_c within k of n_a or n_b, and for each such n_c the n_d within k of n_b
// (direct) or n_a (exchange), in ascending column order, under the
// reference's predicate
//   H entry <=> (support or support_ex) and (r12_allowed or l_eq or l_eq_ex)
//   S entry <=> (support and l_eq) or (support_ex and l_eq_ex).
// Rows with |n_a - n_b| >= k (most of a large block) take the matrix-core
// path (fill_row_staged): per channel the direct and exchange sums are small
// matrix products, placed in CSR by a prefix sum over the candidates n_c.
// The other rows walk the candidates 64 at a time and ballots give each
// entry its CSR position.  The angular factors are per
// (row channel, column channel) coefficient vectors:
//   cdir[k] = ang_k_LS(k, confs, L)               if |ang| >= 5e-16 else 0
//   cex[k]  = (-1)^(l_c+l_d+L) ang_k_LS(k, confs_ex, L)  (same threshold)
// and r12_allowed uses the 5e-15 threshold of :171-178.
#include "bs2e_internal.h"
#include "channel_walk.h"

// Every product and sum below is rounded on its own, as the reference's
// Fortran evaluates them: no fused multiply-adds.
#pragma clang fp contract(off)

namespace bs2e {

typedef double dbl2 __attribute__((ext_vector_type(2)));

constexpr int kCThreads = 256;
constexpr int kWave = 64;
// Fill workgroups per XCD band, for launches of at least kFillBandRows rows.
// Measured on the c3 set (L = 0..4, l <= 4; 1.77 M matrix-core rows) in four
// processes, each against its own placement of the outputs: bands of 512
// fill in 21.4-24.8 ms, within 0.4 ms of the best of 32 to 16384 in each; one
// eighth of the grid per XCD 24.0-24.9 ms, no remapping 22.3-26.8 ms.  On
// smaller sets the bands cost time (the 64 jobs of the serving distribution,
// 26 K-485 K rows each: 174.8 ms of fills against 155.8 ms in launch order),
// so they take the launch order.
constexpr int kFillChunk = 512;
constexpr int64_t kFillBandRows = (int64_t)1 << 20;

struct Cplx {
    double re, im;
};

// complex product and sum without fused multiply-adds, as Fortran evaluates
// double complex arithmetic: (a+bi)(c+di) = (ac-bd) + (ad+bc)i
__device__ __forceinline__ Cplx cmul(Cplx a, Cplx b)
{
    return Cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ Cplx cadd(Cplx a, Cplx b) { return Cplx{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cplx cscale(double s, Cplx a) { return Cplx{s * a.re, s * a.im}; }

__device__ __forceinline__ int pair_of(const int32_t* c_first, int ks, int a, int c)
{
    const int lo = (c - ks + 1) < 1 ? 1 : (c - ks + 1);
    return gload(c_first + c) + (a - lo);
}

// the support window [lo, hi] of pair (a, c): the intervals where both
// splines are non-zero, [max(a,c) + 2 - k, min(a,c) + 1] within [1, n_iv]
__device__ __forceinline__ int window_lo(int a, int c, int ks)
{
    const int lo = (a > c ? a : c) + 2 - ks;
    return lo < 1 ? 1 : lo;
}
__device__ __forceinline__ int window_hi(int a, int c, int n_iv)
{
    const int hi = (a < c ? a : c) + 1;
    return hi > n_iv ? n_iv : hi;
}

// sum_k v[k * sv] coef[k] and sum_k (x[k * sx] y[k * sy]) coef[k], k
// ascending, each operation rounded; four multipoles' operands are loaded
// before they are combined, so that their loads overlap.
__device__ __forceinline__ double kdot2(const double* __restrict__ v, const double* __restrict__ coef, int K)
{
    double t = 0.0;
    int k = 0;
    for (; k + 4 <= K; k += 4) {
        const double v0 = gload(v + k), v1 = gload(v + k + 1), v2 = gload(v + k + 2), v3 = gload(v + k + 3);
        const double c0 = coef[k], c1 = coef[k + 1], c2 = coef[k + 2], c3 = coef[k + 3];
        t = t + v0 * c0;
        t = t + v1 * c1;
        t = t + v2 * c2;
        t = t + v3 * c3;
    }
    for (; k < K; ++k) t = t + gload(v + k) * coef[k];
    return t;
}
__device__ __forceinline__ double kdot3(const double* __restrict__ x, unsigned sx, const double* __restrict__ y,
                                        unsigned sy, const double* __restrict__ coef, int K)
{
    double t = 0.0;
    int k = 0;
    for (; k + 4 <= K; k += 4) {
        const unsigned u = (unsigned)k;
        const double x0 = gload(x + u * sx), x1 = gload(x + (u + 1) * sx), x2 = gload(x + (u + 2) * sx),
                     x3 = gload(x + (u + 3) * sx);
        const double y0 = gload(y + u * sy), y1 = gload(y + (u + 1) * sy), y2 = gload(y + (u + 2) * sy),
                     y3 = gload(y + (u + 3) * sy);
        const double c0 = coef[k], c1 = coef[k + 1], c2 = coef[k + 2], c3 = coef[k + 3];
        t = t + (x0 * y0) * c0;
        t = t + (x1 * y1) * c1;
        t = t + (x2 * y2) * c2;
        t = t + (x3 * y3) * c3;
    }
    for (; k < K; ++k) t = t + (gload(x + (unsigned)k * sx) * gload(y + (unsigned)k * sy)) * coef[k];
    return t;
}

// sum_k coef[k] R^k[p][q] for p = pair(a1, c1), q = pair(a2, c2).  When the
// windows of p and q do not overlap, R^k[p][q] is the rank-one product the
// fill kernel stored (Ptot_p Qtot_q if p's window lies left of q's, Qtot_p
// Ptot_q otherwise: the same single multiplication, so the same bits), read
// from the small multipole-major factor tables, which stay in L2 and put a
// wavefront's consecutive q on consecutive addresses; overlapping pairs read
// the tensor.  coef is the wavefront's copy of the channel coefficients in LDS.
template <class BA>
__device__ __forceinline__ double r12_row(const BA& A, int a1, int c1, int a2, int c2,
                                          const double* __restrict__ coef)
{
    const int K = A.K, ks = A.ks;
    const int p = pair_of(A.c_first, ks, a1, c1), q = pair_of(A.c_first, ks, a2, c2);
    const bool left = window_hi(a1, c1, A.n_iv) < window_lo(a2, c2, ks);
    const bool right = window_lo(a1, c1, ks) > window_hi(a2, c2, A.n_iv);
    if (!left && !right) return kdot2(A.R + ((size_t)p * A.n_pairs + q) * K, coef, K);
    const double* x = (left ? A.PtotT : A.QtotT) + p;
    const double* y = (left ? A.QtotT : A.PtotT) + q;
    const unsigned np = (unsigned)A.n_pairs;   // K * n_pairs < 2^31 (checked on the host)
    return kdot3(x, np, y, np, coef, K);
}

// The same sum for a row whose windows never meet (|n_a - n_b| >= k): every
// product is rank-one with a known side; x and y are the p and q factors in
// the wavefront's LDS copy, K consecutive multipoles from a 16-byte boundary,
// as is coef, so that one LDS read brings two multipoles.
__device__ __forceinline__ double r12_staged(const double* __restrict__ x, const double* __restrict__ y,
                                             const double* __restrict__ coef, int K)
{
    double t = 0.0;
    int k = 0;
    for (; k + 4 <= K; k += 4) {
        const dbl2 xa = *(const dbl2*)(x + k), xb = *(const dbl2*)(x + k + 2);
        const dbl2 ya = *(const dbl2*)(y + k), yb = *(const dbl2*)(y + k + 2);
        const dbl2 ca = *(const dbl2*)(coef + k), cb = *(const dbl2*)(coef + k + 2);
        t = t + (xa.x * ya.x) * ca.x;
        t = t + (xa.y * ya.y) * ca.y;
        t = t + (xb.x * yb.x) * cb.x;
        t = t + (xb.y * yb.y) * cb.y;
    }
    for (; k < K; ++k) t = t + (x[k] * y[k]) * coef[k];
    return t;
}

// element (r, c), 1-based, of a complex row-major n_b x n_b matrix
__device__ __forceinline__ Cplx mat(const double* M, int n_b, int r, int c)
{
    const size_t o = 2 * ((size_t)(r - 1) * n_b + (c - 1));
    const dbl2 v = gload((const dbl2*)(M + o));
    return Cplx{v.x, v.y};
}

// The one-electron parts of H and S at entry (n_a, n_b | n_c, n_d)
// (mat_els.f90:664-678 S_mat_neq, :697-715 H_1p_neq): the direct products
// where l_eq holds, the exchange products with the sign (-1)^(l_c+l_d+L)
// (odd_sign) where l_eq_ex holds, each operation rounded on its own.
template <class BA>
__device__ __forceinline__ void one_electron(const BA& A, int n_a, int n_bb, int n_c, int n_d, int l_a, int l_b,
                                             bool l_eq, bool l_eq_ex, bool odd_sign, Cplx& h1, Cplx& s1)
{
    const size_t nn = (size_t)A.n_b * A.n_b;
    const double* Ha = A.Hl + 2 * nn * l_a;
    const double* Hb = A.Hl + 2 * nn * l_b;
    const double sgn = odd_sign ? -1.0 : 1.0;
    h1 = Cplx{0.0, 0.0};
    s1 = Cplx{0.0, 0.0};
    if (l_eq) {
        h1 = cadd(cadd(h1, cmul(mat(Ha, A.n_b, n_a, n_c), mat(A.S, A.n_b, n_bb, n_d))),
                  cmul(mat(Hb, A.n_b, n_bb, n_d), mat(A.S, A.n_b, n_a, n_c)));
        s1 = cadd(s1, cmul(mat(A.S, A.n_b, n_a, n_c), mat(A.S, A.n_b, n_bb, n_d)));
    }
    if (l_eq_ex) {
        const Cplx x = cadd(cmul(mat(Ha, A.n_b, n_a, n_d), mat(A.S, A.n_b, n_bb, n_c)),
                            cmul(mat(Hb, A.n_b, n_bb, n_c), mat(A.S, A.n_b, n_a, n_d)));
        h1 = cadd(h1, cscale(sgn, x));
        s1 = cadd(s1, cmul(cscale(sgn, mat(A.S, A.n_b, n_a, n_d)), mat(A.S, A.n_b, n_bb, n_c)));
    }
}

__device__ __forceinline__ int64_t uniform64(int64_t x)
{
    const int lo = __builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const int hi = __builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
}

// LDS of one fill wavefront (doubles), in rows of KP = K rounded up to even:
// the channel pair's direct and exchange coefficients (2 rows), then the q
// factors of a row whose windows never meet, [2][L][KP] with L = 2k - 1:
// direct term pairs (d, n_b), exchange term pairs (d, n_a), for the d within
// k of n_b / n_a.
__host__ __device__ constexpr int fill_kp(int K) { return (K + 1) & ~1; }
// (plus 64 bytes for the channel walk's marks)
__host__ __device__ constexpr int fill_wave_lds(int ks, int K) { return 2 * fill_kp(K) * (1 + (2 * ks - 1)) + 8; }


typedef const __attribute__((address_space(4))) BlockArgs CBlockArgs;

// The fill of one configuration row i of block A by one wavefront, in
// general: a channel walk with ballot-placed entries.  wl is the wavefront's
// LDS (fill_wave_lds(lds_ks, lds_K) words).
__device__ __forceinline__ void fill_row_general(const CBlockArgs& A, int i, int n_a, int n_bb, int l_a, int l_b,
                                                 int ch_i, int lane, double* wl, int lds_ks, int lds_K)
{
    double* sc = wl;                                                       // coefficients: direct, exchange
    double* sy = sc + 2 * fill_kp(lds_K);                                  // staged q factors
    unsigned char* own = (unsigned char*)(sc + fill_wave_lds(lds_ks, lds_K) - 8);   // walk marks
    const int ks = A.ks, K = A.K;
    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int64_t nH = 0, nS = 0;
    const int64_t baseH = uniform64(A.ptrH[i]) - A.winH, baseS = uniform64(A.ptrS[i]) - A.winS;
    // |n_a - n_b| >= k: the direct pairs (n_a, n_c), (n_d, n_b) and the
    // exchange pairs (n_b, n_c), (n_d, n_a) have disjoint windows, on a side
    // fixed by the row: n_a < n_b puts every direct p left of its q
    // (Ptot_p Qtot_q) and every exchange p right of its q (Qtot_p Ptot_q).
    // The staged q factors serve the p side too: the window totals of (a, c)
    // and (c, a) are the same sums of the same local moments (one unordered
    // local pair per interval), so the direct p = (n_a, n_c) reads the
    // exchange q run at (n_c, n_a), and the exchange p = (n_b, n_c) the
    // direct q run at (n_c, n_b).
    const bool staged = n_bb - n_a >= ks || n_a - n_bb >= ks;
    const bool a_low = n_bb > n_a;
    const int L = 2 * ks - 1, KP = fill_kp(K);
    const unsigned np = (unsigned)A.n_pairs;   // K * n_pairs < 2^31 (checked on the host)
    const int dlo_dir = n_bb - ks + 1 < 1 ? 1 : n_bb - ks + 1, dlo_ex = n_a - ks + 1 < 1 ? 1 : n_a - ks + 1;
    if (staged) {
        // the q of the direct term are the pairs of column n_b, those of the
        // exchange term the pairs of column n_a: runs of consecutive pairs
        const int q_dir = A.c_first[n_bb], len_dir = A.c_first[n_bb + 1] - q_dir;
        const int q_ex = A.c_first[n_a], len_ex = A.c_first[n_a + 1] - q_ex;
        const double* src_dir = a_low ? A.QtotT : A.PtotT;
        const double* src_ex = a_low ? A.PtotT : A.QtotT;
        for (int e = lane; e < 2 * K * L; e += kWave) {   // read along the runs, store multipole-fastest
            const int t = e >= K * L, r = e - t * K * L, k = r / L, dl = r - k * L;
            if (dl < (t ? len_ex : len_dir))
                sy[(t * L + dl) * KP + k] = gload((t ? src_ex : src_dir) + ((unsigned)k * np + (unsigned)((t ? q_ex : q_dir) + dl)));
        }
    }

    // the channel table and flags through the scalar cache: a vector load
    // here would wait for the previous channel's stores
    typedef const __attribute__((address_space(4))) BlockChan CChan;
    typedef const __attribute__((address_space(4))) int32_t CInt;
    for (int ch = 0; ch < A.n_chan; ++ch) {
        const CChan& cjr = ((CChan*)A.chan)[ch];
        const BlockChan cj{cjr.l1, cjr.l2, cjr.start, 0};
        const int l_c = cj.l1, l_d = cj.l2;
        const int fl = ((CInt*)A.flags)[ch_i * A.n_chan + ch];
        const bool dir_any = fl & 1, ex_any = (fl >> 1) & 1;
        const bool l_eq = (l_a == l_c) && (l_b == l_d);
        const bool l_eq_ex = (l_a == l_d) && (l_b == l_c);
        if (!dir_any && !ex_any && !l_eq && !l_eq_ex) continue;   // no entry can exist in this channel
        {
            // the wavefront's lanes read the previous channel's copy in program
            // order before any lane overwrites it (one wavefront, in-order LDS)
            const double* cg = A.coef + ((size_t)ch_i * A.n_chan + ch) * 2 * K;
            for (int e = lane; e < 2 * K; e += kWave) sc[e < K ? e : e - K + KP] = gload(cg + e);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        const double* cd = sc;
        const double* ce = sc + KP;
        const BlockSeg* sg = A.seg + (size_t)ch * (A.n_b + 1);
        walk_channel<true>(n_a, n_bb, ks, A.n_b, sg, lane, own, [&](bool active, int n_c, int n_d, int col, bool in_dir_c,
                                                          bool in_ex_c) {
            const int j = cj.start + col;
            if (!A.full && j < i) active = false;
            const bool support = in_dir_c && iabs_d(n_bb - n_d) < ks;
            const bool support_ex = in_ex_c && iabs_d(n_a - n_d) < ks;
            const bool r12 = (dir_any && support) || (ex_any && support_ex);
            const bool entry = active && (support || support_ex) && (r12 || l_eq || l_eq_ex);
            const bool one_el = active && ((support && l_eq) || (support_ex && l_eq_ex));
            const unsigned long long mH = __ballot(entry);
            const unsigned long long mS = __ballot(one_el);
            {
                if (entry) {
                    Cplx h{0.0, 0.0};
                    if (r12) {
                        double res = 0.0;
                        // R(a,b,c,d) = R(a,d,c,b): the key [n_a, n_d, n_c, n_b] puts the
                        // wavefront's consecutive n_d on consecutive columns of one row
                        if (support)
                            res = res + (staged ? r12_staged(sy + (L + n_c - dlo_ex) * KP, sy + (n_d - dlo_dir) * KP,
                                                                     cd, K)
                                                        : r12_row(A, n_a, n_c, n_d, n_bb, cd));
                        // exchange R(a,b,d,c) = R(b,a,c,d) = R(b,d,c,a): row pair(n_b, n_c),
                        // consecutive n_d on consecutive columns
                        if (support_ex)
                            res = res + (staged ? r12_staged(sy + (n_c - dlo_dir) * KP, sy + (L + n_d - dlo_ex) * KP,
                                                                     ce, K)
                                                        : r12_row(A, n_bb, n_c, n_d, n_a, ce));
                        h.re = res;
                    }
                    if (one_el) {
                        Cplx h1, s1;
                        one_electron(A, n_a, n_bb, n_c, n_d, l_a, l_b, l_eq, l_eq_ex, ((l_c + l_d + A.L) & 1) != 0,
                                     h1, s1);
                        h = cadd(h, h1);
                        const int64_t ps = baseS + nS + __popcll(mS & lt_mask);
                        if (ps < A.limS) {
                            gstore(A.ixS + ps, (int64_t)(j + 1));
                            gstore((dbl2*)(A.dS + 2 * ps), dbl2{s1.re, s1.im});
                        }
                    }
                    const int64_t ph = baseH + nH + __popcll(mH & lt_mask);
                    if (ph < A.limH) {
                        gstore(A.ixH + ph, (int64_t)(j + 1));
                        gstore((dbl2*)(A.dH + 2 * ph), dbl2{h.re, h.im});
                    }
                }
            }
            nH += __popcll(mH);
            nS += __popcll(mS);
        });
    }
}

// Inclusive prefix sum over the wavefront's 64 lanes: within each row of 16
// lanes by DPP row shifts (a lane shifted in from outside the row adds 0),
// then the totals of the rows before, read from lanes 15, 31 and 47.
__device__ __forceinline__ int wave_incl_sum(int v, int lane)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);   // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);   // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);   // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);   // row_shr:8
    const int t0 = __builtin_amdgcn_readlane(v, 15), t1 = __builtin_amdgcn_readlane(v, 31);
    const int t2 = __builtin_amdgcn_readlane(v, 47);
    const int row = lane >> 4;
    return v + (row > 0 ? t0 : 0) + (row > 1 ? t1 : 0) + (row > 2 ? t2 : 0);
}

// The fill of the staged rows (|n_a - n_b| >= k) on the FP64 matrix cores.
// In such a row every R^k product is rank one with a side fixed by the row,
// and all its factors come from two runs of window totals: run 1, the pairs
// (d, n_b) for d in D1 (within k of n_b), and run 2, the pairs (d, n_a) for
// d in D2 (within k of n_a); the totals of (a, c) and (c, a) are the same
// sums.  Per column channel the direct term of entry (n_c, n_d) (support:
// n_c in D2, n_d in D1) and the exchange term (support_ex: n_c in D1, n_d in
// D2) are
//   D[n_c][n_d] = sum_k run2_k(n_c) cdir_k run1_k(n_d),
//   E[n_c][n_d] = sum_k run1_k(n_c) cex_k run2_k(n_d),
// each a block of one matrix product with inner dimension K, taken in 16 x 16
// tiles on v_mfma_f64_16x16x4 (NT x NT tiles per block, 2k - 1 <= 16 NT)
// from the runs in LDS, multipole-major and zero-padded to whole tiles and to
// a multiple of four multipoles.  When D1 and D2 overlap (k <= |n_a - n_b| <
// 2k - 1) an entry can have both supports; a third block, E at the direct
// block's coordinates (the runs read at shifted offsets), gives such an entry
// both terms in one lane.  A candidate n_c's H entries are one run of n_d
// (the union of its direct and exchange windows, as far as the channel's
// flags allow, clipped to the channel's n2 range and, for upper triangles, to
// j >= i), its S entries another, and a prefix sum over the candidates
// (lanes, in column order) places the runs in the row's CSR range, so an
// entry's position follows from its tile coordinates.  The products are
// summed in the matrix cores' order with fused multiply-adds, and the
// coefficient is applied to the p factor: the terms differ from the
// reference's (x y) c in the last bits.  The one-electron parts are the
// walked fill's, bit for bit.
__host__ __device__ constexpr int sep_kp(int K) { return 4 * ((K + 3) / 4); }
constexpr int kSepChunk = 4;   // channels whose n2 ranges and coefficients are staged at a time
constexpr int kSepSteps = 4;   // 4-multipole steps unrolled together
// LDS of one wavefront (8-byte words, 16-byte aligned parts): the two runs
// [2][KP][16 NT]; per staged channel the candidates' n2 ranges
// [kSepChunk][64] {off, lo | hi << 16} and the coefficients [kSepChunk][2][KP]
// (direct, exchange; zero past K); the current channel's candidate runs [64]
// {H lo * 64 + H len, H offset | S offset << 16, S lo * 64 + S len, column
// base} (16 bytes each)
// The staged fill's tiling: NT matrix-core tiles per window side (2k - 1 <=
// 16 NT), or kSepEdge for order 9 (2k - 1 = 17): one tile per side on the
// matrix cores and the 17th row and column as dot products, one entry per
// lane (two tiles per side would spend three quarters of their products
// outside the window); its runs hold 17 doubles per multipole.
constexpr int kSepEdge = 3;
__host__ __device__ constexpr int sep_ld(int nt) { return nt == kSepEdge ? 17 : 16 * nt; }
__host__ __device__ constexpr int sep_wave_lds(int K, int nt)
{
    return 2 * sep_kp(K) * sep_ld(nt) + kSepChunk * 64 + kSepChunk * 2 * sep_kp(K) + 2 * 64;
}

// X: the row's two windows may overlap (k <= |n_a - n_b| < 2k - 1), which
// takes the third block; rows further apart run without it (fewer registers,
// more wavefronts per SIMD at orders >= 9).
template <int NT, bool X>
__device__ __forceinline__ void fill_row_staged(const CBlockArgs& A, int i, int n_a, int n_bb, int l_a, int l_b,
                                                int ch_i, int lane, double* wl)
{
    typedef double d4_t __attribute__((ext_vector_type(4)));
    typedef int int4v __attribute__((ext_vector_type(4)));
    typedef int int2v __attribute__((ext_vector_type(2)));
    constexpr bool EDGE = NT == kSepEdge;   // one tile per side plus the 17th row and column
    constexpr int NM = EDGE ? 1 : NT;        // matrix-core tiles per side
    constexpr int LD = sep_ld(NT);   // doubles per multipole row of a run
    const int ks = A.ks, K = A.K, N = A.n_b;
    const int KP = sep_kp(K);
    double* run1 = wl;                                                  // [KP][LD]
    double* run2 = run1 + KP * LD;                                      // [KP][LD]
    int2v* sseg = (int2v*)(run2 + KP * LD) + lane;                      // [kSepChunk][64], this lane's column
    double* scoef = run2 + KP * LD + kSepChunk * kWave;                 // [kSepChunk][2][KP]
    int4v* sinfo = (int4v*)(scoef + kSepChunk * 2 * KP);                // [64]
    const bool a_low = n_bb > n_a;
    const int d1lo = n_bb - ks + 1 < 1 ? 1 : n_bb - ks + 1, d1hi = n_bb + ks - 1 > N ? N : n_bb + ks - 1;
    const int d2lo = n_a - ks + 1 < 1 ? 1 : n_a - ks + 1, d2hi = n_a + ks - 1 > N ? N : n_a + ks - 1;
    const int len1 = d1hi - d1lo + 1, len2 = d2hi - d2lo + 1;
    {
        // n_a < n_b: a direct p = (n_a, n_c) lies left of its q = (n_d, n_b)
        // (Ptot_p Qtot_q) and an exchange p = (n_b, n_c) right of its q =
        // (n_d, n_a) (Qtot_p Ptot_q): run 1 holds Qtot, run 2 Ptot; the
        // other way round for n_a > n_b
        const double* src1 = (a_low ? A.QtotT : A.PtotT) + gload(A.c_first + n_bb);
        const double* src2 = (a_low ? A.PtotT : A.QtotT) + gload(A.c_first + n_a);
        const unsigned np = (unsigned)A.n_pairs;   // K * n_pairs < 2^31 (checked on the host)
        // consecutive pairs on consecutive lanes; four loads in flight per
        // lane before their LDS writes
        for (int e0 = 0; e0 < 2 * KP * LD; e0 += 4 * kWave) {
            double v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int e = e0 + j * kWave + lane;
                const int r = e >= KP * LD, f = e - r * KP * LD, kk = f / LD, o = f - kk * LD;
                v[j] = 0.0;
                if (e < 2 * KP * LD && kk < K && o < (r ? len2 : len1))
                    v[j] = gload((r ? src2 : src1) + ((unsigned)kk * np + (unsigned)o));
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (e0 + j * kWave + lane < 2 * KP * LD) run1[e0 + j * kWave + lane] = v[j];   // run2 follows run1
        }
    }
    const int li = lane & 15, lk = lane >> 4;
    // candidates n_c in column order, one per lane: the lower range (D2 when
    // n_a < n_b), then the upper, as one run when they meet
    const int lo_lo = a_low ? d2lo : d1lo, lo_hi = a_low ? d2hi : d1hi;
    const int up_lo = a_low ? d1lo : d2lo, up_hi = a_low ? d1hi : d2hi;
    const bool merged = up_lo <= lo_hi + 1;
    const int lo_len = merged ? (up_hi > lo_hi ? up_hi : lo_hi) - lo_lo + 1 : lo_hi - lo_lo + 1;
    const int n_cand = merged ? lo_len : lo_len + (up_hi - up_lo + 1);
    const bool cand = lane < n_cand;
    const int my_c = lane < lo_len ? lo_lo + lane : up_lo + (lane - lo_len);
    const bool in2 = cand && iabs_d(my_c - n_a) < ks, in1 = cand && iabs_d(my_c - n_bb) < ks;
    // the first candidate lane of the direct block's rows (D2) and of the
    // exchange block's rows (D1)
    const int dir_base = a_low ? 0 : merged ? d2lo - d1lo : lo_len;
    const int ex_base = a_low ? (merged ? d1lo - d2lo : lo_len) : 0;
    const bool both = X && (d1lo > d2lo ? d1lo : d2lo) <= (d1hi < d2hi ? d1hi : d2hi);   // D1 and D2 overlap
    const int shift = d2lo - d1lo;   // the third block reads run 1 at m + shift and run 2 at n - shift
    const int64_t baseH = uniform64(A.ptrH[i]) - A.winH, baseS = uniform64(A.ptrS[i]) - A.winS;
    int64_t nH = 0, nS = 0;
    // Vector loads complete in order with the stores issued before them, so
    // a vector load after a channel's stores would wait for those: the
    // candidates' n2 ranges and the coefficients of kSepChunk channels are
    // staged in LDS before their stores, and the channel table and flags come
    // through the scalar cache.
    const int4v* segc = (const int4v*)A.seg + my_c;   // BlockSeg {off, lo, hi, pad} of (channel, n_c)
    typedef const __attribute__((address_space(4))) BlockChan CChan;
    typedef const __attribute__((address_space(4))) int32_t CInt;
    for (int ch = 0; ch < A.n_chan; ++ch) {
        const int slot = ch % kSepChunk;
        if (slot == 0) {   // all loads in flight before the first LDS write
            int4v qq[kSepChunk];
#pragma unroll
            for (int c = 0; c < kSepChunk; ++c) {   // (past the last channel: the last one again)
                const int cc = ch + c < A.n_chan ? ch + c : A.n_chan - 1;
                qq[c] = cand ? gload(segc + (size_t)cc * (N + 1)) : int4v{0, 1, 0, 0};
            }
#pragma unroll
            for (int c = 0; c < kSepChunk; ++c) sseg[c * kWave] = int2v{qq[c].x, (qq[c].y & 0xffff) | (qq[c].z << 16)};
            // [c][t][kk] from the coefficient table's [ch_i][ch + c][t][k]
            // (2 KP <= 128: two elements per lane and channel)
            const double* cgc = A.coef + ((size_t)ch_i * A.n_chan + ch) * 2 * K;
#pragma unroll 1
            for (int c0 = 0; c0 < kSepChunk; c0 += 2) {   // two channels' loads in flight at a time
                double v[2][2];
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        const int e = j * kWave + lane, t = e >= KP, kk = e - t * KP;
                        v[c][j] = 0.0;
                        if (e < 2 * KP && ch + c0 + c < A.n_chan && kk < K)
                            v[c][j] = gload(cgc + (size_t)(2 * (c0 + c) + t) * K + kk);
                    }
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        if (j * kWave + lane < 2 * KP) scoef[(c0 + c) * 2 * KP + j * kWave + lane] = v[c][j];
            }
            wave_lds_sync();   // the runs and the coefficients are read across lanes
        }
        const int2v qs = sseg[slot * kWave];
        const int qoff = qs.x, qlo = qs.y & 0xffff, qhi = qs.y >> 16;   // n_b < 2^15 (checked on the host)
        const CChan& cjr = ((CChan*)A.chan)[ch];
        const int c_l1 = cjr.l1, c_l2 = cjr.l2, c_start = cjr.start;
        const int fl = ((CInt*)A.flags)[ch_i * A.n_chan + ch];
        const bool dir_any = fl & 1, ex_any = (fl >> 1) & 1;
        const bool l_eq = (l_a == c_l1) && (l_b == c_l2);
        const bool l_eq_ex = (l_a == c_l2) && (l_b == c_l1);
        const bool hd = dir_any || l_eq || l_eq_ex, he = ex_any || l_eq || l_eq_ex;
        if (!hd && !he) continue;   // no entry can exist in this channel
        // this lane's candidate: its direct window D1 (if n_c is in D2) and
        // exchange window D2 (if n_c is in D1), clipped; the H run is the
        // union of those the channel's flags admit, the S run likewise (two
        // windows of one candidate overlap, so a union is one run); the
        // column of (n_c, n_d) is colb + n_d
        int hlo = 1, hlen = 0, slo = 1, slen = 0, colb = 0;
        if (cand) {
            int clo = qlo;
            if (!A.full) {   // column j = start + off + (n_d - lo2) >= i
                const int t = i - c_start - qoff + qlo;
                clo = t > clo ? t : clo;
            }
            const int dlo = d1lo > clo ? d1lo : clo, dhi = d1hi < qhi ? d1hi : qhi;
            const int elo = d2lo > clo ? d2lo : clo, ehi = d2hi < qhi ? d2hi : qhi;
            const bool dok = in2 && dhi >= dlo, eok = in1 && ehi >= elo;
            int a0 = 1 << 20, a1 = -1, b0 = 1 << 20, b1 = -1;
            if (dok && hd) a0 = dlo, a1 = dhi;
            if (eok && he) a0 = elo < a0 ? elo : a0, a1 = ehi > a1 ? ehi : a1;
            if (dok && l_eq) b0 = dlo, b1 = dhi;
            if (eok && l_eq_ex) b0 = elo < b0 ? elo : b0, b1 = ehi > b1 ? ehi : b1;
            if (a1 >= a0) hlo = a0, hlen = a1 - a0 + 1;
            if (b1 >= b0) slo = b0, slen = b1 - b0 + 1;
            colb = c_start + qoff - qlo;
        }
        // H and S counts in one prefix sum (a row's entries in one channel
        // number fewer than 2^16: at most 62 runs of at most 61)
        const int cHS = hlen | (slen << 16);
        const int inc = wave_incl_sum(cHS, lane);
        const int tot = __builtin_amdgcn_readlane(inc, kWave - 1);
        const int totH = tot & 0xffff, totS = tot >> 16;
        if (totH == 0) continue;
        // the candidates' runs for the tile lanes (the previous channel's
        // reads of this array precede these writes: one wavefront, in-order LDS)
        sinfo[lane] = int4v{hlo * 64 + hlen, inc - cHS, slo * 64 + slen, colb};
        wave_lds_sync();
        const bool odd = ((c_l1 + c_l2 + A.L) & 1) != 0;
        // the blocks in one pass over the multipoles: the direct block's
        // operands are run 2 x cdir (rows) and run 1 (columns), the exchange
        // block's run 1 x cex and run 2, so one LDS read of each run serves
        // both; the third block (overlapping windows only) reads the runs at
        // shifted offsets
        d4_t accd[NM][NM], acce[NM][NM], accx[NM][NM];
#pragma unroll
        for (int tm = 0; tm < NM; ++tm)
#pragma unroll
            for (int tn = 0; tn < NM; ++tn) accd[tm][tn] = acce[tm][tn] = accx[tm][tn] = d4_t{0.0, 0.0, 0.0, 0.0};
        if (dir_any || ex_any) {   // (a block without r12 is computed and not read)
            const double* cdir = scoef + slot * 2 * KP;
            const double* cex = cdir + KP;
            // one step: multipoles kk0 .. kk0 + 3, this lane's kk0 + lk
            auto step = [&](int kk) {
                double v1[NM], v2[NM];
#pragma unroll
                for (int t = 0; t < NM; ++t) {
                    v1[t] = run1[kk * LD + 16 * t + li];
                    v2[t] = run2[kk * LD + 16 * t + li];
                }
                const double cd = cdir[kk], ce = cex[kk];
#pragma unroll
                for (int tm = 0; tm < NM; ++tm)
#pragma unroll
                    for (int tn = 0; tn < NM; ++tn) {
                        accd[tm][tn] = __builtin_amdgcn_mfma_f64_16x16x4f64(v2[tm] * cd, v1[tn], accd[tm][tn], 0, 0, 0);
                        acce[tm][tn] = __builtin_amdgcn_mfma_f64_16x16x4f64(v1[tm] * ce, v2[tn], acce[tm][tn], 0, 0, 0);
                    }
            };
            auto step_x = [&](int kk) {   // E at direct coordinates: run 1 at n_c - d1lo, run 2 at n_d - d2lo
                double x1[NM], x2[NM];
#pragma unroll
                for (int t = 0; t < NM; ++t) {
                    const int a = 16 * t + li + shift, c = 16 * t + li - shift;
                    x1[t] = a >= 0 && a < LD ? run1[kk * LD + a] : 0.0;
                    x2[t] = c >= 0 && c < LD ? run2[kk * LD + c] : 0.0;
                }
                const double ce = cex[kk];
#pragma unroll
                for (int tm = 0; tm < NM; ++tm)
#pragma unroll
                    for (int tn = 0; tn < NM; ++tn)
                        accx[tm][tn] = __builtin_amdgcn_mfma_f64_16x16x4f64(x1[tm] * ce, x2[tn], accx[tm][tn], 0, 0, 0);
            };
            int kk0 = 0;
            for (; kk0 + 4 * kSepSteps <= KP; kk0 += 4 * kSepSteps) {   // whole groups of steps, unrolled
#pragma unroll
                for (int j = 0; j < kSepSteps; ++j) step(kk0 + 4 * j + lk);
            }
            for (; kk0 < KP; kk0 += 4) step(kk0 + lk);
            if (both)
                for (int k0 = 0; k0 < KP; k0 += 4) step_x(k0 + lk);
        }
        // result row lk + 4 r (candidate n_c), column li (n_d) of each tile.
        // The channel's outputs start at CSR positions hbase, sbase (uniform
        // pointers); an entry's position within them is a 32-bit offset.
        const int64_t hbase = baseH + nH, sbase = baseS + nS;
        const int64_t hroom = A.limH - hbase, sroom = A.limS - sbase;   // entries the outputs hold from there
        const int hlim = hroom < 0 ? 0 : hroom > 0x7fffffff ? 0x7fffffff : (int)hroom;
        const int slim = sroom < 0 ? 0 : sroom > 0x7fffffff ? 0x7fffffff : (int)sroom;
        int64_t* const ixh = A.ixH + hbase;
        dbl2* const dh = (dbl2*)A.dH + hbase;
        int64_t* const ixs = A.ixS + sbase;
        dbl2* const ds = (dbl2*)A.dS + sbase;
        // the direct block emits every entry with the direct support (adding
        // the exchange term where it also has that support), the exchange
        // block those with the exchange support alone
#pragma unroll 1
        for (int term = 0; term < 2; ++term) {
            const bool dir = term == 0;
            if (dir ? !(hd || (he && both)) : !he) continue;   // no entry of this block's kind
            const int rows = dir ? len2 : len1, cols = dir ? len1 : len2;
            const int row_lo = dir ? d2lo : d1lo, col_lo = dir ? d1lo : d2lo;
            const int cbase = dir ? dir_base : ex_base;
            int4v infos[NM][4];   // the runs of this lane's tile rows, read before any store
#pragma unroll
            for (int tm = 0; tm < NM; ++tm)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = 16 * tm + lk + 4 * r;
                    infos[tm][r] = sinfo[cbase + m < kWave ? cbase + m : kWave - 1];
                }
#pragma unroll
            for (int tm = 0; tm < NM; ++tm)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = 16 * tm + lk + 4 * r;
                    const int4v info = infos[tm][r];
                    const int h_lo = info.x >> 6, h_len = info.x & 63, s_lo = info.z >> 6;
                    const int n_c = row_lo + m;
#pragma unroll
                    for (int tn = 0; tn < NM; ++tn) {
                        const int n_d = col_lo + 16 * tn + li;
                        const int o = n_d - h_lo;
                        // the other support of this entry
                        const bool other = dir ? iabs_d(n_c - n_bb) < ks && iabs_d(n_d - n_a) < ks
                                               : iabs_d(n_c - n_a) < ks && iabs_d(n_d - n_bb) < ks;
                        // (a candidate's H run can pass this block's columns
                        // where its other window continues)
                        if (m < rows && 16 * tn + li < cols && o >= 0 && o < h_len && (dir || !other)) {
                            const bool sup = dir, sex = dir ? other : true;
                            const bool r12 = (dir_any && sup) || (ex_any && sex);
                            double res = 0.0;
                            if (r12) {
                                res = res + (dir ? accd[tm][tn][r] : acce[tm][tn][r]);
                                if (dir && other) res = res + accx[tm][tn][r];
                            }
                            Cplx h{res, 0.0};
                            const int64_t col = info.w + n_d + 1;
                            if ((sup && l_eq) || (sex && l_eq_ex)) {
                                Cplx h1, s1;
                                one_electron(A, n_a, n_bb, n_c, n_d, l_a, l_b, l_eq, l_eq_ex, odd, h1, s1);
                                h = cadd(h, h1);
                                const unsigned ps = (unsigned)((info.y >> 16) + (n_d - s_lo));
                                if ((int)ps < slim) {
                                    gstore(ixs + ps, col);
                                    gstore(ds + ps, dbl2{s1.re, s1.im});
                                }
                            }
                            const unsigned ph = (unsigned)((info.y & 0xffff) + o);
                            if ((int)ph < hlim) {
                                gstore(ixh + ph, col);
                                gstore(dh + ph, dbl2{h.re, h.im});
                            }
                        }
                    }
                }
            if constexpr (EDGE) {
                // the block's 17th row (m = 16, n = 0..16) and column (m =
                // 0..15, n = 16), one entry per lane: the block's sum (and the
                // direct block's exchange term at its coordinates) as fused
                // multiply-adds over the multipoles in order, then the entry as
                // in the tiles
                const int m = lane < 17 ? 16 : lane - 17, nn = lane < 17 ? lane : 16;
                if (lane < 33 && m < rows && nn < cols) {
                    const int4v info = sinfo[cbase + m < kWave ? cbase + m : kWave - 1];
                    const int h_lo = info.x >> 6, h_len = info.x & 63, s_lo = info.z >> 6;
                    const int n_c = row_lo + m, n_d = col_lo + nn;
                    const int o = n_d - h_lo;
                    const bool other = dir ? iabs_d(n_c - n_bb) < ks && iabs_d(n_d - n_a) < ks
                                           : iabs_d(n_c - n_a) < ks && iabs_d(n_d - n_bb) < ks;
                    if (o >= 0 && o < h_len && (dir || !other)) {
                        const bool sup = dir, sex = dir ? other : true;
                        const bool r12 = (dir_any && sup) || (ex_any && sex);
                        double res = 0.0;
                        if (r12) {
                            const double* cdir = scoef + slot * 2 * KP;
                            const double* cex = cdir + KP;
                            const bool third = X && dir && other && both;
                            const int xa = m + shift, xc = nn - shift;
                            double v = 0.0, vx = 0.0;
                            for (int k = 0; k < KP; ++k) {   // (runs and coefficients zero past K)
                                if (dir) v = fma(run2[k * LD + m] * cdir[k], run1[k * LD + nn], v);
                                else v = fma(run1[k * LD + m] * cex[k], run2[k * LD + nn], v);
                                if (third) {
                                    const double x1 = xa >= 0 && xa < LD ? run1[k * LD + xa] : 0.0;
                                    const double x2 = xc >= 0 && xc < LD ? run2[k * LD + xc] : 0.0;
                                    vx = fma(x1 * cex[k], x2, vx);
                                }
                            }
                            res = res + v;
                            if (dir && other) res = res + vx;
                        }
                        Cplx h{res, 0.0};
                        const int64_t col = info.w + n_d + 1;
                        if ((sup && l_eq) || (sex && l_eq_ex)) {
                            Cplx h1, s1;
                            one_electron(A, n_a, n_bb, n_c, n_d, l_a, l_b, l_eq, l_eq_ex, odd, h1, s1);
                            h = cadd(h, h1);
                            const unsigned ps = (unsigned)((info.y >> 16) + (n_d - s_lo));
                            if ((int)ps < slim) {
                                gstore(ixs + ps, col);
                                gstore(ds + ps, dbl2{s1.re, s1.im});
                            }
                        }
                        const unsigned ph = (unsigned)((info.y & 0xffff) + o);
                        if ((int)ph < hlim) {
                            gstore(ixh + ph, col);
                            gstore(dh + ph, dbl2{h.re, h.im});
                        }
                    }
                }
            }
        }
        nH += totH;
        nS += totS;
    }
}

// The fill of a row with |n_a - n_b| < k in tiles of cells.  The candidates
// n_c (within k of n_a or n_b) form one range U = [u0, u0 + m) (the two
// windows D2 = [n_a - k + 1, n_a + k - 1] and D1 = [n_b - k + 1, n_b + k - 1]
// overlap), and so do the n_d of every entry, so the row's entries lie in the
// grid U x U: cell (n_c, n_d) has the direct support when n_c is in D2 and
// n_d in D1, the exchange support when n_c is in D1 and n_d in D2.  The grid
// is taken 64 cells at a time, one per lane; each lane loads its cell's
// direct and exchange multipole vectors once (the tensor's element where the
// pairs' windows overlap, else the product of the window totals, which is the
// tensor's element bit for bit) into registers and keeps them while the
// column channels of a chunk are summed over, so that each tensor element is
// read once per chunk of channels rather than once per channel, and the
// loads of a tile are all in flight together.  Per channel, a candidate's H
// entries are one run of n_d (the union of its direct and exchange windows
// the channel's flags admit, clipped to the channel's n2 range and, for
// upper triangles, to j >= i), its S entries another; a prefix sum over the
// candidates, done for the chunk's channels before the tiles, gives each run
// its CSR offset.  The sums are the walked fill's (each product and sum
// rounded, multipoles ascending), so the entries are bit for bit the same.
// KM >= K multipoles are held per term; m <= 64 (k <= 16).
constexpr int kTChunk = 4;   // column channels per chunk
// ones for the second factor of a tensor entry read in a tile (the product
// with 1.0 is the entry, so every lane takes the same loads and products)
__device__ const double kOnes[64] = {
    1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
    1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
__host__ __device__ constexpr int tiled_kp(int km) { return (km + 1) & ~1; }
// LDS of one wavefront (8-byte words): per chunk channel the candidates'
// runs [kTChunk][64] {H lo | H len << 16, H offset | S offset << 16,
// S lo | S len << 16, column base} (16 bytes each), the coefficients
// [kTChunk][2][KP] (direct, exchange), and the channel's first H and S
// entries and flags [kTChunk][3]
__host__ __device__ constexpr int tiled_wave_lds(int km)
{
    return kTChunk * 64 * 2 + kTChunk * 2 * tiled_kp(km) + kTChunk * 3;
}

template <int KM>
__device__ __forceinline__ void fill_row_tiled(const CBlockArgs& A, int i, int n_a, int n_bb, int l_a, int l_b,
                                               int ch_i, int lane, double* wl)
{
    typedef int int4v __attribute__((ext_vector_type(4)));
    constexpr int KP = tiled_kp(KM);
    const int ks = A.ks, K = A.K, N = A.n_b;
    int4v* sinfo = (int4v*)wl;                                   // [kTChunk][64]
    double* scoef = wl + kTChunk * 64 * 2;                       // [kTChunk][2][KP]
    int64_t* schan = (int64_t*)(scoef + kTChunk * 2 * KP);       // [kTChunk][3]: H base, S base, flags
    const int lo_ = (n_a < n_bb ? n_a : n_bb) - ks + 1, hi_ = (n_a > n_bb ? n_a : n_bb) + ks - 1;
    const int u0 = lo_ < 1 ? 1 : lo_, u1 = hi_ > N ? N : hi_;
    const int m = u1 - u0 + 1;
    const int d1lo = n_bb - ks + 1 < 1 ? 1 : n_bb - ks + 1, d1hi = n_bb + ks - 1 > N ? N : n_bb + ks - 1;
    const int d2lo = n_a - ks + 1 < 1 ? 1 : n_a - ks + 1, d2hi = n_a + ks - 1 > N ? N : n_a + ks - 1;
    const int64_t baseH = uniform64(A.ptrH[i]) - A.winH, baseS = uniform64(A.ptrS[i]) - A.winS;
    const unsigned np = (unsigned)A.n_pairs;   // K * n_pairs < 2^31 (checked on the host)
    // this lane's candidate (the chunk passes) and its supports
    const int my_c = u0 + lane;
    const bool cand = lane < m;
    const bool c_in2 = cand && my_c >= d2lo && my_c <= d2hi, c_in1 = cand && my_c >= d1lo && my_c <= d1hi;
    const int4v* segc = (const int4v*)A.seg + (cand ? my_c : 0);   // BlockSeg {off, lo, hi, pad} of (channel, n_c)
    const int n_cells = m * m;
    // first(n) - lo(n) of the row's two splines and of this lane's candidate
    // (pair indices below); the tables' addresses in scalar registers
    const int fa = __builtin_amdgcn_readfirstlane(gload(A.c_first + n_a)) - (n_a - ks + 1 < 1 ? 1 : n_a - ks + 1);
    const int fb = __builtin_amdgcn_readfirstlane(gload(A.c_first + n_bb)) - (n_bb - ks + 1 < 1 ? 1 : n_bb - ks + 1);
    const int f_my = cand ? gload(A.c_first + my_c) - (my_c - ks + 1 < 1 ? 1 : my_c - ks + 1) : 0;
    const double* const tR = (const double*)uniform64((int64_t)A.R);
    const double* const tP = (const double*)uniform64((int64_t)A.PtotT);
    const double* const tQ = (const double*)uniform64((int64_t)A.QtotT);
    int64_t nH = 0, nS = 0;
    typedef const __attribute__((address_space(4))) BlockChan CChan;
    typedef const __attribute__((address_space(4))) int32_t CInt;
    for (int c0 = 0; c0 < A.n_chan; c0 += kTChunk) {
        const int nc = A.n_chan - c0 < kTChunk ? A.n_chan - c0 : kTChunk;
        // the chunk's runs and coefficients, every load before any LDS write
        int4v qq[kTChunk];
#pragma unroll
        for (int c = 0; c < kTChunk; ++c)
            qq[c] = cand && c < nc ? gload(segc + (size_t)(c0 + c) * (N + 1)) : int4v{0, 1, 0, 0};
        double cv[kTChunk][2];
        const double* cgc = A.coef + ((size_t)ch_i * A.n_chan + c0) * 2 * K;
#pragma unroll
        for (int c = 0; c < kTChunk; ++c)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int e = j * kWave + lane, t = e >= KP, kk = e - t * KP;
                cv[c][j] = (c < nc && e < 2 * KP && kk < K) ? gload(cgc + (size_t)(2 * c + t) * K + kk) : 0.0;
            }
        bool any_r12 = false;
#pragma unroll
        for (int c = 0; c < kTChunk; ++c) {
#pragma unroll
            for (int j = 0; j < 2; ++j)
                if (j * kWave + lane < 2 * KP) scoef[c * 2 * KP + j * kWave + lane] = cv[c][j];
            if (c >= nc) continue;
            const int ch = c0 + c;
            const CChan& cjr = ((CChan*)A.chan)[ch];
            const int c_l1 = cjr.l1, c_l2 = cjr.l2, c_start = cjr.start;
            const int fl = ((CInt*)A.flags)[ch_i * A.n_chan + ch];
            const bool dir_any = fl & 1, ex_any = (fl >> 1) & 1;
            const bool l_eq = (l_a == c_l1) && (l_b == c_l2);
            const bool l_eq_ex = (l_a == c_l2) && (l_b == c_l1);
            const bool hd = dir_any || l_eq || l_eq_ex, he = ex_any || l_eq || l_eq_ex;
            any_r12 |= (fl & 3) != 0;
            if (lane == 0) {   // the channel's first entries (this lane's candidate is not needed)
                schan[3 * c] = baseH + nH;
                schan[3 * c + 1] = baseS + nS;
                schan[3 * c + 2] = fl | (l_eq ? 4 : 0) | (l_eq_ex ? 8 : 0) | ((((c_l1 + c_l2 + A.L) & 1) != 0) ? 16 : 0);
            }
            const int qoff = qq[c].x, qlo = qq[c].y, qhi = qq[c].z;
            int hlo = 1, hlen = 0, slo = 1, slen = 0, colb = 0;
            if (cand && (hd || he)) {
                int clo = qlo;
                if (!A.full) {   // column j = start + off + (n_d - lo2) >= i
                    const int t = i - c_start - qoff + qlo;
                    clo = t > clo ? t : clo;
                }
                const int dlo = d1lo > clo ? d1lo : clo, dhi = d1hi < qhi ? d1hi : qhi;
                const int elo = d2lo > clo ? d2lo : clo, ehi = d2hi < qhi ? d2hi : qhi;
                const bool dok = c_in2 && dhi >= dlo, eok = c_in1 && ehi >= elo;
                int a0 = 1 << 20, a1 = -1, b0 = 1 << 20, b1 = -1;
                if (dok && hd) a0 = dlo, a1 = dhi;
                if (eok && he) a0 = elo < a0 ? elo : a0, a1 = ehi > a1 ? ehi : a1;
                if (dok && l_eq) b0 = dlo, b1 = dhi;
                if (eok && l_eq_ex) b0 = elo < b0 ? elo : b0, b1 = ehi > b1 ? ehi : b1;
                if (a1 >= a0) hlo = a0, hlen = a1 - a0 + 1;
                if (b1 >= b0) slo = b0, slen = b1 - b0 + 1;
                colb = c_start + qoff - qlo;
            }
            // H and S counts in one prefix sum (at most 64 runs of at most
            // 64 entries: below 2^16 each)
            const int cHS = hlen | (slen << 16);
            const int inc = wave_incl_sum(cHS, lane);
            const int tot = __builtin_amdgcn_readlane(inc, kWave - 1);
            const int ex = inc - cHS;
            sinfo[c * kWave + lane] = int4v{hlo | (hlen << 16), ex, slo | (slen << 16), colb};
            nH += tot & 0xffff;
            nS += tot >> 16;
        }
        wave_lds_sync();   // the runs, coefficients and channel entries are read across lanes
        for (int e0 = 0; e0 < n_cells; e0 += kWave) {
            const int e = e0 + lane;
            const int n_c = u0 + e / m, n_d = u0 + e % m;
            const bool live = e < n_cells;
            const bool sd = live && n_c >= d2lo && n_c <= d2hi && n_d >= d1lo && n_d <= d1hi;
            const bool se = live && n_c >= d1lo && n_c <= d1hi && n_d >= d2lo && n_d <= d2hi;
            // the cell's direct vector R^k[pair(n_a, n_c)][pair(n_d, n_b)] and
            // exchange vector R^k[pair(n_b, n_c)][pair(n_d, n_a)]
            double rd[KM], re[KM];
            // v[k] = R^k[pair(a1, n_c)][pair(n_d, a2)], k < K.  Straight-line
            // loads: every lane issues the same KM pairs (an index past K - 1
            // reads entry K - 1 again, a lane that is off reads the tensor's
            // first entries) and all of them are in flight before the first
            // product.  A pair of overlapping windows reads its row of the
            // tensor and ones (the product is the entry itself), a rank-one
            // pair its two factor columns.  Past K - 1 the values are finite
            // and meet coefficients of +0; an off lane's vector is not summed
            // (its cell has no two-electron part in any channel of the chunk).
            // pair(a1, n_c) = first(n_c) + a1 - lo(n_c), pair(n_d, a2) =
            // first(a2) + n_d - lo(a2), lo(c) = max(c - k + 1, 1)
            const int cc = live ? n_c : u0, dd = live ? n_d : u0;
            const int fc = __shfl(f_my, cc - u0, kWave);
            auto load_vec = [&](double(&v)[KM], bool on, int a1, int a2, int fq) {
                const int p = fc + a1, q = fq + dd;
                const bool left = window_hi(a1, cc, A.n_iv) < window_lo(dd, a2, ks);
                const bool right = window_lo(a1, cc, ks) > window_hi(dd, a2, A.n_iv);
                const bool nr = !left && !right;
                const double* x = !on ? tR : nr ? tR + ((size_t)p * A.n_pairs + q) * K : (left ? tP : tQ) + p;
                const double* y = !on || nr ? kOnes : (left ? tQ : tP) + q;
                const unsigned st = !on || nr ? 1u : np;
                double a[KM], b[KM];
#pragma unroll
                for (int k = 0; k < KM; ++k) {
                    const unsigned kk = (unsigned)(k < K ? k : K - 1) * st;
                    a[k] = gload(x + kk);
                    b[k] = gload(y + kk);
                }
#pragma unroll
                for (int k = 0; k < KM; ++k) v[k] = a[k] * b[k];
            };
            load_vec(rd, any_r12 && sd, n_a, n_bb, fb);
            __builtin_amdgcn_sched_barrier(0);   // the direct vector's loads issued before the exchange vector's
            load_vec(re, any_r12 && se, n_bb, n_a, fa);
            // the chunk's two-electron sums first (the vectors are dead after
            // them), then the one-electron parts and the stores; the sums run
            // over all KM terms without a branch (past K - 1 the coefficients
            // are +0: the products are zeros, which leave a sum begun at +0
            // as it is)
            double res[kTChunk];
#pragma unroll
            for (int c = 0; c < kTChunk; ++c) {
                res[c] = 0.0;
                if (c >= nc) continue;
                const int fl = (int)schan[3 * c + 2];
                const bool r12 = ((fl & 1) && sd) || (((fl >> 1) & 1) && se);
                if (!r12) continue;
                const double* cd = scoef + c * 2 * KP;
                const double* ce = cd + KP;
                double r = 0.0;
                if (sd) {
                    double t = 0.0;
#pragma unroll
                    for (int k = 0; k < KM; ++k) t = t + rd[k] * cd[k];
                    r = r + t;
                }
                if (se) {
                    double t = 0.0;
#pragma unroll
                    for (int k = 0; k < KM; ++k) t = t + re[k] * ce[k];
                    r = r + t;
                }
                res[c] = r;
            }
#pragma unroll
            for (int c = 0; c < kTChunk; ++c) {
                if (c >= nc) break;
                const int4v info = sinfo[c * kWave + (live ? n_c - u0 : 0)];
                const int h_lo = info.x & 0xffff, h_len = info.x >> 16;
                const int s_lo = info.z & 0xffff, s_len = info.z >> 16;
                const int oh = n_d - h_lo, os = n_d - s_lo;
                if (!live || oh < 0 || oh >= h_len) continue;   // no H entry here (and then no S entry)
                const int fl = (int)schan[3 * c + 2];
                // the sum is the entry's real part where r12 holds, zero otherwise
                Cplx h{res[c], 0.0};
                const int64_t col = (int64_t)info.w + n_d + 1;
                if (os >= 0 && os < s_len) {
                    Cplx h1, s1;
                    one_electron(A, n_a, n_bb, n_c, n_d, l_a, l_b, (fl >> 2) & 1, (fl >> 3) & 1, (fl >> 4) & 1, h1,
                                 s1);
                    h = cadd(h, h1);
                    const int64_t ps = schan[3 * c + 1] + (info.y >> 16) + os;
                    if (ps < A.limS) {
                        gstore(A.ixS + ps, col);
                        gstore((dbl2*)(A.dS + 2 * ps), dbl2{s1.re, s1.im});
                    }
                }
                const int64_t ph = schan[3 * c] + (info.y & 0xffff) + oh;
                if (ph < A.limH) {
                    gstore(A.ixH + ph, col);
                    gstore((dbl2*)(A.dH + 2 * ph), dbl2{h.re, h.im});
                }
            }
        }
        wave_lds_sync();   // the next chunk overwrites the runs and coefficients
    }
}

// One wavefront per configuration row of any of n_blocks symmetry blocks:
// wavefront w takes global row rows[w] (w < n), which belongs to the last
// block b with row0[b] <= row and is row row - row0[b] of it; each block's
// tables, inputs and outputs are in tab[b].  SEP: the rows are staged ones
// (|n_a - n_b| >= k; NT tiles per window side, 2k - 1 <= 16 NT), taken on the
// matrix cores; otherwise NT > 1: tiles of cells holding NT >= K multipoles
// (3k - 2 <= 64), NT = 1: the channel walk.
// The LDS (sep_wave_lds or fill_wave_lds words per wavefront) is dynamic; ks
// and K are the blocks' common spline order and multipole count.  Entries go
// to the row's CSR range from the count pass, taken relative to the first
// entry the outputs hold (winH, winS: the host launches only the rows of
// that window); stores beyond the outputs (limH, limS entries) are dropped,
// so a count that disagreed with the fill could not write past them.
// (the matrix-core launch with one tile per side is held to 128 vector
// registers, four wavefronts per SIMD, as many as its LDS allows; so are the
// tiles of at most 12 multipoles, while those of 18 hold 150 registers, three
// wavefronts per SIMD: at 128 they spill and run 6 % longer)
template <bool SEP, int NT, bool X = true>
__global__ __launch_bounds__(kCThreads, (SEP && (NT == 1 || NT == kSepEdge)) || (!SEP && NT > 1 && NT <= 12) ? 4 : (SEP && !X) || (!SEP && NT > 12) ? 3 : 1) void kc_block(const BlockArgs* __restrict__ tab,
                                                                             const int64_t* __restrict__ row0,
                                                                             int n_blocks,
                                                                             const int32_t* __restrict__ rows,
                                                                             int64_t n, int lds_ks, int lds_K)
{
    extern __shared__ __align__(16) double lds_fill[];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);
    // in a large launch the XCDs take adjacent bands of kFillChunk workgroups
    // (4 kFillChunk rows), so that each L2 and its address translation see
    // one band of the outputs at a time while the bands in flight stay close
    // together
    const int64_t lb = n >= kFillBandRows ? xcd_chunk_block(blockIdx.x, gridDim.x, kFillChunk) : (int64_t)blockIdx.x;
    const int64_t w = lb * (kCThreads / kWave) + wave;
    if (w >= n) return;   // whole wave exits together
    const int64_t g = rows[w];
    int blo = 0, bhi = n_blocks - 1;
    while (blo < bhi) {
        const int mid = (blo + bhi + 1) >> 1;
        if (row0[mid] <= g) blo = mid;
        else bhi = mid - 1;
    }
    const int b = __builtin_amdgcn_readfirstlane(blo);
    // The descriptor is read through the constant address space (the host
    // writes it before the launch; nothing in the launch writes it), so the
    // compiler may load a field again where it is used instead of keeping all
    // of them in scalar registers and spilling them to vector lanes.
    const CBlockArgs& A = ((const CBlockArgs*)tab)[b];
    const int i = (int)(g - row0[b]);
    const int32_t* ci = A.cfg + 2 * (size_t)i;
    const int cn = __builtin_amdgcn_readfirstlane(ci[0]), cl = __builtin_amdgcn_readfirstlane(ci[1]);
    const int n_a = cn & 0xffff, n_bb = cn >> 16, l_a = cl & 0xffff, l_b = cl >> 16;
    const int ch_i = __builtin_amdgcn_readfirstlane(A.cfg_chan[i]);
    if constexpr (SEP)
        fill_row_staged<NT, X>(A, i, n_a, n_bb, l_a, l_b, ch_i, lane, lds_fill + (size_t)wave * sep_wave_lds(lds_K, NT));
    else if constexpr (NT > 1)
        fill_row_tiled<NT>(A, i, n_a, n_bb, l_a, l_b, ch_i, lane, lds_fill + (size_t)wave * tiled_wave_lds(NT));
    else
        fill_row_general(A, i, n_a, n_bb, l_a, l_b, ch_i, lane, lds_fill + (size_t)wave * fill_wave_lds(lds_ks, lds_K),
                         lds_ks, lds_K);
}

// Row counts of the count pass, in closed form.  Within one column channel,
// candidate n_c (one per lane, within k of n_a or n_b, with configurations in
// the channel) owns two runs of n_d inside its n2 range: D, the n_d within k
// of n_b when |n_c - n_a| < k (direct support), and E, the n_d within k of
// n_a when |n_c - n_b| < k (exchange support); with full false both start at
// the first column j >= i.  The walk's predicates are unions of the runs:
//   H entry    D u E if l_eq or l_eq_ex, else (D if dir_any) u (E if ex_any)
//   S entry    (D if l_eq) u (E if l_eq_ex)
// and the reference's count_nnz (hamiltonian.f90:386-413), which tests the
// exchange term only where the direct support fails:
//   H          (D if dir_any) + (E \ D if ex_any)
//   S          (D if l_eq) + (E \ D if l_eq_ex)
__device__ __forceinline__ int run_len(int lo, int hi) { return hi >= lo ? hi - lo + 1 : 0; }

constexpr int kCountBatch = 8;   // channels whose n2 ranges one count step loads together
__global__ __launch_bounds__(kCThreads) void kc_count(const BlockArgs* __restrict__ tab, const int64_t* __restrict__ row0,
                                                      int n_blocks, int64_t n_rows)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);
    const int64_t g = (int64_t)blockIdx.x * (kCThreads / kWave) + wave;
    if (g >= n_rows) return;   // whole wave exits together
    int blo = 0, bhi = n_blocks - 1;
    while (blo < bhi) {
        const int mid = (blo + bhi + 1) >> 1;
        if (row0[mid] <= g) blo = mid;
        else bhi = mid - 1;
    }
    const int b = __builtin_amdgcn_readfirstlane(blo);
    // the descriptor, channel table and flags through the scalar cache (the
    // host writes them before the launch)
    typedef const __attribute__((address_space(4))) BlockChan CChan;
    typedef const __attribute__((address_space(4))) int32_t CInt;
    const CBlockArgs& A = ((const CBlockArgs*)tab)[b];
    const int i = (int)(g - row0[b]);
    const int ks = A.ks, n_b = A.n_b, n_chan = A.n_chan;
    const int32_t* ci = A.cfg + 2 * (size_t)i;
    const int cn = __builtin_amdgcn_readfirstlane(ci[0]), cl = __builtin_amdgcn_readfirstlane(ci[1]);
    const int n_a = cn & 0xffff, n_bb = cn >> 16, l_a = cl & 0xffff, l_b = cl >> 16;
    const int ch_i = __builtin_amdgcn_readfirstlane(A.cfg_chan[i]);
    const CandRanges cr = candidate_ranges(n_a, n_bb, ks, n_b);
    const bool cand = lane < cr.n1 + cr.n2;
    const int n_c = lane < cr.n1 ? cr.r1lo + lane : cr.r2lo + (lane - cr.n1);
    const bool in_dir_c = cand && iabs_d(n_c - n_a) < ks, in_ex_c = cand && iabs_d(n_c - n_bb) < ks;
    typedef int int4v __attribute__((ext_vector_type(4)));
    const int4v* segc = (const int4v*)A.seg + (cand ? n_c : 1);   // BlockSeg {off, lo, hi, pad}; (a valid column
                                                                   // without a candidate)
    int h = 0, s = 0, rh = 0, rs = 0;
    // kCountBatch channels at a time, their n2 ranges loaded together (past
    // the last channel: the last one again, not counted)
    for (int ch0 = 0; ch0 < n_chan; ch0 += kCountBatch) {
        int4v sgv[kCountBatch];
#pragma unroll
        for (int j = 0; j < kCountBatch; ++j) {
            const int ch = ch0 + j < n_chan ? ch0 + j : n_chan - 1;
            sgv[j] = gload(segc + (size_t)ch * (n_b + 1));
        }
#pragma unroll
        for (int j = 0; j < kCountBatch; ++j) {
            const int ch = ch0 + j;
            if (ch >= n_chan) break;
            const CChan& cj = ((CChan*)A.chan)[ch];
            const int fl = ((CInt*)A.flags)[ch_i * n_chan + ch];
            const bool dir_any = fl & 1, ex_any = (fl >> 1) & 1;
            const bool l_eq = (l_a == cj.l1) && (l_b == cj.l2);
            const bool l_eq_ex = (l_a == cj.l2) && (l_b == cj.l1);
            if (!dir_any && !ex_any && !l_eq && !l_eq_ex) continue;   // no entry can exist in this channel
            const int sg_off = sgv[j].x, sg_lo = sgv[j].y, sg_hi = sgv[j].z;
            if (!cand || sg_hi < sg_lo) continue;
            int lo = sg_lo;
            if (!A.full) {   // column j = start + off + (n_d - lo2) >= i
                const int t = i - cj.start - sg_off + sg_lo;
                lo = t > lo ? t : lo;
            }
            const int dlo = in_dir_c ? (lo > n_bb - ks + 1 ? lo : n_bb - ks + 1) : 1;
            const int dhi = in_dir_c ? (sg_hi < n_bb + ks - 1 ? sg_hi : n_bb + ks - 1) : 0;
            const int elo = in_ex_c ? (lo > n_a - ks + 1 ? lo : n_a - ks + 1) : 1;
            const int ehi = in_ex_c ? (sg_hi < n_a + ks - 1 ? sg_hi : n_a + ks - 1) : 0;
            const int D = run_len(dlo, dhi), E = run_len(elo, ehi);
            const int DE = run_len(dlo > elo ? dlo : elo, dhi < ehi ? dhi : ehi);
            const bool hd = dir_any || l_eq || l_eq_ex, he = ex_any || l_eq || l_eq_ex;
            h += (hd ? D : 0) + (he ? E : 0) - (hd && he ? DE : 0);
            s += (l_eq ? D : 0) + (l_eq_ex ? E : 0) - (l_eq && l_eq_ex ? DE : 0);
            rh += (dir_any ? D : 0) + (ex_any ? E - DE : 0);
            rs += (l_eq ? D : 0) + (l_eq_ex ? E - DE : 0);
        }
    }
    for (int off = kWave / 2; off > 0; off >>= 1) {
        h += __shfl_xor(h, off, kWave);
        s += __shfl_xor(s, off, kWave);
        rh += __shfl_xor(rh, off, kWave);
        rs += __shfl_xor(rs, off, kWave);
    }
    if (lane == 0) {
        A.cntH[i] = h;
        A.cntS[i] = s;
        A.refH[i] = rh;
        A.refS[i] = rs;
    }
}

// One thread per element: HlT/ST take the transposes of the n_l column-major
// complex matrices Hl and of S; PtotT/QtotT the transposes of Ptot/Qtot.
__global__ __launch_bounds__(256) void kc_prepare(const double* __restrict__ Hl, const double* __restrict__ S, int n_l,
                                                  int n_b, const double* __restrict__ Ptot,
                                                  const double* __restrict__ Qtot, int n_pairs, int K,
                                                  double* __restrict__ HlT, double* __restrict__ ST,
                                                  double* __restrict__ PtotT, double* __restrict__ QtotT)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t nn = (int64_t)n_b * n_b, nm = (int64_t)(n_l + 1) * nn, nf = (int64_t)n_pairs * K;
    if (e < nm) {
        const int64_t l = e / nn, rc = e - l * nn;
        const int64_t r = rc / n_b, c = rc - r * n_b;          // destination (r, c), row-major
        const double* src = (l < n_l ? Hl + 2 * l * nn : S) + 2 * (c * n_b + r);
        double* dst = (l < n_l ? HlT + 2 * l * nn : ST) + 2 * rc;
        dst[0] = src[0];
        dst[1] = src[1];
    } else if (e < nm + 2 * nf) {
        const int64_t f = e - nm, w = f / nf, kq = f - w * nf;
        const int64_t k = kq / n_pairs, q = kq - k * n_pairs;  // destination [k][q]
        (w == 0 ? PtotT : QtotT)[kq] = (w == 0 ? Ptot : Qtot)[q * K + k];
    }
}

hipError_t launch_block_prepare(const double* Hl, const double* S, int n_l, int n_b, const double* Ptot,
                                const double* Qtot, int n_pairs, int K, double* HlT, double* ST, double* PtotT,
                                double* QtotT, hipStream_t s)
{
    const int64_t n = (int64_t)(n_l + 1) * n_b * n_b + 2 * (int64_t)n_pairs * K;
    hipLaunchKernelGGL(kc_prepare, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Hl, S, n_l, n_b, Ptot, Qtot,
                       n_pairs, K, HlT, ST, PtotT, QtotT);
    return hipGetLastError();
}

hipError_t launch_block_count(const BlockArgs* tab, const int64_t* row0, int n_blocks, int64_t n_rows, hipStream_t s)
{
    const int64_t threads = n_rows * kWave;
    hipLaunchKernelGGL(kc_count, dim3((unsigned)((threads + kCThreads - 1) / kCThreads)), dim3(kCThreads), 0, s,
                       tab, row0, n_blocks, n_rows);
    return hipGetLastError();
}

// Exclusive prefix sums of the count pass's row counts on the device (the
// fill's row offsets: ptr[r] = sum of cnt[0 .. r) for H, and at ptr + n for
// S), so that a captured graph can count and fill without the host: tiles of
// kScanTile rows are summed and scanned in place (kc_scan_tiles), the tile
// totals scanned by one workgroup (kc_scan_totals), and each tile's offset
// added (kc_scan_add).  tmp holds 2 * ceil(n / kScanTile) int64.
constexpr int kScanTile = 1024;

__device__ __forceinline__ int64_t wave_incl_scan(int64_t v, int lane)
{
    for (int off = 1; off < kWave; off <<= 1) {
        const int64_t y = __shfl_up(v, off, kWave);
        if (lane >= off) v += y;
    }
    return v;
}

// workgroup-wide exclusive scan of one value per thread (256 threads)
__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t* sw, int64_t* total)
{
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int64_t incl = wave_incl_scan(v, lane);
    if (lane == kWave - 1) sw[wave] = incl;
    __syncthreads();
    int64_t before = 0, all = 0;
    for (int w = 0; w < kCThreads / kWave; ++w) {
        before += w < wave ? sw[w] : 0;
        all += sw[w];
    }
    __syncthreads();
    *total = all;
    return before + incl - v;
}

__global__ __launch_bounds__(kCThreads) void kc_scan_tiles(const int64_t* __restrict__ cnt, int64_t* __restrict__ ptr,
                                                           int64_t n, int64_t* __restrict__ tile_tot)
{
    __shared__ int64_t sw[kCThreads / kWave];
    const int64_t n_tiles = (n + kScanTile - 1) / kScanTile;
    const int which = (int)(blockIdx.x / n_tiles);        // 0: H, 1: S
    const int64_t tile = blockIdx.x - which * n_tiles;
    const int64_t* c = cnt + which * n;
    int64_t* o = ptr + which * n;
    const int per = kScanTile / kCThreads;                 // rows per thread
    const int64_t r0 = tile * kScanTile + (int64_t)threadIdx.x * per;
    int64_t v[kScanTile / kCThreads];
    int64_t mine = 0;
#pragma unroll
    for (int j = 0; j < per; ++j) {
        v[j] = r0 + j < n ? c[r0 + j] : 0;
        mine += v[j];
    }
    int64_t total = 0;
    int64_t run = block_excl_scan(mine, sw, &total);
#pragma unroll
    for (int j = 0; j < per; ++j) {
        if (r0 + j < n) o[r0 + j] = run;
        run += v[j];
    }
    if (threadIdx.x == 0) tile_tot[which * n_tiles + tile] = total;
}

__global__ __launch_bounds__(kCThreads) void kc_scan_totals(int64_t* __restrict__ tile_tot, int64_t n_tiles)
{
    __shared__ int64_t sw[kCThreads / kWave];
    for (int which = 0; which < 2; ++which) {
        int64_t* t = tile_tot + which * n_tiles;
        int64_t carry = 0;
        for (int64_t base = 0; base < n_tiles; base += kCThreads) {
            const int64_t e = base + threadIdx.x;
            const int64_t v = e < n_tiles ? t[e] : 0;
            int64_t total = 0;
            const int64_t ex = block_excl_scan(v, sw, &total);
            if (e < n_tiles) t[e] = carry + ex;
            carry += total;
        }
    }
}

__global__ __launch_bounds__(kCThreads) void kc_scan_add(int64_t* __restrict__ ptr, int64_t n,
                                                         const int64_t* __restrict__ tile_tot)
{
    const int64_t n_tiles = (n + kScanTile - 1) / kScanTile;
    const int64_t e = (int64_t)blockIdx.x * kCThreads + threadIdx.x;
    if (e >= 2 * n) return;
    const int which = e >= n;
    const int64_t r = e - which * n;
    ptr[e] += tile_tot[which * n_tiles + r / kScanTile];
}

int64_t row_scan_tmp_elems(int64_t n_rows) { return 2 * ((n_rows + kScanTile - 1) / kScanTile); }

hipError_t launch_row_scan(const int64_t* cnt, int64_t* ptr, int64_t n_rows, int64_t* tmp, hipStream_t s)
{
    if (n_rows <= 0) return hipSuccess;
    const int64_t n_tiles = (n_rows + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(kc_scan_tiles, dim3((unsigned)(2 * n_tiles)), dim3(kCThreads), 0, s, cnt, ptr, n_rows, tmp);
    hipLaunchKernelGGL(kc_scan_totals, dim3(1), dim3(kCThreads), 0, s, tmp, n_tiles);
    hipLaunchKernelGGL(kc_scan_add, dim3((unsigned)((2 * n_rows + kCThreads - 1) / kCThreads)), dim3(kCThreads), 0, s,
                       ptr, n_rows, tmp);
    return hipGetLastError();
}

// Both launches on s, the walked rows first.  (Measured on the c3 set: the
// same fill time as with the walked rows on a second stream beside the
// matrix-core launch, or after it; the two kernels share the memory system.)
hipError_t launch_block_fill(const BlockArgs* tab, const int64_t* row0, int n_blocks, const int32_t* rows_far,
                             int64_t n_far, const int32_t* rows_sep, int64_t n_sep, const int32_t* rows_gen,
                             int64_t n_gen, int ks, int K, hipStream_t s)
{
    // staged tiles per window side: one (k <= 8), one and the 17th row and
    // column (k = 9), two (k >= 10)
    const int nt = 2 * ks - 1 <= 16 ? 1 : 2 * ks - 1 == 17 ? kSepEdge : 2;
    hipError_t e = hipSuccess;
    if (n_gen > 0) {
        const unsigned grid = (unsigned)((n_gen * kWave + kCThreads - 1) / kCThreads);
        // tiles of cells where the multipoles fit the registers, else the walk
        // (KM = 18 holds 150 vector registers, three wavefronts per SIMD;
        // more multipoles would spill)
        const int km = 3 * ks - 2 > kWave ? 0 : K <= 12 ? 12 : K <= 18 ? 18 : 0;
        const void* fn = km == 12 ? (const void*)kc_block<false, 12>
                         : km == 18 ? (const void*)kc_block<false, 18> : (const void*)kc_block<false, 1>;
        const size_t lds = sizeof(double) * (size_t)(kCThreads / kWave) *
                           (size_t)(km ? tiled_wave_lds(km) : fill_wave_lds(ks, K));
        e = allow_lds(fn, lds);   // <= 4 x 32 KiB (k <= 16, K <= 64)
        if (e != hipSuccess) return e;
        if (km == 12)
            hipLaunchKernelGGL((kc_block<false, 12>), dim3(grid), dim3(kCThreads), lds, s, tab, row0, n_blocks,
                               rows_gen, n_gen, ks, K);
        else if (km == 18)
            hipLaunchKernelGGL((kc_block<false, 18>), dim3(grid), dim3(kCThreads), lds, s, tab, row0, n_blocks,
                               rows_gen, n_gen, ks, K);
        else
            hipLaunchKernelGGL((kc_block<false, 1>), dim3(grid), dim3(kCThreads), lds, s, tab, row0, n_blocks,
                               rows_gen, n_gen, ks, K);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    // the staged rows: those whose windows cannot overlap (|n1 - n2| >=
    // 2k - 1, rows_far) without the third block, then the others (rows_sep)
    auto staged = [&](auto kern, const int32_t* rws, int64_t n) {
        const unsigned grid = (unsigned)((n * kWave + kCThreads - 1) / kCThreads);
        const size_t lds = sizeof(double) * (size_t)(kCThreads / kWave) * (size_t)sep_wave_lds(K, nt);
        hipError_t le = allow_lds((const void*)kern, lds);   // <= 4 x 39 KiB (k <= 16, K <= 64)
        if (le != hipSuccess) return le;
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kCThreads), lds, s, tab, row0, n_blocks, rws, n, ks, K);
        return hipGetLastError();
    };
    if (n_far > 0 && e == hipSuccess)
        e = nt == 1 ? staged(kc_block<true, 1, false>, rows_far, n_far)
            : nt == kSepEdge ? staged(kc_block<true, kSepEdge, false>, rows_far, n_far)
                             : staged(kc_block<true, 2, false>, rows_far, n_far);
    if (n_sep > 0 && e == hipSuccess)
        e = nt == 1 ? staged(kc_block<true, 1, true>, rows_sep, n_sep)
            : nt == kSepEdge ? staged(kc_block<true, kSepEdge, true>, rows_sep, n_sep)
                             : staged(kc_block<true, 2, true>, rows_sep, n_sep);
    return e;
}

}  // namespace bs2e
