// channel_walk.h -- the candidate columns of one configuration row within
// one column channel, spread over a wavefront.
//
// Row (n_a, n_b); a column (n_c, n_d) of the channel can only carry an entry
// when its radial pairs share support: |n_a - n_c| < k and |n_b - n_d| < k
// (direct) or |n_a - n_d| < k and |n_b - n_c| < k (exchange).  The n_c
// within k of n_a or n_b each own one or two runs of n_d, clipped to the
// channel's n2 range for that n1 = n_c; the runs are laid end to end across
// the wavefront (prefix sum over lanes) in (n_c, n_d) order, which is
// ascending column order, and visited 64 at a time: visit(active, n_c, n_d,
// column offset within the channel, in_dir_c, in_ex_c) is called by every
// lane of the wavefront once per step (so that it may ballot).
#pragma once

#include "bs2e_internal.h"

namespace bs2e {

__device__ __forceinline__ int iabs_d(int x) { return x < 0 ? -x : x; }

// The n_c candidates of row (n_a, n_b): within k of n_a (direct) or of n_b
// (exchange), ascending, as one or two runs [r1lo, r1lo + n1) and
// [r2lo, r2lo + n2) inside [1, n_b]; n1 + n2 <= 2 (2k - 1) <= 62 for k <= 16.
struct CandRanges {
    int r1lo, n1, r2lo, n2;
};
__device__ __forceinline__ CandRanges candidate_ranges(int n_a, int n_bb, int ks, int n_b)
{
    int r1lo = n_a - ks + 1, r1hi = n_a + ks - 1, r2lo = n_bb - ks + 1, r2hi = n_bb + ks - 1;
    if (r2lo < r1lo) {
        int t0 = r1lo, t1 = r1hi;
        r1lo = r2lo;
        r1hi = r2hi;
        r2lo = t0;
        r2hi = t1;
    }
    if (r2lo <= r1hi + 1) {       // overlapping or adjacent: one range
        r1hi = r2hi > r1hi ? r2hi : r1hi;
        r2lo = 1;
        r2hi = 0;
    }
    r1lo = r1lo < 1 ? 1 : r1lo;
    r1hi = r1hi > n_b ? n_b : r1hi;
    r2lo = r2lo < 1 ? 1 : r2lo;
    r2hi = r2hi > n_b ? n_b : r2hi;
    return CandRanges{r1lo, r1hi >= r1lo ? r1hi - r1lo + 1 : 0, r2lo, r2hi >= r2lo ? r2hi - r2lo + 1 : 0};
}

// Orders a wavefront's LDS writes before its lanes read what other lanes
// wrote (one wavefront: no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// sg: the channel's n2 ranges per n1 (index 1..n_b); k <= 16.  own: 64 bytes
// of the wavefront's LDS for the owner marks (MARKS), else unused: each lane
// then finds its candidate by a binary search over the run starts.
template <bool MARKS, class Visit>
__device__ __forceinline__ void walk_channel(int n_a, int n_bb, int ks, int n_b, const BlockSeg* __restrict__ sg,
                                             int lane, unsigned char* __restrict__ own, Visit&& visit)
{
    constexpr int kWave = 64;
    const CandRanges cr = candidate_ranges(n_a, n_bb, ks, n_b);
    const int r1lo = cr.r1lo, r2lo = cr.r2lo, n1 = cr.n1;
    const int ncand = cr.n1 + cr.n2;   // one per lane
    // lane l takes the l-th candidate n_c and its n_d runs: direct needs
    // |n_b - n_d| < k, exchange |n_a - n_d| < k, both clipped to the
    // channel's n2 range for that n1 = n_c
    int my_c = 0, d1lo = 1, d1hi = 0, d2lo = 1, d2hi = 0, my_soff = 0, my_slo = 0;
    if (lane < ncand) {
        my_c = lane < n1 ? r1lo + lane : r2lo + (lane - n1);
        typedef int int4v __attribute__((ext_vector_type(4)));
        const int4v q = gload((const int4v*)(sg + my_c));   // BlockSeg {off, lo, hi, pad}
        const BlockSeg sgc{q.x, q.y, q.z, q.w};
        if (sgc.hi >= sgc.lo) {
            const bool in_dir_c = iabs_d(my_c - n_a) < ks;
            const bool in_ex_c = iabs_d(my_c - n_bb) < ks;
            int s1lo = 1, s1hi = 0, s2lo = 1, s2hi = 0;
            if (in_dir_c) {
                s1lo = n_bb - ks + 1;
                s1hi = n_bb + ks - 1;
            }
            if (in_ex_c) {
                s2lo = n_a - ks + 1;
                s2hi = n_a + ks - 1;
            }
            if (!in_dir_c) {
                s1lo = s2lo;
                s1hi = s2hi;
                s2lo = 1;
                s2hi = 0;
            } else if (in_ex_c) {
                if (s2lo < s1lo) {
                    int t0 = s1lo, t1 = s1hi;
                    s1lo = s2lo;
                    s1hi = s2hi;
                    s2lo = t0;
                    s2hi = t1;
                }
                if (s2lo <= s1hi + 1) {
                    s1hi = s2hi > s1hi ? s2hi : s1hi;
                    s2lo = 1;
                    s2hi = 0;
                }
            }
            d1lo = s1lo < sgc.lo ? sgc.lo : s1lo;
            d1hi = s1hi > sgc.hi ? sgc.hi : s1hi;
            d2lo = s2lo < sgc.lo ? sgc.lo : s2lo;
            d2hi = s2hi > sgc.hi ? sgc.hi : s2hi;
            my_soff = sgc.off;
            my_slo = sgc.lo;
        }
    }
    const int len1 = d1hi >= d1lo ? d1hi - d1lo + 1 : 0, len2 = d2hi >= d2lo ? d2hi - d2lo + 1 : 0;
    // exclusive prefix over the lanes of the candidates' run lengths
    int incl = len1 + len2;
    for (int off = 1; off < kWave; off <<= 1) {
        const int y = __shfl_up(incl, off, kWave);
        if (lane >= off) incl += y;
    }
    const int total = __shfl(incl, kWave - 1, kWave);
    const int excl = incl - len1 - len2;
    // candidates flattened in (n_c, n_d) order, which is ascending column
    // order.  Lane f of a window [f0, f0 + 64) belongs to the last candidate
    // whose (non-empty) run starts at or before f.  With MARKS the runs
    // starting inside the window mark their first position in own[] with
    // their lane + 1, a ballot of the marks gives each lane the last start at
    // or below it, and a lane before every start continues the run of the
    // window's first position (c0, the owner of the previous window's last
    // lane).
    const bool starts = ncand > lane && len1 + len2 > 0;
    int c0 = 0;
    for (int f0 = 0; f0 < total; f0 += kWave) {
        const int f = f0 + lane;
        int pos = 0;
        if (MARKS) {
            own[lane] = 0;
            wave_lds_sync();
            if (starts && excl >= f0 && excl < f0 + kWave) own[excl - f0] = (unsigned char)(lane + 1);
            wave_lds_sync();
            const int mark = own[lane];
            const unsigned long long marks = __ballot(mark != 0);
            const unsigned long long le = marks & (lane == kWave - 1 ? ~0ull : ((2ull << lane) - 1));
            const int last = 63 - __builtin_clzll(le | 1ull);   // highest mark at or below this lane (0 if none)
            const int owner = __shfl(mark, last, kWave);
            pos = le ? owner - 1 : c0;
            c0 = __shfl(pos, kWave - 1, kWave);
        } else {
            for (int step = kWave / 2; step > 0; step >>= 1) {
                const int cand = pos + step;
                const int e = __shfl(excl, cand < kWave ? cand : kWave - 1, kWave);
                if (cand < ncand && e <= f) pos = cand;
            }
        }
        const int n_c = __shfl(my_c, pos, kWave);
        const int o_lo1 = __shfl(d1lo, pos, kWave), o_len1 = __shfl(len1, pos, kWave);
        const int o_lo2 = __shfl(d2lo, pos, kWave), o_ex = __shfl(excl, pos, kWave);
        const int s_off = __shfl(my_soff, pos, kWave), s_lo = __shfl(my_slo, pos, kWave);
        const int local = f - o_ex;
        const int n_d = local < o_len1 ? o_lo1 + local : o_lo2 + (local - o_len1);
        const bool in_dir_c = iabs_d(n_c - n_a) < ks;
        const bool in_ex_c = iabs_d(n_c - n_bb) < ks;
        visit(f < total, n_c, n_d, s_off + (n_d - s_lo), in_dir_c, in_ex_c);
    }
}

}  // namespace bs2e
