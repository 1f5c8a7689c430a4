// bs2e_internal.h -- shared declarations of the MI355X R^k builder.
//
// Index conventions (the reference's, edvinolo/b-spline-two-e):
//   * spline index s is 1-based, s in [1, n]; basis index b in [1, n_b] is
//     spline b+1 (bspline_tools.f90:33-34, mat_els.f90:208-210);
//   * interval iv in [1, n_iv], breakpoints t(k .. size(t)-k+1)
//     (bspline_tools.f90:36); spline s is supported on iv iff
//     iv <= s <= iv+k-1 (bspline_tools.f90:42-56);
//   * a banded pair (a,c), |a-c| < k, is numbered in the loop order of
//     count_nnz_R_k (sparse_array_tools.f90:347-366): c outer, a inner.
//
// Device layouts (all FP64, row-major, 0-based):
//   Pl, Ql [n_iv][U][K]             local pair moments P^k, Q^k of the U =
//                                   k_s(k_s+1)/2 unordered local spline pairs
//   Pw, Qw, PP, QS [n_pairs][k_s][K]  per-pair window moments P^k, Q^k over the
//                                   pair's support window, their prefix (P) and
//                                   suffix (Q) sums over the window
//   Ptot, Qtot [n_pairs][K]         window totals
//   D  [n_iv][U][U][K]              diagonal-cell integrals (outer local pair,
//                                   inner local pair, multipole); a tensor
//                                   build keeps D + D^T (symmetric in the pairs)
//   R  [n_pairs][n_pairs][K]        the R^k tensor; key [a,b,c,d] lives at
//                                   (pair(a,c)*n_pairs + pair(b,d))*K + k
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace bs2e {

constexpr int kMaxOrder = 20;   // BVALUE_D work arrays are dimension(20) (bspline_tools.f90:163)
constexpr int kMaxGL = 64;
constexpr int kMaxK = 64;       // multipole orders 0..63
// LDS one workgroup may allocate on gfx950 (the whole 160 KiB of a CU).
constexpr size_t kLdsPerWorkgroup = 160 * 1024;
// Lets kernel fn take `bytes` of dynamic LDS (needed beyond 64 KiB).
hipError_t allow_lds(const void* fn, size_t bytes);

struct PairInfo {               // one banded pair (a, c), 1-based basis indices
    int32_t a, c;
    int32_t v0;                 // first interval of the common support window (1-based)
    int32_t m;                  // window length (number of intervals)
};

// Entry of the reference's sparse_6d order: pair J (inner) outer loop, pair I
// (outer) inner loop, intervals v0 .. v0+len-1 (mat_els.f90:251-289).
struct DiagEntry {
    int32_t J, I;
    int32_t v0, len;
    int64_t ptr;                // 0-based first position in sparse_6d order
};

// Host-side description of one basis (knots + quadrature + multipoles).
struct Plan {
    int64_t k_s = 0, n_knots = 0, n = 0, n_b = 0, n_iv = 0;
    int64_t k_gl = 0, max_k = 0, K = 0;
    int64_t n_pairs = 0, nnz4 = 0, nnz6 = 0, U = 0;
    std::vector<double> t, bp, x, w;
    std::vector<PairInfo> pairs;
    std::vector<int64_t> off4;          // per pair: first sparse_4d position
    std::vector<int32_t> c_first;       // first pair index with column c (1-based c, size n_b+2)
    std::vector<DiagEntry> diag;        // sparse_6d order (built on demand)
    int far_variant = -1;               // rank-one fill shape (index into the far variants; -1: default)
    int inner_kc = 0;                   // multipoles per diagonal-cell workgroup (0: the largest that fits)
    std::vector<int32_t> rows;          // near-kernel rows: pairs (a, c) with a <= c (built with the tensor)
    int64_t near_elems = -1;            // tensor elements whose pairs share a support interval (-1: not counted)
};

// Builds the plan; returns an error message or empty string.
std::string make_plan(int64_t k_s, int64_t n_knots, const double* knots, int64_t k_gl, int64_t max_k, Plan& p);
void build_diag_entries(Plan& p);
void build_near_rows(Plan& p);
int64_t count_near_elements(const Plan& p);

// Host numerics shared by the C-ABI (setup, not the hot path).
void gauss_legendre(int64_t N, double* x, double* w);
int64_t generate_grid(int64_t k, int64_t m, int64_t Z, double h_max, double r_max, std::vector<double>& out);
double wigner_3j(int64_t ja, int64_t jb, int64_t jc, int64_t ma, int64_t mb, int64_t mc);
double wigner_6j(int64_t j1, int64_t j2, int64_t j3, int64_t j4, int64_t j5, int64_t j6);
double c_red(int64_t k, int64_t a, int64_t b);   // wigner_tools.f90:107-112 C_red_mat
double ang_k_LS(int64_t k, int64_t l1a, int64_t l1b, int64_t l2a, int64_t l2b, int64_t L);

// Device buffers of one tensor build.
struct DevTables {
    double *t = nullptr, *bp = nullptr, *x = nullptr, *w = nullptr;
    PairInfo* pairs = nullptr;
    double *Pl = nullptr, *Ql = nullptr;
    double *Pw = nullptr, *Qw = nullptr, *PP = nullptr, *QS = nullptr, *Ptot = nullptr, *Qtot = nullptr;
    double* D = nullptr;
    int32_t* rows = nullptr;
    int32_t* c_first = nullptr;
};

// ---- symmetry blocks (angular-channel summation) --------------------------
struct BlockChan {            // one channel (l1, l2) of a block's configurations
    int32_t l1, l2;
    int32_t start, n;         // first configuration (0-based) and count
};

struct BlockSeg {             // configurations of one channel with a given n1
    int32_t off;              // offset of the first one within the channel
    int32_t lo, hi;           // n2 range (hi < lo: none)
    int32_t pad;
};

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2.  xcd_block maps the launch's workgroup b to the logical
// workgroup index that gives XCD (b mod 8) one contiguous range of the n
// logical workgroups, so neighbours that gather the same tables share an L2.
constexpr int kXcds = 8;

// Loads and stores through the global address space: the tables a kernel
// reaches through a descriptor are global memory, and saying so lets the
// compiler use global instead of flat instructions (which also wait on LDS).
template <class T>
__device__ __forceinline__ T gload(const T* p)
{
    return *(const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ void gstore(T* p, T v)
{
    *(__attribute__((address_space(1))) T*)p = v;
}
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t n)
{
    const int64_t x = b % kXcds, j = b / kXcds;
    const int64_t per = n / kXcds, extra = n % kXcds;
    return x * per + (x < extra ? x : extra) + j;
}
// The same in chunks: the j-th workgroup an XCD receives takes logical block
// (j / C) 8C + x C + j % C, so the XCDs work through adjacent bands of C
// logical blocks each (a window of 8C blocks moving through the grid); the
// last, partial window is split as xcd_block splits a grid.  A bijection of
// [0, n).
__device__ __forceinline__ int64_t xcd_chunk_block(int64_t b, int64_t n, int64_t C)
{
    const int64_t R = n / (kXcds * C), x = b % kXcds, j = b / kXcds;
    if (j < R * C) return (j / C) * kXcds * C + x * C + (j % C);
    return R * kXcds * C + xcd_block(b - R * kXcds * C, n - R * kXcds * C);
}

struct BlockArgs {
    int n_config, n_chan, full, L, ks, n_b, n_pairs, K, n_iv;
    const int32_t* cfg;          // 2 per config: n1 | n2 << 16, l1 | l2 << 16
    const int32_t* cfg_chan;     // channel of each config
    const BlockChan* chan;       // per channel
    const BlockSeg* seg;         // [n_chan][n_b + 1]: per n1
    const double* coef;          // [n_chan][n_chan][2][K]: direct, exchange (with its sign)
    const int32_t* flags;        // [n_chan][n_chan]: bit0 direct allowed, bit1 exchange allowed
    const int32_t* c_first;      // pair index of the first pair with column c
    const double* R;             // [n_pairs][n_pairs][K]
    const double *PtotT, *QtotT; // [K][n_pairs] window totals (multipole-major): the rank-one factors of R
    const double* Hl;            // complex, per l an n_b x n_b row-major matrix (the transpose of the caller's)
    const double* S;             // complex, n_b x n_b row-major
    int64_t* cntH;               // per row (count pass)
    int64_t* cntS;
    int64_t* refH;               // per row: the reference's count_nnz predicate
    int64_t* refS;
    const int64_t* ptrH;         // 0-based row starts in the whole set's entries (fill pass)
    const int64_t* ptrS;
    int64_t winH, winS;          // the set's first entry held by the outputs: row i starts at ptr[i] - win
    int64_t limH, limS;          // entries the output arrays hold (fill pass)
    int64_t* ixH;
    double* dH;                  // complex
    int64_t* ixS;
    double* dS;                  // complex
};

// One-electron matrices: spline value / second-derivative tables at the GL
// points (B, B2: [n_iv][k_gl][k]; rw: [n_iv][k_gl][2] r and w) and the
// assembly of S (optional) and H for l = l_min .. l_min+n_l-1.
hipError_t launch_one_body(const Plan& p, const DevTables& d, double* B, double* B2, double* rw, int l_min, int n_l,
                           int Z, int cap_order, double cap_r0, double eta_re, double eta_im, double* S, double* H,
                           hipStream_t s);
// RKH one-electron multipoles r_k_int(0:max_k, n_b, n_b), real, k fastest
hipError_t launch_rk_int(const Plan& p, const DevTables& d, double* B, double* B2, double* rw, int max_k, double r0,
                         double* out, hipStream_t s);

// Dipole blocks: per (row channel, column channel) the terms of dip_mat_neq
// (0: direct electron 1, 1: direct electron 2, 2 and 3: exchange).
struct DipCoef {
    int32_t nz;                 // ang_dip_red > 5e-16: the channel pair carries entries
    int32_t cond[4];            // the term's l-equality holds
    int32_t vkind[4];           // velocity gauge: 0 |l - l'| != 1, 1 l > l', 2 l < l'
    int32_t pad;
    double pf[4];               // (-1)^(..) sqrt((2L1+1)(2L2+1)) 6j(..)
    double cred[4];             // length gauge: C_red(1, l, l')
    double vsq[4];              // velocity gauge: sqrt(l) (l > l') or sqrt(l') (l < l')
    double vm[4];               // velocity gauge: l'+1 (l > l') or l' (l < l')
    double sgn_ex;              // (-1)^(L2 + l_c + l_d)
};

struct DipArgs {
    int n1, n_chan2, ks, n_b, gauge;
    double ang;                 // (-1)^(L1-M1) 3j(L1 1 L2; -M1 q M2)
    const int64_t* cfg1;        // rows: 5 per config
    const int32_t* cfg_chan1;   // channel of each row
    const BlockChan* chan2;     // column channels
    const BlockSeg* seg2;       // [n_chan2][n_b + 1]
    const DipCoef* coef;        // [n_chan1][n_chan2]
    const double *S, *rad1, *rad2;   // complex n_b x n_b row-major (rad: r_mat, or dr_mat and r_inv_mat)
    int64_t* cnt;               // count pass
    const int64_t* ptr;         // fill pass: 0-based row starts
    int64_t lim;                // fill pass: entries the outputs hold
    int64_t* ix;
    double* data;               // complex
};

hipError_t launch_dip_count(const DipArgs& a, hipStream_t s);
hipError_t launch_dip_fill(const DipArgs& a, hipStream_t s);
// the fill's layouts: the caller's column-major S, rad1 and (R2 != NULL) rad2
// transposed to row-major
hipError_t launch_dip_prepare(const double* S, const double* R1, const double* R2, int n_b, double* ST, double* R1T,
                              double* R2T, hipStream_t s);
// radial dipole matrices: gauge 0 -> M1 = r_mat; 1 -> M1 = dr_mat, M2 = r_inv_mat
hipError_t launch_radial_dipole(const Plan& p, const DevTables& d, double* B, double* B1, double* rw, int gauge,
                                double* M1, double* M2, hipStream_t s);

// Count / fill launches over the rows of n_blocks blocks (device table tab,
// device row offsets row0[n_blocks + 1], n_rows = row0[n_blocks]).
hipError_t launch_block_count(const BlockArgs* tab, const int64_t* row0, int n_blocks, int64_t n_rows, hipStream_t s);
// The fill takes the n_gen rows with |n1 - n2| < k listed in rows_gen, then
// the staged rows (on the matrix cores): the n_far with |n1 - n2| >= 2k - 1
// listed in rows_far and the n_sep with k <= |n1 - n2| < 2k - 1 listed in
// rows_sep (global row numbers).
hipError_t launch_block_fill(const BlockArgs* tab, const int64_t* row0, int n_blocks, const int32_t* rows_far,
                             int64_t n_far, const int32_t* rows_sep, int64_t n_sep, const int32_t* rows_gen,
                             int64_t n_gen, int ks, int K, hipStream_t s);
// Exclusive prefix sums of the count pass's per-row counts (cnt: H counts then
// S counts, n_rows each) into ptr (same layout), on the device; tmp holds
// row_scan_tmp_elems(n_rows) int64.
int64_t row_scan_tmp_elems(int64_t n_rows);
hipError_t launch_row_scan(const int64_t* cnt, int64_t* ptr, int64_t n_rows, int64_t* tmp, hipStream_t s);
// The fill's operand layouts: the caller's column-major one-electron matrices
// (n_l of them, then S) to row-major, and the window totals [n_pairs][K] to
// [K][n_pairs].
hipError_t launch_block_prepare(const double* Hl, const double* S, int n_l, int n_b, const double* Ptot,
                                const double* Qtot, int n_pairs, int K, double* HlT, double* ST, double* PtotT,
                                double* QtotT, hipStream_t s);

// orbital_tools.f90:119-216 count_configs: 5 int64 per config (n1, n2, l1, l2,
// eqv); returns the count (writes at most cap configs when out != nullptr).
int64_t count_configs(int64_t L, int64_t term_pi, int64_t max_l_1p, int64_t n_b, int64_t k_spline, int64_t max_n_b,
                      int64_t n_all_l, int64_t l_2_max, int64_t* out, int64_t cap);

// Kernel launchers (each .hip file launches its own kernels).
hipError_t launch_outer_moments(const Plan& p, const DevTables& d, hipStream_t s);           // Pl, Ql
hipError_t launch_diag_cells(const Plan& p, const DevTables& d, hipStream_t s, bool symmetrised);  // D / Dsym
hipError_t launch_symmetrize(const Plan& p, const DevTables& d, hipStream_t s);                    // D -> Dsym
size_t near_lds_bytes(int64_t ks, int64_t K);
constexpr int kNearMaxSeg = 64;   // column segments of one near row (<= 2k - 1 <= 39 for k <= 20)
// mixed precision: Pl, Ql on the FP32 matrix cores from exact integer slices
hipError_t launch_outer_moments_mixed(const Plan& p, const DevTables& d, hipStream_t s);
bool mixed_supported(int64_t ks, int64_t kgl, int64_t K);   // order <= 16, k_GL <= 16
size_t mixed_lds_bytes(int64_t ks, int64_t K);
hipError_t launch_pair_windows(const Plan& p, const DevTables& d, hipStream_t s);            // Pw/Qw + sums
hipError_t launch_tensor_far(const Plan& p, const DevTables& d, double* R, hipStream_t s);   // rank-one part
int far_variant_count();
// Zeros over n16 16-byte elements at dst, one store per thread (the store
// ceiling the rank-one fill is measured against).
hipError_t launch_store_fill(void* dst, int64_t n16, bool nt, hipStream_t s);
bool far_variant_valid(const Plan& p, int v);      // 16-byte variants need rows of even length
int far_variant_default(const Plan& p);
hipError_t launch_tensor_near(const Plan& p, const DevTables& d, double* R, hipStream_t s);  // overlapping pairs
size_t outer_lds_bytes(int64_t ks, int64_t kgl, int64_t K);
size_t inner_lds_bytes(int64_t ks, int64_t kgl, int64_t kc);
int inner_kc(int64_t ks, int64_t kgl, int64_t K);   // multipoles per diagonal-cell workgroup (0: unsupported)
hipError_t launch_scatter_4d(const Plan& p, const DevTables& d, const int64_t* d_off4, double* rk, double* rmk,
                             int64_t* iv, int64_t* i, int64_t* j, hipStream_t s);
hipError_t launch_scatter_6d(const Plan& p, const DevTables& d, const DiagEntry* d_diag, int64_t n_diag, double* rdk,
                             int64_t* iv, int64_t* i, int64_t* ip, int64_t* j, int64_t* jp, hipStream_t s);

hipError_t launch_window_sums(const Plan& p, const DevTables& d, hipStream_t s);             // sums of Pw/Qw
hipError_t launch_fill(double* ptr, int64_t n, double v, hipStream_t s);
hipError_t launch_gather_4d(const Plan& p, const DevTables& d, const int32_t* d_c_first, const double* rk,
                            const double* rmk, const int64_t* iv, const int64_t* i, const int64_t* j, int* bad,
                            hipStream_t s);
hipError_t launch_gather_6d(const Plan& p, const DevTables& d, const double* rdk, const int64_t* iv, const int64_t* i,
                            const int64_t* ip, const int64_t* j, const int64_t* jp, int* bad, hipStream_t s);

// Symmetric local pair index of local splines (ta, tc) on one interval.
__host__ __device__ inline int sym_local(int ta, int tc, int ks)
{
    int lo = ta < tc ? ta : tc, hi = ta < tc ? tc : ta;
    return lo * ks - (lo * (lo - 1)) / 2 + (hi - lo);
}

}  // namespace bs2e
