/*
 * bs2e.h -- C ABI of the MI355X (gfx950) two-electron R^k builder.
 *
 * Drop-in boundary for the two-electron radial matrix-element path of
 * edvinolo/b-spline-two-e (Fortran 2008, default integer kind 8).  Every
 * integer crossing this ABI is int64_t (the reference builds with
 * -fdefault-integer-8, CMakeLists.txt:44-57); Fortran logicals are passed as
 * int64_t 0/1.  Index conventions are the reference's: spline/basis/interval
 * indices are 1-based, containers are column-major where the reference's are.
 *
 * Errors: every entry point returns BS2E_OK (0) or a negative status; the
 * message is available from bs2e_last_error(ctx) (or bs2e_last_error(NULL)
 * for calls without a context).  Nothing aborts inside the library: the
 * Fortran shim turns a negative status into `error stop`, as the reference
 * stops on its own errors.
 *
 * Threading: any host thread may call.  A context serialises its device work
 * on one HIP stream.  bs2e_rk_get_val is reentrant and lock-free once the
 * host mirror exists (the reference wraps its hash-map lookup in
 * `!$omp critical`, sparse_array_tools.f90:546-548).
 */
#ifndef BS2E_H
#define BS2E_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BS2E_OK 0
#define BS2E_EINVAL (-1)   /* invalid argument (sizes, orders, null pointers) */
#define BS2E_EDEVICE (-2)  /* HIP runtime error */
#define BS2E_ENOMEM (-3)   /* device or host allocation failed */
#define BS2E_ERANGE (-4)   /* key outside the banded key set (reference: error stop) */
#define BS2E_ESTATE (-5)   /* object not in a state that allows the call */

#define BS2E_PREC_FP64 0   /* plain FP64 path */
/* Mixed precision: the off-diagonal moments (pair densities x radial moments,
 * mat_els.f90:413-428) on the FP32 matrix cores from exact 10-bit integer
 * slices of both operands with power-of-two scales, recombined in FP64 (about
 * 2^-51 of the largest term of a moment's row and column); de Boor values,
 * diagonal cells, window sums and the tensor's products stay FP64.
 * Orders <= 16 with k_GL <= 16. */
#define BS2E_PREC_MIXED 1

typedef struct bs2e_ctx bs2e_ctx; /* one device + one stream */
typedef struct bs2e_rk bs2e_rk;   /* one basis and its device-resident R^k tensor */

/* ---- lifecycle --------------------------------------------------------- */
/* Creates a context on HIP device `device` with its own default-priority
 * stream.  (No reference counterpart: the reference runs on the host.) */
int64_t bs2e_create(bs2e_ctx **ctx, int64_t device);
int64_t bs2e_destroy(bs2e_ctx *ctx);
const char *bs2e_last_error(const bs2e_ctx *ctx);
/* Runs the context's device work on `hip_stream` (a hipStream_t) instead of
 * its own stream; NULL restores the context's stream. */
int64_t bs2e_set_stream(bs2e_ctx *ctx, void *hip_stream);

/* ---- basis set-up ------------------------------------------------------ */
/* grid_tools.f90:6-55 generate_grid(k,m,Z,h_max,r_max,grid).  Returns the
 * number of knots (writes them when knots != NULL and cap suffices), or a
 * negative status. */
int64_t bs2e_generate_grid(int64_t k, int64_t m, int64_t Z, double h_max, double r_max, double *knots, int64_t cap);

/* Container sizes: sparse_array_tools.f90:276-296 count_nnz_4d,
 * :298-324 count_nnz_6d, :347-366 count_nnz_R_k (= n_pairs^2). */
int64_t bs2e_count_nnz(int64_t k, int64_t n_knots, const double *knots, int64_t *nnz4, int64_t *nnz6,
                       int64_t *n_pairs);

/* ---- primitives -------------------------------------------------------- */
/* mat_els.f90:172-182 setup_Slater_integrals(b_splines, max_k, k_GL, r_k,
 * r_m_k, r_d_k): fills the caller's containers (sparse_4d r_k / r_m_k and
 * sparse_6d r_d_k, sparse_array_tools.f90:9-31) in the reference's order.
 * rk, rmk: nnz4 x (max_k+1) column-major; rdk: nnz6 x (max_k+1).  The
 * reference accumulates into containers that init_4d/init_6d zeroed
 * (sparse_array_tools.f90:159-178); this call overwrites them. */
int64_t bs2e_slater_primitives(bs2e_ctx *ctx, int64_t k, int64_t n_knots, const double *knots, int64_t k_gl,
                               int64_t max_k, int64_t nnz4, double *rk, double *rmk, int64_t *iv4, int64_t *i4,
                               int64_t *j4, int64_t nnz6, double *rdk, int64_t *iv6, int64_t *i6, int64_t *ip6,
                               int64_t *j6, int64_t *jp6);

/* ---- one-electron matrices -------------------------------------------- */
/* mat_els.f90:85-118 setup_S(b_splines, k_GL, S): S(i,j) = sum over the
 * intervals holding both splines and their GL points of w B_i B_j; S is
 * n_b x n_b double complex, column-major (interleaved re/im), zero outside
 * the band. */
int64_t bs2e_setup_S(bs2e_ctx *ctx, int64_t k, int64_t n_knots, const double *knots, int64_t k_gl, double *S);
/* mat_els.f90:47-83 setup_H_one_particle(potential, CAP_c, l, b_splines,
 * k_GL, H) for the hydrogenic potential (potentials.f90 V_hydrogenic(Z):
 * 0.5 l(l+1)/r^2 - Z/r) and the absorbing potential of CAP_tools.f90:24-34
 * (-i eta (r - r_0)^order for r >= r_0), for every l in [l_min, l_max]:
 * H holds l_max-l_min+1 matrices of n_b x n_b double complex, column-major. */
int64_t bs2e_setup_H_one_particle(bs2e_ctx *ctx, int64_t k, int64_t n_knots, const double *knots, int64_t k_gl,
                                  int64_t Z, int64_t l_min, int64_t l_max, int64_t cap_order, double cap_r0,
                                  double cap_eta_re, double cap_eta_im, double *H);

/* ---- the R^k tensor ---------------------------------------------------- */
/* sparse_array_tools.f90:452-493 compute_R_K_map(r_d_k, r_k, r_m_k,
 * b_splines, max_k, R) and :495-507 Nd_DOK%init.  Plans the basis and
 * allocates the device tensor R[pair(a,c)][pair(b,d)][k] (FP64) and its
 * work tables; computes nothing yet.  precision: BS2E_PREC_FP64 or
 * BS2E_PREC_MIXED. */
int64_t bs2e_rk_create(bs2e_ctx *ctx, int64_t k, int64_t n_knots, const double *knots, int64_t k_gl, int64_t max_k,
                       int64_t precision, bs2e_rk **out);
/* sparse_array_tools.f90:452-493 compute_R_K_map(r_d_k, r_k, r_m_k,
 * b_splines, max_k, R) from the caller's own containers (as filled by
 * setup_Slater_integrals / bs2e_slater_primitives): uploads them, places each
 * entry by its index arrays, and plans the tensor.  bs2e_rk_compute then runs
 * only the tensor kernel.  BS2E_ERANGE if an entry's indices fall outside the
 * banded, supported set of the basis. */
int64_t bs2e_rk_from_primitives(bs2e_ctx *ctx, int64_t k, int64_t n_knots, const double *knots, int64_t max_k,
                                int64_t nnz4, const double *rk, const double *rmk, const int64_t *iv4,
                                const int64_t *i4, const int64_t *j4, int64_t nnz6, const double *rdk,
                                const int64_t *iv6, const int64_t *i6, const int64_t *ip6, const int64_t *j6,
                                const int64_t *jp6, bs2e_rk **out);
/* Enqueues the whole build (de Boor tables, pair moments, diagonal cells,
 * tensor) on the context's stream.  Allocates nothing. */
int64_t bs2e_rk_compute(bs2e_rk *rk);
int64_t bs2e_rk_synchronize(bs2e_rk *rk);
/* info[0..10] = n_b, n_pairs, K, nnz4, nnz6, n_iv, k, k_GL, U, n_knots,
 * near elements (tensor elements whose pairs share a support interval) */
int64_t bs2e_rk_info(const bs2e_rk *rk, int64_t *info, int64_t n);
/* Device address and element count of the tensor (n_pairs^2 * K doubles). */
int64_t bs2e_rk_device_ptr(const bs2e_rk *rk, void **ptr, int64_t *n_elems);
/* 0-based position of the banded pair (a, c) in the tensor's row/column
 * order (count_nnz_R_k loop order: c outer, a inner), or BS2E_ERANGE. */
int64_t bs2e_rk_pair_index(const bs2e_rk *rk, int64_t a, int64_t c);
/* Copies the tensor to the host: into dst (n_elems doubles) when dst != NULL,
 * and into the handle's host mirror used by bs2e_rk_get_val. */
int64_t bs2e_rk_copy_to_host(bs2e_rk *rk, double *dst, int64_t n_elems);
/* Copies tensor rows row_lo .. row_lo+n_rows-1 (0-based pair indices; each
 * row n_pairs * K doubles) to dst.  (No reference counterpart: a bulk
 * Nd_DOK%get_val over every key of some electron-1 pairs.) */
int64_t bs2e_rk_copy_rows(const bs2e_rk *rk, int64_t row_lo, int64_t n_rows, double *dst);
/* sparse_array_tools.f90:536-555 Nd_DOK%get_val(indices(4), values(0:max_k)):
 * values of key [a,b,c,d].  Reads the host mirror (lock-free); without a
 * mirror it copies the K values from the device.  BS2E_ERANGE for a key
 * outside |a-c| < k, |b-d| < k (the reference error-stops there). */
int64_t bs2e_rk_get_val(const bs2e_rk *rk, const int64_t *indices, double *values, int64_t n_values);
/* Device timeline of the builds computed while profiling is on (HIP events
 * on the build's two streams), in ms since each build's start, averaged over
 * the builds since profiling was switched on or since the previous
 * bs2e_rk_stage_times call: [0] local moments + pair windows done,
 * [1] rank-one fill done, [2] diagonal cells done (second stream), [3] near
 * elements done (second stream), [4] end of the build. */
int64_t bs2e_rk_set_profiling(bs2e_rk *rk, int64_t on);
int64_t bs2e_rk_stage_times(bs2e_rk *rk, double *ms, int64_t n);
/* on != 0: the next bs2e_rk_compute captures the build (both streams) into a
 * hipGraph and every compute then launches that graph, with one launch call
 * instead of five kernels and the stream events; 0 drops the graph.  Not
 * combinable with profiling.  (No reference counterpart.) */
int64_t bs2e_rk_set_graph(bs2e_rk *rk, int64_t on);
/* Autotuner: times, on this handle's own tables, each shape of the rank-one
 * fill (elements per store, rows and elements per workgroup) and each
 * multipole chunk of the diagonal-cell kernel, and keeps the fastest for
 * this handle and, for the process, for every basis of the same order, GL
 * order, multipole count, pair and interval counts (force != 0 re-times a
 * known one).  All choices compute the same values.  (No reference
 * counterpart.) */
int64_t bs2e_rk_autotune(bs2e_rk *rk, int64_t force);
/* out[0] = rank-one fill shape in use, out[1] = multipoles per
 * diagonal-cell workgroup. */
int64_t bs2e_rk_tuning(const bs2e_rk *rk, int64_t *out, int64_t n);
/* Measurement: writes zeros over n_bytes of device memory at dst (16-byte
 * aligned, a positive multiple of 16 bytes) on the context's stream with the
 * library's plain store kernel (one 16-byte store per thread), without and
 * then with the non-temporal hint, each reps times after a warm-up pass, and
 * reports the faster rate in GB/s (HIP events): the store ceiling the tensor's
 * rank-one fill is measured against.  (No reference counterpart.) */
int64_t bs2e_store_ceiling(bs2e_ctx *ctx, void *dst, int64_t n_bytes, int64_t reps, double *gbs);
/* Pins the rank-one fill shape (0 .. the shapes the autotuner times; those
 * with two elements per store need an even n_pairs * K) and the multipoles
 * per diagonal-cell workgroup (1 .. the largest that fits) for this handle's
 * next builds; BS2E_EINVAL outside those.  All choices compute the same
 * values.  (No reference counterpart.) */
int64_t bs2e_rk_set_tuning(bs2e_rk *rk, int64_t shape, int64_t kc);
int64_t bs2e_rk_release(bs2e_rk *rk);
/* No reference counterpart: for callers that cannot allocate device memory
 * (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The whole build of bs2e_rk_create + bs2e_rk_compute (same
 * kernels, same values, bit for bit) into the caller's device buffer R of
 * n_elems = n_pairs^2 * (max_k+1) doubles on the calling thread's current
 * device, enqueued on hip_stream (NULL: the null stream), with the work tables in one static device array of the
 * library (128 MiB per GPU, held until the process ends; no hipMalloc,
 * hipFree or allocating library call).  A basis whose tables exceed it (for
 * instance order 9 with 600 B-splines and k <= 24) is refused with
 * BS2E_EINVAL.  Every call on one GPU shares that array: calls that overlap
 * on one GPU (other streams, other threads) race; calls on one stream are
 * ordered.  The knot-derived tables are uploaded with a wait on hip_stream
 * before the kernels are enqueued; the call returns without waiting for
 * them. */
int64_t bs2e_rk_build_static(int64_t k, int64_t n_knots, const double *knots, int64_t k_gl, int64_t max_k,
                             int64_t precision, double *R, int64_t n_elems, void *hip_stream);

/* ---- angular algebra --------------------------------------------------- */
/* wigner_tools.f90:126-138 ang_k_LS(k, confs, L) with confs(1)%l = (l1a,l1b),
 * confs(2)%l = (l2a,l2b); wigner_tools.f90:30-60 three_j / six_j. */
double bs2e_ang_k_LS(int64_t k, int64_t l1a, int64_t l1b, int64_t l2a, int64_t l2b, int64_t L);
double bs2e_three_j(int64_t ja, int64_t jb, int64_t jc, int64_t ma, int64_t mb, int64_t mc);
double bs2e_six_j(int64_t j1, int64_t j2, int64_t j3, int64_t j4, int64_t j5, int64_t j6);

/* ---- symmetry blocks --------------------------------------------------- */
typedef struct bs2e_block bs2e_block; /* the configurations of one symmetry and its CSR pattern */

/* orbital_tools.f90:119-216 count_configs(term, max_l_1p, n_b, k_spline,
 * max_n_b, n_all_l, l_2_max) with :46-72 consistent: enumerates the
 * two-electron configurations of symmetry (L, term_pi) (term_pi 1 = odd) in
 * the reference's order (l1, l2 <= l1, n1, n2).  configs receives 5 int64 per
 * configuration: n1, n2, l1, l2, eqv (0/1), when non-NULL.  Returns the
 * count, or BS2E_EINVAL when cap is smaller. */
int64_t bs2e_count_configs(int64_t L, int64_t term_pi, int64_t max_l_1p, int64_t n_b, int64_t k_spline,
                           int64_t max_n_b, int64_t n_all_l, int64_t l_2_max, int64_t *configs, int64_t cap);

/* hamiltonian.f90:106-283 construct_block_tensor(H, S, b_splines, term,
 * max_k, R_k, H_sp, S_sp, full), first half: takes the block's
 * configurations (term%configs, as bs2e_count_configs writes them: each
 * (l1,l2) channel contiguous, n1 ascending, n2 consecutive) and counts, on
 * the device, the CSR entries of the Hamiltonian and overlap blocks against
 * the computed tensor of `rk` (full = 0: upper triangle j >= i only).  The
 * tensor must stay alive while the block is used. */
int64_t bs2e_block_create(bs2e_rk *rk, int64_t L, int64_t n_config, const int64_t *configs, int64_t full,
                          bs2e_block **out);
/* sizes[0..4] = nnz of H, nnz of S (the entries the fill writes), the
 * reference's count_nnz(b_splines, term, max_k, full) for H and S
 * (hamiltonian.f90:348-416; it can be smaller than the fill, in which case
 * the reference writes past its allocation), n_config. */
int64_t bs2e_block_sizes(const bs2e_block *b, int64_t *sizes, int64_t n);
/* hamiltonian.f90:106-283 second half: fills the 1-based CSR arrays of H_sp
 * and S_sp (index_ptr: n_config+1, indices/data: nnz; data double complex as
 * interleaved re/im) with mat_els.f90:608-633 c_mat_neq_tens (angular sums
 * of R^k, :552-571 r_12_tens), :697-715 H_1p_neq and :664-678 S_mat_neq.
 * Hl: max_l_1p+1 one-electron blocks H(l)%data, each n_b x n_b complex
 * column-major (l = 0 first); S: n_b x n_b complex column-major. */
int64_t bs2e_block_fill(bs2e_block *b, int64_t max_l_1p, const double *Hl, const double *S, int64_t *ip_H,
                        int64_t *ix_H, double *d_H, int64_t *ip_S, int64_t *ix_S, double *d_S);
int64_t bs2e_block_release(bs2e_block *b);
/* Device time (ms, HIP events on the context's stream) of the block's count
 * kernel (ms[0], at create) and fill kernel (ms[1], at the last fill). */
int64_t bs2e_block_times(const bs2e_block *b, double *ms, int64_t n);

/* RKH_hamiltonian.f90:91-159 setup_r_k_int(b_splines) (with init_g_l,
 * :23-28, for k_GL, and adiabatic_potentials.f90:24-34 r_k): the one-electron
 * multipoles of the Kramers-Henneberger frame, r_k_int(k, i, j) = sum over
 * the intervals holding both splines and their GL points of
 * w r_k(r; r_0) B_i B_j with r_k = r_0^k / r^(k+1) for r > r_0 and
 * r^k / r_0^(k+1) otherwise (the reference's module variables r_0 =
 * E_0 / omega^2 and max_k = 2 max_l_1p become arguments).  r_k_int is real,
 * (max_k+1) x n_b x n_b with k fastest, as r_k_int(0:max_k, n_b, n_b). */
int64_t bs2e_setup_r_k_int(bs2e_ctx *ctx, int64_t k, int64_t n_knots, const double *knots, int64_t k_gl,
                           int64_t max_k, double r_0, double *r_k_int);

/* ---- several symmetry blocks in one launch ------------------------------ */
typedef struct bs2e_blockset bs2e_blockset; /* the blocks of several symmetries against one tensor */

/* hamiltonian.f90:106-283 construct_block_tensor for n_sym symmetries at
 * once, as main_basis_setup.f90:104-115 loops over bas%syms: symmetry s
 * (1-based) has total L[s-1] and n_config[s-1] configurations, stored one
 * symmetry after the other in configs (5 int64 each, count_configs order).
 * One count launch over the configuration rows of every symmetry (uneven
 * channel counts and sizes share the chip through one wavefront per row). */
int64_t bs2e_blockset_create(bs2e_rk *rk, int64_t n_sym, const int64_t *L, const int64_t *n_config,
                             const int64_t *configs, int64_t full, bs2e_blockset **out);
/* sizes[5(s-1) .. 5(s-1)+4] = the bs2e_block_sizes values of symmetry s. */
int64_t bs2e_blockset_sizes(const bs2e_blockset *set, int64_t *sizes, int64_t n);
/* One fill launch over every row of every symmetry (inputs as
 * bs2e_block_fill); the CSR arrays stay on the device until the next fill or
 * the release. */
int64_t bs2e_blockset_fill(bs2e_blockset *set, int64_t max_l_1p, const double *Hl, const double *S);
/* Copies symmetry sym's (1-based) H_sp and S_sp to the host, as
 * bs2e_block_fill writes them (sizes from bs2e_blockset_sizes). */
int64_t bs2e_blockset_get(const bs2e_blockset *set, int64_t sym, int64_t *ip_H, int64_t *ix_H, double *d_H,
                          int64_t *ip_S, int64_t *ix_S, double *d_S);
/* Rows row_lo .. row_lo+n_rows-1 (0-based, within symmetry sym) of H_sp and
 * S_sp: ip_H/ip_S (n_rows+1 entries) are 1-based offsets relative to the
 * window (ip[0] = 1), ix/data hold ip[n_rows]-1 entries each.  With all four
 * of ix_H, d_H, ix_S, d_S NULL only the offsets are written (to size the
 * buffers).  (No reference counterpart: lets a caller read a block larger
 * than its host memory window by window.) */
int64_t bs2e_blockset_get_rows(const bs2e_blockset *set, int64_t sym, int64_t row_lo, int64_t n_rows,
                               int64_t *ip_H, int64_t *ix_H, double *d_H, int64_t *ip_S, int64_t *ix_S,
                               double *d_S);
/* hamiltonian.f90:106-283 construct_block_tensor over a window of rows:
 * global rows row_lo .. row_lo+n_rows-1 (0-based over the set's rows, the
 * symmetries laid end to end in bs2e_blockset_create's order) filled on the
 * device into outputs that hold that window only (grown when a window needs
 * more room, kept otherwise).  Hl and S as bs2e_blockset_fill, or both NULL
 * to reuse the one-electron matrices of the set's last fill (kept on the
 * device).  The window replaces the earlier fill: bs2e_blockset_get_rows
 * then reads rows inside it, bs2e_blockset_get a symmetry it holds whole,
 * bs2e_blockset_refill fills it again.  The entries are those of the whole
 * set's fill, bit for bit.  (The reference holds each symmetry's whole block,
 * main_basis_setup.f90:104-115; a window lets a set larger than the GPU's
 * memory stream through it.) */
int64_t bs2e_blockset_fill_rows(bs2e_blockset *set, int64_t max_l_1p, const double *Hl, const double *S,
                                int64_t row_lo, int64_t n_rows);
/* Output buffers of the set with room for at least nnz_H H and nnz_S S
 * entries (reallocated, dropping what they held, only when smaller), so that
 * a stream of windows allocates once, for its largest window: a later
 * reallocation costs about as much as the driver's clearing of the freed
 * memory.  (No reference counterpart.) */
int64_t bs2e_blockset_reserve(bs2e_blockset *set, int64_t nnz_H, int64_t nnz_S);
/* The count pass's entries per row for global rows row_lo ..
 * row_lo+n_rows-1 (as bs2e_blockset_fill_rows numbers them): cnt_H[r],
 * cnt_S[r], what a window of those rows holds.  (No reference counterpart.) */
int64_t bs2e_blockset_row_counts(const bs2e_blockset *set, int64_t row_lo, int64_t n_rows, int64_t *cnt_H,
                                 int64_t *cnt_S);
/* The fill again, from the one-electron matrices of the last
 * bs2e_blockset_fill (kept on the device): the layouts and the fill launch
 * over the rows of the last fill (all of them, or the last window), waited
 * for.  (No reference counterpart.) */
int64_t bs2e_blockset_refill(bs2e_blockset *set);
/* on != 0: captures the whole branch of the set into one hipGraph: its
 * tensor's two-stream build, the count pass, the device prefix sums of the
 * row counts, the fill's layouts and the fill over every symmetry and channel
 * (one-electron matrices of the last bs2e_blockset_fill, which must have
 * filled the whole set, not a window: BS2E_ESTATE otherwise); 0 drops it.  A
 * new fill or window drops it too.  bs2e_blockset_replay then recomputes tensor and blocks
 * with one launch on the context's stream (asynchronous:
 * bs2e_blockset_synchronize waits).  (No reference counterpart: the reference
 * builds each symmetry by its own call, main_basis_setup.f90:104-115.) */
int64_t bs2e_blockset_set_graph(bs2e_blockset *set, int64_t on);
int64_t bs2e_blockset_replay(bs2e_blockset *set);
int64_t bs2e_blockset_synchronize(bs2e_blockset *set);
/* Device time (ms) of the count launch (ms[0]) and the last fill (ms[1]). */
int64_t bs2e_blockset_times(const bs2e_blockset *set, double *ms, int64_t n);
int64_t bs2e_blockset_release(bs2e_blockset *set);

/* ---- dipole operators -------------------------------------------------- */
/* mat_els.f90:123-170 setup_radial_dip(b_splines, k_GL, gauge, radial_dip)
 * with :348-370 compute_radial_dip_len / :372-390 compute_radial_dip_vel:
 * gauge 0 (length) writes m1 = r_mat(i,j) = sum w r B_i B_j; gauge 1
 * (velocity) writes m1 = dr_mat = sum (0,-1) w B_i B_j' and m2 = r_inv_mat =
 * sum (0,-1) w B_i B_j / r, the sums over the intervals holding both splines
 * and their GL points.  n_b x n_b double complex, column-major. */
int64_t bs2e_setup_radial_dip(bs2e_ctx *ctx, int64_t k, int64_t n_knots, const double *knots, int64_t k_gl,
                              int64_t gauge, double *m1, double *m2);

typedef struct bs2e_dip bs2e_dip; /* the CSR pattern of one dipole block */

/* dipole.f90:8-47 construct_dip_block_tensor(syms, q, b_splines, S,
 * radial_dip, dip_block, compute), first half (with :95-150
 * init_dip_block): rows are the configurations of syms(1) = (L1, M1, pi1),
 * columns those of syms(2), both as bs2e_count_configs writes them; q is the
 * spherical component (-1, 0, 1).  Counts on the device the entries whose
 * radial pairs share support (direct or exchange) and whose angular factor
 * (mat_els.f90:811-828 ang_dip_red) exceeds 5e-16.  A block without a parity
 * change, with a vanishing Wigner-Eckart factor or with compute = 0 has no
 * entries.  k: the spline order (<= 16); n_b: the radial basis size. */
int64_t bs2e_dip_block_create(bs2e_ctx *ctx, int64_t k, int64_t n_b, int64_t L1, int64_t M1, int64_t pi1,
                              int64_t n1, const int64_t *cf1, int64_t L2, int64_t M2, int64_t pi2, int64_t n2,
                              const int64_t *cf2, int64_t q, int64_t compute, bs2e_dip **out);
/* sizes[0..2] = nnz, rows (syms(1)%n_config), columns (syms(2)%n_config). */
int64_t bs2e_dip_block_sizes(const bs2e_dip *d, int64_t *sizes, int64_t n);
/* dipole.f90:33-46, second half: fills the 1-based CSR arrays (ip: rows+1;
 * ix, data: nnz; data double complex interleaved) with ang * dip_mat_neq
 * (mat_els.f90:717-737 with :739-770 dip_red_mat and :772-809
 * dip_red_1p_len / _vel).  gauge 0: rad1 = r_mat; gauge 1: rad1 = dr_mat,
 * rad2 = r_inv_mat (as bs2e_setup_radial_dip writes them); S: the overlap. */
int64_t bs2e_dip_block_fill(bs2e_dip *d, int64_t gauge, const double *S, const double *rad1, const double *rad2,
                            int64_t *ip, int64_t *ix, double *data);
/* Device time (ms) of the count kernel (ms[0]) and the last fill (ms[1]). */
int64_t bs2e_dip_block_times(const bs2e_dip *d, double *ms, int64_t n);
int64_t bs2e_dip_block_release(bs2e_dip *d);

/* ---- build service ----------------------------------------------------- */
/* No reference counterpart: the reference builds one basis per process
 * (main_basis_setup.f90:75-86).  A service owns an HBM budget (one device
 * allocation, sub-allocated without hipMalloc/hipFree) and `slots` pairs of
 * default-priority streams.  Jobs are R^k builds of differing bases (orders,
 * knot vectors, GL orders, multipole counts); a queued job is admitted as
 * soon as a slot is free and its footprint fits the free budget (in arrival
 * order, smaller jobs passing one that does not fit yet), so builds of
 * several jobs overlap on the device.  Each job's tensor is computed by the
 * same kernels as bs2e_rk_compute, so its values do not depend on what else
 * runs.  Calls on one service must come from one thread at a time. */
typedef struct bs2e_server bs2e_server;
int64_t bs2e_server_create(int64_t device, int64_t hbm_bytes, int64_t slots, bs2e_server **out);
int64_t bs2e_server_destroy(bs2e_server *srv);
const char *bs2e_server_last_error(const bs2e_server *srv);
/* Device bytes one build of this basis allocates (tables + tensor). */
int64_t bs2e_footprint(int64_t k, int64_t n_knots, const double *knots, int64_t k_gl, int64_t max_k, int64_t *bytes);
/* Queues a job and admits what fits; returns its ticket (>= 0) or a
 * negative status (BS2E_ENOMEM when the job alone exceeds the budget).  The
 * job is the R^k build alone (plain FP64); see bs2e_server_submit_job. */
int64_t bs2e_server_submit(bs2e_server *srv, int64_t k, int64_t n_knots, const double *knots, int64_t k_gl,
                           int64_t max_k);

/* One request of the service: the two-electron branch of basis_setup
 * (main_basis_setup.f90:75-115) for one basis: the R^k tensor, the
 * one-electron matrices S and H(l) (mat_els.f90:47-118; hydrogenic charge Z,
 * absorbing potential -i eta (r - cap_r0)^cap_order) and the H/S CSR blocks
 * of n_sym symmetries (hamiltonian.f90:106-283), as bs2e_blockset_create /
 * _fill describe them.  n_sym = 0: the tensor only.  The arrays are copied at
 * submission; the blocks' host tables (channel runs, angular coefficients,
 * row lists) are then built on the service's worker threads while the job
 * waits for a slot. */
typedef struct bs2e_job {
    int64_t k;                  /* B-spline order */
    int64_t n_knots;
    const double *knots;
    int64_t k_gl;               /* GL points per knot interval */
    int64_t max_k;              /* multipoles 0 .. max_k */
    int64_t precision;          /* BS2E_PREC_FP64 or BS2E_PREC_MIXED */
    int64_t n_sym;              /* symmetry blocks */
    const int64_t *L;           /* [n_sym] total L of each symmetry */
    const int64_t *n_config;    /* [n_sym] configurations of each symmetry */
    const int64_t *configs;     /* 5 int64 per configuration (n1, n2, l1, l2, eqv), symmetry after symmetry */
    int64_t full;               /* 0: upper triangles only */
    int64_t Z;                  /* nuclear charge of the one-particle Hamiltonian */
    int64_t max_l_1p;           /* H(l) for l = 0 .. max_l_1p; every configuration's l within it */
    int64_t cap_order;
    double cap_r0, cap_eta_re, cap_eta_im;
    /* With configs NULL: [n_sym] pointers, symmetry s's n_config[s]
     * configurations each (no concatenated copy on the caller's side);
     * not read when configs is given. */
    const int64_t *const *sym_configs;
} bs2e_job;

/* Queues a job of the two-electron branch and admits what fits: a job is
 * admitted when a slot is free and its footprint fits (tensor and tables at
 * first; the blocks' exact sizes once its configurations are counted, a pass
 * of a fraction of a millisecond at admission).  Its tensor build, the
 * one-electron matrices and the block fill run one after the other on the
 * slot's streams, beside the other admitted jobs.  The spline tables and
 * one-electron matrices of a basis stay on the device in the service's cache
 * (keyed by knots, GL rule, charge and absorbing potential) and are reused by
 * later jobs on the same basis; entries no running job holds are evicted,
 * least recently used first, when a job needs their memory.  Completion (the
 * status, finished and latency of a job) is the H/S CSR of every symmetry
 * written. */
int64_t bs2e_server_submit_job(bs2e_server *srv, const bs2e_job *job);
/* A finished job's blocks (read them with bs2e_blockset_sizes, _get,
 * _get_rows); they belong to the service and stay valid until the job is
 * retired.  BS2E_ESTATE for a job without blocks. */
int64_t bs2e_server_blocks(bs2e_server *srv, int64_t ticket, bs2e_blockset **set);
/* info[0..3] = cache entries, their device bytes, admissions that found their
 * basis in the cache, admissions that computed it. */
int64_t bs2e_server_cache(const bs2e_server *srv, int64_t *info, int64_t n);
/* Retires finished builds and admits queued jobs; returns the number of
 * jobs still queued or running. */
int64_t bs2e_server_progress(bs2e_server *srv);
/* state: 0 queued, 1 running, 2 done, 3 failed; times[0] = admission and
 * times[1] = completion (the job's last kernel done: the tensor, or the H/S
 * blocks of a job with blocks), ms after submission (-1 when not yet). */
int64_t bs2e_server_status(bs2e_server *srv, int64_t ticket, int64_t *state, double *times);
/* Blocks until the job is done, admitting others meanwhile. */
int64_t bs2e_server_wait(bs2e_server *srv, int64_t ticket);
/* The finished job's tensor (read it with bs2e_rk_copy_to_host, _get_val,
 * _device_ptr or the block calls); it stays valid until the job is retired
 * and belongs to the service (do not bs2e_rk_release it). */
int64_t bs2e_server_tensor(bs2e_server *srv, int64_t ticket, bs2e_rk **rk);
/* Progresses the service and hands out the jobs that finished since the
 * previous call (at most cap): their tickets and their latency (ms from
 * submission to completion; -1 for a job that failed, whose reason
 * bs2e_server_last_error gives and whose status is state 3).  Returns the
 * count. */
int64_t bs2e_server_finished(bs2e_server *srv, int64_t *tickets, double *latency_ms, int64_t cap);
/* Returns a finished job's memory to the budget. */
int64_t bs2e_server_retire(bs2e_server *srv, int64_t ticket);
/* mem[0] = budget bytes, mem[1] = bytes held by admitted or finished jobs. */
int64_t bs2e_server_memory(const bs2e_server *srv, int64_t *mem);

#ifdef __cplusplus
}
#endif

#endif /* BS2E_H */
