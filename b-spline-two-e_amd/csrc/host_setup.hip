// host_setup.hip -- host-side set-up of one basis: knot grid, Gauss-Legendre
// rule, banded pair table and the sparse_6d ordering.  This is O(n_b k^2)
// bookkeeping done once per basis; every integral is computed on the GPU.
#include <array>
#include <cmath>
#include <cstdlib>

#include "bs2e_internal.h"

namespace bs2e {

// grid_tools.f90:6-55 generate_grid: k zeros, m steps of h = 2^-m, geometric
// growth x(1+h) until the step reaches Z*h_max, linear steps of Z*h_max up to
// Z*r_max, k-1 repeats of the last knot, everything divided by Z.  The
// reference fills an array of 2k + m + ceiling(r_max/h) knots (:20); a grid
// that would not fit (r_max too small for the geometric ramp) overruns it
// there, and is refused here: returns -1.
int64_t generate_grid(int64_t k, int64_t m, int64_t Z, double h_max, double r_max, std::vector<double>& out)
{
    const double h = std::ldexp(1.0, (int)-m);
    const double step_cap = h_max * (double)Z;
    const double r_end = (double)Z * r_max;
    const double cap_d = 2.0 * (double)k + (double)m + std::ceil(r_max / h);
    if (!(cap_d < 1e9)) return -1;
    const size_t cap = (size_t)cap_d;
    std::vector<double> g((size_t)k, 0.0);
    g.reserve(cap);
    for (int64_t i = 0; i < m; ++i) g.push_back(g.back() + h);
    for (;;) {
        double cur = g.back();
        double next = cur * (1.0 + h);
        if ((next - cur) >= step_cap) break;
        if (g.size() + 1 > cap) return -1;
        g.push_back(next);
    }
    while (g.back() < r_end) {
        if (g.size() + 1 > cap) return -1;
        g.push_back(g.back() + (double)Z * h_max);
    }
    if (g.size() + (size_t)(k - 1) > cap) return -1;
    double last = g.back();
    for (int64_t i = 0; i < k - 1; ++i) g.push_back(last);
    out.resize(g.size());
    for (size_t i = 0; i < g.size(); ++i) out[i] = g[i] / (double)Z;
    return (int64_t)out.size();
}

// N-point Gauss-Legendre rule on [-1, 1] (the reference calls
// stdlib_quadrature's gauss_legendre through quad_tools.f90:14-27): Newton's
// method on P_N in extended precision from the Tricomi initial guess, weights
// 2 / ((1 - x^2) P_N'(x)^2), rounded once to double.
void gauss_legendre(int64_t N, double* x, double* w)
{
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int64_t i = 0; i < (N + 1) / 2; ++i) {
        long double z = -std::cos(pi * ((long double)i + 0.75L) / ((long double)N + 0.5L));
        long double pN = 0, pNm1 = 0;
        auto legendre = [&](long double zz) {
            long double a = 1.0L, b = zz;
            if (N == 0) { pN = 1.0L; pNm1 = 0.0L; return; }
            for (int64_t l = 2; l <= N; ++l) {
                long double c = ((2 * l - 1) * zz * b - (l - 1) * a) / (long double)l;
                a = b;
                b = c;
            }
            pN = b;
            pNm1 = a;
        };
        for (int it = 0; it < 64; ++it) {
            legendre(z);
            long double dp = (long double)N * (z * pN - pNm1) / (z * z - 1.0L);
            long double dz = pN / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-19L) break;
        }
        legendre(z);
        long double dp = (long double)N * (z * pN - pNm1) / (z * z - 1.0L);
        long double ww = 2.0L / ((1.0L - z * z) * dp * dp);
        x[i] = (double)z;
        x[N - 1 - i] = (double)-z;
        w[i] = w[N - 1 - i] = (double)ww;
    }
    if (N % 2 == 1) x[N / 2] = 0.0;
}

std::string make_plan(int64_t k_s, int64_t n_knots, const double* knots, int64_t k_gl, int64_t max_k, Plan& p)
{
    if (k_s < 2 || k_s > kMaxOrder) return "spline order k must be in [2, 20] (BVALUE_D work arrays are dimension(20))";
    if (k_gl < 1 || k_gl > kMaxGL) return "k_GL must be in [1, 64]";
    if (max_k < 0 || max_k >= kMaxK) return "max_k must be in [0, 63]";
    if (!knots) return "knot vector is NULL";
    p = Plan();
    p.k_s = k_s;
    p.n_knots = n_knots;
    p.n = n_knots - k_s;                          // bspline_tools.f90:33
    p.n_b = p.n - 2;                              // bspline_tools.f90:34
    p.n_iv = n_knots - 2 * k_s + 1;               // size(breakpoints) - 1
    if (p.n_b < 1 || p.n_iv < 1) return "knot vector too short for the spline order";
    p.k_gl = k_gl;
    p.max_k = max_k;
    p.K = max_k + 1;
    p.t.assign(knots, knots + n_knots);
    for (int64_t i = 1; i < n_knots; ++i)
        if (!(p.t[i] >= p.t[i - 1])) return "knot vector must be non-decreasing";
    p.bp.assign(p.t.begin() + (k_s - 1), p.t.begin() + (k_s - 1) + p.n_iv + 1);   // t(k : size(t)-k+1)
    for (int64_t v = 0; v < p.n_iv; ++v)
        if (!(p.bp[v + 1] > p.bp[v])) return "breakpoints must be strictly increasing";
    if (!(p.bp[0] >= 0.0)) return "breakpoints must be non-negative (r >= 0)";
    p.x.resize(k_gl);
    p.w.resize(k_gl);
    gauss_legendre(k_gl, p.x.data(), p.w.data());

    // Banded pairs in count_nnz_R_k order with their common support window.
    const int64_t nb = p.n_b, ks = k_s, niv = p.n_iv;
    p.c_first.assign((size_t)nb + 2, 0);
    p.off4.clear();
    int64_t nnz4 = 0;
    for (int64_t c = 1; c <= nb; ++c) {
        p.c_first[c] = (int32_t)p.pairs.size();
        for (int64_t a = std::max<int64_t>(1, c - ks + 1); a <= std::min<int64_t>(nb, c + ks - 1); ++a) {
            int64_t lo = std::max<int64_t>(std::max(a, c) + 2 - ks, 1);
            int64_t hi = std::min<int64_t>(std::min(a, c) + 1, niv);
            PairInfo pi{(int32_t)a, (int32_t)c, (int32_t)lo, (int32_t)(hi - lo + 1)};
            if (pi.m < 1) return "internal: empty pair window";
            p.pairs.push_back(pi);
            p.off4.push_back(nnz4);
            nnz4 += pi.m;
        }
    }
    p.c_first[nb + 1] = (int32_t)p.pairs.size();
    p.n_pairs = (int64_t)p.pairs.size();
    p.nnz4 = nnz4;
    p.U = ks * (ks + 1) / 2;

    // nnz6 = sum over (J, I) of the overlap of their windows (sparse_array_tools.f90:298-324)
    //      = sum over intervals v of (number of pairs whose window holds v)^2.
    std::vector<int64_t> cover((size_t)niv + 2, 0);
    for (const PairInfo& pi : p.pairs) {
        cover[pi.v0] += 1;
        cover[pi.v0 + pi.m] -= 1;
    }
    int64_t nnz6 = 0, run = 0;
    for (int64_t v = 1; v <= niv; ++v) {
        run += cover[v];
        nnz6 += run * run;
    }
    p.nnz6 = nnz6;
    return std::string();
}

// The sparse_6d order of mat_els.f90:251-289: j_b_p, j_b (pair J = (j_b, j_b_p),
// loop order), i_b_p, i_b (pair I = (i_b, i_b_p), loop order), then ascending
// intervals common to both windows.
void build_diag_entries(Plan& p)
{
    if (!p.diag.empty() || p.nnz6 == 0) return;
    const int64_t nb = p.n_b, ks = p.k_s;
    int64_t ptr = 0;
    for (int64_t J = 0; J < p.n_pairs; ++J) {
        const PairInfo& pj = p.pairs[J];
        int64_t cmin = std::max<int64_t>(1, pj.v0 - 1);
        int64_t cmax = std::min<int64_t>(nb, pj.v0 + pj.m - 1 + ks - 2);
        for (int64_t I = p.c_first[cmin]; I < p.c_first[cmax + 1]; ++I) {
            const PairInfo& pi = p.pairs[I];
            int64_t lo = std::max(pi.v0, pj.v0), hi = std::min(pi.v0 + pi.m, pj.v0 + pj.m) - 1;
            if (hi < lo) continue;
            p.diag.push_back(DiagEntry{(int32_t)J, (int32_t)I, (int32_t)lo, (int32_t)(hi - lo + 1), ptr});
            ptr += hi - lo + 1;
        }
    }
}

// ---------------------------------------------------------------------------
// Angular algebra (wigner_tools.f90:30-60, 107-138).  The reference calls
// GSL's gsl_sf_coupling_3j/6j; here Racah's closed formulas are evaluated
// with exact integer factorial ratios accumulated in long double.
// ---------------------------------------------------------------------------
static long double fact(int64_t n)
{
    // built once by the first call (static initialization is thread-safe: the
    // build service's worker threads compute angular coefficients at once)
    static const std::array<long double, 401> tab = [] {
        std::array<long double, 401> t{};
        t[0] = 1.0L;
        for (int i = 1; i <= 400; ++i) t[i] = t[i - 1] * (long double)i;
        return t;
    }();
    return (n < 0 || n > 400) ? NAN : tab[(size_t)n];
}

static bool triangle(int64_t a, int64_t b, int64_t c)
{
    return a >= 0 && b >= 0 && c >= 0 && c >= std::llabs(a - b) && c <= a + b;
}

static long double delta(int64_t a, int64_t b, int64_t c)
{
    return fact(a + b - c) * fact(a - b + c) * fact(b + c - a) / fact(a + b + c + 1);
}

double wigner_3j(int64_t ja, int64_t jb, int64_t jc, int64_t ma, int64_t mb, int64_t mc)
{
    if (ma + mb + mc != 0 || !triangle(ja, jb, jc)) return 0.0;
    if (std::llabs(ma) > ja || std::llabs(mb) > jb || std::llabs(mc) > jc) return 0.0;
    int64_t tlo = std::max<int64_t>({0, jb - jc - ma, ja - jc + mb});
    int64_t thi = std::min<int64_t>({ja + jb - jc, ja - ma, jb + mb});
    long double s = 0.0L;
    for (int64_t t = tlo; t <= thi; ++t) {
        long double term = 1.0L / (fact(t) * fact(jc - jb + t + ma) * fact(jc - ja + t - mb) * fact(ja + jb - jc - t) *
                                   fact(ja - t - ma) * fact(jb - t + mb));
        s += (t & 1) ? -term : term;
    }
    long double norm = std::sqrt(delta(ja, jb, jc) * fact(ja + ma) * fact(ja - ma) * fact(jb + mb) * fact(jb - mb) *
                                 fact(jc + mc) * fact(jc - mc));
    long double sign = (std::llabs(ja - jb - mc) & 1) ? -1.0L : 1.0L;
    return (double)(sign * norm * s);
}

double wigner_6j(int64_t j1, int64_t j2, int64_t j3, int64_t j4, int64_t j5, int64_t j6)
{
    if (!triangle(j1, j2, j3) || !triangle(j1, j5, j6) || !triangle(j4, j2, j6) || !triangle(j4, j5, j3)) return 0.0;
    int64_t a[4] = {j1 + j2 + j3, j1 + j5 + j6, j4 + j2 + j6, j4 + j5 + j3};
    int64_t b[3] = {j1 + j2 + j4 + j5, j2 + j3 + j5 + j6, j3 + j1 + j6 + j4};
    int64_t tlo = std::max({a[0], a[1], a[2], a[3]});
    int64_t thi = std::min({b[0], b[1], b[2]});
    long double s = 0.0L;
    for (int64_t t = tlo; t <= thi; ++t) {
        long double term = fact(t + 1) / (fact(t - a[0]) * fact(t - a[1]) * fact(t - a[2]) * fact(t - a[3]) *
                                          fact(b[0] - t) * fact(b[1] - t) * fact(b[2] - t));
        s += (t & 1) ? -term : term;
    }
    long double norm = std::sqrt(delta(j1, j2, j3) * delta(j1, j5, j6) * delta(j4, j2, j6) * delta(j4, j5, j3));
    return (double)(norm * s);
}

static double neg1pow(int64_t e) { return (std::llabs(e) & 1) ? -1.0 : 1.0; }

// wigner_tools.f90:107-112 C_red_mat(k, a, b) and :126-138 ang_k_LS.
double c_red(int64_t k, int64_t a, int64_t b)
{
    return neg1pow(a) * std::sqrt((double)((2 * a + 1) * (2 * b + 1))) * wigner_3j(a, k, b, 0, 0, 0);
}

double ang_k_LS(int64_t k, int64_t l1a, int64_t l1b, int64_t l2a, int64_t l2b, int64_t L)
{
    if (((l1a + k + l2a) & 1) || ((l1b + k + l2b) & 1)) return 0.0;
    return neg1pow(l1b + l2a + L) * wigner_6j(l1a, l1b, L, l2b, l2a, k) * c_red(k, l1a, l2a) * c_red(k, l1b, l2b);
}

}  // namespace bs2e
