// Lowest eigenpair of a symmetric tridiagonal matrix in FP64, replacing the
// MKL dstevx('V','A') call of lanczos_solver.f90:175-211 (the reference then
// keeps only ev(1) and z(:,1), lanczos_solver.f90:67-68).
//
// Method: Sturm-sequence multisection for the smallest eigenvalue to full
// precision, then inverse iteration with a partially pivoted tridiagonal LU
// for its eigenvector.  Copies of the same Ritz value that plain Lanczos can
// produce (loss of orthogonality) form a tight cluster; inverse iteration
// then returns a unit vector inside that cluster's span, which maps back to
// the same Hilbert-space eigenvector.
#include <algorithm>
#include <cmath>
#include <vector>

namespace edg {

// Number of eigenvalues strictly below each of the kPts points x[k].  The
// pivot recurrence carries one division per element; the points' recurrences
// are independent, so they run interleaved (and in vector registers), at
// about the cost of one.
constexpr int kPts = 4;
static void sturm_counts(int n, const double *d, const double *e2, const double *x, int *cnt)
{
    const double tiny = 1e-300;
    double q[kPts];
    int c[kPts];
    for (int k = 0; k < kPts; ++k) {
        q[k] = d[0] - x[k];
        c[k] = q[k] < 0;
    }
    for (int i = 1; i < n; ++i)
        for (int k = 0; k < kPts; ++k) {
            const double qk = std::fabs(q[k]) < tiny ? -tiny : q[k];
            q[k] = d[i] - x[k] - e2[i - 1] / qk;
            c[k] += q[k] < 0;
        }
    for (int k = 0; k < kPts; ++k) cnt[k] = c[k];
}

// d[n], e[n-1].  Returns 0 on success, fills *lambda and z[n] (unit norm,
// largest-magnitude component positive).
int tridiag_lowest(int n, const double *d, const double *e, double *lambda, double *z)
{
    if (n < 1) return 1;
    if (n == 1) {
        *lambda = d[0];
        z[0] = 1.0;
        return 0;
    }
    std::vector<double> e2(n - 1);
    double lo = d[0], hi = d[0];
    for (int i = 0; i < n; ++i) {
        const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i < n - 1 ? std::fabs(e[i]) : 0.0);
        lo = std::min(lo, d[i] - r);
        hi = std::max(hi, d[i] + r);
    }
    for (int i = 0; i < n - 1; ++i) e2[i] = e[i] * e[i];
    const double scale = std::max(std::fabs(lo), std::fabs(hi));
    // multisection: kPts interior points cut [lo, hi] into kPts + 1 parts
    for (int it = 0; it < 200; ++it) {
        if (hi - lo <= 4.0 * 2.220446049250313e-16 * scale) break;
        double x[kPts];
        int cnt[kPts];
        for (int k = 0; k < kPts; ++k) x[k] = lo + (hi - lo) * (double)(k + 1) / (double)(kPts + 1);
        if (x[0] <= lo || x[kPts - 1] >= hi) break;
        sturm_counts(n, d, e2.data(), x, cnt);
        int k = 0;
        while (k < kPts && cnt[k] < 1) ++k;  // first point with an eigenvalue below it
        if (k < kPts) hi = x[k];
        if (k > 0) lo = x[k - 1];
    }
    const double lam = 0.5 * (lo + hi);
    *lambda = lam;

    // inverse iteration on (T - sigma I), sigma just below lam
    const double sigma = lam - 8.0 * 2.220446049250313e-16 * (scale > 0 ? scale : 1.0);
    // LU with partial pivoting of the tridiagonal (T - sigma I): rows have
    // up to 3 nonzeros after pivoting (u0 diag, u1, u2 super-diagonals).
    std::vector<double> u0(n), u1(n), u2(n), l(n);
    std::vector<int> piv(n);
    {
        std::vector<double> a(n), b(n), c(n);  // sub, diag, super
        for (int i = 0; i < n; ++i) {
            b[i] = d[i] - sigma;
            a[i] = i > 0 ? e[i - 1] : 0.0;
            c[i] = i < n - 1 ? e[i] : 0.0;
        }
        double rd = b[0], rs = c[0], rs2 = 0.0;  // current row: diag, super1, super2
        for (int i = 0; i < n - 1; ++i) {
            const double sub = a[i + 1];
            double nd = b[i + 1], ns = c[i + 1];
            if (std::fabs(sub) > std::fabs(rd)) {
                // swap rows i and i+1
                piv[i] = 1;
                const double m = rd / sub;
                u0[i] = sub;
                u1[i] = nd;
                u2[i] = ns;
                l[i] = m;
                rd = rs - m * nd;
                rs = rs2 - m * ns;
                rs2 = 0.0;
            } else {
                piv[i] = 0;
                const double m = (rd != 0.0) ? sub / rd : 0.0;
                u0[i] = rd;
                u1[i] = rs;
                u2[i] = rs2;
                l[i] = m;
                rd = nd - m * rs;
                rs = ns - m * rs2;
                rs2 = 0.0;
            }
        }
        u0[n - 1] = rd;
        u1[n - 1] = 0.0;
        u2[n - 1] = 0.0;
        const double eps = 1e-300 + 2.220446049250313e-16 * (scale > 0 ? scale : 1.0) * 1e-3;
        for (int i = 0; i < n; ++i)
            if (std::fabs(u0[i]) < eps) u0[i] = (u0[i] < 0 ? -eps : eps);
    }
    std::vector<double> v(n, 1.0 / std::sqrt((double)n));
    for (int i = 0; i < n; ++i) v[i] *= 1.0 + 1e-3 * std::sin(1.0 + i);  // generic start
    for (int it = 0; it < 4; ++it) {
        // forward: apply L^{-1} with pivoting
        for (int i = 0; i < n - 1; ++i) {
            if (piv[i]) std::swap(v[i], v[i + 1]);
            v[i + 1] -= l[i] * v[i];
        }
        // back substitution U v = w
        for (int i = n - 1; i >= 0; --i) {
            double s = v[i];
            if (i + 1 < n) s -= u1[i] * v[i + 1];
            if (i + 2 < n) s -= u2[i] * v[i + 2];
            v[i] = s / u0[i];
        }
        double nr = 0.0;
        for (int i = 0; i < n; ++i) nr += v[i] * v[i];
        nr = std::sqrt(nr);
        if (!(nr > 0.0) || !std::isfinite(nr)) return 2;
        for (int i = 0; i < n; ++i) v[i] /= nr;
    }
    int imax = 0;
    for (int i = 1; i < n; ++i)
        if (std::fabs(v[i]) > std::fabs(v[imax])) imax = i;
    const double sg = v[imax] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) z[i] = sg * v[i];
    return 0;
}

}  // namespace edg
