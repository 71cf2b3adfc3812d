// linalg.hpp - dense FP64 linear algebra for the host SCF driver (no Eigen).
//
// Column-major storage throughout (element (r, c) of an n x n matrix at r + c*n), as the
// reference's Eigen::MatrixXd.
//   gemm        C = alpha op(A) op(B) + beta C, OpenMP over output columns
//   eigh        symmetric eigensolver: Householder tridiagonalization followed by the implicit
//               QL iteration with Wilkinson shifts (Golub & Van Loan, "Matrix Computations",
//               4th ed., Alg. 8.3.1 and Sec. 8.3.5; Bowdler, Martin, Reinsch & Wilkinson,
//               Numer. Math. 11 (1968) 293).  Eigenvalues ascending, eigenvectors in columns.
//   lstsq_qrcp  least-squares solve by Householder QR with column pivoting (Golub & Van Loan
//               Alg. 5.4.1), columns beyond the numerical rank set to zero.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <limits>
#include <numeric>
#include <vector>

namespace hfx_la {

inline void gemm(bool ta, bool tb, int m, int n, int k, double alpha, const double* A, int lda, const double* B,
                 int ldb, double beta, double* C, int ldc) {
    // op(A) is m x k, op(B) is k x n.  Transposed operands are copied first so the inner loop
    // is a contiguous axpy over rows of C.
    std::vector<double> At, Bt;
    const double* a = A;
    int la = lda;
    if (ta) {
        At.resize((size_t)m * k);
        for (int i = 0; i < m; ++i)
            for (int p = 0; p < k; ++p) At[i + (size_t)p * m] = A[p + (size_t)i * lda];
        a = At.data();
        la = m;
    }
    const double* b = B;
    int lb = ldb;
    if (tb) {
        Bt.resize((size_t)k * n);
        for (int p = 0; p < k; ++p)
            for (int j = 0; j < n; ++j) Bt[p + (size_t)j * k] = B[j + (size_t)p * ldb];
        b = Bt.data();
        lb = k;
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j) {
        double* c = C + (size_t)j * ldc;
        if (beta == 0.0)
            for (int i = 0; i < m; ++i) c[i] = 0.0;
        else if (beta != 1.0)
            for (int i = 0; i < m; ++i) c[i] *= beta;
        for (int p = 0; p < k; ++p) {
            const double s = alpha * b[p + (size_t)j * lb];
            if (s == 0.0) continue;
            const double* ac = a + (size_t)p * la;
            for (int i = 0; i < m; ++i) c[i] += s * ac[i];
        }
    }
}

// C = A * B for square n x n matrices.
inline void mm(int n, const double* A, const double* B, double* C, bool ta = false, bool tb = false) {
    gemm(ta, tb, n, n, n, 1.0, A, n, B, n, 0.0, C, n);
}

// Symmetric eigensolver.  A (n x n, col-major, symmetric) is not modified.
inline void eigh(int n, const double* A, double* w, double* V) {
    if (n <= 0) return;
    std::vector<double> a(A, A + (size_t)n * n);  // working copy, col-major
    std::vector<double> d(n), e(n, 0.0);
    // Householder reflectors v_k (stored in column k below the subdiagonal) and tau_k.
    std::vector<double> tau(n, 0.0);
    std::vector<double> p(n), wv(n);
    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1;  // length of x = a(k+1:n, k)
        double* x = &a[(k + 1) + (size_t)k * n];
        double sig = 0.0;
        for (int i = 1; i < m; ++i) sig += x[i] * x[i];
        if (sig == 0.0) {
            tau[k] = 0.0;
            continue;
        }
        const double x0 = x[0];
        const double nrm = std::sqrt(x0 * x0 + sig);
        const double alpha = (x0 >= 0.0) ? -nrm : nrm;
        const double v0 = x0 - alpha;
        // v = [v0, x1..], tau = 2 / (v^T v)
        const double vtv = v0 * v0 + sig;
        const double t = 2.0 / vtv;
        x[0] = v0;
        tau[k] = t;
        // p = t * A22 v  (A22 = a(k+1:n, k+1:n), symmetric, full storage kept up to date)
        double* A22 = &a[(k + 1) + (size_t)(k + 1) * n];
#pragma omp parallel for schedule(static) if (m > 128)
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            for (int j = 0; j < m; ++j) s += A22[j + (size_t)i * n] * x[j];  // column i == row i
            p[i] = t * s;
        }
        double pv = 0.0;
        for (int i = 0; i < m; ++i) pv += p[i] * x[i];
        const double hk = 0.5 * t * pv;
        for (int i = 0; i < m; ++i) wv[i] = p[i] - hk * x[i];
#pragma omp parallel for schedule(static) if (m > 128)
        for (int j = 0; j < m; ++j) {
            double* col = A22 + (size_t)j * n;
            const double xj = x[j], wj = wv[j];
            for (int i = 0; i < m; ++i) col[i] -= x[i] * wj + wv[i] * xj;
        }
        e[k] = alpha;  // subdiagonal entry after the reflection
    }
    for (int i = 0; i < n; ++i) d[i] = a[i + (size_t)i * n];
    for (int k = 0; k + 1 < n; ++k)
        if (k + 2 >= n || tau[k] == 0.0) e[k] = a[(k + 1) + (size_t)k * n];
    // Accumulate Q = H_0 H_1 ... H_{n-3}, stored row-major in z for the rotation phase.
    std::vector<double> q((size_t)n * n, 0.0);  // col-major
    for (int i = 0; i < n; ++i) q[i + (size_t)i * n] = 1.0;
    for (int k = n - 3; k >= 0; --k) {
        if (tau[k] == 0.0) continue;
        const int m = n - k - 1;
        const double* v = &a[(k + 1) + (size_t)k * n];
        const double t = tau[k];
#pragma omp parallel for schedule(static) if (m > 128)
        for (int j = k + 1; j < n; ++j) {
            double* col = &q[(k + 1) + (size_t)j * n];
            double s = 0.0;
            for (int i = 0; i < m; ++i) s += v[i] * col[i];
            s *= t;
            for (int i = 0; i < m; ++i) col[i] -= s * v[i];
        }
    }
    std::vector<double> z((size_t)n * n);  // row-major copy of q
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) z[(size_t)r * n + c] = q[r + (size_t)c * n];

    // Implicit QL on (d, e); e[i] couples d[i] and d[i+1].
    std::vector<double> cs(n), sn(n);
    const double eps = std::numeric_limits<double>::epsilon();
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        int m;
        do {
            for (m = l; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= eps * dd) break;
            }
            if (m != l) {
                if (++iter > 200) break;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? std::fabs(r) : -std::fabs(r)));
                double s = 1.0, c = 1.0, pp = 0.0;
                int i;
                int lo = l;
                bool early = false;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i];
                    const double b = c * e[i];
                    r = std::hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= pp;
                        e[m] = 0.0;
                        early = true;
                        lo = i + 1;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - pp;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    pp = s * r;
                    d[i + 1] = g + pp;
                    g = c * r - b;
                    cs[i] = c;
                    sn[i] = s;
                }
                // Apply the sweep's rotations (i = m-1 .. lo) to every row of z.
                const int ilo = early ? lo : l;
#pragma omp parallel for schedule(static) if (n > 128)
                for (int row = 0; row < n; ++row) {
                    double* zr = &z[(size_t)row * n];
                    for (int ii = m - 1; ii >= ilo; --ii) {
                        const double f = zr[ii + 1];
                        zr[ii + 1] = sn[ii] * zr[ii] + cs[ii] * f;
                        zr[ii] = cs[ii] * zr[ii] - sn[ii] * f;
                    }
                }
                if (early) continue;
                d[l] -= pp;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    // Sort ascending.
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
    for (int c = 0; c < n; ++c) {
        w[c] = d[idx[c]];
        for (int r = 0; r < n; ++r) V[r + (size_t)c * n] = z[(size_t)r * n + idx[c]];
    }
}

// Solve min ||B x - b|| for a small dense B (rows x cols, col-major) by QR with column
// pivoting; columns beyond the numerical rank get zero coefficients.
inline std::vector<double> lstsq_qrcp(int rows, int cols, std::vector<double> B, std::vector<double> b) {
    std::vector<int> perm(cols);
    std::iota(perm.begin(), perm.end(), 0);
    std::vector<double> cn(cols);
    for (int j = 0; j < cols; ++j) {
        double s = 0;
        for (int i = 0; i < rows; ++i) s += B[i + (size_t)j * rows] * B[i + (size_t)j * rows];
        cn[j] = s;
    }
    const int kmax = std::min(rows, cols);
    double maxpiv = 0.0;
    int rank = kmax;
    const double thr = std::numeric_limits<double>::epsilon() * kmax;
    for (int k = 0; k < kmax; ++k) {
        int piv = k;
        for (int j = k + 1; j < cols; ++j)
            if (cn[j] > cn[piv]) piv = j;
        if (piv != k) {
            for (int i = 0; i < rows; ++i) std::swap(B[i + (size_t)k * rows], B[i + (size_t)piv * rows]);
            std::swap(cn[k], cn[piv]);
            std::swap(perm[k], perm[piv]);
        }
        double s = 0.0;
        for (int i = k; i < rows; ++i) s += B[i + (size_t)k * rows] * B[i + (size_t)k * rows];
        const double nrm = std::sqrt(s);
        if (k == 0) maxpiv = nrm;
        if (nrm <= thr * maxpiv || nrm == 0.0) {
            rank = k;
            break;
        }
        const double x0 = B[k + (size_t)k * rows];
        const double alpha = (x0 >= 0.0) ? -nrm : nrm;
        std::vector<double> v(rows - k);
        v[0] = x0 - alpha;
        for (int i = k + 1; i < rows; ++i) v[i - k] = B[i + (size_t)k * rows];
        double vtv = 0.0;
        for (double t : v) vtv += t * t;
        const double tau = (vtv > 0.0) ? 2.0 / vtv : 0.0;
        for (int j = k; j < cols; ++j) {
            double dot = 0.0;
            for (int i = k; i < rows; ++i) dot += v[i - k] * B[i + (size_t)j * rows];
            dot *= tau;
            for (int i = k; i < rows; ++i) B[i + (size_t)j * rows] -= dot * v[i - k];
        }
        double dot = 0.0;
        for (int i = k; i < rows; ++i) dot += v[i - k] * b[i];
        dot *= tau;
        for (int i = k; i < rows; ++i) b[i] -= dot * v[i - k];
        for (int j = k + 1; j < cols; ++j) {  // downdate column norms
            const double r = B[k + (size_t)j * rows];
            cn[j] = std::max(0.0, cn[j] - r * r);
        }
    }
    std::vector<double> y(cols, 0.0);
    for (int k = rank - 1; k >= 0; --k) {
        double s = b[k];
        for (int j = k + 1; j < rank; ++j) s -= B[k + (size_t)j * rows] * y[j];
        y[k] = s / B[k + (size_t)k * rows];
    }
    std::vector<double> x(cols, 0.0);
    for (int k = 0; k < cols; ++k) x[perm[k]] = y[k];
    return x;
}

}  // namespace hfx_la
