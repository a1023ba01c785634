// Dense symmetric eigensolver for the SCF driver (host, FP64): Householder
// reduction to tridiagonal form followed by the implicit-shift QL iteration
// (Wilkinson & Reinsch, Handbook for Automatic Computation II, 1971).  The
// reference uses faer's SelfAdjointEigendecomposition (src/scf.rs:607-705);
// only eigenvalues and the spanned subspaces matter downstream.
#include "linalg.h"

#include <cmath>
#include <stdexcept>
#include <vector>

namespace librint {

void eigh(int n, const double *Ain, double *w, double *V)
{
    std::vector<double> a(Ain, Ain + (size_t)n * n);
    std::vector<double> d(n), e(n);
    auto A = [&](int i, int j) -> double & { return a[(size_t)i * n + j]; };
    // Householder tridiagonalisation, accumulating the transform in a
    for (int i = n - 1; i > 0; i--) {
        const int l = i - 1;
        double h = 0.0, scale = 0.0;
        if (l > 0) {
            for (int k = 0; k <= l; k++) scale += std::fabs(A(i, k));
            if (scale == 0.0) {
                e[i] = A(i, l);
            } else {
                for (int k = 0; k <= l; k++) {
                    A(i, k) /= scale;
                    h += A(i, k) * A(i, k);
                }
                double f = A(i, l);
                double g = (f >= 0.0 ? -std::sqrt(h) : std::sqrt(h));
                e[i] = scale * g;
                h -= f * g;
                A(i, l) = f - g;
                f = 0.0;
                for (int j = 0; j <= l; j++) {
                    A(j, i) = A(i, j) / h;
                    g = 0.0;
                    for (int k = 0; k <= j; k++) g += A(j, k) * A(i, k);
                    for (int k = j + 1; k <= l; k++) g += A(k, j) * A(i, k);
                    e[j] = g / h;
                    f += e[j] * A(i, j);
                }
                const double hh = f / (h + h);
                for (int j = 0; j <= l; j++) {
                    f = A(i, j);
                    e[j] = g = e[j] - hh * f;
                    for (int k = 0; k <= j; k++) A(j, k) -= (f * e[k] + g * A(i, k));
                }
            }
        } else {
            e[i] = A(i, l);
        }
        d[i] = h;
    }
    d[0] = 0.0;
    e[0] = 0.0;
    for (int i = 0; i < n; i++) {
        const int l = i - 1;
        if (d[i] != 0.0) {
            for (int j = 0; j <= l; j++) {
                double g = 0.0;
                for (int k = 0; k <= l; k++) g += A(i, k) * A(k, j);
                for (int k = 0; k <= l; k++) A(k, j) -= g * A(k, i);
            }
        }
        d[i] = A(i, i);
        A(i, i) = 1.0;
        for (int j = 0; j <= l; j++) A(j, i) = A(i, j) = 0.0;
    }
    // z: rows are the columns of the accumulated transform (contiguous updates)
    std::vector<double> z((size_t)n * n);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) z[(size_t)i * n + k] = A(k, i);
    for (int i = 1; i < n; i++) e[i - 1] = e[i];
    if (n > 0) e[n - 1] = 0.0;
    for (int l = 0; l < n; l++) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; m++) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= 1e-16 * dd) break;
            }
            if (m != l) {
                if (iter++ == 100) throw std::runtime_error("eigh: QL iteration did not converge");
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                for (i = m - 1; i >= l; i--) {
                    double f = s * e[i];
                    const double b = c * e[i];
                    e[i + 1] = (r = std::hypot(f, g));
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    d[i + 1] = g + (p = s * r);
                    g = c * r - b;
                    double *zi = &z[(size_t)i * n], *zi1 = &z[(size_t)(i + 1) * n];
                    for (int k = 0; k < n; k++) {
                        f = zi1[k];
                        zi1[k] = s * zi[k] + c * f;
                        zi[k] = c * zi[k] - s * f;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    // sort ascending; V[k*n + i] = component k of eigenvector i
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    for (int i = 0; i < n; i++) {
        int m = i;
        for (int j = i + 1; j < n; j++)
            if (d[idx[j]] < d[idx[m]]) m = j;
        std::swap(idx[i], idx[m]);
    }
    for (int i = 0; i < n; i++) {
        w[i] = d[idx[i]];
        const double *zr = &z[(size_t)idx[i] * n];
        for (int k = 0; k < n; k++) V[(size_t)k * n + i] = zr[k];
    }
}

void gemm_nn(int n, const double *A, const double *B, double *C)
{
    for (int i = 0; i < n; i++) {
        double *c = C + (size_t)i * n;
        for (int j = 0; j < n; j++) c[j] = 0.0;
        for (int k = 0; k < n; k++) {
            const double a = A[(size_t)i * n + k];
            if (a == 0.0) continue;
            const double *b = B + (size_t)k * n;
            for (int j = 0; j < n; j++) c[j] += a * b[j];
        }
    }
}

void gemm_tn(int n, const double *A, const double *B, double *C)
{
    for (size_t i = 0; i < (size_t)n * n; i++) C[i] = 0.0;
    for (int k = 0; k < n; k++) {
        const double *a = A + (size_t)k * n;
        const double *b = B + (size_t)k * n;
        for (int i = 0; i < n; i++) {
            const double ai = a[i];
            if (ai == 0.0) continue;
            double *c = C + (size_t)i * n;
            for (int j = 0; j < n; j++) c[j] += ai * b[j];
        }
    }
}

}  // namespace librint
