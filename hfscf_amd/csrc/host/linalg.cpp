#include "linalg.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace hfscf {

void gemm(int n, const double* A, bool ta, const double* B, bool tb, double* C) {
    // Materialise transposes so the inner loop is unit-stride.
    std::vector<double> At, Bt;
    const double* a = A;
    const double* b = B;
    if (ta) {
        At.resize((size_t)n * n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) At[(size_t)i * n + j] = A[(size_t)j * n + i];
        a = At.data();
    }
    if (tb) {
        Bt.resize((size_t)n * n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) Bt[(size_t)i * n + j] = B[(size_t)j * n + i];
        b = Bt.data();
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double* c = C + (size_t)i * n;
        for (int j = 0; j < n; ++j) c[j] = 0.0;
        for (int k = 0; k < n; ++k) {
            const double aik = a[(size_t)i * n + k];
            const double* bk = b + (size_t)k * n;
            for (int j = 0; j < n; ++j) c[j] += aik * bk[j];
        }
    }
}

void eigh(int n, const double* Ain, double* w, double* V) {
    std::vector<double> A(Ain, Ain + (size_t)n * n);
    std::vector<double> Q((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
    auto a = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };

    // Householder reduction to tridiagonal form: A <- H A H, Q <- Q H.
    std::vector<double> v(n), p(n), wv(n);
    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1;  // length of the column below the diagonal
        double alpha = 0.0;
        for (int i = 0; i < m; ++i) alpha += a(k + 1 + i, k) * a(k + 1 + i, k);
        alpha = std::sqrt(alpha);
        if (alpha == 0.0) continue;
        const double x0 = a(k + 1, k);
        if (x0 > 0) alpha = -alpha;  // H x = alpha e1 with alpha of opposite sign to x0
        for (int i = 0; i < m; ++i) v[i] = a(k + 1 + i, k);
        v[0] -= alpha;
        double vnorm2 = 0.0;
        for (int i = 0; i < m; ++i) vnorm2 += v[i] * v[i];
        if (vnorm2 == 0.0) continue;
        const double beta = 2.0 / vnorm2;
        // p = beta * A22 v
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            for (int j = 0; j < m; ++j) s += a(k + 1 + i, k + 1 + j) * v[j];
            p[i] = beta * s;
        }
        double pv = 0.0;
        for (int i = 0; i < m; ++i) pv += p[i] * v[i];
        for (int i = 0; i < m; ++i) wv[i] = p[i] - 0.5 * beta * pv * v[i];
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) a(k + 1 + i, k + 1 + j) -= v[i] * wv[j] + wv[i] * v[j];
        a(k + 1, k) = a(k, k + 1) = alpha;
        for (int i = 1; i < m; ++i) a(k + 1 + i, k) = a(k, k + 1 + i) = 0.0;
        // Q <- Q (I - beta v v^T) on columns k+1..n-1
        for (int r = 0; r < n; ++r) {
            double s = 0.0;
            for (int j = 0; j < m; ++j) s += Q[(size_t)r * n + k + 1 + j] * v[j];
            s *= beta;
            for (int j = 0; j < m; ++j) Q[(size_t)r * n + k + 1 + j] -= s * v[j];
        }
    }
    std::vector<double> d(n), e(n, 0.0);
    for (int i = 0; i < n; ++i) d[i] = a(i, i);
    for (int i = 0; i + 1 < n; ++i) e[i] = a(i + 1, i);  // e[i] couples d[i] and d[i+1]

    // Implicit QL iterations with shifts; rotations accumulated into Q's columns.
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m + 1 < n; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= 1e-16 * dd) break;
            }
            if (m == l) break;
            if (++iter > 60) throw std::runtime_error("eigh: no convergence");
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
            double s = 1.0, c = 1.0, pp = 0.0;
            int i = m - 1;
            bool early = false;
            for (; i >= l; --i) {
                double f = s * e[i];
                double b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= pp;
                    e[m] = 0.0;
                    early = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - pp;
                r = (d[i] - g) * s + 2.0 * c * b;
                pp = s * r;
                d[i + 1] = g + pp;
                g = c * r - b;
                for (int k = 0; k < n; ++k) {
                    double* q = &Q[(size_t)k * n];
                    const double t = q[i + 1];
                    q[i + 1] = s * q[i] + c * t;
                    q[i] = c * q[i] - s * t;
                }
            }
            if (early) continue;
            d[l] -= pp;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return d[x] < d[y]; });
    for (int j = 0; j < n; ++j) {
        w[j] = d[order[j]];
        for (int i = 0; i < n; ++i) V[(size_t)i * n + j] = Q[(size_t)i * n + order[j]];
    }
}

std::vector<double> qr_solve(int n, std::vector<double> A, std::vector<double> b) {
    auto a = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };
    for (int k = 0; k < n; ++k) {
        double norm = 0.0;
        for (int i = k; i < n; ++i) norm += a(i, k) * a(i, k);
        norm = std::sqrt(norm);
        if (norm == 0.0) continue;
        const double alpha = a(k, k) > 0 ? -norm : norm;
        std::vector<double> v(n - k);
        for (int i = k; i < n; ++i) v[i - k] = a(i, k);
        v[0] -= alpha;
        double vn = 0.0;
        for (double x : v) vn += x * x;
        if (vn == 0.0) continue;
        const double beta = 2.0 / vn;
        for (int j = k; j < n; ++j) {
            double s = 0.0;
            for (int i = k; i < n; ++i) s += v[i - k] * a(i, j);
            s *= beta;
            for (int i = k; i < n; ++i) a(i, j) -= s * v[i - k];
        }
        double s = 0.0;
        for (int i = k; i < n; ++i) s += v[i - k] * b[i];
        s *= beta;
        for (int i = k; i < n; ++i) b[i] -= s * v[i - k];
    }
    std::vector<double> x(n, 0.0);
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < n; ++j) s -= a(i, j) * x[j];
        x[i] = a(i, i) != 0.0 ? s / a(i, i) : 0.0;
    }
    return x;
}

}  // namespace hfscf
