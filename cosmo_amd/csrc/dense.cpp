// dense.cpp -- exact diagonalisation of small operators on the host (the reference's
// n <= limit branch, include/xsm/DEXHVcsb.cxx:72-108, which uses Eigen's
// SelfAdjointEigenSolver).  Householder reduction to tridiagonal form followed by the
// implicit QL iteration with Wilkinson-style shifts (Golub & Van Loan, ch. 8), written
// for a row-major symmetric matrix; eigenvectors accumulated.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "internal.h"

namespace cosmo {

// A (n x n, row-major) is overwritten by the eigenvectors (column j = vector j); w = values
static void tridiag_householder(int n, std::vector<double> &A, std::vector<double> &d, std::vector<double> &e) {
  d.assign(n, 0.0);
  e.assign(n, 0.0);
  for (int i = n - 1; i > 0; --i) {
    const int l = i - 1;
    double h = 0.0, scale = 0.0;
    if (l > 0) {
      for (int k = 0; k <= l; ++k) scale += std::fabs(A[(size_t)i * n + k]);
      if (scale == 0.0) {
        e[i] = A[(size_t)i * n + l];
      } else {
        for (int k = 0; k <= l; ++k) {
          A[(size_t)i * n + k] /= scale;
          h += A[(size_t)i * n + k] * A[(size_t)i * n + k];
        }
        double f = A[(size_t)i * n + l];
        double g = (f >= 0.0 ? -std::sqrt(h) : std::sqrt(h));
        e[i] = scale * g;
        h -= f * g;
        A[(size_t)i * n + l] = f - g;
        f = 0.0;
        for (int j = 0; j <= l; ++j) {
          A[(size_t)j * n + i] = A[(size_t)i * n + j] / h;
          g = 0.0;
          for (int k = 0; k <= j; ++k) g += A[(size_t)j * n + k] * A[(size_t)i * n + k];
          for (int k = j + 1; k <= l; ++k) g += A[(size_t)k * n + j] * A[(size_t)i * n + k];
          e[j] = g / h;
          f += e[j] * A[(size_t)i * n + j];
        }
        const double hh = f / (h + h);
        for (int j = 0; j <= l; ++j) {
          f = A[(size_t)i * n + j];
          e[j] = g = e[j] - hh * f;
          for (int k = 0; k <= j; ++k) A[(size_t)j * n + k] -= (f * e[k] + g * A[(size_t)i * n + k]);
        }
      }
    } else {
      e[i] = A[(size_t)i * n + l];
    }
    d[i] = h;
  }
  d[0] = 0.0;
  e[0] = 0.0;
  for (int i = 0; i < n; ++i) {
    const int l = i - 1;
    if (d[i] != 0.0) {
      for (int j = 0; j <= l; ++j) {
        double g = 0.0;
        for (int k = 0; k <= l; ++k) g += A[(size_t)i * n + k] * A[(size_t)k * n + j];
        for (int k = 0; k <= l; ++k) A[(size_t)k * n + j] -= g * A[(size_t)k * n + i];
      }
    }
    d[i] = A[(size_t)i * n + i];
    A[(size_t)i * n + i] = 1.0;
    for (int j = 0; j <= l; ++j) A[(size_t)j * n + i] = A[(size_t)i * n + j] = 0.0;
  }
}

static bool ql_implicit(int n, std::vector<double> &d, std::vector<double> &e, std::vector<double> &Z) {
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  if (n) e[n - 1] = 0.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1e-16 * dd) break;
      }
      if (m != l) {
        if (iter++ == 60) return false;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
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
          for (int k = 0; k < n; ++k) {
            f = Z[(size_t)k * n + i + 1];
            Z[(size_t)k * n + i + 1] = s * Z[(size_t)k * n + i] + c * f;
            Z[(size_t)k * n + i] = c * Z[(size_t)k * n + i] - s * f;
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  return true;
}

// eigenvalues ascending in w, eigenvectors as columns of V (row-major n x n)
bool dense_symmetric_eigen(int n, const std::vector<double> &A, std::vector<double> &w, std::vector<double> &V) {
  std::vector<double> Z(A), d, e;
  tridiag_householder(n, Z, d, e);
  if (!ql_implicit(n, d, e, Z)) return false;
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
  w.resize(n);
  V.assign((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    w[j] = d[idx[j]];
    for (int k = 0; k < n; ++k) V[(size_t)k * n + j] = Z[(size_t)k * n + idx[j]];
  }
  return true;
}

}  // namespace cosmo

extern "C" int cosmo_op_dense_eigen(cosmo_op op, double *evals, double *evecs) {
  using namespace cosmo;
  const int64_t n = cosmo_op_size(op);
  if (n < 0) return n;
  if (n > 20000) return fail(COSMO_EUNSUP, "exact diagonalisation is for small operators (n <= 20000)");
  std::vector<int64_t> cnt(n), rp(n + 1, 0);
  int rc = cosmo_op_upper_rows(op, 0, n, cnt.data(), nullptr, nullptr, nullptr);
  if (rc) return rc;
  for (int64_t r = 0; r < n; ++r) rp[r + 1] = rp[r] + cnt[r];
  std::vector<int64_t> cols(rp[n]);
  std::vector<double> vals(rp[n]);
  rc = cosmo_op_upper_rows(op, 0, n, cnt.data(), rp.data(), cols.data(), vals.data());
  if (rc) return rc;
  std::vector<double> A((size_t)n * n, 0.0), w, V;
  for (int64_t r = 0; r < n; ++r)
    for (int64_t e = rp[r]; e < rp[r + 1]; ++e) A[(size_t)r * n + cols[e]] = A[(size_t)cols[e] * n + r] = vals[e];
  if (!dense_symmetric_eigen((int)n, A, w, V)) return fail(COSMO_ECONV, "QL iteration did not converge");
  std::copy(w.begin(), w.end(), evals);
  if (evecs)
    for (int64_t j = 0; j < n; ++j)
      for (int64_t k = 0; k < n; ++k) evecs[(size_t)j * n + k] = V[(size_t)k * n + j];
  return COSMO_OK;
}
