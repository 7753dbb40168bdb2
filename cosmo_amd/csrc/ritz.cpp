// ritz.cpp -- the lowest eigenpairs of the small dense symmetric Rayleigh-Ritz matrix of the
// Davidson iteration (DiagonalizeS, include/kdav/fsdavidson.cxx:21-98, which calls a
// Householder + QL solver for all pairs).  Only the `want` lowest pairs are used by the
// iteration, so this computes just those:
//   1. Householder reduction A = Q T Q^T to tridiagonal T, reflectors kept (Q not formed);
//   2. all eigenvalues of T by implicit symmetric QR steps with the Wilkinson shift
//      (Golub & Van Loan, Matrix Computations, sec. 8.3), no vectors: O(k^2);
//   3. their eigenvectors of T by inverse iteration on the pivoted tridiagonal LU of
//      T - lambda I, re-orthogonalised against the previous vectors of the same cluster;
//   4. back-transformation by the reflectors.
// Cost O(k^3) for step 1 and O(k^2 * want) for the rest, against a several times larger
// O(k^3) when all eigenvectors are accumulated.
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace cosmo {

namespace {

struct Tridiag {
  int n = 0;
  std::vector<double> d, e;      // diagonal, sub-diagonal (e[i] couples i and i+1)
  std::vector<double> v, tau;    // reflector j: v at v[j*n + (j+1) ..], scale tau[j]
};

// A (n x n, row-major, both triangles) is overwritten
void reduce(int n, std::vector<double> &A, Tridiag &T) {
  T.n = n;
  T.d.assign(n, 0.0);
  T.e.assign(n > 1 ? n - 1 : 0, 0.0);
  T.v.assign((size_t)n * n, 0.0);
  T.tau.assign(n, 0.0);
  std::vector<double> p(n), w(n);
  for (int j = 0; j + 2 < n; ++j) {
    const int m = n - j - 1;  // length of the column below the diagonal
    double *vj = &T.v[(size_t)j * n + j + 1];
    double alpha = 0.0;
    for (int i = 0; i < m; ++i) {
      vj[i] = A[(size_t)(j + 1 + i) * n + j];
      alpha += vj[i] * vj[i];
    }
    alpha = std::sqrt(alpha);
    T.d[j] = A[(size_t)j * n + j];
    if (alpha == 0.0) {  // column already reduced
      T.e[j] = 0.0;
      continue;
    }
    const double beta = vj[0] >= 0.0 ? -alpha : alpha;  // avoids cancellation in v0
    vj[0] -= beta;
    double vv = 0.0;
    for (int i = 0; i < m; ++i) vv += vj[i] * vj[i];
    const double tau = 2.0 / vv;
    T.tau[j] = tau;
    T.e[j] = beta;
    // A22 <- H A22 H with H = I - tau v v^T: p = tau A22 v, w = p - (tau/2)(p.v) v,
    // A22 -= v w^T + w v^T
    double pv = 0.0;
    for (int r = 0; r < m; ++r) {
      const double *ar = &A[(size_t)(j + 1 + r) * n + j + 1];
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;  // four chains: the sum is latency-bound
      int c = 0;
      for (; c + 4 <= m; c += 4) {
        s0 += ar[c] * vj[c];
        s1 += ar[c + 1] * vj[c + 1];
        s2 += ar[c + 2] * vj[c + 2];
        s3 += ar[c + 3] * vj[c + 3];
      }
      for (; c < m; ++c) s0 += ar[c] * vj[c];
      p[r] = tau * ((s0 + s1) + (s2 + s3));
      pv += p[r] * vj[r];
    }
    const double K = 0.5 * tau * pv;
    for (int r = 0; r < m; ++r) w[r] = p[r] - K * vj[r];
    for (int r = 0; r < m; ++r) {
      double *ar = &A[(size_t)(j + 1 + r) * n + j + 1];
      for (int c = 0; c < m; ++c) ar[c] -= vj[r] * w[c] + w[r] * vj[c];
    }
  }
  if (n >= 2) {
    T.d[n - 2] = A[(size_t)(n - 2) * n + n - 2];
    T.e[n - 2] = A[(size_t)(n - 1) * n + n - 2];
  }
  if (n >= 1) T.d[n - 1] = A[(size_t)(n - 1) * n + n - 1];
}

// All eigenvalues of T (unsorted) by implicit QR steps on unreduced diagonal blocks: each
// step rotates in planes (i, i+1) from the top of the block down, the first rotation set
// by the shifted leading column, the next ones chasing the bulge off the bottom; a
// coupling below eps * (|d_i| + |d_i+1|) is dropped.  Returns false if the cap on steps
// is reached.
bool tridiag_eigenvalues(const Tridiag &T, std::vector<double> &ev) {
  const int n = T.n;
  std::vector<double> a(T.d), b(T.e);
  const double eps = std::numeric_limits<double>::epsilon();
  int hi = n - 1, steps = 0;
  while (hi > 0) {
    if (std::fabs(b[hi - 1]) <= eps * (std::fabs(a[hi - 1]) + std::fabs(a[hi]))) {
      b[hi - 1] = 0.0;
      --hi;
      continue;
    }
    int lo = hi - 1;
    while (lo > 0 && std::fabs(b[lo - 1]) > eps * (std::fabs(a[lo - 1]) + std::fabs(a[lo]))) --lo;
    if (lo > 0) b[lo - 1] = 0.0;
    if (++steps > 40 * n) return false;
    // Wilkinson shift: the eigenvalue of the trailing 2 x 2 closer to a[hi]
    const double dd = 0.5 * (a[hi - 1] - a[hi]), bh = b[hi - 1];
    const double den = dd + (dd >= 0 ? 1.0 : -1.0) * std::hypot(dd, bh);
    const double mu = a[hi] - (den != 0.0 ? bh * bh / den : 0.0);
    double x = a[lo] - mu, z = b[lo], bulge = 0.0;
    for (int k = lo; k < hi; ++k) {
      if (k > lo) {
        x = b[k - 1];
        z = bulge;
      }
      // sqrt(x^2 + z^2) directly: std::hypot's overflow guard costs more than the whole
      // rotation, and |x|, |z| are far from the overflow range here
      const double r = (std::fabs(x) < 1e150 && std::fabs(z) < 1e150) ? std::sqrt(x * x + z * z) : std::hypot(x, z);
      const double c = r != 0.0 ? x / r : 1.0, s = r != 0.0 ? z / r : 0.0;
      if (k > lo) b[k - 1] = r;
      const double app = a[k], aqq = a[k + 1], apq = b[k];
      a[k] = c * c * app + 2.0 * c * s * apq + s * s * aqq;
      a[k + 1] = s * s * app - 2.0 * c * s * apq + c * c * aqq;
      b[k] = c * s * (aqq - app) + (c * c - s * s) * apq;
      if (k + 1 < hi) {
        bulge = s * b[k + 1];
        b[k + 1] *= c;
      }
    }
  }
  ev = a;
  return true;
}

// y <- (T - s I)^-1 y by Gaussian elimination with partial pivoting on the tridiagonal;
// pivots below `tiny` are raised to it (s is an eigenvalue to working accuracy)
struct SolveWork {
  std::vector<double> u0, u1, u2, l;  // rows of U (diagonal, two super-diagonals), multipliers
  std::vector<char> swp;              // row i swapped with row i+1
};

void tri_solve(const Tridiag &T, double s, double tiny, std::vector<double> &y, SolveWork &W) {
  const int n = T.n;
  W.u0.assign(n, 0.0);
  W.u1.assign(n, 0.0);
  W.u2.assign(n, 0.0);
  W.l.assign(n, 0.0);
  W.swp.assign(n, 0);
  double *u0 = W.u0.data(), *u1 = W.u1.data(), *u2 = W.u2.data(), *l = W.l.data();
  char *swp = W.swp.data();
  double a = T.d[0] - s, b = n > 1 ? T.e[0] : 0.0;  // current row's diagonal and super
  for (int i = 0; i + 1 < n; ++i) {
    const double c = T.e[i];                                  // sub-diagonal below row i
    const double dn = T.d[i + 1] - s, en = i + 2 < n ? T.e[i + 1] : 0.0;  // next row
    if (std::fabs(a) >= std::fabs(c)) {
      if (a == 0.0) a = tiny;
      const double m = c / a;
      l[i] = m;
      u0[i] = a;
      u1[i] = b;
      u2[i] = 0.0;
      a = dn - m * b;
      b = en;
    } else {  // swap rows i and i+1
      swp[i] = 1;
      const double m = a / c;
      l[i] = m;
      u0[i] = c;
      u1[i] = dn;
      u2[i] = en;
      a = b - m * dn;
      b = -m * en;
    }
  }
  if (std::fabs(a) < tiny) a = a < 0 ? -tiny : tiny;
  u0[n - 1] = a;
  // forward: apply the row swaps and multipliers to y
  for (int i = 0; i + 1 < n; ++i) {
    if (swp[i]) std::swap(y[i], y[i + 1]);
    y[i + 1] -= l[i] * y[i];
  }
  // backward
  for (int i = n - 1; i >= 0; --i) {
    double t = y[i];
    if (i + 1 < n) t -= u1[i] * y[i + 1];
    if (i + 2 < n) t -= u2[i] * y[i + 2];
    double p = u0[i];
    if (std::fabs(p) < tiny) p = p < 0 ? -tiny : tiny;
    y[i] = t / p;
  }
}

// Euclidean norm, scaled so that inverse-iteration growth (up to 1/tiny) cannot overflow
double norm2(const std::vector<double> &y) {
  double m = 0.0;
  for (double x : y) m = std::max(m, std::fabs(x));
  if (!(m > 0.0) || !std::isfinite(m)) return m;
  double s = 0.0;
  for (double x : y) s += (x / m) * (x / m);
  return m * std::sqrt(s);
}

}  // namespace

// Lowest `want` eigenpairs of the symmetric k x k matrix A (row-major): w[i] ascending,
// V[i*k + r] = component r of eigenvector i.  Returns false if an eigenvector could not be
// made orthogonal to its cluster (not seen in practice; the caller then falls back).
bool ritz_lowest(int k, const std::vector<double> &A0, int want, std::vector<double> &w, std::vector<double> &V) {
  want = std::min(want, k);
  w.assign(want, 0.0);
  V.assign((size_t)want * k, 0.0);
  if (k <= 0 || want <= 0) return true;
  std::vector<double> A(A0);
  Tridiag T;
  reduce(k, A, T);
  // scale of T (largest absolute row sum)
  double tn = 0.0;
  for (int i = 0; i < k; ++i) {
    const double r = (i > 0 ? std::fabs(T.e[i - 1]) : 0.0) + (i + 1 < k ? std::fabs(T.e[i]) : 0.0);
    tn = std::max(tn, std::fabs(T.d[i]) + r);
  }
  const double eps = std::numeric_limits<double>::epsilon();
  const double tiny = std::max(tn * eps, 1e-280);  // pivot floor; growth per solve stays finite
  std::vector<double> all;
  if (!tridiag_eigenvalues(T, all)) return false;
  std::sort(all.begin(), all.end());
  for (int i = 0; i < want; ++i) w[i] = all[i];
  // eigenvectors of T, clusters (gap <= 1e-3 |T|) re-orthogonalised
  std::vector<std::vector<double>> Y(want, std::vector<double>(k));
  SolveWork work;
  int cstart = 0;
  for (int i = 0; i < want; ++i) {
    if (i > 0 && w[i] - w[i - 1] > 1e-3 * tn) cstart = i;
    std::vector<double> &y = Y[i];
    for (int r = 0; r < k; ++r) y[r] = 1.0 + 0.5 * std::sin(1.0 + 0.7 * r + 1.3 * i);  // fixed start
    double nrm = 0.0;
    for (int it = 0; it < 4; ++it) {
      tri_solve(T, w[i], tiny, y, work);
      for (int c = cstart; c < i; ++c) {
        double dot = 0.0;
        for (int r = 0; r < k; ++r) dot += Y[c][r] * y[r];
        for (int r = 0; r < k; ++r) y[r] -= dot * Y[c][r];
      }
      nrm = norm2(y);
      if (!(nrm > 0.0) || !std::isfinite(nrm)) return false;
      for (int r = 0; r < k; ++r) y[r] /= nrm;
    }
  }
  // back-transformation: x = H_0 H_1 ... H_{k-3} y
  for (int i = 0; i < want; ++i) {
    std::vector<double> &x = Y[i];
    for (int j = k - 3; j >= 0; --j) {
      if (T.tau[j] == 0.0) continue;
      const double *vj = &T.v[(size_t)j * k + j + 1];
      double dot = 0.0;
      for (int r = 0; r < k - j - 1; ++r) dot += vj[r] * x[j + 1 + r];
      dot *= T.tau[j];
      for (int r = 0; r < k - j - 1; ++r) x[j + 1 + r] -= dot * vj[r];
    }
    std::copy(x.begin(), x.end(), V.begin() + (size_t)i * k);
  }
  return true;
}

}  // namespace cosmo
