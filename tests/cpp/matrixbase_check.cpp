// Exercises cosmo::GpuMatrix (include/cosmo_matrixbase.hpp) through the members that
// MyMatrixDiagonal<SomeMatrix> (include/kdav/MandD.cxx:7-42) uses: copy construction,
// size(), ReadDiagonal(double*), TimesMatrix(double**, double**, int) and
// TimesMatrix(double*, double*, int) -- the in-process boundary a CoSMo build would bind.
//
// Usage: matrixbase_check <sys.xml> <sysname> <m> <out.bin>
// Writes [u64 n][u64 m][n diag][m*n X][m*n Y pointer form][m*n Y contiguous form, from a
// copy of the matrix]; X[v][i] = sin(0.37 i + 1.3 v).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "cosmo_hip.h"
#include "cosmo_matrixbase.hpp"

// the operations MyMatrixDiagonal performs on its matrix, written against the contract only
template <class M>
static void drive(const M &A, int m, std::vector<double> &diag, std::vector<double> &X, std::vector<double> &Y1,
                  std::vector<double> &Y2) {
  M B(A);  // MyMatrixDiagonal holds a copy
  const size_t n = B.size();
  diag.assign(n, 0.0);
  B.ReadDiagonal(diag.data());
  X.assign(n * m, 0.0);
  for (int v = 0; v < m; ++v)
    for (size_t i = 0; i < n; ++i) X[v * n + i] = std::sin(0.37 * double(i) + 1.3 * double(v));
  Y1.assign(n * m, -1.0);
  Y2.assign(n * m, -1.0);
  std::vector<double *> y(m), x(m);
  for (int v = 0; v < m; ++v) {
    y[v] = Y1.data() + v * n;
    x[v] = X.data() + v * n;
  }
  A.TimesMatrix(y.data(), x.data(), m);
  B.TimesMatrix(Y2.data(), X.data(), m);
}

int main(int argc, char **argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: matrixbase_check <sys.xml> <sysname> <m> <out.bin>\n");
    return 2;
  }
  cosmo_system s;
  cosmo_op op;
  if (cosmo_system_from_xml(argv[1], argv[2], &s) || cosmo_op_create(s, COSMO_OP_STORED, &op)) {
    fprintf(stderr, "setup: %s\n", cosmo_last_error());
    return 1;
  }
  const int m = atoi(argv[3]);
  std::vector<double> diag, X, Y1, Y2;
  {
    cosmo::GpuMatrix A(op, COSMO_PREC_FP64, /*owns=*/true);
    drive(A, m, diag, X, Y1, Y2);
  }
  cosmo_system_destroy(s);
  FILE *f = fopen(argv[4], "wb");
  if (!f) return 1;
  const uint64_t n = diag.size(), mm = m;
  fwrite(&n, 8, 1, f);
  fwrite(&mm, 8, 1, f);
  fwrite(diag.data(), 8, n, f);
  fwrite(X.data(), 8, n * m, f);
  fwrite(Y1.data(), 8, n * m, f);
  fwrite(Y2.data(), 8, n * m, f);
  return fclose(f) ? 1 : 0;
}
