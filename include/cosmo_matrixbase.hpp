// cosmo_matrixbase.hpp -- header-only C++ adapter that lets the reference's own Davidson
// drive the GPU operator.
//
// The reference plugs matrices into its eigensolver through MyMatrixDiagonal<SomeMatrix>
// (include/kdav/MandD.cxx:7-42), which needs from SomeMatrix: a copy constructor, size(),
// ReadDiagonal(double*), TimesMatrix(double**, double**, int) and
// TimesMatrix(double*, double*, int) -- the contract SparseMatrixCSB and SparseMatrix
// fulfil (include/tav/csbmatrix.cxx:623-675, include/tav/sparse.cxx:273-318).  GpuMatrix
// offers exactly that over a cosmo_op, so
//     DavidsonDiagonalization("sys.HH.EE", cosmo::GpuMatrix(op), bloc, iters, nstate);
// runs the reference algorithm with every H.V on the MI355X.  Semantics follow
// MatB::MatrixBase::TimesMatrix (include/kdav/matrixbase.h:16-18): y is overwritten with
// H x for m vectors, x is read-only, the call is complete on return.  A failing call
// aborts with the library's message, as the reference's FatalError does.
#ifndef COSMO_MATRIXBASE_HPP
#define COSMO_MATRIXBASE_HPP

#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "cosmo_hip.h"

namespace cosmo {

class GpuMatrix {
 public:
  // shares the operator (copies are shallow, as MyMatrixDiagonal's copy of the matrix is)
  explicit GpuMatrix(cosmo_op op, int precision = COSMO_PREC_FP64, bool owns = false)
      : op_(op, [owns](cosmo_op o) {
          if (owns) cosmo_op_destroy(o);
        }),
        precision_(precision) {}

  size_t size() const { return (size_t)cosmo_op_size(op_.get()); }

  size_t ReadDiagonal(double *d) const {
    check(cosmo_op_diagonal(op_.get(), d));
    return size();
  }

  int TimesMatrix(double **y, double **x, int m) const {
    check(cosmo_op_apply(op_.get(), y, x, m, precision_));
    return 0;
  }

  // contiguous form: y and x hold m consecutive n-vectors
  int TimesMatrix(double *y, double *x, int m) const {
    const size_t n = size();
    std::vector<double *> yy(m), xx(m);
    for (int i = 0; i < m; ++i) {
      yy[i] = y + n * i;
      xx[i] = x + n * i;
    }
    return TimesMatrix(yy.data(), xx.data(), m);
  }

  cosmo_op handle() const { return op_.get(); }

 private:
  static void check(int rc) {
    if (rc) {
      fprintf(stderr, "Error: %s\n", cosmo_last_error());
      exit(1);
    }
  }
  std::shared_ptr<cosmo_op_s> op_;
  int precision_;
};

}  // namespace cosmo

#endif  // COSMO_MATRIXBASE_HPP
