// dense.hip -- exact diagonalisation of small operators on the GPU: the reference's
// n <= limit branch (include/xsm/DEXHVcsb.cxx:72-108, Eigen's SelfAdjointEigenSolver on
// the dense H).  The upper-triangle rows are scattered into a dense column-major FP64
// matrix in HBM and rocSOLVER's divide-and-conquer symmetric eigensolver (dsyevd) returns
// all eigenvalues ascending with their eigenvectors.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <vector>

#include "op.h"

namespace cosmo {

// A[c * n + r] = A[r * n + c] = v for every upper-triangle entry (r, c) of rows [0, n)
__global__ void k_dense_scatter(int64_t n, const int64_t *__restrict__ rowptr, const int64_t *__restrict__ cols,
                                const double *__restrict__ vals, double *__restrict__ A) {
  const int64_t r = blockIdx.x;
  if (r >= n) return;
  for (int64_t e = rowptr[r] + threadIdx.x; e < rowptr[r + 1]; e += blockDim.x) {
    const int64_t c = cols[e];
    A[c * n + r] = vals[e];
    A[r * n + c] = vals[e];
  }
}

}  // namespace cosmo

using namespace cosmo;

extern "C" int cosmo_op_dense_eigen(cosmo_op op, double *evals, double *evecs) {
  const int64_t n = cosmo_op_size(op);
  if (n < 0) return (int)n;
  if (!evals) return fail(COSMO_EINVAL, "cosmo_op_dense_eigen: null evals");
  if (n > 20000) return fail(COSMO_EUNSUP, "exact diagonalisation is for small operators (n <= 20000)");
  if (n == 0) return COSMO_OK;
  // upper-triangle rows on the host (the operator's own row source), then to the device
  std::vector<int64_t> cnt(n), rp(n + 1, 0);
  int rc = cosmo_op_upper_rows(op, 0, n, cnt.data(), nullptr, nullptr, nullptr);
  if (rc) return rc;
  for (int64_t r = 0; r < n; ++r) rp[r + 1] = rp[r] + cnt[r];
  const int64_t nnz = rp[n];
  std::vector<int64_t> cols(std::max<int64_t>(nnz, 1));
  std::vector<double> vals(std::max<int64_t>(nnz, 1));
  rc = cosmo_op_upper_rows(op, 0, n, cnt.data(), rp.data(), cols.data(), vals.data());
  if (rc) return rc;
  hipStream_t st = lib_stream();
  int64_t *drp = nullptr, *dcols = nullptr;
  double *dvals = nullptr, *A = nullptr, *D = nullptr, *E = nullptr;
  rocblas_int *dinfo = nullptr;
  rocblas_handle h = nullptr;
  auto release = [&]() {
    (void)hipStreamSynchronize(st);
    if (h) rocblas_destroy_handle(h);
    for (void *p : {(void *)drp, (void *)dcols, (void *)dvals, (void *)A, (void *)D, (void *)E, (void *)dinfo})
      dfree(p);
  };
  const size_t nn = (size_t)n * n;
  if (dmalloc(&drp, sizeof(int64_t) * (n + 1), st) || dmalloc(&dcols, sizeof(int64_t) * cols.size(), st) ||
      dmalloc(&dvals, sizeof(double) * vals.size(), st) || dmalloc(&A, sizeof(double) * nn, st) ||
      dmalloc(&D, sizeof(double) * n, st) || dmalloc(&E, sizeof(double) * n, st) ||
      dmalloc(&dinfo, sizeof(rocblas_int), st)) {
    release();
    return fail(COSMO_ENOMEM, "cosmo_op_dense_eigen: device allocation failed");
  }
  if (hipMemcpyAsync(drp, rp.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(dcols, cols.data(), sizeof(int64_t) * cols.size(), hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(dvals, vals.data(), sizeof(double) * vals.size(), hipMemcpyHostToDevice, st) ||
      hipMemsetAsync(A, 0, sizeof(double) * nn, st)) {
    release();
    return fail(COSMO_EHIP, "cosmo_op_dense_eigen: upload failed");
  }
  hipLaunchKernelGGL(k_dense_scatter, dim3((unsigned)n), dim3(64), 0, st, n, drp, dcols, dvals, A);
  if (hipGetLastError() != hipSuccess || rocblas_create_handle(&h) != rocblas_status_success ||
      rocblas_set_stream(h, st) != rocblas_status_success) {
    release();
    return fail(COSMO_EHIP, "cosmo_op_dense_eigen: scatter or rocBLAS handle failed");
  }
  const rocblas_status s = rocsolver_dsyevd(h, evecs ? rocblas_evect_original : rocblas_evect_none,
                                            rocblas_fill_upper, (rocblas_int)n, A, (rocblas_int)n, D, E, dinfo);
  rocblas_int info = 0;
  const bool ok = s == rocblas_status_success &&
                  hipMemcpyAsync(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost, st) == hipSuccess &&
                  hipMemcpyAsync(evals, D, sizeof(double) * n, hipMemcpyDeviceToHost, st) == hipSuccess &&
                  (!evecs || hipMemcpyAsync(evecs, A, sizeof(double) * nn, hipMemcpyDeviceToHost, st) == hipSuccess) &&
                  hipStreamSynchronize(st) == hipSuccess;
  release();
  if (!ok) return fail(COSMO_EHIP, "cosmo_op_dense_eigen: rocsolver_dsyevd failed");
  if (info != 0) return fail(COSMO_ECONV, "cosmo_op_dense_eigen: the eigensolver did not converge");
  return COSMO_OK;
}
