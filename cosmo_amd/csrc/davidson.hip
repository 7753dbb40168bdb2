// davidson.hip -- GPU-resident block Davidson for the lowest eigenpairs of a cosmo_op.
//
// Control logic restates DVDSON as used by the reference (include/kdav/fsdavidson.cxx
// dvdrvr2 :615-775, Test_Convergence :212-281, Add_Vectors_To_Basis :426-516,
// Collapse_* :99-121/:517-539, setup_2d :540-611; defaults from
// include/kdav/cdavidson.cpp:86-110 and include/kdav/davidson.cxx:167-223).  The n-long
// work (H.X, Gram blocks, basis rotations, residuals, preconditioning) stays on the GPU;
// the (lim x lim) Rayleigh-Ritz problem is solved in FP64 on the host.
//
// Orthogonalisation is blocked: the new vectors are projected against the existing
// basis with two classical Gram-Schmidt sweeps (B^T V and V - B W as two tall-skinny
// passes) and then orthonormalised among themselves by a pivot-dropping Cholesky QR,
// repeated once in the FP64 phase (not while H.X runs in FP32).  This spans the same space as the reference's per-vector repeated CGS
// (Orthogonalize_Basis :313-390) in a few passes over B instead of one per new vector.
//
// Mixed precision (COSMO_PREC_FP32): H.X runs on the FP32 values with FP32 vectors until
// every wanted residual drops under mixed_switch * max(1, |lambda|); the basis is then
// collapsed to the Ritz vectors, H.X is recomputed in FP64 and the iteration continues
// in FP64 to the reference thresholds (the FP64 refinement).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "op.h"

namespace cosmo {
bool dense_symmetric_eigen(int n, const std::vector<double> &A, std::vector<double> &w, std::vector<double> &V);

#define GRAM_ROWS 32
#define GRAM_MAXW 160  // p + q limit of one Gram launch

// partial[g][a*q+b] = sum over the rows of block g of A[:,a] * Bm[:,b]
__global__ __launch_bounds__(256) void k_gram(int64_t n, const double *__restrict__ A, int64_t lda, int p,
                                              const double *__restrict__ Bm, int64_t ldb, int q,
                                              double *__restrict__ partial) {
  __shared__ double sa[GRAM_ROWS * GRAM_MAXW];
  double acc[4] = {0, 0, 0, 0};
  const int pq = p * q;
  const int w = p + q;
  for (int64_t base = (int64_t)blockIdx.x * GRAM_ROWS; base < n; base += (int64_t)gridDim.x * GRAM_ROWS) {
    const int rows = (int)min<int64_t>(GRAM_ROWS, n - base);
    __syncthreads();
    for (int t = threadIdx.x; t < GRAM_ROWS * w; t += blockDim.x) {
      int col = t / GRAM_ROWS, r = t % GRAM_ROWS;
      double v = 0.0;
      if (r < rows) v = (col < p) ? A[(int64_t)col * lda + base + r] : Bm[(int64_t)(col - p) * ldb + base + r];
      sa[col * GRAM_ROWS + r] = v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      int o = threadIdx.x + k * 256;
      if (o < pq) {
        int a = o / q, b = o % q;
        const double *x = sa + a * GRAM_ROWS, *y = sa + (p + b) * GRAM_ROWS;
        double s = 0.0;
#pragma unroll 8
        for (int r = 0; r < GRAM_ROWS; ++r) s += x[r] * y[r];
        acc[k] += s;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int o = threadIdx.x + k * 256;
    if (o < pq) partial[(int64_t)blockIdx.x * pq + o] = acc[k];
  }
}

// Sum of the G Gram partials in two fixed-order stages (deterministic): stage one sums
// GRAM_SEG contiguous segments of blocks in parallel, stage two the segment sums.
#define GRAM_SEG 32
__global__ void k_reduce_partial(int G, int pq, const double *__restrict__ partial, double *__restrict__ seg) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= pq) return;
  const int per = (G + GRAM_SEG - 1) / GRAM_SEG;
  const int g0 = blockIdx.y * per, g1 = min(G, g0 + per);
  double s = 0.0;
  for (int g = g0; g < g1; ++g) s += partial[(int64_t)g * pq + o];
  seg[(int64_t)blockIdx.y * pq + o] = s;
}
__global__ void k_reduce_segments(int pq, const double *__restrict__ seg, double *__restrict__ out) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= pq) return;
  double s = 0.0;
  for (int k = 0; k < GRAM_SEG; ++k) s += seg[(int64_t)k * pq + o];
  out[o] = s;
}

// Y[:, j] = sum_k A1[:,k] C1[k*q+j] (+ A2[:,k] C2[k*q+j]), j < q <= 16, per row
template <int Q>
__global__ __launch_bounds__(256) void k_gemm(int64_t n, const double *__restrict__ A1, const double *__restrict__ A2,
                                              int64_t lda, int p, const double *__restrict__ C1,
                                              const double *__restrict__ C2, int q, double *__restrict__ Y,
                                              int64_t ldy) {
  __shared__ double c1[128 * Q], c2[128 * Q];
  for (int t = threadIdx.x; t < p * q; t += blockDim.x) {
    int k = t / q, j = t % q;
    c1[k * Q + j] = C1[t];
    c2[k * Q + j] = C2 ? C2[t] : 0.0;
  }
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) acc[j] = 0.0;
  for (int k = 0; k < p; ++k) {
    double a = A1[(int64_t)k * lda + i];
#pragma unroll
    for (int j = 0; j < Q; ++j) acc[j] += a * c1[k * Q + j];
    if (A2) {
      double b = A2[(int64_t)k * lda + i];
#pragma unroll
      for (int j = 0; j < Q; ++j) acc[j] += b * c2[k * Q + j];
    }
  }
#pragma unroll
  for (int j = 0; j < Q; ++j)
    if (j < q) Y[(int64_t)j * ldy + i] = acc[j];
}

// V[:, j] -= sum_k B[:,k] W[k*q+j]
template <int Q>
__global__ __launch_bounds__(256) void k_project_out(int64_t n, const double *__restrict__ B, int64_t ldb, int p,
                                                     const double *__restrict__ W, int q, double *__restrict__ V,
                                                     int64_t ldv) {
  __shared__ double w[128 * Q];
  for (int t = threadIdx.x; t < p * q; t += blockDim.x) w[(t / q) * Q + t % q] = W[t];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) acc[j] = 0.0;
  for (int k = 0; k < p; ++k) {
    double b = B[(int64_t)k * ldb + i];
#pragma unroll
    for (int j = 0; j < Q; ++j) acc[j] += b * w[k * Q + j];
  }
#pragma unroll
  for (int j = 0; j < Q; ++j)
    if (j < q) V[(int64_t)j * ldv + i] -= acc[j];
}

// diagonal preconditioner: r /= (lambda - D_ii), |.| floored at 1e-13 (fsdavidson.cxx:498-510)
__global__ void k_precond(int64_t n, int q, double *__restrict__ R, int64_t ldr, const double *__restrict__ lam,
                          const double *__restrict__ D) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * q) return;
  int64_t i = t % n;
  int j = (int)(t / n);
  double dg = lam[j] - D[i];
  double *r = R + (int64_t)j * ldr + i;
  if (fabs(dg) > 1.0e-13) *r = *r / dg;
  else *r = *r / 1.0e-13;
}

__global__ void k_scale_cols(int64_t n, int q, double *__restrict__ V, int64_t ldv, const double *__restrict__ sc) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * q) return;
  V[(t / n) * ldv + t % n] *= sc[t / n];
}

// column-major B[:, c0 .. c0+m) -> interleaved X (n x ldx), as double or float
// (columns m .. ldx-1 of X are zero padding)
template <typename T>
__global__ void k_cols_to_rows(int64_t n, int m, const double *__restrict__ B, int64_t ldb, T *__restrict__ X, int ldx) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * ldx) return;
  int64_t i = t / ldx;
  int j = (int)(t % ldx);
  X[i * ldx + j] = j < m ? (T)B[(int64_t)j * ldb + i] : (T)0;
}
template <typename T>
__global__ void k_rows_to_cols(int64_t n, int m, const T *__restrict__ Y, int ldy, double *__restrict__ B, int64_t ldb) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * m) return;
  int j = (int)(t / n);
  int64_t i = t % n;
  B[(int64_t)j * ldb + i] = (double)Y[i * ldy + j];
}

// ------------------------------------------------------------------------------------
struct Dav {
  Op *op = nullptr;
  hipStream_t st = nullptr;
  int64_t n = 0;
  int lim = 0, nume = 0, mblock = 0;
  double *B = nullptr, *AB = nullptr, *R = nullptr, *T = nullptr, *Xd = nullptr, *Yd = nullptr;
  float *Xf = nullptr, *Yf = nullptr;
  double *partial = nullptr, *segs = nullptr, *small = nullptr, *small2 = nullptr, *lam = nullptr;
  int G = 0;
  double ms_apply = 0, ms_ortho = 0, ms_rr = 0;
  int matvecs = 0, matvecs64 = 0;
  ~Dav() {
    dfree(B);
    dfree(AB);
    dfree(R);
    dfree(T);
    dfree(Xd);
    dfree(Yd);
    dfree(Xf);
    dfree(Yf);
    dfree(partial);
    dfree(segs);
    dfree(small);
    dfree(small2);
    dfree(lam);
  }
  double *col(double *base, int j) { return base + (int64_t)j * n; }
};

static int dav_alloc(Dav &D) {
  const int64_t n = D.n;
  HIPCHK(dmalloc(&D.B, sizeof(double) * n * D.lim));
  HIPCHK(dmalloc(&D.AB, sizeof(double) * n * D.lim));
  HIPCHK(dmalloc(&D.R, sizeof(double) * n * std::max(D.nume, D.mblock)));
  HIPCHK(dmalloc(&D.T, sizeof(double) * n * std::max(D.lim, 16)));
  int ldmax = 4;
  while (ldmax < D.mblock) ldmax <<= 1;
  HIPCHK(dmalloc(&D.Xd, sizeof(double) * n * ldmax));
  HIPCHK(dmalloc(&D.Yd, sizeof(double) * n * ldmax));
  HIPCHK(dmalloc(&D.Xf, sizeof(float) * n * ldmax));
  HIPCHK(dmalloc(&D.Yf, sizeof(float) * n * ldmax));
  D.G = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n + GRAM_ROWS - 1) / GRAM_ROWS));
  HIPCHK(dmalloc(&D.partial, sizeof(double) * (size_t)D.G * 1024));
  HIPCHK(dmalloc(&D.segs, sizeof(double) * (size_t)GRAM_SEG * 1024));
  HIPCHK(dmalloc(&D.small, sizeof(double) * 8192));
  HIPCHK(dmalloc(&D.small2, sizeof(double) * 8192));
  HIPCHK(dmalloc(&D.lam, sizeof(double) * 256));
  HIPCHK(hipMemsetAsync(D.B, 0, sizeof(double) * n * D.lim, D.st));
  HIPCHK(hipMemsetAsync(D.AB, 0, sizeof(double) * n * D.lim, D.st));
  return COSMO_OK;
}

// out (p x q, row-major a*q+b) = A[:, 0..p)^T Bm[:, 0..q), host copy
static int gram(Dav &D, const double *A, int p, const double *Bm, int q, std::vector<double> &out) {
  out.assign((size_t)p * q, 0.0);
  if (!p || !q) return COSMO_OK;
  // split p so that p_chunk*q <= 1024 and p_chunk+q <= GRAM_MAXW
  int pc = std::max(1, std::min({p, 1024 / q, GRAM_MAXW - q}));
  std::vector<double> tmp;
  for (int a0 = 0; a0 < p; a0 += pc) {
    int pp = std::min(pc, p - a0);
    hipLaunchKernelGGL(k_gram, dim3(D.G), dim3(256), 0, D.st, D.n, A + (int64_t)a0 * D.n, D.n, pp, Bm, D.n, q,
                       D.partial);
    hipLaunchKernelGGL(k_reduce_partial, dim3((pp * q + 255) / 256, GRAM_SEG), dim3(256), 0, D.st, D.G, pp * q,
                       D.partial, D.segs);
    hipLaunchKernelGGL(k_reduce_segments, dim3((pp * q + 255) / 256), dim3(256), 0, D.st, pp * q, D.segs, D.small);
    HIPCHK(hipGetLastError());
    tmp.resize((size_t)pp * q);
    HIPCHK(hipMemcpyAsync(tmp.data(), D.small, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost, D.st));
    HIPCHK(hipStreamSynchronize(D.st));
    std::copy(tmp.begin(), tmp.end(), out.begin() + (size_t)a0 * q);
  }
  return COSMO_OK;
}

// Y[:, 0..q) = A1 C1 (+ A2 C2); C row-major p x q on host; q <= 16 per launch
static int gemm(Dav &D, const double *A1, const double *A2, int p, const std::vector<double> &C1,
                const std::vector<double> *C2, int q, double *Y) {
  if (p > 128) return fail(COSMO_EUNSUP, "basis wider than 128 vectors");
  for (int j0 = 0; j0 < q; j0 += 16) {
    int qq = std::min(16, q - j0);
    std::vector<double> c1((size_t)p * qq), c2;
    for (int k = 0; k < p; ++k)
      for (int j = 0; j < qq; ++j) c1[k * qq + j] = C1[(size_t)k * q + j0 + j];
    if (C2) {
      c2.resize((size_t)p * qq);
      for (int k = 0; k < p; ++k)
        for (int j = 0; j < qq; ++j) c2[k * qq + j] = (*C2)[(size_t)k * q + j0 + j];
    }
    HIPCHK(hipMemcpyAsync(D.small, c1.data(), sizeof(double) * c1.size(), hipMemcpyHostToDevice, D.st));
    if (C2) HIPCHK(hipMemcpyAsync(D.small2, c2.data(), sizeof(double) * c2.size(), hipMemcpyHostToDevice, D.st));
    dim3 g((unsigned)((D.n + 255) / 256)), b(256);
    double *Yj = Y + (int64_t)j0 * D.n;
    const double *s2 = C2 ? D.small2 : nullptr;
    if (qq <= 4) hipLaunchKernelGGL(k_gemm<4>, g, b, 0, D.st, D.n, A1, C2 ? A2 : nullptr, D.n, p, D.small, s2, qq, Yj, D.n);
    else if (qq <= 8) hipLaunchKernelGGL(k_gemm<8>, g, b, 0, D.st, D.n, A1, C2 ? A2 : nullptr, D.n, p, D.small, s2, qq, Yj, D.n);
    else hipLaunchKernelGGL(k_gemm<16>, g, b, 0, D.st, D.n, A1, C2 ? A2 : nullptr, D.n, p, D.small, s2, qq, Yj, D.n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(D.st));  // small buffers are reused by the next chunk
  }
  return COSMO_OK;
}

static int project_out(Dav &D, const double *Bm, int p, const std::vector<double> &W, int q, double *V) {
  if (!p || !q) return COSMO_OK;
  for (int j0 = 0; j0 < q; j0 += 16) {
    int qq = std::min(16, q - j0);
    for (int k0 = 0; k0 < p; k0 += 128) {
      int pp = std::min(128, p - k0);
      std::vector<double> w((size_t)pp * qq);
      for (int k = 0; k < pp; ++k)
        for (int j = 0; j < qq; ++j) w[k * qq + j] = W[(size_t)(k0 + k) * q + j0 + j];
      HIPCHK(hipMemcpyAsync(D.small, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice, D.st));
      dim3 g((unsigned)((D.n + 255) / 256)), b(256);
      const double *Bk = Bm + (int64_t)k0 * D.n;
      double *Vj = V + (int64_t)j0 * D.n;
      if (qq <= 4) hipLaunchKernelGGL(k_project_out<4>, g, b, 0, D.st, D.n, Bk, D.n, pp, D.small, qq, Vj, D.n);
      else if (qq <= 8) hipLaunchKernelGGL(k_project_out<8>, g, b, 0, D.st, D.n, Bk, D.n, pp, D.small, qq, Vj, D.n);
      else hipLaunchKernelGGL(k_project_out<16>, g, b, 0, D.st, D.n, Bk, D.n, pp, D.small, qq, Vj, D.n);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(D.st));
    }
  }
  return COSMO_OK;
}

// AB[:, c0 .. c0+m) = H B[:, c0 .. c0+m)
static int apply_block(Dav &D, int c0, int m, int prec) {
  auto t0 = std::chrono::steady_clock::now();
  for (int j0 = 0; j0 < m; j0 += D.mblock) {
    int mm = std::min(D.mblock, m - j0);
    // leading dimension padded to a power of two of at least 16 bytes (vectorised H.X)
    int ld = (prec == COSMO_PREC_FP32) ? 4 : 2;
    while (ld < mm) ld <<= 1;
    int64_t tot = D.n * mm, totp = D.n * ld;
    dim3 g((unsigned)((tot + 255) / 256)), gp((unsigned)((totp + 255) / 256)), b(256);
    int rc;
    if (prec == COSMO_PREC_FP32) {
      hipLaunchKernelGGL(k_cols_to_rows<float>, gp, b, 0, D.st, D.n, mm, D.col(D.B, c0 + j0), D.n, D.Xf, ld);
      rc = op_apply_dev(*D.op, D.Xf, D.Yf, mm, ld, ld, COSMO_PREC_FP32, D.st);
      if (rc) return rc;
      hipLaunchKernelGGL(k_rows_to_cols<float>, g, b, 0, D.st, D.n, mm, D.Yf, ld, D.col(D.AB, c0 + j0), D.n);
    } else {
      hipLaunchKernelGGL(k_cols_to_rows<double>, gp, b, 0, D.st, D.n, mm, D.col(D.B, c0 + j0), D.n, D.Xd, ld);
      rc = op_apply_dev(*D.op, D.Xd, D.Yd, mm, ld, ld, COSMO_PREC_FP64, D.st);
      if (rc) return rc;
      hipLaunchKernelGGL(k_rows_to_cols<double>, g, b, 0, D.st, D.n, mm, D.Yd, ld, D.col(D.AB, c0 + j0), D.n);
      D.matvecs64 += mm;
    }
    HIPCHK(hipGetLastError());
    D.matvecs += mm;
  }
  HIPCHK(hipStreamSynchronize(D.st));
  D.ms_apply += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return COSMO_OK;
}

// symmetric eigen-decomposition (cyclic Jacobi), ascending, vectors as rows of V
static void sym_eig(int k, const std::vector<double> &A0, std::vector<double> &w, std::vector<double> &V) {
  std::vector<double> A(A0);
  V.assign((size_t)k * k, 0.0);
  for (int i = 0; i < k; ++i) V[i * k + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) {
        tot += A[i * k + j] * A[i * k + j];
        if (i != j) off += A[i * k + j] * A[i * k + j];
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < k; ++p)
      for (int q = p + 1; q < k; ++q) {
        double apq = A[p * k + q];
        if (std::fabs(apq) < 1e-300) continue;
        double theta = (A[q * k + q] - A[p * k + p]) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int r = 0; r < k; ++r) {
          double arp = A[r * k + p], arq = A[r * k + q];
          A[r * k + p] = c * arp - s * arq;
          A[r * k + q] = s * arp + c * arq;
        }
        for (int r = 0; r < k; ++r) {
          double apr = A[p * k + r], aqr = A[q * k + r];
          A[p * k + r] = c * apr - s * aqr;
          A[q * k + r] = s * apr + c * aqr;
        }
        for (int r = 0; r < k; ++r) {  // V holds eigenvectors as columns during the sweep
          double vrp = V[r * k + p], vrq = V[r * k + q];
          V[r * k + p] = c * vrp - s * vrq;
          V[r * k + q] = s * vrp + c * vrq;
        }
      }
  }
  std::vector<int> idx(k);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return A[a * k + a] < A[b * k + b]; });
  w.resize(k);
  std::vector<double> Vr((size_t)k * k);
  for (int e = 0; e < k; ++e) {
    w[e] = A[idx[e] * k + idx[e]];
    for (int r = 0; r < k; ++r) Vr[(size_t)e * k + r] = V[r * k + idx[e]];
  }
  V.swap(Vr);
}

// Orthonormalise V (q columns at Vp) against B[:, 0..p) and among themselves, in
// `rounds` rounds of Gram-Schmidt passes plus Cholesky QR (two for FP64 accuracy; one
// while H.V itself is FP32, whose vectors are only good to single precision).
// Returns the number of kept vectors (compacted to the front); restart=true when the
// overlap with the basis does not decrease (the reference's restart condition).
static int orthonormalise(Dav &D, int p, double *Vp, int q, double ortho, bool &restart, int &kept,
                          int rounds = 2) {
  auto t0 = std::chrono::steady_clock::now();
  restart = false;
  kept = q;
  std::vector<double> W;
  if (q > 0) {
    // unit columns first: the absolute thresholds below then mean what they say even when
    // the preconditioner has blown a vector up by up to 1e13
    std::vector<double> G0;
    int rc0 = gram(D, Vp, q, Vp, q, G0);
    if (rc0) return rc0;
    std::vector<double> sc(q);
    for (int j = 0; j < q; ++j) {
      double nr = std::sqrt(std::max(G0[(size_t)j * q + j], 0.0));
      sc[j] = nr > 0 ? 1.0 / nr : 0.0;
    }
    HIPCHK(hipMemcpyAsync(D.small, sc.data(), sizeof(double) * q, hipMemcpyHostToDevice, D.st));
    hipLaunchKernelGGL(k_scale_cols, dim3((unsigned)((D.n * q + 255) / 256)), dim3(256), 0, D.st, D.n, q, Vp, D.n,
                       D.small);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(D.st));
  }
  for (int round = 0; round < rounds && kept > 0; ++round) {
    double dprev = 1e7;
    for (int pass = 0; pass < 4; ++pass) {
      if (!p) break;
      int rc = gram(D, D.B, p, Vp, kept, W);
      if (rc) return rc;
      double dcur = 0.0;
      for (double x : W) dcur = std::max(dcur, std::fabs(x));
      rc = project_out(D, D.B, p, W, kept, Vp);
      if (rc) return rc;
      if (dcur < ortho) break;
      if (dcur > dprev) {
        restart = true;
        break;
      }
      dprev = dcur;
    }
    if (restart) break;
    // Cholesky QR in order, dropping vectors whose remaining norm is < 1e-14
    std::vector<double> G;
    int rc = gram(D, Vp, kept, Vp, kept, G);
    if (rc) return rc;
    const int k = kept;
    std::vector<double> Rm((size_t)k * k, 0.0);
    std::vector<int> keep;
    for (int j = 0; j < k; ++j) {
      // r_ij = (G_ij - sum_{l<i} r_li r_lj) / r_ii over kept i
      std::vector<double> rj(k, 0.0);
      double d = G[j * k + j];
      for (int ii : keep) {
        double s = G[ii * k + j];
        for (int l : keep) {
          if (l == ii) break;
          s -= Rm[l * k + ii] * Rm[l * k + j];
        }
        s /= Rm[ii * k + ii];
        Rm[ii * k + j] = s;
        d -= s * s;
      }
      double nrm = d > 0 ? std::sqrt(d) : 0.0;
      double orig = std::sqrt(std::max(G[j * k + j], 0.0));
      if (nrm < 1e-14 || nrm < 1e-10 * orig) continue;
      Rm[j * k + j] = nrm;
      keep.push_back(j);
    }
    // T = R^{-1} restricted to kept vectors (upper triangular in kept order)
    const int kk = (int)keep.size();
    std::vector<double> Rk((size_t)kk * kk, 0.0), Tinv((size_t)kk * kk, 0.0);
    for (int a = 0; a < kk; ++a)
      for (int b = a; b < kk; ++b) Rk[a * kk + b] = Rm[keep[a] * k + keep[b]];
    for (int b = 0; b < kk; ++b) {
      Tinv[b * kk + b] = 1.0 / Rk[b * kk + b];
      for (int a = b - 1; a >= 0; --a) {
        double s = 0.0;
        for (int l = a + 1; l <= b; ++l) s += Rk[a * kk + l] * Tinv[l * kk + b];
        Tinv[a * kk + b] = -s / Rk[a * kk + a];
      }
    }
    // gather kept columns into T scratch, then V <- Vk * Tinv
    for (int a = 0; a < kk; ++a)
      HIPCHK(hipMemcpyAsync(D.T + (int64_t)a * D.n, Vp + (int64_t)keep[a] * D.n, sizeof(double) * D.n,
                            hipMemcpyDeviceToDevice, D.st));
    if (kk) {
      rc = gemm(D, D.T, nullptr, kk, Tinv, nullptr, kk, Vp);
      if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(D.st));
    kept = kk;
  }
  D.ms_ortho += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return COSMO_OK;
}

// B[:, 0..nume) = B[:, 0..kp) * Svec^T (Svec rows = Ritz vectors of length >= kp); same for AB
static int collapse(Dav &D, const std::vector<std::vector<double>> &svec, int kp) {
  const int q = (int)svec.size();
  std::vector<double> C((size_t)kp * q);
  for (int k = 0; k < kp; ++k)
    for (int j = 0; j < q; ++j) C[(size_t)k * q + j] = k < (int)svec[j].size() ? svec[j][k] : 0.0;
  int rc = gemm(D, D.B, nullptr, kp, C, nullptr, q, D.T);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(D.B, D.T, sizeof(double) * D.n * q, hipMemcpyDeviceToDevice, D.st));
  rc = gemm(D, D.AB, nullptr, kp, C, nullptr, q, D.T);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(D.AB, D.T, sizeof(double) * D.n * q, hipMemcpyDeviceToDevice, D.st));
  HIPCHK(hipStreamSynchronize(D.st));
  return COSMO_OK;
}

// packed lower triangle of S (s[i*(i+1)/2 + j], j <= i) -> eigenpairs (DiagonalizeS :21-98)
static void diagonalize_S(const std::vector<double> &s, int k, int nume, std::vector<double> &eigs,
                          std::vector<std::vector<double>> &svec) {
  std::vector<double> A((size_t)k * k);
  for (int i = 0; i < k; ++i)
    for (int j = 0; j <= i; ++j) A[i * k + j] = A[j * k + i] = s[(size_t)i * (i + 1) / 2 + j];
  // Householder + implicit QL (O(k^3), k <= 128); cyclic Jacobi only if QL fails
  std::vector<double> w, V, Vc;
  if (dense_symmetric_eigen(k, A, w, Vc)) {
    V.resize((size_t)k * k);
    for (int e = 0; e < k; ++e)
      for (int r = 0; r < k; ++r) V[(size_t)e * k + r] = Vc[(size_t)r * k + e];
  } else {
    sym_eig(k, A, w, V);
  }
  for (int i = 0; i < k; ++i) eigs[i] = w[i];
  svec.assign(nume, std::vector<double>(k, 0.0));
  for (int e = 0; e < nume && e < k; ++e)
    for (int r = 0; r < k; ++r) svec[e][r] = V[(size_t)e * k + r];
}

// S columns for basis vectors [c0, c0+m): s[j*(j+1)/2 + i] = b_i . ab_j, i <= j
static int update_S(Dav &D, std::vector<double> &s, int c0, int m) {
  auto t0 = std::chrono::steady_clock::now();
  std::vector<double> W;
  int rc = gram(D, D.B, c0 + m, D.col(D.AB, c0), m, W);
  if (rc) return rc;
  for (int j = 0; j < m; ++j) {
    int jj = c0 + j;
    for (int i = 0; i <= jj; ++i) s[(size_t)jj * (jj + 1) / 2 + i] = W[(size_t)i * m + j];
  }
  D.ms_rr += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return COSMO_OK;
}

}  // namespace cosmo

using namespace cosmo;

extern "C" {

int cosmo_dav_defaults(cosmo_dav_params *p, int nume) {
  if (!p || nume <= 0) return fail(COSMO_EINVAL, "cosmo_dav_defaults: bad args");
  memset(p, 0, sizeof(*p));
  p->nume = nume;
  p->mblock = nume;
  p->lim = nume + 20;
  p->maxiter = std::max(nume * 40, 200);
  p->crite = 1e-15;
  p->critc = 1e-12;
  p->critr = 1e-8;
  p->ortho = 1e-9;
  p->precision = COSMO_PREC_FP64;
  p->mixed_switch = 0.0;
  p->ninit = 0;
  p->verbose = 0;
  p->strict = 1;
  return COSMO_OK;
}

int cosmo_davidson(cosmo_op op, const cosmo_dav_params *pp, const double *init, double *evals, double *dE,
                   double *resid, double *evecs, cosmo_dav_stats *stats) {
  if (!op || !pp || !evals) return fail(COSMO_EINVAL, "cosmo_davidson: null argument");
  cosmo_dav_params p = *pp;
  const int64_t n = op->n;
  if (p.nume <= 0 || p.nume > n) return fail(COSMO_EINVAL, "cosmo_davidson: nume out of range");
  if (p.mblock <= 0) p.mblock = p.nume;
  if (p.mblock > 64) return fail(COSMO_EUNSUP, "cosmo_davidson: block larger than 64");
  if (p.lim <= 0) p.lim = p.nume + 20;
  if (p.lim > n) p.lim = (int)n;
  if (p.lim > 128) return fail(COSMO_EUNSUP, "cosmo_davidson: basis limit above 128");
  if (p.nume > p.lim || (p.nume == p.lim && p.nume != n)) return fail(COSMO_EINVAL, "cosmo_davidson: lim too small");
  if (p.maxiter <= 0) p.maxiter = std::max(p.nume * 40, 200);
  const bool mixed = p.precision == COSMO_PREC_FP32;
  if (mixed && !op->is_virtual && !op->csr.v32) return fail(COSMO_EINVAL, "mixed precision needs FP32 values");
  const double switch_tol = p.mixed_switch > 0 ? p.mixed_switch : 1e-5;
  hipSetDevice(op->device);
  auto tstart = std::chrono::steady_clock::now();
  Dav D;
  D.op = op;
  D.st = lib_stream();
  note_stream(*op, D.st);
  D.n = n;
  D.lim = p.lim;
  D.nume = p.nume;
  D.mblock = std::min(p.mblock, p.nume);
  int rc = dav_alloc(D);
  if (rc) return rc;
  const int nume = p.nume, lim = p.lim;
  double *dD = op->csr.diag;  // device diagonal
  // ---- setup_2d: initial basis
  int niv = 0;
  if (init && p.ninit > 0) {
    niv = std::min(p.ninit, lim);
    for (int j = 0; j < niv; ++j)
      HIPCHK(hipMemcpyAsync(D.col(D.B, j), init + (int64_t)j * n, sizeof(double) * n, hipMemcpyHostToDevice, D.st));
    // caller vectors need not be orthonormal (Ritz vectors from an .EE file already are,
    // and stay as they are up to rounding); dependent ones are dropped
    bool rs;
    int kept;
    rc = orthonormalise(D, 0, D.col(D.B, 0), niv, 1e-9, rs, kept);
    if (rc) return rc;
    niv = kept;
  }
  if (niv < nume) {
    // unit vectors at the lowest diagonal elements; ties keep the lower index
    std::vector<int64_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    const std::vector<double> &dg = op->hdiag;
    std::partial_sort(idx.begin(), idx.begin() + nume, idx.end(), [&](int64_t a, int64_t b) {
      return dg[a] < dg[b] || (dg[a] == dg[b] && a < b);
    });
    const double one = 1.0;
    HIPCHK(hipMemsetAsync(D.col(D.B, niv), 0, sizeof(double) * n * (nume - niv), D.st));
    for (int j = niv; j < nume; ++j)
      HIPCHK(hipMemcpyAsync(D.col(D.B, j) + idx[j], &one, sizeof(double), hipMemcpyHostToDevice, D.st));
    HIPCHK(hipStreamSynchronize(D.st));
    int added = nume - niv;
    if (niv > 0) {
      bool rs;
      rc = orthonormalise(D, niv, D.col(D.B, niv), nume - niv, 1e-9, rs, added);
      if (rc) return rc;
    }
    niv += added;
  }
  int prec = mixed ? COSMO_PREC_FP32 : COSMO_PREC_FP64;
  rc = apply_block(D, 0, niv, prec);
  if (rc) return rc;
  std::vector<double> s((size_t)lim * (lim + 1) / 2, 0.0);
  rc = update_S(D, s, 0, niv);
  if (rc) return rc;
  int kpass = niv, nncv = niv;
  std::vector<double> eigs(lim, 0.0), old(nume, 0.0);
  for (int i = 0; i < nume; ++i) eigs[i] = 0.0;
  std::vector<std::vector<double>> svec;
  std::vector<int> icv(nume, 0), incv(nume, 0);
  std::vector<double> rnorm(nume, 0.0);
  bool first = true, omin_to = false, fp64_phase = !mixed;
  int nloops = 0, ierr = 0, restarts = 0;
  // mixed precision: collapse to the Ritz vectors, recompute H.B in FP64, restart
  auto switch_fp64 = [&]() -> int {
    // keep the whole FP32-grown basis; it is re-orthonormalised in FP64 (vectors added
    // with a single round may have drifted, which matters once the basis nears the whole
    // space), then H.B and S are recomputed in FP64
    bool rs0;
    int kept0 = kpass;
    int r = orthonormalise(D, 0, D.B, kpass, p.ortho, rs0, kept0, 2);
    if (r) return r;
    kpass = kept0;
    r = apply_block(D, 0, kpass, COSMO_PREC_FP64);
    if (r) return r;
    std::fill(s.begin(), s.end(), 0.0);
    r = update_S(D, s, 0, kpass);
    if (r) return r;
    nncv = kpass;
    std::fill(icv.begin(), icv.end(), 0);
    for (int i = 0; i < nume; ++i) eigs[i] = 1.e+30;
    fp64_phase = true;
    omin_to = false;
    first = true;
    ++nloops;
    return COSMO_OK;
  };
  for (int verify = 0;; ++verify) {
  while (true) {
    for (int i = 0; i < nume; ++i) old[i] = eigs[i];
    diagonalize_S(s, kpass, nume, eigs, svec);
    // ---- Test_Convergence
    bool done = false;
    {
      int nnce = 0;
      for (int i = 0; i < nume; ++i)
        if (std::fabs(old[i] - eigs[i]) >= p.crite) ++nnce;
      if (!nnce) done = true;
      else {
        std::vector<int> ind;
        std::vector<double> rowlast;
        for (int i = 0; i < nume; ++i)
          if (icv[i] == 0) {
            double tmax = std::fabs(svec[i][kpass - 1]);
            for (int l = 1; l <= nncv - 1; ++l) tmax = std::max(tmax, std::fabs(svec[i][kpass - l - 1]));
            if (tmax < p.critc && !p.strict) icv[i] = 1;
            else {
              ind.push_back(i + 1);
              rowlast.push_back(tmax);
            }
          }
        nncv = (int)ind.size();
        if (!nncv) done = true;
        for (int i = 0; i < nncv; ++i) {  // selection by decreasing coefficient, first max wins
          int im = i;
          for (int j = i + 1; j < nncv; ++j)
            if (rowlast[j] > rowlast[im]) im = j;
          std::swap(rowlast[i], rowlast[im]);
          std::swap(ind[i], ind[im]);
          incv[i] = ind[i];
        }
      }
    }
    if (done || kpass >= n) {
      omin_to = true;
      if (!fp64_phase) {  // FP32 phase finished: hand over to the FP64 refinement
        rc = switch_fp64();
        if (rc) return rc;
        if (nloops >= p.maxiter) break;
        continue;
      }
      break;
    }
    if (kpass == lim) {
      rc = collapse(D, svec, lim);
      if (rc) return rc;
      std::fill(s.begin(), s.end(), 0.0);
      for (int i = 0; i < nume; ++i) s[(size_t)i * (i + 1) / 2 + i] = eigs[i];
      for (int i = 0; i < nume; ++i) {
        svec[i].assign(lim, 0.0);
        svec[i][i] = 1.0;
      }
      kpass = nume;
      ++restarts;
    }
    // ---- Add_Vectors_To_Basis
    {
      const int limadd = std::min(lim, D.mblock + kpass);
      const int ncand = nncv;
      // residuals of the candidates into R: r = AB s - lambda B s
      std::vector<double> C1((size_t)kpass * ncand), C2((size_t)kpass * ncand);
      for (int j = 0; j < ncand; ++j) {
        int e = incv[j] - 1;
        for (int k = 0; k < kpass; ++k) {
          C1[(size_t)k * ncand + j] = svec[e][k];
          C2[(size_t)k * ncand + j] = -eigs[e] * svec[e][k];
        }
      }
      auto tr = std::chrono::steady_clock::now();
      rc = gemm(D, D.AB, D.B, kpass, C1, &C2, ncand, D.R);
      if (rc) return rc;
      std::vector<double> G;
      rc = gram(D, D.R, ncand, D.R, ncand, G);
      if (rc) return rc;
      D.ms_rr += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tr).count();
      int nadded = 0, icvc = 0;
      bool all_conv = false;
      std::vector<int> addidx;
      for (int j = 0; j < ncand; ++j) {
        int e = incv[j] - 1;
        double ss = std::sqrt(std::max(G[(size_t)j * ncand + j], 0.0));
        rnorm[e] = ss;
        double thr = fp64_phase ? p.critr : std::max(p.critr, switch_tol * std::max(1.0, std::fabs(eigs[e])));
        if (ss < thr) {
          ++icvc;
          icv[e] = 1;
          if (icvc < ncand) continue;
          all_conv = true;
          break;
        }
        addidx.push_back(j);
        ++nadded;
        if (nadded + kpass == limadd) break;
      }
      if (all_conv) {
        omin_to = true;
        if (!fp64_phase) {
          rc = switch_fp64();
          if (rc) return rc;
          if (nloops >= p.maxiter) break;
          continue;
        }
        break;
      }
      nncv = nadded;
      std::vector<double> lam(nadded);
      for (int a = 0; a < nadded; ++a) {
        int j = addidx[a];
        lam[a] = eigs[incv[j] - 1];
        incv[a] = incv[j];
        HIPCHK(hipMemcpyAsync(D.col(D.B, kpass + a), D.col(D.R, j), sizeof(double) * n, hipMemcpyDeviceToDevice, D.st));
      }
      if (nadded) {
        HIPCHK(hipMemcpyAsync(D.lam, lam.data(), sizeof(double) * nadded, hipMemcpyHostToDevice, D.st));
        int64_t tot = n * nadded;
        hipLaunchKernelGGL(k_precond, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, D.st, n, nadded,
                           D.col(D.B, kpass), n, D.lam, dD);
        HIPCHK(hipGetLastError());
      }
      bool restart = false;
      int kept = nadded;
      rc = orthonormalise(D, kpass, D.col(D.B, kpass), nadded, p.ortho, restart, kept, fp64_phase ? 2 : 1);
      if (rc) return rc;
      nncv = kept;
      if (!restart) first = true;
      else if (first) {
        first = false;
        rc = collapse(D, svec, kpass);
        if (rc) return rc;
        std::fill(s.begin(), s.end(), 0.0);
        for (int i = 0; i < nume; ++i) s[(size_t)i * (i + 1) / 2 + i] = eigs[i];
        for (int i = 0; i < nume; ++i) eigs[i] = 1.e+30;
        kpass = nume;
        ++restarts;
        continue;
      } else {
        ierr += 1024;
        omin_to = true;
        break;
      }
      if (nncv == 0) {
        // every candidate was linearly dependent on the basis: nothing to add
        omin_to = true;
        if (!fp64_phase) {
          rc = switch_fp64();
          if (rc) return rc;
          if (nloops >= p.maxiter) break;
          continue;
        }
        break;
      }
    }
    // ---- Add_Vectors_To_AB_S
    rc = apply_block(D, kpass, nncv, fp64_phase ? COSMO_PREC_FP64 : COSMO_PREC_FP32);
    if (rc) return rc;
    rc = update_S(D, s, kpass, nncv);
    if (rc) return rc;
    kpass += nncv;
    ++nloops;
    if (p.verbose)
      fprintf(stderr, "davidson it %d kpass %d nncv %d e0 %.10f %s\n", nloops, kpass, nncv, eigs[0],
              fp64_phase ? "fp64" : "fp32");
    if (nloops >= p.maxiter) break;
  }
  if (!omin_to) kpass -= nncv;
  for (int i = 0; i < nume; ++i) old[i] = std::fabs(old[i] - eigs[i]);
  if (kpass != nume) {
    // refresh the Ritz vectors for the final kpass (S may have grown before the exit)
    diagonalize_S(s, kpass, nume, eigs, svec);
    rc = collapse(D, svec, kpass);
    if (rc) return rc;
  }
  // final residuals ||AB_i - lambda_i B_i||
  {
    std::vector<double> C1((size_t)nume * nume, 0.0), C2((size_t)nume * nume, 0.0);
    for (int i = 0; i < nume; ++i) {
      C1[(size_t)i * nume + i] = 1.0;
      C2[(size_t)i * nume + i] = -eigs[i];
    }
    rc = gemm(D, D.AB, D.B, nume, C1, &C2, nume, D.R);
    if (rc) return rc;
    std::vector<double> G;
    rc = gram(D, D.R, nume, D.R, nume, G);
    if (rc) return rc;
    for (int i = 0; i < nume; ++i) rnorm[i] = std::sqrt(std::max(G[(size_t)i * nume + i], 0.0));
  }
  // strict: a pair accepted on an earlier residual may have drifted above critr while the
  // basis kept growing; reopen such pairs and iterate again from the Ritz vectors
  bool reopen = false;
  if (p.strict && fp64_phase && nloops < p.maxiter && ierr == 0 && verify < 20)
    for (int i = 0; i < nume; ++i)
      if (rnorm[i] > p.critr) {
        icv[i] = 0;
        reopen = true;
      }
  if (!reopen) break;
  std::fill(s.begin(), s.end(), 0.0);
  for (int i = 0; i < nume; ++i) s[(size_t)i * (i + 1) / 2 + i] = eigs[i];
  for (int i = 0; i < nume; ++i) eigs[i] = 1.e+30;  // the eigenvalue-change test must not fire
  kpass = nume;
  nncv = nume;
  omin_to = false;
  first = true;
  }
  for (int i = 0; i < nume; ++i) {
    evals[i] = eigs[i];
    if (dE) dE[i] = old[i];
    if (resid) resid[i] = rnorm[i];
  }
  if (evecs)
    for (int i = 0; i < nume; ++i)
      HIPCHK(hipMemcpyAsync(evecs + (int64_t)i * n, D.col(D.B, i), sizeof(double) * n, hipMemcpyDeviceToHost, D.st));
  HIPCHK(hipStreamSynchronize(D.st));
  if (stats) {
    stats->iterations = nloops;
    stats->matvecs = D.matvecs;
    stats->matvecs_fp64 = D.matvecs64;
    stats->restarts = restarts;
    stats->ierr = ierr;
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tstart).count();
    stats->ms_apply = D.ms_apply;
    stats->ms_ortho = D.ms_ortho;
    stats->ms_rr = D.ms_rr;
  }
  if (nloops >= p.maxiter && !omin_to) return fail(COSMO_ECONV, "Davidson reached maxiter");
  return COSMO_OK;
}

// .EE layout: [u32 count] then per pair [f64 E][f64 dE][f64 resid][u64 n][n f64]
int cosmo_write_ee(const char *path, int64_t n, int count, const double *evals, const double *dE, const double *resid,
                   const double *evecs) {
  if (!path || count < 0 || !evals) return fail(COSMO_EINVAL, "cosmo_write_ee: bad args");
  FILE *f = fopen(path, "wb");
  if (!f) return fail(COSMO_EIO, std::string("can not open output ") + path);
  uint32_t c = (uint32_t)count;
  fwrite(&c, 4, 1, f);
  uint64_t nn = (uint64_t)n;
  std::vector<double> zero(evecs ? 0 : n, 0.0);
  for (int i = 0; i < count; ++i) {
    double e = evals[i], d = dE ? dE[i] : 0.0, r = resid ? resid[i] : 0.0;
    fwrite(&e, 8, 1, f);
    fwrite(&d, 8, 1, f);
    fwrite(&r, 8, 1, f);
    fwrite(&nn, 8, 1, f);
    fwrite(evecs ? evecs + (int64_t)i * n : zero.data(), 8, n, f);
  }
  if (fclose(f)) return fail(COSMO_EIO, "write failed");
  return COSMO_OK;
}

int cosmo_read_ee(const char *path, int64_t n, int maxcount, int *count, double *evals, double *evecs) {
  FILE *f = fopen(path, "rb");
  if (!f) return fail(COSMO_EIO, std::string("cannot open ") + path);
  uint32_t c = 0;
  if (fread(&c, 4, 1, f) != 1) {
    fclose(f);
    return fail(COSMO_EIO, "truncated .EE");
  }
  int got = 0;
  for (uint32_t i = 0; i < c && (int)i < maxcount; ++i) {
    double e, d, r;
    uint64_t nn;
    if (fread(&e, 8, 1, f) != 1 || fread(&d, 8, 1, f) != 1 || fread(&r, 8, 1, f) != 1 || fread(&nn, 8, 1, f) != 1) {
      fclose(f);
      return fail(COSMO_EIO, "truncated .EE");
    }
    if ((int64_t)nn != n) {
      fclose(f);
      return fail(COSMO_EIO, "Dimension mismatch matrix n=" + std::to_string(n) + " file n=" + std::to_string(nn));
    }
    if (evals) evals[i] = e;
    if (evecs) {
      if (fread(evecs + (int64_t)i * n, 8, n, f) != (size_t)n) {
        fclose(f);
        return fail(COSMO_EIO, "truncated .EE vector");
      }
    } else
      fseek(f, 8 * n, SEEK_CUR);
    ++got;
  }
  fclose(f);
  if (count) *count = got;
  return COSMO_OK;
}

}  // extern "C"
