// dav_wg.h -- block Davidson for one small system per workgroup: every batched solve whose
// dimension, basis limit and block fit the bounds below runs to completion inside a single
// launch, one workgroup per solve.
//
// The iteration restates the host driver (cosmo_davidson in davidson.hip, itself dvdrvr2 of
// include/kdav/fsdavidson.cxx:615-775) step for step: unit start vectors at the lowest
// diagonal elements, the eigenvalue-change and Ritz-coefficient convergence test and the
// candidate order (Test_Convergence :212-281), residuals and the diagonal preconditioner
// (Add_Vectors_To_Basis :426-516, floor 1e-13), blocked Gram-Schmidt passes with a
// pivot-dropping Cholesky QR, restarts by collapse (:99-121, :517-539), the switch from the
// FP32 product to FP64 refinement, and the strict re-opening of pairs whose residual
// drifted.  Every n-long operation is spread over the workgroup; the (kpass x kpass)
// Rayleigh-Ritz problem is solved in LDS by parallel cyclic Jacobi, warm-started from the
// previous Ritz vectors (the basis only grows between restarts, so the previous vectors
// nearly diagonalise the new matrix).  Decisions are taken by thread 0 and published
// through LDS, so control flow is uniform over the workgroup and every barrier is reached
// by all of its threads.  Basis, image and residual blocks are FP64 and column-major
// (ld n); H.X runs on a row-major copy of the block (one cache line per stored entry).
//
// Only plain HIP constructs are used (threadIdx, blockDim, __syncthreads, __shared__), and
// every parallel loop is a stride loop over its work items, with no intra-wave shuffles or
// atomics: the kernel's result does not depend on the workgroup size.

#include "jacobi.h"

#define DWG_THREADS 512
#define DWG_LIM 48     // basis limit (lim) bound
#define DWG_MB 16      // nume and mblock bound
#define DWG_NMAX 32768 // dimension bound

struct DwgSolve {
  int64_t n;
  const int64_t *rowptr;  // full symmetric rows (both triangles), as the stored operator
  const uint32_t *cols;
  const double *v64;
  const float *v32;
  const double *diag;
  double *B, *AB, *R;  // n x lim each, column-major
  double *X, *Y;       // n x DWG_MB each, row-major staging of H.X
  int nume, lim, mblock, maxiter;
  double critr;
  double *evals, *dE, *resid;  // nume each
  int *info;                   // iterations, matvecs, matvecs_fp64, restarts, ierr, converged
};

struct DwgParams {
  double crite, critc, ortho, switch_tol;
  int strict, mixed;
};

struct DwgShared {
  double S[DWG_LIM * DWG_LIM];       // Rayleigh matrix B^T H B (full, row-major)
  double A[DWG_LIM * DWG_LIM];       // Jacobi matrix; R factor of the Cholesky QR
  double V[DWG_LIM * DWG_LIM];       // Ritz vectors in the basis: V[r * DWG_LIM + e]
  double W2[2 * DWG_LIM * DWG_LIM];  // Jacobi second buffers / Gram partial sums / T = R^-1
  double G[DWG_LIM * DWG_LIM];       // Gram results, row-major (P x M)
  double Wt[DWG_LIM * DWG_LIM];      // overlaps B^T V of an orthogonalisation pass
  double C1[DWG_LIM * DWG_MB], C2[DWG_LIM * DWG_MB];
  double eigs[DWG_LIM], old[DWG_MB], rnorm[DWG_MB], lam[DWG_MB], base[DWG_LIM];
  double jc[DWG_LIM / 2], js[DWG_LIM / 2];
  double minv[DWG_THREADS];
  int mini[DWG_THREADS];
  int partner[DWG_LIM], pairidx[DWG_LIM], order[DWG_LIM], keep[DWG_LIM];
  int icv[DWG_MB], incv[DWG_MB], add[DWG_MB], chosen[DWG_MB];
  // control state, written by thread 0 between barriers
  int kpass, nncv, kv, fp64, first, omin_to, nloops, restarts, ierr, verify;
  int done, flag, kept, restart, ncand, nadded, matvecs, matvecs64, rot;
  double dprev, dctl;
};

// ---- parallel building blocks (every thread of the workgroup calls them) ----------------

// G[a * M + b] = sum_r X[r + a n] Y[r + b n], a < P <= DWG_LIM, b < M <= DWG_MB.  Threads own
// 4 x 4 output blocks over interleaved row groups; the groups' sums are added in group order.
__device__ void dwg_gram(DwgShared &sh, int64_t n, const double *X, int P, const double *Y, int M) {
  const int tid = threadIdx.x;
  if (P <= 0 || M <= 0) return;  // uniform: every thread sees the same P, M
  const int nbA = (P + 3) / 4, nbB = (M + 3) / 4, nblk = nbA * nbB;
  int RG = 256 / nblk;
  if (RG < 1) RG = 1;
  const int T = nblk * RG;
  double *part = sh.W2;
  for (int t = tid; t < T; t += blockDim.x) {
    const int blk = t % nblk, g = t / nblk;
    const int a0 = (blk / nbB) * 4, b0 = (blk % nbB) * 4;
    const double *xa[4], *yb[4];
    for (int i = 0; i < 4; ++i) {
      xa[i] = X + (int64_t)min(a0 + i, P - 1) * n;
      yb[i] = Y + (int64_t)min(b0 + i, M - 1) * n;
    }
    double acc[16];
    for (int i = 0; i < 16; ++i) acc[i] = 0.0;
    for (int64_t r = g; r < n; r += RG) {
      double x[4], y[4];
      for (int i = 0; i < 4; ++i) {
        x[i] = xa[i][r];
        y[i] = yb[i][r];
      }
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) acc[i * 4 + j] += x[i] * y[j];
    }
    for (int i = 0; i < 16; ++i) part[(int64_t)t * 16 + i] = acc[i];
  }
  __syncthreads();
  for (int o = tid; o < P * M; o += blockDim.x) {
    const int a = o / M, b = o % M;
    const int blk = (a / 4) * nbB + b / 4, ij = (a % 4) * 4 + b % 4;
    double s = 0.0;
    for (int g = 0; g < RG; ++g) s += part[(int64_t)(g * nblk + blk) * 16 + ij];
    sh.G[o] = s;
  }
  __syncthreads();
}

// Y[:, j] = X1 C1[:, j] (+ X2 C2[:, j]) for j < q <= DWG_MB, k < p; C row-major p x DWG_MB in
// LDS.  Row by row, all reads of a row before its writes, so Y may be X1 or X2 when q <= p
// covers only columns already read (callers use distinct arrays or one chunk).
__device__ void dwg_combine(const DwgShared &sh, int64_t n, const double *X1, const double *X2, int p, int q,
                            double *Yc) {
  for (int64_t r = threadIdx.x; r < n; r += blockDim.x) {
    double acc[DWG_MB];
    for (int j = 0; j < DWG_MB; ++j) acc[j] = 0.0;
    for (int k = 0; k < p; ++k) {
      const double a = X1[r + (int64_t)k * n];
      for (int j = 0; j < DWG_MB; ++j) acc[j] += a * sh.C1[k * DWG_MB + j];
      if (X2) {
        const double b = X2[r + (int64_t)k * n];
        for (int j = 0; j < DWG_MB; ++j) acc[j] += b * sh.C2[k * DWG_MB + j];
      }
    }
    for (int j = 0; j < DWG_MB; ++j)
      if (j < q) Yc[r + (int64_t)j * n] = acc[j];
  }
  __syncthreads();
}

// AB[:, c0 .. c0+m) = H B[:, c0 .. c0+m) through the row-major staging, in chunks of DWG_MB
__device__ void dwg_apply(DwgShared &sh, const DwgSolve &s, int c0, int m, bool fp64) {
  const int64_t n = s.n;
  for (int j0 = 0; j0 < m; j0 += DWG_MB) {
    const int mm = min(DWG_MB, m - j0);
    for (int64_t t = threadIdx.x; t < n * mm; t += blockDim.x) {
      const int64_t r = t / mm;
      const int j = (int)(t % mm);
      s.X[r * DWG_MB + j] = s.B[r + (int64_t)(c0 + j0 + j) * n];
    }
    __syncthreads();
    for (int64_t r = threadIdx.x; r < n; r += blockDim.x) {
      const int64_t e0 = s.rowptr[r], e1 = s.rowptr[r + 1];
      if (fp64) {
        double acc[DWG_MB];
        for (int j = 0; j < DWG_MB; ++j) acc[j] = 0.0;
        int64_t e = e0;
        for (; e + 4 <= e1; e += 4) {  // four entries per step: independent loads in flight
          double h[4];
          const double *x[4];
          for (int u = 0; u < 4; ++u) {
            h[u] = s.v64[e + u];
            x[u] = s.X + (int64_t)s.cols[e + u] * DWG_MB;
          }
          for (int u = 0; u < 4; ++u)
            for (int j = 0; j < DWG_MB; ++j)
              if (j < mm) acc[j] += h[u] * x[u][j];
        }
        for (; e < e1; ++e) {
          const double h = s.v64[e];
          const double *x = s.X + (int64_t)s.cols[e] * DWG_MB;
          for (int j = 0; j < DWG_MB; ++j)
            if (j < mm) acc[j] += h * x[j];
        }
        for (int j = 0; j < DWG_MB; ++j)
          if (j < mm) s.Y[r * DWG_MB + j] = acc[j];
      } else {
        float acc[DWG_MB];
        for (int j = 0; j < DWG_MB; ++j) acc[j] = 0.0f;
        int64_t e = e0;
        for (; e + 4 <= e1; e += 4) {
          float h[4];
          const double *x[4];
          for (int u = 0; u < 4; ++u) {
            h[u] = s.v32[e + u];
            x[u] = s.X + (int64_t)s.cols[e + u] * DWG_MB;
          }
          for (int u = 0; u < 4; ++u)
            for (int j = 0; j < DWG_MB; ++j)
              if (j < mm) acc[j] += h[u] * (float)x[u][j];
        }
        for (; e < e1; ++e) {
          const float h = s.v32[e];
          const double *x = s.X + (int64_t)s.cols[e] * DWG_MB;
          for (int j = 0; j < DWG_MB; ++j)
            if (j < mm) acc[j] += h * (float)x[j];
        }
        for (int j = 0; j < DWG_MB; ++j)
          if (j < mm) s.Y[r * DWG_MB + j] = (double)acc[j];
      }
    }
    __syncthreads();
    for (int64_t t = threadIdx.x; t < n * mm; t += blockDim.x) {
      const int j = (int)(t / n);
      const int64_t r = t % n;
      s.AB[r + (int64_t)(c0 + j0 + j) * n] = s.Y[r * DWG_MB + j];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    sh.matvecs += m;
    if (fp64) sh.matvecs64 += m;
  }
  __syncthreads();
}

// S rows/columns of basis vectors [c0, c0+m): S[i][j] = b_i . ab_j, i < c0+m
__device__ void dwg_update_S(DwgShared &sh, const DwgSolve &s, int c0, int m) {
  for (int j0 = 0; j0 < m; j0 += DWG_MB) {
    const int mm = min(DWG_MB, m - j0);
    dwg_gram(sh, s.n, s.B, c0 + j0 + mm, s.AB + (int64_t)(c0 + j0) * s.n, mm);
    for (int o = threadIdx.x; o < (c0 + j0 + mm) * mm; o += blockDim.x) {
      const int i = o / mm, j = c0 + j0 + o % mm;
      if (i <= j) {
        sh.S[i * DWG_LIM + j] = sh.G[o];
        sh.S[j * DWG_LIM + i] = sh.G[o];
      }
    }
    __syncthreads();
  }
}

__device__ void dwg_eigen(DwgShared &sh, int k) { lds_jacobi<DWG_LIM>(sh, k); }

// component r of Ritz vector i (in the current basis)
__device__ inline double dwg_svec(const DwgShared &sh, int i, int r) { return sh.V[r * DWG_LIM + sh.order[i]]; }

// B[:, 0..q) = B[:, 0..kp) svec^T and the same for AB (q = nume Ritz vectors)
__device__ void dwg_collapse(DwgShared &sh, const DwgSolve &s, int kp, int q) {
  for (int o = threadIdx.x; o < kp * DWG_MB; o += blockDim.x) {
    const int k = o / DWG_MB, j = o % DWG_MB;
    sh.C1[o] = j < q ? dwg_svec(sh, j, k) : 0.0;
  }
  __syncthreads();
  dwg_combine(sh, s.n, s.B, nullptr, kp, q, s.R);
  for (int64_t t = threadIdx.x; t < s.n * q; t += blockDim.x) s.B[t] = s.R[t];
  __syncthreads();
  dwg_combine(sh, s.n, s.AB, nullptr, kp, q, s.R);
  for (int64_t t = threadIdx.x; t < s.n * q; t += blockDim.x) s.AB[t] = s.R[t];
  __syncthreads();
}

// S = diag(eigs[0..q)), V = I: the state right after a collapse
__device__ void dwg_reset_S(DwgShared &sh, int q) {
  for (int o = threadIdx.x; o < DWG_LIM * DWG_LIM; o += blockDim.x) {
    const int r = o / DWG_LIM, c = o % DWG_LIM;
    sh.S[o] = (r == c && r < q) ? sh.eigs[r] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) sh.kv = 0;
  __syncthreads();
}

// Orthonormalise the q vectors at B[:, p .. p+q) against B[:, 0..p) and among themselves
// (same passes and thresholds as the host driver's FP64 path); sh.kept, sh.restart
__device__ void dwg_orthonormalise(DwgShared &sh, const DwgSolve &s, int p, int q, double ortho, int rounds) {
  const int tid = threadIdx.x;
  const int64_t n = s.n;
  double *Vp = s.B + (int64_t)p * n;
  if (tid == 0) {
    sh.kept = q;
    sh.restart = 0;
  }
  __syncthreads();
  for (int round = 0; round < rounds && sh.kept > 0; ++round) {
    const int k = sh.kept;
    if (tid == 0) sh.dprev = 1e7;
    __syncthreads();
    for (int pass = 0; pass < 4; ++pass) {
      // G = [B V]^T V, chunked over V's columns; W = rows < p, Gv = rows p.. (kept in A)
      for (int j0 = 0; j0 < k; j0 += DWG_MB) {
        const int mm = min(DWG_MB, k - j0);
        dwg_gram(sh, n, s.B, p + k, Vp + (int64_t)j0 * n, mm);
        for (int o = tid; o < (p + k) * mm; o += blockDim.x) {
          const int a = o / mm, j = j0 + o % mm;
          if (a < p) sh.Wt[a * DWG_LIM + j] = sh.G[o];  // W
          else sh.A[(a - p) * DWG_LIM + j] = sh.G[o];   // Gv
        }
        __syncthreads();
      }
      if (pass == 0 && tid == 0)
        for (int j = 0; j < k; ++j) sh.base[j] = sqrt(fmax(sh.A[j * DWG_LIM + j], 0.0));
      __syncthreads();
      if (p == 0) break;
      if (tid == 0) {
        double dcur = 0.0;
        for (int j = 0; j < k; ++j) {
          const double nv = sqrt(fmax(sh.A[j * DWG_LIM + j], 0.0));
          if (nv > 0)
            for (int a = 0; a < p; ++a) dcur = fmax(dcur, fabs(sh.Wt[a * DWG_LIM + j]) / nv);
        }
        sh.flag = dcur < ortho ? 1 : (dcur > sh.dprev ? 2 : 0);
        sh.dprev = dcur;
      }
      __syncthreads();
      // V -= B W, chunked over V's columns (W chunk staged in C1)
      for (int j0 = 0; j0 < k; j0 += DWG_MB) {
        const int mm = min(DWG_MB, k - j0);
        for (int o = tid; o < p * DWG_MB; o += blockDim.x) {
          const int a = o / DWG_MB, j = o % DWG_MB;
          sh.C1[o] = j < mm ? sh.Wt[a * DWG_LIM + j0 + j] : 0.0;
        }
        __syncthreads();
        for (int64_t r = tid; r < n; r += blockDim.x) {
          double acc[DWG_MB];
          for (int j = 0; j < DWG_MB; ++j) acc[j] = 0.0;
          for (int a = 0; a < p; ++a) {
            const double b = s.B[r + (int64_t)a * n];
            for (int j = 0; j < DWG_MB; ++j) acc[j] += b * sh.C1[a * DWG_MB + j];
          }
          for (int j = 0; j < DWG_MB; ++j)
            if (j < mm) Vp[r + (int64_t)(j0 + j) * n] -= acc[j];
        }
        __syncthreads();
      }
      // Gv -= W^T W (B orthonormal)
      for (int o = tid; o < k * k; o += blockDim.x) {
        const int i = o / k, j = o % k;
        double t = 0.0;
        for (int a = 0; a < p; ++a) t += sh.Wt[a * DWG_LIM + i] * sh.Wt[a * DWG_LIM + j];
        sh.A[i * DWG_LIM + j] -= t;
      }
      __syncthreads();
      const int flag = sh.flag;
      __syncthreads();
      if (flag == 1) break;
      if (flag == 2) {
        if (tid == 0) sh.restart = 1;
        __syncthreads();
        break;
      }
    }
    if (sh.restart) break;
    // Cholesky QR of Gv in order (thread 0): R factor and T = R^-1 in the two halves of W2
    double *Rm = sh.W2 + DWG_LIM * DWG_LIM, *T = sh.W2;
    if (tid == 0) {
      for (int o = 0; o < k * k; ++o) Rm[o] = 0.0;
      int nk = 0;
      for (int j = 0; j < k; ++j) {
        double d = sh.A[j * DWG_LIM + j];
        for (int ii = 0; ii < nk; ++ii) {
          const int i = sh.keep[ii];
          double t = sh.A[i * DWG_LIM + j];
          for (int ll = 0; ll < ii; ++ll) {
            const int l = sh.keep[ll];
            t -= Rm[l * k + i] * Rm[l * k + j];
          }
          t /= Rm[i * k + i];
          Rm[i * k + j] = t;
          d -= t * t;
        }
        const double nrm = d > 0 ? sqrt(d) : 0.0;
        const double orig = sqrt(fmax(sh.A[j * DWG_LIM + j], 0.0));
        if (nrm <= 1e-14 * sh.base[j] || nrm < 1e-10 * orig) continue;
        Rm[j * k + j] = nrm;
        sh.keep[nk++] = j;
      }
      // T = inverse of R restricted to the kept vectors (upper triangular, kept order)
      for (int o = 0; o < nk * nk; ++o) T[o] = 0.0;
      for (int b = 0; b < nk; ++b) {
        T[b * nk + b] = 1.0 / Rm[sh.keep[b] * k + sh.keep[b]];
        for (int a = b - 1; a >= 0; --a) {
          double t = 0.0;
          for (int l = a + 1; l <= b; ++l) t += Rm[sh.keep[a] * k + sh.keep[l]] * T[l * nk + b];
          T[a * nk + b] = -t / Rm[sh.keep[a] * k + sh.keep[a]];
        }
      }
      sh.kept = nk;
    }
    __syncthreads();
    const int kk = sh.kept;
    // V[:, 0..kk) = V[:, keep] T, through R (the kept columns are read before any write)
    for (int j0 = 0; j0 < kk; j0 += DWG_MB) {
      const int mm = min(DWG_MB, kk - j0);
      for (int64_t r = tid; r < n; r += blockDim.x) {
        double acc[DWG_MB];
        for (int j = 0; j < DWG_MB; ++j) acc[j] = 0.0;
        for (int a = 0; a < kk; ++a) {
          const double v = Vp[r + (int64_t)sh.keep[a] * n];
          for (int j = 0; j < DWG_MB; ++j)
            if (j < mm && a <= j0 + j) acc[j] += v * T[a * kk + j0 + j];
        }
        for (int j = 0; j < DWG_MB; ++j)
          if (j < mm) s.R[r + (int64_t)(j0 + j) * n] = acc[j];
      }
      __syncthreads();
    }
    for (int64_t t = tid; t < n * kk; t += blockDim.x) Vp[t] = s.R[t];
    __syncthreads();
  }
}

// the whole solve of workgroup blockIdx.x
__global__ __launch_bounds__(DWG_THREADS) void k_dav_wg(const DwgSolve *__restrict__ solves, DwgParams prm) {
  __shared__ DwgShared sh;
  const DwgSolve s = solves[blockIdx.x];
  const int tid = threadIdx.x;
  const int64_t n = s.n;
  const int nume = s.nume, lim = s.lim;
  if (tid == 0) {
    sh.kv = 0;
    sh.nloops = 0;
    sh.restarts = 0;
    sh.ierr = 0;
    sh.matvecs = 0;
    sh.matvecs64 = 0;
    sh.fp64 = prm.mixed ? 0 : 1;
    sh.first = 1;
    sh.omin_to = 0;
  }
  // ---- start vectors: unit vectors at the nume lowest diagonal elements (ties: lower index)
  for (int c = 0; c < nume; ++c) {
    double best = 1e308;
    int bi = 0x7fffffff;
    for (int64_t r = tid; r < n; r += blockDim.x) {
      bool taken = false;
      for (int l = 0; l < c; ++l) taken |= (sh.chosen[l] == (int)r);
      const double v = s.diag[r];
      if (!taken && (v < best || (v == best && (int)r < bi))) {
        best = v;
        bi = (int)r;
      }
    }
    sh.minv[tid] = best;
    sh.mini[tid] = bi;
    __syncthreads();
    if (tid == 0) {
      for (int t = 1; t < (int)blockDim.x; ++t)
        if (sh.minv[t] < best || (sh.minv[t] == best && sh.mini[t] < bi)) {
          best = sh.minv[t];
          bi = sh.mini[t];
        }
      sh.chosen[c] = bi;
    }
    __syncthreads();
  }
  for (int64_t t = tid; t < n * nume; t += blockDim.x) s.B[t] = (t % n == sh.chosen[t / n]) ? 1.0 : 0.0;
  __syncthreads();
  dwg_apply(sh, s, 0, nume, sh.fp64);
  for (int o = tid; o < DWG_LIM * DWG_LIM; o += blockDim.x) sh.S[o] = 0.0;
  __syncthreads();
  dwg_update_S(sh, s, 0, nume);
  if (tid == 0) {
    sh.kpass = nume;
    sh.nncv = nume;
    for (int i = 0; i < DWG_LIM; ++i) sh.eigs[i] = 0.0;
    for (int i = 0; i < nume; ++i) {
      sh.icv[i] = 0;
      sh.old[i] = 0.0;
      sh.rnorm[i] = 0.0;
    }
  }
  __syncthreads();
  for (int verify = 0;; ++verify) {
    while (true) {
      // ---- Rayleigh-Ritz and Test_Convergence
      if (tid == 0)
        for (int i = 0; i < nume; ++i) sh.old[i] = sh.eigs[i];
      __syncthreads();
      dwg_eigen(sh, sh.kpass);
      if (tid == 0) {
        const int kpass = sh.kpass;
        int done = 0, nnce = 0;
        for (int i = 0; i < nume; ++i)
          if (fabs(sh.old[i] - sh.eigs[i]) >= prm.crite) ++nnce;
        if (!nnce) done = 1;
        else {
          int ind[DWG_MB];
          double rowlast[DWG_MB];
          int nc = 0;
          for (int i = 0; i < nume; ++i)
            if (sh.icv[i] == 0) {
              double tmax = fabs(dwg_svec(sh, i, kpass - 1));
              for (int l = 1; l <= min(sh.nncv, kpass) - 1; ++l)
                tmax = fmax(tmax, fabs(dwg_svec(sh, i, kpass - l - 1)));
              if (tmax < prm.critc && !prm.strict) sh.icv[i] = 1;
              else {
                ind[nc] = i + 1;
                rowlast[nc] = tmax;
                ++nc;
              }
            }
          sh.nncv = nc;
          if (!nc) done = 1;
          for (int i = 0; i < nc; ++i) {  // decreasing coefficient, first maximum wins
            int im = i;
            for (int j = i + 1; j < nc; ++j)
              if (rowlast[j] > rowlast[im]) im = j;
            const double tr = rowlast[i];
            rowlast[i] = rowlast[im];
            rowlast[im] = tr;
            const int ti = ind[i];
            ind[i] = ind[im];
            ind[im] = ti;
            sh.incv[i] = ind[i];
          }
        }
        // 0: go on, 1: phase finished (switch or stop), 3: collapse first
        sh.done = (done || kpass >= n) ? 1 : (kpass == lim ? 3 : 0);
      }
      __syncthreads();
      if (sh.done == 1) {
        if (tid == 0) sh.omin_to = 1;
        __syncthreads();
        if (!sh.fp64) {
          // switch to FP64 refinement: re-orthonormalise the basis, H.B and S in FP64
          dwg_orthonormalise(sh, s, 0, sh.kpass, prm.ortho, 2);
          if (tid == 0) sh.kpass = sh.kept;
          __syncthreads();
          dwg_apply(sh, s, 0, sh.kpass, true);
          for (int o = tid; o < DWG_LIM * DWG_LIM; o += blockDim.x) sh.S[o] = 0.0;
          __syncthreads();
          dwg_update_S(sh, s, 0, sh.kpass);
          if (tid == 0) {
            sh.nncv = sh.kpass;
            for (int i = 0; i < nume; ++i) {
              sh.icv[i] = 0;
              sh.eigs[i] = 1e30;
            }
            sh.kv = 0;
            sh.fp64 = 1;
            sh.omin_to = 0;
            sh.first = 1;
            ++sh.nloops;
          }
          __syncthreads();
          if (sh.nloops >= s.maxiter) break;
          continue;
        }
        break;
      }
      if (sh.done == 3) {
        dwg_collapse(sh, s, lim, nume);
        dwg_reset_S(sh, nume);
        if (tid == 0) {
          for (int e = 0; e < nume; ++e) sh.order[e] = e;
          sh.kpass = nume;
          ++sh.restarts;
        }
        __syncthreads();
        // V = I of size nume for the residuals below
        for (int o = tid; o < nume * nume; o += blockDim.x)
          sh.V[(o / nume) * DWG_LIM + o % nume] = (o / nume == o % nume) ? 1.0 : 0.0;
        __syncthreads();
      }
      // ---- residuals of the candidates: R = AB s - lambda B s
      const int kpass = sh.kpass, ncand = sh.nncv;
      for (int o = tid; o < kpass * DWG_MB; o += blockDim.x) {
        const int k = o / DWG_MB, j = o % DWG_MB;
        double v = 0.0, l = 0.0;
        if (j < ncand) {
          const int e = sh.incv[j] - 1;
          v = dwg_svec(sh, e, k);
          l = sh.eigs[e];
        }
        sh.C1[o] = v;
        sh.C2[o] = -l * v;
      }
      __syncthreads();
      dwg_combine(sh, n, s.AB, s.B, kpass, ncand, s.R);
      dwg_gram(sh, n, s.R, ncand, s.R, ncand);
      if (tid == 0) {
        const int limadd = min(lim, s.mblock + kpass);
        int nadded = 0, icvc = 0, all_conv = 0;
        for (int j = 0; j < ncand; ++j) {
          const int e = sh.incv[j] - 1;
          const double ss = sqrt(fmax(sh.G[j * ncand + j], 0.0));
          sh.rnorm[e] = ss;
          const double thr = sh.fp64 ? s.critr : fmax(s.critr, prm.switch_tol * fmax(1.0, fabs(sh.eigs[e])));
          if (ss < thr) {
            ++icvc;
            sh.icv[e] = 1;
            if (icvc < ncand) continue;
            all_conv = 1;
            break;
          }
          sh.add[nadded++] = j;
          if (nadded + kpass == limadd) break;
        }
        sh.done = all_conv;
        sh.nadded = nadded;
        for (int a = 0; a < nadded; ++a) sh.lam[a] = sh.eigs[sh.incv[sh.add[a]] - 1];
        for (int a = 0; a < nadded; ++a) sh.incv[a] = sh.incv[sh.add[a]];
        sh.nncv = nadded;
      }
      __syncthreads();
      if (sh.done) {
        if (tid == 0) sh.omin_to = 1;
        __syncthreads();
        if (!sh.fp64) {
          dwg_orthonormalise(sh, s, 0, sh.kpass, prm.ortho, 2);
          if (tid == 0) sh.kpass = sh.kept;
          __syncthreads();
          dwg_apply(sh, s, 0, sh.kpass, true);
          for (int o = tid; o < DWG_LIM * DWG_LIM; o += blockDim.x) sh.S[o] = 0.0;
          __syncthreads();
          dwg_update_S(sh, s, 0, sh.kpass);
          if (tid == 0) {
            sh.nncv = sh.kpass;
            for (int i = 0; i < nume; ++i) {
              sh.icv[i] = 0;
              sh.eigs[i] = 1e30;
            }
            sh.kv = 0;
            sh.fp64 = 1;
            sh.omin_to = 0;
            sh.first = 1;
            ++sh.nloops;
          }
          __syncthreads();
          if (sh.nloops >= s.maxiter) break;
          continue;
        }
        break;
      }
      // ---- new vectors: preconditioned residuals into B[:, kpass ..)
      const int nadded = sh.nadded;
      for (int64_t t = tid; t < n * nadded; t += blockDim.x) {
        const int a = (int)(t / n);
        const int64_t r = t % n;
        const double dg = sh.lam[a] - s.diag[r];
        const double v = s.R[r + (int64_t)sh.add[a] * n];
        s.B[r + (int64_t)(kpass + a) * n] = fabs(dg) > 1.0e-13 ? v / dg : v / 1.0e-13;
      }
      __syncthreads();
      dwg_orthonormalise(sh, s, kpass, nadded, prm.ortho, sh.fp64 ? 2 : 1);
      if (tid == 0) {
        sh.nncv = sh.kept;
        sh.done = 0;
        if (!sh.restart) sh.first = 1;
        else if (sh.first) {
          sh.first = 0;
          sh.done = 2;  // collapse to the Ritz vectors and restart
        } else {
          sh.ierr += 1024;
          sh.omin_to = 1;
          sh.done = 1;
        }
        if (!sh.done && sh.nncv == 0) {
          sh.omin_to = 1;
          sh.done = 3;  // nothing to add
        }
      }
      __syncthreads();
      if (sh.done == 1) break;
      if (sh.done == 2) {
        dwg_collapse(sh, s, kpass, nume);
        dwg_reset_S(sh, nume);
        if (tid == 0) {
          for (int i = 0; i < nume; ++i) sh.eigs[i] = 1e30;
          sh.kpass = nume;
          ++sh.restarts;
        }
        __syncthreads();
        continue;
      }
      if (sh.done == 3) {
        if (!sh.fp64) {
          dwg_orthonormalise(sh, s, 0, sh.kpass, prm.ortho, 2);
          if (tid == 0) sh.kpass = sh.kept;
          __syncthreads();
          dwg_apply(sh, s, 0, sh.kpass, true);
          for (int o = tid; o < DWG_LIM * DWG_LIM; o += blockDim.x) sh.S[o] = 0.0;
          __syncthreads();
          dwg_update_S(sh, s, 0, sh.kpass);
          if (tid == 0) {
            sh.nncv = sh.kpass;
            for (int i = 0; i < nume; ++i) {
              sh.icv[i] = 0;
              sh.eigs[i] = 1e30;
            }
            sh.kv = 0;
            sh.fp64 = 1;
            sh.omin_to = 0;
            sh.first = 1;
            ++sh.nloops;
          }
          __syncthreads();
          if (sh.nloops >= s.maxiter) break;
          continue;
        }
        break;
      }
      // ---- Add_Vectors_To_AB_S
      dwg_apply(sh, s, kpass, sh.nncv, sh.fp64);
      dwg_update_S(sh, s, kpass, sh.nncv);
      if (tid == 0) {
        sh.kpass = kpass + sh.nncv;
        ++sh.nloops;
      }
      __syncthreads();
      if (sh.nloops >= s.maxiter) break;
    }
    // ---- final Ritz vectors and residuals
    if (tid == 0) {
      if (!sh.omin_to) sh.kpass -= sh.nncv;
      for (int i = 0; i < nume; ++i) sh.old[i] = fabs(sh.old[i] - sh.eigs[i]);
    }
    __syncthreads();
    dwg_eigen(sh, sh.kpass);  // always rotate to the Ritz vectors (see the host driver)
    dwg_collapse(sh, s, sh.kpass, nume);
    for (int o = tid; o < nume * DWG_MB; o += blockDim.x) {
      const int k = o / DWG_MB, j = o % DWG_MB;
      sh.C1[o] = (k == j) ? 1.0 : 0.0;
      sh.C2[o] = (k == j) ? -sh.eigs[j] : 0.0;
    }
    __syncthreads();
    dwg_combine(sh, n, s.AB, s.B, nume, nume, s.R);
    dwg_gram(sh, n, s.R, nume, s.R, nume);
    if (tid == 0) {
      for (int i = 0; i < nume; ++i) sh.rnorm[i] = sqrt(fmax(sh.G[i * nume + i], 0.0));
      int reopen = 0;
      if (prm.strict && sh.fp64 && sh.nloops < s.maxiter && sh.ierr == 0 && verify < 20)
        for (int i = 0; i < nume; ++i)
          if (sh.rnorm[i] > s.critr) {
            sh.icv[i] = 0;
            reopen = 1;
          }
      sh.flag = reopen;
    }
    __syncthreads();
    if (!sh.flag) break;
    dwg_reset_S(sh, nume);
    if (tid == 0) {
      for (int i = 0; i < nume; ++i) sh.eigs[i] = 1e30;
      sh.kpass = nume;
      sh.nncv = nume;
      sh.omin_to = 0;
      sh.first = 1;
    }
    __syncthreads();
  }
  if (tid == 0) {
    for (int i = 0; i < nume; ++i) {
      s.evals[i] = sh.eigs[i];
      s.dE[i] = sh.old[i];
      s.resid[i] = sh.rnorm[i];
    }
    s.info[0] = sh.nloops;
    s.info[1] = sh.matvecs;
    s.info[2] = sh.matvecs64;
    s.info[3] = sh.restarts;
    s.info[4] = sh.ierr;
    // converged: stopped before maxiter and, under the strict test, every returned
    // residual below critr (an orthogonalisation failure or the reopen cap counts as not
    // converged)
    int conv = (sh.nloops >= s.maxiter && !sh.omin_to) ? 0 : 1;
    if (prm.strict)
      for (int i = 0; i < nume; ++i)
        if (!(sh.rnorm[i] <= s.critr)) conv = 0;
    s.info[5] = conv;
  }
}
