// dav_wg.h -- block Davidson for batched solves, all of them in one persistent launch.
//
// Every solve is run to completion by a team of G workgroups (G = 1 for the small ones).
// Workgroups take seats in the launch's seat order (solves by decreasing team size), each
// seat being one rank of one solve, and take the next seat when their solve ends.  Seats
// are handed out in order, so every solve whose seats have all been taken has all of its
// workgroups running: a team waits for its last members only until earlier solves end,
// and the launch cannot deadlock while the largest team fits the device at once.
//
// The iteration restates the host driver (cosmo_davidson in davidson.hip, itself dvdrvr2 of
// include/kdav/fsdavidson.cxx:615-775) step for step: unit start vectors at the lowest
// diagonal elements, the eigenvalue-change and Ritz-coefficient convergence test and the
// candidate order (Test_Convergence :212-281), residuals and the diagonal preconditioner
// (Add_Vectors_To_Basis :426-516, floor 1e-13), blocked Gram-Schmidt passes with a
// pivot-dropping Cholesky QR, restarts by collapse (:99-121, :517-539), the switch from the
// FP32 product to FP64 refinement, and the strict re-opening of pairs whose residual
// drifted.  Basis, image and residual blocks are FP64 and column-major (ld n).
//
// Rows are split among the team in contiguous ranges, and every n-long step is row-local
// except three, which exchange data through global memory between team barriers: H.X (the
// block to multiply is staged row-major, each workgroup writing its own rows, and gathered
// by all), the Gram blocks (one partial block per workgroup, summed by every workgroup in
// rank order, so that all hold the same bits) and the choice of start vectors (each
// workgroup's lowest diagonal elements, merged by every workgroup).  Each workgroup then
// solves the same (kpass x kpass) Rayleigh-Ritz problem in its own LDS by parallel cyclic
// Jacobi, warm-started from the previous Ritz vectors, and takes the same decisions (thread
// 0, published through LDS), so control flow is uniform over the team and every barrier is
// reached by all of its threads.  Staging and partial buffers alternate between two copies
// with the barrier count, so one barrier per exchange suffices.  A team barrier that waits
// longer than the launch's limit marks the solve failed instead of waiting on.
//
// Results depend on the team size and on the lanes per row of H.X (fixed functions of the
// system), never on the other solves of the launch or on timing.

#include "jacobi.h"

#define DWG_THREADS 512
#define DWG_LIM 48                    // basis limit (lim) bound
#define DWG_MB 16                     // nume and mblock bound
#define DWG_GMAX 128                  // team size bound
#define DWG_NMAX ((int64_t)1 << 30)   // dimension bound (row numbers of the start vectors are int)
#define DWG_PART (DWG_LIM * DWG_MB)   // doubles of one workgroup's partial block
#define DWG_ABORTED 2048              // ierr bit: a team barrier timed out

struct DwgSolve {
  int64_t n;
  const int64_t *rowptr;  // full symmetric rows (both triangles), as the stored operator
  const uint32_t *cols;
  const double *v64;
  const float *v32;
  const double *diag;
  double *B, *AB, *R;  // n x lim each, column-major
  double *X;           // two staging buffers of n x DWG_MB doubles (the block to multiply)
  double *part;        // two buffers of G x DWG_PART doubles (the team's partial blocks)
  unsigned *bar;       // team barrier counter, zero at launch
  unsigned *claim;     // the launch's seat counter
  int G;               // workgroups in the team
  int lanes;           // lanes per row in H.X: a power of two, 1..64
  int seat0;           // first seat of the solve; seats ascend with the array index
  int nume, lim, mblock, maxiter;
  double critr;
  double *evals, *dE, *resid;  // nume each
  int *info;                   // iterations, matvecs, matvecs_fp64, restarts, ierr, converged
};

struct DwgParams {
  double crite, critc, ortho, switch_tol;
  int strict, mixed;
  uint64_t spin_ticks;  // longest wait at a team barrier once the team is seated, wall-clock ticks
};

struct DwgShared {
  double S[DWG_LIM * DWG_LIM];       // Rayleigh matrix B^T H B (full, row-major)
  double A[DWG_LIM * DWG_LIM];       // Jacobi matrix; R factor of the Cholesky QR
  double V[DWG_LIM * DWG_LIM];       // Ritz vectors in the basis: V[r * DWG_LIM + e]
  double W2[2 * DWG_LIM * DWG_LIM];  // Jacobi second buffers / T = R^-1
  double G[DWG_LIM * DWG_MB];        // Gram results, row-major (P x M)
  double Wt[DWG_LIM * DWG_LIM];      // overlaps B^T V of an orthogonalisation pass
  double C1[DWG_LIM * DWG_MB], C2[DWG_LIM * DWG_MB];
  double eigs[DWG_LIM], old[DWG_MB], rnorm[DWG_MB], lam[DWG_MB], base[DWG_LIM];
  double jc[DWG_LIM / 2], js[DWG_LIM / 2];
  double minv[DWG_THREADS];
  int mini[DWG_THREADS];
  int partner[DWG_LIM], pairidx[DWG_LIM], order[DWG_LIM], keep[DWG_LIM];
  int icv[DWG_MB], incv[DWG_MB], add[DWG_MB], chosen[DWG_MB];
  DwgSolve sol;   // the solve this workgroup runs
  DwgParams prm;
  // team: own rows [r0, r1), rank, barriers passed, timed out
  int64_t r0, r1;
  int rank, phase, aborted, seat;
  // control state, written by thread 0 between barriers
  int kpass, nncv, kv, fp64, first, omin_to, nloops, restarts, ierr, verify;
  int done, flag, kept, restart, ncand, nadded, matvecs, matvecs64, rot;
  double dprev, dctl;
};

// ---- team building blocks (every thread of every workgroup of the team calls them) -------

// Team barrier: returns once all G workgroups have arrived (the writes before it visible to
// the whole team), or with sh.aborted set after waiting prm.spin_ticks from the moment every
// rank of the team held its seat (until then teammates may still be finishing earlier
// solves); later calls of an aborted workgroup wait no more.
__device__ void dwg_sync(DwgShared &sh, const DwgSolve &s, const DwgParams &prm) {
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s.G > 1 && !sh.aborted) {
      __threadfence();
      const unsigned target = (unsigned)s.G * (unsigned)(sh.phase + 1);
      const unsigned seated = (unsigned)(s.seat0 + s.G);
      atomicAdd(s.bar, 1u);
      bool all_seated = false;
      uint64_t t0 = 0;
      while (__hip_atomic_load(s.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (!all_seated) {
          all_seated = __hip_atomic_load(s.claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= seated;
          t0 = wall_clock64();
        } else if (wall_clock64() - t0 > prm.spin_ticks) {
          sh.aborted = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      __threadfence();
    }
    ++sh.phase;
  }
  __syncthreads();
}

// Sum of 16 values over the wave's 64 lanes by recursive halving: afterwards lane l with
// l % 4 == 0 holds in v[0] the total of value 8 b5 + 4 b4 + 2 b3 + b2 (bits of l).
__device__ inline void dwg_wave_sum16(double (&v)[16], int lane) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const bool hi = lane & 32;
    const double snd = hi ? v[i] : v[i + 8], kp = hi ? v[i + 8] : v[i];
    v[i] = kp + __shfl_xor(snd, 32);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const bool hi = lane & 16;
    const double snd = hi ? v[i] : v[i + 4], kp = hi ? v[i + 4] : v[i];
    v[i] = kp + __shfl_xor(snd, 16);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const bool hi = lane & 8;
    const double snd = hi ? v[i] : v[i + 2], kp = hi ? v[i + 2] : v[i];
    v[i] = kp + __shfl_xor(snd, 8);
  }
  {
    const bool hi = lane & 4;
    const double snd = hi ? v[0] : v[1], kp = hi ? v[1] : v[0];
    v[0] = kp + __shfl_xor(snd, 4);
  }
  v[0] += __shfl_xor(v[0], 2);
  v[0] += __shfl_xor(v[0], 1);
}

// sh.G[a * M + b] = sum over all n rows of X[r + a n] Y[r + b n], a < P <= DWG_LIM,
// b < M <= DWG_MB.  A wave owns a 4 x 4 output block over the workgroup's rows (lanes on
// consecutive rows); the team's partial blocks are added in rank order.
__device__ __attribute__((noinline)) void dwg_gram(DwgShared &sh, const DwgSolve &s, const DwgParams &prm, const double *X, int P,
                         const double *Y, int M) {
  if (P <= 0 || M <= 0) return;  // uniform: every rank sees the same P, M
  const int nbA = (P + 3) / 4, nbB = (M + 3) / 4, nblk = nbA * nbB;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int64_t n = s.n, r0 = sh.r0, r1 = sh.r1;
  const int par = sh.phase & 1;
  double *dst = s.G > 1 ? s.part + ((int64_t)par * s.G + sh.rank) * DWG_PART : sh.G;
  for (int blk = wave; blk < nblk; blk += nwave) {
    const int a0 = (blk / nbB) * 4, b0 = (blk % nbB) * 4;
    const double *xa[4], *yb[4];
    for (int i = 0; i < 4; ++i) {
      xa[i] = X + (int64_t)min(a0 + i, P - 1) * n;
      yb[i] = Y + (int64_t)min(b0 + i, M - 1) * n;
    }
    double acc[16];
    for (int i = 0; i < 16; ++i) acc[i] = 0.0;
#pragma unroll 2
    for (int64_t r = r0 + lane; r < r1; r += 64) {
      double x[4], y[4];
      for (int i = 0; i < 4; ++i) {
        x[i] = xa[i][r];
        y[i] = yb[i][r];
      }
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) acc[i * 4 + j] += x[i] * y[j];
    }
    dwg_wave_sum16(acc, lane);
    if ((lane & 3) == 0) {
      const int ij = ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 + ((lane >> 2) & 1);
      const int a = a0 + ij / 4, b = b0 + ij % 4;
      if (a < P && b < M) dst[a * M + b] = acc[0];
    }
  }
  dwg_sync(sh, s, prm);
  if (s.G > 1) {
    const double *src = s.part + (int64_t)par * s.G * DWG_PART;
    for (int o = threadIdx.x; o < P * M; o += blockDim.x) {
      double t = 0.0;
      for (int g = 0; g < s.G; ++g) t += src[(int64_t)g * DWG_PART + o];
      sh.G[o] = t;
    }
    __syncthreads();
  }
}

// own rows of Y[:, j] = X1 C1[:, j] (+ X2 C2[:, j]) for j < q <= DWG_MB, k < p; C row-major
// p x DWG_MB in LDS.  Row by row, all reads of a row before its writes, so Y may be X1 or X2
// when q <= p covers only columns already read (callers use distinct arrays or one chunk).
__device__ __attribute__((noinline)) void dwg_combine(const DwgShared &sh, const DwgSolve &s, const double *X1, const double *X2, int p,
                            int q, double *Yc) {
  const int64_t n = s.n, r1 = sh.r1;
#pragma unroll 1
  for (int64_t r = sh.r0 + threadIdx.x; r < r1; r += blockDim.x) {
    double acc[DWG_MB];
    for (int j = 0; j < DWG_MB; ++j) acc[j] = 0.0;
    int k = 0;
#pragma unroll 1
    for (; k + 4 <= p; k += 4) {  // four columns' loads in flight
      double a[4], b[4];
      for (int u = 0; u < 4; ++u) a[u] = X1[r + (int64_t)(k + u) * n];
      if (X2)
        for (int u = 0; u < 4; ++u) b[u] = X2[r + (int64_t)(k + u) * n];
      for (int u = 0; u < 4; ++u) {
        for (int j = 0; j < DWG_MB; ++j) acc[j] += a[u] * sh.C1[(k + u) * DWG_MB + j];
        if (X2)
          for (int j = 0; j < DWG_MB; ++j) acc[j] += b[u] * sh.C2[(k + u) * DWG_MB + j];
      }
    }
#pragma unroll 1
    for (; k < p; ++k) {
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

// own rows of dst[:, 0..q) = src[:, 0..q)
__device__ __attribute__((noinline)) void dwg_copy_cols(const DwgShared &sh, const DwgSolve &s, double *dst, const double *src, int q) {
  const int64_t n = s.n, r0 = sh.r0, m = sh.r1 - sh.r0;
  for (int64_t t = threadIdx.x; t < m * q; t += blockDim.x) {
    const int64_t o = (t / m) * n + r0 + t % m;
    dst[o] = src[o];
  }
  __syncthreads();
}

// own rows of the staging buffer: xs[r * PW + j] = B[r + (c + j) n] for j < mm, 0 up to PW
template <typename T, int PW>
__device__ __attribute__((noinline)) void dwg_stage(const DwgShared &sh, const DwgSolve &s, T *xs, int c, int mm) {
  const int64_t n = s.n, r0 = sh.r0, m = sh.r1 - sh.r0;
  for (int64_t t = threadIdx.x; t < m * PW; t += blockDim.x) {
    const int64_t r = r0 + t / PW;
    const int j = (int)(t % PW);
    xs[r * PW + j] = j < mm ? (T)s.B[r + (int64_t)(c + j) * n] : (T)0;
  }
}

// own rows of AB[:, c .. c+mm) = H (staged block): s.lanes lanes per row, each taking every
// lanes-th entry, then summed across the lanes by butterfly (all lanes end with the same
// bits); lane j % lanes writes column j
template <typename T, int PW>
__device__ __attribute__((noinline)) void dwg_hx_rows(const DwgShared &sh, const DwgSolve &s, const T *xs, const T *vals, int c, int mm) {
  const int L = s.lanes;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int rpw = 64 / L, sub = lane & (L - 1);
  const int64_t n = s.n, r1 = sh.r1;
#pragma unroll 1
  for (int64_t rb = sh.r0 + (int64_t)wave * rpw; rb < r1; rb += (int64_t)nwave * rpw) {
    const int64_t r = rb + lane / L;
    int64_t e0 = 0, e1 = 0;
    if (r < r1) {
      e0 = s.rowptr[r];
      e1 = s.rowptr[r + 1];
    }
    T acc[PW];
    for (int j = 0; j < PW; ++j) acc[j] = (T)0;
    // U entries per step, all their loads issued before the first product: the gathers
    // are latency-bound with one workgroup per CU
    constexpr int U = PW * sizeof(T) <= 16 ? 8 : (PW * sizeof(T) <= 32 ? 4 : 2);
    int64_t e = e0 + sub;
#pragma unroll 1
    for (; e + (U - 1) * L < e1; e += U * L) {
      T h[U], x[U][PW];
      uint32_t c[U];
      for (int u = 0; u < U; ++u) {
        h[u] = vals[e + u * L];
        c[u] = s.cols[e + u * L];
      }
      for (int u = 0; u < U; ++u)
        for (int j = 0; j < PW; ++j) x[u][j] = xs[(int64_t)c[u] * PW + j];
      for (int u = 0; u < U; ++u)
        for (int j = 0; j < PW; ++j) acc[j] += h[u] * x[u][j];
    }
#pragma unroll 1
    for (; e < e1; e += L) {
      const T hh = vals[e];
      const T *xx = xs + (int64_t)s.cols[e] * PW;
      for (int j = 0; j < PW; ++j) acc[j] += hh * xx[j];
    }
#pragma unroll 1
    for (int m = L >> 1; m >= 1; m >>= 1)
      for (int j = 0; j < PW; ++j) acc[j] += __shfl_xor(acc[j], m);
    if (r < r1)
      for (int j = 0; j < PW; ++j)
        if (j < mm && (j & (L - 1)) == sub) s.AB[r + (int64_t)(c + j) * n] = (double)acc[j];
  }
}

template <typename T, int PW>
__device__ void dwg_apply_chunk(DwgShared &sh, const DwgSolve &s, const DwgParams &prm, const T *vals, int c,
                                int mm) {
  T *xs = (T *)(s.X + (int64_t)(sh.phase & 1) * s.n * DWG_MB);
  dwg_stage<T, PW>(sh, s, xs, c, mm);
  dwg_sync(sh, s, prm);
  dwg_hx_rows<T, PW>(sh, s, xs, vals, c, mm);
  __syncthreads();
}

// AB[:, c0 .. c0+m) = H B[:, c0 .. c0+m), in chunks of DWG_MB columns
__device__ void dwg_apply(DwgShared &sh, const DwgSolve &s, const DwgParams &prm, int c0, int m, bool fp64) {
  for (int j0 = 0; j0 < m; j0 += DWG_MB) {
    const int mm = min(DWG_MB, m - j0), c = c0 + j0;
    if (fp64) {
      if (mm <= 4) dwg_apply_chunk<double, 4>(sh, s, prm, s.v64, c, mm);
      else if (mm <= 8) dwg_apply_chunk<double, 8>(sh, s, prm, s.v64, c, mm);
      else dwg_apply_chunk<double, 16>(sh, s, prm, s.v64, c, mm);
    } else {
      if (mm <= 4) dwg_apply_chunk<float, 4>(sh, s, prm, s.v32, c, mm);
      else if (mm <= 8) dwg_apply_chunk<float, 8>(sh, s, prm, s.v32, c, mm);
      else dwg_apply_chunk<float, 16>(sh, s, prm, s.v32, c, mm);
    }
  }
  if (threadIdx.x == 0) {
    sh.matvecs += m;
    if (fp64) sh.matvecs64 += m;
  }
  __syncthreads();
}

// S rows/columns of basis vectors [c0, c0+m): S[i][j] = b_i . ab_j, i < c0+m
__device__ void dwg_update_S(DwgShared &sh, const DwgSolve &s, const DwgParams &prm, int c0, int m) {
  for (int j0 = 0; j0 < m; j0 += DWG_MB) {
    const int mm = min(DWG_MB, m - j0);
    dwg_gram(sh, s, prm, s.B, c0 + j0 + mm, s.AB + (int64_t)(c0 + j0) * s.n, mm);
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

__device__ __attribute__((noinline)) void dwg_eigen(DwgShared &sh, int k) { lds_jacobi<DWG_LIM>(sh, k); }

// component r of Ritz vector i (in the current basis)
__device__ inline double dwg_svec(const DwgShared &sh, int i, int r) { return sh.V[r * DWG_LIM + sh.order[i]]; }

// B[:, 0..q) = B[:, 0..kp) svec^T and the same for AB (q = nume Ritz vectors)
__device__ void dwg_collapse(DwgShared &sh, const DwgSolve &s, int kp, int q) {
  for (int o = threadIdx.x; o < kp * DWG_MB; o += blockDim.x) {
    const int k = o / DWG_MB, j = o % DWG_MB;
    sh.C1[o] = j < q ? dwg_svec(sh, j, k) : 0.0;
  }
  __syncthreads();
  dwg_combine(sh, s, s.B, nullptr, kp, q, s.R);
  dwg_copy_cols(sh, s, s.B, s.R, q);
  dwg_combine(sh, s, s.AB, nullptr, kp, q, s.R);
  dwg_copy_cols(sh, s, s.AB, s.R, q);
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
__device__ __attribute__((noinline)) void dwg_orthonormalise(DwgShared &sh, const DwgSolve &s, const DwgParams &prm, int p, int q,
                                   int rounds) {
  const int tid = threadIdx.x;
  const int64_t n = s.n, r0 = sh.r0, r1 = sh.r1;
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
        dwg_gram(sh, s, prm, s.B, p + k, Vp + (int64_t)j0 * n, mm);
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
        sh.flag = dcur < prm.ortho ? 1 : (dcur > sh.dprev ? 2 : 0);
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
#pragma unroll 1
        for (int64_t r = r0 + tid; r < r1; r += blockDim.x) {
          double acc[DWG_MB];
          for (int j = 0; j < DWG_MB; ++j) acc[j] = 0.0;
          int a = 0;
#pragma unroll 1
          for (; a + 4 <= p; a += 4) {  // four columns' loads in flight
            double b[4];
            for (int u = 0; u < 4; ++u) b[u] = s.B[r + (int64_t)(a + u) * n];
            for (int u = 0; u < 4; ++u)
              for (int j = 0; j < DWG_MB; ++j) acc[j] += b[u] * sh.C1[(a + u) * DWG_MB + j];
          }
#pragma unroll 1
          for (; a < p; ++a) {
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
#pragma unroll 1
      for (int64_t r = r0 + tid; r < r1; r += blockDim.x) {
        double acc[DWG_MB];
        for (int j = 0; j < DWG_MB; ++j) acc[j] = 0.0;
#pragma unroll 1
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
    dwg_copy_cols(sh, s, Vp, s.R, kk);
  }
}

// Start vectors: the nume lowest diagonal elements (ties: lower index), chosen by every
// rank from the team's per-rank lists; sh.chosen.
__device__ void dwg_start_rows(DwgShared &sh, const DwgSolve &s, const DwgParams &prm) {
  const int tid = threadIdx.x, nume = s.nume;
  const int64_t r0 = sh.r0, r1 = sh.r1;
  for (int c = 0; c < nume; ++c) {  // own rows: repeated selection
    double best = 1e308;
    int bi = 0x7fffffff;
    for (int64_t r = r0 + tid; r < r1; r += blockDim.x) {
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
      sh.base[c] = best;
    }
    __syncthreads();
  }
  if (s.G == 1) return;
  const int par = sh.phase & 1;
  double *mine = s.part + ((int64_t)par * s.G + sh.rank) * DWG_PART;
  for (int c = tid; c < nume; c += blockDim.x) {
    mine[2 * c] = sh.base[c];
    mine[2 * c + 1] = (double)sh.chosen[c];
  }
  dwg_sync(sh, s, prm);
  if (tid == 0) {
    const double *all = s.part + (int64_t)par * s.G * DWG_PART;
    for (int c = 0; c < nume; ++c) {
      double best = 1e308;
      int bi = 0x7fffffff;
      for (int g = 0; g < s.G; ++g)
        for (int l = 0; l < nume; ++l) {
          const double v = all[(int64_t)g * DWG_PART + 2 * l];
          const int i = (int)all[(int64_t)g * DWG_PART + 2 * l + 1];
          if (i == 0x7fffffff) continue;
          bool taken = false;
          for (int u = 0; u < c; ++u) taken |= (sh.add[u] == i);
          if (!taken && (v < best || (v == best && i < bi))) {
            best = v;
            bi = i;
          }
        }
      sh.add[c] = bi;
    }
    for (int c = 0; c < nume; ++c) sh.chosen[c] = sh.add[c];
  }
  __syncthreads();
}

// switch from the FP32 product to FP64 refinement: collapse the basis to the wanted Ritz
// vectors (current in sh.V at every switch point), re-orthonormalise them, H.B and S in
// FP64; the caller breaks out when the iteration cap is reached
__device__ void dwg_to_fp64(DwgShared &sh, const DwgSolve &s, const DwgParams &prm) {
  const int tid = threadIdx.x;
  if (sh.kpass > s.nume) {
    dwg_collapse(sh, s, sh.kpass, s.nume);
    if (tid == 0) sh.kpass = s.nume;
    __syncthreads();
  }
  dwg_orthonormalise(sh, s, prm, 0, sh.kpass, 2);
  if (tid == 0) sh.kpass = sh.kept;
  __syncthreads();
  dwg_apply(sh, s, prm, 0, sh.kpass, true);
  for (int o = tid; o < DWG_LIM * DWG_LIM; o += blockDim.x) sh.S[o] = 0.0;
  __syncthreads();
  dwg_update_S(sh, s, prm, 0, sh.kpass);
  if (tid == 0) {
    sh.nncv = sh.kpass;
    for (int i = 0; i < s.nume; ++i) {
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
}

// the whole solve, as rank sh.rank of its team (sh.r0, sh.r1, sh.phase, sh.aborted set)
__device__ void dwg_solve(DwgShared &sh, const DwgSolve &s, const DwgParams &prm) {
  const int tid = threadIdx.x;
  const int64_t n = s.n, r0 = sh.r0, r1 = sh.r1;
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
  __syncthreads();
  dwg_start_rows(sh, s, prm);
  for (int c = 0; c < nume; ++c)
    for (int64_t r = r0 + tid; r < r1; r += blockDim.x) s.B[r + (int64_t)c * n] = (r == sh.chosen[c]) ? 1.0 : 0.0;
  __syncthreads();
  dwg_apply(sh, s, prm, 0, nume, sh.fp64);
  for (int o = tid; o < DWG_LIM * DWG_LIM; o += blockDim.x) sh.S[o] = 0.0;
  __syncthreads();
  dwg_update_S(sh, s, prm, 0, nume);
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
  for (int verify = 0; !sh.aborted; ++verify) {
    while (!sh.aborted) {
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
          dwg_to_fp64(sh, s, prm);
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
      dwg_combine(sh, s, s.AB, s.B, kpass, ncand, s.R);
      dwg_gram(sh, s, prm, s.R, ncand, s.R, ncand);
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
          dwg_to_fp64(sh, s, prm);
          if (sh.nloops >= s.maxiter) break;
          continue;
        }
        break;
      }
      // ---- new vectors: preconditioned residuals into B[:, kpass ..)
      const int nadded = sh.nadded;
      for (int a = 0; a < nadded; ++a) {
        const double lam = sh.lam[a];
        const double *Ra = s.R + (int64_t)sh.add[a] * n;
        double *Ba = s.B + (int64_t)(kpass + a) * n;
        for (int64_t r = r0 + tid; r < r1; r += blockDim.x) {
          const double dg = lam - s.diag[r];
          const double v = Ra[r];
          Ba[r] = fabs(dg) > 1.0e-13 ? v / dg : v / 1.0e-13;
        }
      }
      __syncthreads();
      dwg_orthonormalise(sh, s, prm, kpass, nadded, sh.fp64 ? 2 : 1);
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
          dwg_to_fp64(sh, s, prm);
          if (sh.nloops >= s.maxiter) break;
          continue;
        }
        break;
      }
      // ---- Add_Vectors_To_AB_S
      dwg_apply(sh, s, prm, kpass, sh.nncv, sh.fp64);
      dwg_update_S(sh, s, prm, kpass, sh.nncv);
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
    dwg_combine(sh, s, s.AB, s.B, nume, nume, s.R);
    dwg_gram(sh, s, prm, s.R, nume, s.R, nume);
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
  if (tid == 0 && sh.rank == 0) {
    if (sh.aborted) sh.ierr |= DWG_ABORTED;
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
    // residual below critr (an orthogonalisation failure, the reopen cap or a timed-out
    // team barrier counts as not converged)
    int conv = (sh.nloops >= s.maxiter && !sh.omin_to) ? 0 : 1;
    if (prm.strict)
      for (int i = 0; i < nume; ++i)
        if (!(sh.rnorm[i] <= s.critr)) conv = 0;
    if (sh.aborted) conv = 0;
    s.info[5] = conv;
  }
}

// Persistent loop of one workgroup: take the next seat (one atomic per workgroup), run that
// rank of its solve, until the seats run out.  solves[] ascends in seat0.
__device__ void dwg_run(DwgShared &sh, const DwgSolve *__restrict__ solves, int nsolves, int nseats,
                        unsigned *claim, const DwgParams &prm) {
  if (threadIdx.x == 0) sh.prm = prm;
  for (;;) {
    if (threadIdx.x == 0) sh.seat = (int)atomicAdd(claim, 1u);
    __syncthreads();
    const int seat = sh.seat;
    if (seat >= nseats) return;
    int lo = 0, hi = nsolves - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (solves[mid].seat0 <= seat) lo = mid;
      else hi = mid - 1;
    }
    if (threadIdx.x == 0) {
      sh.sol = solves[lo];
      const DwgSolve &s = sh.sol;
      const int64_t chunk = ((s.n + s.G - 1) / s.G + 63) / 64 * 64;
      sh.rank = seat - s.seat0;
      sh.r0 = min(s.n, (int64_t)sh.rank * chunk);
      sh.r1 = min(s.n, sh.r0 + chunk);
      sh.phase = 0;
      sh.aborted = 0;
    }
    __syncthreads();
    dwg_solve(sh, sh.sol, sh.prm);
    __syncthreads();
  }
}

__global__ __launch_bounds__(DWG_THREADS) void k_dav_wg(const DwgSolve *__restrict__ solves, int nsolves, int nseats,
                                                         unsigned *claim, DwgParams prm) {
  __shared__ DwgShared sh;
  dwg_run(sh, solves, nsolves, nseats, claim, prm);
}
