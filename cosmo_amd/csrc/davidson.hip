// davidson.hip -- GPU-resident block Davidson for the lowest eigenpairs of a cosmo_op.
//
// Control logic restates DVDSON as used by the reference (include/kdav/fsdavidson.cxx
// dvdrvr2 :615-775, Test_Convergence :212-281, Add_Vectors_To_Basis :426-516,
// Collapse_* :99-121/:517-539, setup_2d :540-611; defaults from
// include/kdav/cdavidson.cpp:86-110 and include/kdav/davidson.cxx:167-223).  The n-long
// work (H.X, Gram blocks, basis rotations, residuals, preconditioning) stays on the GPU;
// the (lim x lim) Rayleigh-Ritz problem is solved in FP64 on the host.
//
// Orthogonalisation is blocked: the new vectors V sit right after the basis B in storage,
// so one Gram pass gives both B^T V and V^T V; classical Gram-Schmidt sweeps (V - B W)
// repeat until the overlap is below the threshold, and a pivot-dropping Cholesky QR of
// V^T V - W^T W orthonormalises V among themselves; the whole is repeated once in the
// FP64 phase (not while H.X runs in FP32).  This spans the same space as the
// reference's per-vector repeated CGS (Orthogonalize_Basis :313-390) in a few passes
// over B instead of one per new vector.  Gram blocks run on the FP64 matrix cores.
//
// Mixed precision (COSMO_PREC_FP32): H.X runs on the FP32 values with FP32 vectors, and
// the basis, its image and the work blocks are stored in FP32 (all sums in FP64), until
// every wanted residual drops under mixed_switch * max(1, |lambda|); the basis is then
// collapsed to the wanted Ritz vectors, widened to FP64 and re-orthonormalised, H.X is
// recomputed in FP64 and the iteration continues in FP64 to the reference thresholds (the
// FP64 refinement).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "op.h"

namespace cosmo {
bool ritz_lowest(int k, const std::vector<double> &A0, int want, std::vector<double> &w, std::vector<double> &V);


#define GRAM_MAXP 128   // A columns per launch (8 output tiles of 16)
#define GRAM_MAXQ 16    // Bm columns per launch (one output tile)
#define GRAM_SLAB 32    // rows per wave per step: 4 lane groups x 8 rows
#define GRAM_LDS 33     // LDS column stride of a staged 32-row tile (doubles)
#define GRAM_BLOCKS 1024
#define GRAM_SMALL 8192   // doubles in the small device buffer (Gram results, coefficients)
#define GRAM_UPLOAD 4096  // doubles per pinned upload slot (two 128 x 16 coefficient blocks)

// partial[g][a*q+b] = sum over the rows that block g visits of A[:,a] * Bm[:,b], on the
// FP64 matrix cores (v_mfma_f64_16x16x4f64: 16 x 16 output tile, K = 4 rows per
// instruction).  Each wave takes 32-row slabs round-robin over the grid.  A slab of 16
// columns is read coalesced (32 consecutive rows of a column per half-wave), widened
// exactly to FP64 and staged in the wave's own LDS tile; lane (i, kg) = (lane % 16,
// lane / 16) then feeds column i at rows kg*8 + j, j = 0..7, as the 8 K steps.  The Bm
// tile is staged once per slab, the A tiles one after another.  The four waves of a block
// are summed in a fixed order through LDS, so the result does not depend on timing.
// Columns past p (q) are read clamped and their outputs dropped; rows past n read as zero.
typedef double gram_d4 __attribute__((ext_vector_type(4)));

// orders this wave's LDS accesses: every earlier LDS access has completed, and the
// compiler keeps memory accesses on their side
__device__ __forceinline__ void wave_lds_fence() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// 16 columns x 32 rows of a slab: lane (row, half) = (lane % 32, lane / 32) reads its row
// of columns 2s + half, s = 0..7 (two whole column segments per instruction).  p0 points
// at the lane's row in column 0 of the tile, p1 at the same row in column `half`; the
// column offsets are uniform, and columns past ncol read column ncol - 1 (or 0), unused.
template <typename V>
__device__ __forceinline__ void gram_load(V (&v)[8], const V *p0, const V *p1, int64_t ld, int ncol) {
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int c = 2 * s;
    const V *src = (c + 1 < ncol) ? p1 : p0;
    v[s] = src[(c < ncol) ? (int64_t)c * ld : 0];
  }
}
template <typename V>
__device__ __forceinline__ void gram_store(double *__restrict__ tile, const V (&v)[8], int64_t r0, int64_t n,
                                           int lane) {
  const int row = lane & 31, half = lane >> 5;
  const bool ok = r0 + row < n;
#pragma unroll
  for (int s = 0; s < 8; ++s) tile[(2 * s + half) * GRAM_LDS + row] = ok ? (double)v[s] : 0.0;
}

template <typename V, int PT>
__global__ __launch_bounds__(256) void k_gram(int64_t n, const V *__restrict__ A, int64_t lda, int p,
                                              const V *__restrict__ Bm, int64_t ldb, int q,
                                              double *__restrict__ partial) {
  __shared__ double red[PT * 256];
  __shared__ double stage[4][2][16 * GRAM_LDS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = lane & 15, kg = lane >> 4;
  double *tb = stage[wave][0], *ta = stage[wave][1];
  gram_d4 acc[PT];
#pragma unroll
  for (int t = 0; t < PT; ++t) acc[t] = gram_d4{0.0, 0.0, 0.0, 0.0};
  // software pipeline: the next A tile (and, during the last tile, the next slab's Bm and
  // first A tile) is in flight while the current one is multiplied; rows are clamped, so
  // a read past the last slab stays inside the arrays and is never used
  const int64_t nslab = (n + GRAM_SLAB - 1) / GRAM_SLAB, stride = (int64_t)gridDim.x * 4;
  int64_t s = (int64_t)blockIdx.x * 4 + wave;
  const int half = lane >> 5;
  auto row_of = [&](int64_t r0) { return min(r0 + (lane & 31), n - 1); };
  V vb[8], vc[8];
  {
    const int64_t r = row_of(s * GRAM_SLAB);
    gram_load<V>(vb, Bm + r, Bm + r + half * ldb, ldb, q);
    gram_load<V>(vc, A + r, A + r + half * lda, lda, p);
  }
  for (; s < nslab; s += stride) {
    const int64_t r0 = s * GRAM_SLAB, r = row_of(r0), rn = row_of(r0 + stride * GRAM_SLAB);
    wave_lds_fence();
    gram_store<V>(tb, vb, r0, n, lane);
#pragma unroll
    for (int t = 0; t < PT; ++t) {
      gram_store<V>(ta, vc, r0, n, lane);
      if (t + 1 < PT) {
        const V *p0 = A + (int64_t)(t + 1) * 16 * lda + r;
        gram_load<V>(vc, p0, p0 + half * lda, lda, p - (t + 1) * 16);
      } else {
        gram_load<V>(vb, Bm + rn, Bm + rn + half * ldb, ldb, q);
        gram_load<V>(vc, A + rn, A + rn + half * lda, lda, p);
      }
      wave_lds_fence();
#pragma unroll
      for (int j = 0; j < 8; ++j)
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ta[i * GRAM_LDS + kg * 8 + j], tb[i * GRAM_LDS + kg * 8 + j],
                                                      acc[t], 0, 0, 0);
      wave_lds_fence();
    }
  }
  // fixed-order block sum: ((w3 + w2) + w1) + w0
  for (int w = 3; w >= 1; --w) {
    if (wave == w) {
#pragma unroll
      for (int t = 0; t < PT; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          double x = acc[t][v];
          if (w < 3) x += red[t * 256 + v * 64 + lane];
          red[t * 256 + v * 64 + lane] = x;
        }
    }
    __syncthreads();
  }
  if (wave == 0) {
    const int pq = p * q;
    // the FP64 16x16 result puts tile row 4*v + kg, column i in register v of lane
    // (i, kg): element v of acc[t] is G[t*16 + 4*v + kg][i]
#pragma unroll
    for (int t = 0; t < PT; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int a = t * 16 + 4 * v + kg;
        if (a < p && i < q) partial[(int64_t)blockIdx.x * pq + a * q + i] = acc[t][v] + red[t * 256 + v * 64 + lane];
      }
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
template <int Q, typename V>
__global__ __launch_bounds__(256) void k_gemm(int64_t n, const V *__restrict__ A1, const V *__restrict__ A2,
                                              int64_t lda, int p, const double *__restrict__ C1,
                                              const double *__restrict__ C2, int q, V *__restrict__ Y,
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
    double a = (double)A1[(int64_t)k * lda + i];
#pragma unroll
    for (int j = 0; j < Q; ++j) acc[j] += a * c1[k * Q + j];
    if (A2) {
      double b = (double)A2[(int64_t)k * lda + i];
#pragma unroll
      for (int j = 0; j < Q; ++j) acc[j] += b * c2[k * Q + j];
    }
  }
#pragma unroll
  for (int j = 0; j < Q; ++j)
    if (j < q) Y[(int64_t)j * ldy + i] = (V)acc[j];
}

// X[:, j] -= sum_k B[:,k] W[k*q+j]
template <int Q, typename V>
__global__ __launch_bounds__(256) void k_project_out(int64_t n, const V *__restrict__ B, int64_t ldb, int p,
                                                     const double *__restrict__ W, int q, V *__restrict__ X,
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
    double b = (double)B[(int64_t)k * ldb + i];
#pragma unroll
    for (int j = 0; j < Q; ++j) acc[j] += b * w[k * Q + j];
  }
#pragma unroll
  for (int j = 0; j < Q; ++j)
    if (j < q) X[(int64_t)j * ldv + i] = (V)((double)X[(int64_t)j * ldv + i] - acc[j]);
}

// dst[0..count) = src[0..count): coefficients read by one block from the host's pinned
// buffer through its device mapping (a launch costs the host far less than a copy call)
__global__ __launch_bounds__(256) void k_fetch(const double *__restrict__ src, double *__restrict__ dst, int count) {
  for (int t = threadIdx.x; t < count; t += blockDim.x) dst[t] = src[t];
}

// Up to 16 column indices (and one value per column) passed by value to a launch.
struct ColList {
  int c[16];
  double s[16];
};

// dst[:, a] = src[:, L.c[a]] for a = blockIdx.y < 16 (column-major, ld n)
template <typename V>
__global__ void k_copy_cols(int64_t n, const V *__restrict__ src, ColList L, V *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a = blockIdx.y;
  dst[(int64_t)a * n + i] = src[(int64_t)L.c[a] * n + i];
}

// dst[:, a] = src[:, L.c[a]] / (L.s[a] - D_ii), |.| of the denominator floored at 1e-13
// (diagonal preconditioner, fsdavidson.cxx:498-510): the residual of pair a moved into
// the basis and preconditioned in one pass
template <typename V>
__global__ void k_precond_cols(int64_t n, const V *__restrict__ src, ColList L, V *__restrict__ dst,
                               const double *__restrict__ D) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a = blockIdx.y;
  const double dg = L.s[a] - D[i];
  const double r = (double)src[(int64_t)L.c[a] * n + i];
  dst[(int64_t)a * n + i] = (V)(fabs(dg) > 1.0e-13 ? r / dg : r / 1.0e-13);
}

// element-wise copy between the FP32 and FP64 basis storage (n x m, column-major, ld n)
template <typename D, typename S>
__global__ void k_convert(int64_t len, const S *__restrict__ src, D *__restrict__ dst) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < len) dst[t] = (D)src[t];
}

// column-major B[:, c0 .. c0+m) -> interleaved X (n x ldx), as double or float
// (columns m .. ldx-1 of X are zero padding)
template <typename T, typename V>
__global__ void k_cols_to_rows(int64_t n, int m, const V *__restrict__ B, int64_t ldb, T *__restrict__ X, int ldx) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * ldx) return;
  int64_t i = t / ldx;
  int j = (int)(t % ldx);
  X[i * ldx + j] = j < m ? (T)B[(int64_t)j * ldb + i] : (T)0;
}
template <typename T, typename V>
__global__ void k_rows_to_cols(int64_t n, int m, const T *__restrict__ Y, int ldy, V *__restrict__ B, int64_t ldb) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * m) return;
  int j = (int)(t / n);
  int64_t i = t % n;
  B[(int64_t)j * ldb + i] = (V)Y[i * ldy + j];
}

// Pinned, device-mapped host staging (Gram read-back, then two upload slots), one per
// running solve, from a library-wide pool: a solve takes one and gives it back when it
// ends.  Buffers are kept until the process ends (freeing pinned memory synchronises the
// device, which would stall concurrent solves), so the pool holds as many as solves ever
// ran at once, whichever threads ran them.  Coherent (fine-grained): the Gram sums the
// device writes are visible to the host as soon as the stream has drained.
static std::mutex g_stage_mu;
static std::vector<double *> g_stage_free;
static double *stage_acquire() {
  {
    std::lock_guard<std::mutex> lk(g_stage_mu);
    if (!g_stage_free.empty()) {
      double *p = g_stage_free.back();
      g_stage_free.pop_back();
      return p;
    }
  }
  double *p = nullptr;
  if (hipHostMalloc((void **)&p, sizeof(double) * (GRAM_SMALL + 2 * GRAM_UPLOAD),
                    hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
    return nullptr;
  return p;
}
static void stage_release(double *p) {
  std::lock_guard<std::mutex> lk(g_stage_mu);
  g_stage_free.push_back(p);
}

// ------------------------------------------------------------------------------------
struct Dav {
  Op *op = nullptr;
  hipStream_t st = nullptr;
  int64_t n = 0;
  int lim = 0, nume = 0, mblock = 0;
  double *B = nullptr, *AB = nullptr, *R = nullptr, *T = nullptr, *Xd = nullptr, *Yd = nullptr;
  float *Xf = nullptr, *Yf = nullptr;
  // While H.X runs in FP32 the basis, its H image and the work blocks are kept as FP32
  // copies (half the bytes for every pass over them; products still accumulate in FP64);
  // es is the element size of the current phase and Bv/ABv/Rv/Tv point at its storage.
  float *B32 = nullptr, *AB32 = nullptr, *R32 = nullptr, *T32 = nullptr;
  int es = 8;
  void *Bv = nullptr, *ABv = nullptr, *Rv = nullptr, *Tv = nullptr;
  double *partial = nullptr, *segs = nullptr, *small = nullptr;
  // pinned host staging (one buffer of the library-wide pool per running solve): GRAM_SMALL doubles of Gram read-back,
  // then two upload slots of GRAM_UPLOAD doubles used alternately
  double *hbuf = nullptr, *hdev = nullptr;  // hdev: the device's address of hbuf
  bool up_busy[2] = {false, false};
  int up_next = 0;
  int G = 0;
  double ms_apply = 0, ms_ortho = 0, ms_rr = 0;
  int matvecs = 0, matvecs64 = 0;
  int verbose = 0;
  ~Dav() {
    for (void *p : {(void *)B, (void *)AB, (void *)R, (void *)T, (void *)Xd, (void *)Yd, (void *)Xf, (void *)Yf,
                    (void *)B32, (void *)AB32, (void *)R32, (void *)T32, (void *)partial, (void *)segs,
                    (void *)small})
      dfree(p);
    if (hbuf) {
      // work queued before an early error return may still write the buffer
      if (st) (void)hipStreamSynchronize(st);
      stage_release(hbuf);
    }
  }
  double *col(double *base, int j) { return base + (int64_t)j * n; }
  void *cv(void *base, int j) const { return (char *)base + (int64_t)j * n * es; }
  void use64() {
    es = 8;
    Bv = B;
    ABv = AB;
    Rv = R;
    Tv = T;
  }
  void use32() {
    es = 4;
    Bv = B32;
    ABv = AB32;
    Rv = R32;
    Tv = T32;
  }
};

static int dav_alloc(Dav &D, bool mixed) {
  const int64_t n = D.n;
  const int nr = std::max(D.nume, D.mblock), nt = std::max(D.lim, 16);
  HIPCHK(dmalloc(&D.B, sizeof(double) * n * D.lim));
  HIPCHK(dmalloc(&D.AB, sizeof(double) * n * D.lim));
  HIPCHK(dmalloc(&D.R, sizeof(double) * n * nr));
  HIPCHK(dmalloc(&D.T, sizeof(double) * n * nt));
  if (mixed) {
    HIPCHK(dmalloc(&D.B32, sizeof(float) * n * D.lim));
    HIPCHK(dmalloc(&D.AB32, sizeof(float) * n * D.lim));
    HIPCHK(dmalloc(&D.R32, sizeof(float) * n * nr));
    HIPCHK(dmalloc(&D.T32, sizeof(float) * n * nt));
  }
  int ldmax = 4;
  while (ldmax < D.mblock) ldmax <<= 1;
  HIPCHK(dmalloc(&D.Xd, sizeof(double) * n * ldmax));
  HIPCHK(dmalloc(&D.Yd, sizeof(double) * n * ldmax));
  HIPCHK(dmalloc(&D.Xf, sizeof(float) * n * ldmax));
  HIPCHK(dmalloc(&D.Yf, sizeof(float) * n * ldmax));
  D.G = (int)std::min<int64_t>(GRAM_BLOCKS, std::max<int64_t>(1, (n + 4 * GRAM_SLAB - 1) / (4 * GRAM_SLAB)));
  HIPCHK(dmalloc(&D.partial, sizeof(double) * (size_t)D.G * GRAM_MAXP * GRAM_MAXQ));
  HIPCHK(dmalloc(&D.segs, sizeof(double) * (size_t)GRAM_SEG * GRAM_MAXP * GRAM_MAXQ));
  HIPCHK(dmalloc(&D.small, sizeof(double) * GRAM_SMALL));
  D.hbuf = stage_acquire();
  if (!D.hbuf) return fail(COSMO_ENOMEM, "pinned staging buffer: hipHostMalloc failed");
  HIPCHK(hipHostGetDevicePointer((void **)&D.hdev, D.hbuf, 0));
  HIPCHK(hipMemsetAsync(D.B, 0, sizeof(double) * n * D.lim, D.st));
  HIPCHK(hipMemsetAsync(D.AB, 0, sizeof(double) * n * D.lim, D.st));
  D.use64();
  return COSMO_OK;
}

// waits for the stream; every pinned upload slot is then free
static int dsync(Dav &D) {
  // poll rather than block: the Davidson loop waits on the device several times per
  // iteration, and a blocking wait adds tens of microseconds of wake-up latency each time
  hipError_t e;
  while ((e = hipStreamQuery(D.st)) == hipErrorNotReady) std::this_thread::yield();
  HIPCHK(e);
  D.up_busy[0] = D.up_busy[1] = false;
  return COSMO_OK;
}
#define DSYNC(D)                \
  do {                          \
    int rc_ = dsync(D);         \
    if (rc_) return rc_;        \
  } while (0)

// the pinned slot for the next host->device upload; waits for the stream only if the
// slot's previous upload may not have run yet (uploads need no synchronisation of their own)
static int upload_slot(Dav &D, double *&slot) {
  const int k = D.up_next;
  D.up_next ^= 1;
  if (D.up_busy[k]) DSYNC(D);
  D.up_busy[k] = true;
  slot = D.hbuf + GRAM_SMALL + (size_t)k * GRAM_UPLOAD;
  return COSMO_OK;
}
// queue the copy of `count` doubles of an upload slot (filled by the host) to D.small
static int upload(Dav &D, const double *slot, size_t count) {
  hipLaunchKernelGGL(k_fetch, dim3(1), dim3(256), 0, D.st, D.hdev + (slot - D.hbuf), D.small, (int)count);
  HIPCHK(hipGetLastError());
  return COSMO_OK;
}

// copy m basis columns between the FP64 and FP32 storage (dst/src element types by es)
static int convert_cols(Dav &D, const void *src, int src_es, void *dst, int m) {
  const int64_t len = D.n * m;
  if (!len) return COSMO_OK;
  dim3 g((unsigned)((len + 255) / 256)), b(256);
  if (src_es == 8) hipLaunchKernelGGL((k_convert<float, double>), g, b, 0, D.st, len, (const double *)src, (float *)dst);
  else hipLaunchKernelGGL((k_convert<double, float>), g, b, 0, D.st, len, (const float *)src, (double *)dst);
  HIPCHK(hipGetLastError());
  return COSMO_OK;
}

template <typename V>
static void launch_gram(Dav &D, const V *A, int pp, const V *Bm, int qq) {
  dim3 g(D.G), b(256);
  switch ((pp + 15) / 16) {
    case 1: hipLaunchKernelGGL((k_gram<V, 1>), g, b, 0, D.st, D.n, A, D.n, pp, Bm, D.n, qq, D.partial); break;
    case 2: hipLaunchKernelGGL((k_gram<V, 2>), g, b, 0, D.st, D.n, A, D.n, pp, Bm, D.n, qq, D.partial); break;
    case 3: hipLaunchKernelGGL((k_gram<V, 3>), g, b, 0, D.st, D.n, A, D.n, pp, Bm, D.n, qq, D.partial); break;
    case 4: hipLaunchKernelGGL((k_gram<V, 4>), g, b, 0, D.st, D.n, A, D.n, pp, Bm, D.n, qq, D.partial); break;
    case 5: hipLaunchKernelGGL((k_gram<V, 5>), g, b, 0, D.st, D.n, A, D.n, pp, Bm, D.n, qq, D.partial); break;
    case 6: hipLaunchKernelGGL((k_gram<V, 6>), g, b, 0, D.st, D.n, A, D.n, pp, Bm, D.n, qq, D.partial); break;
    case 7: hipLaunchKernelGGL((k_gram<V, 7>), g, b, 0, D.st, D.n, A, D.n, pp, Bm, D.n, qq, D.partial); break;
    default: hipLaunchKernelGGL((k_gram<V, 8>), g, b, 0, D.st, D.n, A, D.n, pp, Bm, D.n, qq, D.partial); break;
  }
}

// out (p x q, row-major a*q+b) = A[:, 0..p)^T Bm[:, 0..q), host copy; A, Bm in the
// current phase's storage.  Launches cover at most 128 x 16 outputs; their sums are
// gathered in the small device buffer and copied back with one synchronisation per fill.
static int gram(Dav &D, const void *A, int p, const void *Bm, int q, std::vector<double> &out) {
  out.assign((size_t)p * q, 0.0);
  if (!p || !q) return COSMO_OK;
  struct Piece {
    int a0, b0, pp, qq;
    size_t off;
  };
  std::vector<Piece> pend;
  size_t used = 0;
  auto flush = [&]() -> int {
    if (pend.empty()) return COSMO_OK;
    DSYNC(D);  // the sums were written straight into the pinned buffer
    for (const Piece &c : pend)
      for (int a = 0; a < c.pp; ++a)
        for (int b = 0; b < c.qq; ++b) out[(size_t)(c.a0 + a) * q + c.b0 + b] = D.hbuf[c.off + (size_t)a * c.qq + b];
    pend.clear();
    used = 0;
    return COSMO_OK;
  };
  for (int b0 = 0; b0 < q; b0 += GRAM_MAXQ) {
    const int qq = std::min(GRAM_MAXQ, q - b0);
    const void *Bb = (const char *)Bm + (int64_t)b0 * D.n * D.es;
    for (int a0 = 0; a0 < p; a0 += GRAM_MAXP) {
      const int pp = std::min(GRAM_MAXP, p - a0), pq = pp * qq;
      if (used + pq > GRAM_SMALL) {
        int rc = flush();
        if (rc) return rc;
      }
      const void *Aa = (const char *)A + (int64_t)a0 * D.n * D.es;
      if (D.es == 4) launch_gram<float>(D, (const float *)Aa, pp, (const float *)Bb, qq);
      else launch_gram<double>(D, (const double *)Aa, pp, (const double *)Bb, qq);
      hipLaunchKernelGGL(k_reduce_partial, dim3((pq + 255) / 256, GRAM_SEG), dim3(256), 0, D.st, D.G, pq, D.partial,
                         D.segs);
      hipLaunchKernelGGL(k_reduce_segments, dim3((pq + 255) / 256), dim3(256), 0, D.st, pq, D.segs, D.hdev + used);
      HIPCHK(hipGetLastError());
      pend.push_back({a0, b0, pp, qq, used});
      used += pq;
    }
  }
  return flush();
}

template <typename V>
static void launch_gemm(Dav &D, int qq, const V *A1, const V *A2, int p, const double *s2, V *Yj) {
  dim3 g((unsigned)((D.n + 255) / 256)), b(256);
  if (qq <= 4) hipLaunchKernelGGL((k_gemm<4, V>), g, b, 0, D.st, D.n, A1, A2, D.n, p, D.small, s2, qq, Yj, D.n);
  else if (qq <= 8) hipLaunchKernelGGL((k_gemm<8, V>), g, b, 0, D.st, D.n, A1, A2, D.n, p, D.small, s2, qq, Yj, D.n);
  else hipLaunchKernelGGL((k_gemm<16, V>), g, b, 0, D.st, D.n, A1, A2, D.n, p, D.small, s2, qq, Yj, D.n);
}

// Y[:, 0..q) = A1 C1 (+ A2 C2); C row-major p x q on host; q <= 16 per launch
static int gemm(Dav &D, const void *A1, const void *A2, int p, const std::vector<double> &C1,
                const std::vector<double> *C2, int q, void *Y) {
  if (p > 128) return fail(COSMO_EUNSUP, "basis wider than 128 vectors");
  for (int j0 = 0; j0 < q; j0 += 16) {
    // C1 then C2 chunks, packed in the pinned buffer and sent in one copy
    const int qq = std::min(16, q - j0);
    const size_t pc = (size_t)p * qq;
    double *c1;
    int rc = upload_slot(D, c1);
    if (rc) return rc;
    double *c2 = c1 + pc;
    for (int k = 0; k < p; ++k)
      for (int j = 0; j < qq; ++j) c1[k * qq + j] = C1[(size_t)k * q + j0 + j];
    if (C2)
      for (int k = 0; k < p; ++k)
        for (int j = 0; j < qq; ++j) c2[k * qq + j] = (*C2)[(size_t)k * q + j0 + j];
    rc = upload(D, c1, pc * (C2 ? 2 : 1));
    if (rc) return rc;
    const double *s2 = C2 ? D.small + pc : nullptr;
    void *Yj = D.cv(Y, j0);
    if (D.es == 4)
      launch_gemm<float>(D, qq, (const float *)A1, C2 ? (const float *)A2 : nullptr, p, s2, (float *)Yj);
    else
      launch_gemm<double>(D, qq, (const double *)A1, C2 ? (const double *)A2 : nullptr, p, s2, (double *)Yj);
    HIPCHK(hipGetLastError());
  }
  return COSMO_OK;
}

template <typename V>
static void launch_project(Dav &D, int qq, const V *Bk, int pp, V *Vj) {
  dim3 g((unsigned)((D.n + 255) / 256)), b(256);
  if (qq <= 4) hipLaunchKernelGGL((k_project_out<4, V>), g, b, 0, D.st, D.n, Bk, D.n, pp, D.small, qq, Vj, D.n);
  else if (qq <= 8) hipLaunchKernelGGL((k_project_out<8, V>), g, b, 0, D.st, D.n, Bk, D.n, pp, D.small, qq, Vj, D.n);
  else hipLaunchKernelGGL((k_project_out<16, V>), g, b, 0, D.st, D.n, Bk, D.n, pp, D.small, qq, Vj, D.n);
}

static int project_out(Dav &D, const void *Bm, int p, const std::vector<double> &W, int q, void *V) {
  if (!p || !q) return COSMO_OK;
  for (int j0 = 0; j0 < q; j0 += 16) {
    int qq = std::min(16, q - j0);
    for (int k0 = 0; k0 < p; k0 += 128) {
      int pp = std::min(128, p - k0);
      double *w;
      int rc = upload_slot(D, w);
      if (rc) return rc;
      for (int k = 0; k < pp; ++k)
        for (int j = 0; j < qq; ++j) w[k * qq + j] = W[(size_t)(k0 + k) * q + j0 + j];
      rc = upload(D, w, (size_t)pp * qq);
      if (rc) return rc;
      const void *Bk = (const char *)Bm + (int64_t)k0 * D.n * D.es;
      void *Vj = D.cv(V, j0);
      if (D.es == 4) launch_project<float>(D, qq, (const float *)Bk, pp, (float *)Vj);
      else launch_project<double>(D, qq, (const double *)Bk, pp, (double *)Vj);
      HIPCHK(hipGetLastError());
    }
  }
  return COSMO_OK;
}

// AB[:, c0 .. c0+m) = H B[:, c0 .. c0+m) in the current phase's storage; prec must match
// it (FP32 storage with the FP32 product, FP64 storage with the FP64 product)
static int apply_block(Dav &D, int c0, int m, int prec) {
  if ((prec == COSMO_PREC_FP32) != (D.es == 4)) return fail(COSMO_EINVAL, "apply_block: storage/precision mismatch");
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
      hipLaunchKernelGGL((k_cols_to_rows<float, float>), gp, b, 0, D.st, D.n, mm, (const float *)D.cv(D.Bv, c0 + j0),
                         D.n, D.Xf, ld);
      rc = op_apply_dev(*D.op, D.Xf, D.Yf, mm, ld, ld, COSMO_PREC_FP32, D.st);
      if (rc) return rc;
      hipLaunchKernelGGL((k_rows_to_cols<float, float>), g, b, 0, D.st, D.n, mm, D.Yf, ld,
                         (float *)D.cv(D.ABv, c0 + j0), D.n);
    } else {
      hipLaunchKernelGGL((k_cols_to_rows<double, double>), gp, b, 0, D.st, D.n, mm,
                         (const double *)D.cv(D.Bv, c0 + j0), D.n, D.Xd, ld);
      rc = op_apply_dev(*D.op, D.Xd, D.Yd, mm, ld, ld, COSMO_PREC_FP64, D.st);
      if (rc) return rc;
      hipLaunchKernelGGL((k_rows_to_cols<double, double>), g, b, 0, D.st, D.n, mm, D.Yd, ld,
                         (double *)D.cv(D.ABv, c0 + j0), D.n);
      D.matvecs64 += mm;
    }
    HIPCHK(hipGetLastError());
    D.matvecs += mm;
  }
  DSYNC(D);  // keeps the stage times honest (one short wait per block)
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

// Orthonormalise V (q columns at Vp, right after B[:, 0..p)) against B and among
// themselves, in `rounds` rounds of Gram-Schmidt passes plus Cholesky QR (two for FP64
// accuracy; one while H.V itself is FP32, whose vectors are only good to single
// precision).  All in
// the current phase's storage; with FP32 storage the thresholds are those of single
// precision (overlap 3e-6, a dependent vector keeps < 1e-5 of its norm).
// Returns the number of kept vectors (compacted to the front); restart=true when the
// overlap with the basis does not decrease (the reference's restart condition).
static int orthonormalise(Dav &D, int p, void *Vp, int q, double ortho, bool &restart, int &kept,
                          int rounds = 2) {
  const bool f32 = D.es == 4;
  if (f32) ortho = std::max(ortho, 3e-6);
  const double drop_rel = f32 ? 1e-5 : 1e-10;
  auto t0 = std::chrono::steady_clock::now();
  restart = false;
  kept = q;
  // V follows the basis in storage, so one Gram pass gives both B^T V and V^T V; all
  // thresholds are relative to the column norms (the preconditioner can scale a vector
  // by up to 1e13), so no separate normalisation pass is needed
  if (q > 0 && Vp != D.cv(D.Bv, p)) return fail(COSMO_EINVAL, "orthonormalise: vectors must follow the basis");
  std::vector<double> G, W, Gv;
  for (int round = 0; round < rounds && kept > 0; ++round) {
    const int k = kept;
    double dprev = 1e7;
    std::vector<double> base(k, 0.0);  // column norms at the start of the round
    for (int pass = 0; pass < 4; ++pass) {
      int rc = gram(D, D.Bv, p + k, Vp, k, G);
      if (rc) return rc;
      W.assign(G.begin(), G.begin() + (size_t)p * k);
      Gv.assign(G.begin() + (size_t)p * k, G.end());
      if (pass == 0)
        for (int j = 0; j < k; ++j) base[j] = std::sqrt(std::max(Gv[(size_t)j * k + j], 0.0));
      if (!p) break;
      double dcur = 0.0;
      for (int j = 0; j < k; ++j) {
        const double nv = std::sqrt(std::max(Gv[(size_t)j * k + j], 0.0));
        if (nv > 0)
          for (int a = 0; a < p; ++a) dcur = std::max(dcur, std::fabs(W[(size_t)a * k + j]) / nv);
      }
      rc = project_out(D, D.Bv, p, W, k, Vp);
      if (rc) return rc;
      // V^T V after the projection, B orthonormal: V^T V - W^T W
      for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) {
          double s = 0.0;
          for (int a = 0; a < p; ++a) s += W[(size_t)a * k + i] * W[(size_t)a * k + j];
          Gv[(size_t)i * k + j] -= s;
        }
      if (dcur < ortho) break;
      if (dcur > dprev) {
        if (D.verbose > 1) fprintf(stderr, "davidson ortho: restart at pass %d overlap %.3e > %.3e\n", pass, dcur, dprev);
        restart = true;
        break;
      }
      dprev = dcur;
    }
    if (restart) break;
    // Cholesky QR of V^T V in order, dropping vectors whose remaining norm is below 1e-14
    // of their norm at the start of the round, or below drop_rel of the projected norm
    std::vector<double> Rm((size_t)k * k, 0.0);
    std::vector<int> keep;
    for (int j = 0; j < k; ++j) {
      // r_ij = (G_ij - sum_{l<i} r_li r_lj) / r_ii over kept i
      double d = Gv[(size_t)j * k + j];
      for (int ii : keep) {
        double s = Gv[(size_t)ii * k + j];
        for (int l : keep) {
          if (l == ii) break;
          s -= Rm[l * k + ii] * Rm[l * k + j];
        }
        s /= Rm[ii * k + ii];
        Rm[ii * k + j] = s;
        d -= s * s;
      }
      double nrm = d > 0 ? std::sqrt(d) : 0.0;
      double orig = std::sqrt(std::max(Gv[(size_t)j * k + j], 0.0));
      if (nrm <= 1e-14 * base[j] || nrm < drop_rel * orig) {
        if (D.verbose > 1)
          fprintf(stderr, "davidson ortho: drop vector %d remaining %.3e start %.3e projected %.3e\n", j, nrm, base[j],
                  orig);
        continue;
      }
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
    for (int a0 = 0; a0 < kk; a0 += 16) {
      ColList L;
      const int m = std::min(16, kk - a0);
      for (int a = 0; a < m; ++a) L.c[a] = keep[a0 + a];
      dim3 g((unsigned)((D.n + 255) / 256), m);
      if (D.es == 4)
        hipLaunchKernelGGL(k_copy_cols<float>, g, dim3(256), 0, D.st, D.n, (const float *)Vp, L,
                           (float *)D.cv(D.Tv, a0));
      else
        hipLaunchKernelGGL(k_copy_cols<double>, g, dim3(256), 0, D.st, D.n, (const double *)Vp, L,
                           (double *)D.cv(D.Tv, a0));
      HIPCHK(hipGetLastError());
    }
    if (kk) {
      int rc = gemm(D, D.Tv, nullptr, kk, Tinv, nullptr, kk, Vp);
      if (rc) return rc;
    }
    kept = kk;
  }
  DSYNC(D);
  D.ms_ortho += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return COSMO_OK;
}

// B[:, 0..nume) = B[:, 0..kp) * Svec^T (Svec rows = Ritz vectors of length >= kp); same for AB
static int collapse(Dav &D, const std::vector<std::vector<double>> &svec, int kp) {
  const int q = (int)svec.size();
  std::vector<double> C((size_t)kp * q);
  for (int k = 0; k < kp; ++k)
    for (int j = 0; j < q; ++j) C[(size_t)k * q + j] = k < (int)svec[j].size() ? svec[j][k] : 0.0;
  int rc = gemm(D, D.Bv, nullptr, kp, C, nullptr, q, D.Tv);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(D.Bv, D.Tv, (size_t)D.es * D.n * q, hipMemcpyDeviceToDevice, D.st));
  rc = gemm(D, D.ABv, nullptr, kp, C, nullptr, q, D.Tv);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(D.ABv, D.Tv, (size_t)D.es * D.n * q, hipMemcpyDeviceToDevice, D.st));
  DSYNC(D);
  return COSMO_OK;
}

// packed lower triangle of S (s[i*(i+1)/2 + j], j <= i) -> eigenpairs (DiagonalizeS :21-98)
static void diagonalize_S(const std::vector<double> &s, int k, int nume, std::vector<double> &eigs,
                          std::vector<std::vector<double>> &svec) {
  std::vector<double> A((size_t)k * k);
  for (int i = 0; i < k; ++i)
    for (int j = 0; j <= i; ++j) A[i * k + j] = A[j * k + i] = s[(size_t)i * (i + 1) / 2 + j];
  // only the nume lowest pairs are used: Householder + tridiagonal QR eigenvalues + inverse
  // iteration (ritz.cpp); cyclic Jacobi on the whole matrix if that ever fails
  const int want = std::min(nume, k);
  std::vector<double> w, V;
  svec.assign(nume, std::vector<double>(k, 0.0));
  if (ritz_lowest(k, A, want, w, V)) {
    for (int e = 0; e < want; ++e) {
      eigs[e] = w[e];
      for (int r = 0; r < k; ++r) svec[e][r] = V[(size_t)e * k + r];
    }
  } else {
    sym_eig(k, A, w, V);
    for (int e = 0; e < want; ++e) {
      eigs[e] = w[e];
      for (int r = 0; r < k; ++r) svec[e][r] = V[(size_t)e * k + r];
    }
  }
}

// S columns for basis vectors [c0, c0+m): s[j*(j+1)/2 + i] = b_i . ab_j, i <= j
static int update_S(Dav &D, std::vector<double> &s, int c0, int m) {
  auto t0 = std::chrono::steady_clock::now();
  std::vector<double> W;
  int rc = gram(D, D.Bv, c0 + m, D.cv(D.ABv, c0), m, W);
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
  D.verbose = p.verbose;
  int rc = dav_alloc(D, mixed);
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
    DSYNC(D);
    int added = nume - niv;
    if (niv > 0) {
      bool rs;
      rc = orthonormalise(D, niv, D.col(D.B, niv), nume - niv, 1e-9, rs, added);
      if (rc) return rc;
    }
    niv += added;
  }
  int prec = mixed ? COSMO_PREC_FP32 : COSMO_PREC_FP64;
  if (mixed) {  // the FP32 phase works on FP32 copies of the basis
    rc = convert_cols(D, D.B, 8, D.B32, niv);
    if (rc) return rc;
    D.use32();
  }
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
  // mixed precision: widen the basis to FP64, recompute H.B in FP64, continue in FP64
  auto switch_fp64 = [&]() -> int {
    // collapse the FP32-grown basis to the nume wanted Ritz vectors (H.X is recomputed in
    // FP64 for these only: on the fp 4p4n 10-pair solve 319 FP64 products instead of 372,
    // 0.98 s instead of 1.02 s; keeping the whole basis or 32 / 48 vectors was slower), re-
    // orthonormalise them in FP64, then recompute H.B and S in FP64
    if (kpass > nume) {
      int r0 = collapse(D, svec, kpass);
      if (r0) return r0;
      kpass = nume;
    }
    int r = convert_cols(D, D.B32, 4, D.B, kpass);
    if (r) return r;
    D.use64();
    bool rs0;
    int kept0 = kpass;
    r = orthonormalise(D, 0, D.B, kpass, p.ortho, rs0, kept0, 2);
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
      if (p.verbose > 1) fprintf(stderr, "davidson: converged by eigenvalue/coefficient test at it %d\n", nloops);
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
      rc = gemm(D, D.ABv, D.Bv, kpass, C1, &C2, ncand, D.Rv);
      if (rc) return rc;
      std::vector<double> G;
      rc = gram(D, D.Rv, ncand, D.Rv, ncand, G);
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
      if (p.verbose > 1) {
        fprintf(stderr, "davidson it %d kpass %d candidates", nloops, kpass);
        for (int j = 0; j < ncand; ++j) fprintf(stderr, " %d:%.2e", incv[j] - 1, rnorm[incv[j] - 1]);
        fprintf(stderr, "\n");
      }
      nncv = nadded;
      for (int a0 = 0; a0 < nadded; a0 += 16) {
        ColList L;
        const int m = std::min(16, nadded - a0);
        for (int a = 0; a < m; ++a) {
          int j = addidx[a0 + a];
          L.c[a] = j;
          L.s[a] = eigs[incv[j] - 1];
        }
        dim3 g((unsigned)((n + 255) / 256), m);
        if (D.es == 4)
          hipLaunchKernelGGL(k_precond_cols<float>, g, dim3(256), 0, D.st, n, (const float *)D.Rv, L,
                             (float *)D.cv(D.Bv, kpass + a0), dD);
        else
          hipLaunchKernelGGL(k_precond_cols<double>, g, dim3(256), 0, D.st, n, (const double *)D.Rv, L,
                             (double *)D.cv(D.Bv, kpass + a0), dD);
        HIPCHK(hipGetLastError());
      }
      for (int a = 0; a < nadded; ++a) incv[a] = incv[addidx[a]];
      bool restart = false;
      int kept = nadded;
      rc = orthonormalise(D, kpass, D.cv(D.Bv, kpass), nadded, p.ortho, restart, kept, fp64_phase ? 2 : 1);
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
        if (p.verbose > 1) fprintf(stderr, "davidson: second restart in a row at it %d (ierr 1024)\n", nloops);
        ierr += 1024;
        omin_to = true;
        break;
      }
      if (nncv == 0) {
        if (p.verbose > 1) fprintf(stderr, "davidson: every candidate dependent at it %d\n", nloops);
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
  if (D.es == 4) {  // stopped inside the FP32 phase: finish on FP64 copies of its basis
    rc = convert_cols(D, D.B32, 4, D.B, kpass);
    if (!rc) rc = convert_cols(D, D.AB32, 4, D.AB, kpass);
    if (rc) return rc;
    D.use64();
  }
  if (!omin_to) kpass -= nncv;
  for (int i = 0; i < nume; ++i) old[i] = std::fabs(old[i] - eigs[i]);
  {
    // refresh the Ritz vectors for the final kpass (S may have grown before the exit).
    // The reference skips this when kpass == nume (fsdavidson.cxx:755-758), which leaves
    // the start vectors in place of the Ritz vectors when nume == n; rotating always is
    // exact when the basis already is the Ritz basis (S diagonal after a collapse).
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
  if (p.verbose > 1) {
    fprintf(stderr, "davidson: exit check it %d verify %d ierr %d residuals", nloops, verify, ierr);
    for (int i = 0; i < nume; ++i) fprintf(stderr, " %.2e", rnorm[i]);
    fprintf(stderr, "%s\n", reopen ? " (reopened)" : "");
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
  DSYNC(D);
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
  if (p.strict) {
    // every exit (orthogonalisation failure, every candidate dependent, the reopen cap)
    // reports the pairs whose final residual is not below critr
    int bad = 0;
    double worst = 0.0;
    for (int i = 0; i < nume; ++i)
      if (!(rnorm[i] <= p.critr)) {
        ++bad;
        worst = std::max(worst, rnorm[i]);
      }
    if (bad)
      return fail(COSMO_ECONV, "Davidson stopped with " + std::to_string(bad) + " pair(s) above critr (worst residual " +
                                   std::to_string(worst) + ", ierr " + std::to_string(ierr) + ")");
  }
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
