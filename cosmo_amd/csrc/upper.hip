// upper.hip -- upper-triangle rows (the content of a .HH or CSB file, or of
// cosmo_op_create_upper's arrays) -> the operator's full symmetric rows, on the GPU.
//
// Row r of the result holds the lower entries (r', r), r' < r, in ascending r', then its
// own upper entries in ascending column order: the same rows, bit for bit, as the
// transposition done on the host.  Lower entries reach their rows through atomic slot
// counters and are then sorted per row in LDS (their keys are unique), so the result does
// not depend on the order of the atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "op.h"

namespace cosmo {

namespace {

constexpr int SYM_SORT_CAP = 4096;  // lower entries of one row sorted in LDS

// one thread per row: checks the columns (in [r, n), strictly ascending), counts the
// row's live entries (the diagonal always, others unless |h| < 1e-14 is dropped) and adds
// one to the lower count of each off-diagonal column
__global__ __launch_bounds__(256) void k_sym_count(int64_t n, const int64_t *__restrict__ urp,
                                                   const int64_t *__restrict__ uc, const double *__restrict__ uv,
                                                   int keep, int32_t *__restrict__ upcnt, int32_t *__restrict__ lowcnt,
                                                   int *__restrict__ err) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  int live = 0;
  int64_t prev = r - 1;
  for (int64_t e = urp[r]; e < urp[r + 1]; ++e) {
    const int64_t c = uc[e];
    if (c <= prev || c >= n) {
      atomicOr(err, 1);
      return;
    }
    prev = c;
    if (c != r && !keep && fabs(uv[e]) < 1e-14) continue;
    ++live;
    if (c != r) atomicAdd(&lowcnt[c], 1);
  }
  upcnt[r] = live;
}

// one thread per row: lower entries into the rows of their columns (slots in any order),
// own upper entries after the lower part, diagonal out
__global__ __launch_bounds__(256) void k_sym_scatter(int64_t n, const int64_t *__restrict__ urp,
                                                     const int64_t *__restrict__ uc, const double *__restrict__ uv,
                                                     int keep, const int64_t *__restrict__ frp,
                                                     const int32_t *__restrict__ lowcnt, int32_t *__restrict__ lowfill,
                                                     uint32_t *__restrict__ fc, double *__restrict__ fv,
                                                     double *__restrict__ diag) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  int64_t q = frp[r] + lowcnt[r];
  double d = 0.0;
  for (int64_t e = urp[r]; e < urp[r + 1]; ++e) {
    const int64_t c = uc[e];
    const double v = uv[e];
    if (c == r) d = v;
    else if (!keep && fabs(v) < 1e-14) continue;
    fc[q] = (uint32_t)c;
    fv[q++] = v;
    if (c != r) {
      const int64_t s = frp[c] + atomicAdd(&lowfill[c], 1);
      fc[s] = (uint32_t)r;
      fv[s] = v;
    }
  }
  diag[r] = d;
}

// one workgroup per row (grid-stride): bitonic sort of the lower part by column
__global__ __launch_bounds__(256) void k_sym_sort(int64_t n, const int64_t *__restrict__ frp,
                                                  const int32_t *__restrict__ lowcnt, uint32_t *__restrict__ fc,
                                                  double *__restrict__ fv) {
  __shared__ uint32_t key[SYM_SORT_CAP];
  __shared__ double val[SYM_SORT_CAP];
  for (int64_t r = blockIdx.x; r < n; r += gridDim.x) {
    const int L = lowcnt[r];
    if (L < 2) continue;  // uniform across the workgroup
    const int64_t base = frp[r];
    int P = 2;
    while (P < L) P <<= 1;
    for (int i = threadIdx.x; i < P; i += 256) {
      key[i] = i < L ? fc[base + i] : 0xffffffffu;
      val[i] = i < L ? fv[base + i] : 0.0;
    }
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < P; i += 256) {
          const int l = i ^ j;
          if (l > i) {
            const bool up = (i & k) == 0;
            if ((key[i] > key[l]) == up) {
              const uint32_t tk = key[i];
              key[i] = key[l];
              key[l] = tk;
              const double tv = val[i];
              val[i] = val[l];
              val[l] = tv;
            }
          }
        }
        __syncthreads();
      }
    for (int i = threadIdx.x; i < L; i += 256) {
      fc[base + i] = key[i];
      fv[base + i] = val[i];
    }
    __syncthreads();  // LDS reused by the next row
  }
}

__global__ __launch_bounds__(256) void k_to_f32(int64_t cnt, const double *__restrict__ src, float *__restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * 256)
    dst[i] = (float)src[i];
}

}  // namespace

// Host -> device through two pinned staging buffers: the copy from pageable memory into
// one overlaps the DMA out of the other.
int h2d_staged(void *dst, const void *src, size_t bytes, hipStream_t st) {
  constexpr size_t CH = (size_t)32 << 20;
  if (bytes <= CH) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return COSMO_OK;
  }
  void *pin[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool used[2] = {false, false};
  int rc = COSMO_OK;
  for (int i = 0; i < 2 && !rc; ++i) {
    if (hipHostMalloc(&pin[i], CH, hipHostMallocDefault) != hipSuccess || hipEventCreate(&ev[i]) != hipSuccess)
      rc = fail(COSMO_ENOMEM, "pinned staging buffer");
  }
  size_t off = 0;
  for (int i = 0; !rc && off < bytes; off += CH, i ^= 1) {
    const size_t len = std::min(CH, bytes - off);
    if (used[i] && hipEventSynchronize(ev[i]) != hipSuccess) rc = fail(COSMO_EHIP, "staged upload wait");
    if (rc) break;
    std::memcpy(pin[i], (const char *)src + off, len);
    if (hipMemcpyAsync((char *)dst + off, pin[i], len, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipEventRecord(ev[i], st) != hipSuccess)
      rc = fail(COSMO_EHIP, "staged upload");
    used[i] = true;
  }
  if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = fail(COSMO_EHIP, "staged upload sync");
  for (int i = 0; i < 2; ++i) {
    if (ev[i]) (void)hipEventDestroy(ev[i]);
    if (pin[i]) (void)hipHostFree(pin[i]);
  }
  return rc;
}

// Full symmetric rows on the device from host upper rows; COSMO_EUNSUP (nothing changed)
// when a row's lower part exceeds the LDS sort, so the caller can use the host path.
int upper_to_full_dev(Op &op, int64_t n, const int64_t *rowptr, const int64_t *cols, const double *vals,
                      int flags, hipStream_t st) {
  const int keep = (flags & COSMO_OP_KEEP_SMALL) ? 1 : 0;
  const int64_t E = rowptr[n] - rowptr[0];
  if (rowptr[0] != 0) return fail(COSMO_EINVAL, "upper rows must start at offset 0");
  int64_t *urp = nullptr, *uc = nullptr;
  double *uv = nullptr;
  int32_t *upcnt = nullptr, *lowcnt = nullptr;
  int *err = nullptr;
  auto cleanup = [&]() {
    dfree(urp);
    dfree(uc);
    dfree(uv);
    dfree(upcnt);
    dfree(lowcnt);
    dfree(err);
  };
  if (dmalloc(&urp, sizeof(int64_t) * (n + 1)) != hipSuccess ||
      dmalloc(&uc, sizeof(int64_t) * std::max<int64_t>(E, 1)) != hipSuccess ||
      dmalloc(&uv, sizeof(double) * std::max<int64_t>(E, 1)) != hipSuccess ||
      dmalloc(&upcnt, sizeof(int32_t) * std::max<int64_t>(n, 1)) != hipSuccess ||
      dmalloc(&lowcnt, sizeof(int32_t) * std::max<int64_t>(n, 1)) != hipSuccess || dmalloc(&err, sizeof(int)) != hipSuccess) {
    cleanup();
    return fail(COSMO_ENOMEM, "upper rows: device buffers");
  }
  int rc = h2d_staged(urp, rowptr, sizeof(int64_t) * (n + 1), st);
  if (!rc) rc = h2d_staged(uc, cols, sizeof(int64_t) * E, st);
  if (!rc) rc = h2d_staged(uv, vals, sizeof(double) * E, st);
  if (rc) {
    cleanup();
    return rc;
  }
  const unsigned grid = (unsigned)std::max<int64_t>(1, (n + 255) / 256);
  // every queued call's status is folded into e; the synchronisation below reports it
  hipError_t e = hipMemsetAsync(lowcnt, 0, sizeof(int32_t) * std::max<int64_t>(n, 1), st);
  auto also = [&](hipError_t x) {
    if (e == hipSuccess) e = x;
  };
  also(hipMemsetAsync(err, 0, sizeof(int), st));
  if (n) hipLaunchKernelGGL(k_sym_count, dim3(grid), dim3(256), 0, st, n, urp, uc, uv, keep, upcnt, lowcnt, err);
  also(hipGetLastError());
  std::vector<int32_t> hu(n), hl(n);
  int herr = 0;
  also(hipMemcpyAsync(hu.data(), upcnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
  also(hipMemcpyAsync(hl.data(), lowcnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
  also(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st));
  also(hipStreamSynchronize(st));
  if (e != hipSuccess) {
    cleanup();
    return fail(COSMO_EHIP, "upper rows: count pass");
  }
  if (herr) {
    cleanup();
    return fail(COSMO_EINVAL, "upper rows must hold strictly ascending columns in [row, n)");
  }
  int maxlow = 0;
  for (int64_t r = 0; r < n; ++r) maxlow = std::max(maxlow, hl[r]);
  if (maxlow > SYM_SORT_CAP) {
    cleanup();
    return COSMO_EUNSUP;
  }
  std::vector<int64_t> frp(n + 1, 0);
  int64_t nup = 0;
  for (int64_t r = 0; r < n; ++r) {
    frp[r + 1] = frp[r] + hu[r] + hl[r];
    nup += hu[r];
  }
  const int64_t nnz = frp[n];
  CsrDev &M = op.csr;
  M.n = n;
  M.nnz = nnz;
  M.nnz_upper = nup;
  if (dmalloc(&M.rowptr, sizeof(int64_t) * (n + 1)) != hipSuccess ||
      dmalloc(&M.cols, sizeof(uint32_t) * std::max<int64_t>(nnz, 1)) != hipSuccess ||
      dmalloc(&M.v64, sizeof(double) * std::max<int64_t>(nnz, 1)) != hipSuccess ||
      dmalloc(&M.diag, sizeof(double) * std::max<int64_t>(n, 1)) != hipSuccess) {
    cleanup();
    return fail(COSMO_ENOMEM, "upper rows: operator arrays");
  }
  also(hipMemcpyAsync(M.rowptr, frp.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, st));
  also(hipMemsetAsync(upcnt, 0, sizeof(int32_t) * std::max<int64_t>(n, 1), st));  // reused as the lower fill counters
  if (n) {
    hipLaunchKernelGGL(k_sym_scatter, dim3(grid), dim3(256), 0, st, n, urp, uc, uv, keep, M.rowptr, lowcnt, upcnt,
                       M.cols, M.v64, M.diag);
    hipLaunchKernelGGL(k_sym_sort, dim3((unsigned)std::min<int64_t>(n, 65536)), dim3(256), 0, st, n, M.rowptr,
                       lowcnt, M.cols, M.v64);
  }
  if (!(flags & COSMO_OP_NO_FP32)) {
    if (dmalloc(&M.v32, sizeof(float) * std::max<int64_t>(nnz, 1)) != hipSuccess) {
      cleanup();
      return fail(COSMO_ENOMEM, "upper rows: FP32 values");
    }
    if (nnz) hipLaunchKernelGGL(k_to_f32, dim3(4096), dim3(256), 0, st, nnz, M.v64, M.v32);
  }
  also(hipGetLastError());
  op.hdiag.assign(n, 0.0);
  if (n) also(hipMemcpyAsync(op.hdiag.data(), M.diag, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  also(hipStreamSynchronize(st));
  cleanup();
  if (e != hipSuccess) return fail(COSMO_EHIP, std::string("upper rows: assembly: ") + hipGetErrorString(e));
  return COSMO_OK;
}

}  // namespace cosmo
