// hamiltonian.hip -- GPU Hamiltonian builder, stored H.X and on-the-fly ("virtual") H.X.
//
// A basis state is a 64-bit occupation mask (bit g = orbital g; neutrons first).  One
// wavefront owns one bra row and generates every hop of the one- and two-body operators
// from it, in both directions, so each row holds the full symmetric row of H:
//   * two-body moves (a,b)->(c,d) with {c,d} empty: the single forward term of the pair,
//     or, for kets below the bra, the forward term of the ket applied to the ket;
//   * one-body moves a->c (same species, a spectator s carries the two-body term): the
//     sum over spectators s in ascending order -- the reference's term order, since the
//     keys (sort(a,s), sort(c,s)) ascend with s (include/sm/cioperators.cxx:46-58 walks
//     the operator map in key order, include/tav/stensor.cxx:59-79);
//   * the diagonal: one-body terms over occupied orbitals, then pair terms in key order.
// Phases are parities of occupied orbitals passed, exactly the running counts of
// Annihilation / Creation (include/sm/operators.cxx:121-187), here as popcounts.
// Ket indices come from rank tables instead of a binary search over the basis
// (include/sm/ManyBodyStates.cxx:92-130): index = offset(neutron config) + rank of the
// proton config within its (M, parity) sector.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "device.h"
#include "op.h"

namespace cosmo {

__constant__ uint32_t c_binom[33][33];

// Device and stream are per host thread: independent solves driven from several threads
// (batches, the server's workers) each run on their own default-priority stream and
// overlap on the device.
// A thread's stream comes from a per-device pool and goes back to it (drained) when the
// thread releases it; library streams are never destroyed, so a stream handle recorded
// by an operator always names the same stream, and short-lived worker threads (one set
// per batch call) do not create and destroy streams each time.
static std::mutex g_mu;
static thread_local int t_device = 0;
static thread_local hipStream_t t_stream[64] = {};
static bool g_binom_done[64] = {};
static std::vector<hipStream_t> g_stream_pool[64];

hipStream_t lib_stream() {
  int d = t_device;
  if (!t_stream[d]) {
    {
      std::lock_guard<std::mutex> lk(g_mu);
      if (!g_stream_pool[d].empty()) {
        t_stream[d] = g_stream_pool[d].back();
        g_stream_pool[d].pop_back();
      }
    }
    if (!t_stream[d]) {
      hipSetDevice(d);
      hipStreamCreateWithFlags(&t_stream[d], hipStreamNonBlocking);
    }
  }
  return t_stream[d];
}

static bool g_pool_done[64] = {};

hipError_t dmalloc_void(void **p, size_t bytes, hipStream_t st) {
  const int d = t_device;
  if (!g_pool_done[d]) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_pool_done[d]) {
      hipMemPool_t pool;
      if (hipDeviceGetDefaultMemPool(&pool, d) == hipSuccess) {
        uint64_t keep = UINT64_MAX;  // freed blocks stay in the pool for the next solve
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
      }
      g_pool_done[d] = true;
    }
  }
  return hipMallocAsync(p, bytes ? bytes : 1, st ? st : lib_stream());
}

void dfree(void *p) {
  if (p) (void)hipFreeAsync(p, lib_stream());
}

void release_thread_streams() {
  for (int d = 0; d < 64; ++d)
    if (t_stream[d]) {
      hipSetDevice(d);
      hipStreamSynchronize(t_stream[d]);
      std::lock_guard<std::mutex> lk(g_mu);
      g_stream_pool[d].push_back(t_stream[d]);
      t_stream[d] = nullptr;
    }
}

static int ensure_binom() {
  std::lock_guard<std::mutex> lk(g_mu);
  int d = t_device;
  if (g_binom_done[d]) return COSMO_OK;
  uint32_t b[33][33];
  for (int n = 0; n <= 32; ++n)
    for (int k = 0; k <= 32; ++k) b[n][k] = (uint32_t)binom(n, k);
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_binom), b, sizeof(b)));
  HIPCHK(hipDeviceSynchronize());
  g_binom_done[d] = true;
  return COSMO_OK;
}

// ------------------------------------------------------------------------------------
// device helpers
__device__ __forceinline__ uint32_t colex32(uint32_t m) {
  uint32_t r = 0;
  int i = 1;
  while (m) {
    int p = __ffs(m) - 1;
    r += c_binom[p][i];
    ++i;
    m &= m - 1;
  }
  return r;
}

__device__ __forceinline__ int pid_of(int a, int b) { return b * (b - 1) / 2 + a; }  // a < b

__device__ __forceinline__ uint64_t below(int p) { return (p >= 64) ? ~0ull : ((1ull << p) - 1ull); }

// parity of the two-body term (annihilate p0<p1, create p2<p3) acting on state S
__device__ __forceinline__ int phase2(uint64_t S, int p0, int p1, int p2, int p3) {
  int ph = __popcll(S & below(p0)) + __popcll(S & below(p1)) - 1;
  uint64_t I = S & ~(1ull << p0) & ~(1ull << p1);
  ph += __popcll(I & below(p2)) + __popcll(I & below(p3));
  return ph & 1;
}
// parity of the one-body term (annihilate a, create c) acting on S
__device__ __forceinline__ int phase1(uint64_t S, int a, int c) {
  int ph = __popcll(S & below(a));
  uint64_t I = S & ~(1ull << a);
  ph += __popcll(I & below(c));
  return ph & 1;
}

__device__ __forceinline__ bool vexists(const DevTables &T, int pa, int pc) {
  int64_t k = (int64_t)pa * T.P + pc;
  return (T.vexist[k >> 5] >> (k & 31)) & 1u;
}

// basis index of state K, or -1 when K is outside the basis
__device__ __forceinline__ int64_t ket_index(const DevTables &T, uint64_t K) {
  uint32_t nm = (uint32_t)(K & ((T.nn >= 64) ? ~0ull : ((1ull << T.nn) - 1ull)));
  uint32_t pm = (uint32_t)(K >> T.nn);
  if (__popc(nm) != T.nv || __popc(pm) != T.zv) return -1;  // species not conserved
  int ni = T.ncfg_index[colex32(nm)];
  if (ni < 0) return -1;
  uint32_t pc = colex32(pm);
  if (T.pcfg_sec[pc] != T.nsector[ni]) return -1;
  const int32_t r = T.pcfg_rank[pc];
  if (!T.krank) return T.noff[ni] + r;
  // truncated basis: the position of rank r among the block's kept ranks, if kept
  int64_t lo = T.kstart[ni], hi = T.kstart[ni + 1];
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (T.krank[mid] < r) lo = mid + 1;
    else hi = mid;
  }
  if (lo == T.kstart[ni + 1] || T.krank[lo] != r) return -1;
  return T.noff[ni] + (lo - T.kstart[ni]);
}

__device__ __forceinline__ uint64_t bra_state(const DevTables &T, int64_t row) {
  int lo = 0, hi = T.nnconf - 1;
  while (lo < hi) {  // last block with noff <= row
    int mid = (lo + hi + 1) >> 1;
    if (T.noff[mid] <= row) lo = mid; else hi = mid - 1;
  }
  int sec = T.nsector[lo];
  const int64_t r = T.krank ? T.krank[T.kstart[lo] + (row - T.noff[lo])] : row - T.noff[lo];
  uint32_t pm = T.pmask[T.psec_start[sec] + r];
  return (uint64_t)T.nmask[lo] | ((uint64_t)pm << T.nn);
}

#define MAXPART 24
#define MAXPAIRS (MAXPART * (MAXPART - 1) / 2)

struct RowScratch {
  int occ[64];
  int pa[MAXPAIRS];
  int fpref[MAXPAIRS + 1];
  int rpref[MAXPAIRS + 1];
  int opref[MAXPART + 1];
  int lpref[MAXPART + 1];
  int N, NP;
};

// Generate every entry of the full symmetric row of bra state B (row index `row`).
// Each entry is delivered to add(col, value) by exactly one lane of the wavefront.
template <class Add>
__device__ __forceinline__ void gen_row(const DevTables &T, uint64_t B, int64_t row, RowScratch &S,
                                        Add &add) {
  const int lane = threadIdx.x & 63;
  // occupied orbitals in ascending order
  if (lane < T.norb && ((B >> lane) & 1ull)) S.occ[__popcll(B & below(lane))] = lane;
  if (lane == 0) {
    S.N = __popcll(B);
    S.NP = S.N * (S.N - 1) / 2;
  }
  __syncthreads();
  const int N = S.N, NP = S.NP;
  // occupied pairs in key order; lane i fills pairs (occ[i], occ[j>i])
  if (lane < N) {
    int q = lane * (2 * N - lane - 1) / 2;
    for (int j = lane + 1; j < N; ++j) S.pa[q++] = pid_of(S.occ[lane], S.occ[j]);
  }
  __syncthreads();
  if (lane == 0) {
    S.fpref[0] = 0;
    S.rpref[0] = 0;
    for (int q = 0; q < NP; ++q) {
      int p = S.pa[q];
      S.fpref[q + 1] = S.fpref[q] + (T.fstart[p + 1] - T.fstart[p]);
      S.rpref[q + 1] = S.rpref[q] + (T.rstart[p + 1] - T.rstart[p]);
    }
    S.opref[0] = 0;
    S.lpref[0] = 0;
    for (int i = 0; i < N; ++i) {
      int a = S.occ[i];
      S.opref[i + 1] = S.opref[i] + (T.ob_hi_start[a + 1] - T.ob_hi_start[a]);
      S.lpref[i + 1] = S.lpref[i] + (T.ob_lo_start[a + 1] - T.ob_lo_start[a]);
    }
  }
  __syncthreads();
  const int P = T.P;
  // two-body moves to kets above the bra
  const int Tf = S.fpref[NP];
  for (int t = lane; t < Tf; t += 64) {
    int lo = 0, hi = NP - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (S.fpref[mid] <= t) lo = mid; else hi = mid - 1;
    }
    int pa = S.pa[lo];
    int pc = T.fterm[T.fstart[pa] + (t - S.fpref[lo])];
    int c = T.pair_a[pc], d = T.pair_b[pc];
    uint64_t cd = (1ull << c) | (1ull << d);
    if (B & cd) continue;
    int a = T.pair_a[pa], b = T.pair_b[pa];
    uint64_t K = (B & ~((1ull << a) | (1ull << b))) | cd;
    int64_t col = ket_index(T, K);
    if (col < 0) continue;
    double v = T.vdense[(int64_t)pa * P + pc];
    add(col, phase2(B, a, b, c, d) ? -v : v);
  }
  // two-body moves to kets below the bra: the bra's pair is the created one
  const int Tr = S.rpref[NP];
  for (int t = lane; t < Tr; t += 64) {
    int lo = 0, hi = NP - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (S.rpref[mid] <= t) lo = mid; else hi = mid - 1;
    }
    int pc = S.pa[lo];
    int pa = T.rterm[T.rstart[pc] + (t - S.rpref[lo])];
    int a = T.pair_a[pa], b = T.pair_b[pa];
    uint64_t ab = (1ull << a) | (1ull << b);
    if (B & ab) continue;
    int c = T.pair_a[pc], d = T.pair_b[pc];
    uint64_t K = (B & ~((1ull << c) | (1ull << d))) | ab;
    int64_t col = ket_index(T, K);
    if (col < 0) continue;
    double v = T.vdense[(int64_t)pa * P + pc];
    add(col, phase2(K, a, b, c, d) ? -v : v);
  }
  // one-body moves a -> c above the bra (spectator sums in ascending order)
  const int To = S.opref[N];
  for (int t = lane; t < To; t += 64) {
    int i = 0;
    while (S.opref[i + 1] <= t) ++i;
    int a = S.occ[i];
    int c = T.ob_hi[T.ob_hi_start[a] + (t - S.opref[i])];
    if ((B >> c) & 1ull) continue;
    uint64_t K = (B & ~(1ull << a)) | (1ull << c);
    double acc = 0.0;
    bool any = false;
    if (T.esp_exist[a * T.norb + c]) {
      double e = T.esp[a * T.norb + c];
      acc = acc + (phase1(B, a, c) ? -e : e);
      any = true;
    }
    for (int j = 0; j < N; ++j) {
      int s = S.occ[j];
      if (s == a) continue;
      int p0 = min(a, s), p1 = max(a, s), p2 = min(c, s), p3 = max(c, s);
      int pa = pid_of(p0, p1), pc = pid_of(p2, p3);
      if (!vexists(T, pa, pc)) continue;
      double v = T.vdense[(int64_t)pa * P + pc];
      acc = acc + (phase2(B, p0, p1, p2, p3) ? -v : v);
      any = true;
    }
    if (!any) continue;
    int64_t col = ket_index(T, K);
    if (col >= 0) add(col, acc);
  }
  // one-body moves below the bra: occupied c, empty a < c; the ket's forward sum
  const int Tl = S.lpref[N];
  for (int t = lane; t < Tl; t += 64) {
    int i = 0;
    while (S.lpref[i + 1] <= t) ++i;
    int c = S.occ[i];
    int a = T.ob_lo[T.ob_lo_start[c] + (t - S.lpref[i])];
    if ((B >> a) & 1ull) continue;
    uint64_t K = (B & ~(1ull << c)) | (1ull << a);
    double acc = 0.0;
    bool any = false;
    if (T.esp_exist[a * T.norb + c]) {
      double e = T.esp[a * T.norb + c];
      acc = acc + (phase1(K, a, c) ? -e : e);
      any = true;
    }
    for (int j = 0; j < N; ++j) {  // spectators of K other than a: occ(B) without c
      int s = S.occ[j];
      if (s == c) continue;
      int p0 = min(a, s), p1 = max(a, s), p2 = min(c, s), p3 = max(c, s);
      int pa = pid_of(p0, p1), pc = pid_of(p2, p3);
      if (!vexists(T, pa, pc)) continue;
      double v = T.vdense[(int64_t)pa * P + pc];
      acc = acc + (phase2(K, p0, p1, p2, p3) ? -v : v);
      any = true;
    }
    if (!any) continue;
    int64_t col = ket_index(T, K);
    if (col >= 0) add(col, acc);
  }
  // diagonal: one-body terms over occupied orbitals, then pair terms in key order
  if (lane == 0) {
    double acc = 0.0;
    bool any = false;
    for (int j = 0; j < N; ++j) {
      int s = S.occ[j];
      if (T.esp_exist[s * T.norb + s]) {
        acc = acc + T.esp[s * T.norb + s];
        any = true;
      }
    }
    for (int q = 0; q < NP; ++q) {
      int p = S.pa[q];
      if (!vexists(T, p, p)) continue;
      acc = acc + T.vdense[(int64_t)p * P + p];
      any = true;
    }
    if (any) add(row, acc);
  }
}

// ------------------------------------------------------------------------------------
// stored build: pass 0 counts the kept entries of each row, pass 1 sorts them by column
// and writes them (CSR, both triangles) with FP64 and FP32 values.
template <int CAP>
__global__ __launch_bounds__(64) void k_build_rows(DevTables T, int64_t row0, int64_t nrows, int pass,
                                                  int keep_small, const int64_t *__restrict__ rowptr,
                                                  uint32_t *__restrict__ cols, double *__restrict__ v64,
                                                  float *__restrict__ v32, int32_t *__restrict__ counts,
                                                  double *__restrict__ diag, int *__restrict__ err) {
  __shared__ RowScratch S;
  __shared__ uint32_t s_col[CAP];
  __shared__ double s_val[CAP];
  __shared__ int s_n;
  const int64_t row = row0 + blockIdx.x;
  if (row >= row0 + nrows) return;
  const int lane = threadIdx.x;
  const uint64_t B = bra_state(T, row);
  if (lane == 0) s_n = 0;
  __syncthreads();
  auto add = [&](int64_t col, double v) {
    int at = atomicAdd(&s_n, 1);
    if (at < CAP) {
      s_col[at] = (uint32_t)col;
      s_val[at] = v;
    }
  };
  gen_row(T, B, row, S, add);
  __syncthreads();
  const int n = s_n;
  if (n > CAP) {
    if (lane == 0) atomicMax(err, n);
    return;
  }
  auto kept = [&](int i) { return keep_small || s_col[i] == (uint32_t)row || fabs(s_val[i]) >= 1e-14; };
  if (pass == 0) {
    int c = 0;
    for (int i = lane; i < n; i += 64) c += kept(i) ? 1 : 0;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if (lane == 0) counts[row] = c;
    return;
  }
  // bitonic sort of (col, val) by col, padded to a power of two
  int p2 = 1;
  while (p2 < n) p2 <<= 1;
  for (int i = n + lane; i < p2; i += 64) s_col[i] = 0xffffffffu;
  __syncthreads();
  for (int k = 2; k <= p2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < p2; i += 64) {
        int ixj = i ^ j;
        if (ixj > i) {
          bool up = (i & k) == 0;
          uint32_t ci = s_col[i], cj = s_col[ixj];
          if ((ci > cj) == up) {
            s_col[i] = cj;
            s_col[ixj] = ci;
            double t = s_val[i];
            s_val[i] = s_val[ixj];
            s_val[ixj] = t;
          }
        }
      }
      __syncthreads();
    }
  }
  int64_t out = rowptr[row];
  for (int base = 0; base < n; base += 64) {
    int i = base + lane;
    bool k = (i < n) && kept(i);
    uint64_t m = __ballot(k);
    int pos = __popcll(m & below(lane));
    if (k) {
      cols[out + pos] = s_col[i];
      v64[out + pos] = s_val[i];
      if (v32) v32[out + pos] = (float)s_val[i];
      if (s_col[i] == (uint32_t)row) diag[row] = s_val[i];
    }
    out += __popcll(m);
  }
}

// ------------------------------------------------------------------------------------
// stored H.X: one wavefront per row; lanes = (entry slot, vector).  X, Y row-major n x ld.
template <typename VT, typename XT, typename AT, int K>
__global__ __launch_bounds__(256) void k_spmm(int64_t n, const int64_t *__restrict__ rowptr,
                                              const uint32_t *__restrict__ cols, const VT *__restrict__ vals,
                                              const XT *__restrict__ X, XT *__restrict__ Y, int m, int ldx,
                                              int ldy) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n) return;
  constexpr int SLOTS = 64 / K;
  const int slot = lane / K, v = lane % K;
  const bool act = v < m;
  const int64_t s = rowptr[row], e = rowptr[row + 1];
  AT acc = 0;
#pragma unroll 4
  for (int64_t i = s + slot; i < e; i += SLOTS) {
    uint32_t c = cols[i];
    AT h = (AT)vals[i];
    if (act) acc += h * (AT)X[(int64_t)c * ldx + v];
  }
#pragma unroll
  for (int off = K; off < 64; off <<= 1) acc += __shfl_xor(acc, off);
  if (slot == 0 && act) Y[row * ldy + v] = (XT)acc;
}

// Stored H.X, vectorised form: X, Y row-major with ld = LPE * (16 / sizeof(T)) so one
// X row is LPE 16-byte pieces.  A wavefront owns a row; per 64-entry chunk every lane
// loads one (col, value) pair (coalesced, prefetched one chunk ahead), then LPE steps
// broadcast the entries to lane groups of LPE lanes, each lane gathering one 16-byte
// piece of its entry's X row.  Partial sums are reduced across groups at the row end.
// Workgroups are remapped so that each XCD (dispatched round-robin) takes chunks of 64
// consecutive workgroups: neighbouring rows touch neighbouring X rows, which then share
// that XCD's L2, and the eight XCDs work on nearby chunks at once, so the X rows they
// gather overlap in the shared Infinity Cache.  A workgroup holds 16 consecutive rows
// (1024 threads), so they also share one CU's L1.  Measured at fp 4p4n, k=16, FP32 /
// FP64, same bits in every case: 4 / 8 / 16 rows per workgroup 7.60 / 7.14 / 6.89 ms and
// 11.31 / 11.57 / 11.15 ms with one contiguous eighth of the rows per XCD; 16 rows with
// chunks of 16 / 64 / 512 / 4096 workgroups 6.79 / 6.74 / 6.76 / 6.85 ms (FP32); 2 or 4
// rows per wave in sequence (32 or 64 rows per workgroup) 7.55 / 8.22 ms.
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  using type = float4;
  static constexpr int N = 4;
};
template <>
struct Vec16<double> {
  using type = double2;
  static constexpr int N = 2;
};

__device__ __forceinline__ int64_t xcd_row_block(int remap) {
  const int64_t G = gridDim.x, i = blockIdx.x;
  if (!remap) return i;
  // XCD x = i % 8 takes chunks x, x + 8, x + 16, ... of C workgroups; the last
  // G mod 8C workgroups, dispatched last, keep their own index (a bijection on [0, G))
  constexpr int64_t C = 64;
  const int64_t G8 = (G / (8 * C)) * (8 * C);
  if (i >= G8) return i;
  const int64_t x = i % 8, j = i / 8;
  return ((j / C) * 8 + x) * C + j % C;
}

template <typename T, int LPE>
__global__ __launch_bounds__(1024) void k_spmm_vec(int64_t n, const int64_t *__restrict__ rowptr,
                                                  const uint32_t *__restrict__ cols, const T *__restrict__ vals,
                                                  const T *__restrict__ X, T *__restrict__ Y, int m, int ld,
                                                  int remap) {
  using V = typename Vec16<T>::type;
  constexpr int VPL = Vec16<T>::N;
  constexpr int EPS = 64 / LPE;
  const int lane = threadIdx.x & 63;
  const int64_t row = xcd_row_block(remap) * 16 + (threadIdx.x >> 6);
  if (row >= n) return;
  const int grp = lane / LPE, sub = lane % LPE;
  const int64_t s = rowptr[row], e = rowptr[row + 1];
  T acc[VPL];
#pragma unroll
  for (int q = 0; q < VPL; ++q) acc[q] = 0;
  uint32_t c = 0;
  T h = 0;
  if (s + lane < e) {
    c = __builtin_nontemporal_load(cols + s + lane);
    h = __builtin_nontemporal_load(vals + s + lane);
  }
  for (int64_t base = s; base < e; base += 64) {
    const int64_t nidx = base + 64 + lane;
    uint32_t cn = 0;
    T hn = 0;
    if (nidx < e) {
      cn = __builtin_nontemporal_load(cols + nidx);
      hn = __builtin_nontemporal_load(vals + nidx);
    }
    const int cnt = (int)min<int64_t>(64, e - base);
#pragma unroll
    for (int st = 0; st < LPE; ++st) {
      const int j = st * EPS + grp;
      const uint32_t cj = __shfl(c, j);
      const T hj = __shfl(h, j);
      if (j < cnt) {
        const V xv = *reinterpret_cast<const V *>(X + (int64_t)cj * ld + sub * VPL);
        const T *xs = reinterpret_cast<const T *>(&xv);
#pragma unroll
        for (int q = 0; q < VPL; ++q) acc[q] += hj * xs[q];
      }
    }
    c = cn;
    h = hn;
  }
#pragma unroll
  for (int q = 0; q < VPL; ++q)
    for (int off = LPE; off < 64; off <<= 1) acc[q] += __shfl_xor(acc[q], off);
  if (grp == 0) {
    T *y = Y + row * ld + sub * VPL;
    if ((sub + 1) * VPL <= m) {
      V o;
      T *os = reinterpret_cast<T *>(&o);
#pragma unroll
      for (int q = 0; q < VPL; ++q) os[q] = acc[q];
      *reinterpret_cast<V *>(y) = o;
    } else {
#pragma unroll
      for (int q = 0; q < VPL; ++q)
        if (sub * VPL + q < m) y[q] = acc[q];
    }
  }
}

// FP64 product summed per output element over the full row in ascending column order,
// one multiply then one add per entry (no contraction): the summation order of
// SparseMatrixCSB::TimesMatrix for a symmetric matrix (include/tav/csbmatrix.cxx:623-663,
// CSBBlock::TimesMatrix/TimesMatrixTranspose :280-328), so the result is bit-identical.
__global__ __launch_bounds__(256) void k_spmm_ordered(int64_t n, const int64_t *__restrict__ rowptr,
                                                      const uint32_t *__restrict__ cols,
                                                      const double *__restrict__ vals,
                                                      const double *__restrict__ X, double *__restrict__ Y,
                                                      int m, int ldx, int ldy) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = t / m;
  const int v = (int)(t % m);
  if (row >= n) return;
  double acc = 0.0;
  for (int64_t i = rowptr[row]; i < rowptr[row + 1]; ++i) {
    double p = vals[i] * X[(int64_t)cols[i] * ldx + v];
    acc = acc + p;
  }
  Y[row * ldy + v] = acc;
}

// ------------------------------------------------------------------------------------
// virtual H.X: rows are regenerated on the fly and applied at once; full symmetric rows
// make it a pure gather.  Every lane keeps K accumulators for the entries it generated.
template <typename XT, typename AT, int K>
__global__ __launch_bounds__(64) void k_virtual_apply(DevTables T, int64_t n, const XT *__restrict__ X,
                                                     XT *__restrict__ Y, int m, int ldx, int ldy) {
  __shared__ RowScratch S;
  const int64_t row = blockIdx.x;
  if (row >= n) return;
  const int lane = threadIdx.x;
  const uint64_t B = bra_state(T, row);
  AT acc[K];
#pragma unroll
  for (int v = 0; v < K; ++v) acc[v] = 0;
  auto add = [&](int64_t col, double h) {
    const XT *x = X + col * ldx;
    const AT hh = (AT)h;
#pragma unroll
    for (int v = 0; v < K; ++v)
      if (v < m) acc[v] += hh * (AT)x[v];
  };
  gen_row(T, B, row, S, add);
#pragma unroll
  for (int v = 0; v < K; ++v) {
    AT a = acc[v];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
    acc[v] = a;
  }
  if (lane < m && lane < K) {
    AT out = 0;
#pragma unroll
    for (int v = 0; v < K; ++v)
      if (v == lane) out = acc[v];
    Y[row * ldy + lane] = (XT)out;
  }
}

// number of rows whose stored entries include the diagonal (columns ascending: binary
// search per row); one integer atomic per block
__global__ __launch_bounds__(256) void k_count_diag(int64_t n, const int64_t *__restrict__ rowptr,
                                                    const uint32_t *__restrict__ cols,
                                                    unsigned long long *__restrict__ count) {
  __shared__ unsigned int part[256];
  unsigned int c = 0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = rowptr[r], hi = rowptr[r + 1];
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (cols[mid] < (uint32_t)r) lo = mid + 1;
      else hi = mid;
    }
    c += (lo < rowptr[r + 1] && cols[lo] == (uint32_t)r);
  }
  part[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(count, (unsigned long long)part[0]);
}

__global__ void k_pack_cast(int64_t n, int m, const double *__restrict__ src, int lds, float *__restrict__ dst,
                            int ldd) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * m) return;
  int64_t r = t / m;
  int v = (int)(t % m);
  dst[r * ldd + v] = (float)src[r * lds + v];
}
__global__ void k_unpack_cast(int64_t n, int m, const float *__restrict__ src, int lds, double *__restrict__ dst,
                              int ldd) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * m) return;
  int64_t r = t / m;
  int v = (int)(t % m);
  dst[r * ldd + v] = (double)src[r * lds + v];
}

// ------------------------------------------------------------------------------------
// host side: tables
int upload_tables(const System &s, DevTablesOwner &o, hipStream_t st) {
  int rc = ensure_binom();
  if (rc) return rc;
  const int norb = (int)s.orb.size();
  const int P = norb * (norb - 1) / 2;
  auto pid = [](int a, int b) { return b * (b - 1) / 2 + a; };
  std::vector<uint8_t> pair_a(std::max(P, 1)), pair_b(std::max(P, 1));
  for (int b = 1; b < norb; ++b)
    for (int a = 0; a < b; ++a) pair_a[pid(a, b)] = (uint8_t)a, pair_b[pid(a, b)] = (uint8_t)b;
  std::vector<double> vdense((size_t)P * P, 0.0);
  std::vector<uint32_t> vexist((size_t)P * P / 32 + 1, 0u);
  std::vector<std::vector<int32_t>> fw(P), rv(P);
  std::vector<std::vector<uint8_t>> obhi(norb), oblo(norb);
  auto add_partner = [&](int a, int c) {  // a < c
    if (std::find(obhi[a].begin(), obhi[a].end(), (uint8_t)c) == obhi[a].end()) obhi[a].push_back((uint8_t)c);
    if (std::find(oblo[c].begin(), oblo[c].end(), (uint8_t)a) == oblo[c].end()) oblo[c].push_back((uint8_t)a);
  };
  const size_t nv = s.vspx_v.size();
  for (size_t t = 0; t < nv; ++t) {
    const uint16_t *k = &s.vspx_k[4 * t];
    // terms that are not forward-ordered can only reach kets below the bra, which the
    // reference's forward search never finds (PPMatrix.cxx:615-619): they are dropped
    if (k[2] < k[0] || (k[2] == k[0] && k[3] < k[1])) continue;
    int pa = pid(k[0], k[1]), pc = pid(k[2], k[3]);
    size_t key = (size_t)pa * P + pc;
    vdense[key] = s.vspx_v[t];
    vexist[key >> 5] |= 1u << (key & 31);
    int shared = (k[2] == k[0] || k[2] == k[1]) + (k[3] == k[0] || k[3] == k[1]);
    if (shared == 0) {
      fw[pa].push_back(pc);
      rv[pc].push_back(pa);
    } else if (shared == 1) {
      int sp = (k[2] == k[0] || k[2] == k[1]) ? k[2] : k[3];
      int a = (k[0] == sp) ? k[1] : k[0];
      int c = (k[2] == sp) ? k[3] : k[2];
      if (a < c) add_partner(a, c);
    }
  }
  std::vector<double> esp((size_t)norb * norb, 0.0);
  std::vector<uint8_t> espx((size_t)norb * norb, 0);
  for (size_t t = 0; t < s.esp_v.size(); ++t) {
    int a = s.esp_k[2 * t], c = s.esp_k[2 * t + 1];
    if (c < a) continue;  // not forward: never reaches a ket at or above the bra
    esp[a * norb + c] = s.esp_v[t];
    espx[a * norb + c] = 1;
    if (a < c) add_partner(a, c);
  }
  for (auto &v : obhi) std::sort(v.begin(), v.end());
  for (auto &v : oblo) std::sort(v.begin(), v.end());
  auto flat32 = [](const std::vector<std::vector<int32_t>> &v, std::vector<int32_t> &start, std::vector<int32_t> &d) {
    start.assign(1, 0);
    for (auto &x : v) {
      d.insert(d.end(), x.begin(), x.end());
      start.push_back((int32_t)d.size());
    }
    if (d.empty()) d.push_back(0);
  };
  auto flat8 = [](const std::vector<std::vector<uint8_t>> &v, std::vector<int32_t> &start, std::vector<uint8_t> &d) {
    start.assign(1, 0);
    for (auto &x : v) {
      d.insert(d.end(), x.begin(), x.end());
      start.push_back((int32_t)d.size());
    }
    if (d.empty()) d.push_back(0);
  };
  std::vector<int32_t> fstart, fterm, rstart, rterm, ohs, ols;
  std::vector<uint8_t> oh, ol;
  flat32(fw, fstart, fterm);
  flat32(rv, rstart, rterm);
  flat8(obhi, ohs, oh);
  flat8(oblo, ols, ol);
  std::vector<int32_t> pcfg_sec(s.pcfg_rank.size(), -1);
  for (size_t sec = 0; sec + 1 < s.psec_start.size(); ++sec)
    for (int64_t q = s.psec_start[sec]; q < s.psec_start[sec + 1]; ++q) pcfg_sec[colex_rank(s.pmask[q])] = (int32_t)sec;

  // one blob, 256-byte aligned pieces
  struct Piece {
    const void *src;
    size_t bytes;
    size_t off;
  };
  std::vector<Piece> pcs;
  size_t total = 0;
  auto put = [&](const void *p, size_t b) {
    size_t off = total;
    pcs.push_back({p, b, off});
    total += (b + 255) & ~size_t(255);
    return off;
  };
  size_t o_nmask = put(s.nmask.data(), s.nmask.size() * 4), o_noff = put(s.noff.data(), s.noff.size() * 8),
         o_nsec = put(s.nsector.data(), s.nsector.size() * 4), o_pmask = put(s.pmask.data(), s.pmask.size() * 4),
         o_psec = put(s.psec_start.data(), s.psec_start.size() * 8),
         o_prank = put(s.pcfg_rank.data(), s.pcfg_rank.size() * 4), o_psecof = put(pcfg_sec.data(), pcfg_sec.size() * 4),
         o_nidx = put(s.ncfg_index.data(), s.ncfg_index.size() * 4), o_vd = put(vdense.data(), vdense.size() * 8),
         o_vx = put(vexist.data(), vexist.size() * 4), o_fs = put(fstart.data(), fstart.size() * 4),
         o_ft = put(fterm.data(), fterm.size() * 4), o_rs = put(rstart.data(), rstart.size() * 4),
         o_rt = put(rterm.data(), rterm.size() * 4), o_ohs = put(ohs.data(), ohs.size() * 4),
         o_oh = put(oh.data(), oh.size()), o_ols = put(ols.data(), ols.size() * 4), o_ol = put(ol.data(), ol.size()),
         o_esp = put(esp.data(), esp.size() * 8), o_espx = put(espx.data(), espx.size()),
         o_pa = put(pair_a.data(), pair_a.size()), o_pb = put(pair_b.data(), pair_b.size()),
         o_kr = put(s.krank.data(), s.krank.size() * 4), o_ks = put(s.kstart.data(), s.kstart.size() * 8);
  std::vector<uint8_t> host(total, 0);
  for (auto &p : pcs)
    if (p.bytes) memcpy(host.data() + p.off, p.src, p.bytes);
  void *blob = nullptr;
  HIPCHK(dmalloc(&blob, total));
  HIPCHK(hipMemcpyAsync(blob, host.data(), total, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  auto at = [&](size_t off) { return (const void *)((const char *)blob + off); };
  DevTables &T = o.t;
  T.nn = s.nn;
  T.np = s.np;
  T.nv = s.nv;
  T.zv = s.zv;
  T.norb = norb;
  T.P = P;
  T.nnconf = (int)s.nmask.size();
  T.nmask = (const uint32_t *)at(o_nmask);
  T.noff = (const int64_t *)at(o_noff);
  T.nsector = (const int32_t *)at(o_nsec);
  T.pmask = (const uint32_t *)at(o_pmask);
  T.psec_start = (const int64_t *)at(o_psec);
  T.pcfg_rank = (const int32_t *)at(o_prank);
  T.pcfg_sec = (const int32_t *)at(o_psecof);
  T.ncfg_index = (const int32_t *)at(o_nidx);
  T.krank = s.truncated() ? (const int32_t *)at(o_kr) : nullptr;
  T.kstart = s.truncated() ? (const int64_t *)at(o_ks) : nullptr;
  T.vdense = (const double *)at(o_vd);
  T.vexist = (const uint32_t *)at(o_vx);
  T.fstart = (const int32_t *)at(o_fs);
  T.fterm = (const int32_t *)at(o_ft);
  T.rstart = (const int32_t *)at(o_rs);
  T.rterm = (const int32_t *)at(o_rt);
  T.ob_hi_start = (const int32_t *)at(o_ohs);
  T.ob_hi = (const uint8_t *)at(o_oh);
  T.ob_lo_start = (const int32_t *)at(o_ols);
  T.ob_lo = (const uint8_t *)at(o_ol);
  T.esp = (const double *)at(o_esp);
  T.esp_exist = (const uint8_t *)at(o_espx);
  T.pair_a = (const uint8_t *)at(o_pa);
  T.pair_b = (const uint8_t *)at(o_pb);
  o.blob = blob;
  o.bytes = total;
  return COSMO_OK;
}

void free_tables(DevTablesOwner &o) {
  if (o.blob) dfree(o.blob);
  o.blob = nullptr;
}

// ------------------------------------------------------------------------------------
// operator build
template <int CAP>
static void launch_build(const DevTables &T, int64_t n, int pass, int keep, CsrDev &M, int32_t *counts, int *err,
                         hipStream_t st) {
  const int64_t chunk = 1 << 20;
  for (int64_t r0 = 0; r0 < n; r0 += chunk) {
    int64_t nr = std::min(chunk, n - r0);
    hipLaunchKernelGGL((k_build_rows<CAP>), dim3((unsigned)nr), dim3(64), 0, st, T, r0, nr, pass, keep, M.rowptr,
                       M.cols, M.v64, M.v32, counts, M.diag, err);
  }
}

int build_stored(Op &op, const System &s, int flags, hipStream_t st) {
  const int64_t n = s.dim;
  const int keep = (flags & COSMO_OP_KEEP_SMALL) ? 1 : 0;
  int32_t *counts = nullptr;
  int *err = nullptr;
  HIPCHK(dmalloc(&counts, sizeof(int32_t) * (n + 1)));
  HIPCHK(dmalloc(&err, sizeof(int)));
  HIPCHK(hipMemsetAsync(err, 0, sizeof(int), st));
  CsrDev &M = op.csr;
  int cap = 2048;
  launch_build<2048>(op.tab.t, n, 0, keep, M, counts, err, st);
  int herr = 0;
  HIPCHK(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (herr > 0) {
    if (herr > 8192) {
      dfree(counts);
      dfree(err);
      return fail(COSMO_EUNSUP, "Hamiltonian row longer than 8192 entries");
    }
    cap = 8192;
    HIPCHK(hipMemsetAsync(err, 0, sizeof(int), st));
    launch_build<8192>(op.tab.t, n, 0, keep, M, counts, err, st);
  }
  std::vector<int32_t> hc(n);
  HIPCHK(hipMemcpyAsync(hc.data(), counts, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::vector<int64_t> rp(n + 1, 0);
  for (int64_t r = 0; r < n; ++r) rp[r + 1] = rp[r] + hc[r];
  const int64_t nnz = rp[n];
  M.n = n;
  M.nnz = nnz;
  HIPCHK(dmalloc(&M.rowptr, sizeof(int64_t) * (n + 1)));
  HIPCHK(dmalloc(&M.cols, sizeof(uint32_t) * std::max<int64_t>(nnz, 1)));
  HIPCHK(dmalloc(&M.v64, sizeof(double) * std::max<int64_t>(nnz, 1)));
  if (!(flags & COSMO_OP_NO_FP32)) HIPCHK(dmalloc(&M.v32, sizeof(float) * std::max<int64_t>(nnz, 1)));
  HIPCHK(dmalloc(&M.diag, sizeof(double) * std::max<int64_t>(n, 1)));
  HIPCHK(hipMemsetAsync(M.diag, 0, sizeof(double) * n, st));
  HIPCHK(hipMemcpyAsync(M.rowptr, rp.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, st));
  if (cap == 2048) launch_build<2048>(op.tab.t, n, 1, keep, M, counts, err, st);
  else launch_build<8192>(op.tab.t, n, 1, keep, M, counts, err, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  dfree(counts);
  dfree(err);
  // upper-triangle count: off-diagonal entries come in pairs, plus the diagonal entries
  // actually stored (a row whose terms give no diagonal element stores none)
  unsigned long long *nd = nullptr;
  HIPCHK(dmalloc(&nd, sizeof(unsigned long long)));
  HIPCHK(hipMemsetAsync(nd, 0, sizeof(unsigned long long), st));
  hipLaunchKernelGGL(k_count_diag, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)), dim3(256), 0, st, n,
                     (const int64_t *)M.rowptr, (const uint32_t *)M.cols, nd);
  HIPCHK(hipGetLastError());
  unsigned long long ndiag = 0;
  HIPCHK(hipMemcpyAsync(&ndiag, nd, sizeof(ndiag), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  dfree(nd);
  M.nnz_upper = (nnz + (int64_t)ndiag) / 2;
  return COSMO_OK;
}

// ------------------------------------------------------------------------------------
// products
int g_spmm_remap = 1;  // XCD-chunked row ranges (COSMO_SPMM_REMAP=0 turns it off)

// vectorised kernel when X and Y share a power-of-two leading dimension of >= 16 bytes
template <typename T>
static bool launch_spmm_vec(const CsrDev &M, const T *vals, const T *X, T *Y, int m, int ld, hipStream_t st) {
  constexpr int VPL = 16 / sizeof(T);
  if (ld < VPL || (ld & (ld - 1)) || ld > 64 || m > ld) return false;
  if ((reinterpret_cast<uintptr_t>(X) | reinterpret_cast<uintptr_t>(Y)) & 15) return false;
  const int lpe = ld / VPL;
  dim3 g((unsigned)((M.n + 15) / 16)), b(1024);
  const int rm = g_spmm_remap;
  switch (lpe) {
    case 1: hipLaunchKernelGGL((k_spmm_vec<T, 1>), g, b, 0, st, M.n, M.rowptr, M.cols, vals, X, Y, m, ld, rm); break;
    case 2: hipLaunchKernelGGL((k_spmm_vec<T, 2>), g, b, 0, st, M.n, M.rowptr, M.cols, vals, X, Y, m, ld, rm); break;
    case 4: hipLaunchKernelGGL((k_spmm_vec<T, 4>), g, b, 0, st, M.n, M.rowptr, M.cols, vals, X, Y, m, ld, rm); break;
    case 8: hipLaunchKernelGGL((k_spmm_vec<T, 8>), g, b, 0, st, M.n, M.rowptr, M.cols, vals, X, Y, m, ld, rm); break;
    case 16: hipLaunchKernelGGL((k_spmm_vec<T, 16>), g, b, 0, st, M.n, M.rowptr, M.cols, vals, X, Y, m, ld, rm); break;
    case 32: hipLaunchKernelGGL((k_spmm_vec<T, 32>), g, b, 0, st, M.n, M.rowptr, M.cols, vals, X, Y, m, ld, rm); break;
    default: return false;
  }
  return true;
}

template <typename VT, typename XT, typename AT>
static int launch_spmm_t(const CsrDev &M, const VT *vals, const XT *X, XT *Y, int m, int ldx, int ldy,
                         hipStream_t st) {
  if constexpr (std::is_same<VT, XT>::value) {
    if (ldx == ldy && launch_spmm_vec<XT>(M, vals, X, Y, m, ldx, st)) {
      HIPCHK(hipGetLastError());
      return COSMO_OK;
    }
  }
  dim3 g((unsigned)((M.n + 3) / 4)), b(256);
  if (m <= 1) hipLaunchKernelGGL((k_spmm<VT, XT, AT, 1>), g, b, 0, st, M.n, M.rowptr, M.cols, vals, X, Y, m, ldx, ldy);
  else if (m <= 2) hipLaunchKernelGGL((k_spmm<VT, XT, AT, 2>), g, b, 0, st, M.n, M.rowptr, M.cols, vals, X, Y, m, ldx, ldy);
  else if (m <= 4) hipLaunchKernelGGL((k_spmm<VT, XT, AT, 4>), g, b, 0, st, M.n, M.rowptr, M.cols, vals, X, Y, m, ldx, ldy);
  else if (m <= 8) hipLaunchKernelGGL((k_spmm<VT, XT, AT, 8>), g, b, 0, st, M.n, M.rowptr, M.cols, vals, X, Y, m, ldx, ldy);
  else if (m <= 16) hipLaunchKernelGGL((k_spmm<VT, XT, AT, 16>), g, b, 0, st, M.n, M.rowptr, M.cols, vals, X, Y, m, ldx, ldy);
  else if (m <= 32) hipLaunchKernelGGL((k_spmm<VT, XT, AT, 32>), g, b, 0, st, M.n, M.rowptr, M.cols, vals, X, Y, m, ldx, ldy);
  else hipLaunchKernelGGL((k_spmm<VT, XT, AT, 64>), g, b, 0, st, M.n, M.rowptr, M.cols, vals, X, Y, m, ldx, ldy);
  HIPCHK(hipGetLastError());
  return COSMO_OK;
}

template <typename XT, typename AT>
static int launch_virtual_t(const DevTables &T, int64_t n, const XT *X, XT *Y, int m, int ldx, int ldy,
                            hipStream_t st) {
  dim3 g((unsigned)n), b(64);
  if (m <= 1) hipLaunchKernelGGL((k_virtual_apply<XT, AT, 1>), g, b, 0, st, T, n, X, Y, m, ldx, ldy);
  else if (m <= 2) hipLaunchKernelGGL((k_virtual_apply<XT, AT, 2>), g, b, 0, st, T, n, X, Y, m, ldx, ldy);
  else if (m <= 4) hipLaunchKernelGGL((k_virtual_apply<XT, AT, 4>), g, b, 0, st, T, n, X, Y, m, ldx, ldy);
  else if (m <= 8) hipLaunchKernelGGL((k_virtual_apply<XT, AT, 8>), g, b, 0, st, T, n, X, Y, m, ldx, ldy);
  else hipLaunchKernelGGL((k_virtual_apply<XT, AT, 16>), g, b, 0, st, T, n, X, Y, m, ldx, ldy);
  HIPCHK(hipGetLastError());
  return COSMO_OK;
}

// Y = H X on device, X/Y row-major with leading dimensions; m <= 64 (virtual: <= 16)
int op_apply_dev(Op &op, const void *X, void *Y, int m, int ldx, int ldy, int prec, hipStream_t st) {
  if (m <= 0) return COSMO_OK;
  static bool env_done = false;
  if (!env_done) {
    if (const char *e = getenv("COSMO_SPMM_REMAP")) g_spmm_remap = atoi(e);
    if (const char *e = getenv("COSMO_SELL")) g_sell_off = atoi(e) == 0;
    if (const char *e = getenv("COSMO_SELL_CHUNK")) g_sell_chunk = std::max(1, atoi(e));
    if (const char *e = getenv("COSMO_SELL_STATS")) g_sell_stats = atoi(e);
    if (const char *e = getenv("COSMO_SELL_SPLIT")) g_sell_split = atoi(e) != 0;
    if (const char *e = getenv("COSMO_SELL_NB")) g_sell_nb = atoi(e);
    env_done = true;
  }
  if (op.fact && prec != COSMO_PREC_FP64_ORDERED && ldx == ldy) {
    const int vpl = prec == COSMO_PREC_FP32 ? 4 : 2;
    if (ldx % vpl == 0 && !(ldx & (ldx - 1))) return fact_apply(op, X, Y, m, ldx, prec, st);
  }
  if (op.is_virtual) {
    for (int v0 = 0; v0 < m; v0 += 16) {
      int mm = std::min(16, m - v0);
      int rc;
      if (prec == COSMO_PREC_FP32)
        rc = launch_virtual_t<float, float>(op.tab.t, op.n, (const float *)X + v0, (float *)Y + v0, mm, ldx, ldy, st);
      else
        rc = launch_virtual_t<double, double>(op.tab.t, op.n, (const double *)X + v0, (double *)Y + v0, mm, ldx, ldy,
                                              st);
      if (rc) return rc;
    }
    return COSMO_OK;
  }
  if (ldx == ldy && m <= ldx && prec != COSMO_PREC_FP64_ORDERED) {
    const int rc = sell_apply(op, X, Y, m, ldx, prec, st);
    if (rc != COSMO_EUNSUP) return rc;
  }
  const CsrDev &M = op.csr;
  for (int v0 = 0; v0 < m; v0 += 64) {
    int mm = std::min(64, m - v0);
    int rc = COSMO_OK;
    if (prec == COSMO_PREC_FP32) {
      if (!M.v32) return fail(COSMO_EINVAL, "operator built without FP32 values");
      rc = launch_spmm_t<float, float, float>(M, M.v32, (const float *)X + v0, (float *)Y + v0, mm, ldx, ldy, st);
    } else if (prec == COSMO_PREC_FP64) {
      rc = launch_spmm_t<double, double, double>(M, M.v64, (const double *)X + v0, (double *)Y + v0, mm, ldx, ldy,
                                                 st);
    } else if (prec == COSMO_PREC_FP64_ORDERED) {
      int64_t tot = M.n * mm;
      hipLaunchKernelGGL(k_spmm_ordered, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, M.n, M.rowptr, M.cols,
                         M.v64, (const double *)X + v0, (double *)Y + v0, mm, ldx, ldy);
      HIPCHK(hipGetLastError());
    } else
      return fail(COSMO_EINVAL, "unknown precision");
    if (rc) return rc;
  }
  return COSMO_OK;
}

int cast_pack(int64_t n, int m, const double *src, int lds, float *dst, int ldd, hipStream_t st) {
  int64_t tot = n * m;
  if (!tot) return COSMO_OK;
  hipLaunchKernelGGL(k_pack_cast, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n, m, src, lds, dst, ldd);
  HIPCHK(hipGetLastError());
  return COSMO_OK;
}
int cast_unpack(int64_t n, int m, const float *src, int lds, double *dst, int ldd, hipStream_t st) {
  int64_t tot = n * m;
  if (!tot) return COSMO_OK;
  hipLaunchKernelGGL(k_unpack_cast, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n, m, src, lds, dst, ldd);
  HIPCHK(hipGetLastError());
  return COSMO_OK;
}

int set_device_impl(int d) {
  int cnt = 0;
  HIPCHK(hipGetDeviceCount(&cnt));
  if (d < 0 || d >= cnt || d >= 64) return fail(COSMO_EINVAL, "cosmo_set_device: no such device");
  HIPCHK(hipSetDevice(d));
  t_device = d;
  return COSMO_OK;
}

int current_device() { return t_device; }

}  // namespace cosmo

namespace cosmo {
// diagonal of H for the virtual operator: one thread per row, the reference's order
// (one-body terms over occupied orbitals, then pair terms in key order)
__global__ void k_diag_rows(DevTables T, int64_t n, double *__restrict__ diag) {
  int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  uint64_t B = bra_state(T, row);
  int occ[64];
  int N = 0;
  for (uint64_t m = B; m; m &= m - 1) occ[N++] = __ffsll((unsigned long long)m) - 1;
  double acc = 0.0;
  for (int j = 0; j < N; ++j)
    if (T.esp_exist[occ[j] * T.norb + occ[j]]) acc = acc + T.esp[occ[j] * T.norb + occ[j]];
  for (int i = 0; i < N; ++i)
    for (int j = i + 1; j < N; ++j) {
      int p = pid_of(occ[i], occ[j]);
      if (vexists(T, p, p)) acc = acc + T.vdense[(int64_t)p * T.P + p];
    }
  diag[row] = acc;
}

int virtual_diag(const DevTables &T, int64_t n, double *d_diag, hipStream_t st) {
  if (n <= 0) return COSMO_OK;
  hipLaunchKernelGGL(k_diag_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, T, n, d_diag);
  HIPCHK(hipGetLastError());
  return COSMO_OK;
}
}  // namespace cosmo

namespace cosmo {
// exact device bytes of a stored operator: the builder's count pass without the fill
int stored_footprint(const System &s, int flags, int64_t *bytes) {
  hipStream_t st = lib_stream();
  DevTablesOwner tab;
  int rc = upload_tables(s, tab, st);
  if (rc) return rc;
  const int64_t n = s.dim;
  int32_t *counts = nullptr;
  int *err = nullptr;
  CsrDev M;
  if (dmalloc(&counts, sizeof(int32_t) * (n + 1)) != hipSuccess || dmalloc(&err, sizeof(int)) != hipSuccess) {
    dfree(counts);
    free_tables(tab);
    return fail(COSMO_ENOMEM, "footprint count buffers");
  }
  hipMemsetAsync(err, 0, sizeof(int), st);
  const int keep = (flags & COSMO_OP_KEEP_SMALL) ? 1 : 0;
  launch_build<2048>(tab.t, n, 0, keep, M, counts, err, st);
  int herr = 0;
  hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st);
  hipStreamSynchronize(st);
  if (herr > 0) {
    hipMemsetAsync(err, 0, sizeof(int), st);
    launch_build<8192>(tab.t, n, 0, keep, M, counts, err, st);
  }
  std::vector<int32_t> hc(n);
  hipMemcpyAsync(hc.data(), counts, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st);
  hipError_t e = hipStreamSynchronize(st);
  dfree(counts);
  dfree(err);
  int64_t nnz = 0;
  for (int64_t r = 0; r < n; ++r) nnz += hc[r];
  *bytes = (int64_t)tab.bytes + 8 * (n + 1) + 8 * n + nnz * (4 + 8 + ((flags & COSMO_OP_NO_FP32) ? 0 : 4)) +
           sell_footprint_bound(n, nnz);
  free_tables(tab);
  if (e != hipSuccess) return fail(COSMO_EHIP, hipGetErrorString(e));
  return COSMO_OK;
}
}  // namespace cosmo
