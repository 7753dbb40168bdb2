// factored.hip -- virtual H.X in proton-neutron product form (no H stored).
//
// The basis is blocked by neutron configuration nu; the rows of a block are the proton
// configurations pi of one (M, parity) proton sector.  With E_ca = a+_c a_a,
//   H = Hn (x) 1 + 1 (x) Hp + sum_{a,c,b,d} W[a][c][b][d] E^n_ca E^p_db,
// where Hn / Hp hold every one- and two-body term of one species and W the pn two-body
// terms.  Neutron orbitals are the low bits of a state, so the fermion sign of a pn move
// is the product of a neutron sign (popcounts inside nu) and a proton sign (inside pi).
// For block nu the product becomes
//   Y_nu = sum_nu' Hn[nu,nu'] X_nu' + Hp_s X_nu                      (class products, MFMA)
//        + sum_{a in nu} W_aa . E^p X_nu                               (self step)
//        + sum_{a->c} phase_n(nu,a,c) sum_{b,d} W_ac[b,d] E^p_db X_nu' (one step per hop)
// with nu' = nu - a + c.  Each step stages the X rows of block nu' (contiguous: one tile)
// in LDS and applies, per row pi, the proton hops b -> d whose (delta m, delta parity)
// matches the neutron hop -- the proton hop lists depend only on the sector and that key,
// are L2-resident, and are staged in LDS once per key (steps are sorted by key).  One
// workgroup owns one block (and one column slice of the vectors); its lane groups own
// rows and keep the sums in registers; the next step's tile is loaded into registers
// while the current one is consumed.  The single-species parts are dense products per
// (M_n, parity) class, one launch of k_class_mm on the matrix cores for all classes,
// reading X in place (each class is an S*ld x nb matrix whose columns are its blocks,
// multiplied by Hn_class, and every block times the sector's Hp); the step kernel adds
// both results in its epilogue.  Classes beyond that kernel's limits use a gathered
// class-contiguous copy of X and rocBLAS instead.
//
// Hn, Hp and W are produced by the same row generator as the stored operator (the
// species-only subsystems built with the GPU builder for Hn and Hp; W is the pn part of
// the forward two-body table as the builder reads it, include/sm/spherical2sp.cxx:441-442
// ordering), so the product equals the stored operator's up to the summation order.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <memory>
#include <vector>

#include "op.h"

namespace cosmo {

namespace {

constexpr int FACT_THREADS = 512;

template <typename T>
struct Piece16;
template <>
struct Piece16<float> {
  using type = float4;
  static constexpr int N = 4;
};
template <>
struct Piece16<double> {
  using type = double2;
  static constexpr int N = 2;
};

// step code: a | c << 5 | phase_n << 10 | self << 11 | key << 12
struct FactArgs {
  int nblk, nn, np2, NK;
  int hop_regs;  // every (sector, key) hop list fits the register prefetch
  int wneg;      // offset (in elements) of the negated W slice behind the W slice
  // LDS byte offsets (the W slice and its negation sit at 0): X tile, step table, key
  // table, hop row pointers, hop words
  int lds_tile, lds_st, lds_key, lds_hptr, lds_hw;
  const int64_t *noff;          // [nblk+1] first row of each block
  const int32_t *nsector;       // [nblk]
  const uint32_t *nmask;        // [nblk]
  const int64_t *poff;          // [nblk] first row in the class-contiguous layout
  const int64_t *qoff;          // [nblk] row of (proton 0, this block) in the proton-major layout
  const int32_t *qnb;           // [nblk] blocks of its class (row stride of the proton-major layout)
  const int32_t *step_start;    // [nblk+1]
  const int32_t *step_tb;       // target block of each step
  const uint32_t *step_code;
  const int64_t *hk_pbeg;       // [nsec*NK] start of the row pointers of (sector, key)
  const int32_t *hk_ptr;        // row pointers into hk_words
  const uint32_t *hk_words;     // target row << 16 | 4 * (b*np + d + sign * wneg)
};

struct StepInfo {
  uint32_t code;
  uint32_t t0;  // first row of the target block
  int S;        // rows of the target block
  int pad;
};

struct KeyInfo {
  int pb;    // row pointers of (sector, key) start at hk_ptr[pb]
  int base;  // first hop word
  int cnt;   // hop words
  int pad;
};

constexpr int FACT_HOP_REGS = 12;  // hop words per thread held for the next key

// workgroups are dealt round-robin over the 8 XCDs: XCD x = i % 8 takes chunks x, x + 8,
// x + 16, ... of 4 consecutive blocks, so neighbouring blocks share its L2 while the eight
// XCDs work on nearby blocks and share X rows in the Infinity Cache; the last nblk mod 32
// workgroups, dispatched last, keep their own index (a bijection on [0, nblk)).  Measured
// at fp 4p4n, k=16 FP32 / FP64, same bits: one contiguous eighth of the blocks per XCD
// 3.72 / 7.39 ms, chunks of 1 / 2 / 4 / 16 blocks 3.60 / 3.59 / 3.56 / 3.63 ms (FP32).
__device__ __forceinline__ int xcd_block(int nblk) {
  constexpr int C = 4;
  const int i = blockIdx.x, G8 = (nblk / (8 * C)) * (8 * C);
  if (i >= G8) return i;
  const int x = i % 8, j = i / 8;
  return ((j / C) * 8 + x) * C + j % C;
}

// pn part of Y for one block and one column slice (the neutron-only and proton-only
// products are in Yp, added in the epilogue).  Per step the LDS holds the X tile of the
// target block, the signed W slice of the neutron hop and the proton hop lists of the
// step's key.  The next step's X tile and W slice, and the hop lists of the next key,
// are loaded into registers while the current step computes; the block's step and key
// tables are read into LDS once.
template <typename T, int LPE, int RMAX>
__global__ __launch_bounds__(FACT_THREADS) void k_fact(FactArgs A, const T *__restrict__ W,
                                                       const T *__restrict__ X, const T *__restrict__ Yp,
                                                       const T *__restrict__ Yq, T *__restrict__ Y, int ld, int m) {
  constexpr int VPL = Piece16<T>::N;
  // native vectors (not HIP's union vector types, which keep a register array in scratch)
  typedef T NV __attribute__((ext_vector_type(VPL)));
  constexpr int G = FACT_THREADS / LPE;  // lane groups (one row each per pass)
  constexpr int SW = LPE * VPL;          // vector columns of one slice
  constexpr int PP = (RMAX * G + FACT_THREADS) / FACT_THREADS;  // row pointers per thread
  extern __shared__ __align__(16) unsigned char lds[];
  T *wsl = reinterpret_cast<T *>(lds);
  NV *tile = reinterpret_cast<NV *>(lds + A.lds_tile);
  StepInfo *sst = reinterpret_cast<StepInfo *>(lds + A.lds_st);
  KeyInfo *skey = reinterpret_cast<KeyInfo *>(lds + A.lds_key);
  int *hptr = reinterpret_cast<int *>(lds + A.lds_hptr);
  uint32_t *hw = reinterpret_cast<uint32_t *>(lds + A.lds_hw);

  const int blk = xcd_block(A.nblk);
  const int col0 = blockIdx.y * SW;
  if (col0 >= m) return;
  const int tid = threadIdx.x, g = tid / LPE, sub = tid % LPE;
  const int64_t r0 = A.noff[blk];
  const int S = (int)(A.noff[blk + 1] - r0);
  const int sec = A.nsector[blk];
  const uint32_t nu = A.nmask[blk];
  const int st0 = A.step_start[blk], nst = A.step_start[blk + 1] - st0;

  // block header: step table and key table of the sector
  for (int j = tid; j < nst; j += FACT_THREADS) {
    const int tb = A.step_tb[st0 + j];
    const int64_t t0 = A.noff[tb];
    sst[j] = StepInfo{A.step_code[st0 + j], (uint32_t)t0, (int)(A.noff[tb + 1] - t0), 0};
  }
  for (int k = tid; k < A.NK; k += FACT_THREADS) {
    const int pb = (int)A.hk_pbeg[(int64_t)sec * A.NK + k];
    const int base = A.hk_ptr[pb];
    skey[k] = KeyInfo{pb, base, A.hk_ptr[pb + S] - base, 0};
  }
  __syncthreads();

  T acc[RMAX][VPL];
#pragma unroll
  for (int i = 0; i < RMAX; ++i)
#pragma unroll
    for (int v = 0; v < VPL; ++v) acc[i][v] = 0;

  NV pre[RMAX];
  T wp[2];
  uint32_t hreg[FACT_HOP_REGS];
  int preg[PP];
  // Prefetch loads are unconditional (addresses clamped into range, surplus lanes load a
  // duplicate): a load under a divergent branch is waited for at the end of that branch,
  // which would serialise the prefetch instead of overlapping it with the step's compute.
  auto prefetch = [&](const StepInfo &si) {
    const int npc = si.S * LPE;
#pragma unroll
    for (int j = 0; j < RMAX; ++j) {
      const int p = min(tid + j * FACT_THREADS, npc - 1);
      pre[j] = *reinterpret_cast<const NV *>(X + ((int64_t)si.t0 + p / LPE) * ld + col0 + (p % LPE) * VPL);
    }
    if (!((si.code >> 11) & 1u)) {
      const T *src = W + (int64_t)((si.code & 31) * A.nn + ((si.code >> 5) & 31)) * A.np2;
#pragma unroll
      for (int j = 0; j < 2; ++j) wp[j] = src[min(tid + j * FACT_THREADS, A.np2 - 1)];
    } else {  // the block itself: W summed over its occupied neutron orbitals a (a -> a)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int t = min(tid + j * FACT_THREADS, A.np2 - 1);
        T w = 0;
        for (uint32_t bb = nu; bb; bb &= bb - 1) {
          const int a = __ffs(bb) - 1;
          w += W[(int64_t)(a * A.nn + a) * A.np2 + t];
        }
        wp[j] = w;
      }
    }
  };
  auto hop_prefetch = [&](int key) {
    const KeyInfo ki = skey[key];
#pragma unroll
    for (int j = 0; j < PP; ++j) preg[j] = A.hk_ptr[ki.pb + min(tid + j * FACT_THREADS, S)] - ki.base;
    const int last = max(ki.cnt - 1, 0);  // hk_words holds one spare entry at its end
#pragma unroll
    for (int j = 0; j < FACT_HOP_REGS; ++j) hreg[j] = A.hk_words[ki.base + min(tid + j * FACT_THREADS, last)];
  };
  int cur_key = -1;
  prefetch(sst[0]);
  if (A.hop_regs) hop_prefetch((int)(sst[0].code >> 12));
  for (int i = 0; i < nst; ++i) {
    const StepInfo si = sst[i];
    const int key = (int)(si.code >> 12);
    const bool self = (si.code >> 11) & 1u;
    __syncthreads();  // the previous step is done with the LDS
    {
      const int npc = si.S * LPE;
#pragma unroll
      for (int j = 0; j < RMAX; ++j) {
        const int p = tid + j * FACT_THREADS;
        if (p < npc) tile[p] = pre[j];
      }
    }
    {
      const bool neg = !self && ((si.code >> 10) & 1u);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int t = tid + j * FACT_THREADS;
        if (t < A.np2) {
          wsl[t] = neg ? -wp[j] : wp[j];
          wsl[A.wneg + t] = neg ? wp[j] : -wp[j];
        }
      }
    }
    if (key != cur_key) {
      if (A.hop_regs) {
        const int cnt = skey[key].cnt;
#pragma unroll
        for (int j = 0; j < PP; ++j) {
          const int r = tid + j * FACT_THREADS;
          if (r <= S) hptr[r] = preg[j];
        }
#pragma unroll
        for (int j = 0; j < FACT_HOP_REGS; ++j) {
          const int e = tid + j * FACT_THREADS;
          if (e < cnt) hw[e] = hreg[j];
        }
      } else {  // lists beyond the register budget: staged straight from memory
        const KeyInfo ki = skey[key];
        for (int r = tid; r <= S; r += FACT_THREADS) hptr[r] = A.hk_ptr[ki.pb + r] - ki.base;
        for (int e = tid; e < ki.cnt; e += FACT_THREADS) hw[e] = A.hk_words[ki.base + e];
      }
      cur_key = key;
    }
    __syncthreads();
    if (i + 1 < nst) prefetch(sst[i + 1]);  // in flight while this step computes
    if (A.hop_regs && i + 1 < nst) {
      const int nk = (int)(sst[i + 1].code >> 12);
      if (nk != key) hop_prefetch(nk);
    }
    // hop word w: target row in the high half; in the low half the byte offset of its
    // (signed) W entry for floats, half of it for doubles -- the W slices start at LDS
    // address 0, so the offset is the address
    typedef const __attribute__((address_space(3))) unsigned char lds_byte;
    const uint32_t lbase = (uint32_t)(uintptr_t)(lds_byte *)lds;
    const uint32_t tl = lbase + (uint32_t)A.lds_tile + sub * 16;
    auto wval = [&](uint32_t w) -> T {
      return *(const __attribute__((address_space(3))) T *)(lbase + (sizeof(T) == 8 ? (w & 0xffffu) << 1 : (w & 0xffffu)));
    };
    auto xrow = [&](uint32_t w) -> NV {
      return *(const __attribute__((address_space(3))) NV *)(tl + (w >> 16) * (LPE * 16));
    };
#pragma unroll
    for (int j = 0; j < RMAX; ++j) {
      const int r = g + j * G;
      if (r < S) {
        int e = hptr[r];
        const int e1 = hptr[r + 1];
        for (; e + 4 <= e1; e += 4) {
          uint32_t w[4];
          T h[4];
          NV xv[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) w[u] = hw[e + u];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            h[u] = wval(w[u]);
            xv[u] = xrow(w[u]);
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < VPL; ++v) acc[j][v] += h[u] * xv[u][v];
        }
        for (; e < e1; ++e) {
          const uint32_t w = hw[e];
          const T h = wval(w);
          const NV xv = xrow(w);
#pragma unroll
          for (int v = 0; v < VPL; ++v) acc[j][v] += h * xv[v];
        }
      }
    }
  }
  // epilogue: add the single-species products and store the columns below m
  const int64_t pr0 = A.poff[blk], qr0 = A.qoff[blk], qs = A.qnb[blk];
#pragma unroll
  for (int j = 0; j < RMAX; ++j) {
    const int r = g + j * G;
    if (r < S) {
      const int c = col0 + sub * VPL;
      const NV yn = *reinterpret_cast<const NV *>(Yp + (pr0 + r) * ld + c) +
                    *reinterpret_cast<const NV *>(Yq + (qr0 + r * qs) * ld + c);
      T *y = Y + (r0 + r) * ld + c;
      if (c + VPL <= m) {
        NV o;
#pragma unroll
        for (int v = 0; v < VPL; ++v) o[v] = acc[j][v] + yn[v];
        *reinterpret_cast<NV *>(y) = o;
      } else {
#pragma unroll
        for (int v = 0; v < VPL; ++v)
          if (c + v < m) y[v] = acc[j][v] + yn[v];
      }
    }
  }
}

// ---- single-species products on the matrix cores, read straight from X -----------------
// Per (M_n, parity) class c with nb blocks of S rows (block i of the class starts at basis
// row off[i]) two dense products, each written D = H^T X' with H square and column-major:
//   neutron (kind 0): K = nb, H = Hn_c;  X'(k, col) = X[off[k] ld + col], col < S ld
//                     (col = proton row * ld + vector element);  D(row, col) -> Yn[off[row] ld + col]
//   proton  (kind 1): K = S,  H = Hp_c;  X'(k, col) = X[(off[col / ld] + k) ld + col % ld],
//                     col < nb ld;       D(row, col) -> Yq[(off[col / ld] + row) ld + col % ld]
// so both outputs are in X's own row order and the step kernel adds them in its epilogue.
// One launch covers every product of every class: 64 x 64 output tiles (row tiles fastest,
// so neighbouring workgroups share X columns), 256 threads as 2 x 2 wavefronts of 32 x 32,
// each wavefront four 16x16x4 MFMA accumulators; K advances 16 at a time through one LDS
// buffer per operand, the next slab's loads in flight while the current one is
// multiplied.  The output's contiguous dimension is the MFMA column (lane % 16), so 16
// lanes write 16 consecutive elements; H is uploaded transposed, so both slabs are read
// along the output dimension (coalesced, conflict-free LDS stores).  Results equal the
// dense products up to summation order.  Measured at fp 4p4n k=16 (58 GFLOP per apply):
// FP32 0.74 ms, FP64 1.28 ms, against 2.1 / 3.7 ms of rocBLAS kernels (spread over four
// streams) plus the gather, 3.69 / 7.21 ms per whole apply against 3.60 / 7.15 ms.
constexpr int MM_T = 64;   // output tile edge
constexpr int MM_K = 16;   // K per LDS slab
constexpr int MM_P = 16;   // LDS row padding (elements): the four k rows of an MFMA operand read
                           // fall on disjoint banks
constexpr int MM_NLD = 7;  // tile tables for ld = 1, 2, 4, ..., 64
constexpr int MM_MAXK = 1024;     // largest class (blocks or proton rows) the kernel takes
constexpr int MM_MAXPROB = 512;   // products per launch (two per class)
constexpr int MM_XC = 16;         // consecutive tiles per XCD chunk

struct MMProb {
  int kind, K, S, nb;
  int tr;                  // row tiles (ceil(K / 64))
  int tstart[MM_NLD + 1];  // first workgroup of this product for ld = 1 << li (the last entry unused)
  const int64_t *off;      // [nb] first basis row of each block of the class
  const double *H64;       // H transposed: H64[k * K + row] = H(k, row), contiguous in row
  const float *H32;
};

template <typename T>
struct MmAcc;
template <>
struct MmAcc<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  // 16x16x4 FP32 result: register v of lane l holds row 4 * (l / 16) + v, column l % 16
  static __device__ __forceinline__ int row(int kq, int v) { return 4 * kq + v; }
  static __device__ __forceinline__ type mma(float a, float b, type c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
template <>
struct MmAcc<double> {
  typedef double type __attribute__((ext_vector_type(4)));
  // 16x16x4 FP64 result: register v of lane l holds row 4 * v + l / 16, column l % 16
  static __device__ __forceinline__ int row(int kq, int v) { return 4 * v + kq; }
  static __device__ __forceinline__ type mma(double a, double b, type c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

// the tile (row0, col0) of one product; KIND is the species (see above)
template <typename T, int KIND>
__device__ __forceinline__ void class_mm_tile(const MMProb &pr, int row0, int col0, int ld,
                                              const T *__restrict__ X, T *__restrict__ Yn, T *__restrict__ Yq,
                                              T (*Hs)[MM_T + MM_P], T (*Xs)[MM_T + MM_P], const int *soff) {
  constexpr int BK = MM_K, NJ = BK / 4;  // slab elements per thread of each operand
  typedef typename MmAcc<T>::type Acc;
  // global (not flat) loads: a flat load also counts as an LDS access, so the waits for
  // the LDS operands below would wait for the prefetched slab as well
  typedef const __attribute__((address_space(1))) T *GP;
  const int K = pr.K, ncol = (KIND == 0 ? pr.S : pr.nb) * ld;
  const GP Hg = (GP)(sizeof(T) == 8 ? (const T *)pr.H64 : (const T *)pr.H32), Xg = (GP)X;
  const int tid = threadIdx.x;
  // load roles: H slab element (k = tid % 16, row = tid / 16 + 16 j); X slab element
  // (col = tid % 64, k = tid / 64 + 4 j)
  const int xc = tid & 63, xk = tid >> 6;
  const int gcol = col0 + xc;
  const bool col_ok = gcol < ncol;
  int64_t xbase = 0;  // proton kind: X row of (block gcol / ld, proton row 0), element gcol % ld
  if (KIND == 1 && col_ok) xbase = pr.off[gcol / ld] * ld + gcol % ld;
  T hv[NJ], xv[NJ];
  bool hok[NJ], xok[NJ];
  int nk[NJ];  // neutron kind: the blocks' first rows of the next slab's k rows
  // slab elements of both operands: (k = tid / 64 + 4 j, row or column = tid % 64), so 64
  // lanes read 64 consecutive elements and store them to consecutive LDS words
  const int hr2 = row0 + xc;
  auto koff = [&](int k0) {
    if (KIND == 0)
#pragma unroll
      for (int j = 0; j < NJ; ++j) nk[j] = soff[min(k0 + xk + 4 * j, K - 1)];
  };
  // raw loads from clamped addresses; out-of-range elements are zeroed when stored
  auto load = [&](int k0) {
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int k = k0 + xk + 4 * j;
      hok[j] = k < K && hr2 < K;
      hv[j] = Hg[hok[j] ? (int64_t)k * K + hr2 : 0];
    }
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int k = k0 + xk + 4 * j;
      xok[j] = col_ok && k < K;
      const int64_t a = KIND == 0 ? (int64_t)nk[j] * ld + gcol : xbase + (int64_t)k * ld;
      xv[j] = Xg[xok[j] ? a : 0];
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int j = 0; j < NJ; ++j) Hs[xk + 4 * j][xc] = hok[j] ? hv[j] : T(0);
#pragma unroll
    for (int j = 0; j < NJ; ++j) Xs[xk + 4 * j][xc] = xok[j] ? xv[j] : T(0);
  };
  const int lane = tid & 63, w = tid >> 6, i = lane & 15, kq = lane >> 4;
  const int wr = (w & 1) * 32, wc = (w >> 1) * 32;
  Acc acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = Acc{0, 0, 0, 0};
  koff(0);
  load(0);
  store();
  koff(BK);
  __syncthreads();
  for (int k0 = 0; k0 < K; k0 += BK) {
    const bool more = k0 + BK < K;
    load(more ? k0 + BK : k0);  // unconditional: the loads fly during the products below
    __builtin_amdgcn_sched_barrier(0);  // and are issued before them
    T a[NJ][2], b[NJ][2];
#pragma unroll
    for (int kk = 0; kk < NJ; ++kk)
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        a[kk][s] = Hs[4 * kk + kq][wr + 16 * s + i];
        b[kk][s] = Xs[4 * kk + kq][wc + 16 * s + i];
      }
    koff(k0 + 2 * BK);  // read with the operands, used by the next load
#pragma unroll
    for (int kk = 0; kk < NJ; ++kk)
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int u = 0; u < 2; ++u) acc[s][u] = MmAcc<T>::mma(a[kk][s], b[kk][u], acc[s][u]);
    __builtin_amdgcn_sched_barrier(0);  // the products before the store that waits for the loads
    // one LDS buffer (two measured slower for FP64: fewer workgroups fit a CU): every
    // wavefront has read it before the next slab is stored.  The store is unconditional
    // (after the last slab it stores a copy nobody reads): a store under a branch would
    // pull the loads above into that branch, after the products.
    __syncthreads();
    store();
    __syncthreads();
  }
  // write D rows < K, columns < ncol
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int col = col0 + wc + 16 * u + i;
      if (col >= ncol) continue;
      // neutron kind: the column's offset in a block; proton kind: its (block, element) base
      const int64_t cb = KIND == 0 ? (int64_t)col : pr.off[col / ld] * ld + col % ld;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = row0 + wr + 16 * s + MmAcc<T>::row(kq, v);
        if (row >= K) continue;
        if (KIND == 0) Yn[(int64_t)soff[row] * ld + cb] = acc[s][u][v];
        else Yq[cb + (int64_t)row * ld] = acc[s][u][v];
      }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_class_mm(const MMProb *__restrict__ P, const int *__restrict__ first,
                                                  int nprob, int li, int ld, const T *__restrict__ X,
                                                  T *__restrict__ Yn, T *__restrict__ Yq) {
  __shared__ T Hs[MM_K][MM_T + MM_P];  // [k][output row]
  __shared__ T Xs[MM_K][MM_T + MM_P];  // [k][output column]
  __shared__ int soff[MM_MAXK];           // neutron kind: the first rows of the class's K blocks
  // the products' first tiles, read before the first H slab is stored over them
  int *sfirst = reinterpret_cast<int *>(&Hs[0][0]);
  static_assert(sizeof(Hs) >= sizeof(int) * MM_MAXPROB, "product table");
  const int tid = threadIdx.x;
  // tile order: the eight XCDs take chunks of MM_XC consecutive tiles in turn (row tiles are
  // fastest, so one XCD's L2 serves the X columns of its chunk to all its row tiles)
  int g = blockIdx.x;
  {
    const int G = gridDim.x, G8 = (G / (8 * MM_XC)) * (8 * MM_XC);
    if (g < G8) {
      const int x = g % 8, j = g / 8;
      g = ((j / MM_XC) * 8 + x) * MM_XC + j % MM_XC;
    }
  }
  // the product of this tile: the last one whose first tile is not after it
  for (int p = tid; p < nprob; p += 256) sfirst[p] = first[p];
  __syncthreads();
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sfirst[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  const MMProb pr = P[lo];
  const int t = g - sfirst[lo];
  const int row0 = (t % pr.tr) * MM_T, col0 = (t / pr.tr) * MM_T;
  if (pr.kind == 0)
    for (int k = tid; k < pr.K; k += 256) soff[k] = (int)pr.off[k];
  __syncthreads();  // soff written, sfirst read
  if (pr.kind == 0) class_mm_tile<T, 0>(pr, row0, col0, ld, X, Yn, Yq, Hs, Xs, soff);
  else class_mm_tile<T, 1>(pr, row0, col0, ld, X, Yn, Yq, Hs, Xs, soff);
}

// class-contiguous copies of X: neutron-major (block b's rows at row poff[b], for the Hn
// product) and proton-major (row pi of block b at row qoff[b] + pi * qnb[b], for Hp)
template <typename T>
__global__ __launch_bounds__(256) void k_class_gather(int nblk, const int64_t *__restrict__ noff,
                                                      const int64_t *__restrict__ poff,
                                                      const int64_t *__restrict__ qoff, const int32_t *__restrict__ qnb,
                                                      const T *__restrict__ X, T *__restrict__ Xp, T *__restrict__ Xq,
                                                      int ld) {
  const int b = blockIdx.x;
  if (b >= nblk) return;
  const int64_t len = (noff[b + 1] - noff[b]) * ld;
  const T *src = X + noff[b] * ld;
  T *dst = Xp + poff[b] * ld;
  const int64_t q0 = qoff[b], qs = qnb[b];
  for (int64_t i = threadIdx.x; i < len; i += 256) {
    const T x = src[i];
    dst[i] = x;
    Xq[(q0 + (i / ld) * qs) * ld + i % ld] = x;
  }
}

}  // namespace

// ------------------------------------------------------------------------------------
struct FactOp {
  FactArgs A{};
  std::vector<void *> owned;  // device allocations
  const double *W64 = nullptr;
  const float *W32 = nullptr;
  int maxS = 0, hopcap = 0, np2 = 0, maxsteps = 0;
  // one (M_n, parity) neutron class: nb blocks of S rows, contiguous from row prow of the
  // class-contiguous layout; dense Hn (nb x nb) and the dense Hp of its proton sector (S x S)
  struct Cls {
    int nb = 0, S = 0;
    int64_t prow = 0;
    const double *H64 = nullptr, *P64 = nullptr;
    const float *H32 = nullptr, *P32 = nullptr;
  };
  std::vector<Cls> cls;
  void *Xp = nullptr, *Yp = nullptr, *Xq = nullptr, *Yq = nullptr;
  size_t p_bytes = 0;
  // the class products of k_class_mm: one table entry per (class, species), with the
  // first workgroup of each for every ld = 1 << li; k_fact then reads Yp, Yq in X's row
  // order (An: A with poff = qoff = the block's first row and qnb = 1)
  const MMProb *mm = nullptr;
  const int *mm_first[MM_NLD] = {};  // tstart of every product for each ld, contiguous
  int nprob = 0;
  int mm_total[MM_NLD] = {};
  FactArgs An{};
  bool use_mm = true;  // COSMO_CLASS_MM=0: the gathered copies and rocBLAS instead
  // The class GEMMs are independent of each other: they are spread over FACT_AUX streams
  // (one rocBLAS handle each), forked from and joined back into the caller's stream.
  static constexpr int FACT_AUX = 4;
  rocblas_handle blas[FACT_AUX] = {};
  hipStream_t aux[FACT_AUX] = {};
  hipEvent_t fork = nullptr, join[FACT_AUX] = {};
  std::map<int, int> tuned;  // (element size, ld) -> k_fact slice width (lanes per row)
  // The whole apply (gather, forked GEMMs, k_fact) captured as a HIP graph per (element
  // size, ld, m, X, Y) once the same call repeats, then replayed by one launch.
  static constexpr size_t FACT_GRAPHS = 64;
  using GraphKey = std::tuple<int, int, int, const void *, const void *>;
  std::map<GraphKey, hipGraphExec_t> graphs;
  std::map<GraphKey, int> seen;  // calls per key; -1: capture failed, stay eager
  void drop_graphs() {
    for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
    graphs.clear();
    seen.clear();
  }
  int64_t n = 0;
  size_t bytes = 0;
  ~FactOp() {
    drop_graphs();
    for (void *p : owned) dfree(p);
    dfree(Xp);
    dfree(Yp);
    dfree(Xq);
    dfree(Yq);
    for (int j = 0; j < FACT_AUX; ++j) {
      if (blas[j]) rocblas_destroy_handle(blas[j]);
      if (join[j]) (void)hipEventDestroy(join[j]);
      if (aux[j]) (void)hipStreamDestroy(aux[j]);
    }
    if (fork) (void)hipEventDestroy(fork);
  }
  template <typename T>
  int put(const std::vector<T> &h, const T **out) {
    void *d = nullptr;
    const size_t b = std::max<size_t>(h.size(), 1) * sizeof(T);
    if (dmalloc(&d, b) != hipSuccess) return fail(COSMO_ENOMEM, "factorised operator: table allocation failed");
    owned.push_back(d);
    bytes += b;
    if (!h.empty() && (hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, lib_stream()) !=
                           hipSuccess ||
                       hipStreamSynchronize(lib_stream()) != hipSuccess))
      return fail(COSMO_EHIP, "factorised operator: table upload failed");
    *out = (const T *)d;
    return COSMO_OK;
  }
};

// Full symmetric rows of a species-only subsystem (upper rows from the GPU builder).
static int subsystem_rows(const System &sub, std::vector<int64_t> &rp, std::vector<int64_t> &cols,
                          std::vector<double> &vals) {
  cosmo_system_s tmp;
  static_cast<System &>(tmp) = sub;
  cosmo_op op = nullptr;
  int rc = cosmo_op_create(&tmp, COSMO_OP_NO_FP32 | COSMO_OP_KEEP_SMALL, &op);
  if (rc) return rc;
  const int64_t n = cosmo_op_size(op);
  std::vector<int64_t> cnt(n), up(n + 1, 0);
  rc = cosmo_op_upper_rows(op, 0, n, cnt.data(), nullptr, nullptr, nullptr);
  for (int64_t r = 0; r < n && !rc; ++r) up[r + 1] = up[r] + cnt[r];
  std::vector<int64_t> uc(up[n]);
  std::vector<double> uv(up[n]);
  if (!rc) rc = cosmo_op_upper_rows(op, 0, n, cnt.data(), up.data(), uc.data(), uv.data());
  cosmo_op_destroy(op);
  if (rc) return rc;
  std::vector<int64_t> deg(n, 0);
  for (int64_t r = 0; r < n; ++r)
    for (int64_t e = up[r]; e < up[r + 1]; ++e) {
      ++deg[r];
      if (uc[e] != r) ++deg[uc[e]];
    }
  rp.assign(n + 1, 0);
  for (int64_t r = 0; r < n; ++r) rp[r + 1] = rp[r] + deg[r];
  cols.assign(rp[n], 0);
  vals.assign(rp[n], 0.0);
  std::vector<int64_t> fill(rp.begin(), rp.end() - 1);
  for (int64_t r = 0; r < n; ++r)
    for (int64_t e = up[r]; e < up[r + 1]; ++e) {
      const int64_t c = uc[e];
      cols[fill[r]] = c;
      vals[fill[r]++] = uv[e];
      if (c != r) {
        cols[fill[c]] = r;
        vals[fill[c]++] = uv[e];
      }
    }
  return COSMO_OK;
}

static inline int popc(uint32_t x) { return __builtin_popcount(x); }
static inline uint32_t below32(int i) { return (1u << i) - 1u; }

// Upper bound of the device bytes of a product-form operator of s together with the
// class copies of X and Y its applies allocate at the widest Davidson pitch (16 FP64
// vectors); -1 when s has no product form (one species, or terms that move a particle
// between species, or sectors or dense class blocks beyond what build_factored takes).
int64_t factored_footprint(const System &s) {
  const int nn = s.nn, np = s.np, nblk = (int)s.nmask.size();
  if (s.nv < 1 || s.zv < 1 || nn > 31 || np > 31 || s.truncated()) return -1;
  for (size_t t = 0; t < s.esp_v.size(); ++t)
    if ((s.esp_k[2 * t] < nn) != (s.esp_k[2 * t + 1] < nn)) return -1;
  for (size_t t = 0; t < s.vspx_v.size(); ++t) {
    const uint16_t *k = &s.vspx_k[4 * t];
    if ((k[0] < nn) + (k[1] < nn) != (k[2] < nn) + (k[3] < nn)) return -1;
  }
  const int64_t n = s.dim;
  const int nsec = (int)s.psec_start.size() - 1;
  int64_t maxS = 0, dense = 0, hops = 0;
  std::vector<int64_t> nb(nsec, 0);
  for (int b = 0; b < nblk; ++b) {
    maxS = std::max<int64_t>(maxS, s.noff[b + 1] - s.noff[b]);
    if (s.nsector[b] >= 0 && s.nsector[b] < nsec) ++nb[s.nsector[b]];
  }
  if (maxS > 4096) return -1;
  const int m2 = 2 * 31 + 1;
  for (int sec = 0; sec < nsec; ++sec) {
    const int64_t S = s.psec_start[sec + 1] - s.psec_start[sec];
    dense += (S * S + nb[sec] * nb[sec]) * 12;  // Hp and Hn, FP64 + FP32
    hops += S * ((int64_t)s.zv * (np - s.zv + 1) + 2 * m2 + 2) * 4 + 16;
  }
  if (dense > ((int64_t)4 << 30)) return -1;
  int64_t wneg = 16;
  while (wneg < (int64_t)np * np) wneg <<= 1;
  const int64_t W = (int64_t)nn * nn * wneg * 2 * 12;
  const int64_t steps = (int64_t)nblk * (1 + (int64_t)s.nv * (nn - s.nv + 1)) * 12 + 64 * (int64_t)nblk;
  const int64_t tables = (int64_t)(nn + np) * (nn + np) * (nn + np) * (nn + np) * 8 + ((int64_t)1 << 20);
  return 4 * n * 16 * 8 + 8 * n + dense + hops + W + steps + tables;
}

int build_factored(Op &op, const System &s) {
  const int nn = s.nn, np = s.np;
  const int nblk = (int)s.nmask.size();
  if (s.nv < 1 || s.zv < 1) return fail(COSMO_EUNSUP, "factorised operator needs valence neutrons and protons");
  // the product form needs every block to be its whole proton sector
  if (s.truncated()) return fail(COSMO_EUNSUP, "factorised operator: truncated bases use the stored or row form");
  if (nn > 31 || np > 31) return fail(COSMO_EUNSUP, "factorised operator: more than 31 orbitals of a species");
  // every term must keep the neutron and proton counts
  for (size_t t = 0; t < s.esp_v.size(); ++t)
    if ((s.esp_k[2 * t] < nn) != (s.esp_k[2 * t + 1] < nn))
      return fail(COSMO_EUNSUP, "factorised operator: one-body term changes the species");
  for (size_t t = 0; t < s.vspx_v.size(); ++t) {
    const uint16_t *k = &s.vspx_k[4 * t];
    if ((k[0] < nn) + (k[1] < nn) != (k[2] < nn) + (k[3] < nn))
      return fail(COSMO_EUNSUP, "factorised operator: two-body term changes the species");
  }
  auto F = std::make_unique<FactOp>();
  FactOp &f = *F;
  f.n = s.dim;
  f.np2 = np * np;
  const int nsec = (int)s.psec_start.size() - 1;
  int maxS = 0;
  for (int b = 0; b < nblk; ++b) maxS = std::max<int>(maxS, (int)(s.noff[b + 1] - s.noff[b]));
  if (maxS > 4096) return fail(COSMO_EUNSUP, "factorised operator: proton sector above 4096 states");
  f.maxS = maxS;
  // keys: (proton delta 2m, delta parity)
  int m2max = 1;
  for (int i = nn; i < nn + np; ++i) m2max = std::max(m2max, std::abs(s.orb[i].m2));
  const int dmin = -2 * m2max;
  const int NK = 2 * (2 * m2max + 1);
  auto key_of = [&](int dm2, int dpar) { return ((dm2 - dmin) / 2) * 2 + dpar; };
  std::vector<int32_t> psec_of(s.pcfg_rank.size(), -1);
  for (int sec = 0; sec < nsec; ++sec)
    for (int64_t q = s.psec_start[sec]; q < s.psec_start[sec + 1]; ++q) psec_of[colex_rank(s.pmask[q])] = sec;

  // ---- proton hop lists per (sector, key), CSR over the sector's rows
  std::vector<int64_t> hk_pbeg((size_t)nsec * NK);
  std::vector<int32_t> hk_ptr;
  std::vector<uint32_t> hk_words;
  int wneg = 16;  // the negated W slice follows at a power of two >= np*np
  while (wneg < np * np) wneg <<= 1;
  int hopcap = 0;
  for (int sec = 0; sec < nsec; ++sec) {
    const int64_t qa = s.psec_start[sec], qb = s.psec_start[sec + 1];
    std::vector<std::vector<std::vector<uint32_t>>> rows(NK, std::vector<std::vector<uint32_t>>(qb - qa));
    for (int64_t q = qa; q < qb; ++q) {
      const uint32_t P = s.pmask[q];
      for (uint32_t bb = P; bb; bb &= bb - 1) {
        const int b = __builtin_ctz(bb);
        for (int d = 0; d < np; ++d) {
          if (d != b && ((P >> d) & 1u)) continue;
          const uint32_t P2 = (P & ~(1u << b)) | (1u << d);
          if (psec_of[colex_rank(P2)] < 0) continue;
          const int t = s.pcfg_rank[colex_rank(P2)];
          const int sign = (popc(P & below32(b)) + popc((P & ~(1u << b)) & below32(d))) & 1;
          const Orbital &ob = s.orb[nn + b], &od = s.orb[nn + d];
          const int key = key_of(od.m2 - ob.m2, ob.par ^ od.par);
          rows[key][q - qa].push_back((uint32_t)t << 16 | (uint32_t)(4 * (b * np + d + sign * wneg)));
        }
      }
    }
    for (int k = 0; k < NK; ++k) {
      hk_pbeg[(size_t)sec * NK + k] = (int64_t)hk_ptr.size();
      const size_t base = hk_words.size();
      for (auto &r : rows[k]) {
        hk_ptr.push_back((int32_t)hk_words.size());
        hk_words.insert(hk_words.end(), r.begin(), r.end());
      }
      hk_ptr.push_back((int32_t)hk_words.size());
      hopcap = std::max<int>(hopcap, (int)(hk_words.size() - base));
    }
  }
  hk_words.push_back(0);  // spare entry: clamped prefetch reads of an empty last list
  f.hopcap = hopcap;

  // ---- pn table W[a*nn + c][b*np + d] from the forward two-body keys and their mirrors
  std::vector<double> W((size_t)nn * nn * f.np2, 0.0);
  for (size_t t = 0; t < s.vspx_v.size(); ++t) {
    const uint16_t *k = &s.vspx_k[4 * t];
    if (k[2] < k[0] || (k[2] == k[0] && k[3] < k[1])) continue;  // not forward: dropped
    if (!(k[0] < nn && k[1] >= nn)) continue;                     // not a pn pair
    const int a = k[0], b = k[1] - nn, c = k[2], d = k[3] - nn;
    W[(size_t)(a * nn + c) * f.np2 + b * np + d] = s.vspx_v[t];
    W[(size_t)(c * nn + a) * f.np2 + d * np + b] = s.vspx_v[t];
  }
  std::vector<uint8_t> wnz((size_t)nn * nn, 0);
  for (int ac = 0; ac < nn * nn; ++ac)
    for (int bd = 0; bd < f.np2; ++bd)
      if (W[(size_t)ac * f.np2 + bd] != 0.0) wnz[ac] = 1;

  // ---- steps per block: the block itself, then every neutron hop a -> c, sorted by key
  const int key0 = key_of(0, 0);
  std::vector<int32_t> step_start(nblk + 1, 0), step_tb;
  std::vector<uint32_t> step_code;
  std::vector<std::pair<int, std::pair<int32_t, uint32_t>>> st;
  for (int blk = 0; blk < nblk; ++blk) {
    const uint32_t nu = s.nmask[blk];
    st.clear();
    st.push_back({key0 * 2, {blk, (1u << 11) | (uint32_t)key0 << 12}});
    for (uint32_t aa = nu; aa; aa &= aa - 1) {
      const int a = __builtin_ctz(aa);
      for (int c = 0; c < nn; ++c) {
        if (((nu >> c) & 1u) || !wnz[a * nn + c]) continue;
        const uint32_t nu2 = (nu & ~(1u << a)) | (1u << c);
        const int tb = s.ncfg_index[colex_rank(nu2)];
        if (tb < 0) continue;
        const int dm2 = s.orb[a].m2 - s.orb[c].m2;  // the proton side changes 2m by this
        if (dm2 < dmin || dm2 > -dmin) continue;
        const int key = key_of(dm2, s.orb[a].par ^ s.orb[c].par);
        const int ph = (popc(nu & below32(a)) + popc((nu & ~(1u << a)) & below32(c))) & 1;
        st.push_back({key * 2 + 1, {tb, (uint32_t)a | (uint32_t)c << 5 | (uint32_t)ph << 10 | (uint32_t)key << 12}});
      }
    }
    std::stable_sort(st.begin(), st.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    for (auto &x : st) {
      step_tb.push_back(x.second.first);
      step_code.push_back(x.second.second);
    }
    step_start[blk + 1] = (int32_t)step_tb.size();
    f.maxsteps = std::max<int>(f.maxsteps, (int)st.size());
  }

  // ---- Hp: dense per proton sector, from the proton-only subsystem (no neutrons)
  std::vector<std::vector<double>> hp_dense(nsec);
  size_t dense_bytes = 0;
  for (int sec = 0; sec < nsec; ++sec) {
    const size_t S = (size_t)(s.psec_start[sec + 1] - s.psec_start[sec]);
    dense_bytes += S * S * 12;
  }
  for (int sec = 0; sec < nsec; ++sec) {
    const int64_t qa = s.psec_start[sec], S = s.psec_start[sec + 1] - qa;
    if (dense_bytes > ((size_t)2 << 30)) break;  // reported below
    int M = 0, par = 0;
    for (uint32_t bb = s.pmask[qa]; bb; bb &= bb - 1) {
      M += s.orb[nn + __builtin_ctz(bb)].m2;
      par ^= s.orb[nn + __builtin_ctz(bb)].par;
    }
    System sub = s;
    sub.nv = 0;
    sub.twoM = M;
    sub.parity = par;
    sub.build_basis();
    // the subsystem enumerates the proton configurations exactly as s does: its one block
    // (no neutrons) must be this sector
    if (sub.dim != S || sub.nsector.size() != 1 || sub.nsector[0] != sec || sub.pmask != s.pmask)
      return fail(COSMO_EINVAL, "factorised operator: proton subsystem differs from its sector");
    std::vector<int64_t> rp, cc;
    std::vector<double> vv;
    int rc = subsystem_rows(sub, rp, cc, vv);
    if (rc) return rc;
    hp_dense[sec].assign((size_t)S * S, 0.0);
    for (int64_t r = 0; r < S; ++r)
      for (int64_t e = rp[r]; e < rp[r + 1]; ++e) hp_dense[sec][(size_t)cc[e] * S + r] = vv[e];
  }

  // ---- Hn: dense per (M_n, parity) class, and the class-contiguous block order
  std::map<std::pair<int, int>, std::vector<int>> byclass;
  for (int b = 0; b < nblk; ++b) {
    int M = 0, par = 0;
    for (uint32_t bb = s.nmask[b]; bb; bb &= bb - 1) {
      M += s.orb[__builtin_ctz(bb)].m2;
      par ^= s.orb[__builtin_ctz(bb)].par;
    }
    byclass[{M, par}].push_back(b);
  }
  for (auto &kv : byclass) dense_bytes += kv.second.size() * kv.second.size() * 12;
  if (dense_bytes > ((size_t)2 << 30))
    return fail(COSMO_EUNSUP, "factorised operator: single-species blocks too large for dense products");
  std::vector<int64_t> poff(nblk, 0), qoff(nblk, 0);
  std::vector<int32_t> qnb(nblk, 0);
  std::vector<int64_t> cls_off;  // the blocks' first rows, class after class
  std::vector<std::vector<double>> cls_hn, cls_hp;  // each class's dense Hn and Hp (column-major)
  std::vector<size_t> cls_off_at;
  int64_t prow = 0;
  for (auto &kv : byclass) {
    System sub = s;
    sub.zv = 0;
    sub.twoM = kv.first.first;
    sub.parity = kv.first.second;
    sub.build_basis();
    const int N = (int)kv.second.size();
    bool same = sub.dim == N;
    for (int i = 0; same && i < N; ++i) same = sub.nmask[i] == s.nmask[kv.second[i]];
    if (!same) return fail(COSMO_EINVAL, "factorised operator: neutron subsystem differs from its class");
    std::vector<int64_t> rp, cc;
    std::vector<double> vv;
    int rc = subsystem_rows(sub, rp, cc, vv);
    if (rc) return rc;
    std::vector<double> Hd((size_t)N * N, 0.0);
    for (int r = 0; r < N; ++r)
      for (int64_t e = rp[r]; e < rp[r + 1]; ++e) Hd[(size_t)cc[e] * N + r] = vv[e];
    std::vector<float> Hf(Hd.begin(), Hd.end());
    cls_hn.push_back(Hd);
    cls_hp.push_back(hp_dense[s.nsector[kv.second[0]]]);
    FactOp::Cls c;
    c.nb = N;
    c.S = (int)(s.noff[kv.second[0] + 1] - s.noff[kv.second[0]]);
    c.prow = prow;
    cls_off_at.push_back(cls_off.size());
    for (int i = 0; i < N; ++i) {
      const int b = kv.second[i];
      poff[b] = prow + (int64_t)i * c.S;
      qoff[b] = prow + i;
      qnb[b] = N;
      cls_off.push_back(s.noff[b]);
    }
    prow += (int64_t)N * c.S;
    const int sec = s.nsector[kv.second[0]];
    std::vector<float> Pf(hp_dense[sec].begin(), hp_dense[sec].end());
    if ((rc = f.put(Hd, &c.H64)) || (rc = f.put(Hf, &c.H32)) || (rc = f.put(hp_dense[sec], &c.P64)) ||
        (rc = f.put(Pf, &c.P32)))
      return rc;
    f.cls.push_back(c);
  }

  // ---- upload
  std::vector<float> W32(W.begin(), W.end());
  FactArgs &A = f.A;
  A.nblk = nblk;
  A.nn = nn;
  A.np2 = f.np2;
  A.NK = NK;
  A.wneg = wneg;
  int rc;
  if ((rc = f.put(s.noff, &A.noff)) || (rc = f.put(s.nsector, &A.nsector)) || (rc = f.put(s.nmask, &A.nmask)) ||
      (rc = f.put(poff, &A.poff)) || (rc = f.put(qoff, &A.qoff)) || (rc = f.put(qnb, &A.qnb)) || (rc = f.put(step_start, &A.step_start)) || (rc = f.put(step_tb, &A.step_tb)) ||
      (rc = f.put(step_code, &A.step_code)) || (rc = f.put(hk_pbeg, &A.hk_pbeg)) || (rc = f.put(hk_ptr, &A.hk_ptr)) ||
      (rc = f.put(hk_words, &A.hk_words)) || (rc = f.put(W, &f.W64)) || (rc = f.put(W32, &f.W32)))
    return rc;
  // class products: tables for k_class_mm and the row-order epilogue offsets of k_fact
  {
    const int64_t *doff = nullptr;
    const int32_t *ones = nullptr;
    if ((rc = f.put(cls_off, &doff)) || (rc = f.put(std::vector<int32_t>(nblk, 1), &ones))) return rc;
    std::vector<MMProb> mp;
    for (size_t ci = 0; ci < f.cls.size(); ++ci) {
      const FactOp::Cls &c = f.cls[ci];
      for (int kind = 0; kind < 2; ++kind) {
        // the class matrix transposed (k_class_mm reads H(k, row) along row)
        const std::vector<double> &Hd = kind == 0 ? cls_hn[ci] : cls_hp[ci];
        const int K = kind == 0 ? c.nb : c.S;
        std::vector<double> T64((size_t)K * K);
        for (int r = 0; r < K; ++r)
          for (int k = 0; k < K; ++k) T64[(size_t)k * K + r] = Hd[(size_t)r * K + k];
        std::vector<float> T32(T64.begin(), T64.end());
        const double *t64 = nullptr;
        const float *t32 = nullptr;
        if ((rc = f.put(T64, &t64)) || (rc = f.put(T32, &t32))) return rc;
        MMProb q{};
        q.kind = kind;
        q.K = kind == 0 ? c.nb : c.S;
        q.S = c.S;
        q.nb = c.nb;
        q.tr = (q.K + MM_T - 1) / MM_T;
        q.off = doff + cls_off_at[ci];
        q.H64 = t64;
        q.H32 = t32;
        mp.push_back(q);
      }
    }
    for (int li = 0; li < MM_NLD; ++li) {
      int64_t t = 0;
      for (MMProb &q : mp) {
        q.tstart[li] = (int)t;
        const int64_t ncol = (int64_t)(q.kind == 0 ? q.S : q.nb) << li;
        t += (int64_t)q.tr * ((ncol + MM_T - 1) / MM_T);
      }
      if (t > INT32_MAX) return fail(COSMO_EUNSUP, "factorised operator: too many class-product tiles");
      f.mm_total[li] = (int)t;
    }
    f.nprob = (int)mp.size();
    if ((rc = f.put(mp, &f.mm))) return rc;
    for (int li = 0; li < MM_NLD; ++li) {
      std::vector<int> fi;
      for (const MMProb &q : mp) fi.push_back(q.tstart[li]);
      if ((rc = f.put(fi, &f.mm_first[li]))) return rc;
    }
    int maxk = 0;
    for (const MMProb &q : mp) maxk = std::max(maxk, q.kind == 0 ? q.K : 0);
    // beyond these the gathered copies and rocBLAS do the class products
    if (maxk > MM_MAXK || f.nprob > MM_MAXPROB || s.noff[nblk] >= INT32_MAX) f.use_mm = false;
    f.An = A;
    f.An.poff = A.noff;
    f.An.qoff = A.noff;
    f.An.qnb = ones;
    if (const char *e = getenv("COSMO_CLASS_MM")) f.use_mm = f.use_mm && atoi(e) != 0;
  }
  HIPCHK(hipEventCreateWithFlags(&f.fork, hipEventDisableTiming));
  for (int j = 0; j < FactOp::FACT_AUX; ++j) {
    HIPCHK(hipStreamCreateWithFlags(&f.aux[j], hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&f.join[j], hipEventDisableTiming));
    if (rocblas_create_handle(&f.blas[j]) != rocblas_status_success ||
        rocblas_set_stream(f.blas[j], f.aux[j]) != rocblas_status_success)
      return fail(COSMO_EHIP, "factorised operator: rocblas handle setup failed");
  }
  op.fact = F.release();
  return COSMO_OK;
}

void free_factored(Op &op) {
  delete static_cast<FactOp *>(op.fact);
  op.fact = nullptr;
}

size_t factored_bytes(const Op &op) {
  const FactOp *f = static_cast<const FactOp *>(op.fact);
  return f ? f->bytes + (f->use_mm ? 2 : 4) * f->p_bytes : 0;
}

template <typename T, int LPE>
static bool launch_fact_r(int rmax, dim3 g, size_t lds, hipStream_t st, const FactArgs &A, const T *W, const T *X,
                          const T *Yp, const T *Yq, T *Y, int ld, int m) {
#define FACT_R(R)                                                                                    \
  if (rmax <= R) {                                                                                   \
    auto kern = k_fact<T, LPE, R>;                                                                   \
    if (lds > 65536)                                                                                 \
      hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
    hipLaunchKernelGGL(kern, g, dim3(FACT_THREADS), lds, st, A, W, X, Yp, Yq, Y, ld, m);            \
    return true;                                                                                     \
  }
  FACT_R(1)
  FACT_R(2)
  FACT_R(3)
  FACT_R(4)
  FACT_R(5)
  FACT_R(6)
  FACT_R(8)
  FACT_R(12)
  FACT_R(16)
#undef FACT_R
  return false;
}

static inline int align16(size_t b) { return (int)((b + 15) & ~size_t(15)); }

// k_fact launch for one slice width (lpe lanes of 16 bytes per row); 1 = launched,
// 0 = this width does not fit, <0 = error
template <typename T>
static int launch_fact(const FactOp &f, int lpe, const T *X, const T *Yp, const T *Yq, T *Y, int m, int ld,
                       hipStream_t st) {
  constexpr int VPL = Piece16<T>::N;
  if (lpe * VPL > ld || (size_t)f.maxS * lpe * 16 > (lpe == 8 ? 96 : 64) * 1024) return 0;
  const int G = FACT_THREADS / lpe;
  const int rmax = (f.maxS + G - 1) / G;
  const int slices = ld / (lpe * VPL);
  FactArgs A = f.use_mm ? f.An : f.A;
  A.lds_tile = align16((size_t)2 * A.wneg * sizeof(T));
  A.lds_st = A.lds_tile + align16((size_t)f.maxS * lpe * 16);
  A.lds_key = A.lds_st + align16((size_t)f.maxsteps * sizeof(StepInfo));
  A.lds_hptr = A.lds_key + align16((size_t)A.NK * sizeof(KeyInfo));
  A.lds_hw = A.lds_hptr + align16((size_t)(f.maxS + 1) * 4);
  const size_t lds = (size_t)A.lds_hw + (size_t)f.hopcap * 4;
  A.hop_regs = f.hopcap <= FACT_HOP_REGS * FACT_THREADS;
  if (lds > 160 * 1024 || f.np2 > 2 * FACT_THREADS) return 0;
  const T *W = sizeof(T) == 8 ? (const T *)f.W64 : (const T *)f.W32;
  dim3 g((unsigned)A.nblk, (unsigned)slices);
  bool ok;
  if (lpe == 8) ok = launch_fact_r<T, 8>(rmax, g, lds, st, A, W, X, Yp, Yq, Y, ld, m);
  else if (lpe == 4) ok = launch_fact_r<T, 4>(rmax, g, lds, st, A, W, X, Yp, Yq, Y, ld, m);
  else if (lpe == 2) ok = launch_fact_r<T, 2>(rmax, g, lds, st, A, W, X, Yp, Yq, Y, ld, m);
  else ok = launch_fact_r<T, 1>(rmax, g, lds, st, A, W, X, Yp, Yq, Y, ld, m);
  if (!ok) return 0;
  HIPCHK(hipGetLastError());
  return 1;
}

// the class products of every class, Yp (neutron) and Yq (proton) in X's row order
template <typename T>
static int launch_class_mm(const FactOp &f, const T *X, int ld, T *Yp, T *Yq, hipStream_t st) {
  const int li = __builtin_ctz((unsigned)ld);
  if (li >= MM_NLD) return fail(COSMO_EUNSUP, "factorised apply: row pitch above 64");
  if (f.mm_total[li] > 0)
    hipLaunchKernelGGL(k_class_mm<T>, dim3((unsigned)f.mm_total[li]), dim3(256), 0, st, f.mm, f.mm_first[li],
                       f.nprob, li, ld, X, Yp, Yq);
  HIPCHK(hipGetLastError());
  return COSMO_OK;
}

// Everything one apply puts on the stream (no waits, no allocation), up to the join of the
// forked GEMMs; k_fact follows with the chosen slice width.
template <typename T>
static int fact_enqueue_single(FactOp &f, const T *X, int ld, hipStream_t st) {
  T *Xp = (T *)f.Xp, *Yp = (T *)f.Yp, *Xq = (T *)f.Xq, *Yq = (T *)f.Yq;
  if (f.use_mm) return launch_class_mm<T>(f, X, ld, Yp, Yq, st);
  hipLaunchKernelGGL(k_class_gather<T>, dim3((unsigned)f.A.nblk), dim3(256), 0, st, f.A.nblk, f.A.noff, f.A.poff,
                     f.A.qoff, f.A.qnb, X, Xp, Xq, ld);
  HIPCHK(hipGetLastError());
  // fork: the GEMMs start once the gather is done
  HIPCHK(hipEventRecord(f.fork, st));
  for (int j = 0; j < FactOp::FACT_AUX; ++j) HIPCHK(hipStreamWaitEvent(f.aux[j], f.fork, 0));
  const T one = 1, zero = 0;
  const auto N_ = rocblas_operation_none;
  int gi = 0;  // GEMMs dealt round-robin over the auxiliary streams
  for (auto &c : f.cls) {
    const rocblas_int L = (rocblas_int)((int64_t)c.S * ld), N = c.nb, S = c.S;
    const rocblas_int M = (rocblas_int)((int64_t)c.nb * ld);
    T *xc = Xp + c.prow * ld, *yc = Yp + c.prow * ld, *xq = Xq + c.prow * ld, *yq = Yq + c.prow * ld;
    rocblas_handle h1 = f.blas[gi++ % FactOp::FACT_AUX], h2 = f.blas[gi++ % FactOp::FACT_AUX];
    rocblas_status b1, b2;
    if constexpr (sizeof(T) == 8) {
      b1 = rocblas_dgemm(h1, N_, N_, L, N, N, &one, xc, L, c.H64, N, &zero, yc, L);
      b2 = rocblas_dgemm(h2, N_, N_, M, S, S, &one, xq, M, c.P64, S, &zero, yq, M);
    } else {
      b1 = rocblas_sgemm(h1, N_, N_, L, N, N, &one, xc, L, c.H32, N, &zero, yc, L);
      b2 = rocblas_sgemm(h2, N_, N_, M, S, S, &one, xq, M, c.P32, S, &zero, yq, M);
    }
    if (b1 != rocblas_status_success || b2 != rocblas_status_success)
      return fail(COSMO_EHIP, "factorised apply: rocblas gemm failed");
  }
  // join: k_fact reads every class's Yp and Yq
  for (int j = 0; j < FactOp::FACT_AUX; ++j) {
    HIPCHK(hipEventRecord(f.join[j], f.aux[j]));
    HIPCHK(hipStreamWaitEvent(st, f.join[j], 0));
  }
  return COSMO_OK;
}

// Captures one apply (gather, forked GEMMs, k_fact) on st into an executable graph.
// False, with the stream out of capture mode and nothing kept, when the capture fails or
// holds a node other than kernels, event waits and empty joins (a memset node in
// particular: they replay unreliably here).
template <typename T>
static bool fact_capture(FactOp &f, int lpe, const T *X, const T *Yp, const T *Yq, T *Y, int m, int ld,
                         hipStream_t st, hipGraphExec_t *out) {
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  int r = fact_enqueue_single<T>(f, X, ld, st);
  if (r == COSMO_OK) r = launch_fact<T>(f, lpe, X, Yp, Yq, Y, m, ld, st) == 1 ? COSMO_OK : COSMO_EHIP;
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(st, &graph);
  bool ok = r == COSMO_OK && ec == hipSuccess && graph;
  if (ok) {
    size_t nn = 0;
    ok = hipGraphGetNodes(graph, nullptr, &nn) == hipSuccess && nn > 0;
    std::vector<hipGraphNode_t> nodes(nn);
    if (ok) ok = hipGraphGetNodes(graph, nodes.data(), &nn) == hipSuccess;
    for (size_t i = 0; ok && i < nn; ++i) {
      hipGraphNodeType t;
      ok = hipGraphNodeGetType(nodes[i], &t) == hipSuccess &&
           (t == hipGraphNodeTypeKernel || t == hipGraphNodeTypeEmpty || t == hipGraphNodeTypeWaitEvent ||
            t == hipGraphNodeTypeEventRecord);
    }
  }
  if (ok) ok = hipGraphInstantiate(out, graph, nullptr, nullptr, 0) == hipSuccess;
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipGetLastError();
  return ok;
}

template <typename T>
static int fact_apply_t(Op &op, const T *X, T *Y, int m, int ld, hipStream_t st) {
  FactOp &f = *static_cast<FactOp *>(op.fact);
  constexpr int VPL = Piece16<T>::N;
  if (ld % VPL || (ld & (ld - 1)) || ((uintptr_t)X | (uintptr_t)Y) % 16)
    return fail(COSMO_EUNSUP, "factorised apply: leading dimension must be a power of two of >= 16 bytes");
  // single-species parts on the class-contiguous copy of X:
  //   Yp_class = Xp_class Hn_class   (neutron-major: S*ld x nb, one column per block)
  //   Yq_class = Xq_class Hp_sector  (proton-major: nb*ld x S, one column per proton row)
  const size_t need = (size_t)f.n * ld * sizeof(T);
  if (f.p_bytes < need) {
    dfree(f.Xp);
    dfree(f.Yp);
    dfree(f.Xq);
    dfree(f.Yq);
    f.Xp = f.Yp = f.Xq = f.Yq = nullptr;
    f.p_bytes = 0;
    if (!f.use_mm) {  // the class-contiguous copies of X
      HIPCHK(dmalloc(&f.Xp, need, st));
      HIPCHK(dmalloc(&f.Xq, need, st));
    }
    HIPCHK(dmalloc(&f.Yp, need, st));
    HIPCHK(dmalloc(&f.Yq, need, st));
    f.p_bytes = need;
    f.drop_graphs();  // captured graphs point at the old buffers
  }
  const T *Yp = (const T *)f.Yp, *Yq = (const T *)f.Yq;
  // Slice width (1, 2, 4 or 8 lanes of 16 bytes per row): tuned once per (precision, ld)
  // by timing every width that fits; the widths give bit-identical results (each output
  // element is summed by one lane in the same hop order), so the choice only decides speed.
  const int key = (int)sizeof(T) * 4096 + ld;
  auto it = f.tuned.find(key);
  if (it == f.tuned.end()) {
    const int e = fact_enqueue_single<T>(f, X, ld, st);
    if (e != COSMO_OK) return e;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int best = 0;
    float best_ms = 1e30f;
    for (int lpe = 8; lpe >= 1; lpe >>= 1) {
      float ms = 1e30f;
      for (int rep = 0; rep < 2; ++rep) {
        HIPCHK(hipEventRecord(e0, st));
        const int r = launch_fact<T>(f, lpe, X, Yp, Yq, Y, m, ld, st);
        if (r < 0) return r;
        if (r == 0) break;
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        ms = std::min(ms, t);
      }
      if (ms < best_ms) {
        best_ms = ms;
        best = lpe;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!best) return fail(COSMO_EUNSUP, "factorised apply: proton sector too large for one workgroup");
    f.tuned.emplace(key, best);
    // the timed launches all wrote Y; one more with the chosen width leaves the same bits
    const int r = launch_fact<T>(f, best, X, Yp, Yq, Y, m, ld, st);
    return r < 0 ? r : COSMO_OK;
  }
  const int lpe = it->second;
  // Repeated calls on the same buffers replay one captured graph: one host launch instead
  // of the gather, the ~2 GEMMs per class and k_fact.  The second call with a key captures;
  // a capture that fails, or records anything but kernels and event nodes, leaves that key
  // on the eager path.
  const FactOp::GraphKey gk{(int)sizeof(T), ld, m, (const void *)X, (const void *)Y};
  auto g = f.graphs.find(gk);
  if (g != f.graphs.end()) {
    HIPCHK(hipGraphLaunch(g->second, st));
    return COSMO_OK;
  }
  if (f.seen.size() >= 4 * FactOp::FACT_GRAPHS && !f.seen.count(gk)) f.seen.clear();  // bounded bookkeeping
  int &seen = f.seen[gk];
  if (seen >= 1 && f.graphs.size() < FactOp::FACT_GRAPHS) {
    hipGraphExec_t exec = nullptr;
    if (fact_capture<T>(f, lpe, X, Yp, Yq, Y, m, ld, st, &exec)) {
      f.graphs.emplace(gk, exec);
      HIPCHK(hipGraphLaunch(exec, st));
      return COSMO_OK;
    }
    seen = -1;
  }
  if (seen >= 0) ++seen;
  const int e = fact_enqueue_single<T>(f, X, ld, st);
  if (e != COSMO_OK) return e;
  const int r = launch_fact<T>(f, lpe, X, Yp, Yq, Y, m, ld, st);
  if (r < 0) return r;
  if (r == 0) return fail(COSMO_EUNSUP, "factorised apply: block tables exceed the LDS");
  return COSMO_OK;
}

// ---- the product without allocation -------------------------------------------------
// For callers that cannot allocate (code under HIP graph capture, tight serving loops) the
// class products' outputs (Yp, then Yq, each n * ld elements rounded up to 256 bytes) live
// in a file-scope device array of FACT_SCRATCH_BYTES, one per GPU, held until the process
// ends.  Every call on a GPU shares it, so calls that overlap on one GPU race.  Nothing in
// it needs clearing: the class products write every element the step kernel reads.
constexpr size_t FACT_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[FACT_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64] = {};  // per device: the array's device address

// the array's address on device dev (current): looked up once per device
static int scratch_address(int dev, void **out) {
  if (dev < 0 || dev >= 64) return fail(COSMO_EINVAL, "scratch: device index");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(COSMO_EHIP, std::string("scratch: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return COSMO_OK;
}

template <typename T>
static int fact_apply_noalloc_t(Op &op, const T *X, T *Y, int m, int ld, hipStream_t st) {
  FactOp &f = *static_cast<FactOp *>(op.fact);
  constexpr int VPL = Piece16<T>::N;
  if (ld % VPL || (ld & (ld - 1)) || ((uintptr_t)X | (uintptr_t)Y) % 16)
    return fail(COSMO_EUNSUP, "factorised apply: leading dimension must be a power of two of >= 16 bytes");
  if (!f.use_mm)
    return fail(COSMO_EUNSUP, "cosmo_op_apply_device_noalloc: this operator's class products need the allocating path");
  const size_t half = ((size_t)f.n * ld * sizeof(T) + 255) & ~(size_t)255;
  if (2 * half > FACT_SCRATCH_BYTES)
    return fail(COSMO_EINVAL, "cosmo_op_apply_device_noalloc: needs " + std::to_string(2 * half) +
                                  " bytes of scratch, above the maximum of " + std::to_string(FACT_SCRATCH_BYTES));
  void *base = nullptr;
  int rc = scratch_address(op.device, &base);
  if (rc) return rc;
  T *Yp = (T *)base, *Yq = (T *)((char *)base + half);
  rc = launch_class_mm<T>(f, X, ld, Yp, Yq, st);
  if (rc) return rc;
  // the slice width tuned by the allocating path for this (precision, ld), else the widest
  // that fits: every width gives the same bits
  const auto it = f.tuned.find((int)sizeof(T) * 4096 + ld);
  if (it != f.tuned.end()) {
    const int r = launch_fact<T>(f, it->second, X, Yp, Yq, Y, m, ld, st);
    if (r != 0) return r < 0 ? r : COSMO_OK;
  }
  for (int lpe = 8; lpe >= 1; lpe >>= 1) {
    const int r = launch_fact<T>(f, lpe, X, Yp, Yq, Y, m, ld, st);
    if (r != 0) return r < 0 ? r : COSMO_OK;
  }
  return fail(COSMO_EUNSUP, "factorised apply: proton sector too large for one workgroup");
}

int fact_apply_noalloc(Op &op, const void *X, void *Y, int m, int ld, int prec, hipStream_t st) {
  if (prec == COSMO_PREC_FP32) return fact_apply_noalloc_t<float>(op, (const float *)X, (float *)Y, m, ld, st);
  if (prec == COSMO_PREC_FP64) return fact_apply_noalloc_t<double>(op, (const double *)X, (double *)Y, m, ld, st);
  return fail(COSMO_EUNSUP, "cosmo_op_apply_device_noalloc: precision must be COSMO_PREC_FP32 or COSMO_PREC_FP64");
}

int fact_apply(Op &op, const void *X, void *Y, int m, int ld, int prec, hipStream_t st) {
  if (prec == COSMO_PREC_FP32) return fact_apply_t<float>(op, (const float *)X, (float *)Y, m, ld, st);
  return fact_apply_t<double>(op, (const double *)X, (double *)Y, m, ld, st);
}

}  // namespace cosmo
