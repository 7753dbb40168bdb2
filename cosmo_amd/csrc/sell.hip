// sell.hip -- stored H.X on row blocks with the X tiles staged in LDS.
//
// The reference multiplies its CSB blocks one after another, each block with the slice of
// x it needs (SparseMatrixCSB::TimesMatrix, include/tav/csbmatrix.cxx:623-663, blocks of
// CSBBlock :128-336).  Here the blocks are the basis's own neutron-configuration blocks
// (consecutive ones merged up to SELL_MAXROWS rows): the rows of one neutron
// configuration, and as columns the rows of another.  Block (b, t) is non-empty only when
// the neutron configurations of b and t are one or two moves apart, so a row block meets
// tens to a hundred or so column tiles, and each entry's X row is one of the tile's rows.
//
// Entries go to one of three parts, each multiplied in ascending column order within a row
// (deterministic: the same bits on every call):
//  * the proton-spectator part (entries between two neutron configurations with the same
//    proton configuration) on proton-major tiles, k_pm_apply, which writes every row of Y;
//  * value panels: per row block and each of its column tiles in ascending order, per slice
//    of 64 rows (one wavefront) an ELL panel: slot j of the slice's lane l is entry
//    sl_base[q] + 64 j + l, holding the column inside the tile (u16) and the value (FP32
//    and FP64 copies); a slice's panel is as long as its longest row in t, shorter rows are
//    padded with zero values pointing at a zero row kept after each staged tile;
//  * coded panels: the same layout for the (row block, tile) pairs whose entries take at
//    most SELL_DICT - 1 distinct non-zero values, with one 32-bit word per slot: the column
//    in the low half, above it the index into the pair's dictionary (entry 0 is 0.0).  Off a row block's own tile
//    the proton-neutron entries of a pair all move the same neutron, so their values are
//    the interaction's for that move, with signs: a hundred or so values.
// Lanes take the rows of a block in the order of their numbers of entries (lrow), so the
// rows of a 64-row slice have similar panel lengths and little padding.
//
// Panel product: one workgroup per row block, one lane per row holding its whole Y row in
// registers.  Per column tile: the tile's X rows (contiguous in X) and, when coded, the
// pair's dictionary go from memory straight into one of two LDS buffers
// (global_load_lds, no registers involved) one tile ahead, and the first SELL_PF slots of
// each lane's next panel are loaded into registers, so the loads of the next tile fly
// while the current one is multiplied out of LDS; one wait and one barrier per tile.  The
// coded kernel multiplies the block's value pairs first, one at a time, while its first
// coded tile lands.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "op.h"

namespace cosmo {

#define SELL_MAXROWS 576  // rows of a row block: nine wavefronts, one lane per row
#define SELL_PF 4         // entry slots per lane loaded one tile ahead (value panels)
#define SELL_PFC 4        // the same for coded panels, whose slots take one register each
#define SELL_MIN_N 262144  // automatic layout: smaller operators keep the row kernel
#define SELL_AFTER 256     // automatic layout: built once the row kernel has made this many products
#define SELL_DICT 256     // entries of a pair's value dictionary (entry 0 is 0.0)
#define SELL_HASH 1024    // LDS hash slots that collect a pair's distinct values
#define PM_PF 4           // slots per group in k_pm_apply (the slot arrays carry 4 groups of slack)
static_assert(SELL_PFC >= SELL_PF, "panel slot arrays carry SELL_PFC * 64 slots of slack for both kinds");

// one set of panels over the row blocks: per row block a subset of the column tiles it
// meets (ascending), per (row block, tile, 64-row slice) an ELL panel
struct Panels {
  int maxtiles = 0;
  int64_t K = 0, Q = 0, E = 0, Ealloc = 0;
  int32_t *nbr_start = nullptr;  // [nblk+1] into nbr_tile
  int32_t *nbr_tile = nullptr;   // [K] column tiles of each row block, ascending
  int64_t *qstart = nullptr;     // [nblk] first slice of the block: slice (k, w) is qstart + k * nslices + w
  int64_t *sl_base = nullptr;    // [Q+1] first entry slot of each slice (a multiple of 64)
  uint16_t *cols = nullptr;      // [Ealloc] column inside the tile
  float *v32 = nullptr;          // value panels: [Ealloc]
  double *v64 = nullptr;
  bool coded = false;            // coded panels: per slot column | dictionary index << 16
  uint32_t *pk32 = nullptr;      // [Ealloc] slots with the FP32 dictionaries' indices
  uint32_t *pk64 = nullptr;      // [Ealloc] slots with the FP64 dictionaries' indices
  float *dict32 = nullptr;       // [dict_total]
  double *dict64 = nullptr;
  int32_t *dict_off = nullptr;   // [K] first entry of each pair's dictionary (a multiple of 4)
  int32_t *dict_n4 = nullptr;    // [K] its length in groups of 4 entries
  int64_t dict_total = 0;
  uint32_t *pair_c0 = nullptr;   // [K] first X row of each pair's column tile
  int32_t *pair_rows = nullptr;  // [K] its rows
  int4 *items = nullptr;         // [2 nblk] the item of each row block
};

struct Sell {
  int nblk = 0, maxrows = 0, nwaves = 0;
  int64_t *boff = nullptr;       // [nblk+1] first row of each block
  uint16_t *lrow = nullptr;      // [n] row of block b taken by lane l: boff[b] + lrow[boff[b] + l]
  Panels P[2];                   // value panels, coded panels
  size_t bytes = 0;
  std::map<int, int> chunk;      // (element size, pitch) -> 4 * XCD chunk + tile buffers, chosen by timing
  // Proton-spectator part: the entries between two different neutron configurations with
  // the same proton configuration (neutron two-move terms, and one-move terms whose proton
  // configuration is unchanged), multiplied on proton-major tiles by k_pm_apply and left
  // out of the panels above.  A proton-major tile is one proton configuration with every
  // neutron configuration of its class (the neutron configurations whose sector holds it):
  // these entries never leave it.  Units: (class, proton rank, rows of the class from i0).
  bool split = false;
  int pm_units = 0, pm_maxC = 0, pm_nwaves = 0;
  int64_t pm_Q = 0, pm_E = 0, pm_Ealloc = 0;
  uint64_t *rid = nullptr;       // [n] (neutron block << 32) | proton configuration of each row
  int32_t *cls_local = nullptr;  // [neutron blocks] index of the block within its class
  uint32_t *pm_rows = nullptr;   // class members' first rows, class after class
  int4 *pm_items = nullptr;      // [2 units]: (first member, members C, proton rank, i0), (rows, first slice)
  int64_t *pm_base = nullptr;    // [pm_Q + 1] first slot of each slice (a multiple of 64)
  uint16_t *pm_cols = nullptr;   // [pm_Ealloc] 1 + member index of the column (0: the zero row)
  float *pm_v32 = nullptr;
  double *pm_v64 = nullptr;
};

// an entry (r, c) of the proton-spectator part: the same proton configuration, another
// neutron configuration
__device__ __forceinline__ bool pm_entry(const uint64_t *__restrict__ rid, uint64_t ra, uint32_t c) {
  if (!rid) return false;
  const uint64_t rb = rid[c];
  return (uint32_t)ra == (uint32_t)rb && (ra >> 32) != (rb >> 32);
}

#define SELL_ORDER_MAX 64  // entries of a lane in one panel that panel_order places

// The slot of each of a lane's entries in one panel.  In the product, lanes l, l+4, l+8 and
// l+12 read the same piece position of their X rows at once (the pieces rotate with the
// lane), so their reads fall in different LDS bank groups when their rows inside the tile
// differ modulo 4.  At slot j a lane therefore takes, while it has one, an entry whose row
// inside the tile is (j + l / 4) modulo 4, else its first entry left in column order.
// slot[k] receives the slot of the lane's k-th entry (proton-spectator entries skipped);
// false when the lane has more than SELL_ORDER_MAX entries there (column order then).
__device__ inline bool panel_order(const uint32_t *__restrict__ cols, const uint64_t *__restrict__ rid, uint64_t ra,
                                   int64_t e0, int64_t e1, uint32_t cs, int lane, uint8_t *slot) {
  uint8_t cls[SELL_ORDER_MAX];
  int cnt = 0;
  for (int64_t e = e0; e < e1; ++e) {
    if (pm_entry(rid, ra, cols[e])) continue;
    if (cnt == SELL_ORDER_MAX) return false;
    cls[cnt++] = (uint8_t)((cols[e] - cs) & 3u);
  }
  uint64_t used = 0;
  for (int j = 0; j < cnt; ++j) {
    const int want = (j + (lane >> 2)) & 3;
    int pick = -1;
    for (int k = 0; k < cnt && pick < 0; ++k)
      if (!((used >> k) & 1ull) && cls[k] == want) pick = k;
    for (int k = 0; k < cnt && pick < 0; ++k)
      if (!((used >> k) & 1ull)) pick = k;
    used |= 1ull << pick;
    slot[pick] = (uint8_t)j;
  }
  return true;
}

// A row block's item is two int4: (first pair, tiles, rows, 0) and (first row, first
// slice, 0, 0); pairs are the (row block, column tile) pairs in block-major order.
struct SellView {
  int maxtiles, nthr;
  const int4 *items;
  const uint32_t *pair_c0;       // per pair: the tile's first X row
  const int32_t *pair_rows;      // and its rows
  const int64_t *sl_base;
  const uint16_t *cols;
  const int32_t *dict_off;       // coded panels: per pair its dictionary
  const int32_t *dict_n4;
  const uint32_t *packed;        // coded panels: column | dictionary index << 16
  const uint16_t *lrow;
};

// the row of the row block starting at r0 that lane (thread) l takes
__device__ __forceinline__ int64_t lane_row(const uint16_t *__restrict__ lrow, int64_t r0, int l) {
  return r0 + (lrow ? (int64_t)lrow[r0 + l] : (int64_t)l);
}

// ---- build -------------------------------------------------------------------------
// column tiles met by row block b: a bitmap over tiles in LDS from the block's entries
// (contiguous in the CSR), then either its population (pass 0) or the ascending list
__global__ __launch_bounds__(256) void k_sell_tiles(int pass, const int64_t *__restrict__ boff,
                                                    const int64_t *__restrict__ rowptr,
                                                    const uint32_t *__restrict__ cols,
                                                    const int32_t *__restrict__ blk_of, int nblk,
                                                    int32_t *__restrict__ count,
                                                    const int32_t *__restrict__ nbr_start,
                                                    int32_t *__restrict__ nbr_tile,
                                                    const uint64_t *__restrict__ rid) {
  extern __shared__ uint32_t bm[];
  const int b = blockIdx.x, nw = (nblk + 31) >> 5;
  for (int w = threadIdx.x; w < nw; w += blockDim.x) bm[w] = 0u;
  __syncthreads();
  if (!rid) {
    const int64_t e0 = rowptr[boff[b]], e1 = rowptr[boff[b + 1]];
    for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
      const int t = blk_of[cols[e]];
      atomicOr(&bm[t >> 5], 1u << (t & 31));
    }
  } else {  // one thread per row, the proton-spectator entries left out
    for (int64_t r = boff[b] + threadIdx.x; r < boff[b + 1]; r += blockDim.x) {
      const uint64_t ra = rid[r];
      for (int64_t e = rowptr[r]; e < rowptr[r + 1]; ++e) {
        const uint32_t c = cols[e];
        if (pm_entry(rid, ra, c)) continue;
        const int t = blk_of[c];
        atomicOr(&bm[t >> 5], 1u << (t & 31));
      }
    }
  }
  __syncthreads();
  if (pass == 0) {
    __shared__ int tot;
    if (threadIdx.x == 0) tot = 0;
    __syncthreads();
    int c = 0;
    for (int w = threadIdx.x; w < nw; w += blockDim.x) c += __popc(bm[w]);
    atomicAdd(&tot, c);
    __syncthreads();
    if (threadIdx.x == 0) count[b] = tot;
  } else if (threadIdx.x == 0) {
    int o = nbr_start[b];
    for (int w = 0; w < nw; ++w)
      for (uint32_t m = bm[w]; m; m &= m - 1) nbr_tile[o++] = w * 32 + __ffs(m) - 1;
  }
}

// pass 0: per slice and tile, the longest row (entries of the row in the tile), written
// as slot counts to len[q]; pass 1: the entries into their slots.  One lane per row, every
// lane walks the block's tiles in order with a cursor into its (column-sorted) row.
__global__ __launch_bounds__(SELL_MAXROWS) void k_sell_slots(
    int pass, const int64_t *__restrict__ boff, const int64_t *__restrict__ rowptr,
    const uint32_t *__restrict__ cols, const double *__restrict__ v64, const float *__restrict__ v32,
    const int32_t *__restrict__ nbr_start, const int32_t *__restrict__ nbr_tile,
    const int64_t *__restrict__ qstart, int64_t *__restrict__ len, const int64_t *__restrict__ sl_base,
    uint16_t *__restrict__ scol, double *__restrict__ s64, float *__restrict__ s32,
    const uint64_t *__restrict__ rid, const uint16_t *__restrict__ lrow) {
  const int b = blockIdx.x;
  const int64_t r0 = boff[b];
  const int S = (int)(boff[b + 1] - r0), ns = (S + 63) >> 6;
  const int rho = threadIdx.x, lane = rho & 63, w = rho >> 6;
  if (w >= ns) return;  // no barrier below: whole idle wavefronts leave
  const bool on = rho < S;
  const int64_t rr = on ? lane_row(lrow, r0, rho) : 0;
  int64_t e = on ? rowptr[rr] : 0;
  const int64_t ee = on ? rowptr[rr + 1] : 0;
  const uint64_t ra = (on && rid) ? rid[rr] : 0;
  const int k0 = nbr_start[b], k1 = nbr_start[b + 1];
  for (int k = k0; k < k1; ++k) {
    const int t = nbr_tile[k];
    const uint32_t cs = (uint32_t)boff[t], ce = (uint32_t)boff[t + 1];
    const int64_t q = qstart[b] + (int64_t)(k - k0) * ns + w;
    int c = 0;
    while (e < ee && cols[e] < cs) ++e;  // entries in tiles of the other panel set
    if (pass == 0) {
      for (; e < ee && cols[e] < ce; ++e)
        if (!pm_entry(rid, ra, cols[e])) ++c;
      for (int off = 32; off > 0; off >>= 1) c = max(c, __shfl_xor(c, off));
      if (lane == 0) len[q] = c;
    } else {
      const int64_t base = sl_base[q] + lane;
      int64_t e_end = e;
      while (e_end < ee && cols[e_end] < ce) ++e_end;
      uint8_t slot[SELL_ORDER_MAX];
      const bool ordered = panel_order(cols, rid, ra, e, e_end, cs, lane, slot);
      for (; e < e_end; ++e) {
        if (pm_entry(rid, ra, cols[e])) continue;
        const int64_t p = base + (int64_t)(ordered ? slot[c] : c) * 64;
        if (scol) scol[p] = (uint16_t)(cols[e] - cs);
        if (s64) s64[p] = v64[e];
        if (s32) s32[p] = v32[e];
        ++c;
      }
    }
  }
}

// the proton-spectator part's panels: one workgroup per unit, one lane per row of the
// unit (member i0 + lane of the class at the unit's proton rank); pass 0 writes each
// slice's longest row to len[q], pass 1 the entries (in column order) into their slots
__global__ __launch_bounds__(SELL_MAXROWS) void k_pm_slots(
    int pass, const int4 *__restrict__ items, const uint32_t *__restrict__ pm_rows, const uint64_t *__restrict__ rid,
    const int32_t *__restrict__ cls_local, const int64_t *__restrict__ rowptr, const uint32_t *__restrict__ cols,
    const double *__restrict__ v64, const float *__restrict__ v32, int64_t *__restrict__ len,
    const int64_t *__restrict__ base, uint16_t *__restrict__ scol, double *__restrict__ s64,
    float *__restrict__ s32) {
  const int4 u0 = items[2 * blockIdx.x], u1 = items[2 * blockIdx.x + 1];
  const int S = u1.x, ns = (S + 63) >> 6;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (w >= ns) return;
  const bool on = tid < S;
  const int64_t r = on ? (int64_t)pm_rows[u0.x + u0.w + tid] + u0.z : 0;
  int64_t e = on ? rowptr[r] : 0;
  const int64_t ee = on ? rowptr[r + 1] : 0;
  const uint64_t ra = on ? rid[r] : 0;
  const int64_t q = (int64_t)u1.y + w;
  int c = 0;
  if (pass == 0) {
    for (; e < ee; ++e)
      if (pm_entry(rid, ra, cols[e])) ++c;
    for (int off = 32; off > 0; off >>= 1) c = max(c, __shfl_xor(c, off));
    if (lane == 0) len[q] = c;
  } else {
    const int64_t b0 = base[q] + lane;
    for (; e < ee; ++e) {
      const uint32_t col = cols[e];
      if (!pm_entry(rid, ra, col)) continue;
      const int64_t p = b0 + (int64_t)c * 64;
      if (scol) scol[p] = (uint16_t)(1 + cls_local[rid[col] >> 32]);
      if (s64) s64[p] = v64[e];
      if (s32) s32[p] = v32[e];
      ++c;
    }
  }
}

// layout statistics: the distinct tile rows the (row block, column tile) pairs reference,
// summed into *out (one workgroup per row block, a bitmap of the tile's rows in LDS)
__global__ __launch_bounds__(SELL_MAXROWS) void k_sell_distinct(const int64_t *__restrict__ boff,
                                                               const int64_t *__restrict__ rowptr,
                                                               const uint32_t *__restrict__ cols,
                                                               const int32_t *__restrict__ nbr_start,
                                                               const int32_t *__restrict__ nbr_tile,
                                                               unsigned long long *__restrict__ out,
                                                               const uint64_t *__restrict__ rid) {
  constexpr int NW = SELL_MAXROWS / 32;
  __shared__ uint32_t bm[NW];
  __shared__ int tot;
  const int b = blockIdx.x, rho = threadIdx.x;
  const int64_t r0 = boff[b];
  const bool on = rho < (int)(boff[b + 1] - r0);
  int64_t e = on ? rowptr[r0 + rho] : 0;  // lane order does not matter here
  const int64_t ee = on ? rowptr[r0 + rho + 1] : 0;
  const uint64_t ra = (on && rid) ? rid[r0 + rho] : 0;
  if (rho == 0) tot = 0;
  for (int k = nbr_start[b]; k < nbr_start[b + 1]; ++k) {
    if (rho < NW) bm[rho] = 0u;
    __syncthreads();
    const int t = nbr_tile[k];
    const uint32_t cs = (uint32_t)boff[t], ce = (uint32_t)boff[t + 1];
    while (e < ee && cols[e] < cs) ++e;
    for (; e < ee && cols[e] < ce; ++e)
      if (!pm_entry(rid, ra, cols[e])) atomicOr(&bm[(cols[e] - cs) >> 5], 1u << ((cols[e] - cs) & 31));
    __syncthreads();
    if (rho < NW) atomicAdd(&tot, __popc(bm[rho]));
    __syncthreads();
  }
  if (rho == 0) atomicAdd(out, (unsigned long long)tot);
}

// per row: its stored entries outside the proton-spectator part (the lane order's key)
__global__ void k_row_weight(int64_t n, const int64_t *__restrict__ rowptr, const uint32_t *__restrict__ cols,
                             const uint64_t *__restrict__ rid, int32_t *__restrict__ wt) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int c = 0;
  const uint64_t ra = rid ? rid[r] : 0;
  for (int64_t e = rowptr[r]; e < rowptr[r + 1]; ++e)
    if (!pm_entry(rid, ra, cols[e])) ++c;
  wt[r] = c;
}

// bit patterns of values as hash keys (never all ones: that is a NaN)
__device__ __forceinline__ unsigned int val_bits(float v) { return __float_as_uint(v); }
__device__ __forceinline__ unsigned long long val_bits(double v) {
  return (unsigned long long)__double_as_longlong(v);
}
__device__ __forceinline__ float bits_val(unsigned int k, float) { return __uint_as_float(k); }
__device__ __forceinline__ double bits_val(unsigned long long k, double) { return __longlong_as_double((long long)k); }
__device__ __forceinline__ uint32_t key_hash(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  return (uint32_t)k;
}

// The distinct non-zero values of each (row block, tile) pair of a panel set, collected in
// an LDS hash table (one workgroup per row block, one lane per row, pairs in turn).
// pass 0: count[k] = their number, or SELL_DICT when there are more than SELL_DICT - 1;
// pass 1 (pairs of at most SELL_DICT - 1 values): the pair's dictionary at dict_off[k]
// (entry 0 is 0.0, then the values in hash-table order) and every slot's word (column
// inside the tile | index << 16); padding slots keep the zero row and index 0.
template <typename T, typename KT>
__global__ __launch_bounds__(SELL_MAXROWS) void k_sell_dict(
    int pass, const int64_t *__restrict__ boff, const int64_t *__restrict__ rowptr,
    const uint32_t *__restrict__ cols, const T *__restrict__ vals, const uint64_t *__restrict__ rid,
    const uint16_t *__restrict__ lrow, const int32_t *__restrict__ nbr_start, const int32_t *__restrict__ nbr_tile,
    int32_t *__restrict__ count, const int64_t *__restrict__ qstart, const int64_t *__restrict__ sl_base,
    const int32_t *__restrict__ dict_off, T *__restrict__ dict, uint32_t *__restrict__ packed) {
  __shared__ KT tab[SELL_HASH];
  __shared__ uint8_t tcode[SELL_HASH];
  __shared__ int nkeys, over;
  const KT EMPTY = ~(KT)0;
  const int b = blockIdx.x;
  const int64_t r0 = boff[b];
  const int S = (int)(boff[b + 1] - r0), ns = (S + 63) >> 6;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const bool on = tid < S;
  const int64_t rr = on ? lane_row(lrow, r0, tid) : 0;
  int64_t e = on ? rowptr[rr] : 0;
  const int64_t ee = on ? rowptr[rr + 1] : 0;
  const uint64_t ra = (on && rid) ? rid[rr] : 0;
  const int k0 = nbr_start[b], k1 = nbr_start[b + 1];
  for (int k = k0; k < k1; ++k) {
    for (int i = tid; i < SELL_HASH; i += blockDim.x) tab[i] = EMPTY;
    if (tid == 0) nkeys = 0, over = 0;
    __syncthreads();
    const int t = nbr_tile[k];
    const uint32_t cs = (uint32_t)boff[t], ce = (uint32_t)boff[t + 1];
    while (e < ee && cols[e] < cs) ++e;  // entries in tiles of the other panel set
    const int64_t e_first = e;
    for (; e < ee && cols[e] < ce; ++e) {
      if (pm_entry(rid, ra, cols[e])) continue;
      const KT key = val_bits(vals[e]);
      if (key == 0) continue;  // +0.0 is entry 0
      uint32_t h = key_hash(key) & (SELL_HASH - 1);
      int probe = 0;
      for (; probe < SELL_HASH; ++probe) {  // bounded: a full table ends the search
        const KT old = atomicCAS(&tab[h], EMPTY, key);
        if (old == EMPTY) {
          atomicAdd(&nkeys, 1);
          break;
        }
        if (old == key) break;
        h = (h + 1) & (SELL_HASH - 1);
      }
      if (probe == SELL_HASH) over = 1;
    }
    __syncthreads();
    if (pass == 0) {
      if (tid == 0) count[k] = (over || nkeys >= SELL_DICT) ? SELL_DICT : nkeys;
      __syncthreads();  // the table is cleared for the next pair only after this read
      continue;
    }
    if (w == 0) {  // codes 1, 2, ... in table order
      const int d0 = dict_off[k];
      int run = 0;
      for (int c0 = 0; c0 < SELL_HASH; c0 += 64) {
        const KT key = tab[c0 + lane];
        const bool occ = key != EMPTY;
        const uint64_t mask = __ballot(occ);
        const int pre = run + __popcll(mask & ((1ull << lane) - 1ull));
        if (occ) {
          tcode[c0 + lane] = (uint8_t)(1 + pre);
          dict[d0 + 1 + pre] = bits_val(key, (T)0);
        }
        run += __popcll(mask);
      }
      if (lane == 0) dict[d0] = (T)0;
    }
    __syncthreads();
    if (on) {
      const int64_t base = sl_base[qstart[b] + (int64_t)(k - k0) * ns + w] + lane;
      uint8_t slot[SELL_ORDER_MAX];
      const bool ordered = panel_order(cols, rid, ra, e_first, e, cs, lane, slot);
      int c = 0;
      for (int64_t x = e_first; x < e; ++x) {
        if (pm_entry(rid, ra, cols[x])) continue;
        const KT key = val_bits(vals[x]);
        uint32_t cd = 0;
        if (key != 0) {  // present: every key of the pair was inserted above
          uint32_t h = key_hash(key) & (SELL_HASH - 1);
          while (tab[h] != key) h = (h + 1) & (SELL_HASH - 1);
          cd = tcode[h];
        }
        packed[base + (int64_t)(ordered ? slot[c] : c) * 64] = (cols[x] - cs) | (cd << 16);
        ++c;
      }
    }
    __syncthreads();
  }
}

// ---- product -----------------------------------------------------------------------
// 16-byte pieces as native vectors (an array of HIP's union vector types is kept in
// scratch memory by the compiler; arrays of these stay in registers)
template <typename T>
struct V16t;
template <>
struct V16t<float> {
  typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct V16t<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};

// XCD-aware order: the eight XCDs take chunks of C consecutive row blocks in turn, so each
// XCD works through neighbouring blocks (which share column tiles in its L2) and the XCDs
// stay close together in the basis (sharing X in the Infinity Cache)
__device__ __forceinline__ int sell_block(int chunk) {
  const int G = gridDim.x, i = blockIdx.x;
  const int G8 = (G / (8 * chunk)) * (8 * chunk);
  if (chunk <= 1 || i >= G8) return i;
  const int x = i % 8, j = i / 8;
  return ((j / chunk) * 8 + x) * chunk + j % chunk;
}

// LDS bytes of a panel product's pair table and slice offsets (rounded to 16)
__host__ __device__ inline size_t sell_meta_bytes(int maxtiles, int nwaves) {
  return (16 * (size_t)maxtiles + 2 * ((size_t)maxtiles * nwaves + 1) + 15) & ~(size_t)15;
}

// A: the panels multiplied through the pipeline, with vals their slot values (CODED
// false) or their dictionaries (CODED true); B and bvals (CODED only): the value panels,
// multiplied first, pair by pair
template <typename T, int LD, bool CODED, int NB>
__device__ __forceinline__ void sell_apply_body(const SellView &A, const T *__restrict__ vals, const SellView &B,
                                                const T *__restrict__ bvals, const T *__restrict__ X,
                                                T *__restrict__ Y, int m, int chunk, int accum) {
  typedef typename V16t<T>::type V;
  typedef typename std::conditional<CODED, uint32_t, T>::type W;  // what a slot holds
  constexpr int PF = CODED ? SELL_PFC : SELL_PF;                   // slots per prefetch group
  constexpr int VPR = (int)(LD * sizeof(T) / 16);  // 16-byte pieces per X row
  constexpr int VW = (int)(16 / sizeof(T));        // elements per piece
  extern __shared__ __align__(16) unsigned char lds_raw[];
  // the row block's record: the only load before the tile table
  const int b = sell_block(chunk);
  const int4 d0 = A.items[2 * b], d1 = A.items[2 * b + 1];
  const int k0 = d0.x, K = d0.y, S = d0.z;
  const int64_t r0 = d1.x, qb = d1.y;
  const int ns = (S + 63) >> 6;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nthr = blockDim.x, wc = min(w, ns - 1);
  const bool wave_on = w < ns;
  // LDS: NB tile buffers of nthr + 1 rows (row nthr stays zero: the padding slots' row;
  // the direct loads fill rows [0, nthr)), then the block's tile table (first X row,
  // 16-byte pieces, dictionary offset and length) and its slice offsets (64-slot units from
  // the block's first slot), then (coded) NB dictionary buffers
  const int zrow = A.nthr;
  V *buf0 = reinterpret_cast<V *>(lds_raw), *buf1 = buf0 + (size_t)(NB - 1) * (zrow + 1) * VPR;
  uint32_t *mt_c0 = reinterpret_cast<uint32_t *>(buf0 + (size_t)NB * (zrow + 1) * VPR);
  int32_t *mt_np = reinterpret_cast<int32_t *>(mt_c0 + A.maxtiles);
  int32_t *mt_doff = mt_np + A.maxtiles;
  int32_t *mt_dn = mt_doff + A.maxtiles;
  uint16_t *mt_off = reinterpret_cast<uint16_t *>(mt_dn + A.maxtiles);
  T *dict0 = reinterpret_cast<T *>(reinterpret_cast<unsigned char *>(mt_c0) +
                                   sell_meta_bytes(A.maxtiles, A.nthr >> 6));
  T *dict1 = dict0 + (NB - 1) * SELL_DICT;
  const int64_t slot0 = A.sl_base[qb];
  for (int i = tid; i < K; i += nthr) {
    mt_c0[i] = A.pair_c0[k0 + i];
    mt_np[i] = A.pair_rows[k0 + i] * VPR;
    if (CODED) {
      mt_doff[i] = A.dict_off[k0 + i];
      mt_dn[i] = A.dict_n4[k0 + i];
    }
  }
  for (int i = tid; i <= K * ns; i += nthr) mt_off[i] = (uint16_t)((A.sl_base[qb + i] - slot0) >> 6);
  if (tid < NB * VPR) buf0[(size_t)(tid / VPR) * (zrow + 1) * VPR + (size_t)zrow * VPR + tid % VPR] = V{};
  // instruction p of an entry reads piece (p + lane) % VPR of its X row into accumulator
  // block p: lanes on random rows then hit all LDS banks (a plain piece-p read from rows
  // 64 bytes apart reaches a quarter of them); the rotation is undone when Y is written
  int rot[VPR];
#pragma unroll
  for (int p = 0; p < VPR; ++p) rot[p] = (p + lane) & (VPR - 1);
  const int64_t myrow = lane_row(A.lrow, r0, min(tid, S - 1));
  T acc[LD];
#pragma unroll
  for (int i = 0; i < LD; ++i) acc[i] = (T)0;
  if (accum && tid < S) {  // Y already holds the proton-spectator part of the row
    const V *yr = reinterpret_cast<const V *>(Y + myrow * LD);
#pragma unroll
    for (int p = 0; p < VPR; ++p) {
      const V u = yr[rot[p]];
#pragma unroll
      for (int i = 0; i < VW; ++i) acc[p * VW + i] = u[i];
    }
  }
  // tile loads straight into LDS (global_load_lds_dwordx4: lane i of a wavefront writes
  // 16 bytes at the wavefront's LDS base + 16 i): piece idx = tid + p * nthr of the tile's
  // contiguous X rows goes to LDS piece idx; pieces past the tile load the tile's last
  // piece again (one cached line per wavefront) into rows no slot refers to
  auto dma = [&](uint32_t c0, int np, V *dst) {
    const V *src = reinterpret_cast<const V *>(X + (int64_t)c0 * LD);
#pragma unroll
    for (int p = 0; p < VPR; ++p)
      __builtin_amdgcn_global_load_lds((const void *)(src + min(tid + p * nthr, np - 1)),
                                       (__attribute__((address_space(3))) void *)(dst + (w * 64 + p * nthr)), 16, 0,
                                       0);
  };
  // a coded pair's dictionary (n4 groups of 4 entries from entry d0) into a dictionary
  // buffer, by wavefront 0: pieces past its end load its last piece again
  auto dma_dict = [&](int d0, int n4, T *dd) {
    if (CODED && w == 0) {
      const V *src = reinterpret_cast<const V *>(vals + d0);
      const int np = n4 * 4 / VW;
#pragma unroll
      for (int i = 0; i < SELL_DICT / VW / 64; ++i)
        __builtin_amdgcn_global_load_lds((const void *)(src + min(i * 64 + lane, np - 1)),
                                         (__attribute__((address_space(3))) void *)(reinterpret_cast<V *>(dd) + i * 64),
                                         16, 0, 0);
    }
  };
  auto dma_tile = [&](int kk, V *dst, T *dd) {
    dma(mt_c0[kk], mt_np[kk], dst);
    if (CODED) dma_dict(mt_doff[kk], mt_dn[kk], dd);
  };
  // the panel of this wavefront's slice for step kk (the same for the whole wavefront:
  // held in scalar registers); slots [j0, j0 + PF) of a lane are read unconditionally
  // (the slot arrays carry SELL_PFC * 64 slots of slack), so the compiler never makes a
  // load wait for the next tile's loads at a branch
  auto slice_of = [&](int kk, int64_t &base, int &L) {
    const int o = __builtin_amdgcn_readfirstlane((int)mt_off[kk * ns + wc]);
    base = slot0 + (int64_t)o * 64 + lane;
    L = __builtin_amdgcn_readfirstlane((int)mt_off[kk * ns + wc + 1]) - o;
  };
  auto load_slots = [&](int64_t base, int L, int j0, uint32_t (&c)[PF], W (&h)[PF]) {
    (void)L;
#pragma unroll
    for (int j = 0; j < PF; ++j) {
      const int64_t p = base + (int64_t)(j0 + j) * 64;
      if constexpr (CODED) {
        c[j] = A.packed[p];
        (void)h;
      } else {
        c[j] = A.cols[p];
        h[j] = vals[p];
      }
    }
  };
  auto fma_x = [&](const V *tile, uint32_t c, T h) {
    const V *xr = tile + c * VPR;
#pragma unroll
    for (int p = 0; p < VPR; ++p) {
      const V u = xr[rot[p]];
#pragma unroll
      for (int i = 0; i < VW; ++i) acc[p * VW + i] += h * u[i];
    }
  };
  auto fma_row = [&](const V *tile, const T *dict, uint32_t c, W hw) {
    if constexpr (CODED) fma_x(tile, c & 0xffffu, dict[c >> 16]), (void)hw;
    else fma_x(tile, c, hw), (void)dict;
  };
  // One tile step: the next tile's X pieces and the first slot group of its panel are
  // loaded into the "next" register set while the current tile is multiplied; long
  // panels alternate between two groups (one in flight while the other is used).  The
  // step loop is unrolled by two so the sets trade roles instead of being copied: a copy
  // of a register whose load is in flight makes the compiler wait for every load issued
  // before it, the next tile's included.
  uint32_t c0s[PF], c1s[PF], cx[PF];
  W h0s[PF], h1s[PF], hx[PF];
  int64_t base0 = 0, base1 = 0;
  int L0 = 0, L1 = 0;
  // slots j0 + j past the panel's end L are skipped here at the use (a branch uniform over
  // the wavefront), not at the load, so the loads stay in flight until the group is reached
  auto group = [&](const V *cur, const T *cd, const uint32_t (&c)[PF], const W (&h)[PF], int j0, int L) {
#pragma unroll
    for (int j = 0; j < PF; ++j)
      if (j0 + j < L) fma_row(cur, cd, c[j], h[j]);
  };
  auto step = [&](int kk, const V *cur, const T *cd, V *nxt, T *nd, uint32_t (&ca)[PF], W (&ha)[PF],
                  int64_t base_a, int La, uint32_t (&cb)[PF], W (&hb)[PF], int64_t &base_b, int &Lb) {
    // the next step's loads are issued unconditionally (the last step reloads its own tile
    // and slots, unused): a load under a branch would make the first use after the join
    // wait for every load in flight, the next tile's included
    const bool more = kk + 1 < K;
    const int kn = more ? kk + 1 : kk;
    slice_of(kn, base_b, Lb);
    load_slots(base_b, Lb, 0, cb, hb);
    if (NB == 2 && more) dma_tile(kn, nxt, nd);  // the buffers read in the previous step: free since its barrier
    if (wave_on && La > 0) {
      group(cur, cd, ca, ha, 0, La);
      if (PF < La) {
        load_slots(base_a, La, PF, cx, hx);
        for (int j0 = PF; j0 < La; j0 += 2 * PF) {
          if (j0 + PF < La) load_slots(base_a, La, j0 + PF, ca, ha);
          group(cur, cd, cx, hx, j0, La);
          if (j0 + PF >= La) break;
          if (j0 + 2 * PF < La) load_slots(base_a, La, j0 + 2 * PF, cx, hx);
          group(cur, cd, ca, ha, j0 + PF, La);
        }
      }
    }
    if (NB == 1) {  // one buffer: the next tile is loaded once every wavefront is done with it
      __syncthreads();
      if (more) dma_tile(kn, nxt, nd);
    }
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this wavefront's direct loads have landed
    __syncthreads();
  };
  // the first pipelined tile (and its dictionary) straight from memory, beside the table's
  // loads; with two buffers it lands while the value pairs (coded kernel) are multiplied in
  // the other one
  auto first_tile = [&]() {
    dma(A.pair_c0[k0], A.pair_rows[k0] * VPR, buf0);
    if (CODED) dma_dict(A.dict_off[k0], A.dict_n4[k0], dict0);
  };
  if (NB == 2 && K > 0) first_tile();
  if constexpr (CODED) {
    // value pairs one at a time in buf1 (buf0 with one buffer): tile, wait, multiply; slot
    // groups of SELL_PF with the next group in flight
    const int4 e0 = B.items[2 * b], e1 = B.items[2 * b + 1];
    const int bk0 = e0.x, bK = e0.y;
    uint32_t cva[SELL_PF], cvb[SELL_PF];
    T hva[SELL_PF], hvb[SELL_PF];
    auto vload = [&](int64_t base, int j0, uint32_t (&c)[SELL_PF], T (&h)[SELL_PF]) {
#pragma unroll
      for (int j = 0; j < SELL_PF; ++j) {
        const int64_t p = base + (int64_t)(j0 + j) * 64;
        c[j] = B.cols[p];
        h[j] = bvals[p];
      }
    };
    auto vgroup = [&](const uint32_t (&c)[SELL_PF], const T (&h)[SELL_PF], int j0, int L) {
#pragma unroll
      for (int j = 0; j < SELL_PF; ++j)
        if (j0 + j < L) fma_x(buf1, c[j], h[j]);
    };
    for (int i = 0; i < bK; ++i) {
      const int64_t q = (int64_t)e1.y + (int64_t)i * ns + wc;
      const int64_t sb = B.sl_base[q];
      const int L = __builtin_amdgcn_readfirstlane((int)((B.sl_base[q + 1] - sb) >> 6));
      const int64_t base = sb + lane;
      dma(B.pair_c0[bk0 + i], B.pair_rows[bk0 + i] * VPR, buf1);
      vload(base, 0, cva, hva);
      __builtin_amdgcn_s_waitcnt(0x0f70);
      __syncthreads();
      if (wave_on)
        for (int j0 = 0; j0 < L; j0 += 2 * SELL_PF) {
          if (j0 + SELL_PF < L) vload(base, j0 + SELL_PF, cvb, hvb);
          vgroup(cva, hva, j0, L);
          if (j0 + SELL_PF >= L) break;
          if (j0 + 2 * SELL_PF < L) vload(base, j0 + 2 * SELL_PF, cva, hva);
          vgroup(cvb, hvb, j0 + SELL_PF, L);
        }
      __syncthreads();  // buf1 is loaded again
    }
  }
  if (K > 0) {
    // the first slots; the barrier after them also publishes the table
    if (NB == 1) first_tile();
    const int64_t q0 = A.sl_base[qb + wc];
    base0 = q0 + lane;
    L0 = __builtin_amdgcn_readfirstlane((int)((A.sl_base[qb + wc + 1] - q0) >> 6));
    load_slots(base0, L0, 0, c0s, h0s);
    __builtin_amdgcn_s_waitcnt(0x0f70);
    __syncthreads();
    for (int kk = 0; kk < K; kk += 2) {
      step(kk, buf0, dict0, buf1, dict1, c0s, h0s, base0, L0, c1s, h1s, base1, L1);
      if (kk + 1 < K) step(kk + 1, buf1, dict1, buf0, dict0, c1s, h1s, base1, L1, c0s, h0s, base0, L0);
    }
  }
  if (tid < S) {
    T *y = Y + myrow * LD;
#pragma unroll
    for (int p = 0; p < VPR; ++p) {
      // accumulator block p holds piece rot[p] of the row
      if ((rot[p] + 1) * VW <= m) {
        V u;
#pragma unroll
        for (int i = 0; i < VW; ++i) u[i] = acc[p * VW + i];
        reinterpret_cast<V *>(y)[rot[p]] = u;
      } else {
#pragma unroll
        for (int i = 0; i < VW; ++i)
          if (rot[p] * VW + i < m) y[rot[p] * VW + i] = acc[p * VW + i];
      }
    }
  }
}

template <typename T, int LD, bool CODED, int NB>
__global__ __launch_bounds__(SELL_MAXROWS) void k_sell_apply(SellView A, const T *__restrict__ vals, SellView B,
                                                            const T *__restrict__ bvals, const T *__restrict__ X,
                                                            T *__restrict__ Y, int m, int chunk, int accum) {
  sell_apply_body<T, LD, CODED, NB>(A, vals, B, bvals, X, Y, m, chunk, accum);
}

// Proton-spectator part of Y = H X: one workgroup per unit (a proton configuration and
// up to SELL_MAXROWS neutron configurations of its class), one lane per row.  The X rows
// of the whole proton-major tile (every member of the class at the unit's proton rank:
// one 16-byte-aligned row per member, scattered through X) are gathered straight into LDS
// (global_load_lds) behind a zero row; each lane then runs its ELL panel (16-bit member
// index + 1, value) against them and writes its whole Y row.  Every row of the operator
// belongs to exactly one unit, so this launch writes every row of Y once.
struct PmView {
  const int4 *items;
  const uint32_t *rows;
  const int64_t *base;
  const uint16_t *cols;
};

template <typename T, int LD>
__global__ __launch_bounds__(SELL_MAXROWS) void k_pm_apply(PmView P, const T *__restrict__ vals,
                                                          const T *__restrict__ X, T *__restrict__ Y, int m,
                                                          int tile_pieces) {
  typedef typename V16t<T>::type V;
  constexpr int VPR = (int)(LD * sizeof(T) / 16);
  constexpr int VW = (int)(16 / sizeof(T));
  constexpr int PF = 4;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int u = sell_block(8);
  const int4 u0 = P.items[2 * u], u1 = P.items[2 * u + 1];
  const int C = u0.y, i0 = u0.w, S = u1.x;
  const uint32_t rho = (uint32_t)u0.z;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nthr = blockDim.x;
  const int ns = (S + 63) >> 6, wc = min(w, ns - 1);
  // LDS: the zero row, then the tile (tile_pieces pieces: C rows rounded up to whole
  // passes of the workgroup), then the members' X rows
  V *tile = reinterpret_cast<V *>(lds_raw);
  uint32_t *xrow = reinterpret_cast<uint32_t *>(tile + VPR + tile_pieces);
  for (int j = tid; j < C; j += nthr) xrow[j] = P.rows[u0.x + j] + rho;
  if (tid < VPR) tile[tid] = V{};
  // this lane's slice: slots base + 64 j, j < L; the first group is loaded before the tile
  const int64_t sb = P.base[u1.y + wc];
  const int L = __builtin_amdgcn_readfirstlane((int)((P.base[u1.y + wc + 1] - sb) >> 6));
  const int64_t base = sb + lane;
  uint32_t ca[PF], cb[PF];
  T ha[PF], hb[PF];
  auto load = [&](int j0, uint32_t (&c)[PF], T (&h)[PF]) {
#pragma unroll
    for (int j = 0; j < PF; ++j) {  // the slot arrays carry 4 * PF * 64 slots of slack
      c[j] = P.cols[base + (int64_t)(j0 + j) * 64];
      h[j] = vals[base + (int64_t)(j0 + j) * 64];
    }
  };
  load(0, ca, ha);
  __syncthreads();
  const int np = C * VPR;
  for (int b0 = 0; b0 < np; b0 += nthr) {
    const int idx = min(b0 + tid, np - 1), j = idx / VPR, p = idx - j * VPR;
    const V *src = reinterpret_cast<const V *>(X + (int64_t)xrow[j] * LD) + p;
    __builtin_amdgcn_global_load_lds((const void *)src,
                                     (__attribute__((address_space(3))) void *)(tile + VPR + b0 + w * 64), 16, 0, 0);
  }
  int rot[VPR];
#pragma unroll
  for (int p = 0; p < VPR; ++p) rot[p] = (p + lane) & (VPR - 1);
  T acc[LD];
#pragma unroll
  for (int i = 0; i < LD; ++i) acc[i] = (T)0;
  __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): the gathered rows have landed
  __syncthreads();
  if (w >= ns) return;  // no barrier below
  auto group = [&](const uint32_t (&c)[PF], const T (&h)[PF], int j0) {
#pragma unroll
    for (int j = 0; j < PF; ++j)
      if (j0 + j < L) {
        const V *xr = tile + c[j] * VPR;
#pragma unroll
        for (int p = 0; p < VPR; ++p) {
          const V x = xr[rot[p]];
#pragma unroll
          for (int i = 0; i < VW; ++i) acc[p * VW + i] += h[j] * x[i];
        }
      }
  };
  for (int j0 = 0; j0 < L; j0 += 2 * PF) {
    load(j0 + PF, cb, hb);
    group(ca, ha, j0);
    if (j0 + PF >= L) break;
    load(j0 + 2 * PF, ca, ha);
    group(cb, hb, j0 + PF);
  }
  if (tid < S) {
    T *y = Y + (int64_t)xrow[i0 + tid] * LD;
#pragma unroll
    for (int p = 0; p < VPR; ++p) {
      if ((rot[p] + 1) * VW <= m) {
        V v;
#pragma unroll
        for (int i = 0; i < VW; ++i) v[i] = acc[p * VW + i];
        reinterpret_cast<V *>(y)[rot[p]] = v;
      } else {
#pragma unroll
        for (int i = 0; i < VW; ++i)
          if (rot[p] * VW + i < m) y[rot[p] * VW + i] = acc[p * VW + i];
      }
    }
  }
}

// ---- host side ------------------------------------------------------------------------
static std::mutex g_sell_mu;

void free_sell(Op &op) {
  Sell *s = (Sell *)op.sell;
  if (!s) return;
  for (Panels &P : s->P)
    for (void *p : {(void *)P.nbr_start, (void *)P.nbr_tile, (void *)P.qstart, (void *)P.sl_base, (void *)P.cols,
                    (void *)P.v32, (void *)P.v64, (void *)P.pk32, (void *)P.pk64, (void *)P.dict32,
                    (void *)P.dict64, (void *)P.dict_off, (void *)P.dict_n4, (void *)P.pair_c0,
                    (void *)P.pair_rows, (void *)P.items})
      dfree(p);
  for (void *p : {(void *)s->boff, (void *)s->lrow, (void *)s->rid, (void *)s->cls_local, (void *)s->pm_rows,
                  (void *)s->pm_items, (void *)s->pm_base, (void *)s->pm_cols, (void *)s->pm_v32, (void *)s->pm_v64})
    dfree(p);
  delete s;
  op.sell = nullptr;
}

size_t sell_bytes(const Op &op) { return op.sell ? ((const Sell *)op.sell)->bytes : 0; }

// row blocks: the operator's row partition (neutron-configuration blocks) with
// consecutive blocks merged while they fit SELL_MAXROWS rows and longer ones split; an
// operator without a partition gets blocks of SELL_MAXROWS rows
static std::vector<int64_t> sell_partition(const Op &op) {
  const int64_t n = op.n;
  std::vector<int64_t> p = op.part;
  if (p.size() < 2 || p.front() != 0 || p.back() != n) {
    p.clear();
    for (int64_t r = 0; r < n; r += SELL_MAXROWS) p.push_back(r);
    p.push_back(n);
  }
  std::vector<int64_t> out{0};  // out.back(): first row of the block being formed
  for (size_t i = 0; i + 1 < p.size(); ++i) {
    const int64_t a = p[i], e = p[i + 1];
    if (e - a > SELL_MAXROWS) {  // an oversized block: equal pieces of its own
      if (a > out.back()) out.push_back(a);
      const int64_t pieces = (e - a + SELL_MAXROWS - 1) / SELL_MAXROWS, step = (e - a + pieces - 1) / pieces;
      for (int64_t x = a + step; x < e; x += step) out.push_back(x);
      out.push_back(e);
      continue;
    }
    if (e - out.back() > SELL_MAXROWS) out.push_back(a);  // [a, e) starts a new block
  }
  if (out.back() != n) out.push_back(n);
  return out;
}


// LDS bytes of k_pm_apply at a row pitch of `row` bytes: zero row, tile, member rows
static size_t pm_lds_bytes(const Sell &s, size_t row) {
  const int64_t vpr = (int64_t)row / 16, nthr = s.pm_nwaves * 64;
  const int64_t pieces = (s.pm_maxC * vpr + nthr - 1) / nthr * nthr;
  return 16 * (size_t)(vpr + pieces) + 4 * (size_t)s.pm_maxC;
}

// The proton-spectator split of an operator built from a full basis: classes are the
// neutron blocks that share a proton sector (so rows of a class line up proton rank by
// proton rank).  Fills s.rid, s.cls_local, s.pm_rows, s.pm_items and the unit counts;
// returns false (no split) when the operator carries no basis structure, a class's blocks
// differ in length, or a tile at the widest pitch would not fit LDS.
static bool plan_pm(Op &op, Sell &s, hipStream_t st, int &rc) {
  rc = COSMO_OK;
  const std::vector<int64_t> &P = op.part;
  const std::vector<int64_t> &B = op.part_pbase;
  if (!g_sell_split || P.size() < 2 || B.size() + 1 != P.size() || P.back() != op.n) return false;
  const int nb = (int)B.size();
  std::map<int64_t, std::vector<int>> cls;
  for (int v = 0; v < nb; ++v) cls[B[v]].push_back(v);
  int cmax = 0;
  for (auto &kv : cls) {
    const int64_t S = P[kv.second[0] + 1] - P[kv.second[0]];
    for (int v : kv.second)
      if (P[v + 1] - P[v] != S) return false;
    cmax = std::max<int>(cmax, (int)kv.second.size());
  }
  if (cmax >= 65535) return false;
  s.pm_maxC = cmax;
  s.pm_nwaves = (std::min(cmax, SELL_MAXROWS) + 63) / 64;
  if (pm_lds_bytes(s, 16 * sizeof(double)) > 160 * 1024) return false;
  std::vector<uint64_t> hrid(op.n);
  std::vector<int32_t> loc(nb);
  std::vector<uint32_t> rows;
  std::vector<int4> items;
  int64_t q = 0;
  for (auto &kv : cls) {
    const std::vector<int> &mem = kv.second;
    const int C = (int)mem.size(), first = (int)rows.size();
    const int64_t S = P[mem[0] + 1] - P[mem[0]];
    for (int i = 0; i < C; ++i) {
      const int v = mem[i];
      loc[v] = i;
      rows.push_back((uint32_t)P[v]);
      for (int64_t r = P[v]; r < P[v + 1]; ++r) hrid[r] = ((uint64_t)v << 32) | (uint64_t)(B[v] + (r - P[v]));
    }
    for (int64_t rho = 0; rho < S; ++rho)
      for (int i0 = 0; i0 < C; i0 += SELL_MAXROWS) {
        const int rows_u = std::min(SELL_MAXROWS, C - i0);
        items.push_back(make_int4(first, C, (int)rho, i0));
        items.push_back(make_int4(rows_u, (int)q, 0, 0));
        q += (rows_u + 63) / 64;
        if (q >= INT32_MAX || rho >= INT32_MAX) return false;
      }
  }
  s.pm_units = (int)(items.size() / 2);
  s.pm_Q = q;
  if (dmalloc(&s.rid, sizeof(uint64_t) * op.n) || dmalloc(&s.cls_local, sizeof(int32_t) * nb) ||
      dmalloc(&s.pm_rows, sizeof(uint32_t) * rows.size()) || dmalloc(&s.pm_items, sizeof(int4) * items.size())) {
    rc = fail(COSMO_ENOMEM, "tiled layout: allocation failed");
    return false;
  }
  if (hipMemcpyAsync(s.rid, hrid.data(), sizeof(uint64_t) * op.n, hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(s.cls_local, loc.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(s.pm_rows, rows.data(), sizeof(uint32_t) * rows.size(), hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(s.pm_items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice, st) ||
      hipStreamSynchronize(st)) {
    rc = fail(COSMO_EHIP, "tiled layout: upload failed");
    return false;
  }
  s.bytes += sizeof(uint64_t) * op.n + sizeof(int32_t) * nb + sizeof(uint32_t) * rows.size() +
             sizeof(int4) * items.size();
  return true;
}

// the proton-spectator panels: lengths, slot offsets, columns
static int build_pm(const Op &op, Sell &s, hipStream_t st) {
  const CsrDev &M = op.csr;
  const int thr = s.pm_nwaves * 64;
  int64_t *len = nullptr;
  if (dmalloc(&len, sizeof(int64_t) * std::max<int64_t>(s.pm_Q, 1)) || dmalloc(&s.pm_base, sizeof(int64_t) * (s.pm_Q + 1)))
    return fail(COSMO_ENOMEM, "tiled layout: allocation failed");
  hipLaunchKernelGGL(k_pm_slots, dim3(s.pm_units), dim3(thr), 0, st, 0, (const int4 *)s.pm_items,
                     (const uint32_t *)s.pm_rows, (const uint64_t *)s.rid, (const int32_t *)s.cls_local,
                     (const int64_t *)M.rowptr, (const uint32_t *)M.cols, (const double *)nullptr,
                     (const float *)nullptr, len, (const int64_t *)nullptr, (uint16_t *)nullptr, (double *)nullptr,
                     (float *)nullptr);
  HIPCHK(hipGetLastError());
  std::vector<int64_t> hl(s.pm_Q), hb(s.pm_Q + 1, 0);
  HIPCHK(hipMemcpyAsync(hl.data(), len, sizeof(int64_t) * s.pm_Q, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  dfree(len);
  for (int64_t q = 0; q < s.pm_Q; ++q) hb[q + 1] = hb[q] + 64 * hl[q];
  s.pm_E = hb[s.pm_Q];
  s.pm_Ealloc = s.pm_E + 4 * PM_PF * 64;
  if (dmalloc(&s.pm_cols, sizeof(uint16_t) * s.pm_Ealloc)) return fail(COSMO_ENOMEM, "tiled layout: allocation failed");
  HIPCHK(hipMemcpyAsync(s.pm_base, hb.data(), sizeof(int64_t) * (s.pm_Q + 1), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(s.pm_cols, 0, sizeof(uint16_t) * s.pm_Ealloc, st));  // padding: the zero row
  hipLaunchKernelGGL(k_pm_slots, dim3(s.pm_units), dim3(thr), 0, st, 1, (const int4 *)s.pm_items,
                     (const uint32_t *)s.pm_rows, (const uint64_t *)s.rid, (const int32_t *)s.cls_local,
                     (const int64_t *)M.rowptr, (const uint32_t *)M.cols, (const double *)nullptr,
                     (const float *)nullptr, (int64_t *)nullptr, (const int64_t *)s.pm_base, s.pm_cols,
                     (double *)nullptr, (float *)nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  s.bytes += sizeof(int64_t) * (s.pm_Q + 1) + sizeof(uint16_t) * s.pm_Ealloc;
  return COSMO_OK;
}

// one panel set from its per-block tile lists (hs: first pair of each block, ht: tiles):
// panel lengths, slot offsets and columns, the pair table and the block items
static int build_panels(Sell &s, Panels &P, const std::vector<int64_t> &hb, const std::vector<int32_t> &hs,
                        const std::vector<int32_t> &ht, const CsrDev &M, const uint64_t *rid, hipStream_t st) {
  const int nblk = s.nblk;
  P.K = hs[nblk];
  std::vector<int64_t> hq(nblk);
  int64_t Q = 0;
  for (int b = 0; b < nblk; ++b) {
    hq[b] = Q;
    Q += (int64_t)(hs[b + 1] - hs[b]) * ((hb[b + 1] - hb[b] + 63) / 64);
  }
  if (Q >= INT32_MAX) return fail(COSMO_EUNSUP, "tiled layout: too many slices");
  P.Q = Q;
  int64_t *len = nullptr;
  if (dmalloc(&P.nbr_start, sizeof(int32_t) * (nblk + 1)) ||
      dmalloc(&P.nbr_tile, sizeof(int32_t) * std::max<int64_t>(P.K, 1)) ||
      dmalloc(&P.qstart, sizeof(int64_t) * nblk) || dmalloc(&P.sl_base, sizeof(int64_t) * (Q + 1)) ||
      dmalloc(&len, sizeof(int64_t) * std::max<int64_t>(Q, 1)))
    return fail(COSMO_ENOMEM, "tiled layout: allocation failed");
  HIPCHK(hipMemcpyAsync(P.nbr_start, hs.data(), sizeof(int32_t) * (nblk + 1), hipMemcpyHostToDevice, st));
  if (P.K) HIPCHK(hipMemcpyAsync(P.nbr_tile, ht.data(), sizeof(int32_t) * P.K, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(P.qstart, hq.data(), sizeof(int64_t) * nblk, hipMemcpyHostToDevice, st));
  const int thr = s.nwaves * 64;
  hipLaunchKernelGGL(k_sell_slots, dim3(nblk), dim3(thr), 0, st, 0, (const int64_t *)s.boff,
                     (const int64_t *)M.rowptr, (const uint32_t *)M.cols, (const double *)nullptr,
                     (const float *)nullptr, (const int32_t *)P.nbr_start, (const int32_t *)P.nbr_tile,
                     (const int64_t *)P.qstart, len, (const int64_t *)nullptr, (uint16_t *)nullptr,
                     (double *)nullptr, (float *)nullptr, rid, (const uint16_t *)s.lrow);
  HIPCHK(hipGetLastError());
  std::vector<int64_t> hl(std::max<int64_t>(Q, 1)), hbase(Q + 1, 0);
  if (Q) HIPCHK(hipMemcpyAsync(hl.data(), len, sizeof(int64_t) * Q, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  dfree(len);
  for (int64_t q = 0; q < Q; ++q) hbase[q + 1] = hbase[q] + 64 * hl[q];
  P.E = hbase[Q];
  // the product keeps each block's slice offsets in LDS as 16-bit counts of 64 slots
  for (int b = 0; b < nblk; ++b) {
    P.maxtiles = std::max(P.maxtiles, hs[b + 1] - hs[b]);
    const int64_t qe = hq[b] + (int64_t)(hs[b + 1] - hs[b]) * ((hb[b + 1] - hb[b] + 63) / 64);
    if ((hbase[qe] - hbase[hq[b]]) / 64 > 65535) return fail(COSMO_EUNSUP, "a row block too long for the tiled product");
  }
  P.Ealloc = P.E + 64 * SELL_PFC;  // a lane reads up to SELL_PFC slots past its slice's panel
  if (dmalloc(&P.cols, sizeof(uint16_t) * P.Ealloc)) return fail(COSMO_ENOMEM, "tiled layout: allocation failed");
  HIPCHK(hipMemcpyAsync(P.sl_base, hbase.data(), sizeof(int64_t) * (Q + 1), hipMemcpyHostToDevice, st));
  // padding slots point at the zero row kept after every staged tile (index nwaves * 64)
  HIPCHK(hipMemsetD16Async((hipDeviceptr_t)P.cols, (unsigned short)(s.nwaves * 64), P.Ealloc, st));
  hipLaunchKernelGGL(k_sell_slots, dim3(nblk), dim3(thr), 0, st, 1, (const int64_t *)s.boff,
                     (const int64_t *)M.rowptr, (const uint32_t *)M.cols, (const double *)nullptr,
                     (const float *)nullptr, (const int32_t *)P.nbr_start, (const int32_t *)P.nbr_tile,
                     (const int64_t *)P.qstart, (int64_t *)nullptr, (const int64_t *)P.sl_base, P.cols,
                     (double *)nullptr, (float *)nullptr, rid, (const uint16_t *)s.lrow);
  HIPCHK(hipGetLastError());
  // the product's per-pair tile table and whole-block items (rows and slices held as int32)
  std::vector<uint32_t> pc(std::max<int64_t>(P.K, 1), 0);
  std::vector<int32_t> pr(std::max<int64_t>(P.K, 1), 0);
  for (int64_t k = 0; k < P.K; ++k) {
    const int t = ht[k];
    pc[k] = (uint32_t)hb[t];
    pr[k] = (int32_t)(hb[t + 1] - hb[t]);
  }
  std::vector<int4> it(2 * (size_t)nblk);
  for (int b = 0; b < nblk; ++b) {
    it[2 * b] = make_int4(hs[b], hs[b + 1] - hs[b], (int)(hb[b + 1] - hb[b]), 0);
    it[2 * b + 1] = make_int4((int)hb[b], (int)hq[b], 0, 0);
  }
  if (dmalloc(&P.pair_c0, sizeof(uint32_t) * pc.size()) || dmalloc(&P.pair_rows, sizeof(int32_t) * pr.size()) ||
      dmalloc(&P.items, sizeof(int4) * it.size()))
    return fail(COSMO_ENOMEM, "tiled layout: allocation failed");
  HIPCHK(hipMemcpyAsync(P.pair_c0, pc.data(), sizeof(uint32_t) * pc.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(P.pair_rows, pr.data(), sizeof(int32_t) * pr.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(P.items, it.data(), sizeof(int4) * it.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  s.bytes += sizeof(int32_t) * (nblk + 1 + P.K) + sizeof(int64_t) * (nblk + Q + 1) + (size_t)P.Ealloc * 2 +
             8 * (size_t)P.K + 32 * (size_t)nblk;
  return COSMO_OK;
}

// layout statistics of one panel set on stderr
static void panel_stats(const char *what, const Sell &s, const Panels &P, const std::vector<int64_t> &hb,
                        const std::vector<int32_t> &ht, const CsrDev &M, const uint64_t *rid, hipStream_t st) {
  const int nblk = s.nblk;
  const int64_t n = hb[nblk];
  int64_t staged = 0;  // every (row block, column tile) pair loads the tile's rows
  for (int64_t k = 0; k < P.K; ++k) staged += hb[ht[k] + 1] - hb[ht[k]];
  unsigned long long *dd = nullptr, distinct = 0;
  if (P.K && !dmalloc(&dd, sizeof(unsigned long long))) {
    if (!hipMemsetAsync(dd, 0, sizeof(unsigned long long), st)) {
      hipLaunchKernelGGL(k_sell_distinct, dim3(nblk), dim3(s.nwaves * 64), 0, st, (const int64_t *)s.boff,
                         (const int64_t *)M.rowptr, (const uint32_t *)M.cols, (const int32_t *)P.nbr_start,
                         (const int32_t *)P.nbr_tile, dd, rid);
      if (hipMemcpyAsync(&distinct, dd, sizeof(distinct), hipMemcpyDeviceToHost, st) || hipStreamSynchronize(st))
        distinct = 0;
    }
    dfree(dd);
  }
  fprintf(stderr,
          "[sell] %s panels: pairs=%lld (<= %d per block) slices=%lld slots=%lld staged X rows=%lld (%.3f n) "
          "referenced=%llu (%.3f n)\n",
          what, (long long)P.K, P.maxtiles, (long long)P.Q, (long long)P.E, (long long)staged,
          (double)staged / (double)n, distinct, (double)distinct / (double)n);
  if (P.coded)
    fprintf(stderr, "[sell] %s panels: dictionary entries %lld (%.1f per pair)\n", what, (long long)P.dict_total,
            P.K ? (double)P.dict_total / (double)P.K : 0.0);
}

static int build_sell(Op &op, hipStream_t st) {
  const CsrDev &M = op.csr;
  const int64_t n = op.n;
  if (n >= INT32_MAX) return fail(COSMO_EUNSUP, "tiled layout: too many rows");
  std::vector<int64_t> hb = sell_partition(op);
  const int nblk = (int)hb.size() - 1;
  if (nblk <= 0) return fail(COSMO_EINVAL, "empty operator");
  if (nblk > (1 << 17)) return fail(COSMO_EUNSUP, "too many row blocks for the tiled product");
  int maxrows = 0;
  for (int b = 0; b < nblk; ++b) maxrows = std::max<int>(maxrows, (int)(hb[b + 1] - hb[b]));
  auto *s = new Sell();
  s->nblk = nblk;
  s->maxrows = maxrows;
  s->nwaves = (maxrows + 63) / 64;
  int32_t *blk_of = nullptr, *cnt = nullptr, *all_start = nullptr, *all_tile = nullptr, *dcnt = nullptr;
  auto bail = [&](int rc) {
    for (void *p : {(void *)blk_of, (void *)cnt, (void *)all_start, (void *)all_tile, (void *)dcnt}) dfree(p);
    op.sell = s;
    free_sell(op);
    return rc;
  };
  {
    int rc = COSMO_OK;
    s->split = plan_pm(op, *s, st, rc);
    if (rc) return bail(rc);
    if (s->split && (rc = build_pm(op, *s, st))) return bail(rc);
  }
  const uint64_t *rid = s->split ? s->rid : nullptr;
  // lane order: the rows of each block by their numbers of entries outside the
  // proton-spectator part, most first (ties in row order)
  {
    int32_t *wt = nullptr;
    if (dmalloc(&wt, sizeof(int32_t) * n) || dmalloc(&s->lrow, sizeof(uint16_t) * n)) {
      dfree(wt);
      return bail(fail(COSMO_ENOMEM, "tiled layout: allocation failed"));
    }
    hipLaunchKernelGGL(k_row_weight, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
                       (const int64_t *)M.rowptr, (const uint32_t *)M.cols, rid, wt);
    std::vector<int32_t> hw(n);
    if (hipGetLastError() || hipMemcpyAsync(hw.data(), wt, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st) ||
        hipStreamSynchronize(st)) {
      dfree(wt);
      return bail(fail(COSMO_EHIP, "tiled layout: row weights failed"));
    }
    dfree(wt);
    std::vector<uint16_t> hl(n);
    std::vector<int> idx;
    for (int b = 0; b < nblk; ++b) {
      const int S = (int)(hb[b + 1] - hb[b]);
      idx.resize(S);
      for (int i = 0; i < S; ++i) idx[i] = i;
      const int32_t *wb = hw.data() + hb[b];
      std::stable_sort(idx.begin(), idx.end(), [&](int a, int c) { return wb[a] > wb[c]; });
      for (int i = 0; i < S; ++i) hl[hb[b] + i] = (uint16_t)idx[i];
    }
    HIPCHK(hipMemcpyAsync(s->lrow, hl.data(), sizeof(uint16_t) * n, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    s->bytes += sizeof(uint16_t) * n;
  }
  // every (row block, column tile) pair with entries outside the proton-spectator part
  std::vector<int32_t> hblk(n);
  for (int b = 0; b < nblk; ++b)
    for (int64_t r = hb[b]; r < hb[b + 1]; ++r) hblk[r] = b;
  if (dmalloc(&s->boff, sizeof(int64_t) * (nblk + 1)) || dmalloc(&blk_of, sizeof(int32_t) * n) ||
      dmalloc(&cnt, sizeof(int32_t) * nblk) || dmalloc(&all_start, sizeof(int32_t) * (nblk + 1)))
    return bail(fail(COSMO_ENOMEM, "tiled layout: allocation failed"));
  HIPCHK(hipMemcpyAsync(s->boff, hb.data(), sizeof(int64_t) * (nblk + 1), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(blk_of, hblk.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
  const size_t bm_bytes = sizeof(uint32_t) * (size_t)((nblk + 31) / 32);
  hipLaunchKernelGGL(k_sell_tiles, dim3(nblk), dim3(256), bm_bytes, st, 0, (const int64_t *)s->boff,
                     (const int64_t *)M.rowptr, (const uint32_t *)M.cols, (const int32_t *)blk_of, nblk, cnt,
                     (const int32_t *)nullptr, (int32_t *)nullptr, rid);
  HIPCHK(hipGetLastError());
  std::vector<int32_t> hc(nblk), hs(nblk + 1, 0);
  HIPCHK(hipMemcpyAsync(hc.data(), cnt, sizeof(int32_t) * nblk, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int b = 0; b < nblk; ++b) hs[b + 1] = hs[b] + hc[b];
  const int64_t K = hs[nblk];
  if (dmalloc(&all_tile, sizeof(int32_t) * std::max<int64_t>(K, 1)) ||
      dmalloc(&dcnt, sizeof(int32_t) * std::max<int64_t>(K, 1)))
    return bail(fail(COSMO_ENOMEM, "tiled layout: allocation failed"));
  HIPCHK(hipMemcpyAsync(all_start, hs.data(), sizeof(int32_t) * (nblk + 1), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_sell_tiles, dim3(nblk), dim3(256), bm_bytes, st, 1, (const int64_t *)s->boff,
                     (const int64_t *)M.rowptr, (const uint32_t *)M.cols, (const int32_t *)blk_of, nblk, cnt,
                     (const int32_t *)all_start, all_tile, rid);
  HIPCHK(hipGetLastError());
  // the distinct FP64 values of each pair (an FP32 copy can only have fewer)
  hipLaunchKernelGGL((k_sell_dict<double, unsigned long long>), dim3(nblk), dim3(s->nwaves * 64), 0, st, 0,
                     (const int64_t *)s->boff, (const int64_t *)M.rowptr, (const uint32_t *)M.cols,
                     (const double *)M.v64, rid, (const uint16_t *)s->lrow, (const int32_t *)all_start,
                     (const int32_t *)all_tile, dcnt, (const int64_t *)nullptr, (const int64_t *)nullptr,
                     (const int32_t *)nullptr, (double *)nullptr, (uint32_t *)nullptr);
  HIPCHK(hipGetLastError());
  std::vector<int32_t> ht(std::max<int64_t>(K, 1)), hd(std::max<int64_t>(K, 1), SELL_DICT);
  if (K) {
    HIPCHK(hipMemcpyAsync(ht.data(), all_tile, sizeof(int32_t) * K, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hd.data(), dcnt, sizeof(int32_t) * K, hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  for (void *p : {(void *)blk_of, (void *)cnt, (void *)all_start, (void *)all_tile, (void *)dcnt}) dfree(p);
  blk_of = cnt = all_start = all_tile = dcnt = nullptr;
  // pairs of at most SELL_DICT - 1 distinct non-zero values go to the coded panels
  std::vector<int32_t> hs0(nblk + 1, 0), hs1(nblk + 1, 0), ht0, ht1, hn1;
  for (int b = 0; b < nblk; ++b) {
    for (int k = hs[b]; k < hs[b + 1]; ++k) {
      if (hd[k] < SELL_DICT) ht1.push_back(ht[k]), hn1.push_back(hd[k]);
      else ht0.push_back(ht[k]);
    }
    hs0[b + 1] = (int32_t)ht0.size();
    hs1[b + 1] = (int32_t)ht1.size();
  }
  int rc = build_panels(*s, s->P[0], hb, hs0, ht0, M, rid, st);
  if (!rc && !ht1.empty()) rc = build_panels(*s, s->P[1], hb, hs1, ht1, M, rid, st);
  if (rc) return bail(rc);
  Panels &C = s->P[1];
  if (C.K) {
    C.coded = true;
    std::vector<int32_t> off(C.K), n4(C.K);
    int64_t tot = 0;
    for (int64_t k = 0; k < C.K; ++k) {
      off[k] = (int32_t)tot;
      n4[k] = (hn1[k] + 1 + 3) / 4;  // entry 0 (0.0) and the values, in groups of 4
      tot += 4 * (int64_t)n4[k];
    }
    if (tot >= INT32_MAX) return bail(fail(COSMO_EUNSUP, "tiled layout: dictionaries too large"));
    C.dict_total = tot;
    if (dmalloc(&C.dict_off, sizeof(int32_t) * C.K) || dmalloc(&C.dict_n4, sizeof(int32_t) * C.K))
      return bail(fail(COSMO_ENOMEM, "tiled layout: allocation failed"));
    HIPCHK(hipMemcpyAsync(C.dict_off, off.data(), sizeof(int32_t) * C.K, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(C.dict_n4, n4.data(), sizeof(int32_t) * C.K, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    s->bytes += 8 * (size_t)C.K;
  }
  s->bytes += sizeof(int64_t) * (nblk + 1);
  if (g_sell_stats) {
    fprintf(stderr, "[sell] n=%lld blocks=%d (rows <= %d) stored=%lld\n", (long long)n, nblk, maxrows,
            (long long)M.nnz);
    panel_stats("value", *s, s->P[0], hb, ht0, M, rid, st);
    if (C.K) panel_stats("coded", *s, C, hb, ht1, M, rid, st);
    if (s->split)
      fprintf(stderr, "[sell] proton-spectator part: units=%d (class rows <= %d) slices=%lld slots=%lld\n",
              s->pm_units, s->pm_maxC, (long long)s->pm_Q, (long long)s->pm_E);
  }
  op.sell = s;
  return COSMO_OK;
}

// the slot values of one precision, filled on its first product (the FP32 and the FP64
// copies are only kept for the precisions in use): values of the value panels and of the
// proton-spectator part, codes and dictionaries of the coded panels
static int sell_values(Op &op, bool f32, hipStream_t st) {
  Sell &s = *(Sell *)op.sell;
  Panels &V = s.P[0], &C = s.P[1];
  if (f32 ? V.v32 != nullptr : V.v64 != nullptr) return COSMO_OK;
  const CsrDev &M = op.csr;
  const uint64_t *rid = s.split ? (const uint64_t *)s.rid : (const uint64_t *)nullptr;
  const size_t es = f32 ? sizeof(float) : sizeof(double);
  void *dst = nullptr;
  if (dmalloc_void(&dst, es * V.Ealloc, st)) return fail(COSMO_ENOMEM, "tiled layout: value allocation failed");
  HIPCHK(hipMemsetAsync(dst, 0, es * V.Ealloc, st));
  if (V.K)
    hipLaunchKernelGGL(k_sell_slots, dim3(s.nblk), dim3(s.nwaves * 64), 0, st, 1, (const int64_t *)s.boff,
                       (const int64_t *)M.rowptr, (const uint32_t *)M.cols, (const double *)M.v64,
                       (const float *)M.v32, (const int32_t *)V.nbr_start, (const int32_t *)V.nbr_tile,
                       (const int64_t *)V.qstart, (int64_t *)nullptr, (const int64_t *)V.sl_base, (uint16_t *)nullptr,
                       f32 ? (double *)nullptr : (double *)dst, f32 ? (float *)dst : (float *)nullptr, rid,
                       (const uint16_t *)s.lrow);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  if (f32) V.v32 = (float *)dst;
  else V.v64 = (double *)dst;
  s.bytes += es * (size_t)V.Ealloc;
  if (C.K) {
    uint32_t *code = nullptr;
    void *dict = nullptr;
    if (dmalloc(&code, sizeof(uint32_t) * C.Ealloc) || dmalloc_void(&dict, es * C.dict_total, st)) {
      dfree(code);
      return fail(COSMO_ENOMEM, "tiled layout: code allocation failed");
    }
    // padding slots: the zero row, dictionary entry 0 (the value 0.0)
    HIPCHK(hipMemsetD32Async((hipDeviceptr_t)code, (int)(s.nwaves * 64), C.Ealloc, st));
    HIPCHK(hipMemsetAsync(dict, 0, es * C.dict_total, st));
    if (f32)
      hipLaunchKernelGGL((k_sell_dict<float, unsigned int>), dim3(s.nblk), dim3(s.nwaves * 64), 0, st, 1,
                         (const int64_t *)s.boff, (const int64_t *)M.rowptr, (const uint32_t *)M.cols,
                         (const float *)M.v32, rid, (const uint16_t *)s.lrow, (const int32_t *)C.nbr_start,
                         (const int32_t *)C.nbr_tile, (int32_t *)nullptr, (const int64_t *)C.qstart,
                         (const int64_t *)C.sl_base, (const int32_t *)C.dict_off, (float *)dict, code);
    else
      hipLaunchKernelGGL((k_sell_dict<double, unsigned long long>), dim3(s.nblk), dim3(s.nwaves * 64), 0, st, 1,
                         (const int64_t *)s.boff, (const int64_t *)M.rowptr, (const uint32_t *)M.cols,
                         (const double *)M.v64, rid, (const uint16_t *)s.lrow, (const int32_t *)C.nbr_start,
                         (const int32_t *)C.nbr_tile, (int32_t *)nullptr, (const int64_t *)C.qstart,
                         (const int64_t *)C.sl_base, (const int32_t *)C.dict_off, (double *)dict, code);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    if (f32) C.pk32 = code, C.dict32 = (float *)dict;
    else C.pk64 = code, C.dict64 = (double *)dict;
    s.bytes += sizeof(uint32_t) * (size_t)C.Ealloc + es * (size_t)C.dict_total;
  }
  if (!s.split) return COSMO_OK;
  void *pd = nullptr;
  if (dmalloc_void(&pd, es * s.pm_Ealloc, st)) return fail(COSMO_ENOMEM, "tiled layout: value allocation failed");
  HIPCHK(hipMemsetAsync(pd, 0, es * s.pm_Ealloc, st));
  hipLaunchKernelGGL(k_pm_slots, dim3(s.pm_units), dim3(s.pm_nwaves * 64), 0, st, 1, (const int4 *)s.pm_items,
                     (const uint32_t *)s.pm_rows, (const uint64_t *)s.rid, (const int32_t *)s.cls_local,
                     (const int64_t *)M.rowptr, (const uint32_t *)M.cols, (const double *)M.v64,
                     (const float *)M.v32, (int64_t *)nullptr, (const int64_t *)s.pm_base, (uint16_t *)nullptr,
                     f32 ? (double *)nullptr : (double *)pd, f32 ? (float *)pd : (float *)nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  if (f32) s.pm_v32 = (float *)pd;
  else s.pm_v64 = (double *)pd;
  s.bytes += es * (size_t)s.pm_Ealloc;
  return COSMO_OK;
}

// device bytes the tiled layout of an operator with n rows and nnz stored entries can
// take with both value copies (the padding of the panels stays below 1.7 x on the
// systems measured; codes and dictionaries where pairs are coded): the server's admission
// counts it beside the rows
int64_t sell_footprint_bound(int64_t n, int64_t nnz) {
  if (n < SELL_MIN_N) return 0;
  return (int64_t)(1.7 * (double)nnz) * (2 + 4 + 8 + 2) + 64 * n + 256 * n;
}

int g_sell_chunk = 0;  // XCD chunk of row blocks (COSMO_SELL_CHUNK; 0: chosen by timing)
int g_sell_off = 0;    // COSMO_SELL=0: the row kernel instead
int g_sell_stats = 0;  // COSMO_SELL_STATS=1: the layout's sizes on stderr when it is built
int g_sell_split = 1;  // COSMO_SELL_SPLIT=0: the proton-spectator entries stay in the panels

template <typename T>
static const T *pick(const float *a, const double *b) {
  if constexpr (sizeof(T) == 4) return (const T *)a;
  else return (const T *)b;
}

static SellView panel_view(const Sell &s, const Panels &P, const uint32_t *packed) {
  return SellView{P.maxtiles, s.nwaves * 64, P.items, P.pair_c0, P.pair_rows, P.sl_base, P.cols, P.dict_off,
                  P.dict_n4, packed, s.lrow};
}

template <typename T, int LD>
static size_t panels_lds(const Sell &s, const Panels &P, int nb) {
  return nb * (size_t)(s.nwaves * 64 + 1) * LD * sizeof(T) + sell_meta_bytes(P.maxtiles, s.nwaves) +
         (P.coded ? nb * SELL_DICT * sizeof(T) : 0);
}

int g_sell_nb = 0;  // COSMO_SELL_NB: tile buffers per workgroup of the panel product (1, 2; 0: by timing)

template <typename T, int LD, bool CODED, int NB>
static void launch_panel_kernel(const Sell &s, size_t lds, const SellView &a, const T *avals, const SellView &b,
                                const T *bvals, const T *X, T *Y, int m, int chunk, int accum, hipStream_t st) {
  auto kern = k_sell_apply<T, LD, CODED, NB>;
  static bool attr_done = false;
  if (!attr_done) {
    (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr_done = true;
  }
  hipLaunchKernelGGL(kern, dim3(s.nblk), dim3(s.nwaves * 64), lds, st, a, avals, b, bvals, X, Y, m, chunk, accum);
}

// one product: the proton-spectator part (when split) writes every row of Y; then one
// panel launch adds the rest: the coded kernel (value pairs, then coded pairs) when any
// pair is coded, else the value panels alone
template <typename T, int LD>
static int launch_sell(const Sell &s, const T *X, T *Y, int m, int chunk, int nb, hipStream_t st) {
  const Panels &V = s.P[0], &C = s.P[1];
  const bool coded = C.K > 0;
  const size_t lds = coded ? panels_lds<T, LD>(s, C, nb) : panels_lds<T, LD>(s, V, nb);
  if (lds > 160 * 1024) return COSMO_EUNSUP;  // tiles and table do not fit one CU's LDS
  int accum = 0;
  if (s.split) {
    auto pkern = k_pm_apply<T, LD>;
    static bool attr_done = false;
    if (!attr_done) {
      (void)hipFuncSetAttribute((const void *)pkern, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      attr_done = true;
    }
    PmView pv{s.pm_items, s.pm_rows, s.pm_base, s.pm_cols};
    const int vpr = (int)(LD * sizeof(T) / 16), nthr = s.pm_nwaves * 64;
    const int pieces = (s.pm_maxC * vpr + nthr - 1) / nthr * nthr;
    hipLaunchKernelGGL(pkern, dim3(s.pm_units), dim3(nthr), pm_lds_bytes(s, LD * sizeof(T)), st, pv,
                       pick<T>(s.pm_v32, s.pm_v64), X, Y, m, pieces);
    HIPCHK(hipGetLastError());
    accum = 1;
  }
  const SellView vv = panel_view(s, V, nullptr);
  const T *vvals = pick<T>(V.v32, V.v64);
  if (coded) {
    const SellView cv = panel_view(s, C, sizeof(T) == 4 ? C.pk32 : C.pk64);
    const T *dict = pick<T>(C.dict32, C.dict64);
    if (nb == 1) launch_panel_kernel<T, LD, true, 1>(s, lds, cv, dict, vv, vvals, X, Y, m, chunk, accum, st);
    else launch_panel_kernel<T, LD, true, 2>(s, lds, cv, dict, vv, vvals, X, Y, m, chunk, accum, st);
  } else {
    if (nb == 1) launch_panel_kernel<T, LD, false, 1>(s, lds, vv, vvals, vv, vvals, X, Y, m, chunk, accum, st);
    else launch_panel_kernel<T, LD, false, 2>(s, lds, vv, vvals, vv, vvals, X, Y, m, chunk, accum, st);
  }
  HIPCHK(hipGetLastError());
  return COSMO_OK;
}

// One product with the XCD chunk and tile buffers of this (precision, pitch): on the first
// call every chunk of SELL_CHUNKS with one and with two tile buffers is timed (twice each)
// and the fastest kept.  One buffer leaves room in LDS for more workgroups per CU, two
// overlap each tile's load with the previous tile's products.  Neither changes a row's
// summation, so every launch writes the same Y.
static constexpr int SELL_CHUNKS[] = {1, 2, 4, 8, 16, 32};
template <typename T, int LD>
static int launch_sell_tuned(Sell &s, const T *X, T *Y, int m, hipStream_t st) {
  const int nb_only = (g_sell_nb == 1 || g_sell_nb == 2) ? g_sell_nb : 0;
  if (g_sell_chunk > 0 && nb_only) return launch_sell<T, LD>(s, X, Y, m, g_sell_chunk, nb_only, st);
  const int key = (int)sizeof(T) * 4096 + LD;
  int chunk, nb;
  {
    std::lock_guard<std::mutex> lk(g_sell_mu);
    auto it = s.chunk.find(key);
    if (it != s.chunk.end()) {
      chunk = it->second / 4;
      nb = it->second % 4;
    } else {
      hipEvent_t e0, e1;
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
      float best_ms = 1e30f;
      chunk = 0;
      nb = 2;
      for (int b : {1, 2}) {
        if (nb_only && b != nb_only) continue;
        for (int c : SELL_CHUNKS) {
          if (g_sell_chunk > 0 && c != g_sell_chunk) continue;
          for (int rep = 0; rep < 2; ++rep) {
            HIPCHK(hipEventRecord(e0, st));
            const int rc = launch_sell<T, LD>(s, X, Y, m, c, b, st);
            if (rc == COSMO_EUNSUP) break;  // this buffer count does not fit LDS
            if (rc) {
              (void)hipEventDestroy(e0);
              (void)hipEventDestroy(e1);
              return rc;
            }
            HIPCHK(hipEventRecord(e1, st));
            HIPCHK(hipEventSynchronize(e1));
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, e0, e1));
            if (t < best_ms) best_ms = t, chunk = c, nb = b;
          }
        }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      if (chunk == 0) return COSMO_EUNSUP;  // nothing fits LDS: the caller falls back
      s.chunk.emplace(key, chunk * 4 + nb);
      if (g_sell_stats)
        fprintf(stderr, "[sell] XCD chunk %d, %d tile buffer(s) for %d-byte values at pitch %d (%.3f ms)\n", chunk, nb,
                (int)sizeof(T), LD, best_ms);
      // the fastest configuration writes Y last, from the same X
      return launch_sell<T, LD>(s, X, Y, m, chunk, nb, st);
    }
  }
  return launch_sell<T, LD>(s, X, Y, m, chunk, nb, st);
}

// Y = H X through the tiled layout when it applies (stored operator, row pitch ld with
// 16-byte X rows up to 16 elements); COSMO_EUNSUP otherwise (the caller falls back).
int sell_apply(Op &op, const void *X, void *Y, int m, int ld, int prec, hipStream_t st) {
  if (op.is_virtual || !op.csr.rowptr || op.layout == 1) return COSMO_EUNSUP;
  // automatic layout: the tiled layout costs about 0.6 us per row to build (0.8 s at
  // n = 1.4 M) and pays back about 2 ns per row and product at k = 8 from n of a few
  // hundred thousand, some 400 products; below 262,144 rows it is no faster than the row
  // kernel.  So it is built only for large operators that have already made SELL_AFTER
  // products (one Davidson solve makes a few hundred): one-off solves stay on the row
  // kernel, operators in long or repeated use switch.
  if (op.layout == 0) {
    if (g_sell_off || op.n < SELL_MIN_N) return COSMO_EUNSUP;
    if (!op.sell && __atomic_add_fetch(&op.row_products, 1, __ATOMIC_RELAXED) < SELL_AFTER) return COSMO_EUNSUP;
  }
  const bool f32 = prec == COSMO_PREC_FP32;
  if (!(f32 || prec == COSMO_PREC_FP64)) return COSMO_EUNSUP;
  if (f32 && !(ld == 4 || ld == 8 || ld == 16)) return COSMO_EUNSUP;
  if (!f32 && !(ld == 2 || ld == 4 || ld == 8 || ld == 16)) return COSMO_EUNSUP;
  if (f32 && !op.csr.v32) return COSMO_EUNSUP;
  {
    std::lock_guard<std::mutex> lk(g_sell_mu);
    if (!op.sell) {
      const int rc = build_sell(op, st);
      if (rc) return rc;
    }
    const int rc = sell_values(op, f32, st);
    if (rc) return rc;
  }
  Sell &s = *(Sell *)op.sell;
  if (f32) {
    const float *x = (const float *)X;
    float *y = (float *)Y;
    if (ld == 4) return launch_sell_tuned<float, 4>(s, x, y, m, st);
    if (ld == 8) return launch_sell_tuned<float, 8>(s, x, y, m, st);
    return launch_sell_tuned<float, 16>(s, x, y, m, st);
  }
  const double *x = (const double *)X;
  double *y = (double *)Y;
  if (ld == 2) return launch_sell_tuned<double, 2>(s, x, y, m, st);
  if (ld == 4) return launch_sell_tuned<double, 4>(s, x, y, m, st);
  if (ld == 8) return launch_sell_tuned<double, 8>(s, x, y, m, st);
  return launch_sell_tuned<double, 16>(s, x, y, m, st);
}

}  // namespace cosmo
