// op.h -- the device-resident operator object behind cosmo_op.
#pragma once
#include <vector>

#include "device.h"

namespace cosmo {

// full symmetric rows (both triangles), columns ascending within each row
struct CsrDev {
  int64_t n = 0, nnz = 0, nnz_upper = 0;
  int64_t *rowptr = nullptr;
  uint32_t *cols = nullptr;
  double *v64 = nullptr;
  float *v32 = nullptr;
  double *diag = nullptr;
};

struct Op {
  int device = 0;
  int64_t n = 0;
  bool is_virtual = false;
  int flags = 0;
  CsrDev csr;
  DevTablesOwner tab;          // builder tables (kept for the virtual apply)
  std::vector<double> hdiag;   // host copy of the diagonal (MatrixBase::Diagonal)
  // staging for the host-pointer apply
  void *sx = nullptr, *sy = nullptr, *sx32 = nullptr, *sy32 = nullptr;
  size_t s_elems = 0;
  void *fact = nullptr;        // product-form tables (COSMO_OP_FACTORED)
  void *sell = nullptr;        // tiled layout of the stored rows (built on the first apply)
  std::vector<int64_t> part;   // row partition: first row of each neutron configuration
  // per neutron configuration of `part`: the index of the first configuration of its proton
  // sector in a numbering of all proton configurations, so block b's row r holds proton
  // configuration part_pbase[b] + (r - part[b]) (empty: unknown or a truncated basis)
  std::vector<int64_t> part_pbase;
  hipStream_t home = nullptr;  // library stream of the creating thread
  bool foreign = false;        // used on another stream: destroy synchronises the device
  int layout = 0;              // stored product: 0 automatic, 1 rows, 2 tiled (sell.hip)
  int64_t row_products = 0;    // automatic layout: products made with the row kernel so far
  double build_ms = 0, last_apply_ms = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
};

int build_stored(Op &op, const System &s, int flags, hipStream_t st);
// host upper rows -> full symmetric rows on the device (COSMO_EUNSUP: too long for it)
int upper_to_full_dev(Op &op, int64_t n, const int64_t *rowptr, const int64_t *cols, const double *vals,
                      int flags, hipStream_t st);
// host -> device copy through pinned double buffers
int h2d_staged(void *dst, const void *src, size_t bytes, hipStream_t st);
int build_factored(Op &op, const System &s);
int64_t factored_footprint(const System &s);  // -1: no product form
void free_factored(Op &op);
size_t factored_bytes(const Op &op);
int fact_apply(Op &op, const void *X, void *Y, int m, int ld, int prec, hipStream_t st);
int fact_apply_noalloc(Op &op, const void *X, void *Y, int m, int ld, int prec, hipStream_t st);
int op_apply_dev(Op &op, const void *X, void *Y, int m, int ldx, int ldy, int prec, hipStream_t st);
// stored H.X through the tiled layout (sell.hip); COSMO_EUNSUP where it does not apply
int sell_apply(Op &op, const void *X, void *Y, int m, int ld, int prec, hipStream_t st);
void free_sell(Op &op);
size_t sell_bytes(const Op &op);
int64_t sell_footprint_bound(int64_t n, int64_t nnz);
extern int g_sell_chunk, g_sell_off, g_sell_stats, g_sell_split, g_sell_nb;
int op_diag_dev(Op &op, double *d_diag, hipStream_t st);
int cast_pack(int64_t n, int m, const double *src, int lds, float *dst, int ldd, hipStream_t st);
int cast_unpack(int64_t n, int m, const float *src, int lds, double *dst, int ldd, hipStream_t st);
int set_device_impl(int d);
// record that op is about to be used on stream st
inline void note_stream(Op &op, hipStream_t st) {
  if (st != op.home) op.foreign = true;
}
int current_device();

}  // namespace cosmo

struct cosmo_op_s : cosmo::Op {};
