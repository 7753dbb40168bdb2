// op.hip -- C ABI of the device-resident Hamiltonian operator (cosmo_op_*).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "op.h"

namespace cosmo {
int virtual_diag(const DevTables &T, int64_t n, double *d_diag, hipStream_t st);

static void free_op(Op *op) {
  if (!op) return;
  hipSetDevice(op->device);
  CsrDev &M = op->csr;
  dfree(M.rowptr);
  dfree(M.cols);
  dfree(M.v64);
  dfree(M.v32);
  dfree(M.diag);
  free_tables(op->tab);
  free_factored(*op);
  free_sell(*op);
  dfree(op->sx);
  dfree(op->sy);
  dfree(op->sx32);
  dfree(op->sy32);
  if (op->e0) hipEventDestroy(op->e0);
  if (op->e1) hipEventDestroy(op->e1);
}

static int fetch_diag(Op &op, hipStream_t st) {
  op.hdiag.assign(op.n, 0.0);
  if (op.n == 0) return COSMO_OK;
  HIPCHK(hipMemcpyAsync(op.hdiag.data(), op.csr.diag, sizeof(double) * op.n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return COSMO_OK;
}

static int ensure_staging(Op &op, size_t elems) {
  if (elems <= op.s_elems) return COSMO_OK;
  dfree(op.sx);
  dfree(op.sy);
  dfree(op.sx32);
  dfree(op.sy32);
  op.sx = op.sy = op.sx32 = op.sy32 = nullptr;
  HIPCHK(dmalloc(&op.sx, elems * sizeof(double)));
  HIPCHK(dmalloc(&op.sy, elems * sizeof(double)));
  HIPCHK(dmalloc(&op.sx32, elems * sizeof(float)));
  HIPCHK(dmalloc(&op.sy32, elems * sizeof(float)));
  op.s_elems = elems;
  return COSMO_OK;
}

// Upper-triangle rows (cols ascending per row) -> full symmetric rows on the device.  The
// GPU assembles them (upper.hip); this host transposition remains for matrices whose
// columns are too long for its LDS sort.
static int upload_upper(Op &op, int64_t n, const int64_t *rowptr, const int64_t *cols, const double *vals,
                        int flags, hipStream_t st) {
  {
    const int rc = upper_to_full_dev(op, n, rowptr, cols, vals, flags, st);
    if (rc != COSMO_EUNSUP) return rc;
  }
  const bool keep = flags & COSMO_OP_KEEP_SMALL;
  std::vector<int64_t> cnt(n + 1, 0);
  for (int64_t r = 0; r < n; ++r)
    for (int64_t e = rowptr[r]; e < rowptr[r + 1]; ++e) {
      int64_t c = cols[e];
      if (c < r || c >= n) return fail(COSMO_EINVAL, "upper rows must hold columns in [row, n)");
      if (!keep && c != r && std::fabs(vals[e]) < 1e-14) continue;
      cnt[r + 1]++;
      if (c != r) cnt[c + 1]++;
    }
  for (int64_t r = 0; r < n; ++r) cnt[r + 1] += cnt[r];
  const int64_t nnz = cnt[n];
  std::vector<uint32_t> hc(std::max<int64_t>(nnz, 1));
  std::vector<double> hv(std::max<int64_t>(nnz, 1));
  std::vector<double> hd(n, 0.0);
  std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
  // lower entries of row c come from rows r < c, visited in ascending r: they land first,
  // in ascending column order; then the row's own upper entries, also ascending
  for (int64_t r = 0; r < n; ++r)
    for (int64_t e = rowptr[r]; e < rowptr[r + 1]; ++e) {
      int64_t c = cols[e];
      if (!keep && c != r && std::fabs(vals[e]) < 1e-14) continue;
      if (c == r) hd[r] = vals[e];
      if (c != r) {
        hc[fill[c]] = (uint32_t)r;
        hv[fill[c]++] = vals[e];
      }
    }
  // the upper parts go after the lower parts of each row
  std::vector<int64_t> up(n, 0);
  for (int64_t r = 0; r < n; ++r) up[r] = fill[r];
  for (int64_t r = 0; r < n; ++r)
    for (int64_t e = rowptr[r]; e < rowptr[r + 1]; ++e) {
      int64_t c = cols[e];
      if (!keep && c != r && std::fabs(vals[e]) < 1e-14) continue;
      hc[up[r]] = (uint32_t)c;
      hv[up[r]++] = vals[e];
    }
  CsrDev &M = op.csr;
  M.n = n;
  M.nnz = nnz;
  int64_t nup = 0;
  for (int64_t r = 0; r < n; ++r) nup += up[r] - fill[r];
  M.nnz_upper = nup;
  HIPCHK(dmalloc(&M.rowptr, sizeof(int64_t) * (n + 1)));
  HIPCHK(dmalloc(&M.cols, sizeof(uint32_t) * hc.size()));
  HIPCHK(dmalloc(&M.v64, sizeof(double) * hv.size()));
  HIPCHK(dmalloc(&M.diag, sizeof(double) * std::max<int64_t>(n, 1)));
  HIPCHK(hipMemcpyAsync(M.rowptr, cnt.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(M.cols, hc.data(), sizeof(uint32_t) * hc.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(M.v64, hv.data(), sizeof(double) * hv.size(), hipMemcpyHostToDevice, st));
  if (n) HIPCHK(hipMemcpyAsync(M.diag, hd.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  if (!(flags & COSMO_OP_NO_FP32)) {
    std::vector<float> hf(hv.size());
    for (size_t i = 0; i < hv.size(); ++i) hf[i] = (float)hv[i];
    HIPCHK(dmalloc(&M.v32, sizeof(float) * hf.size()));
    HIPCHK(hipMemcpyAsync(M.v32, hf.data(), sizeof(float) * hf.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  HIPCHK(hipStreamSynchronize(st));
  op.hdiag = hd;
  return COSMO_OK;
}

// Upper rows from the blocks of a symmetric CSB (upper triangle), fetched block-row major
// in ascending block column by get(I, J, nnz, rows, cols, vals) (0 = ok, else the error to
// return).  Each row's entries are sorted by column (blocks from PPMatrix2CSBBlk are
// already in (row, col) order); a repeated column, an index outside its block and an
// entry below the diagonal are errors.
template <class Get>
static int csb_upper_rows(int64_t n, int64_t br, int64_t bc, int64_t nbr, int64_t nbc, Get get,
                          std::vector<int64_t> &rp, std::vector<int64_t> &cols, std::vector<double> &vals,
                          std::string &why) {
  rp.assign(n + 1, 0);
  cols.clear();
  vals.clear();
  std::vector<int64_t> R, C, ord;
  std::vector<double> V;
  for (int64_t I = 0; I < nbr; ++I) {
    R.clear();
    C.clear();
    V.clear();
    for (int64_t J = 0; J < nbc; ++J) {
      uint32_t nnz = 0;
      const uint16_t *rr = nullptr, *cc = nullptr;
      const double *vv = nullptr;
      const int rc = get(I, J, nnz, rr, cc, vv);
      if (rc) return rc;
      for (uint32_t e = 0; e < nnz; ++e) {
        const int64_t r = I * br + rr[e], c = J * bc + cc[e];
        if (rr[e] >= br || cc[e] >= bc || r >= n || c >= n) {
          why = "index outside its block";
          return COSMO_EINVAL;
        }
        if (c < r) {
          why = "entry below the diagonal";
          return COSMO_EINVAL;
        }
        R.push_back(r);
        C.push_back(c);
        V.push_back(vv[e]);
      }
    }
    const int64_t r0 = I * br, r1 = std::min<int64_t>(n, r0 + br);
    std::vector<int64_t> pos(r1 - r0 + 1, 0);
    for (int64_t r : R) pos[r - r0 + 1]++;
    for (int64_t i = 0; i < r1 - r0; ++i) pos[i + 1] += pos[i];
    ord.resize(R.size());
    {
      std::vector<int64_t> at(pos.begin(), pos.end() - 1);
      for (size_t e = 0; e < R.size(); ++e) ord[at[R[e] - r0]++] = (int64_t)e;
    }
    const int64_t base = (int64_t)cols.size();
    cols.resize(base + R.size());
    vals.resize(base + R.size());
    for (int64_t i = 0; i < r1 - r0; ++i) {
      auto b0 = ord.begin() + pos[i], b1 = ord.begin() + pos[i + 1];
      std::stable_sort(b0, b1, [&](int64_t x, int64_t y) { return C[x] < C[y]; });
      for (int64_t k = pos[i]; k < pos[i + 1]; ++k) {
        cols[base + k] = C[ord[k]];
        vals[base + k] = V[ord[k]];
        if (k > pos[i] && cols[base + k] == cols[base + k - 1]) {
          why = "a column repeated in a row";
          return COSMO_EINVAL;
        }
      }
      rp[r0 + i + 1] = base + pos[i + 1];
    }
  }
  return COSMO_OK;
}

}  // namespace cosmo

using namespace cosmo;

extern "C" {

int cosmo_set_device(int device) { return set_device_impl(device); }

int cosmo_op_create(cosmo_system s, int flags, cosmo_op *out) {
  if (!s || !out) return fail(COSMO_EINVAL, "cosmo_op_create: null argument");
  if (s->nv + s->zv > 24) return fail(COSMO_EUNSUP, "cosmo_op_create: more than 24 valence particles");
  if (s->dim >= (int64_t)0xffffffffLL) return fail(COSMO_EUNSUP, "cosmo_op_create: dimension beyond 2^32-1");
  if (s->dim <= 0)
    return fail(COSMO_EINVAL, "cosmo_op_create: empty basis (no many-body state has the system's M and parity)");
  auto *op = new cosmo_op_s();
  op->device = current_device();
  op->n = s->dim;
  op->flags = flags;
  op->is_virtual = (flags & (COSMO_OP_VIRTUAL | COSMO_OP_FACTORED)) != 0;
  hipStream_t st = lib_stream();
  op->home = st;
  auto t0 = std::chrono::steady_clock::now();
  int rc = upload_tables(*s, op->tab, st);
  if (!rc && (flags & COSMO_OP_FACTORED)) rc = build_factored(*op, *s);
  if (!rc) {
    if (op->is_virtual) {
      op->csr.n = op->n;
      if (dmalloc(&op->csr.diag, sizeof(double) * std::max<int64_t>(op->n, 1)) != hipSuccess)
        rc = fail(COSMO_ENOMEM, "diagonal allocation failed");
      else
        rc = virtual_diag(op->tab.t, op->n, op->csr.diag, st);
    } else {
      rc = build_stored(*op, *s, flags, st);
      op->part = s->noff;  // row blocks of the tiled product
      if (!s->truncated()) {
        op->part_pbase.resize(s->nsector.size());
        for (size_t b = 0; b < s->nsector.size(); ++b) op->part_pbase[b] = s->psec_start[s->nsector[b]];
      }
    }
  }
  if (!rc) rc = fetch_diag(*op, st);
  if (rc) {
    free_op(op);
    delete op;
    return rc;
  }
  op->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  hipEventCreate(&op->e0);
  hipEventCreate(&op->e1);
  *out = op;
  return COSMO_OK;
}

int cosmo_op_create_upper(int64_t n, const int64_t *rowptr, const int64_t *cols, const double *vals, int flags,
                          cosmo_op *out) {
  if (!out || n < 0 || (n && (!rowptr || !cols || !vals))) return fail(COSMO_EINVAL, "cosmo_op_create_upper: bad args");
  if (n >= (int64_t)0xffffffffLL) return fail(COSMO_EUNSUP, "dimension beyond 2^32-1");
  // the row offsets bound every read of cols / vals, on the host and in the kernels that
  // assemble the full rows: checked before anything is allocated or uploaded
  if (n && rowptr[0] != 0) return fail(COSMO_EINVAL, "cosmo_op_create_upper: rowptr[0] must be 0");
  for (int64_t r = 0; r < n; ++r)
    if (rowptr[r + 1] < rowptr[r]) return fail(COSMO_EINVAL, "cosmo_op_create_upper: rowptr decreases at row " +
                                                              std::to_string(r));
  auto *op = new cosmo_op_s();
  op->device = current_device();
  op->n = n;
  op->flags = flags;
  hipStream_t st = lib_stream();
  op->home = st;
  auto t0 = std::chrono::steady_clock::now();
  int rc = upload_upper(*op, n, rowptr, cols, vals, flags, st);
  if (rc) {
    free_op(op);
    delete op;
    return rc;
  }
  op->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  hipEventCreate(&op->e0);
  hipEventCreate(&op->e1);
  *out = op;
  return COSMO_OK;
}

int cosmo_op_read_hh(const char *path, int flags, cosmo_op *out) {
  if (!path || !out) return fail(COSMO_EINVAL, "cosmo_op_read_hh: null argument");
  FILE *f = fopen(path, "rb");
  if (!f) return fail(COSMO_EIO, std::string("File ") + path + " is not found");
  std::vector<char> iobuf((size_t)16 << 20);
  setvbuf(f, iobuf.data(), _IOFBF, iobuf.size());
  auto bad = [&](const char *m) {
    fclose(f);
    return fail(COSMO_EIO, std::string(path) + ": " + m);
  };
  if (fseeko(f, 0, SEEK_END) != 0) return bad("cannot size the file");
  const int64_t fsize = ftello(f);
  if (fseeko(f, 0, SEEK_SET) != 0) return bad("cannot rewind the file");
  uint64_t n = 0;
  if (fsize < 8 || fread(&n, 8, 1, f) != 1) return bad("truncated .HH header");
  if (n > (uint64_t)(fsize - 8) / 4) return bad("truncated .HH file (rows)");
  // [u64 n] then per row [u32 nz][nz u64 col][nz f64]: the entry count follows from the size
  const int64_t rest = fsize - 8 - 4 * (int64_t)n;
  if (rest % 16) return bad("truncated or malformed .HH file");
  const int64_t E = rest / 16;
  std::vector<int64_t> rp(n + 1, 0), cols(E);
  std::vector<double> vals(E);
  int64_t off = 0;
  for (uint64_t r = 0; r < n; ++r) {
    uint32_t nz = 0;
    if (fread(&nz, 4, 1, f) != 1) return bad("truncated .HH row header");
    if (off + (int64_t)nz > E) return bad("row lengths exceed the file");
    static_assert(sizeof(int64_t) == sizeof(uint64_t), "column words");
    if (fread(cols.data() + off, 8, nz, f) != nz || fread(vals.data() + off, 8, nz, f) != nz)
      return bad("truncated .HH row");
    off += nz;
    rp[r + 1] = off;
  }
  fclose(f);
  if (off != E) return fail(COSMO_EIO, std::string(path) + ": trailing bytes after the last row");
  return cosmo_op_create_upper((int64_t)n, rp.data(), cols.data(), vals.data(), flags | COSMO_OP_KEEP_SMALL, out);
}

int cosmo_op_destroy(cosmo_op op) {
  if (!op) return COSMO_OK;
  hipSetDevice(op->device);
  // stream-ordered frees follow the work of this thread's stream; anything the operator
  // did on other streams must be finished first
  if (op->foreign || op->home != lib_stream()) hipDeviceSynchronize();
  free_op(op);
  delete op;
  return COSMO_OK;
}

int64_t cosmo_op_size(cosmo_op op) { return op ? op->n : COSMO_EINVAL; }

int cosmo_op_nnz(cosmo_op op, int64_t *nnz_upper, int64_t *nnz_full) {
  if (!op) return fail(COSMO_EINVAL, "null operator");
  if (nnz_upper) *nnz_upper = op->csr.nnz_upper;
  if (nnz_full) *nnz_full = op->csr.nnz;
  return COSMO_OK;
}

int64_t cosmo_op_bytes(cosmo_op op) {
  if (!op) return COSMO_EINVAL;
  const CsrDev &M = op->csr;
  int64_t b = (int64_t)op->tab.bytes + 8 * M.n;
  if (!op->is_virtual) b += 8 * (M.n + 1) + M.nnz * (4 + 8 + (M.v32 ? 4 : 0));
  return b + (int64_t)factored_bytes(*op) + (int64_t)sell_bytes(*op);
}

int cosmo_op_diagonal(cosmo_op op, double *d) {
  if (!op || !d) return fail(COSMO_EINVAL, "null argument");
  std::copy(op->hdiag.begin(), op->hdiag.end(), d);
  return COSMO_OK;
}

int cosmo_op_apply_device(cosmo_op op, const void *X, void *Y, int m, int ld, int precision, void *stream) {
  if (!op || (!X && m) || (!Y && m) || m < 0 || ld < m) return fail(COSMO_EINVAL, "cosmo_op_apply_device: bad args");
  hipSetDevice(op->device);
  hipStream_t st = stream ? (hipStream_t)stream : lib_stream();
  note_stream(*op, st);
  hipEventRecord(op->e0, st);
  int rc = op_apply_dev(*op, X, Y, m, ld, ld, precision, st);
  hipEventRecord(op->e1, st);
  return rc;
}

int cosmo_op_apply_device_noalloc(cosmo_op op, const void *X, void *Y, int m, int ld, int precision, void *stream) {
  if (!op || (!X && m) || (!Y && m) || m < 0 || ld < m)
    return fail(COSMO_EINVAL, "cosmo_op_apply_device_noalloc: bad args");
  if (!op->fact)
    return fail(COSMO_EUNSUP, "cosmo_op_apply_device_noalloc: product-form operators (COSMO_OP_FACTORED) only");
  if (m == 0) return COSMO_OK;
  hipSetDevice(op->device);
  // the caller's stream, else the null stream: no library stream is created here
  hipStream_t st = (hipStream_t)stream;
  note_stream(*op, st);
  return fact_apply_noalloc(*op, X, Y, m, ld, precision, st);
}

int cosmo_op_apply(cosmo_op op, double *const *y, const double *const *x, int m, int precision) {
  if (!op || m < 0 || (m && (!x || !y))) return fail(COSMO_EINVAL, "cosmo_op_apply: bad args");
  if (m == 0) return COSMO_OK;
  hipSetDevice(op->device);
  hipStream_t st = lib_stream();
  note_stream(*op, st);
  const int64_t n = op->n;
  // the product-form and the tiled stored apply want a power-of-two row pitch of at least
  // 16 bytes
  int ld = m;
  if (op->fact || (m <= 16 && precision != COSMO_PREC_FP64_ORDERED)) {
    ld = 4;
    while (ld < m) ld <<= 1;
  }
  int rc = ensure_staging(*op, (size_t)n * ld);
  if (rc) return rc;
  // rows padded to ld on the host, so each direction is one contiguous copy
  std::vector<double> h((size_t)n * ld, 0.0);
  for (int v = 0; v < m; ++v)
    for (int64_t r = 0; r < n; ++r) h[(size_t)r * ld + v] = x[v][r];
  HIPCHK(hipMemcpyAsync(op->sx, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, st));
  if (precision == COSMO_PREC_FP32) {
    rc = cast_pack(n, m, (const double *)op->sx, ld, (float *)op->sx32, ld, st);
    if (rc) return rc;
    HIPCHK(hipEventRecord(op->e0, st));
    rc = op_apply_dev(*op, op->sx32, op->sy32, m, ld, ld, precision, st);
    HIPCHK(hipEventRecord(op->e1, st));
    if (rc) return rc;
    rc = cast_unpack(n, m, (const float *)op->sy32, ld, (double *)op->sy, ld, st);
  } else {
    HIPCHK(hipEventRecord(op->e0, st));
    rc = op_apply_dev(*op, op->sx, op->sy, m, ld, ld, precision, st);
    HIPCHK(hipEventRecord(op->e1, st));
  }
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(h.data(), op->sy, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  float ms = 0;
  hipEventElapsedTime(&ms, op->e0, op->e1);
  op->last_apply_ms = ms;
  for (int v = 0; v < m; ++v)
    for (int64_t r = 0; r < n; ++r) y[v][r] = h[(size_t)r * ld + v];
  return COSMO_OK;
}

int cosmo_op_upper_rows(cosmo_op op, int64_t r0, int64_t r1, int64_t *counts, const int64_t *rowptr, int64_t *cols,
                        double *vals) {
  if (!op || !counts || r0 < 0 || r1 > op->n || r1 < r0) return fail(COSMO_EINVAL, "cosmo_op_upper_rows: bad args");
  if (op->is_virtual) return fail(COSMO_EUNSUP, "virtual operator has no stored rows");
  hipSetDevice(op->device);
  hipStream_t st = lib_stream();
  const CsrDev &M = op->csr;
  std::vector<int64_t> rp(r1 - r0 + 1);
  if (r1 == r0) return COSMO_OK;
  HIPCHK(hipMemcpyAsync(rp.data(), M.rowptr + r0, sizeof(int64_t) * rp.size(), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const int64_t e0 = rp[0], e1 = rp.back();
  std::vector<uint32_t> hc(e1 - e0);
  std::vector<double> hv(cols ? e1 - e0 : 0);  // counts alone need no values
  if (e1 > e0) {
    HIPCHK(hipMemcpyAsync(hc.data(), M.cols + e0, sizeof(uint32_t) * hc.size(), hipMemcpyDeviceToHost, st));
    if (cols) HIPCHK(hipMemcpyAsync(hv.data(), M.v64 + e0, sizeof(double) * hv.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  for (int64_t r = r0; r < r1; ++r) {
    int64_t c = 0;
    int64_t o = rowptr ? rowptr[r - r0] : 0;
    for (int64_t e = rp[r - r0]; e < rp[r - r0 + 1]; ++e) {
      uint32_t col = hc[e - e0];
      if ((int64_t)col < r) continue;
      if (cols) {
        cols[o + c] = col;
        vals[o + c] = hv[e - e0];
      }
      ++c;
    }
    counts[r - r0] = c;
  }
  return COSMO_OK;
}

int cosmo_op_write_hh(cosmo_op op, const char *path) {
  if (!op || !path) return fail(COSMO_EINVAL, "null argument");
  if (op->is_virtual) return fail(COSMO_EUNSUP, "virtual operator has no stored rows");
  hipSetDevice(op->device);
  hipStream_t st = lib_stream();
  const CsrDev &M = op->csr;
  FILE *f = fopen(path, "wb");
  if (!f) return fail(COSMO_EIO, std::string("cannot write ") + path);
  auto bad = [&](const std::string &m) {
    fclose(f);
    return fail(COSMO_EIO, m);
  };
  uint64_t n = (uint64_t)op->n;
  fwrite(&n, 8, 1, f);
  // per chunk of rows: one copy of the stored rows, their upper parts (the stored rows
  // are ascending, so the upper part is the tail from the diagonal on) laid out in one
  // buffer and written at once
  const int64_t chunk = 1 << 16;
  std::vector<int64_t> rp;
  std::vector<uint32_t> hc;
  std::vector<double> hv;
  std::vector<char> buf;
  for (int64_t r0 = 0; r0 < op->n; r0 += chunk) {
    const int64_t r1 = std::min<int64_t>(op->n, r0 + chunk);
    rp.resize(r1 - r0 + 1);
    if (hipMemcpyAsync(rp.data(), M.rowptr + r0, sizeof(int64_t) * rp.size(), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return bad("reading the stored rows failed");
    const int64_t e0 = rp[0], e1 = rp.back();
    hc.resize(e1 - e0);
    hv.resize(e1 - e0);
    if (e1 > e0 &&
        (hipMemcpyAsync(hc.data(), M.cols + e0, sizeof(uint32_t) * hc.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
         hipMemcpyAsync(hv.data(), M.v64 + e0, sizeof(double) * hv.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
         hipStreamSynchronize(st) != hipSuccess))
      return bad("reading the stored rows failed");
    buf.resize((size_t)(4 * (r1 - r0) + 16 * (e1 - e0)));
    char *w = buf.data();
    for (int64_t r = r0; r < r1; ++r) {
      int64_t a = rp[r - r0] - e0;
      const int64_t b = rp[r - r0 + 1] - e0;
      while (a < b && (int64_t)hc[a] < r) ++a;
      const uint32_t nz = (uint32_t)(b - a);
      std::memcpy(w, &nz, 4);
      w += 4;
      for (int64_t e = a; e < b; ++e) {
        const uint64_t c = hc[e];
        std::memcpy(w, &c, 8);
        w += 8;
      }
      std::memcpy(w, hv.data() + a, 8 * (size_t)nz);
      w += 8 * (size_t)nz;
    }
    const size_t len = (size_t)(w - buf.data());
    if (fwrite(buf.data(), 1, len, f) != len) return bad(std::string("write failed: ") + path);
  }
  if (fclose(f)) return fail(COSMO_EIO, std::string("write failed: ") + path);
  return COSMO_OK;
}

// CSB file: the layout of SparseMatrixCSB::Write (include/tav/csbmatrix.cxx:553-576) and
// CSBBlock::Write (:246-254).  Header [bool symmetric][u32 n][u16 BR][u16 BC][u16 nBR]
// [u16 nBC], then every block, block-row major: [bool isDiagonal][u32 nnz][nnz u16 col]
// [nnz u16 row][nnz f64].  Symmetric files hold the upper triangle: block (I, J), J >= I,
// has the entries of rows [I*BR, (I+1)*BR) and columns [J*BC, (J+1)*BC) with col >= row,
// in ascending (row, col) order (PPMatrix2CSBBlk, include/sm/PPMatrix.cxx:571-656).
int cosmo_op_write_csb(cosmo_op op, const char *path, int block) {
  if (!op || !path || block < 0 || block > 65535) return fail(COSMO_EINVAL, "cosmo_op_write_csb: bad args");
  if (op->is_virtual) return fail(COSMO_EUNSUP, "virtual operator has no stored rows");
  const int64_t n = op->n;
  if (n > (int64_t)0xffffffffLL) return fail(COSMO_EUNSUP, "CSB file: dimension beyond 2^32-1");
  // default block size as DEXHVcsb picks it (include/xsm/DEXHVcsb.cxx:63)
  const int64_t br = block ? block : (int64_t)(std::sqrt(1. * (double)n) + 1);
  if (br > 65535) return fail(COSMO_EUNSUP, "CSB file: block size beyond 65535");
  const int64_t nb = (n + br - 1) / br;
  if (nb > 65535) return fail(COSMO_EUNSUP, "CSB file: more than 65535 block rows");
  FILE *f = fopen(path, "wb");
  if (!f) return fail(COSMO_EIO, std::string("cannot write ") + path);
  const bool sym = true;
  const uint32_t n32 = (uint32_t)n;
  const uint16_t b16 = (uint16_t)br, nb16 = (uint16_t)nb;
  fwrite(&sym, sizeof(bool), 1, f);
  fwrite(&n32, 4, 1, f);
  fwrite(&b16, 2, 1, f);
  fwrite(&b16, 2, 1, f);
  fwrite(&nb16, 2, 1, f);
  fwrite(&nb16, 2, 1, f);
  hipSetDevice(op->device);
  hipStream_t st = lib_stream();
  const CsrDev &M = op->csr;
  auto bad = [&](const std::string &m) {
    fclose(f);
    return fail(COSMO_EIO, m);
  };
  // per block row: one copy of its stored rows, their upper parts (the tail of each
  // ascending row from the diagonal on) grouped by block column, and every block of the
  // block row laid out in one buffer and written at once
  std::vector<int64_t> rp, start(nb + 1), at(nb);
  std::vector<uint32_t> hc;
  std::vector<double> hv;
  std::vector<int64_t> first;  // first upper entry of each row of the block row
  std::vector<uint16_t> bc, brow;
  std::vector<double> bv;
  std::vector<char> buf;
  for (int64_t I = 0; I < nb; ++I) {
    const int64_t r0 = I * br, r1 = std::min(n, r0 + br);
    rp.resize(r1 - r0 + 1);
    if (hipMemcpyAsync(rp.data(), M.rowptr + r0, sizeof(int64_t) * rp.size(), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return bad("reading the stored rows failed");
    const int64_t e0 = rp[0], e1 = rp.back();
    hc.resize(e1 - e0);
    hv.resize(e1 - e0);
    if (e1 > e0 &&
        (hipMemcpyAsync(hc.data(), M.cols + e0, sizeof(uint32_t) * hc.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
         hipMemcpyAsync(hv.data(), M.v64 + e0, sizeof(double) * hv.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
         hipStreamSynchronize(st) != hipSuccess))
      return bad("reading the stored rows failed");
    first.resize(r1 - r0);
    std::fill(start.begin(), start.end(), 0);
    int64_t nup = 0;
    for (int64_t i = 0; i < r1 - r0; ++i) {
      int64_t a = rp[i] - e0;
      const int64_t b = rp[i + 1] - e0;
      while (a < b && (int64_t)hc[a] < r0 + i) ++a;
      first[i] = a;
      for (int64_t e = a; e < b; ++e) start[hc[e] / br + 1]++;
      nup += b - a;
    }
    for (int64_t J = 0; J < nb; ++J) start[J + 1] += start[J];
    bc.resize(nup);
    brow.resize(nup);
    bv.resize(nup);
    std::copy(start.begin(), start.end() - 1, at.begin());
    for (int64_t i = 0; i < r1 - r0; ++i)
      for (int64_t e = first[i]; e < rp[i + 1] - e0; ++e) {
        const int64_t J = hc[e] / br, q = at[J]++;
        bc[q] = (uint16_t)(hc[e] - J * br);
        brow[q] = (uint16_t)i;
        bv[q] = hv[e];
      }
    buf.resize((size_t)(nb * (sizeof(bool) + 4) + nup * 12));
    char *w = buf.data();
    for (int64_t J = 0; J < nb; ++J) {
      const bool diag = J == I;
      const uint32_t nnz = (uint32_t)(start[J + 1] - start[J]);
      std::memcpy(w, &diag, sizeof(bool));
      w += sizeof(bool);
      std::memcpy(w, &nnz, 4);
      w += 4;
      std::memcpy(w, bc.data() + start[J], 2 * (size_t)nnz);
      w += 2 * (size_t)nnz;
      std::memcpy(w, brow.data() + start[J], 2 * (size_t)nnz);
      w += 2 * (size_t)nnz;
      std::memcpy(w, bv.data() + start[J], 8 * (size_t)nnz);
      w += 8 * (size_t)nnz;
    }
    const size_t len = (size_t)(w - buf.data());
    if (fwrite(buf.data(), 1, len, f) != len) return bad(std::string("write failed: ") + path);
  }
  if (fclose(f)) return fail(COSMO_EIO, std::string("write failed: ") + path);
  return COSMO_OK;
}

// Symmetric CSB file -> stored operator (SparseMatrixCSB::Read, csbmatrix.cxx:524-552).
// A diagonal entry the file lacks (dropped below 1e-14 by the builder) reads as 0.
int cosmo_op_read_csb(const char *path, int flags, cosmo_op *out) {
  if (!path || !out) return fail(COSMO_EINVAL, "cosmo_op_read_csb: null argument");
  FILE *f = fopen(path, "rb");
  if (!f) return fail(COSMO_EIO, std::string("File ") + path + " is not found");
  bool sym = false;
  uint32_t n32 = 0;
  uint16_t br = 0, bcl = 0, nbr = 0, nbc = 0;
  auto bad = [&](const std::string &m) {
    fclose(f);
    return fail(COSMO_EIO, std::string("CSB file ") + path + ": " + m);
  };
  if (fread(&sym, sizeof(bool), 1, f) != 1 || fread(&n32, 4, 1, f) != 1 || fread(&br, 2, 1, f) != 1 ||
      fread(&bcl, 2, 1, f) != 1 || fread(&nbr, 2, 1, f) != 1 || fread(&nbc, 2, 1, f) != 1)
    return bad("truncated header");
  const int64_t n = n32;
  if (!sym) {
    fclose(f);
    return fail(COSMO_EUNSUP, std::string("CSB file ") + path + ": not symmetric (H is stored as its upper triangle)");
  }
  if (br == 0 || bcl == 0 || (int64_t)nbr * br < n || (int64_t)nbc * bcl < n ||
      (n > 0 && ((int64_t)(nbr - 1) * br >= n || (int64_t)(nbc - 1) * bcl >= n)))
    return bad("block sizes do not match the dimension");
  std::vector<uint16_t> bc16, br16;
  std::vector<double> bv;
  auto get = [&](int64_t, int64_t, uint32_t &nnz, const uint16_t *&rr, const uint16_t *&cc,
                 const double *&vv) -> int {
    bool isd = false;
    if (fread(&isd, sizeof(bool), 1, f) != 1 || fread(&nnz, 4, 1, f) != 1) return COSMO_EIO;
    bc16.resize(nnz);
    br16.resize(nnz);
    bv.resize(nnz);
    if (fread(bc16.data(), 2, nnz, f) != nnz || fread(br16.data(), 2, nnz, f) != nnz ||
        fread(bv.data(), 8, nnz, f) != nnz)
      return COSMO_EIO;
    rr = br16.data();
    cc = bc16.data();
    vv = bv.data();
    return COSMO_OK;
  };
  std::vector<int64_t> rp, cols;
  std::vector<double> vals;
  std::string why;
  const int rc = csb_upper_rows(n, br, bcl, nbr, nbc, get, rp, cols, vals, why);
  if (rc) return bad(why.empty() ? "truncated block" : why);
  const int extra = fgetc(f);
  fclose(f);
  if (extra != EOF) return fail(COSMO_EIO, std::string("CSB file ") + path + ": trailing bytes");
  return cosmo_op_create_upper(n, rp.data(), cols.data(), vals.data(), flags, out);
}

// Symmetric CSB blocks held in memory (SparseMatrixCSB, csbmatrix.cxx:359-377) -> stored
// operator.  Block (i, j) is entry i * nbc + j of the arrays; upper blocks only.
int cosmo_op_create_csb(int64_t n, int br, int bc, int nbr, int nbc, const uint32_t *nnz,
                        const uint16_t *const *rows, const uint16_t *const *cols, const double *const *vals,
                        int flags, cosmo_op *out) {
  if (!out || n < 0 || br <= 0 || bc <= 0 || nbr < 0 || nbc < 0 || (nbr && nbc && !nnz))
    return fail(COSMO_EINVAL, "cosmo_op_create_csb: bad args");
  if ((int64_t)nbr * br < n || (int64_t)nbc * bc < n)
    return fail(COSMO_EINVAL, "cosmo_op_create_csb: blocks do not cover the dimension");
  auto get = [&](int64_t I, int64_t J, uint32_t &k, const uint16_t *&rr, const uint16_t *&cc,
                 const double *&vv) -> int {
    const int64_t b = I * nbc + J;
    k = nnz[b];
    if (k && (!rows || !cols || !vals || !rows[b] || !cols[b] || !vals[b])) return COSMO_EINVAL;
    rr = k ? rows[b] : nullptr;
    cc = k ? cols[b] : nullptr;
    vv = k ? vals[b] : nullptr;
    return COSMO_OK;
  };
  std::vector<int64_t> rp, cl;
  std::vector<double> vl;
  std::string why;
  const int rc = csb_upper_rows(n, br, bc, nbr, nbc, get, rp, cl, vl, why);
  if (rc) return fail(rc, "cosmo_op_create_csb: " + (why.empty() ? std::string("missing block arrays") : why));
  return cosmo_op_create_upper(n, rp.data(), cl.data(), vl.data(), flags, out);
}

int cosmo_op_set_layout(cosmo_op op, int layout) {
  if (!op || layout < 0 || layout > 2) return fail(COSMO_EINVAL, "cosmo_op_set_layout: bad args");
  op->layout = layout;
  return COSMO_OK;
}


int cosmo_op_timings(cosmo_op op, double *build_ms, double *last_apply_ms) {
  if (!op) return fail(COSMO_EINVAL, "null operator");
  if (build_ms) *build_ms = op->build_ms;
  if (last_apply_ms) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, op->e0, op->e1) == hipSuccess) op->last_apply_ms = ms;
    *last_apply_ms = op->last_apply_ms;
  }
  return COSMO_OK;
}

}  // extern "C"
