/*
 * cosmo_hip.h -- C ABI of the MI355X configuration-interaction hot path.
 *
 * Plain C types only (pointers, sizes, status codes).  Every entry point returns 0 on
 * success and a negative COSMO_E* code on failure; cosmo_last_error() then holds a
 * message.  Reference citations are file:line into the CoSMo source tree; each entry
 * point names the reference interface it replaces.
 *
 * Objects:
 *   cosmo_system  host-side model space: m-scheme orbital table, one- and two-body
 *                 m-scheme operators (ESP, VSPX) and the basis specification
 *                 (Nv, Zv, 2Jz, parity).  Replaces SingleParticleStates +
 *                 Many_Body_States + the STensor pair fed to PPMatrix.
 *   cosmo_op      device-resident symmetric Hamiltonian: stored rows in FP64 and FP32
 *                 (both triangles, built on the GPU) or the on-the-fly "virtual" form.
 *                 Replaces SparseMatrixCSB / SparseMatrix / PPMatrix behind
 *                 MatB::MatrixBase (include/kdav/matrixbase.h:8-25).
 *   cosmo_batch   many independent solves run together on one device.
 *   cosmo_server  request loop with admission by HBM footprint (continuous batching).
 */
#ifndef COSMO_HIP_H
#define COSMO_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define COSMO_OK 0
#define COSMO_EINVAL (-1)   /* bad argument                                   */
#define COSMO_ENOMEM (-2)   /* device or host allocation failed / too large   */
#define COSMO_EHIP (-3)     /* HIP runtime error                              */
#define COSMO_EIO (-4)      /* file missing, truncated or malformed           */
#define COSMO_EUNSUP (-5)   /* valid input this build does not handle          */
#define COSMO_ECONV (-6)    /* eigensolver stopped without convergence         */

typedef struct cosmo_system_s *cosmo_system;
typedef struct cosmo_op_s *cosmo_op;
typedef struct cosmo_batch_s *cosmo_batch;
typedef struct cosmo_server_s *cosmo_server;

/* ---- library ----------------------------------------------------------------------- */
const char *cosmo_last_error(void);
/* Select the HIP device used by later calls of this thread (default 0). */
int cosmo_set_device(int device);

/* ---- system setup ------------------------------------------------------------------ */
/* XML system file -> orbital table, m-scheme interaction and basis specification.
 * Replaces SM::Make(sps, <valencespace>) (include/xsm/xSingleParticleStates.cxx:26-104),
 * xReadSPInt -> ReadOxbashINT / MakeESPO / PPInteraction2sp
 * (include/xsm/xReadSPInt.cxx:24-374, include/sm/sphericalint.cxx:145-246,
 * include/sm/MakeSPO.cxx:195-202, include/sm/spherical2sp.cxx:790-814) and the system
 * attributes read by MakeMBS2 (include/xsm/xManyBodyStates.cxx:22-90).  Interaction
 * files are looked up in the XML's directory, then in $SMINTPATH
 * (include/av/iofile.cxx:62-120).  The <system> node is chosen by name when sysname is
 * non-NULL, else the last <system>. */
int cosmo_system_from_xml(const char *xml_path, const char *sysname, cosmo_system *out);

/* Raw tables.  Orbital g (0..nsps-1) has level, 2j, 2m, 2tz (+1 neutron, -1 proton),
 * n, l; neutrons must precede protons.  ESP keys (nesp x 2: annihilate, create), VSPX
 * keys (nvspx x 4: annihilate b0<b1, create b2<b3), both in reference key order.       */
int cosmo_system_create(int nsps, const int *level, const int *j2, const int *m2, const int *tz2,
                        const int *nq, const int *lq, int nv, int zv, int twoM, int parity,
                        int64_t nesp, const uint16_t *esp_keys, const double *esp_vals,
                        int64_t nvspx, const uint16_t *vspx_keys, const double *vspx_vals,
                        cosmo_system *out);
int cosmo_system_destroy(cosmo_system s);
/* M-scheme dimension n of the basis (Many_Body_States::n). */
int64_t cosmo_system_dim(cosmo_system s);
/* Number of particles N = Nv + Zv (Many_Body_States::N) and orbitals. */
int cosmo_system_info(cosmo_system s, int *nparticles, int *nsps, int *nn, int *np, int64_t *nesp,
                      int64_t *nvspx);
/* Basis in reference order, n x N uint16 occupation lists: the rows of the .mbs file
 * written by XsysmbsOMP (xmlshell/xmbs/XsysmbsOMP.cpp:403-429). */
int cosmo_system_basis(cosmo_system s, uint16_t *out);
/* Copy of the m-scheme operators (for checks and dumps). */
int cosmo_system_tensors(cosmo_system s, uint16_t *esp_keys, double *esp_vals, uint16_t *vspx_keys,
                         double *vspx_vals);
/* Orbital table copy (nsps entries each). */
int cosmo_system_orbitals(cosmo_system s, int *level, int *j2, int *m2, int *tz2);
/* Spin labels of eigenvectors (count host n-vectors, column-major): jm_jp[v] = <J-J+>,
 * tm_tp[v] = <T-T+> (either may be NULL), computed on the GPU as |J+ z|^2, |T+ z|^2.
 * Replaces PPMatrix::Sqr of the J+ / T+ operators in XSHLJT
 * (include/xsm/XSHLJT.cxx:38-142, include/sm/PPMatrix.cxx:400-411). */
int cosmo_system_labels(cosmo_system s, const double *evecs, int count, double *jm_jp, double *tm_tp);
/* 2J from <J-J+> and 2M: sqrt(4x + 2M(2M+2) + 1) - 1 rounded, or -1 when J + M is not an
 * integer to within 0.1 (include/sm/QN.cxx:40-55). */
int cosmo_twice_spin(double x, int twoM);
/* .mbs file I/O (include/sm/ManyBodyStates.cxx:199-246).  The basis is the system's: the
 * full M-scheme basis, or the one its <system> truncates by oscillator quanta (hwmax,
 * <hwrejection HWMAX HWMIN>) and occupation rejections (<rejection NMIN NMAX><orbital
 * index type/>) as XsysmbsOMP does (xmlshell/xmbs/XsysmbsOMP.cpp:212-303), in the same order.
 * cosmo_system_check_mbs fails with COSMO_EUNSUP when the file holds another basis. */
int cosmo_system_write_mbs(cosmo_system s, const char *path);
int cosmo_system_check_mbs(cosmo_system s, const char *path);

/* ---- operator ---------------------------------------------------------------------- */
#define COSMO_OP_STORED 0x0      /* rows stored in HBM (FP64 + FP32 values)          */
#define COSMO_OP_VIRTUAL 0x1     /* no stored rows: H.X regenerated on the fly        */
#define COSMO_OP_KEEP_SMALL 0x2  /* keep |h| < 1e-14 (the .HH content); default drops */
                                 /* them as the CSB builder does (PPMatrix.cxx:621)   */
#define COSMO_OP_NO_FP32 0x4     /* stored: skip the FP32 value copy                   */
#define COSMO_OP_FACTORED 0x8    /* virtual, in proton-neutron product form: H.X from  */
                                 /* per-species tables (no H in HBM); FP32 and FP64    */
                                 /* applies use it, FP64_ORDERED uses the row generator */

#define COSMO_PREC_FP64 0        /* FP64 values, FP64 vectors, FP64 accumulation       */
#define COSMO_PREC_FP32 1        /* FP32 values, FP32 vectors, FP32 accumulation       */
#define COSMO_PREC_FP64_ORDERED 2 /* FP64, each y[r] summed over the full row in      */
                                  /* ascending column order without fused multiply-add: */
                                  /* bit-identical to SparseMatrixCSB::TimesMatrix      */
                                  /* (include/tav/csbmatrix.cxx:623-663) on the same H  */

/* GPU Hamiltonian builder: replaces PPMatrix2CSBBlk (include/sm/PPMatrix.cxx:571-656)
 * and PPMatrix::Write's row generation (:413-465).  Rows hold every hop of the
 * one-/two-body operators with the reference's fermion phases and (row, col) indices;
 * values are summed in the reference's term order. */
int cosmo_op_create(cosmo_system s, int flags, cosmo_op *out);
/* Upload an upper-triangle row matrix (the .HH content: per row, cols ascending with the
 * diagonal first).  Replaces SparseMatrix<double>::Read (include/tav/sparse.cxx:401-445)
 * feeding davidson_vfile.  Entries with |h| < 1e-14 are kept unless dropped by flags. */
int cosmo_op_create_upper(int64_t n, const int64_t *rowptr, const int64_t *cols, const double *vals,
                          int flags, cosmo_op *out);
/* Read a .HH file into an operator (rows read on the host, then uploaded as by
 * cosmo_op_create_upper; |h| < 1e-14 kept, as the file holds them). */
int cosmo_op_read_hh(const char *path, int flags, cosmo_op *out);
/* Read a symmetric CSB file (SparseMatrixCSB::Read, include/tav/csbmatrix.cxx:524-552;
 * layout under cosmo_op_write_csb) into a stored operator; flags as cosmo_op_create_upper.
 * A diagonal entry missing from the file reads as 0.  Non-symmetric files: COSMO_EUNSUP. */
int cosmo_op_read_csb(const char *path, int flags, cosmo_op *out);
/* The same from CSB blocks in memory: the arrays of a symmetric SparseMatrixCSB
 * (include/tav/csbmatrix.cxx:128-377: BlockRow, BlockCol, numBlockRow, numBlockCol and per
 * block nnz, indRow, indCol, elem).  Block (i, j) is entry i * nblock_cols + j of nnz, rows,
 * cols and vals (pointers may be NULL where nnz is 0); only blocks on and above the
 * diagonal hold entries, each entry with col >= row.  In-block order is free (each row's
 * entries are sorted by column); a repeated entry is an error.  flags as
 * cosmo_op_create_upper.  Replaces handing a SparseMatrixCSB to MyMatrixDiagonal. */
int cosmo_op_create_csb(int64_t n, int block_rows, int block_cols, int nblock_rows, int nblock_cols,
                        const uint32_t *nnz, const uint16_t *const *rows, const uint16_t *const *cols,
                        const double *const *vals, int flags, cosmo_op *out);
int cosmo_op_destroy(cosmo_op op);
/* MatrixBase::Size (include/kdav/matrixbase.h:20). */
int64_t cosmo_op_size(cosmo_op op);
/* Stored upper-triangle and full entry counts (0 for a virtual operator). */
int cosmo_op_nnz(cosmo_op op, int64_t *nnz_upper, int64_t *nnz_full);
/* Exact device bytes a stored operator of s would take (the builder's counting pass on
 * the GPU, without the fill); used by the server's admission control. */
int cosmo_op_footprint(cosmo_system s, int flags, int64_t *bytes);
/* Device bytes held by the operator. */
int64_t cosmo_op_bytes(cosmo_op op);
/* MatrixBase::Diagonal / SparseMatrixCSB::ReadDiagonal (include/tav/csbmatrix.cxx:393-419). */
int cosmo_op_diagonal(cosmo_op op, double *d);
/* MatrixBase::TimesMatrix(ftyp** y, ftyp** x, ntyp m) (include/kdav/matrixbase.h:16-18,
 * called at include/kdav/fsdavidson.cxx:181): y[v] = H x[v] for m host n-vectors; y is
 * overwritten, x is read-only, complete on return.  precision: COSMO_PREC_*. */
int cosmo_op_apply(cosmo_op op, double *const *y, const double *const *x, int m, int precision);
/* Device-resident form: X and Y are n x m row-major (vector index fastest) device
 * arrays of double (FP64 modes) or float (FP32 mode) with leading dimension ld >= m.
 * stream may be NULL (the library's stream). */
int cosmo_op_apply_device(cosmo_op op, const void *X, void *Y, int m, int ld, int precision,
                          void *stream);
/* No counterpart in the reference.  cosmo_op_apply_device for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops): the same product, with
 * the same bits, allocating nothing (no hipMalloc/hipFree, no allocating library call, no
 * tuning, no graph capture).  Product-form operators (COSMO_OP_FACTORED) only, FP32 or
 * FP64 (not FP64_ORDERED), ld a power of two of >= 16 bytes.  The class products' scratch
 * is a 64 MiB device array per GPU held until the process ends: calls needing more than
 * 2 * n * ld * element size bytes (each half rounded up to 256) fail with COSMO_EINVAL.
 * Calls that overlap on one GPU share that array and race.  stream NULL: the null stream. */
int cosmo_op_apply_device_noalloc(cosmo_op op, const void *X, void *Y, int m, int ld, int precision,
                                  void *stream);
/* Upper-triangle rows [r0, r1) as stored (cols ascending, diagonal first): counts[r-r0]
 * always; cols/vals filled when non-NULL at offsets given by rowptr (exclusive prefix of
 * counts, relative to r0).  This is what XHH+JJ writes to .HH (PPMatrix.cxx:413-465). */
int cosmo_op_upper_rows(cosmo_op op, int64_t r0, int64_t r1, int64_t *counts, const int64_t *rowptr,
                        int64_t *cols, double *vals);
/* Write the .HH file of the operator (include/sm/PPMatrix.cxx:413-465 layout). */
int cosmo_op_write_hh(cosmo_op op, const char *path);
/* Write the operator's upper triangle as a CSB file, the layout of SparseMatrixCSB::Write
 * (include/tav/csbmatrix.cxx:553-576, CSBBlock::Write :246-254): [bool symmetric=1]
 * [u32 n][u16 BR][u16 BC][u16 nBR][u16 nBC], then all nBR x nBC blocks, block-row major:
 * [bool isDiagonal][u32 nnz][nnz u16 col % BC][nnz u16 row % BR][nnz f64], entries in
 * ascending (row, col) order, blocks below the diagonal empty (PPMatrix2CSBBlk,
 * include/sm/PPMatrix.cxx:571-656).  block = 0 uses uint16(sqrt(n)) + 1 as DEXHVcsb does
 * (include/xsm/DEXHVcsb.cxx:63).  An operator built with the default 1e-14 filter writes
 * the reference's file. */
int cosmo_op_write_csb(cosmo_op op, const char *path, int block);
/* All eigenpairs of a small operator (n <= 20000), ascending; evecs n x n column-major
 * (vector j at evecs + j*n) or NULL.  The exact-diagonalisation branch of the reference
 * for n <= limit (include/xsm/DEXHVcsb.cxx:72-108), computed on the GPU: the rows are
 * scattered into a dense FP64 matrix in HBM and solved by rocSOLVER (dsyevd). */
int cosmo_op_dense_eigen(cosmo_op op, double *evals, double *evecs);
/* Kernel of the stored FP32 / FP64 product: 0 automatic (the row kernel, then, for n >=
 * 262144, the tiled layout from the operator's 257th product on: it takes about 0.6 us per
 * row to build and repays that over some 400 products), 1 the row kernel, 2 the tiled
 * layout at any size from the next product on.  The tiled layout (row blocks of whole
 * neutron configurations, X tiles staged in LDS, proton-spectator tiles, value
 * dictionaries) is built on the first product that uses it and kept with the operator.  No reference counterpart:
 * a kernel choice of this build; every choice sums each row in ascending column order
 * within its kernel's partial sums and is deterministic. */
int cosmo_op_set_layout(cosmo_op op, int layout);
/* Last apply/build timings in milliseconds (device events). */
int cosmo_op_timings(cosmo_op op, double *build_ms, double *last_apply_ms);

/* ---- block Davidson ---------------------------------------------------------------- */
/* Parameters mirror DavidsonDiagonalization / Davidson::SetDefault
 * (include/kdav/davidson.cxx:167-223, include/kdav/cdavidson.cpp:86-110):
 * nume = number of lowest pairs, mblock = block size (may be < nume here),
 * lim = basis limit (nume+20), maxiter, crite/critc/critr/ortho thresholds.            */
typedef struct {
  int nume;
  int mblock;
  int lim;
  int maxiter;
  double crite, critc, critr, ortho;
  int precision;       /* COSMO_PREC_FP64: all-FP64.  COSMO_PREC_FP32: mixed (FP32 H.V
                          and bulk orthogonalisation, FP64 Rayleigh-Ritz, residuals and a
                          final FP64 refinement)                                          */
  double mixed_switch; /* mixed: residual (relative to |lambda|) at which the FP32 phase
                          hands over to FP64 refinement; 0 picks the default 1e-5        */
  int ninit;           /* number of initial vectors given (0: unit vectors at the lowest
                          diagonal elements, fsdavidson.cxx:540-611)                    */
  int verbose;
  int strict;          /* 1 (default): a pair counts as converged only when its residual
                          is below critr.  0: the reference's test, which also accepts a
                          pair whose last Ritz coefficients fall under critc and keeps it
                          converged for good (fsdavidson.cxx:243-256) -- that test can
                          end with unconverged pairs                                      */
} cosmo_dav_params;

typedef struct {
  int iterations;      /* dvdrvr2 loop count                                  */
  int matvecs;         /* H.V vector applications (nmv)                       */
  int matvecs_fp64;    /* of which in FP64                                    */
  int restarts;
  int ierr;            /* 1024: orthogonalisation failure (fsdavidson.cxx:721) */
  /* host wall time per stage; work is queued without waiting where nothing is read back,
     so a stage's time can include device work queued by the stage before it            */
  double ms_total, ms_apply, ms_ortho, ms_rr;
} cosmo_dav_stats;

/* Fill p with the reference defaults for nume pairs (block = nume). */
int cosmo_dav_defaults(cosmo_dav_params *p, int nume);
/* Lowest nume eigenpairs of op.  Outputs (host): evals[nume], dE[nume] (|delta lambda|
 * of the last step), resid[nume] (||H x - lambda x||), evecs (n x nume, column-major,
 * may be NULL).  init: ninit host n-vectors (column-major) or NULL.  Replaces
 * Davidson::Calculate + dvdrvr2 (include/kdav/cdavidson.cpp:220-303,
 * include/kdav/fsdavidson.cxx:615-775).  Returns COSMO_ECONV when maxiter was reached
 * (results are still written). */
int cosmo_davidson(cosmo_op op, const cosmo_dav_params *p, const double *init, double *evals, double *dE,
                   double *resid, double *evecs, cosmo_dav_stats *stats);
/* .EE file I/O (include/kdav/cdavidson.cpp:327-368, :386-427). */
int cosmo_write_ee(const char *path, int64_t n, int count, const double *evals, const double *dE,
                   const double *resid, const double *evecs);
int cosmo_read_ee(const char *path, int64_t n, int maxcount, int *count, double *evals, double *evecs);

/* ---- batched solves ------------------------------------------------------------------ */
/* Independent systems solved together (DavidsonDiagonalization per system,
 * include/kdav/davidson.cxx:167-223, with a wider basis: lim = the larger of p->lim,
 * nume + max(20, nume) and min(nume + 40, 45), at most n; at least 4000 iterations).  Solves of dimension <= the single-launch bound (default
 * 16384, at most 2^30) with lim <= 48 and nume, block <= 16 on stored operators run to
 * completion in ONE persistent kernel launch, each by a team of workgroups (one per 2048
 * rows, at most 128); the others run concurrently from worker threads, each on its own
 * stream with the host-driven Davidson.  Results land per solve (evals and resid:
 * nume_per_solve[i] entries each, concatenated).  COSMO_ECONV: a solve stopped above its
 * residual target; COSMO_EHIP: a team of the launch waited more than 10 s at a barrier
 * with all its workgroups seated (the launch then ends without waiting further). */
int cosmo_batch_create(int nsys, const cosmo_system *systems, int flags, cosmo_batch *out);
int cosmo_batch_solve(cosmo_batch b, const cosmo_dav_params *p, const int *nume_per_solve, double *evals,
                      double *resid, cosmo_dav_stats *stats);
/* Dimension bound of the single-launch path (0 turns it off; 1 << 30 puts every solve
 * within the kernel's bounds in the one launch).  No reference counterpart: a batching
 * knob of this build. */
int cosmo_batch_set_single_launch(cosmo_batch b, int64_t nmax);
int cosmo_batch_destroy(cosmo_batch b);

/* ---- serving ----------------------------------------------------------------------------- */
typedef struct {
  int64_t hbm_budget;  /* bytes the server may hold for operators + Davidson blocks     */
  int max_active;      /* concurrent solves on the device                              */
  int precision;       /* COSMO_PREC_FP64 or COSMO_PREC_FP32 (mixed)                   */
} cosmo_server_params;

typedef struct {
  int64_t id;
  int64_t n;
  int nume;
  int status;          /* COSMO_OK / error code                                        */
  double t_submit, t_admit, t_done; /* seconds since server start                      */
  double evals[64];
  double resid[64];
} cosmo_result;

int cosmo_server_create(const cosmo_server_params *p, cosmo_server *out);
/* Queue a request: the system (ownership stays with the caller until the request
 * completes), pairs wanted and residual tolerance.  Returns the request id in *id.  A
 * request whose stored H and Davidson blocks exceed hbm_budget runs on the product-form
 * operator (COSMO_OP_FACTORED, no H in HBM) when that fits; otherwise COSMO_ENOMEM. */
int cosmo_server_submit(cosmo_server s, cosmo_system sys, int nume, double tol, int64_t *id);
/* Run admission + one scheduling round; returns the number of results written (<= cap).*/
int cosmo_server_poll(cosmo_server s, cosmo_result *results, int cap);
/* Block until every submitted request has finished; results as in poll. */
int cosmo_server_drain(cosmo_server s, cosmo_result *results, int cap, int *written);
int cosmo_server_destroy(cosmo_server s);

#ifdef __cplusplus
}
#endif
#endif /* COSMO_HIP_H */
