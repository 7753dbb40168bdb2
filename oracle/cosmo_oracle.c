/*
 * cosmo_oracle.c -- CPU restatement of the CoSMo configuration-interaction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product library: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The product
 * never links or calls it.
 *
 * Every function restates one piece of the reference algorithm from its published
 * behaviour (file:line citations are into the CoSMo source tree).  The implementation is
 * deliberately the plain list-walking one (sorted occupation lists, binary search), not
 * the bit-string / rank-table scheme of the GPU builder, so that agreement between the
 * two is evidence rather than a tautology.
 *
 * Pinning (how each function is checked, see tests/test_oracle_*.py):
 *   orc_cg            closed-form Racah formula; checked against textbook values,
 *                     orthogonality sums and the reference VSPX dump (oracle/_ref).
 *   orc_jm_vspx       against the reference's own VSPX/ESP tensors (ref_harness dumpint),
 *                     keys exact, values <= 1e-13 abs (CG by closed form vs recurrence).
 *   orc_basis         byte-exact against the reference .mbs (XsysmbsOMP), closed-form
 *                     dimensions, brute force on tiny spaces.
 *   orc_h_rows        against the reference .HH (XHH+JJ): row structure exact, values
 *                     <= 1e-12 abs; invariants (Hermiticity, J^2 multiplets).
 *   orc_csb_times     against the reference's SparseMatrixCSB::TimesMatrix output
 *                     (ref_harness spmm) on the same H, and dense products.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------ */
/* Clebsch-Gordan coefficient <j1 m1 j2 m2 | J M>, all arguments doubled.
 * Reference computes it from a Schulten-Gordon 3j recurrence
 * (include/tav/ThreeJSymbolM.cxx:167-492, ClebschGordan at :487-492); this restates the
 * same quantity by Racah's closed form in long double.                                  */
static long double lfact(int n) {
  long double r = 1.0L;
  for (int i = 2; i <= n; ++i) r *= (long double)i;
  return r;
}

double orc_cg(int j1, int m1, int j2, int m2, int J, int M) {
  if (m1 + m2 != M) return 0.0;
  if (J < abs(j1 - j2) || J > j1 + j2) return 0.0;
  if (((j1 + j2 + J) & 1) != 0) return 0.0;
  if (abs(m1) > j1 || abs(m2) > j2 || abs(M) > J) return 0.0;
  if (((j1 + m1) & 1) || ((j2 + m2) & 1) || ((J + M) & 1)) return 0.0;
  int a = (j1 + j2 - J) / 2, b = (j1 - j2 + J) / 2, c = (-j1 + j2 + J) / 2, d = (j1 + j2 + J) / 2 + 1;
  long double pre = (long double)(J + 1) * lfact(a) * lfact(b) * lfact(c) / lfact(d);
  pre *= lfact((j1 + m1) / 2) * lfact((j1 - m1) / 2) * lfact((j2 + m2) / 2) * lfact((j2 - m2) / 2) *
         lfact((J + M) / 2) * lfact((J - M) / 2);
  long double s = 0.0L;
  for (int k = 0; k < 64; ++k) {
    int e1 = a - k, e2 = (j1 - m1) / 2 - k, e3 = (j2 + m2) / 2 - k;
    int e4 = (J - j2 + m1) / 2 + k, e5 = (J - j1 - m2) / 2 + k;
    if (e1 < 0 || e2 < 0 || e3 < 0) break;
    if (e4 < 0 || e5 < 0) continue;
    long double t = 1.0L / (lfact(k) * lfact(e1) * lfact(e2) * lfact(e3) * lfact(e4) * lfact(e5));
    s += (k & 1) ? -t : t;
  }
  return (double)(sqrtl(pre) * s);
}

/* ------------------------------------------------------------------------------------ */
/* J-coupled -> m-scheme two-body operator.  Restates AddPPVsp
 * (include/sm/AddPPVsp.cxx:32-146) driven by PPInteraction2sp
 * (include/sm/spherical2sp.cxx:790-814): for every JV key [l0 l1 l2 l3 J T] in map order
 * and every (a0,a1,a2,a3) of those levels in ascending sps order with conserved Jz, Tz,
 *   h = CG(J) CG(J) CG(T) CG(T) (skipped if |h| < 1e-9) * V / sqrt((1+d01)(1+d23)),
 *   sqrt(2) for a pn pair of distinct levels with equal n, j, parity,
 *   b = sorted pairs, phase ^= 2j&1 per swap, pairs forward-ordered,
 *   h /= 2 if the level pairs are equal, h *= 2 if the m-pairs are equal,
 *   VSP[b] += +-h   (contributions to one key summed in generation order).
 * Output entries sorted by key; returns the count (or -needed if cap is too small).    */
typedef struct {
  uint16_t k[4];
  double v;
  long seq;
} vcontrib;

static int cmp_vcontrib(const void *pa, const void *pb) {
  const vcontrib *a = (const vcontrib *)pa, *b = (const vcontrib *)pb;
  for (int i = 0; i < 4; ++i)
    if (a->k[i] != b->k[i]) return a->k[i] < b->k[i] ? -1 : 1;
  return a->seq < b->seq ? -1 : (a->seq > b->seq ? 1 : 0);
}

long orc_jm_vspx(int nsps, const int *level, const int *j2, const int *m2, const int *tz2,
                 const int *t2, const int *nq, const int *par, long njv, const int *jvk,
                 const double *jvv, long cap, uint16_t *okeys, double *ovals) {
  long ncap = 1024, nc = 0;
  vcontrib *cs = (vcontrib *)malloc(sizeof(vcontrib) * ncap);
  for (long q = 0; q < njv; ++q) {
    const int *ijk = jvk + 6 * q;
    double V = jvv[q];
    for (int a0 = 0; a0 < nsps; ++a0) {
      if (level[a0] != ijk[0]) continue;
      for (int a1 = 0; a1 < nsps; ++a1) {
        if (level[a1] != ijk[1]) continue;
        if ((j2[a0] & 1) && a1 == a0) continue;
        for (int a2 = 0; a2 < nsps; ++a2) {
          if (level[a2] != ijk[2]) continue;
          for (int a3 = 0; a3 < nsps; ++a3) {
            if (level[a3] != ijk[3]) continue;
            if ((j2[a2] & 1) && a2 == a3) continue;
            if (m2[a0] + m2[a1] != m2[a3] + m2[a2]) continue;
            if (tz2[a0] + tz2[a1] != tz2[a3] + tz2[a2]) continue;
            double h = orc_cg(j2[a0], m2[a0], j2[a1], m2[a1], 2 * ijk[4], m2[a0] + m2[a1]);
            h *= orc_cg(j2[a2], m2[a2], j2[a3], m2[a3], 2 * ijk[4], m2[a2] + m2[a3]);
            h *= orc_cg(t2[a0], tz2[a0], t2[a1], tz2[a1], 2 * ijk[5], tz2[a0] + tz2[a1]);
            h *= orc_cg(t2[a2], tz2[a2], t2[a3], tz2[a3], 2 * ijk[5], tz2[a2] + tz2[a3]);
            if (fabs(h) < 1e-9) continue;
            h = h * V / sqrt(((double)(ijk[0] == ijk[1]) + 1.) * ((double)(ijk[2] == ijk[3]) + 1.));
            if (ijk[0] != ijk[1] && nq[a0] == nq[a1] && j2[a0] == j2[a1] && tz2[a0] == -tz2[a1] &&
                par[a0] == par[a1])
              h *= sqrt(2.);
            if (ijk[2] != ijk[3] && nq[a2] == nq[a3] && j2[a2] == j2[a3] && tz2[a2] == -tz2[a3] &&
                par[a2] == par[a3])
              h *= sqrt(2.);
            int b[4] = {a0, a1, a2, a3};
            int phase = 0;
            if (b[0] > b[1]) { int t = b[0]; b[0] = b[1]; b[1] = t; phase ^= j2[a0] & 1; }
            if (b[2] > b[3]) { int t = b[2]; b[2] = b[3]; b[3] = t; phase ^= j2[a2] & 1; }
            if (b[2] < b[0] || (b[2] == b[0] && b[3] < b[1])) {
              int t0 = b[0], t1 = b[1];
              b[0] = b[2]; b[1] = b[3]; b[2] = t0; b[3] = t1;
            }
            if (ijk[0] == ijk[2] && ijk[1] == ijk[3]) h /= 2.0;
            if (b[0] == b[2] && b[1] == b[3]) h *= 2.0;
            if (nc == ncap) { ncap *= 2; cs = (vcontrib *)realloc(cs, sizeof(vcontrib) * ncap); }
            for (int i = 0; i < 4; ++i) cs[nc].k[i] = (uint16_t)b[i];
            cs[nc].v = phase ? -h : h;
            cs[nc].seq = nc;
            ++nc;
          }
        }
      }
    }
  }
  qsort(cs, nc, sizeof(vcontrib), cmp_vcontrib);
  long nout = 0;
  for (long i = 0; i < nc;) {
    long j = i;
    double acc = 0.0;
    while (j < nc && memcmp(cs[j].k, cs[i].k, sizeof(cs[i].k)) == 0) { acc += cs[j].v; ++j; }
    if (nout < cap) {
      memcpy(okeys + 4 * nout, cs[i].k, sizeof(cs[i].k));
      ovals[nout] = acc;
    }
    ++nout;
    i = j;
  }
  free(cs);
  return nout <= cap ? nout : -nout;
}

/* ------------------------------------------------------------------------------------ */
/* M-scheme basis in the reference order.  The reference writes all sorted occupation
 * lists with sum(2m) = 2Jz, the given neutron/proton counts and parity, sorted
 * lexicographically (xmlshell/xmbs/XsysmbsOMP.cpp:403-429, MakeMBS.cxx:186-410).  Here:
 * an odometer over neutron combinations (orbitals [0,nn)) in lexicographic order, and for
 * each an odometer over proton combinations ([nn, nn+np)), keeping matching states.
 * Output n x (Nv+Zv) uint16; returns n (or -n if cap too small).                       */
static int next_comb(int *c, int k, int lo, int hi) {
  /* next k-combination of [lo,hi) in lexicographic order; 0 when exhausted */
  int i = k - 1;
  while (i >= 0 && c[i] == hi - k + i) --i;
  if (i < 0) return 0;
  c[i]++;
  for (int j = i + 1; j < k; ++j) c[j] = c[j - 1] + 1;
  (void)lo;
  return 1;
}

long orc_basis(int nn, int np, const int *m2, const int *par, int Nv, int Zv, int twoM, int parity,
               long cap, uint16_t *out) {
  int A = Nv + Zv;
  int cn[64], cp[64];
  long n = 0;
  for (int i = 0; i < Nv; ++i) cn[i] = i;
  int more_n = (Nv <= nn);
  while (more_n) {
    int mn = 0, pn = 0;
    for (int i = 0; i < Nv; ++i) { mn += m2[cn[i]]; pn ^= par[cn[i]]; }
    for (int i = 0; i < Zv; ++i) cp[i] = nn + i;
    int more_p = (Zv <= np);
    while (more_p) {
      int mp = 0, pp = 0;
      for (int i = 0; i < Zv; ++i) { mp += m2[cp[i]]; pp ^= par[cp[i]]; }
      if (mn + mp == twoM && (pn ^ pp) == parity) {
        if (n < cap) {
          for (int i = 0; i < Nv; ++i) out[n * A + i] = (uint16_t)cn[i];
          for (int i = 0; i < Zv; ++i) out[n * A + Nv + i] = (uint16_t)cp[i];
        }
        ++n;
      }
      more_p = Zv > 0 ? next_comb(cp, Zv, nn, nn + np) : 0;
    }
    more_n = Nv > 0 ? next_comb(cn, Nv, 0, nn) : 0;
  }
  return n <= cap ? n : -n;
}

/* ------------------------------------------------------------------------------------ */
/* Fermion operator application.  Restates Annihilation / Creation
 * (include/sm/operators.cxx:121-187): phase ^= parity of the number of particles passed
 * in the running list, false when the move is impossible.                              */
static int annihilate(int *phase, int Nbra, const uint16_t *bra, int Nop, const uint16_t *op,
                      uint16_t *ket) {
  int nnf = 0, nni = 0;
  if (Nop > Nbra) return 0;
  for (int nno = 0; nno < Nop; ++nno) {
    while (nni < Nbra && bra[nni] < op[nno]) ket[nnf++] = bra[nni++];
    if (nni == Nbra || bra[nni] != op[nno]) return 0;
    ++nni;
    *phase ^= (nnf & 1);
  }
  while (nni < Nbra) ket[nnf++] = bra[nni++];
  return 1;
}

static int create(int *phase, int Nbra, const uint16_t *bra, int Nop, const uint16_t *op,
                  uint16_t *ket) {
  int nnf = 0, nni = 0, nno = 0, Nket = Nbra + Nop;
  while (nnf < Nket) {
    if (nni == Nbra) { ket[nnf++] = op[nno++]; *phase ^= (nni & 1); continue; }
    if (nno == Nop) { ket[nnf++] = bra[nni++]; continue; }
    if (bra[nni] < op[nno]) { ket[nnf++] = bra[nni++]; continue; }
    if (bra[nni] > op[nno]) { ket[nnf++] = op[nno++]; *phase ^= (nni & 1); continue; }
    return 0;
  }
  return 1;
}

static int cmp_list(int N, const uint16_t *a, const uint16_t *b) {
  for (int i = 0; i < N; ++i) {
    if (a[i] < b[i]) return -1;
    if (a[i] > b[i]) return 1;
  }
  return 0;
}

/* binary search of a sorted occupation list (include/sm/ManyBodyStates.cxx:92-130);
 * returns -1 when absent */
static long locate(int N, long n, const uint16_t *basis, const uint16_t *a, long lo) {
  long hi = n - 1;
  if (lo > hi) return -1;
  if (cmp_list(N, a, basis + lo * N) < 0 || cmp_list(N, a, basis + hi * N) > 0) return -1;
  while (lo <= hi) {
    long mid = lo + (hi - lo) / 2;
    int c = cmp_list(N, a, basis + mid * N);
    if (c == 0) return mid;
    if (c > 0) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

typedef struct {
  uint16_t ket[32];
  double v;
  int seq;
} ycontrib;

static int g_N_for_sort;
static int cmp_ycontrib(const void *pa, const void *pb) {
  const ycontrib *a = (const ycontrib *)pa, *b = (const ycontrib *)pb;
  int c = cmp_list(g_N_for_sort, a->ket, b->ket);
  if (c) return c;
  return a->seq < b->seq ? -1 : (a->seq > b->seq ? 1 : 0);
}

/* One bra row of the upper-triangle H.  Restates ActFermiOperator over the operator list
 * {ESP, VSPX} (include/sm/cioperators.cxx:36-62, PPMatrix.cxx:612-613): terms in map
 * order, Y[ket] += +-v merged per ket in term order; then kets ascending, located in the
 * basis from a lower bound (PPMatrix.cxx:615-629).  drop_small applies the CSB filter
 * |h| < 1e-14 (PPMatrix.cxx:621); without it the row is the .HH row (PPMatrix.cxx:432-440).
 * Returns the row length; cols/vals filled when non-NULL.                               */
static long h_row(int N, long n, const uint16_t *basis, long bra, long nesp, const uint16_t *espk,
                  const double *espv, long nv, const uint16_t *vk, const double *vv, const long *pstart,
                  int norb, int drop_small, ycontrib **ybuf, long *ycap, long *cols, double *vals) {
  const uint16_t *b = basis + bra * N;
  uint16_t mid[32];
  long ny = 0;
  int seq = 0;
  ycontrib *Y = *ybuf;
  /* one-body terms: annihilate espk[2t], create espk[2t+1] */
  for (long t = 0; t < nesp; ++t) {
    int ph = 0;
    if (!annihilate(&ph, N, b, 1, espk + 2 * t, mid)) continue;
    uint16_t ket[32];
    if (!create(&ph, N - 1, mid, 1, espk + 2 * t + 1, ket)) continue;
    if (ny == *ycap) { *ycap *= 2; Y = *ybuf = (ycontrib *)realloc(Y, sizeof(ycontrib) * (*ycap)); }
    memcpy(Y[ny].ket, ket, N * sizeof(uint16_t));
    Y[ny].v = ph ? -espv[t] : espv[t];
    Y[ny].seq = seq++;
    ++ny;
  }
  /* two-body terms, visited grouped by their annihilated pair in key order */
  for (int i0 = 0; i0 < N; ++i0)
    for (int i1 = i0 + 1; i1 < N; ++i1) {
      long p = (long)b[i0] * norb + b[i1];
      for (long t = pstart[p]; t < pstart[p + 1]; ++t) {
        int ph = 0;
        if (!annihilate(&ph, N, b, 2, vk + 4 * t, mid)) continue;
        uint16_t ket[32];
        if (!create(&ph, N - 2, mid, 2, vk + 4 * t + 2, ket)) continue;
        if (ny == *ycap) { *ycap *= 2; Y = *ybuf = (ycontrib *)realloc(Y, sizeof(ycontrib) * (*ycap)); }
        memcpy(Y[ny].ket, ket, N * sizeof(uint16_t));
        Y[ny].v = ph ? -vv[t] : vv[t];
        Y[ny].seq = seq++;
        ++ny;
      }
    }
  g_N_for_sort = N;
  qsort(Y, ny, sizeof(ycontrib), cmp_ycontrib);
  long cnt = 0, lo = bra;
  for (long i = 0; i < ny;) {
    long j = i;
    double acc = 0.0;
    while (j < ny && cmp_list(N, Y[j].ket, Y[i].ket) == 0) { acc += Y[j].v; ++j; }
    long ket = locate(N, n, basis, Y[i].ket, lo);
    i = j;
    if (ket < 0) continue;
    if (drop_small && fabs(acc) < 1e-14) continue;
    lo = ket;
    if (cols) { cols[cnt] = ket; vals[cnt] = acc; }
    ++cnt;
  }
  return cnt;
}

/* pstart[p] for p = b0*norb + b1 over the sorted VSPX keys: VSPX key order groups terms
 * by their annihilated pair (b0,b1), so the group of one pair is a contiguous range.    */
static long *pair_starts(int norb, long nv, const uint16_t *vk) {
  long np = (long)norb * norb;
  long *ps = (long *)calloc(np + 1, sizeof(long));
  for (long t = 0; t < nv; ++t) ps[(long)vk[4 * t] * norb + vk[4 * t + 1] + 1]++;
  for (long p = 0; p < np; ++p) ps[p + 1] += ps[p];
  return ps;
}

/* Rows [r0,r1) of the upper triangle: rowcount[r-r0] always; when cols != NULL also
 * the entries at offsets rowptr[r-r0] (exclusive prefix given by the caller).           */
long orc_h_rows(int N, long n, const uint16_t *basis, int norb, long nesp, const uint16_t *espk,
                const double *espv, long nv, const uint16_t *vk, const double *vv, long r0, long r1,
                int drop_small, long *rowcount, const long *rowptr, long *cols, double *vals) {
  long *ps = pair_starts(norb, nv, vk);
  long total = 0;
#pragma omp parallel reduction(+ : total)
  {
    long ycap = 4096;
    ycontrib *Y = (ycontrib *)malloc(sizeof(ycontrib) * ycap);
#pragma omp for schedule(dynamic, 64)
    for (long r = r0; r < r1; ++r) {
      long off = rowptr ? rowptr[r - r0] : 0;
      long c = h_row(N, n, basis, r, nesp, espk, espv, nv, vk, vv, ps, norb, drop_small, &Y, &ycap,
                     cols ? cols + off : NULL, vals ? vals + off : NULL);
      rowcount[r - r0] = c;
      total += c;
    }
    free(Y);
  }
  free(ps);
  return total;
}

/* ------------------------------------------------------------------------------------ */
/* CSB product.  Restates SparseMatrixCSB<double>::TimesMatrix for a symmetric matrix
 * (include/tav/csbmatrix.cxx:623-663 with CSBBlock::TimesMatrix / TimesMatrixTranspose
 * :280-328): BR = BC = br; y zeroed; for every block row j and block column i, blocks
 * with i < j are applied transposed from the stored upper block (i,j), blocks with i >= j
 * directly, the diagonal block with both triangles.  Within a block entries are in
 * (row, col) order.  Input: upper-triangle rows in CSR (rowptr/cols/vals, cols ascending).
 * x, y are m separate n-vectors (column-major n x m).                                  */
typedef struct {
  long start, count;
} blk;

void orc_csb_times(long n, int br, const long *rowptr, const long *cols, const double *vals, int m,
                   const double *x, double *y) {
  long nb = (n + br - 1) / br;
  /* scatter entries into blocks (row-major per block, as the builder writes them) */
  long *bcount = (long *)calloc(nb * nb, sizeof(long));
  for (long r = 0; r < n; ++r)
    for (long e = rowptr[r]; e < rowptr[r + 1]; ++e) bcount[(r / br) * nb + cols[e] / br]++;
  long *bstart = (long *)malloc(sizeof(long) * (nb * nb + 1));
  bstart[0] = 0;
  for (long b = 0; b < nb * nb; ++b) bstart[b + 1] = bstart[b] + bcount[b];
  long nnz = bstart[nb * nb];
  int *lr = (int *)malloc(sizeof(int) * nnz), *lc = (int *)malloc(sizeof(int) * nnz);
  double *lv = (double *)malloc(sizeof(double) * nnz);
  long *fill = (long *)calloc(nb * nb, sizeof(long));
  for (long r = 0; r < n; ++r)
    for (long e = rowptr[r]; e < rowptr[r + 1]; ++e) {
      long bidx = (r / br) * nb + cols[e] / br;
      long at = bstart[bidx] + fill[bidx]++;
      lr[at] = (int)(r % br);
      lc[at] = (int)(cols[e] % br);
      lv[at] = vals[e];
    }
  for (int v = 0; v < m; ++v) memset(y + (long)v * n, 0, sizeof(double) * n);
#pragma omp parallel for schedule(dynamic, 1)
  for (long j = 0; j < nb; ++j) {
    for (long i = 0; i < nb; ++i) {
      if (i < j) {
        long bidx = i * nb + j;
        for (int v = 0; v < m; ++v) {
          const double *xv = x + (long)v * n + i * br;
          double *yv = y + (long)v * n + j * br;
          for (long e = bstart[bidx]; e < bstart[bidx + 1]; ++e) yv[lc[e]] += lv[e] * xv[lr[e]];
        }
      } else {
        long bidx = j * nb + i;
        int diag = (i == j);
        for (int v = 0; v < m; ++v) {
          const double *xv = x + (long)v * n + i * br;
          double *yv = y + (long)v * n + j * br;
          const double *xr = x + (long)v * n + j * br;
          double *yc = y + (long)v * n + i * br;
          for (long e = bstart[bidx]; e < bstart[bidx + 1]; ++e) {
            yv[lr[e]] += lv[e] * xv[lc[e]];
            if (diag && lr[e] != lc[e]) yc[lc[e]] += lv[e] * xr[lr[e]];
          }
        }
      }
    }
  }
  free(bcount); free(bstart); free(lr); free(lc); free(lv); free(fill);
}

/* .HH row-format product: restates SparseMatrix<double>::TimesMatrix with type=1
 * (include/tav/sparse.cxx:273-291): per vector, rows in order, y[i] += h x[col] and, off
 * the diagonal, y[col] += h x[i].                                                       */
void orc_hh_times(long n, const long *rowptr, const long *cols, const double *vals, int m,
                  const double *x, double *y) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int v = 0; v < m; ++v) {
    const double *xv = x + (long)v * n;
    double *yv = y + (long)v * n;
    for (long i = 0; i < n; ++i) yv[i] = 0.0;
    for (long i = 0; i < n; ++i)
      for (long e = rowptr[i]; e < rowptr[i + 1]; ++e) {
        yv[i] += vals[e] * xv[cols[e]];
        if (cols[e] != i) yv[cols[e]] += vals[e] * xv[i];
      }
  }
}

/* CPU baseline form of the CSB product timed on the host: block rows in parallel with
 * the same block traversal, vectors innermost over an n x m row-major (interleaved)
 * layout.  Used only to time the reference algorithm on this machine's cores.          */
typedef struct {
  long nb, br, nnz;
  long *bstart;
  uint16_t *lr, *lc;
  double *lv;
} orc_csb;

void *orc_csb_build(long n, int br, const long *rowptr, const long *cols, const double *vals) {
  orc_csb *A = (orc_csb *)calloc(1, sizeof(orc_csb));
  long nb = (n + br - 1) / br;
  A->nb = nb; A->br = br;
  long *bcount = (long *)calloc(nb * nb, sizeof(long));
  for (long r = 0; r < n; ++r)
    for (long e = rowptr[r]; e < rowptr[r + 1]; ++e) bcount[(r / br) * nb + cols[e] / br]++;
  A->bstart = (long *)malloc(sizeof(long) * (nb * nb + 1));
  A->bstart[0] = 0;
  for (long b = 0; b < nb * nb; ++b) A->bstart[b + 1] = A->bstart[b] + bcount[b];
  A->nnz = A->bstart[nb * nb];
  A->lr = (uint16_t *)malloc(2 * A->nnz);
  A->lc = (uint16_t *)malloc(2 * A->nnz);
  A->lv = (double *)malloc(8 * A->nnz);
  memset(bcount, 0, sizeof(long) * nb * nb);
  for (long r = 0; r < n; ++r)
    for (long e = rowptr[r]; e < rowptr[r + 1]; ++e) {
      long bidx = (r / br) * nb + cols[e] / br;
      long at = A->bstart[bidx] + bcount[bidx]++;
      A->lr[at] = (uint16_t)(r % br);
      A->lc[at] = (uint16_t)(cols[e] % br);
      A->lv[at] = vals[e];
    }
  free(bcount);
  return A;
}

void orc_csb_free(void *p) {
  orc_csb *A = (orc_csb *)p;
  free(A->bstart); free(A->lr); free(A->lc); free(A->lv); free(A);
}

void orc_csb_apply(void *p, long n, int m, const double *x, double *y) {
  orc_csb *A = (orc_csb *)p;
  long nb = A->nb, br = A->br;
  memset(y, 0, sizeof(double) * n * m);
#pragma omp parallel for schedule(dynamic, 1)
  for (long j = 0; j < nb; ++j)
    for (long i = 0; i < nb; ++i) {
      if (i < j) {
        long b = i * nb + j;
        for (long e = A->bstart[b]; e < A->bstart[b + 1]; ++e) {
          const double *xs = x + (i * br + A->lr[e]) * m;
          double *ys = y + (j * br + A->lc[e]) * m;
          double h = A->lv[e];
          for (int v = 0; v < m; ++v) ys[v] += h * xs[v];
        }
      } else {
        long b = j * nb + i;
        for (long e = A->bstart[b]; e < A->bstart[b + 1]; ++e) {
          long r = j * br + A->lr[e], c = i * br + A->lc[e];
          double h = A->lv[e];
          for (int v = 0; v < m; ++v) y[r * m + v] += h * x[c * m + v];
          if (i == j && r != c)
            for (int v = 0; v < m; ++v) y[c * m + v] += h * x[r * m + v];
        }
      }
    }
}
