// labels.hip -- spin and isospin of eigenvectors: <J-J+> = |J+ z|^2 and <T-T+> = |T+ z|^2
// computed on the GPU (the reference's XSHLJT, include/xsm/XSHLJT.cxx:38-142, evaluates
// PPMatrix::Sqr of the J+ / T+ one-body operators, include/sm/PPMatrix.cxx:400-411).
// J+ = sum_g c_g a+_{g'} a_g with g' = same level and species, m+1, c_g = sqrt(j(j+1) -
// m(m+1)); T+ = sum a+_{neutron} a_{proton} over equal (level, m).  The images live in
// the basis with M+1 (resp. one more neutron), built with the same rank tables.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "op.h"

namespace cosmo {

__device__ __forceinline__ uint32_t colex32_l(const uint32_t *binom, uint32_t m) {
  uint32_t r = 0;
  int i = 1;
  while (m) {
    int p = __ffs(m) - 1;
    r += binom[p * 33 + i];
    ++i;
    m &= m - 1;
  }
  return r;
}

struct KetTables {
  int nn, nv, zv;
  const int32_t *ncfg_index, *pcfg_rank, *pcfg_sec, *nsector;
  const int64_t *noff;
};

struct BraTables {
  int nn, nnconf;
  const uint32_t *nmask, *pmask;
  const int64_t *noff, *psec_start;
  const int32_t *nsector;
  const int32_t *krank;  // truncated source basis: kept sector ranks per block (else null)
  const int64_t *kstart;
};

// |J+ z|^2 and |T+ z|^2 run over every state the raising operator reaches, as the
// reference's PPMatrix::Sqr does, so the target basis is the untruncated one
static void untruncate(System &t) {
  t.hw_cut = false;
  t.rejections.clear();
}

// partner[g] = orbital reached by the raising operator from g (-1 none); coef[g] its factor
__global__ void k_raise(BraTables S, KetTables T, const uint32_t *binom, int64_t n, const int16_t *partner,
                        const double *coef, int norb, const double *__restrict__ Z, int count,
                        double *__restrict__ Y, int64_t ny) {
  int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  int lo = 0, hi = S.nnconf - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (S.noff[mid] <= row) lo = mid; else hi = mid - 1;
  }
  const int64_t pr = S.krank ? S.krank[S.kstart[lo] + (row - S.noff[lo])] : row - S.noff[lo];
  uint32_t pm = S.pmask[S.psec_start[S.nsector[lo]] + pr];
  uint64_t B = (uint64_t)S.nmask[lo] | ((uint64_t)pm << S.nn);
  for (int g = 0; g < norb; ++g) {
    if (!((B >> g) & 1ull)) continue;
    int h = partner[g];
    if (h < 0 || ((B >> h) & 1ull)) continue;
    uint64_t K = (B & ~(1ull << g)) | (1ull << h);
    uint32_t nm = (uint32_t)(K & ((1ull << T.nn) - 1ull)), km = (uint32_t)(K >> T.nn);
    if (__popc(nm) != T.nv || __popc(km) != T.zv) continue;
    int ni = T.ncfg_index[colex32_l(binom, nm)];
    if (ni < 0) continue;
    uint32_t pc = colex32_l(binom, km);
    if (T.pcfg_sec[pc] != T.nsector[ni]) continue;
    int64_t col = T.noff[ni] + T.pcfg_rank[pc];
    int ph = __popcll(B & ((1ull << g) - 1ull)) + __popcll((B & ~(1ull << g)) & ((1ull << h) - 1ull));
    double c = (ph & 1) ? -coef[g] : coef[g];
    for (int v = 0; v < count; ++v) atomicAdd(&Y[(int64_t)v * ny + col], c * Z[(int64_t)v * n + row]);
  }
}

template <typename T>
static int up(std::vector<void *> &owned, const std::vector<T> &h, const T **out) {
  void *d = nullptr;
  size_t b = std::max<size_t>(h.size(), 1) * sizeof(T);
  HIPCHK(dmalloc(&d, b));
  owned.push_back(d);
  if (!h.empty()) HIPCHK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  *out = (const T *)d;
  return COSMO_OK;
}

static int raise_norms(const System &s, const System &t, const std::vector<int16_t> &partner,
                       const std::vector<double> &coef, const double *Z, int count, double *out) {
  std::vector<void *> owned;
  auto cleanup = [&]() {
    for (void *p : owned) dfree(p);
  };
  if (t.dim == 0) {
    for (int v = 0; v < count; ++v) out[v] = 0.0;
    return COSMO_OK;
  }
  std::vector<uint32_t> b(33 * 33);
  for (int i = 0; i <= 32; ++i)
    for (int k = 0; k <= 32; ++k) b[i * 33 + k] = (uint32_t)binom(i, k);
  std::vector<int32_t> psec(t.pcfg_rank.size(), -1);
  for (size_t sec = 0; sec + 1 < t.psec_start.size(); ++sec)
    for (int64_t q = t.psec_start[sec]; q < t.psec_start[sec + 1]; ++q) psec[colex_rank(t.pmask[q])] = (int32_t)sec;
  BraTables S{};
  KetTables T{};
  const uint32_t *db;
  const int16_t *dp;
  const double *dc;
  int rc = 0;
  rc |= up(owned, b, &db);
  rc |= up(owned, partner, &dp);
  rc |= up(owned, coef, &dc);
  S.nn = s.nn;
  S.nnconf = (int)s.nmask.size();
  rc |= up(owned, s.nmask, &S.nmask);
  rc |= up(owned, s.pmask, &S.pmask);
  rc |= up(owned, s.noff, &S.noff);
  rc |= up(owned, s.psec_start, &S.psec_start);
  rc |= up(owned, s.nsector, &S.nsector);
  if (s.truncated()) {
    rc |= up(owned, s.krank, &S.krank);
    rc |= up(owned, s.kstart, &S.kstart);
  }
  T.nn = t.nn;
  T.nv = t.nv;
  T.zv = t.zv;
  rc |= up(owned, t.ncfg_index, &T.ncfg_index);
  rc |= up(owned, t.pcfg_rank, &T.pcfg_rank);
  rc |= up(owned, psec, &T.pcfg_sec);
  rc |= up(owned, t.nsector, &T.nsector);
  rc |= up(owned, t.noff, &T.noff);
  double *dZ = nullptr, *dY = nullptr;
  if (!rc) {
    if (dmalloc(&dZ, sizeof(double) * s.dim * count) != hipSuccess ||
        dmalloc(&dY, sizeof(double) * t.dim * count) != hipSuccess)
      rc = fail(COSMO_ENOMEM, "label buffers");
  }
  if (!rc) {
    hipMemcpy(dZ, Z, sizeof(double) * s.dim * count, hipMemcpyHostToDevice);
    hipMemset(dY, 0, sizeof(double) * t.dim * count);
    hipLaunchKernelGGL(k_raise, dim3((unsigned)((s.dim + 255) / 256)), dim3(256), 0, 0, S, T, db, s.dim, dp, dc,
                       (int)s.orb.size(), dZ, count, dY, t.dim);
    std::vector<double> y((size_t)t.dim * count);
    if (hipMemcpy(y.data(), dY, sizeof(double) * y.size(), hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(COSMO_EHIP, "label kernel failed");
    for (int v = 0; v < count && !rc; ++v) {
      double acc = 0.0;
      for (int64_t i = 0; i < t.dim; ++i) acc += y[(size_t)v * t.dim + i] * y[(size_t)v * t.dim + i];
      out[v] = acc;
    }
  }
  dfree(dZ);
  dfree(dY);
  cleanup();
  return rc;
}

}  // namespace cosmo

using namespace cosmo;

extern "C" int cosmo_system_labels(cosmo_system s, const double *evecs, int count, double *jm_jp, double *tm_tp) {
  if (!s || !evecs || count <= 0) return fail(COSMO_EINVAL, "cosmo_system_labels: bad args");
  const int norb = (int)s->orb.size();
  if (jm_jp) {
    System t = *s;
    t.twoM = s->twoM + 2;
    untruncate(t);
    t.build_basis();
    std::vector<int16_t> partner(norb, -1);
    std::vector<double> coef(norb, 0.0);
    for (int g = 0; g < norb; ++g)
      for (int h = 0; h < norb; ++h) {
        const Orbital &a = s->orb[g], &b = s->orb[h];
        if (a.level == b.level && a.tz2 == b.tz2 && b.m2 == a.m2 + 2) {
          partner[g] = (int16_t)h;
          coef[g] = 0.5 * std::sqrt((double)(a.j2 - a.m2) * (double)(a.j2 + a.m2 + 2));
        }
      }
    int rc = raise_norms(*s, t, partner, coef, evecs, count, jm_jp);
    if (rc) return rc;
  }
  if (tm_tp) {
    System t = *s;
    untruncate(t);
    t.nv = s->nv + 1;
    t.zv = s->zv - 1;
    if (t.zv < 0) {
      for (int v = 0; v < count; ++v) tm_tp[v] = 0.0;
    } else {
      t.build_basis();
      std::vector<int16_t> partner(norb, -1);
      std::vector<double> coef(norb, 0.0);
      for (int g = 0; g < norb; ++g)
        for (int h = 0; h < norb; ++h) {
          const Orbital &a = s->orb[g], &b = s->orb[h];
          if (a.tz2 < 0 && b.tz2 > 0 && a.level == b.level && a.m2 == b.m2) {
            partner[g] = (int16_t)h;
            coef[g] = 1.0;
          }
        }
      int rc = raise_norms(*s, t, partner, coef, evecs, count, tm_tp);
      if (rc) return rc;
    }
  }
  return COSMO_OK;
}

// twice the spin from <J-J+> and twice its projection (include/sm/QN.cxx:40-55)
extern "C" int cosmo_twice_spin(double x, int twoM) {
  const double s2 = std::sqrt(4.0 * x + (double)twoM * (twoM + 2) + 1.0) - 1.0;
  // J + M must be an integer to within 0.1, else the spin is rejected as -1
  const double jm = 0.5 * (s2 + twoM);
  if (!(std::fabs(jm - std::round(jm)) < 0.1)) return -1;
  return (int)std::lround(s2);
}
