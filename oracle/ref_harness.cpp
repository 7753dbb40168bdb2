// Harness that drives the reference's own CSB branch (include/xsm/DEXHVcsb.cxx:60-71)
// without its Eigen-only dense branch.  Built by oracle/Makefile against the reference
// headers where they lie; output goes to oracle/_ref/ only.  Test infrastructure: it is
// the checker / CPU baseline, never the product.
//
// Modes:
//   dumpint <sys.xml> <out.bin>            m-scheme ESP and VSPX tensors after xReadSPInt
//   csb     <sys.xml> <out.EE> <bloc> <iters> <nstate>
//           builds SparseMatrixCSB via PPMatrix2CSBBlk, times it, runs the reference
//           Davidson on it (DavidsonDiagonalization), prints timings as JSON on stdout
//   csbwrite <sys.xml> <out.csb>
//           builds CSB and writes it with SparseMatrixCSB::Write
//   csbtimes <file.csb> <m> <reps>
//           reads a CSB file with SparseMatrixCSB::Read and times TimesMatrix on m vectors
//   spmm    <sys.xml> <m> <reps> <out.bin>
//           builds CSB, applies TimesMatrix to m deterministic vectors reps times,
//           writes n, m, X, Y and prints the mean time per call as JSON
#include <cstdint>
typedef uint16_t spsint;
typedef unsigned long long uint_nbasis;
#include <av/argparser.cxx>
#include <av/statusclass.cxx>
#include <xsm/xManyBodyStates.cxx>
#include <xsm/xsphericalint.cxx>
#include <sm/MakeSPO.cxx>
#include <sm/spherical2sp.cxx>
#include <tav/sparse.cxx>
#include <tav/csbmatrix.cxx>
#include <xsm/xReadSPInt.cxx>
#include <sm/PPMatrix.cxx>
#include <kdav/davidson.cxx>
#include <chrono>
#include <cstdio>
#include <vector>

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Setup {
  SM::SingleParticleStates sps;
  SM::Many_Body_States st;
  tav::STensor<spsint, double> ESP{2};
  tav::STensor<spsint, double> VSPX{4};
  XMLNode main, sys;
};

static void load(Setup &s, const char *xml) {
  s.main = XMLNode::openFileHelper(xml, xsm::cosmo);
  SM::Make(s.sps, s.main.getChildNode(xsm::valencespace));
  s.sys = xml::SelectNode(s.main, xsm::system, "name");
  SM::Get(s.st, s.main, s.sys);
  xReadSPInt(s.ESP, s.VSPX, s.sys, s.main, s.sps);
}

static void dump_tensor(FILE *f, tav::STensor<spsint, double> &T, uint32_t rank) {
  uint64_t cnt = T.size();
  fwrite(&rank, 4, 1, f);
  fwrite(&cnt, 8, 1, f);
  for (auto it = T.begin(); it != T.end(); ++it) {
    for (uint32_t r = 0; r < rank; ++r) { uint16_t k = it->first[r]; fwrite(&k, 2, 1, f); }
    double v = it->second;
    fwrite(&v, 8, 1, f);
  }
}

// csbtimes <file.csb> <m> <reps>: the reference's SparseMatrixCSB::Read of a CSB file,
// then its TimesMatrix (csbmatrix.cxx:623-675) on m deterministic vectors, timed
static int csbtimes(const char *path, int m, int reps) {
  tav::SparseMatrixCSB<double, uint32_t, uint16_t> FHH(1, 1, true);
  double t0 = now_s();
  FHH.Read(path);
  const double t_read = now_s() - t0;
  const uint64_t n = FHH.n;
  uint64_t nnz = 0;
  for (int i = 0; i < FHH.numBlockRow; ++i)
    for (int j = 0; j < FHH.numBlockCol; ++j) nnz += FHH.Blocks[i][j].nnz;
  std::vector<double> X(n * m), Y(n * m);
  for (int v = 0; v < m; ++v)
    for (uint64_t i = 0; i < n; ++i) X[v * n + i] = std::sin(0.37 * double(i) + 1.3 * double(v));
  FHH.TimesMatrix(Y.data(), X.data(), m);  // warm-up
  t0 = now_s();
  for (int r = 0; r < reps; ++r) FHH.TimesMatrix(Y.data(), X.data(), m);
  const double t_mv = (now_s() - t0) / reps;
  printf("{\"n\": %llu, \"nnz_upper\": %llu, \"t_read\": %.6f, \"t_spmm\": %.6f, \"m\": %d, \"reps\": %d}\n",
         (unsigned long long)n, (unsigned long long)nnz, t_read, t_mv, m, reps);
  return 0;
}

// hhtimes <file.HH> <m> <out.bin>: the reference's .HH reader and its symmetric row product
// (SparseMatrix<double>::Read with type = 1 as davidson_vfile sets it, TimesMatrix on
// pointer arrays, include/tav/sparse.cxx:273-291,401-445) on m deterministic vectors
static int hhtimes(const char *path, int m, const char *out) {
  tav::SparseMatrix<double> FHH;
  if (!FHH.Read(path)) return 1;
  FHH.type = 1;
  const uint64_t n = FHH.n;
  std::vector<double> X(n * m), Y(n * m);
  for (int v = 0; v < m; ++v)
    for (uint64_t i = 0; i < n; ++i) X[v * n + i] = std::sin(0.37 * double(i) + 1.3 * double(v));
  std::vector<double *> xp(m), yp(m);
  for (int v = 0; v < m; ++v) xp[v] = X.data() + v * n, yp[v] = Y.data() + v * n;
  FHH.TimesMatrix(yp.data(), xp.data(), (unsigned)m);
  FILE *f = fopen(out, "wb");
  uint64_t mm = m;
  fwrite(&n, 8, 1, f); fwrite(&mm, 8, 1, f);
  fwrite(Y.data(), 8, n * m, f);
  fclose(f);
  printf("{\"n\": %llu, \"m\": %d}\n", (unsigned long long)n, m);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: ref_harness dumpint|csb|csbwrite|spmm|csbtimes|hhtimes|davtraj <file> ...\n");
    return 2;
  }
  std::string mode = argv[1];
  if (mode == "csbtimes") return argc == 5 ? csbtimes(argv[2], atoi(argv[3]), atoi(argv[4])) : 2;
  if (mode == "hhtimes") return argc == 5 ? hhtimes(argv[2], atoi(argv[3]), argv[4]) : 2;
  Setup s;
  load(s, argv[2]);
  if (mode == "dumpint") {
    FILE *f = fopen(argv[3], "wb");
    dump_tensor(f, s.ESP, 2);
    dump_tensor(f, s.VSPX, 4);
    fclose(f);
    printf("{\"n\": %llu, \"esp\": %zu, \"vspx\": %zu}\n", (unsigned long long)s.st.n, s.ESP.size(), s.VSPX.size());
    return 0;
  }
  SM::PPMatrix HH(s.st);
  HH.push_back(&s.ESP);
  HH.push_back(&s.VSPX);
  uint16_t br = std::sqrt(1. * s.st.n) + 1;
  tav::SparseMatrixCSB<double, uint32_t, uint16_t> FHH(s.st.n, br, br, true);
  double t0 = now_s();
  SM::PPMatrix2CSBBlk(FHH, HH);
  double t_build = now_s() - t0;
  uint64_t nnz = 0;
  for (int i = 0; i < FHH.numBlockRow; ++i)
    for (int j = 0; j < FHH.numBlockCol; ++j) nnz += FHH.Blocks[i][j].nnz;
  if (mode == "csbwrite") {  // the reference's own CSB file writer (csbmatrix.cxx:553-576)
    FHH.Write(argv[3]);
    printf("{\"n\": %llu, \"nnz_upper\": %llu, \"block_rows\": %d}\n", (unsigned long long)s.st.n,
           (unsigned long long)nnz, (int)br);
    return 0;
  }
  if (mode == "csb") {
    int bloc = atoi(argv[4]), iters = atoi(argv[5]), nstate = atoi(argv[6]);
    if (bloc < nstate) bloc = nstate;
    t0 = now_s();
    DavidsonDiagonalization(argv[3], FHH, bloc, iters, nstate);
    double t_dav = now_s() - t0;
    printf("{\"n\": %llu, \"nnz_upper\": %llu, \"block_rows\": %d, \"t_build\": %.6f, \"t_davidson\": %.6f}\n",
           (unsigned long long)s.st.n, (unsigned long long)nnz, (int)br, t_build, t_dav);
    return 0;
  }
  if (mode == "davtraj") {
    // davtraj <sys.xml> <bloc> <iters> <nstate>: the reference Davidson on the CSB matrix,
    // set up as DavidsonDiagonalization sets it up (include/kdav/davidson.cxx:167-223:
    // numemax = mblock = bloc, crite 1e-15, critc 1e-12, critr 1e-8, ortho 1e-9,
    // maxiter max(40 bloc, 200) when 0 is given, lim = bloc + 20), then its own matrix-
    // vector count and error flag next to every pair's eigenvalue, |dE| and residual
    int bloc = atoi(argv[3]), iters = atoi(argv[4]), nstate = atoi(argv[5]);
    if (bloc < nstate) bloc = nstate;
    MyMatrixDiagonal<tav::SparseMatrixCSB<double, uint32_t, uint16_t>> *diag =
        new MyMatrixDiagonal<tav::SparseMatrixCSB<double, uint32_t, uint16_t>>(FHH);
    Davidson *dav = new Davidson(diag, bloc, 1);
    dav->crite = 1e-15; dav->critc = 1e-12; dav->critr = 1e-8; dav->ortho = 1e-9;
    dav->silence = 1;
    if (!iters) iters = MAX(dav->GetNumeMax() * 40, 200);
    dav->maxiter = iters;
    dav->SetLimMax(dav->GetNumeMax() + 20);
    dav->mblock = bloc;
    dav->Calculate();
    printf("{\"n\": %llu, \"bloc\": %d, \"maxiter\": %d, \"nmv\": %lld, \"ierr\": %lld, \"pairs\": [",
           (unsigned long long)s.st.n, bloc, iters, (long long)dav->GetMatrixVectProd(), (long long)dav->GetIerr());
    for (int i = 0; i < bloc; ++i) {
      eig_val e;
      dav->EigenValue(e, i);
      printf("%s[%.17g, %.17g, %.17g]", i ? ", " : "", e.Eigenvalue, e.Eigval_differences, e.Residual);
    }
    printf("]}\n");
    return 0;
  }
  if (mode == "spmm") {
    int m = atoi(argv[3]), reps = atoi(argv[4]);
    uint64_t n = s.st.n;
    std::vector<double> X(n * m), Y(n * m);
    // deterministic input: x[v][i] = sin(0.37*i + 1.3*v) (both sides compute the same doubles)
    for (int v = 0; v < m; ++v)
      for (uint64_t i = 0; i < n; ++i) X[v * n + i] = std::sin(0.37 * double(i) + 1.3 * double(v));
    FHH.TimesMatrix(Y.data(), X.data(), m);  // warm-up
    t0 = now_s();
    for (int r = 0; r < reps; ++r) FHH.TimesMatrix(Y.data(), X.data(), m);
    double t_mv = (now_s() - t0) / reps;
    FILE *f = fopen(argv[5], "wb");
    uint64_t mm = m;
    fwrite(&n, 8, 1, f); fwrite(&mm, 8, 1, f);
    fwrite(X.data(), 8, n * m, f); fwrite(Y.data(), 8, n * m, f);
    fclose(f);
    printf("{\"n\": %llu, \"nnz_upper\": %llu, \"block_rows\": %d, \"t_build\": %.6f, \"t_spmm\": %.6f, \"m\": %d}\n",
           (unsigned long long)n, (unsigned long long)nnz, (int)br, t_build, t_mv, m);
    return 0;
  }
  fprintf(stderr, "unknown mode %s\n", mode.c_str());
  return 2;
}
