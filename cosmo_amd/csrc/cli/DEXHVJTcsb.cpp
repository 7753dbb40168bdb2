// DEXHVJTcsb <xml> [-i iters] [-n states] [-b block] [-f init.EE] [-o out.xml] [-l limit]
// Builds H on the GPU, diagonalises it (block Davidson when n > limit > 0, exact dense
// solve otherwise), writes <sys>.HH.EE and labels the states with J and T into the XML
// (<state> / <wavefunction> nodes).  Command line of davidson/DEXHVJTcsb.cpp:36-77, driver
// include/xsm/DEXHVcsb.cxx:32-113, labelling include/xsm/XSHLJT.cxx:38-142 and
// include/xsm/xState.cxx:113-150.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <sstream>

#include "cli_common.h"

static std::string spin_str(int twice) {
  if (twice % 2 == 0) return std::to_string(twice / 2);
  return std::to_string(twice) + "/2";
}

static std::string num(double x) {
  std::ostringstream o;
  o.precision(10);
  o << x;
  return o.str();
}

int main(int argc, char **argv) {
  Args a = parse_args(argc, argv, {{"i", "iterations"}, {"n", "number"}, {"b", "block"}, {"f", "file"},
                                   {"o", "output"}, {"l", "limit"}});
  if (a.help || a.pos.size() < 2) {
    fprintf(stderr,
            "Davidson or exact diagonalization of the Hamiltonian built on the GPU, followed by spin analysis\n"
            "usage: DEXHVJTcsb <system.xml> [-i iterations] [-n states] [-b block] [-f init.EE] [-o out.xml] "
            "[-l limit]\n");
    return a.help ? 0 : 1;
  }
  int iters = flag_int(a, "i", 0), nstate = flag_int(a, "n", 0), bloc = flag_int(a, "b", 10);
  int limit = flag_int(a, "l", 1000);
  if (bloc < nstate) bloc = nstate;
  const std::string xml = a.pos[1];
  cosmo::XNode doc;
  std::string err;
  if (!cosmo::xml_parse_file(xml, doc, err)) return fprintf(stderr, "Error: %s\n", err.c_str()), 1;
  cosmo::XNode *root = doc.child("cosmo");
  cosmo::XNode *sysn = root ? select_system(root) : nullptr;
  if (!sysn || !sysn->get("name")) return fprintf(stderr, "Error: no <system name=...>\n"), 1;
  const std::string sysname = sysn->get("name");
  cosmo_system s;
  if (cosmo_system_from_xml(xml.c_str(), sysname.c_str(), &s)) return die("system");
  std::string dir = xml.find('/') == std::string::npos ? "." : xml.substr(0, xml.find_last_of('/'));
  cosmo::XNode *mbs = sysn->child("ManyBodyStates");
  if (mbs && mbs->get("name")) {
    std::string p = mbs->get("name");
    if (p[0] != '/' && !std::ifstream(p)) p = dir + "/" + p;
    if (!std::ifstream(p)) {  // named but missing: generate it again
      if (cosmo_system_write_mbs(s, p.c_str())) return die("write .mbs");
    } else if (cosmo_system_check_mbs(s, p.c_str()))
      return die("many-body states");
  }
  const int64_t n = cosmo_system_dim(s);
  const std::string eef = sysname + ".HH.EE";
  fprintf(stderr, "generating %s.HH\ndimension %lld\n", sysname.c_str(), (long long)n);
  cosmo_op op;
  if (cosmo_op_create(s, 0, &op)) return die("Hamiltonian build");
  int count;
  std::vector<double> ev, de, rs, vec;
  if (n > limit && limit > 0) {
    if (bloc > n) bloc = (int)n;
    cosmo_dav_params p;
    cosmo_dav_defaults(&p, bloc);
    if (iters == -1) p.maxiter = 1 << 30;
    else if (iters > 0) p.maxiter = iters;
    std::vector<double> init;
    if (a.flags.count("f") && std::ifstream(a.flags["f"])) {
      init.assign((size_t)n * bloc, 0.0);
      int got = 0;
      if (cosmo_read_ee(a.flags["f"].c_str(), n, bloc, &got, nullptr, init.data())) return die("initial vectors");
      p.ninit = got;
    }
    fprintf(stderr, "Davidson: Block: %d Iterations: %d Eigenvalues: %d\n", bloc, p.maxiter, nstate ? nstate : bloc);
    ev.assign(bloc, 0.0);
    de.assign(bloc, 0.0);
    rs.assign(bloc, 0.0);
    vec.assign((size_t)n * bloc, 0.0);
    cosmo_dav_stats st;
    int rc = cosmo_davidson(op, &p, init.empty() ? nullptr : init.data(), ev.data(), de.data(), rs.data(),
                            vec.data(), &st);
    if (rc && rc != COSMO_ECONV) return die("davidson");
    // the reference writes its .EE whatever the residuals; so does this tool, with a warning
    if (rc == COSMO_ECONV) fprintf(stderr, "D: warning: %s\n", cosmo_last_error());
    count = nstate ? nstate : bloc;
    for (int i = 0; i < bloc; ++i) printf("%.6g %.6g %.6g\n", ev[i], de[i], rs[i]);
  } else {
    printf("Matrix is small for davidson. Using exact diagonalization\n");
    ev.assign(n, 0.0);
    vec.assign((size_t)n * n, 0.0);
    if (cosmo_op_dense_eigen(op, ev.data(), vec.data())) return die("exact diagonalization");
    de.assign(n, 0.0);
    rs.assign(n, 0.0);
    count = (int)n;
    for (int i = 0; i < (nstate ? std::min<int64_t>(nstate, n) : n); ++i) printf("%.6g ", ev[i]);
    printf("\n");
  }
  if (cosmo_write_ee(eef.c_str(), n, count, ev.data(), de.data(), rs.data(), vec.data())) return die("write .EE");
  // spin / isospin labels of the saved states
  int lstates = count;
  if (nstate > 0 && nstate < lstates) lstates = nstate;
  std::vector<double> jj(lstates), tt(lstates);
  if (cosmo_system_labels(s, vec.data(), lstates, jj.data(), tt.data())) return die("labels");
  int nparticles, nsps, nn, np;
  int64_t ne, nvx;
  cosmo_system_info(s, &nparticles, &nsps, &nn, &np, &ne, &nvx);
  int twoM = 0;
  {
    std::string jz = sysn->get("Jz") ? sysn->get("Jz") : "0";
    size_t sl = jz.find('/');
    twoM = sl == std::string::npos ? 2 * atoi(jz.c_str()) : atoi(jz.substr(0, sl).c_str());
  }
  int nv = sysn->get("Nv") ? atoi(sysn->get("Nv")) : 0, zv = sysn->get("Zv") ? atoi(sysn->get("Zv")) : 0;
  int twoTz = nv - zv;
  const char *par = sysn->get("P") ? sysn->get("P") : "+";
  for (int i = 0; i < lstates; ++i) {
    int J2 = cosmo_twice_spin(jj[i], twoM), T2 = cosmo_twice_spin(tt[i], twoTz);
    fprintf(stderr, "State: %d E= %.6f JJ= %.6f TT= %.6f\n", i, ev[i], jj[i], tt[i]);
    cosmo::XNode *st = nullptr;
    for (auto &k : sysn->kids)
      if (k.name == "state" && k.get("J") && k.get("T") && k.get("E") && spin_str(J2) == k.get("J") &&
          spin_str(T2) == k.get("T") && std::fabs(atof(k.get("E")) - ev[i]) < 1e-6)
        st = &k;
    if (!st) {
      cosmo::XNode k;
      k.name = "state";
      k.set("name", spin_str(J2));
      k.set("J", spin_str(J2));
      k.set("P", par[0] == '-' ? "-" : "+");
      k.set("T", spin_str(T2));
      k.set("E", num(ev[i]));
      sysn->kids.push_back(k);
      st = &sysn->kids.back();
    }
    bool have = false;
    for (auto &w : st->kids)
      if (w.name == "wavefunction" && w.get("file") && eef == w.get("file") && w.get("id") && atoi(w.get("id")) == i)
        have = true;
    if (!have) {
      cosmo::XNode w;
      w.name = "wavefunction";
      w.set("Jz", spin_str(twoM));
      w.set("Tz", spin_str(twoTz));
      w.set("file", eef);
      w.set("id", std::to_string(i));
      st->kids.push_back(w);
    }
    st->set("dE", num(de[i]));
  }
  // states ordered by energy after the other children
  std::stable_sort(sysn->kids.begin(), sysn->kids.end(), [](const cosmo::XNode &x, const cosmo::XNode &y) {
    bool sx = x.name == "state", sy = y.name == "state";
    if (sx != sy) return !sx;
    if (!sx) return false;
    return atof(x.get("E")) < atof(y.get("E"));
  });
  cosmo::XNode *mat = sysn->child("matrix");
  if (!mat) {
    cosmo::XNode m;
    m.name = "matrix";
    sysn->kids.insert(sysn->kids.begin(), m);
    mat = sysn->child("matrix");
  }
  mat->set("name", sysname + ".HH");
  mat->set("format", "pp");
  std::string out = a.flags.count("o") ? a.flags["o"] : xml;
  if (!cosmo::xml_write_file(out, doc)) return fprintf(stderr, "Error: cannot write %s\n", out.c_str()), 1;
  cosmo_op_destroy(op);
  cosmo_system_destroy(s);
  return 0;
}
