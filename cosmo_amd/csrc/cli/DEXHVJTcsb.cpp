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
#include "states_xml.h"

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
  // the matrix node the labelling reads (the reference expects it in the input)
  cosmo::XNode *mat = sysn->child("matrix", -1);
  if (!mat) {
    cosmo::XNode m;
    m.name = "matrix";
    sysn->kids.insert(sysn->kids.begin(), m);
    mat = sysn->child("matrix", -1);
  }
  mat->set("name", sysname + ".HH");
  mat->set("format", "pp");
  // spin / isospin labels, J / T filters, JJ / TT shifts and the state nodes (XSHLJT over
  // the saved states)
  int lstates = count;
  if (nstate > 0 && nstate < lstates) lstates = nstate;
  if (label_states(s, sysn, eef, lstates, ev.data(), de.data(), vec.data())) return 1;
  std::string out = a.flags.count("o") ? a.flags["o"] : xml;
  if (!cosmo::xml_write_file(out, doc)) return fprintf(stderr, "Error: cannot write %s\n", out.c_str()), 1;
  cosmo_op_destroy(op);
  cosmo_system_destroy(s);
  return 0;
}
