// davidson_vfile <file.HH | system.xml> [-i iters] [-n states] [-b block] [-f init.EE]
// Block Davidson on the GPU for a Hamiltonian stored in a .HH file; writes <file>.EE
// (include/kdav/cdavidson.cpp:327-368).  Command line of davidson/davidson_vfile.cpp:28-75:
// block >= states is forced, the lowest `block` pairs are computed and `states` of them
// (all when 0) saved.
#include "cli_common.h"

int main(int argc, char **argv) {
  Args a = parse_args(argc, argv, {{"i", "iterations"}, {"n", "number"}, {"b", "block"}, {"f", "file"}});
  if (a.help || a.pos.size() < 2) {
    fprintf(stderr,
            "Davidson diagonalization of a matrix stored in a file.\n"
            "usage: davidson_vfile <file.HH | system.xml> [-i iterations] [-n states] [-b block] [-f init.EE]\n");
    return a.help ? 0 : 1;
  }
  std::string sysfile = a.pos[1];
  if (sysfile.back() != 'H') {  // an XML file: take the last <matrix name=...>
    cosmo::XNode doc;
    std::string err;
    if (!cosmo::xml_parse_file(sysfile, doc, err)) return fprintf(stderr, "Error: %s\n", err.c_str()), 1;
    cosmo::XNode *root = doc.child("cosmo");
    cosmo::XNode *sysn = root ? select_system(root) : nullptr;
    cosmo::XNode *mat = sysn ? sysn->child("matrix", -1) : nullptr;
    if (!mat || !mat->get("name")) return fprintf(stderr, "Error: You need Hamiltonian matrix\n"), 1;
    sysfile = mat->get("name");
    fprintf(stderr, "Diagonalizing file %s\n", sysfile.c_str());
  }
  int iters = flag_int(a, "i", 0), nstate = flag_int(a, "n", 0), bloc = flag_int(a, "b", 10);
  if (bloc < nstate) bloc = nstate;
  fprintf(stderr, "Work with: %s\n", sysfile.c_str());
  cosmo_op op;
  if (cosmo_op_read_hh(sysfile.c_str(), 0, &op)) return die("read .HH");
  const int64_t n = cosmo_op_size(op);
  if (bloc > n) bloc = (int)n;
  cosmo_dav_params p;
  cosmo_dav_defaults(&p, bloc);
  if (iters == -1) p.maxiter = 1 << 30;
  else if (iters > 0) p.maxiter = iters;
  std::vector<double> init;
  if (a.flags.count("f")) {
    init.assign((size_t)n * bloc, 0.0);
    int got = 0;
    if (cosmo_read_ee(a.flags["f"].c_str(), n, bloc, &got, nullptr, init.data())) return die("initial vectors");
    p.ninit = got;
    fprintf(stderr, "Number of input vectors %d\n", got);
  }
  fprintf(stderr, "Size: %lld\nDavidson: Block: %d Iterations: %d Eigenvalues: %d\n", (long long)n, bloc, p.maxiter,
          nstate ? nstate : bloc);
  std::vector<double> ev(bloc), de(bloc), rs(bloc), vec((size_t)n * bloc);
  cosmo_dav_stats st;
  int rc = cosmo_davidson(op, &p, init.empty() ? nullptr : init.data(), ev.data(), de.data(), rs.data(), vec.data(), &st);
  if (rc && rc != COSMO_ECONV) return die("davidson");
  // the reference writes its .EE whatever the residuals; so does this tool, with a warning
  if (rc == COSMO_ECONV) fprintf(stderr, "D: warning: %s\n", cosmo_last_error());
  const std::string outf = sysfile + ".EE";
  const int save = nstate ? nstate : bloc;
  fprintf(stderr, " writing output to %s\n", outf.c_str());
  if (cosmo_write_ee(outf.c_str(), n, save, ev.data(), de.data(), rs.data(), vec.data())) return die("write .EE");
  for (int i = 0; i < bloc; ++i) printf("%.6g %.6g %.6g\n", ev[i], de[i], rs[i]);
  fprintf(stderr, "my lowest eigenvalue is %f\n", ev[0]);
  cosmo_op_destroy(op);
  return 0;
}
