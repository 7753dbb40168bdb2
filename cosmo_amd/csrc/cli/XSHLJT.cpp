// XSHLJT <xml> [-o out.xml] [-n states]
// Spin and isospin analysis of the states saved in <matrix name>.EE and their entry in the
// XML state database: command line of xmlshell/XSHLJT.cpp:34-53, analysis of
// include/xsm/XSHLJT.cxx:38-142 (J / T filters of the system, JJ / TT shifts of the matrix
// removed from the energies), <J-J+> and <T-T+> on the GPU.  Up to -n states (default 10);
// when -n does not cut the file's count and it holds more than 20 states the reference
// asks how many to process: here all are processed.
#include <cstdint>
#include <cstring>
#include <fstream>

#include "cli_common.h"
#include "states_xml.h"

// the .EE records [u32 count] then per state [f64 E][f64 dE][f64 resid][u64 n][n f64]
// (include/kdav/cdavidson.cpp:327-368), up to max states of dimension n
static bool read_ee_records(const std::string &path, int64_t n, int max, std::vector<double> &E,
                            std::vector<double> &dE, std::vector<double> &V, int &found) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return fprintf(stderr, "Error: cannot open %s\n", path.c_str()), false;
  uint32_t c = 0;
  if (!f.read((char *)&c, 4)) return fprintf(stderr, "Error: truncated %s\n", path.c_str()), false;
  found = (int)c;
  const int k = std::min<int>((int)c, max);
  E.assign(k, 0.0);
  dE.assign(k, 0.0);
  V.assign((size_t)n * k, 0.0);
  for (int i = 0; i < k; ++i) {
    double rec[3];
    uint64_t len = 0;
    if (!f.read((char *)rec, sizeof rec) || !f.read((char *)&len, 8))
      return fprintf(stderr, "Error: truncated %s\n", path.c_str()), false;
    if ((int64_t)len != n)
      return fprintf(stderr, "Error: %s holds vectors of length %llu, the basis has %lld\n", path.c_str(),
                     (unsigned long long)len, (long long)n),
             false;
    E[i] = rec[0];
    dE[i] = rec[1];
    if (!f.read((char *)(V.data() + (size_t)n * i), sizeof(double) * n))
      return fprintf(stderr, "Error: truncated %s\n", path.c_str()), false;
  }
  return true;
}

int main(int argc, char **argv) {
  Args a = parse_args(argc, argv, {{"o", "output"}, {"n", "number"}});
  if (a.help || a.pos.size() < 2) {
    fprintf(stderr,
            "Determine spins and isospins of states and populate xml datables\n"
            "usage: XSHLJT <system.xml> [-o out.xml] [-n states]\n");
    return a.help ? 0 : 1;
  }
  const std::string xml = a.pos[1];
  const int lstates = flag_int(a, "n", 10);
  cosmo::XNode doc;
  std::string err;
  if (!cosmo::xml_parse_file(xml, doc, err)) return fprintf(stderr, "Error: %s\n", err.c_str()), 1;
  cosmo::XNode *root = doc.child("cosmo");
  cosmo::XNode *sysn = root ? select_system(root) : nullptr;
  if (!sysn || !sysn->get("name")) return fprintf(stderr, "Error: no <system name=...>\n"), 1;
  cosmo::XNode *mat = sysn->child("matrix", -1);
  if (!mat) return fprintf(stderr, "Error: You need Hamiltonian matrix\n"), 1;
  if (!mat->get("name")) return fprintf(stderr, "Error: the <matrix> node has no name attribute\n"), 1;
  cosmo_system s;
  if (cosmo_system_from_xml(xml.c_str(), sysn->get("name"), &s)) return die("system");
  const int64_t n = cosmo_system_dim(s);
  const std::string eef = std::string(mat->get("name")) + ".EE";
  printf("reading file %s\n", eef.c_str());
  std::vector<double> E, dE, V;
  int found = 0;
  // the count first, then as many records as will be processed
  if (!read_ee_records(eef, n, 0, E, dE, V, found)) return 1;
  fprintf(stderr, "states found %d\n", found);
  int count = found;
  if (lstates > 0 && lstates < count) count = lstates;
  else if (count > 20)
    fprintf(stderr, "There are %d states; all are processed (this tool asks no questions)\n", count);
  fprintf(stderr, "processing states %d\n", count);
  if (!read_ee_records(eef, n, count, E, dE, V, found)) return 1;
  if (label_states(s, sysn, eef, count, E.data(), dE.data(), V.data())) return 1;
  const std::string out = a.flags.count("o") ? a.flags["o"] : xml;
  if (!cosmo::xml_write_file(out, doc)) return fprintf(stderr, "Error: cannot write %s\n", out.c_str()), 1;
  cosmo_system_destroy(s);
  return 0;
}
