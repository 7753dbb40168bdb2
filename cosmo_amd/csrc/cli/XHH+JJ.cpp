// XHH+JJ <xml> [-o out.xml] : build the many-body Hamiltonian of the XML system on the GPU
// and write it as <sys>.HH (row format of include/sm/PPMatrix.cxx:413-465); records
// <matrix name="<sys>.HH" format="pp"> in the XML (include/xsm/xReadSPInt.cxx:39-52).
// Command line of xmlshell/XHH+JJ.cpp:39-43.
#include <chrono>
#include <fstream>

#include "cli_common.h"

int main(int argc, char **argv) {
  Args a = parse_args(argc, argv, {{"o", "output"}});
  if (a.help || a.pos.size() < 2) {
    fprintf(stderr, "Create a Hamiltonian matrix and save to file.\nusage: XHH+JJ <system.xml> [-o output.xml]\n");
    return a.help ? 0 : 1;
  }
  const std::string xml = a.pos[1];
  cosmo::XNode doc;
  std::string err;
  if (!cosmo::xml_parse_file(xml, doc, err)) {
    fprintf(stderr, "Error: %s\n", err.c_str());
    return 1;
  }
  cosmo::XNode *root = doc.child("cosmo");
  if (!root) return fprintf(stderr, "Error: no <cosmo> root\n"), 1;
  cosmo::XNode *sysn = select_system(root);
  if (!sysn || !sysn->get("name")) return fprintf(stderr, "Error: no <system name=...>\n"), 1;
  const std::string sysname = sysn->get("name");
  cosmo_system s;
  if (cosmo_system_from_xml(xml.c_str(), sysname.c_str(), &s)) return die("system");
  // basis: an existing .mbs must be the full basis; otherwise write one (as MakeMBS2 does)
  cosmo::XNode *mbs = sysn->child("ManyBodyStates");
  std::string dir = xml.find('/') == std::string::npos ? "." : xml.substr(0, xml.find_last_of('/'));
  if (mbs && mbs->get("name")) {
    std::string p = mbs->get("name");
    if (p[0] != '/' && !std::ifstream(p)) p = dir + "/" + p;
    if (!std::ifstream(p)) {  // named but missing: generate it again
      if (cosmo_system_write_mbs(s, p.c_str())) return die("write .mbs");
    } else if (cosmo_system_check_mbs(s, p.c_str()))
      return die("many-body states");
  } else {
    std::string p = sysname + ".mbs";
    if (cosmo_system_write_mbs(s, p.c_str())) return die("write .mbs");
    cosmo::XNode n;
    n.name = "ManyBodyStates";
    n.decl = true;
    n.set("name", p);
    n.set("size", std::to_string(cosmo_system_dim(s)));
    n.set("creator", "MakeMBS2");
    sysn->kids.insert(sysn->kids.begin(), n);
  }
  const std::string hname = sysname + ".HH";
  fprintf(stderr, "generating Hamiltonian %s\ndimension %lld\n", hname.c_str(), (long long)cosmo_system_dim(s));
  auto t0 = std::chrono::steady_clock::now();
  cosmo_op op;
  if (cosmo_op_create(s, COSMO_OP_KEEP_SMALL | COSMO_OP_NO_FP32, &op)) return die("Hamiltonian build");
  if (cosmo_op_write_hh(op, hname.c_str())) return die("write .HH");
  double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  int64_t nu = 0;
  cosmo_op_nnz(op, &nu, nullptr);
  fprintf(stderr, "WH: %lld upper entries in %.2f sec\n", (long long)nu, dt);
  cosmo::XNode *mat = sysn->child("matrix");
  if (!mat) {
    cosmo::XNode m;
    m.name = "matrix";
    sysn->kids.push_back(m);
    mat = &sysn->kids.back();
  }
  mat->set("name", hname);
  mat->set("format", "pp");
  std::string out = a.flags.count("o") ? a.flags["o"] : xml;
  if (!cosmo::xml_write_file(out, doc)) return fprintf(stderr, "Error: cannot write %s\n", out.c_str()), 1;
  cosmo_op_destroy(op);
  cosmo_system_destroy(s);
  return 0;
}
