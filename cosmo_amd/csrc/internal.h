// internal.h -- shared host-side definitions of the cosmo_hip library.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/cosmo_hip.h"

namespace cosmo {

// error state (thread-local message, set by every failing entry point)
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

// m-scheme orbital: neutrons first (tz2 = +1), then protons (tz2 = -1)
struct Orbital {
  int level, j2, m2, tz2, n, l, par;
};

// One XML element (the tiny DOM used for CoSMo system files).
struct XNode {
  std::string name;                                      // "" for text, "?" for <?..?>
  std::vector<std::pair<std::string, std::string>> attr;  // in file order
  std::string text;
  std::vector<XNode> kids;
  bool decl = false;  // written as <?name ...?> (CoSMo's ManyBodyStates node)
  const char *get(const std::string &k) const;
  void set(const std::string &k, const std::string &v);
  XNode *child(const std::string &n, int which = 0);     // which < 0: last
  int count(const std::string &n) const;
};
bool xml_parse_file(const std::string &path, XNode &root, std::string &err);
bool xml_write_file(const std::string &path, const XNode &root);

struct System {
  std::vector<Orbital> orb;
  int nn = 0, np = 0;        // neutron / proton orbital counts
  int nv = 0, zv = 0;        // valence neutrons / protons
  int twoM = 0, parity = 0;  // 2*Jz, 0 '+' / 1 '-'
  std::string name;
  // m-scheme operators, reference key order
  std::vector<uint16_t> esp_k;  // 2 per entry
  std::vector<double> esp_v;
  std::vector<uint16_t> vspx_k;  // 4 per entry
  std::vector<double> vspx_v;
  // basis structure: neutron configurations in lexicographic order, each with the
  // proton sector matching (M - M_n, P ^ P_n); proton configurations grouped by sector
  std::vector<uint32_t> nmask;     // neutron configs with a non-empty proton sector
  std::vector<int64_t> noff;       // basis offset of each (size nmask+1)
  std::vector<int32_t> nsector;    // proton sector of each
  std::vector<uint32_t> pmask;     // proton configs, sector-major, lex within a sector
  std::vector<int64_t> psec_start; // sector -> first index in pmask (size nsec+1)
  std::vector<int32_t> pcfg_rank;  // colex(pmask) -> rank within its sector (-1 none)
  std::vector<int32_t> ncfg_index; // colex(nmask) -> index into nmask (-1 none)
  int64_t dim = 0;
  // basis truncation of XsysmbsOMP (xmlshell/xmbs/XsysmbsOMP.cpp:212-303): oscillator
  // quanta and occupation rejections.  When any is set, a block holds only its kept proton
  // configurations: krank[kstart[b] .. kstart[b+1]) are their ranks within the sector,
  // ascending (kstart empty: the full M-scheme basis)
  struct Rejection {
    std::vector<int> nlev, plev;  // neutron / proton levels summed (orbital index - 1)
    int nmin = 0, nmax = 0;
  };
  bool hw_cut = false;
  int hwmin = 0, hwmax = 0;
  std::vector<Rejection> rejections;
  std::vector<int32_t> krank;
  std::vector<int64_t> kstart;
  bool truncated() const { return !kstart.empty(); }
  // sector rank of the proton configuration of basis row `row` in block b
  int64_t prow(int b, int64_t row) const {
    return truncated() ? krank[kstart[b] + (row - noff[b])] : row - noff[b];
  }
  void build_basis();
};

double clebsch_gordan(int j1, int m1, int j2, int m2, int J, int M);
uint64_t binom(int n, int k);
uint64_t colex_rank(uint64_t mask);

}  // namespace cosmo

struct cosmo_system_s : cosmo::System {};
