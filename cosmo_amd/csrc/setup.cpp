// setup.cpp -- host-side model-space setup for the cosmo_hip library:
//   CoSMo XML system files, Oxbash interaction files, the m-scheme orbital table,
//   J-coupled -> m-scheme conversion of the interaction, and the M-scheme basis.
// Everything here is cheap (milliseconds); the Hamiltonian itself is built on the GPU.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

#include "internal.h"

namespace cosmo {

static thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }
int fail(int code, const std::string &m) {
  g_err = m;
  return code;
}

// ------------------------------------------------------------------------------------
// minimal XML DOM
const char *XNode::get(const std::string &k) const {
  for (auto &a : attr)
    if (a.first == k) return a.second.c_str();
  return nullptr;
}
void XNode::set(const std::string &k, const std::string &v) {
  for (auto &a : attr)
    if (a.first == k) {
      a.second = v;
      return;
    }
  attr.emplace_back(k, v);
}
XNode *XNode::child(const std::string &n, int which) {
  XNode *last = nullptr;
  int c = 0;
  for (auto &k : kids)
    if (k.name == n) {
      if (c == which) return &k;
      last = &k;
      ++c;
    }
  return which < 0 ? last : nullptr;
}
int XNode::count(const std::string &n) const {
  int c = 0;
  for (auto &k : kids) c += (k.name == n);
  return c;
}

static std::string unescape(const std::string &s) {
  std::string o;
  for (size_t i = 0; i < s.size(); ++i) {
    if (s[i] == '&') {
      size_t e = s.find(';', i);
      if (e != std::string::npos) {
        std::string ent = s.substr(i + 1, e - i - 1);
        if (ent == "lt") o += '<';
        else if (ent == "gt") o += '>';
        else if (ent == "amp") o += '&';
        else if (ent == "quot") o += '"';
        else if (ent == "apos") o += '\'';
        else o += s.substr(i, e - i + 1);
        i = e;
        continue;
      }
    }
    o += s[i];
  }
  return o;
}
static std::string escape(const std::string &s) {
  std::string o;
  for (char c : s) {
    if (c == '<') o += "&lt;";
    else if (c == '>') o += "&gt;";
    else if (c == '&') o += "&amp;";
    else if (c == '"') o += "&quot;";
    else o += c;
  }
  return o;
}

namespace {
struct XParser {
  const std::string &s;
  size_t p = 0;
  std::string err;
  explicit XParser(const std::string &str) : s(str) {}
  void ws() {
    while (p < s.size() && isspace((unsigned char)s[p])) ++p;
  }
  std::string name() {
    size_t b = p;
    while (p < s.size() && (isalnum((unsigned char)s[p]) || strchr("_-.:+", s[p]))) ++p;
    return s.substr(b, p - b);
  }
  bool attrs(XNode &n, const char *end) {
    for (;;) {
      ws();
      if (p >= s.size()) return false;
      if (s.compare(p, strlen(end), end) == 0) return true;
      if (s[p] == '/' && strcmp(end, ">") == 0) return true;
      std::string k = name();
      if (k.empty()) return false;
      ws();
      if (p >= s.size() || s[p] != '=') {
        n.attr.emplace_back(k, "");
        continue;
      }
      ++p;
      ws();
      char q = s[p];
      if (q != '"' && q != '\'') return false;
      size_t e = s.find(q, p + 1);
      if (e == std::string::npos) return false;
      n.attr.emplace_back(k, unescape(s.substr(p + 1, e - p - 1)));
      p = e + 1;
    }
  }
  // parse children of 'parent' until </parent.name> (or EOF for the document node)
  bool content(XNode &parent) {
    for (;;) {
      size_t lt = s.find('<', p);
      std::string text = s.substr(p, (lt == std::string::npos ? s.size() : lt) - p);
      if (text.find_first_not_of(" \t\r\n") != std::string::npos) {
        XNode t;
        t.text = unescape(text);
        parent.kids.push_back(t);
      }
      if (lt == std::string::npos) {
        p = s.size();
        return parent.name.empty();
      }
      p = lt;
      if (s.compare(p, 4, "<!--") == 0) {
        size_t e = s.find("-->", p);
        if (e == std::string::npos) return false;
        p = e + 3;
        continue;
      }
      if (s.compare(p, 9, "<!DOCTYPE") == 0 || s.compare(p, 9, "<![CDATA[") == 0) {
        size_t e = s.find('>', p);
        if (e == std::string::npos) return false;
        p = e + 1;
        continue;
      }
      if (s.compare(p, 2, "<?") == 0) {
        p += 2;
        XNode n;
        n.name = name();
        n.decl = true;
        if (!attrs(n, "?>")) return false;
        p += 2;
        if (n.name != "xml") parent.kids.push_back(n);
        continue;
      }
      if (s.compare(p, 2, "</") == 0) {
        p += 2;
        std::string nm = name();
        ws();
        if (p < s.size() && s[p] == '>') ++p;
        if (nm != parent.name) {
          err = "mismatched </" + nm + "> in <" + parent.name + ">";
          return false;
        }
        return true;
      }
      ++p;
      XNode n;
      n.name = name();
      if (n.name.empty()) return false;
      if (!attrs(n, ">")) return false;
      if (s[p] == '/') {
        p += 2;
        parent.kids.push_back(n);
        continue;
      }
      ++p;
      if (!content(n)) return false;
      parent.kids.push_back(n);
    }
  }
};
}  // namespace

bool xml_parse_file(const std::string &path, XNode &root, std::string &err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  std::string s = ss.str();
  XParser P(s);
  root = XNode();
  if (!P.content(root)) {
    err = "malformed XML in " + path + (P.err.empty() ? "" : ": " + P.err);
    return false;
  }
  return true;
}

static void xml_emit(std::ostream &o, const XNode &n, int depth) {
  std::string ind(depth, '\t');
  if (n.name.empty()) {
    o << ind << escape(n.text) << "\n";
    return;
  }
  o << ind << (n.decl ? "<?" : "<") << n.name;
  for (auto &a : n.attr) o << " " << a.first << "=\"" << escape(a.second) << "\"";
  if (n.decl) {
    o << "?>\n";
    return;
  }
  if (n.kids.empty()) {
    o << "/>\n";
    return;
  }
  if (n.kids.size() == 1 && n.kids[0].name.empty()) {
    o << ">" << escape(n.kids[0].text) << "</" << n.name << ">\n";
    return;
  }
  o << ">\n";
  for (auto &k : n.kids) xml_emit(o, k, depth + 1);
  o << ind << "</" << n.name << ">\n";
}

bool xml_write_file(const std::string &path, const XNode &root) {
  std::ofstream o(path, std::ios::binary);
  if (!o) return false;
  o << "<?xml version=\"1.0\" encoding=\"ISO-8859-1\"?>\n";
  for (auto &k : root.kids) xml_emit(o, k, 0);
  return (bool)o;
}

// ------------------------------------------------------------------------------------
// angular momentum algebra
uint64_t binom(int n, int k) {
  if (k < 0 || n < 0 || k > n) return 0;
  uint64_t r = 1;
  for (int i = 1; i <= k; ++i) r = r * (uint64_t)(n - k + i) / (uint64_t)i;
  return r;
}

uint64_t colex_rank(uint64_t mask) {
  uint64_t r = 0;
  int i = 1;
  while (mask) {
    int p = __builtin_ctzll(mask);
    r += binom(p, i);
    ++i;
    mask &= mask - 1;
  }
  return r;
}

// <j1 m1 j2 m2 | J M>, doubled arguments, Racah's closed form in long double.
double clebsch_gordan(int j1, int m1, int j2, int m2, int J, int M) {
  if (m1 + m2 != M || J < std::abs(j1 - j2) || J > j1 + j2 || ((j1 + j2 + J) & 1)) return 0.0;
  if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(M) > J) return 0.0;
  if (((j1 + m1) & 1) || ((j2 + m2) & 1) || ((J + M) & 1)) return 0.0;
  auto f = [](int n) {
    long double r = 1.0L;
    for (int i = 2; i <= n; ++i) r *= i;
    return r;
  };
  const int a = (j1 + j2 - J) / 2, b = (j1 - j2 + J) / 2, c = (J + j2 - j1) / 2;
  long double pre = (long double)(J + 1) * f(a) * f(b) * f(c) / f((j1 + j2 + J) / 2 + 1);
  pre *= f((j1 + m1) / 2) * f((j1 - m1) / 2) * f((j2 + m2) / 2) * f((j2 - m2) / 2) * f((J + M) / 2) *
         f((J - M) / 2);
  long double sum = 0.0L;
  for (int k = 0;; ++k) {
    const int d1 = a - k, d2 = (j1 - m1) / 2 - k, d3 = (j2 + m2) / 2 - k;
    const int d4 = (J - j2 + m1) / 2 + k, d5 = (J - j1 - m2) / 2 + k;
    if (d1 < 0 || d2 < 0 || d3 < 0) break;
    if (d4 < 0 || d5 < 0) continue;
    long double t = 1.0L / (f(k) * f(d1) * f(d2) * f(d3) * f(d4) * f(d5));
    sum += (k & 1) ? -t : t;
  }
  return (double)(std::sqrt(pre) * sum);
}

// ------------------------------------------------------------------------------------
// Oxbash interaction reader.  Numbers are parsed at float precision as the reference's
// std::stof does (include/sm/sphericalint.cxx:186-237), then promoted to double.
struct JKey {
  int k[6];
  bool operator<(const JKey &o) const { return std::lexicographical_compare(k, k + 6, o.k, o.k + 6); }
};

static bool is_comment(const std::string &line) {
  size_t p = line.find_first_not_of(' ');
  if (p == std::string::npos) return false;
  return strchr("#!%/", line[p]) != nullptr && line[p] != '\0';
}

static std::vector<std::string> split_tokens(const std::string &s) {
  std::vector<std::string> t;
  std::string cur;
  for (char c : s) {
    if (isspace((unsigned char)c) || c == ',') {
      if (!cur.empty()) t.push_back(cur), cur.clear();
    } else
      cur += c;
  }
  if (!cur.empty()) t.push_back(cur);
  return t;
}

static int read_oxbash(const std::string &path, int av, double spescale, double vscale, std::vector<double> &spe,
                       std::map<JKey, double> &jv) {
  std::ifstream f(path);
  if (!f) return fail(COSMO_EIO, "interaction file not found: " + path);
  std::string line;
  int lc = 0;
  std::map<JKey, double> tmp;
  while (std::getline(f, line)) {
    if (line.empty()) continue;
    size_t b = line.find_first_not_of(" \t");
    if (b == std::string::npos) continue;
    line = line.substr(b);
    if (is_comment(line)) continue;
    auto tok = split_tokens(line);
    if (tok.empty()) continue;
    if (lc == 0) {
      int nl = atoi(tok[0].c_str());
      int words = (int)tok.size() - 1 - (nl < 0 ? 3 : 0);
      if (words < 0) return fail(COSMO_EIO, "bad first line in " + path);
      if ((int)spe.size() < words) spe.resize(words, 0.0);
      for (int i = 0; i < words; ++i) spe[i] += (double)strtof(tok[i + 1].c_str(), nullptr) * spescale;
      if (nl < 0) {
        double acore = strtof(tok[words + 1].c_str(), nullptr);
        double ascale = strtof(tok[words + 2].c_str(), nullptr);
        double spow = strtof(tok[words + 3].c_str(), nullptr);
        if (av >= 0) vscale *= std::pow(ascale / (av + acore), spow);
      }
    } else {
      if (tok.size() < 7) {
        ++lc;
        continue;
      }
      JKey k;
      k.k[2] = atoi(tok[0].c_str()) - 1;
      k.k[3] = atoi(tok[1].c_str()) - 1;
      k.k[0] = atoi(tok[2].c_str()) - 1;
      k.k[1] = atoi(tok[3].c_str()) - 1;
      k.k[4] = atoi(tok[4].c_str());
      k.k[5] = atoi(tok[5].c_str());
      tmp[k] += vscale * (double)strtof(tok[6].c_str(), nullptr);
    }
    ++lc;
  }
  // duplicate TBMEs (same J T and unordered pairs, either orientation) are zeroed after the
  // first occurrence in key order (include/sm/sphericalint.cxx:43-84)
  std::vector<std::pair<JKey, double>> items(tmp.begin(), tmp.end());
  for (size_t i = 0; i < items.size(); ++i) {
    const int *a0 = items[i].first.k;
    int a[4] = {std::min(a0[0], a0[1]), std::max(a0[0], a0[1]), std::min(a0[2], a0[3]), std::max(a0[2], a0[3])};
    for (size_t j = i + 1; j < items.size(); ++j) {
      const int *b0 = items[j].first.k;
      if (a0[4] != b0[4] || a0[5] != b0[5]) continue;
      int bb[4] = {std::min(b0[0], b0[1]), std::max(b0[0], b0[1]), std::min(b0[2], b0[3]), std::max(b0[2], b0[3])};
      bool bad = (a[0] == bb[0] && a[1] == bb[1] && a[2] == bb[2] && a[3] == bb[3]) ||
                 (a[0] == bb[2] && a[1] == bb[3] && a[2] == bb[0] && a[3] == bb[1]);
      if (bad) items[j].second = 0.0;
    }
  }
  for (auto &it : items) jv[it.first] += it.second;
  return COSMO_OK;
}

// J-coupled -> m-scheme two-body operator: for each JV key in order and each (a0..a3) of
// the key's levels in ascending orbital order with conserved Jz and Tz, accumulate the
// coupled amplitude into the forward-ordered m-scheme key (include/sm/AddPPVsp.cxx:32-146).
static void jm_convert(const std::vector<Orbital> &o, const std::map<JKey, double> &jv, std::vector<uint16_t> &keys,
                       std::vector<double> &vals) {
  std::map<std::array<uint16_t, 4>, double> acc;
  const int ns = (int)o.size();
  for (auto &kv : jv) {
    const int *ijk = kv.first.k;
    const double V = kv.second;
    for (int a0 = 0; a0 < ns; ++a0) {
      if (o[a0].level != ijk[0]) continue;
      for (int a1 = 0; a1 < ns; ++a1) {
        if (o[a1].level != ijk[1]) continue;
        if ((o[a0].j2 & 1) && a1 == a0) continue;
        for (int a2 = 0; a2 < ns; ++a2) {
          if (o[a2].level != ijk[2]) continue;
          for (int a3 = 0; a3 < ns; ++a3) {
            if (o[a3].level != ijk[3]) continue;
            if ((o[a2].j2 & 1) && a2 == a3) continue;
            if (o[a0].m2 + o[a1].m2 != o[a3].m2 + o[a2].m2) continue;
            if (o[a0].tz2 + o[a1].tz2 != o[a3].tz2 + o[a2].tz2) continue;
            double h = clebsch_gordan(o[a0].j2, o[a0].m2, o[a1].j2, o[a1].m2, 2 * ijk[4], o[a0].m2 + o[a1].m2);
            h *= clebsch_gordan(o[a2].j2, o[a2].m2, o[a3].j2, o[a3].m2, 2 * ijk[4], o[a2].m2 + o[a3].m2);
            h *= clebsch_gordan(1, o[a0].tz2, 1, o[a1].tz2, 2 * ijk[5], o[a0].tz2 + o[a1].tz2);
            h *= clebsch_gordan(1, o[a2].tz2, 1, o[a3].tz2, 2 * ijk[5], o[a2].tz2 + o[a3].tz2);
            if (std::fabs(h) < 1e-9) continue;
            h = h * V / std::sqrt((double(ijk[0] == ijk[1]) + 1.) * (double(ijk[2] == ijk[3]) + 1.));
            if (ijk[0] != ijk[1] && o[a0].n == o[a1].n && o[a0].j2 == o[a1].j2 && o[a0].tz2 == -o[a1].tz2 &&
                o[a0].par == o[a1].par)
              h *= std::sqrt(2.);
            if (ijk[2] != ijk[3] && o[a2].n == o[a3].n && o[a2].j2 == o[a3].j2 && o[a2].tz2 == -o[a3].tz2 &&
                o[a2].par == o[a3].par)
              h *= std::sqrt(2.);
            std::array<uint16_t, 4> b = {(uint16_t)a0, (uint16_t)a1, (uint16_t)a2, (uint16_t)a3};
            bool phase = false;
            if (b[0] > b[1]) std::swap(b[0], b[1]), phase ^= (o[a0].j2 & 1);
            if (b[2] > b[3]) std::swap(b[2], b[3]), phase ^= (o[a2].j2 & 1);
            if (b[2] < b[0] || (b[2] == b[0] && b[3] < b[1])) std::swap(b[0], b[2]), std::swap(b[1], b[3]);
            if (ijk[0] == ijk[2] && ijk[1] == ijk[3]) h /= 2.0;
            if (b[0] == b[2] && b[1] == b[3]) h *= 2.0;
            acc[b] += phase ? -h : h;
          }
        }
      }
    }
  }
  keys.clear();
  vals.clear();
  for (auto &kv : acc) {
    keys.insert(keys.end(), kv.first.begin(), kv.first.end());
    vals.push_back(kv.second);
  }
}

// scale * J-J+ (isospin: T-T+) added to ESP and VSPX as <matrix JJ=".." TT=".."> does
// (include/xsm/xReadSPInt.cxx:274-305):
//  * the raising operator's one-body coefficients, keyed [annihilate, create]: J+ moves m
//    to m+1 inside a level and species, sqrt((2j-2m)(2j+2m+2))/2; T+ moves a proton to the
//    neutron orbital of equal n, l, j, m and parity, with T = 1 (SPJPlus / SPTPlus,
//    include/sm/spherical2sp.cxx:623-699);
//  * its square, every (i, j) pair of raising terms in key order: the one-body part where
//    the created orbitals agree, and the two-body key (create_j, annihilate_i, annihilate_j,
//    create_i) with both pairs sorted, each swap flipping the sign, products below 1e-8
//    skipped (SPSqrM, include/sm/cioperators.cxx:200-240);
//  * keys whose created pair precedes the annihilated one removed in map order, where the
//    entry after a removal at the front is not examined again (StripHermitianOperator,
//    include/sm/cioperators.cxx:274-287);
//  * every value times scale, then added key by key into ESP and VSPX (STensor *= and +=,
//    include/tav/stensor.cxx:259-307).
// Two-body keys that annihilate or create one orbital twice never act on a state; they are
// dropped from VSPX after the merge.
template <size_t R>
static void strip_backward(std::vector<std::pair<std::array<uint16_t, R>, double>> &v) {
  const size_t na = R / 2;
  auto backward = [&](const std::array<uint16_t, R> &k) {
    for (size_t i = 0; i < na; ++i)
      if (k[na + i] != k[i]) return k[na + i] < k[i];
    return false;
  };
  size_t i = 0;
  while (i < v.size()) {
    if (backward(v[i].first)) {
      v.erase(v.begin() + i);
      i = 1;
    } else
      ++i;
  }
}

static void add_raising_square(const std::vector<Orbital> &o, bool isospin, double scale, std::vector<uint16_t> &ek,
                               std::vector<double> &ev, std::vector<uint16_t> &vk, std::vector<double> &vv) {
  const int ns = (int)o.size();
  std::map<std::array<uint16_t, 2>, double> jo;
  for (int p0 = 0; p0 < ns; ++p0)
    for (int p1 = p0 + 1; p1 < ns; ++p1) {
      const Orbital &a = o[p0], &b = o[p1];
      int qa, qb, top;  // raised projection (2m or 2tz) and its total (2j or 2t)
      if (!isospin) {
        if (a.level != b.level || a.tz2 != b.tz2) continue;
        qa = a.m2, qb = b.m2, top = a.j2;
      } else {
        if (a.m2 != b.m2 || a.n != b.n || a.par != b.par || a.l != b.l || a.j2 != b.j2) continue;
        qa = a.tz2, qb = b.tz2, top = 1;
      }
      if (qa + 2 == qb) jo[{(uint16_t)p0, (uint16_t)p1}] = std::sqrt((top - qa) * (top + qa + 2.)) / 2.0;
      else if (qa - 2 == qb) jo[{(uint16_t)p1, (uint16_t)p0}] = std::sqrt((top - qb) * (top + qb + 2.)) / 2.0;
    }
  std::map<std::array<uint16_t, 2>, double> jsp;
  std::map<std::array<uint16_t, 4>, double> vsp;
  for (auto &it : jo)
    for (auto &jt : jo) {
      double h = it.second * jt.second;
      if (std::fabs(h) < 1e-8) continue;
      if (jt.first[1] == it.first[1]) jsp[{jt.first[0], it.first[0]}] += h;
      std::array<uint16_t, 4> aa = {jt.first[1], it.first[0], jt.first[0], it.first[1]};
      if (aa[0] > aa[1]) std::swap(aa[0], aa[1]), h = -h;
      if (aa[2] > aa[3]) std::swap(aa[2], aa[3]), h = -h;
      vsp[aa] += h;
    }
  std::vector<std::pair<std::array<uint16_t, 2>, double>> j1(jsp.begin(), jsp.end());
  std::vector<std::pair<std::array<uint16_t, 4>, double>> j2(vsp.begin(), vsp.end());
  strip_backward(j1);
  strip_backward(j2);
  std::map<std::array<uint16_t, 2>, double> e;
  for (size_t t = 0; t < ev.size(); ++t) e[{ek[2 * t], ek[2 * t + 1]}] = ev[t];
  for (auto &kv : j1) e[kv.first] += kv.second * scale;
  std::map<std::array<uint16_t, 4>, double> v;
  for (size_t t = 0; t < vv.size(); ++t) v[{vk[4 * t], vk[4 * t + 1], vk[4 * t + 2], vk[4 * t + 3]}] = vv[t];
  for (auto &kv : j2) v[kv.first] += kv.second * scale;
  ek.clear();
  ev.clear();
  for (auto &kv : e) {
    ek.insert(ek.end(), kv.first.begin(), kv.first.end());
    ev.push_back(kv.second);
  }
  vk.clear();
  vv.clear();
  for (auto &kv : v) {
    if (kv.first[0] == kv.first[1] || kv.first[2] == kv.first[3]) continue;
    vk.insert(vk.end(), kv.first.begin(), kv.first.end());
    vv.push_back(kv.second);
  }
}

// ------------------------------------------------------------------------------------
// basis structure.  The reference basis is the lexicographic order of sorted occupation
// lists (neutrons numbered first), which is (neutron list, then proton list) order.
// Neutron configurations are enumerated in lexicographic order; for each, the proton
// configurations of the matching (M, parity) sector follow in lexicographic order.
static void lex_combinations(int norb, int k, std::vector<uint32_t> &out) {
  out.clear();
  if (k > norb) return;
  std::vector<int> c(k);
  for (int i = 0; i < k; ++i) c[i] = i;
  for (;;) {
    uint32_t m = 0;
    for (int i = 0; i < k; ++i) m |= 1u << c[i];
    out.push_back(m);
    int i = k - 1;
    while (i >= 0 && c[i] == norb - k + i) --i;
    if (i < 0) break;
    ++c[i];
    for (int j = i + 1; j < k; ++j) c[j] = c[j - 1] + 1;
  }
}

void System::build_basis() {
  std::vector<uint32_t> ncomb, pcomb;
  lex_combinations(nn, nv, ncomb);
  lex_combinations(np, zv, pcomb);
  auto mpar = [&](uint32_t mask, int base, int &M, int &P) {
    M = 0;
    P = 0;
    while (mask) {
      int b = __builtin_ctz(mask);
      M += orb[base + b].m2;
      P ^= orb[base + b].par;
      mask &= mask - 1;
    }
  };
  // proton sectors keyed by (M, P)
  std::map<std::pair<int, int>, std::vector<uint32_t>> sec;
  for (uint32_t m : pcomb) {
    int M, P;
    mpar(m, nn, M, P);
    sec[{M, P}].push_back(m);
  }
  std::map<std::pair<int, int>, int> secid;
  pmask.clear();
  psec_start.assign(1, 0);
  pcfg_rank.assign(binom(np, zv), -1);
  for (auto &kv : sec) {
    secid[kv.first] = (int)psec_start.size() - 1;
    int r = 0;
    for (uint32_t m : kv.second) {
      pcfg_rank[colex_rank(m)] = r++;
      pmask.push_back(m);
    }
    psec_start.push_back((int64_t)pmask.size());
  }
  nmask.clear();
  noff.assign(1, 0);
  nsector.clear();
  ncfg_index.assign(binom(nn, nv), -1);
  krank.clear();
  kstart.clear();
  const bool cut = hw_cut || !rejections.empty();
  // a configuration's signature: its oscillator quanta, then its occupation of each
  // rejection's levels (the constraints only see these sums)
  const int nr = (int)rejections.size();
  auto signature = [&](uint32_t mask, int base, bool neutron) {
    std::vector<int> g(nr + 1, 0);
    for (uint32_t x = mask; x; x &= x - 1) {
      const Orbital &o = orb[base + __builtin_ctz(x)];
      g[0] += 2 * o.n + o.l;
      for (int r = 0; r < nr; ++r) {
        const std::vector<int> &lv = neutron ? rejections[r].nlev : rejections[r].plev;
        if (std::find(lv.begin(), lv.end(), o.level) != lv.end()) ++g[r + 1];
      }
    }
    return g;
  };
  // core quanta as CoreQuanta counts them (include/sm/SingleParticleStates.cxx:155-168):
  // the first nv (zv) orbitals of the species in table order (its sort has an empty range)
  int ncore = 0, pcore = 0;
  for (int i = 0; i < nv && i < nn; ++i) ncore += 2 * orb[i].n + orb[i].l;
  for (int i = 0; i < zv && i < np; ++i) pcore += 2 * orb[nn + i].n + orb[nn + i].l;
  // per sector: the proton configurations grouped by signature, ranks ascending
  std::vector<std::map<std::vector<int>, std::vector<int32_t>>> groups;
  if (cut) {
    groups.resize(psec_start.size() - 1);
    for (size_t s = 0; s + 1 < psec_start.size(); ++s)
      for (int64_t q = psec_start[s]; q < psec_start[s + 1]; ++q)
        groups[s][signature(pmask[q], nn, false)].push_back((int32_t)(q - psec_start[s]));
  }
  std::vector<int32_t> keep;
  for (uint32_t m : ncomb) {
    int M, P;
    mpar(m, 0, M, P);
    auto it = secid.find({twoM - M, parity ^ P});
    if (it == secid.end()) continue;
    int s = it->second;
    int64_t size = psec_start[s + 1] - psec_start[s];
    if (cut) {
      // the reference's partition filters: neutron and proton quanta each within hwmax of
      // their core values (Partition::MakeQ), the total within [hwmin, hwmax] above both
      // (QuantaMinMaxPurge), every rejection's occupation within [NMIN, NMAX]
      // (TestRestriction, xmlshell/xmbs/PartitionClass.cxx:23-39)
      const std::vector<int> gn = signature(m, 0, true);
      keep.clear();
      if (!hw_cut || (gn[0] >= ncore && gn[0] <= ncore + hwmax)) {
        for (auto &g : groups[s]) {
          const std::vector<int> &gp = g.first;
          bool ok = true;
          if (hw_cut) {
            const int q = gn[0] + gp[0];
            ok = gp[0] >= pcore && gp[0] <= pcore + hwmax && q >= ncore + pcore + hwmin &&
                 q <= ncore + pcore + hwmax;
          }
          for (int r = 0; ok && r < nr; ++r) {
            const int occ = gn[r + 1] + gp[r + 1];
            ok = occ >= rejections[r].nmin && occ <= rejections[r].nmax;
          }
          if (ok) keep.insert(keep.end(), g.second.begin(), g.second.end());
        }
      }
      if (keep.empty()) continue;
      std::sort(keep.begin(), keep.end());
      if (kstart.empty()) kstart.push_back(0);
      krank.insert(krank.end(), keep.begin(), keep.end());
      kstart.push_back((int64_t)krank.size());
      size = (int64_t)keep.size();
    }
    ncfg_index[colex_rank(m)] = (int32_t)nmask.size();
    nmask.push_back(m);
    nsector.push_back(s);
    noff.push_back(noff.back() + size);
  }
  // a truncation that keeps nothing still marks the basis as truncated
  if (cut && kstart.empty()) kstart.push_back(0);
  dim = noff.back();
}

// value `name` of a node as ReadXMLData reads it (include/xml/xmlParser.cxx:156-177): the
// attribute, else the text of a child element of that name; present with text that does
// not parse as a number gives 0 (a stream extraction that fails leaves the 0 in place).
// False when the node is absent or has neither.
static bool matrix_term(XNode *node, const char *name, double &x) {
  x = 0.0;
  if (!node) return false;
  const char *t = node->get(name);
  std::string txt;
  if (t) txt = t;
  else if (XNode *c = node->child(name)) {
    for (auto &k : c->kids)
      if (k.name.empty()) txt += k.text;
  } else
    return false;
  char *end = nullptr;
  const double v = strtod(txt.c_str(), &end);
  if (end != txt.c_str()) x = v;
  return true;
}

// integer value read the same way (attribute, else child text)
static bool node_int(XNode *node, const char *name, int &x) {
  double v;
  if (!matrix_term(node, name, v)) return false;
  x = (int)v;
  return true;
}

// the basis truncation of a <system> (xmlshell/xmbs/XsysmbsOMP.cpp:212-303): hwmax (an
// attribute or element of the system, hwmin 0), or <hwrejection HWMAX HWMIN> (the first of
// them); every <rejection NMIN NMAX> with its <orbital index type>: the levels of that XML
// orbital index for both species ("pn"), protons ("p"), or neutrons ("n", and when the type
// is absent).  An orbital index without a level of that species is refused.
static int read_truncation(XNode *sysn, System &s) {
  s.hw_cut = false;
  s.rejections.clear();
  if (node_int(sysn, "hwmax", s.hwmax)) {
    s.hwmin = 0;
    s.hw_cut = true;
  }
  if (XNode *h = sysn->child("hwrejection")) {
    if (sysn->count("hwrejection") > 1) fprintf(stderr, "Only one rejection for hw is used \n");
    if (!node_int(h, "HWMAX", s.hwmax) || !node_int(h, "HWMIN", s.hwmin))
      return fail(COSMO_EIO, "<hwrejection> needs HWMAX and HWMIN");
    s.hw_cut = true;
  }
  auto has_level = [&](int lev, bool neutron) {
    for (int i = neutron ? 0 : s.nn; i < (neutron ? s.nn : s.nn + s.np); ++i)
      if (s.orb[i].level == lev) return true;
    return false;
  };
  for (auto &k : sysn->kids) {
    if (k.name != "rejection") continue;
    System::Rejection r;
    if (!node_int(&k, "NMIN", r.nmin) || !node_int(&k, "NMAX", r.nmax))
      return fail(COSMO_EIO, "<rejection> needs NMIN and NMAX");
    for (auto &o : k.kids) {
      if (o.name != "orbital") continue;
      int ix;
      if (!node_int(&o, "index", ix)) return fail(COSMO_EIO, "<rejection><orbital> without index");
      const int lev = ix - 1;
      const std::string type = o.get("type") ? o.get("type") : "n";
      const bool n = type == "pn" || type == "n", p = type == "pn" || type == "p";
      if (!n && !p) return fail(COSMO_EIO, "<rejection><orbital> type must be pn, p or n");
      if ((n && !has_level(lev, true)) || (p && !has_level(lev, false)))
        return fail(COSMO_EUNSUP, "<rejection> names orbital index " + std::to_string(ix) +
                                      " for a species without that level");
      if (n) r.nlev.push_back(lev);
      if (p) r.plev.push_back(lev);
    }
    s.rejections.push_back(r);
  }
  return COSMO_OK;
}

static std::string dir_of(const std::string &p) {
  size_t s = p.find_last_of('/');
  return s == std::string::npos ? std::string(".") : p.substr(0, s);
}

static bool file_exists(const std::string &p) {
  std::ifstream f(p);
  return (bool)f;
}

// interaction file lookup: as given, next to the XML, then along $SMINTPATH, each also
// with ".int" appended (include/av/iofile.cxx:62-120)
static std::string find_int(const std::string &name, const std::string &xmldir) {
  std::vector<std::string> cand = {name, xmldir + "/" + name};
  if (const char *sp = getenv("SMINTPATH")) {
    std::string s = sp, cur;
    for (char c : s + ":") {
      if (c == ':' || c == ';') {
        if (!cur.empty()) cand.push_back(cur + (cur.back() == '/' ? "" : "/") + name);
        cur.clear();
      } else
        cur += c;
    }
  }
  for (auto &c : cand)
    if (file_exists(c)) return c;
  for (auto &c : cand)
    if (file_exists(c + ".int")) return c + ".int";
  return "";
}

static int parse_half(const char *s, int &out) {
  // "5/2" -> 5, "2" -> 4, "-1/2" -> -1, "0.5" -> 1
  if (!s) return COSMO_EINVAL;
  std::string t = s;
  size_t sl = t.find('/');
  if (sl != std::string::npos) {
    int num = atoi(t.substr(0, sl).c_str()), den = atoi(t.substr(sl + 1).c_str());
    if (den == 2) out = num;
    else if (den == 1) out = 2 * num;
    else return COSMO_EINVAL;
    return COSMO_OK;
  }
  double v = atof(t.c_str());
  out = (int)std::lround(2.0 * v);
  if (std::fabs(out - 2.0 * v) > 0.01) return COSMO_EINVAL;
  return COSMO_OK;
}

}  // namespace cosmo

using namespace cosmo;

extern "C" {

const char *cosmo_last_error(void) { return g_err.c_str(); }

int cosmo_system_create(int nsps, const int *level, const int *j2, const int *m2, const int *tz2, const int *nq,
                        const int *lq, int nv, int zv, int twoM, int parity, int64_t nesp, const uint16_t *esp_keys,
                        const double *esp_vals, int64_t nvspx, const uint16_t *vspx_keys, const double *vspx_vals,
                        cosmo_system *out) {
  if (!out || nsps <= 0 || nsps > 64 || nv < 0 || zv < 0) return fail(COSMO_EINVAL, "cosmo_system_create: bad sizes");
  auto *s = new cosmo_system_s();
  int nn = 0, np = 0;
  for (int i = 0; i < nsps; ++i) {
    Orbital o{level[i], j2[i], m2[i], tz2[i], nq ? nq[i] : 0, lq ? lq[i] : 0, lq ? (lq[i] & 1) : 0};
    if (o.tz2 > 0) {
      if (np) {
        delete s;
        return fail(COSMO_EINVAL, "cosmo_system_create: neutrons must precede protons");
      }
      ++nn;
    } else
      ++np;
    s->orb.push_back(o);
  }
  if (nn > 32 || np > 32) {
    delete s;
    return fail(COSMO_EUNSUP, "cosmo_system_create: more than 32 orbitals of one species");
  }
  s->nn = nn;
  s->np = np;
  s->nv = nv;
  s->zv = zv;
  s->twoM = twoM;
  s->parity = parity & 1;
  s->esp_k.assign(esp_keys, esp_keys + 2 * nesp);
  s->esp_v.assign(esp_vals, esp_vals + nesp);
  s->vspx_k.assign(vspx_keys, vspx_keys + 4 * nvspx);
  s->vspx_v.assign(vspx_vals, vspx_vals + nvspx);
  for (int64_t t = 0; t < nvspx; ++t) {
    const uint16_t *k = vspx_keys + 4 * t;
    if (k[0] >= k[1] || k[2] >= k[3] || k[3] >= nsps || k[1] >= nsps) {
      delete s;
      return fail(COSMO_EINVAL, "cosmo_system_create: VSPX keys must be sorted pairs of valid orbitals");
    }
  }
  s->build_basis();
  *out = s;
  return COSMO_OK;
}

int cosmo_system_from_xml(const char *xml_path, const char *sysname, cosmo_system *out) {
  if (!xml_path || !out) return fail(COSMO_EINVAL, "cosmo_system_from_xml: null argument");
  XNode doc;
  std::string err;
  if (!xml_parse_file(xml_path, doc, err)) return fail(COSMO_EIO, err);
  XNode *root = doc.child("cosmo");
  if (!root) return fail(COSMO_EIO, std::string("no <cosmo> root in ") + xml_path);
  XNode *vs = root->child("valencespace");
  if (!vs) return fail(COSMO_EIO, "no <valencespace>");
  // orbitals sorted by index; neutrons (all levels) first, then protons
  struct Lev {
    int index, n, l, j2, par;
    std::string type;
  };
  std::vector<Lev> lev;
  for (auto &k : vs->kids) {
    if (k.name != "orbital") continue;
    Lev L{};
    if (!k.get("index") || !k.get("j")) return fail(COSMO_EIO, "orbital without index or j");
    L.index = atoi(k.get("index"));
    L.n = k.get("n") ? atoi(k.get("n")) : 0;
    L.l = k.get("l") ? atoi(k.get("l")) : 0;
    if (parse_half(k.get("j"), L.j2)) return fail(COSMO_EIO, "bad orbital j");
    L.par = k.get("P") ? (k.get("P")[0] == '-' || atoi(k.get("P")) != 0) : (L.l & 1);
    L.type = k.get("type") ? k.get("type") : "pn";
    lev.push_back(L);
  }
  if (lev.empty()) return fail(COSMO_EIO, "no orbitals defined in xml file");
  std::sort(lev.begin(), lev.end(), [](const Lev &a, const Lev &b) { return a.index < b.index; });
  std::vector<Orbital> neu, pro;
  for (auto &L : lev)
    for (int m = -L.j2; m <= L.j2; m += 2) {
      Orbital o{L.index - 1, L.j2, m, 1, L.n, L.l, L.par};
      if (L.type == "pn" || L.type == "n") neu.push_back(o);
      if (L.type == "pn" || L.type == "p") {
        o.tz2 = -1;
        pro.push_back(o);
      }
    }
  // <system>
  XNode *sysn = nullptr;
  if (sysname) {
    for (auto &k : root->kids)
      if (k.name == "system" && k.get("name") && std::string(k.get("name")) == sysname) sysn = &k;
  } else
    sysn = root->child("system", -1);
  if (!sysn) return fail(COSMO_EIO, "no <system> node");
  auto need = [&](const char *a) -> const char * { return sysn->get(a); };
  if (!need("name")) return fail(COSMO_EIO, "system without name attribute");
  int av = need("Av") ? atoi(need("Av")) : -1;
  int nv, zv;
  if (need("Nv") && need("Zv")) {
    nv = atoi(need("Nv"));
    zv = atoi(need("Zv"));
    if (av < 0) av = nv + zv;
  } else if (av >= 0 && need("Tz")) {
    // valence Tz = (N - Z) / 2 (include/xsm/xManyBodyStates.cxx:47-55)
    int tz2;
    if (parse_half(need("Tz"), tz2) || ((av + tz2) & 1) || std::abs(tz2) > av)
      return fail(COSMO_EIO, "bad Tz for the valence particle number");
    zv = (av - tz2) / 2;
    nv = (av + tz2) / 2;
  } else
    return fail(COSMO_EIO, "system needs Nv and Zv (or Av and Tz); interactive prompts are not supported");
  int twoM = av & 1;
  if (need("Jz") && parse_half(need("Jz"), twoM)) return fail(COSMO_EIO, "bad Jz");
  int parity = need("P") ? (need("P")[0] == '-') : 0;
  // interaction: <hamiltonian><file><name>..</name><scale rank=..>x</scale></file>
  XNode *ham = root->child("hamiltonian");
  if (!ham) return fail(COSMO_EIO, "no <hamiltonian>");
  // what xReadSPInt would add beyond the Oxbash files and JJ/TT (include/xsm/xReadSPInt.cxx:
  // 61-77 JISP, 106-270 tag files, 306-372 Lawson, Coulomb, Delta) is refused rather than
  // ignored, so no input builds a different H than the reference's without an error
  XNode *mat = sysn->child("matrix");
  for (const char *k : {"jispfile", "tagfile"})
    if (ham->count(k)) return fail(COSMO_EUNSUP, std::string("<hamiltonian><") + k + "> interactions are not supported");
  for (const char *k : {"Lawson", "Coulomb", "Delta"}) {
    double x;
    if (matrix_term(mat, k, x) || matrix_term(ham, k, x))
      return fail(COSMO_EUNSUP, std::string("the ") + k + " Hamiltonian term is not supported");
  }
  std::vector<double> spe;
  std::map<JKey, double> jv;
  std::string xdir = dir_of(xml_path);
  int nfiles = 0;
  for (auto &fk : ham->kids) {
    if (fk.name != "file") continue;
    double sc[10];
    for (double &v : sc) v = 1.0;
    std::string fname;
    for (auto &c : fk.kids) {
      if (c.name == "name" && !c.kids.empty()) fname = c.kids[0].text;
      if (c.name == "scale" && c.get("rank") && !c.kids.empty()) {
        int r = atoi(c.get("rank"));
        if (r >= 0 && r < 10) sc[r] *= atof(c.kids[0].text.c_str());
      }
    }
    fname.erase(0, fname.find_first_not_of(" \t\r\n"));
    fname.erase(fname.find_last_not_of(" \t\r\n") + 1);
    std::string path = find_int(fname, xdir);
    if (path.empty()) return fail(COSMO_EIO, "interaction file not found: " + fname);
    int rc = read_oxbash(path, av, sc[1], sc[2], spe, jv);
    if (rc) return rc;
    ++nfiles;
  }
  if (!nfiles) return fail(COSMO_EUNSUP, "only Oxbash <file> interactions are supported");
  auto *s = new cosmo_system_s();
  s->orb = neu;
  s->orb.insert(s->orb.end(), pro.begin(), pro.end());
  s->nn = (int)neu.size();
  s->np = (int)pro.size();
  if (s->nn > 32 || s->np > 32) {
    delete s;
    return fail(COSMO_EUNSUP, "more than 32 orbitals of one species");
  }
  s->nv = nv;
  s->zv = zv;
  s->twoM = twoM;
  s->parity = parity;
  s->name = need("name");
  // one-body operator: ESP(i,i) = e[level(i)] (include/sm/MakeSPO.cxx:195-202)
  for (size_t i = 0; i < s->orb.size(); ++i) {
    s->esp_k.push_back((uint16_t)i);
    s->esp_k.push_back((uint16_t)i);
    int l = s->orb[i].level;
    s->esp_v.push_back(l < (int)spe.size() ? 0.0 + spe[l] : 0.0);
  }
  jm_convert(s->orb, jv, s->vspx_k, s->vspx_v);
  // <system><matrix> add-ons, in the reference's order: JJ, then TT
  for (int t = 0; t < 2; ++t) {
    double scale = 0.0;
    if (matrix_term(mat, t ? "TT" : "JJ", scale))
      add_raising_square(s->orb, t == 1, scale, s->esp_k, s->esp_v, s->vspx_k, s->vspx_v);
  }
  {
    const int rc = read_truncation(sysn, *s);
    if (rc) {
      delete s;
      return rc;
    }
  }
  s->build_basis();
  *out = s;
  return COSMO_OK;
}

int cosmo_system_destroy(cosmo_system s) {
  delete s;
  return COSMO_OK;
}

int64_t cosmo_system_dim(cosmo_system s) { return s ? s->dim : COSMO_EINVAL; }

int cosmo_system_info(cosmo_system s, int *nparticles, int *nsps, int *nn, int *np, int64_t *nesp, int64_t *nvspx) {
  if (!s) return fail(COSMO_EINVAL, "null system");
  if (nparticles) *nparticles = s->nv + s->zv;
  if (nsps) *nsps = (int)s->orb.size();
  if (nn) *nn = s->nn;
  if (np) *np = s->np;
  if (nesp) *nesp = (int64_t)s->esp_v.size();
  if (nvspx) *nvspx = (int64_t)s->vspx_v.size();
  return COSMO_OK;
}

int cosmo_system_basis(cosmo_system s, uint16_t *out) {
  if (!s || !out) return fail(COSMO_EINVAL, "null argument");
  const int A = s->nv + s->zv;
  for (size_t i = 0; i < s->nmask.size(); ++i) {
    const int64_t q0 = s->psec_start[s->nsector[i]];
    for (int64_t r = s->noff[i]; r < s->noff[i + 1]; ++r) {
      uint16_t *row = out + r * A;
      int c = 0;
      for (uint32_t m = s->nmask[i]; m; m &= m - 1) row[c++] = (uint16_t)__builtin_ctz(m);
      for (uint32_t m = s->pmask[q0 + s->prow((int)i, r)]; m; m &= m - 1)
        row[c++] = (uint16_t)(s->nn + __builtin_ctz(m));
    }
  }
  return COSMO_OK;
}

int cosmo_system_tensors(cosmo_system s, uint16_t *ek, double *ev, uint16_t *vk, double *vv) {
  if (!s) return fail(COSMO_EINVAL, "null system");
  if (ek) std::copy(s->esp_k.begin(), s->esp_k.end(), ek);
  if (ev) std::copy(s->esp_v.begin(), s->esp_v.end(), ev);
  if (vk) std::copy(s->vspx_k.begin(), s->vspx_k.end(), vk);
  if (vv) std::copy(s->vspx_v.begin(), s->vspx_v.end(), vv);
  return COSMO_OK;
}

int cosmo_system_orbitals(cosmo_system s, int *level, int *j2, int *m2, int *tz2) {
  if (!s) return fail(COSMO_EINVAL, "null system");
  for (size_t i = 0; i < s->orb.size(); ++i) {
    if (level) level[i] = s->orb[i].level;
    if (j2) j2[i] = s->orb[i].j2;
    if (m2) m2[i] = s->orb[i].m2;
    if (tz2) tz2[i] = s->orb[i].tz2;
  }
  return COSMO_OK;
}

int cosmo_system_write_mbs(cosmo_system s, const char *path) {
  if (!s || !path) return fail(COSMO_EINVAL, "null argument");
  const int A = s->nv + s->zv;
  std::vector<uint16_t> z((size_t)s->dim * A);
  cosmo_system_basis(s, z.data());
  FILE *f = fopen(path, "wb");
  if (!f) return fail(COSMO_EIO, std::string("cannot write ") + path);
  int32_t N = A;
  uint64_t n = (uint64_t)s->dim;
  fwrite(&N, 4, 1, f);
  fwrite(&n, 8, 1, f);
  fwrite(z.data(), 2, z.size(), f);
  fclose(f);
  return COSMO_OK;
}

int cosmo_system_check_mbs(cosmo_system s, const char *path) {
  if (!s || !path) return fail(COSMO_EINVAL, "null argument");
  FILE *f = fopen(path, "rb");
  if (!f) return fail(COSMO_EIO, std::string("cannot open ") + path);
  int32_t N = 0;
  uint64_t n = 0;
  if (fread(&N, 4, 1, f) != 1 || fread(&n, 8, 1, f) != 1) {
    fclose(f);
    return fail(COSMO_EIO, "truncated .mbs header");
  }
  const int A = s->nv + s->zv;
  if (N != A) {
    fclose(f);
    return fail(COSMO_EUNSUP, ".mbs particle number differs from the system");
  }
  std::vector<uint16_t> z((size_t)n * A);
  size_t got = fread(z.data(), 2, z.size(), f);
  int extra = fgetc(f);
  fclose(f);
  if (got != z.size() || extra != EOF) return fail(COSMO_EIO, "Read:ManyBodyStates file-size mismatch");
  const char *what = s->truncated() ? "the system's truncated basis" : "the full M-scheme basis";
  if ((int64_t)n != s->dim) return fail(COSMO_EUNSUP, std::string(".mbs basis size differs from ") + what);
  std::vector<uint16_t> mine(z.size());
  cosmo_system_basis(s, mine.data());
  if (mine != z) return fail(COSMO_EUNSUP, std::string(".mbs basis differs from ") + what + " in reference order");
  return COSMO_OK;
}

}  // extern "C"
