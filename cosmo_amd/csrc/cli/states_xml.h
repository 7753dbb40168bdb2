// states_xml.h -- the state database step of the command-line tools: spin and isospin
// labels of eigenstates, the system's J / T filters, removal of the matrix's JJ / TT shifts
// and the <state> / <wavefunction> nodes.  Follows XSHLJT (include/xsm/XSHLJT.cxx:38-142),
// AddState / SearchState (include/xsm/xState.cxx:71-92, 113-148) and OrganizeStates
// (:198-244), with <J-J+>, <T-T+> from the GPU (cosmo_system_labels).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "cli_common.h"

// a spin in the reference's text form: "2", "3/2", "-1/2"
inline std::string spin_text(int twice) {
  if (twice % 2 == 0) return std::to_string(twice / 2);
  return std::to_string(twice) + "/2";
}

// twice a spin given as "3/2", "2" or "1.5" (the reference's spin(const char *)); false
// when it is neither an integer nor a half-integer
inline bool parse_spin(const char *s, int &twice) {
  if (!s || !*s) return false;
  std::string t = s;
  const size_t sl = t.find('/');
  if (sl != std::string::npos) {
    if (t.compare(sl + 1, std::string::npos, "2") != 0) return false;
    twice = atoi(t.substr(0, sl).c_str());
    return true;
  }
  const double x = atof(t.c_str());
  twice = (int)(x >= 0 ? 2.0 * x + 0.5 : 2.0 * x - 0.5);
  return std::fabs(twice - 2.0 * x) <= 0.01;
}

// the reference's double attributes (6 significant digits)
inline std::string g6(double x) {
  char b[64];
  snprintf(b, sizeof b, "%g", x);
  return b;
}

// a term of <matrix>: attribute, else child element text (xml::getXMLData); 0 when absent
inline double matrix_scale(cosmo::XNode *mat, const char *k) {
  if (!mat) return 0.0;
  if (const char *v = mat->get(k)) return atof(v);
  if (cosmo::XNode *c = mat->child(k)) return atof(c->text.c_str());
  return 0.0;
}

// AddState: the state with these labels and energy within 0.001 (a new one otherwise),
// then its wavefunction (file, id) unless the first wavefunction of that file already has
// this id
inline cosmo::XNode *add_state(cosmo::XNode *sysn, int J2, int T2, bool neg_parity, double E,
                               const std::string &file, int id, int twoM, int twoTz) {
  cosmo::XNode *st = nullptr;
  for (auto &k : sysn->kids) {
    if (k.name != "state") continue;
    int j2, t2;
    const char *je = k.get("J"), *te = k.get("T"), *ee = k.get("E");
    if (!je || !te || !ee || !parse_spin(je, j2) || !parse_spin(te, t2)) break;  // as the search's exception
    if (j2 != J2 || t2 != T2 || std::fabs(atof(ee) - E) > 0.001) continue;
    st = &k;
    break;
  }
  if (!st) {
    cosmo::XNode k;
    k.name = "state";
    k.set("J", spin_text(J2));
    k.set("P", neg_parity ? "-" : "+");
    k.set("T", spin_text(T2));
    k.set("E", g6(E));
    sysn->kids.push_back(k);
    st = &sysn->kids.back();
  }
  for (auto &w : st->kids)
    if (w.name == "wavefunction" && w.get("file") && file == w.get("file")) {
      if (w.get("id") && atoi(w.get("id")) == id) return st;
      break;
    }
  cosmo::XNode w;
  w.name = "wavefunction";
  w.set("Jz", spin_text(twoM));
  w.set("Tz", spin_text(twoTz));
  w.set("file", file);
  w.set("id", std::to_string(id));
  st->kids.push_back(w);
  return st;
}

// OrganizeStates: states by (printed) energy, named J P(k) by their rank among states of
// the same J and parity, Ex above the lowest, moved after the system's other children
inline int organize_states(cosmo::XNode *sysn) {
  std::vector<std::pair<double, int>> e;
  std::vector<cosmo::XNode> states, rest;
  for (auto &k : sysn->kids) {
    if (k.name == "state") {
      if (!k.get("E")) return fprintf(stderr, "Error: <state> without E\n"), 1;
      e.push_back({atof(k.get("E")), (int)states.size()});
      states.push_back(k);
    } else
      rest.push_back(k);
  }
  std::sort(e.begin(), e.end());
  std::map<std::pair<std::string, std::string>, int> order;
  for (auto &x : e) {
    cosmo::XNode &k = states[x.second];
    const std::string J = k.get("J") ? k.get("J") : "??", P = k.get("P") ? k.get("P") : "?";
    const int c = ++order[{J, P}];
    k.set("name", J + P[0] + "(" + std::to_string(c) + ")");
    k.set("Ex", g6(x.first - e[0].first));
    rest.push_back(k);
  }
  sysn->kids.swap(rest);
  return 0;
}

// XSHLJT over count states (energies, |dE| and host vectors, column-major n x count) kept
// in file eefile: labels, filters, shifts, state nodes.  0 or an error already reported.
inline int label_states(cosmo_system s, cosmo::XNode *sysn, const std::string &eefile, int count,
                        const double *evals, const double *dE, const double *evecs) {
  cosmo::XNode *mat = sysn->child("matrix", -1);
  if (!mat) return fprintf(stderr, "Error: You need Hamiltonian matrix\n"), 1;
  if (!mat->get("name")) return fprintf(stderr, "Error: the <matrix> node has no name attribute\n"), 1;
  if (!mat->get("format") || std::string(mat->get("format")) != "pp")
    return fprintf(stderr, "Error: matrix is not in pp format\n"), 1;
  int xJ = -1, xT = -1;  // the system's J / T: only states with these are processed
  if (sysn->get("J") && !parse_spin(sysn->get("J"), xJ)) return fprintf(stderr, "Error: bad system J\n"), 1;
  if (sysn->get("T") && !parse_spin(sysn->get("T"), xT)) return fprintf(stderr, "Error: bad system T\n"), 1;
  if (xJ != -1) fprintf(stderr, "States only with J=%s are processed \n", spin_text(xJ).c_str());
  if (xT != -1) fprintf(stderr, "States only with T=%s are processed \n", spin_text(xT).c_str());
  const double jscale = matrix_scale(mat, "JJ"), tscale = matrix_scale(mat, "TT");
  fprintf(stderr, "J+J- scaling is %g\nT+T- scaling is %g\n", jscale, tscale);
  int nparticles, nsps, nn, np;
  int64_t ne, nvx;
  if (cosmo_system_info(s, &nparticles, &nsps, &nn, &np, &ne, &nvx)) return die("system");
  int twoM = nparticles & 1, twoTz = 0;
  if (sysn->get("Jz") && !parse_spin(sysn->get("Jz"), twoM)) return fprintf(stderr, "Error: bad Jz\n"), 1;
  // Tz of the system node (physical (N - Z) / 2), else from the valence numbers
  if (sysn->get("Tz")) {
    if (!parse_spin(sysn->get("Tz"), twoTz)) return fprintf(stderr, "Error: bad Tz\n"), 1;
  } else {
    twoTz = (sysn->get("Nv") ? atoi(sysn->get("Nv")) : 0) - (sysn->get("Zv") ? atoi(sysn->get("Zv")) : 0);
  }
  const bool neg = sysn->get("P") && sysn->get("P")[0] == '-';
  std::vector<double> jj(std::max(count, 1)), tt(std::max(count, 1));
  if (count > 0 && cosmo_system_labels(s, evecs, count, jj.data(), tt.data())) return die("labels");
  for (int i = 0; i < count; ++i) {
    double E = evals[i];
    fprintf(stderr, "State: %d E= %g JJ= %g\n", i, E, jj[i]);
    const int J2 = cosmo_twice_spin(jj[i], twoM);
    if (xJ != -1 && xJ != J2) continue;
    if (std::fabs(jscale) > 1e-8) {
      const double sh = jscale * ((J2 - twoM) * (J2 + twoM + 2) / 4.0);
      E -= sh;
      fprintf(stderr, "Applying shift js=%g s=%g\n", jscale, sh);
    }
    fprintf(stderr, " TT= %g \n", tt[i]);
    const int T2 = cosmo_twice_spin(tt[i], twoTz);
    if (xT != -1 && xT != T2) continue;
    if (std::fabs(tscale) > 1e-8) {
      const double sh = tscale * ((T2 - twoTz) * (T2 + twoTz + 2) / 4.0);
      E -= sh;
      fprintf(stderr, "Applying shift ts=%g s=%g\n", tscale, sh);
    }
    cosmo::XNode *st = add_state(sysn, J2, T2, neg, E, eefile, i, twoM, twoTz);
    st->set("dE", g6(dE ? dE[i] : 0.0));
  }
  return organize_states(sysn);
}
