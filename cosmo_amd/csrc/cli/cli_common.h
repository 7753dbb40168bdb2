// cli_common.h -- shared pieces of the command-line tools: a small POSIX/GNU flag parser
// with the flag surface of the reference's av::argparser (include/av/argparser.cxx:76-251)
// and helpers around the cosmo_hip C ABI.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "../internal.h"

struct Args {
  std::vector<std::string> pos;              // positional arguments (pos[0] = program)
  std::map<std::string, std::string> flags;  // canonical short name -> value ("" if none)
  bool help = false;
};

// spec: short name -> long name; every flag takes one value
inline Args parse_args(int argc, char **argv, const std::map<std::string, std::string> &spec) {
  Args a;
  a.pos.push_back(argv[0]);
  for (int i = 1; i < argc; ++i) {
    std::string s = argv[i];
    if (s == "-h" || s == "--help") {
      a.help = true;
      continue;
    }
    std::string key;
    if (s.size() > 2 && s[0] == '-' && s[1] == '-') {
      for (auto &kv : spec)
        if (kv.second == s.substr(2)) key = kv.first;
    } else if (s.size() > 1 && s[0] == '-' && !(isdigit((unsigned char)s[1]))) {
      if (spec.count(s.substr(1))) key = s.substr(1);
    } else {
      a.pos.push_back(s);
      continue;
    }
    if (key.empty()) {
      fprintf(stderr, "Warning: unknown flag %s is ignored\n", s.c_str());
      continue;
    }
    a.flags[key] = (i + 1 < argc) ? argv[++i] : "";
  }
  return a;
}

inline int flag_int(const Args &a, const char *k, int def) {
  auto it = a.flags.find(k);
  return it == a.flags.end() ? def : atoi(it->second.c_str());
}

inline int die(const char *what) {
  fprintf(stderr, "Error: %s: %s\n", what, cosmo_last_error());
  return 1;
}

// <system> node chosen by name attribute (the last one when several have no match)
inline cosmo::XNode *select_system(cosmo::XNode *root) { return root->child("system", -1); }
