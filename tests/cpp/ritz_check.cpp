// Test driver for ritz_lowest (cosmo_amd/csrc/ritz.cpp): reads cases from a binary file
// (int32 count; per case int32 k, int32 want, k*k float64 row-major) and writes, per case,
// int32 ok, want float64 eigenvalues and want*k float64 eigenvector components.
#include <cstdint>
#include <cstdio>
#include <vector>

namespace cosmo {
bool ritz_lowest(int k, const std::vector<double> &A0, int want, std::vector<double> &w, std::vector<double> &V);
}

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  FILE *in = fopen(argv[1], "rb"), *out = fopen(argv[2], "wb");
  if (!in || !out) return 3;
  int32_t count = 0;
  if (fread(&count, 4, 1, in) != 1) return 4;
  for (int c = 0; c < count; ++c) {
    int32_t k = 0, want = 0;
    if (fread(&k, 4, 1, in) != 1 || fread(&want, 4, 1, in) != 1) return 4;
    std::vector<double> A((size_t)k * k), w, V;
    if (fread(A.data(), 8, A.size(), in) != A.size()) return 4;
    const int32_t ok = cosmo::ritz_lowest(k, A, want, w, V) ? 1 : 0;
    fwrite(&ok, 4, 1, out);
    w.resize(want);
    V.resize((size_t)want * k);
    fwrite(w.data(), 8, w.size(), out);
    fwrite(V.data(), 8, V.size(), out);
  }
  fclose(in);
  fclose(out);
  return 0;
}
