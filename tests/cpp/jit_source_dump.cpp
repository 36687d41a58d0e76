// Reads a binary site program (text) and prints the run-time specialised
// kernel source the library generates for it (test helper).
// Input: n_slots, then per slot: k_others n_clusters weight offsets...
#include <cstdio>
#include <fstream>
#include <iostream>

#include "../../casmcode_clexmonte_amd/csrc/clexmc_jit.h"

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  std::ifstream in(argv[1]);
  int ns = 0;
  in >> ns;
  std::vector<clexmc::BinSub> subs(1);
  clexmc::BinSub &B = subs[0];
  std::vector<int> off;
  for (int j = 0; j < ns; ++j) {
    int ko = 0, n = 0;
    double w = 0.0;
    in >> ko >> n >> w;
    B.slot_k[j] = ko;
    B.slot_n[j] = n;
    B.slot_begin[j] = (int)off.size();
    B.slot_w[j] = (float)w;
    for (int e = 0; e < n * ko; ++e) {
      int x = 0;
      in >> x;
      off.push_back(x);
    }
  }
  std::cout << clexmc::jit_source(1, {1}, subs, off, ns);
  return 0;
}
