// Reads a site program (text) and prints the run-time specialised kernel
// source the library generates for it (test helper).
//   binary:  "bin" n_slots, then per slot: k_others n_clusters weight offsets...
//   general: "gen" nf n_entries, then per entry: k table_offset (offset radix)*k
// A second argument "1" asks for the 2-bit packed image variant.
#include <fstream>
#include <iostream>
#include <string>

#include "../../casmcode_clexmonte_amd/csrc/clexmc_jit.h"

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  std::ifstream in(argv[1]);
  std::string mode;
  in >> mode;
  std::vector<clexmc::BinSub> subs(1);
  std::vector<clexmc::GenSub> gsubs(1);
  std::vector<int> off, prog;
  int ns = 0;
  if (mode == "bin") {
    in >> ns;
    clexmc::BinSub &B = subs[0];
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
  } else {
    clexmc::GenSub &G = gsubs[0];
    in >> G.nf >> G.n_entries;
    G.begin = 0;
    for (int e = 0; e < G.n_entries; ++e) {
      int k = 0, toff = 0;
      in >> k >> toff;
      prog.push_back(k);
      prog.push_back(toff);
      for (int i = 0; i < 2 * k; ++i) {
        int x = 0;
        in >> x;
        prog.push_back(x);
      }
    }
  }
  // optional second argument: 1 = 2-bit packed image (table path)
  const bool packed = argc > 2 && std::string(argv[2]) == "1";
  std::cout << clexmc::jit_source(1, {1}, mode == "bin", subs, off, ns, gsubs, prog, 1, packed, mode == "bin" ? 2 : 1,
                                  512);
  return 0;
}
