"""CPU check of the run-time specialised fields: the straight-line code the
library generates for a site program (constant offsets) must give the same
results as the interpreted site program, on random images.

* binary: the integer cluster-product counts K and the FP32 field
  F = fma(w_j, K_j, F) (the generic kernel's rule: per orbit slot, X = sum
  over clusters of the XOR of the other sites, K = 2X - n for an odd number
  of other sites, n - 2X for an even number), whole and as the two lane-pair
  partial sums followed by the finish;
* general (table path): G_f = sum over clusters of V[toff + idx * nf + f],
  idx = sum of occupant x radix multiplier of the other sites, in entry
  order (bit-identical FP32 sums).

The generated JitField is compiled for the host with g++."""
import os
import subprocess

import numpy as np
import pytest

from casmcode_clexmonte_amd.system import FlatCE
from casmcode_clexmonte_amd.workloads import fcc_binary_ce, fcc_ordering_ce

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DUMP = os.path.join(ROOT, "tests", "cpp", "jit_source_dump")

HOST = r'''
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>
#define __device__
#define __forceinline__ inline
constexpr int kMaxSlots = 8;
constexpr int kMaxFunc = 3;
struct FieldCtx {};
struct HostImg { const uint8_t *p; int get(int a) const { return p[a]; } };
JITFIELD
int main(int, char **argv) {
  std::ifstream in(argv[1]); std::string mode; in >> mode; int ns; in >> ns;
  std::vector<int> ko(ns), n(ns), beg(ns); std::vector<float> w(ns); std::vector<int> off;
  for (int j = 0; j < ns; ++j) { double ww; in >> ko[j] >> n[j] >> ww; w[j] = (float)ww; beg[j] = (int)off.size();
    for (int e = 0; e < n[j] * ko[j]; ++e) { int x; in >> x; off.push_back(x); } }
  const int P = PADDIM, H = HALO, L = P - 2 * H; std::vector<uint8_t> img(P * P * P); srand(7);
  int bad = 0;
  for (int trial = 0; trial < 3000; ++trial) {
    for (auto &v : img) v = (uint8_t)(rand() & 1);
    const int pa = (H + rand() % L) + P * (H + rand() % L) + P * P * (H + rand() % L);
    int Kg[8] = {0}; float Fg = 0.f;
    for (int j = 0; j < ns; ++j) { if (n[j] <= 0) continue; int X = 0;
      for (int e = 0; e < n[j]; ++e) { int x = 0; for (int i = 0; i < ko[j]; ++i) x ^= img[pa + off[beg[j] + e * ko[j] + i]]; X += x; }
      const int K = (ko[j] & 1) ? (2 * X - n[j]) : (n[j] - 2 * X); Kg[j] = K; Fg = std::fma(w[j], (float)K, Fg); }
    int Kj[8]; HostImg im{img.data()}; FieldCtx c;
    const float Fj = JitField::bin(c, im, 0, pa, img[pa], Kj);
    for (int j = 0; j < 8; ++j) bad += Kj[j] != Kg[j];
    bad += Fj != Fg;
    // lane-pair partial sums, combined, then the finish
    int X0[8], X1[8], X[8], K2[8];
    JitField::bin_x(c, im, 0, pa, img[pa], 0, X0);
    JitField::bin_x(c, im, 0, pa, img[pa], 1, X1);
    for (int j = 0; j < 8; ++j) X[j] = X0[j] + X1[j];
    const float F2 = JitField::bin_finish(c, 0, X, K2);
    for (int j = 0; j < 8; ++j) bad += K2[j] != Kg[j];
    bad += F2 != Fg;
  }
  printf("%d\n", bad); return 0;
}
'''


def _program(flat, L):
    """The library's binary site program for sublattice 0 of an L^3 FCC
    supercell: per orbit slot (orbits with >= 1 site, in orbit order), the
    padded offsets of the other sites of each cluster containing the site."""
    h = int(flat.halo)
    P = L + 2 * h
    lines = []
    slots = [a for a in range(flat.n_orbits) if flat.orbit_k[a] >= 1]
    w = np.asarray(flat.bin_orbit_weight)
    for a in slots:
        ko = int(flat.orbit_k[a]) - 1
        offs, n = [], 0
        for e in range(flat.flower_begin[0], flat.flower_begin[1]):
            if flat.entry_orbit[e] != a:
                continue
            n += 1
            for q in range(flat.entry_site_begin[e], flat.entry_site_begin[e + 1]):
                b2, di, dj, dk = flat.entry_sites[q]
                offs.append(int(b2 * P ** 3 + dk * P * P + dj * P + di))
        lines.append(f"{ko} {n} {float(w[a])!r} " + " ".join(map(str, offs)))
    return P, h, f"bin {len(slots)}\n" + "\n".join(lines) + "\n"


@pytest.mark.parametrize("ce", [fcc_binary_ce, fcc_ordering_ce])
def test_generated_field_matches_interpreted_program(ce, tmp_path):
    if not os.path.exists(DUMP):
        pytest.skip("tests/cpp/jit_source_dump not built (run __graft_entry__.build())")
    flat = FlatCE(ce())
    P, h, prog = _program(flat, 12)
    pf = tmp_path / "prog.txt"
    pf.write_text(prog)
    src = subprocess.run([DUMP, str(pf)], capture_output=True, text=True, check=True).stdout
    i = src.index("struct JitField {")
    j = src.index("};\n}  // namespace clexmc")
    host = HOST.replace("JITFIELD", src[i:j + 2]).replace("PADDIM", str(P)).replace("HALO", str(h))
    cpp = tmp_path / "check.cpp"
    cpp.write_text(host)
    exe = tmp_path / "check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", str(exe), str(cpp)], check=True, capture_output=True)
    r = subprocess.run([str(exe), str(pf)], capture_output=True, text=True, check=True)
    assert int(r.stdout.strip()) == 0


HOST_GEN = r'''
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>
#define __device__
#define __forceinline__ inline
constexpr int kMaxSlots = 8;
constexpr int kMaxFunc = 3;
struct FieldCtx {};
struct HostImg { const uint8_t *p; int get(int a) const { return p[a]; } };
JITFIELD
int main(int, char **argv) {
  std::ifstream in(argv[1]); std::string mode; in >> mode; int nf, ne; in >> nf >> ne;
  std::vector<int> k(ne), toff(ne), beg(ne); std::vector<int> om;
  for (int e = 0; e < ne; ++e) { in >> k[e] >> toff[e]; beg[e] = (int)om.size();
    for (int i = 0; i < 2 * k[e]; ++i) { int x; in >> x; om.push_back(x); } }
  const int P = PADDIM, H = HALO, L = P - 2 * H, NOCC = NOCCUPANTS; std::vector<uint8_t> img(P * P * P); srand(11);
  std::vector<float> tab(TABSIZE); for (auto &t : tab) t = (float)(rand() % 20001 - 10000) * 1e-5f;
  int bad = 0;
  for (int trial = 0; trial < 3000; ++trial) {
    for (auto &v : img) v = (uint8_t)(rand() % NOCC);
    const int pa = (H + rand() % L) + P * (H + rand() % L) + P * P * (H + rand() % L);
    float Gg[3] = {0.f, 0.f, 0.f};
    for (int e = 0; e < ne; ++e) { int idx = 0;
      for (int i = 0; i < k[e]; ++i) idx += img[pa + om[beg[e] + 2 * i]] * om[beg[e] + 2 * i + 1];
      const float *V = tab.data() + toff[e] + idx * nf; Gg[0] += V[0]; if (nf > 1) Gg[1] += V[1]; if (nf > 2) Gg[2] += V[2]; }
    float Gj[3]; HostImg im{img.data()}; FieldCtx c;
    JitField::gen(c, im, 0, pa, tab.data(), Gj);
    for (int f = 0; f < 3; ++f) bad += Gj[f] != Gg[f];
  }
  printf("%d\n", bad); return 0;
}
'''


def _gen_program(flat, L):
    """The library's general site program for sublattice 0 of an L^3
    supercell (single-sublattice lattices): per flower entry, k, its table
    offset and (padded offset, radix multiplier) of each other site."""
    h = int(flat.halo)
    P = L + 2 * h
    nf = int(flat.n_occ[0]) - 1
    lines = []
    n = 0
    tab_max = 0
    for e in range(flat.flower_begin[0], flat.flower_begin[1]):
        sites = range(flat.entry_site_begin[e], flat.entry_site_begin[e + 1])
        toff = int(flat.entry_table[e])
        mul, parts = 1, []
        for q in sites:
            b2, di, dj, dk = flat.entry_sites[q]
            parts.append((int(b2 * P ** 3 + dk * P * P + dj * P + di), mul))
            mul *= int(flat.n_occ[b2])
        tab_max = max(tab_max, toff + mul * nf)
        lines.append(f"{len(parts)} {toff} " + " ".join(f"{o} {m}" for o, m in parts))
        n += 1
    return P, h, nf, tab_max, f"gen {nf} {n}\n" + "\n".join(lines) + "\n"


HOST_GEN_P2 = HOST_GEN.replace(
    "struct HostImg { const uint8_t *p; int get(int a) const { return p[a]; } };",
    "#define __builtin_amdgcn_ubfe(x, off, w) ((uint32_t)(x) >> (off) & ((1u << (w)) - 1u))\n"
    "struct HostImg { const uint8_t *p; const uint32_t *w; int get(int a) const { return p[a]; } };").replace(
    "HostImg im{img.data()};", "std::vector<uint32_t> words((img.size() + 15) / 16, 0u);\n"
    "    for (size_t s = 0; s < img.size(); ++s) words[s >> 4] |= (uint32_t)img[s] << (2 * (s & 15));\n"
    "    HostImg im{img.data(), words.data()};")


@pytest.mark.parametrize("packed", [False, True])
def test_generated_general_field_matches_interpreted_program(tmp_path, packed):
    """Table path, int8 image (one load per read) and 2-bit packed image
    (per-residue word pointers and shifts, one bit-field extract per read)."""
    if not os.path.exists(DUMP):
        pytest.skip("tests/cpp/jit_source_dump not built (run __graft_entry__.build())")
    from casmcode_clexmonte_amd.system import BasisSet, ClusterExpansion, fcc_prim
    prim = fcc_prim(1.0, ("A", "B", "C"))
    bs = BasisSet(prim, {2: 1.05, 3: 0.75}, "chebychev")
    flat = FlatCE(ClusterExpansion(bs, np.random.default_rng(2).normal(0, 0.05, bs.n_functions)))
    P, h, nf, tab_max, prog = _gen_program(flat, 9)
    pf = tmp_path / "prog.txt"
    pf.write_text(prog)
    src = subprocess.run([DUMP, str(pf), "1" if packed else "0"], capture_output=True, text=True,
                         check=True).stdout
    assert ("__builtin_amdgcn_ubfe" in src) == packed
    i = src.index("struct JitField {")
    j = src.index("};\n}  // namespace clexmc")
    host = ((HOST_GEN_P2 if packed else HOST_GEN).replace("JITFIELD", src[i:j + 2]).replace("PADDIM", str(P)).replace("HALO", str(h))
            .replace("NOCCUPANTS", "3").replace("TABSIZE", str(tab_max + 8)))
    cpp = tmp_path / "check.cpp"
    cpp.write_text(host)
    exe = tmp_path / "check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", str(exe), str(cpp)], check=True, capture_output=True)
    r = subprocess.run([str(exe), str(pf)], capture_output=True, text=True, check=True)
    assert int(r.stdout.strip()) == 0
