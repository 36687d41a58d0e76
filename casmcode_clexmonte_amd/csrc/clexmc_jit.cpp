// Run-time specialisation of the binary Metropolis kernel (see clexmc_jit.h).
#include "clexmc_jit.h"

#include <hip/hiprtc.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <sstream>

#include "clexmc_jit_src.inc"  // kJitHeaderNames / kJitHeaderTexts: the device headers as text

namespace clexmc {

// Lane-pair split of one sublattice's binary site program: every orbit slot
// goes wholly to one lane of the pair, choosing among all 2^n assignments the
// one whose busier lane has the least work, counted as 2 per distinct LDS
// neighbour read plus 1 per XOR / add, plus 48 on lane 0 for the proposal's
// random number, acceptance and update.  Orbits whose clusters share
// neighbours (pairs and the triplets built on them) can then stay on one
// lane, so the pair reads each neighbour about once.
static std::vector<int> split_slots(const BinSub &B, const std::vector<int> &bin_off, int n_slots) {
  std::vector<int> part(kMaxSlots, 0);
  std::vector<int> active;
  for (int j = 0; j < n_slots; ++j)
    if (B.slot_n[j] > 0 && B.slot_k[j] > 0) active.push_back(j);
  const int na = (int)active.size();
  if (na == 0 || na > 16) return part;
  long best = -1;
  int best_mask = 0;
  for (int mask = 0; mask < (1 << na); ++mask) {
    long cost[2] = {48, 0};
    std::vector<int> offs[2];
    for (int q = 0; q < na; ++q) {
      const int j = active[q], p = (mask >> q) & 1;
      const int n = B.slot_n[j], ko = B.slot_k[j];
      cost[p] += (long)n * ko;
      for (int e = 0; e < n * ko; ++e) offs[p].push_back(bin_off[B.slot_begin[j] + e]);
    }
    for (int p = 0; p < 2; ++p) {
      std::sort(offs[p].begin(), offs[p].end());
      cost[p] += 2 * (long)(std::unique(offs[p].begin(), offs[p].end()) - offs[p].begin());
    }
    const long c = std::max(cost[0], cost[1]);
    if (best < 0 || c < best) {
      best = c;
      best_mask = mask;
    }
  }
  for (int q = 0; q < na; ++q) part[active[q]] = (best_mask >> q) & 1;
  return part;
}

// `base` = pa + minoff with every read at base + d, d >= 0: an empty asm
// statement makes base opaque, so the compiler cannot fold minoff back into
// the offsets (negative offsets do not fit the LDS instructions' unsigned
// immediate and cost one address add per read)
static std::string base_line(int minoff) {
  return "      int base = pa + (" + std::to_string(minoff) + ");\n      asm volatile(\"\" : \"+v\"(base));\n";
}

static std::string hexf(float v) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%af", (double)v);
  return buf;
}

// The field policy struct `name` for one supercell's site programs (inside
// namespace clexmc).
static void emit_cube(std::ostringstream &o, const BinSub &B, const std::vector<int> &bin_off, int n_slots,
                      const CubeSpec &cs);

static void emit_field(std::ostringstream &o, const std::string &name, int n_sublat, const std::vector<int> &variable,
                       bool binary, const std::vector<BinSub> &subs, const std::vector<int> &bin_off, int n_slots,
                       const std::vector<GenSub> &gsubs, const std::vector<int> &gen_prog, bool packed,
                       const CubeSpec *cube = nullptr) {
  o << "struct " << name << " {\n"
    << "  template <class IMG>\n"
    << "  __device__ __forceinline__ static float bin(const FieldCtx &, const IMG &img, int b, int pa, int o,\n"
    << "                                              int (&K)[kMaxSlots]) {\n"
    << "    (void)img; (void)b; (void)pa; (void)o;\n";
  for (int j = 0; j < kMaxSlots; ++j) o << "    K[" << j << "] = 0;\n";
  o << "    float F = 0.0f;\n";
  for (int b = 0; binary && b < n_sublat; ++b) {
    if (!variable[b]) continue;
    const BinSub &B = subs[b];
    int minoff = 0;
    for (int j = 0; j < n_slots; ++j) {
      const int n = B.slot_n[j], ko = B.slot_k[j];
      for (int e = 0; e < n * ko; ++e) minoff = std::min(minoff, bin_off[B.slot_begin[j] + e]);
    }
    // base + (off - minoff) == pa + off, with every immediate offset >= 0
    o << "    if (b == " << b << ") {\n" << base_line(minoff);
    for (int j = 0; j < n_slots; ++j) {
      const int n = B.slot_n[j], ko = B.slot_k[j];
      if (n <= 0) continue;
      o << "      {\n        int X = 0;\n";
      for (int e = 0; e < n && ko > 0; ++e) {
        o << "        X += ";
        for (int i = 0; i < ko; ++i) {
          if (i) o << " ^ ";
          o << "img.get(base + " << (bin_off[B.slot_begin[j] + e * ko + i] - minoff) << ")";
        }
        o << ";\n";
      }
      if (ko & 1)
        o << "        const int Kj = 2 * X - " << n << ";\n";
      else
        o << "        const int Kj = " << n << " - 2 * X;\n";
      o << "        K[" << j << "] = Kj;\n        F = __builtin_fmaf(" << hexf(B.slot_w[j]) << ", (float)Kj, F);\n"
        << "      }\n";
    }
    o << "      return F;\n    }\n";
  }
  o << "    return F;\n  }\n";
  // lane-pair partial sums (orbit slots split between the two lanes) and the finish
  o << "  template <class IMG>\n"
    << "  __device__ __forceinline__ static void bin_x(const FieldCtx &, const IMG &img, int b, int pa, int o, int part,\n"
    << "                                               int (&X)[kMaxSlots]) {\n"
    << "    (void)img; (void)b; (void)pa; (void)o; (void)part;\n";
  for (int j = 0; j < kMaxSlots; ++j) o << "    X[" << j << "] = 0;\n";
  for (int b = 0; binary && b < n_sublat; ++b) {
    if (!variable[b]) continue;
    const BinSub &B = subs[b];
    int minoff = 0;
    for (int j = 0; j < n_slots; ++j)
      for (int e = 0; e < B.slot_n[j] * B.slot_k[j]; ++e) minoff = std::min(minoff, bin_off[B.slot_begin[j] + e]);
    o << "    if (b == " << b << ") {\n" << base_line(minoff);
    const std::vector<int> slot_part = split_slots(B, bin_off, n_slots);
    for (int part = 0; part < 2; ++part) {
      o << "      if (part == " << part << ") {\n";
      for (int j = 0; j < n_slots; ++j) {
        const int n = B.slot_n[j], ko = B.slot_k[j];
        if (n <= 0 || ko <= 0 || slot_part[j] != part) continue;
        for (int e = 0; e < n; ++e) {
          o << "        X[" << j << "] += ";
          for (int i = 0; i < ko; ++i) {
            if (i) o << " ^ ";
            o << "img.get(base + " << (bin_off[B.slot_begin[j] + e * ko + i] - minoff) << ")";
          }
          o << ";\n";
        }
      }
      o << "      }\n";
    }
    o << "      return;\n    }\n";
  }
  o << "  }\n";
  o << "  __device__ __forceinline__ static float bin_finish(const FieldCtx &, int b, const int (&X)[kMaxSlots],\n"
    << "                                                     int (&K)[kMaxSlots]) {\n"
    << "    (void)b; (void)X;\n";
  for (int j = 0; j < kMaxSlots; ++j) o << "    K[" << j << "] = 0;\n";
  o << "    float F = 0.0f;\n";
  for (int b = 0; binary && b < n_sublat; ++b) {
    if (!variable[b]) continue;
    const BinSub &B = subs[b];
    o << "    if (b == " << b << ") {\n";
    for (int j = 0; j < n_slots; ++j) {
      const int n = B.slot_n[j], ko = B.slot_k[j];
      if (n <= 0) continue;
      if (ko & 1)
        o << "      K[" << j << "] = 2 * X[" << j << "] - " << n << ";\n";
      else
        o << "      K[" << j << "] = " << n << " - 2 * X[" << j << "];\n";
      o << "      F = __builtin_fmaf(" << hexf(B.slot_w[j]) << ", (float)K[" << j << "], F);\n";
    }
    o << "      return F;\n    }\n";
  }
  o << "    return F;\n  }\n";
  // general table path: same entry order and accumulation as general_field
  o << "  template <class IMG>\n"
    << "  __device__ __forceinline__ static void gen(const FieldCtx &, const IMG &img, int b, int pa, const float *tab,\n"
    << "                                             float (&G)[kMaxFunc]) {\n"
    << "    (void)img; (void)b; (void)pa; (void)tab;\n"
    << "    G[0] = G[1] = G[2] = 0.0f;\n";
  for (int b = 0; !binary && b < n_sublat; ++b) {
    if (!variable[b]) continue;
    const GenSub &S = gsubs[b];
    int minoff = 0;
    {
      int q = S.begin;
      for (int e = 0; e < S.n_entries; ++e) {
        const int k = gen_prog[q];
        q += 2;
        for (int i = 0; i < k; ++i, q += 2) minoff = std::min(minoff, gen_prog[q]);
      }
    }
    o << "    if (b == " << b << ") {\n" << base_line(minoff);
    if (packed) {
      // 2-bit image, 16 sites per word: a read at base + d takes word
      // (base >> 4) + (d >> 4) + carry and bit field ((base & 15) + (d & 15)) & 15,
      // where carry and field depend only on d & 15 -- so one word pointer and
      // one shift per residue used, and each read is a load with a constant
      // offset and one bit-field extract
      bool used[16] = {};
      int q = S.begin;
      for (int e = 0; e < S.n_entries; ++e) {
        const int k = gen_prog[q];
        q += 2;
        for (int i = 0; i < k; ++i, q += 2) used[(gen_prog[q] - minoff) & 15] = true;
      }
      o << "      const uint32_t *wb_ = img.w + (base >> 4);\n      const int rb_ = base & 15;\n";
      for (int r = 0; r < 16; ++r)
        if (used[r])
          o << "      const uint32_t *w" << r << "_ = wb_ + ((rb_ + " << r << ") >> 4);\n"
            << "      const uint32_t s" << r << "_ = (uint32_t)((rb_ + " << r << ") & 15) << 1;\n";
    }
    int q = S.begin;
    for (int e = 0; e < S.n_entries; ++e) {
      const int k = gen_prog[q], toff = gen_prog[q + 1];
      q += 2;
      o << "      {\n        const int idx = 0";
      for (int i = 0; i < k; ++i, q += 2) {
        const int d = gen_prog[q] - minoff;
        if (packed)
          o << " + (int)__builtin_amdgcn_ubfe(w" << (d & 15) << "_[" << (d >> 4) << "], s" << (d & 15) << "_, 2u) * "
            << gen_prog[q + 1];
        else
          o << " + img.get(base + " << d << ") * " << gen_prog[q + 1];
      }
      o << ";\n        const float *V = tab + " << toff << " + idx * " << S.nf << ";\n";
      for (int f = 0; f < S.nf && f < kMaxFunc; ++f) o << "        G[" << f << "] += V[" << f << "];\n";
      o << "      }\n";
    }
    o << "      return;\n    }\n";
  }
  o << "  }\n";
  if (cube && binary && n_sublat == 1)
    emit_cube(o, subs[0], bin_off, n_slots, *cube);
  else
    o << "  static constexpr bool kCube = false;\n  struct Cube {};\n";
  o << "};\n";
}

// Cube cell q (x + S y + S^2 z, S = 3 + 2h) is bit q & 31 of register q >> 5
// (binary occupants: one bit per cell, so the cube takes few registers).
static std::string cube_bit(int q) {
  return "__builtin_amdgcn_ubfe(cb.r[" + std::to_string(q >> 5) + "], " + std::to_string(q & 31) + "u, 1u)";
}

static void emit_cube(std::ostringstream &o, const BinSub &B, const std::vector<int> &bin_off, int n_slots,
                      const CubeSpec &cs) {
  const int h = cs.h, S = 3 + 2 * h, n_cells = S * S * S, nr = (n_cells + 31) / 32;
  // cell offsets (dx, dy, dz) of the padded LDS offsets (|d| <= h)
  auto decode = [&](int off, int &dx, int &dy, int &dz) {
    dz = (int)std::lround((double)off / cs.P12);
    const int r = off - dz * cs.P12;
    dy = (int)std::lround((double)r / cs.P1);
    dx = r - dy * cs.P1;
  };
  o << "  static constexpr bool kCube = true;\n"
    << "  static constexpr int kCubeBlocks[3] = {" << cs.nb[0] << ", " << cs.nb[1] << ", " << cs.nb[2] << "};\n"
    << "  struct Cube {\n    uint32_t r[" << nr << "];\n  };\n";
  // load: every cube cell from the padded image (halos valid at launch start)
  o << "  __device__ __forceinline__ static void cube_load(Cube &cb, const ImgU8 &img, int base) {\n"
    << "    asm volatile(\"\" : \"+v\"(base));\n";
  for (int w = 0; w < nr; ++w) {
    o << "    cb.r[" << w << "] = 0u";
    for (int t = 0; t < 32; ++t) {
      const int q = 32 * w + t;
      if (q >= n_cells) break;
      const int x = q % S, y = (q / S) % S, z = q / (S * S);
      o << " | ((uint32_t)img.p[base + " << (z * cs.P12 + y * cs.P1 + x) << "] << " << t << ")";
    }
    o << ";\n";
  }
  o << "  }\n";
  // one pass over the 27 classes: per class the field of the lane's site from
  // the cube, the proposal (site callback: draws, accepts, accumulates), the
  // lane's own byte and LDS image updated, a barrier, then the other lanes'
  // sites of the class read back into the cube.  ax/ay/az: padded
  // coordinates (times pitch) of the block origins at offsets -1, 0, +1.
  // Empty asm statements at each class start mark the cube and the block
  // origins as changed: otherwise the compiler shares bit extractions and
  // refresh addresses among classes (and hoists them out of the pass loop)
  // and keeps hundreds of values alive, spilling to scratch memory.
  o << "  template <class SITE>\n"
    << "  __device__ __forceinline__ static void cube_pass(Cube &cb, const ImgU8 &img, const int (&ax)[3],\n"
    << "                                                   const int (&ay)[3], const int (&az)[3], SITE &&site) {\n"
    << "    int x0 = ax[0], x1 = ax[1], x2 = ax[2], y0 = ay[0], y1 = ay[1], y2 = ay[2], z0 = az[0], z1 = az[1],"
    << " z2 = az[2];\n";
  std::string fence = "      asm volatile(\"\" :";
  for (int w = 0; w < nr; ++w) fence += std::string(w ? "," : "") + " \"+v\"(cb.r[" + std::to_string(w) + "])";
  fence += ");\n      asm volatile(\"\" : \"+v\"(x0), \"+v\"(x1), \"+v\"(x2), \"+v\"(y0), \"+v\"(y1), "
           "\"+v\"(y2), \"+v\"(z0), \"+v\"(z1), \"+v\"(z2));\n";
  const char *xs[3] = {"x0", "x1", "x2"}, *ys[3] = {"y0", "y1", "y2"}, *zs[3] = {"z0", "z1", "z2"};
  for (int c = 0; c < 27; ++c) {
    const int r1 = cs.res[c][0], r2 = cs.res[c][1], r3 = cs.res[c][2];
    const int P = (h + r1) + S * (h + r2) + S * S * (h + r3);
    o << "    {\n" << fence << "      int K[kMaxSlots];\n";
    for (int j = 0; j < kMaxSlots; ++j) o << "      K[" << j << "] = 0;\n";
    o << "      float F = 0.0f;\n";
    for (int j = 0; j < n_slots; ++j) {
      const int n = B.slot_n[j], ko = B.slot_k[j];
      if (n <= 0) continue;
      o << "      {\n        int X = 0;\n";
      if (ko == 1) {
        // pairs: the orbit's neighbour bits counted word by word (mask, popcount)
        std::map<int, uint32_t> masks;
        for (int e = 0; e < n; ++e) {
          int dx, dy, dz;
          decode(bin_off[B.slot_begin[j] + e], dx, dy, dz);
          const int q = P + dx + S * dy + S * S * dz;
          masks[q >> 5] |= 1u << (q & 31);
        }
        for (auto const &wm : masks)
          o << "        X += __builtin_popcount(cb.r[" << wm.first << "] & " << wm.second << "u);\n";
      }
      for (int e = 0; e < n && ko > 1; ++e) {
        o << "        X += ";
        for (int i = 0; i < ko; ++i) {
          int dx, dy, dz;
          decode(bin_off[B.slot_begin[j] + e * ko + i], dx, dy, dz);
          const int q = P + dx + S * dy + S * S * dz;
          if (i) o << " ^ ";
          o << "(int)" << cube_bit(q);
        }
        o << ";\n";
      }
      if (ko & 1)
        o << "        const int Kj = 2 * X - " << n << ";\n";
      else
        o << "        const int Kj = " << n << " - 2 * X;\n";
      o << "        K[" << j << "] = Kj;\n        F = __builtin_fmaf(" << hexf(B.slot_w[j]) << ", (float)Kj, F);\n"
        << "      }\n";
    }
    const int own = r3 * cs.P12 + r2 * cs.P1 + r1;
    o << "      const int o = (int)" << cube_bit(P) << ";\n"
      << "      if (site(F, K, o)) {\n"
      << "        cb.r[" << (P >> 5) << "] ^= " << (1u << (P & 31)) << "u;\n"
      << "        img.set(z1 + y1 + x1 + " << own << ", o, 1 - o);\n"
      << "      }\n      __syncthreads();\n";
    for (int bk = -1; bk <= 1; ++bk)
      for (int bj = -1; bj <= 1; ++bj)
        for (int bi = -1; bi <= 1; ++bi) {
          if (bi == 0 && bj == 0 && bk == 0) continue;
          const int X = h + 3 * bi + r1, Y = h + 3 * bj + r2, Z = h + 3 * bk + r3;
          if (X < 0 || X >= S || Y < 0 || Y >= S || Z < 0 || Z >= S) continue;
          const int q = X + S * Y + S * S * Z;
          o << "      cb.r[" << (q >> 5) << "] = (cb.r[" << (q >> 5) << "] & " << ~(1u << (q & 31))
            << "u) | ((uint32_t)img.p[" << zs[bk + 1] << " + " << ys[bj + 1] << " + " << xs[bi + 1] << " + " << own
            << "] << " << (q & 31) << ");\n";
        }
    o << "    }\n";
  }
  o << "  }\n";
}

std::string jit_source(int n_sublat, const std::vector<int> &variable, bool binary, const std::vector<BinSub> &subs,
                       const std::vector<int> &bin_off, int n_slots, const std::vector<GenSub> &gsubs,
                       const std::vector<int> &gen_prog, int kind, bool packed, int lanes, int max_threads,
                       const CubeSpec *cube) {
  std::ostringstream o;
  // the semi-grand canonical, unpacked, one-lane variant also carries the N-fold kernel
  const bool with_nfold = kind == 1 && !packed && lanes == 1;
  o << "#include \"clexmc_metropolis.h\"\n" << (with_nfold ? "#include \"clexmc_nfold.h\"\n" : "")
    << "namespace clexmc {\n";
  emit_field(o, "JitField", n_sublat, variable, binary, subs, bin_off, n_slots, gsubs, gen_prog, packed,
             kind == 1 && !packed && lanes == 1 ? cube : nullptr);
  o << "}  // namespace clexmc\n";
  o << "extern \"C\" __global__ __launch_bounds__(" << (lanes == 2 || max_threads > 512 ? 1024 : 512) << ") void "
    << jit_kernel_name(kind, packed, lanes) << "(clexmc::RunArgs a) { clexmc::metropolis_body<"
    << (binary ? "true" : "false") << ", " << kind << ", " << (packed ? "true" : "false") << ", clexmc::JitField, "
    << lanes << ">(a, clexmc::BlockRole{(int)blockIdx.x, a.chain_map, a.slabs}); }\n";
  if (cube && binary && n_sublat == 1 && kind == 1 && !packed && lanes == 1)
    o << "extern \"C\" __global__ __launch_bounds__(" << (max_threads > 512 ? 1024 : 512) << ") void "
      << kJitCubeName << "(clexmc::RunArgs a) { clexmc::metropolis_body<true, 1, false, clexmc::JitField, 1, true>"
      << "(a, clexmc::BlockRole{(int)blockIdx.x, a.chain_map, a.slabs}); }\n";
  if (with_nfold)
    o << "extern \"C\" __global__ __launch_bounds__(256) void " << kJitNfoldName
      << "(clexmc::NfoldArgs na) { clexmc::nfold_body<" << (binary ? "true" : "false")
      << ", clexmc::JitField>(na); }\n";
  return o.str();
}

std::string jit_source_fused(int n_sublat, const std::vector<int> &variable, bool binary, int n_slots,
                             const std::vector<JitShape> &shapes, int kind, int max_threads) {
  std::ostringstream o;
  o << "#include \"clexmc_metropolis.h\"\nnamespace clexmc {\n";
  for (size_t g = 0; g < shapes.size(); ++g) {
    const JitShape &J = shapes[g];
    emit_field(o, "JitField" + std::to_string(g), n_sublat, variable, binary, *J.subs, *J.bin_off, n_slots, *J.gsubs,
               *J.gen_prog, J.packed);
  }
  o << "}  // namespace clexmc\n";
  // the group of a block: the last group whose first block is at or below it
  o << "extern \"C\" __global__ __launch_bounds__(" << (max_threads > 512 ? 1024 : 512) << ") void "
    << kJitFusedName << "(clexmc::RunArgs a, clexmc::FusedGroups t) {\n"
    << "  const int blk = (int)blockIdx.x;\n  int g = 0;\n"
    << "  for (int q = 1; q < " << shapes.size() << "; ++q) g += blk >= t.block_begin[q] ? 1 : 0;\n"
    << "  const clexmc::BlockRole role{blk - t.block_begin[g], t.chain_map[g], t.slabs[g]};\n"
    << "  switch (g) {\n";
  for (size_t g = 0; g < shapes.size(); ++g)
    o << "    case " << g << ": clexmc::metropolis_body<" << (binary ? "true" : "false") << ", " << kind << ", "
      << (shapes[g].packed ? "true" : "false") << ", clexmc::JitField" << g << ", 1>(a, role); break;\n";
  o << "    default: break;\n  }\n}\n";
  return o.str();
}

const char *jit_kernel_name(int kind, bool packed, int lanes_per_site) {
  // kind 0 canonical (c), 1 semi-grand canonical (s), 2 with multiswaps (m)
  static const char *names[3][2][2] = {{{"clexmc_jit_c_u8", "clexmc_jit_c_u8_x2"}, {"clexmc_jit_c_p2", "clexmc_jit_c_p2_x2"}},
                                       {{"clexmc_jit_s_u8", "clexmc_jit_s_u8_x2"}, {"clexmc_jit_s_p2", "clexmc_jit_s_p2_x2"}},
                                       {{"clexmc_jit_m_u8", "clexmc_jit_m_u8_x2"}, {"clexmc_jit_m_p2", "clexmc_jit_m_p2_x2"}}};
  return names[kind < 0 ? 0 : (kind > 2 ? 2 : kind)][packed ? 1 : 0][lanes_per_site == 2 ? 1 : 0];
}

// code objects of the sources compiled so far in this process (a handle or
// supercell with the same site programs and variant reuses them)
static std::mutex g_code_mutex;
static std::map<std::string, std::vector<char>> g_code_cache;

hipError_t jit_compile_load(const std::string &src, hipModule_t *mod, std::string *log) {
  {
    std::lock_guard<std::mutex> lock(g_code_mutex);
    auto it = g_code_cache.find(src);
    if (it != g_code_cache.end()) return hipModuleLoadData(mod, it->second.data());
  }
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "clexmc_jit.hip", kJitNumHeaders, kJitHeaderTexts,
                          kJitHeaderNames) != HIPRTC_SUCCESS)
    return hipErrorInvalidValue;
  const char *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  size_t ls = 0;
  hiprtcGetProgramLogSize(prog, &ls);
  if (log && ls > 1) {
    log->assign(ls, '\0');
    hiprtcGetProgramLog(prog, &(*log)[0]);
  }
  if (rc != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return hipErrorInvalidImage;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  std::vector<char> code(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  hipError_t e = hipModuleLoadData(mod, code.data());
  if (e == hipSuccess) {
    std::lock_guard<std::mutex> lock(g_code_mutex);
    g_code_cache.emplace(src, std::move(code));
  }
  return e;
}

}  // namespace clexmc
