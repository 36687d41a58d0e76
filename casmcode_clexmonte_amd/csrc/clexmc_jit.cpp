// Run-time specialisation of the binary Metropolis kernel (see clexmc_jit.h).
#include "clexmc_jit.h"

#include <hip/hiprtc.h>

#include <algorithm>
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
static void emit_field(std::ostringstream &o, const std::string &name, int n_sublat, const std::vector<int> &variable,
                       bool binary, const std::vector<BinSub> &subs, const std::vector<int> &bin_off, int n_slots,
                       const std::vector<GenSub> &gsubs, const std::vector<int> &gen_prog, bool packed) {
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
  o << "  }\n};\n";
}

std::string jit_source(int n_sublat, const std::vector<int> &variable, bool binary, const std::vector<BinSub> &subs,
                       const std::vector<int> &bin_off, int n_slots, const std::vector<GenSub> &gsubs,
                       const std::vector<int> &gen_prog, int kind, bool packed, int lanes, int max_threads) {
  std::ostringstream o;
  // the semi-grand canonical, unpacked, one-lane variant also carries the N-fold kernel
  const bool with_nfold = kind == 1 && !packed && lanes == 1;
  o << "#include \"clexmc_metropolis.h\"\n" << (with_nfold ? "#include \"clexmc_nfold.h\"\n" : "")
    << "namespace clexmc {\n";
  emit_field(o, "JitField", n_sublat, variable, binary, subs, bin_off, n_slots, gsubs, gen_prog, packed);
  o << "}  // namespace clexmc\n";
  o << "extern \"C\" __global__ __launch_bounds__(" << (lanes == 2 || max_threads > 512 ? 1024 : 512) << ") void "
    << jit_kernel_name(kind, packed, lanes) << "(clexmc::RunArgs a) { clexmc::metropolis_body<"
    << (binary ? "true" : "false") << ", " << kind << ", " << (packed ? "true" : "false") << ", clexmc::JitField, "
    << lanes << ">(a, clexmc::BlockRole{(int)blockIdx.x, a.chain_map, a.slabs}); }\n";
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
