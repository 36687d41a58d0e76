// Run-time specialisation of the binary Metropolis kernel to one supercell
// shape: the site programs (padded LDS offsets of every other site of every
// cluster, per orbit) become compile-time constants, so each neighbour read
// is one LDS load with an immediate offset and no per-read address or
// scalar-load work.  Compiled with hipRTC for gfx950 on first use.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "clexmc_device.h"

namespace clexmc {

// Register-cube sweep of the semi-grand canonical binary kernel (one
// sublattice, classes = the 27 residues of the cells mod 3, one lane per 3x3x3
// block of cells, the class lists in block order): each lane keeps the
// (3 + 2h)^3 cells around its block in registers, computes every field from
// them and refreshes, after each class, the cells of that class the other
// lanes changed.
struct CubeSpec {
  int nb[3];        // blocks per axis (L / 3)
  int h;            // halo: the largest cell offset of a cluster
  int P1, P12;      // padded row and plane pitch of the LDS image
  int res[27][3];   // cell residues mod 3 of each class, in class order
};

// Source of the specialised kernels for the site programs of one supercell:
// the binary programs (subs[n_sublat], offsets as laid out by the library)
// when `binary`, else the general table programs (gsubs[n_sublat], gen_prog:
// per cluster k, table offset, then k (offset, radix multiplier) pairs).
// One kernel per (kind, packed, lanes per site) variant; `max_threads` (512 or
// 1024) is its launch bound.
// `cube` (semi-grand canonical, unpacked, one-lane variant only; may be
// nullptr) adds the register-cube sweep.
std::string jit_source(int n_sublat, const std::vector<int> &variable, bool binary, const std::vector<BinSub> &subs,
                       const std::vector<int> &bin_off, int n_slots, const std::vector<GenSub> &gsubs,
                       const std::vector<int> &gen_prog, int kind, bool packed, int lanes, int max_threads,
                       const CubeSpec *cube = nullptr);

// One supercell shape of a fused launch: its site programs (as for
// jit_source) and image format.
struct JitShape {
  const std::vector<BinSub> *subs;
  const std::vector<int> *bin_off;
  const std::vector<GenSub> *gsubs;
  const std::vector<int> *gen_prog;
  bool packed;
};

// Source of one kernel for chains on several supercell shapes (one launch):
// a field struct per shape and a block -> group dispatch (FusedGroups), one
// lane per site; `max_threads` sets the launch bound as for jit_source.
std::string jit_source_fused(int n_sublat, const std::vector<int> &variable, bool binary, int n_slots,
                             const std::vector<JitShape> &shapes, int kind, int max_threads);
constexpr const char *kJitFusedName = "clexmc_jit_fused";

// Kernel name for (kind, packed, lanes per site) in a module built from
// jit_source (two lanes per site: binary expansions only).
const char *jit_kernel_name(int kind, bool packed, int lanes_per_site);

// The register-cube kernel in the module of the semi-grand canonical,
// unpacked, one-lane variant of a supercell with a CubeSpec.
constexpr const char *kJitCubeName = "clexmc_jit_s_u8_cube";

// The N-fold kernel in the module of the semi-grand canonical, unpacked,
// one-lane variant.
constexpr const char *kJitNfoldName = "clexmc_jit_nfold";

// Compiles `src` (hipRTC, gfx950) and loads it as a module; code objects are
// cached per source text for the life of the process.
hipError_t jit_compile_load(const std::string &src, hipModule_t *mod, std::string *log);

}  // namespace clexmc
