// Device-side data structures and helpers shared by the clexmc kernels.
//
// Layout in device memory (HBM):
//  * per chain: int8 occupation (N sites, linear index l = b*n_cells + cell);
//  * per supercell: padded image geometry, independent-set classes as padded
//    LDS addresses, halo copy lists and, per sublattice, the "site program":
//    the padded LDS offsets of every other site of every cluster that
//    contains the site, grouped per orbit (binary path) or per cluster with
//    its table offset (general path);
//  * system: flower tables (FP32 for the kernels, FP64 for correlations).
// In LDS (one workgroup per chain): the padded occupation image
//   [b][k + h][j + h][i + h] (int8), so every neighbour of every site is a
//   constant offset away; h is the largest cell offset of any cluster.
// Read-only tables are read through the constant address space so that the
// uniform per-cluster offsets become scalar (SMEM) loads.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

namespace clexmc {

constexpr int kMaxSublat = 8;
constexpr int kMaxOcc = 4;
constexpr int kMaxFunc = 3;
constexpr int kMaxOrbits = 64;   // correlations kernel / exact counts storage
constexpr int kMaxSlots = 8;     // binary path: orbits with >= 1 site (registers)
constexpr int kMaxSitesPerProp = 4;
constexpr int kMaxMswapSites = 8;    // sites changed by one multiswap
constexpr int kMaxMswapParts = 4;    // (unit, from, to, count) parts of one multiswap type
constexpr int kMaxMswapTypes = 256;
// LDS of the multiswap phase: pool counts [kMaxSublat][kMaxOcc], 8 control
// words and the picked sites
constexpr int kMswapLdsBytes = ((kMaxSublat * kMaxOcc + 8 + kMaxMswapSites) * 4 + 15) & ~15;
constexpr int kMaxAutoRounds = 1024; // canonical matchings per independent set and pass (automatic)
constexpr int kAotMaxThreads = 512;  // launch bound of the ahead-of-time Metropolis kernels
constexpr int kMaxJitThreads = 1024; // launch bound of run-time specialised kernels for large classes

// Automatic canonical rounds of one independent set.  A random matching of
// the set's n sites pairs a given site with each other site with probability
// 1/n, so a round makes m / (2n) attempts between different occupants on
// average, m = n^2 - sum_o n_o^2 (twice the set's unordered pairs of
// different occupants).  Giving every set with m > 0 ceil(2 N n / M) rounds,
// M = sum of m over the sets and N = number of mutable sites, makes about N
// such attempts per pass in total, spread over the sets in proportion to
// their pairs of different occupants: the reference's steps per pass, each a
// swap of two different occupants.  The counts depend only on occupant counts
// per set, which swaps inside the sets conserve, so the class sweep remains a
// fixed composition of stationary moves.
// Automatic pair-phase rounds: independent sets never exchange occupants
// among themselves, so with automatic rounds a pass also runs enough pair
// phases (disjoint non-interacting pairs across sets and sublattices) for
// about n_mutable / 4 pair attempts, at most kMaxPairRounds.
constexpr int kMaxPairRounds = 64;
__host__ __device__ inline int auto_pair_rounds(int n_mutable, int n_pairs, int n_phases) {
  if (n_phases <= 0 || n_pairs <= 0) return 0;
  const long long per = (n_pairs + n_phases - 1) / n_phases;
  const long long r = (n_mutable + 4 * per - 1) / (4 * per);
  return r < 1 ? 1 : (r > kMaxPairRounds ? kMaxPairRounds : (int)r);
}

__host__ __device__ inline int auto_canonical_rounds(int n, long long m, long long n_mutable, long long m_total) {
  if (n < 2 || m <= 0 || m_total <= 0) return 0;
  const long long r = (2 * n_mutable * n + m_total - 1) / m_total;
  return r > kMaxAutoRounds ? kMaxAutoRounds : (int)r;
}
// learned proposal: local-environment MLP (n_env <= kEnvMax inputs, kHidden units)
constexpr int kEnvMax = 48;
constexpr int kHidden = 64;
// The network output z is a flip-energy score in units of kMlpEnergyUnit
// (eV); a chain at inverse temperature beta uses the logit beta*unit*z, so
// one network serves every temperature of a batch.
constexpr float kMlpEnergyUnit = 0.1f;

template <typename T>
using cptr = const __attribute__((address_space(4))) T *;
template <typename T>
__device__ __forceinline__ cptr<T> cst(const T *p) {
  return (cptr<T>)(p);
}

struct DevSystem {
  int n_sublat, n_occ_max, n_func_max, n_orbits, n_corr, binary;
  int n_occ[kMaxSublat];
  int asym[kMaxSublat];         // asymmetric unit (canonical swaps stay within one)
  int n_func[kMaxSublat];
  float site_func[kMaxSublat][kMaxFunc][kMaxOcc];
  // flower (general table path, delta batch)
  const int *flower_begin;      // [n_sublat+1]
  const int *entry_k;           // [n_entries]
  const int *entry_site_begin;  // [n_entries+1]
  const int4 *entry_sites;      // [n_entry_sites] (b, di, dj, dk)
  const int *entry_mul;         // [n_entry_sites] radix multiplier of each other site
  const int *entry_table;       // [n_entries] value offset into flower_tables
  const float *flower_tables;
  // binary path: slots = orbits with >= 1 site
  int n_slots;
  int slot_orbit[kMaxSlots];
  double bin_w64[kMaxOrbits];   // J_a / m_a per orbit (energy from exact counts)
  // correlations
  const int *orbit_k_all;       // [n_orbits]
  const int *orbit_mult;
  const int *orbit_nfunc;
  const int *orbit_func_begin;
  const int *corr_orbit_begin;  // [n_orbits+1]
  const int *corr_elem_begin;   // [n_orbits+1]
  const int4 *corr_elem_sites;
  const int *orbit_table_offset;
  const double *orbit_tables;
  const int *func_orbit;        // [n_corr] orbit of each function
};

// Binary site program of one sublattice (per supercell).
struct BinSub {
  int slot_k[kMaxSlots];      // other sites per cluster (0: point)
  int slot_n[kMaxSlots];      // clusters of the orbit containing the site
  int slot_begin[kMaxSlots];  // first offset (index into bin_off)
  int slot_len[kMaxSlots];    // clusters after padding to a multiple of 4
  int slot_dummy[kMaxSlots];  // padding clusters (all offsets 0 = the site itself)
  float slot_w[kMaxSlots];    // J_a / m_a
};

// General site program of one sublattice: gen_prog[begin .. begin+len) holds,
// per cluster: k, table_offset, then k pairs (offset, radix multiplier).
struct GenSub {
  int begin, n_entries, nf, pad;
};

struct DevSupercell {
  int L1, L2, L3, h;
  int P1, P2, P3, S;      // padded dims, S = P1*P2*P3 bytes per sublattice
  int n_cells, n_sites, pad_bytes;
  int n_classes;
  const int *class_begin;  // [n_classes+1]
  const int *class_sublat; // [n_classes]
  const int *class_addr;   // padded LDS address of each class site
  const int *halo_begin;   // [n_classes+1]
  const int2 *halo;        // (src, dst) padded addresses
  const int *halo_site;    // [n_class_sites+1] first halo entry of each class site
  const int *site_tpos;    // [n_sites] class-list position of each site (-1: fixed site)
  const int *entry_off;    // [n_entry_sites] padded relative offset (general path)
  const BinSub *bin_sub;   // [n_sublat]
  const int *bin_off;      // binary site programs
  const GenSub *gen_sub;   // [n_sublat]
  const int *gen_prog;     // general site programs
  int n_pair_phases;
  const int *pair_begin;   // [n_pair_phases+1]
  const int2 *pair_addr;   // padded addresses of both sites
  const int *pair_sublat;  // [n_pairs] sublattices of the two sites: b | (b2 << 8)
  const int *pair_halo_begin; // [n_pair_phases+1]
  const int2 *pair_halo;
  int max_class;           // largest class size
  const int *env_off;      // [n_sublat][kEnvMax] padded offsets of the environment sites
  // interaction neighbours per sublattice (the site itself included): the
  // sites whose change of occupation changes a flip energy of the site
  int nb_max;              // largest list
  const int *nb_begin;     // [n_sublat+1]
  const int4 *nb_cell;     // (sublattice, di, dj, dk) cell offsets
  int nb_distinct;         // 1: no two offsets of a list reach the same site of this supercell
  // slabs: largest |dk| between two sites that share a cluster, and per class
  // the index (relative to class_begin) of its first site with cell k >= K,
  // K = 0..L3 (class sites are in linear-index order, so k is sorted)
  int kreach;
  const int *class_kidx;   // [n_classes][L3+1]
};

// Learned-proposal network (device copy).  FP32 master weights are kept by
// the host-side trainer; inference uses the BF16 copy of W1.
struct MlpDev {
  int n_env;                       // inputs used (<= kEnvMax), rest zero
  float eps;                       // weight floor: w = eps + sigmoid(logit)
  float b2;
  int pad;
  float b1[kHidden];
  float W2[kHidden];
  uint16_t W1b[kHidden * kEnvMax]; // bf16 bits, [hidden][input]
};

// Slab launches: what one workgroup has added to its chain since the launch
// started, recorded at each sampled pass and at the launch's last pass
// (dI: binary exact-count slots, then occupant counts, attempts, accepts).
constexpr int kSlabMax = 64;       // slabs per chain
constexpr int kSlabEpochPasses = 4; // passes per slab launch
constexpr int kSlabRecords = kSlabEpochPasses + 1;
struct SlabRecord {
  double dE;
  long long dI[kMaxSlots + kMaxOcc + 2];
};

// Read-only per-chain parameters (constant during a launch).
struct ChainParams {
  int sc;
  int occ_offset;          // into the occupation buffer
  float beta;
  int pad0;
  double beta64;
  float chem[kMaxSublat * kMaxOcc];
  uint32_t key0, key1;
};

// Mutable per-chain state.
struct ChainState {
  unsigned long long draw_base;
  unsigned long long pass_base;
  double E;                // running formation energy per supercell (FP64)
  long long n_attempt, n_accept;
  long long counts[kMaxOcc];
  // FP64 observable accumulators over recorded samples
  double sumE, sumE2, sumN1, sumN12, sumEN1;
  long long n_samples_acc;
  unsigned long long draw_hi;  // slab launches: largest Philox draw counter used
  unsigned long long flags;    // kChainFlag* conditions raised by a kernel, cleared by the host
};

// Whole-chain Metropolis launches hold the chain's running totals in LDS for
// the launch: thread 0 copies them in at the start, updates them at every
// bookkeeping step (LDS instead of a chain of dependent global
// read-modify-writes while the workgroup waits at a barrier) and writes them
// back at the end.
struct ChainLds {
  double E, sumE, sumE2, sumN1, sumN12, sumEN1;
  long long counts[kMaxOcc];
  long long n_attempt, n_accept, n_samples_acc;
  int sample_count, pad_;
  long long C[kMaxOrbits];  // binary: exact correlation counts
};
constexpr int kChainLdsBytes = (int)((sizeof(ChainLds) + 15) & ~(size_t)15);

// N-fold way: the chain's total rate fell below the smallest the kernel
// accepts (no event can be drawn reliably); the run stopped early.
constexpr unsigned long long kChainFlagNfoldStalled = 1ull;
// Multiswaps: a workgroup-wide search found fewer candidate sites than the
// kernel's pool counts (never expected: the counts follow every move); the
// attempt was rejected.
constexpr unsigned long long kChainFlagMswapCounts = 2ull;

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies with the
// published multipliers and Weyl key increments.  Each product is written as
// one 64-bit multiply: on gfx950 that is a single v_mad_u64_u32 instead of a
// v_mul_hi_u32 / v_mul_lo_u32 pair.
// ---------------------------------------------------------------------------
__host__ __device__ inline uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t m0 = 0xD2511F53u, m1 = 0xCD9E8D57u;
    const uint64_t p0 = (uint64_t)m0 * c.x, p1 = (uint64_t)m1 * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// 24 random bits -> float in [0,1) (site and occupant choices)
__device__ inline float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// Acceptance with probability p from a 32-bit uniform word x: the test
// u < p for u = (x + y / 2^32) / 2^32 with y a second uniform word.  With
// t = p * 2^32, x < floor(t) accepts and x > floor(t) rejects; only
// x == floor(t) (chance 2^-32) calls more() for y and accepts if
// y < (t - floor(t)) * 2^32.  The acceptance probability is p (to FP32
// rounding of p) plus at most 2^-64, for every p down to the smallest
// normal float (1.2e-38, beta dE ~ 87); a single 24-bit uniform would accept
// any change with probability at least 2^-24.  p >= 1 accepts; p == 0 or NaN
// rejects.
// The common path is branch-free (compares and selects); only the tie takes
// a branch.
template <class MORE>
__host__ __device__ inline bool accept_prob(float p, uint32_t x, MORE more) {
  const float t = p * 4294967296.0f;
  const bool pos = t > 0.0f;                  // false for 0 and NaN
  const bool all = t >= 4294967296.0f;        // p >= 1
  const bool part = pos && !all;
  const uint32_t ti = part ? (uint32_t)t : 0u;
  bool acc = all || x < ti;
  if (part && x == ti) acc = (float)more() < (t - (float)ti) * 4294967296.0f;
  return acc;
}

}  // namespace clexmc
