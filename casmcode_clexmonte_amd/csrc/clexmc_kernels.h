// Kernel argument blocks and launch wrappers (internal to the library).
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stddef.h>
#endif

#include "clexmc_device.h"

namespace clexmc {

struct RunArgs {
  const DevSystem *sys;
  const DevSupercell *scs;
  const ChainParams *params;
  ChainState *state;
  int8_t *occ;
  long long *binC;     // [n_chains][n_orbits] exact binary correlation counts
  double *sampE;       // [n_chains][sample_cap]
  int *sampN;          // [n_chains][sample_cap][kMaxOcc]
  int *sample_count;   // [n_chains]
  int sample_cap;
  int n_passes;
  int sample_period;
  int rounds;
  int n_orbits;
  int n_tab;           // general path: flower table floats staged in LDS
  int global_swaps;    // canonical: serial swaps of uniformly random site pairs per pass
  const int *chain_map; // workgroup -> chain (nullptr: workgroup i runs chain i)
  // slabs > 1: workgroup w runs slab w % slabs of chain w / slabs; it updates
  // only the sites of its slab's planes (minus a frozen band of kreach
  // planes), records what it added to the chain at the sampled passes and at
  // the end (slab_rec), and leaves samples, chain totals and the pass / draw
  // counters to slab_post_kernel
  int slabs;
  SlabRecord *slab_rec;  // [n_chains][kSlabMax][kSlabRecords]
  // semi-grand canonical multiswaps (n_mswaps > 0; the sweep is then the
  // fixed-composition canonical one): after every pass, mswap_per_pass serial
  // attempts of a multiswap type drawn uniformly; type t has the parts
  // mswap_parts[mswap_begin[t] .. mswap_begin[t + 1]) = (asymmetric unit,
  // occupant from, occupant to, count), and mswap_rev[t] is the type that
  // undoes it
  int n_mswaps;
  int mswap_per_pass;
  const int4 *mswap_parts;
  const int *mswap_begin;
  const int *mswap_rev;
};

// A fused launch of chains on several supercell shapes (one specialised
// kernel for all of them): group g owns blocks [block_begin[g],
// block_begin[g + 1]), maps its workgroups to chains through chain_map[g]
// and runs slabs[g] workgroups per chain.  The RunArgs fields chain_map and
// slabs are not used by such a launch.
constexpr int kMaxFusedGroups = 16;
struct FusedGroups {
  int block_begin[kMaxFusedGroups + 1];
  int slabs[kMaxFusedGroups];
  const int *chain_map[kMaxFusedGroups];
};

struct DeltaArgs {
  const DevSystem *sys;
  const DevSupercell *scs;
  const ChainParams *params;
  const int8_t *occ;
  int n_prop;
  int sites_per_prop;
  int include_chem;
  const int *chain;
  const int *sites;
  const int *new_occ;
  float *dE;
};

struct CorrArgs {
  const DevSystem *sys;
  const DevSupercell *scs;
  const ChainParams *params;
  const int8_t *occ;
  int chain0;            // first chain; blockIdx.y selects chain0 + y
  int n_corr;
  double *out;           // [n_chains][n_corr]
};

struct LearnArgs {
  const MlpDev *mlp;
  int rounds_per_group;
};

struct CollectArgs {
  const DevSystem *sys;
  const DevSupercell *scs;
  const ChainParams *params;
  const int8_t *occ;
  const int4 *env_sites;   // [n_sublat][kEnvMax] (b, di, dj, dk)
  int n_env;
  int n_per_chain;
  uint32_t seed;
  int include_chem;
  int8_t *X;               // [n_chains * n_per_chain][kEnvMax]
  float *y;                // [n_chains * n_per_chain]
  float *t;                // [n_chains * n_per_chain] logit scale beta * kMlpEnergyUnit
};

struct TrainArgs {
  int n;
  int n_env;
  const int8_t *X;         // [n][kEnvMax]
  const float *y;          // [n]
  const float *t;          // [n] logit scale of each sample (nullptr: 1)
  const float *W1, *b1, *W2, *b2;   // views into the FP32 parameter vector
  float *H;                // [kHidden][n] pre-activations
  float *dl;               // [n]
  float *loss;             // [n]
  float *grad;             // [n_params]
};

struct AdamArgs {
  int n_params;
  float *p, *m, *v;
  const float *grad;
  float lr, b1m, b2m, b1t, b2t;  // b1t = b1m^t, b2t = b2m^t
};

// N-fold way (rejection-free) semi-grand canonical runs.
// Per block: [0] sum w, [1] sum w E, [2] sum w E^2, [3 + o] sum w N_o,
// [7 + t] sum w N_o N_q (o <= q in row order: t = 0 for (0,0), 1 (0,1), 2
// (0,2), 3 (0,3), 4 (1,1), 5 (1,2), 6 (1,3), 7 (2,2), 8 (2,3), 9 (3,3)),
// [17 + o] sum w E N_o; E the energy per supercell, N_o the sites holding
// occupant o, w the residence weight of the state.
constexpr int kNfoldBlockSums = 21;
// LDS bytes of the N-fold kernel's neighbour-list starts ([kMaxSublat + 1] ints)
constexpr int kNfoldNbBeginBytes = ((kMaxSublat + 1) * 4 + 15) & ~15;
struct NfoldArgs {
  RunArgs a;               // system, supercells, chain params / state, occupations, exact counts
  int n_events;            // events per chain in this launch
  int events_per_block;    // events per output record
  int n_blocks;            // n_events / events_per_block
  double *blocks;          // [n_chains][n_blocks][kNfoldBlockSums]
  // sampling by N-fold time: n_time_samples > 0 samples per chain at times
  // 0, sample_dt, 2 sample_dt, ... (passes: expected Metropolis steps per
  // mutable site) into the sample storage of RunArgs; the chain stops after
  // its last sample (n_events is then an upper bound)
  int n_time_samples;
  double sample_dt;
};

// clexmc_delta_energy_batch_noalloc: proposals per call, and the layout of
// its file-scope staging array (per proposal: chain, up to kMaxSitesPerProp
// sites and new occupants, one dE), 40 bytes per proposal, 2.5 MiB in all.
constexpr int kNoAllocMaxProposals = 65536;
constexpr size_t kNoAllocScratchBytes = (size_t)kNoAllocMaxProposals * (4 + 8 * kMaxSitesPerProp + 4);

#ifndef __HIPCC_RTC__
// Device address of the staging array on the current device, looked up once
// per device (hipGetSymbolAddress) and kept for the life of the process.
hipError_t noalloc_scratch(int device, unsigned char **ptr);
hipError_t launch_learned(const RunArgs &args, const LearnArgs &la, int kind, int weights, int n_chains,
                          int threads, size_t lds, hipStream_t st, bool launch);
size_t learned_lds_bytes(int pad_bytes);
hipError_t launch_mlp_infer(const MlpDev *mlp, int n, const int8_t *X, float *out, hipStream_t st);
hipError_t launch_metropolis_split(const RunArgs &args, int kind, bool packed, int n_chains, int threads, size_t lds,
                                   hipStream_t st, bool launch);
hipError_t launch_lds_probe(int mode, int iters, int blocks, unsigned *sink, hipStream_t st);
hipError_t launch_wave_pick(int n_rows, const float *v, const float *u, float *scan_out, int *pick_out,
                            hipStream_t st);
hipError_t launch_mlp_collect(const CollectArgs &a, int n_chains, hipStream_t st);
hipError_t launch_mlp_train_step(const TrainArgs &a, hipStream_t st);
hipError_t launch_mlp_adam(const AdamArgs &a, hipStream_t st);
hipError_t launch_mlp_publish(const float *p, MlpDev *m, int n_env, float eps, hipStream_t st);

hipError_t launch_metropolis(const RunArgs &args, bool binary, int kind, bool packed, int n_chains, int threads,
                             size_t lds, hipStream_t st, bool launch);
hipError_t launch_slab_post(const RunArgs &a, int n_chains, hipStream_t st);
hipError_t launch_delta_batch(const DeltaArgs &a, hipStream_t st);
hipError_t launch_corr(const CorrArgs &a, int n_corr, int n_chains, hipStream_t st);
hipError_t launch_philox(int n, const uint4 *ctr, const uint2 *key, uint4 *out, hipStream_t st);
size_t metropolis_lds_bytes(int pad_sites, int n_tab, bool binary, int n_func_max, int max_class, bool packed);
hipError_t launch_nfold(const NfoldArgs &a, bool binary, int n_chains, int threads, size_t lds, hipStream_t st);
// nb_total: neighbour-list entries of the supercell, all sublattices
size_t nfold_lds_bytes(int pad_sites, int n_sites, int nb_max, int nb_total);

#endif  // __HIPCC_RTC__

}  // namespace clexmc
