// Kernel argument blocks and launch wrappers (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

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

hipError_t launch_metropolis(const RunArgs &args, bool binary, int kind, int n_chains, int threads,
                             size_t lds, hipStream_t st);
hipError_t launch_delta_batch(const DeltaArgs &a, hipStream_t st);
hipError_t launch_corr(const CorrArgs &a, int n_corr, int n_chains, hipStream_t st);
hipError_t launch_philox(int n, const uint4 *ctr, const uint2 *key, uint4 *out, hipStream_t st);
size_t metropolis_lds_bytes(int pad_bytes, int n_tab, bool binary, int threads, int max_class);

}  // namespace clexmc
