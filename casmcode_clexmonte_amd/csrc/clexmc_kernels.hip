// gfx950 kernels of the clexmc Metropolis hot path.
//
//  metropolis_kernel   one workgroup per chain; the chain's occupation lives
//                      in LDS as a halo-padded image (int8 per site, or 2 bits
//                      per site packed 16 to a 32-bit word when the int8 image
//                      does not fit: up to 4 occupants); sites are updated
//                      class by class (independent sets: no cluster contains
//                      two sites of one class), so all proposals of a class
//                      are evaluated and applied concurrently and exactly.
//                      Binary (+-1) expansions use integer cluster-product
//                      counts per orbit (K_a, exact), FP32 dE = dphi*sum w_a K_a
//                      and exact int64 correlation counts C_a per chain; other
//                      expansions use FP32 flower tables (staged in LDS) and
//                      an FP64 running energy.  Philox4x32-10 supplies the
//                      uniforms; lane accumulators are reduced per wavefront
//                      with cross-lane shuffles.
//  delta_batch_kernel  per-proposal dE (BaseMontePotential::occ_delta_per_supercell)
//  corr_kernel         full per-supercell correlations in FP64
//  philox_kernel       device Philox for known-answer checks
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <type_traits>

#include "clexmc_device.h"
#include "clexmc_kernels.h"
#include "clexmc_metropolis.h"

namespace clexmc {

// ---------------------------------------------------------------------------
// Persistent Metropolis kernel (generic site programs)
// ---------------------------------------------------------------------------
template <bool BIN, int KIND, bool PACK>
__global__ __launch_bounds__(512) void metropolis_kernel(RunArgs args) {
  metropolis_body<BIN, KIND, PACK, GenericField>(args, BlockRole{(int)blockIdx.x, args.chain_map, args.slabs});
}

// binary path, two lanes per site in the class sweeps (up to 1024 threads)
template <int KIND, bool PACK>
__global__ __launch_bounds__(1024) void metropolis_split_kernel(RunArgs args) {
  metropolis_body<true, KIND, PACK, GenericField, 2>(args, BlockRole{(int)blockIdx.x, args.chain_map, args.slabs});
}

// ---------------------------------------------------------------------------
// Per-proposal dE in FP32 on unpadded occupation (periodic wrap)
// ---------------------------------------------------------------------------
__global__ void delta_batch_kernel(DeltaArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n_prop) return;
  const DevSystem *__restrict__ sys = a.sys;
  const int chain = a.chain[p];
  const ChainParams &cp = a.params[chain];
  const DevSupercell &sc = a.scs[cp.sc];
  const int8_t *__restrict__ occ = a.occ + cp.occ_offset;
  const int m = a.sites_per_prop;
  const int *sites = a.sites + p * m;
  const int *nocc = a.new_occ + p * m;
  float dE = 0.0f;
  for (int i = 0; i < m; ++i) {
    const int l = sites[i];
    const int b = l / sc.n_cells;
    const int c = l - b * sc.n_cells;
    const int i0 = c % sc.L1, j0 = (c / sc.L1) % sc.L2, k0 = c / (sc.L1 * sc.L2);
    // current occupant of l (with earlier changes of this proposal applied)
    int o = occ[l];
    for (int j = 0; j < i; ++j)
      if (sites[j] == l) o = nocc[j];
    const int nw = nocc[i];
    if (nw == o) continue;
    float G[kMaxFunc] = {0.f, 0.f, 0.f};
    const int nf = sys->n_func[b];
    for (int e = sys->flower_begin[b]; e < sys->flower_begin[b + 1]; ++e) {
      int idx = 0;
      for (int q = sys->entry_site_begin[e]; q < sys->entry_site_begin[e + 1]; ++q) {
        const int4 d = sys->entry_sites[q];
        int ii = (i0 + d.y) % sc.L1, jj = (j0 + d.z) % sc.L2, kk = (k0 + d.w) % sc.L3;
        if (ii < 0) ii += sc.L1;
        if (jj < 0) jj += sc.L2;
        if (kk < 0) kk += sc.L3;
        const int lq = d.x * sc.n_cells + ii + sc.L1 * (jj + sc.L2 * kk);
        int ov = occ[lq];
        for (int j = 0; j < i; ++j)
          if (sites[j] == lq) ov = nocc[j];
        idx += ov * sys->entry_mul[q];
      }
      const float *V = sys->flower_tables + sys->entry_table[e] + idx * nf;
      for (int f = 0; f < nf; ++f) G[f] += V[f];
    }
    float d = 0.0f;
    for (int f = 0; f < nf; ++f) d += (sys->site_func[b][f][nw] - sys->site_func[b][f][o]) * G[f];
    if (a.include_chem) d -= cp.chem[b * kMaxOcc + nw] - cp.chem[b * kMaxOcc + o];
    dE += d;
  }
  a.dE[p] = dE;
}

// ---------------------------------------------------------------------------
// Full correlations: one block per (function, chain), FP64 deterministic
// reduction
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void corr_kernel(CorrArgs a) {
  __shared__ double red[256];
  const int f = blockIdx.x;
  const int chain = a.chain0 + blockIdx.y;
  const DevSystem *__restrict__ sys = a.sys;
  const ChainParams &cp = a.params[chain];
  const DevSupercell &sc = a.scs[cp.sc];
  const int8_t *__restrict__ occ = a.occ + cp.occ_offset;
  const int orb = sys->func_orbit[f];
  const int k = sys->orbit_k_all[orb];
  const int nf = sys->orbit_nfunc[orb];
  const int fl = f - sys->orbit_func_begin[orb];
  const double *T = sys->orbit_tables + sys->orbit_table_offset[orb];
  const int ne = sys->corr_orbit_begin[orb + 1] - sys->corr_orbit_begin[orb];
  const int4 *es = sys->corr_elem_sites + sys->corr_elem_begin[orb];
  const long long total = (long long)ne * sc.n_cells;
  double acc = 0.0;
  for (long long w = threadIdx.x; w < total; w += blockDim.x) {
    const int e = (int)(w / sc.n_cells);
    const int c = (int)(w - (long long)e * sc.n_cells);
    const int i0 = c % sc.L1, j0 = (c / sc.L1) % sc.L2, k0 = c / (sc.L1 * sc.L2);
    int tau = 0, mul = 1;
    for (int q = 0; q < k; ++q) {
      const int4 d = es[e * k + q];
      int ii = (i0 + d.y) % sc.L1, jj = (j0 + d.z) % sc.L2, kk = (k0 + d.w) % sc.L3;
      if (ii < 0) ii += sc.L1;
      if (jj < 0) jj += sc.L2;
      if (kk < 0) kk += sc.L3;
      tau += occ[d.x * sc.n_cells + ii + sc.L1 * (jj + sc.L2 * kk)] * mul;
      mul *= sys->n_occ[d.x];
    }
    acc += T[tau * nf + fl];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.out[(long long)blockIdx.y * a.n_corr + f] = red[0] / (double)sys->orbit_mult[orb];
}

__global__ void philox_kernel(int n, const uint4 *ctr, const uint2 *key, uint4 *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = philox4x32_10(ctr[i], key[i].x, key[i].y);
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
// launch == false: only set the kernel's dynamic LDS limit (not a stream
// operation, so it stays outside graph capture)
template <bool BIN, int KIND, bool PACK>
static hipError_t launch_metro(const RunArgs &args, int n_chains, int threads, size_t lds, hipStream_t st,
                               bool launch) {
  auto kern = metropolis_kernel<BIN, KIND, PACK>;
  if (!launch) return hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(kern, dim3(n_chains), dim3(threads), lds, st, args);
  return hipGetLastError();
}

template <int KIND, bool PACK>
static hipError_t launch_split(const RunArgs &args, int n_chains, int threads, size_t lds, hipStream_t st,
                               bool launch) {
  auto kern = metropolis_split_kernel<KIND, PACK>;
  if (!launch) return hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(kern, dim3(n_chains), dim3(threads), lds, st, args);
  return hipGetLastError();
}

hipError_t launch_metropolis_split(const RunArgs &args, int kind, bool packed, int n_chains, int threads, size_t lds,
                                   hipStream_t st, bool launch) {
  if (packed)
    return kind == 0 ? launch_split<0, true>(args, n_chains, threads, lds, st, launch)
                     : launch_split<1, true>(args, n_chains, threads, lds, st, launch);
  return kind == 0 ? launch_split<0, false>(args, n_chains, threads, lds, st, launch)
                   : launch_split<1, false>(args, n_chains, threads, lds, st, launch);
}

hipError_t launch_metropolis(const RunArgs &args, bool binary, int kind, bool packed, int n_chains, int threads,
                             size_t lds, hipStream_t st, bool launch) {
  // kind 0 canonical, 1 semi-grand canonical, 2 semi-grand canonical with multiswaps
  if (packed) {
    if (binary)
      return kind == 0   ? launch_metro<true, 0, true>(args, n_chains, threads, lds, st, launch)
             : kind == 1 ? launch_metro<true, 1, true>(args, n_chains, threads, lds, st, launch)
                         : launch_metro<true, 2, true>(args, n_chains, threads, lds, st, launch);
    return kind == 0   ? launch_metro<false, 0, true>(args, n_chains, threads, lds, st, launch)
           : kind == 1 ? launch_metro<false, 1, true>(args, n_chains, threads, lds, st, launch)
                       : launch_metro<false, 2, true>(args, n_chains, threads, lds, st, launch);
  }
  if (binary)
    return kind == 0   ? launch_metro<true, 0, false>(args, n_chains, threads, lds, st, launch)
           : kind == 1 ? launch_metro<true, 1, false>(args, n_chains, threads, lds, st, launch)
                       : launch_metro<true, 2, false>(args, n_chains, threads, lds, st, launch);
  return kind == 0   ? launch_metro<false, 0, false>(args, n_chains, threads, lds, st, launch)
         : kind == 1 ? launch_metro<false, 1, false>(args, n_chains, threads, lds, st, launch)
                     : launch_metro<false, 2, false>(args, n_chains, threads, lds, st, launch);
}

// ---------------------------------------------------------------------------
// After a slab launch (RunArgs::slabs > 1), one thread per chain: sums the
// slabs' records in pass order (what each slab had added since the launch
// started, at every sampled pass and at the last pass), writes the samples,
// then the chain's energy, counts and exact binary counts, and advances the
// pass and draw counters, which the slabs read but do not change.
// ---------------------------------------------------------------------------
__global__ void slab_post_kernel(RunArgs a, int n_chains) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_chains) return;
  const int chain = a.chain_map ? a.chain_map[i] : i;
  const DevSystem *sys = a.sys;
  ChainState *cs = a.state + chain;
  const bool bin = sys->binary != 0;
  long long *C = a.binC + (long long)chain * a.n_orbits;
  const double E0 = cs->E;
  const unsigned long long pass0 = cs->pass_base;
  constexpr int nI = kMaxSlots + kMaxOcc + 2;
  long long tot[nI];
  double totE = 0.0;
  int j = 0;
  for (int p = 0; p < a.n_passes; ++p) {
    const bool due = ((pass0 + (unsigned long long)p + 1ull) % (unsigned long long)a.sample_period) == 0ull;
    if (!due && p != a.n_passes - 1) continue;
    for (int q = 0; q < nI; ++q) tot[q] = 0;
    totE = 0.0;
    for (int s = 0; s < a.slabs; ++s) {
      const SlabRecord &r = a.slab_rec[((long long)chain * kSlabMax + s) * kSlabRecords + j];
      totE += r.dE;
      for (int q = 0; q < nI; ++q) tot[q] += r.dI[q];
    }
    ++j;
    if (!due) continue;
    double E;
    if (bin) {
      E = 0.0;
      for (int o = 0; o < a.n_orbits; ++o) {
        long long c = C[o];
        for (int q = 0; q < sys->n_slots; ++q)
          if (sys->slot_orbit[q] == o) c += tot[q];
        E += sys->bin_w64[o] * (double)c;
      }
    } else {
      E = E0 + totE;
    }
    long long cnt[kMaxOcc];
    for (int q = 0; q < kMaxOcc; ++q) cnt[q] = cs->counts[q] + tot[kMaxSlots + q];
    const int idx = a.sample_count[chain];
    if (idx < a.sample_cap) {
      const long long o = (long long)chain * a.sample_cap + idx;
      a.sampE[o] = E;
      for (int q = 0; q < kMaxOcc; ++q) a.sampN[o * kMaxOcc + q] = (int)cnt[q];
      a.sample_count[chain] = idx + 1;
    }
    const double n1 = (double)cnt[1];
    cs->sumE += E;
    cs->sumE2 += E * E;
    cs->sumN1 += n1;
    cs->sumN12 += n1 * n1;
    cs->sumEN1 += E * n1;
    cs->n_samples_acc += 1;
  }
  // the last record holds the launch's totals
  if (bin) {
    for (int q = 0; q < sys->n_slots; ++q) C[sys->slot_orbit[q]] += tot[q];
    double E = 0.0;
    for (int o = 0; o < a.n_orbits; ++o) E += sys->bin_w64[o] * (double)C[o];
    cs->E = E;
  } else {
    cs->E = E0 + totE;
  }
  for (int q = 0; q < kMaxOcc; ++q) cs->counts[q] += tot[kMaxSlots + q];
  cs->n_attempt += tot[kMaxSlots + kMaxOcc];
  cs->n_accept += tot[kMaxSlots + kMaxOcc + 1];
  cs->pass_base += (unsigned long long)a.n_passes;
  const unsigned long long next = cs->draw_hi + 1ull;
  if (next > cs->draw_base) cs->draw_base = next;
}

hipError_t launch_slab_post(const RunArgs &a, int n_chains, hipStream_t st) {
  if (n_chains <= 0) return hipSuccess;
  hipLaunchKernelGGL(slab_post_kernel, dim3((n_chains + 63) / 64), dim3(64), 0, st, a, n_chains);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// LDS read-throughput probe: every lane reads LDS in a loop and folds the
// values into a checksum (written out so that nothing is optimised away).
// mode 0: byte gathers, lanes of a wavefront 3 bytes apart (the stride of
// independent-set class sites along a row) at 8 offsets per iteration;
// mode 1: 16-byte reads, lanes 16 bytes apart (bank-conflict free).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(512) void lds_probe_kernel(int mode, int iters, unsigned *sink) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int tid = threadIdx.x;
  const int bytes = 32768;
  for (int i = tid; i < bytes / 4; i += blockDim.x) ((unsigned *)lds)[i] = (unsigned)(i * 2654435761u);
  __syncthreads();
  unsigned acc = 0u;
  if (mode == 0) {
    const int base = (tid & 63) * 3 + (tid >> 6) * 211;
    for (int it = 0; it < iters; ++it) {
      const int o = (it * 97) & 4095;
#pragma unroll
      for (int k = 0; k < 8; ++k) acc += lds[(base + o + k * 401) & (bytes - 1)];
    }
  } else {
    const uint4 *v = (const uint4 *)lds;
    const int base = tid & 511;
    for (int it = 0; it < iters; ++it) {
      const int o = (it * 64) & 1023;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const uint4 x = v[(base + o + k * 256) & 2047];
        acc += x.x ^ x.y ^ x.z ^ x.w;
      }
    }
  }
  sink[blockIdx.x * blockDim.x + tid] = acc;
}

hipError_t launch_lds_probe(int mode, int iters, int blocks, unsigned *sink, hipStream_t st) {
  hipLaunchKernelGGL(lds_probe_kernel, dim3(blocks), dim3(512), 32768, st, mode, iters, sink);
  return hipGetLastError();
}

// Staging array of clexmc_delta_energy_batch_noalloc: each GPU holds its own
// copy at its own address until the process ends.  Host code never passes
// the name itself to a kernel (in host code it is a stand-in, not a device
// address); it uses the address hipGetSymbolAddress returns for the device.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch(int device, unsigned char **ptr) {
  static std::mutex mu;
  static std::map<int, unsigned char *> by_device;
  std::lock_guard<std::mutex> lock(mu);
  auto it = by_device.find(device);
  if (it != by_device.end()) {
    *ptr = it->second;
    return hipSuccess;
  }
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  by_device[device] = (unsigned char *)p;
  *ptr = (unsigned char *)p;
  return hipSuccess;
}

hipError_t launch_delta_batch(const DeltaArgs &a, hipStream_t st) {
  const int threads = 128;
  const int blocks = (a.n_prop + threads - 1) / threads;
  hipLaunchKernelGGL(delta_batch_kernel, dim3(blocks), dim3(threads), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_corr(const CorrArgs &a, int n_corr, int n_chains, hipStream_t st) {
  hipLaunchKernelGGL(corr_kernel, dim3(n_corr, n_chains), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_philox(int n, const uint4 *ctr, const uint2 *key, uint4 *out, hipStream_t st) {
  hipLaunchKernelGGL(philox_kernel, dim3((n + 63) / 64), dim3(64), 0, st, n, ctr, key, out);
  return hipGetLastError();
}

size_t metropolis_lds_bytes(int pad_sites, int n_tab, bool binary, int n_func_max, int max_class, bool packed) {
  size_t img = (size_t)image_bytes(pad_sites, packed);
  size_t tab = binary ? 0 : (size_t)((n_tab * 4 + 15) & ~15);
  const int n_islot = (binary ? kMaxSlots : 0) + kMaxOcc + 2;
  size_t red = (size_t)n_islot * 16 * sizeof(int) + 16 * sizeof(double) + 16 + 128 + kMswapLdsBytes +
               kChainLdsBytes;
  const int nfs = binary ? 1 : n_func_max;
  size_t scratch = (size_t)nfs * max_class * sizeof(float) + (binary ? (size_t)kMaxSlots * max_class : 0);
  return img + tab + red + scratch + 64;
}

}  // namespace clexmc
