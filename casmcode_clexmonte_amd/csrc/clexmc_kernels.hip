// gfx950 kernels of the clexmc Metropolis hot path.
//
//  metropolis_kernel   one workgroup per chain; the chain's int8 occupation
//                      lives in LDS as a halo-padded image; sites are updated
//                      class by class (independent sets: no cluster contains
//                      two sites of one class), so all proposals of a class
//                      are evaluated and applied concurrently and exactly.
//                      Binary (+-1) expansions use integer cluster-product
//                      counts per orbit (K_a, exact), FP32 dE = dphi*sum w_a K_a
//                      and exact int64 correlation counts C_a per chain; other
//                      expansions use FP32 flower tables and an FP64 running
//                      energy.  Philox4x32-10 supplies the uniforms.
//  delta_batch_kernel  per-proposal dE (BaseMontePotential::occ_delta_per_supercell)
//  corr_kernel         full per-supercell correlations in FP64
//  philox_kernel       device Philox for known-answer checks
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clexmc_device.h"
#include "clexmc_kernels.h"

namespace clexmc {

// ---------------------------------------------------------------------------
// Per-lane uniform stream: one Philox block yields four uniforms.
// Counter = (lane, draw_lo, draw_hi, chain); key = chain seed.
// ---------------------------------------------------------------------------
struct LaneRng {
  uint4 buf;
  int left;
  unsigned long long draw;
  uint32_t lane, chain, k0, k1;

  __device__ __forceinline__ float next() {
    if (left == 0) {
      buf = philox4x32_10(make_uint4(lane, (uint32_t)draw, (uint32_t)(draw >> 32), chain), k0, k1);
      ++draw;
      left = 4;
    }
    uint32_t x;
    switch (left) {
      case 4: x = buf.x; break;
      case 3: x = buf.y; break;
      case 2: x = buf.z; break;
      default: x = buf.w; break;
    }
    --left;
    return u01(x);
  }
};

// block-uniform random integer in [0, n) for (pass, phase, round)
__device__ __forceinline__ uint32_t uniform_int(uint32_t n, unsigned long long pass, uint32_t phase,
                                                uint32_t chain, uint32_t k0, uint32_t k1) {
  uint4 r = philox4x32_10(make_uint4(0xFFFFFFFFu, (uint32_t)pass, phase, chain ^ 0x9E3779B9u), k0, k1);
  unsigned long long x = ((unsigned long long)r.x << 32) | r.y;
  return n ? (uint32_t)(x % n) : 0u;
}

// ---------------------------------------------------------------------------
// Binary path: K_a(s) = sum over instances of orbit a containing s of
// prod phi(other sites) = (-1)^(k-1) * (n_a - 2 * sum XOR(other bits)).
// ---------------------------------------------------------------------------
template <int NORB>
__device__ __forceinline__ float binary_field(const int8_t *__restrict__ img, int pa, int b,
                                              const DevSystem *__restrict__ sys,
                                              const DevSupercell &sc, int n_orbits, int (&K)[NORB]) {
  float F = 0.0f;
  const int2 *rng = sys->bin_range + b * n_orbits;
  const int *__restrict__ esb = sys->entry_site_begin;
  const int *__restrict__ off = sc.entry_off;
#pragma unroll
  for (int a = 0; a < NORB; ++a) {
    K[a] = 0;
    if (a < n_orbits) {
      const int2 r = rng[a];
      const int n = r.y - r.x;
      if (n > 0) {
        const int ko = sys->orbit_k[a] - 1;
        int X = 0;
        if (ko == 1) {
          const int q0 = esb[r.x];
          for (int q = 0; q < n; ++q) X += img[pa + off[q0 + q]];
        } else if (ko == 2) {
          const int q0 = esb[r.x];
          for (int q = 0; q < n; ++q) X += img[pa + off[q0 + 2 * q]] ^ img[pa + off[q0 + 2 * q + 1]];
        } else if (ko > 2) {
          for (int e = r.x; e < r.y; ++e) {
            int x = 0;
            for (int q = esb[e]; q < esb[e + 1]; ++q) x ^= img[pa + off[q]];
            X += x;
          }
        }
        const int Ka = (ko & 1) ? -(n - 2 * X) : (n - 2 * X);
        K[a] = Ka;
        F += sys->bin_w[a] * (float)Ka;
      }
    }
  }
  return F;
}

// General path: G_f(s) = sum_e V_e[tau_others][f] from the flower tables.
__device__ __forceinline__ void general_field(const int8_t *__restrict__ img, int pa, int b,
                                              const DevSystem *__restrict__ sys,
                                              const DevSupercell &sc, float (&G)[kMaxFunc]) {
  G[0] = G[1] = G[2] = 0.0f;
  const int nf = sys->n_func[b];
  const int e0 = sys->flower_begin[b], e1 = sys->flower_begin[b + 1];
  const int *__restrict__ esb = sys->entry_site_begin;
  const int *__restrict__ off = sc.entry_off;
  const int *__restrict__ mul = sys->entry_mul;
  for (int e = e0; e < e1; ++e) {
    int idx = 0;
    for (int q = esb[e]; q < esb[e + 1]; ++q) idx += img[pa + off[q]] * mul[q];
    const float *V = sys->flower_tables + sys->entry_table[e] + idx * nf;
    G[0] += V[0];
    if (nf > 1) G[1] += V[1];
    if (nf > 2) G[2] += V[2];
  }
}

__device__ __forceinline__ float general_dE(const DevSystem *__restrict__ sys, int b, int o, int nw,
                                            const float (&G)[kMaxFunc]) {
  const int nf = sys->n_func[b];
  float d = (sys->site_func[b][0][nw] - sys->site_func[b][0][o]) * G[0];
  if (nf > 1) d += (sys->site_func[b][1][nw] - sys->site_func[b][1][o]) * G[1];
  if (nf > 2) d += (sys->site_func[b][2][nw] - sys->site_func[b][2][o]) * G[2];
  return d;
}

__device__ __forceinline__ bool metropolis_accept(float dE, float beta, float u) {
  return dE <= 0.0f || u < __expf(-beta * dE);
}

// ---------------------------------------------------------------------------
// Persistent Metropolis kernel
// ---------------------------------------------------------------------------
template <bool BIN, int KIND>
__global__ __launch_bounds__(512) void metropolis_kernel(RunArgs args) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int chain = blockIdx.x;
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;
  const DevSystem *__restrict__ sys = args.sys;
  DevChain *ch = args.chains + chain;
  const DevSupercell sc = args.scs[ch->sc];
  const int n_orbits = sys->n_orbits;
  const int nocc_max = sys->n_occ_max;

  int8_t *img = (int8_t *)smem;
  const int img_bytes = (sc.pad_bytes + 15) & ~15;
  // reduction slots: int slots [n_islot][nthr], then double slot [nthr]
  const int n_islot = (BIN ? n_orbits : 0) + kMaxOcc + 2;
  int *islot = (int *)(smem + img_bytes);
  double *dslot = (double *)(islot + n_islot * nthr);
  int *isum = (int *)(dslot + nthr);             // [n_islot]
  double *dsum = (double *)(isum + ((n_islot + 1) & ~1));  // [1]
  unsigned char *scratch = (unsigned char *)(dsum + 2);
  // canonical class scratch: F/G per class site and K (int8) per orbit
  float *cF = (float *)scratch;                               // [kMaxFunc][max_class]
  int8_t *cK = (int8_t *)(cF + kMaxFunc * sc.max_class);      // [NORB][max_class]

  // ---- load the padded image ----
  const int8_t *__restrict__ occG = args.occ + ch->occ_offset;
  {
    const int P1 = sc.P1, P12 = sc.P1 * sc.P2, S = sc.S, h = sc.h;
    for (int p = tid; p < sc.pad_bytes; p += nthr) {
      int b = p / S;
      int r = p - b * S;
      int pk = r / P12;
      int r2 = r - pk * P12;
      int pj = r2 / P1;
      int pi = r2 - pj * P1;
      int i = ((pi - h) % sc.L1 + sc.L1) % sc.L1;
      int j = ((pj - h) % sc.L2 + sc.L2) % sc.L2;
      int k = ((pk - h) % sc.L3 + sc.L3) % sc.L3;
      img[p] = occG[b * sc.n_cells + i + sc.L1 * (j + sc.L2 * k)];
    }
  }
  __syncthreads();

  const float beta = ch->beta;
  LaneRng rng;
  rng.left = 0;
  rng.draw = ch->draw_base;
  rng.lane = (uint32_t)tid;
  rng.chain = (uint32_t)chain;
  rng.k0 = ch->key0;
  rng.k1 = ch->key1;

  // per-lane accumulators
  int accC[kMaxOrbits];
#pragma unroll
  for (int a = 0; a < kMaxOrbits; ++a) accC[a] = 0;
  int accN[kMaxOcc] = {0, 0, 0, 0};
  int accAtt = 0, accAcc = 0;
  double accE = 0.0;

  const unsigned long long pass0 = ch->pass_base;
  const int rounds = args.rounds < 1 ? 1 : args.rounds;

  for (int pass = 0; pass < args.n_passes; ++pass) {
    const unsigned long long gpass = pass0 + (unsigned long long)pass;
    for (int c = 0; c < sc.n_classes; ++c) {
      const int b = sc.class_sublat[c];
      const int cb = sc.class_begin[c], ce = sc.class_begin[c + 1];
      if (KIND == 1) {
        // ---- semi-grand canonical: every site of the class proposes a change
        for (int t = cb + tid; t < ce; t += nthr) {
          const int pa = sc.class_addr[t];
          const int o = img[pa];
          float u = rng.next();
          if (BIN) {
            int K[kMaxOrbits];
            const float F = binary_field<kMaxOrbits>(img, pa, b, sys, sc, n_orbits, K);
            const int nw = 1 - o;
            const float dphi = o ? -2.0f : 2.0f;
            const float dE = dphi * F - (ch->chem[b * kMaxOcc + nw] - ch->chem[b * kMaxOcc + o]);
            const bool acc = metropolis_accept(dE, beta, u);
            accAtt += 1;
            if (acc) {
              img[pa] = (int8_t)nw;
              accAcc += 1;
              const int s = o ? -2 : 2;
#pragma unroll
              for (int a = 0; a < kMaxOrbits; ++a) accC[a] += s * K[a];
              accN[nw] += 1;
              accN[o] -= 1;
            }
          } else {
            float G[kMaxFunc];
            general_field(img, pa, b, sys, sc, G);
            const int no = sys->n_occ[b];
            float u2 = rng.next();
            int r = (int)(u2 * (float)(no - 1));
            if (r > no - 2) r = no - 2;
            const int nw = r >= o ? r + 1 : r;
            const float dEf = general_dE(sys, b, o, nw, G);
            const float dE = dEf - (ch->chem[b * kMaxOcc + nw] - ch->chem[b * kMaxOcc + o]);
            const bool acc = metropolis_accept(dE, beta, u);
            accAtt += 1;
            if (acc) {
              img[pa] = (int8_t)nw;
              accAcc += 1;
              accE += (double)dEf;
              accN[nw] += 1;
              accN[o] -= 1;
            }
          }
        }
        __syncthreads();
      } else {
        // ---- canonical: fields of all class sites, then random matchings
        const int n = ce - cb;
        for (int t = tid; t < n; t += nthr) {
          const int pa = sc.class_addr[cb + t];
          if (BIN) {
            int K[kMaxOrbits];
            cF[t] = binary_field<kMaxOrbits>(img, pa, b, sys, sc, n_orbits, K);
#pragma unroll
            for (int a = 0; a < kMaxOrbits; ++a)
              if (a < n_orbits) cK[a * sc.max_class + t] = (int8_t)K[a];
          } else {
            float G[kMaxFunc];
            general_field(img, pa, b, sys, sc, G);
            cF[t] = G[0];
            cF[sc.max_class + t] = G[1];
            cF[2 * sc.max_class + t] = G[2];
          }
        }
        __syncthreads();
        for (int rd = 0; rd < rounds; ++rd) {
          const uint32_t r = uniform_int((uint32_t)n, gpass, (uint32_t)(c * rounds + rd), (uint32_t)chain,
                                         ch->key0, ch->key1);
          for (int p = tid; p < n; p += nthr) {
            int q = (int)r - p;
            if (q < 0) q += n;
            if (p >= q) continue;  // q == p: unpaired; p > q: the partner leads
            const int pa = sc.class_addr[cb + p], qa = sc.class_addr[cb + q];
            const int op = img[pa], oq = img[qa];
            if (op == oq) continue;
            float u = rng.next();
            float dE;
            if (BIN) {
              const float dp = op ? -2.0f : 2.0f;
              dE = dp * cF[p] - dp * cF[q];
            } else {
              float Gp[kMaxFunc] = {cF[p], cF[sc.max_class + p], cF[2 * sc.max_class + p]};
              float Gq[kMaxFunc] = {cF[q], cF[sc.max_class + q], cF[2 * sc.max_class + q]};
              dE = general_dE(sys, b, op, oq, Gp) + general_dE(sys, b, oq, op, Gq);
            }
            const bool acc = metropolis_accept(dE, beta, u);
            accAtt += 1;
            if (acc) {
              img[pa] = (int8_t)oq;
              img[qa] = (int8_t)op;
              accAcc += 1;
              if (BIN) {
                const int sp = op ? -2 : 2;
#pragma unroll
                for (int a = 0; a < kMaxOrbits; ++a)
                  if (a < n_orbits)
                    accC[a] += sp * ((int)cK[a * sc.max_class + p] - (int)cK[a * sc.max_class + q]);
              } else {
                accE += (double)dE;
              }
            }
          }
          __syncthreads();
        }
      }
      // ---- halo copies of this class
      for (int t = sc.halo_begin[c] + tid; t < sc.halo_begin[c + 1]; t += nthr) {
        const int2 hc = sc.halo[t];
        img[hc.y] = img[hc.x];
      }
      __syncthreads();
    }

    // ---- canonical cross-class mixing: one randomly chosen pair phase
    if (KIND == 0 && sc.n_pair_phases > 0) {
      const uint32_t ph = uniform_int((uint32_t)sc.n_pair_phases, gpass, 0xFFFFFFF0u, (uint32_t)chain,
                                      ch->key0, ch->key1);
      const int pb = sc.pair_begin[ph], pe = sc.pair_begin[ph + 1];
      for (int t = pb + tid; t < pe; t += nthr) {
        const int2 pq = sc.pair_addr[t];
        const int b = sc.pair_sublat[t];
        const int op = img[pq.x], oq = img[pq.y];
        if (op == oq) continue;
        float u = rng.next();
        float dE;
        int Kp[kMaxOrbits], Kq[kMaxOrbits];
        if (BIN) {
          const float Fp = binary_field<kMaxOrbits>(img, pq.x, b, sys, sc, n_orbits, Kp);
          const float Fq = binary_field<kMaxOrbits>(img, pq.y, b, sys, sc, n_orbits, Kq);
          const float dp = op ? -2.0f : 2.0f;
          dE = dp * Fp - dp * Fq;
        } else {
          float Gp[kMaxFunc], Gq[kMaxFunc];
          general_field(img, pq.x, b, sys, sc, Gp);
          general_field(img, pq.y, b, sys, sc, Gq);
          dE = general_dE(sys, b, op, oq, Gp) + general_dE(sys, b, oq, op, Gq);
        }
        const bool acc = metropolis_accept(dE, beta, u);
        accAtt += 1;
        if (acc) {
          img[pq.x] = (int8_t)oq;
          img[pq.y] = (int8_t)op;
          accAcc += 1;
          if (BIN) {
            const int sp = op ? -2 : 2;
#pragma unroll
            for (int a = 0; a < kMaxOrbits; ++a) accC[a] += sp * (Kp[a] - Kq[a]);
          } else {
            accE += (double)dE;
          }
        }
      }
      __syncthreads();
      for (int t = sc.pair_halo_begin[ph] + tid; t < sc.pair_halo_begin[ph + 1]; t += nthr) {
        const int2 hc = sc.pair_halo[t];
        img[hc.y] = img[hc.x];
      }
      __syncthreads();
    }

    // ---- reduction of lane accumulators and sampling
    const bool sample_due = ((gpass + 1) % (unsigned long long)args.sample_period) == 0ull;
    const bool flush = sample_due || ((pass & 31) == 31) || (pass == args.n_passes - 1);
    if (flush) {
      int s = 0;
      if (BIN) {
#pragma unroll
        for (int a = 0; a < kMaxOrbits; ++a)
          if (a < n_orbits) islot[(s + a) * nthr + tid] = accC[a];
        s += n_orbits;
      }
#pragma unroll
      for (int o = 0; o < kMaxOcc; ++o) islot[(s + o) * nthr + tid] = accN[o];
      s += kMaxOcc;
      islot[s * nthr + tid] = accAtt;
      islot[(s + 1) * nthr + tid] = accAcc;
      dslot[tid] = accE;
#pragma unroll
      for (int a = 0; a < kMaxOrbits; ++a) accC[a] = 0;
#pragma unroll
      for (int o = 0; o < kMaxOcc; ++o) accN[o] = 0;
      accAtt = accAcc = 0;
      accE = 0.0;
      __syncthreads();
      if (tid < n_islot) {
        long long acc = 0;
        for (int l = 0; l < nthr; ++l) acc += islot[tid * nthr + l];
        isum[tid] = (int)acc;
      } else if (tid == n_islot) {
        double acc = 0.0;
        for (int l = 0; l < nthr; ++l) acc += dslot[l];
        dsum[0] = acc;
      }
      __syncthreads();
      if (tid == 0) {
        int s2 = 0;
        double E;
        if (BIN) {
          long long *C = args.binC + (long long)chain * n_orbits;
          E = 0.0;
          for (int a = 0; a < n_orbits; ++a) {
            C[a] += isum[a];
            E += sys->bin_w64[a] * (double)C[a];
          }
          s2 = n_orbits;
        } else {
          E = ch->E + dsum[0];
        }
        ch->E = E;
        for (int o = 0; o < kMaxOcc; ++o) ch->counts[o] += isum[s2 + o];
        ch->n_attempt += isum[s2 + kMaxOcc];
        ch->n_accept += isum[s2 + kMaxOcc + 1];
        if (sample_due) {
          const int idx = args.sample_count[chain];
          if (idx < args.sample_cap) {
            const long long o = (long long)chain * args.sample_cap + idx;
            args.sampE[o] = E;
            for (int q = 0; q < kMaxOcc; ++q) args.sampN[o * kMaxOcc + q] = (int)ch->counts[q];
            args.sample_count[chain] = idx + 1;
          }
          const double n1 = (double)ch->counts[1];
          ch->sumE += E;
          ch->sumE2 += E * E;
          ch->sumN1 += n1;
          ch->sumN12 += n1 * n1;
          ch->sumEN1 += E * n1;
          ch->n_samples_acc += 1;
        }
      }
      __syncthreads();
    }
    (void)nocc_max;
  }

  // ---- write back occupation and RNG state
  {
    int8_t *occW = args.occ + ch->occ_offset;
    const int P1 = sc.P1, P12 = sc.P1 * sc.P2, h = sc.h;
    for (int l = tid; l < sc.n_sites; l += nthr) {
      int b = l / sc.n_cells;
      int cc = l - b * sc.n_cells;
      int i = cc % sc.L1;
      int j = (cc / sc.L1) % sc.L2;
      int k = cc / (sc.L1 * sc.L2);
      occW[l] = img[b * sc.S + (k + h) * P12 + (j + h) * P1 + (i + h)];
    }
  }
  // all lanes advance by the same bound so the next launch starts fresh
  __shared__ unsigned long long max_draw;
  if (tid == 0) max_draw = 0ull;
  __syncthreads();
  atomicMax(&max_draw, rng.draw);
  __syncthreads();
  if (tid == 0) {
    ch->draw_base = max_draw + 1ull;
    ch->pass_base = pass0 + (unsigned long long)args.n_passes;
  }
}

// ---------------------------------------------------------------------------
// Per-proposal dE in FP32 on unpadded occupation (periodic wrap)
// ---------------------------------------------------------------------------
__global__ void delta_batch_kernel(DeltaArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n_prop) return;
  const DevSystem *__restrict__ sys = a.sys;
  const int chain = a.chain[p];
  const DevChain &ch = a.chains[chain];
  const DevSupercell &sc = a.scs[ch.sc];
  const int8_t *__restrict__ occ = a.occ + ch.occ_offset;
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
    if (a.include_chem) d -= ch.chem[b * kMaxOcc + nw] - ch.chem[b * kMaxOcc + o];
    dE += d;
  }
  a.dE[p] = dE;
}

// ---------------------------------------------------------------------------
// Full correlations: one block per (function), FP64 deterministic reduction
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void corr_kernel(CorrArgs a) {
  __shared__ double red[256];
  const int f = blockIdx.x;
  const DevSystem *__restrict__ sys = a.sys;
  const DevChain &ch = a.chains[a.chain];
  const DevSupercell &sc = a.scs[ch.sc];
  const int8_t *__restrict__ occ = a.occ + ch.occ_offset;
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
  if (threadIdx.x == 0) a.out[f] = red[0] / (double)sys->orbit_mult[orb];
}

__global__ void philox_kernel(int n, const uint4 *ctr, const uint2 *key, uint4 *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = philox4x32_10(ctr[i], key[i].x, key[i].y);
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
template <bool BIN, int KIND>
static hipError_t launch_metro(const RunArgs &args, int n_chains, int threads, size_t lds, hipStream_t st) {
  auto kern = metropolis_kernel<BIN, KIND>;
  hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(n_chains), dim3(threads), lds, st, args);
  return hipGetLastError();
}

hipError_t launch_metropolis(const RunArgs &args, bool binary, int kind, int n_chains, int threads,
                             size_t lds, hipStream_t st) {
  if (binary) {
    return kind == 0 ? launch_metro<true, 0>(args, n_chains, threads, lds, st)
                     : launch_metro<true, 1>(args, n_chains, threads, lds, st);
  }
  return kind == 0 ? launch_metro<false, 0>(args, n_chains, threads, lds, st)
                   : launch_metro<false, 1>(args, n_chains, threads, lds, st);
}

hipError_t launch_delta_batch(const DeltaArgs &a, hipStream_t st) {
  const int threads = 128;
  const int blocks = (a.n_prop + threads - 1) / threads;
  hipLaunchKernelGGL(delta_batch_kernel, dim3(blocks), dim3(threads), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_corr(const CorrArgs &a, int n_corr, hipStream_t st) {
  hipLaunchKernelGGL(corr_kernel, dim3(n_corr), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_philox(int n, const uint4 *ctr, const uint2 *key, uint4 *out, hipStream_t st) {
  hipLaunchKernelGGL(philox_kernel, dim3((n + 63) / 64), dim3(64), 0, st, n, ctr, key, out);
  return hipGetLastError();
}

size_t metropolis_lds_bytes(int pad_bytes, int n_orbits, bool binary, int threads, int max_class) {
  size_t img = (size_t)((pad_bytes + 15) & ~15);
  int n_islot = (binary ? n_orbits : 0) + kMaxOcc + 2;
  size_t slots = (size_t)n_islot * threads * sizeof(int) + (size_t)threads * sizeof(double);
  size_t sums = (size_t)((n_islot + 1) & ~1) * sizeof(int) + 2 * sizeof(double);
  size_t scratch = (size_t)kMaxFunc * max_class * sizeof(float) + (size_t)kMaxOrbits * max_class;
  return img + slots + sums + scratch + 64;
}

}  // namespace clexmc
