// Device code of the N-fold way (rejection-free, Bortz-Kalos-Lebowitz) semi-grand canonical
// sampling: the reference's nfold method
// (include/casm/clexmonte/nfold/nfold_impl.hh:18-98: events are the
// semi-grand canonical changes of single sites; rates
// src/casm/clexmonte/nfold/nfold_events.cc:28-62: 1 if dE <= 0, else
// exp(-beta dE), dE the semi-grand potential change).
//
// The events are those of the plain semi-grand canonical proposal (a site
// and a uniformly chosen other occupant): a site's rate is the mean
// Metropolis acceptance of its changes, (1/(n_occ-1)) sum_o' min(1,
// exp(-beta dOmega)); binary expansions have one change per site.
//
// One workgroup per chain.  LDS holds the padded int8 image, the FP32 rate
// of every site and FP64 sums of the rates over groups of 64 sites.  Per
// event:
//  * wavefront 0 forms the total rate R (FP64), draws a 53-bit uniform u,
//    finds the group by a scan over per-lane chunks of group sums, then the
//    site by a scan of the group's 64 rates (FP64 wavefront scans), so a rate
//    is drawn in proportion to its value however small it is next to R;
//  * the current state is weighted by its expected residence time
//    w = n_mutable / R in Metropolis steps (the jump chain visits states in
//    proportion to pi(x) R(x); weights 1/R recover pi);
//  * one lane applies the change (the new occupant drawn in proportion to
//    its rate; binary: exact integer cluster counts, general: FP64 energy;
//    every periodic copy of the site), then the lanes recompute the rates of
//    the site and of its interaction neighbours (each neighbour once), and
//    the sum of every group holding one of them is formed again from its 64
//    rates.  Group sums are never updated by differences, so they carry no
//    cancellation error when rates of very different size come and go;
//  * if R falls below 1e-30 (every rate of the chain within a few decades of
//    FP32 underflow, beta dE > ~69 for all changes) the chain stops and
//    records the condition in its state; the host reports it as an error.
//
// nfold_body<BIN, FIELD>: FIELD is GenericField (interpreted site programs,
// clexmc_nfold.hip) or the run-time specialised JitField (clexmc_jit.cpp);
// both sum the clusters in the same order, so the results are identical.
#pragma once
#include "clexmc_metropolis.h"

namespace clexmc {

__host__ __device__ inline size_t nfold_align16(size_t x) { return (x + 15) & ~(size_t)15; }

// x of the lane that DPP control CTRL names (0 where it names none, and in
// the rows ROW_MASK leaves out), the two halves moved separately
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_d(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROW_MASK, 0xf, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Inclusive prefix sum (FP64) over the 64 lanes of a full wave: shifted adds
// within rows of 16 lanes (DPP row_shr 1, 2, 4, 8), then row 0's total into
// row 1 and row 2's into row 3 (row_bcast 15), then rows 0-1's total into
// rows 2-3 (row_bcast 31).  VALU only: no LDS round trips.
__device__ __forceinline__ double wave_scan_incl_d(double v) {
  v += dpp_d<0x111, 0xf>(v);
  v += dpp_d<0x112, 0xf>(v);
  v += dpp_d<0x114, 0xf>(v);
  v += dpp_d<0x118, 0xf>(v);
  v += dpp_d<0x142, 0xa>(v);
  v += dpp_d<0x143, 0xc>(v);
  return v;
}

// x of lane l (l the same in every lane)
__device__ __forceinline__ double lane_d(double x, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// first lane whose inclusive prefix exceeds t; if rounding leaves none, the
// last lane with a positive value (-1 if there is none)
__device__ __forceinline__ int pick_lane(double incl, double v, double t) {
  const unsigned long long m = __ballot(incl > t && v > 0.0);
  if (m) return __ffsll((long long)m) - 1;
  const unsigned long long m2 = __ballot(v > 0.0);
  return m2 ? 63 - __clzll((long long)m2) : -1;
}

// uniform double in [0, 1) from 53 random bits of two words
__device__ __forceinline__ double u01_53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// sum of the 64 FP32 rates of a group in FP64, in a fixed order (four
// interleaved partial sums of 16), so every lane that forms it gets the same
// value
__device__ __forceinline__ double group_rate_sum(const float *rate, int g) {
  const float4 *r4 = (const float4 *)(rate + (size_t)g * 64);
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const float4 v = r4[q];
    a0 += (double)v.x;
    a1 += (double)v.y;
    a2 += (double)v.z;
    a3 += (double)v.w;
  }
  return (a0 + a1) + (a2 + a3);
}

// smallest total rate a chain may reach before it stops with an error
constexpr double kNfoldMinTotalRate = 1e-30;

template <bool BIN, class FIELD>
__device__ __forceinline__ void nfold_body(const NfoldArgs &na) {
  extern __shared__ __align__(16) unsigned char smem[];
  const RunArgs &args = na.a;
  const int chain = blockIdx.x;
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int n_waves = nthr >> 6;
  cptr<DevSystem> sys = cst(args.sys);
  cptr<ChainParams> cp = cst(args.params + chain);
  ChainState *cs = args.state + chain;
  cptr<DevSupercell> sc = cst(args.scs + cp->sc);

  const int L1 = sc->L1, L2 = sc->L2, L3 = sc->L3, h = sc->h;
  const int P1 = sc->P1, P12 = sc->P1 * sc->P2, S = sc->S;
  const int n_cells = sc->n_cells, n_sites = sc->n_sites, pad_bytes = sc->pad_bytes;
  const int n_slots = sys->n_slots;
  const int n_groups = (n_sites + 63) / 64;

  ImgU8 img;
  img.p = (uint8_t *)smem;
  size_t off = nfold_align16((size_t)image_bytes(pad_bytes, false));
  float *rate = (float *)(smem + off);  // [n_groups * 64], zero past n_sites
  off += nfold_align16((size_t)n_groups * 64 * 4);
  double *G = (double *)(smem + off);   // [n_groups]
  off += nfold_align16((size_t)n_groups * 8);
  int *aff = (int *)(smem + off);    // linear indices of the sites whose rates change
  off += nfold_align16((size_t)sc->nb_max * 4);
  int *affp = (int *)(smem + off);   // their padded addresses
  off += nfold_align16((size_t)sc->nb_max * 4);
  int *sel = (int *)(smem + off);
  off += 16;
  // the interaction-neighbour offsets of every sublattice, staged once: an
  // event reads them right after its flip
  int *nbB = (int *)(smem + off);     // [n_sublat + 1] list starts
  off += kNfoldNbBeginBytes;
  int4 *nbL = (int4 *)(smem + off);   // (sublattice, di, dj, dk) per neighbour
  off += nfold_align16((size_t)(sc->nb_begin[sys->n_sublat]) * 16);
  // occupant counts and chemical potentials per sublattice: the rate of a
  // neighbour reads them for its sublattice (a per-lane index)
  int *noccL = (int *)(smem + off);           // [kMaxSublat]
  float *chemL = (float *)(noccL + kMaxSublat);  // [kMaxSublat][kMaxOcc]

  // ---- padded image, neighbour lists
  const int8_t *__restrict__ occG = args.occ + cp->occ_offset;
  for (int p = tid; p < pad_bytes; p += nthr) {
    const int b = p / S;
    const int r = p - b * S;
    const int pk = r / P12, pj = (r - pk * P12) / P1, pi = r - pk * P12 - pj * P1;
    const int i = pad_wrap(pi, h, L1), j = pad_wrap(pj, h, L2), k = pad_wrap(pk, h, L3);
    img.p[p] = (uint8_t)occG[b * n_cells + i + L1 * (j + L2 * k)];
  }
  {
    const int n_sub = sys->n_sublat;
    const int nb_tot = sc->nb_begin[n_sub];
    for (int t = tid; t < nb_tot; t += nthr) nbL[t] = sc->nb_cell[t];
    if (tid <= n_sub) nbB[tid] = sc->nb_begin[tid];
    if (tid < n_sub) noccL[tid] = sys->n_occ[tid];
    if (tid < n_sub * kMaxOcc) chemL[tid] = cp->chem[tid];
  }
  __syncthreads();
  const int P3 = S / P12;

  const float beta = cp->beta;
  cptr<int> bin_off = cst(sc->bin_off);
  auto main_addr = [&](int l) {
    const int b = l / n_cells;
    const int c = l - b * n_cells;
    const int i = c % L1, j = (c / L1) % L2, k = c / (L1 * L2);
    return b * S + (k + h) * P12 + (j + h) * P1 + (i + h);
  };
  const float *tab = sys->flower_tables;
  cptr<int> gen_prog = cst(sc->gen_prog);
  const FieldCtx fctx{sc, bin_off, gen_prog, n_slots};
  // rate of one change o -> nw of a general site with cluster sums G
  auto change_rate = [&](int b, int o, int nw, const float (&G)[kMaxFunc]) -> float {
    const float dE = general_dE(sys, b, o, nw, G) - (chemL[b * kMaxOcc + nw] - chemL[b * kMaxOcc + o]);
    return dE <= 0.0f ? 1.0f : __expf(-beta * dE);
  };
  // rate of the site at padded address pa on sublattice b (0 for fixed
  // sites); binary: its cluster sums K (which do not depend on the site itself)
  auto site_rate_at = [&](int b, int pa, int (&K)[kMaxSlots]) -> float {
    const int no = noccL[b];
    if (no < 2) return 0.0f;
    const int o = img.get(pa);
    if (BIN) {
      const float F = FIELD::bin(fctx, img, b, pa, o, K);
      const float chem0 = chemL[b * kMaxOcc + 0], chem1 = chemL[b * kMaxOcc + 1];
      const float dE = (o ? -2.0f : 2.0f) * F - (o ? (chem0 - chem1) : (chem1 - chem0));
      return dE <= 0.0f ? 1.0f : __expf(-beta * dE);
    }
    float G[kMaxFunc];
    FIELD::gen(fctx, img, b, pa, tab, G);
    float r = 0.0f;
    for (int q = 0; q < no; ++q)
      if (q != o) r += change_rate(b, o, q, G);
    return r / (float)(no - 1);
  };
  auto site_rate = [&](int l) -> float {
    int K[kMaxSlots];
    return site_rate_at(l / n_cells, main_addr(l), K);
  };
  for (int l = tid; l < n_groups * 64; l += nthr) rate[l] = l < n_sites ? site_rate(l) : 0.0f;
  __syncthreads();
  for (int g = tid; g < n_groups; g += nthr) G[g] = group_rate_sum(rate, g);
  __syncthreads();

  // per-chain running quantities (lane 0 of the workgroup)
  long long C[kMaxSlots] = {0, 0, 0, 0, 0, 0, 0, 0};  // binary: cluster-count changes per slot
  long long cnt[kMaxOcc];
  for (int q = 0; q < kMaxOcc; ++q) cnt[q] = cs->counts[q];
  double E_run = cs->E;  // general: running energy
  // residence-weighted moments of the current block (kNfoldBlockSums layout)
  double mom[kNfoldBlockSums];
#pragma unroll
  for (int q = 0; q < kNfoldBlockSums; ++q) mom[q] = 0.0;
  long long n_done = 0;
  double n_mutable = 0.0;
  for (int b = 0; b < sys->n_sublat; ++b)
    if (sys->n_occ[b] >= 2) n_mutable += (double)n_cells;
  // binary: the weight of each slot's orbit, read once (the energy of a
  // state is formed at every event)
  double wslot[kMaxSlots];
#pragma unroll
  for (int j = 0; j < kMaxSlots; ++j) wslot[j] = (BIN && j < n_slots) ? sys->bin_w64[sys->slot_orbit[j]] : 0.0;
  double E0 = 0.0;  // binary: energy of the exact counts at launch
  if (BIN) {
    const long long *C0 = args.binC + (long long)chain * args.n_orbits;
    for (int a = 0; a < args.n_orbits; ++a) E0 += sys->bin_w64[a] * (double)C0[a];
  }
  const unsigned long long ev0 = cs->pass_base;
  // sampling by N-fold time (na.n_time_samples > 0): samples at times
  // 0, dt, 2 dt, ... (passes since the launch) of the state occupied then
  const int n_tsamp = na.n_time_samples;
  // (the sample storage is only passed in for time sampling)
  const int s_base = n_tsamp > 0 ? args.sample_count[chain] : 0;
  double T_now = 0.0;
  int n_rec = 0;
  if (tid == 0) sel[3] = 0;

  for (int ev = 0; ev < na.n_events; ++ev) {
    // ---- select an event (wavefront 0)
    if (wave == 0) {
      const int chunk = (n_groups + 63) / 64;
      const int g0 = min(lane * chunk, n_groups), g1 = min(g0 + chunk, n_groups);
      double ls = 0.0;
      for (int g = g0; g < g1; ++g) ls += G[g];
      const double incl = wave_scan_incl_d(ls);
      const double R = lane_d(incl, 63);
      const unsigned long long gev = ev0 + (unsigned long long)ev;
      const uint4 r = philox4x32_10(make_uint4(0x4E464C44u, (uint32_t)gev, (uint32_t)(gev >> 32), (uint32_t)chain),
                                    cp->key0, cp->key1);
      const double t = u01_53(r.x, r.y) * R;
      int site = -1;
      const bool live = R >= kNfoldMinTotalRate;
      const int L = live ? pick_lane(incl, ls, t) : -1;
      if (L >= 0) {
        const double before = lane_d(incl, L) - lane_d(ls, L);
        const int gl0 = __builtin_amdgcn_readlane(g0, L), gl1 = __builtin_amdgcn_readlane(g1, L);
        const double gv = (gl0 + lane < gl1) ? G[gl0 + lane] : 0.0;
        const double gi = wave_scan_incl_d(gv);
        const double t2 = t - before;
        const int lg = pick_lane(gi, gv, t2);
        if (lg >= 0) {
          const int g = gl0 + lg;
          const double before2 = lane_d(gi, lg) - lane_d(gv, lg);
          const double rv = (double)rate[g * 64 + lane];
          const double ri = wave_scan_incl_d(rv);
          const int ls2 = pick_lane(ri, rv, t2 - before2);
          if (ls2 >= 0) site = g * 64 + ls2;
        }
      }
      if (lane == 0) {
        // a chain whose total rate vanished (or a group sum without a
        // positive rate, which exact group sums rule out) stops: -2
        sel[0] = site >= 0 ? site : -2;
        sel[1] = (int)r.z;  // the new occupant (general sites): 53-bit uniform
        sel[2] = (int)r.w;
        if (site >= 0) {
          // weight of the current state: expected Metropolis steps spent in it
          double E = E0;
          if (BIN) {
#pragma unroll
            for (int j = 0; j < kMaxSlots; ++j)
              if (j < n_slots) E += wslot[j] * (double)C[j];
          } else {
            E = E_run;
          }
          const double w = n_mutable / (double)R;
          // sampling by N-fold time: every sample time in this state's
          // residence interval [T, T + w / n_mutable) (in passes) samples it
          if (n_tsamp > 0) {
            const double t1 = T_now + w / n_mutable;
            while (n_rec < n_tsamp && (double)n_rec * na.sample_dt < t1) {
              const long long so = (long long)chain * args.sample_cap + (s_base + n_rec);
              args.sampE[so] = E;
#pragma unroll
              for (int q = 0; q < kMaxOcc; ++q) args.sampN[so * kMaxOcc + q] = (int)cnt[q];
              ++n_rec;
            }
            T_now = t1;
            if (n_rec >= n_tsamp) sel[3] = 1;  // every sample taken: stop before this event
          }
          mom[0] += w;
          mom[1] += w * E;
          mom[2] += w * E * E;
          int t = 0;
#pragma unroll
          for (int o = 0; o < kMaxOcc; ++o) {
            const double wn = w * (double)cnt[o];
            mom[3 + o] += wn;
            mom[17 + o] += wn * E;
#pragma unroll
            for (int q = o; q < kMaxOcc; ++q, ++t) mom[7 + t] += wn * (double)cnt[q];
          }
        }
      }
    }
    __syncthreads();
    const int site = __builtin_amdgcn_readfirstlane(sel[0]);  // the same in every lane
    if (site < 0) {
      // uniform across the workgroup: record the stall for the host
      if (tid == 0) cs->flags |= kChainFlagNfoldStalled;
      break;
    }
    if (sel[3]) break;  // time sampling complete (uniform across the workgroup)

    // ---- apply the change (one lane); binary: its cluster sums K come with
    // its new rate below; general: the new occupant in proportion to its rate
    int o_flip = 0;
    if (tid == 0) {
      const int b = site / n_cells;
      const int pa = main_addr(site);
      const int o = img.get(pa);
      o_flip = o;
      int nw = 1 - o;
      if (!BIN) {
        const int no = noccL[b];
        float G[kMaxFunc];
        FIELD::gen(fctx, img, b, pa, tab, G);
        double rq[kMaxOcc] = {0.0, 0.0, 0.0, 0.0}, tot = 0.0;
        for (int q = 0; q < no; ++q)
          if (q != o) {
            rq[q] = (double)change_rate(b, o, q, G);
            tot += rq[q];
          }
        const double t = u01_53((uint32_t)sel[1], (uint32_t)sel[2]) * tot;
        double acc = 0.0;
        nw = -1;
        for (int q = 0; q < no; ++q) {
          if (q == o) continue;
          acc += rq[q];
          if (nw < 0 && t < acc) nw = q;
        }
        if (nw < 0) nw = (o + 1 < no) ? no - 1 : no - 2;  // rounding: the last other occupant
        E_run += (double)general_dE(sys, b, o, nw, G);
      }
      cnt[o] -= 1;
      cnt[nw] += 1;
      // the site and every periodic copy of it in the padded image: padded
      // coordinate (i + h) + d L for every d that stays inside [0, L + 2h)
      {
        const int c = site - b * n_cells;
        const int ci = c % L1, cj = (c / L1) % L2, ck = c / (L1 * L2);
        if (h <= L1 && h <= L2 && h <= L3) {
          // d in {0, +1 (i < h), -1 (i >= L - h)}: at most three per axis
          const int qi[3] = {ci + h, ci + h + L1, ci + h - L1}, qj[3] = {cj + h, cj + h + L2, cj + h - L2},
                    qk[3] = {ck + h, ck + h + L3, ck + h - L3};
          const bool vi[3] = {true, ci < h, ci >= L1 - h}, vj[3] = {true, cj < h, cj >= L2 - h},
                     vk[3] = {true, ck < h, ck >= L3 - h};
#pragma unroll
          for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int e = 0; e < 3; ++e)
#pragma unroll
              for (int f = 0; f < 3; ++f)
                if (vk[a] && vj[e] && vi[f]) img.p[b * S + qk[a] * P12 + qj[e] * P1 + qi[f]] = (uint8_t)nw;
        } else {
        const int m1 = (h + L1 - 1) / L1, m2 = (h + L2 - 1) / L2, m3 = (h + L3 - 1) / L3;
        for (int dk = -m3; dk <= m3; ++dk) {
          const int qk = ck + h + dk * L3;
          if (qk < 0 || qk >= P3) continue;
          for (int dj = -m2; dj <= m2; ++dj) {
            const int qj = cj + h + dj * L2;
            if (qj < 0 || qj >= P12 / P1) continue;
            for (int di = -m1; di <= m1; ++di) {
              const int qi = ci + h + di * L1;
              if (qi < 0 || qi >= P1) continue;
              img.p[b * S + qk * P12 + qj * P1 + qi] = (uint8_t)nw;
            }
          }
        }
        }
      }
      ++n_done;
    }
    // ---- rates of the site and its interaction neighbours
    {
      const int b = site / n_cells;
      const int c = site - b * n_cells;
      const int i = c % L1, j = (c / L1) % L2, k = c / (L1 * L2);
      const int nb0 = nbB[b], nb1 = nbB[b + 1];
      const bool distinct = sc->nb_distinct != 0;  // no duplicate sites: no list in LDS, no search
      if (!distinct)
        for (int t = tid; t < nb1 - nb0; t += nthr) {
          const int4 o = nbL[nb0 + t];
          // (neighbour offsets are within the halo width h)
          const int it = pad_wrap(i + o.y + h, h, L1), jt = pad_wrap(j + o.z + h, h, L2), kt = pad_wrap(k + o.w + h, h, L3);
          aff[t] = o.x * n_cells + it + L1 * (jt + L2 * kt);
          affp[t] = o.x * S + (kt + h) * P12 + (jt + h) * P1 + (it + h);
        }
      __syncthreads();  // the flip (thread 0) and the list before any rate is read
      for (int t = tid; t < nb1 - nb0; t += nthr) {
        int lt, pt;
        if (distinct) {
          const int4 o = nbL[nb0 + t];
          // (neighbour offsets are within the halo width h)
          const int it = pad_wrap(i + o.y + h, h, L1), jt = pad_wrap(j + o.z + h, h, L2), kt = pad_wrap(k + o.w + h, h, L3);
          lt = o.x * n_cells + it + L1 * (jt + L2 * kt);
          pt = o.x * S + (kt + h) * P12 + (jt + h) * P1 + (it + h);
        } else {
          lt = aff[t];
          pt = affp[t];
        }
        bool dup = false;
        if (!distinct)
          for (int q = 0; q < t; ++q) dup = dup || (aff[q] == lt);
        if (distinct) aff[t] = lt;  // the groups to sum again below
        if (!dup) {
          int K[kMaxSlots];
          rate[lt] = site_rate_at(nbL[nb0 + t].x, pt, K);
          if (BIN && t == 0) {
            // the flipped site itself (first in the list, lane 0): exact cluster-count change
            const int sk = o_flip ? -2 : 2;
#pragma unroll
            for (int a = 0; a < kMaxSlots; ++a) C[a] += (long long)sk * K[a];
          }
        }
      }
      __syncthreads();
      // the groups of the changed rates, summed again from their 64 rates
      // (several lanes may form one group's sum: they write the same value)
      for (int t = tid; t < nb1 - nb0; t += nthr) {
        const int g = aff[t] >> 6;
        G[g] = group_rate_sum(rate, g);
      }
      __syncthreads();
    }
    if (tid == 0 && ((ev + 1) % na.events_per_block) == 0) {
      const int blk = (ev + 1) / na.events_per_block - 1;
      if (blk < na.n_blocks) {
        double *out = na.blocks + ((long long)chain * na.n_blocks + blk) * kNfoldBlockSums;
#pragma unroll
        for (int q = 0; q < kNfoldBlockSums; ++q) out[q] = mom[q];
      }
#pragma unroll
      for (int q = 0; q < kNfoldBlockSums; ++q) mom[q] = 0.0;
    }
  }

  // ---- write back: occupation, exact counts, energy, counters
  int8_t *occW = args.occ + cp->occ_offset;
  for (int l = tid; l < n_sites; l += nthr) occW[l] = (int8_t)img.get(main_addr(l));
  if (tid == 0) {
    if (BIN) {
      long long *Cg = args.binC + (long long)chain * args.n_orbits;
      for (int j = 0; j < n_slots; ++j) Cg[sys->slot_orbit[j]] += C[j];
      double E = 0.0;
      for (int a = 0; a < args.n_orbits; ++a) E += sys->bin_w64[a] * (double)Cg[a];
      cs->E = E;
    } else {
      cs->E = E_run;
    }
    for (int q = 0; q < kMaxOcc; ++q) cs->counts[q] = cnt[q];
    if (n_tsamp > 0) args.sample_count[chain] = s_base + n_rec;
    cs->n_attempt += n_done;
    cs->n_accept += n_done;
    cs->pass_base = ev0 + (unsigned long long)na.n_events;
  }
}

}  // namespace clexmc
