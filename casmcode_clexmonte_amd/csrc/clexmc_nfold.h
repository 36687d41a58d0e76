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
// One workgroup per chain.  LDS holds the padded int8 image, the rate of
// every site and its sums over groups of 64 sites.  Per event:
//  * wavefront 0 forms the total rate R, draws u, finds the group by a scan
//    over per-lane chunks of group sums, then the site by a scan of the
//    group's 64 rates (all prefix sums are wavefront scans);
//  * the current state is weighted by its expected residence time
//    w = n_mutable / R in Metropolis steps (the jump chain visits states in
//    proportion to pi(x) R(x); weights 1/R recover pi);
//  * one lane applies the change (the new occupant drawn in proportion to
//    its rate; binary: exact integer cluster counts, general: FP64 energy;
//    every periodic copy of the site), then the lanes recompute the rates of the site and
//    of its interaction neighbours (each neighbour once) and add the changes
//    to the group sums with LDS float atomics;
//  * every 128 events the group sums are summed again from the rates, so
//    rounding from the atomic updates does not accumulate.
//
// nfold_body<BIN, FIELD>: FIELD is GenericField (interpreted site programs,
// clexmc_nfold.hip) or the run-time specialised JitField (clexmc_jit.cpp);
// both sum the clusters in the same order, so the results are identical.
#pragma once
#include "clexmc_metropolis.h"

namespace clexmc {

__host__ __device__ inline size_t nfold_align16(size_t x) { return (x + 15) & ~(size_t)15; }


__device__ __forceinline__ float wave_scan_incl_f(float v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const float y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  return v;
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// first lane whose inclusive prefix exceeds t; if rounding leaves none, the
// last lane with a positive value (-1 if there is none)
__device__ __forceinline__ int pick_lane(float incl, float v, float t) {
  const unsigned long long m = __ballot(incl > t && v > 0.0f);
  if (m) return __ffsll((long long)m) - 1;
  const unsigned long long m2 = __ballot(v > 0.0f);
  return m2 ? 63 - __clzll((long long)m2) : -1;
}




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
  float *rate = (float *)(smem + off);
  off += nfold_align16((size_t)n_sites * 4);
  float *G = (float *)(smem + off);
  off += nfold_align16((size_t)n_groups * 4);
  int *aff = (int *)(smem + off);    // linear indices of the sites whose rates change
  off += nfold_align16((size_t)sc->nb_max * 4);
  int *affp = (int *)(smem + off);   // their padded addresses
  off += nfold_align16((size_t)sc->nb_max * 4);
  int *sel = (int *)(smem + off);

  // ---- padded image
  const int8_t *__restrict__ occG = args.occ + cp->occ_offset;
  for (int p = tid; p < pad_bytes; p += nthr) {
    const int b = p / S;
    const int r = p - b * S;
    const int pk = r / P12, pj = (r - pk * P12) / P1, pi = r - pk * P12 - pj * P1;
    const int i = ((pi - h) % L1 + L1) % L1, j = ((pj - h) % L2 + L2) % L2, k = ((pk - h) % L3 + L3) % L3;
    img.p[p] = (uint8_t)occG[b * n_cells + i + L1 * (j + L2 * k)];
  }
  __syncthreads();

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
    const float dE = general_dE(sys, b, o, nw, G) - (cp->chem[b * kMaxOcc + nw] - cp->chem[b * kMaxOcc + o]);
    return dE <= 0.0f ? 1.0f : __expf(-beta * dE);
  };
  // rate of the site at padded address pa on sublattice b (0 for fixed
  // sites); binary: its cluster sums K (which do not depend on the site itself)
  auto site_rate_at = [&](int b, int pa, int (&K)[kMaxSlots]) -> float {
    const int no = sys->n_occ[b];
    if (no < 2) return 0.0f;
    const int o = img.get(pa);
    if (BIN) {
      const float F = FIELD::bin(fctx, img, b, pa, o, K);
      const float chem0 = cp->chem[b * kMaxOcc + 0], chem1 = cp->chem[b * kMaxOcc + 1];
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
  auto refresh_groups = [&]() {
    for (int g = wave; g < n_groups; g += n_waves) {
      const int l = g * 64 + lane;
      const float s = wave_sum_f(l < n_sites ? rate[l] : 0.0f);
      if (lane == 0) G[g] = s;
    }
  };

  for (int l = tid; l < n_sites; l += nthr) rate[l] = site_rate(l);
  __syncthreads();
  refresh_groups();
  __syncthreads();

  // per-chain running quantities (lane 0 of the workgroup)
  long long C[kMaxSlots] = {0, 0, 0, 0, 0, 0, 0, 0};  // binary: cluster-count changes per slot
  long long cnt[kMaxOcc];
  for (int q = 0; q < kMaxOcc; ++q) cnt[q] = cs->counts[q];
  double E_run = cs->E;  // general: running energy
  double W = 0.0, WE = 0.0, WN = 0.0, WE2 = 0.0, WN2 = 0.0, WEN = 0.0;
  long long n_done = 0;
  double n_mutable = 0.0;
  for (int b = 0; b < sys->n_sublat; ++b)
    if (sys->n_occ[b] >= 2) n_mutable += (double)n_cells;
  double E0 = 0.0;  // binary: energy of the exact counts at launch
  if (BIN) {
    const long long *C0 = args.binC + (long long)chain * args.n_orbits;
    for (int a = 0; a < args.n_orbits; ++a) E0 += sys->bin_w64[a] * (double)C0[a];
  }
  const unsigned long long ev0 = cs->pass_base;

  bool retried = false;
  for (int ev = 0; ev < na.n_events; ++ev) {
    // ---- select an event (wavefront 0)
    if (wave == 0) {
      const int chunk = (n_groups + 63) / 64;
      const int g0 = min(lane * chunk, n_groups), g1 = min(g0 + chunk, n_groups);
      float ls = 0.0f;
      for (int g = g0; g < g1; ++g) ls += G[g];
      const float incl = wave_scan_incl_f(ls, lane);
      const float R = __shfl(incl, 63, 64);
      const unsigned long long gev = ev0 + (unsigned long long)ev;
      const uint4 r = philox4x32_10(make_uint4(0x4E464C44u, (uint32_t)gev, (uint32_t)(gev >> 32), (uint32_t)chain),
                                    cp->key0, cp->key1);
      const float t = u01(r.x) * R;
      int site = -1;
      const int L = R > 0.0f ? pick_lane(incl, ls, t) : -1;
      if (L >= 0) {
        const float before = __shfl(incl, L, 64) - __shfl(ls, L, 64);
        const int gl0 = __shfl(g0, L, 64), gl1 = __shfl(g1, L, 64);
        const float gv = (gl0 + lane < gl1) ? G[gl0 + lane] : 0.0f;
        const float gi = wave_scan_incl_f(gv, lane);
        const float t2 = t - before;
        const int lg = pick_lane(gi, gv, t2);
        if (lg >= 0) {
          const int g = gl0 + lg;
          const float before2 = __shfl(gi, lg, 64) - __shfl(gv, lg, 64);
          const int l = g * 64 + lane;
          const float rv = l < n_sites ? rate[l] : 0.0f;
          const float ri = wave_scan_incl_f(rv, lane);
          const int ls2 = pick_lane(ri, rv, t2 - before2);
          if (ls2 >= 0) site = g * 64 + ls2;
        }
      }
      if (lane == 0) {
        // a group sum left positive by rounding with no positive rate in the
        // group: sum the groups again and draw once more
        sel[0] = (site < 0 && R > 0.0f && !retried) ? -2 : site;
        sel[1] = __float_as_int(u01(r.y));  // draws the new occupant (general sites)
        if (site >= 0) {
          // weight of the current state: expected Metropolis steps spent in it
          double E = E0;
          if (BIN)
            for (int j = 0; j < n_slots; ++j) E += sys->bin_w64[sys->slot_orbit[j]] * (double)C[j];
          else
            E = E_run;
          const double w = n_mutable / (double)R;
          const double x = (double)cnt[1];
          W += w;
          WE += w * E;
          WN += w * x;
          WE2 += w * E * E;
          WN2 += w * x * x;
          WEN += w * E * x;
        }
      }
    }
    __syncthreads();
    const int site = sel[0];
    if (site == -2) {
      refresh_groups();
      __syncthreads();
      retried = true;
      --ev;
      continue;
    }
    retried = false;
    if (site < 0) break;  // no event has a positive rate (uniform across the workgroup)

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
        const int no = sys->n_occ[b];
        float G[kMaxFunc];
        FIELD::gen(fctx, img, b, pa, tab, G);
        float rq[kMaxOcc] = {0.0f, 0.0f, 0.0f, 0.0f}, tot = 0.0f;
        for (int q = 0; q < no; ++q)
          if (q != o) {
            rq[q] = change_rate(b, o, q, G);
            tot += rq[q];
          }
        const float t = __int_as_float(sel[1]) * tot;
        float acc = 0.0f;
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
      // the site and every periodic copy of it (the halo list of its class position)
      img.p[pa] = (uint8_t)nw;
      const int tp = sc->site_tpos[site];
      for (int q = sc->halo_site[tp]; q < sc->halo_site[tp + 1]; ++q) img.p[sc->halo[q].y] = (uint8_t)nw;
      ++n_done;
    }
    // ---- rates of the site and its interaction neighbours
    {
      const int b = site / n_cells;
      const int c = site - b * n_cells;
      const int i = c % L1, j = (c / L1) % L2, k = c / (L1 * L2);
      const int nb0 = sc->nb_begin[b], nb1 = sc->nb_begin[b + 1];
      const bool distinct = sc->nb_distinct != 0;  // no duplicate sites: no list in LDS, no search
      if (!distinct)
        for (int t = tid; t < nb1 - nb0; t += nthr) {
          const int4 o = sc->nb_cell[nb0 + t];
          const int it = ((i + o.y) % L1 + L1) % L1, jt = ((j + o.z) % L2 + L2) % L2, kt = ((k + o.w) % L3 + L3) % L3;
          aff[t] = o.x * n_cells + it + L1 * (jt + L2 * kt);
          affp[t] = o.x * S + (kt + h) * P12 + (jt + h) * P1 + (it + h);
        }
      __syncthreads();  // the flip (thread 0) and the list before any rate is read
      for (int t = tid; t < nb1 - nb0; t += nthr) {
        int lt, pt;
        if (distinct) {
          const int4 o = sc->nb_cell[nb0 + t];
          const int it = ((i + o.y) % L1 + L1) % L1, jt = ((j + o.z) % L2 + L2) % L2, kt = ((k + o.w) % L3 + L3) % L3;
          lt = o.x * n_cells + it + L1 * (jt + L2 * kt);
          pt = o.x * S + (kt + h) * P12 + (jt + h) * P1 + (it + h);
        } else {
          lt = aff[t];
          pt = affp[t];
        }
        bool dup = false;
        if (!distinct)
          for (int q = 0; q < t; ++q) dup = dup || (aff[q] == lt);
        if (!dup) {
          int K[kMaxSlots];
          const float nr = site_rate_at(sc->nb_cell[nb0 + t].x, pt, K);
          const float d = nr - rate[lt];
          rate[lt] = nr;
          atomicAdd(&G[lt >> 6], d);
          if (BIN && t == 0) {
            // the flipped site itself (first in the list, lane 0): exact cluster-count change
            const int sk = o_flip ? -2 : 2;
#pragma unroll
            for (int a = 0; a < kMaxSlots; ++a) C[a] += (long long)sk * K[a];
          }
        }
      }
      __syncthreads();
    }
    if ((ev & 127) == 127) {
      refresh_groups();
      __syncthreads();
    }
    if (tid == 0 && ((ev + 1) % na.events_per_block) == 0) {
      const int blk = (ev + 1) / na.events_per_block - 1;
      if (blk < na.n_blocks) {
        double *out = na.blocks + ((long long)chain * na.n_blocks + blk) * kNfoldBlockSums;
        out[0] = W;
        out[1] = WE;
        out[2] = WN;
        out[3] = WE2;
        out[4] = WN2;
        out[5] = WEN;
      }
      W = WE = WN = WE2 = WN2 = WEN = 0.0;
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
    cs->n_attempt += n_done;
    cs->n_accept += n_done;
    cs->pass_base = ev0 + (unsigned long long)na.n_events;
  }
}

}  // namespace clexmc
