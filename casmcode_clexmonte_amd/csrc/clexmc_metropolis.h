// Device code of the persistent Metropolis kernel, shared by the library's
// ahead-of-time build (generic site programs) and the run-time compiled
// kernels specialised to one supercell shape (constant site offsets).
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include <type_traits>

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
    --left;
    return u01(word(left));
  }
  // one 32-bit word of the stream (acceptance tests)
  __device__ __forceinline__ uint32_t next_u32() {
    if (left == 0) {
      buf = philox4x32_10(make_uint4(lane, (uint32_t)draw, (uint32_t)(draw >> 32), chain), k0, k1);
      ++draw;
      left = 4;
    }
    --left;
    return word(left);
  }
  // word 3 - r of the block (r = words left after this one): selects, not
  // branches, so lanes whose counters differ stay converged
  __device__ __forceinline__ uint32_t word(int r) const {
    const uint32_t b0 = 0u - (uint32_t)(r & 1), b1 = 0u - (uint32_t)((r >> 1) & 1);
    const uint32_t lo = (buf.z & b0) | (buf.w & ~b0);  // r = 1 : 0
    const uint32_t hi = (buf.x & b0) | (buf.y & ~b0);  // r = 3 : 2
    return (hi & b1) | (lo & ~b1);
  }
};

// block-uniform random integer in [0, n) for (pass, phase)
__device__ __forceinline__ uint32_t uniform_int(uint32_t n, unsigned long long pass, uint32_t phase,
                                                uint32_t chain, uint32_t k0, uint32_t k1) {
  uint4 r = philox4x32_10(make_uint4(0xFFFFFFFFu, (uint32_t)pass, phase, chain ^ 0x9E3779B9u), k0, k1);
  unsigned long long x = ((unsigned long long)r.x << 32) | r.y;
  return n ? (uint32_t)(x % n) : 0u;
}

// Sum over the 64 lanes of a full wave, returned to every lane.  Four DPP
// row shifts (row_shr 1, 2, 4, 8; lanes shifted in from outside a row of 16
// add 0) leave each row's sum in the row's last lane, and the four row sums
// are read out as scalars: VALU work only, where a butterfly of ds_bpermute
// shuffles waits on six LDS round trips per value.
__device__ __forceinline__ int wave_sum(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);
  return __builtin_amdgcn_readlane(v, 15) + __builtin_amdgcn_readlane(v, 31) + __builtin_amdgcn_readlane(v, 47) +
         __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---------------------------------------------------------------------------
// LDS occupation image accessors.  A site is written only by the lane that
// owns it in the current phase (sites of a class share no cluster), but with
// packing 16 sites share a word, so packed writes are atomic XORs of the
// changed bits (LDS atomics); the other sites' bits are untouched.
// ---------------------------------------------------------------------------
struct ImgU8 {
  uint8_t *p;
  __device__ __forceinline__ int get(int a) const { return p[a]; }
  __device__ __forceinline__ void set(int a, int old, int nw) const {
    (void)old;
    p[a] = (uint8_t)nw;
  }
  __device__ __forceinline__ void copy(int dst, int src) const { p[dst] = p[src]; }
};

struct ImgP2 {
  uint32_t *w;
  __device__ __forceinline__ int get(int a) const {
    return (int)__builtin_amdgcn_ubfe(w[a >> 4], (uint32_t)((a & 15) << 1), 2u);
  }
  __device__ __forceinline__ void set(int a, int old, int nw) const {
    atomicXor(&w[a >> 4], (uint32_t)(old ^ nw) << ((a & 15) << 1));
  }
  __device__ __forceinline__ void copy(int dst, int src) const {
    const int v = get(src), o = get(dst);
    if (v != o) set(dst, o, v);
  }
};

// Mixed-radix position (b, pk, pj, pi) of an index into a P1 x P2 x P3 box
// per sublattice, advanced by a fixed stride with carries instead of
// divisions (image loads and stores visit indices t, t + s, t + 2s, ...).
struct PadWalk {
  int pi, pj, pk, b;
  int si, sj, sk, sb;
  int P1, P2, P3;
  __device__ __forceinline__ PadWalk(int p, int stride, int P1_, int P2_, int P3_) : P1(P1_), P2(P2_), P3(P3_) {
    const int P12 = P1 * P2, S = P12 * P3;
    b = p / S;
    int r = p - b * S;
    pk = r / P12;
    r -= pk * P12;
    pj = r / P1;
    pi = r - pj * P1;
    sb = stride / S;
    r = stride - sb * S;
    sk = r / P12;
    r -= sk * P12;
    sj = r / P1;
    si = r - sj * P1;
  }
  __device__ __forceinline__ void step() {
    pi += si;
    int c = pi >= P1 ? 1 : 0;
    pi -= c * P1;
    pj += sj + c;
    c = pj >= P2 ? 1 : 0;
    pj -= c * P2;
    pk += sk + c;
    c = pk >= P3 ? 1 : 0;
    pk -= c * P3;
    b += sb + c;
  }
};

// the cell coordinate of padded coordinate q (halo h, period L): q - h in [0, L)
// (q in [0, L + 2h), so for h <= L one period either way is enough: two
// selects; h and L are the same in every lane).  Bytes past L + 2h in a row
// (the unused part of a padded row pitch) are clamped to a valid cell: their
// image value is never read, but the global read of it must stay in bounds.
__device__ __forceinline__ int pad_wrap(int q, int h, int L) {
  int x = q - h;
  if (h <= L) {
    x = x < 0 ? x + L : x;
    x = x >= L ? x - L : x;
    return x < L ? x : L - 1;
  }
  x %= L;
  return x < 0 ? x + L : x;
}

// bytes of the LDS image for pad_sites padded sites, a multiple of 16 so that
// the arrays placed after it keep their natural (up to 16-byte) alignment
__host__ __device__ inline int image_bytes(int pad_sites, bool packed) {
  return packed ? ((pad_sites + 63) / 64) * 16 : ((pad_sites + 15) & ~15);
}

// ---------------------------------------------------------------------------
// Binary path: K_a(s) = sum over clusters of orbit a containing s of
// prod phi(other sites) = (-1)^(k-1) * (n_a - 2 * sum XOR(other bits)).
// The site program lists the padded offsets of the other sites per orbit,
// padded to multiples of 4 clusters with offset-0 entries (the site itself):
// an even number of copies of the site XORs to 0, an odd number contributes
// the site's own bit, removed below.
// ---------------------------------------------------------------------------
template <class IMG>
__device__ __forceinline__ float binary_field(const IMG &img, int pa, int o, cptr<BinSub> bs, cptr<int> off,
                                              int n_slots, int (&K)[kMaxSlots]) {
  float F = 0.0f;
#pragma unroll
  for (int j = 0; j < kMaxSlots; ++j) {
    K[j] = 0;
    if (j < n_slots) {
      const int n = bs->slot_n[j];
      if (n > 0) {
        const int ko = bs->slot_k[j];
        const int q0 = bs->slot_begin[j];
        const int len = bs->slot_len[j];
        int X = 0;
        if (ko == 1) {
#pragma unroll 2
          for (int e = 0; e < len; e += 4) {
            const int q = q0 + e;
            const int b0 = img.get(pa + off[q]);
            const int b1 = img.get(pa + off[q + 1]);
            const int b2 = img.get(pa + off[q + 2]);
            const int b3 = img.get(pa + off[q + 3]);
            X += (b0 + b1) + (b2 + b3);
          }
        } else if (ko == 2) {
#pragma unroll 2
          for (int e = 0; e < len; e += 4) {
            const int q = q0 + 2 * e;
            const int b0 = img.get(pa + off[q]);
            const int b1 = img.get(pa + off[q + 1]);
            const int b2 = img.get(pa + off[q + 2]);
            const int b3 = img.get(pa + off[q + 3]);
            const int b4 = img.get(pa + off[q + 4]);
            const int b5 = img.get(pa + off[q + 5]);
            const int b6 = img.get(pa + off[q + 6]);
            const int b7 = img.get(pa + off[q + 7]);
            X += ((b0 ^ b1) + (b2 ^ b3)) + ((b4 ^ b5) + (b6 ^ b7));
          }
        } else if (ko >= 3) {
          for (int e = 0; e < len; ++e) {
            int x = 0;
            for (int q = q0 + e * ko; q < q0 + (e + 1) * ko; ++q) x ^= img.get(pa + off[q]);
            X += x;
          }
        }
        if (ko & 1) X -= bs->slot_dummy[j] * o;
        const int Kj = (ko & 1) ? (2 * X - n) : (n - 2 * X);
        K[j] = Kj;
        F = __builtin_fmaf(bs->slot_w[j], (float)Kj, F);
      }
    }
  }
  return F;
}

// Partial cluster sums of the binary field for lane pairs: part p of nparts
// takes a contiguous share of each orbit's cluster list (in chunks of 4, the
// padding unit); part 0 also removes the padding's self copies.  Summing the
// parts gives binary_field's X exactly.
template <class IMG>
__device__ __forceinline__ void binary_partial_x(const IMG &img, int pa, int o, cptr<BinSub> bs, cptr<int> off,
                                                 int n_slots, int part, int nparts, int (&X)[kMaxSlots]) {
#pragma unroll
  for (int j = 0; j < kMaxSlots; ++j) {
    X[j] = 0;
    if (j < n_slots) {
      const int n = bs->slot_n[j];
      if (n > 0) {
        const int ko = bs->slot_k[j];
        const int q0 = bs->slot_begin[j];
        const int chunks = bs->slot_len[j] / 4;
        const int c0 = (chunks * part) / nparts, c1 = (chunks * (part + 1)) / nparts;
        int x = 0;
        for (int e = 4 * c0; e < 4 * c1; ++e) {
          int y = 0;
          for (int q = q0 + e * ko; q < q0 + (e + 1) * ko; ++q) y ^= img.get(pa + off[q]);
          x += y;
        }
        if (part == 0 && (ko & 1)) x -= bs->slot_dummy[j] * o;
        X[j] = x;
      }
    }
  }
}

// K and F from complete cluster sums X (the same arithmetic as binary_field)
__device__ __forceinline__ float binary_finish(cptr<BinSub> bs, int n_slots, const int (&X)[kMaxSlots],
                                               int (&K)[kMaxSlots]) {
  float F = 0.0f;
#pragma unroll
  for (int j = 0; j < kMaxSlots; ++j) {
    K[j] = 0;
    if (j < n_slots) {
      const int n = bs->slot_n[j];
      if (n > 0) {
        const int ko = bs->slot_k[j];
        const int Kj = (ko & 1) ? (2 * X[j] - n) : (n - 2 * X[j]);
        K[j] = Kj;
        F = __builtin_fmaf(bs->slot_w[j], (float)Kj, F);
      }
    }
  }
  return F;
}

// value of the partner lane (lane ^ 1) by DPP quad permutation [1, 0, 3, 2]
__device__ __forceinline__ int pair_swap(int v) { return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false); }

// General path: G_f(s) = sum_e V_e[tau_others][f] from the flower tables.
template <class IMG>
__device__ __forceinline__ void general_field(const IMG &img, int pa, cptr<GenSub> gs, cptr<int> prog,
                                              const float *tab, float (&G)[kMaxFunc]) {
  G[0] = G[1] = G[2] = 0.0f;
  const int nf = gs->nf;
  const int n = gs->n_entries;
  int q = gs->begin;
  for (int e = 0; e < n; ++e) {
    const int k = prog[q];
    const int toff = prog[q + 1];
    q += 2;
    int idx = 0;
    for (int i = 0; i < k; ++i) {
      idx += img.get(pa + prog[q]) * prog[q + 1];
      q += 2;
    }
    const float *V = tab + toff + idx * nf;
    G[0] += V[0];
    if (nf > 1) G[1] += V[1];
    if (nf > 2) G[2] += V[2];
  }
}

__device__ __forceinline__ float general_dE(cptr<DevSystem> sys, int b, int o, int nw, const float (&G)[kMaxFunc]) {
  const int nf = sys->n_func[b];
  float d = (sys->site_func[b][0][nw] - sys->site_func[b][0][o]) * G[0];
  if (nf > 1) d += (sys->site_func[b][1][nw] - sys->site_func[b][1][o]) * G[1];
  if (nf > 2) d += (sys->site_func[b][2][nw] - sys->site_func[b][2][o]) * G[2];
  return d;
}

// dE <= 0 accepts (p = 2 by a select, so no branch; exp(-beta dE) could round
// below 1 for dE just below 0)
__device__ __forceinline__ bool metropolis_accept(float dE, float beta, uint32_t x, LaneRng &rng) {
  const float p = dE <= 0.0f ? 2.0f : __expf(-beta * dE);
  return accept_prob(p, x, [&]() { return rng.next_u32(); });
}

// Binary field policy.  FieldCtx carries what the runtime site programs
// need; GenericField interprets them, a specialised policy (compiled at run
// time for one supercell shape) has the offsets built in and ignores it.
struct FieldCtx {
  cptr<DevSupercell> sc;
  cptr<int> bin_off;
  cptr<int> gen_prog;
  int n_slots;
};

struct GenericField {
  static constexpr bool kCube = false;
  struct Cube {};
  template <class IMG>
  __device__ __forceinline__ static float bin(const FieldCtx &c, const IMG &img, int b, int pa, int o,
                                              int (&K)[kMaxSlots]) {
    return binary_field(img, pa, o, cst(c.sc->bin_sub + b), c.bin_off, c.n_slots, K);
  }
  template <class IMG>
  __device__ __forceinline__ static void bin_x(const FieldCtx &c, const IMG &img, int b, int pa, int o, int part,
                                               int (&X)[kMaxSlots]) {
    binary_partial_x(img, pa, o, cst(c.sc->bin_sub + b), c.bin_off, c.n_slots, part, 2, X);
  }
  __device__ __forceinline__ static float bin_finish(const FieldCtx &c, int b, const int (&X)[kMaxSlots],
                                                     int (&K)[kMaxSlots]) {
    return binary_finish(cst(c.sc->bin_sub + b), c.n_slots, X, K);
  }
  template <class IMG>
  __device__ __forceinline__ static void gen(const FieldCtx &c, const IMG &img, int b, int pa, const float *tab,
                                             float (&G)[kMaxFunc]) {
    general_field(img, pa, cst(c.sc->gen_sub + b), c.gen_prog, tab, G);
  }
};

// ---------------------------------------------------------------------------
// Persistent Metropolis kernel
// ---------------------------------------------------------------------------
// SPLIT == 2 (binary path): the two lanes of a pair share one site's field in
// the class sweeps (each sums half of every orbit's clusters, combined with a
// DPP lane swap), which doubles the wavefronts per chain.
// KIND: 0 canonical, 1 semi-grand canonical, 2 semi-grand canonical with
// multiswaps (the canonical sweep followed by the multiswap attempts; a kind
// of its own so that the other kernels carry none of its registers).
//
// `role`: this workgroup's block index within its launch group, the group's
// chain map (nullptr: block -> chain directly) and its slabs per chain.  A
// launch of one group passes (blockIdx.x, args.chain_map, args.slabs); a
// fused launch of several supercell groups passes each group's own.
struct BlockRole {
  int block;
  const int *chain_map;
  int slabs;
};

// CUBE: the register-cube sweep instead of the class loop (the field policy
// must provide it; semi-grand canonical binary one-lane unpacked only).
template <bool BIN, int KIND, bool PACK, class FIELD, int SPLIT = 1, bool CUBE = false>
__device__ __forceinline__ void metropolis_body(const RunArgs &args, const BlockRole role) {
  extern __shared__ __align__(16) unsigned char smem[];
  // slabs > 1: this workgroup runs slab `slab` of its chain
  const int n_slabs = role.slabs > 1 ? role.slabs : 1;
  const int wg = role.block / n_slabs, slab = role.block - wg * n_slabs;
  const int chain = role.chain_map ? role.chain_map[wg] : wg;
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
  const int n_classes = sc->n_classes, max_class = sc->max_class;
  const int n_slots = sys->n_slots;

  // ---- LDS layout ----
  using Img = typename std::conditional<PACK, ImgP2, ImgU8>::type;
  Img img;
  if constexpr (PACK) img.w = (uint32_t *)smem;
  else img.p = (uint8_t *)smem;
  size_t off_b = (size_t)image_bytes(pad_bytes, PACK);
  float *tabL = (float *)(smem + off_b);
  const int n_tab = BIN ? 0 : args.n_tab;
  off_b += (size_t)((n_tab * 4 + 15) & ~15);
  const int n_islot = (BIN ? kMaxSlots : 0) + kMaxOcc + 2;
  int *redI = (int *)(smem + off_b);                     // [n_islot][n_waves]
  off_b += (size_t)n_islot * 16 * sizeof(int);
  double *redD = (double *)(smem + off_b);               // [n_waves]
  off_b += 16 * sizeof(double);
  unsigned long long *maxDraw = (unsigned long long *)(smem + off_b);
  off_b += 16;
  // slab launches: this workgroup's totals since the launch started
  long long *slabI = (long long *)(smem + off_b);        // [kMaxSlots + kMaxOcc + 2]
  double *slabE = (double *)(slabI + kMaxSlots + kMaxOcc + 2);
  off_b += 128;
  // multiswaps: occupant counts per (asymmetric unit, occupant), the search
  // request / result words and the sites picked by the current attempt
  int *poolN = (int *)(smem + off_b);                    // [kMaxSublat][kMaxOcc]
  int *msw = poolN + kMaxSublat * kMaxOcc;               // [8] control + [kMaxMswapSites] picks
  off_b += kMswapLdsBytes;
  // whole-chain launches: the chain's running totals (thread 0 only)
  ChainLds *cl = (ChainLds *)(smem + off_b);
  off_b += kChainLdsBytes;
  const int nfs = BIN ? 1 : sys->n_func_max;              // field values kept per class site
  float *cF = (float *)(smem + off_b);                   // [nfs][max_class]
  int8_t *cK = (int8_t *)(cF + nfs * max_class);         // [kMaxSlots][max_class] (binary)

  // ---- load the padded image (and general tables) ----
  const int8_t *__restrict__ occG = args.occ + cp->occ_offset;
  const int P2 = P12 / P1, P3 = S / P12;
  auto walk_occ = [&](const PadWalk &w) -> int {
    return occG[w.b * n_cells + pad_wrap(w.pi, h, L1) + L1 * (pad_wrap(w.pj, h, L2) + L2 * pad_wrap(w.pk, h, L3))];
  };
  if constexpr (PACK) {
    // each thread assembles whole words (16 padded sites)
    const int n_words = (pad_bytes + 15) / 16;
    PadWalk ws(16 * tid, 16 * nthr, P1, P2, P3);
    for (int wd = tid; wd < n_words; wd += nthr, ws.step()) {
      PadWalk w = ws;
      w.si = 1;
      w.sj = w.sk = w.sb = 0;
      uint32_t word = 0u;
      for (int q = 0; q < 16; ++q, w.step()) {
        const int p = wd * 16 + q;
        if (p < pad_bytes) word |= (uint32_t)(walk_occ(w) & 3) << (2 * q);
      }
      img.w[wd] = word;
    }
  } else {
    // batches of 8 global loads in flight per lane before their LDS stores
    // (one load at a time left a chain of global-memory latencies per lane)
    constexpr int kBatch = 8;
    PadWalk w(tid, nthr, P1, P2, P3);
    for (int p0 = tid; p0 < pad_bytes; p0 += kBatch * nthr) {
      int v[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u, w.step()) v[u] = p0 + u * nthr < pad_bytes ? walk_occ(w) : 0;
#pragma unroll
      for (int u = 0; u < kBatch; ++u)
        if (p0 + u * nthr < pad_bytes) img.p[p0 + u * nthr] = (uint8_t)v[u];
    }
  }
  if (!BIN) {
    const float *__restrict__ tg = sys->flower_tables;
    constexpr int kBatch = 8;
    for (int p0 = tid; p0 < n_tab; p0 += kBatch * nthr) {
      float v[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) v[u] = p0 + u * nthr < n_tab ? tg[p0 + u * nthr] : 0.0f;
#pragma unroll
      for (int u = 0; u < kBatch; ++u)
        if (p0 + u * nthr < n_tab) tabL[p0 + u * nthr] = v[u];
    }
  }
  if (tid == 0) {
    *maxDraw = 0ull;
    for (int q = 0; q < kMaxSlots + kMaxOcc + 2; ++q) slabI[q] = 0;
    *slabE = 0.0;
    if (n_slabs == 1) {
      cl->E = cs->E;
      cl->sumE = cs->sumE;
      cl->sumE2 = cs->sumE2;
      cl->sumN1 = cs->sumN1;
      cl->sumN12 = cs->sumN12;
      cl->sumEN1 = cs->sumEN1;
      for (int o = 0; o < kMaxOcc; ++o) cl->counts[o] = cs->counts[o];
      cl->n_attempt = cs->n_attempt;
      cl->n_accept = cs->n_accept;
      cl->n_samples_acc = cs->n_samples_acc;
      cl->sample_count = args.sample_count ? args.sample_count[chain] : 0;
      if (BIN) {
        const long long *C = args.binC + (long long)chain * args.n_orbits;
        for (int a = 0; a < args.n_orbits; ++a) cl->C[a] = C[a];
      }
    }
  }
  // (multiswap runs: the fixed-composition sweep, kind 2)
  const bool mswaps = KIND == 2 && args.n_mswaps > 0 && n_slabs == 1;
  if (mswaps) {
    if (tid < kMaxSublat * kMaxOcc) poolN[tid] = 0;
    __syncthreads();
    for (int l = tid; l < n_sites; l += nthr) {
      const int b = l / n_cells;
      if (sys->n_occ[b] >= 2) atomicAdd(&poolN[sys->asym[b] * kMaxOcc + occG[l]], 1);
    }
  }
  __syncthreads();

  const float beta = cp->beta;
  LaneRng rng;
  rng.left = 0;
  rng.draw = cs->draw_base;
  rng.lane = (uint32_t)tid | ((uint32_t)slab << 16);
  rng.chain = (uint32_t)chain;
  const uint32_t rchain = (uint32_t)chain | ((uint32_t)slab << 24);  // block-uniform draws of this workgroup
  rng.k0 = cp->key0;
  rng.k1 = cp->key1;

  // per-lane accumulators (cluster-count changes are +-2 times sums of at
  // most a few hundred cluster products: 24-bit multiplies are exact)
  int accC[kMaxSlots];
#pragma unroll
  for (int a = 0; a < kMaxSlots; ++a) accC[a] = 0;
  int accN[kMaxOcc] = {0, 0, 0, 0};
  int accAtt = 0, accAcc = 0;
  double accE = 0.0;
  int n_records = 0;  // slab launches (thread 0): records written

  const unsigned long long pass0 = cs->pass_base;
  // passes until the next sampled one (pass index, 0 = this launch's first)
  int to_sample = (int)((unsigned long long)(args.sample_period - 1) -
                        pass0 % (unsigned long long)args.sample_period);
  // ---- slab of this workgroup: cell planes [ka, ka + U) (mod L3) are
  // updated; the slab's top kreach planes stay frozen, so the updated sites of
  // two slabs are more than kreach planes apart and share no cluster, and
  // every site this workgroup reads is its own or frozen for the launch.  The
  // slab boundaries move by a random plane offset per launch (drawn from the
  // chain's key and pass counter, not from the state).
  int ka = 0, U = L3;
  if (n_slabs > 1) {
    // phase word 0xFFFF0000: above the class-matching words (class << 12 | round,
    // below 2^24) and below the pair-phase words (0xFFFFFFF0 - round, round < 64)
    const int k0 = (int)uniform_int((uint32_t)L3, pass0, 0xFFFF0000u, (uint32_t)chain, cp->key0, cp->key1);
    const int s0 = (int)(((long long)slab * L3) / n_slabs), s1 = (int)(((long long)(slab + 1) * L3) / n_slabs);
    ka = (k0 + s0) % L3;
    U = s1 - s0 - sc->kreach;
  }
  // class sites of this workgroup: list positions [r0, r0 + n0) and [r1, r1 + n1)
  struct Ranges {
    int r0, n0, r1, n1;
    __device__ __forceinline__ int at(int v) const { return v < n0 ? r0 + v : r1 + (v - n0); }
    __device__ __forceinline__ int n() const { return n0 + n1; }
  };
  auto ranges = [&](int c) -> Ranges {
    const int cb = sc->class_begin[c], ce = sc->class_begin[c + 1];
    if (n_slabs == 1) return Ranges{cb, ce - cb, ce, 0};
    cptr<int> kx = cst(sc->class_kidx + c * (L3 + 1));
    const int kb = ka + U;
    if (kb <= L3) return Ranges{cb + kx[ka], kx[kb] - kx[ka], ce, 0};
    return Ranges{cb + kx[ka], kx[L3] - kx[ka], cb, kx[kb - L3]};
  };
  // whether the site at padded main address a lies on an updated plane
  auto updated = [&](int a) -> bool {
    if (n_slabs == 1) return true;
    const int k = (a % S) / P12 - h;
    int d = k - ka;
    if (d < 0) d += L3;
    return d < U;
  };
  // the site at padded main address a and every periodic copy of it in the
  // halo: old -> nw (serial moves that change sites one at a time)
  auto set_all = [&](int a, int old, int nw) {
    const int b = a / S;
    const int r = a - b * S;
    const int pk = r / P12, pj = (r - pk * P12) / P1, pi = r - pk * P12 - pj * P1;
    const int m1 = (h + L1 - 1) / L1, m2 = (h + L2 - 1) / L2, m3 = (h + L3 - 1) / L3;
    for (int dk = -m3; dk <= m3; ++dk) {
      const int qk = pk + dk * L3;
      if (qk < 0 || qk >= P3) continue;
      for (int dj = -m2; dj <= m2; ++dj) {
        const int qj = pj + dj * L2;
        if (qj < 0 || qj >= P2) continue;
        for (int di = -m1; di <= m1; ++di) {
          const int qi = pi + di * L1;
          if (qi < 0 || qi >= P1) continue;
          img.set(b * S + qk * P12 + qj * P1 + qi, old, nw);
        }
      }
    }
  };
  const bool auto_rounds = args.rounds <= 0;
  const int rounds = auto_rounds ? 0 : args.rounds;
  cptr<int> bin_off = cst(sc->bin_off);
  cptr<int> gen_prog = cst(sc->gen_prog);
  const FieldCtx fctx{sc, bin_off, gen_prog, n_slots};

  // register-cube sweep: each lane keeps the cells around its 3x3x3 block
  // (the specialised semi-grand canonical binary kernel of a supercell whose
  // classes are the residues mod 3, lane t = block t)
  constexpr bool kCubePath = CUBE && BIN && KIND == 1 && !PACK && SPLIT == 1 && FIELD::kCube;
  constexpr bool use_cube = kCubePath;  // the host never runs this kernel as slabs
  [[maybe_unused]] typename FIELD::Cube cube;
  int cax[3] = {0, 0, 0}, cay[3] = {0, 0, 0}, caz[3] = {0, 0, 0};
  if constexpr (kCubePath) {
    if (use_cube) {
      const int n1 = FIELD::kCubeBlocks[0], n2 = FIELD::kCubeBlocks[1], n3 = FIELD::kCubeBlocks[2];
      const int bi = tid % n1, bj = (tid / n1) % n2, bk = tid / (n1 * n2);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        cax[d] = ((bi + d - 1 + n1) % n1) * 3 + h;
        cay[d] = (((bj + d - 1 + n2) % n2) * 3 + h) * P1;
        caz[d] = (((bk + d - 1 + n3) % n3) * 3 + h) * P12;
      }
      // cube cell (0,0,0): the block origin minus the halo, in the padded image
      FIELD::cube_load(cube, img, (bk * 3) * P12 + (bj * 3) * P1 + bi * 3);
    }
  }

  // n^2 - sum_o n_o^2 over the occupants of one independent set (block-uniform)
  auto class_mixed = [&](const Ranges &R) -> long long {
    const int n = R.n();
    int cnt[kMaxOcc] = {0, 0, 0, 0};
    for (int t = tid; t < n; t += nthr) {
      const int o = img.get(sc->class_addr[R.at(t)]);
#pragma unroll
      for (int q = 0; q < kMaxOcc; ++q) cnt[q] += (o == q) ? 1 : 0;
    }
#pragma unroll
    for (int q = 0; q < kMaxOcc; ++q) {
      const int v = wave_sum(cnt[q]);
      if (lane == 0) redI[q * n_waves + wave] = v;
    }
    __syncthreads();
    long long sq = 0;
    for (int q = 0; q < kMaxOcc; ++q) {
      long long m = 0;
      for (int w = 0; w < n_waves; ++w) m += redI[q * n_waves + w];
      sq += m * m;
    }
    __syncthreads();  // every lane has read redI before it is written again
    return (long long)n * n - sq;
  };

  for (int pass = 0; pass < args.n_passes; ++pass) {
    const unsigned long long gpass = pass0 + (unsigned long long)pass;
    long long m_total = 0;
    int n_mut = sc->class_begin[n_classes];  // mutable sites this workgroup updates
    if (n_slabs > 1) {
      n_mut = 0;
      for (int c = 0; c < n_classes; ++c) n_mut += ranges(c).n();
    }
    if (KIND != 1 && auto_rounds)
      for (int c = 0; c < n_classes; ++c) m_total += class_mixed(ranges(c));
    if constexpr (kCubePath) {
      if (use_cube) {
        // register-cube sweep (one sublattice, 27 residue classes, lane =
        // block): the same proposals, draws and updates as the class loop
        const float chem0 = cp->chem[0], chem1 = cp->chem[1];
        auto site = [&](float F, const int (&K)[kMaxSlots], int o) -> bool {
          const uint32_t u = rng.next_u32();
          const float dphi = o ? -2.0f : 2.0f;
          const float dE = dphi * F - (o ? (chem0 - chem1) : (chem1 - chem0));
          const bool acc = metropolis_accept(dE, beta, u, rng);
          accAtt += 1;
          const int sk = acc ? (o ? -2 : 2) : 0;
#pragma unroll
          for (int a = 0; a < kMaxSlots; ++a) accC[a] += __mul24(sk, K[a]);
          accN[1] += sk / 2;
          accN[0] -= sk / 2;
          accAcc += acc ? 1 : 0;
          // keep the running sums in place: without this the compiler defers
          // the 27 classes' additions to the end of the pass and keeps every
          // class's cluster sums alive until then
          asm volatile("" : "+v"(accC[0]), "+v"(accC[1]), "+v"(accC[2]), "+v"(accC[3]), "+v"(accC[4]),
                       "+v"(accC[5]), "+v"(accC[6]), "+v"(accC[7]), "+v"(accN[0]), "+v"(accN[1]), "+v"(accAtt),
                       "+v"(accAcc));
          return acc;
        };
        FIELD::cube_pass(cube, img, cax, cay, caz, site);
      }
    }
    for (int c = 0; c < n_classes && !use_cube; ++c) {
      const int b = sc->class_sublat[c];
      const Ranges R = ranges(c);
      const int nR = R.n();
      // this lane's first halo copy of the class, loaded before the sweep so
      // that its global-memory latency overlaps the sweep
      const int hb0 = sc->halo_begin[c] + tid, hb1 = n_slabs == 1 ? sc->halo_begin[c + 1] : 0;
      const int2 hc0 = hb0 < hb1 ? sc->halo[hb0] : make_int2(0, 0);
      cptr<BinSub> bs = cst(sc->bin_sub + b);
      cptr<GenSub> gs = cst(sc->gen_sub + b);
      const float chem0 = cp->chem[b * kMaxOcc + 0], chem1 = cp->chem[b * kMaxOcc + 1];
      if (KIND == 1) {
        // ---- semi-grand canonical: every site of the class proposes a change
        if constexpr (BIN && SPLIT == 2) {
          const int part = tid & 1;
          for (int v = tid >> 1; v < nR; v += (nthr >> 1)) {
            const int pa = sc->class_addr[R.at(v)];
            const int o = img.get(pa);
            int X[kMaxSlots], K[kMaxSlots];
            FIELD::bin_x(fctx, img, b, pa, o, part, X);
#pragma unroll
            for (int a = 0; a < kMaxSlots; ++a) X[a] += pair_swap(X[a]);
            const float F = FIELD::bin_finish(fctx, b, X, K);
            if (part == 0) {
              const uint32_t u = rng.next_u32();
              const float dphi = o ? -2.0f : 2.0f;
              const float dE = dphi * F - (o ? (chem0 - chem1) : (chem1 - chem0));
              const bool acc = metropolis_accept(dE, beta, u, rng);
              accAtt += 1;
              const int sk = acc ? (o ? -2 : 2) : 0;
#pragma unroll
              for (int a = 0; a < kMaxSlots; ++a) accC[a] += __mul24(sk, K[a]);
              accN[1] += sk / 2;
              accN[0] -= sk / 2;
              accAcc += acc ? 1 : 0;
              if (acc) img.set(pa, o, 1 - o);
            }
          }
        } else
        for (int v = tid; v < nR; v += nthr) {
          const int pa = sc->class_addr[R.at(v)];
          const int o = img.get(pa);
          const uint32_t u = rng.next_u32();
          if (BIN) {
            int K[kMaxSlots];
            const float F = FIELD::bin(fctx, img, b, pa, o, K);
            const float dphi = o ? -2.0f : 2.0f;
            const float dE = dphi * F - (o ? (chem0 - chem1) : (chem1 - chem0));
            const bool acc = metropolis_accept(dE, beta, u, rng);
            accAtt += 1;
            const int sk = acc ? (o ? -2 : 2) : 0;
#pragma unroll
            for (int a = 0; a < kMaxSlots; ++a) accC[a] += __mul24(sk, K[a]);
            accN[1] += sk / 2;
            accN[0] -= sk / 2;
            accAcc += acc ? 1 : 0;
            if (acc) img.set(pa, o, 1 - o);
          } else {
            float G[kMaxFunc];
            FIELD::gen(fctx, img, b, pa, tabL, G);
            const int no = sys->n_occ[b];
            const float u2 = rng.next();
            int r = (int)(u2 * (float)(no - 1));
            if (r > no - 2) r = no - 2;
            const int nw = r >= o ? r + 1 : r;
            const float dEf = general_dE(sys, b, o, nw, G);
            const float dE = dEf - (cp->chem[b * kMaxOcc + nw] - cp->chem[b * kMaxOcc + o]);
            const bool acc = metropolis_accept(dE, beta, u, rng);
            accAtt += 1;
            if (acc) {
              img.set(pa, o, nw);
              accAcc += 1;
              accE += (double)dEf;
#pragma unroll
              for (int q = 0; q < kMaxOcc; ++q) accN[q] += (q == nw) - (q == o);
            }
          }
        }
        __syncthreads();
      } else {
        // ---- canonical: fields of all class sites, then random matchings
        const int n = nR;
        const int n_rounds = auto_rounds ? auto_canonical_rounds(n, class_mixed(R), n_mut, m_total) : rounds;
        if (n_rounds > 0) {
        if constexpr (BIN && SPLIT == 2) {
          const int part = tid & 1;
          for (int t = tid >> 1; t < n; t += (nthr >> 1)) {
            const int pa = sc->class_addr[R.at(t)];
            int X[kMaxSlots], K[kMaxSlots];
            FIELD::bin_x(fctx, img, b, pa, img.get(pa), part, X);
#pragma unroll
            for (int a = 0; a < kMaxSlots; ++a) X[a] += pair_swap(X[a]);
            const float F = FIELD::bin_finish(fctx, b, X, K);
            if (part == 0) {
              cF[t] = F;
#pragma unroll
              for (int a = 0; a < kMaxSlots; ++a) cK[a * max_class + t] = (int8_t)K[a];
            }
          }
        } else
        for (int t = tid; t < n; t += nthr) {
          const int pa = sc->class_addr[R.at(t)];
          if (BIN) {
            int K[kMaxSlots];
            cF[t] = FIELD::bin(fctx, img, b, pa, img.get(pa), K);
#pragma unroll
            for (int a = 0; a < kMaxSlots; ++a) cK[a * max_class + t] = (int8_t)K[a];
          } else {
            float G[kMaxFunc];
            FIELD::gen(fctx, img, b, pa, tabL, G);
            cF[t] = G[0];
            if (nfs > 1) cF[max_class + t] = G[1];
            if (nfs > 2) cF[2 * max_class + t] = G[2];
          }
        }
        __syncthreads();
        for (int rd = 0; rd < n_rounds; ++rd) {
          const uint32_t r = uniform_int((uint32_t)n, gpass, ((uint32_t)c << 12) | (uint32_t)rd, rchain,
                                         cp->key0, cp->key1);
          for (int p = tid; p < n; p += nthr) {
            int q = (int)r - p;
            if (q < 0) q += n;
            if (p >= q) continue;  // q == p: unpaired; p > q: the partner leads
            const int pa = sc->class_addr[R.at(p)], qa = sc->class_addr[R.at(q)];
            const int op = img.get(pa), oq = img.get(qa);
            if (op == oq) continue;
            const uint32_t u = rng.next_u32();
            float dE;
            if (BIN) {
              const float dp = op ? -2.0f : 2.0f;
              dE = dp * cF[p] - dp * cF[q];
            } else {
              float Gp[kMaxFunc] = {cF[p], nfs > 1 ? cF[max_class + p] : 0.f, nfs > 2 ? cF[2 * max_class + p] : 0.f};
              float Gq[kMaxFunc] = {cF[q], nfs > 1 ? cF[max_class + q] : 0.f, nfs > 2 ? cF[2 * max_class + q] : 0.f};
              dE = general_dE(sys, b, op, oq, Gp) + general_dE(sys, b, oq, op, Gq);
            }
            const bool acc = metropolis_accept(dE, beta, u, rng);
            accAtt += 1;
            if (acc) {
              img.set(pa, op, oq);
              img.set(qa, oq, op);
              accAcc += 1;
              if (BIN) {
                const int sp = op ? -2 : 2;
#pragma unroll
                for (int a = 0; a < kMaxSlots; ++a)
                  accC[a] += __mul24(sp, (int)cK[a * max_class + p] - (int)cK[a * max_class + q]);
              } else {
                accE += (double)dE;
              }
            }
          }
          __syncthreads();
        }
        }  // n_rounds > 0
      }
      // ---- halo copies of this class (slabs: of this workgroup's sites only;
      // the other sites are unchanged in this image)
      if (n_slabs == 1) {
        if (hb0 < hb1) img.copy(hc0.y, hc0.x);
        for (int t = hb0 + nthr; t < hb1; t += nthr) {
          const int2 hc = sc->halo[t];
          img.copy(hc.y, hc.x);
        }
      } else {
        const int a0 = sc->halo_site[R.r0], a1 = sc->halo_site[R.r0 + R.n0];
        const int b0 = sc->halo_site[R.r1], b1 = sc->halo_site[R.r1 + R.n1];
        const int na = a1 - a0, nh = na + (b1 - b0);
        for (int t = tid; t < nh; t += nthr) {
          const int2 hc = sc->halo[t < na ? a0 + t : b0 + (t - na)];
          img.copy(hc.y, hc.x);
        }
      }
      __syncthreads();
    }

    // ---- canonical cross-class mixing: randomly chosen pair phases (one per
    // pass, or with automatic rounds about n_mutable / 4 pair attempts)
    const int n_pair_rounds =
        (KIND != 1 && auto_rounds && sc->n_pair_phases > 0)
            ? auto_pair_rounds(sc->class_begin[n_classes], sc->pair_begin[sc->n_pair_phases], sc->n_pair_phases)
            : 1;
    for (int pr = 0; KIND != 1 && sc->n_pair_phases > 0 && pr < n_pair_rounds; ++pr) {
      const uint32_t ph = uniform_int((uint32_t)sc->n_pair_phases, gpass, 0xFFFFFFF0u - (uint32_t)pr,
                                      (uint32_t)chain, cp->key0, cp->key1);
      const int pb = sc->pair_begin[ph], pe = sc->pair_begin[ph + 1];
      for (int t = pb + tid; t < pe; t += nthr) {
        const int2 pq = sc->pair_addr[t];
        if (!updated(pq.x) || !updated(pq.y)) continue;
        const int bpq = sc->pair_sublat[t];
        const int bp = bpq & 0xFF, bq = bpq >> 8;
        const int op = img.get(pq.x), oq = img.get(pq.y);
        if (op == oq) continue;
        const uint32_t u = rng.next_u32();
        float dE;
        int Kp[kMaxSlots], Kq[kMaxSlots];
        if (BIN) {
          const float Fp = FIELD::bin(fctx, img, bp, pq.x, op, Kp);
          const float Fq = FIELD::bin(fctx, img, bq, pq.y, oq, Kq);
          const float dp = op ? -2.0f : 2.0f;
          dE = dp * Fp - dp * Fq;
        } else {
          float Gp[kMaxFunc], Gq[kMaxFunc];
          FIELD::gen(fctx, img, bp, pq.x, tabL, Gp);
          FIELD::gen(fctx, img, bq, pq.y, tabL, Gq);
          dE = general_dE(sys, bp, op, oq, Gp) + general_dE(sys, bq, oq, op, Gq);
        }
        const bool acc = metropolis_accept(dE, beta, u, rng);
        accAtt += 1;
        if (acc) {
          img.set(pq.x, op, oq);
          img.set(pq.y, oq, op);
          accAcc += 1;
          if (BIN) {
            const int sp = op ? -2 : 2;
#pragma unroll
            for (int a = 0; a < kMaxSlots; ++a) accC[a] += __mul24(sp, Kp[a] - Kq[a]);
          } else {
            accE += (double)dE;
          }
        }
      }
      __syncthreads();
      for (int t = sc->pair_halo_begin[ph] + tid; t < sc->pair_halo_begin[ph + 1]; t += nthr) {
        const int2 hc = sc->pair_halo[t];
        img.copy(hc.y, hc.x);
      }
      __syncthreads();
    }

    // ---- canonical global swaps: uniformly random site pairs anywhere in the
    // supercell (the two sites may share clusters), evaluated serially by one
    // lane as two sequential flips on the image (every periodic copy of the
    // first site is updated before the second site's field is taken).  These
    // are the reference's unrestricted pair moves; they make every supercell
    // ergodic, however small, for the canonical ensemble.
    if (KIND != 1 && args.global_swaps > 0 && n_slabs == 1) {
      if (tid == 0) {
        const int n_var = sc->class_begin[n_classes];
        for (int g = 0; g < args.global_swaps; ++g) {
          const float u1 = rng.next(), u2 = rng.next();
          const uint32_t u3 = rng.next_u32();
          const int is = min((int)(u1 * (float)n_var), n_var - 1);
          const int it = min((int)(u2 * (float)n_var), n_var - 1);
          const int sa = sc->class_addr[is], ta = sc->class_addr[it];
          if (sa == ta) continue;
          const int bsa = sa / S, bta = ta / S;
          if (sys->n_occ[bsa] != sys->n_occ[bta] || sys->asym[bsa] != sys->asym[bta]) continue;
          const int os = img.get(sa), ot = img.get(ta);
          if (os == ot) continue;
          float dE1, dE2;
          int Ks[kMaxSlots], Kt[kMaxSlots];
          if (BIN) {
            dE1 = (os ? -2.0f : 2.0f) * FIELD::bin(fctx, img, bsa, sa, os, Ks);
          } else {
            float G[kMaxFunc];
            FIELD::gen(fctx, img, bsa, sa, tabL, G);
            dE1 = general_dE(sys, bsa, os, ot, G);
          }
          set_all(sa, os, ot);
          if (BIN) {
            dE2 = (ot ? -2.0f : 2.0f) * FIELD::bin(fctx, img, bta, ta, ot, Kt);
          } else {
            float G[kMaxFunc];
            FIELD::gen(fctx, img, bta, ta, tabL, G);
            dE2 = general_dE(sys, bta, ot, os, G);
          }
          const float dE = dE1 + dE2;
          accAtt += 1;
          if (metropolis_accept(dE, beta, u3, rng)) {
            set_all(ta, ot, os);
            accAcc += 1;
            if (BIN) {
              const int ss = os ? -2 : 2, st = ot ? -2 : 2;
#pragma unroll
              for (int a = 0; a < kMaxSlots; ++a) accC[a] += ss * Ks[a] + st * Kt[a];
            } else {
              accE += (double)dE;
            }
          } else {
            set_all(sa, ot, os);
          }
        }
      }
      __syncthreads();
    }

    // ---- semi-grand canonical multiswaps (SemiGrandCanonicalCalculator.cc:
    // 104-115 with the system's multiswaps, System_json_io.cc:955-1006).  An
    // attempt draws a type uniformly and, for each of its parts (unit u,
    // occupant from -> to, count c), c distinct sites uniformly among the
    // sites of u holding `from` and not yet picked: by rejection from all
    // sites of u (up to 32 draws per site), else by a workgroup-wide search
    // for a uniformly drawn k-th such site (both uniform).  Thread 0 then
    // changes the sites one after another (each site's field sees the earlier
    // changes, every periodic copy written) and accepts with
    // min(1, exp(-beta dOmega) q_rev / q_fwd), where q of a move is
    // (1 / n_types) prod_parts 1 / C(r, c), r the part's pool size (minus the
    // sites earlier parts took from the same pool) in the state the move
    // starts from and the reverse type's parts for q_rev.  A part whose pool
    // is too small makes the attempt a rejection.
    if (KIND == 2 && mswaps) {
      const int n_ty = args.n_mswaps;
      auto unit_sites = [&](int u) -> int {
        int m = 0;
        for (int b = 0; b < sys->n_sublat; ++b) m += (sys->asym[b] == u && sys->n_occ[b] >= 2) ? 1 : 0;
        return m * n_cells;
      };
      auto unit_addr = [&](int u, int j) -> int {
        int q = j / n_cells;
        const int cell = j - q * n_cells;
        int bb = 0;
        for (int b = 0; b < sys->n_sublat; ++b)
          if (sys->asym[b] == u && sys->n_occ[b] >= 2) {
            if (q == 0) {
              bb = b;
              break;
            }
            --q;
          }
        const int i = cell % L1, jj = (cell / L1) % L2, k = cell / (L1 * L2);
        return bb * S + (k + h) * P12 + (jj + h) * P1 + (i + h);
      };
      auto binom = [](int r, int c) -> double {
        double v = 1.0;
        for (int t = 0; t < c; ++t) v *= (double)(r - t) / (double)(t + 1);
        return v;
      };
      const int per_pass = args.mswap_per_pass > 0 ? args.mswap_per_pass : sc->class_begin[n_classes];
      // thread 0's attempt: type, parts cursor, picked sites and target occupants
      int ty = 0, part = 0, in_part = 0, n_pick = 0;
      int pick_a[kMaxMswapSites], pick_to[kMaxMswapSites];
      double fwd = 1.0;
      bool failed = false;  // a search found no site: the pool counts disagree with the image
      for (int g = 0; g < per_pass; ++g) {
        if (tid == 0) {
          // exactly uniform type: rejection of the top of the 32-bit range
          const unsigned long long lim = 4294967296ull - (4294967296ull % (unsigned long long)n_ty);
          uint32_t x = rng.next_u32();
          for (int t = 0; t < 64 && (unsigned long long)x >= lim; ++t) x = rng.next_u32();
          ty = (int)(x % (uint32_t)n_ty);
          part = args.mswap_begin[ty];
          in_part = 0;
          n_pick = 0;
          fwd = 1.0;
          failed = false;
          msw[0] = 0;
        }
        for (;;) {
          if (tid == 0) {
            int state = failed ? 3 : 2;  // 2: all sites picked, 3: impossible, 1: search requested
            while (!failed && part < args.mswap_begin[ty + 1]) {
              const int4 pt = args.mswap_parts[part];  // (unit, from, to, count)
              int taken = 0;
              for (int q = 0; q < n_pick; ++q)
                taken += (sys->asym[pick_a[q] / S] == pt.x && img.get(pick_a[q]) == pt.y) ? 1 : 0;
              const int r = poolN[pt.x * kMaxOcc + pt.y] - taken;
              if (in_part == 0) {
                if (r < pt.w) {
                  state = 3;
                  break;
                }
                fwd *= binom(r, pt.w);
              }
              if (in_part == pt.w) {
                ++part;
                in_part = 0;
                continue;
              }
              const int nu = unit_sites(pt.x);
              int a = -1;
              for (int t = 0; t < 32 && a < 0; ++t) {
                const int j = (int)(((unsigned long long)rng.next_u32() * (unsigned long long)nu) >> 32);
                const int aj = unit_addr(pt.x, j);
                bool fresh = img.get(aj) == pt.y;
                for (int q = 0; q < n_pick; ++q) fresh = fresh && pick_a[q] != aj;
                if (fresh) a = aj;
              }
              if (a < 0) {
                // uniform k among the r remaining candidates: exact integer draw
                const unsigned long long lim = 4294967296ull - (4294967296ull % (unsigned long long)r);
                uint32_t x = rng.next_u32();
                for (int t = 0; t < 64 && (unsigned long long)x >= lim; ++t) x = rng.next_u32();
                msw[1] = pt.x;
                msw[2] = pt.y;
                msw[3] = (int)(x % (uint32_t)r);
                msw[4] = -1;
                state = 1;
                break;
              }
              pick_a[n_pick] = a;
              pick_to[n_pick] = pt.z;
              ++n_pick;
              ++in_part;
            }
            msw[0] = state;
            msw[5] = n_pick;
            for (int q = 0; q < n_pick; ++q) msw[8 + q] = pick_a[q];
          }
          __syncthreads();
          const int state = msw[0];
          if (state != 1) break;
          // workgroup-wide search: the k-th site (in unit order) of unit u
          // holding `from` that is not picked yet
          {
            const int u = msw[1], of = msw[2], kk = msw[3], np = msw[5];
            const int nu = unit_sites(u);
            const int chunk = (nu + nthr - 1) / nthr;
            const int j0 = min(tid * chunk, nu), j1 = min(j0 + chunk, nu);
            auto cand = [&](int j) -> int {
              const int aj = unit_addr(u, j);
              if (img.get(aj) != of) return -1;
              for (int q = 0; q < np; ++q)
                if (msw[8 + q] == aj) return -1;
              return aj;
            };
            int cnt = 0;
            for (int j = j0; j < j1; ++j) cnt += cand(j) >= 0 ? 1 : 0;
            int incl = cnt;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
              const int y = __shfl_up(incl, o, 64);
              if (lane >= o) incl += y;
            }
            if (lane == 63) redI[wave] = incl;
            __syncthreads();
            int before = incl - cnt;
            for (int w = 0; w < wave; ++w) before += redI[w];
            if (kk >= before && kk < before + cnt) {
              int left = kk - before;
              for (int j = j0; j < j1; ++j) {
                const int aj = cand(j);
                if (aj >= 0) {
                  if (left == 0) {
                    msw[4] = aj;
                    break;
                  }
                  --left;
                }
              }
            }
            __syncthreads();
          }
          if (tid == 0) {
            if (msw[4] < 0) {
              failed = true;
              cs->flags |= kChainFlagMswapCounts;
            } else {
              const int4 pt = args.mswap_parts[part];
              pick_a[n_pick] = msw[4];
              pick_to[n_pick] = pt.z;
              ++n_pick;
              ++in_part;
            }
          }
        }
        if (tid == 0) {
          accAtt += 1;
          if (msw[0] == 2) {
            // sequential changes and the semi-grand energy change
            float dEf = 0.0f, dchem = 0.0f;
            int tC[kMaxSlots];
#pragma unroll
            for (int a = 0; a < kMaxSlots; ++a) tC[a] = 0;
            int pick_from[kMaxMswapSites];
            for (int q = 0; q < n_pick; ++q) {
              const int a = pick_a[q], b = a / S, o = img.get(a), nw = pick_to[q];
              pick_from[q] = o;
              if (BIN) {
                int K[kMaxSlots];
                const float F = FIELD::bin(fctx, img, b, a, o, K);
                dEf += (o ? -2.0f : 2.0f) * F;
                const int sk = o ? -2 : 2;
#pragma unroll
                for (int t = 0; t < kMaxSlots; ++t) tC[t] += sk * K[t];
              } else {
                float G[kMaxFunc];
                FIELD::gen(fctx, img, b, a, tabL, G);
                dEf += general_dE(sys, b, o, nw, G);
              }
              dchem += cp->chem[b * kMaxOcc + nw] - cp->chem[b * kMaxOcc + o];
              set_all(a, o, nw);
            }
            // reverse move's pool sizes in the new state
            double rev = 1.0;
            bool ok = true;
            const int rt = args.mswap_rev[ty];
            for (int p2 = args.mswap_begin[rt]; p2 < args.mswap_begin[rt + 1]; ++p2) {
              const int4 pt = args.mswap_parts[p2];
              int ny = poolN[pt.x * kMaxOcc + pt.y];
              for (int p1 = args.mswap_begin[ty]; p1 < args.mswap_begin[ty + 1]; ++p1) {
                const int4 f = args.mswap_parts[p1];
                if (f.x == pt.x) ny += (f.z == pt.y ? f.w : 0) - (f.y == pt.y ? f.w : 0);
              }
              for (int p3 = args.mswap_begin[rt]; p3 < p2; ++p3) {
                const int4 e = args.mswap_parts[p3];
                if (e.x == pt.x && e.y == pt.y) ny -= e.w;
              }
              if (ny < pt.w) ok = false;
              else rev *= binom(ny, pt.w);
            }
            const double pacc = ok ? (double)__expf(-beta * (dEf - dchem)) * fwd / rev : 0.0;
            const bool acc = accept_prob((float)fmin(pacc, 2.0), rng.next_u32(), [&]() { return rng.next_u32(); });
            if (acc) {
              accAcc += 1;
              if (BIN) {
#pragma unroll
                for (int t = 0; t < kMaxSlots; ++t) accC[t] += tC[t];
              } else {
                accE += (double)dEf;
              }
              for (int q = 0; q < n_pick; ++q) {
                const int u = sys->asym[pick_a[q] / S];
                poolN[u * kMaxOcc + pick_from[q]] -= 1;
                poolN[u * kMaxOcc + pick_to[q]] += 1;
#pragma unroll
                for (int o = 0; o < kMaxOcc; ++o) accN[o] += (o == pick_to[q]) - (o == pick_from[q]);
              }
            } else {
              for (int q = n_pick - 1; q >= 0; --q) set_all(pick_a[q], pick_to[q], pick_from[q]);
            }
          }
        }
        __syncthreads();
      }
    }

    // ---- reduction of lane accumulators and sampling
    // a sample is due after pass gpass when (gpass + 1) % sample_period == 0:
    // a countdown set from one 64-bit remainder at the launch start
    const bool due = to_sample == 0;
    to_sample = due ? args.sample_period - 1 : to_sample - 1;
    const bool sample_due = n_slabs == 1 && due;
    const bool record_due = n_slabs > 1 && (due || pass == args.n_passes - 1);
    const bool flush = sample_due || record_due || ((pass & 31) == 31) || (pass == args.n_passes - 1);
    if (flush) {
      // every wave sum first (independent DPP chains the compiler can
      // interleave), then lane q stores this wave's total of slot q
      int vals[kMaxSlots + kMaxOcc + 2];
      int s = 0;
      if (BIN) {
#pragma unroll
        for (int a = 0; a < kMaxSlots; ++a) {
          vals[s + a] = wave_sum(accC[a]);
          accC[a] = 0;
        }
        s += kMaxSlots;
      }
#pragma unroll
      for (int o = 0; o < kMaxOcc; ++o) {
        vals[s + o] = wave_sum(accN[o]);
        accN[o] = 0;
      }
      s += kMaxOcc;
      vals[s] = wave_sum(accAtt);
      vals[s + 1] = wave_sum(accAcc);
      accAtt = accAcc = 0;
      {
        int mine = 0;
#pragma unroll
        for (int q = 0; q < kMaxSlots + kMaxOcc + 2; ++q)
          if (q < n_islot) mine = lane == q ? vals[q] : mine;
        if (lane < n_islot) redI[lane * n_waves + wave] = mine;
      }
      if (!BIN) {
        const double v = wave_sum(accE);
        if (lane == 0) redD[wave] = v;
        accE = 0.0;
      }
      __syncthreads();
      if (wave == 0) {
        // lane q adds up slot q over the waves; thread 0 reads the totals
        int part = 0;
        if (lane < n_islot)
          for (int w = 0; w < n_waves; ++w) part += redI[lane * n_waves + w];
        int sumI[kMaxSlots + kMaxOcc + 2];
#pragma unroll
        for (int q = 0; q < kMaxSlots + kMaxOcc + 2; ++q) sumI[q] = q < n_islot ? __builtin_amdgcn_readlane(part, q) : 0;
        const int s2 = BIN ? kMaxSlots : 0;
        if (n_slabs == 1) {
          // lane a adds the slots of orbit a to its exact count, lane o the
          // change of occupant o's count (one LDS update per lane)
          if (BIN && lane < args.n_orbits) {
            long long d = 0;
            for (int j = 0; j < n_slots; ++j) d += sys->slot_orbit[j] == lane ? (long long)sumI[j] : 0ll;
            cl->C[lane] += d;
          }
          if (lane < kMaxOcc) {
            int dn = sumI[s2];
#pragma unroll
            for (int o = 1; o < kMaxOcc; ++o) dn = lane == o ? sumI[s2 + o] : dn;
            cl->counts[lane] += dn;
          }
        }
      if (lane == 0) {
        double E;
        if (n_slabs > 1) {
          // running totals of this workgroup since the launch started; the
          // chain's totals and samples are formed by slab_post_kernel
          for (int q = 0; q < n_islot; ++q) slabI[q] += sumI[q];
          if (!BIN)
            for (int w = 0; w < n_waves; ++w) *slabE += redD[w];
          if (record_due) {
            SlabRecord *r = args.slab_rec + ((long long)chain * kSlabMax + slab) * kSlabRecords + n_records;
            r->dE = *slabE;
            const int base = BIN ? 0 : kMaxSlots;  // dI layout: slots, counts, attempts, accepts
            for (int q = 0; q < base; ++q) r->dI[q] = 0;
            for (int q = 0; q < n_islot; ++q) r->dI[base + q] = slabI[q];
            ++n_records;
          }
        } else {
        if (BIN) {
          E = 0.0;
          for (int a = 0; a < args.n_orbits; ++a) E += sys->bin_w64[a] * (double)cl->C[a];
        } else {
          double dsum = 0.0;
          for (int w = 0; w < n_waves; ++w) dsum += redD[w];
          E = cl->E + dsum;
        }
        cl->E = E;
        cl->n_attempt += sumI[s2 + kMaxOcc];
        cl->n_accept += sumI[s2 + kMaxOcc + 1];
        if (sample_due) {
          const int idx = cl->sample_count;
          if (idx < args.sample_cap) {
            const long long o = (long long)chain * args.sample_cap + idx;
            args.sampE[o] = E;
            for (int q = 0; q < kMaxOcc; ++q) args.sampN[o * kMaxOcc + q] = (int)cl->counts[q];
            cl->sample_count = idx + 1;
          }
          const double n1 = (double)cl->counts[1];
          cl->sumE += E;
          cl->sumE2 += E * E;
          cl->sumN1 += n1;
          cl->sumN12 += n1 * n1;
          cl->sumEN1 += E * n1;
          cl->n_samples_acc += 1;
        }
        }  // n_slabs == 1
      }  // lane 0
      }  // wave 0
      __syncthreads();
    }
  }

  // ---- write back occupation and RNG state
  {
    int8_t *occW = args.occ + cp->occ_offset;
    if (n_slabs == 1) {
      PadWalk w(tid, nthr, L1, L2, L3);
      for (int l = tid; l < n_sites; l += nthr, w.step())
        occW[l] = (int8_t)img.get(w.b * S + (w.pk + h) * P12 + (w.pj + h) * P1 + (w.pi + h));
    } else {
      // the sites of the updated planes only (one or two plane ranges per sublattice)
      const int plane = L1 * L2, n_sub = n_cells > 0 ? n_sites / n_cells : 0;
      const int kb = ka + U;
      const int lo0 = ka, hi0 = kb <= L3 ? kb : L3, hi1 = kb <= L3 ? 0 : kb - L3;
      const int n0 = (hi0 - lo0) * plane, n1 = hi1 * plane, per = n0 + n1;
      for (int x = tid; x < n_sub * per; x += nthr) {
        const int b = x / per;
        int r = x - b * per;
        const int cc = r < n0 ? lo0 * plane + r : r - n0;
        const int i = cc % L1, j = (cc / L1) % L2, k = cc / plane;
        occW[b * n_cells + cc] = (int8_t)img.get(b * S + (k + h) * P12 + (j + h) * P1 + (i + h));
      }
    }
  }
  // all lanes advance by the same bound so the next launch starts fresh
  atomicMax(maxDraw, rng.draw);
  __syncthreads();
  if (tid == 0) {
    if (n_slabs > 1) {
      // the chain's counters advance in slab_post_kernel, after every slab
      // of the launch has read them
      atomicMax(&cs->draw_hi, *maxDraw);
    } else {
      cs->draw_base = *maxDraw + 1ull;
      cs->pass_base = pass0 + (unsigned long long)args.n_passes;
      cs->E = cl->E;
      cs->sumE = cl->sumE;
      cs->sumE2 = cl->sumE2;
      cs->sumN1 = cl->sumN1;
      cs->sumN12 = cl->sumN12;
      cs->sumEN1 = cl->sumEN1;
      for (int o = 0; o < kMaxOcc; ++o) cs->counts[o] = cl->counts[o];
      cs->n_attempt = cl->n_attempt;
      cs->n_accept = cl->n_accept;
      cs->n_samples_acc = cl->n_samples_acc;
      if (args.sample_count) args.sample_count[chain] = cl->sample_count;
      if (BIN) {
        long long *C = args.binC + (long long)chain * args.n_orbits;
        for (int a = 0; a < args.n_orbits; ++a) C[a] = cl->C[a];
      }
    }
  }
}

}  // namespace clexmc
