// Learned proposal for the gfx950 Metropolis path.
//
// Network: a local-environment MLP with n_env (<= 48) inputs, 64 ReLU hidden
// units and one output logit.  Input k of site s is +1 if environment site k
// holds the same occupant as s and -1 otherwise (0 for padding), so the
// input of the flipped site is the negated vector and its hidden
// pre-activation is 2 b1 - h: one GEMM gives the logits of both occupations.
// Inference runs on MFMA (v_mfma_f32_16x16x16_bf16): per wavefront a
// [64 sites x 48] x [48 x 64] product with BF16 inputs/weights and FP32
// accumulation.  Training runs FP32 forward/backward kernels on samples drawn
// from the chains' current (plain-sampler) configurations.
//
// The network output z is a score of the site's flip in energy units; a
// chain at inverse temperature beta uses the logit t*z with
// t = beta * kMlpEnergyUnit, so the trained weights are shared by all chain
// temperatures.  Training labels are the heat-bath (Barker) flip
// probabilities sigmoid(-beta e_s) of sites of plain-sampler configurations.
//
// Proposal (exact Metropolis-Hastings): a class of an independent-set
// partition is split into groups of 16 sites (one per lane, a DPP row of a
// wavefront: four groups per wavefront; the grouping is a random rotation
// chosen independently of the occupation).  Within a group the rounds make
// sequential proposals, the four groups of a wavefront side by side:
// canonical, an A site a and a B site b with probabilities w_a / Z_A and
// w_b / Z_B; semi-grand canonical, one site s with probability w_s / Z.
// Weights (WEIGHTS template argument):
//   0 uniform  w = 1;
//   1 learned  w = eps + sigmoid(t z(env));
//   2 informed w = eps + sigmoid(-beta e_s), the exact Barker weight of the
//     site's flip energy (no network: the ablation of the learned weights).
// Sites of one class share no cluster, so dE = e_a + e_b is exact and, after
// an accepted move, each moved site just swaps its (w, w_flipped) and negates
// its flip energy; Z_A, Z_B are updated in O(1) and the reverse-proposal
// probability is exact.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clexmc_device.h"
#include "clexmc_kernels.h"

namespace clexmc {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_of_sign(int v) {
  return v > 0 ? (short)0x3F80 : (v < 0 ? (short)0xBF80 : (short)0);
}

struct MlpFrags {
  short4_t B[3][4];  // B operand: lane l holds W1[16 tn + l%16][16 tk + 4 (l/16) + 0..3]
  float b1[4], W2[4];
  float b2, eps;
  int n_env;
};

__device__ __forceinline__ void load_frags(cptr<MlpDev> m, int lane, MlpFrags &F) {
  const int g = lane >> 4, c = lane & 15;
#pragma unroll
  for (int tk = 0; tk < 3; ++tk)
#pragma unroll
    for (int tn = 0; tn < 4; ++tn) {
      const int row = 16 * tn + c;
      const int col = 16 * tk + 4 * g;
      short4_t v;
      v[0] = (short)m->W1b[row * kEnvMax + col + 0];
      v[1] = (short)m->W1b[row * kEnvMax + col + 1];
      v[2] = (short)m->W1b[row * kEnvMax + col + 2];
      v[3] = (short)m->W1b[row * kEnvMax + col + 3];
      F.B[tk][tn] = v;
    }
#pragma unroll
  for (int tn = 0; tn < 4; ++tn) {
    F.b1[tn] = m->b1[16 * tn + c];
    F.W2[tn] = m->W2[16 * tn + c];
  }
  F.b2 = m->b2;
  F.eps = m->eps;
  F.n_env = m->n_env;
}

// Network outputs of the 64 sites of a wavefront for their current
// occupation (outC) and the flipped one (outN).  fm_lo/fm_hi hold each lane's
// own site's environment as bits (bit k set: environment site k holds the
// same occupant); inputs k >= n_env are zero.  Results are written to LDS
// scratch by lanes with l%16 == 0.
__device__ __forceinline__ void mlp_wave(const MlpFrags &F, uint32_t fm_lo, uint32_t fm_hi, int lane, float *outC,
                                         float *outN) {
  const int g = lane >> 4, c = lane & 15;
#pragma unroll
  for (int tm = 0; tm < 4; ++tm) {
    const int s = 16 * tm + c;
    const uint32_t lo = (uint32_t)__shfl((int)fm_lo, s, 64);
    const uint32_t hi = (uint32_t)__shfl((int)fm_hi, s, 64);
    short4_t A[3];
#pragma unroll
    for (int tk = 0; tk < 3; ++tk) {
      short4_t v;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k = 16 * tk + 4 * g + q;
        const uint32_t bit = k < 32 ? (lo >> k) & 1u : (hi >> (k - 32)) & 1u;
        v[q] = k < F.n_env ? (bit ? (short)0x3F80 : (short)0xBF80) : (short)0;
      }
      A[tk] = v;
    }
    float pc[4] = {0.f, 0.f, 0.f, 0.f}, pn[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int tn = 0; tn < 4; ++tn) {
      float4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int tk = 0; tk < 3; ++tk) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[tk], F.B[tk][tn], acc, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float hc = acc[r] + F.b1[tn];
        const float hn = F.b1[tn] - acc[r];
        pc[r] += F.W2[tn] * fmaxf(hc, 0.f);
        pn[r] += F.W2[tn] * fmaxf(hn, 0.f);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
      for (int o = 8; o > 0; o >>= 1) {
        pc[r] += __shfl_xor(pc[r], o, 64);
        pn[r] += __shfl_xor(pn[r], o, 64);
      }
    }
    if (c == 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        outC[16 * tm + 4 * g + r] = pc[r] + F.b2;
        outN[16 * tm + 4 * g + r] = pn[r] + F.b2;
      }
    }
  }
}

__device__ __forceinline__ float sigmoidf_(float x) { return 1.0f / (1.0f + __expf(-x)); }

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Inclusive prefix sum over the 64 lanes of a wavefront with DPP: row shifts
// by 1, 2, 4, 8 (rows of 16 lanes; lanes without a source add 0), then the
// row-15 and row-31 broadcasts that carry row totals into the next rows.
__device__ __forceinline__ float wave_scan_incl(float x) {
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x111, 0xF, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x112, 0xF, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x114, 0xF, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x118, 0xF, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x142, 0xA, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x143, 0xC, 0xF, false));
  return x;
}

__device__ __forceinline__ float lane_f(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// Weighted choice of a lane: v >= 0 per lane, scan = wave_scan_incl(v),
// t = u * total.  Returns the first lane whose inclusive sum exceeds t among
// lanes with v > 0 (the last such lane if rounding leaves none), -1 if all v
// are 0.  The result is wavefront-uniform.
__device__ __forceinline__ int wave_pick(float v, float scan, float t) {
  unsigned long long m = __ballot(scan > t && v > 0.f);
  if (m == 0ull) {
    m = __ballot(v > 0.f);
    if (m == 0ull) return -1;
    return 63 - __clzll((long long)m);
  }
  return __ffsll((long long)m) - 1;
}

// Inclusive prefix sum within each row of 16 lanes (DPP row shifts by 1, 2,
// 4, 8; lanes without a source add 0).
__device__ __forceinline__ float row_scan_incl(float x) {
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x111, 0xF, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x112, 0xF, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x114, 0xF, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x118, 0xF, 0xF, false));
  return x;
}

// The row's total (the inclusive sum of its last lane) in every lane of the
// row; rb is the row's first lane.
__device__ __forceinline__ float row_total(float scan, int rb) {
  const float z0 = lane_f(scan, 15), z1 = lane_f(scan, 31), z2 = lane_f(scan, 47), z3 = lane_f(scan, 63);
  return rb == 0 ? z0 : (rb == 16 ? z1 : (rb == 32 ? z2 : z3));
}

// Weighted choice within a row (lanes rb .. rb + 15): v >= 0 per lane, scan
// = row_scan_incl(v), t = u * the row's total.  The first lane of the row
// whose inclusive sum exceeds t among lanes with v > 0 (the last such lane
// if rounding leaves none), -1 if all v of the row are 0; the same in every
// lane of the row.  Every lane must take part (two ballots).
__device__ __forceinline__ int row_pick(float v, float scan, float t, int rb) {
  const unsigned long long m = __ballot(scan > t && v > 0.f);
  const unsigned long long m2 = __ballot(v > 0.f);
  const uint32_t mr = (uint32_t)(m >> rb) & 0xFFFFu, mr2 = (uint32_t)(m2 >> rb) & 0xFFFFu;
  if (mr) return rb + __ffs((int)mr) - 1;
  return mr2 ? rb + 31 - __clz((int)mr2) : -1;
}

// binary field and exact integer cluster-product counts (same as the plain kernel)
__device__ __forceinline__ float binary_field_l(const uint8_t *img, int pa, int o, cptr<BinSub> bs, cptr<int> off,
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
          for (int e = 0; e < len; e += 4) {
            const int q = q0 + e;
            X += (img[pa + off[q]] + img[pa + off[q + 1]]) + (img[pa + off[q + 2]] + img[pa + off[q + 3]]);
          }
        } else if (ko == 2) {
          for (int e = 0; e < len; e += 4) {
            const int q = q0 + 2 * e;
            X += ((img[pa + off[q]] ^ img[pa + off[q + 1]]) + (img[pa + off[q + 2]] ^ img[pa + off[q + 3]])) +
                 ((img[pa + off[q + 4]] ^ img[pa + off[q + 5]]) + (img[pa + off[q + 6]] ^ img[pa + off[q + 7]]));
          }
        } else if (ko >= 3) {
          for (int e = 0; e < len; ++e) {
            int x = 0;
            for (int q = q0 + e * ko; q < q0 + (e + 1) * ko; ++q) x ^= img[pa + off[q]];
            X += x;
          }
        }
        if (ko & 1) X -= bs->slot_dummy[j] * o;
        const int Kj = (ko & 1) ? (2 * X - n) : (n - 2 * X);
        K[j] = Kj;
        F += bs->slot_w[j] * (float)Kj;
      }
    }
  }
  return F;
}

// ---------------------------------------------------------------------------
// Grouped-proposal Metropolis kernel (binary path).  KIND 0 canonical, 1 SGC;
// WEIGHTS 0 uniform, 1 learned network, 2 informed (exact Barker weights).
// ---------------------------------------------------------------------------
template <int KIND, int WEIGHTS>
__global__ __launch_bounds__(512) void learned_kernel(RunArgs args, LearnArgs la) {
  extern __shared__ __align__(16) unsigned char smem[];
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
  cptr<MlpDev> mlp = cst(la.mlp);

  const int L1 = sc->L1, L2 = sc->L2, L3 = sc->L3, h = sc->h;
  const int P1 = sc->P1, P12 = sc->P1 * sc->P2, S = sc->S;
  const int n_cells = sc->n_cells, n_sites = sc->n_sites, pad_bytes = sc->pad_bytes;
  const int n_classes = sc->n_classes;
  const int n_slots = sys->n_slots;

  uint8_t *img = (uint8_t *)smem;
  size_t off_b = (size_t)((pad_bytes + 15) & ~15);
  int *redI = (int *)(smem + off_b);  // [kMaxSlots + 4][16]
  off_b += (size_t)(kMaxSlots + 4) * 16 * sizeof(int);
  unsigned long long *maxDraw = (unsigned long long *)(smem + off_b);
  off_b += 16;
  float *logitC = (float *)(smem + off_b);  // [n_waves][64]
  float *logitN = logitC + 16 * 64;         // [n_waves][64]

  const int8_t *__restrict__ occG = args.occ + cp->occ_offset;
  for (int p = tid; p < pad_bytes; p += nthr) {
    int b = p / S;
    int r = p - b * S;
    int pk = r / P12;
    int r2 = r - pk * P12;
    int pj = r2 / P1;
    int pi = r2 - pj * P1;
    int i = ((pi - h) % L1 + L1) % L1;
    int j = ((pj - h) % L2 + L2) % L2;
    int k = ((pk - h) % L3 + L3) % L3;
    img[p] = (uint8_t)occG[b * n_cells + i + L1 * (j + L2 * k)];
  }
  if (tid == 0) *maxDraw = 0ull;
  __syncthreads();

  constexpr bool USE_MLP = WEIGHTS == 1;
  MlpFrags F;
  if (USE_MLP) load_frags(mlp, lane, F);
  const float beta = cp->beta;
  const float tsc = beta * kMlpEnergyUnit;
  const float eps_w = mlp->eps;
  // per-lane and per-wave random streams
  unsigned long long draw = cs->draw_base;
  uint32_t k0 = cp->key0, k1 = cp->key1;
  draw = ((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(draw >> 32)) << 32) |
         (unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)draw);
  const int wave_u = __builtin_amdgcn_readfirstlane(wave);

  int accC[kMaxSlots];
#pragma unroll
  for (int a = 0; a < kMaxSlots; ++a) accC[a] = 0;
  int accN1 = 0, accAtt = 0, accAcc = 0;  // per lane (pair phases, composition; a row's first lane
                                           // counts the row's grouped proposals)

  const unsigned long long pass0 = cs->pass_base;
  cptr<int> bin_off = cst(sc->bin_off);
  const int R = la.rounds_per_group < 1 ? 1 : la.rounds_per_group;
  float *myC = logitC + wave_u * 64;
  float *myN = logitN + wave_u * 64;

  for (int pass = 0; pass < args.n_passes; ++pass) {
    const unsigned long long gpass = pass0 + (unsigned long long)pass;
    for (int c = 0; c < n_classes; ++c) {
      const int b = sc->class_sublat[c];
      const int cb = sc->class_begin[c], ce = sc->class_begin[c + 1];
      const int n = ce - cb;
      cptr<BinSub> bs = cst(sc->bin_sub + b);
      cptr<int> env = cst(sc->env_off + b * kEnvMax);
      const float chem0 = cp->chem[b * kMaxOcc + 0], chem1 = cp->chem[b * kMaxOcc + 1];
      // random rotation of the class order (independent of the occupation)
      uint4 rr = philox4x32_10(make_uint4(0xFFFFFFFDu, (uint32_t)gpass, (uint32_t)c, (uint32_t)chain), k0, k1);
      const int rot = n ? (int)(rr.x % (uint32_t)n) : 0;
      const int n_groups = (n + 63) / 64;
      for (int grp = wave_u; grp < n_groups; grp += n_waves) {
        const int p = grp * 64 + lane;
        const bool valid = p < n;
        int q = p + rot;
        if (q >= n) q -= n;
        const int pa = valid ? sc->class_addr[cb + q] : sc->class_addr[cb];
        const int o0 = valid ? (int)img[pa] : 0;
        int o = o0;
        int K[kMaxSlots];
        const float Fl = binary_field_l(img, pa, o, bs, bin_off, n_slots, K);
        float e = (o ? -2.0f : 2.0f) * Fl;  // flip energy of the site
        if (KIND == 1) e -= o ? (chem0 - chem1) : (chem1 - chem0);
        float w = 1.0f, wn = 1.0f;
        if (USE_MLP) {
          uint32_t fm_lo = 0u, fm_hi = 0u;
          for (int kk = 0; kk < F.n_env; ++kk) {
            const uint32_t same = (int)img[pa + env[kk]] == o ? 1u : 0u;
            if (kk < 32) fm_lo |= same << kk;
            else fm_hi |= same << (kk - 32);
          }
          mlp_wave(F, fm_lo, fm_hi, lane, myC, myN);
          __builtin_amdgcn_s_waitcnt(0);
          __builtin_amdgcn_wave_barrier();
          w = F.eps + sigmoidf_(tsc * myC[lane]);
          wn = F.eps + sigmoidf_(tsc * myN[lane]);
        } else if (WEIGHTS == 2) {
          w = eps_w + sigmoidf_(-beta * e);
          wn = eps_w + sigmoidf_(beta * e);
        }
        if (!valid) {
          w = 0.0f;
          wn = 0.0f;
        }
        // The wavefront's 64 sites are four independent groups of 16, its
        // DPP rows: every round, each row with positive weights makes one
        // proposal among its own sites (the rows' sites share no cluster, so
        // their proposals are exact and concurrent).  Uniforms of 16 rounds
        // are drawn lane-parallel (lane 16 r + k holds row r's round k) and
        // moved to the row's lanes per round.
        const int rb = lane & 48;  // first lane of this lane's row
        uint4 ul = make_uint4(0u, 0u, 0u, 0u);
        for (int rd = 0; rd < R; ++rd) {
          const int slot = rd & 15;
          if (slot == 0) {
            const unsigned long long dl = draw + (unsigned long long)lane;
            ul = philox4x32_10(make_uint4(0xFFFF0000u | (uint32_t)wave_u, (uint32_t)dl, (uint32_t)(dl >> 32),
                                          (uint32_t)chain), k0, k1);
            draw += 64ull;
          }
          const int src = rb | slot;
          const uint32_t x1 = (uint32_t)__shfl((int)ul.x, src, 64);
          const uint32_t x2 = (uint32_t)__shfl((int)ul.y, src, 64);
          const uint32_t x3 = (uint32_t)__shfl((int)ul.z, src, 64);
          // the fourth word refines an acceptance test whose 32-bit word ties
          const uint32_t x4 = (uint32_t)__shfl((int)ul.w, src, 64);
          auto more = [&]() { return x4; };
          const float u1 = u01(x1);
          if (KIND == 0) {
            const float vA = (o == 0) ? w : 0.0f;
            const float vB = (o == 1) ? w : 0.0f;
            const float sA = row_scan_incl(vA), sB = row_scan_incl(vB);
            const float ZA = row_total(sA, rb), ZB = row_total(sB, rb);
            const bool live = ZA > 0.0f && ZB > 0.0f;
            const int la_ = row_pick(vA, sA, u1 * ZA, rb), lb_ = row_pick(vB, sB, u01(x2) * ZB, rb);
            const bool prop = live && la_ >= 0 && lb_ >= 0;
            // the picked lanes' values (every lane takes part in the shuffles)
            const int pa_ = prop ? la_ : lane, pb_ = prop ? lb_ : lane;
            const float wa = __shfl(w, pa_, 64), wna = __shfl(wn, pa_, 64), ea = __shfl(e, pa_, 64);
            const float wb = __shfl(w, pb_, 64), wnb = __shfl(wn, pb_, 64), eb = __shfl(e, pb_, 64);
            if (prop) {
              const float dE = ea + eb;
              const float ZA2 = ZA - wa + wnb, ZB2 = ZB - wb + wna;
              const float ratio = __fdividef(wnb * wna * ZA * ZB, wa * wb * ZA2 * ZB2);
              const bool acc = accept_prob(__expf(-beta * dE) * ratio, x3, more);
              if (lane == rb) {
                accAtt += 1;
                accAcc += acc ? 1 : 0;
              }
              if (acc && (lane == la_ || lane == lb_)) {
                o ^= 1;
                e = -e;
                const float t = w;
                w = wn;
                wn = t;
              }
            }
          } else {
            const float sv = row_scan_incl(w);
            const float Z = row_total(sv, rb);
            const int ls = row_pick(w, sv, u1 * Z, rb);
            const bool prop = Z > 0.0f && ls >= 0;
            const int ps = prop ? ls : lane;
            const float ws = __shfl(w, ps, 64), wns = __shfl(wn, ps, 64), es = __shfl(e, ps, 64);
            if (prop) {
              const float Z2 = Z - ws + wns;
              const float ratio = __fdividef(wns * Z, ws * Z2);
              const bool acc = accept_prob(__expf(-beta * es) * ratio, x2, more);
              if (lane == rb) {
                accAtt += 1;
                accAcc += acc ? 1 : 0;
              }
              if (acc && lane == ls) {
                o ^= 1;
                e = -e;
                const float t = w;
                w = wn;
                wn = t;
              }
            }
          }
        }
        // sites of one class share no cluster: apply the group's net changes
        if (valid && o != o0) {
          img[pa] = (uint8_t)o;
          const int sk = o0 ? -2 : 2;
#pragma unroll
          for (int a = 0; a < kMaxSlots; ++a) accC[a] += __mul24(sk, K[a]);  // small operands: exact
          if (KIND == 1) accN1 += o0 ? -1 : 1;
        }
      }
      __syncthreads();
      for (int t = sc->halo_begin[c] + tid; t < sc->halo_begin[c + 1]; t += nthr) {
        const int2 hc = sc->halo[t];
        img[hc.y] = img[hc.x];
      }
      __syncthreads();
    }

    // canonical cross-class mixing with the plain pair phases: as many per
    // pass as the plain kernel's automatic count (about n_mutable / 4 pair
    // attempts), each phase drawn at random
    const int n_pair_rounds =
        (KIND == 0 && sc->n_pair_phases > 0)
            ? auto_pair_rounds(sc->class_begin[n_classes], sc->pair_begin[sc->n_pair_phases], sc->n_pair_phases)
            : 0;
    for (int pr = 0; pr < n_pair_rounds; ++pr) {
      const uint32_t ph = philox4x32_10(make_uint4(0xFFFFFFFFu, (uint32_t)gpass, 0xFFFFFFF0u - (uint32_t)pr,
                                                   (uint32_t)chain ^ 0x9E3779B9u),
                                        k0, k1)
                              .x %
                          (uint32_t)sc->n_pair_phases;
      const int pb = sc->pair_begin[ph], pe = sc->pair_begin[ph + 1];
      for (int t = pb + tid; t < pe; t += nthr) {
        const int2 pq = sc->pair_addr[t];
        const int bpq = sc->pair_sublat[t];
        const int op = img[pq.x], oq = img[pq.y];
        if (op == oq) continue;
        const uint4 ur = philox4x32_10(make_uint4((uint32_t)tid, (uint32_t)draw, (uint32_t)(draw >> 32) | 0x80000000u,
                                                 (uint32_t)chain), k0, k1);
        int Kp[kMaxSlots], Kq[kMaxSlots];
        const float Fp = binary_field_l(img, pq.x, op, cst(sc->bin_sub + (bpq & 0xFF)), bin_off, n_slots, Kp);
        const float Fq = binary_field_l(img, pq.y, oq, cst(sc->bin_sub + (bpq >> 8)), bin_off, n_slots, Kq);
        const float dp = op ? -2.0f : 2.0f;
        const float dE = dp * Fp - dp * Fq;
        const bool acc = dE <= 0.0f || accept_prob(__expf(-beta * dE), ur.x, [&]() { return ur.y; });
        accAtt += 1;
        if (acc) {
          img[pq.x] = (uint8_t)oq;
          img[pq.y] = (uint8_t)op;
          accAcc += 1;
          const int sp = op ? -2 : 2;
#pragma unroll
          for (int a = 0; a < kMaxSlots; ++a) accC[a] += __mul24(sp, Kp[a] - Kq[a]);
        }
      }
      ++draw;
      __syncthreads();
      for (int t = sc->pair_halo_begin[ph] + tid; t < sc->pair_halo_begin[ph + 1]; t += nthr) {
        const int2 hc = sc->pair_halo[t];
        img[hc.y] = img[hc.x];
      }
      __syncthreads();
    }

    // canonical serial random-pair swaps, as in the plain kernel: two
    // uniformly random mutable sites of one asymmetric unit anywhere in the
    // supercell, evaluated by one lane as two sequential flips with every
    // periodic copy updated.  Group proposals never leave a class and pair
    // phases may be empty on small cells; these keep every cell ergodic.
    if (KIND == 0 && args.global_swaps > 0) {
      if (tid == 0) {
        const int n_var = sc->class_begin[n_classes];
        auto set_all = [&](int a, int nw) {
          const int b = a / S;
          const int r = a - b * S;
          const int pk = r / P12, pj = (r - pk * P12) / P1, pi = r - pk * P12 - pj * P1;
          const int P2 = P12 / P1, P3 = S / P12;
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
                img[b * S + qk * P12 + qj * P1 + qi] = (uint8_t)nw;
              }
            }
          }
        };
        for (int g = 0; g < args.global_swaps; ++g) {
          const uint4 ug = philox4x32_10(
              make_uint4(0xFFFFFFFEu, (uint32_t)gpass, (uint32_t)g, (uint32_t)chain ^ 0x5bd1e995u), k0, k1);
          const int is = min((int)(u01(ug.x) * (float)n_var), n_var - 1);
          const int it = min((int)(u01(ug.y) * (float)n_var), n_var - 1);
          const int sa = sc->class_addr[is], ta = sc->class_addr[it];
          if (sa == ta) continue;
          const int bsa = sa / S, bta = ta / S;
          if (sys->n_occ[bsa] != sys->n_occ[bta] || sys->asym[bsa] != sys->asym[bta]) continue;
          const int os = img[sa], ot = img[ta];
          if (os == ot) continue;
          int Ks[kMaxSlots], Kt[kMaxSlots];
          const float dE1 = (os ? -2.0f : 2.0f) * binary_field_l(img, sa, os, cst(sc->bin_sub + bsa), bin_off,
                                                                  n_slots, Ks);
          set_all(sa, ot);
          const float dE2 = (ot ? -2.0f : 2.0f) * binary_field_l(img, ta, ot, cst(sc->bin_sub + bta), bin_off,
                                                                  n_slots, Kt);
          const float dE = dE1 + dE2;
          accAtt += 1;
          if (dE <= 0.0f || accept_prob(__expf(-beta * dE), ug.z, [&]() { return ug.w; })) {
            set_all(ta, os);
            accAcc += 1;
            const int ss = os ? -2 : 2, st = ot ? -2 : 2;
#pragma unroll
            for (int a = 0; a < kMaxSlots; ++a) accC[a] += ss * Ks[a] + st * Kt[a];
          } else {
            set_all(sa, os);
          }
        }
      }
      __syncthreads();
    }

    // reduction and sampling
    const bool sample_due = ((gpass + 1) % (unsigned long long)args.sample_period) == 0ull;
    const bool flush = sample_due || ((pass & 31) == 31) || (pass == args.n_passes - 1);
    if (flush) {
#pragma unroll
      for (int a = 0; a < kMaxSlots; ++a) {
        const int v = wave_sum_i(accC[a]);
        if (lane == 0) redI[a * 16 + wave] = v;
        accC[a] = 0;
      }
      {
        const int v0 = wave_sum_i(accN1), v1 = wave_sum_i(accAtt), v2 = wave_sum_i(accAcc);
        if (lane == 0) {
          redI[(kMaxSlots + 0) * 16 + wave] = v0;
          redI[(kMaxSlots + 1) * 16 + wave] = v1;
          redI[(kMaxSlots + 2) * 16 + wave] = v2;
        }
        accN1 = accAtt = accAcc = 0;
      }
      __syncthreads();
      if (tid == 0) {
        int sum[kMaxSlots + 3];
        for (int q2 = 0; q2 < kMaxSlots + 3; ++q2) {
          int acc = 0;
          for (int w2 = 0; w2 < n_waves; ++w2) acc += redI[q2 * 16 + w2];
          sum[q2] = acc;
        }
        long long *C = args.binC + (long long)chain * args.n_orbits;
        for (int j = 0; j < n_slots; ++j) C[sys->slot_orbit[j]] += sum[j];
        double E = 0.0;
        for (int a = 0; a < args.n_orbits; ++a) E += sys->bin_w64[a] * (double)C[a];
        cs->E = E;
        cs->counts[1] += sum[kMaxSlots];
        cs->counts[0] -= sum[kMaxSlots];
        cs->n_attempt += sum[kMaxSlots + 1];
        cs->n_accept += sum[kMaxSlots + 2];
        if (sample_due) {
          const int idx = args.sample_count[chain];
          if (idx < args.sample_cap) {
            const long long o2 = (long long)chain * args.sample_cap + idx;
            args.sampE[o2] = E;
            for (int q3 = 0; q3 < kMaxOcc; ++q3) args.sampN[o2 * kMaxOcc + q3] = (int)cs->counts[q3];
            args.sample_count[chain] = idx + 1;
          }
          const double n1 = (double)cs->counts[1];
          cs->sumE += E;
          cs->sumE2 += E * E;
          cs->sumN1 += n1;
          cs->sumN12 += n1 * n1;
          cs->sumEN1 += E * n1;
          cs->n_samples_acc += 1;
        }
      }
      __syncthreads();
    }
  }
  {
    int8_t *occW = args.occ + cp->occ_offset;
    for (int l = tid; l < n_sites; l += nthr) {
      int b = l / n_cells;
      int cc = l - b * n_cells;
      int i = cc % L1;
      int j = (cc / L1) % L2;
      int k = cc / (L1 * L2);
      occW[l] = (int8_t)img[b * S + (k + h) * P12 + (j + h) * P1 + (i + h)];
    }
  }
  atomicMax(maxDraw, draw);
  __syncthreads();
  if (tid == 0) {
    cs->draw_base = *maxDraw + 1ull;
    cs->pass_base = pass0 + (unsigned long long)args.n_passes;
  }
}

// ---------------------------------------------------------------------------
// Diagnostic: MFMA inference on given +-1 feature rows (parity with the CPU)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void mlp_infer_kernel(const MlpDev *mlp, int n, const int8_t *X, float *out) {
  __shared__ float sC[64], sN[64];
  const int lane = threadIdx.x;
  const int base = blockIdx.x * 64;
  MlpFrags F;
  load_frags(cst(mlp), lane, F);
  // row of this lane as environment bits (input k > 0 <=> bit k set)
  uint32_t lo = 0u, hi = 0u;
  const int row = base + lane;
  if (row < n)
    for (int k = 0; k < F.n_env; ++k) {
      const uint32_t bit = X[(long long)row * kEnvMax + k] > 0 ? 1u : 0u;
      if (k < 32) lo |= bit << k;
      else hi |= bit << (k - 32);
    }
  mlp_wave(F, lo, hi, lane, sC, sN);
  __syncthreads();
  if (base + lane < n) {
    out[2 * (base + lane)] = sC[lane];
    out[2 * (base + lane) + 1] = sN[lane];
  }
}

// ---------------------------------------------------------------------------
// Diagnostic: the wavefront scan and weighted lane choice of the grouped
// proposals on given weights (known-answer checks against the CPU)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void wave_pick_kernel(int n_rows, const float *v, const float *u, float *scan_out,
                                                       int *pick_out) {
  const int row = blockIdx.x;
  const int lane = threadIdx.x;
  if (row >= n_rows) return;
  const float x = v[(long long)row * 64 + lane];
  const float sc = wave_scan_incl(x);
  const float total = lane_f(sc, 63);
  const int pk = wave_pick(x, sc, u[row] * total);
  scan_out[(long long)row * 64 + lane] = sc;
  if (lane == 0) pick_out[row] = pk;
}

hipError_t launch_wave_pick(int n_rows, const float *v, const float *u, float *scan_out, int *pick_out,
                            hipStream_t st) {
  hipLaunchKernelGGL(wave_pick_kernel, dim3(n_rows), dim3(64), 0, st, n_rows, v, u, scan_out, pick_out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Training data from the chains' current configurations: features of random
// variable sites, the heat-bath flip probability sigmoid(-beta e_s) as a
// soft label (e_s: flip energy incl. the chemical term when include_chem)
// and the chain's logit scale t = beta * kMlpEnergyUnit.
// ---------------------------------------------------------------------------
// Flip energy of variable site l (binary: to the other occupant) from the
// flower tables, including the chemical term when include_chem.
__device__ float collect_flip_energy(const CollectArgs &a, const ChainParams &cp, const DevSupercell &sc,
                                     const int8_t *__restrict__ occ, int l) {
  const DevSystem *__restrict__ sys = a.sys;
  const int b = l / sc.n_cells;
  const int c = l - b * sc.n_cells;
  const int i0 = c % sc.L1, j0 = (c / sc.L1) % sc.L2, k0 = c / (sc.L1 * sc.L2);
  const int o = occ[l];
  const int nw = 1 - o;
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
      idx += occ[d.x * sc.n_cells + ii + sc.L1 * (jj + sc.L2 * kk)] * sys->entry_mul[q];
    }
    const float *V = sys->flower_tables + sys->entry_table[e] + idx * nf;
    for (int f = 0; f < nf; ++f) G[f] += V[f];
  }
  float dE = 0.0f;
  for (int f = 0; f < nf; ++f) dE += (sys->site_func[b][f][nw] - sys->site_func[b][f][o]) * G[f];
  if (a.include_chem) dE -= cp.chem[b * kMaxOcc + nw] - cp.chem[b * kMaxOcc + o];
  return dE;
}

// Sample t of a chain: 8 random variable sites are drawn; even t keep the
// first, odd t the one with the largest flip probability.  Rare mobile
// environments (the sites a weighted proposal should find) thus make up
// about half of the training set even in well-ordered configurations.
__global__ void mlp_collect_kernel(CollectArgs a) {
  const int chain = blockIdx.y;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_per_chain) return;
  const DevSystem *__restrict__ sys = a.sys;
  const ChainParams &cp = a.params[chain];
  const DevSupercell &sc = a.scs[cp.sc];
  const int8_t *__restrict__ occ = a.occ + cp.occ_offset;
  const uint4 r0 = philox4x32_10(make_uint4((uint32_t)t, a.seed, 0xC011EC7u, (uint32_t)chain), cp.key0, cp.key1);
  const uint4 r1 = philox4x32_10(make_uint4((uint32_t)t, a.seed, 0xC011EC8u, (uint32_t)chain), cp.key0, cp.key1);
  const uint32_t rr[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
  const int n_cand = (t & 1) ? 8 : 1;
  int l = -1;
  float best_y = -1.0f;
  for (int m = 0; m < n_cand; ++m) {
    int lc = (int)(rr[m] % (uint32_t)sc.n_sites);
    for (int tries = 0; tries < sc.n_sites && sys->n_occ[lc / sc.n_cells] < 2; ++tries) lc = (lc + 1) % sc.n_sites;
    const float dE = collect_flip_energy(a, cp, sc, occ, lc);
    const float y = 1.0f / (1.0f + __expf(cp.beta * dE));
    if (y > best_y) {
      best_y = y;
      l = lc;
    }
  }
  const int b = l / sc.n_cells;
  const int c = l - b * sc.n_cells;
  const int i0 = c % sc.L1, j0 = (c / sc.L1) % sc.L2, k0 = c / (sc.L1 * sc.L2);
  const int o = occ[l];
  const long long out = (long long)chain * a.n_per_chain + t;
  for (int k = 0; k < kEnvMax; ++k) {
    int8_t f = 0;
    if (k < a.n_env) {
      const int4 d = a.env_sites[b * kEnvMax + k];
      int ii = (i0 + d.y) % sc.L1, jj = (j0 + d.z) % sc.L2, kk = (k0 + d.w) % sc.L3;
      if (ii < 0) ii += sc.L1;
      if (jj < 0) jj += sc.L2;
      if (kk < 0) kk += sc.L3;
      f = occ[d.x * sc.n_cells + ii + sc.L1 * (jj + sc.L2 * kk)] == o ? 1 : -1;
    }
    a.X[out * kEnvMax + k] = f;
  }
  a.y[out] = best_y;
  a.t[out] = cp.beta * kMlpEnergyUnit;
}

// ---------------------------------------------------------------------------
// Training: FP32 forward (per sample), gradients (per parameter), Adam.
// Loss: binary cross-entropy of sigmoid(t z) against the soft label y,
// averaged over the batch (z: network output, t: the sample's logit scale).
// ---------------------------------------------------------------------------
__global__ void mlp_forward_train_kernel(TrainArgs a) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.n) return;
  const float *W1 = a.W1;  // [64][48]
  float z = a.b2[0];
  for (int j = 0; j < kHidden; ++j) {
    float hj = a.b1[j];
    for (int k = 0; k < a.n_env; ++k) hj += W1[j * kEnvMax + k] * (float)a.X[(long long)s * kEnvMax + k];
    a.H[(long long)j * a.n + s] = hj;
    z += a.W2[j] * fmaxf(hj, 0.f);
  }
  const float ts = a.t ? a.t[s] : 1.0f;
  const float logit = ts * z;
  const float p = 1.0f / (1.0f + __expf(-logit));
  a.dl[s] = ts * (p - a.y[s]);  // d loss / d z
  const float y = a.y[s];
  const float lp = logit >= 0 ? -log1pf(__expf(-logit)) : logit - log1pf(__expf(logit));      // log p
  const float lq = logit >= 0 ? -logit - log1pf(__expf(-logit)) : -log1pf(__expf(logit));     // log(1-p)
  a.loss[s] = -(y * lp + (1.0f - y) * lq);
}

__global__ void mlp_grad_kernel(TrainArgs a) {
  // params: W1 [64*48], b1 [64], W2 [64], b2 [1]
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nW1 = kHidden * kEnvMax;
  if (t >= nW1 + 2 * kHidden + 1) return;
  const float inv = 1.0f / (float)a.n;
  double g = 0.0;
  if (t < nW1) {
    const int j = t / kEnvMax, k = t % kEnvMax;
    if (k < a.n_env) {
      const float w2 = a.W2[j];
      for (int s = 0; s < a.n; ++s) {
        const float hj = a.H[(long long)j * a.n + s];
        if (hj > 0.f) g += (double)(a.dl[s] * w2 * (float)a.X[(long long)s * kEnvMax + k]);
      }
    }
  } else if (t < nW1 + kHidden) {
    const int j = t - nW1;
    const float w2 = a.W2[j];
    for (int s = 0; s < a.n; ++s)
      if (a.H[(long long)j * a.n + s] > 0.f) g += (double)(a.dl[s] * w2);
  } else if (t < nW1 + 2 * kHidden) {
    const int j = t - nW1 - kHidden;
    for (int s = 0; s < a.n; ++s) g += (double)(a.dl[s] * fmaxf(a.H[(long long)j * a.n + s], 0.f));
  } else {
    for (int s = 0; s < a.n; ++s) g += (double)a.dl[s];
  }
  a.grad[t] = (float)(g * inv);
}

__global__ void mlp_adam_kernel(AdamArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_params) return;
  const float g = a.grad[t];
  const float m = a.b1m * a.m[t] + (1.0f - a.b1m) * g;
  const float v = a.b2m * a.v[t] + (1.0f - a.b2m) * g * g;
  a.m[t] = m;
  a.v[t] = v;
  const float mh = m / (1.0f - a.b1t), vh = v / (1.0f - a.b2t);
  a.p[t] -= a.lr * mh / (sqrtf(vh) + 1e-8f);
}

// FP32 master parameters -> device inference copy (BF16 W1, round to nearest even)
__global__ void mlp_publish_kernel(const float *p, MlpDev *m, int n_env, float eps) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nW1 = kHidden * kEnvMax;
  if (t < nW1) {
    const int k = t % kEnvMax;
    float x = k < n_env ? p[t] : 0.0f;
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    m->W1b[t] = (uint16_t)(u >> 16);
  } else if (t < nW1 + kHidden) {
    m->b1[t - nW1] = p[t];
  } else if (t < nW1 + 2 * kHidden) {
    m->W2[t - nW1 - kHidden] = p[t];
  } else if (t == nW1 + 2 * kHidden) {
    m->b2 = p[t];
    m->n_env = n_env;
    m->eps = eps;
  }
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
template <int KIND, int WEIGHTS>
static hipError_t launch_l(const RunArgs &args, const LearnArgs &la, int n_chains, int threads, size_t lds,
                           hipStream_t st, bool launch) {
  auto kern = learned_kernel<KIND, WEIGHTS>;
  if (!launch) return hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(kern, dim3(n_chains), dim3(threads), lds, st, args, la);
  return hipGetLastError();
}

hipError_t launch_learned(const RunArgs &args, const LearnArgs &la, int kind, int weights, int n_chains,
                          int threads, size_t lds, hipStream_t st, bool launch) {
  if (kind == 0) {
    if (weights == 1) return launch_l<0, 1>(args, la, n_chains, threads, lds, st, launch);
    if (weights == 2) return launch_l<0, 2>(args, la, n_chains, threads, lds, st, launch);
    return launch_l<0, 0>(args, la, n_chains, threads, lds, st, launch);
  }
  if (weights == 1) return launch_l<1, 1>(args, la, n_chains, threads, lds, st, launch);
  if (weights == 2) return launch_l<1, 2>(args, la, n_chains, threads, lds, st, launch);
  return launch_l<1, 0>(args, la, n_chains, threads, lds, st, launch);
}

size_t learned_lds_bytes(int pad_bytes) {
  return (size_t)((pad_bytes + 15) & ~15) + (size_t)(kMaxSlots + 4) * 16 * sizeof(int) + 16 +
         2 * 16 * 64 * sizeof(float) + 64;
}

hipError_t launch_mlp_infer(const MlpDev *mlp, int n, const int8_t *X, float *out, hipStream_t st) {
  hipLaunchKernelGGL(mlp_infer_kernel, dim3((n + 63) / 64), dim3(64), 0, st, mlp, n, X, out);
  return hipGetLastError();
}

hipError_t launch_mlp_collect(const CollectArgs &a, int n_chains, hipStream_t st) {
  hipLaunchKernelGGL(mlp_collect_kernel, dim3((a.n_per_chain + 127) / 128, n_chains), dim3(128), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_mlp_train_step(const TrainArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(mlp_forward_train_kernel, dim3((a.n + 127) / 128), dim3(128), 0, st, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int np = kHidden * kEnvMax + 2 * kHidden + 1;
  hipLaunchKernelGGL(mlp_grad_kernel, dim3((np + 63) / 64), dim3(64), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_mlp_adam(const AdamArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((a.n_params + 255) / 256), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_mlp_publish(const float *p, MlpDev *m, int n_env, float eps, hipStream_t st) {
  const int np = kHidden * kEnvMax + 2 * kHidden + 1;
  hipLaunchKernelGGL(mlp_publish_kernel, dim3((np + 255) / 256), dim3(256), 0, st, p, m, n_env, eps);
  return hipGetLastError();
}

}  // namespace clexmc
