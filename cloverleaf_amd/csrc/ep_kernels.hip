// HIP kernels for one Embedding Propagation training step on gfx950 (MI355X).
//
// Step = k_step_begin -> k_sample -> k_scan -> k_forward -> k_emit ->
//        {radix sort -> k_seg_adam -> k_pieces -> k_long_combine}
//     or {k_atomic_scatter -> k_adam_touched}
//
// Every kernel reads the per-step values (batch offset, batch size, learning rate,
// key count) from the device-resident StepState, so the whole step can be captured
// once in a hipGraph and replayed for every batch of a pass.  Grids are sized from
// host-known capacities and every kernel exits early (or grid-strides) against the
// device-side counts.
#include "ep_device.h"
#include "ep_kernels.h"
#include "ep_loss.h"

namespace ep {

static constexpr int NW = kBlockWaves;   // waves per 256-thread block
static constexpr uint32_t PIECE = kPieceEntries;  // entries per piece of a long segment
static constexpr int SR_UNROLL = 16;       // adjoint rows in flight per wave in the scatter
static constexpr uint32_t HASH_EMPTY = 0xFFFFFFFFu;
static constexpr uint32_t HASH_FBITS = 27;
static constexpr uint32_t HASH_FMASK = (1u << HASH_FBITS) - 1u;

// Independent partial sums of the gathers: row u of each group of U in-flight rows adds
// into partial u % kGatherSums, and the partials are added at the end (blocked
// summation: the FP32 rounding error of a mean over n rows grows with n / kGatherSums
// instead of n, which keeps means over thousands of rows within 1e-5 of FP64).
static constexpr int kGatherSums = 4;

// ------------------------------------------------------------------------------------
// Helpers
// ------------------------------------------------------------------------------------

// Gradient key of row f and the row of a key (see StepConfig: a multiplicative
// permutation modulo key_m).
// x * y mod m for x, y < m = 2^b - 1 (x * y < 2^2b): two folds and one subtraction.
__device__ inline uint32_t mul_mod_mersenne(uint32_t x, uint32_t y, uint32_t m, uint32_t b) {
  uint64_t p = (uint64_t)x * y;
  p = (p & m) + (p >> b);  // < 2^(b+1)
  p = (p & m) + (p >> b);  // <= m + 1
  return (uint32_t)(p >= m ? p - m : p);
}
__device__ inline uint32_t key_of_row(const StepConfig& c, uint32_t f) {
  return mul_mod_mersenne(f, c.key_a, c.key_m, c.key_bits);
}
__device__ inline uint32_t row_of_key(const StepConfig& c, uint32_t k) {
  return mul_mod_mersenne(k, c.key_ainv, c.key_m, c.key_bits);
}

// Adjoint row `slot` (whole-wave slice layout): FP32 rows in c.adj, or, with c.adjh,
// FP16 rows scaled by a power of two 2^e >= max |g| of the row (exact scaling, so the
// only rounding is the FP16 one: 2^-11 relative to the row's largest value).
template <int D>
__device__ inline void store_adj_row(const StepConfig& c, size_t slot, int lane, const float (&g)[RowShape<D>::PL]) {
  if (c.adjh) {
    constexpr int PL = RowShape<D>::PL;
    float mx = 0.f;
#pragma unroll
    for (int q = 0; q < PL; ++q)
      if (lane * PL + q < D) mx = fmaxf(mx, fabsf(g[q]));
    mx = wave_max_nonneg(mx);
    int ex = 0;
    (void)frexpf(mx, &ex);  // mx = m 2^ex, m in [0.5, 1) (ex = 0 for mx = 0)
    ex = ex > 127 ? 127 : ex;
    store_slice_f16<D>(c.adjh + slot * D, lane, g, ldexpf(1.0f, -ex));
    if (lane == 0) c.ascale[slot] = ldexpf(1.0f, ex);
  } else {
    store_slice<D>(c.adj + slot * D, lane, g);
  }
}

template <int D, bool H>
__device__ inline void load_adj_row(const StepConfig& c, size_t slot, int lane, float (&r)[RowShape<D>::PL]) {
  if constexpr (H) load_slice_f16<D>(c.adjh + slot * D, lane, r, c.ascale[slot]);
  else load_slice<D>(c.adj + slot * D, lane, r);
}

// First index i in [lo, hi) with K[i] >= x (hi if none), K sorted; 64-ary search by
// the whole wave.
__device__ inline uint32_t wave_lower_bound(const uint32_t* __restrict__ K, uint32_t lo, uint32_t hi, uint32_t x) {
  const uint32_t lane = lane_id();
  while (hi > lo) {
    const uint32_t span = hi - lo;
    const uint32_t step = (span + 63u) / 64u;
    const uint64_t off = (uint64_t)lane * step;
    const bool valid = off < span;
    const uint32_t kv = valid ? K[lo + (uint32_t)off] : 0u;
    const uint64_t m = __ballot(valid && kv >= x);
    if (step == 1u) return m ? lo + (uint32_t)(__ffsll((unsigned long long)m) - 1) : hi;
    if (m == 0ull) {
      lo = lo + ((span - 1u) / step) * step + 1u;  // past the last sample (all < x)
      continue;
    }
    const uint32_t f = (uint32_t)(__ffsll((unsigned long long)m) - 1);
    hi = lo + f * step;                        // this sample is >= x
    lo = f > 0u ? lo + (f - 1u) * step + 1u : lo;  // the one before is < x
  }
  return lo;
}

// Accumulate rows table[fids[fs..fe)] into acc (lane slice), U row loads in flight.
template <int D, int U, typename T>
__device__ inline void gather_rows(const T* __restrict__ table, const uint32_t* __restrict__ fids,
                                   uint64_t fs, uint64_t fe, int lane, float (&acc)[RowShape<D>::PL]) {
  constexpr int PL = RowShape<D>::PL;
  constexpr int A = U < kGatherSums ? U : kGatherSums;
  float part[A][PL];
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int q = 0; q < PL; ++q) part[i][q] = 0.f;
  for (uint64_t base = fs; base < fe; base += EP_WAVE) {
    const int n = (int)((fe - base) < (uint64_t)EP_WAVE ? (fe - base) : (uint64_t)EP_WAVE);
    const uint32_t myf = lane < n ? fids[base + lane] : 0u;
    for (int j = 0; j < n; j += U) {
      float r[U][PL];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int t = (j + u < n) ? (j + u) : (n - 1);
        const uint32_t f = __shfl(myf, t, EP_WAVE);
        load_slice<D>(table + (size_t)f * D, lane, r[u]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (j + u < n) {
#pragma unroll
          for (int q = 0; q < PL; ++q) part[u % A][q] += r[u][q];
        }
      }
    }
  }
#pragma unroll
  for (int i = 1; i < A; ++i)
#pragma unroll
    for (int q = 0; q < PL; ++q) part[0][q] += part[i][q];
#pragma unroll
  for (int q = 0; q < PL; ++q) acc[q] += part[0][q];
}

// 256-thread block exclusive scan of one value per thread; returns the block total in *total.
__device__ inline uint32_t block_exclusive_scan(uint32_t x, uint32_t* s_wave, uint32_t* total) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t inc = wave_inclusive_scan_u32(x);
  if (lane == 63) s_wave[w] = inc;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    const uint32_t sw = s_wave[i];
    if (i < w) pre += sw;
    tot += sw;
  }
  __syncthreads();
  *total = tot;
  return pre + inc - x;
}

// First index p in [lo, hi) with K[p] != f (K sorted ascending and K[i] == f before p),
// or hi.  64-ary search: all lanes probe at once, ~log64(span) dependent rounds.
__device__ inline uint32_t wave_upper_bound(const uint32_t* __restrict__ K, uint32_t lo, uint32_t hi, uint32_t f) {
  const uint32_t lane = lane_id();
  while (true) {
    const uint32_t span = hi - lo;
    if (span == 0) return lo;
    const uint32_t step = (span + 63u) / 64u;
    const uint64_t off = (uint64_t)lane * step;
    const bool valid = off < span;
    const uint32_t kv = valid ? K[lo + (uint32_t)off] : f;
    const uint64_t m = __ballot(valid && kv != f);
    if (m == 0ull) {
      const uint32_t lmax = (span - 1u) / step;
      lo = lo + lmax * step + 1u;
      if (step == 1u) return lo;
      continue;
    }
    const uint32_t L = (uint32_t)(__ffsll((unsigned long long)m) - 1);
    if (step == 1u) return lo + L;
    hi = lo + L * step;
    lo = L > 0 ? lo + (L - 1u) * step + 1u : lo;
  }
}

__device__ inline uint32_t hash_slot(uint32_t f, uint32_t mask) {
  return (f * 0x9E3779B1u) & mask;
}

// Insert (f, c) keeping the minimum construction index per feature.  Entries pack
// (c << 27) | f, so for a fixed f the minimum packed value has the minimum c.
__device__ inline void hash_insert_min(uint32_t* h, uint32_t mask, uint32_t f, uint32_t c) {
  const uint32_t kv = (c << HASH_FBITS) | f;
  uint32_t i = hash_slot(f, mask);
  while (true) {
    const uint32_t old = atomicCAS(&h[i], HASH_EMPTY, kv);
    if (old == HASH_EMPTY) return;
    if ((old & HASH_FMASK) == f) {
      atomicMin(&h[i], kv);
      return;
    }
    i = (i + 1u) & mask;
  }
}

__device__ inline uint32_t hash_lookup(const uint32_t* h, uint32_t mask, uint32_t f) {
  uint32_t i = hash_slot(f, mask);
  while (true) {
    const uint32_t val = h[i];
    if (val == HASH_EMPTY) return 0xFFu;
    if ((val & HASH_FMASK) == f) return val >> HASH_FBITS;
    i = (i + 1u) & mask;
  }
}

// ------------------------------------------------------------------------------------
// k_step_begin: advance the batch cursor, LR schedule and Adam bias corrections.
// LRScheduler::cos_decay (scheduler.rs:39-47) as built at mod.rs:132-137, in f32.
// ------------------------------------------------------------------------------------
__device__ inline float scheduled_lr(const StepConfig& c, uint64_t cur) {
  const float alpha = c.alpha;
  const float mn = alpha / 100.0f;
  const uint64_t warm = (uint64_t)((float)c.total_updates / 5.0f);
  if (cur > warm) {
    const float ratio = (float)cur / (float)c.total_updates;
    return mn + 0.5f * (alpha - mn) * (1.0f + cosf(3.14159265358979323846f * ratio));
  }
  const float ratio = (float)cur / (float)warm;
  return mn + ratio * (alpha - mn);
}

// The group of a step covers up to c.batch anchors, cut into reference batches of
// c.rbatch; batch k of the group takes step number next_step + k (mod.rs:234 fetch_add
// per batch) and its scheduled learning rate.
__global__ void k_step_begin(StepConfig c, StepState* st, float lr_override) {
  uint32_t* dt = &st->dtot[0][0];
  for (int i = threadIdx.x; i < 4 * 2048; i += blockDim.x) dt[i] = 0u;
  if (threadIdx.x == 0) {
    const uint32_t sip = st->step_in_pass + 1u;  // host writes 0xFFFFFFFF at pass start
    st->step_in_pass = sip;
    const uint64_t base64 = (uint64_t)sip * c.batch;
    const uint32_t base = (uint32_t)(base64 < c.n_batch_total ? base64 : c.n_batch_total);
    st->anchor_base = base;
    const uint32_t rem = c.n_batch_total - base;
    const uint32_t bn = rem < c.batch ? rem : c.batch;
    st->batch_n = bn;
    uint32_t nb = (bn + c.rbatch - 1u) / c.rbatch;
    if (nb > kMaxGroup) nb = kMaxGroup;
    st->n_batches = nb;
    const uint64_t cur = st->next_step;
    st->next_step = cur + nb;
    st->step0 = cur;
    for (uint32_t k = 0; k < kMaxGroup; ++k) {
      st->lr[k] = k < nb ? (lr_override >= 0.f ? lr_override : scheduled_lr(c, cur + k)) : 0.f;
      // LRScheduler::exp_decay(0, noise, decay) (mod.rs:140-147, scheduler.rs:49-51); with no
      // schedule length (explicit test steps) the sd stays at `noise`
      float sd = 0.f;
      if (c.noise > 1e-9f && k < nb)
        sd = c.total_updates ? fmaxf(c.noise * powf(c.noise_decay, (float)(cur + k)), 0.f) : c.noise;
      st->noise_sd[k] = sd;
    }
    const float t = (float)st->pass + 1.0f;  // optimizer.rs:50: t = pass + 1
    st->bc1 = 1.0f - powf(0.9f, t);
    st->bc2 = 1.0f - powf(0.999f, t);
    st->n_keys = 0u;
    st->n_real = 0u;
    st->n_unique = 0u;
    st->n_long = 0u;
    st->n_pieces = 0u;
    st->n_touched = 0u;
    st->n_short = 0u;
    st->stat_examples += bn;
  }
}

// ------------------------------------------------------------------------------------
// k_sample: one wavefront per anchor.  Neighbour sample: S of deg without replacement
// (Floyd's algorithm; picks in LDS, the duplicate test is one ballot), all
// neighbours when deg <= S; K negatives uniform with replacement from the pool; R_e =
// feature occurrences of the example.  model.rs:354-377 + utils.rs:193-216 (Fixed(S))
// and node_sampler.rs:185-188; the PPR loss takes random-walk endpoints instead of
// neighbours.  Draw i of a purpose is word i%4 of Philox block i/4.
// ------------------------------------------------------------------------------------
__device__ inline uint32_t draw_word(uint32_t pos, uint32_t pass, uint32_t purpose, uint32_t i, uint32_t k0,
                                     uint32_t k1) {
  const u32x4 r = philox4x32(u32x4{pos, pass, purpose, i >> 2}, k0, k1);
  const uint32_t q = i & 3u;
  return q == 0 ? r.x : q == 1 ? r.y : q == 2 ? r.z : r.w;
}

// The walk of random_walk (loss.rs:193-226, node_sampler.rs:109-140): from v, at most 10
// steps to a uniform neighbour; from the second step on it stops with probability
// restart_p at a node != v; it also stops at a node without neighbours.  Step i draws
// words r0 + 2i (neighbour) and r0 + 2i + 1 (stop).  Returns the node it ends on.
__device__ inline uint32_t walk_end(const StepConfig& c, uint32_t v, float restart_p, uint32_t pos, uint32_t pass,
                                    uint32_t purpose, uint32_t r0, uint32_t k0, uint32_t k1) {
  uint32_t node = v;
  for (uint32_t i = 1;; ++i) {
    const uint64_t e0 = c.indptr[node];
    const uint32_t dn = (uint32_t)(c.indptr[node + 1] - e0);
    if (dn == 0u || i > 10u) break;
    node = c.indices[e0 + uniform_below(draw_word(pos, pass, purpose, r0 + 2u * i, k0, k1), dn)];
    if (i > 1u && uniform01(draw_word(pos, pass, purpose, r0 + 2u * i + 1u, k0, k1)) < restart_p && node != v)
      break;
  }
  return node;
}

// K negatives of one example (one wave).  RandomWalkHardSampler (node_sampler.rs:83-107):
// with H = min(hard_negatives, K) > 0, up to 2H walks (restart 0.25) run in order and
// each endpoint that is neither v nor one of v's neighbours and not already taken becomes
// a negative, until H are taken; lane j runs walk j and the order is kept by ballots.
// The rest are uniform with replacement from the pool (node_sampler.rs:102-105).
__device__ inline void sample_negatives(const StepConfig& c, uint32_t v, uint64_t s0, uint32_t deg, uint32_t pos,
                                        uint32_t pass, uint32_t k0, uint32_t k1, uint32_t* negs, int lane) {
  uint32_t taken = 0;
  if (c.hard_negs) {
    const uint32_t H = c.hard_negs;
    bool ok = false;
    uint32_t cand = v;
    if ((uint32_t)lane < 2u * H) {
      cand = walk_end(c, v, 0.25f, pos, pass, RNG_HARD_WALK, (uint32_t)lane * 32u, k0, k1);
      ok = cand != v;
    }
    // drop endpoints adjacent to v: the wave reads v's edges 64 at a time
    if (wave_any(ok)) {
      for (uint32_t j0 = 0; j0 < deg; j0 += 64u) {
        const uint32_t e = j0 + (uint32_t)lane < deg ? c.indices[s0 + j0 + lane] : v;
        const uint32_t nj = min(64u, deg - j0);
        for (uint32_t t = 0; t < nj; ++t) ok &= (uint32_t)__shfl((int)e, (int)t) != cand;
      }
    }
    // keep the first walk of each endpoint
    bool dup = false;
    for (uint32_t t = 0; t < 2u * H; ++t) {
      const uint32_t ct = (uint32_t)__shfl((int)cand, (int)t);
      const bool okt = __shfl((int)ok, (int)t) != 0;
      dup |= okt && (t < (uint32_t)lane) && ct == cand;
    }
    const bool keep = ok && !dup;
    const uint64_t m = __ballot(keep);
    const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (keep && rank < H) negs[rank] = cand;
    taken = min((uint32_t)__popcll(m), H);
  }
  for (uint32_t k = taken + lane; k < c.K; k += 64u)
    negs[k] = c.negpool[uniform_below(draw_word(pos, pass, RNG_NEGATIVES, k, k0, k1), c.n_negpool)];
}

// CDFtoP (graph.rs:630-661): probability of edge j of the row starting at s0.
__device__ inline float edge_prob(const StepConfig& c, uint64_t s0, uint32_t j) {
  return j == 0u ? c.cdf[s0] : c.cdf[s0 + j] - c.cdf[s0 + j - 1u];
}

// Order-preserving map of a float to an unsigned key (larger float, larger key).
__device__ inline uint32_t float_order_key(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// weighted_reservoir_sample (utils.rs:147-161) keeps the n items with the largest
// u^(1/w); with w_j = scale * p_j * deg the common factor leaves the order of
// log(u_j) / p_j unchanged, so that is the key (u_j: draw j of the neighbour stream).
__device__ inline uint32_t ares_key(const StepConfig& c, uint64_t s0, uint32_t j, uint32_t pos, uint32_t pass,
                                    uint32_t k0, uint32_t k1) {
  const float u = uniform01(draw_word(pos, pass, RNG_NEIGHBOURS, j, k0, k1));
  return float_order_key(logf(u) / edge_prob(c, s0, j));
}

// One wave: the S of deg edges with the largest keys, by an MSB-first radix select
// (8-bit digits, per-wave histogram in LDS; the keys are recomputed each pass), ties
// at the threshold taken in edge order.  Writes the neighbours to out[0..S) in edge
// order and their weights scale * p_j * deg to w[0..S).
__device__ inline void ares_select(const StepConfig& c, uint64_t s0, uint32_t deg, uint32_t S, float scale,
                                   uint32_t pos, uint32_t pass, uint32_t k0, uint32_t k1, uint32_t* hist,
                                   uint32_t* out, float* w, int lane) {
  uint32_t prefix = 0u, pmask = 0u, need = S;
  for (int shift = 24; shift >= 0; shift -= 8) {
    for (int i = lane; i < 256; i += 64) hist[i] = 0u;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t j = lane; j < deg; j += 64u) {
      const uint32_t k = ares_key(c, s0, j, pos, pass, k0, k1);
      if ((k & pmask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1u);
    }
    __builtin_amdgcn_wave_barrier();
    // lane l owns digits 4l..4l+3; above = elements with a larger digit
    uint32_t h[4], own = 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q) { h[q] = hist[4 * lane + q]; own += h[q]; }
    uint32_t suf = own;  // inclusive suffix sum over lanes >= l
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t t = (uint32_t)__shfl_down((int)suf, off, EP_WAVE);
      if (lane + off < 64) suf += t;
    }
    uint32_t above = suf - own, dsel = 0u, asel = 0u;
    bool found = false;
#pragma unroll
    for (int q = 3; q >= 0; --q) {
      if (!found && above < need && need <= above + h[q]) {
        found = true;
        dsel = 4u * (uint32_t)lane + (uint32_t)q;
        asel = above;
      }
      above += h[q];
    }
    const uint64_t m = __ballot(found);
    const int src = __ffsll((unsigned long long)m) - 1;
    dsel = (uint32_t)__shfl((int)dsel, src);
    asel = (uint32_t)__shfl((int)asel, src);
    need -= asel;
    prefix |= dsel << shift;
    pmask |= 0xFFu << shift;
    __builtin_amdgcn_wave_barrier();
  }
  uint32_t taken = 0u, eq_seen = 0u;
  for (uint32_t j0 = 0; j0 < deg; j0 += 64u) {
    const uint32_t j = j0 + (uint32_t)lane;
    uint32_t k = 0u;
    if (j < deg) k = ares_key(c, s0, j, pos, pass, k0, k1);
    const bool eq = j < deg && k == prefix;
    const uint64_t me = __ballot(eq);
    const uint32_t eq_rank = eq_seen + (uint32_t)__popcll(me & ((1ull << lane) - 1ull));
    const bool sel = j < deg && (k > prefix || (eq && eq_rank < need));
    const uint64_t ms = __ballot(sel);
    if (sel) {
      const uint32_t at = taken + (uint32_t)__popcll(ms & ((1ull << lane) - 1ull));
      out[at] = c.indices[s0 + j];
      if (w) w[at] = scale * edge_prob(c, s0, j) * (float)deg;
    }
    taken += (uint32_t)__popcll(ms);
    eq_seen += (uint32_t)__popcll(me);
  }
}

// max_features selection of one node of an example (utils.rs:193-216 with at_least_one,
// model.rs:264-269): bit i of mask = position i of the node's feature list is kept, and
// scale is added to the weight of every kept occurrence.  Fixed(k), n > k: Floyd's
// k-subset; Probability(p): every position independently with probability p, one uniform
// position when none is kept (scale (1 - p^n) / p, as the reference writes it).  Stream:
// purpose RNG_FEATURES | (e << 4), e = 0 (v), 1..cnt (neighbours), then the negatives.
struct FeatSel {
  uint64_t mask;
  float scale;
};
__device__ inline FeatSel feature_select(const StepConfig& c, uint32_t n, uint32_t pos, uint32_t pass, uint32_t e,
                                         uint32_t k0, uint32_t k1) {
  const uint64_t all = n >= 64u ? ~0ull : ((1ull << n) - 1ull);
  const uint32_t purpose = (uint32_t)RNG_FEATURES | (e << 4);
  if (c.mf_kind == 1) {
    if (n <= c.mf_k) return FeatSel{all, 1.f};
    uint64_t m = 0;
    for (uint32_t i = n - c.mf_k; i < n; ++i) {
      const uint32_t t = uniform_below(draw_word(pos, pass, purpose, i, k0, k1), i + 1u);
      m |= ((m >> t) & 1ull) ? (1ull << i) : (1ull << t);
    }
    return FeatSel{m, (float)n / (float)c.mf_k};
  }
  uint64_t m = 0;
  for (uint32_t i0 = 0; i0 < n; i0 += 4u) {
    const u32x4 r = philox4x32(u32x4{pos, pass, purpose, i0 >> 2}, k0, k1);
    const uint32_t w4[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
    for (uint32_t q = 0; q < 4u; ++q)
      if (i0 + q < n && uniform01(w4[q]) < c.mf_p) m |= 1ull << (i0 + q);
  }
  if (m) return FeatSel{m, 1.f / c.mf_p};
  const uint32_t t = uniform_below(draw_word(pos, pass, purpose, 64u, k0, k1), n);
  return FeatSel{1ull << t, (1.f - powf(c.mf_p, (float)n)) / c.mf_p};
}

// Number of kept features of a node (the popcount of its selection).
__device__ inline uint32_t feature_count(const StepConfig& c, uint32_t n, uint32_t pos, uint32_t pass, uint32_t e,
                                         uint32_t k0, uint32_t k1) {
  if (c.mf_kind == 0) return n;
  if (c.mf_kind == 1) return min(n, c.mf_k);
  return (uint32_t)__popcll(feature_select(c, n, pos, pass, e, k0, k1).mask);
}

// Position of the q-th set bit (q < popcount).
__device__ inline uint32_t nth_set_bit(uint64_t m, uint32_t q) {
  for (uint32_t i = 0; i < q; ++i) m &= m - 1ull;
  return (uint32_t)__ffsll((unsigned long long)m) - 1u;
}

__global__ __launch_bounds__(256) void k_sample(StepConfig c, StepState* st) {
  const int lane = threadIdx.x & 63;
  const uint32_t b = blockIdx.x * NW + (threadIdx.x >> 6);
  if (b >= c.batch) return;
  const uint32_t bn = st->batch_n;
  if (b >= bn) {
    if (lane == 0) {
      c.rcount[b] = 0u;
      c.nbr_cnt[b] = 0u;
    }
    return;
  }
  const uint32_t pos = st->anchor_base + b;
  const uint32_t pass = st->pass;
  const uint32_t v = c.perm[pos];
  const uint64_t s0 = c.indptr[v], s1 = c.indptr[v + 1];
  const uint32_t deg = (uint32_t)(s1 - s0);
  uint32_t* out = c.nbr + (size_t)b * c.S;
  uint32_t* negs = c.neg + (size_t)b * c.K;
  const uint32_t k0 = (uint32_t)c.seed, k1 = (uint32_t)(c.seed >> 32);
  uint32_t cnt;
  if (!c.explicit_samples && c.loss_kind == 5) {
    // Loss::PPR positives (loss.rs:191-226): K random walks from v of at most 10 steps;
    // after the first step a walk stops with probability restart_p at a node != v; a walk
    // that ends on v falls back to a uniform neighbour of v; no neighbours: v itself.
    uint32_t endp = 0;
    bool some = false;
    if ((uint32_t)lane < c.K) {
      const uint32_t r0 = (uint32_t)lane * 32u;
      const uint32_t node = walk_end(c, v, c.loss_p2, pos, pass, RNG_WALK, r0, k0, k1);
      if (node != v) {
        endp = node;
        some = true;
      } else if (deg > 0u) {
        endp = c.indices[s0 + uniform_below(draw_word(pos, pass, RNG_WALK, r0 + 30u, k0, k1), deg)];
        some = true;
      }
    }
    const uint64_t m = __ballot(some);
    if (some) out[__popcll(m & ((1ull << lane) - 1ull))] = endp;
    cnt = (uint32_t)__popcll(m);
    if (cnt == 0u) {
      if (lane == 0) out[0] = v;
      cnt = 1u;
    }
    sample_negatives(c, v, s0, deg, pos, pass, k0, k1, negs, lane);
    if (lane == 0) c.nbr_cnt[b] = cnt;
  } else if (!c.explicit_samples && c.pnbrs) {
    // Sample::Probability(p).sample(deg, true) (utils.rs:203-210) and the reservoirs of
    // model.rs:363-377: every edge is kept with probability p (a Binomial count with a
    // uniform subset of that size), scale 1/p; none kept: one uniform edge, scale
    // (1 - p^deg) / p.  With weighted sampling and averaging the kept count is then
    // re-drawn by the edge-weight reservoir.  List and weights go to v's CSR offset.
    uint32_t* o2 = c.nbr_sel + s0;
    float* w2 = c.nbr_rows ? c.nbr_selw + s0 : nullptr;
    uint32_t k = 0;
    float scale = 1.f / c.mn_p;
    for (uint32_t j0 = 0; j0 < deg; j0 += 64u) {
      const uint32_t j = j0 + (uint32_t)lane;
      const bool keep = j < deg && uniform01(draw_word(pos, pass, RNG_NBR_KEEP, j, k0, k1)) < c.mn_p;
      const uint64_t m = __ballot(keep);
      if (keep) {
        const uint32_t at = k + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        o2[at] = c.indices[s0 + j];
        if (w2) w2[at] = c.wavg ? scale * edge_prob(c, s0, j) * (float)deg : scale;
      }
      k += (uint32_t)__popcll(m);
    }
    if (deg > 0u && k == 0u) {
      const uint32_t t = uniform_below(draw_word(pos, pass, RNG_NBR_KEEP, deg, k0, k1), deg);
      scale = (1.f - powf(c.mn_p, (float)deg)) / c.mn_p;
      if (lane == 0) {
        o2[0] = c.indices[s0 + t];
        if (w2) w2[0] = c.wavg ? scale * edge_prob(c, s0, t) * (float)deg : scale;
      }
      k = 1u;
    }
    if (c.wavg && c.wsample && k < deg) {
      __shared__ uint32_t s_hist2[NW][256];
      ares_select(c, s0, deg, k, scale, pos, pass, k0, k1, s_hist2[threadIdx.x >> 6], o2, w2, lane);
    }
    cnt = k;
    sample_negatives(c, v, s0, deg, pos, pass, k0, k1, negs, lane);
    if (lane == 0) c.nbr_cnt[b] = cnt;
  } else if (!c.explicit_samples) {
    if (c.all_nbrs) {
      cnt = deg;
    } else if (deg <= c.S) {
      cnt = deg;
      for (uint32_t j = lane; j < deg; j += 64u) out[j] = c.indices[s0 + j];
      if (c.wavg)  // Fixed(S) with deg <= S: scale deg/deg = 1
        for (uint32_t j = lane; j < deg; j += 64u) c.nbr_w[(size_t)b * c.S + j] = edge_prob(c, s0, j) * (float)deg;
    } else if (c.wavg && c.wsample) {
      cnt = c.S;
      __shared__ uint32_t s_hist[NW][256];
      ares_select(c, s0, deg, c.S, (float)deg / (float)c.S, pos, pass, k0, k1, s_hist[threadIdx.x >> 6], out,
                  c.nbr_w + (size_t)b * c.S, lane);
    } else {
      cnt = c.S;
      // Floyd: for i = 0..S-1 draw t in [0, deg-S+i]; take deg-S+i instead when t was
      // already picked.  Up to 64 picks: lane i draws pick i's t (all draws at once), then
      // the picks are settled in order, each membership test one ballot over the lanes
      // already settled.  More picks: one at a time, kept in LDS.
      if (c.S <= 64u) {
        const uint32_t hi = deg - c.S + (uint32_t)lane;
        const uint32_t t = (uint32_t)lane < c.S
                               ? uniform_below(draw_word(pos, pass, RNG_NEIGHBOURS, (uint32_t)lane, k0, k1), hi + 1u)
                               : 0u;
        uint32_t pick = t;
        for (uint32_t i = 1; i < c.S; ++i) {
          const uint32_t ti = (uint32_t)__builtin_amdgcn_readlane((int)t, (int)i);
          const bool dup = __ballot((uint32_t)lane < i && pick == ti) != 0ull;
          if ((uint32_t)lane == i && dup) pick = deg - c.S + i;
        }
        if ((uint32_t)lane < c.S) {
          out[lane] = c.indices[s0 + pick];
          if (c.wavg) c.nbr_w[(size_t)b * c.S + lane] = (float)deg / (float)c.S * edge_prob(c, s0, pick) * (float)deg;
        }
      } else {
        extern __shared__ uint32_t s_pick[];
        uint32_t* pk = s_pick + (threadIdx.x >> 6) * c.S;
        for (uint32_t i = 0; i < c.S; ++i) {
          const uint32_t hi = deg - c.S + i;
          const uint32_t t = uniform_below(draw_word(pos, pass, RNG_NEIGHBOURS, i, k0, k1), hi + 1u);
          bool dup = false;
          for (uint32_t q = lane; q < i; q += 64u) dup |= (pk[q] == t);
          const bool any = wave_any(dup);
          if (lane == 0) pk[i] = any ? hi : t;
          __builtin_amdgcn_wave_barrier();
        }
        for (uint32_t j = lane; j < c.S; j += 64u) out[j] = c.indices[s0 + pk[j]];
        if (c.wavg) {
          const float scale = (float)deg / (float)c.S;
          for (uint32_t j = lane; j < c.S; j += 64u)
            c.nbr_w[(size_t)b * c.S + j] = scale * edge_prob(c, s0, pk[j]) * (float)deg;
        }
      }
    }
    sample_negatives(c, v, s0, deg, pos, pass, k0, k1, negs, lane);
    if (lane == 0) c.nbr_cnt[b] = cnt;
  } else {
    cnt = c.nbr_cnt[b];
  }
  uint32_t R = 0;
  if (cnt > 0u) {
    const uint32_t* pn = nbr_list(c, b, s0);
    const uint32_t E = 1u + cnt + c.K;  // v, neighbours, negatives
    for (uint32_t e = lane; e < E; e += 64u) {
      const uint32_t u = e == 0u ? v : (e <= cnt ? pn[e - 1u] : negs[e - 1u - cnt]);
      const uint64_t fs = c.findptr[u];
      const uint32_t n = (uint32_t)(c.findptr[u + 1] - fs);
      if (e < c.ent_cap) {  // the forward stages these lists without the node lookups
        c.ex_fs[(size_t)b * c.ent_cap + e] = fs;
        c.ex_fn[(size_t)b * c.ent_cap + e] = n;
      }
      R += feature_count(c, n, pos, pass, e, k0, k1);
    }
    R = wave_sum_u32(R);
  }
  if (lane == 0) c.rcount[b] = R;
}

// ------------------------------------------------------------------------------------
// k_scan: exclusive scan of R_e over the batch -> key offsets; total -> st->n_keys.
// Attention model: also the scan of the attention blocks per example (1 + cnt + K for
// an example with neighbours) -> blkoff.
// ------------------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_scan(StepConfig c, StepState* st) {
  __shared__ uint32_t s_w[16], s_w2[16], s_w3[16];
  __shared__ uint32_t s_carry, s_carry2, s_carry3;
  __shared__ unsigned long long s_nodes;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const bool att = c.model == 1;
  const bool nrows = c.nbr_rows && (c.all_nbrs || c.pnbrs);  // per-neighbour rows at scanned offsets
  if (tid == 0) { s_carry = 0u; s_carry2 = 0u; s_carry3 = 0u; s_nodes = 0ull; }
  __syncthreads();
  uint64_t nodes = 0;
  // rounds of 1024 x IT anchors: thread t takes IT consecutive ones (a running sum in
  // registers), then one block scan of the thread sums
  constexpr uint32_t IT = 8;
  for (uint32_t base = 0; base < c.batch; base += 1024u * IT) {
    uint32_t x[IT], y[IT], z[IT];
    uint32_t sx = 0, sy = 0, sz = 0;
#pragma unroll
    for (uint32_t q = 0; q < IT; ++q) {
      const uint32_t i = base + (uint32_t)tid * IT + q;
      x[q] = i < c.batch ? c.rcount[i] : 0u;
      y[q] = z[q] = 0u;
      if (i < c.batch && i < st->batch_n) {
        const uint32_t cn = c.nbr_cnt[i];
        if (cn > 0u) {
          nodes += cn + c.K;
          y[q] = 1u + cn + c.K;
          z[q] = cn;
        }
      }
      sx += x[q];
      sy += y[q];
      sz += z[q];
    }
    uint32_t inc = sx, inc2 = sy, inc3 = sz;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t t1 = __shfl_up(inc, off, EP_WAVE);
      const uint32_t t2 = __shfl_up(inc2, off, EP_WAVE);
      const uint32_t t3 = __shfl_up(inc3, off, EP_WAVE);
      if (lane >= off) { inc += t1; inc2 += t2; inc3 += t3; }
    }
    if (lane == 63) { s_w[w] = inc; s_w2[w] = inc2; s_w3[w] = inc3; }
    __syncthreads();
    uint32_t pre = s_carry, tot = 0, pre2 = s_carry2, tot2 = 0, pre3 = s_carry3, tot3 = 0;
    for (int q = 0; q < 16; ++q) {
      if (q < w) { pre += s_w[q]; pre2 += s_w2[q]; pre3 += s_w3[q]; }
      tot += s_w[q];
      tot2 += s_w2[q];
      tot3 += s_w3[q];
    }
    uint32_t o1 = pre + inc - sx, o2 = pre2 + inc2 - sy, o3 = pre3 + inc3 - sz;  // this thread's first offsets
#pragma unroll
    for (uint32_t q = 0; q < IT; ++q) {
      const uint32_t i = base + (uint32_t)tid * IT + q;
      if (i < c.batch) {
        c.keyoff[i] = o1;
        if (att) c.blkoff[i] = o2;
        if (nrows) c.nbroff[i] = o3;
      }
      o1 += x[q];
      o2 += y[q];
      o3 += z[q];
    }
    __syncthreads();
    if (tid == 0) { s_carry += tot; s_carry2 += tot2; s_carry3 += tot3; }
    __syncthreads();
  }
  atomicAdd(&s_nodes, (unsigned long long)nodes);
  __syncthreads();
  if (tid == 0) {
    c.keyoff[c.batch] = s_carry;
    st->n_keys = s_carry;
    st->stat_rows += s_carry;
    st->stat_nodes += s_nodes;
    if (s_carry > c.key_cap) st->err = 1u;  // capacity error (checked by the host)
    if (att) {
      // blocks past the capacity are not run (and their keys not emitted)
      c.blkoff[c.batch] = s_carry2 <= c.blk_cap ? s_carry2 : 0u;
      if (s_carry2 > c.blk_cap) { st->err = 1u; st->n_keys = c.key_cap + 1u; }
    }
  }
}

// Accumulate rows table[ids[a..b)] (ids in LDS) into acc, U row loads in flight.
template <int D, int U, typename T>
__device__ inline void gather_lds(const T* __restrict__ table, const uint32_t* ids, uint32_t a, uint32_t b, int lane,
                                  float (&acc)[RowShape<D>::PL]) {
  constexpr int PL = RowShape<D>::PL;
  constexpr int A = U < kGatherSums ? U : kGatherSums;
  float part[A][PL];
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int q = 0; q < PL; ++q) part[i][q] = 0.f;
  for (uint32_t j = a; j < b; j += U) {
    float r[U][PL];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t t = (j + u < b) ? (j + u) : (b - 1u);
      load_slice<D>(table + (size_t)ids[t] * D, lane, r[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (j + u < b) {
#pragma unroll
        for (int q = 0; q < PL; ++q) part[u % A][q] += r[u][q];
      }
    }
  }
#pragma unroll
  for (int i = 1; i < A; ++i)
#pragma unroll
    for (int q = 0; q < PL; ++q) part[0][q] += part[i][q];
#pragma unroll
  for (int q = 0; q < PL; ++q) acc[q] += part[0][q];
}

// Accumulate the rows of the features kept by a max_features mask (bit i: fids[fs + i]),
// U row loads in flight.
template <int D, int U, typename T>
__device__ inline void gather_masked(const T* __restrict__ table, const uint32_t* __restrict__ fids, uint64_t fs,
                                     uint64_t mask, int lane, float (&acc)[RowShape<D>::PL]) {
  constexpr int PL = RowShape<D>::PL;
  while (mask) {
    uint32_t f[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      ok[u] = mask != 0ull;
      const uint32_t i = ok[u] ? (uint32_t)__ffsll((unsigned long long)mask) - 1u : 0u;
      mask &= mask - 1ull;
      f[u] = fids[fs + i];
    }
    float r[U][PL];
#pragma unroll
    for (int u = 0; u < U; ++u) load_slice<D>(table + (size_t)(ok[u] ? f[u] : f[0]) * D, lane, r[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (ok[u]) {
#pragma unroll
        for (int q = 0; q < PL; ++q) acc[q] += r[u][q];
      }
    }
  }
}

// Wide gather: lane groups of LW = D / VW lanes each load one row with 16-byte loads
// (VW = 4 FP32 or 8 BF16 values per lane), RPW = 64 / LW rows per wave instruction and U
// rows in flight per group; the group partial sums are added across groups and lanes of
// group 0 write scale * sum to dst[0..D) (LDS).  Used where LW is a power of two.
template <typename T> struct WideVec;
template <> struct WideVec<float> { static constexpr int VW = 4; };
template <> struct WideVec<uint16_t> { static constexpr int VW = 8; };

template <int D, typename T>
struct WideShape {
  static constexpr int VW = WideVec<T>::VW;
  static constexpr int LW = D / VW;
  static constexpr bool OK = (D % VW == 0) && LW >= 1 && LW <= 64 && ((LW & (LW - 1)) == 0);
  static constexpr int RPW = OK ? 64 / LW : 1;
};

// Zero ids after the forward's staged occurrence list: a wide gather round reads the ids
// of up to (U - 1) * RPW rows past its range without a bounds check (the next
// construction's ids, or these: valid rows either way) and adds only the rows inside
// it.  U is at most 2x(8 for RPW < 4, else 4) rows in flight per lane group.
__host__ __device__ constexpr uint32_t ids_pad_entries(uint32_t d, bool bf16) {
  const uint32_t vw = bf16 ? 8u : 4u, lw = d / vw;
  if (d % vw != 0u || lw < 1u || lw > 64u || (lw & (lw - 1u)) != 0u) return 0u;
  const uint32_t rpw = 64u / lw, u = (rpw >= 4u ? 4u : 8u) * 2u;
  return (u - 1u) * rpw + 1u;
}

template <typename T>
__device__ inline void load_wide(const T* __restrict__ p, float (&v)[WideVec<T>::VW]);
template <>
__device__ inline void load_wide<float>(const float* __restrict__ p, float (&v)[4]) {
  const float4 t = *reinterpret_cast<const float4*>(p);
  v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
}
template <>
__device__ inline void load_wide<uint16_t>(const uint16_t* __restrict__ p, float (&v)[8]) {
  const uint4 t = *reinterpret_cast<const uint4*>(p);
  const uint32_t w[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[2 * k] = __uint_as_float(w[k] << 16);
    v[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
  }
}

template <int D, int U, typename T>
__device__ inline void gather_lds_wide(const T* __restrict__ table, const uint32_t* ids, uint32_t a, uint32_t b,
                                       int lane, float* dst, float scale) {
  using WS = WideShape<D, T>;
  constexpr int VW = WS::VW, LW = WS::LW, RPW = WS::RPW;
  static_assert((U - 1) * RPW < (int)ids_pad_entries(D, sizeof(T) == 2), "a round reads at most into the zero padding");
  const int g = lane / LW, lg = lane % LW;
  constexpr int A = U < kGatherSums ? U : kGatherSums;
  // The row stride is hidden from the optimiser, so that each row address is one
  // 32 x 32 + 64-bit multiply-add of the id (a shift of a zero-extended id ties every id
  // read to one register pair and serialises the LDS reads of a round).
  uint32_t rb = (uint32_t)(D * sizeof(T));
  asm volatile("" : "+s"(rb));
  const char* tl = reinterpret_cast<const char*>(table + lg * VW);
  float part[A][VW];
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = 0; k < VW; ++k) part[i][k] = 0.f;
  for (uint32_t j = a + (uint32_t)g; j < b; j += (uint32_t)(RPW * U)) {
    uint32_t id[U];
#pragma unroll
    for (int u = 0; u < U; ++u) id[u] = ids[j + (uint32_t)(u * RPW)];
    float r[U][VW];
#pragma unroll
    for (int u = 0; u < U; ++u) load_wide<T>(reinterpret_cast<const T*>(tl + (uint64_t)id[u] * rb), r[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (j + (uint32_t)(u * RPW) < b) {
#pragma unroll
        for (int k = 0; k < VW; ++k) part[u % A][k] += r[u][k];
      }
    }
  }
  float acc[VW];
#pragma unroll
  for (int k = 0; k < VW; ++k) {
    acc[k] = part[0][k];
#pragma unroll
    for (int i = 1; i < A; ++i) acc[k] += part[i][k];
  }
#pragma unroll
  for (int off = LW; off < 64; off <<= 1) {
#pragma unroll
    for (int k = 0; k < VW; ++k) acc[k] += __shfl_xor(acc[k], off, EP_WAVE);
  }
  if (g == 0) {
#pragma unroll
    for (int k = 0; k < VW; ++k) dst[lg * VW + k] = acc[k] * scale;
  }
}

// Weight w_j of neighbour j of example b before the feature scale, and its own adjoint
// row: 1 for PPR walk endpoints (loss.rs:171-177); with weighted averaging the sampler's
// value (Fixed) or p_j * deg (every neighbour); else the max_nodes scale, deg / S when
// subsampled and 1 otherwise (model.rs:355-361).
__device__ inline float nbr_weight(const StepConfig& c, uint32_t b, uint32_t j, uint64_t s0, uint32_t deg) {
  if (c.loss_kind == 5) return 1.f;
  if (c.pnbrs) return c.nbr_selw[s0 + j];
  if (c.wavg) return c.all_nbrs ? edge_prob(c, s0, j) * (float)deg : c.nbr_w[(size_t)b * c.S + j];
  if (c.all_nbrs || deg <= c.S) return 1.f;
  return (float)deg / (float)c.S;
}
__device__ inline uint32_t nbr_row(const StepConfig& c, uint32_t b, uint32_t j) {
  return c.nrow_base + ((c.all_nbrs || c.pnbrs) ? c.nbroff[b] : b * c.S) + j;
}

// Loss, adjoints and eligibility of one example for the losses other than margin (block
// of k_forward; xc holds a, p, n_1..n_K, sn the occurrence totals).
template <int D>
__device__ inline void forward_other_loss(const StepConfig& c, StepState* st, uint32_t b, uint32_t deg, const float* xc,
                                          const float* sn, uint8_t* s_el, float* stash) {
  constexpr int PL = RowShape<D>::PL;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t K = c.K, NC = K + 2u;
  const int base = lane * PL;
  __shared__ float s_nrm[32], s_dot[32], s_ca[32], s_cns[32], s_cna[32], s_cp[2], s_scale;
  __shared__ uint32_t s_reach;
  __shared__ int s_contrib2;
  for (uint32_t i = w; i < NC; i += NW) {  // |x_i| and a.x_i
    float sq = 0.f, dt = 0.f;
#pragma unroll
    for (int q = 0; q < PL; ++q) {
      if (base + q < D) {
        const float x = xc[i * D + base + q];
        sq += x * x;
        dt += xc[base + q] * x;
      }
    }
    sq = wave_sum(sq);
    dt = wave_sum(dt);
    if (lane == 0) { s_nrm[i] = sqrtf(sq); s_dot[i] = dt; }
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t reach = 0;
    const float L = other_loss(c.loss_kind, c.gamma, c.loss_p2, K, s_nrm, s_dot[1], s_dot + 2, s_ca, s_cp, s_cns,
                               s_cna, &reach);
    float div = 1.f;
    if (c.lw_kind == 1) div = logf(1.f + (float)deg);
    else if (c.lw_kind == 2) div = powf((float)deg, c.lw_exp);
    const float Lw = L / div;
    const int contrib = (Lw > 0.f) && !isnan(Lw);
    s_contrib2 = contrib;
    s_reach = contrib ? reach : 0u;
    s_scale = 1.f / div;
    c.loss[b] = contrib ? Lw : 0.f;
    if (c.pass_loss) c.pass_loss[st->anchor_base + b] = contrib ? Lw : 0.f;
  }
  __syncthreads();
  if (c.xout) {
    float* xo = c.xout + (size_t)b * NC * D;
    for (uint32_t t = tid; t < NC * D; t += blockDim.x) xo[t] = xc[t];
  }
  if (!c.train) return;
  const uint32_t reach = s_reach;
  const float scale = s_scale;
  for (uint32_t ci = w; ci < NC; ci += NW) {
    bool e = (reach >> ci) & 1u;
    float g[PL];
#pragma unroll
    for (int q = 0; q < PL; ++q) g[q] = 0.f;
    if (e) {
#pragma unroll
      for (int q = 0; q < PL; ++q) {
        if (base + q < D) {
          const float av = xc[base + q];
          float x;
          if (ci == 0u) {
            x = 0.f;
            for (uint32_t y = 0; y < NC; ++y) x += s_ca[y] * xc[y * D + base + q];
          } else if (ci == 1u) {
            x = s_cp[0] * av + s_cp[1] * xc[D + base + q];
          } else {
            x = s_cns[ci - 2u] * xc[ci * D + base + q] + s_cna[ci - 2u] * av;
          }
          g[q] = x * scale;
        }
      }
      bool bad = false;
#pragma unroll
      for (int q = 0; q < PL; ++q) bad |= (base + q < D) && !isfinite(g[q]);
      if (wave_any(bad)) e = false;
    }
    const float co = 1.0f / sn[ci];
    if (e) {
#pragma unroll
      for (int q = 0; q < PL; ++q) g[q] *= co;
      store_adj_row<D>(c, (size_t)b * NC + ci, lane, g);
      if (c.nbr_rows && ci == 1u) {
#pragma unroll
        for (int q = 0; q < PL; ++q)
          if (base + q < D) stash[base + q] = g[q];
      }
    }
    if (lane == 0) {
      c.elig[(size_t)b * NC + ci] = e ? 1 : 0;
      c.coef[(size_t)b * NC + ci] = co;
      s_el[ci] = e ? 1 : 0;
    }
  }
}

// 32-bit words of the float part of the forward kernel's LDS (xc, part, dist, sn),
// rounded up to an even count so the 64-bit entry array after it is aligned.
__host__ __device__ inline uint32_t forward_float_words(uint32_t NC, uint32_t K, uint32_t D) {
  const uint32_t w = (NC + NW) * D + (K + 1u) + NC;
  return (w + 1u) & ~1u;
}

// ------------------------------------------------------------------------------------
// k_forward: one 256-thread block per anchor.
//   gather + mean of h(v), ~h(v), h(u_1..K)     (model.rs:257-301, 342-416, 448-463)
//   margin loss (loss.rs:59-73) + loss weighting (mod.rs:190-200)
//   adjoints dL/dx of every construction, eligibility (finite, in the loss graph)
// Neighbour features of ~h(v) are split over the 4 waves (fixed-order combine);
// the 1+K single-node constructions are dealt round-robin to the waves.
// ------------------------------------------------------------------------------------
// WX multiplies the rows in flight of the wide gathers (more bytes in flight per wave,
// more VGPRs, so fewer waves); the partial sums take the rows in the same order for
// every WX, so the variants give identical bits and the tuner keeps the faster one.
template <int D, typename T, int WX>
__global__ __launch_bounds__(256) void k_forward(StepConfig c, StepState* st) {
  constexpr int PL = RowShape<D>::PL;
  constexpr int U = 8;
  constexpr int GU = PL <= 2 ? 16 : 8;  // rows in flight per wave on the staged path
  constexpr int WU = (WideShape<D, T>::RPW >= 4 ? 4 : 8) * WX;  // rows in flight per lane group (wide)
  const uint32_t b = blockIdx.x;
  if (b >= st->batch_n) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t K = c.K, NC = K + 2u;
  extern __shared__ float sm[];
  float* xc = sm;                 // NC*D
  float* part = xc + NC * D;      // NW*D
  float* dist = part + NW * D;    // K+1
  float* sn = dist + K + 1;       // NC occurrence counts
  __shared__ float s_scale, s_invA;
  __shared__ float s_r[32];  // 1/D_0, then invA/D_i for the active negatives (0 otherwise)
  __shared__ int s_contrib;
  __shared__ uint32_t s_active;
  __shared__ uint8_t s_el[32];
  const T* table = (const T*)c.table;
  // the feature lists of the example's entries, read before anything they wait on
  uint64_t x_fs = 0;
  uint32_t x_fn = 0;
  if ((uint32_t)tid < c.ent_cap) {
    x_fs = c.ex_fs[(size_t)b * c.ent_cap + tid];
    x_fn = c.ex_fn[(size_t)b * c.ent_cap + tid];
  }
  const uint32_t v = c.perm[st->anchor_base + b];
  const uint32_t cnt = c.nbr_cnt[b];
  if (cnt == 0u) {  // no neighbourhood: no example
    if (tid == 0) {
      c.loss[b] = 0.f;
      if (c.pass_loss) c.pass_loss[st->anchor_base + b] = 0.f;
    }
    if (tid < (int)NC) c.elig[(size_t)b * NC + tid] = 0;
    return;
  }
  const uint64_t vs = c.indptr[v];
  const uint32_t deg = (uint32_t)(c.indptr[v + 1] - vs);
  const uint32_t* pn = nbr_list(c, b, vs);
  const int base = lane * PL;
  const uint32_t apos = st->anchor_base + b, apass = st->pass;
  const uint32_t k0 = (uint32_t)c.seed, k1 = (uint32_t)(c.seed >> 32);
  // read now for the key emission at the end, so no global load waits there (a load
  // after the adjoint and key stores would wait for those stores too)
  const uint32_t kb = c.keyoff[b];
  const bool keys_fit = st->n_keys <= c.key_cap;

  __shared__ uint32_t s_wave[NW];
  __shared__ uint32_t s_np[NW];
  __shared__ float s_npw[NW];
  // Staged path: the example's occurrence list is copied to LDS first (the entries'
  // feature lists from k_sample, then the feature ids, one block-wide round trip each), so
  // every wave streams rows from a contiguous id range with GU loads in flight.  Entries: e = 0 (v), 1..cnt
  // (neighbours), cnt+1..cnt+K (negatives).
  const uint32_t E = 1u + cnt + K;
  const uint32_t Re = c.rcount[b];
  const bool staged = (E <= c.ent_cap) && (Re <= c.ids_cap);
  uint64_t* ent_fs = (uint64_t*)(sm + forward_float_words(NC, K, D));
  uint64_t* ent_mask = ent_fs + c.ent_cap;  // max_features selections
  uint32_t* ent_off = (uint32_t*)(ent_mask + c.sel_cap);
  float* ent_sc = (float*)(ent_off + c.ent_cap + 1);
  uint32_t* ids = (uint32_t*)(ent_sc + c.sel_cap);
  if (staged) {
    uint32_t ne = 0;
    if ((uint32_t)tid < E) {
      const uint64_t fs = x_fs;
      ne = x_fn;
      ent_fs[tid] = fs;
      if (c.mf_kind) {
        const FeatSel fsel = feature_select(c, ne, apos, apass, (uint32_t)tid, k0, k1);
        ent_mask[tid] = fsel.mask;
        ent_sc[tid] = fsel.scale;
        ne = (uint32_t)__popcll(fsel.mask);
      }
    }
    uint32_t tot;
    const uint32_t excl = block_exclusive_scan(ne, s_wave, &tot);
    if ((uint32_t)tid < E) ent_off[tid] = excl;
    if (tid == 0) ent_off[E] = tot;
    __syncthreads();
    constexpr int EU = 16;  // entries per wave on the entry-by-entry path
    if (!c.mf_kind && E <= (uint32_t)(NW * EU)) {
      // entry by entry: wave w loads the lists of entries w, w + NW, ..., one id per lane
      // (64 at a time), every first load in flight before the first store
      uint32_t f[EU];
#pragma unroll
      for (int u = 0; u < EU; ++u) {
        const uint32_t e = (uint32_t)w + (uint32_t)(u * NW);
        f[u] = 0u;
        if (e < E && (uint32_t)lane < ent_off[e + 1] - ent_off[e]) f[u] = c.fids[ent_fs[e] + (uint32_t)lane];
      }
#pragma unroll
      for (int u = 0; u < EU; ++u) {
        const uint32_t e = (uint32_t)w + (uint32_t)(u * NW);
        if (e < E) {
          const uint32_t a = ent_off[e], n = ent_off[e + 1] - a;
          if ((uint32_t)lane < n) ids[a + (uint32_t)lane] = f[u];
          for (uint32_t q = (uint32_t)lane + 64u; q < n; q += 64u) ids[a + q] = c.fids[ent_fs[e] + q];
        }
      }
    } else {
      // IDS_U feature ids per thread in flight: positions past tot repeat the last one
      constexpr uint32_t IDS_U = 8;
      for (uint32_t o0 = tid; o0 < tot; o0 += IDS_U * blockDim.x) {
        uint64_t src[IDS_U];
#pragma unroll
        for (uint32_t u = 0; u < IDS_U; ++u) {
          const uint32_t o = min(o0 + u * blockDim.x, tot - 1u);
          uint32_t lo = 0, hi = E;  // last entry e with ent_off[e] <= o
          while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (ent_off[mid] <= o) lo = mid; else hi = mid;
          }
          const uint32_t q = o - ent_off[lo];
          src[u] = ent_fs[lo] + (c.mf_kind ? nth_set_bit(ent_mask[lo], q) : q);
        }
        uint32_t f[IDS_U];
#pragma unroll
        for (uint32_t u = 0; u < IDS_U; ++u) f[u] = c.fids[src[u]];
#pragma unroll
        for (uint32_t u = 0; u < IDS_U; ++u)
          if (o0 + u * blockDim.x < tot) ids[o0 + u * blockDim.x] = f[u];
      }
    }
    for (uint32_t t = tid; t < ids_pad_entries(D, sizeof(T) == 2); t += blockDim.x) ids[tot + t] = 0u;
    __syncthreads();
    // ~h(v): the pooled neighbour occurrences split evenly over the waves
    const uint32_t p0 = ent_off[1], p1 = ent_off[1 + cnt];
    const uint32_t per = (p1 - p0 + NW - 1) / NW;
    const uint32_t a0 = min(p1, p0 + w * per), a1 = min(p1, a0 + per);
    if (c.nbr_rows) {
      // neighbour j's occurrences count w_j + s_j (s_j its max_features scale, else 1),
      // neighbours dealt round-robin
      float acc[PL];
#pragma unroll
      for (int q = 0; q < PL; ++q) acc[q] = 0.f;
      float nw = 0.f;
      for (uint32_t j = w; j < cnt; j += NW) {
        const float cj = nbr_weight(c, b, j, vs, deg) + (c.mf_kind ? ent_sc[1 + j] : 1.f);
        float t[PL];
#pragma unroll
        for (int q = 0; q < PL; ++q) t[q] = 0.f;
        gather_lds<D, GU>(table, ids, ent_off[1 + j], ent_off[2 + j], lane, t);
#pragma unroll
        for (int q = 0; q < PL; ++q) acc[q] += cj * t[q];
        nw += cj * (float)(ent_off[2 + j] - ent_off[1 + j]);
      }
#pragma unroll
      for (int q = 0; q < PL; ++q)
        if (base + q < D) part[w * D + base + q] = acc[q];
      if (lane == 0) s_npw[w] = nw;
    } else if constexpr (WideShape<D, T>::OK) {
      gather_lds_wide<D, WU, T>(table, ids, a0, a1, lane, part + w * D, 1.0f);
    } else {
      float acc[PL];
#pragma unroll
      for (int q = 0; q < PL; ++q) acc[q] = 0.f;
      gather_lds<D, GU>(table, ids, a0, a1, lane, acc);
#pragma unroll
      for (int q = 0; q < PL; ++q)
        if (base + q < D) part[w * D + base + q] = acc[q];
    }
    if (tid == 0) sn[1] = (float)(p1 - p0);
    // single-node constructions: i = 0 -> h(v) (c=0), i >= 1 -> h(u_i) (c=i+1)
    for (uint32_t i = w; i < K + 1u; i += NW) {
      const uint32_t e = i == 0u ? 0u : cnt + i;
      const uint32_t ci = i == 0u ? 0u : i + 1u;
      const float n = (float)(ent_off[e + 1] - ent_off[e]);
      if constexpr (WideShape<D, T>::OK) {
        gather_lds_wide<D, WU, T>(table, ids, ent_off[e], ent_off[e + 1], lane, xc + ci * D, 1.0f / n);
      } else {
        float a2[PL];
#pragma unroll
        for (int q = 0; q < PL; ++q) a2[q] = 0.f;
        gather_lds<D, GU>(table, ids, ent_off[e], ent_off[e + 1], lane, a2);
#pragma unroll
        for (int q = 0; q < PL; ++q)
          if (base + q < D) xc[ci * D + base + q] = a2[q] / n;
      }
      if (lane == 0) sn[ci] = n;
    }
    __syncthreads();
  } else {
    // Fallback for examples larger than the LDS stage: neighbour nodes dealt round-robin.
    float acc[PL];
#pragma unroll
    for (int q = 0; q < PL; ++q) acc[q] = 0.f;
    uint32_t np = 0;
    float nw = 0.f;
    for (uint32_t j = w; j < cnt; j += NW) {
      const uint32_t u = pn[j];
      const uint64_t fs = c.findptr[u], fe = c.findptr[u + 1];
      if (c.nbr_rows) {
        float cj = nbr_weight(c, b, j, vs, deg);
        float t[PL];
#pragma unroll
        for (int q = 0; q < PL; ++q) t[q] = 0.f;
        uint32_t r = (uint32_t)(fe - fs);
        if (c.mf_kind) {
          const FeatSel fsel = feature_select(c, r, apos, apass, 1u + j, k0, k1);
          gather_masked<D, U>(table, c.fids, fs, fsel.mask, lane, t);
          r = (uint32_t)__popcll(fsel.mask);
          cj += fsel.scale;
        } else {
          gather_rows<D, U>(table, c.fids, fs, fe, lane, t);
          cj += 1.f;
        }
#pragma unroll
        for (int q = 0; q < PL; ++q) acc[q] += cj * t[q];
        nw += cj * (float)r;
        np += r;
      } else {
        gather_rows<D, U>(table, c.fids, fs, fe, lane, acc);
        np += (uint32_t)(fe - fs);
      }
    }
    if (lane == 0) s_npw[w] = nw;
#pragma unroll
    for (int q = 0; q < PL; ++q)
      if (base + q < D) part[w * D + base + q] = acc[q];
    if (lane == 0) s_np[w] = np;
    for (uint32_t i = w; i < K + 1u; i += NW) {
      const uint32_t node = i == 0u ? v : c.neg[(size_t)b * K + i - 1u];
      const uint32_t ci = i == 0u ? 0u : i + 1u;
      float a2[PL];
#pragma unroll
      for (int q = 0; q < PL; ++q) a2[q] = 0.f;
      const uint64_t fs = c.findptr[node], fe = c.findptr[node + 1];
      float n = (float)(fe - fs);
      if (c.mf_kind) {
        const FeatSel fsel = feature_select(c, (uint32_t)(fe - fs), apos, apass, i == 0u ? 0u : cnt + i, k0, k1);
        gather_masked<D, U>(table, c.fids, fs, fsel.mask, lane, a2);
        n = (float)__popcll(fsel.mask);
      } else {
        gather_rows<D, U>(table, c.fids, fs, fe, lane, a2);
      }
#pragma unroll
      for (int q = 0; q < PL; ++q)
        if (base + q < D) xc[ci * D + base + q] = a2[q] / n;
      if (lane == 0) sn[ci] = n;
    }
    __syncthreads();
    if (tid == 0) sn[1] = (float)(s_np[0] + s_np[1] + s_np[2] + s_np[3]);
    __syncthreads();
  }
  if (c.nbr_rows) {
    if (tid == 0) sn[1] = ((s_npw[0] + s_npw[1]) + s_npw[2]) + s_npw[3];
    __syncthreads();
  }
  {
    const float ntot = sn[1];
    for (int t = tid; t < D; t += blockDim.x)
      xc[D + t] = (((part[t] + part[D + t]) + part[2 * D + t]) + part[3 * D + t]) / ntot;
  }
  __syncthreads();
  if (c.loss_kind != 0 && c.loss_kind != 5) {
    forward_other_loss<D>(c, st, b, deg, xc, sn, s_el, part);
    if (!c.train) return;
  } else {
  // distances: D0 = |p - a|, D_i = |p - n_i|
  for (uint32_t i = w; i < K + 1u; i += NW) {
    const float* other = (i == 0u) ? xc : xc + (i + 1u) * D;
    float sq = 0.f;
#pragma unroll
    for (int q = 0; q < PL; ++q) {
      if (base + q < D) {
        const float d = xc[D + base + q] - other[base + q];
        sq += d * d;
      }
    }
    sq = wave_sum_dpp(sq);
    if (lane == 0) dist[i] = sqrtf(sq);
  }
  __syncthreads();
  if (w == 0) {
    // lane i - 1 tests negative i (K <= 29) and keeps its adjoint coefficient, so the
    // adjoints below multiply instead of dividing
    const float d1 = c.gamma + dist[0];
    const uint32_t ii = (uint32_t)lane + 1u;
    const bool on = ii <= K && d1 - dist[ii] > 0.f;
    const uint32_t act_w = (uint32_t)__ballot(on);
    const uint32_t nA_w = (uint32_t)__popc(act_w);
    if (ii <= K) s_r[ii] = on ? (1.f / (float)nA_w) / dist[ii] : 0.f;
    if (lane == 0) s_r[0] = 1.f / dist[0];
  }
  if (tid == 0) {
    const float d1 = c.gamma + dist[0];
    float sum = 0.f;
    uint32_t nA = 0, act = 0;
    for (uint32_t i = 1; i <= K; ++i) {
      const float l = d1 - dist[i];
      if (l > 0.f) {
        sum += l;
        ++nA;
        act |= 1u << (i - 1u);
      }
    }
    const float L = nA ? sum / (float)nA : 0.f;
    float div = 1.f;
    if (c.lw_kind == 1) div = logf(1.f + (float)deg);
    else if (c.lw_kind == 2) div = powf((float)deg, c.lw_exp);
    const float Lw = L / div;
    const int contrib = (Lw > 0.f) && !isnan(Lw);
    s_contrib = contrib;
    s_active = act;
    s_invA = nA ? 1.f / (float)nA : 0.f;
    s_scale = 1.f / div;
    c.loss[b] = contrib ? Lw : 0.f;
    if (c.pass_loss) c.pass_loss[apos] = contrib ? Lw : 0.f;
  }
  __syncthreads();
  if (c.xout) {
    float* xo = c.xout + (size_t)b * NC * D;
    for (uint32_t t = tid; t < NC * D; t += blockDim.x) xo[t] = xc[t];
  }
  if (!c.train) return;
  const int contrib = s_contrib;
  const uint32_t act = s_active;
  const float invA = s_invA, scale = s_scale, D0 = dist[0];
  for (uint32_t ci = w; ci < NC; ci += NW) {
    bool e = false;
    float g[PL];
#pragma unroll
    for (int q = 0; q < PL; ++q) g[q] = 0.f;
    if (contrib) {
      if (ci == 0u) {
        e = true;
        const float r0 = s_r[0];
#pragma unroll
        for (int q = 0; q < PL; ++q)
          if (base + q < D) g[q] = (xc[base + q] - xc[D + base + q]) * r0 * scale;
      } else if (ci == 1u) {
        e = true;
        const float r0 = s_r[0];
        float p[PL];
#pragma unroll
        for (int q = 0; q < PL; ++q) {
          p[q] = base + q < D ? xc[D + base + q] : 0.f;
          g[q] = base + q < D ? (p[q] - xc[base + q]) * r0 : 0.f;
        }
        for (uint32_t m = act; m != 0u; m &= m - 1u) {  // the active negatives, in order
          const uint32_t i = (uint32_t)__ffs(m);        // bit i - 1 set: negative i
          const float ri = s_r[i];
#pragma unroll
          for (int q = 0; q < PL; ++q)
            if (base + q < D) g[q] -= (p[q] - xc[(i + 1u) * D + base + q]) * ri;
        }
#pragma unroll
        for (int q = 0; q < PL; ++q) g[q] *= scale;
      } else {
        const uint32_t i = ci - 1u;
        if (act & (1u << (i - 1u))) {
          e = true;
          const float ri = s_r[i];
#pragma unroll
          for (int q = 0; q < PL; ++q)
            if (base + q < D) g[q] = (xc[D + base + q] - xc[ci * D + base + q]) * ri * scale;
        }
      }
      if (e) {
        bool bad = false;
#pragma unroll
        for (int q = 0; q < PL; ++q) bad |= (base + q < D) && !isfinite(g[q]);
        if (wave_any(bad)) e = false;
      }
    }
    // stored pre-scaled by the occurrence coefficient 1/n_c (count/total of the mean),
    // so a feature's gradient through this construction is (multiplicity) x adj
    const float co = 1.0f / sn[ci];
    if (e) {
#pragma unroll
      for (int q = 0; q < PL; ++q) g[q] *= co;
      store_adj_row<D>(c, (size_t)b * NC + ci, lane, g);
      if (c.nbr_rows && ci == 1u) {
#pragma unroll
        for (int q = 0; q < PL; ++q)
          if (base + q < D) part[base + q] = g[q];
      }
    }
    if (lane == 0) {
      c.elig[(size_t)b * NC + ci] = e ? 1 : 0;
      c.coef[(size_t)b * NC + ci] = co;
      s_el[ci] = e ? 1 : 0;
    }
  }
  }
  if (c.nbr_rows) {
    // neighbour j's adjoint row is (dL/dp / sum of counts) * (w_j + s_j)
    __syncthreads();
    if (s_el[1]) {
      for (uint32_t j = w; j < cnt; j += NW) {
        float sj = 1.f;
        if (c.mf_kind) {
          if (staged) {
            sj = ent_sc[1 + j];
          } else {
            const uint32_t u = pn[j];
            sj = feature_select(c, (uint32_t)(c.findptr[u + 1] - c.findptr[u]), apos, apass, 1u + j, k0, k1).scale;
          }
        }
        const float cj = nbr_weight(c, b, j, vs, deg) + sj;
        float g[PL];
#pragma unroll
        for (int q = 0; q < PL; ++q) g[q] = (base + q < D) ? part[base + q] * cj : 0.f;
        store_adj_row<D>(c, (size_t)nbr_row(c, b, j), lane, g);
      }
    }
  }
  // Key emission for staged examples (k_emit handles the rest): one (feature, slot) key
  // per occurrence, sentinel unless the construction is eligible and, under first_leaf,
  // the feature's first eligible construction (mod.rs:347-352, 366-367).  The
  // first-leaf hash lives in the construction / partial-sum region of the LDS, dead by
  // now; examples with more occurrences than fuse_max (5/8 of its entries) are left to
  // k_emit.
  if (!(staged && c.fuse_emit)) return;
  if (c.grad_mode == 0 && Re > c.fuse_max) return;
  if (!keys_fit) return;
  __syncthreads();
  const uint32_t tot = ent_off[E];
  const bool first_leaf = c.grad_mode == 0;
  uint32_t* hsh = reinterpret_cast<uint32_t*>(sm);
  const uint32_t hmask = c.fuse_hash - 1u;
  // the entry of each occurrence: one byte each (E <= 256), written entry by entry after
  // the hash where the region holds them, else a search of the entry offsets
  uint8_t* oent = reinterpret_cast<uint8_t*>(hsh + c.fuse_hash);
  const bool ent_bytes = (size_t)c.fuse_hash * 4u + tot <= (size_t)forward_float_words(NC, K, D) * 4u;
  if (ent_bytes)
    for (uint32_t e = (uint32_t)w; e < E; e += NW)
      for (uint32_t o = ent_off[e] + (uint32_t)lane; o < ent_off[e + 1]; o += 64u) oent[o] = (uint8_t)e;
  auto entry_of = [&](uint32_t o) -> uint32_t {
    if (ent_bytes) return oent[o];
    uint32_t lo = 0, hi = E;  // last entry e with ent_off[e] <= o
    while (hi - lo > 1u) {
      const uint32_t mid = (lo + hi) >> 1;
      if (ent_off[mid] <= o) lo = mid; else hi = mid;
    }
    return lo;
  };
  auto cons_of_entry = [&](uint32_t e) -> uint32_t { return e == 0u ? 0u : (e <= cnt ? 1u : 2u + (e - 1u - cnt)); };
  if (first_leaf) {
    for (uint32_t i = tid; i < c.fuse_hash; i += blockDim.x) hsh[i] = HASH_EMPTY;
    __syncthreads();
    for (uint32_t o = tid; o < tot; o += blockDim.x) {
      const uint32_t ci = cons_of_entry(entry_of(o));
      if (s_el[ci]) hash_insert_min(hsh, hmask, ids[o], ci);
    }
  }
  __syncthreads();
  // two loops, so that the common one holds no global load (nbr_row reads one): its
  // stores are then never waited for inside the loop
  const uint32_t sbatch = (b / c.rbatch) << kSlotBatchShift;
  if (c.nbr_rows) {
    for (uint32_t o = tid; o < tot; o += blockDim.x) {
      const uint32_t e = entry_of(o);
      const uint32_t ci = cons_of_entry(e);
      const uint32_t f = ids[o];
      uint32_t key = c.sentinel;
      if (s_el[ci] && (!first_leaf || hash_lookup(hsh, hmask, f) == ci)) key = key_of_row(c, f);
      c.keys[0][kb + o] = key;
      c.slots[0][kb + o] = sbatch | (ci == 1u ? nbr_row(c, b, e - 1u) : b * NC + ci);
    }
  } else {
    for (uint32_t o = tid; o < tot; o += blockDim.x) {
      const uint32_t ci = cons_of_entry(entry_of(o));
      const uint32_t f = ids[o];
      uint32_t key = c.sentinel;
      if (s_el[ci] && (!first_leaf || hash_lookup(hsh, hmask, f) == ci)) key = key_of_row(c, f);
      c.keys[0][kb + o] = key;
      c.slots[0][kb + o] = sbatch | (b * NC + ci);
    }
  }
}

// ------------------------------------------------------------------------------------
// k_emit: one block per anchor.  Writes one (feature, slot) key per feature occurrence,
// in construction order a, p, n_1..n_K, at keyoff[b].  Occurrences of ineligible
// constructions, and (first_leaf) occurrences outside the feature's first eligible
// construction (mod.rs:347-352, 366-367), get the sentinel key.  Attention model: the
// slot is the occurrence's key position (its gradient row), and in h(v) / h(u_k) only
// the first occurrence of a feature carries one.
// ------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_emit(StepConfig c, StepState* st) {
  const uint32_t b = blockIdx.x;
  if (b >= st->batch_n) return;
  if (st->n_keys > c.key_cap) return;
  const uint32_t cnt = c.nbr_cnt[b];
  if (cnt == 0u) return;
  const int tid = threadIdx.x;
  const uint32_t K = c.K, NC = K + 2u;
  if (c.fuse_emit && 1u + cnt + K <= c.ent_cap && c.rcount[b] <= c.ids_cap &&
      (c.grad_mode != 0 || c.rcount[b] <= c.fuse_max))
    return;  // emitted by k_forward
  extern __shared__ uint32_t hsh[];
  __shared__ uint32_t s_wave[NW];
  __shared__ uint8_t s_el[32];
  __shared__ uint32_t s_off[257];
  __shared__ uint32_t s_node[256];
  __shared__ uint8_t s_ci[256];
  __shared__ uint64_t s_mask[256];  // max_features selections
  const uint32_t H = c.hash_entries, hmask = H - 1u;
  const uint32_t apos = st->anchor_base + b, apass = st->pass;
  const uint32_t k0 = (uint32_t)c.seed, k1 = (uint32_t)(c.seed >> 32);
  const bool first_leaf = c.grad_mode == 0;
  const bool att = c.model == 1;
  const uint32_t v = c.perm[st->anchor_base + b];
  const uint32_t* pn = nbr_list(c, b, c.indptr[v]);
  if (tid < (int)NC) s_el[tid] = c.elig[(size_t)b * NC + tid];
  if (first_leaf)
    for (uint32_t i = tid; i < H; i += blockDim.x) hsh[i] = HASH_EMPTY;
  __syncthreads();
  const uint32_t E = 1u + cnt + K;
  const uint32_t kbase = c.keyoff[b];
  const uint32_t slot0 = b * NC;
  for (int phase = first_leaf ? 0 : 1; phase < 2; ++phase) {
    uint32_t running = 0;
    for (uint32_t cs = 0; cs < E; cs += blockDim.x) {
      const uint32_t e = cs + tid;
      uint32_t ne = 0, node = 0;
      uint8_t ci = 0;
      if (e < E) {
        if (e == 0u) { node = v; ci = 0; }
        else if (e <= cnt) { node = pn[e - 1u]; ci = 1; }
        else { node = c.neg[(size_t)b * K + (e - cnt - 1u)]; ci = (uint8_t)(2u + (e - cnt - 1u)); }
        ne = (uint32_t)(c.findptr[node + 1] - c.findptr[node]);
        if (c.mf_kind) {
          const uint64_t m = feature_select(c, ne, apos, apass, e, k0, k1).mask;
          s_mask[tid] = m;
          ne = (uint32_t)__popcll(m);
        }
      }
      uint32_t total;
      const uint32_t excl = block_exclusive_scan(ne, s_wave, &total);
      s_off[tid] = excl;
      s_node[tid] = node;
      s_ci[tid] = ci;
      if (tid == 0) s_off[256] = total;
      __syncthreads();
      const uint32_t m = (E - cs) < blockDim.x ? (E - cs) : blockDim.x;
      for (uint32_t o = tid; o < total; o += blockDim.x) {
        // entry j with s_off[j] <= o < s_off[j+1] (entries beyond m have zero length)
        uint32_t lo = 0, hi = m;  // find last j with s_off[j] <= o
        while (hi - lo > 1u) {
          const uint32_t mid = (lo + hi) >> 1;
          if (s_off[mid] <= o) lo = mid; else hi = mid;
        }
        const uint32_t j = lo;
        const uint32_t nd = s_node[j];
        const uint32_t cj = s_ci[j];
        const uint64_t nfs = c.findptr[nd];
        const uint32_t qpos = o - s_off[j];
        const uint32_t f = c.fids[nfs + (c.mf_kind ? nth_set_bit(s_mask[j], qpos) : qpos)];
        const bool ok = s_el[cj] != 0;
        if (phase == 0) {
          if (ok) hash_insert_min(hsh, hmask, f, cj);
        } else {
          bool emit = ok && (!first_leaf || hash_lookup(hsh, hmask, f) == cj);
          const uint32_t pos = kbase + running + o;
          if (att && emit && cj != 1u && c.att_win == 0xFFFFFFFFu) {
            // h(v), h(u_k) under Full attention: one attention item (and gradient row)
            // per distinct feature, at the feature's first position in the node's list
            // (Sliding attention keeps every listed occurrence)
            for (uint32_t t = 0; t < qpos && emit; ++t) emit = c.fids[nfs + t] != f;
          }
          c.keys[0][pos] = emit ? key_of_row(c, f) : c.sentinel;
          // mean model: the slot of the construction's adjoint; attention: the
          // occurrence's own gradient row
          const uint32_t row = att ? pos : ((c.nbr_rows && cj == 1u) ? nbr_row(c, b, cs + j - 1u) : slot0 + cj);
          c.slots[0][pos] = ((b / c.rbatch) << kSlotBatchShift) | row;
        }
      }
      running += total;
      __syncthreads();
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------
// LSD radix sort of (key, slot) pairs with the element count read from device memory.
// RB-bit digits (RB <= 11).  Each block owns a contiguous range, processed in tiles of
// 256 x RS_ITEMS elements loaded up front.  Stable: within a round of 256 elements the
// ranks follow thread order (wave-level match via RB ballots); rounds, tiles and blocks
// are ordered through per-block digit bases.
// ------------------------------------------------------------------------------------
static constexpr int RS_ITEMS = 8;

__device__ inline void block_range(uint32_t n, int G, uint32_t blk, uint32_t* lo, uint32_t* hi) {
  const uint64_t per = (((uint64_t)n + G - 1) / G + 255u) / 256u * 256u;
  const uint64_t l = (uint64_t)blk * per;
  const uint64_t h = l + per;
  *lo = (uint32_t)(l < n ? l : n);
  *hi = (uint32_t)(h < n ? h : n);
}

// Lanes of the wave holding the same digit as this lane (active lanes only).
template <int RB>
__device__ inline uint64_t digit_peers(bool act, uint32_t d) {
  uint64_t peers = __ballot(act);
#pragma unroll
  for (int bit = 0; bit < RB; ++bit) {
    const bool set = (d >> bit) & 1u;
    const uint64_t bb = __ballot(act && set);
    peers &= set ? bb : ~bb;
  }
  return act ? peers : 0ull;
}

// Pass 0 reads the emitted keys (st->n_keys, sentinels included) and drops the sentinels;
// later passes read the st->n_real compacted keys.
template <int RB>
__global__ __launch_bounds__(256) void k_rs_hist(const uint32_t* __restrict__ kin, StepState* st, uint32_t* ghist,
                                                 int pass, int shift, int G, uint32_t sent) {
  constexpr uint32_t BINS = 1u << RB, MASK = BINS - 1u;
  __shared__ uint32_t h[BINS];
  const int tid = threadIdx.x, lane = tid & 63;
  uint32_t lo, hi;
  block_range(pass == 0 ? st->n_keys : st->n_real, G, blockIdx.x, &lo, &hi);
  for (uint32_t d = tid; d < BINS; d += 256u) h[d] = 0u;
  __syncthreads();
  for (uint32_t t0 = lo; t0 < hi; t0 += 256u * RS_ITEMS) {
    uint32_t k[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
      const uint32_t i = t0 + r * 256u + tid;
      k[r] = i < hi ? kin[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
      const uint32_t i = t0 + r * 256u + tid;
      const bool act = i < hi && k[r] != sent;
      const uint32_t d = (k[r] >> shift) & MASK;
      // one digit for the whole wave (runs of a hot row): one add of the count; else an
      // LDS add per key (the LDS serialises only the lanes that share an address)
      const uint64_t am = __ballot(act);
      const uint32_t d0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)d);
      if (__ballot(act && d == d0) == am) {
        if (am && lane == __ffsll((unsigned long long)am) - 1) atomicAdd(&h[d0], (uint32_t)__popcll(am));
      } else if (act) {
        atomicAdd(&h[d], 1u);
      }
    }
  }
  __syncthreads();
  for (uint32_t d = tid; d < BINS; d += 256u) {
    ghist[(size_t)blockIdx.x * BINS + d] = h[d];  // block-major: coalesced
    if (h[d]) atomicAdd(&st->dtot[pass][d], h[d]);
  }
}

// Turns the block-major counts into scatter bases: base(blk, d) = sum_{d' < d} total(d')
// + sum_{b < blk} count(b, d).  One wave per digit; lane l owns G/64 consecutive source
// blocks (their loads are independent), then a wave scan.
template <int RB>
__global__ __launch_bounds__(256) void k_rs_scan(uint32_t* ghist, StepState* st, int pass, int G) {
  constexpr uint32_t BINS = 1u << RB;
  const int lane = threadIdx.x & 63;
  const uint32_t d = blockIdx.x * NW + (threadIdx.x >> 6);
  if (d >= BINS) return;
  uint32_t pre = 0;
  for (uint32_t q = lane; q < d; q += 64u) pre += st->dtot[pass][q];
  pre = wave_sum_u32(pre);
  if (d == BINS - 1u && lane == 0) st->n_real = pre + st->dtot[pass][d];
  const uint32_t per = (uint32_t)G / 64u;  // G is a multiple of 256
  const uint32_t b0 = (uint32_t)lane * per;
  uint32_t loc = 0;
  for (uint32_t i = 0; i < per; ++i) loc += ghist[(size_t)(b0 + i) * BINS + d];
  uint32_t inc = loc;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(inc, off, EP_WAVE);
    if (lane >= off) inc += y;
  }
  uint32_t run = pre + inc - loc;
  for (uint32_t i = 0; i < per; ++i) {
    const size_t idx = (size_t)(b0 + i) * BINS + d;
    const uint32_t x = ghist[idx];
    ghist[idx] = run;
    run += x;
  }
}

// Scatter of one digit pass, a tile of 256 x RS_SCATTER_ITEMS elements at a time: wave w
// takes the tile's w-th quarter (coalesced loads), ranks its elements by digit in
// element order with its own LDS counters, the tile's digit starts come from a block
// scan, the tile is reordered by digit in LDS, and the threads then write it out in
// order, so each digit's elements leave as one contiguous run instead of one scattered
// store each.  Stable: the tile order is (wave, item, lane) = element order, and tiles
// and blocks follow each other through the per-block digit bases.
static constexpr int RS_SCATTER_ITEMS = 16;

template <int RB>
__global__ __launch_bounds__(256) void k_rs_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                    StepState* st, const uint32_t* __restrict__ ghist, int pass,
                                                    int shift, int G, uint32_t sent) {
  constexpr uint32_t BINS = 1u << RB, MASK = BINS - 1u;
  constexpr int IT = RS_SCATTER_ITEMS;
  constexpr uint32_t TILE = 256u * IT;
  constexpr uint32_t WSPAN = 64u * IT;                          // elements of one wave in a tile
  constexpr uint32_t DPT = BINS >= 256u ? BINS / 256u : 1u;     // digits per thread in the scans
  __shared__ uint32_t run[BINS];       // output position of the block's next element of each digit
  __shared__ uint32_t wcnt[NW][BINS];  // per-wave digit counts, then their prefix over the waves
  __shared__ uint32_t toff[BINS];      // start of each digit in the reordered tile
  __shared__ uint32_t s_wave[NW];
  __shared__ uint32_t lk[TILE], lv[TILE];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  uint32_t lo, hi;
  block_range(pass == 0 ? st->n_keys : st->n_real, G, blockIdx.x, &lo, &hi);
  for (uint32_t d = tid; d < BINS; d += 256u) {
    run[d] = ghist[(size_t)blockIdx.x * BINS + d];
#pragma unroll
    for (int q = 0; q < NW; ++q) wcnt[q][d] = 0u;
  }
  __syncthreads();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (uint32_t t0 = lo; t0 < hi; t0 += TILE) {
    uint32_t k[IT], val[IT], rk[IT];
#pragma unroll
    for (int q = 0; q < IT; ++q) {
      const uint32_t i = t0 + (uint32_t)w * WSPAN + (uint32_t)q * 64u + (uint32_t)lane;
      k[q] = i < hi ? kin[i] : sent;  // sent: inactive (pass 0 drops the sentinels)
      val[q] = i < hi ? vin[i] : 0u;
    }
    // ranks within the wave's quarter: count of earlier elements with the same digit
#pragma unroll
    for (int q = 0; q < IT; ++q) {
      const bool act = k[q] != sent;
      const uint32_t d = (k[q] >> shift) & MASK;
      // the lanes sharing this lane's digit: all active lanes when the wave holds one digit
      // (runs of a hot row), else matched bit by bit
      const uint64_t am = __ballot(act);
      const uint32_t d0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)d);
      const uint64_t peers = __ballot(act && d == d0) == am ? (act ? am : 0ull) : digit_peers<RB>(act, d);
      const uint32_t before = act ? wcnt[w][d] : 0u;
      rk[q] = before + (uint32_t)__popcll(peers & lt);
      if (act && (peers >> lane) == 1ull) wcnt[w][d] = before + (uint32_t)__popcll(peers);  // last lane of the group
      __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit: prefix of the wave counts (in place) and the tile count; tile starts by a
    // block scan over the digits (thread tid owns digits [tid*DPT, tid*DPT + DPT))
    uint32_t tt[DPT];
    uint32_t tsum = 0u;
#pragma unroll
    for (uint32_t j = 0; j < DPT; ++j) {
      const uint32_t d = (uint32_t)tid * DPT + j;
      uint32_t acc = 0u;
      if (d < BINS) {
#pragma unroll
        for (int q = 0; q < NW; ++q) {
          const uint32_t x = wcnt[q][d];
          wcnt[q][d] = acc;
          acc += x;
        }
      }
      tt[j] = acc;
      tsum += acc;
    }
    uint32_t ntile = 0u;
    uint32_t excl = block_exclusive_scan(tsum, s_wave, &ntile);
#pragma unroll
    for (uint32_t j = 0; j < DPT; ++j) {
      const uint32_t d = (uint32_t)tid * DPT + j;
      if (d < BINS) toff[d] = excl;
      excl += tt[j];
    }
    __syncthreads();
    // reorder the tile by digit in LDS
#pragma unroll
    for (int q = 0; q < IT; ++q) {
      if (k[q] != sent) {
        const uint32_t d = (k[q] >> shift) & MASK;
        const uint32_t pos = toff[d] + wcnt[w][d] + rk[q];
        lk[pos] = k[q];
        lv[pos] = val[q];
      }
    }
    __syncthreads();
    // write it out in order: consecutive threads, consecutive positions of a digit's run
    for (uint32_t j = (uint32_t)tid; j < ntile; j += 256u) {
      const uint32_t kk = lk[j];
      const uint32_t d = (kk >> shift) & MASK;
      const uint32_t o = run[d] + (j - toff[d]);
      kout[o] = kk;
      vout[o] = lv[j];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < DPT; ++j) {
      const uint32_t d = (uint32_t)tid * DPT + j;
      if (d < BINS) {
        run[d] += tt[j];
#pragma unroll
        for (int q = 0; q < NW; ++q) wcnt[q][d] = 0u;
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------
// Sparse lazy Adam (optimizer.rs:43-81): skip the row if its gradient has any NaN
// (line 56); m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
// theta -= lr (m/bc1) / (sqrt(v/bc2) + eps).  bf16 modes: theta is taken to FP32 (from
// the persistent master when there is one, else from the bf16 row itself), updated, and
// written back to the bf16 table with stochastic rounding (hash bits keyed by row,
// column and step).
// The bias corrections enter as the step's reciprocals ibc = 1/bc (computed once per
// kernel), the square root and the quotient through the hardware v_sqrt_f32 /
// v_rcp_f32 (1 ulp each): a few ulp from the IEEE divisions of the formula, at a
// quarter of their instructions (the update applies one Adam step per batch of the
// group to every value of every touched row).
// ------------------------------------------------------------------------------------
__device__ inline void adam_math(float& th, float& m, float& v, float g, float lr, float ibc1, float ibc2) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  m = b1 * m + (1.0f - b1) * g;
  v = b2 * v + (1.0f - b2) * (g * g);
  const float den = __builtin_amdgcn_sqrtf(v * ibc2) + eps;
  th -= lr * (m * ibc1) * __builtin_amdgcn_rcpf(den);
}

// Per-step scalars of the update, read once per kernel (before any store, so the
// compiler keeps them in registers instead of reloading them after each store).
struct StepScalars {
  float lr0;         // learning rate of the group's first batch
  float ibc1, ibc2;  // reciprocals of the Adam bias corrections 1 - beta^t
  uint32_t sr_step;  // step number keying the stochastic-rounding bits
  uint32_t step0;    // step number of the group's first batch (noise key)
  float sd0;         // noise sd of the group's first batch
};
__device__ inline StepScalars step_scalars(const StepState* st) {
  return StepScalars{st->lr[0], 1.0f / st->bc1, 1.0f / st->bc2, (uint32_t)st->next_step, (uint32_t)st->step0,
                     st->noise_sd[0]};
}

// 16 stochastic-rounding bits for column 4*col4 + q of row f at this step.
__device__ inline uint32_t sr_bits(const StepConfig& c, const StepScalars& sv, uint32_t f, uint32_t col4, int q) {
  const uint32_t r0 = sr_key(f, col4, sv.sr_step, (uint32_t)c.seed);
  const uint32_t w = (q >> 1) == 0 ? r0 : hash32(r0 ^ 0x27D4EB2Fu);
  return (q & 1) ? (w >> 16) : (w & 0xffffu);
}

// Gaussian gradient noise (mod.rs:238-247): N(0,1) for columns 4*col4 .. 4*col4+3 of row
// f at step `step`, two Box-Muller pairs from one Philox4x32 block keyed by (row, column
// block, step).  The reference draws from XorShift seeded with seed + batch + feature;
// the stream differs, the distribution (independent standard normals) is the same.
__device__ inline void noise4(const StepConfig& c, uint32_t f, uint32_t col4, uint32_t step, float (&z)[4]) {
  const u32x4 r = philox4x32(u32x4{f, col4, step, RNG_NOISE}, (uint32_t)c.seed, (uint32_t)(c.seed >> 32));
  const float two_pi = 6.28318530717958647692f;
  const float ra = sqrtf(-2.0f * logf(((float)(r.x >> 8) + 1.0f) * (1.0f / 16777216.0f)));
  const float rb = sqrtf(-2.0f * logf(((float)(r.z >> 8) + 1.0f) * (1.0f / 16777216.0f)));
  const float ta = two_pi * ((float)(r.y >> 8) * (1.0f / 16777216.0f));
  const float tb = two_pi * ((float)(r.w >> 8) * (1.0f / 16777216.0f));
  z[0] = ra * cosf(ta);
  z[1] = ra * sinf(ta);
  z[2] = rb * cosf(tb);
  z[3] = rb * sinf(tb);
}

// Adds sd * N(0,1) to a whole-wave slice (lane owns PL consecutive columns).
template <int D>
__device__ inline void add_noise_slice(const StepConfig& c, uint32_t f, uint32_t step, float sd, int lane,
                                       float (&g)[RowShape<D>::PL]) {
  constexpr int PL = RowShape<D>::PL;
  const int base = lane * PL;
#pragma unroll
  for (int q = 0; q < PL; ++q) {
    const int col = base + q;
    if (col < D && (q == 0 || (col & 3) == 0)) {
      float z[4];
      noise4(c, f, (uint32_t)col >> 2, step, z);
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (q + k < PL && ((col + k) >> 2) == (col >> 2) && col + k < D) g[q + k] += sd * z[(col + k) & 3];
    }
  }
}

// Whole-wave row layout (lane owns RowShape<D>::PL values).
template <int D>
__device__ inline void adam_row(const StepConfig& c, const StepScalars& sv, uint32_t f,
                                const float (&g)[RowShape<D>::PL], int lane) {
  constexpr int PL = RowShape<D>::PL;
  const int base = lane * PL;
  bool bad = false;
#pragma unroll
  for (int q = 0; q < PL; ++q) bad |= (base + q < D) && isnan(g[q]);
  if (wave_any(bad)) return;
  const size_t off = (size_t)f * D;
  const size_t ooff = (size_t)f * c.ostride;  // optimizer record [master |] m | v
  float m[PL], v[PL], th[PL];
  load_slice<D>(c.mom + ooff, lane, m);
  load_slice<D>(c.var + ooff, lane, v);
  if (c.master) load_slice<D>(c.master + ooff, lane, th);
  else if (c.bf16) load_slice<D>((const uint16_t*)c.table + off, lane, th);
  else load_slice<D>((const float*)c.table + off, lane, th);
#pragma unroll
  for (int q = 0; q < PL; ++q) adam_math(th[q], m[q], v[q], g[q], sv.lr0, sv.ibc1, sv.ibc2);
  store_slice<D>(c.mom + ooff, lane, m);
  store_slice<D>(c.var + ooff, lane, v);
  if (c.bf16) {
    if (c.master) store_slice<D>(c.master + ooff, lane, th);
    uint16_t* row = (uint16_t*)c.table + off;
#pragma unroll
    for (int q = 0; q < PL; ++q)
      if (base + q < D) row[base + q] = f32_to_bf16_sr(th[q], sr_bits(c, sv, f, (uint32_t)((base + q) >> 2), (base + q) & 3));
  } else {
    store_slice<D>((float*)c.table + off, lane, th);
  }
}

// ------------------------------------------------------------------------------------
// Batch merge + update over the sorted keys (mod.rs:223-233 merge, then the update of
// each batch of the group in batch order).  The sorted keys form rows (same row id)
// made of runs (same row and batch).
//   k_seg_plan    one wave per 64-entry chunk; owns the rows that start in it.  Rows of
//                 at most kSegShort entries become short-row records (first entry, row,
//                 run-start mask, run batches); longer rows, and a last row running past
//                 the 128-entry window, become long rows cut into pieces of at most
//                 kPieceEntries entries of one run.
//   k_seg_update  short rows, balanced over lane groups: theta/m/v loaded once, each run
//                 reduced (adjoint rows in entry order) and applied as one Adam step
//                 with its batch's learning rate, the row stored once.
//   k_pieces      partial sums of the pieces.
//   k_long_combine  one wave per long row: pieces added in order, one Adam step per run.
// grad_out != 0 writes each run's merged gradient to G/ufeat (ufeat = row | batch << 27)
// instead of updating (test hook).
// ------------------------------------------------------------------------------------

// Pieces of a run of `len` entries of a long row: none for runs of at most PIECE entries
// (summed by the combine kernel itself), else pieces of long_piece_len(len) entries: PIECE,
// or for hot runs a multiple of PIECE that keeps a run within kRunPieces pieces, so the
// combine kernel's one wave per row adds at most kRunPieces partial rows per run (fewer
// than len / 16 pieces in every case).
static constexpr uint32_t kRunPieces = 64;
__device__ inline uint32_t long_piece_len(uint32_t len) {
  const uint32_t per = (len + kRunPieces - 1u) / kRunPieces;
  return per <= PIECE ? PIECE : (per + PIECE - 1u) / PIECE * PIECE;
}
__device__ inline uint32_t long_run_pieces(uint32_t len) {
  if (len <= PIECE) return 0u;
  const uint32_t pl = long_piece_len(len);
  return (len + pl - 1u) / pl;
}

// Pieces of the runs of one long row, starting at piece pb; run_at(j) gives run j's
// absolute [start, end) and batch.  Lanes lane0, lane0 + stride, ... write the pieces.
template <typename RunAt>
__device__ inline void write_long_row(const StepConfig& c, uint32_t idx, uint32_t st0, uint32_t pb, uint32_t nruns,
                                      RunAt run_at, uint32_t lane0, uint32_t stride) {
  if (lane0 == 0u) {
    c.long_st[idx] = st0;
    c.long_pbase[idx] = pb;
    c.long_len[idx] = nruns;
  }
  uint32_t p = pb;
  for (uint32_t j = 0; j < nruns; ++j) {
    uint32_t a, z, bq;
    run_at(j, &a, &z, &bq);
    const uint32_t np = long_run_pieces(z - a);
    const uint32_t pl = long_piece_len(z - a);
    if (lane0 == 0u) {
      c.long_runs[(size_t)idx * kMaxGroup + j] = np | (bq << 24);
      c.long_rlen[(size_t)idx * kMaxGroup + j] = z - a;
    }
    for (uint32_t q = lane0; q < np; q += stride) {
      const uint32_t ps = a + q * pl;
      if (p + q < c.piece_cap) {
        c.piece_st[p + q] = ps;
        c.piece_info[p + q] = (z - ps < pl ? z - ps : pl) | (bq << 29);
      }
    }
    p += np;
  }
}

// Batch boundaries of a long row [lo, hi) whose batches rise: lane group g (8 lanes) finds
// the first index whose batch is >= g + 1 by an 8-ary search, all groups at once (one
// dependent load per level where one search per batch would take PM in sequence).  Groups
// g >= pm return lo.
__device__ inline uint32_t group_batch_bounds(const uint32_t* __restrict__ Ss, uint32_t lo, uint32_t hi, uint32_t pm,
                                              uint32_t lane) {
  const uint32_t g = lane >> 3, gl = lane & 7u, q = g + 1u;
  uint32_t a = lo, z = q <= pm ? hi : lo;
  while (true) {
    const bool act = z > a;
    if (__ballot(act) == 0ull) break;
    const uint32_t span = z - a;
    const uint32_t step = act ? (span + 7u) / 8u : 1u;
    const uint32_t off = gl * step;
    const bool valid = act && off < span;
    const uint32_t bt = valid ? Ss[a + off] >> kSlotBatchShift : 0u;
    const uint32_t gm = (uint32_t)(__ballot(valid && bt >= q) >> (g * 8u)) & 0xffu;
    if (act) {
      if (step == 1u) {
        a = gm ? a + (uint32_t)(__ffs(gm) - 1) : z;
        z = a;
      } else if (gm == 0u) {
        a = a + ((span - 1u) / step) * step + 1u;  // past the last sample (all below)
      } else {
        const uint32_t f = (uint32_t)(__ffs(gm) - 1);
        z = a + f * step;                    // this sample is in batch >= q
        a = f > 0u ? a + (f - 1u) * step + 1u : a;  // the one before is not
      }
    }
  }
  return a;
}

__global__ __launch_bounds__(256) void k_seg_plan(StepConfig c, StepState* st, const uint32_t* __restrict__ Ks,
                                                  const uint32_t* __restrict__ Ss) {
  __shared__ uint32_t s_run[NW][132];  // run starts (chunk-relative); s_run[nrun] = window end
  __shared__ uint8_t s_bat[NW][132];   // batch of each run
  __shared__ uint32_t s_row[NW][66];   // first run of each owned row; s_row[nrow] = nrun
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint32_t gw = blockIdx.x * NW + w;
  const uint32_t nwaves = gridDim.x * NW;
  const uint32_t n = st->n_keys <= c.key_cap ? st->n_real : 0u;  // sorted, sentinels dropped
  const uint32_t SENT = c.sentinel, PM = c.pmask;
  const uint32_t nchunks = (n + 63u) / 64u;
  const uint32_t R = c.n_regions, rpw = R / nwaves;
  const uint32_t cpr = (nchunks + R - 1u) / R;  // chunks per region
  const uint32_t cbeg = min(nchunks, gw * rpw * cpr), cend = min(nchunks, (gw + 1u) * rpw * cpr);
  if (gw == 0u && lane == 0) st->reg_next = 0u;
  uint32_t reg = gw * rpw;  // region of the chunk being classified
  size_t sbase = (size_t)reg * cpr * 64u, mbase = (size_t)reg * cpr * 2u;
  uint32_t ns = 0, nm = 0;  // records written to region reg
  // Rows (permuted ids; SENT is above every row) and batches of the window's two halves,
  // loaded one chunk ahead: the next chunk's first half is this chunk's second half, its
  // second half is loaded while this chunk is classified.  nprev / npb: the entry before
  // the chunk (the previous chunk's last).
  uint32_t nr0 = SENT, nb0 = 0u, nr1 = SENT, nb1 = 0u, nprev = 0u, npb = 0u;
  auto prime = [&](uint32_t ch) {  // the window of chunk ch and the entry before it
    const uint32_t i0 = ch * 64u;
    const uint32_t j0 = i0 + lane, j1 = i0 + 64u + lane;
    nr0 = j0 < n ? Ks[j0] : SENT;
    nr1 = j1 < n ? Ks[j1] : SENT;
    nb0 = j0 < n ? Ss[j0] >> kSlotBatchShift : 0u;
    nb1 = j1 < n ? Ss[j1] >> kSlotBatchShift : 0u;
    nprev = i0 > 0u ? Ks[i0 - 1u] : ~__shfl(nr0, 0, EP_WAVE);
    npb = i0 > 0u ? Ss[i0 - 1u] >> kSlotBatchShift : 0u;
  };
  if (cbeg < cend) prime(cbeg);
  for (uint32_t chunk = cbeg; chunk < cend; ++chunk) {
    const uint32_t i0 = chunk * 64u;
    if (chunk / cpr != reg) {  // the next region: close this one (and any jumped over)
      if (lane == 0) {
        c.scount[reg] = ns;
        c.mcount[reg] = nm;
        for (uint32_t r = reg + 1u; r < chunk / cpr; ++r) {
          c.scount[r] = 0u;
          c.mcount[r] = 0u;
        }
      }
      reg = chunk / cpr;
      sbase = (size_t)reg * cpr * 64u;
      mbase = (size_t)reg * cpr * 2u;
      ns = nm = 0u;
    }
    const uint32_t r0 = nr0, r1 = nr1, b0 = nb0, b1 = nb1;
    uint32_t prev = __shfl_up(r0, 1, EP_WAVE), pb0 = __shfl_up(b0, 1, EP_WAVE);
    if (lane == 0) {
      prev = nprev;
      pb0 = npb;
    }
    const uint32_t last0 = __shfl(r0, 63, EP_WAVE), lastb0 = __shfl(b0, 63, EP_WAVE);
    nprev = last0;
    npb = lastb0;
    nr0 = r1;
    nb0 = b1;
    {
      const uint32_t jn = i0 + 128u + lane;  // the next chunk's second half
      nr1 = (chunk + 1u < cend && jn < n) ? Ks[jn] : SENT;
      nb1 = (chunk + 1u < cend && jn < n) ? Ss[jn] >> kSlotBatchShift : 0u;
    }
    uint32_t prev1 = __shfl_up(r1, 1, EP_WAVE), pb1 = __shfl_up(b1, 1, EP_WAVE);
    if (lane == 0) {
      prev1 = last0;
      pb1 = lastb0;
    }
    const bool is_start = (r0 != SENT) && (r0 != prev);
    const uint64_t starts = __ballot(is_start);
    // a row end to find (key of the row + 1, 0: none) and where to go on from
    uint32_t find = 0u, jump = 0u;
    bool deferred = false;
    int ql = 0;
    uint32_t fl = 0u;
    if (starts == 0ull) {
      // inside a row that started before this chunk (another wave's, or a deferred one):
      // when it also fills the next chunk, find its end and go on from there
      const uint32_t key = (uint32_t)__builtin_amdgcn_readfirstlane((int)r0);
      if (key != SENT && (uint32_t)__builtin_amdgcn_readlane((int)r1, 63) == key) find = key + 1u;
    } else {
      const int first = __ffsll((unsigned long long)starts) - 1;
      ql = 63 - __clzll((unsigned long long)starts);
      fl = (uint32_t)__builtin_amdgcn_readlane((int)r0, ql);
      const uint64_t m0 = __ballot(lane > ql && r0 != fl);
      const uint64_t m1 = __ballot(r1 != fl);
      deferred = (m0 == 0ull) && (m1 == 0ull);
      if (deferred) find = fl + 1u;
      const uint32_t wend = m0 ? (uint32_t)(__ffsll((unsigned long long)m0) - 1)
                               : (m1 ? 64u + (uint32_t)(__ffsll((unsigned long long)m1) - 1) : (uint32_t)ql);
      // runs of the rows handled in the window: [first, wend)
      const uint64_t fs0 = __ballot(lane >= first && (uint32_t)lane < wend && (r0 != prev || b0 != pb0));
      const uint64_t fs1 = __ballot(64u + (uint32_t)lane < wend && (r1 != prev1 || b1 != pb1));
      const uint32_t n0 = (uint32_t)__popcll(fs0);
      const uint32_t nrun = n0 + (uint32_t)__popcll(fs1);
      if ((fs0 >> lane) & 1ull) {
        const uint32_t i = (uint32_t)__popcll(fs0 & lt);
        s_run[w][i] = (uint32_t)lane;
        s_bat[w][i] = (uint8_t)b0;
      }
      if ((fs1 >> lane) & 1ull) {
        const uint32_t i = n0 + (uint32_t)__popcll(fs1 & lt);
        s_run[w][i] = 64u + (uint32_t)lane;
        s_bat[w][i] = (uint8_t)b1;
      }
      const uint64_t rowm = deferred ? (starts & ~(1ull << ql)) : starts;
      const uint32_t nrow = (uint32_t)__popcll(rowm);
      if ((rowm >> lane) & 1ull) s_row[w][__popcll(rowm & lt)] = (uint32_t)__popcll(fs0 & lt);
      if (lane == 0) {
        s_run[w][nrun] = wend;
        s_row[w][nrow] = nrun;
      }
      __builtin_amdgcn_wave_barrier();
      // classify the owned rows of the window: short records or medium records
      const bool own = (uint32_t)lane < nrow;
      const uint32_t ra = own ? s_row[w][lane] : 0u;
      const uint32_t rz = own ? s_row[w][lane + 1] : 0u;
      const uint32_t ea = own ? s_run[w][ra] : 0u;
      const uint32_t ez = own ? s_run[w][rz] : 0u;
      const bool med = own && (ez - ea) > kSegShort;
      const uint64_t mmask = __ballot(med);
      const uint64_t smask = __ballot(own && !med);
      const uint32_t f = row_of_key(c, __shfl(r0, (int)ea, EP_WAVE));  // row starts lie in [0, 64)
      if (own && !med) {
        uint32_t mask = 0u, meta = ez - ea;
        for (uint32_t j = ra; j < rz; ++j) {
          mask |= 1u << (s_run[w][j] - ea);
          meta |= (uint32_t)s_bat[w][j] << (6u + 3u * (j - ra));
        }
        c.srec[sbase + ns + (uint32_t)__popcll(smask & lt)] = make_uint4(i0 + ea, f, mask, meta);
      }
      if (med) c.mrec[mbase + nm + (uint32_t)__popcll(mmask & lt)] = make_uint2(i0 + ea, ez - ea);
      ns += (uint32_t)__popcll(smask);
      nm += (uint32_t)__popcll(mmask);
      __builtin_amdgcn_wave_barrier();
    }
    if (find) {
      // the end of the row that fills the rest of the window (both search paths share this
      // one call); the chunks before the one holding it start no row
      const uint32_t hi = wave_lower_bound(Ks, i0 + 128u, n, find);
      if (hi / 64u > chunk + 1u) jump = min(hi / 64u, cend);
      if (deferred) {
        // the chunk's last row runs past the window: register it as long, its runs found by
        // the batch boundaries (its values rise batch by batch)
        const uint32_t lo = i0 + (uint32_t)ql;
        const uint32_t gb = group_batch_bounds(Ss, lo, hi, PM, (uint32_t)lane);
        auto bnd = [&](uint32_t b) -> uint32_t {  // b = 0..PM + 1
          return b == 0u ? lo : (b > PM ? hi : (uint32_t)__builtin_amdgcn_readlane((int)gb, (int)((b - 1u) * 8u)));
        };
        uint32_t pcs = 0, nr = 0;
        for (uint32_t b = 0; b <= PM; ++b) {
          if (bnd(b + 1) > bnd(b)) ++nr;
          pcs += long_run_pieces(bnd(b + 1) - bnd(b));
        }
        uint32_t idx = 0, pb = 0;
        if (lane == 0) {
          idx = atomicAdd(&st->n_long, 1u);
          pb = atomicAdd(&st->n_pieces, pcs);
        }
        idx = (uint32_t)__builtin_amdgcn_readfirstlane((int)idx);
        pb = (uint32_t)__builtin_amdgcn_readfirstlane((int)pb);
        if (idx < c.long_cap) {
          auto run_at = [&](uint32_t j, uint32_t* a, uint32_t* z, uint32_t* bq) {
            uint32_t k = 0, seen = 0;  // j-th nonempty batch run
            for (uint32_t b = 0; b <= PM; ++b)
              if (bnd(b + 1) > bnd(b)) { if (seen == j) k = b; ++seen; }
            *a = bnd(k);
            *z = bnd(k + 1);
            *bq = k;
          };
          write_long_row(c, idx, lo, pb, nr, run_at, (uint32_t)lane, 64u);
        }
      }
    }
    if (jump) {
      if (jump < cend) prime(jump);
      chunk = jump - 1u;
    }
  }
  if (lane == 0 && cbeg < cend) {
    c.scount[reg] = ns;
    c.mcount[reg] = nm;
    for (uint32_t r = reg + 1u; r <= (cend - 1u) / cpr; ++r) {  // regions jumped over at the end
      c.scount[r] = 0u;
      c.mcount[r] = 0u;
    }
  }
}

// Short rows: lane groups of L = D/VPL lanes (VPL values per lane as VPL/4 float4
// columns strided by L), G = 64/L rows per wave, U adjoint rows in flight per group.
// Every (VPL, U) adds each run's rows in entry order: identical bits, so the tuner picks
// the fastest.
// NZ: the gradient-noise path (noise > 0), built for one tile per width only: its
// Philox and transcendental code would raise every tile's VGPRs (and cut its waves).
template <int D, int VPL, int U, bool H, bool NZ = false>
__global__ __launch_bounds__(256) void k_seg_update(StepConfig c, StepState* st, const uint32_t* __restrict__ Ss) {
  static_assert(D % VPL == 0 && VPL % 4 == 0 && D / VPL <= 64, "float4 lane slices");
  constexpr int L = D / VPL;      // lanes per row
  constexpr int G = 64 / L;       // rows in flight per wave
  constexpr int NQ = VPL / 4;     // float4 columns per lane
  constexpr int NS = ((int)kSegShort + L - 1) / L;  // slots held per lane
  __shared__ float s_lr[kMaxGroup];
  const int tid = threadIdx.x, lane = tid & 63;
  const int grp = lane / L, lg = lane % L;
  const bool gact = grp < G;
  const uint64_t gmask = (L == 64) ? ~0ull : (((1ull << L) - 1ull) << ((grp * L) & 63));
  // plan regions holding chunks
  const uint32_t n = st->n_keys <= c.key_cap ? st->n_real : 0u;
  const uint32_t nchunks = (n + 63u) / 64u;
  const uint32_t cpr = (nchunks + c.n_regions - 1u) / c.n_regions;
  const uint32_t nreg = cpr ? (nchunks + cpr - 1u) / cpr : 0u;
  const StepScalars sv = step_scalars(st);
  __shared__ float s_sd[kMaxGroup];
  if (tid < (int)kMaxGroup) {
    s_lr[tid] = st->lr[tid];
    s_sd[tid] = st->noise_sd[tid];
  }
  __syncthreads();
  const bool noisy = c.noise > 1e-9f;
  uint32_t done = 0;
  // wave gw takes region gw, then (only when there are more regions than waves) regions
  // waves, waves + 1, ... from the queue: small steps make no atomics at all
  const uint32_t nwv = gridDim.x * NW;
  uint32_t reg = blockIdx.x * NW + (uint32_t)(tid >> 6);
  for (;;) {
    if (reg >= nreg) break;
    const uint32_t nrec = c.scount[reg];
    const uint4* recs = c.srec + (size_t)reg * cpr * 64u;
    // Software pipeline over the region's rows: iteration k issues, back to back, the row's
    // theta/m/v, the next row's slots (its record arrived one iteration earlier), the
    // record two rows ahead and the row's adjoint rows (its slots arrived one iteration
    // earlier), so each row costs one memory latency instead of the chain record ->
    // slots -> adjoint rows.  The arithmetic and its order are those of one row at a time.
    auto load_rec = [&](uint32_t jj) -> uint4 {
      return (gact && jj < nrec) ? recs[jj] : make_uint4(0u, 0u, 0u, 0u);
    };
    // the row's slots, spread over the group's lanes (entry e at lane e % L, register e / L)
    // (raw sorted values, batch bits included), loaded unconditionally at an in-bounds
    // index (entry 0 of the record past its length) and masked where used: a load inside
    // a branch whose value is modified at the join makes the compiler wait there for
    // every load in flight, the row's theta/m/v included, which undoes the pipeline.
    auto load_slots = [&](const uint4& r, uint32_t (&out)[NS]) {
  #pragma unroll
      for (int i = 0; i < NS; ++i) {
        const uint32_t e = (uint32_t)(lg + i * L);
        out[i] = Ss[r.x + (e < (r.w & 63u) ? e : 0u)];
      }
    };
    uint4 rec_next = load_rec((uint32_t)grp);
    uint32_t sl_next[NS];
    load_slots(rec_next, sl_next);
    uint4 rec_after = load_rec((uint32_t)(G + grp));
    for (uint32_t j0 = 0; j0 < nrec; j0 += (uint32_t)G) {
      const uint32_t j = j0 + (uint32_t)grp;
      const bool has = gact && j < nrec;
      uint4 rec = rec_next;
      if constexpr (G == 1) {
        // one row per wave: its record is wave-uniform (scalar registers and branches)
        rec = make_uint4(__builtin_amdgcn_readfirstlane(rec.x), __builtin_amdgcn_readfirstlane(rec.y),
                         __builtin_amdgcn_readfirstlane(rec.z), __builtin_amdgcn_readfirstlane(rec.w));
      }
      uint32_t sl[NS];
  #pragma unroll
      for (int i = 0; i < NS; ++i) sl[i] = sl_next[i];
      const uint32_t f = rec.y, rmask = rec.z;
      const uint32_t len = rec.w & 63u;
      // theta/m/v, the same loads in every precision (row 0 stands in for an empty record
      // slot and is never stored; grad_out mode loads and ignores them): the sources are
      // chosen by wave-uniform selects, and the load a precision does not need reads the
      // m row again from cache, so no branch join waits for these loads.
      float4 m4[NQ], v4[NQ], t4[NQ];
      uint2 tb[NQ];  // BF16 table without master: the raw row, widened after the first run's sum
      const bool widen = c.bf16 && !c.master;
      const size_t roff = (size_t)f * D;
      const size_t ooff = (size_t)f * c.ostride;  // optimizer record [master |] m | v
      const float* tsrc = c.master ? c.master + ooff : (c.bf16 ? c.mom + ooff : (const float*)c.table + roff);
      const uint16_t* bsrc = widen ? (const uint16_t*)c.table + roff : (const uint16_t*)(c.mom + ooff);
  #pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const uint32_t col = (uint32_t)(q * L + lg) * 4u;
        m4[q] = *reinterpret_cast<const float4*>(c.mom + ooff + col);
        v4[q] = *reinterpret_cast<const float4*>(c.var + ooff + col);
        t4[q] = *reinterpret_cast<const float4*>(tsrc + col);
        tb[q] = *reinterpret_cast<const uint2*>(bsrc + (widen ? col : 2u * col));
      }
      rec_next = rec_after;
      load_slots(rec_next, sl_next);
      rec_after = load_rec(j0 + (uint32_t)(2 * G) + (uint32_t)grp);
      const uint32_t nruns = has ? (uint32_t)__popc(rmask) : 0u;
      uint32_t maxruns = nruns;
      if constexpr (G > 1) {
  #pragma unroll
        for (int off = 32; off > 0; off >>= 1) maxruns = max(maxruns, (uint32_t)__shfl_xor(maxruns, off, EP_WAVE));
      }
      bool upd = false;
      uint32_t rest = rmask;  // run starts not yet consumed
      for (uint32_t rq = 0; rq < maxruns; ++rq) {
        const bool hr = rq < nruns;
        const uint32_t a = hr ? (uint32_t)__builtin_ctz(rest) : 0u;
        rest = hr ? (rest & (rest - 1u)) : rest;
        const uint32_t z = hr ? (rest ? (uint32_t)__builtin_ctz(rest) : len) : 0u;
        const uint32_t bq = (rec.w >> (6u + 3u * rq)) & 7u;
        uint32_t maxlen = z - a;
        if constexpr (G > 1) {
  #pragma unroll
          for (int off = 32; off > 0; off >>= 1) maxlen = max(maxlen, (uint32_t)__shfl_xor(maxlen, off, EP_WAVE));
        }
        float4 acc[NQ];
  #pragma unroll
        for (int q = 0; q < NQ; ++q) acc[q] = make_float4(0.f, 0.f, 0.f, 0.f);
        for (uint32_t u0 = 0; u0 < maxlen; u0 += U) {
          // entries past the run skip their loads (a wave-uniform branch when G == 1: most
          // runs of short rows are 1-3 entries long, whose loads all issue before a wait)
          float4 r[U][NQ];
  #pragma unroll
          for (int u = 0; u < U; ++u) {
            const uint32_t e = a + u0 + (uint32_t)u;
            const bool valid = hr && e < z;
            uint32_t slot;
            if constexpr (G == 1 && NS == 1) {
              slot = __builtin_amdgcn_readlane(sl[0], (int)(e & 63u));  // entry e sits in lane e
            } else {
              const uint32_t ei = e / (uint32_t)L;
              uint32_t pick = sl[0];
  #pragma unroll
              for (int i = 1; i < NS; ++i) pick = (ei == (uint32_t)i) ? sl[i] : pick;
              slot = __shfl(pick, grp * L + (int)(e % (uint32_t)L), EP_WAVE);
            }
            slot &= kSlotMask;
            if constexpr (H) {
              const float sc = valid ? c.ascale[slot] : 0.f;
  #pragma unroll
              for (int q = 0; q < NQ; ++q) {
                const uint32_t col = (uint32_t)(q * L + lg) * 4u;
                const uint2 t = valid ? *reinterpret_cast<const uint2*>(c.adjh + (size_t)slot * D + col) : make_uint2(0u, 0u);
                r[u][q] = make_float4(sc * f16_to_f32((uint16_t)(t.x & 0xffffu)), sc * f16_to_f32((uint16_t)(t.x >> 16)),
                                      sc * f16_to_f32((uint16_t)(t.y & 0xffffu)), sc * f16_to_f32((uint16_t)(t.y >> 16)));
              }
            } else {
  #pragma unroll
              for (int q = 0; q < NQ; ++q) {
                const uint32_t col = (uint32_t)(q * L + lg) * 4u;
                r[u][q] = valid ? *reinterpret_cast<const float4*>(c.adj + (size_t)slot * D + col)
                                : make_float4(0.f, 0.f, 0.f, 0.f);
              }
            }
          }
  #pragma unroll
          for (int u = 0; u < U; ++u) {
  #pragma unroll
            for (int q = 0; q < NQ; ++q) {
              acc[q].x += r[u][q].x; acc[q].y += r[u][q].y; acc[q].z += r[u][q].z; acc[q].w += r[u][q].w;
            }
          }
        }
        if (rq == 0u && widen) {  // the raw BF16 row has arrived with the first run's rows
  #pragma unroll
          for (int q = 0; q < NQ; ++q) {
            uint32_t bx = tb[q].x, by = tb[q].y;
            // an empty register fence: keeps the widening (and the wait for the row) here,
            // after the first run's adjoint loads, instead of hoisted in front of them
            asm volatile("" : "+v"(bx), "+v"(by));
            t4[q] = make_float4(__uint_as_float(bx << 16), __uint_as_float(bx & 0xffff0000u),
                                __uint_as_float(by << 16), __uint_as_float(by & 0xffff0000u));
          }
        }
        if (NZ && noisy && hr) {  // mod.rs:238-247: noise on the batch's merged gradient row
          const float sd = s_sd[bq];
  #pragma unroll
          for (int q = 0; q < NQ; ++q) {
            float z[4];
            noise4(c, f, (uint32_t)(q * L + lg), sv.step0 + bq, z);
            acc[q].x += sd * z[0]; acc[q].y += sd * z[1]; acc[q].z += sd * z[2]; acc[q].w += sd * z[3];
          }
        }
        if (c.grad_out) {
          uint32_t ub = 0;
          const uint64_t heads = __ballot(hr && lg == 0);
          if (lane == 0 && heads) ub = atomicAdd(&st->n_unique, (uint32_t)__popcll(heads));
          ub = __shfl(ub, 0, EP_WAVE);
          const uint32_t u = ub + (uint32_t)__popcll(heads & ((grp * L == 0) ? 0ull : (~0ull >> (64 - grp * L))));
          if (hr) {
  #pragma unroll
            for (int q = 0; q < NQ; ++q)
              *reinterpret_cast<float4*>(c.G + (size_t)u * D + (q * L + lg) * 4) = acc[q];
            if (lg == 0) c.ufeat[u] = f | (bq << 27);
          }
          continue;
        }
        bool bad = false;
  #pragma unroll
        for (int q = 0; q < NQ; ++q) bad |= isnan(acc[q].x) || isnan(acc[q].y) || isnan(acc[q].z) || isnan(acc[q].w);
        const bool skip = (__ballot(hr && bad) & gmask) != 0ull;  // optimizer.rs:56, per batch
        if (hr && !skip) {
          const float lr = s_lr[bq];
  #pragma unroll
          for (int q = 0; q < NQ; ++q) {
            adam_math(t4[q].x, m4[q].x, v4[q].x, acc[q].x, lr, sv.ibc1, sv.ibc2);
            adam_math(t4[q].y, m4[q].y, v4[q].y, acc[q].y, lr, sv.ibc1, sv.ibc2);
            adam_math(t4[q].z, m4[q].z, v4[q].z, acc[q].z, lr, sv.ibc1, sv.ibc2);
            adam_math(t4[q].w, m4[q].w, v4[q].w, acc[q].w, lr, sv.ibc1, sv.ibc2);
          }
          upd = true;
        }
      }
      if (upd) {
        if (lg == 0) ++done;
  #pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const uint32_t col = (uint32_t)(q * L + lg) * 4u;
          *reinterpret_cast<float4*>(c.mom + ooff + col) = m4[q];
          *reinterpret_cast<float4*>(c.var + ooff + col) = v4[q];
          if (c.bf16) {
            if (c.master) *reinterpret_cast<float4*>(c.master + ooff + col) = t4[q];
            const uint32_t h0 = sr_key(f, col >> 2, sv.sr_step, (uint32_t)c.seed);
            const uint32_t h1 = hash32(h0 ^ 0x27D4EB2Fu);
            uint2 pk2;
            pk2.x = (uint32_t)f32_to_bf16_sr(t4[q].x, h0 & 0xffffu) | ((uint32_t)f32_to_bf16_sr(t4[q].y, h0 >> 16) << 16);
            pk2.y = (uint32_t)f32_to_bf16_sr(t4[q].z, h1 & 0xffffu) | ((uint32_t)f32_to_bf16_sr(t4[q].w, h1 >> 16) << 16);
            *reinterpret_cast<uint2*>((uint16_t*)c.table + roff + col) = pk2;
          } else {
            *reinterpret_cast<float4*>((float*)c.table + roff + col) = t4[q];
          }
        }
      }
    }
    if (nreg <= nwv) break;
    uint32_t nx = 0u;
    if (lane == 0) nx = atomicAdd(&st->reg_next, 1u);
    reg = nwv + (uint32_t)__shfl((int)nx, 0, EP_WAVE);
  }
  done = wave_sum_u32(done);
  if (lane == 0 && done) {
    atomicAdd((unsigned long long*)&st->stat_unique, (unsigned long long)done);
    atomicAdd((unsigned long long*)&st->stat_short, (unsigned long long)done);
  }
}

// Loads row f's optimizer operands (whole-wave slice layout).
template <int D>
__device__ inline void load_row_state(const StepConfig& c, uint32_t f, int lane, float (&th)[RowShape<D>::PL],
                                      float (&m)[RowShape<D>::PL], float (&v)[RowShape<D>::PL]) {
  const size_t off = (size_t)f * D, ooff = (size_t)f * c.ostride;
  load_slice<D>(c.mom + ooff, lane, m);
  load_slice<D>(c.var + ooff, lane, v);
  if (c.master) load_slice<D>(c.master + ooff, lane, th);
  else if (c.bf16) load_slice<D>((const uint16_t*)c.table + off, lane, th);
  else load_slice<D>((const float*)c.table + off, lane, th);
}

// Stores row f's optimizer operands; bf16 tables get the stochastically rounded theta.
template <int D>
__device__ inline void store_row_state(const StepConfig& c, const StepScalars& sv, uint32_t f, int lane,
                                       const float (&th)[RowShape<D>::PL], const float (&m)[RowShape<D>::PL],
                                       const float (&v)[RowShape<D>::PL]) {
  constexpr int PL = RowShape<D>::PL;
  const int base = lane * PL;
  const size_t off = (size_t)f * D, ooff = (size_t)f * c.ostride;
  store_slice<D>(c.mom + ooff, lane, m);
  store_slice<D>(c.var + ooff, lane, v);
  if (c.bf16) {
    if (c.master) store_slice<D>(c.master + ooff, lane, th);
    uint16_t* row = (uint16_t*)c.table + off;
#pragma unroll
    for (int q = 0; q < PL; ++q)
      if (base + q < D)
        row[base + q] = f32_to_bf16_sr(th[q], sr_bits(c, sv, f, (uint32_t)((base + q) >> 2), (base + q) & 3));
  } else {
    store_slice<D>((float*)c.table + off, lane, th);
  }
}

// Applies one run's merged gradient (NaN rows skipped, optimizer.rs:56) or, in grad_out
// mode, writes it as a (row, batch) record.  Returns whether theta changed.
template <int D>
__device__ inline bool apply_run(const StepConfig& c, StepState* st, const StepScalars& sv, float lr, float sd,
                                 uint32_t f, uint32_t bq, int lane, float (&acc)[RowShape<D>::PL],
                                 float (&th)[RowShape<D>::PL], float (&m)[RowShape<D>::PL], float (&v)[RowShape<D>::PL]) {
  constexpr int PL = RowShape<D>::PL;
  const int base = lane * PL;
  if (c.noise > 1e-9f) add_noise_slice<D>(c, f, sv.step0 + bq, sd, lane, acc);  // mod.rs:238-247
  if (c.grad_out) {
    uint32_t ui = 0;
    if (lane == 0) ui = atomicAdd(&st->n_unique, 1u);
    ui = __shfl(ui, 0, EP_WAVE);
    store_slice<D>(c.G + (size_t)ui * D, lane, acc);
    if (lane == 0) c.ufeat[ui] = f | (bq << 27);
    return false;
  }
  bool bad = false;
#pragma unroll
  for (int q = 0; q < PL; ++q) bad |= (base + q < D) && isnan(acc[q]);
  if (wave_any(bad)) return false;
#pragma unroll
  for (int q = 0; q < PL; ++q) adam_math(th[q], m[q], v[q], acc[q], lr, sv.ibc1, sv.ibc2);
  return true;
}

// acc += the first n of the U rows r[0..U), added as a pairwise tree: a long reduction
// done U rows at a time then has a rounding error that grows with its length / U
// instead of its length (U a power of two; rows past n count as zero).
template <int U, int PL>
__device__ inline void add_rows_tree(float (&acc)[PL], float (&r)[U][PL], uint32_t n) {
  static_assert((U & (U - 1)) == 0, "U is a power of two");
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int q = 0; q < PL; ++q) r[u][q] = (uint32_t)u < n ? r[u][q] : 0.f;
#pragma unroll
  for (int w = 1; w < U; w *= 2)
#pragma unroll
    for (int u = 0; u + w < U; u += 2 * w)
#pragma unroll
      for (int q = 0; q < PL; ++q) r[u][q] += r[u + w][q];
#pragma unroll
  for (int q = 0; q < PL; ++q) acc[q] += r[0][q];
}

// Partial sums of the pieces of long rows (each piece: <= kPieceEntries entries of one
// run), then the medium rows of this wave's plan region: one wave per row (<= 128
// entries), its runs reduced from the adjoint rows and applied in batch order.
template <int D, bool H>
__global__ __launch_bounds__(256) void k_pieces(StepConfig c, StepState* st, const uint32_t* __restrict__ Ks,
                                                const uint32_t* __restrict__ Ss) {
  constexpr int PL = RowShape<D>::PL;
  constexpr int U = SR_UNROLL;
  __shared__ float s_lr[kMaxGroup], s_sd[kMaxGroup];
  const int lane = threadIdx.x & 63;
  const uint32_t gw = blockIdx.x * NW + (threadIdx.x >> 6);
  const uint32_t nwaves = gridDim.x * NW;
  if (threadIdx.x < kMaxGroup) {
    s_lr[threadIdx.x] = st->lr[threadIdx.x];
    s_sd[threadIdx.x] = st->noise_sd[threadIdx.x];
  }
  __syncthreads();
  const uint32_t np = st->n_pieces < c.piece_cap ? st->n_pieces : c.piece_cap;
  // Batch by batch: a piece reads adjoint rows of its batch only, so every wave working
  // on the same batch keeps the rows being read within one batch's share of the adjoint
  // buffer (which the memory-side cache can hold) instead of all of them.
  for (uint32_t bq = 0; bq <= c.pmask; ++bq)
  for (uint32_t gp = gw; gp < np; gp += nwaves) {
    const uint32_t info = c.piece_info[gp];
    if ((info >> 29) != bq) continue;
    const uint32_t s0 = c.piece_st[gp];
    const uint32_t m = info & 0x1fffffffu;  // entries: a multiple of PIECE at most
    float acc[PL];
#pragma unroll
    for (int q = 0; q < PL; ++q) acc[q] = 0.f;
    for (uint32_t c0 = 0; c0 < m; c0 += 64u) {  // 64 slots per lane-held chunk
      const uint32_t mm = m - c0 < 64u ? m - c0 : 64u;
      const uint32_t sj = (uint32_t)lane < mm ? (Ss[s0 + c0 + lane] & kSlotMask) : 0u;
      for (uint32_t t0 = 0; t0 < mm; t0 += U) {
        float r[U][PL];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint32_t t = (t0 + u < mm) ? (t0 + u) : (mm - 1u);
          const uint32_t s = __shfl(sj, (int)t, EP_WAVE);
          load_adj_row<D, H>(c, s, lane, r[u]);
        }
        add_rows_tree<U, PL>(acc, r, mm - t0);
      }
    }
    store_slice<D>(c.part + (size_t)gp * D, lane, acc);
  }
  // medium rows of the plan regions gw, gw + nwaves, ...
  const uint32_t n = st->n_keys <= c.key_cap ? st->n_real : 0u;
  const uint32_t nchunks = (n + 63u) / 64u;
  const uint32_t cpr = (nchunks + c.n_regions - 1u) / c.n_regions;
  const uint32_t nreg = cpr ? (nchunks + cpr - 1u) / cpr : 0u;
  const StepScalars sv = step_scalars(st);
  uint32_t done = 0;
  for (uint32_t reg = gw; reg < nreg; reg += nwaves) {
    const uint32_t nmed = c.mcount[reg];
    const uint2* recs = c.mrec + (size_t)reg * cpr * 2u;
    for (uint32_t k = 0; k < nmed; ++k) {
      const uint2 rec = recs[k];
      const uint32_t start = rec.x, len = rec.y;  // 32 < len <= 128
      const uint32_t e0 = (uint32_t)lane, e1 = 64u + (uint32_t)lane;
      const uint32_t kk0 = e0 < len ? Ss[start + e0] >> kSlotBatchShift : 0u;  // batches
      const uint32_t kk1 = e1 < len ? Ss[start + e1] >> kSlotBatchShift : 0u;
      const uint32_t ss0 = e0 < len ? (Ss[start + e0] & kSlotMask) : 0u;
      const uint32_t ss1 = e1 < len ? (Ss[start + e1] & kSlotMask) : 0u;
      const uint32_t f = row_of_key(c, Ks[start]);
      float th[PL], mo[PL], va[PL];
      if (!c.grad_out) load_row_state<D>(c, f, lane, th, mo, va);
      const uint32_t last0 = __shfl(kk0, 63, EP_WAVE);
      uint32_t p0 = __shfl_up(kk0, 1, EP_WAVE);
      uint32_t p1 = __shfl_up(kk1, 1, EP_WAVE);
      if (lane == 0) { p0 = ~kk0; p1 = last0; }
      uint64_t rs0 = __ballot(e0 < len && kk0 != p0);  // run starts
      uint64_t rs1 = __ballot(e1 < len && kk1 != p1);
      bool upd = false;
      while (rs0 | rs1) {
        const uint32_t a = rs0 ? (uint32_t)(__ffsll((unsigned long long)rs0) - 1) : 64u + (uint32_t)(__ffsll((unsigned long long)rs1) - 1);
        if (rs0) rs0 &= rs0 - 1ull; else rs1 &= rs1 - 1ull;
        const uint32_t z = rs0 ? (uint32_t)(__ffsll((unsigned long long)rs0) - 1)
                               : (rs1 ? 64u + (uint32_t)(__ffsll((unsigned long long)rs1) - 1) : len);
        const uint32_t ka = __shfl(kk0, (int)(a & 63u), EP_WAVE);
        const uint32_t kb = __shfl(kk1, (int)(a & 63u), EP_WAVE);
        const uint32_t bq = a < 64u ? ka : kb;
        float acc[PL];
  #pragma unroll
        for (int q = 0; q < PL; ++q) acc[q] = 0.f;
        for (uint32_t t0 = a; t0 < z; t0 += U) {
          float r[U][PL];
  #pragma unroll
          for (int u = 0; u < U; ++u) {
            const uint32_t t = (t0 + u < z) ? (t0 + u) : (z - 1u);
            const uint32_t sa = __shfl(ss0, (int)(t & 63u), EP_WAVE);
            const uint32_t sb = __shfl(ss1, (int)(t & 63u), EP_WAVE);
            load_adj_row<D, H>(c, t < 64u ? sa : sb, lane, r[u]);
          }
          add_rows_tree<U, PL>(acc, r, z - t0);
        }
        upd |= apply_run<D>(c, st, sv, s_lr[bq], s_sd[bq], f, bq, lane, acc, th, mo, va);
      }
      if (upd) {
        store_row_state<D>(c, sv, f, lane, th, mo, va);
        ++done;
      }
    }
  }
  if (lane == 0 && done) atomicAdd((unsigned long long*)&st->stat_unique, (unsigned long long)done);
}

// One wave per long row: loads theta/m/v once, adds the pieces of each run in order and
// applies one Adam step per run (its batch's learning rate), stores the row once.
template <int D, bool H>
__global__ __launch_bounds__(256) void k_long_combine(StepConfig c, StepState* st, const uint32_t* __restrict__ Ks,
                                                      const uint32_t* __restrict__ Ss) {
  constexpr int PL = RowShape<D>::PL;
  // 8 rows in flight where the other row kernels keep 16: at d = 256 the kernel then runs
  // 4 waves per SIMD instead of 3 (its runs hold a few pieces each)
  constexpr int U = PL >= 4 ? 8 : SR_UNROLL;
  const int lane = threadIdx.x & 63;
  const uint32_t gw = blockIdx.x * NW + (threadIdx.x >> 6);
  const uint32_t nwaves = gridDim.x * NW;
  const uint32_t nl = st->n_long < c.long_cap ? st->n_long : c.long_cap;
  const StepScalars sv = step_scalars(st);
  __shared__ float s_lr[kMaxGroup], s_sd[kMaxGroup];
  if (threadIdx.x < kMaxGroup) {
    s_lr[threadIdx.x] = st->lr[threadIdx.x];
    s_sd[threadIdx.x] = st->noise_sd[threadIdx.x];
  }
  __syncthreads();
  uint32_t done = 0;
  for (uint32_t idx = gw; idx < nl; idx += nwaves) {
    const uint32_t pb0 = c.long_pbase[idx];
    const uint32_t nruns = c.long_len[idx];
    const uint32_t f = row_of_key(c, Ks[c.long_st[idx]]);
    float th[PL], m[PL], v[PL];
    if (!c.grad_out) load_row_state<D>(c, f, lane, th, m, v);
    bool upd = false;
    uint32_t pb = pb0;
    uint32_t cur = c.long_st[idx];  // first entry of the current run
    for (uint32_t jr = 0; jr < nruns && jr < kMaxGroup; ++jr) {
      const uint32_t rinfo = c.long_runs[(size_t)idx * kMaxGroup + jr];
      const uint32_t rlen = c.long_rlen[(size_t)idx * kMaxGroup + jr];
      const uint32_t bq = (rinfo >> 24) & 7u;
      uint32_t pcs = rinfo & 0xffffffu;
      if (pb >= c.piece_cap) pcs = 0u;
      else if (pcs > c.piece_cap - pb) pcs = c.piece_cap - pb;
      float acc[PL];
#pragma unroll
      for (int q = 0; q < PL; ++q) acc[q] = 0.f;
      if ((rinfo & 0xffffffu) == 0u) {
        // a run of at most PIECE entries: its adjoint rows in entry order
        const uint32_t sj = (uint32_t)lane < rlen ? (Ss[cur + lane] & kSlotMask) : 0u;
        for (uint32_t t0 = 0; t0 < rlen; t0 += U) {
          float r[U][PL];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const uint32_t t = (t0 + u < rlen) ? (t0 + u) : (rlen - 1u);
            load_adj_row<D, H>(c, __shfl(sj, (int)t, EP_WAVE), lane, r[u]);
          }
          add_rows_tree<U, PL>(acc, r, rlen - t0);
        }
      }
      for (uint32_t j0 = 0; j0 < pcs; j0 += U) {
        float r[U][PL];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint32_t jj = (j0 + u < pcs) ? j0 + u : pcs - 1u;
          load_slice<D>(c.part + (size_t)(pb + jj) * D, lane, r[u]);
        }
        add_rows_tree<U, PL>(acc, r, pcs - j0);
      }
      pb += rinfo & 0xffffffu;
      cur += rlen;
      upd |= apply_run<D>(c, st, sv, s_lr[bq], s_sd[bq], f, bq, lane, acc, th, m, v);
    }
    if (upd) {
      store_row_state<D>(c, sv, f, lane, th, m, v);
      ++done;
    }
  }
  if (lane == 0 && done) atomicAdd((unsigned long long*)&st->stat_unique, (unsigned long long)done);
}

// ------------------------------------------------------------------------------------
// Atomic fast path: scatter the (pre-scaled) adjoint rows straight into a dense FP32 gradient table with
// float atomics (emission order, no sort), then Adam over the touched rows, which also
// re-zeroes their gradient rows and flags.  Differs from the sorted path only in the
// order of the FP32 additions.
// ------------------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(256) void k_atomic_scatter(StepConfig c, StepState* st) {
  constexpr int PL = RowShape<D>::PL;
  constexpr int U = 8;
  const int lane = threadIdx.x & 63;
  const int base = lane * PL;
  const uint32_t gw = blockIdx.x * NW + (threadIdx.x >> 6);
  const uint32_t nwaves = gridDim.x * NW;
  const uint32_t n = st->n_keys <= c.key_cap ? st->n_keys : 0u;
  const uint32_t SENT = c.sentinel;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (uint64_t chunk = gw; chunk * 64u < n; chunk += nwaves) {
    const uint32_t i0 = (uint32_t)(chunk * 64u);
    const uint32_t j = i0 + lane;
    const bool valid = j < n;
    const uint32_t kf = valid ? c.keys[0][j] : SENT;
    const uint32_t s = valid ? (c.slots[0][j] & kSlotMask) : 0u;
    const bool real = kf != SENT;
    const uint32_t f = real ? row_of_key(c, kf) : 0u;
    const bool newt = real && atomicCAS(&c.flag[f], 0u, 1u) == 0u;
    const uint64_t nm = __ballot(newt);
    uint32_t tb = 0;
    if (lane == 0 && nm) tb = atomicAdd(&st->n_touched, (uint32_t)__popcll(nm));
    tb = __shfl(tb, 0, EP_WAVE);
    if (newt) c.touched[tb + (uint32_t)__popcll(nm & lt)] = f;
    const uint64_t rm = __ballot(real);
    if (!rm) continue;
    const int m = 64 - __clzll((unsigned long long)rm);  // entries up to the last real one
    for (int t0 = 0; t0 < m; t0 += U) {
      float r[U][PL];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int t = (t0 + u < m) ? (t0 + u) : (m - 1);
        const uint32_t st_ = __shfl(s, t, EP_WAVE);
        load_slice<D>(c.adj + (size_t)st_ * D, lane, r[u]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int t = t0 + u;
        if (t < m && ((rm >> t) & 1ull)) {
          const uint32_t ft = __shfl(f, t, EP_WAVE);
          float* grow = c.Gd + (size_t)ft * D;
#pragma unroll
          for (int q = 0; q < PL; ++q)
            if (base + q < D) atomicAdd(grow + base + q, r[u][q]);
        }
      }
    }
  }
}

template <int D>
__global__ __launch_bounds__(256) void k_adam_touched(StepConfig c, StepState* st) {
  constexpr int PL = RowShape<D>::PL;
  const int lane = threadIdx.x & 63;
  const uint32_t gw = blockIdx.x * NW + (threadIdx.x >> 6);
  const uint32_t nwaves = gridDim.x * NW;
  const uint32_t nt = st->n_touched;
  const StepScalars sv = step_scalars(st);
  if (gw == 0 && lane == 0) atomicAdd((unsigned long long*)&st->stat_unique, (unsigned long long)nt);
  for (uint32_t u = gw; u < nt; u += nwaves) {
    const uint32_t f = c.touched[u];
    float* grow = c.Gd + (size_t)f * D;
    float g[PL], z[PL];
    load_slice<D>(grow, lane, g);
#pragma unroll
    for (int q = 0; q < PL; ++q) z[q] = 0.f;
    store_slice<D>(grow, lane, z);
    if (lane == 0) c.flag[f] = 0u;
    if (c.noise > 1e-9f) add_noise_slice<D>(c, f, sv.step0, sv.sd0, lane, g);  // mod.rs:238-247
    adam_row<D>(c, sv, f, g, lane);
  }
}

// ------------------------------------------------------------------------------------
// Random init (mod.rs:381-392): x = 2u - 1 per value, then each row divided by its L2 norm.
// ------------------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(256) void k_init_table(StepConfig c, uint64_t seed) {
  constexpr int PL = RowShape<D>::PL;
  const int lane = threadIdx.x & 63;
  const int base = lane * PL;
  const uint64_t gw = (uint64_t)blockIdx.x * NW + (threadIdx.x >> 6);
  const uint64_t nwaves = (uint64_t)gridDim.x * NW;
  for (uint64_t row = gw; row < c.n_vocab; row += nwaves) {
    float x[PL];
    float sq = 0.f;
#pragma unroll
    for (int q0 = 0; q0 < PL; q0 += 4) {
      // one Philox block per 4 values of the lane's slice
      const u32x4 r0 = philox4x32(u32x4{(uint32_t)row, (uint32_t)(row >> 32), (uint32_t)(lane + 64 * (q0 >> 2)),
                                        RNG_INIT},
                                  (uint32_t)seed, (uint32_t)(seed >> 32));
      const uint32_t rw[4] = {r0.x, r0.y, r0.z, r0.w};
#pragma unroll
      for (int q = q0; q < PL && q < q0 + 4; ++q) {
        x[q] = (base + q < D) ? (2.0f * uniform01(rw[q & 3]) - 1.0f) : 0.f;
        sq += x[q] * x[q];
      }
    }
    const float nrm = sqrtf(wave_sum(sq));
#pragma unroll
    for (int q = 0; q < PL; ++q) x[q] = x[q] / nrm;
    if (c.bf16) {
      if (c.master) store_slice<D>(c.master + row * c.ostride, lane, x);
      uint16_t* t = (uint16_t*)c.table + row * D;
#pragma unroll
      for (int q = 0; q < PL; ++q)
        if (base + q < D) t[base + q] = f32_to_bf16_rne(x[q]);
    } else {
      store_slice<D>((float*)c.table + row * D, lane, x);
    }
  }
}

// dst[r*D + j] = bf16(src[r*stride + j]) for rows r < n/D.
__global__ void k_fill_bf16(const float* __restrict__ src, uint16_t* __restrict__ dst, uint64_t n, uint32_t D,
                            uint32_t stride) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = f32_to_bf16_rne(src[(i / D) * stride + i % D]);
}

// ------------------------------------------------------------------------------------
// Averaged inference aggregate (aggregator.rs:34-47): node row = mean of its feature rows.
// ------------------------------------------------------------------------------------
template <int D, typename T>
__global__ __launch_bounds__(256) void k_embed_avg(const T* __restrict__ table, const uint64_t* __restrict__ findptr,
                                                   const uint32_t* __restrict__ fids, uint64_t n_nodes,
                                                   float* __restrict__ out) {
  constexpr int PL = RowShape<D>::PL;
  const int lane = threadIdx.x & 63;
  const uint64_t gw = (uint64_t)blockIdx.x * NW + (threadIdx.x >> 6);
  const uint64_t nwaves = (uint64_t)gridDim.x * NW;
  for (uint64_t node = gw; node < n_nodes; node += nwaves) {
    float acc[PL];
#pragma unroll
    for (int q = 0; q < PL; ++q) acc[q] = 0.f;
    const uint64_t fs = findptr[node], fe = findptr[node + 1];
    gather_rows<D, 8>(table, fids, fs, fe, lane, acc);
    const float n = (float)(fe - fs);
#pragma unroll
    for (int q = 0; q < PL; ++q) acc[q] = fe > fs ? acc[q] / n : 0.f;
    store_slice<D>(out + node * D, lane, acc);
  }
}

// WeightedAggregator::construct (aggregator.rs:98-116): node = sum_f w_f E_f / sum_f w_f
// over the listed features, w_f = alpha / (alpha + p_w[f]) (SIF weights from unigram
// probabilities); nodes without features keep a zero row.
template <int D, typename T>
__global__ __launch_bounds__(256) void k_embed_weighted(const T* __restrict__ table, const uint64_t* __restrict__ findptr,
                                                        const uint32_t* __restrict__ fids, const float* __restrict__ p_w,
                                                        float alpha, uint64_t n_nodes, float* __restrict__ out) {
  constexpr int PL = RowShape<D>::PL;
  constexpr int U = 8;
  const int lane = threadIdx.x & 63;
  const uint64_t gw = (uint64_t)blockIdx.x * NW + (threadIdx.x >> 6);
  const uint64_t nwaves = (uint64_t)gridDim.x * NW;
  for (uint64_t node = gw; node < n_nodes; node += nwaves) {
    float acc[PL];
#pragma unroll
    for (int q = 0; q < PL; ++q) acc[q] = 0.f;
    float wsum = 0.f;
    const uint64_t fs = findptr[node], fe = findptr[node + 1];
    for (uint64_t base = fs; base < fe; base += EP_WAVE) {
      const int n = (int)((fe - base) < (uint64_t)EP_WAVE ? (fe - base) : (uint64_t)EP_WAVE);
      const uint32_t myf = lane < n ? fids[base + lane] : 0u;
      const float myw = lane < n ? alpha / (alpha + p_w[myf]) : 0.f;
      for (int j = 0; j < n; j += U) {
        float r[U][PL], w[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int t = (j + u < n) ? (j + u) : (n - 1);
          const uint32_t f = __shfl(myf, t, EP_WAVE);
          w[u] = __shfl(myw, t, EP_WAVE);
          load_slice<D>(table + (size_t)f * D, lane, r[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (j + u < n) {
            wsum += w[u];
#pragma unroll
            for (int q = 0; q < PL; ++q) acc[q] += w[u] * r[u][q];
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < PL; ++q) acc[q] = fe > fs ? acc[q] / wsum : 0.f;
    store_slice<D>(out + node * D, lane, acc);
  }
}

// ------------------------------------------------------------------------------------
// k_align: NeighborhoodAligner::align (alignment.rs:25-75), one wave per node v of
// [first, first + count): out = (1 - a) * sum_j p_j E[u_j] + a * E[v] over the first
// min(deg, max_nb) edges (max_nb 0: all; the probabilities are not renormalised), p_j the
// edge probability (CDFtoP of the row's CDF), a = alpha, or min(1 / ln(deg + 1), 1) with
// the full degree when adaptive.  Each lane adds its columns in edge order; U rows in
// flight.  A node without edges gets a * E[v] (the reference indexes weights[0] there).
// ------------------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(256) void k_align(const float* __restrict__ emb, const uint64_t* __restrict__ indptr,
                                               const uint32_t* __restrict__ indices, const float* __restrict__ cdf,
                                               float alpha, int adaptive, uint32_t max_nb, uint64_t first,
                                               uint64_t count, float* __restrict__ out) {
  constexpr int PL = RowShape<D>::PL;
  constexpr int U = 8;
  const int lane = threadIdx.x & 63;
  const uint64_t gw = (uint64_t)blockIdx.x * NW + (threadIdx.x >> 6);
  const uint64_t nwaves = (uint64_t)gridDim.x * NW;
  for (uint64_t i = gw; i < count; i += nwaves) {
    const uint64_t v = first + i;
    const uint64_t s0 = indptr[v];
    const uint32_t deg = (uint32_t)(indptr[v + 1] - s0);
    const uint32_t m = max_nb ? min(deg, max_nb) : deg;
    float acc[PL];
#pragma unroll
    for (int q = 0; q < PL; ++q) acc[q] = 0.f;
    for (uint32_t base = 0; base < m; base += 64u) {
      const int n = (int)min(64u, m - base);
      const uint32_t e = base + (uint32_t)lane;
      const uint32_t myu = lane < n ? indices[s0 + e] : 0u;
      const float myp = lane < n ? (e == 0u ? cdf[s0] : cdf[s0 + e] - cdf[s0 + e - 1u]) : 0.f;
      for (int j = 0; j < n; j += U) {
        float r[U][PL], w[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int t = (j + u < n) ? (j + u) : (n - 1);
          const uint32_t uu = (uint32_t)__shfl((int)myu, t, EP_WAVE);
          w[u] = __shfl(myp, t, EP_WAVE);
          load_slice<D>(emb + (size_t)uu * D, lane, r[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (j + u < n) {
#pragma unroll
            for (int q = 0; q < PL; ++q) acc[q] += w[u] * r[u][q];
          }
        }
      }
    }
    const float a = adaptive ? fminf(1.f / logf((float)deg + 1.f), 1.f) : alpha;
    float o[PL];
    load_slice<D>(emb + (size_t)v * D, lane, o);
#pragma unroll
    for (int q = 0; q < PL; ++q) acc[q] = acc[q] * (1.f - a) + o[q] * a;
    store_slice<D>(out + (size_t)i * D, lane, acc);
  }
}

// ------------------------------------------------------------------------------------
// Dispatch
// ------------------------------------------------------------------------------------
#define EP_DIMS(X) X(8) X(16) X(32) X(64) X(100) X(128) X(256)

bool dims_supported(int d) {
#define EP_CASE(DD) if (d == DD) return true;
  EP_DIMS(EP_CASE)
#undef EP_CASE
  return false;
}

#define EP_ROW_DIMS(X) EP_DIMS(X) X(384) X(512) X(768)

bool row_dims_supported(int d) {
#define EP_CASE(DD) if (d == DD) return true;
  EP_ROW_DIMS(EP_CASE)
#undef EP_CASE
  return false;
}

// Table-row kernels: the mean model's widths and the attention row widths.
#define EP_ROW_DISPATCH(d, BODY)             \
  switch (d) {                               \
    case 8: { constexpr int D = 8; BODY; } break;     \
    case 16: { constexpr int D = 16; BODY; } break;   \
    case 32: { constexpr int D = 32; BODY; } break;   \
    case 64: { constexpr int D = 64; BODY; } break;   \
    case 100: { constexpr int D = 100; BODY; } break; \
    case 128: { constexpr int D = 128; BODY; } break; \
    case 256: { constexpr int D = 256; BODY; } break; \
    case 384: { constexpr int D = 384; BODY; } break; \
    case 512: { constexpr int D = 512; BODY; } break; \
    case 768: { constexpr int D = 768; BODY; } break; \
    default: return hipErrorInvalidValue;    \
  }

#define EP_DISPATCH(d, BODY)                 \
  switch (d) {                               \
    case 8: { constexpr int D = 8; BODY; } break;     \
    case 16: { constexpr int D = 16; BODY; } break;   \
    case 32: { constexpr int D = 32; BODY; } break;   \
    case 64: { constexpr int D = 64; BODY; } break;   \
    case 100: { constexpr int D = 100; BODY; } break; \
    case 128: { constexpr int D = 128; BODY; } break; \
    case 256: { constexpr int D = 256; BODY; } break; \
    default: return hipErrorInvalidValue;    \
  }

hipError_t launch_step_begin(const StepConfig& c, StepState* st, hipStream_t s) {
  hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(256), 0, s, c, st, -1.0f);
  return hipGetLastError();
}

hipError_t launch_step_begin_lr(const StepConfig& c, StepState* st, float lr, hipStream_t s) {
  hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(256), 0, s, c, st, lr);
  return hipGetLastError();
}

hipError_t launch_sample(const StepConfig& c, StepState* st, hipStream_t s) {
  hipLaunchKernelGGL(k_sample, dim3((c.batch + NW - 1) / NW), dim3(256), (size_t)NW * c.S * sizeof(uint32_t), s, c, st);
  return hipGetLastError();
}

hipError_t launch_scan(const StepConfig& c, StepState* st, hipStream_t s) {
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, c, st);
  return hipGetLastError();
}

size_t forward_lds_bytes(const StepConfig& c) {
  const uint32_t NC = c.K + 2;
  return (size_t)forward_float_words(NC, c.K, c.dims) * 4 + (size_t)c.ent_cap * 8 + (size_t)c.sel_cap * 12 +
         (size_t)(c.ent_cap + 1) * 4 + ((size_t)c.ids_cap + ids_pad_entries(c.dims, c.bf16 != 0)) * 4;
}

hipError_t launch_forward(const StepConfig& c, StepState* st, hipStream_t s) {
  const size_t lds = forward_lds_bytes(c);
  EP_DISPATCH(c.dims, {
    const bool x2 = c.fwd_x == 2;
    if (c.bf16) {
      if constexpr (WideShape<D, uint16_t>::OK) {
        if (x2) hipLaunchKernelGGL((k_forward<D, uint16_t, 2>), dim3(c.batch), dim3(256), lds, s, c, st);
        else hipLaunchKernelGGL((k_forward<D, uint16_t, 1>), dim3(c.batch), dim3(256), lds, s, c, st);
      } else {
        hipLaunchKernelGGL((k_forward<D, uint16_t, 1>), dim3(c.batch), dim3(256), lds, s, c, st);
      }
    } else {
      if constexpr (WideShape<D, float>::OK) {
        if (x2) hipLaunchKernelGGL((k_forward<D, float, 2>), dim3(c.batch), dim3(256), lds, s, c, st);
        else hipLaunchKernelGGL((k_forward<D, float, 1>), dim3(c.batch), dim3(256), lds, s, c, st);
      } else {
        hipLaunchKernelGGL((k_forward<D, float, 1>), dim3(c.batch), dim3(256), lds, s, c, st);
      }
    }
  });
  return hipGetLastError();
}

hipError_t launch_emit(const StepConfig& c, StepState* st, hipStream_t s) {
  const size_t lds = c.grad_mode == 0 ? (size_t)c.hash_entries * sizeof(uint32_t) : 0;
  hipLaunchKernelGGL(k_emit, dim3(c.batch), dim3(256), lds, s, c, st);
  return hipGetLastError();
}

// Digit plan: passes = ceil(bits / 11), RB = ceil(bits / passes).
static void radix_plan(int bits, int* passes, int* rb) {
  *passes = (bits + 10) / 11;
  *rb = (bits + *passes - 1) / *passes;
}

// Blocks of 256 threads of kernel k resident on the current device at once, at most cap.
template <typename Kernel>
static int resident_blocks(Kernel k, int cap) {
  int dev = 0, ncu = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return cap;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0) return cap;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, 256, 0) != hipSuccess || per <= 0) return cap;
  return per * ncu < cap ? per * ncu : cap;
}

template <int RB>
static void launch_sort_rb(const StepConfig& c, StepState* st, hipStream_t s, int passes, int* cur) {
  // whole rounds of resident scatter blocks (at 9-bit digits its tile and counters take
  // 44 KB of LDS, three blocks per CU): a last, partial round would leave part of the
  // device idle; the three kernels of a pass share the block partition
  const int res = resident_blocks(k_rs_scatter<RB>, c.sort_grid);
  const int G = res < c.sort_grid ? (c.sort_grid / res) * res : c.sort_grid;
  for (int p = 0; p < passes; ++p) {
    const int shift = p * RB;
    hipLaunchKernelGGL((k_rs_hist<RB>), dim3(G), dim3(256), 0, s, c.keys[*cur], st, c.ghist, p, shift, G,
                       c.sentinel);
    hipLaunchKernelGGL((k_rs_scan<RB>), dim3(((1u << RB) + NW - 1u) / NW), dim3(256), 0, s, c.ghist, st, p, G);
    hipLaunchKernelGGL((k_rs_scatter<RB>), dim3(G), dim3(256), 0, s, c.keys[*cur], c.slots[*cur], c.keys[1 - *cur],
                       c.slots[1 - *cur], st, c.ghist, p, shift, G, c.sentinel);
    *cur = 1 - *cur;
  }
}

hipError_t launch_sort(const StepConfig& c, StepState* st, hipStream_t s, int* out_buf) {
  int passes, rb, cur = 0;
  radix_plan(c.sort_bits, &passes, &rb);
  switch (rb) {
    case 1: launch_sort_rb<1>(c, st, s, passes, &cur); break;
    case 2: launch_sort_rb<2>(c, st, s, passes, &cur); break;
    case 3: launch_sort_rb<3>(c, st, s, passes, &cur); break;
    case 4: launch_sort_rb<4>(c, st, s, passes, &cur); break;
    case 5: launch_sort_rb<5>(c, st, s, passes, &cur); break;
    case 6: launch_sort_rb<6>(c, st, s, passes, &cur); break;
    case 7: launch_sort_rb<7>(c, st, s, passes, &cur); break;
    case 8: launch_sort_rb<8>(c, st, s, passes, &cur); break;
    case 9: launch_sort_rb<9>(c, st, s, passes, &cur); break;
    case 10: launch_sort_rb<10>(c, st, s, passes, &cur); break;
    case 11: launch_sort_rb<11>(c, st, s, passes, &cur); break;
    default: return hipErrorInvalidValue;
  }
  *out_buf = cur;
  return hipGetLastError();
}

// Values-per-lane variants of k_seg_update available for a row width (VPL | D, D/VPL <= 64).
bool seg_vpl_supported(int d, int vpl) {
  if (vpl != 4 && vpl != 8 && vpl != 12 && vpl != 16) return false;
  return d % vpl == 0 && d / vpl <= 64 && (d % 4 == 0);
}

bool fwd_variant_supported(int d, int bf16, int x) {
  if (x == 1) return true;
  if (x != 2) return false;
  bool ok = false;
  EP_DISPATCH(d, { ok = bf16 ? (WideShape<D, uint16_t>::OK) : (WideShape<D, float>::OK); });
  return ok;
}

bool seg_variant_supported(int d, int vpl, int unroll) {
  if (!seg_vpl_supported(d, vpl)) return false;
  if (vpl == 12) return unroll == 2 || unroll == 4;
  const int nq = vpl / 4;
  return unroll == 8 / nq || unroll == 16 / nq || (nq <= 2 && unroll == 4 / nq);
}

template <int D, int VPL, int U, bool H>
static bool try_seg_update(const StepConfig& c, StepState* st, int buf, hipStream_t s, int grid) {
  if constexpr (D % VPL == 0 && D / VPL <= 64) {
    if (c.seg_vpl == VPL && c.seg_unroll == U) {
      hipLaunchKernelGGL((k_seg_update<D, VPL, U, H>), dim3(grid), dim3(256), 0, s, c, st, c.slots[buf]);
      return true;
    }
  }
  return false;
}

template <int D, bool H>
static void launch_seg_update_variant(const StepConfig& c, StepState* st, int buf, hipStream_t s, int grid) {
  if (c.noise > 1e-9f) {  // the noise path: one tile (every tile gives the same results)
    if constexpr (D % 4 == 0 && D / 4 <= 64)
      hipLaunchKernelGGL((k_seg_update<D, 4, 8, H, true>), dim3(grid), dim3(256), 0, s, c, st, c.slots[buf]);
    else
      hipLaunchKernelGGL((k_seg_update<D, 16, 2, H, true>), dim3(grid), dim3(256), 0, s, c, st, c.slots[buf]);
    return;
  }
  if (try_seg_update<D, 16, 2, H>(c, st, buf, s, grid) || try_seg_update<D, 16, 4, H>(c, st, buf, s, grid) ||
      try_seg_update<D, 12, 2, H>(c, st, buf, s, grid) || try_seg_update<D, 12, 4, H>(c, st, buf, s, grid) ||
      try_seg_update<D, 8, 4, H>(c, st, buf, s, grid) || try_seg_update<D, 8, 8, H>(c, st, buf, s, grid) ||
      try_seg_update<D, 4, 16, H>(c, st, buf, s, grid) || try_seg_update<D, 4, 8, H>(c, st, buf, s, grid) ||
      try_seg_update<D, 4, 4, H>(c, st, buf, s, grid) || try_seg_update<D, 8, 2, H>(c, st, buf, s, grid))
    return;
  // a width no variant divides: the widest built form
  if constexpr (D % 16 == 0 && D / 16 <= 64)
    hipLaunchKernelGGL((k_seg_update<D, 16, 2, H>), dim3(grid), dim3(256), 0, s, c, st, c.slots[buf]);
  else
    hipLaunchKernelGGL((k_seg_update<D, 4, 8, H>), dim3(grid), dim3(256), 0, s, c, st, c.slots[buf]);
}


template <int D, bool H>
static void launch_row_kernels(const StepConfig& c, StepState* st, int buf, hipStream_t s, int grid, hipEvent_t* ev) {
  launch_seg_update_variant<D, H>(c, st, buf, s, grid);
  if (ev) (void)hipEventRecord(ev[1], s);
  // k_pieces sweeps the pieces batch by batch: one round of resident blocks keeps every
  // wave on the same batch
  const int pgrid = resident_blocks(k_pieces<D, H>, grid);
  hipLaunchKernelGGL((k_pieces<D, H>), dim3(pgrid), dim3(256), 0, s, c, st, c.keys[buf], c.slots[buf]);
  if (ev) (void)hipEventRecord(ev[2], s);
  // hot rows dealt to one round of resident blocks: no late blocks holding the last rows
  const int lgrid = resident_blocks(k_long_combine<D, H>, grid);
  hipLaunchKernelGGL((k_long_combine<D, H>), dim3(lgrid), dim3(256), 0, s, c, st, c.keys[buf], c.slots[buf]);
  if (ev) (void)hipEventRecord(ev[3], s);
}

hipError_t launch_scatter_sorted(const StepConfig& c, StepState* st, int buf, hipStream_t s, hipEvent_t* ev) {
  const int grid = c.unique_grid;
  hipLaunchKernelGGL(k_seg_plan, dim3(grid), dim3(256), 0, s, c, st, c.keys[buf], c.slots[buf]);
  if (ev) (void)hipEventRecord(ev[0], s);
  if (c.adjh) {
    // FP16 adjoint rows: the mean model's widths
    EP_DISPATCH(c.dims, { (launch_row_kernels<D, true>)(c, st, buf, s, grid, ev); });
  } else {
    EP_ROW_DISPATCH(c.dims, { (launch_row_kernels<D, false>)(c, st, buf, s, grid, ev); });
  }
  return hipGetLastError();
}

hipError_t launch_atomic_scatter(const StepConfig& c, StepState* st, hipStream_t s) {
  EP_DISPATCH(c.dims, { hipLaunchKernelGGL((k_atomic_scatter<D>), dim3(c.unique_grid), dim3(256), 0, s, c, st); });
  return hipGetLastError();
}

hipError_t launch_adam_touched(const StepConfig& c, StepState* st, hipStream_t s) {
  EP_DISPATCH(c.dims, { hipLaunchKernelGGL((k_adam_touched<D>), dim3(c.unique_grid), dim3(256), 0, s, c, st); });
  return hipGetLastError();
}

hipError_t launch_init_table(const StepConfig& c, uint64_t seed, hipStream_t s) {
  EP_ROW_DISPATCH(c.dims, { hipLaunchKernelGGL((k_init_table<D>), dim3(2048), dim3(256), 0, s, c, seed); });
  return hipGetLastError();
}

hipError_t launch_fill_bf16(const float* src, uint16_t* dst, uint64_t n, uint32_t D, uint32_t stride, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_bf16, dim3(2048), dim3(256), 0, s, src, dst, n, D, stride);
  return hipGetLastError();
}

hipError_t launch_embed_weighted(const StepConfig& c, const uint64_t* findptr, const uint32_t* fids,
                                 const float* p_w, float alpha, uint64_t n_nodes, float* out, hipStream_t s) {
  EP_ROW_DISPATCH(c.dims, {
    if (c.bf16)
      hipLaunchKernelGGL((k_embed_weighted<D, uint16_t>), dim3(2048), dim3(256), 0, s, (const uint16_t*)c.table,
                         findptr, fids, p_w, alpha, n_nodes, out);
    else
      hipLaunchKernelGGL((k_embed_weighted<D, float>), dim3(2048), dim3(256), 0, s, (const float*)c.table, findptr,
                         fids, p_w, alpha, n_nodes, out);
  });
  return hipGetLastError();
}

hipError_t launch_align(int dims, const float* emb, const uint64_t* indptr, const uint32_t* indices, const float* cdf,
                        float alpha, int adaptive, uint32_t max_nb, uint64_t first, uint64_t count, float* out,
                        hipStream_t s) {
  const uint64_t want = (count + NW - 1) / NW + 1;
  const unsigned grid = (unsigned)(want < 4096u ? want : 4096u);
  EP_ROW_DISPATCH(dims, {
    hipLaunchKernelGGL((k_align<D>), dim3(grid), dim3(256), 0, s, emb, indptr, indices, cdf, alpha, adaptive, max_nb,
                       first, count, out);
  });
  return hipGetLastError();
}

hipError_t launch_embed_averaged(const StepConfig& c, const uint64_t* findptr, const uint32_t* fids,
                                 uint64_t n_nodes, float* out, hipStream_t s) {
  EP_ROW_DISPATCH(c.dims, {
    if (c.bf16)
      hipLaunchKernelGGL((k_embed_avg<D, uint16_t>), dim3(2048), dim3(256), 0, s, (const uint16_t*)c.table, findptr,
                         fids, n_nodes, out);
    else
      hipLaunchKernelGGL((k_embed_avg<D, float>), dim3(2048), dim3(256), 0, s, (const float*)c.table, findptr, fids,
                         n_nodes, out);
  });
  return hipGetLastError();
}

}  // namespace ep
