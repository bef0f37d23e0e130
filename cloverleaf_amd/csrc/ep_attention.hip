// Attention aggregator of the EP step (AttentionFeatureModel, src/algos/ep/model.rs:158-249,
// 306-335, 419-446; attention_mean, src/algos/ep/attention.rs:100-131, Full attention
// 171-191, softmax 253-296), gfx950 / wave64.
//
// A table row is [Q_1..Q_H | K_1..K_H | V_1..V_H] (attention.rs:38-66): H heads of d_k
// query and key values and d_model values, width W = H (2 d_k + d_model).
//
// An example's attention blocks are numbered from blkoff[b]: block 0 is h(v), blocks
// 1..cnt are the sampled neighbours of ~h(v) (one attention per neighbour, averaged),
// blocks cnt+1..cnt+K are the negatives h(u_k).  One wavefront per block of n <= 64
// items:
//   h(v), h(u_k)  one item per distinct feature of the node, count 2 x multiplicity
//                 (model.rs:267-268 adds 1 + 1 per occurrence);
//   neighbour     the node's listed features, count = multiplicity in the node
//                 (model.rs:426-443).
// Per head: S_ij = (Q_i . K_j) m_ij / sqrt(d_k), m_ij = c_i c_j when >= 1 (else 1),
// A = row softmax(S) (exact exp), w_j = sum_i A_ij, out_h = (1/n) sum_j w_j V_j; the
// block output is the mean over heads; one item returns V of head 0 (attention.rs:113-115).
//
// The n x n products run on the matrix cores, FP32 in and out
// (v_mfma_f32_32x32x2f32: 32x32 tiles, two k values per instruction, so a d_k <= 32
// contraction is 16 instructions per tile).  Items are padded to 32 or 64 with zero
// rows; padded scores are masked.  Lane l of a tile holds column l % 32 and the 16 rows
// 8 (r / 4) + 4 (l / 32) + r % 4 of its registers r.  Two layouts of the scores:
//   K Q^T  (lane = query i, registers = keys j): the row softmax needs only the partner
//          lane l ^ 32; the probabilities there are the A operand of dQ = dT K;
//   Q K^T  (lane = key j, registers = queries i): column sums w_j need only the
//          partner lane; the probabilities there are the A operand of dK = dT^T Q.
//
// k_att_forward   block outputs
// k_att_loss      one wave per example: a, p = mean of the neighbour blocks, n_k; margin
//                 loss, weighting, adjoints, eligibility (as the mean model's forward)
// k_att_backward  recomputes each eligible block and writes one gradient row of width W
//                 per item at the slot of the item's feature occurrence (key position);
//                 the sorted segment reduce sums a feature's occurrence rows.
#include <algorithm>

#include "ep_device.h"
#include "ep_kernels.h"
#include "ep_loss.h"

namespace ep {

namespace {

constexpr int NW = kBlockWaves;  // waves per 256-thread block (k_att_loss)
constexpr int kAttMaxItems = 64;
constexpr int kTileStride = 34;  // LDS row stride (floats) of the Q / K operand tiles (8-byte rows)

typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ inline f32x16 mfma32(float a, float b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

// Row of register r (0..15) of a 32x32 MFMA result tile held by lane l.
__device__ inline int mrow(int r, int l) { return 8 * (r >> 2) + 4 * (l >> 5) + (r & 3); }

// e^x for the softmax, where x = score - row max <= 0 or -inf: 2^(x log2 e) on the hardware
// exponential (about 1 ulp; results below 2^-126 flush to 0, next to the row max's 1).
__device__ inline float softmax_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }

// 4 x 4 transpose inside every group of 4 lanes (DPP quad permutes, all lanes active):
// lane 4k + m ends with x[j] = what lane 4k + j held in x[m].
template <int CTRL>
__device__ inline float dpp_quad(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ inline void quad_transpose(float& x0, float& x1, float& x2, float& x3) {
  const uint32_t m = lane_id() & 3u;
  const bool b0 = (m & 1u) != 0u, b1 = (m & 2u) != 0u;
  // swap across lane bit 0 (quad_perm [1,0,3,2]) the registers whose bit 0 differs
  float t0 = dpp_quad<0xB1>(b0 ? x0 : x1);
  float t1 = dpp_quad<0xB1>(b0 ? x2 : x3);
  if (b0) { x0 = t0; x2 = t1; } else { x1 = t0; x3 = t1; }
  // then across lane bit 1 (quad_perm [2,3,0,1]) the registers whose bit 1 differs
  t0 = dpp_quad<0x4E>(b1 ? x0 : x2);
  t1 = dpp_quad<0x4E>(b1 ? x1 : x3);
  if (b1) { x0 = t0; x1 = t1; } else { x2 = t0; x3 = t1; }
}

// Store columns [col0, col0 + d_k) of a 32 x 32 result tile (rows: item 32 T + row) into the
// gradient rows of its items: registers 4q .. 4q + 3 hold rows 8q + 4 hf + 0..3 of column
// l % 32; after a quad transpose lane 4k + m holds row 8q + 4 hf + m, columns 4k .. 4k + 3,
// one 16-byte store (row offsets are multiples of 4 floats).
template <int DK, typename RowFn>
__device__ inline void store_tile_cols(f32x16 v, int T, uint32_t n, uint32_t col0, RowFn slot_row) {
  const int l = (int)lane_id(), hf = l >> 5, c32 = l & 31;
  const int col = c32 & ~3;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    float x0 = v[4 * q], x1 = v[4 * q + 1], x2 = v[4 * q + 2], x3 = v[4 * q + 3];
    quad_transpose(x0, x1, x2, x3);
    const uint32_t r = 32u * (uint32_t)T + 8u * (uint32_t)q + 4u * (uint32_t)hf + (uint32_t)(c32 & 3);
    if (col < DK && r < n) *reinterpret_cast<float4*>(slot_row(r) + col0 + col) = make_float4(x0, x1, x2, x3);
  }
}

__device__ inline float count_factor(float ci, float cj) {
  const float mm = ci * cj;
  return mm >= 1.0f ? mm : 1.0f;  // attention.rs:183-186
}

template <typename T>
__device__ inline float ld1(const T* p);
template <>
__device__ inline float ld1<float>(const float* p) { return *p; }
template <>
__device__ inline float ld1<uint16_t>(const uint16_t* p) { return bf16_to_f32(*p); }

// Four consecutive values (offset a multiple of 4).
template <typename T>
__device__ inline float4 ld4(const T* p);
template <>
__device__ inline float4 ld4<float>(const float* p) { return *reinterpret_cast<const float4*>(p); }
template <>
__device__ inline float4 ld4<uint16_t>(const uint16_t* p) {
  const uint2 t = *reinterpret_cast<const uint2*>(p);
  return make_float4(__uint_as_float(t.x << 16), __uint_as_float(t.x & 0xffff0000u), __uint_as_float(t.y << 16),
                     __uint_as_float(t.y & 0xffff0000u));
}

// Eight consecutive values (offset a multiple of 8): two 16-byte loads of FP32, one of BF16.
template <typename T>
__device__ inline void ld8(const T* p, float (&v)[8]);
template <>
__device__ inline void ld8<float>(const float* p, float (&v)[8]) {
  const float4 a = *reinterpret_cast<const float4*>(p);
  const float4 b = *reinterpret_cast<const float4*>(p + 4);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
template <>
__device__ inline void ld8<uint16_t>(const uint16_t* p, float (&v)[8]) {
  const uint4 t = *reinterpret_cast<const uint4*>(p);
  const uint32_t w[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[2 * k] = __uint_as_float(w[k] << 16);
    v[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
  }
}

// LDS of one block's wave (floats / words).  The per-item arrays come first; the last
// region holds the Q / K operand tiles (backward) or the value accumulators (forward).
struct AttLds {
  float* M;      // [64] row max of the scaled scores
  float* iZ;     // [64] 1 / softmax denominator
  float* r;      // [64] sum_j A_ij u_j
  uint32_t* f;   // [64] feature of each item
  float* cnt;    // [64] count of each item
  uint32_t* q;   // [64] position of each item in the node's feature list
  uint32_t* raw; // [64] the node's listed features
  float* u;      // [64][H] dL/dw_jh
  float* w;      // [64][H] column sums w_jh
  float* g;      // [att_dm] block adjoint
  float* Q;      // [64][kTileStride]: Q_h of every item (backward B operand)
  float* K;      // [64][kTileStride]: K_h of every item
  float* acc;    // [H * att_dm] value sweep (forward; aliases Q / K)
  uint32_t win;  // Sliding attention: entries with |i - j| <= win exist; ~0u: Full
  // Random attention: row masks of the current head (nullptr otherwise) and the stream
  // of the block: example position rpos in the pass, pass, block index re in the example
  uint64_t* rmask;
  int rkind;
  uint32_t rk;
  float rp;
  uint32_t rpos, rpass, re, k0, k1;
};

// Entry (i, j) of a block of n items exists: both are items and, under Sliding
// attention, |i - j| <= window_size (compute_sliding_attention_matrix, attention.rs:194-218),
// under Random attention j is one of row i's sampled columns.
// Attention modes of the block kernels (compile-time): Full, Sliding, Random.
enum : int { AM_FULL = 0, AM_SLIDING = 1, AM_RANDOM = 2 };

template <int AM>
__device__ inline bool att_pair(const AttLds& L, uint32_t i, uint32_t j, uint32_t n) {
  if (!(i < n && j < n)) return false;
  if constexpr (AM == AM_RANDOM) return ((L.rmask[i] >> j) & 1ull) != 0ull;
  if constexpr (AM == AM_SLIDING) return (i > j ? i - j : j - i) <= L.win;
  return true;
}

__device__ inline uint32_t att_word(const AttLds& L, uint32_t purpose, uint32_t i) {
  const u32x4 r = philox4x32(u32x4{L.rpos, L.rpass, purpose, i >> 2}, L.k0, L.k1);
  const uint32_t q = i & 3u;
  return q == 0 ? r.x : q == 1 ? r.y : q == 2 ? r.z : r.w;
}

// Random attention for head h of a block of n items (compute_random_attention_matrix,
// attention.rs:221-250): (k, scale) = num_features.sample(n, true) once per head (Fixed:
// min(K, n), n / k; Probability: a Binomial count of n keeps, at least one, 1 / p or
// (1 - p^n) / p), then row i attends to k distinct columns chosen uniformly (Floyd's
// algorithm, lane i); every item's q, k and v are scaled by `scale`, which is returned.
// Fills L.rmask (the wave is the block).
__device__ inline float att_random_head(const AttLds& L, int h, uint32_t n) {
  const uint32_t lane = lane_id();
  const uint32_t tag = (L.re << 8) | ((uint32_t)h << 4);
  uint32_t k;
  float scale;
  if (L.rkind == 1) {
    k = min(L.rk, n);
    scale = (float)n / (float)k;
  } else {
    const bool keep = lane < n && uniform01(att_word(L, RNG_ATT_COUNT | tag, lane)) < L.rp;
    k = (uint32_t)__popcll(__ballot(keep));
    scale = 1.f / L.rp;
    if (k == 0u) {
      k = 1u;
      scale = (1.f - powf(L.rp, (float)n)) / L.rp;
    }
  }
  uint64_t m = 0ull;
  if (lane < n) {
    if (k >= n) {
      m = n >= 64u ? ~0ull : ((1ull << n) - 1ull);
    } else {
      for (uint32_t t = n - k; t < n; ++t) {
        const uint32_t r = uniform_below(att_word(L, RNG_ATT_COLS | tag, lane * 64u + t), t + 1u);
        m |= ((m >> r) & 1ull) ? (1ull << t) : (1ull << r);
      }
    }
  }
  L.rmask[lane] = m;
  __syncthreads();
  return scale;
}

constexpr int kMaxHeads = 16;

// Per-item arrays.  Random attention adds the 64 row masks in front; the backward works
// one head at a time and keeps u for that head only (no w).  LDS per block bounds the
// blocks per CU of the backward: at H = 4, d = 128 it takes 19968 bytes, 8 blocks per CU
// (the VGPR bound), where 20992 would allow 7.
__host__ __device__ inline size_t att_lds_small(int dm, int H, bool rnd, bool backward) {
  return (rnd ? 2 * (size_t)kAttMaxItems : 0) + 7 * (size_t)kAttMaxItems +
         (backward ? (size_t)kAttMaxItems : 2 * (size_t)kAttMaxItems * H) + (size_t)((dm + 3) & ~3);
}

// forward: the value accumulators; backward: the operand tiles
__host__ __device__ inline size_t att_lds_floats(int dm, int H, bool backward, bool rnd) {
  return att_lds_small(dm, H, rnd, backward) + (backward ? 2 * (size_t)kAttMaxItems * kTileStride : (size_t)H * dm);
}

__device__ inline AttLds att_lds(float* sm, int dm, int H, uint32_t win, bool rnd, bool backward = false) {
  AttLds L;
  L.win = win;
  L.rmask = nullptr;  // set by the training kernels under Random attention
  L.rkind = 0;
  L.rk = 0u;
  L.rp = 0.f;
  L.rpos = L.rpass = L.re = L.k0 = L.k1 = 0u;
  L.M = sm + (rnd ? 2 * kAttMaxItems : 0);  // after the 64 row masks (Random attention)
  L.iZ = L.M + kAttMaxItems;
  L.r = L.iZ + kAttMaxItems;
  L.f = (uint32_t*)(L.r + kAttMaxItems);
  L.cnt = (float*)(L.f + kAttMaxItems);
  L.q = (uint32_t*)(L.cnt + kAttMaxItems);
  L.raw = L.q + kAttMaxItems;
  L.u = (float*)(L.raw + kAttMaxItems);
  L.w = backward ? nullptr : L.u + kAttMaxItems * H;
  L.g = backward ? L.u + kAttMaxItems : L.w + kAttMaxItems * H;
  L.Q = sm + att_lds_small(dm, H, rnd, backward);  // 16-byte aligned: the small part is a multiple of 4 floats
  L.K = L.Q + kAttMaxItems * kTileStride;
  L.acc = L.Q;
  return L;
}

// The block a wave works on.
struct AttBlock {
  uint32_t b;      // example in the step
  uint32_t ci;     // construction: 0 a, 1 p, 2+k negative k
  uint32_t node;
  uint32_t n;      // items
  uint32_t slot0;  // key position of the node's first feature occurrence
  bool neighbour;  // block inside ~h(v)
};

// Block g's example, node, slot base (from the block map of k_att_map) and items (into
// LDS).  Returns false for waves past the step's blocks.
__device__ inline bool att_setup(const StepConfig& c, const StepState* st, uint32_t g, const AttLds& L,
                                 AttBlock& B) {
  const uint32_t lane = lane_id();
  const uint32_t nb = c.blkoff[c.batch];
  if (g >= nb) return false;
  const uint32_t meta = c.bmeta[g];
  B.b = meta >> 5;
  B.ci = meta & 31u;
  B.neighbour = B.ci == 1u;
  B.node = c.bnode[g];
  B.slot0 = c.bslot[g];
  const uint64_t fs = c.findptr[B.node];
  const uint32_t nraw = (uint32_t)(c.findptr[B.node + 1] - fs);  // <= 64 (checked by the host)
  const uint32_t fq = lane < nraw ? c.fids[fs + lane] : 0xFFFFFFFFu;
  L.raw[lane] = fq;
  __syncthreads();
  uint32_t mult = 0;
  bool first = true;
  for (uint32_t t = 0; t < nraw; ++t) {
    const bool eq = L.raw[t] == fq;
    mult += eq ? 1u : 0u;
    first = first && !(eq && t < lane);
  }
  const bool sliding = c.att_win != 0xFFFFFFFFu;
  const bool nocount = sliding || c.att_rkind != 0;  // Sliding and Random attention: no count scaling
  if (B.neighbour || sliding) {
    // listed features in order (a repeated feature is repeated); Sliding attention does
    // not scale scores by counts, and count 0 gives the factor 1
    B.n = nraw;
    if (lane >= nraw) L.cnt[lane] = 0.f;  // padded items
    if (lane < nraw) {
      L.f[lane] = fq;
      L.cnt[lane] = nocount ? 0.f : (float)mult;
      L.q[lane] = lane;
    }
  } else {
    const uint64_t m = __ballot(lane < nraw && first);
    B.n = (uint32_t)__popcll(m);
    if (lane >= B.n) L.cnt[lane] = 0.f;  // padded items
    if (lane < nraw && first) {
      const uint32_t i = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
      L.f[i] = fq;
      L.cnt[i] = nocount ? 0.f : 2.0f * (float)mult;
      L.q[i] = lane;
    }
  }
  __syncthreads();
  return true;
}

// Elements 16 hf .. 16 hf + 15 of one head's d_k-wide part (Q_h or K_h) of an item for
// lane half hf as FP32 (zero past d_k and for padded items).
template <int DK, typename T>
__device__ inline void load16(const T* __restrict__ part, bool valid, int hf, float (&x)[16]) {
#pragma unroll
  for (int t = 0; t < 16; ++t) x[t] = 0.f;
  if (!valid) return;
  const int k0 = 16 * hf;
  if constexpr (DK % 8 == 0) {
#pragma unroll
    for (int t = 0; t < 16; t += 8) {
      if (k0 + t < DK) {
        float a[8];
        ld8<T>(part + k0 + t, a);
#pragma unroll
        for (int e = 0; e < 8; ++e) x[t + e] = a[e];
      }
    }
  } else {
#pragma unroll
    for (int t = 0; t < 16; t += 4) {
      if (k0 + t < DK) {
        const float4 a = ld4<T>(part + k0 + t);
        x[t] = a.x; x[t + 1] = a.y; x[t + 2] = a.z; x[t + 3] = a.w;
      }
    }
  }
}

// Score-tile operand of one lane: elements 16 hf .. 16 hf + 15 of an item's Q_h or K_h
// (zero past d_k and for padded items).  FP32 table: 16 floats, the A / B values of 16
// v_mfma_f32_32x32x2f32.  BF16 table: the same 16 values as 8 packed pairs, the operands
// of 2 v_mfma_f32_32x32x16_bf16; the values are bf16 already, so every product is the
// same exact FP32 number and only the FP32 accumulation order differs.
template <typename T>
struct QKop;
template <>
struct QKop<float> {
  float v[16];
  __device__ float at(int t) const { return v[t]; }
};
template <>
struct QKop<uint16_t> {
  uint32_t v[8];
  __device__ float at(int t) const {
    return __uint_as_float((t & 1) ? (v[t >> 1] & 0xffff0000u) : (v[t >> 1] << 16));
  }
};

template <int DK>
__device__ inline void load_op(const float* __restrict__ part, bool valid, int hf, QKop<float>& x) {
  load16<DK, float>(part, valid, hf, x.v);
}
template <int DK>
__device__ inline void load_op(const uint16_t* __restrict__ part, bool valid, int hf, QKop<uint16_t>& x) {
#pragma unroll
  for (int e = 0; e < 8; ++e) x.v[e] = 0u;
  if (!valid) return;
  const int k0 = 16 * hf;
  if constexpr (DK % 8 == 0) {
#pragma unroll
    for (int t = 0; t < 16; t += 8) {
      if (k0 + t < DK) {
        const uint4 r = *reinterpret_cast<const uint4*>(part + k0 + t);
        x.v[t / 2] = r.x; x.v[t / 2 + 1] = r.y; x.v[t / 2 + 2] = r.z; x.v[t / 2 + 3] = r.w;
      }
    }
  } else {
#pragma unroll
    for (int t = 0; t < 16; t += 4) {
      if (k0 + t < DK) {
        const uint2 r = *reinterpret_cast<const uint2*>(part + k0 + t);
        x.v[t / 2] = r.x; x.v[t / 2 + 1] = r.y;
      }
    }
  }
}

// Q_h and K_h operands of item i for lane half hf.
template <int DK, typename T>
__device__ inline void load_qk(const T* __restrict__ row, bool valid, int h, int H, int hf, QKop<T>& q,
                               QKop<T>& k) {
  load_op<DK>(row + h * DK, valid, hf, q);
  load_op<DK>(row + H * DK + h * DK, valid, hf, k);
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// acc + A B^T for 32 x d_k tiles A and B (row: lane % 32 of the operand).  Lane half hf
// holds k-slots 16 hf .. 16 hf + 15 of both operands, the same slots in both, so any
// assignment of slots to the instructions' k positions gives the same sum; slots at or
// past d_k are zero and are skipped.
template <int DK>
__device__ inline f32x16 score_tile(const QKop<float>& a, const QKop<float>& b, f32x16 acc) {
  constexpr int TN = DK < 16 ? DK : 16;
#pragma unroll
  for (int t = 0; t < TN; ++t) acc = mfma32(a.v[t], b.v[t], acc);
  return acc;
}
template <int DK>
__device__ inline f32x16 score_tile(const QKop<uint16_t>& a, const QKop<uint16_t>& b, f32x16 acc) {
  constexpr int SN = DK <= 8 ? 1 : 2;
#pragma unroll
  for (int s = 0; s < SN; ++s) {
    const uint4 ra = make_uint4(a.v[4 * s], a.v[4 * s + 1], a.v[4 * s + 2], a.v[4 * s + 3]);
    const uint4 rb = make_uint4(b.v[4 * s], b.v[4 * s + 1], b.v[4 * s + 2], b.v[4 * s + 3]);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, ra), __builtin_bit_cast(bf16x8, rb),
                                                  acc, 0, 0, 0);
  }
  return acc;
}

// Row statistics of query tile I from the K Q^T tiles: scaled, masked scores into T,
// returns the row max and 1 / denominator of query i = 32 I + l % 32 (both halves).
template <int AM>
__device__ inline void row_softmax_stats(f32x16 (&T)[2], int NT, uint32_t n, int I, float ci, const AttLds& L,
                                         float isq, float* mx_out, float* iz_out) {
  const int l = (int)lane_id();
  const uint32_t i = 32u * I + (l & 31);
  float mx = -INFINITY;
  #pragma unroll
  for (int J = 0; J < 2; ++J) {
    if (J >= NT) break;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const uint32_t j = 32u * J + mrow(r, l);
      float s = T[J][r] * count_factor(ci, L.cnt[j]) * isq;
      s = att_pair<AM>(L, i, j, n) ? s : -INFINITY;
      T[J][r] = s;
      mx = fmaxf(mx, s);
    }
  }
  mx = fmaxf(mx, __shfl_xor(mx, 32, EP_WAVE));
  const float m0 = (i < n) ? mx : 0.f;
  float z = 0.f;
  #pragma unroll
  for (int J = 0; J < 2; ++J) {
    if (J >= NT) break;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const float e = softmax_exp(T[J][r] - m0);  // masked scores give exp(-inf) = 0
      T[J][r] = e;
      z += e;
    }
  }
  z += __shfl_xor(z, 32, EP_WAVE);
  *mx_out = m0;
  *iz_out = (i < n) ? 1.0f / z : 0.f;
}

}  // namespace

// ------------------------------------------------------------------------------------
// k_att_map: one wave per example; for each of its attention blocks the node, the key
// position of the node's first feature occurrence (the example's occurrences are listed
// node by node in block order from keyoff[b]) and example << 5 | construction.
// ------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_att_map(StepConfig c, StepState* st) {
  const uint32_t lane = lane_id();
  const uint32_t b = blockIdx.x * NW + (threadIdx.x >> 6);
  if (b >= c.batch || b >= st->batch_n) return;
  const uint32_t cnt = c.nbr_cnt[b];
  if (cnt == 0u) return;
  const uint32_t K = c.K, E = 1u + cnt + K;
  const uint32_t g0 = c.blkoff[b];
  if (g0 + E > c.blkoff[c.batch]) return;  // over capacity: the step reports an error
  const uint32_t v = c.perm[st->anchor_base + b];
  const uint32_t* pn = nbr_list(c, b, c.indptr[v]);
  uint32_t carry = c.keyoff[b];
  for (uint32_t j0 = 0; j0 < E; j0 += 64u) {
    const uint32_t j = j0 + lane;
    uint32_t node = 0, len = 0, ci = 0;
    if (j < E) {
      node = j == 0u ? v : (j <= cnt ? pn[j - 1u] : c.neg[(size_t)b * K + (j - 1u - cnt)]);
      len = (uint32_t)(c.findptr[node + 1] - c.findptr[node]);
      ci = j == 0u ? 0u : (j <= cnt ? 1u : 2u + (j - 1u - cnt));
    }
    uint32_t inc = len;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(inc, off, EP_WAVE);
      if ((int)lane >= off) inc += y;
    }
    if (j < E) {
      c.bnode[g0 + j] = node;
      c.bslot[g0 + j] = carry + inc - len;
      c.bmeta[g0 + j] = (b << 5) | ci;
    }
    carry += __shfl(inc, 63, EP_WAVE);
  }
}

// ------------------------------------------------------------------------------------
// Output of one attention block whose n items (features f, counts cnt) are in LDS:
// per head scores on the matrix cores, softmax statistics and column sums w_jh (LDS),
// each head's Q / K loaded at its start; then one sweep over the items' value parts
// (all heads of an item are contiguous): out = sum_j sum_h w_jh V_jh / (n H).
template <int DK, typename T, int AM>
__device__ inline void att_block_output(const T* __restrict__ table, uint32_t W, int H, int dm, const AttLds& L,
                                        uint32_t n, float* __restrict__ out) {
  const int l = (int)lane_id(), hf = l >> 5, c32 = l & 31;
  const uint32_t vbase = 2u * H * DK;  // first V column
  if (n == 1u) {  // attention.rs:113-115: the value vector of head 0
    const T* row = table + (size_t)L.f[0] * W + vbase;
    for (int t = l; t < dm; t += 64) out[t] = ld1<T>(row + t);
    return;
  }
  const float isq = 1.0f / sqrtf((float)DK);
  const int NT = (int)((n + 31u) / 32u);
  // Each head's Q / K operands load at the head's start: prefetching the next head's
  // holds 64 more VGPRs and drops the kernel from 3 to 2 waves per SIMD, which measured
  // slower.
  QKop<T> qa[2], ka[2];
  for (int h = 0; h < H; ++h) {
    // Random attention: this head's row masks and item scale (scores scale by its square)
    float hs = 1.f;
    if constexpr (AM == AM_RANDOM) hs = att_random_head(L, h, n);
    const float isq_h = isq * hs * hs;
#pragma unroll
    for (int I = 0; I < 2; ++I) {
      const uint32_t i = 32u * I + c32;
      const bool ok = I < NT && i < n;
      load_qk<DK, T>(table + (size_t)(ok ? L.f[i] : 0u) * W, ok, h, H, hf, qa[I], ka[I]);
    }
    // softmax statistics per query row from the K Q^T tiles
#pragma unroll
    for (int I = 0; I < 2; ++I) {
      if (I >= NT) break;
      f32x16 Tt[2];
#pragma unroll
      for (int J = 0; J < 2; ++J) {
        if (J >= NT) break;
        Tt[J] = score_tile<DK>(ka[J], qa[I], f32x16{});
      }
      const uint32_t i = 32u * I + c32;
      const float ci = i < n ? L.cnt[i] : 0.f;
      float mx, iz;
      row_softmax_stats<AM>(Tt, NT, n, I, ci, L, isq_h, &mx, &iz);
      if (hf == 0) { L.M[i] = mx; L.iZ[i] = iz; }
    }
    __syncthreads();
    // column sums w_jh from the Q K^T tiles
#pragma unroll
    for (int J = 0; J < 2; ++J) {
      if (J >= NT) break;
      const uint32_t j = 32u * J + c32;
      const float cj = j < n ? L.cnt[j] : 0.f;
      float wj = 0.f;
#pragma unroll
      for (int I = 0; I < 2; ++I) {
        if (I >= NT) break;
        const f32x16 a = score_tile<DK>(qa[I], ka[J], f32x16{});
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const uint32_t i = 32u * I + mrow(r, l);
          const float sc = a[r] * count_factor(L.cnt[i], cj) * isq_h;
          wj += att_pair<AM>(L, i, j, n) ? softmax_exp(sc - L.M[i]) * L.iZ[i] : 0.f;
        }
      }
      wj += __shfl_xor(wj, 32, EP_WAVE);
      if (hf == 0) L.w[j * H + h] = j < n ? wj * hs : 0.f;
    }
    __syncthreads();  // the next head rewrites the row statistics read above
  }
  // value sweep: lane l covers 8 consecutive values p of an item's value part
  const int VP = H * dm;
  float* sacc = L.acc;
  for (int c0 = 0; c0 < VP; c0 += 512) {
    const int p = c0 + 8 * l;
    const bool pv = p < VP;
    const int hh = pv ? p / dm : 0;
    float a8[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (uint32_t j0 = 0; j0 < n; j0 += 4u) {
      float v[4][8];
      float w[4];
      bool okk[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t j = j0 + (uint32_t)u;
        okk[u] = pv && j < n;
        w[u] = okk[u] ? L.w[j * H + hh] : 0.f;
        const T* row = table + (size_t)L.f[okk[u] ? j : 0u] * W + vbase + (okk[u] ? p : 0);
        ld8<T>(row, v[u]);  // an in-range row when !okk (not added)
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int e = 0; e < 8; ++e) a8[e] += okk[u] ? w[u] * v[u][e] : 0.f;
    }
    if (pv) {
      *reinterpret_cast<float4*>(sacc + p) = make_float4(a8[0], a8[1], a8[2], a8[3]);
      *reinterpret_cast<float4*>(sacc + p + 4) = make_float4(a8[4], a8[5], a8[6], a8[7]);
    }
  }
  __syncthreads();
  const float scale = 1.0f / ((float)n * (float)H);
  for (int d = l; d < dm; d += 64) {
    float a = 0.f;
    for (int h = 0; h < H; ++h) a += sacc[h * dm + d];
    out[d] = a * scale;
  }
}


// Random attention stream of block g (example B.b) for the training kernels.
__device__ inline void att_random_setup(const StepConfig& c, const StepState* st, uint32_t g, const AttBlock& B,
                                        float* sm, AttLds& L) {
  if (!c.att_rkind) return;
  L.rmask = (uint64_t*)sm;
  L.rkind = c.att_rkind;
  L.rk = c.att_rk;
  L.rp = c.att_rp;
  L.rpos = st->anchor_base + B.b;
  L.rpass = st->pass;
  L.re = g - c.blkoff[B.b];
  L.k0 = (uint32_t)c.seed;
  L.k1 = (uint32_t)(c.seed >> 32);
}

// ------------------------------------------------------------------------------------
// k_att_forward: one wave (64-thread block) per attention block.
// ------------------------------------------------------------------------------------
template <int DK, typename T, int AM>
__global__ __launch_bounds__(64) void k_att_forward(StepConfig c, StepState* st) {
  extern __shared__ float sm[];
  const int H = c.att_heads, dm = c.att_dm;
  AttLds L = att_lds(sm, dm, H, c.att_win, AM == AM_RANDOM);
  AttBlock B;
  if (!att_setup(c, st, blockIdx.x, L, B)) return;
  att_random_setup(c, st, blockIdx.x, B, sm, L);
  if (lane_id() == 0) c.bitems[blockIdx.x] = B.n;
  att_block_output<DK, T, AM>((const T*)c.table, (uint32_t)c.dims, H, dm, L, B.n,
                               c.bout + (size_t)blockIdx.x * dm);
}

// ------------------------------------------------------------------------------------
// k_att_embed: AttentionAggregator::construct (aggregator.rs:137-153) for inference, one
// wave per node: Full attention over the node's listed features with count 1 each.
// Nodes without features keep a zero row.
// ------------------------------------------------------------------------------------
template <int DK, typename T, int AM>
__global__ __launch_bounds__(64) void k_att_embed(const T* __restrict__ table, uint32_t W, int H, int dm,
                                                  const uint64_t* __restrict__ findptr,
                                                  const uint32_t* __restrict__ fids, uint64_t n_nodes,
                                                  uint32_t win, float* __restrict__ out) {
  extern __shared__ float sm[];
  const AttLds L = att_lds(sm, dm, H, win, false);
  const uint32_t lane = lane_id();
  for (uint64_t node = blockIdx.x; node < n_nodes; node += gridDim.x) {
    const uint64_t fs = findptr[node];
    const uint32_t n = (uint32_t)(findptr[node + 1] - fs);  // <= 64 (checked by the host)
    float* o = out + node * (uint64_t)dm;
    if (n == 0u) {
      for (int t = lane; t < dm; t += 64) o[t] = 0.f;
      continue;
    }
    L.f[lane] = lane < n ? fids[fs + lane] : 0u;
    L.cnt[lane] = lane < n ? 1.0f : 0.f;
    __syncthreads();
    att_block_output<DK, T, AM>(table, W, H, dm, L, n, o);
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------
// k_att_loss: one wave per example.  a = block 0, p = mean of the neighbour blocks,
// n_k = the negative blocks; margin loss (loss.rs:59-73), weighting (mod.rs:190-200),
// adjoints of every construction and their eligibility (finite, in the loss graph).
// Block adjoints: a and n_k get dL/dx, each neighbour block dL/dp / cnt.
// ------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_att_loss(StepConfig c, StepState* st) {
  const uint32_t lane = lane_id();
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t b = blockIdx.x * NW + w;
  const uint32_t K = c.K, NC = K + 2u;
  const int dm = c.att_dm;
  const bool inb = b < c.batch && b < st->batch_n;
  const uint32_t cnt = inb ? c.nbr_cnt[b] : 0u;
  const uint32_t g0 = inb ? c.blkoff[b] : 0u;
  // blocks over capacity (the step reports an error) count as no example
  const bool ex = inb && cnt > 0u && g0 + 1u + cnt + K <= c.blkoff[c.batch];
  {
    // attention work of the block's examples: sum over attention blocks of items^2 and
    // items, one atomic pair per 256-thread block
    __shared__ unsigned long long s_stat[2][NW];
    uint32_t s2 = 0, s1 = 0;
    if (ex)
      for (uint32_t j = lane; j < 1u + cnt + K; j += 64u) {
        const uint32_t nj = c.bitems[g0 + j];
        s2 += nj * nj;
        s1 += nj;
      }
    s2 = wave_sum_u32(s2);
    s1 = wave_sum_u32(s1);
    if (lane == 0) { s_stat[0][w] = s2; s_stat[1][w] = s1; }
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long a2 = 0, a1 = 0;
      for (int q = 0; q < NW; ++q) { a2 += s_stat[0][q]; a1 += s_stat[1][q]; }
      if (a1) {
        atomicAdd((unsigned long long*)&st->stat_att_n2, a2);
        atomicAdd((unsigned long long*)&st->stat_att_n, a1);
      }
    }
  }
  if (!inb) return;
  if (!ex) {  // no neighbourhood: no example
    if (lane == 0) {
      c.loss[b] = 0.f;
      if (c.pass_loss) c.pass_loss[st->anchor_base + b] = 0.f;
    }
    if (lane < NC) c.elig[(size_t)b * NC + lane] = 0;
    return;
  }
  const uint32_t v = c.perm[st->anchor_base + b];
  const uint32_t deg = (uint32_t)(c.indptr[v + 1] - c.indptr[v]);
  float a[4], p[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int t = (int)lane + 64 * r;
    a[r] = p[r] = 0.f;
    if (t < dm) {
      a[r] = c.bout[(size_t)g0 * dm + t];
      float s = 0.f;
      for (uint32_t j = 0; j < cnt; ++j) s += c.bout[(size_t)(g0 + 1u + j) * dm + t];
      p[r] = s / (float)cnt;
    }
  }
  auto neg_val = [&](uint32_t k, int r) -> float {
    const int t = (int)lane + 64 * r;
    return t < dm ? c.bout[(size_t)(g0 + 1u + cnt + k) * dm + t] : 0.f;
  };
  if (c.loss_kind != 0 && c.loss_kind != 5) {
    // losses other than margin: norms and dots, coefficients (lane 0), adjoints
    __shared__ float s_nrm[NW][32], s_dan[NW][32], s_ca[NW][32], s_cns[NW][32], s_cna[NW][32], s_cp[NW][2];
    float* nrm = s_nrm[w];
    float* dan = s_dan[w];
    float* ca = s_ca[w];
    float* cns = s_cns[w];
    float* cna = s_cna[w];
    float* cp = s_cp[w];
    float qa = 0.f, qp = 0.f, dap = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) { qa += a[r] * a[r]; qp += p[r] * p[r]; dap += a[r] * p[r]; }
    qa = wave_sum(qa); qp = wave_sum(qp); dap = wave_sum(dap);
    if (lane == 0) { nrm[0] = sqrtf(qa); nrm[1] = sqrtf(qp); }
    for (uint32_t k = 0; k < K; ++k) {
      float qn = 0.f, dn = 0.f;
#pragma unroll
      for (int r = 0; r < 4; ++r) { const float x = neg_val(k, r); qn += x * x; dn += a[r] * x; }
      qn = wave_sum(qn); dn = wave_sum(dn);
      if (lane == 0) { nrm[2 + k] = sqrtf(qn); dan[k] = dn; }
    }
    __builtin_amdgcn_wave_barrier();
    float Lo = 0.f;
    uint32_t reach = 0;
    if (lane == 0) Lo = other_loss(c.loss_kind, c.gamma, c.loss_p2, K, nrm, dap, dan, ca, cp, cns, cna, &reach);
    __builtin_amdgcn_wave_barrier();
    Lo = __shfl(Lo, 0, EP_WAVE);
    reach = __shfl(reach, 0, EP_WAVE);
    float div = 1.f;
    if (c.lw_kind == 1) div = logf(1.f + (float)deg);
    else if (c.lw_kind == 2) div = powf((float)deg, c.lw_exp);
    const float Lw = Lo / div;
    const bool contrib = (Lw > 0.f) && !isnan(Lw);
    if (!contrib) reach = 0u;
    if (lane == 0) {
      c.loss[b] = contrib ? Lw : 0.f;
      if (c.pass_loss) c.pass_loss[st->anchor_base + b] = contrib ? Lw : 0.f;
    }
    if (c.xout) {
      float* xo = c.xout + (size_t)b * NC * dm;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int t = (int)lane + 64 * r;
        if (t < dm) {
          xo[t] = a[r];
          xo[dm + t] = p[r];
          for (uint32_t k = 0; k < K; ++k) xo[(size_t)(2u + k) * dm + t] = neg_val(k, r);
        }
      }
    }
    if (!c.train) return;
    const float scale = 1.f / div;
    for (uint32_t ci = 0; ci < NC; ++ci) {
      bool e = (reach >> ci) & 1u;
      float g[4];
      bool ok = true;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float x;
        if (ci == 0u) {
          x = ca[0] * a[r] + ca[1] * p[r];
          for (uint32_t k = 0; k < K; ++k) x += ca[2 + k] * neg_val(k, r);
        } else if (ci == 1u) {
          x = cp[0] * a[r] + cp[1] * p[r];
        } else {
          x = cns[ci - 2u] * neg_val(ci - 2u, r) + cna[ci - 2u] * a[r];
        }
        g[r] = x * scale;
        if ((int)lane + 64 * r < dm) ok = ok && isfinite(g[r]);
      }
      e = e && !wave_any(!ok);
      if (e) {
        // block adjoints: a and n_k directly, each neighbour block dL/dp / cnt
        const uint32_t first = ci == 0u ? g0 : (ci == 1u ? g0 + 1u : g0 + 1u + cnt + (ci - 2u));
        const uint32_t nb = ci == 1u ? cnt : 1u;
        const float f = ci == 1u ? 1.0f / (float)cnt : 1.0f;
        for (uint32_t j = 0; j < nb; ++j) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int t = (int)lane + 64 * r;
            if (t < dm) c.badj[(size_t)(first + j) * dm + t] = g[r] * f;
          }
        }
      }
      if (lane == 0) c.elig[(size_t)b * NC + ci] = e ? 1 : 0;
    }
    return;
  }
  float sq = 0.f;
#pragma unroll
  for (int r = 0; r < 4; ++r) sq += (p[r] - a[r]) * (p[r] - a[r]);
  const float D0 = sqrtf(wave_sum(sq));
  // distances to the negatives, lane k keeps D_k
  float Dk_mine = 0.f;
  for (uint32_t k = 0; k < K; ++k) {
    float s2 = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float d = p[r] - neg_val(k, r);
      s2 += d * d;
    }
    const float Dk = sqrtf(wave_sum(s2));
    if (lane == k) Dk_mine = Dk;
  }
  const float d1 = c.gamma + D0;
  const float lk = d1 - Dk_mine;
  const bool actk = lane < K && lk > 0.f;
  const uint64_t actm = __ballot(actk);
  const uint32_t nA = (uint32_t)__popcll(actm);
  // sum of active losses in lane order (the reference adds them in negative order)
  float sum = 0.f;
  for (uint32_t k = 0; k < K; ++k) {
    const float lkk = __shfl(lk, (int)k, EP_WAVE);
    if ((actm >> k) & 1ull) sum += lkk;
  }
  const float Lm = nA ? sum / (float)nA : 0.f;
  float div = 1.f;
  if (c.lw_kind == 1) div = logf(1.f + (float)deg);
  else if (c.lw_kind == 2) div = powf((float)deg, c.lw_exp);
  const float Lw = Lm / div;
  const bool contrib = (Lw > 0.f) && !isnan(Lw);
  if (lane == 0) {
    c.loss[b] = contrib ? Lw : 0.f;
    if (c.pass_loss) c.pass_loss[st->anchor_base + b] = contrib ? Lw : 0.f;
  }
  if (c.xout) {
    float* xo = c.xout + (size_t)b * NC * dm;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int t = (int)lane + 64 * r;
      if (t < dm) {
        xo[t] = a[r];
        xo[dm + t] = p[r];
        for (uint32_t k = 0; k < K; ++k) xo[(size_t)(2u + k) * dm + t] = neg_val(k, r);
      }
    }
  }
  if (!c.train) return;
  const float invA = nA ? 1.f / (float)nA : 0.f, scale = 1.f / div;
  // a
  {
    bool e = contrib;
    float g[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      g[r] = (a[r] - p[r]) / D0 * scale;
      if ((int)lane + 64 * r < dm) e = e && isfinite(g[r]);
    }
    e = contrib && !wave_any(contrib && !e);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int t = (int)lane + 64 * r;
      if (e && t < dm) c.badj[(size_t)g0 * dm + t] = g[r];
    }
    if (lane == 0) c.elig[(size_t)b * NC] = e ? 1 : 0;
  }
  // p (each neighbour block gets dL/dp / cnt)
  {
    float g[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) g[r] = (p[r] - a[r]) / D0;
    for (uint32_t k = 0; k < K; ++k) {
      if (!((actm >> k) & 1ull)) continue;
      const float Dk = __shfl(Dk_mine, (int)k, EP_WAVE);
#pragma unroll
      for (int r = 0; r < 4; ++r) g[r] -= (p[r] - neg_val(k, r)) / Dk * invA;
    }
    bool ok = true;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      g[r] *= scale;
      if ((int)lane + 64 * r < dm) ok = ok && isfinite(g[r]);
    }
    const bool e = contrib && !wave_any(!ok);
    const float inv_cnt = 1.0f / (float)cnt;
    if (e) {
      for (uint32_t j = 0; j < cnt; ++j) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int t = (int)lane + 64 * r;
          if (t < dm) c.badj[(size_t)(g0 + 1u + j) * dm + t] = g[r] * inv_cnt;
        }
      }
    }
    if (lane == 0) c.elig[(size_t)b * NC + 1] = e ? 1 : 0;
  }
  // negatives
  for (uint32_t k = 0; k < K; ++k) {
    const bool act = contrib && ((actm >> k) & 1ull);
    bool e = act;
    if (act) {
      const float Dk = __shfl(Dk_mine, (int)k, EP_WAVE);
      float g[4];
      bool ok = true;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        g[r] = (p[r] - neg_val(k, r)) / Dk * invA * scale;
        if ((int)lane + 64 * r < dm) ok = ok && isfinite(g[r]);
      }
      e = !wave_any(!ok);
      if (e) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int t = (int)lane + 64 * r;
          if (t < dm) c.badj[(size_t)(g0 + 1u + cnt + k) * dm + t] = g[r];
        }
      }
    } else {
      (void)__shfl(Dk_mine, (int)k, EP_WAVE);  // keep the wave's shuffles uniform
    }
    if (lane == 0) c.elig[(size_t)b * NC + 2u + k] = e ? 1 : 0;
  }
}

// ------------------------------------------------------------------------------------
// k_att_backward: one wave per attention block of an eligible construction.  Writes the
// gradient row [dQ | dK | dV] of every item at adj[slot * W], slot = the key position
// of the item's feature occurrence.  Per head, with g the block adjoint:
//   u_j = (V_j . g) / (n H)          dL/dw_j
//   dS_ij = A_ij (u_j - r_i), r_i = sum_k A_ik u_k,   dT = dS m / sqrt(d_k)
//   dQ = dT K,  dK = dT^T Q  (matrix cores),  dV_j = w_j g / (n H)
// ------------------------------------------------------------------------------------
template <int DK, typename T, int AM>
__global__ __launch_bounds__(64) void k_att_backward(StepConfig c, StepState* st) {
  extern __shared__ float sm[];
  const int H = c.att_heads, dm = c.att_dm;
  const uint32_t W = (uint32_t)c.dims;
  AttLds L = att_lds(sm, dm, H, c.att_win, AM == AM_RANDOM, true);
  if (st->n_keys > c.key_cap) return;
  AttBlock B;
  if (!att_setup(c, st, blockIdx.x, L, B)) return;
  att_random_setup(c, st, blockIdx.x, B, sm, L);
  const uint32_t NC = c.K + 2u;
  if (!c.elig[(size_t)B.b * NC + B.ci]) return;
  const int l = (int)lane_id(), hf = l >> 5, c32 = l & 31;
  const T* table = (const T*)c.table;
  const uint32_t n = B.n;
  for (int t = l; t < dm; t += 64) L.g[t] = c.badj[(size_t)blockIdx.x * dm + t];
  __syncthreads();
  const uint32_t qoff = 2u * H * DK;  // first V column
  if (n == 1u) {
    float* grow = c.adj + (size_t)(B.slot0 + L.q[0]) * W;
    for (uint32_t t = l; t < W; t += 64u) {
      const uint32_t vt = t - qoff;
      grow[t] = (t >= qoff && vt < (uint32_t)dm) ? L.g[vt] : 0.f;
    }
    return;
  }
  const float isq = 1.0f / sqrtf((float)DK);
  const float scale = 1.0f / ((float)n * (float)H);
  const int NT = (int)((n + 31u) / 32u);
  auto slot_row = [&](uint32_t i) -> float* { return c.adj + (size_t)(B.slot0 + L.q[i]) * W; };
  for (int h = 0; h < H; ++h) {
    float hs = 1.f;  // Random attention: this head's row masks and item scale
    if constexpr (AM == AM_RANDOM) hs = att_random_head(L, h, n);
    const float isq_h = isq * hs * hs;
    QKop<T> qa[2], ka[2];
#pragma unroll
    for (int I = 0; I < 2; ++I) {
      const uint32_t i = 32u * I + c32;
      const bool ok = I < NT && i < n;
      load_qk<DK, T>(table + (size_t)(ok ? L.f[i] : 0u) * W, ok, h, H, hf, qa[I], ka[I]);
      if (I < NT) {
#pragma unroll
        for (int t = 0; t < 16; t += 2) {
          *reinterpret_cast<float2*>(L.Q + (size_t)i * kTileStride + 16 * hf + t) =
              make_float2(qa[I].at(t), qa[I].at(t + 1));
          *reinterpret_cast<float2*>(L.K + (size_t)i * kTileStride + 16 * hf + t) =
              make_float2(ka[I].at(t), ka[I].at(t + 1));
        }
      }
    }
    // u_j: 8-lane groups, item j0 + 8 e + l / 8 per group; each lane loads 16 bytes (4
    // FP32 or 8 BF16 values), so one load instruction reads 8 contiguous 128-byte segments;
    // two items per group and a 128-value chunk of each are in flight (padded items read
    // row 0, unused)
    {
      constexpr int VE = std::is_same<T, uint16_t>::value ? 8 : 4;  // values per lane and load
      constexpr int NU = 128 / (8 * VE);                            // loads per lane and chunk
      const int gi = l >> 3, li = l & 7;
      for (uint32_t j0 = 0; j0 < 32u * (uint32_t)NT; j0 += 16u) {
        float p[2] = {0.f, 0.f};
        for (int t0 = 0; t0 < dm; t0 += 128) {
          float vv[2][NU][VE];
#pragma unroll
          for (int e = 0; e < 2; ++e) {
            const uint32_t j = j0 + 8u * (uint32_t)e + (uint32_t)gi;
            const T* vrow = table + (size_t)(j < n ? L.f[j] : 0u) * W + qoff + (size_t)h * dm;
#pragma unroll
            for (int u = 0; u < NU; ++u) {
              const int t = t0 + VE * li + 8 * VE * u;  // d_model is a multiple of 8
              const bool ok = j < n && t < dm;
              if constexpr (VE == 8) {
                if (ok) {
                  ld8<T>(vrow + t, vv[e][u]);
                } else {
#pragma unroll
                  for (int x = 0; x < VE; ++x) vv[e][u][x] = 0.f;
                }
              } else {
                const float4 v4 = ok ? ld4<T>(vrow + t) : make_float4(0.f, 0.f, 0.f, 0.f);
                vv[e][u][0] = v4.x; vv[e][u][1] = v4.y; vv[e][u][2] = v4.z; vv[e][u][3] = v4.w;
              }
            }
          }
#pragma unroll
          for (int e = 0; e < 2; ++e) {
#pragma unroll
            for (int u = 0; u < NU; ++u) {
              const int t = t0 + VE * li + 8 * VE * u;
              if (t < dm) {
#pragma unroll
                for (int x = 0; x < VE; x += 4) {
                  const float4 g4 = *reinterpret_cast<const float4*>(L.g + t + x);
                  p[e] += vv[e][u][x] * g4.x + vv[e][u][x + 1] * g4.y + vv[e][u][x + 2] * g4.z + vv[e][u][x + 3] * g4.w;
                }
              }
            }
          }
        }
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          float q = p[e];
          q += __shfl_xor(q, 1, EP_WAVE);
          q += __shfl_xor(q, 2, EP_WAVE);
          q += __shfl_xor(q, 4, EP_WAVE);
          const uint32_t j = j0 + 8u * (uint32_t)e + (uint32_t)gi;
          if (li == 0) L.u[j] = j < n ? q * scale * hs : 0.f;
        }
      }
    }
    __syncthreads();
    // query tiles (K Q^T layout): probabilities, r_i, dT, dQ = dT K
    #pragma unroll
    for (int I = 0; I < 2; ++I) {
      if (I >= NT) break;
      f32x16 Tt[2];
      #pragma unroll
      for (int J = 0; J < 2; ++J) {
        if (J >= NT) break;
        Tt[J] = score_tile<DK>(ka[J], qa[I], f32x16{});
      }
      const uint32_t i = 32u * I + c32;
      const float ci = i < n ? L.cnt[i] : 0.f;
      float mx, iz;
      row_softmax_stats<AM>(Tt, NT, n, I, ci, L, isq_h, &mx, &iz);
      float ri = 0.f;
      #pragma unroll
      for (int J = 0; J < 2; ++J) {
        if (J >= NT) break;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const uint32_t j = 32u * J + mrow(r, l);
          Tt[J][r] *= iz;  // A_ij
          ri += Tt[J][r] * L.u[j];
        }
      }
      ri += __shfl_xor(ri, 32, EP_WAVE);
      if (hf == 0) { L.M[i] = mx; L.iZ[i] = iz; L.r[i] = ri; }
      f32x16 dq = {};
      #pragma unroll
      for (int J = 0; J < 2; ++J) {
        if (J >= NT) break;
#pragma unroll
        for (int t = 0; t < 16; ++t) {
          const uint32_t j = 32u * J + mrow(t, l);
          const float dT = Tt[J][t] * (L.u[j] - ri) * count_factor(ci, L.cnt[j]) * isq_h;
          dq = mfma32(dT, L.K[(size_t)j * kTileStride + c32], dq);
        }
      }
      store_tile_cols<DK>(dq, I, n, (uint32_t)(h * DK), slot_row);
    }
    __syncthreads();
    // key tiles (Q K^T layout): w_j, dT, dK = dT^T Q, dV_j = w_j g / (n H)
    #pragma unroll
    for (int J = 0; J < 2; ++J) {
      if (J >= NT) break;
      const uint32_t j = 32u * J + c32;
      const float cj = j < n ? L.cnt[j] : 0.f;
      const float uj = j < n ? L.u[j] : 0.f;
      float wj = 0.f;
      f32x16 dk = {};
      #pragma unroll
      for (int I = 0; I < 2; ++I) {
        if (I >= NT) break;
        f32x16 a = score_tile<DK>(qa[I], ka[J], f32x16{});
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const uint32_t i = 32u * I + mrow(r, l);
          const float f = count_factor(L.cnt[i], cj) * isq_h;
          const float A = att_pair<AM>(L, i, j, n) ? softmax_exp(a[r] * f - L.M[i]) * L.iZ[i] : 0.f;
          wj += A;
          a[r] = A * (uj - L.r[i]) * f;  // dT_ij
        }
#pragma unroll
        for (int t = 0; t < 16; ++t) dk = mfma32(a[t], L.Q[(size_t)(32u * I + mrow(t, l)) * kTileStride + c32], dk);
      }
      wj += __shfl_xor(wj, 32, EP_WAVE);
      store_tile_cols<DK>(dk, J, n, (uint32_t)(H * DK + h * DK), slot_row);
      // dV rows, two items per pass: lane half hf writes item 32 J + jj + hf, 4 values per
      // lane (d_model and the row offsets are multiples of 8 floats, g is 16-byte aligned)
      const uint32_t jn = min(32u, n - 32u * J);
      for (uint32_t jj = 0; jj < jn; jj += 2u) {
        const uint32_t jx = jj + (uint32_t)hf;
        const float cw = __shfl(wj, (int)min(jx, 31u), EP_WAVE) * scale * hs;
        if (jx < jn) {
          float* vr = slot_row(32u * J + jx) + qoff + (size_t)h * dm;
          for (int t = 4 * c32; t < dm; t += 128) {
            const float4 g4 = *reinterpret_cast<const float4*>(L.g + t);
            *reinterpret_cast<float4*>(vr + t) = make_float4(cw * g4.x, cw * g4.y, cw * g4.z, cw * g4.w);
          }
        }
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------------------
bool att_shape_supported(int heads, int dk, int dm) {
  return heads >= 1 && heads <= kMaxHeads && (dk == 4 || dk == 8 || dk == 16 || dk == 32) && dm >= 8 && dm <= 256 &&
         dm % 8 == 0;
}

// Attention mode of the block kernels: Random (max_features), Sliding (window), Full.
#define EP_ATT_MODE(rkind, win, BODY)                                   \
  if (rkind) {                                                          \
    constexpr int AM = AM_RANDOM;                                       \
    BODY;                                                               \
  } else if ((win) != 0xFFFFFFFFu) {                                    \
    constexpr int AM = AM_SLIDING;                                      \
    BODY;                                                               \
  } else {                                                              \
    constexpr int AM = AM_FULL;                                         \
    BODY;                                                               \
  }

#define EP_ATT_DISPATCH(dk, BODY)                      \
  switch (dk) {                                        \
    case 4: { constexpr int DK = 4; BODY; } break;     \
    case 8: { constexpr int DK = 8; BODY; } break;     \
    case 16: { constexpr int DK = 16; BODY; } break;   \
    case 32: { constexpr int DK = 32; BODY; } break;   \
    default: return hipErrorInvalidValue;              \
  }

hipError_t launch_att_forward(const StepConfig& c, StepState* st, hipStream_t s) {
  const size_t lds = att_lds_floats(c.att_dm, c.att_heads, false, c.att_rkind != 0) * 4;
  hipLaunchKernelGGL(k_att_map, dim3((c.batch + NW - 1) / NW), dim3(256), 0, s, c, st);
  EP_ATT_DISPATCH(c.att_dk, {
    EP_ATT_MODE(c.att_rkind, c.att_win, {
      if (c.bf16) hipLaunchKernelGGL((k_att_forward<DK, uint16_t, AM>), dim3(c.blk_cap), dim3(64), lds, s, c, st);
      else hipLaunchKernelGGL((k_att_forward<DK, float, AM>), dim3(c.blk_cap), dim3(64), lds, s, c, st);
    });
  });
  hipLaunchKernelGGL(k_att_loss, dim3((c.batch + NW - 1) / NW), dim3(256), 0, s, c, st);
  return hipGetLastError();
}

hipError_t launch_att_embed(const StepConfig& c, const uint64_t* findptr, const uint32_t* fids, uint64_t n_nodes,
                            float* out, hipStream_t s) {
  const size_t lds = att_lds_floats(c.att_dm, c.att_heads, false, false) * 4;
  const unsigned grid = (unsigned)std::min<uint64_t>(n_nodes, 1u << 20);
  // inference: Full or Sliding (AttentionAggregator has no Random attention, lib.rs:2028-2033)
  EP_ATT_DISPATCH(c.att_dk, {
    if (c.att_win != 0xFFFFFFFFu) {
      constexpr int AM = AM_SLIDING;
      if (c.bf16)
        hipLaunchKernelGGL((k_att_embed<DK, uint16_t, AM>), dim3(grid), dim3(64), lds, s, (const uint16_t*)c.table,
                           (uint32_t)c.dims, c.att_heads, c.att_dm, findptr, fids, n_nodes, c.att_win, out);
      else
        hipLaunchKernelGGL((k_att_embed<DK, float, AM>), dim3(grid), dim3(64), lds, s, (const float*)c.table,
                           (uint32_t)c.dims, c.att_heads, c.att_dm, findptr, fids, n_nodes, c.att_win, out);
    } else {
      constexpr int AM = AM_FULL;
      if (c.bf16)
        hipLaunchKernelGGL((k_att_embed<DK, uint16_t, AM>), dim3(grid), dim3(64), lds, s, (const uint16_t*)c.table,
                           (uint32_t)c.dims, c.att_heads, c.att_dm, findptr, fids, n_nodes, c.att_win, out);
      else
        hipLaunchKernelGGL((k_att_embed<DK, float, AM>), dim3(grid), dim3(64), lds, s, (const float*)c.table,
                           (uint32_t)c.dims, c.att_heads, c.att_dm, findptr, fids, n_nodes, c.att_win, out);
    }
  });
  return hipGetLastError();
}

hipError_t launch_att_backward(const StepConfig& c, StepState* st, hipStream_t s) {
  const size_t lds = att_lds_floats(c.att_dm, c.att_heads, true, c.att_rkind != 0) * 4;
  EP_ATT_DISPATCH(c.att_dk, {
    EP_ATT_MODE(c.att_rkind, c.att_win, {
      if (c.bf16) hipLaunchKernelGGL((k_att_backward<DK, uint16_t, AM>), dim3(c.blk_cap), dim3(64), lds, s, c, st);
      else hipLaunchKernelGGL((k_att_backward<DK, float, AM>), dim3(c.blk_cap), dim3(64), lds, s, c, st);
    });
  });
  return hipGetLastError();
}

}  // namespace ep
