// Attention aggregator of the EP step (AttentionFeatureModel, src/algos/ep/model.rs:158-249,
// 306-335, 419-446; attention_mean, src/algos/ep/attention.rs:100-131, Full attention
// 171-191, softmax 253-296), gfx950 / wave64.
//
// A table row is [Q_1..Q_H | K_1..K_H | V_1..V_H] (attention.rs:38-66): H heads of d_k
// query and key values and d_model values, width W = H (2 d_k + d_model).
//
// An example's attention blocks are numbered from blkoff[b]: block 0 is h(v), blocks
// 1..cnt are the sampled neighbours of ~h(v) (one attention per neighbour, averaged),
// blocks cnt+1..cnt+K are the negatives h(u_k).  One wavefront per block, one item per
// lane (at most 64 items):
//   h(v), h(u_k)  one item per distinct feature of the node, count 2 x multiplicity
//                 (model.rs:267-268 adds 1 + 1 per occurrence);
//   neighbour     the node's listed features, count = multiplicity in the node
//                 (model.rs:426-443).
// Per head: S_ij = (Q_i . K_j) m_ij / sqrt(d_k), m_ij = c_i c_j when >= 1 (else 1),
// A = row softmax(S) (exact exp), w_j = sum_i A_ij, out_h = (1/n) sum_j w_j V_j; the
// block output is the mean over heads; one item returns V of head 0 (attention.rs:113-115).
//
// k_att_forward   block outputs
// k_att_loss      one wave per example: a, p = mean of the neighbour blocks, n_k; margin
//                 loss, weighting, adjoints, eligibility (as the mean model's forward)
// k_att_backward  recomputes each eligible block and writes one gradient row of width W
//                 per item at the slot of the item's feature occurrence (key position);
//                 the sorted segment reduce sums a feature's occurrence rows.
#include "ep_device.h"
#include "ep_kernels.h"

namespace ep {

namespace {

constexpr int NW = kBlockWaves;  // waves per 256-thread block (k_att_loss)
constexpr int kAttMaxItems = 64;
constexpr int kAttLdsStride = 65;  // row stride of the n x n score tile (conflict-free columns)

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

// LDS of one block's wave (floats / words).
struct AttLds {
  float* K;      // [64][DK]: K_h (forward, backward), then Q_h (backward dK pass)
  float* A;      // [64][65]: scores, probabilities, then dT
  float* invZ;   // [64]
  float* u;      // [64]
  float* g;      // [att_dm] block adjoint
  uint32_t* f;   // [64] feature of each item
  float* cnt;    // [64] count of each item
  uint32_t* q;   // [64] position of each item in the node's feature list
  uint32_t* raw; // [64] the node's listed features
};

template <int DK>
__device__ inline size_t att_lds_words(int dm) {
  return (size_t)kAttMaxItems * DK + (size_t)kAttMaxItems * kAttLdsStride + 2 * kAttMaxItems + (size_t)dm +
         4 * kAttMaxItems;
}

template <int DK>
__device__ inline AttLds att_lds(float* sm, int dm) {
  AttLds L;
  L.K = sm;
  L.A = L.K + kAttMaxItems * DK;
  L.invZ = L.A + kAttMaxItems * kAttLdsStride;
  L.u = L.invZ + kAttMaxItems;
  L.g = L.u + kAttMaxItems;
  L.f = (uint32_t*)(L.g + dm);
  L.cnt = (float*)(L.f + kAttMaxItems);
  L.q = (uint32_t*)(L.cnt + kAttMaxItems);
  L.raw = L.q + kAttMaxItems;
  return L;
}

// The block a wave works on.
struct AttBlock {
  uint32_t b;      // example in the step
  uint32_t ci;     // construction: 0 a, 1 p, 2+k negative k
  uint32_t cnt;    // sampled neighbours of the example
  uint32_t node;
  uint32_t n;      // items
  uint32_t slot0;  // key position of the node's first feature occurrence
  bool neighbour;  // block inside ~h(v)
};

// Finds block g's example (last b with blkoff[b] <= g), node, items (into LDS) and slot
// base.  Returns false for waves past the step's blocks.
__device__ inline bool att_setup(const StepConfig& c, const StepState* st, uint32_t g, const AttLds& L,
                                 AttBlock& B) {
  const uint32_t lane = lane_id();
  const uint32_t nb = c.blkoff[c.batch];
  if (g >= nb) return false;
  uint32_t lo = 0, hi = c.batch;  // blkoff[lo] <= g < blkoff[hi]
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (c.blkoff[mid] <= g) lo = mid; else hi = mid;
  }
  B.b = lo;
  const uint32_t j = g - c.blkoff[lo];
  B.cnt = c.nbr_cnt[lo];
  const uint32_t v = c.perm[st->anchor_base + lo];
  const uint32_t* pn = c.all_nbrs ? c.indices + c.indptr[v] : c.nbr + (size_t)lo * c.S;
  auto node_of = [&](uint32_t t) -> uint32_t {
    return t == 0u ? v : (t <= B.cnt ? pn[t - 1u] : c.neg[(size_t)lo * c.K + (t - 1u - B.cnt)]);
  };
  B.node = node_of(j);
  B.neighbour = j >= 1u && j <= B.cnt;
  B.ci = j == 0u ? 0u : (B.neighbour ? 1u : 2u + (j - 1u - B.cnt));
  // key position: the example's occurrences are listed node by node in block order
  uint32_t pre = 0;
  for (uint32_t t0 = 0; t0 < j; t0 += 64u) {
    const uint32_t t = t0 + lane;
    uint32_t x = 0;
    if (t < j) {
      const uint32_t u = node_of(t);
      x = (uint32_t)(c.findptr[u + 1] - c.findptr[u]);
    }
    pre += wave_sum_u32(x);
  }
  B.slot0 = c.keyoff[lo] + pre;
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
  if (B.neighbour) {
    B.n = nraw;
    if (lane < nraw) {
      L.f[lane] = fq;
      L.cnt[lane] = (float)mult;
      L.q[lane] = lane;
    }
  } else {
    const uint64_t m = __ballot(lane < nraw && first);
    B.n = (uint32_t)__popcll(m);
    if (lane < nraw && first) {
      const uint32_t i = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
      L.f[i] = fq;
      L.cnt[i] = 2.0f * (float)mult;
      L.q[i] = lane;
    }
  }
  __syncthreads();
  return true;
}

// Row i (lane i < n) of the head's attention into L.A[i][*]: scores, then the softmax
// probabilities.  K_h must be in L.K.
template <int DK>
__device__ inline void att_row_softmax(const AttLds& L, uint32_t n, uint32_t i, const float (&q)[DK], float ci,
                                       float inv_sqrt_dk) {
  float* Ai = L.A + (size_t)i * kAttLdsStride;
  float mx = -INFINITY;
  for (uint32_t j = 0; j < n; ++j) {
    const float* Kj = L.K + (size_t)j * DK;
    float s = 0.f;
#pragma unroll
    for (int t = 0; t < DK; t += 4) {
      const float4 k4 = *reinterpret_cast<const float4*>(Kj + t);
      s += q[t] * k4.x;
      s += q[t + 1] * k4.y;
      s += q[t + 2] * k4.z;
      s += q[t + 3] * k4.w;
    }
    const float mm = ci * L.cnt[j];
    if (mm >= 1.0f) s *= mm;
    s *= inv_sqrt_dk;
    Ai[j] = s;
    mx = fmaxf(mx, s);
  }
  float z = 0.f;
  for (uint32_t j = 0; j < n; ++j) {
    const float e = expf(Ai[j] - mx);
    Ai[j] = e;
    z += e;
  }
  const float iz = 1.0f / z;
  for (uint32_t j = 0; j < n; ++j) Ai[j] *= iz;
}

template <int DK, typename T>
__device__ inline void load_qk(const T* __restrict__ row, int h, int H, float (&q)[DK], float* kdst) {
#pragma unroll
  for (int t = 0; t < DK; t += 4) {
    const float4 a = ld4<T>(row + h * DK + t);
    q[t] = a.x; q[t + 1] = a.y; q[t + 2] = a.z; q[t + 3] = a.w;
    const float4 k = ld4<T>(row + H * DK + h * DK + t);
    *reinterpret_cast<float4*>(kdst + t) = k;
  }
}

}  // namespace

// ------------------------------------------------------------------------------------
// k_att_forward: one wave (64-thread block) per attention block.
// ------------------------------------------------------------------------------------
template <int DK, typename T>
__global__ __launch_bounds__(64) void k_att_forward(StepConfig c, StepState* st) {
  extern __shared__ float sm[];
  const int H = c.att_heads, dm = c.att_dm;
  const uint32_t W = (uint32_t)c.dims;
  const AttLds L = att_lds<DK>(sm, dm);
  AttBlock B;
  if (!att_setup(c, st, blockIdx.x, L, B)) return;
  const uint32_t lane = lane_id();
  const T* table = (const T*)c.table;
  float* out = c.bout + (size_t)blockIdx.x * dm;
  const uint32_t n = B.n;
  if (lane == 0) c.bitems[blockIdx.x] = n;
  if (n == 1u) {  // attention.rs:113-115: the value vector of head 0
    const T* row = table + (size_t)L.f[0] * W + 2u * H * DK;
    for (int t = lane; t < dm; t += 64) out[t] = ld1<T>(row + t);
    return;
  }
  const float inv_sqrt_dk = 1.0f / sqrtf((float)DK);
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  const float ci = lane < n ? L.cnt[lane] : 0.f;
  for (int h = 0; h < H; ++h) {
    float q[DK];
    if (lane < n) load_qk<DK, T>(table + (size_t)L.f[lane] * W, h, H, q, L.K + (size_t)lane * DK);
    __syncthreads();
    if (lane < n) att_row_softmax<DK>(L, n, lane, q, ci, inv_sqrt_dk);
    __syncthreads();
    float w = 0.f;
    if (lane < n)
      for (uint32_t i = 0; i < n; ++i) w += L.A[(size_t)i * kAttLdsStride + lane];
    for (uint32_t j = 0; j < n; ++j) {
      const float wj = __shfl(w, (int)j, EP_WAVE);
      const T* vrow = table + (size_t)L.f[j] * W + 2u * H * DK + (uint32_t)h * dm;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int t = (int)lane + 64 * r;
        if (t < dm) acc[r] += wj * ld1<T>(vrow + t);
      }
    }
    __syncthreads();
  }
  const float scale = 1.0f / ((float)n * (float)H);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int t = (int)lane + 64 * r;
    if (t < dm) out[t] = acc[r] * scale;
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
//   dS_ij = A_ij (u_j - sum_k A_ik u_k),   dT = dS m / sqrt(d_k)
//   dQ_i = sum_j dT_ij K_j,  dK_j = sum_i dT_ij Q_i,  dV_j = w_j g / (n H)
// ------------------------------------------------------------------------------------
template <int DK, typename T>
__global__ __launch_bounds__(64) void k_att_backward(StepConfig c, StepState* st) {
  extern __shared__ float sm[];
  const int H = c.att_heads, dm = c.att_dm;
  const uint32_t W = (uint32_t)c.dims;
  const AttLds L = att_lds<DK>(sm, dm);
  if (st->n_keys > c.key_cap) return;
  AttBlock B;
  if (!att_setup(c, st, blockIdx.x, L, B)) return;
  const uint32_t NC = c.K + 2u;
  if (!c.elig[(size_t)B.b * NC + B.ci]) return;
  const uint32_t lane = lane_id();
  const T* table = (const T*)c.table;
  const uint32_t n = B.n;
  for (int t = lane; t < dm; t += 64) L.g[t] = c.badj[(size_t)blockIdx.x * dm + t];
  __syncthreads();
  const uint32_t qoff = 2u * H * DK;  // first V column
  if (n == 1u) {
    float* grow = c.adj + (size_t)(B.slot0 + L.q[0]) * W;
    for (uint32_t t = lane; t < W; t += 64u) {
      const uint32_t vt = t - qoff;
      grow[t] = (t >= qoff && vt < (uint32_t)dm) ? L.g[vt] : 0.f;
    }
    return;
  }
  const float inv_sqrt_dk = 1.0f / sqrtf((float)DK);
  const float scale = 1.0f / ((float)n * (float)H);
  const float ci = lane < n ? L.cnt[lane] : 0.f;
  const uint32_t my_slot = lane < n ? B.slot0 + L.q[lane] : 0u;
  float* my_row = c.adj + (size_t)my_slot * W;
  for (int h = 0; h < H; ++h) {
    float q[DK];
    if (lane < n) {
      const T* row = table + (size_t)L.f[lane] * W;
      load_qk<DK, T>(row, h, H, q, L.K + (size_t)lane * DK);
      // u_i = (V_i,h . g) / (n H)
      const T* vrow = row + qoff + (size_t)h * dm;
      float s = 0.f;
      for (int t = 0; t < dm; t += 4) {
        const float4 v4 = ld4<T>(vrow + t);
        s += v4.x * L.g[t] + v4.y * L.g[t + 1] + v4.z * L.g[t + 2] + v4.w * L.g[t + 3];
      }
      L.u[lane] = s * scale;
    }
    __syncthreads();
    float r_i = 0.f;
    if (lane < n) {
      att_row_softmax<DK>(L, n, lane, q, ci, inv_sqrt_dk);
      const float* Ai = L.A + (size_t)lane * kAttLdsStride;
      for (uint32_t j = 0; j < n; ++j) r_i += Ai[j] * L.u[j];
    }
    __syncthreads();
    float w = 0.f;  // column sums: w_j
    if (lane < n)
      for (uint32_t i = 0; i < n; ++i) w += L.A[(size_t)i * kAttLdsStride + lane];
    __syncthreads();
    // row pass: dT into the tile, dQ_i
    float dq[DK];
#pragma unroll
    for (int t = 0; t < DK; ++t) dq[t] = 0.f;
    if (lane < n) {
      float* Ai = L.A + (size_t)lane * kAttLdsStride;
      for (uint32_t j = 0; j < n; ++j) {
        float mm = ci * L.cnt[j];
        if (!(mm >= 1.0f)) mm = 1.0f;
        const float dT = Ai[j] * (L.u[j] - r_i) * mm * inv_sqrt_dk;
        Ai[j] = dT;
        const float* Kj = L.K + (size_t)j * DK;
#pragma unroll
        for (int t = 0; t < DK; t += 4) {
          const float4 k4 = *reinterpret_cast<const float4*>(Kj + t);
          dq[t] += dT * k4.x;
          dq[t + 1] += dT * k4.y;
          dq[t + 2] += dT * k4.z;
          dq[t + 3] += dT * k4.w;
        }
      }
#pragma unroll
      for (int t = 0; t < DK; t += 4)
        *reinterpret_cast<float4*>(my_row + h * DK + t) = make_float4(dq[t], dq[t + 1], dq[t + 2], dq[t + 3]);
    }
    __syncthreads();
    if (lane < n)
#pragma unroll
      for (int t = 0; t < DK; t += 4)
        *reinterpret_cast<float4*>(L.K + (size_t)lane * DK + t) = make_float4(q[t], q[t + 1], q[t + 2], q[t + 3]);
    __syncthreads();
    // column pass: dK_j = sum_i dT_ij Q_i
    if (lane < n) {
      float dk[DK];
#pragma unroll
      for (int t = 0; t < DK; ++t) dk[t] = 0.f;
      for (uint32_t i = 0; i < n; ++i) {
        const float dT = L.A[(size_t)i * kAttLdsStride + lane];
        const float* Qi = L.K + (size_t)i * DK;
#pragma unroll
        for (int t = 0; t < DK; t += 4) {
          const float4 q4 = *reinterpret_cast<const float4*>(Qi + t);
          dk[t] += dT * q4.x;
          dk[t + 1] += dT * q4.y;
          dk[t + 2] += dT * q4.z;
          dk[t + 3] += dT * q4.w;
        }
      }
#pragma unroll
      for (int t = 0; t < DK; t += 4)
        *reinterpret_cast<float4*>(my_row + H * DK + h * DK + t) = make_float4(dk[t], dk[t + 1], dk[t + 2], dk[t + 3]);
    }
    // dV_j = w_j g / (n H), row by row across the wave
    for (uint32_t j = 0; j < n; ++j) {
      const float cw = __shfl(w, (int)j, EP_WAVE) * scale;
      float* vr = c.adj + (size_t)(B.slot0 + L.q[j]) * W + qoff + (size_t)h * dm;
      for (int t = lane; t < dm; t += 64) vr[t] = cw * L.g[t];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------------------
bool att_shape_supported(int heads, int dk, int dm) {
  return heads >= 1 && heads <= 16 && (dk == 4 || dk == 8 || dk == 16 || dk == 32) && dm >= 4 && dm <= 256 &&
         dm % 4 == 0;
}

static size_t att_lds_bytes(int dk, int dm) {
  return ((size_t)kAttMaxItems * dk + (size_t)kAttMaxItems * kAttLdsStride + 2 * kAttMaxItems + (size_t)dm +
          4 * kAttMaxItems) * 4;
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
  const size_t lds = att_lds_bytes(c.att_dk, c.att_dm);
  EP_ATT_DISPATCH(c.att_dk, {
    if (c.bf16) hipLaunchKernelGGL((k_att_forward<DK, uint16_t>), dim3(c.blk_cap), dim3(64), lds, s, c, st);
    else hipLaunchKernelGGL((k_att_forward<DK, float>), dim3(c.blk_cap), dim3(64), lds, s, c, st);
  });
  hipLaunchKernelGGL(k_att_loss, dim3((c.batch + NW - 1) / NW), dim3(256), 0, s, c, st);
  return hipGetLastError();
}

hipError_t launch_att_backward(const StepConfig& c, StepState* st, hipStream_t s) {
  const size_t lds = att_lds_bytes(c.att_dk, c.att_dm);
  EP_ATT_DISPATCH(c.att_dk, {
    if (c.bf16) hipLaunchKernelGGL((k_att_backward<DK, uint16_t>), dim3(c.blk_cap), dim3(64), lds, s, c, st);
    else hipLaunchKernelGGL((k_att_backward<DK, float>), dim3(c.blk_cap), dim3(64), lds, s, c, st);
  });
  return hipGetLastError();
}

}  // namespace ep
