// Losses of the EP step other than margin, shared by the mean-model forward
// (ep_kernels.hip) and the attention loss kernel (ep_attention.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ep {

// ------------------------------------------------------------------------------------
// Losses other than margin (loss.rs:56-151).  With x^ = x / |x| and the adjoint of the
// normalisation (g - x^ (x^.g)) / |x|, every gradient is a combination of a, p and n_k:
//   starspace  l_k = gamma - (p^.a^ - a^.n_k^), mean over l_k > 0
//   contrastive max(0, pm - p^.a^) and max(0, a^.n_k^ - nm), mean over the positive terms
//   rank       -ln s_pos if s_pos < tau, s = softmax(a.n_1, .., a.n_K, a.p) (exact exp)
//   rankspace  starspace(gamma) + rank(gamma)
// ------------------------------------------------------------------------------------
__device__ inline float other_loss(int kind, float gamma, float p2, uint32_t K, const float* nrm, float dap,
                            const float* dan, float* ca, float* cp, float* cns, float* cna, uint32_t* reached) {
  for (uint32_t i = 0; i < K + 2u; ++i) ca[i] = 0.f;
  for (uint32_t k = 0; k < K; ++k) cns[k] = cna[k] = 0.f;
  cp[0] = cp[1] = 0.f;
  uint32_t rmask = 0;
  float L = 0.f;
  const float na = nrm[0], np = nrm[1];
  const float cpa = dap / (na * np);  // p^.a^
  if (kind == 2 || kind == 4) {       // starspace
    float sum = 0.f, sak = 0.f;
    uint32_t A = 0;
    for (uint32_t k = 0; k < K; ++k) {
      const float cak = dan[k] / (na * nrm[2 + k]);
      const float l = gamma - (cpa - cak);
      if (l > 0.f) { sum += l; sak += cak; ++A; }
    }
    if (A) {
      const float iA = 1.f / (float)A;
      L += sum * iA;
      const float ta = sak * iA - cpa;  // a^.G_a^
      ca[0] += -ta / (na * na);
      ca[1] += -1.f / (np * na);
      cp[0] += -1.f / (na * np);
      cp[1] += cpa / (np * np);
      rmask |= 3u;
      for (uint32_t k = 0; k < K; ++k) {
        const float nk = nrm[2 + k], cak = dan[k] / (na * nk);
        if (gamma - (cpa - cak) > 0.f) {
          ca[2 + k] += iA / (nk * na);
          cna[k] += iA / (na * nk);
          cns[k] += -cak * iA / (nk * nk);
          rmask |= 1u << (2 + k);
        }
      }
    }
  }
  if (kind == 1) {  // contrastive
    const float pos = gamma - cpa;
    const bool P = pos > 0.f;
    float sum = P ? pos : 0.f, sak = 0.f;
    uint32_t m = P ? 1u : 0u;
    for (uint32_t k = 0; k < K; ++k) {
      const float t = dan[k] / (na * nrm[2 + k]) - p2;
      if (t > 0.f) { sum += t; sak += t + p2; ++m; }
    }
    if (m) {
      const float im = 1.f / (float)m;
      L = sum * im;
      const float ta = (sak - (P ? cpa : 0.f)) * im;
      ca[0] += -ta / (na * na);
      rmask |= 1u;
      if (P) {
        ca[1] += -im / (np * na);
        cp[0] += -im / (na * np);
        cp[1] += cpa * im / (np * np);
        rmask |= 2u;
      }
      for (uint32_t k = 0; k < K; ++k) {
        const float nk = nrm[2 + k], cak = dan[k] / (na * nk);
        if (cak - p2 > 0.f) {
          ca[2 + k] += im / (nk * na);
          cna[k] += im / (na * nk);
          cns[k] += -cak * im / (nk * nk);
          rmask |= 1u << (2 + k);
        }
      }
    }
  }
  if (kind == 3 || kind == 4) {  // rank (tau = gamma)
    float mx = dap;
    for (uint32_t k = 0; k < K; ++k) mx = fmaxf(mx, dan[k]);
    float z = expf(dap - mx);
    for (uint32_t k = 0; k < K; ++k) z += expf(dan[k] - mx);
    const float spos = expf(dap - mx) / z;
    if (spos < gamma) {
      L += logf(z) - (dap - mx);  // -ln s_pos
      ca[1] += spos - 1.f;
      cp[0] += spos - 1.f;
      for (uint32_t k = 0; k < K; ++k) {
        const float sk = expf(dan[k] - mx) / z;
        ca[2 + k] += sk;
        cna[k] += sk;
      }
      rmask = (K + 2u >= 32u) ? ~0u : ((1u << (K + 2u)) - 1u);
    }
  }
  *reached = (L > 0.f) ? rmask : 0u;
  return L;
}

}  // namespace ep
