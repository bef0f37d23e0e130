// CPU restatement of the reference's EP training loop, FP32, multithreaded.
//
// TEST INFRASTRUCTURE ONLY: the CPU baseline timed by bench.py and the CPU side of the
// holdout-AUC comparison.  Never linked into the shipped library.
//
// Follows src/algos/ep/mod.rs:92-287 with the reference's parallel structure:
//   * per node, in parallel over the threads (mod.rs:183): h(v), ~h(v), K negatives,
//     margin loss, gradients (mod.rs:183-215);
//   * per-node RNG XorShift seeded with seed + batch_idx + node_id (mod.rs:185);
//   * ~h(v) neighbours by reservoir sampling over the whole row (utils.rs:128-145,
//     model.rs:354-377): O(deg) like the reference;
//   * first-leaf gradient extraction with a per-example hash map (mod.rs:329-378);
//   * per-batch merge of the examples' gradient maps (mod.rs:223-233), thread maps summed
//     in shards by feature id;
//   * sparse Adam over the merged rows, t = pass + 1, the shards in parallel
//     (optimizer.rs:43-81);
//   * cosine LR with 20% warm-up (scheduler.rs:39-47, mod.rs:132-137).
// The mean aggregator with max_features = All and MarginLoss (loss.rs:59-73); with
// --heads H --dk d_k the attention aggregator instead (AttentionFeatureModel,
// model.rs:158-249, attention.rs:100-296, Full attention): rows of H (2 d_k + dims)
// values, gradients by the hand-written adjoint of the attention (the reference takes
// them from its autodiff tape).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

struct XorShift {  // xorshift128 (the reference uses rand_xorshift's XorShiftRng)
  uint32_t x, y, z, w;
  explicit XorShift(uint64_t seed) {
    uint64_t s = seed;
    auto mix = [&]() {
      uint64_t q = (s += 0x9E3779B97F4A7C15ull);
      q = (q ^ (q >> 30)) * 0xBF58476D1CE4E5B9ull;
      q = (q ^ (q >> 27)) * 0x94D049BB133111EBull;
      return q ^ (q >> 31);
    };
    uint64_t a = mix(), b = mix();
    x = (uint32_t)a; y = (uint32_t)(a >> 32); z = (uint32_t)b; w = (uint32_t)(b >> 32) | 1u;
  }
  uint32_t next() {
    uint32_t t = x ^ (x << 11);
    x = y; y = z; z = w;
    w = w ^ (w >> 19) ^ (t ^ (t >> 8));
    return w;
  }
  uint32_t below(uint32_t n) { return (uint32_t)(((uint64_t)next() * n) >> 32); }
  float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

struct Args {
  std::string dir;
  uint64_t nodes = 0;
  uint32_t vocab = 0, dims = 64, batch = 128, negatives = 5, max_nodes = 10, threads = 1, passes = 1;
  uint32_t heads = 0, dk = 0;  // attention model when dk > 0
  float gamma = 0.5f, alpha = 0.01f, budget_s = 0.f;
  uint64_t seed = 20222022ull;
  bool json = false;
  std::string table_out, train_file;
  std::string explicit_dir;  // one deterministic batch on explicit lists (see run_explicit)
  float lr = 0.01f;
  uint32_t pass = 1;
};

template <typename T>
std::vector<T> read_bin(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) { std::fprintf(stderr, "cannot open %s\n", path.c_str()); std::exit(2); }
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::vector<T> v(n / sizeof(T));
  if (n && std::fread(v.data(), 1, n, f) != (size_t)n) { std::fprintf(stderr, "short read %s\n", path.c_str()); std::exit(2); }
  std::fclose(f);
  return v;
}

// Zero-initialised float buffer from calloc: pages are committed on first touch, so the
// Adam moments of a large vocabulary cost memory only for the rows a run updates.
struct FBuf {
  float* p = nullptr;
  size_t n = 0;
  FBuf() = default;
  FBuf(const FBuf&) = delete;
  FBuf& operator=(const FBuf&) = delete;
  ~FBuf() { std::free(p); }
  void zeros(size_t count) {
    std::free(p);
    p = (float*)std::calloc(std::max<size_t>(count, 1), sizeof(float));
    if (!p) { std::fprintf(stderr, "out of memory (%zu floats)\n", count); std::exit(2); }
    n = count;
  }
  void assign(const std::vector<float>& v) {
    zeros(v.size());
    std::copy(v.begin(), v.end(), p);
  }
  float& operator[](size_t i) { return p[i]; }
  const float& operator[](size_t i) const { return p[i]; }
  size_t size() const { return n; }
  float* data() { return p; }
};

struct Model {
  const Args& a;
  const uint64_t* indptr;
  const uint32_t* indices;
  const uint64_t* findptr;
  const uint32_t* fids;
  FBuf table, mom, var;
  std::vector<uint32_t> train;

  // NodeCounts: feature -> count in first-seen order (model.rs:254-278)
  struct Cons {
    std::vector<uint32_t> feats;
    std::vector<float> counts;
    std::unordered_map<uint32_t, uint32_t> pos;
    void add(uint32_t f, float w) {
      auto it = pos.find(f);
      if (it == pos.end()) { pos.emplace(f, (uint32_t)feats.size()); feats.push_back(f); counts.push_back(w); }
      else counts[it->second] += w;
    }
  };

  void collect(Cons& c, uint32_t node, float weight) const {
    // Sample::All for features: scale 1; weight += scale (model.rs:267-268)
    const float w = weight + 1.0f;
    for (uint64_t q = findptr[node]; q < findptr[node + 1]; ++q) c.add(fids[q], w);
  }

  uint32_t width() const { return a.dk ? a.heads * (2 * a.dk + a.dims) : a.dims; }

  // Full attention over items (features f, counts c), attention_mean (attention.rs:100-131):
  // out[dims]; with g != nullptr instead adds each item's gradient row (width W) to gout.
  void attention(const std::vector<uint32_t>& f, const std::vector<float>& c, float* out, const float* g,
                 float* gout) const {
    const uint32_t H = a.heads, dk = a.dk, dm = a.dims, W = width(), n = (uint32_t)f.size();
    const uint32_t vb = 2 * H * dk;
    if (n == 1) {  // attention.rs:113-115
      const float* e = &table[(size_t)f[0] * W + vb];
      if (!g) for (uint32_t d = 0; d < dm; ++d) out[d] = e[d];
      else for (uint32_t d = 0; d < dm; ++d) gout[vb + d] += g[d];
      return;
    }
    const float isq = 1.f / std::sqrt((float)dk);
    const float sc = 1.f / ((float)n * (float)H);
    std::vector<float> A((size_t)n * n), w(n), u(n), r(n);
    if (!g) std::fill(out, out + dm, 0.f);
    for (uint32_t h = 0; h < H; ++h) {
      for (uint32_t i = 0; i < n; ++i) {
        const float* qi = &table[(size_t)f[i] * W + h * dk];
        float mx = -INFINITY;
        for (uint32_t j = 0; j < n; ++j) {
          const float* kj = &table[(size_t)f[j] * W + H * dk + h * dk];
          float s = 0.f;
          for (uint32_t t = 0; t < dk; ++t) s += qi[t] * kj[t];
          const float mm = c[i] * c[j];
          if (mm >= 1.f) s *= mm;
          s *= isq;
          A[(size_t)i * n + j] = s;
          mx = std::max(mx, s);
        }
        float z = 0.f;
        for (uint32_t j = 0; j < n; ++j) { const float e = std::exp(A[(size_t)i * n + j] - mx); A[(size_t)i * n + j] = e; z += e; }
        for (uint32_t j = 0; j < n; ++j) A[(size_t)i * n + j] /= z;
      }
      for (uint32_t j = 0; j < n; ++j) {
        float s = 0.f;
        for (uint32_t i = 0; i < n; ++i) s += A[(size_t)i * n + j];
        w[j] = s;
      }
      if (!g) {
        for (uint32_t j = 0; j < n; ++j) {
          const float* vj = &table[(size_t)f[j] * W + vb + h * dm];
          for (uint32_t d = 0; d < dm; ++d) out[d] += w[j] * vj[d];
        }
        continue;
      }
      for (uint32_t j = 0; j < n; ++j) {
        const float* vj = &table[(size_t)f[j] * W + vb + h * dm];
        float s = 0.f;
        for (uint32_t d = 0; d < dm; ++d) s += vj[d] * g[d];
        u[j] = s * sc;
        float* gv = &gout[(size_t)j * W + vb + h * dm];
        for (uint32_t d = 0; d < dm; ++d) gv[d] += w[j] * sc * g[d];
      }
      for (uint32_t i = 0; i < n; ++i) {
        float s = 0.f;
        for (uint32_t j = 0; j < n; ++j) s += A[(size_t)i * n + j] * u[j];
        r[i] = s;
      }
      for (uint32_t i = 0; i < n; ++i) {
        const float* qi = &table[(size_t)f[i] * W + h * dk];
        float* gq = &gout[(size_t)i * W + h * dk];
        for (uint32_t j = 0; j < n; ++j) {
          float mm = c[i] * c[j];
          if (!(mm >= 1.f)) mm = 1.f;
          const float dT = A[(size_t)i * n + j] * (u[j] - r[i]) * mm * isq;
          const float* kj = &table[(size_t)f[j] * W + H * dk + h * dk];
          float* gk = &gout[(size_t)j * W + H * dk + h * dk];
          for (uint32_t t = 0; t < dk; ++t) { gq[t] += dT * kj[t]; gk[t] += dT * qi[t]; }
        }
      }
    }
    if (!g) for (uint32_t d = 0; d < dm; ++d) out[d] *= sc;
  }

  // Items of a neighbour's attention inside ~h(v): the listed features, count = multiplicity
  void neighbour_items(uint32_t node, std::vector<uint32_t>& f, std::vector<float>& c) const {
    f.assign(fids + findptr[node], fids + findptr[node + 1]);
    c.resize(f.size());
    for (size_t i = 0; i < f.size(); ++i) c[i] = (float)std::count(f.begin(), f.end(), f[i]);
  }

  void mean(const Cons& c, float* out) const {  // mean_embeddings (model.rs:448-463)
    const uint32_t D = a.dims;
    std::fill(out, out + D, 0.f);
    float n = 0.f;
    for (size_t i = 0; i < c.feats.size(); ++i) {
      const float* e = &table[(size_t)c.feats[i] * D];
      const float cnt = c.counts[i];
      for (uint32_t d = 0; d < D; ++d) out[d] += e[d] * cnt;
      n += cnt;
    }
    for (uint32_t d = 0; d < D; ++d) out[d] /= n;
  }

  // One example; returns loss (<= 0 means no gradient) and fills grads (first-leaf rule).
  // Samples as the reference: ~h(v) from Fixed(S).sample(deg) = (min(S,deg), deg/r), all
  // neighbours when deg <= S, else a reservoir over the row (utils.rs:128-145); K
  // negatives uniform with replacement from the train ids (node_sampler.rs:102-105).
  float example(uint32_t v, XorShift& rng, std::unordered_map<uint32_t, std::vector<float>>& grads) const {
    const uint32_t K = a.negatives;
    const uint64_t s0 = indptr[v], s1 = indptr[v + 1];
    const uint32_t deg = (uint32_t)(s1 - s0);
    if (deg == 0) return 0.f;
    const uint32_t S = a.max_nodes;
    const uint32_t r = std::min(S, deg);
    const float scale = (float)deg / (float)r;
    std::vector<uint32_t> nbrs;
    if (deg <= S) {
      nbrs.assign(indices + s0, indices + s1);
    } else {
      nbrs.assign(indices + s0, indices + s0 + S);
      for (uint32_t i = S; i < deg; ++i) {
        const uint32_t idx = rng.below(i);
        if (idx < S) nbrs[idx] = indices[s0 + i];
      }
    }
    std::vector<uint32_t> negs(K);
    for (uint32_t k = 0; k < K; ++k) negs[k] = train[rng.below((uint32_t)train.size())];
    return example_lists(v, nbrs, scale, negs, grads);
  }

  // One example on given neighbour and negative lists (the neighbours' features count
  // scale + 1 each, model.rs:267-268).
  float example_lists(uint32_t v, const std::vector<uint32_t>& nbrs, float scale, const std::vector<uint32_t>& negs,
                      std::unordered_map<uint32_t, std::vector<float>>& grads) const {
    const uint32_t D = a.dims, K = a.negatives;
    if (nbrs.empty()) return 0.f;
    std::vector<Cons> cons(2 + K);
    collect(cons[0], v, 1.0f);
    for (uint32_t u : nbrs) collect(cons[1], u, scale);
    for (uint32_t k = 0; k < K; ++k) collect(cons[2 + k], negs[k], 1.0f);
    std::vector<float> x((2 + K) * (size_t)D);
    const bool att = a.dk != 0;
    if (!att) {
      for (uint32_t c = 0; c < 2 + K; ++c) mean(cons[c], &x[(size_t)c * D]);
    } else {
      // the sampled neighbours again, in order (one attention each, model.rs:419-446)
      attention(cons[0].feats, cons[0].counts, &x[0], nullptr, nullptr);
      std::fill(&x[D], &x[2 * D], 0.f);
      std::vector<uint32_t> nf;
      std::vector<float> nc, tmp(D);
      for (uint32_t u : nbrs) {
        neighbour_items(u, nf, nc);
        attention(nf, nc, tmp.data(), nullptr, nullptr);
        for (uint32_t d = 0; d < D; ++d) x[D + d] += tmp[d];
      }
      for (uint32_t d = 0; d < D; ++d) x[D + d] /= (float)nbrs.size();
      for (uint32_t k = 0; k < K; ++k)
        attention(cons[2 + k].feats, cons[2 + k].counts, &x[(size_t)(2 + k) * D], nullptr, nullptr);
    }
    const float* av = &x[0];
    const float* pv = &x[D];
    auto dist = [&](const float* p, const float* q) {
      float s = 0.f;
      for (uint32_t d = 0; d < D; ++d) { const float t = p[d] - q[d]; s += t * t; }
      return std::sqrt(s);
    };
    const float D0 = dist(pv, av);
    const float d1 = a.gamma + D0;
    std::vector<float> Di(K);
    std::vector<char> act(K, 0);
    float sum = 0.f;
    uint32_t nA = 0;
    for (uint32_t k = 0; k < K; ++k) {
      Di[k] = dist(pv, &x[(size_t)(2 + k) * D]);
      const float l = d1 - Di[k];
      if (l > 0.f) { act[k] = 1; sum += l; ++nA; }
    }
    if (nA == 0) return 0.f;
    const float L = sum / (float)nA;
    if (!(L > 0.f)) return 0.f;
    // adjoints
    std::vector<float> g((2 + K) * (size_t)D, 0.f);
    std::vector<char> elig(2 + K, 0);
    elig[0] = elig[1] = 1;
    for (uint32_t d = 0; d < D; ++d) {
      g[d] = (av[d] - pv[d]) / D0;
      g[D + d] = (pv[d] - av[d]) / D0;
    }
    for (uint32_t k = 0; k < K; ++k) {
      if (!act[k]) continue;
      elig[2 + k] = 1;
      const float* n = &x[(size_t)(2 + k) * D];
      for (uint32_t d = 0; d < D; ++d) {
        const float t = (pv[d] - n[d]) / Di[k] / (float)nA;
        g[D + d] -= t;
        g[(size_t)(2 + k) * D + d] = t;
      }
    }
    // first-leaf extraction (mod.rs:347-378)
    for (uint32_t c = 0; c < 2 + K; ++c) {
      if (!elig[c]) continue;
      const float* gc = &g[(size_t)c * D];
      bool finite = true;
      for (uint32_t d = 0; d < D; ++d) finite &= std::isfinite(gc[d]);
      if (!finite) continue;
      if (att) {
        // leaf gradients of the construction: one row per distinct feature
        const uint32_t W = width();
        std::vector<float> rows((size_t)cons[c].feats.size() * W, 0.f);
        if (c == 1) {
          std::vector<uint32_t> nf;
          std::vector<float> nc, gi(D), tmp;
          for (uint32_t d = 0; d < D; ++d) gi[d] = gc[d] / (float)nbrs.size();
          for (uint32_t u : nbrs) {
            neighbour_items(u, nf, nc);
            tmp.assign((size_t)nf.size() * W, 0.f);
            attention(nf, nc, nullptr, gi.data(), tmp.data());
            for (size_t i = 0; i < nf.size(); ++i) {
              float* dst = &rows[(size_t)cons[1].pos.at(nf[i]) * W];
              for (uint32_t t = 0; t < W; ++t) dst[t] += tmp[i * W + t];
            }
          }
        } else {
          attention(cons[c].feats, cons[c].counts, nullptr, gc, rows.data());
        }
        for (size_t i = 0; i < cons[c].feats.size(); ++i) {
          const uint32_t f = cons[c].feats[i];
          if (grads.count(f)) continue;
          grads.emplace(f, std::vector<float>(rows.begin() + i * W, rows.begin() + (i + 1) * W));
        }
        continue;
      }
      float C = 0.f;
      for (float q : cons[c].counts) C += q;
      for (size_t i = 0; i < cons[c].feats.size(); ++i) {
        const uint32_t f = cons[c].feats[i];
        if (grads.count(f)) continue;
        std::vector<float> gr(D);
        const float w = cons[c].counts[i] / C;
        for (uint32_t d = 0; d < D; ++d) gr[d] = w * gc[d];
        grads.emplace(f, std::move(gr));
      }
    }
    return L;
  }

  void adam(const std::unordered_map<uint32_t, std::vector<float>>& all, float lr, float pass) {
    const uint32_t D = width();
    const float t = pass + 1.f, b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.f - std::pow(b1, t), bc2 = 1.f - std::pow(b2, t);
    for (const auto& kv : all) {
      const std::vector<float>& g = kv.second;
      bool nan = false;
      for (float q : g) nan |= std::isnan(q);
      if (nan) continue;
      float* m = &mom[(size_t)kv.first * D];
      float* v = &var[(size_t)kv.first * D];
      float* e = &table[(size_t)kv.first * D];
      for (uint32_t d = 0; d < D; ++d) {
        m[d] = b1 * m[d] + (1.f - b1) * g[d];
        v[d] = b2 * v[d] + (1.f - b2) * g[d] * g[d];
        e[d] -= lr * ((m[d] / bc1) / (std::sqrt(v[d] / bc2) + eps));
      }
    }
  }
};

float cos_decay(uint64_t step, float alpha, uint64_t total) {
  const uint64_t warm = (uint64_t)((float)total / 5.f);
  const float mn = alpha / 100.f;
  if (step > warm) return mn + 0.5f * (alpha - mn) * (1.f + std::cos(3.14159265f * (float)step / (float)total));
  return mn + ((float)step / (float)warm) * (alpha - mn);
}

template <typename T>
void write_raw(const std::string& path, const T* p, size_t n) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) { std::fprintf(stderr, "cannot write %s\n", path.c_str()); std::exit(2); }
  std::fwrite(p, sizeof(T), n, f);
  std::fclose(f);
}
template <typename T>
void write_bin(const std::string& path, const std::vector<T>& v) { write_raw(path, v.data(), v.size()); }
void write_bin(const std::string& path, FBuf& v) { write_raw(path, v.data(), v.size()); }

// One batch on explicit lists, sequentially: dir holds table.bin (vocab x width f32),
// anchors.bin (n u32), cnt.bin (n u32), nbr.bin (n x max_nodes u32, the first cnt[i]
// used), neg.bin (n x negatives u32).  Every example's first-leaf gradients are merged
// in anchor order (mod.rs:223-233) and one Adam update with --lr and t = --pass + 1 is
// applied (optimizer.rs:43-81).  Writes out_loss.bin (n: the loss of each contributing
// example, else 0), out_grad.bin (merged gradient, dense), out_table.bin, out_m.bin,
// out_v.bin.
int run_explicit(Model& m) {
  const Args& a = m.a;
  const std::string d = a.explicit_dir;
  const uint32_t W = m.width();
  m.table.assign(read_bin<float>(d + "/table.bin"));
  if (m.table.size() != (size_t)a.vocab * W) { std::fprintf(stderr, "table size mismatch\n"); return 2; }
  m.mom.zeros(m.table.size());
  m.var.zeros(m.table.size());
  const auto anchors = read_bin<uint32_t>(d + "/anchors.bin");
  const auto cnt = read_bin<uint32_t>(d + "/cnt.bin");
  const auto nbr = read_bin<uint32_t>(d + "/nbr.bin");
  const auto neg = read_bin<uint32_t>(d + "/neg.bin");
  const size_t n = anchors.size(), S = a.max_nodes, K = a.negatives;
  if (cnt.size() != n || nbr.size() != n * S || neg.size() != n * K) { std::fprintf(stderr, "list size mismatch\n"); return 2; }
  std::vector<float> loss(n, 0.f), grad(m.table.size(), 0.f);
  std::unordered_map<uint32_t, std::vector<float>> all;
  std::vector<uint32_t> order;  // merged rows in first-seen order
  for (size_t i = 0; i < n; ++i) {
    std::vector<uint32_t> nb(nbr.begin() + i * S, nbr.begin() + i * S + cnt[i]);
    std::vector<uint32_t> ng(neg.begin() + i * K, neg.begin() + (i + 1) * K);
    std::unordered_map<uint32_t, std::vector<float>> g;
    const float L = m.example_lists(anchors[i], nb, 1.0f, ng, g);
    if (!(L > 0.f)) continue;
    loss[i] = L;
    for (auto& kv : g) {
      auto it = all.find(kv.first);
      if (it == all.end()) { all.emplace(kv.first, std::move(kv.second)); order.push_back(kv.first); }
      else for (uint32_t t = 0; t < W; ++t) it->second[t] += kv.second[t];
    }
  }
  for (const auto& kv : all) std::copy(kv.second.begin(), kv.second.end(), grad.begin() + (size_t)kv.first * W);
  if (!all.empty()) m.adam(all, a.lr, (float)a.pass);
  write_bin(d + "/out_loss.bin", loss);
  write_bin(d + "/out_grad.bin", grad);
  write_bin(d + "/out_table.bin", m.table);
  write_bin(d + "/out_m.bin", m.mom);
  write_bin(d + "/out_v.bin", m.var);
  return 0;
}


struct RunStats {
  uint64_t examples = 0, batches = 0;
  double seconds = 0.0, train_error = 0.0;
};

// randomize_embedding_store (mod.rs:381-392): x ~ U(-1, 1), then each row / its norm;
// rows drawn in parallel from per-row XorShift streams.
void init_table(Model& m) {
  const Args& a = m.a;
  const uint32_t D = m.width();
  m.table.zeros((size_t)a.vocab * D);
  m.mom.zeros(m.table.size());
  m.var.zeros(m.table.size());
  const uint32_t T = std::max(1u, a.threads);
  std::vector<std::thread> ts;
  for (uint32_t t = 0; t < T; ++t)
    ts.emplace_back([&, t]() {
      for (uint32_t r = t; r < a.vocab; r += T) {
        XorShift init(a.seed ^ (0x9E3779B97F4A7C15ull * (r + 1ull)));
        float* e = &m.table[(size_t)r * D];
        float n = 0.f;
        for (uint32_t d = 0; d < D; ++d) { e[d] = 2.f * init.uniform() - 1.f; n += e[d] * e[d]; }
        n = std::sqrt(n);
        for (uint32_t d = 0; d < D; ++d) e[d] /= n;
      }
    });
  for (auto& t : ts) t.join();
}

// The training loop (mod.rs:156-265): shuffled passes; each batch's examples run in
// parallel on all threads (the per-node par_iter of mod.rs:183), each thread merging its
// examples' first-leaf gradients into its own map; the batch merge then sums the thread
// maps shard by shard (feature id mod threads) and every shard applies its rows' Adam
// step (optimizer.rs:51 runs the update in parallel over rows).  Batches run one after
// the other (the reference also overlaps batches on the rayon pool; with every core
// busy inside each batch the throughput is the same, and no update reads a stale table).
// Stops after the first batch past budget_s.  The clock starts after the table
// initialisation.
RunStats run_training(Model& m) {
  const Args& a = m.a;
  const uint32_t D = m.width();
  init_table(m);
  if (m.train.empty()) { m.train.resize(a.nodes); for (uint64_t i = 0; i < a.nodes; ++i) m.train[i] = (uint32_t)i; }
  std::vector<uint32_t> order = m.train;
  const uint64_t steps_per_pass = (order.size() + a.batch - 1) / a.batch;
  const uint64_t total = steps_per_pass * a.passes;
  const uint32_t T = std::max(1u, a.threads);
  using GradMap = std::unordered_map<uint32_t, std::vector<float>>;
  uint64_t step = 1, examples = 0, batches_done = 0;
  XorShift shuf(a.seed ^ 0xABCDEF);
  const auto t0 = std::chrono::steady_clock::now();
  bool stop = false;
  double err_sum = 0.0;
  auto parallel = [&](auto&& body) {
    std::vector<std::thread> ts;
    for (uint32_t t = 0; t < T; ++t) ts.emplace_back(body, t);
    for (auto& t : ts) t.join();
  };
  for (uint32_t pass = 1; pass <= a.passes && !stop; ++pass) {
    for (size_t i = order.size(); i > 1; --i) std::swap(order[i - 1], order[shuf.below((uint32_t)i)]);
    double errs = 0.0;
    for (uint64_t bi = 0; bi < steps_per_pass && !stop; ++bi) {
      const size_t lo = bi * a.batch, hi = std::min(order.size(), lo + a.batch);
      std::vector<GradMap> local(T);
      std::vector<double> err(T, 0.0);
      std::vector<uint32_t> cnt(T, 0);
      std::atomic<size_t> next{lo};
      parallel([&](uint32_t t) {
        for (;;) {
          const size_t q = next.fetch_add(1);
          if (q >= hi) return;
          const uint32_t v = order[q];
          XorShift rng(a.seed + bi + v);
          GradMap g;
          const float L = m.example(v, rng, g);
          if (!(L > 0.f)) continue;
          ++cnt[t];
          err[t] += L;
          for (auto& kv : g) {
            auto it = local[t].find(kv.first);
            if (it == local[t].end()) local[t].emplace(kv.first, std::move(kv.second));
            else for (uint32_t d = 0; d < D; ++d) it->second[d] += kv.second[d];
          }
        }
      });
      uint32_t contributing = 0;
      double e = 0.0;
      for (uint32_t t = 0; t < T; ++t) { contributing += cnt[t]; e += err[t]; }
      const float lr = cos_decay(step, a.alpha, total);  // mod.rs:234-250
      ++step;
      if (contributing > 0) {
        parallel([&](uint32_t s) {  // merge shard s (mod.rs:223-233), then its Adam rows
          GradMap shard;
          for (uint32_t t = 0; t < T; ++t)
            for (auto& kv : local[t]) {
              if (kv.first % T != s) continue;
              auto it = shard.find(kv.first);
              if (it == shard.end()) shard.emplace(kv.first, std::move(kv.second));
              else for (uint32_t d = 0; d < D; ++d) it->second[d] += kv.second[d];
            }
          m.adam(shard, lr, (float)pass);
        });
      }
      errs += contributing ? e / (double)(hi - lo) : 0.0;
      examples += hi - lo;
      ++batches_done;
      if (a.budget_s > 0) {
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (el > a.budget_s) stop = true;
      }
    }
    err_sum = errs / (double)steps_per_pass;
    if (!a.json) std::fprintf(stderr, "pass %u train error %.5f\n", pass, err_sum);
  }
  RunStats r;
  r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  r.examples = examples;
  r.batches = batches_done;
  r.train_error = err_sum;
  return r;
}

}  // namespace

// The timed CPU baseline for in-process callers (oracle/_build/libep_cpu.so, built with
// EP_CPU_NO_MAIN): trains on the given CSR arrays (mean
// model, or attention with dk > 0) for about budget_s seconds of batches.  out[0] =
// examples, out[1] = seconds, out[2] = batches.  Returns 0.
extern "C" int ep_cpu_bench(const uint64_t* indptr, const uint32_t* indices, const uint64_t* findptr,
                            const uint32_t* fids, uint64_t nodes, uint32_t vocab, uint32_t dims, uint32_t batch,
                            uint32_t negatives, uint32_t max_nodes, float gamma, float alpha, uint32_t heads,
                            uint32_t dk, uint32_t threads, float budget_s, double* out) {
  Args a;
  a.nodes = nodes;
  a.vocab = vocab;
  a.dims = dims;
  a.batch = batch;
  a.negatives = negatives;
  a.max_nodes = max_nodes;
  a.gamma = gamma;
  a.alpha = alpha;
  a.heads = heads;
  a.dk = dk;
  a.threads = threads;
  a.budget_s = budget_s;
  a.json = true;
  Model m{a, indptr, indices, findptr, fids, {}, {}, {}, {}};
  const RunStats r = run_training(m);
  out[0] = (double)r.examples;
  out[1] = r.seconds;
  out[2] = (double)r.batches;
  return 0;
}

namespace {
}  // namespace

#ifndef EP_CPU_NO_MAIN
int main(int argc, char** argv) {
  Args a;
  for (int i = 1; i < argc; ++i) {
    std::string k = argv[i];
    auto nx = [&]() { return std::string(argv[++i]); };
    if (k == "--dir") a.dir = nx();
    else if (k == "--nodes") a.nodes = std::stoull(nx());
    else if (k == "--vocab") a.vocab = (uint32_t)std::stoul(nx());
    else if (k == "--dims") a.dims = (uint32_t)std::stoul(nx());
    else if (k == "--heads") a.heads = (uint32_t)std::stoul(nx());
    else if (k == "--dk") a.dk = (uint32_t)std::stoul(nx());
    else if (k == "--batch") a.batch = (uint32_t)std::stoul(nx());
    else if (k == "--negatives") a.negatives = (uint32_t)std::stoul(nx());
    else if (k == "--max-nodes") a.max_nodes = (uint32_t)std::stoul(nx());
    else if (k == "--gamma") a.gamma = std::stof(nx());
    else if (k == "--alpha") a.alpha = std::stof(nx());
    else if (k == "--threads") a.threads = (uint32_t)std::stoul(nx());
    else if (k == "--passes") a.passes = (uint32_t)std::stoul(nx());
    else if (k == "--budget-s") a.budget_s = std::stof(nx());
    else if (k == "--seed") a.seed = std::stoull(nx());
    else if (k == "--table-out") a.table_out = nx();
    else if (k == "--train") a.train_file = nx();
    else if (k == "--json") a.json = true;
    else if (k == "--explicit") a.explicit_dir = nx();
    else if (k == "--lr") a.lr = std::stof(nx());
    else if (k == "--pass") a.pass = (uint32_t)std::stoul(nx());
    else { std::fprintf(stderr, "unknown argument %s\n", k.c_str()); return 2; }
  }
  const auto indptr = read_bin<uint64_t>(a.dir + "/indptr.bin");
  const auto indices = read_bin<uint32_t>(a.dir + "/indices.bin");
  const auto findptr = read_bin<uint64_t>(a.dir + "/findptr.bin");
  const auto fids = read_bin<uint32_t>(a.dir + "/fids.bin");
  if (indptr.size() != a.nodes + 1 || findptr.size() != a.nodes + 1) { std::fprintf(stderr, "size mismatch\n"); return 2; }
  Model m{a, indptr.data(), indices.data(), findptr.data(), fids.data(), {}, {}, {}, {}};
  if (!a.explicit_dir.empty()) return run_explicit(m);
  if (!a.train_file.empty()) m.train = read_bin<uint32_t>(a.train_file);
  const RunStats r = run_training(m);
  if (!a.table_out.empty()) write_bin(a.table_out, m.table);
  std::printf("{\"examples\": %llu, \"batches\": %llu, \"seconds\": %.4f, \"examples_per_s\": %.2f, "
              "\"threads\": %u, \"train_error\": %.6f}\n",
              (unsigned long long)r.examples, (unsigned long long)r.batches, r.seconds,
              (double)r.examples / r.seconds, a.threads, r.train_error);
  return 0;
}
#endif  // EP_CPU_NO_MAIN
