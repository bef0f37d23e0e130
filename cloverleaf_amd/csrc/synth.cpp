// Seeded synthetic inputs for the EP workloads (host only, multithreaded).
//
//  * Chung-Lu graphs: expected-degree weights w_i = (i + i0)^(-1/(beta-1)); each of
//    m edges picks its endpoints independently with probability proportional to w
//    (inverse-CDF sampling).  i0 is solved so the largest expected degree matches a
//    target.  Self loops are dropped; repeated edges are merged (GraphBuilder's
//    default deduplicate=true, src/graph.rs:120-142), columns sorted within a row.
//  * Bipartite variant: endpoints drawn from two separate weight sets (query, item).
//  * Feature lists: per-node count from 1+Poisson(lam), U{lo..hi} or log-uniform
//    [1, hi]; ids drawn Zipf(s) over a vocabulary (P(k) ~ (k+1)^-s).
//
// Output is deterministic for a seed, independent of the thread count: work is cut
// into fixed-size chunks, each with its own splitmix64 stream.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  uint64_t below(uint64_t n) { return (uint64_t)(((__uint128_t)next() * n) >> 64); }
};

uint64_t stream_seed(uint64_t seed, uint64_t purpose, uint64_t chunk) {
  Rng r(seed ^ (purpose * 0xD1B54A32D192ED03ull));
  r.s += chunk * 0x8CB92BA72F3D8DD7ull;
  return r.next();
}

unsigned n_threads() {
  const char* e = std::getenv("CLOVERLEAF_SYNTH_THREADS");
  unsigned t = e ? (unsigned)std::atoi(e) : std::thread::hardware_concurrency();
  return std::max(1u, std::min(t, 64u));
}

template <typename F>
void parallel_chunks(uint64_t n_chunks, F&& f) {
  std::atomic<uint64_t> next{0};
  const unsigned T = n_threads();
  std::vector<std::thread> ts;
  for (unsigned t = 0; t < T; ++t)
    ts.emplace_back([&]() {
      for (uint64_t c; (c = next.fetch_add(1)) < n_chunks;) f(c);
    });
  for (auto& t : ts) t.join();
}

// Normalised CDF of (i + i0)^(-alpha), i in [0, n).
std::vector<double> powerlaw_cdf(uint64_t n, double alpha, double i0) {
  std::vector<double> cdf(n);
  double acc = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    acc += std::pow((double)i + i0, -alpha);
    cdf[i] = acc;
  }
  for (auto& x : cdf) x /= acc;
  cdf[n - 1] = 1.0;
  return cdf;
}

// i0 such that the top node's share p0 = w0 / sum w gives endpoints*p0 ~= dmax.
double solve_i0(uint64_t n, double alpha, double endpoints, double dmax) {
  if (dmax <= 0) return 1.0;
  auto share = [&](double i0) {
    // sum_{i<n} (i+i0)^-alpha via a coarse integral + exact head
    double s = 0.0;
    const uint64_t head = std::min<uint64_t>(n, 4096);
    for (uint64_t i = 0; i < head; ++i) s += std::pow((double)i + i0, -alpha);
    if (n > head) {
      const double a = (double)head + i0 - 0.5, b = (double)n + i0 - 0.5;
      s += (std::abs(alpha - 1.0) < 1e-12) ? std::log(b / a)
                                           : (std::pow(b, 1.0 - alpha) - std::pow(a, 1.0 - alpha)) / (1.0 - alpha);
    }
    return std::pow(i0, -alpha) / s;
  };
  const double target = dmax / endpoints;
  double lo = 1e-3, hi = (double)n * 10;
  if (share(lo) <= target) return lo;
  for (int it = 0; it < 200; ++it) {
    const double mid = std::sqrt(lo * hi);
    if (share(mid) > target) lo = mid; else hi = mid;
  }
  return std::sqrt(lo * hi);
}

inline uint32_t draw_cdf(const std::vector<double>& cdf, double u) {
  const size_t k = (size_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
  return (uint32_t)(k < cdf.size() ? k : cdf.size() - 1);
}

struct Edge { uint32_t s, t; };

// Build a deduplicated CSR (columns sorted, self loops removed) from an edge list.
void build_csr(uint64_t n, const std::vector<Edge>& edges, uint64_t** indptr_out, uint32_t** indices_out,
               uint64_t* n_entries) {
  std::vector<uint64_t> cnt(n + 1, 0);
  for (const Edge& e : edges) cnt[e.s + 1]++;
  for (uint64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  std::vector<uint32_t> cols(edges.size());
  {
    std::vector<uint64_t> cur(cnt.begin(), cnt.end() - 1);
    for (const Edge& e : edges) cols[cur[e.s]++] = e.t;
  }
  // sort + unique within each row
  std::vector<uint64_t> newdeg(n, 0);
  const uint64_t CH = 16384;
  parallel_chunks((n + CH - 1) / CH, [&](uint64_t c) {
    const uint64_t lo = c * CH, hi = std::min(n, lo + CH);
    for (uint64_t i = lo; i < hi; ++i) {
      uint32_t* b = cols.data() + cnt[i];
      uint32_t* e = cols.data() + cnt[i + 1];
      std::sort(b, e);
      newdeg[i] = (uint64_t)(std::unique(b, e) - b);
    }
  });
  uint64_t* indptr = (uint64_t*)std::malloc((n + 1) * sizeof(uint64_t));
  indptr[0] = 0;
  for (uint64_t i = 0; i < n; ++i) indptr[i + 1] = indptr[i] + newdeg[i];
  uint32_t* indices = (uint32_t*)std::malloc(std::max<uint64_t>(indptr[n], 1) * sizeof(uint32_t));
  parallel_chunks((n + CH - 1) / CH, [&](uint64_t c) {
    const uint64_t lo = c * CH, hi = std::min(n, lo + CH);
    for (uint64_t i = lo; i < hi; ++i)
      std::memcpy(indices + indptr[i], cols.data() + cnt[i], newdeg[i] * sizeof(uint32_t));
  });
  *indptr_out = indptr;
  *indices_out = indices;
  *n_entries = indptr[n];
}

}  // namespace

extern "C" {

void synth_free(void* p) { std::free(p); }

// Unipartite Chung-Lu: n nodes, m sampled edges, exponent beta, largest expected
// degree ~dmax (<= 0: i0 = 1).  directed=0 adds both directions of every edge.
int synth_chung_lu(uint64_t n, uint64_t m, double beta, double dmax, int directed, uint64_t seed,
                   uint64_t** indptr, uint32_t** indices, uint64_t* n_entries) {
  if (n == 0 || n >= (1ull << 32) || beta <= 1.0) return -1;
  const double alpha = 1.0 / (beta - 1.0);
  const double endpoints = directed ? (double)m : 2.0 * (double)m;
  const double i0 = solve_i0(n, alpha, endpoints, dmax);
  const std::vector<double> cdf = powerlaw_cdf(n, alpha, i0);
  const uint64_t CH = 1 << 20;
  const uint64_t nch = (m + CH - 1) / CH;
  std::vector<Edge> edges(directed ? m : 2 * m);
  std::vector<uint8_t> keep(m, 1);
  parallel_chunks(nch, [&](uint64_t c) {
    Rng r(stream_seed(seed, 1, c));
    const uint64_t lo = c * CH, hi = std::min(m, lo + CH);
    for (uint64_t k = lo; k < hi; ++k) {
      const uint32_t a = draw_cdf(cdf, r.uniform());
      const uint32_t b = draw_cdf(cdf, r.uniform());
      if (a == b) keep[k] = 0;
      if (directed) edges[k] = Edge{a, b};
      else { edges[2 * k] = Edge{a, b}; edges[2 * k + 1] = Edge{b, a}; }
    }
  });
  // drop self loops
  size_t w = 0;
  for (uint64_t k = 0; k < m; ++k) {
    if (!keep[k]) continue;
    if (directed) edges[w++] = edges[k];
    else { edges[w++] = edges[2 * k]; edges[w++] = edges[2 * k + 1]; }
  }
  edges.resize(w);
  build_csr(n, edges, indptr, indices, n_entries);
  return 0;
}

// Bipartite Chung-Lu: queries [0, nq), items [nq, nq+ni); m undirected edges.
int synth_bipartite(uint64_t nq, uint64_t ni, uint64_t m, double beta, double dmax, uint64_t seed,
                    uint64_t** indptr, uint32_t** indices, uint64_t* n_entries) {
  const uint64_t n = nq + ni;
  if (nq == 0 || ni == 0 || n >= (1ull << 32) || beta <= 1.0) return -1;
  const double alpha = 1.0 / (beta - 1.0);
  const std::vector<double> cq = powerlaw_cdf(nq, alpha, solve_i0(nq, alpha, (double)m, dmax));
  const std::vector<double> ci = powerlaw_cdf(ni, alpha, solve_i0(ni, alpha, (double)m, dmax));
  const uint64_t CH = 1 << 20;
  std::vector<Edge> edges(2 * m);
  parallel_chunks((m + CH - 1) / CH, [&](uint64_t c) {
    Rng r(stream_seed(seed, 2, c));
    const uint64_t lo = c * CH, hi = std::min(m, lo + CH);
    for (uint64_t k = lo; k < hi; ++k) {
      const uint32_t q = draw_cdf(cq, r.uniform());
      const uint32_t it = (uint32_t)nq + draw_cdf(ci, r.uniform());
      edges[2 * k] = Edge{q, it};
      edges[2 * k + 1] = Edge{it, q};
    }
  });
  build_csr(n, edges, indptr, indices, n_entries);
  return 0;
}

// Community variant (quality workloads): node i is in community i % C.  Each edge picks
// u ~ w; with probability p_in the other endpoint is drawn ~ w within u's community,
// otherwise ~ w over all nodes.  Undirected.
int synth_chung_lu_communities(uint64_t n, uint64_t m, double beta, double dmax, uint32_t C, double p_in,
                               uint64_t seed, uint64_t** indptr, uint32_t** indices, uint64_t* n_entries) {
  if (n == 0 || n >= (1ull << 32) || beta <= 1.0 || C == 0) return -1;
  const double alpha = 1.0 / (beta - 1.0);
  const double i0 = solve_i0(n, alpha, 2.0 * (double)m, dmax);
  const std::vector<double> cdf = powerlaw_cdf(n, alpha, i0);
  // per-community CDFs over members i = c, c + C, ...
  std::vector<std::vector<double>> ccdf(C);
  for (uint32_t c = 0; c < C; ++c) {
    double acc = 0.0;
    for (uint64_t i = c; i < n; i += C) { acc += std::pow((double)i + i0, -alpha); ccdf[c].push_back(acc); }
    for (auto& x : ccdf[c]) x /= acc;
    if (!ccdf[c].empty()) ccdf[c].back() = 1.0;
  }
  const uint64_t CH = 1 << 20;
  std::vector<Edge> edges(2 * m);
  std::vector<uint8_t> keep(m, 1);
  parallel_chunks((m + CH - 1) / CH, [&](uint64_t ch) {
    Rng r(stream_seed(seed, 5, ch));
    const uint64_t lo = ch * CH, hi = std::min(m, lo + CH);
    for (uint64_t k = lo; k < hi; ++k) {
      const uint32_t a = draw_cdf(cdf, r.uniform());
      uint32_t b;
      if (r.uniform() < p_in) {
        const uint32_t c = a % C;
        b = c + C * draw_cdf(ccdf[c], r.uniform());
      } else {
        b = draw_cdf(cdf, r.uniform());
      }
      if (a == b) keep[k] = 0;
      edges[2 * k] = Edge{a, b};
      edges[2 * k + 1] = Edge{b, a};
    }
  });
  size_t wi = 0;
  for (uint64_t k = 0; k < m; ++k) {
    if (!keep[k]) continue;
    edges[wi++] = edges[2 * k];
    edges[wi++] = edges[2 * k + 1];
  }
  edges.resize(wi);
  build_csr(n, edges, indptr, indices, n_entries);
  return 0;
}

// Feature lists.  count_kind: 0 -> 1 + Poisson(p1); 1 -> U{p1..p2}; 2 -> log-uniform [1, p2].
// C > 0: node i draws Zipf ranks shifted by its community's offset (i % C), so each
// community has its own hot features.
int synth_features_ex(uint64_t n, int count_kind, double p1, double p2, uint32_t vocab, double zipf_s, uint64_t seed,
                      uint32_t C, uint64_t** f_indptr, uint32_t** f_ids);

int synth_features(uint64_t n, int count_kind, double p1, double p2, uint32_t vocab, double zipf_s, uint64_t seed,
                   uint64_t** f_indptr, uint32_t** f_ids) {
  return synth_features_ex(n, count_kind, p1, p2, vocab, zipf_s, seed, 0, f_indptr, f_ids);
}

int synth_features_ex(uint64_t n, int count_kind, double p1, double p2, uint32_t vocab, double zipf_s, uint64_t seed,
                      uint32_t C, uint64_t** f_indptr, uint32_t** f_ids) {
  if (n == 0 || vocab == 0) return -1;
  const uint64_t CH = 1 << 16;
  const uint64_t nch = (n + CH - 1) / CH;
  std::vector<uint32_t> cnt(n);
  parallel_chunks(nch, [&](uint64_t c) {
    Rng r(stream_seed(seed, 3, c));
    const uint64_t lo = c * CH, hi = std::min(n, lo + CH);
    for (uint64_t i = lo; i < hi; ++i) {
      uint32_t k = 1;
      if (count_kind == 0) {
        // Poisson by inversion (small mean)
        const double L = std::exp(-p1);
        double p = 1.0;
        uint32_t x = 0;
        do { p *= r.uniform(); ++x; } while (p > L);
        k = 1 + (x - 1);
      } else if (count_kind == 1) {
        const uint64_t lo_ = (uint64_t)p1, hi_ = (uint64_t)p2;
        k = (uint32_t)(lo_ + r.below(hi_ - lo_ + 1));
      } else {
        k = (uint32_t)std::floor(std::exp(r.uniform() * std::log(p2 + 1.0)));
        k = std::max<uint32_t>(1, std::min<uint32_t>(k, (uint32_t)p2));
      }
      cnt[i] = k;
    }
  });
  uint64_t* ip = (uint64_t*)std::malloc((n + 1) * sizeof(uint64_t));
  ip[0] = 0;
  for (uint64_t i = 0; i < n; ++i) ip[i + 1] = ip[i] + cnt[i];
  uint32_t* ids = (uint32_t*)std::malloc(std::max<uint64_t>(ip[n], 1) * sizeof(uint32_t));
  const std::vector<double> cdf = powerlaw_cdf(vocab, zipf_s, 1.0);
  parallel_chunks(nch, [&](uint64_t c) {
    Rng r(stream_seed(seed, 4, c));
    const uint64_t lo = c * CH, hi = std::min(n, lo + CH);
    for (uint64_t i = lo; i < hi; ++i) {
      const uint64_t off = C ? (stream_seed(seed, 6, i % C) % vocab) : 0;
      for (uint64_t q = ip[i]; q < ip[i + 1]; ++q) ids[q] = (uint32_t)((draw_cdf(cdf, r.uniform()) + off) % vocab);
    }
  });
  *f_indptr = ip;
  *f_ids = ids;
  return 0;
}

}  // extern "C"
