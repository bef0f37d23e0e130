#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "cloverleaf_ep.h"

// Drives the graph, feature and embedding file loaders (ep_io.cpp) over the files in the
// directory argv[1]; built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_loaders_sanitized.py, so any memory or UB error aborts the run.
int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  auto at = [&](const char* f) { return dir + "/" + f; };
  // graph
  ep_graph_file* g = nullptr;
  int rc = ep_graph_load_tsv(at("g.tsv").c_str(), 1, 1, 1, 0, &g);
  uint64_t n = 0, e = 0, nb = 0;
  if (rc == 0) {
    ep_graph_file_sizes(g, &n, &e, &nb);
    std::vector<uint64_t> ip(n + 1); std::vector<uint32_t> ix(e ? e : 1); std::vector<float> cdf(e ? e : 1);
    std::vector<char> names(nb ? nb : 1);
    ep_graph_file_copy(g, ip.data(), ix.data(), cdf.data(), names.data());
    ep_graph_file_free(g);
  }
  std::printf("graph rc=%d n=%llu e=%llu\n", rc, (unsigned long long)n, (unsigned long long)e);
  // features: nodes n0..n9 of type u, one known feature
  std::string nb2; std::vector<uint64_t> noff{0};
  for (int i = 0; i < 10; ++i) { nb2 += "u"; noff.push_back(nb2.size()); nb2 += "n" + std::to_string(i); noff.push_back(nb2.size()); }
  std::string fb = "featw0"; std::vector<uint64_t> foff{0, 4, 6};
  for (int kind = 0; kind < 3; ++kind) {
    ep_feature_file* f = nullptr;
    rc = ep_features_load_tsv(at("f.tsv").c_str(), nb2.data(), noff.data(), 10, fb.data(), foff.data(), 1, kind,
                              kind == 2 ? ":" : "feat", &f);
    uint64_t ns = 0, ni = 0, nn = 0, nbytes = 0;
    if (rc == 0) {
      ep_feature_file_sizes(f, &ns, &ni, &nn, &nbytes);
      std::vector<uint32_t> nodes(ns ? ns : 1), ids(ni ? ni : 1); std::vector<uint64_t> ip(ns + 1), off(2 * nn + 1);
      std::vector<char> blob(nbytes ? nbytes : 1);
      ep_feature_file_copy(f, nodes.data(), ip.data(), ids.data(), blob.data(), off.data());
      ep_feature_file_free(f);
    }
    std::printf("features kind %d rc=%d set=%llu ids=%llu new=%llu\n", kind, rc, (unsigned long long)ns,
                (unsigned long long)ni, (unsigned long long)nn);
  }
  // embeddings, with and without a filter, and a malformed file
  const char* tb = "u"; uint64_t toff[2] = {0, 1};
  for (int filt = 0; filt < 2; ++filt) {
    ep_embedding_file* m = nullptr;
    rc = ep_embeddings_load_tsv(at("e.tsv").c_str(), tb, toff, 1, filt, 1, &m);
    uint64_t r = 0, d = 0, b = 0;
    if (rc == 0) {
      ep_embedding_file_sizes(m, &r, &d, &b);
      std::vector<float> t(r * d ? r * d : 1); std::vector<char> blob(b ? b : 1); std::vector<uint64_t> off(2 * r + 1);
      ep_embedding_file_copy(m, t.data(), blob.data(), off.data());
      ep_embedding_file_free(m);
    }
    std::printf("embeddings filter %d rc=%d rows=%llu dims=%llu\n", filt, rc, (unsigned long long)r, (unsigned long long)d);
  }
  for (const char* badf : {"bad1.tsv", "bad2.tsv", "bad3.tsv"}) {
    const std::string bad = at(badf);
    ep_embedding_file* m = nullptr;
    rc = ep_embeddings_load_tsv(bad.c_str(), nullptr, nullptr, 0, 0, 0, &m);
    std::printf("bad %s rc=%d msg=%s\n", badf, rc, ep_graph_file_error());
    if (rc == 0) ep_embedding_file_free(m);
  }
  return 0;
}
