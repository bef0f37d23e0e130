// Native graph file loader (host C++): GraphReader::load (src/io.rs:270-323) with
// CSR::construct_from_edges (src/graph.rs:70-142) and CumCSR::convert (graph.rs:414-422,
// convert_edges_to_cdf 664-691), for inputs with billions of edges that a Python parser
// cannot take.  Plain text or gzip (by the ".gz" suffix, as open_file_for_reading,
// io.rs:200-211).
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <string_view>
#include <unordered_map>
#include <vector>

#include "cloverleaf_ep.h"

namespace {

thread_local std::string g_io_err;

int io_fail(int code, const std::string& msg) {
  g_io_err = msg;
  return code;
}

// Line source over a plain or gzip file; lines lose their "\n" or "\r\n" (BufRead::lines).
struct LineReader {
  FILE* f = nullptr;
  gzFile gz = nullptr;
  std::vector<char> buf = std::vector<char>(1 << 16);
  std::string line;

  bool open(const char* path) {
    const size_t n = std::strlen(path);
    if (n >= 3 && std::strcmp(path + n - 3, ".gz") == 0) {
      gz = gzopen(path, "rb");
      if (gz) gzbuffer(gz, 1 << 20);
      return gz != nullptr;
    }
    f = std::fopen(path, "rb");
    return f != nullptr;
  }
  // false at end of file
  bool next() {
    line.clear();
    bool any = false;
    while (true) {
      const char* r = gz ? gzgets(gz, buf.data(), (int)buf.size()) : std::fgets(buf.data(), (int)buf.size(), f);
      if (!r) break;
      any = true;
      const size_t len = std::strlen(buf.data());
      line.append(buf.data(), len);
      if (len && buf[len - 1] == '\n') break;
    }
    if (!any) return false;
    if (!line.empty() && line.back() == '\n') line.pop_back();
    if (!line.empty() && line.back() == '\r') line.pop_back();
    return true;
  }
  ~LineReader() {
    if (f) std::fclose(f);
    if (gz) gzclose(gz);
  }
};

// str::parse::<f32> (Rust): an optional sign, decimal or exponent notation, inf / infinity /
// nan in any case; nothing else on the field.
bool parse_f32(std::string_view s, float* out) {
  if (s.empty()) return false;
  bool neg = false;
  if (s[0] == '+' || s[0] == '-') {
    neg = s[0] == '-';
    s.remove_prefix(1);
    if (s.empty() || s[0] == '+' || s[0] == '-') return false;
  }
  float v = 0.f;
  const auto r = std::from_chars(s.data(), s.data() + s.size(), v, std::chars_format::general);
  if (r.ec != std::errc() || r.ptr != s.data() + s.size()) {
    std::string low(s);
    for (char& c : low) c = (char)std::tolower((unsigned char)c);
    if (low == "inf" || low == "infinity") v = INFINITY;
    else if (low == "nan") v = NAN;
    else return false;
  }
  *out = neg ? -v : v;
  return true;
}

struct NodeKey {
  std::string type, name;
  bool operator==(const NodeKey& o) const { return type == o.type && name == o.name; }
};
struct NodeKeyHash {
  size_t operator()(const NodeKey& k) const {
    return std::hash<std::string>()(k.type) * 1000003u ^ std::hash<std::string>()(k.name);
  }
};

}  // namespace

struct ep_graph_file {
  std::vector<uint64_t> indptr;
  std::vector<uint32_t> indices;
  std::vector<float> cdf;
  std::vector<NodeKey> names;
};

extern "C" {

int ep_graph_load_tsv(const char* path, int undirected, int weighted, int deduplicate, uint64_t skip_rows,
                      ep_graph_file** out) {
  if (!path || !out) return io_fail(EP_EINVAL, "NULL argument");
  *out = nullptr;
  LineReader rd;
  if (!rd.open(path)) return io_fail(EP_EINVAL, std::string("cannot open ") + path);
  std::unordered_map<NodeKey, uint32_t, NodeKeyHash> vocab;
  std::vector<NodeKey> names;
  struct Edge {
    uint32_t from, to;
    float w;
  };
  std::vector<Edge> edges;
  auto id_of = [&](std::string_view t, std::string_view n) -> uint32_t {
    NodeKey k{std::string(t), std::string(n)};
    auto it = vocab.find(k);
    if (it != vocab.end()) return it->second;
    const uint32_t id = (uint32_t)names.size();
    vocab.emplace(k, id);
    names.push_back(std::move(k));
    return id;
  };
  uint64_t i = 0;
  for (; rd.next(); ++i) {
    if (i < skip_rows) continue;
    std::string_view l(rd.line);
    std::string_view p[5];
    int np = 0;
    size_t s = 0;
    while (true) {
      const size_t t = l.find('\t', s);
      if (np < 5) p[np] = l.substr(s, t == std::string_view::npos ? std::string_view::npos : t - s);
      ++np;
      if (t == std::string_view::npos) break;
      s = t + 1;
    }
    if (np != 5) return io_fail(EP_EINVAL, std::to_string(i) + ": Malformed graph file: Expected 5 fields!");
    float w = 1.f;
    if (weighted && !parse_f32(p[4], &w))
      return io_fail(EP_EINVAL, std::to_string(i) + ": Malformed graph file! invalid float literal - \"" +
                                    std::string(p[4]) + "\"");
    const uint32_t f = id_of(p[0], p[1]);
    const uint32_t t = id_of(p[2], p[3]);
    edges.push_back({f, t, w});
    if (undirected) edges.push_back({t, f, w});
    if (names.size() >= 0xFFFFFFFFull) return io_fail(EP_ECAPACITY, "more than 2^32 - 1 nodes");
  }
  if (deduplicate) {
    // CSR::deduplicate_edges: stable sort by (from, to), weights of a repeated edge added
    // in file order
    std::stable_sort(edges.begin(), edges.end(),
                     [](const Edge& a, const Edge& b) { return a.from != b.from ? a.from < b.from : a.to < b.to; });
    size_t cur = 0;
    for (size_t k = 1; k < edges.size(); ++k) {
      if (edges[k].from == edges[cur].from && edges[k].to == edges[cur].to) edges[cur].w += edges[k].w;
      else edges[++cur] = edges[k];
    }
    if (!edges.empty()) edges.resize(cur + 1);
  }
  auto* g = new ep_graph_file();
  const uint64_t n = names.size();
  g->indptr.assign(n + 1, 0);
  for (const Edge& e : edges) ++g->indptr[e.from + 1];
  for (uint64_t k = 0; k < n; ++k) g->indptr[k + 1] += g->indptr[k];
  g->indices.resize(edges.size());
  g->cdf.resize(edges.size());
  std::vector<uint64_t> fill(g->indptr.begin(), g->indptr.end() - 1);
  for (const Edge& e : edges) {  // counting sort by source: file order within a row
    const uint64_t at = fill[e.from]++;
    g->indices[at] = e.to;
    g->cdf[at] = e.w;
  }
  // convert_edges_to_cdf per row, in f32
  for (uint64_t v = 0; v < n; ++v) {
    float* w = g->cdf.data() + g->indptr[v];
    const uint64_t m = g->indptr[v + 1] - g->indptr[v];
    if (!m) continue;
    float denom = 0.f;
    for (uint64_t k = 0; k < m; ++k) denom += w[k];
    if (denom == 0.f) {
      for (uint64_t k = 0; k < m; ++k) w[k] = 1.f;
      denom = (float)m;
    }
    float acc = 0.f;
    for (uint64_t k = 0; k < m; ++k) {
      acc += w[k];
      w[k] = std::min(acc / denom, 1.0f);
    }
    w[m - 1] = 1.f;
  }
  g->names = std::move(names);
  *out = g;
  return EP_OK;
}

int ep_graph_file_sizes(const ep_graph_file* g, uint64_t* n_nodes, uint64_t* n_entries, uint64_t* name_bytes) {
  if (!g || !n_nodes || !n_entries || !name_bytes) return io_fail(EP_EINVAL, "NULL argument");
  *n_nodes = g->names.size();
  *n_entries = g->indices.size();
  uint64_t b = 0;
  for (const NodeKey& k : g->names) b += k.type.size() + k.name.size() + 2;
  *name_bytes = b;
  return EP_OK;
}

int ep_graph_file_copy(const ep_graph_file* g, uint64_t* indptr, uint32_t* indices, float* cdf, char* names) {
  if (!g || !indptr || (!g->indices.empty() && (!indices || !cdf)) || !names) return io_fail(EP_EINVAL, "NULL argument");
  std::memcpy(indptr, g->indptr.data(), g->indptr.size() * sizeof(uint64_t));
  if (!g->indices.empty()) {
    std::memcpy(indices, g->indices.data(), g->indices.size() * sizeof(uint32_t));
    std::memcpy(cdf, g->cdf.data(), g->cdf.size() * sizeof(float));
  }
  char* o = names;
  for (const NodeKey& k : g->names) {
    std::memcpy(o, k.type.data(), k.type.size());
    o += k.type.size();
    *o++ = '\t';
    std::memcpy(o, k.name.data(), k.name.size());
    o += k.name.size();
    *o++ = '\n';
  }
  return EP_OK;
}

void ep_graph_file_free(ep_graph_file* g) { delete g; }

const char* ep_graph_file_error(void) { return g_io_err.c_str(); }

}  // extern "C"
