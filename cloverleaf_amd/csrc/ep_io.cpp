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
#include <memory>
#include <string>
#include <string_view>
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

// ryu::Buffer::format for an f32 (EmbeddingWriter::format_embedding, io.rs:69-76).  The
// digits are the shortest decimal that reads back to the same f32, the one closest to it
// when several are as short (std::to_chars gives exactly these).  With d the digits
// (n of them) and the value d x 10^k, kk = n + k, the layout is:
//   k >= 0 and kk <= 13   the integer and ".0"               1e12   -> 1000000000000.0
//   0 < kk <= 13          a point after kk digits            123.45
//   -6 < kk <= 0          "0.", -kk zeros, the digits        1e-6   -> 0.000001
//   otherwise             d[.ddd] "e" exponent kk - 1        1e13, 1e-7, 3.4028235e38
// zero as "0.0" (sign kept), NaN as "NaN", infinities as "inf" / "-inf".  Writes at most
// 17 characters to out (no terminator) and returns their count.
size_t format_f32_ryu(float x, char* out) {
  if (std::isnan(x)) { std::memcpy(out, "NaN", 3); return 3; }
  if (std::isinf(x)) {
    if (x < 0.f) { std::memcpy(out, "-inf", 4); return 4; }
    std::memcpy(out, "inf", 3);
    return 3;
  }
  char* p = out;
  if (std::signbit(x)) *p++ = '-';
  const float ax = std::fabs(x);
  if (ax == 0.f) {
    std::memcpy(p, "0.0", 3);
    return (size_t)(p - out) + 3;
  }
  char sci[32];
  const auto r = std::to_chars(sci, sci + sizeof(sci), ax, std::chars_format::scientific);  // d[.ddd]e±XX
  char dig[16];
  int nd = 0;
  const char* q = sci;
  for (; q < r.ptr && *q != 'e'; ++q)
    if (*q != '.') dig[nd++] = *q;
  int e10 = 0;
  std::from_chars(q + 1 + (q[1] == '+' ? 1 : 0), r.ptr, e10);
  const int k = e10 - (nd - 1), kk = e10 + 1;
  if (k >= 0 && kk <= 13) {
    std::memcpy(p, dig, nd);
    p += nd;
    for (int i = 0; i < k; ++i) *p++ = '0';
    *p++ = '.';
    *p++ = '0';
  } else if (kk > 0 && kk <= 13) {
    std::memcpy(p, dig, kk);
    p += kk;
    *p++ = '.';
    std::memcpy(p, dig + kk, nd - kk);
    p += nd - kk;
  } else if (kk > -6 && kk <= 0) {
    *p++ = '0';
    *p++ = '.';
    for (int i = 0; i < -kk; ++i) *p++ = '0';
    std::memcpy(p, dig, nd);
    p += nd;
  } else {
    *p++ = dig[0];
    if (nd > 1) {
      *p++ = '.';
      std::memcpy(p, dig + 1, nd - 1);
      p += nd - 1;
    }
    *p++ = 'e';
    const auto w = std::to_chars(p, p + 8, kk - 1);
    p = w.ptr;
  }
  return (size_t)(p - out);
}

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


// (type, name) -> id table: open addressing over a flat arena of the names' bytes, so a
// lookup hashes two string views and allocates nothing.
class NameTable {
 public:
  void reserve(size_t n) {
    size_t cap = 16;
    while (cap < 2 * n) cap <<= 1;
    if (cap > slots_.size()) rehash(cap);
  }
  size_t size() const { return ents_.size(); }
  // id of (t, n), or ~0u
  uint32_t find(std::string_view t, std::string_view n) const {
    if (slots_.empty()) return ~0u;
    const uint64_t h = hash(t, n);
    for (size_t i = h & mask_;; i = (i + 1) & mask_) {
      const uint32_t e = slots_[i];
      if (e == 0u) return ~0u;
      if (ents_[e - 1].h == h && equal(ents_[e - 1], t, n)) return e - 1;
    }
  }
  // id of (t, n), inserted with the next id when new (*fresh = true)
  uint32_t find_or_insert(std::string_view t, std::string_view n, bool* fresh) {
    if (2 * (ents_.size() + 1) > slots_.size()) rehash(slots_.empty() ? 16 : 2 * slots_.size());
    const uint64_t h = hash(t, n);
    size_t i = h & mask_;
    for (;; i = (i + 1) & mask_) {
      const uint32_t e = slots_[i];
      if (e == 0u) break;
      if (ents_[e - 1].h == h && equal(ents_[e - 1], t, n)) {
        *fresh = false;
        return e - 1;
      }
    }
    const uint32_t id = (uint32_t)ents_.size();
    ents_.push_back(Ent{arena_.size(), (uint32_t)t.size(), (uint32_t)n.size(), h});
    arena_.insert(arena_.end(), t.begin(), t.end());
    arena_.insert(arena_.end(), n.begin(), n.end());
    slots_[i] = id + 1u;
    *fresh = true;
    return id;
  }
  std::string_view type(uint32_t id) const { return {arena_.data() + ents_[id].off, ents_[id].tlen}; }
  std::string_view name(uint32_t id) const {
    return {arena_.data() + ents_[id].off + ents_[id].tlen, ents_[id].nlen};
  }

 private:
  struct Ent {
    uint64_t off;
    uint32_t tlen, nlen;
    uint64_t h;
  };
  static uint64_t hash(std::string_view t, std::string_view n) {
    const uint64_t a = std::hash<std::string_view>()(t), b = std::hash<std::string_view>()(n);
    uint64_t x = a * 0x9E3779B97F4A7C15ull ^ (b + 0x632BE59BD9B4E019ull + (a << 6) + (a >> 2));
    x ^= x >> 29;  // mix the high bits into the slot index
    return x;
  }
  bool equal(const Ent& e, std::string_view t, std::string_view n) const {
    return e.tlen == t.size() && e.nlen == n.size() &&
           (t.empty() || std::memcmp(arena_.data() + e.off, t.data(), t.size()) == 0) &&
           (n.empty() || std::memcmp(arena_.data() + e.off + e.tlen, n.data(), n.size()) == 0);
  }
  void rehash(size_t cap) {
    slots_.assign(cap, 0u);
    mask_ = cap - 1;
    for (uint32_t id = 0; id < (uint32_t)ents_.size(); ++id) {
      size_t i = ents_[id].h & mask_;
      while (slots_[i] != 0u) i = (i + 1) & mask_;
      slots_[i] = id + 1u;
    }
  }
  std::vector<char> arena_;
  std::vector<Ent> ents_;
  std::vector<uint32_t> slots_;
  size_t mask_ = 0;
};

// Byte length of the Unicode White_Space character (char::is_whitespace, what Rust's
// str::split_whitespace splits on) that starts at p, 0 if none.  UTF-8: U+0009-000D, U+0020,
// U+0085, U+00A0, U+1680, U+2000-200A, U+2028, U+2029, U+202F, U+205F, U+3000.
size_t ws_len(const unsigned char* p, const unsigned char* end) {
  const unsigned char c = p[0];
  if (c == ' ' || (c >= 0x09 && c <= 0x0D)) return 1;
  if (c == 0xC2 && end - p >= 2 && (p[1] == 0x85 || p[1] == 0xA0)) return 2;
  if (end - p >= 3) {
    if (c == 0xE1 && p[1] == 0x9A && p[2] == 0x80) return 3;
    if (c == 0xE2 && p[1] == 0x80 && ((p[2] >= 0x80 && p[2] <= 0x8A) || p[2] == 0xA8 || p[2] == 0xA9 || p[2] == 0xAF))
      return 3;
    if (c == 0xE2 && p[1] == 0x81 && p[2] == 0x9F) return 3;
    if (c == 0xE3 && p[1] == 0x80 && p[2] == 0x80) return 3;
  }
  return 0;
}

// Inserts the (type, name) pairs k = 0..n-1 of a name table, blob[off[2k], off[2k+1]) and
// blob[off[2k+1], off[2k+2]), with ids 0..n-1 (a repeated pair is malformed).
bool read_names(const char* blob, const uint64_t* off, uint64_t n, NameTable* out) {
  if (n == 0) return true;
  if (!blob || !off) return false;
  out->reserve(n);
  for (uint64_t k = 0; k < n; ++k) {
    if (off[2 * k] > off[2 * k + 1] || off[2 * k + 1] > off[2 * k + 2]) return false;
    bool fresh;
    out->find_or_insert(std::string_view(blob + off[2 * k], off[2 * k + 1] - off[2 * k]),
                        std::string_view(blob + off[2 * k + 1], off[2 * k + 2] - off[2 * k + 1]), &fresh);
    if (!fresh) return false;
  }
  return true;
}

}  // namespace

struct ep_feature_file {
  std::vector<uint32_t> nodes;     // nodes the file set, ascending id
  std::vector<uint64_t> indptr;    // their feature lists in ids
  std::vector<uint32_t> ids;
  NameTable feats;                 // known features, then those first seen in the file
  uint64_t n_known = 0;            // new features: ids n_known, n_known + 1, ...
};

struct ep_embedding_file {
  NameTable names;          // nodes in file order
  uint64_t dims = 0;
  std::vector<float> table;  // names.size() x dims
};

struct ep_graph_file {
  std::vector<uint64_t> indptr;
  std::vector<uint32_t> indices;
  std::vector<float> cdf;
  NameTable names;
};

extern "C" {

int ep_graph_load_tsv(const char* path, int undirected, int weighted, int deduplicate, uint64_t skip_rows,
                      ep_graph_file** out) {
  if (!path || !out) return io_fail(EP_EINVAL, "NULL argument");
  *out = nullptr;
  LineReader rd;
  if (!rd.open(path)) return io_fail(EP_EINVAL, std::string("cannot open ") + path);
  NameTable names;  // node ids in first-appearance order
  struct Edge {
    uint32_t from, to;
    float w;
  };
  std::vector<Edge> edges;
  auto id_of = [&](std::string_view t, std::string_view n) -> uint32_t {
    bool fresh;
    return names.find_or_insert(t, n, &fresh);
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
  for (uint32_t id = 0; id < (uint32_t)g->names.size(); ++id) b += g->names.type(id).size() + g->names.name(id).size() + 2;
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
  for (uint32_t id = 0; id < (uint32_t)g->names.size(); ++id) {
    const std::string_view kt = g->names.type(id), kn = g->names.name(id);
    std::memcpy(o, kt.data(), kt.size());
    o += kt.size();
    *o++ = '\t';
    std::memcpy(o, kn.data(), kn.size());
    o += kn.size();
    *o++ = '\n';
  }
  return EP_OK;
}

void ep_graph_file_free(ep_graph_file* g) { delete g; }

const char* ep_graph_file_error(void) { return g_io_err.c_str(); }

int ep_features_load_tsv(const char* path, const char* node_blob, const uint64_t* node_off, uint64_t n_nodes,
                         const char* feat_blob, const uint64_t* feat_off, uint64_t n_feats, int ns_kind,
                         const char* ns_value, ep_feature_file** out) {
  if (!path || !out || (ns_kind != EP_FNS_NODE_TYPE && !ns_value)) return io_fail(EP_EINVAL, "NULL argument");
  if (ns_kind < EP_FNS_STATIC || ns_kind > EP_FNS_PREFIX) return io_fail(EP_EINVAL, "unknown feature namespace");
  *out = nullptr;
  if (n_nodes >= 0xFFFFFFFFull || n_feats >= 0xFFFFFFFFull) return io_fail(EP_EINVAL, "too many names");
  auto res = std::make_unique<ep_feature_file>();
  NameTable nodes;
  if (!read_names(node_blob, node_off, n_nodes, &nodes) || !read_names(feat_blob, feat_off, n_feats, &res->feats))
    return io_fail(EP_EINVAL, "malformed name table");
  res->n_known = n_feats;
  const std::string ns = ns_value ? ns_value : "";
  const std::string_view nsv(ns);
  LineReader rd;
  if (!rd.open(path)) return io_fail(EP_EINVAL, std::string("cannot open ") + path);
  // every accepted line appends its list to a log; the last line of a node wins
  std::vector<uint64_t> log_start{0};
  std::vector<uint32_t> log_ids;
  std::vector<uint32_t> last(nodes.size(), 0xFFFFFFFFu);
  uint64_t line_no = 0;
  while (rd.next()) {
    ++line_no;
    const std::string& ln = rd.line;
    // line.split('\t') must give exactly three pieces (lib.rs:1591-1595)
    const size_t t1 = ln.find('\t');
    const size_t t2 = t1 == std::string::npos ? std::string::npos : ln.find('\t', t1 + 1);
    if (t1 == std::string::npos || t2 == std::string::npos || ln.find('\t', t2 + 1) != std::string::npos)
      return io_fail(EP_EINVAL, "Line " + std::to_string(line_no) +
                                    ":Malformed feature line! Need node_type<TAB>name<TAB>f1 f2 ...");
    const std::string_view ntype(ln.data(), t1), nname(ln.data() + t1 + 1, t2 - t1 - 1);
    const uint32_t node = nodes.find(ntype, nname);
    if (node == ~0u) continue;  // unknown node: the line is ignored, its tokens too
    const unsigned char* p = (const unsigned char*)ln.data() + t2 + 1;
    const unsigned char* end = (const unsigned char*)ln.data() + ln.size();
    while (p < end) {
      size_t w;
      while (p < end && (w = ws_len(p, end)) != 0) p += w;
      if (p >= end) break;
      const unsigned char* q = p;
      while (q < end && ws_len(q, end) == 0) ++q;
      const std::string_view tok((const char*)p, (size_t)(q - p));
      p = q;
      // FeatureNamespace::parse (lib.rs:1391-1403), ("feat", token) when it gives none
      std::string_view ft, fn;
      if (ns_kind == EP_FNS_STATIC) {
        ft = nsv;
        fn = tok;
      } else if (ns_kind == EP_FNS_NODE_TYPE) {
        ft = ntype;
        fn = tok;
      } else {
        const size_t idx = nsv.empty() ? 0 : tok.find(nsv);
        if (idx == std::string_view::npos) {
          ft = "feat";
          fn = tok;
        } else {  // the name starts one byte past the delimiter's start, as in the reference
          ft = tok.substr(0, idx);
          fn = idx + 1 <= tok.size() ? tok.substr(idx + 1) : std::string_view();
        }
      }
      if (res->feats.size() >= 0xFFFFFFFEull) return io_fail(EP_EINVAL, "too many features");
      bool fresh;
      log_ids.push_back(res->feats.find_or_insert(ft, fn, &fresh));
    }
    last[node] = (uint32_t)(log_start.size() - 1);
    log_start.push_back(log_ids.size());
  }
  res->indptr.push_back(0);
  for (uint32_t v = 0; v < (uint32_t)last.size(); ++v) {
    if (last[v] == 0xFFFFFFFFu) continue;
    const uint64_t a = log_start[last[v]], z = log_start[last[v] + 1];
    res->nodes.push_back(v);
    res->ids.insert(res->ids.end(), log_ids.begin() + (ptrdiff_t)a, log_ids.begin() + (ptrdiff_t)z);
    res->indptr.push_back(res->ids.size());
  }
  *out = res.release();
  return EP_OK;
}

int ep_feature_file_sizes(const ep_feature_file* f, uint64_t* n_set, uint64_t* n_ids, uint64_t* n_new,
                          uint64_t* new_name_bytes) {
  if (!f || !n_set || !n_ids || !n_new || !new_name_bytes) return io_fail(EP_EINVAL, "NULL argument");
  *n_set = f->nodes.size();
  *n_ids = f->ids.size();
  *n_new = f->feats.size() - f->n_known;
  uint64_t b = 0;
  for (uint32_t id = (uint32_t)f->n_known; id < (uint32_t)f->feats.size(); ++id)
    b += f->feats.type(id).size() + f->feats.name(id).size();
  *new_name_bytes = b;
  return EP_OK;
}

int ep_feature_file_copy(const ep_feature_file* f, uint32_t* nodes, uint64_t* indptr, uint32_t* ids,
                         char* new_blob, uint64_t* new_off) {
  if (!f || !indptr || !new_off || (!f->nodes.empty() && !nodes) || (!f->ids.empty() && !ids) ||
      (f->feats.size() > f->n_known && !new_blob))
    return io_fail(EP_EINVAL, "NULL argument");
  if (!f->nodes.empty()) std::memcpy(nodes, f->nodes.data(), f->nodes.size() * sizeof(uint32_t));
  std::memcpy(indptr, f->indptr.data(), f->indptr.size() * sizeof(uint64_t));
  if (!f->ids.empty()) std::memcpy(ids, f->ids.data(), f->ids.size() * sizeof(uint32_t));
  uint64_t o = 0;
  new_off[0] = 0;
  for (uint64_t k = 0; k + f->n_known < f->feats.size(); ++k) {
    const uint32_t id = (uint32_t)(f->n_known + k);
    const std::string_view t = f->feats.type(id), n = f->feats.name(id);
    std::memcpy(new_blob + o, t.data(), t.size());
    o += t.size();
    new_off[2 * k + 1] = o;
    std::memcpy(new_blob + o, n.data(), n.size());
    o += n.size();
    new_off[2 * k + 2] = o;
  }
  return EP_OK;
}

void ep_feature_file_free(ep_feature_file* f) { delete f; }

int ep_embeddings_load_tsv(const char* path, const char* type_blob, const uint64_t* type_off, uint64_t n_types,
                           int filter, uint64_t skip_rows, ep_embedding_file** out) {
  if (!path || !out || (filter && n_types && (!type_blob || !type_off))) return io_fail(EP_EINVAL, "NULL argument");
  *out = nullptr;
  NameTable types;  // filter types as (type, "")
  const std::string_view none("", 0);
  if (filter) {
    for (uint64_t k = 0; k < n_types; ++k) {
      if (type_off[k] > type_off[k + 1]) return io_fail(EP_EINVAL, "malformed type table");
      bool fresh;
      types.find_or_insert(std::string_view(type_blob + type_off[k], type_off[k + 1] - type_off[k]), none, &fresh);
    }
  }
  LineReader rd;
  if (!rd.open(path)) return io_fail(EP_EINVAL, std::string("cannot open ") + path);
  auto res = std::make_unique<ep_embedding_file>();
  std::vector<float> emb;
  uint64_t line_no = 0;
  while (rd.next()) {
    if (line_no++ < skip_rows) continue;  // RecordReader skips before filtering (io.rs:99)
    const std::string& ln = rd.line;
    const size_t t1 = ln.find('\t');
    // is_node_type (io.rs:226-231): the text before the first tab is a wanted type
    if (filter && (t1 == std::string::npos || types.find(std::string_view(ln.data(), t1), none) == ~0u)) continue;
    // line_to_embedding (io.rs:255-268): three tab fields, "[v1,v2,...]" with the first and
    // last byte dropped, values split at ',' and trimmed, each a whole f32 (fast_float::parse)
    const size_t t2 = t1 == std::string::npos ? std::string::npos : ln.find('\t', t1 + 1);
    bool ok = t2 != std::string::npos && ln.find('\t', t2 + 1) == std::string::npos && ln.size() - t2 - 1 >= 2;
    emb.clear();
    if (ok) {
      const char* a = ln.data() + t2 + 2;
      const char* z = ln.data() + ln.size() - 1;
      while (ok) {
        const char* c = static_cast<const char*>(std::memchr(a, ',', (size_t)(z - a)));
        const char* e = c ? c : z;
        const unsigned char* b = (const unsigned char*)a;
        const unsigned char* q = (const unsigned char*)e;
        size_t w;
        while (b < q && (w = ws_len(b, q)) != 0) b += w;
        while (b < q) {  // trailing White_Space: find the last character boundary
          const unsigned char* r = q - 1;
          while (r > b && (*r & 0xC0) == 0x80) --r;
          if (ws_len(r, q) == (size_t)(q - r)) q = r;
          else break;
        }
        float v;
        ok = parse_f32(std::string_view((const char*)b, (size_t)(q - b)), &v);
        if (ok) emb.push_back(v);
        if (!c) break;
        a = c + 1;
      }
    }
    if (!ok) return io_fail(EP_EINVAL, "Error parsing line: " + ln);
    const uint64_t i = res->names.size();
    if (i == 0) res->dims = emb.size();
    bool fresh;
    res->names.find_or_insert(std::string_view(ln.data(), t1), std::string_view(ln.data() + t1 + 1, t2 - t1 - 1),
                              &fresh);
    if (!fresh) return io_fail(EP_EDUPLICATE, "found duplicate node at " + std::to_string(i) + "!");
    if (emb.size() != res->dims) return io_fail(EP_EINVAL, "Embeddings have different sizes!");
    res->table.insert(res->table.end(), emb.begin(), emb.end());
  }
  *out = res.release();
  return EP_OK;
}

int ep_embedding_file_sizes(const ep_embedding_file* f, uint64_t* n_rows, uint64_t* dims, uint64_t* name_bytes) {
  if (!f || !n_rows || !dims || !name_bytes) return io_fail(EP_EINVAL, "NULL argument");
  *n_rows = f->names.size();
  *dims = f->dims;
  uint64_t b = 0;
  for (uint32_t id = 0; id < (uint32_t)f->names.size(); ++id) b += f->names.type(id).size() + f->names.name(id).size();
  *name_bytes = b;
  return EP_OK;
}

int ep_embedding_file_copy(const ep_embedding_file* f, float* table, char* name_blob, uint64_t* name_off) {
  if (!f || !name_off || (!f->table.empty() && !table) || (f->names.size() && !name_blob))
    return io_fail(EP_EINVAL, "NULL argument");
  if (!f->table.empty()) std::memcpy(table, f->table.data(), f->table.size() * sizeof(float));
  uint64_t o = 0;
  name_off[0] = 0;
  for (uint32_t id = 0; id < (uint32_t)f->names.size(); ++id) {
    const std::string_view t = f->names.type(id), n = f->names.name(id);
    std::memcpy(name_blob + o, t.data(), t.size());
    o += t.size();
    name_off[2 * (uint64_t)id + 1] = o;
    std::memcpy(name_blob + o, n.data(), n.size());
    o += n.size();
    name_off[2 * (uint64_t)id + 2] = o;
  }
  return EP_OK;
}

void ep_embedding_file_free(ep_embedding_file* f) { delete f; }

int ep_embeddings_write_tsv(const char* path, int comp_level, uint64_t n_rows, uint64_t dims, const float* table,
                            const char* name_blob, const uint64_t* name_off) {
  if (!path || (n_rows && (!table || !name_blob || !name_off))) return io_fail(EP_EINVAL, "NULL argument");
  if (comp_level > 9) return io_fail(EP_EINVAL, "comp_level must be in [0, 9] (or negative for the default)");
  const size_t pl = std::strlen(path);
  const bool gzip = pl >= 3 && std::strcmp(path + pl - 3, ".gz") == 0;
  FILE* f = nullptr;
  gzFile gz = nullptr;
  if (gzip) {
    // GzEncoder with Compression::new(level), Compression::fast() (1) by default
    const std::string mode = "wb" + std::to_string(comp_level < 0 ? 1 : comp_level);
    gz = gzopen(path, mode.c_str());
    if (gz) gzbuffer(gz, 1 << 20);
  } else {
    f = std::fopen(path, "wb");
  }
  if (!f && !gz) return io_fail(EP_EINVAL, std::string("cannot open ") + path + " for writing");
  std::string line;
  char num[48];
  int rc = EP_OK;
  for (uint64_t i = 0; i < n_rows && rc == EP_OK; ++i) {
    line.assign(name_blob + name_off[i], name_blob + name_off[i + 1]);  // "node_type\tname"
    line += "\t[";
    const float* row = table + i * dims;
    for (uint64_t j = 0; j < dims; ++j) {
      if (j) line += ',';
      line.append(num, format_f32_ryu(row[j], num));
    }
    line += "]\n";
    const bool ok = gz ? gzwrite(gz, line.data(), (unsigned)line.size()) == (int)line.size()
                       : std::fwrite(line.data(), 1, line.size(), f) == line.size();
    if (!ok) rc = io_fail(EP_EINVAL, std::string("write failed: ") + path);
  }
  if (gz && gzclose(gz) != Z_OK && rc == EP_OK) rc = io_fail(EP_EINVAL, std::string("gzip close failed: ") + path);
  if (f && std::fclose(f) != 0 && rc == EP_OK) rc = io_fail(EP_EINVAL, std::string("close failed: ") + path);
  return rc;
}

int ep_format_f32(float x, char* out) {
  if (!out) return io_fail(EP_EINVAL, "NULL argument");
  const size_t n = format_f32_ryu(x, out);
  out[n] = '\0';
  return (int)n;
}

}  // extern "C"
