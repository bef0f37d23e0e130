// Host runtime of the EP training step: context, device memory plan, training loop,
// hipGraph capture and the C ABI declared in include/cloverleaf_ep.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/cloverleaf_ep.h"
#include "ep_device.h"
#include "ep_kernels.h"

// Step buffers of ep_step_explicit_noalloc: one array per GPU, held from the library's
// load until the process ends.  Host code never uses this name as an address (on the
// host it is a stand-in): the device address comes from hipGetSymbolAddress, once per
// device (scratch_ptr).  Every call on a GPU shares that GPU's array.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define EP_HIP(call)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail(EP_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));                \
  } while (0)

// splitmix64: host-side stream for the train/valid split and per-pass shuffles.
struct HostRng {
  uint64_t s;
  explicit HostRng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  uint64_t below(uint64_t n) { return (uint64_t)(((__uint128_t)next() * n) >> 64); }
};

template <typename T>
void shuffle(std::vector<T>& v, HostRng& r) {
  for (size_t i = v.size(); i > 1; --i) std::swap(v[i - 1], v[r.below(i)]);
}

uint32_t next_pow2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return (uint32_t)p;
}

// Device address of `scratch` on each device, resolved by ep_create with the device
// current (hipGetSymbolAddress gives the current device's copy).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

struct ep_ctx {
  ep_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  unsigned char* scratch = nullptr;  // this device's `scratch` (device address)

  // graph / features
  uint64_t n_nodes = 0, n_entries = 0;
  uint32_t deg_max = 0;
  std::vector<uint32_t> degrees;
  std::vector<uint64_t> h_indptr;  // host copy of the CSR offsets (hooks of max_nodes Probability)
  std::vector<uint32_t> last_anchors, last_cnt;  // the last ep_sample_batch (its weights hook)
  uint64_t* d_indptr = nullptr;
  uint32_t* d_indices = nullptr;
  float* d_cdf = nullptr;        // per-edge CDF (weighted neighbours), ep_set_edge_weights
  std::vector<float> explicit_w;  // neighbour weights for ep_step_explicit (test hook)
  uint64_t f_nodes = 0, f_total = 0;
  uint32_t n_vocab = 0, f_max = 0;
  uint64_t* d_findptr = nullptr;
  uint32_t* d_fids = nullptr;

  // parameters
  void* d_table = nullptr;
  float* d_ostate = nullptr;  // optimizer records: [master |] m | v per row
  bool table_ready = false;

  // split
  std::vector<uint32_t> train, valid;
  bool split_set = false;
  uint32_t* d_train = nullptr;  // negative pool (train ids)
  uint32_t* d_valid = nullptr;
  uint32_t* d_perm = nullptr;   // pass permutation of train ids
  uint32_t* h_perm = nullptr;   // pinned staging
  float* d_pass_loss = nullptr;
  float* d_valid_loss = nullptr;

  // step buffers
  uint32_t cap_batch = 0;
  uint32_t cap_S = 0;
  uint32_t key_cap = 0;
  uint64_t unique_cap = 0;  // merged-gradient records the test hook may need
  ep::StepConfig base{};
  ep::StepState* d_state = nullptr;
  ep::StepState* d_vstate = nullptr;
  uint32_t* d_explicit = nullptr;  // anchors of explicit / sampled test batches
  float* d_xout = nullptr;

  // graph capture
  hipGraphExec_t step_exec = nullptr;
  bool step_exec_ready = false;

  // segment-Adam variant tuner: the first training steps rotate through the supported
  // values-per-lane variants (identical results) and time the scatter with events
  int seg_vpl = 0;                 // chosen variant (values per lane, unroll), 0 while tuning
  int seg_unroll = 0;
  int fwd_x = 0;                   // chosen k_forward variant (rows-in-flight multiplier), 0 while tuning
  std::vector<std::pair<int, int>> tune_cands;
  std::vector<int> tune_fwd;       // k_forward candidates, timed in the same steps
  std::vector<hipEvent_t> tune_ev;  // two per tuning step
  uint32_t tune_i = 0;

  ~ep_ctx() {
    for (hipEvent_t e : tune_ev) (void)hipEventDestroy(e);
    if (step_exec) (void)hipGraphExecDestroy(step_exec);
    for (void* p : allocs) (void)hipFree(p);
    if (h_perm) (void)hipHostFree(h_perm);
    if (stream) (void)hipStreamDestroy(stream);
  }

  template <typename T>
  int alloc(T** p, size_t count) {
    void* q = nullptr;
    const size_t bytes = std::max<size_t>(count * sizeof(T), 16);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(EP_ENOMEM, "hipMalloc(" + std::to_string(bytes) + " bytes) failed: " + hipGetErrorString(e));
    allocs.push_back(q);
    *p = (T*)q;
    return EP_OK;
  }

  template <typename T>
  void release(T*& p) {
    if (!p) return;
    auto it = std::find(allocs.begin(), allocs.end(), (void*)p);
    if (it != allocs.end()) allocs.erase(it);
    (void)hipFree((void*)p);
    p = nullptr;
  }
};

#define EP_TRY(x)            \
  do {                       \
    int r_ = (x);            \
    if (r_ != EP_OK) return r_; \
  } while (0)

namespace {

// Table row width: d_model for the mean model, H (2 d_k + d_model) for attention
// (model.rs:241-243).
uint32_t row_width(const ep_config& cfg) {
  return cfg.attention_dk ? cfg.attention_heads * (2u * cfg.attention_dk + cfg.dims) : cfg.dims;
}

bool ppr_loss(const ep_config& cfg) { return cfg.loss_kind == EP_LOSS_PPR; }

// Positive-list capacity of the step buffers: max_nodes (Fixed), the K walk endpoints
// (PPR), 0 with every neighbour (the CSR rows are read in place).
uint32_t list_cap(const ep_config& cfg) {
  return ppr_loss(cfg) ? cfg.negatives : (cfg.max_nodes_kind == EP_SAMPLE_FIXED ? cfg.max_nodes : 0u);
}

uint32_t neighbour_cap(const ep_ctx* c) {
  if (ppr_loss(c->cfg)) return c->cfg.negatives;
  return c->cfg.max_nodes_kind == EP_SAMPLE_FIXED ? c->cfg.max_nodes : c->deg_max;
}

// weighted_neighbor_averaging applies to ~h(v) from sampled neighbours (not PPR walks).
// The attention model does not average by weight (model.rs:214-218 passes false), and its
// weighted sampling draws over equal weights, which is a uniform sample.
bool weighted_avg(const ep_config& cfg) {
  return cfg.weighted_neighbor_averaging && !ppr_loss(cfg) && !cfg.attention_dk;
}

// max_features as feature dropout (mean model); the attention model takes it as Random
// attention's column sample instead (lib.rs:1288-1300).
bool feature_sampling(const ep_config& cfg) { return cfg.max_features_kind != EP_SAMPLE_ALL && !cfg.attention_dk; }

// max_nodes Probability: variable-length lists stored at the anchors' CSR offsets.
bool prob_nodes(const ep_config& cfg) { return cfg.max_nodes_kind == EP_SAMPLE_PROBABILITY && !ppr_loss(cfg); }

// Width of the neighbour lists of the sample / explicit-step hooks.
uint32_t hook_list_width(const ep_ctx* c) { return prob_nodes(c->cfg) ? std::max<uint32_t>(c->deg_max, 1u) : list_cap(c->cfg); }

// ~h(v) with a count per neighbour (one adjoint row each): weighted averaging or
// max_features sampling (the feature scale differs between neighbours).
bool neighbour_rows(const ep_config& cfg) { return weighted_avg(cfg) || feature_sampling(cfg); }

// FP16 adjoint rows: requested, with a BF16 table, the mean model and the sort scatter.
bool fp16_adj(const ep_config& cfg) {
  return cfg.fp16_adjoints && cfg.precision != EP_PREC_FP32 && !cfg.attention_dk && cfg.scatter == EP_SCATTER_SORT;
}

// Anchors per training step: concurrent_batches reference batches.
uint32_t group_anchors(const ep_ctx* c) { return c->cfg.batch_size * std::max<uint32_t>(1u, c->cfg.concurrent_batches); }

// Worst-case feature occurrences of one example: F_max * (1 + S + K).
uint64_t example_rows_max(const ep_ctx* c) {
  return (uint64_t)c->f_max * (1ull + neighbour_cap(c) + c->cfg.negatives);
}

// Places every per-step array of StepConfig b for batches of up to `batch` anchors and
// sets the step's capacities and grid shapes.  place(&ptr, count) gives an array its
// memory: hipMalloc (ensure_step_buffers) or a carve-out of the no-allocation scratch
// (ep_step_explicit_noalloc).  *keycap_out receives the key capacity.
template <typename Place>
int place_step_buffers(ep_ctx* c, uint32_t batch, ep::StepConfig& b, Place&& place, uint64_t* keycap_out) {
  const ep_config& cfg = c->cfg;
  const uint32_t K = cfg.negatives, NC = K + 2, D = row_width(cfg);
  const bool att = cfg.attention_dk != 0;
  const uint32_t S = list_cap(cfg);
  // key capacity: sum of the largest `batch` example sizes
  uint64_t keycap, nbr_max;  // nbr_max: neighbours of a batch at most
  if (cfg.max_nodes_kind == EP_SAMPLE_FIXED || ppr_loss(cfg)) {
    keycap = (uint64_t)batch * example_rows_max(c);
    nbr_max = (uint64_t)batch * S;
  } else {
    std::vector<uint32_t> d = c->degrees;
    const size_t take = std::min<size_t>(batch, d.size());
    std::partial_sort(d.begin(), d.begin() + take, d.end(), std::greater<uint32_t>());
    uint64_t s = 0;
    for (size_t i = 0; i < take; ++i) s += d[i];
    keycap = (uint64_t)c->f_max * (s + (uint64_t)batch * (1ull + K));
    nbr_max = s;
  }
  // weighted neighbours: one adjoint row per neighbour after the construction rows
  const bool wavg = weighted_avg(cfg);
  const bool nrows = neighbour_rows(cfg) && !att;
  const uint64_t nrow_cap = nrows ? nbr_max : 0;
  if (nrows && (uint64_t)batch * NC + nrow_cap > ep::kSlotMask)
    return fail(EP_ECAPACITY, "batch too large: neighbour adjoint rows exceed 2^29");
  if (keycap >= (1ull << 31)) return fail(EP_ECAPACITY, "batch too large: key capacity exceeds 2^31");
  if ((uint64_t)batch * NC > ep::kSlotMask) return fail(EP_ECAPACITY, "batch too large: slots exceed 2^29");
  // attention: one gradient row per feature occurrence, slot = key position
  if (att && keycap > ep::kSlotMask) return fail(EP_ECAPACITY, "batch too large: attention slots exceed 2^29");
  if (att && c->f_max > 64) return fail(EP_EINVAL, "attention supports at most 64 features per node");
  const uint64_t blk_cap = att ? (uint64_t)batch * (1ull + neighbour_cap(c) + K) : 0;
  if (blk_cap >= (1ull << 31)) return fail(EP_ECAPACITY, "batch too large: attention blocks exceed 2^31");
  const uint64_t hash = next_pow2(2 * std::max<uint64_t>(example_rows_max(c), 1));
  if (cfg.grad_mode == EP_GRAD_FIRST_LEAF && hash > 16384)
    return fail(EP_ECAPACITY, "first_leaf dedup needs <= 8192 feature occurrences per example (F_max*(1+S+K))");

  // long rows run past the plan's 128-entry window (>= 65 entries); their runs of more
  // than kPieceEntries entries are cut into pieces of kPieceEntries, fewer than
  // (run entries) / 16 per run
  const uint32_t long_cap = (uint32_t)(keycap / 65 + 1);
  const uint32_t piece_cap = (uint32_t)(keycap / 16 + 1);
  // persistent row kernels: 2048 blocks of kBlockWaves waves
  const int unique_grid = 2048;
  // radix-sort grid: ~8K keys per block at capacity, a multiple of 256 in [256, 2048]
  const int sort_grid = (int)std::min<uint64_t>(2048, std::max<uint64_t>(256, (keycap / 8192 + 255) / 256 * 256));
  EP_TRY(place(&b.nbr_cnt, batch));
  EP_TRY(place(&b.nbr, (size_t)batch * std::max<uint32_t>(S, 1)));
  EP_TRY(place(&b.neg, (size_t)batch * K));
  EP_TRY(place(&b.rcount, batch));
  const uint32_t ent_cap = (uint32_t)std::min<uint64_t>(256, 1ull + neighbour_cap(c) + K);
  EP_TRY(place(&b.ex_fs, (size_t)batch * ent_cap));
  EP_TRY(place(&b.ex_fn, (size_t)batch * ent_cap));
  EP_TRY(place(&b.keyoff, batch + 1));
  // adjoint rows: mean model one per construction (D wide); attention one gradient row
  // per feature occurrence
  if (fp16_adj(cfg)) {
    EP_TRY(place(&b.adjh, ((size_t)batch * NC + nrow_cap) * D));
    EP_TRY(place(&b.ascale, (size_t)batch * NC + nrow_cap));
  } else {
    EP_TRY(place(&b.adj, att ? (size_t)keycap * D : ((size_t)batch * NC + nrow_cap) * D));
  }
  if (wavg && S) EP_TRY(place(&b.nbr_w, (size_t)batch * S));
  if (nrows && !S) EP_TRY(place(&b.nbroff, batch));
  if (prob_nodes(cfg)) {
    EP_TRY(place(&b.nbr_sel, std::max<uint64_t>(c->n_entries, 1)));
    if (nrows) EP_TRY(place(&b.nbr_selw, std::max<uint64_t>(c->n_entries, 1)));
  }
  b.nrow_base = batch * NC;
  if (att) {
    const uint32_t dm = cfg.dims;
    EP_TRY(place(&b.blkoff, batch + 1));
    EP_TRY(place(&b.bitems, blk_cap));
    EP_TRY(place(&b.bnode, blk_cap));
    EP_TRY(place(&b.bslot, blk_cap));
    EP_TRY(place(&b.bmeta, blk_cap));
    EP_TRY(place(&b.bout, (size_t)blk_cap * dm));
    EP_TRY(place(&b.badj, (size_t)blk_cap * dm));
  }
  b.blk_cap = (uint32_t)blk_cap;
  EP_TRY(place(&b.coef, (size_t)batch * NC));
  EP_TRY(place(&b.elig, (size_t)batch * NC));
  EP_TRY(place(&b.loss, batch));
  for (int i = 0; i < 2; ++i) {
    EP_TRY(place(&b.keys[i], keycap));
    EP_TRY(place(&b.slots[i], keycap));
  }
  EP_TRY(place(&b.ghist, (size_t)2048 * sort_grid));
  EP_TRY(place(&b.long_st, long_cap));
  EP_TRY(place(&b.long_len, long_cap));
  EP_TRY(place(&b.long_pbase, long_cap));
  EP_TRY(place(&b.long_runs, (size_t)long_cap * ep::kMaxGroup));
  EP_TRY(place(&b.long_rlen, (size_t)long_cap * ep::kMaxGroup));
  // plan regions of cpr chunks of 64 keys; up to 64 short and 2 medium rows start in a
  // chunk.  Steps of at least 32 chunks per plan wave get 8 regions per wave, so the
  // update kernel's waves, taking regions from a queue, end within a few percent of each
  // other whatever the residency of its blocks.
  const uint64_t nwaves = (uint64_t)unique_grid * ep::kBlockWaves;
  const uint64_t chunks_max = (keycap + 63) / 64;
  const uint64_t nregions = nwaves * (chunks_max >= nwaves * 32 ? 8 : 1);
  const uint64_t cpr_max = (chunks_max + nregions - 1) / nregions;
  EP_TRY(place(&b.srec, nregions * cpr_max * 64));
  EP_TRY(place(&b.mrec, nregions * cpr_max * 2));
  EP_TRY(place(&b.scount, nregions));
  EP_TRY(place(&b.mcount, nregions));
  b.srec_cap = (uint32_t)(nregions * cpr_max * 64);
  b.mrec_cap = (uint32_t)(nregions * cpr_max * 2);
  b.n_regions = (uint32_t)nregions;
  EP_TRY(place(&b.piece_st, piece_cap));
  EP_TRY(place(&b.piece_info, piece_cap));
  EP_TRY(place(&b.part, (size_t)piece_cap * D));
  b.batch = batch;
  b.S = S;
  b.all_nbrs = (cfg.max_nodes_kind == EP_SAMPLE_ALL && !ppr_loss(cfg)) ? 1 : 0;
  b.key_cap = (uint32_t)keycap;
  b.sort_grid = sort_grid;
  b.long_cap = long_cap;
  b.piece_cap = piece_cap;
  b.hash_entries = (uint32_t)hash;
  b.ent_cap = ent_cap;
  b.sel_cap = feature_sampling(cfg) ? b.ent_cap : 0u;
  b.ids_cap = (uint32_t)std::min<uint64_t>(8192, example_rows_max(c));
  // emission fused into the forward (mean model): its first-leaf hash takes the LDS of the
  // constructions and partial sums, (NC + waves) * D floats, after the adjoints are out
  b.fuse_emit = att ? 0 : 1;
  {
    const uint64_t words = (uint64_t)(NC + ep::kBlockWaves) * D;
    uint32_t fh = 1;
    while ((uint64_t)fh * 2 <= words) fh *= 2;
    b.fuse_hash = fh;
    b.fuse_max = fh / 8 * 5;
  }
  b.unique_grid = unique_grid;
  *keycap_out = keycap;
  return EP_OK;
}

// Nulls every per-step array of a StepConfig (the ones place_step_buffers sets, the
// gradient-output and atomic-path arrays, the constructions hook).
void clear_step_buffers(ep::StepConfig& b) {
  b.nbr_cnt = nullptr; b.nbr = nullptr; b.neg = nullptr; b.rcount = nullptr; b.keyoff = nullptr;
  b.ex_fs = nullptr; b.ex_fn = nullptr;
  b.nbr_w = nullptr; b.nbroff = nullptr; b.nbr_sel = nullptr; b.nbr_selw = nullptr;
  b.adj = nullptr; b.adjh = nullptr; b.ascale = nullptr; b.coef = nullptr; b.elig = nullptr; b.loss = nullptr;
  b.pass_loss = nullptr; b.xout = nullptr;
  b.keys[0] = b.keys[1] = nullptr; b.slots[0] = b.slots[1] = nullptr; b.ghist = nullptr;
  b.G = nullptr; b.ufeat = nullptr;
  b.srec = nullptr; b.mrec = nullptr; b.scount = nullptr; b.mcount = nullptr;
  b.long_st = nullptr; b.long_len = nullptr; b.long_pbase = nullptr; b.long_runs = nullptr; b.long_rlen = nullptr;
  b.piece_st = nullptr; b.piece_info = nullptr; b.part = nullptr;
  b.Gd = nullptr; b.flag = nullptr; b.touched = nullptr;
  b.blkoff = nullptr; b.bitems = nullptr; b.bnode = nullptr; b.bslot = nullptr; b.bmeta = nullptr;
  b.bout = nullptr; b.badj = nullptr;
}

// Sizes every per-step buffer for batches of up to `batch` anchors.
int ensure_step_buffers(ep_ctx* c, uint32_t batch) {
  if (weighted_avg(c->cfg) && !c->d_cdf)
    return fail(EP_ESTATE, "weighted_neighbor_averaging needs the edge weights (ep_set_edge_weights)");
  if (feature_sampling(c->cfg) && c->f_max > 64)
    return fail(EP_EINVAL, "max_features sampling supports at most 64 features per node");
  if (batch <= c->cap_batch && c->base.nbr) return EP_OK;
  if (!c->d_indptr || !c->d_findptr) return fail(EP_ESTATE, "graph and features must be set first");
  const ep_config& cfg = c->cfg;
  const uint32_t D = row_width(cfg);
  // drop old buffers
  ep::StepConfig& b = c->base;
  c->release(b.nbr_cnt); c->release(b.nbr); c->release(b.neg); c->release(b.rcount); c->release(b.keyoff);
  c->release(b.ex_fs); c->release(b.ex_fn);
  c->release(b.nbr_w); c->release(b.nbroff); c->release(b.nbr_sel); c->release(b.nbr_selw);
  c->release(b.adj); c->release(b.adjh); c->release(b.ascale); c->release(b.coef); c->release(b.elig);
  c->release(b.loss);
  c->release(b.keys[0]); c->release(b.keys[1]); c->release(b.slots[0]); c->release(b.slots[1]);
  c->release(b.ghist); c->release(b.G); c->release(b.ufeat); c->release(b.long_st); c->release(b.long_len);
  c->release(b.long_pbase); c->release(b.long_runs); c->release(b.long_rlen); c->release(b.piece_st);
  c->release(b.piece_info);
  c->release(b.part); c->release(b.srec); c->release(b.mrec); c->release(b.scount); c->release(b.mcount);
  c->release(b.blkoff); c->release(b.bitems); c->release(b.bout); c->release(b.badj);
  c->release(b.bnode); c->release(b.bslot); c->release(b.bmeta);
  c->release(c->d_explicit);
  c->release(c->d_xout);
  if (c->step_exec) { (void)hipGraphExecDestroy(c->step_exec); c->step_exec = nullptr; c->step_exec_ready = false; }

  uint64_t keycap = 0;
  EP_TRY(place_step_buffers(c, batch, b, [c](auto** p, size_t n) { return c->alloc(p, n); }, &keycap));
  const uint32_t P = std::max<uint32_t>(1u, cfg.concurrent_batches);
  // merged-gradient records (G, ufeat) only serve the test hook's gradient output (one per
  // (row, batch) run); they are allocated by ep_step_explicit on first use
  c->unique_cap = std::min<uint64_t>((uint64_t)c->n_vocab * P, keycap) + 1;
  EP_TRY(c->alloc(&c->d_explicit, batch));
  if (!b.Gd && cfg.scatter == EP_SCATTER_ATOMIC) {
    EP_TRY(c->alloc(&b.Gd, (size_t)c->n_vocab * D));
    EP_TRY(c->alloc(&b.flag, c->n_vocab));
    EP_TRY(c->alloc(&b.touched, c->n_vocab));
    EP_HIP(hipMemsetAsync(b.Gd, 0, (size_t)c->n_vocab * D * sizeof(float), c->stream));
    EP_HIP(hipMemsetAsync(b.flag, 0, (size_t)c->n_vocab * sizeof(uint32_t), c->stream));
  }
  c->cap_batch = batch;
  c->cap_S = b.S;
  c->key_cap = (uint32_t)keycap;
  return EP_OK;
}

// Fills the parts of the base StepConfig that depend on graph / table / config.
void refresh_base(ep_ctx* c) {
  ep::StepConfig& b = c->base;
  const ep_config& cfg = c->cfg;
  b.indptr = c->d_indptr;
  b.indices = c->d_indices;
  b.findptr = c->d_findptr;
  b.fids = c->d_fids;
  b.n_nodes = c->n_nodes;
  b.n_vocab = c->n_vocab;
  // keys: a permutation of [0, 2^bits - 1) (> every row id)
  int bits = 1;
  while ((1ull << bits) - 1ull < (uint64_t)c->n_vocab) ++bits;
  const uint32_t P = std::max<uint32_t>(1u, cfg.concurrent_batches);
  uint32_t pbits = 0;
  while ((1u << pbits) < P) ++pbits;
  b.pbits = pbits;
  b.pmask = (1u << pbits) - 1u;
  b.sort_bits = bits;  // the batch rides in the sorted value, not in the key
  b.sentinel = (uint32_t)((1ull << bits) - 1ull);
  const uint64_t m = (1ull << bits) - 1ull;
  b.key_m = (uint32_t)m;
  b.key_bits = (uint32_t)bits;
  uint64_t a = m > 2 ? 0x9E3779B1ull % m : 1ull;  // multiplier coprime with m
  auto gcd = [](uint64_t x, uint64_t y) { while (y) { const uint64_t t = x % y; x = y; y = t; } return x; };
  while (m > 2 && (a < 2 || gcd(a, m) != 1)) a = (a + 1) % m;
  // inverse of a modulo m (extended Euclid)
  int64_t t0 = 0, t1 = 1, r0 = (int64_t)m, r1 = (int64_t)(a % std::max<uint64_t>(m, 1));
  while (r1 != 0) {
    const int64_t q = r0 / r1;
    int64_t tmp = t0 - q * t1; t0 = t1; t1 = tmp;
    tmp = r0 - q * r1; r0 = r1; r1 = tmp;
  }
  if (t0 < 0) t0 += (int64_t)m;
  b.key_a = (uint32_t)a;
  b.key_ainv = m > 1 ? (uint32_t)t0 : 1u;
  b.rbatch = cfg.batch_size;
  b.noise = cfg.noise;
  b.noise_decay = 1.0f;
  b.table = c->d_table;
  {
    const uint32_t D = row_width(cfg);
    const bool ms = cfg.precision == EP_PREC_BF16_MASTER;
    b.ostride = ms ? 3 * D : 2 * D;
    b.master = (ms && c->d_ostate) ? c->d_ostate : nullptr;
    b.mom = c->d_ostate ? c->d_ostate + (ms ? D : 0) : nullptr;
    b.var = c->d_ostate ? c->d_ostate + (ms ? 2 * D : D) : nullptr;
  }
  b.bf16 = cfg.precision != EP_PREC_FP32 ? 1 : 0;
  b.dims = (int)row_width(cfg);
  b.model = cfg.attention_dk ? 1 : 0;
  b.att_heads = (int)cfg.attention_heads;
  b.att_dk = (int)cfg.attention_dk;
  b.att_win = cfg.context_window ? cfg.context_window : 0xFFFFFFFFu;  // Sliding window_size, else Full
  b.att_dm = (int)cfg.dims;
  b.K = cfg.negatives;
  b.gamma = cfg.gamma;
  b.loss_kind = cfg.loss_kind;
  b.loss_p2 = cfg.loss_param2;
  b.wavg = weighted_avg(cfg) ? 1 : 0;
  b.wsample = (cfg.weighted_neighbor_sampling && !cfg.attention_dk) ? 1 : 0;
  b.cdf = c->d_cdf;
  b.mf_kind = feature_sampling(cfg) ? cfg.max_features_kind : 0;
  // Random attention: attention with max_features and no context_window (lib.rs:1290-1296)
  const bool rnd = cfg.attention_dk && cfg.context_window == 0 && cfg.max_features_kind != EP_SAMPLE_ALL;
  b.att_rkind = rnd ? cfg.max_features_kind : 0;
  b.att_rk = cfg.max_features;
  b.att_rp = cfg.max_features_p;
  b.mf_k = cfg.max_features;
  b.mf_p = cfg.max_features_p;
  b.nbr_rows = (neighbour_rows(cfg) && !cfg.attention_dk) ? 1 : 0;
  b.pnbrs = prob_nodes(cfg) ? 1 : 0;
  b.mn_p = cfg.max_nodes_p;
  b.hard_negs = std::min(cfg.hard_negatives, cfg.negatives);
  b.lw_kind = cfg.loss_weighting;
  b.lw_exp = cfg.loss_weight_exp;
  b.grad_mode = cfg.grad_mode;
  b.seed = cfg.seed;
  b.alpha = cfg.alpha;
  b.scatter = cfg.scatter;
  b.train = 1;
  b.explicit_samples = 0;
  b.pass_loss = nullptr;
  b.xout = nullptr;
}

// Launches one step.  update == false runs the sorted path in gradient-output mode (the
// merged gradient lands in G/ufeat, nothing is updated); *sorted_buf receives the key
// buffer holding the sorted keys.
// ev (nullable): 9 events recorded at the step start and after the scan, forward, emit,
// sort, plan, short-row, pieces and long-row phases (atomic path: after its scatter and
// its update, the remaining ones at the same point).
// The k_seg_update and k_forward variants of a launch that names none: the configured or
// tuned ones, (4, 8) and 1 while the tuner is choosing (every variant gives the same
// results).
void pick_variant(const ep_ctx* c, ep::StepConfig& sc) {
  if (sc.fwd_x == 0) sc.fwd_x = c->fwd_x ? c->fwd_x : 1;
  if (sc.seg_vpl != 0) return;
  sc.seg_vpl = c->seg_vpl ? c->seg_vpl : 4;
  sc.seg_unroll = c->seg_vpl ? c->seg_unroll : 8;
}

int launch_training_step(ep_ctx* c, const ep::StepConfig& sc0, bool update, float lr_override,
                         hipEvent_t* ev = nullptr, int* sorted_buf = nullptr) {
  hipStream_t s = c->stream;
  ep::StepConfig sc = sc0;
  pick_variant(c, sc);
  auto mark = [&](int i) -> hipError_t { return ev ? hipEventRecord(ev[i], s) : hipSuccess; };
  EP_HIP(mark(0));
  if (lr_override >= 0.f) EP_HIP(ep::launch_step_begin_lr(sc, c->d_state, lr_override, s));
  else EP_HIP(ep::launch_step_begin(sc, c->d_state, s));
  EP_HIP(ep::launch_sample(sc, c->d_state, s));
  EP_HIP(ep::launch_scan(sc, c->d_state, s));
  EP_HIP(mark(1));
  if (sc.model == 1) {
    EP_HIP(ep::launch_att_forward(sc, c->d_state, s));
    EP_HIP(ep::launch_att_backward(sc, c->d_state, s));
  } else {
    EP_HIP(ep::launch_forward(sc, c->d_state, s));
  }
  EP_HIP(mark(2));
  EP_HIP(ep::launch_emit(sc, c->d_state, s));
  EP_HIP(mark(3));
  if (sc.scatter == EP_SCATTER_SORT) {
    int buf = 0;
    EP_HIP(ep::launch_sort(sc, c->d_state, s, &buf));
    if (sorted_buf) *sorted_buf = buf;
    EP_HIP(mark(4));
    ep::StepConfig g = sc;
    g.grad_out = update ? 0 : 1;
    EP_HIP(ep::launch_scatter_sorted(g, c->d_state, buf, s, ev ? ev + 5 : nullptr));
  } else {
    EP_HIP(ep::launch_atomic_scatter(sc, c->d_state, s));
    EP_HIP(mark(4));
    if (update) EP_HIP(ep::launch_adam_touched(sc, c->d_state, s));
    for (int i = 5; i < 9; ++i) EP_HIP(mark(i));
  }
  return EP_OK;
}

int check_state_err(ep_ctx* c) {
  ep::StepState hs;
  EP_HIP(hipMemcpyAsync(&hs, c->d_state, offsetof(ep::StepState, dtot), hipMemcpyDeviceToHost, c->stream));
  EP_HIP(hipStreamSynchronize(c->stream));
  if (hs.err) return fail(EP_ECAPACITY, "key capacity exceeded in a step");
  return EP_OK;
}

int set_pass(ep_ctx* c, ep::StepState* st, uint32_t pass, uint64_t* next_step) {
  // pass and step_in_pass live at fixed offsets; write them with small async copies
  const uint32_t vals[2] = {pass, 0xFFFFFFFFu};
  EP_HIP(hipMemcpyAsync((char*)st + offsetof(ep::StepState, pass), vals, sizeof(vals), hipMemcpyHostToDevice,
                        c->stream));
  if (next_step)
    EP_HIP(hipMemcpyAsync((char*)st + offsetof(ep::StepState, next_step), next_step, sizeof(uint64_t),
                          hipMemcpyHostToDevice, c->stream));
  EP_HIP(hipStreamSynchronize(c->stream));
  return EP_OK;
}

int ensure_split(ep_ctx* c) {
  if (c->split_set) return EP_OK;
  // mod.rs:119-122: shuffle node ids, the last floor(N * valid_pct) are validation
  HostRng rng(c->cfg.seed);
  std::vector<uint32_t> ids(c->n_nodes);
  for (uint64_t i = 0; i < c->n_nodes; ++i) ids[i] = (uint32_t)i;
  shuffle(ids, rng);
  const uint64_t nv = (uint64_t)((float)c->n_nodes * c->cfg.valid_pct);
  c->valid.assign(ids.end() - nv, ids.end());
  ids.resize(c->n_nodes - nv);
  c->train.swap(ids);
  c->split_set = true;
  return EP_OK;
}

int upload_split(ep_ctx* c) {
  c->release(c->d_train);
  c->release(c->d_valid);
  c->release(c->d_perm);
  c->release(c->d_pass_loss);
  c->release(c->d_valid_loss);
  if (c->h_perm) { (void)hipHostFree(c->h_perm); c->h_perm = nullptr; }
  const size_t nt = c->train.size(), nv = c->valid.size();
  EP_TRY(c->alloc(&c->d_train, nt));
  EP_TRY(c->alloc(&c->d_valid, nv));
  EP_TRY(c->alloc(&c->d_perm, nt));
  EP_TRY(c->alloc(&c->d_pass_loss, nt));
  EP_TRY(c->alloc(&c->d_valid_loss, nv));
  EP_HIP(hipHostMalloc((void**)&c->h_perm, std::max<size_t>(nt, 1) * sizeof(uint32_t), hipHostMallocDefault));
  if (nt) EP_HIP(hipMemcpy(c->d_train, c->train.data(), nt * sizeof(uint32_t), hipMemcpyHostToDevice));
  if (nv) EP_HIP(hipMemcpy(c->d_valid, c->valid.data(), nv * sizeof(uint32_t), hipMemcpyHostToDevice));
  if (c->step_exec) { (void)hipGraphExecDestroy(c->step_exec); c->step_exec = nullptr; c->step_exec_ready = false; }
  return EP_OK;
}

// Shuffle the train ids for a pass (mod.rs:173) and stage them to the device.
int upload_pass_perm(ep_ctx* c, HostRng& rng) {
  EP_HIP(hipStreamSynchronize(c->stream));
  std::vector<uint32_t>& t = c->train;
  shuffle(t, rng);
  std::memcpy(c->h_perm, t.data(), t.size() * sizeof(uint32_t));
  if (!t.empty())
    EP_HIP(hipMemcpyAsync(c->d_perm, c->h_perm, t.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  return EP_OK;
}

ep::StepConfig train_config(ep_ctx* c) {
  ep::StepConfig sc = c->base;
  sc.perm = c->d_perm;
  sc.negpool = c->d_train;
  sc.n_negpool = (uint32_t)c->train.size();
  sc.n_batch_total = (uint32_t)c->train.size();
  const uint32_t B = c->cfg.batch_size;
  const uint32_t GA = group_anchors(c);
  sc.steps_per_pass = (uint32_t)((c->train.size() + GA - 1) / GA);  // launches per pass
  sc.total_updates = (uint32_t)((c->train.size() + B - 1) / B) * c->cfg.passes;  // reference batches
  // mod.rs:140-147: sd(step) = noise * decay^step, decay = exp((ln 1e-9 - ln noise) / T), in f32
  if (c->cfg.noise > 1e-9f)
    sc.noise_decay = std::exp((std::log(1e-9f) - std::log(c->cfg.noise)) / (float)sc.total_updates);
  sc.pass_loss = c->d_pass_loss;
  return sc;
}

// Number of tuning rounds; the first round is a warm-up and is not scored.
constexpr uint32_t kTuneRounds = 4;

bool tuning(const ep_ctx* c) { return c->cfg.scatter == EP_SCATTER_SORT && c->seg_vpl == 0; }

// One eager training step with the next candidate variants; after the last tuning step
// the fastest variant of each kernel (median time over the scored rounds) is fixed.  The
// k_seg_update candidates rotate with the step index modulo their count, the k_forward
// candidates modulo theirs; a round is as many steps as the larger count.
int tune_step(ep_ctx* c, const ep::StepConfig& sc) {
  if (c->tune_cands.empty()) {
    for (int v : {16, 12, 8, 4})
      for (int u : {v == 12 ? 2 : 8 / (v / 4), v == 12 ? 4 : 16 / (v / 4)})
        if (ep::seg_variant_supported((int)row_width(c->cfg), v, u)) c->tune_cands.push_back({v, u});
    if (c->tune_cands.empty()) c->tune_cands.push_back({4, 8});
    c->tune_fwd.clear();
    if (c->cfg.forward_variant) {
      c->tune_fwd.push_back(c->cfg.forward_variant);
    } else {
      for (int x : {1, 2})
        if (!c->cfg.attention_dk && ep::fwd_variant_supported((int)row_width(c->cfg), c->cfg.precision != EP_PREC_FP32, x))
          c->tune_fwd.push_back(x);
    }
    if (c->tune_fwd.empty()) c->tune_fwd.push_back(1);
    if (c->tune_cands.size() <= 1 && c->tune_fwd.size() <= 1) {
      c->seg_vpl = c->tune_cands[0].first;
      c->seg_unroll = c->tune_cands[0].second;
      c->fwd_x = c->tune_fwd[0];
      return launch_training_step(c, sc, true, -1.f);
    }
    const size_t ne = 9 * std::max(c->tune_cands.size(), c->tune_fwd.size()) * kTuneRounds;
    c->tune_ev.resize(ne, nullptr);
    for (auto& e : c->tune_ev) EP_HIP(hipEventCreate(&e));
  }
  const uint32_t n = (uint32_t)c->tune_cands.size(), nf = (uint32_t)c->tune_fwd.size();
  const uint32_t per_round = std::max(n, nf);
  const uint32_t i = c->tune_i++;
  ep::StepConfig t = sc;
  t.seg_vpl = c->tune_cands[i % n].first;
  t.seg_unroll = c->tune_cands[i % n].second;
  t.fwd_x = c->tune_fwd[i % nf];
  hipEvent_t evs[9];
  for (int k = 0; k < 9; ++k) evs[k] = c->tune_ev[9 * i + k];
  EP_TRY(launch_training_step(c, t, true, -1.f, evs));
  if (c->tune_i < per_round * kTuneRounds) return EP_OK;
  EP_HIP(hipEventSynchronize(c->tune_ev.back()));
  // median over the scored steps (rounds 1..) of a candidate's phase time: events a -> b
  auto median = [&](uint32_t ncand, uint32_t k, int a, int b, float* out) -> int {
    std::vector<float> ms;
    for (uint32_t j = per_round; j < per_round * kTuneRounds; ++j) {
      if (j % ncand != k) continue;
      float x = 0.f;
      EP_HIP(hipEventElapsedTime(&x, c->tune_ev[9 * j + a], c->tune_ev[9 * j + b]));
      ms.push_back(x);
    }
    std::sort(ms.begin(), ms.end());
    *out = ms.empty() ? INFINITY : ms[ms.size() / 2];
    return EP_OK;
  };
  float best = INFINITY;
  std::pair<int, int> pick = c->tune_cands.back();
  for (uint32_t k = 0; k < n; ++k) {
    float med = 0.f;
    EP_TRY(median(n, k, 5, 6, &med));  // short-row kernel
    if (med < best) { best = med; pick = c->tune_cands[k]; }
  }
  float bestf = INFINITY;
  int pickf = c->tune_fwd[0];
  for (uint32_t k = 0; k < nf; ++k) {
    float med = 0.f;
    EP_TRY(median(nf, k, 1, 2, &med));  // forward
    if (med < bestf) { bestf = med; pickf = c->tune_fwd[k]; }
  }
  c->seg_vpl = pick.first;
  c->seg_unroll = pick.second;
  c->fwd_x = pickf;
  return EP_OK;
}

int run_step(ep_ctx* c, const ep::StepConfig& sc) {
  if (tuning(c)) return tune_step(c, sc);
  if (!c->cfg.use_graph) return launch_training_step(c, sc, true, -1.f);
  if (!c->step_exec_ready) {
    hipGraph_t g = nullptr;
    EP_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int r = launch_training_step(c, sc, true, -1.f);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (r != EP_OK) { if (g) (void)hipGraphDestroy(g); return r; }
    if (e != hipSuccess) return fail(EP_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&c->step_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(EP_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    c->step_exec_ready = true;
  }
  EP_HIP(hipGraphLaunch(c->step_exec, c->stream));
  return EP_OK;
}

int validate_config(const ep_config* cfg) {
  if (cfg->attention_dk == 0) {
    if (!ep::dims_supported((int)cfg->dims)) return fail(EP_EINVAL, "dims must be one of 8,16,32,64,100,128,256");
  } else {
    if (!ep::att_shape_supported((int)cfg->attention_heads, (int)cfg->attention_dk, (int)cfg->dims))
      return fail(EP_EINVAL, "attention: d_k must be 4, 8, 16 or 32, heads 1..16, dims a multiple of 8 in [8, 256]");
    if (!ep::row_dims_supported((int)row_width(*cfg)))
      return fail(EP_EINVAL, "attention: row width heads*(2*d_k+dims) must be one of 8,16,32,64,100,128,256,384,"
                             "512,768");
    if (cfg->scatter != EP_SCATTER_SORT) return fail(EP_EINVAL, "attention needs the sort scatter");
  }
  if (cfg->batch_size == 0) return fail(EP_EINVAL, "batch_size must be > 0");
  if (cfg->negatives == 0 || cfg->negatives > 29) return fail(EP_EINVAL, "negatives must be in [1, 29]");
  if (cfg->passes == 0) return fail(EP_EINVAL, "passes must be > 0");
  if (cfg->max_nodes_kind == EP_SAMPLE_FIXED && (cfg->max_nodes == 0 || cfg->max_nodes > 1024))
    return fail(EP_EINVAL, "max_nodes (Fixed) must be in [1, 1024]");
  if (cfg->max_nodes_kind != EP_SAMPLE_FIXED && cfg->max_nodes_kind != EP_SAMPLE_ALL &&
      cfg->max_nodes_kind != EP_SAMPLE_PROBABILITY)
    return fail(EP_EINVAL, "max_nodes must be All, Fixed or Probability");
  if (cfg->max_nodes_kind == EP_SAMPLE_PROBABILITY && !(cfg->max_nodes_p > 0.f && cfg->max_nodes_p < 1.f))
    return fail(EP_EINVAL, "max_nodes (Probability) must be in (0, 1)");
  if (cfg->max_features_kind != EP_SAMPLE_ALL && cfg->max_features_kind != EP_SAMPLE_FIXED &&
      cfg->max_features_kind != EP_SAMPLE_PROBABILITY)
    return fail(EP_EINVAL, "max_features must be All, Fixed or Probability");
  if (cfg->max_features_kind == EP_SAMPLE_FIXED && cfg->max_features == 0)
    return fail(EP_EINVAL, "max_features (Fixed) must be >= 1");
  if (cfg->max_features_kind == EP_SAMPLE_PROBABILITY && !(cfg->max_features_p > 0.f && cfg->max_features_p < 1.f))
    return fail(EP_EINVAL, "max_features (Probability) must be in (0, 1)");
  if (cfg->loss_kind < EP_LOSS_MARGIN || cfg->loss_kind > EP_LOSS_PPR) return fail(EP_EINVAL, "loss_kind");
  if (!std::isfinite(cfg->gamma) || !std::isfinite(cfg->loss_param2))
    return fail(EP_EINVAL, "loss parameters must be finite");
  if (cfg->loss_kind == EP_LOSS_PPR && !(cfg->loss_param2 >= 0.f && cfg->loss_param2 <= 1.f))
    return fail(EP_EINVAL, "ppr restart_p must be in [0, 1]");
  if (!(cfg->noise >= 0.f) || !std::isfinite(cfg->noise)) return fail(EP_EINVAL, "noise must be finite and >= 0");
  if (cfg->precision != EP_PREC_FP32 && cfg->precision != EP_PREC_BF16 && cfg->precision != EP_PREC_BF16_MASTER)
    return fail(EP_EINVAL, "precision");
  if (cfg->grad_mode != EP_GRAD_FIRST_LEAF && cfg->grad_mode != EP_GRAD_EXACT) return fail(EP_EINVAL, "grad_mode");
  if (cfg->scatter != EP_SCATTER_SORT && cfg->scatter != EP_SCATTER_ATOMIC) return fail(EP_EINVAL, "scatter");
  if (cfg->loss_weighting < EP_LW_NONE || cfg->loss_weighting > EP_LW_EXPONENTIAL)
    return fail(EP_EINVAL, "loss_weighting");
  if (!(cfg->valid_pct >= 0.f && cfg->valid_pct < 1.f)) return fail(EP_EINVAL, "valid_pct must be in [0, 1)");
  if (cfg->concurrent_batches != 1 && cfg->concurrent_batches != 2 && cfg->concurrent_batches != 4 &&
      cfg->concurrent_batches != 8)
    return fail(EP_EINVAL, "concurrent_batches must be 1, 2, 4 or 8");
  if (cfg->concurrent_batches > 1 && cfg->scatter != EP_SCATTER_SORT)
    return fail(EP_EINVAL, "concurrent_batches > 1 needs the sort scatter");
  if (cfg->fp16_adjoints != 0 && cfg->fp16_adjoints != 1) return fail(EP_EINVAL, "fp16_adjoints must be 0 or 1");
  if (cfg->fp16_adjoints && (cfg->precision == EP_PREC_FP32 || cfg->attention_dk || cfg->scatter != EP_SCATTER_SORT))
    return fail(EP_EINVAL, "fp16_adjoints needs a BF16 table (precision bf16 or bf16-master), the mean model and the "
                           "sort scatter");
  if (cfg->forward_variant != 0 && cfg->forward_variant != 1 &&
      !(cfg->forward_variant == 2 && !cfg->attention_dk &&
        ep::fwd_variant_supported((int)row_width(*cfg), cfg->precision != EP_PREC_FP32, 2)))
    return fail(EP_EINVAL, "forward_variant must be 0 (tuned), 1, or 2 for the widths the 16-byte gather serves "
                           "(mean model)");
  if ((cfg->seg_vpl == 0) != (cfg->seg_unroll == 0) ||
      (cfg->seg_vpl != 0 && !ep::seg_variant_supported((int)row_width(*cfg), cfg->seg_vpl, cfg->seg_unroll)))
    return fail(EP_EINVAL, "seg_vpl/seg_unroll must both be 0 (tuned) or a built variant: vpl in {4,8,12,16} "
                           "dividing the row width, unroll in {8,16}/(vpl/4) ({2,4} for vpl 12)");
  return EP_OK;
}

}  // namespace

extern "C" {

const char* ep_last_error(void) { return g_err.c_str(); }

int ep_config_default(ep_config* cfg) {
  if (!cfg) return fail(EP_EINVAL, "cfg is NULL");
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->dims = 100;
  cfg->batch_size = 50;
  cfg->passes = 100;
  cfg->negatives = 1;
  cfg->alpha = 0.9f;
  cfg->gamma = 1.0f;
  cfg->valid_pct = 0.1f;
  cfg->max_nodes_kind = EP_SAMPLE_ALL;
  cfg->max_features_kind = EP_SAMPLE_ALL;
  cfg->loss_weighting = EP_LW_NONE;
  cfg->precision = EP_PREC_FP32;
  cfg->grad_mode = EP_GRAD_FIRST_LEAF;
  cfg->scatter = EP_SCATTER_SORT;
  cfg->use_graph = 0;
  cfg->seed = 20222022ull;
  cfg->device = 0;
  cfg->concurrent_batches = 1;
  cfg->attention_heads = 1;
  return EP_OK;
}

int ep_create(const ep_config* cfg, ep_ctx** out) {
  if (!cfg || !out) return fail(EP_EINVAL, "NULL argument");
  *out = nullptr;
  EP_TRY(validate_config(cfg));
  int ndev = 0;
  EP_HIP(hipGetDeviceCount(&ndev));
  if (cfg->device < 0 || cfg->device >= ndev) return fail(EP_EINVAL, "no such HIP device");
  if (cfg->device >= kMaxDevices) return fail(EP_EINVAL, "device index beyond the supported 64");
  EP_HIP(hipSetDevice(cfg->device));
  void* sp = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[cfg->device]) {
      void* p = nullptr;
      EP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      if (!p) return fail(EP_EHIP, "hipGetSymbolAddress(scratch) returned NULL");
      g_scratch_ptr[cfg->device] = p;
    }
    sp = g_scratch_ptr[cfg->device];
  }
  ep_ctx* c = new ep_ctx();
  c->scratch = (unsigned char*)sp;
  c->cfg = *cfg;
  c->seg_vpl = cfg->seg_vpl;  // 0: tuned during the first training steps
  c->seg_unroll = cfg->seg_unroll;
  c->fwd_x = cfg->forward_variant;  // 0: tuned with the segment tile
  c->device = cfg->device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete c; return fail(EP_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)); }
  int r = c->alloc(&c->d_state, 1);
  if (r == EP_OK) r = c->alloc(&c->d_vstate, 1);
  if (r != EP_OK) { delete c; return r; }
  ep::StepState zero{};
  zero.next_step = 1;
  e = hipMemcpy(c->d_state, &zero, sizeof(zero), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->d_vstate, &zero, sizeof(zero), hipMemcpyHostToDevice);
  if (e != hipSuccess) { delete c; return fail(EP_EHIP, std::string("hipMemcpy: ") + hipGetErrorString(e)); }
  *out = c;
  return EP_OK;
}

void ep_destroy(ep_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
}

int ep_set_graph(ep_ctx* c, uint64_t n_nodes, const uint64_t* indptr, const uint32_t* indices) {
  if (!c || !indptr || (!indices && indptr[n_nodes] > 0)) return fail(EP_EINVAL, "NULL argument");
  if (n_nodes == 0 || n_nodes >= (1ull << 32)) return fail(EP_EINVAL, "n_nodes must be in [1, 2^32)");
  if (indptr[0] != 0) return fail(EP_EINVAL, "indptr[0] must be 0");
  std::vector<uint32_t> deg(n_nodes);
  uint32_t dmax = 0;
  for (uint64_t i = 0; i < n_nodes; ++i) {
    if (indptr[i + 1] < indptr[i]) return fail(EP_EINVAL, "indptr must be non-decreasing");
    const uint64_t d = indptr[i + 1] - indptr[i];
    if (d >= (1ull << 32)) return fail(EP_EINVAL, "degree too large");
    deg[i] = (uint32_t)d;
    dmax = std::max(dmax, deg[i]);
  }
  const uint64_t ne = indptr[n_nodes];
  for (uint64_t i = 0; i < ne; ++i)
    if (indices[i] >= n_nodes) return fail(EP_EINVAL, "edge target out of range");
  EP_HIP(hipSetDevice(c->device));
  c->release(c->d_indptr);
  c->release(c->d_indices);
  c->release(c->d_cdf);
  EP_TRY(c->alloc(&c->d_indptr, n_nodes + 1));
  EP_TRY(c->alloc(&c->d_indices, ne));
  EP_HIP(hipMemcpy(c->d_indptr, indptr, (n_nodes + 1) * sizeof(uint64_t), hipMemcpyHostToDevice));
  if (ne) EP_HIP(hipMemcpy(c->d_indices, indices, ne * sizeof(uint32_t), hipMemcpyHostToDevice));
  c->n_nodes = n_nodes;
  c->n_entries = ne;
  c->deg_max = dmax;
  c->degrees.swap(deg);
  c->h_indptr.assign(indptr, indptr + n_nodes + 1);
  c->cap_batch = 0;  // capacities depend on degrees
  c->split_set = false;
  return EP_OK;
}

int ep_set_edge_weights(ep_ctx* c, const float* cdf) {
  if (!c || (!cdf && c->n_entries)) return fail(EP_EINVAL, "NULL argument");
  if (!c->d_indptr) return fail(EP_ESTATE, "the graph must be set first");
  for (uint64_t i = 0; i < c->n_entries; ++i)
    if (!(cdf[i] >= 0.f) || !std::isfinite(cdf[i])) return fail(EP_EINVAL, "edge CDF values must be finite and >= 0");
  EP_HIP(hipSetDevice(c->device));
  c->release(c->d_cdf);
  EP_TRY(c->alloc(&c->d_cdf, std::max<uint64_t>(c->n_entries, 1)));
  if (c->n_entries) EP_HIP(hipMemcpy(c->d_cdf, cdf, c->n_entries * sizeof(float), hipMemcpyHostToDevice));
  return EP_OK;
}

int ep_set_features(ep_ctx* c, uint64_t n_nodes, uint32_t n_vocab, const uint64_t* f_indptr, const uint32_t* f_ids) {
  if (!c || !f_indptr || !f_ids) return fail(EP_EINVAL, "NULL argument");
  if (c->n_nodes && n_nodes != c->n_nodes) return fail(EP_EINVAL, "features must cover every graph node");
  if (n_vocab == 0 || n_vocab >= (1u << 27) - 1) return fail(EP_EINVAL, "n_vocab must be in [1, 2^27-1)");
  if (f_indptr[0] != 0) return fail(EP_EINVAL, "f_indptr[0] must be 0");
  uint32_t fmax = 0;
  for (uint64_t i = 0; i < n_nodes; ++i) {
    if (f_indptr[i + 1] <= f_indptr[i])
      return fail(EP_EINVAL, "every node needs at least one feature (fill_missing_nodes)");
    fmax = (uint32_t)std::max<uint64_t>(fmax, f_indptr[i + 1] - f_indptr[i]);
  }
  const uint64_t tot = f_indptr[n_nodes];
  for (uint64_t i = 0; i < tot; ++i)
    if (f_ids[i] >= n_vocab) return fail(EP_EINVAL, "feature id out of range");
  EP_HIP(hipSetDevice(c->device));
  c->release(c->d_findptr);
  c->release(c->d_fids);
  EP_TRY(c->alloc(&c->d_findptr, n_nodes + 1));
  EP_TRY(c->alloc(&c->d_fids, tot));
  EP_HIP(hipMemcpy(c->d_findptr, f_indptr, (n_nodes + 1) * sizeof(uint64_t), hipMemcpyHostToDevice));
  EP_HIP(hipMemcpy(c->d_fids, f_ids, tot * sizeof(uint32_t), hipMemcpyHostToDevice));
  if (c->n_vocab != n_vocab) {
    c->release(c->d_table); c->release(c->d_ostate);
    c->release(c->base.Gd); c->release(c->base.flag); c->release(c->base.touched);
    c->table_ready = false;
  }
  c->f_nodes = n_nodes;
  c->f_total = tot;
  c->n_vocab = n_vocab;
  c->f_max = fmax;
  c->cap_batch = 0;
  return EP_OK;
}

int ep_set_table(ep_ctx* c, const float* rows) {
  if (!c) return fail(EP_EINVAL, "NULL ctx");
  if (!c->n_vocab) return fail(EP_ESTATE, "features must be set first");
  EP_HIP(hipSetDevice(c->device));
  const size_t n = (size_t)c->n_vocab * row_width(c->cfg);
  const bool bf16 = c->cfg.precision != EP_PREC_FP32;
  const bool master = c->cfg.precision == EP_PREC_BF16_MASTER;
  const uint32_t D = row_width(c->cfg);
  const size_t ostride = master ? 3 * (size_t)D : 2 * (size_t)D;
  if (!c->d_table) {
    if (bf16) {
      uint16_t* t = nullptr;
      EP_TRY(c->alloc(&t, n));
      c->d_table = t;
    } else {
      float* t = nullptr;
      EP_TRY(c->alloc(&t, n));
      c->d_table = t;
    }
    EP_TRY(c->alloc(&c->d_ostate, (size_t)c->n_vocab * ostride));
  }
  refresh_base(c);
  if (rows && master) {
    // FP32 rows into the master slots of the records, bf16 copy rounded from them
    EP_HIP(hipMemsetAsync(c->d_ostate, 0, (size_t)c->n_vocab * ostride * sizeof(float), c->stream));
    EP_HIP(hipMemcpy2DAsync(c->d_ostate, ostride * sizeof(float), rows, D * sizeof(float), D * sizeof(float),
                            c->n_vocab, hipMemcpyHostToDevice, c->stream));
    EP_HIP(ep::launch_fill_bf16(c->d_ostate, (uint16_t*)c->d_table, n, D, (uint32_t)ostride, c->stream));
  } else if (rows && bf16) {
    // stage the FP32 rows in the (not yet used) optimizer records, round them, then clear
    EP_HIP(hipMemcpyAsync(c->d_ostate, rows, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    EP_HIP(ep::launch_fill_bf16(c->d_ostate, (uint16_t*)c->d_table, n, D, D, c->stream));
    EP_HIP(hipMemsetAsync(c->d_ostate, 0, (size_t)c->n_vocab * ostride * sizeof(float), c->stream));
  } else {
    EP_HIP(hipMemsetAsync(c->d_ostate, 0, (size_t)c->n_vocab * ostride * sizeof(float), c->stream));
    if (rows) EP_HIP(hipMemcpyAsync(c->d_table, rows, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    else EP_HIP(ep::launch_init_table(c->base, c->cfg.seed, c->stream));
  }
  EP_HIP(hipStreamSynchronize(c->stream));
  c->table_ready = true;
  return EP_OK;
}

int ep_get_table(ep_ctx* c, float* rows) {
  if (!c || !rows) return fail(EP_EINVAL, "NULL argument");
  if (!c->table_ready) return fail(EP_ESTATE, "no table");
  EP_HIP(hipSetDevice(c->device));
  const size_t n = (size_t)c->n_vocab * row_width(c->cfg);
  const uint32_t D = row_width(c->cfg);
  EP_HIP(hipStreamSynchronize(c->stream));
  if (c->cfg.precision == EP_PREC_BF16_MASTER) {
    EP_HIP(hipMemcpy2D(rows, D * sizeof(float), c->d_ostate, 3 * D * sizeof(float), D * sizeof(float), c->n_vocab,
                       hipMemcpyDeviceToHost));
  } else if (c->cfg.precision == EP_PREC_BF16) {
    // the bf16 table is the parameter: widen it (exact)
    std::vector<uint16_t> h(n);
    EP_HIP(hipMemcpy(h.data(), c->d_table, n * sizeof(uint16_t), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) {
      const uint32_t u = (uint32_t)h[i] << 16;
      std::memcpy(rows + i, &u, sizeof(float));
    }
  } else {
    EP_HIP(hipMemcpy(rows, c->d_table, n * sizeof(float), hipMemcpyDeviceToHost));
  }
  return EP_OK;
}

int ep_get_table_bf16(ep_ctx* c, uint16_t* rows) {
  if (!c || !rows) return fail(EP_EINVAL, "NULL argument");
  if (!c->table_ready || c->cfg.precision == EP_PREC_FP32) return fail(EP_ESTATE, "no bf16 table");
  EP_HIP(hipSetDevice(c->device));
  EP_HIP(hipStreamSynchronize(c->stream));
  EP_HIP(hipMemcpy(rows, c->d_table, (size_t)c->n_vocab * row_width(c->cfg) * sizeof(uint16_t), hipMemcpyDeviceToHost));
  return EP_OK;
}

int ep_get_adam(ep_ctx* c, float* m, float* v) {
  if (!c || !m || !v) return fail(EP_EINVAL, "NULL argument");
  if (!c->table_ready) return fail(EP_ESTATE, "no table");
  EP_HIP(hipSetDevice(c->device));
  const uint32_t D = row_width(c->cfg);
  const bool ms = c->cfg.precision == EP_PREC_BF16_MASTER;
  const size_t ostride = (ms ? 3 : 2) * (size_t)D;
  EP_HIP(hipStreamSynchronize(c->stream));
  EP_HIP(hipMemcpy2D(m, D * sizeof(float), c->d_ostate + (ms ? D : 0), ostride * sizeof(float), D * sizeof(float),
                     c->n_vocab, hipMemcpyDeviceToHost));
  EP_HIP(hipMemcpy2D(v, D * sizeof(float), c->d_ostate + (ms ? 2 * D : D), ostride * sizeof(float),
                     D * sizeof(float), c->n_vocab, hipMemcpyDeviceToHost));
  return EP_OK;
}

int ep_set_split(ep_ctx* c, const uint32_t* train, uint64_t n_train, const uint32_t* valid, uint64_t n_valid) {
  if (!c || (!train && n_train) || (!valid && n_valid)) return fail(EP_EINVAL, "NULL argument");
  if (!c->n_nodes) return fail(EP_ESTATE, "graph must be set first");
  for (uint64_t i = 0; i < n_train; ++i) if (train[i] >= c->n_nodes) return fail(EP_EINVAL, "train id out of range");
  for (uint64_t i = 0; i < n_valid; ++i) if (valid[i] >= c->n_nodes) return fail(EP_EINVAL, "valid id out of range");
  if (n_train == 0) return fail(EP_EINVAL, "empty train set");
  c->train.assign(train, train + n_train);
  c->valid.assign(valid, valid + n_valid);
  c->split_set = true;
  EP_HIP(hipSetDevice(c->device));
  return upload_split(c);
}

int ep_learn(ep_ctx* c, ep_progress_fn progress, void* user) {
  if (!c) return fail(EP_EINVAL, "NULL ctx");
  if (!c->d_indptr || !c->d_findptr) return fail(EP_ESTATE, "graph and features must be set first");
  if (c->f_nodes != c->n_nodes) return fail(EP_ESTATE, "features must cover every graph node");
  EP_HIP(hipSetDevice(c->device));
  if (!c->table_ready) EP_TRY(ep_set_table(c, nullptr));
  if (!c->split_set) {
    EP_TRY(ensure_split(c));
    EP_TRY(upload_split(c));
  } else if (!c->d_train) {
    EP_TRY(upload_split(c));
  }
  EP_TRY(ensure_step_buffers(c, group_anchors(c)));
  refresh_base(c);
  ep::StepConfig sc = train_config(c);
  const uint32_t B = c->cfg.batch_size;
  const uint32_t spp = sc.steps_per_pass;
  const uint32_t nbatch = (uint32_t)((c->train.size() + B - 1) / B);  // reference batches per pass
  HostRng rng(c->cfg.seed ^ 0x5851F42D4C957F2Dull);
  uint64_t next_step = 1;
  std::vector<float> losses(c->train.size());
  std::vector<float> vloss(c->valid.size());
  // validation config (mod.rs:267-283): forward only, negatives from the valid set, fixed seed - 1
  ep::StepConfig vc = c->base;
  vc.perm = c->d_valid;
  vc.negpool = c->d_valid;
  vc.n_negpool = (uint32_t)c->valid.size();
  vc.n_batch_total = (uint32_t)c->valid.size();
  vc.train = 0;
  vc.lw_kind = 0;  // the validation sums each anchor's unweighted forward loss (mod.rs:272-278)
  vc.seed = c->cfg.seed - 1ull;
  vc.pass_loss = c->d_valid_loss;
  vc.steps_per_pass = (uint32_t)((c->valid.size() + vc.batch - 1) / vc.batch);
  vc.total_updates = 1;
  for (uint32_t pass = 1; pass <= c->cfg.passes; ++pass) {
    EP_TRY(upload_pass_perm(c, rng));
    EP_TRY(set_pass(c, c->d_state, pass, pass == 1 ? &next_step : nullptr));
    for (uint32_t s = 0; s < spp; ++s) EP_TRY(run_step(c, sc));
    EP_TRY(check_state_err(c));
    EP_HIP(hipMemcpy(losses.data(), c->d_pass_loss, losses.size() * sizeof(float), hipMemcpyDeviceToHost));
    // per-batch error = sum of contributing losses / batch size; pass error = mean over batches
    double err = 0.0;
    for (uint32_t s = 0; s < nbatch; ++s) {
      const size_t lo = (size_t)s * B, hi = std::min(losses.size(), lo + B);
      float e = 0.f;
      for (size_t i = lo; i < hi; ++i) e += losses[i];
      float be = e / (float)(hi - lo);
      if (std::isinf(be)) be = 0.f;
      err += be;
    }
    const float train_err = nbatch ? (float)(err / nbatch) : 0.f;
    float valid_err = NAN;
    if (!c->valid.empty()) {
      EP_TRY(set_pass(c, c->d_vstate, 0, nullptr));
      for (uint32_t s = 0; s < vc.steps_per_pass; ++s) {
        EP_HIP(ep::launch_step_begin_lr(vc, c->d_vstate, 0.f, c->stream));
        EP_HIP(ep::launch_sample(vc, c->d_vstate, c->stream));
        if (vc.model == 1) {
          EP_HIP(ep::launch_scan(vc, c->d_vstate, c->stream));
          EP_HIP(ep::launch_att_forward(vc, c->d_vstate, c->stream));
        } else {
          EP_HIP(ep::launch_forward(vc, c->d_vstate, c->stream));
        }
      }
      EP_HIP(hipStreamSynchronize(c->stream));
      EP_HIP(hipMemcpy(vloss.data(), c->d_valid_loss, vloss.size() * sizeof(float), hipMemcpyDeviceToHost));
      double vs = 0.0;
      for (float x : vloss) vs += x;
      valid_err = (float)(vs / vloss.size());
    }
    if (progress) {
      ep::StepState hs;
      EP_HIP(hipMemcpy(&hs, c->d_state, sizeof(hs), hipMemcpyDeviceToHost));
      progress(user, pass, hs.next_step, train_err, valid_err, hs.lr[hs.n_batches ? hs.n_batches - 1 : 0]);
    }
  }
  EP_HIP(hipStreamSynchronize(c->stream));
  return EP_OK;
}

int ep_sample_batch(ep_ctx* c, const uint32_t* anchors, uint32_t n, uint32_t pass, uint32_t* nbr_cnt, uint32_t* nbr,
                    uint32_t* neg) {
  if (!c || !anchors || !nbr_cnt || !nbr || !neg || n == 0) return fail(EP_EINVAL, "bad argument");
  if (c->cfg.max_nodes_kind == EP_SAMPLE_ALL && !ppr_loss(c->cfg))
    return fail(EP_EINVAL, "ep_sample_batch needs max_nodes Fixed or Probability");
  for (uint32_t i = 0; i < n; ++i) if (anchors[i] >= c->n_nodes) return fail(EP_EINVAL, "anchor out of range");
  if (prob_nodes(c->cfg)) {
    std::vector<uint32_t> a(anchors, anchors + n);
    std::sort(a.begin(), a.end());
    if (std::adjacent_find(a.begin(), a.end()) != a.end())
      return fail(EP_EINVAL, "max_nodes Probability: the anchors of a batch must be distinct");
  }
  EP_HIP(hipSetDevice(c->device));
  if (!c->split_set) { EP_TRY(ensure_split(c)); EP_TRY(upload_split(c)); }
  else if (!c->d_train) EP_TRY(upload_split(c));
  EP_TRY(ensure_step_buffers(c, std::max(n, group_anchors(c))));
  refresh_base(c);
  ep::StepConfig sc = c->base;
  sc.perm = c->d_explicit;
  sc.negpool = c->d_train;
  sc.n_negpool = (uint32_t)c->train.size();
  sc.n_batch_total = n;
  sc.batch = n;
  sc.rbatch = n;
  EP_HIP(hipMemcpy(c->d_explicit, anchors, n * sizeof(uint32_t), hipMemcpyHostToDevice));
  EP_TRY(set_pass(c, c->d_vstate, pass, nullptr));
  EP_HIP(ep::launch_step_begin_lr(sc, c->d_vstate, 0.f, c->stream));
  EP_HIP(ep::launch_sample(sc, c->d_vstate, c->stream));
  EP_HIP(hipStreamSynchronize(c->stream));
  EP_HIP(hipMemcpy(nbr_cnt, sc.nbr_cnt, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (sc.pnbrs) {
    // lists at the anchors' CSR offsets, returned n x deg_max
    const uint32_t L = hook_list_width(c);
    std::fill(nbr, nbr + (size_t)n * L, 0u);
    for (uint32_t i = 0; i < n; ++i)
      if (nbr_cnt[i])
        EP_HIP(hipMemcpy(nbr + (size_t)i * L, sc.nbr_sel + c->h_indptr[anchors[i]], nbr_cnt[i] * sizeof(uint32_t),
                         hipMemcpyDeviceToHost));
  } else {
    EP_HIP(hipMemcpy(nbr, sc.nbr, (size_t)n * sc.S * sizeof(uint32_t), hipMemcpyDeviceToHost));
  }
  EP_HIP(hipMemcpy(neg, sc.neg, (size_t)n * sc.K * sizeof(uint32_t), hipMemcpyDeviceToHost));
  c->last_anchors.assign(anchors, anchors + n);
  c->last_cnt.assign(nbr_cnt, nbr_cnt + n);
  return EP_OK;
}

int ep_sample_weights(ep_ctx* c, uint32_t n, float* w) {
  if (!c || !w || n == 0) return fail(EP_EINVAL, "bad argument");
  if (c->base.pnbrs) {
    if (!c->base.nbr_selw) return fail(EP_ESTATE, "no neighbour weights: needs weighted averaging or max_features");
    if (n != c->last_anchors.size()) return fail(EP_EINVAL, "n must be the last sample_batch's size");
    const uint32_t L = hook_list_width(c);
    EP_HIP(hipSetDevice(c->device));
    std::fill(w, w + (size_t)n * L, 0.f);
    for (uint32_t i = 0; i < n; ++i)
      if (c->last_cnt[i])
        EP_HIP(hipMemcpy(w + (size_t)i * L, c->base.nbr_selw + c->h_indptr[c->last_anchors[i]],
                         c->last_cnt[i] * sizeof(float), hipMemcpyDeviceToHost));
    return EP_OK;
  }
  if (!c->base.nbr_w) return fail(EP_ESTATE, "no neighbour weights: weighted_neighbor_averaging with max_nodes Fixed");
  if (n > c->cap_batch) return fail(EP_EINVAL, "n exceeds the sampled batch");
  EP_HIP(hipSetDevice(c->device));
  EP_HIP(hipMemcpy(w, c->base.nbr_w, (size_t)n * c->base.S * sizeof(float), hipMemcpyDeviceToHost));
  return EP_OK;
}

int ep_set_explicit_weights(ep_ctx* c, uint32_t n, const float* w) {
  if (!c || (!w && n)) return fail(EP_EINVAL, "bad argument");
  c->explicit_w.assign(w, w + (size_t)n * hook_list_width(c));
  return EP_OK;
}

int ep_step_explicit(ep_ctx* c, const uint32_t* anchors, uint32_t n, const uint32_t* nbr_cnt, const uint32_t* nbr,
                     const uint32_t* neg, uint32_t pass, float lr, int apply_update, float* loss, float* cons,
                     uint32_t* n_unique, uint32_t* feat, float* grad) {
  if (!c || !anchors || !nbr_cnt || !nbr || !neg || n == 0) return fail(EP_EINVAL, "bad argument");
  if (c->cfg.max_nodes_kind == EP_SAMPLE_ALL && !ppr_loss(c->cfg))
    return fail(EP_EINVAL, "ep_step_explicit needs max_nodes Fixed or Probability");
  if (!(lr >= 0.f)) return fail(EP_EINVAL, "lr must be >= 0");
  if (!c->d_indptr || !c->d_findptr) return fail(EP_ESTATE, "graph and features must be set first");
  const bool pn = prob_nodes(c->cfg);
  const uint32_t S = hook_list_width(c), K = c->cfg.negatives, D = row_width(c->cfg), NC = K + 2;
  if (pn) {
    std::vector<uint32_t> a(anchors, anchors + n);
    std::sort(a.begin(), a.end());
    if (std::adjacent_find(a.begin(), a.end()) != a.end())
      return fail(EP_EINVAL, "max_nodes Probability: the anchors of a batch must be distinct");
  }
  for (uint32_t i = 0; i < n; ++i) {
    if (anchors[i] >= c->n_nodes) return fail(EP_EINVAL, "anchor out of range");
    if (nbr_cnt[i] > S) return fail(EP_EINVAL, "nbr_cnt exceeds max_nodes");
    if (pn && nbr_cnt[i] > c->degrees[anchors[i]]) return fail(EP_EINVAL, "nbr_cnt exceeds the anchor's degree");
    for (uint32_t j = 0; j < nbr_cnt[i]; ++j)
      if (nbr[(size_t)i * S + j] >= c->n_nodes) return fail(EP_EINVAL, "neighbour out of range");
    for (uint32_t k = 0; k < K; ++k)
      if (neg[(size_t)i * K + k] >= c->n_nodes) return fail(EP_EINVAL, "negative out of range");
  }
  EP_HIP(hipSetDevice(c->device));
  if (!c->table_ready) EP_TRY(ep_set_table(c, nullptr));
  // concurrent_batches == 1: the n anchors are one batch; else batches of batch_size,
  // at most concurrent_batches of them
  const uint32_t P = std::max<uint32_t>(1u, c->cfg.concurrent_batches);
  if (P > 1 && (n + c->cfg.batch_size - 1) / c->cfg.batch_size > P)
    return fail(EP_EINVAL, "ep_step_explicit: more than concurrent_batches batches of batch_size anchors");
  EP_TRY(ensure_step_buffers(c, std::max(n, group_anchors(c))));
  refresh_base(c);
  if (cons && !c->d_xout) EP_TRY(c->alloc(&c->d_xout, (size_t)c->cap_batch * NC * D));
  if (!c->base.G) {
    EP_TRY(c->alloc(&c->base.G, (size_t)c->unique_cap * D));
    EP_TRY(c->alloc(&c->base.ufeat, c->unique_cap));
  }
  ep::StepConfig sc = c->base;
  sc.perm = c->d_explicit;
  sc.negpool = c->d_explicit;
  sc.n_negpool = n;
  sc.n_batch_total = n;
  sc.batch = n;
  sc.rbatch = P > 1 ? c->cfg.batch_size : n;
  sc.explicit_samples = 1;
  sc.xout = cons ? c->d_xout : nullptr;
  EP_HIP(hipMemcpy(c->d_explicit, anchors, n * sizeof(uint32_t), hipMemcpyHostToDevice));
  EP_HIP(hipMemcpy(sc.nbr_cnt, nbr_cnt, n * sizeof(uint32_t), hipMemcpyHostToDevice));
  if (pn) {
    // the lists (and weights) go where the sampler puts them: at the anchors' CSR offsets
    if (sc.nbr_rows && c->explicit_w.size() != (size_t)n * S)
      return fail(EP_ESTATE, "max_nodes Probability: set n x deg_max neighbour weights first");
    for (uint32_t i = 0; i < n; ++i) {
      if (!nbr_cnt[i]) continue;
      const uint64_t off = c->h_indptr[anchors[i]];
      EP_HIP(hipMemcpy(sc.nbr_sel + off, nbr + (size_t)i * S, nbr_cnt[i] * sizeof(uint32_t), hipMemcpyHostToDevice));
      if (sc.nbr_rows)
        EP_HIP(hipMemcpy(sc.nbr_selw + off, c->explicit_w.data() + (size_t)i * S, nbr_cnt[i] * sizeof(float),
                         hipMemcpyHostToDevice));
    }
  } else {
    EP_HIP(hipMemcpy(sc.nbr, nbr, (size_t)n * S * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  EP_HIP(hipMemcpy(sc.neg, neg, (size_t)n * K * sizeof(uint32_t), hipMemcpyHostToDevice));
  if (sc.wavg && !pn) {
    if (c->explicit_w.size() != (size_t)n * S)
      return fail(EP_ESTATE, "weighted_neighbor_averaging: set n x max_nodes neighbour weights first");
    EP_HIP(hipMemcpy(sc.nbr_w, c->explicit_w.data(), (size_t)n * S * sizeof(float), hipMemcpyHostToDevice));
  }
  EP_TRY(set_pass(c, c->d_state, pass, nullptr));
  const bool upd = apply_update != 0;
  // run everything up to (but excluding) the update, read the merged gradient, then update
  int buf = 0;
  EP_TRY(launch_training_step(c, sc, false, lr, nullptr, &buf));
  EP_TRY(check_state_err(c));
  ep::StepState hs;
  EP_HIP(hipMemcpy(&hs, c->d_state, offsetof(ep::StepState, dtot), hipMemcpyDeviceToHost));
  if (loss) EP_HIP(hipMemcpy(loss, sc.loss, n * sizeof(float), hipMemcpyDeviceToHost));
  if (cons)
    EP_HIP(hipMemcpy(cons, c->d_xout, (size_t)n * NC * c->cfg.dims * sizeof(float), hipMemcpyDeviceToHost));
  if (sc.scatter == EP_SCATTER_SORT) {
    if (n_unique) *n_unique = hs.n_unique;
    if (feat) EP_HIP(hipMemcpy(feat, sc.ufeat, hs.n_unique * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (grad) EP_HIP(hipMemcpy(grad, sc.G, (size_t)hs.n_unique * D * sizeof(float), hipMemcpyDeviceToHost));
    if (upd) {
      // the update pass re-reads the sorted keys; its long-segment lists start empty
      const uint32_t zero4[5] = {0u, 0u, 0u, 0u, 0u};  // n_unique, n_long, n_pieces, n_touched, n_short
      static_assert(offsetof(ep::StepState, n_short) == offsetof(ep::StepState, n_unique) + 4 * sizeof(uint32_t),
                    "counter layout");
      EP_HIP(hipMemcpy((char*)c->d_state + offsetof(ep::StepState, n_unique), zero4, sizeof(zero4),
                       hipMemcpyHostToDevice));
      pick_variant(c, sc);
      EP_HIP(ep::launch_scatter_sorted(sc, c->d_state, buf, c->stream));
    }
  } else {
    if (n_unique) *n_unique = hs.n_touched;
    std::vector<uint32_t> tf(hs.n_touched);
    if (hs.n_touched)
      EP_HIP(hipMemcpy(tf.data(), sc.touched, hs.n_touched * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (feat && hs.n_touched) std::memcpy(feat, tf.data(), hs.n_touched * sizeof(uint32_t));
    if (grad)
      for (uint32_t i = 0; i < hs.n_touched; ++i)
        EP_HIP(hipMemcpy(grad + (size_t)i * D, sc.Gd + (size_t)tf[i] * D, D * sizeof(float), hipMemcpyDeviceToHost));
    if (upd) {
      EP_HIP(ep::launch_adam_touched(sc, c->d_state, c->stream));
    } else {
      for (uint32_t i = 0; i < hs.n_touched; ++i) {
        EP_HIP(hipMemsetAsync(sc.Gd + (size_t)tf[i] * D, 0, D * sizeof(float), c->stream));
        EP_HIP(hipMemsetAsync(sc.flag + tf[i], 0, sizeof(uint32_t), c->stream));
      }
    }
  }
  EP_HIP(hipStreamSynchronize(c->stream));
  return EP_OK;
}

int ep_step_explicit_noalloc(ep_ctx* c, const uint32_t* anchors, uint32_t n, const uint32_t* nbr_cnt,
                             const uint32_t* nbr, const uint32_t* neg, uint32_t pass, float lr, float* loss) {
  if (!c || !anchors || !nbr_cnt || !nbr || !neg || n == 0) return fail(EP_EINVAL, "bad argument");
  if (c->cfg.max_nodes_kind != EP_SAMPLE_FIXED && !ppr_loss(c->cfg))
    return fail(EP_EINVAL, "ep_step_explicit_noalloc needs max_nodes Fixed (or the ppr loss)");
  if (c->cfg.scatter != EP_SCATTER_SORT) return fail(EP_EINVAL, "ep_step_explicit_noalloc needs the sort scatter");
  if (!(lr >= 0.f)) return fail(EP_EINVAL, "lr must be >= 0");
  if (!c->d_indptr || !c->d_findptr) return fail(EP_ESTATE, "graph and features must be set first");
  if (!c->table_ready) return fail(EP_ESTATE, "the table must be set first (ep_set_table allocates it)");
  if (weighted_avg(c->cfg) && !c->d_cdf)
    return fail(EP_ESTATE, "weighted_neighbor_averaging needs the edge weights (ep_set_edge_weights)");
  const uint32_t S = list_cap(c->cfg), K = c->cfg.negatives;
  for (uint32_t i = 0; i < n; ++i) {
    if (anchors[i] >= c->n_nodes) return fail(EP_EINVAL, "anchor out of range");
    if (nbr_cnt[i] > S) return fail(EP_EINVAL, "nbr_cnt exceeds max_nodes");
    for (uint32_t j = 0; j < nbr_cnt[i]; ++j)
      if (nbr[(size_t)i * S + j] >= c->n_nodes) return fail(EP_EINVAL, "neighbour out of range");
    for (uint32_t k = 0; k < K; ++k)
      if (neg[(size_t)i * K + k] >= c->n_nodes) return fail(EP_EINVAL, "negative out of range");
  }
  const uint32_t P = std::max<uint32_t>(1u, c->cfg.concurrent_batches);
  if (P > 1 && (n + c->cfg.batch_size - 1) / c->cfg.batch_size > P)
    return fail(EP_EINVAL, "ep_step_explicit_noalloc: more than concurrent_batches batches of batch_size anchors");
  const bool wavg = weighted_avg(c->cfg);
  if (wavg && c->explicit_w.size() != (size_t)n * S)
    return fail(EP_ESTATE, "weighted_neighbor_averaging: set n x max_nodes neighbour weights first");
  EP_HIP(hipSetDevice(c->device));
  refresh_base(c);
  ep::StepConfig sc = c->base;
  clear_step_buffers(sc);
  // carve the step's arrays out of this device's scratch, 256-byte aligned
  size_t need = 0;
  auto carve = [&](auto** p, size_t count) -> int {
    const size_t at = (need + 255) & ~(size_t)255;
    need = at + std::max<size_t>(count * sizeof(**p), 16);
    *p = need <= kScratchBytes ? reinterpret_cast<std::remove_reference_t<decltype(*p)>>(c->scratch + at) : nullptr;
    return EP_OK;
  };
  uint32_t* d_anchors = nullptr;
  carve(&d_anchors, n);
  uint64_t keycap = 0;
  EP_TRY(place_step_buffers(c, n, sc, carve, &keycap));
  if (need > kScratchBytes)
    return fail(EP_ECAPACITY, "ep_step_explicit_noalloc: the step needs " + std::to_string(need) +
                                  " bytes of scratch, the library holds " + std::to_string(kScratchBytes));
  sc.perm = d_anchors;
  sc.negpool = d_anchors;
  sc.n_negpool = n;
  sc.n_batch_total = n;
  sc.batch = n;
  sc.rbatch = P > 1 ? c->cfg.batch_size : n;
  sc.explicit_samples = 1;
  EP_HIP(hipMemcpy(d_anchors, anchors, n * sizeof(uint32_t), hipMemcpyHostToDevice));
  EP_HIP(hipMemcpy(sc.nbr_cnt, nbr_cnt, n * sizeof(uint32_t), hipMemcpyHostToDevice));
  if (S) EP_HIP(hipMemcpy(sc.nbr, nbr, (size_t)n * S * sizeof(uint32_t), hipMemcpyHostToDevice));
  EP_HIP(hipMemcpy(sc.neg, neg, (size_t)n * K * sizeof(uint32_t), hipMemcpyHostToDevice));
  if (wavg) EP_HIP(hipMemcpy(sc.nbr_w, c->explicit_w.data(), (size_t)n * S * sizeof(float), hipMemcpyHostToDevice));
  EP_TRY(set_pass(c, c->d_state, pass, nullptr));
  EP_TRY(launch_training_step(c, sc, true, lr));
  EP_TRY(check_state_err(c));
  if (loss) EP_HIP(hipMemcpy(loss, sc.loss, n * sizeof(float), hipMemcpyDeviceToHost));
  EP_HIP(hipStreamSynchronize(c->stream));
  return EP_OK;
}

int ep_bench_steps(ep_ctx* c, uint32_t warmup, uint32_t steps, int instrument, ep_bench_result* out) {
  if (!c || !out || steps == 0) return fail(EP_EINVAL, "bad argument");
  std::memset(out, 0, sizeof(*out));
  EP_HIP(hipSetDevice(c->device));
  if (!c->table_ready) EP_TRY(ep_set_table(c, nullptr));
  if (!c->split_set) { EP_TRY(ensure_split(c)); EP_TRY(upload_split(c)); }
  else if (!c->d_train) EP_TRY(upload_split(c));
  EP_TRY(ensure_step_buffers(c, group_anchors(c)));
  refresh_base(c);
  ep::StepConfig sc = train_config(c);
  sc.pass_loss = c->d_pass_loss;
  HostRng rng(c->cfg.seed ^ 0x2545F4914F6CDD1Dull);
  uint32_t pass = 1;
  uint32_t in_pass = 0;
  uint64_t next_step = 1;
  EP_TRY(upload_pass_perm(c, rng));
  EP_TRY(set_pass(c, c->d_state, pass, &next_step));
  auto advance = [&]() -> int {
    if (in_pass == sc.steps_per_pass) {
      ++pass;
      in_pass = 0;
      EP_TRY(upload_pass_perm(c, rng));
      EP_TRY(set_pass(c, c->d_state, pass, nullptr));
    }
    ++in_pass;
    return EP_OK;
  };
  // warm-up steps; extended while the scatter-variant tuner is still choosing
  for (uint32_t i = 0; i < warmup || tuning(c); ++i) {
    EP_TRY(advance());
    EP_TRY(run_step(c, sc));
  }
  EP_HIP(hipStreamSynchronize(c->stream));
  EP_TRY(check_state_err(c));
  ep::StepState h0, h1;
  EP_HIP(hipMemcpy(&h0, c->d_state, sizeof(h0), hipMemcpyDeviceToHost));
  hipEvent_t e0, e1;
  EP_HIP(hipEventCreate(&e0));
  EP_HIP(hipEventCreate(&e1));
  // timed steps: no host work except at pass boundaries
  EP_HIP(hipEventRecord(e0, c->stream));
  for (uint32_t i = 0; i < steps; ++i) {
    if (in_pass == sc.steps_per_pass) {
      // a pass boundary inside the timed region: reshuffle (counted in the time)
      EP_TRY(advance());
    } else {
      ++in_pass;
    }
    EP_TRY(run_step(c, sc));
  }
  EP_HIP(hipEventRecord(e1, c->stream));
  EP_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  EP_HIP(hipEventElapsedTime(&ms, e0, e1));
  EP_TRY(check_state_err(c));
  EP_HIP(hipMemcpy(&h1, c->d_state, sizeof(h1), hipMemcpyDeviceToHost));
  out->ms_total = ms;
  out->steps = steps;
  out->examples = h1.stat_examples - h0.stat_examples;
  out->rows = h1.stat_rows - h0.stat_rows;
  out->nodes = h1.stat_nodes - h0.stat_nodes;
  out->unique_rows = h1.stat_unique - h0.stat_unique;
  out->att_items_sq = h1.stat_att_n2 - h0.stat_att_n2;
  out->att_items = h1.stat_att_n - h0.stat_att_n;
  out->seg_vpl = c->seg_vpl;
  out->seg_unroll = c->seg_unroll;
  out->fwd_x = c->fwd_x ? c->fwd_x : 1;
  if (instrument) {
    // per-phase timing: the same number of steps launched one by one with events
    std::vector<hipEvent_t> ev(9 * (size_t)steps);
    for (auto& e : ev) EP_HIP(hipEventCreate(&e));
    ep::StepState g0, g1;
    EP_HIP(hipMemcpy(&g0, c->d_state, sizeof(g0), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < steps; ++i) {
      if (in_pass == sc.steps_per_pass) EP_TRY(advance());
      else ++in_pass;
      EP_TRY(launch_training_step(c, sc, true, -1.f, &ev[9 * (size_t)i]));
    }
    EP_HIP(hipStreamSynchronize(c->stream));
    EP_HIP(hipMemcpy(&g1, c->d_state, sizeof(g1), hipMemcpyDeviceToHost));
    double ph[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t i = 0; i < steps; ++i)
      for (int k = 0; k < 8; ++k) {
        float x = 0.f;
        EP_HIP(hipEventElapsedTime(&x, ev[9 * (size_t)i + k], ev[9 * (size_t)i + k + 1]));
        ph[k] += x;
      }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (int k = 0; k < 8; ++k) out->ms_phase[k] = ph[k] / steps;
    out->ms_forward_avg = out->ms_phase[1];
    out->ms_scatter_avg = out->ms_phase[3] + out->ms_phase[4] + out->ms_phase[5] + out->ms_phase[6] + out->ms_phase[7];
    out->short_rows = g1.stat_short - g0.stat_short;
    out->last_long_rows = g1.n_long;
    out->last_pieces = g1.n_pieces;
    out->forward_launches = steps;
    EP_TRY(check_state_err(c));
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return EP_OK;
}

int ep_embed_averaged(ep_ctx* c, uint64_t n_nodes, const uint64_t* f_indptr, const uint32_t* f_ids, float* out) {
  if (!c || !f_indptr || !out) return fail(EP_EINVAL, "NULL argument");
  if (!c->table_ready) return fail(EP_ESTATE, "no table");
  const uint64_t tot = f_indptr[n_nodes];
  for (uint64_t i = 0; i < n_nodes; ++i)
    if (f_indptr[i + 1] < f_indptr[i]) return fail(EP_EINVAL, "f_indptr must be non-decreasing");
  for (uint64_t i = 0; i < tot; ++i)
    if (f_ids[i] >= c->n_vocab) return fail(EP_EINVAL, "feature id out of range");
  EP_HIP(hipSetDevice(c->device));
  uint64_t* dp = nullptr;
  uint32_t* di = nullptr;
  float* dout = nullptr;
  EP_TRY(c->alloc(&dp, n_nodes + 1));
  EP_TRY(c->alloc(&di, tot));
  EP_TRY(c->alloc(&dout, (size_t)n_nodes * row_width(c->cfg)));
  EP_HIP(hipMemcpy(dp, f_indptr, (n_nodes + 1) * sizeof(uint64_t), hipMemcpyHostToDevice));
  if (tot) EP_HIP(hipMemcpy(di, f_ids, tot * sizeof(uint32_t), hipMemcpyHostToDevice));
  refresh_base(c);
  EP_HIP(ep::launch_embed_averaged(c->base, dp, di, n_nodes, dout, c->stream));
  EP_HIP(hipStreamSynchronize(c->stream));
  EP_HIP(hipMemcpy(out, dout, (size_t)n_nodes * row_width(c->cfg) * sizeof(float), hipMemcpyDeviceToHost));
  c->release(dp);
  c->release(di);
  c->release(dout);
  return EP_OK;
}

int ep_embed_weighted(ep_ctx* c, uint64_t n_nodes, const uint64_t* f_indptr, const uint32_t* f_ids,
                      const float* p_w, uint32_t n_p, float alpha, float* out) {
  if (!c || !f_indptr || !out || !p_w) return fail(EP_EINVAL, "NULL argument");
  if (!c->table_ready) return fail(EP_ESTATE, "no table");
  if (!(alpha > 0.f) || !std::isfinite(alpha)) return fail(EP_EINVAL, "alpha must be finite and > 0");
  const uint64_t tot = f_indptr[n_nodes];
  for (uint64_t i = 0; i < n_nodes; ++i)
    if (f_indptr[i + 1] < f_indptr[i]) return fail(EP_EINVAL, "f_indptr must be non-decreasing");
  for (uint64_t i = 0; i < tot; ++i)
    if (f_ids[i] >= c->n_vocab || f_ids[i] >= n_p) return fail(EP_EINVAL, "feature id out of range");
  for (uint32_t i = 0; i < n_p; ++i)
    if (!(p_w[i] >= 0.f)) return fail(EP_EINVAL, "unigram probabilities must be >= 0");
  EP_HIP(hipSetDevice(c->device));
  const size_t W = row_width(c->cfg);
  uint64_t* dp = nullptr;
  uint32_t* di = nullptr;
  float* dw = nullptr;
  float* dout = nullptr;
  EP_TRY(c->alloc(&dp, n_nodes + 1));
  EP_TRY(c->alloc(&di, tot));
  EP_TRY(c->alloc(&dw, n_p));
  EP_TRY(c->alloc(&dout, (size_t)n_nodes * W));
  EP_HIP(hipMemcpy(dp, f_indptr, (n_nodes + 1) * sizeof(uint64_t), hipMemcpyHostToDevice));
  if (tot) EP_HIP(hipMemcpy(di, f_ids, tot * sizeof(uint32_t), hipMemcpyHostToDevice));
  EP_HIP(hipMemcpy(dw, p_w, (size_t)n_p * sizeof(float), hipMemcpyHostToDevice));
  refresh_base(c);
  EP_HIP(ep::launch_embed_weighted(c->base, dp, di, dw, alpha, n_nodes, dout, c->stream));
  EP_HIP(hipStreamSynchronize(c->stream));
  EP_HIP(hipMemcpy(out, dout, (size_t)n_nodes * W * sizeof(float), hipMemcpyDeviceToHost));
  c->release(dp);
  c->release(di);
  c->release(dw);
  c->release(dout);
  return EP_OK;
}

int ep_embed_attention(ep_ctx* c, uint64_t n_nodes, const uint64_t* f_indptr, const uint32_t* f_ids, float* out) {
  if (!c || !f_indptr || !out) return fail(EP_EINVAL, "NULL argument");
  if (!c->table_ready) return fail(EP_ESTATE, "no table");
  if (c->cfg.attention_dk == 0) return fail(EP_ESTATE, "the context is not configured for attention");
  const uint64_t tot = f_indptr[n_nodes];
  for (uint64_t i = 0; i < n_nodes; ++i) {
    if (f_indptr[i + 1] < f_indptr[i]) return fail(EP_EINVAL, "f_indptr must be non-decreasing");
    if (f_indptr[i + 1] - f_indptr[i] > 64) return fail(EP_EINVAL, "attention supports at most 64 features per node");
  }
  for (uint64_t i = 0; i < tot; ++i)
    if (f_ids[i] >= c->n_vocab) return fail(EP_EINVAL, "feature id out of range");
  EP_HIP(hipSetDevice(c->device));
  const size_t dm = c->cfg.dims;
  uint64_t* dp = nullptr;
  uint32_t* di = nullptr;
  float* dout = nullptr;
  EP_TRY(c->alloc(&dp, n_nodes + 1));
  EP_TRY(c->alloc(&di, tot));
  EP_TRY(c->alloc(&dout, (size_t)n_nodes * dm));
  EP_HIP(hipMemcpy(dp, f_indptr, (n_nodes + 1) * sizeof(uint64_t), hipMemcpyHostToDevice));
  if (tot) EP_HIP(hipMemcpy(di, f_ids, tot * sizeof(uint32_t), hipMemcpyHostToDevice));
  refresh_base(c);
  EP_HIP(ep::launch_att_embed(c->base, dp, di, n_nodes, dout, c->stream));
  EP_HIP(hipStreamSynchronize(c->stream));
  EP_HIP(hipMemcpy(out, dout, (size_t)n_nodes * dm * sizeof(float), hipMemcpyDeviceToHost));
  c->release(dp);
  c->release(di);
  c->release(dout);
  return EP_OK;
}

// NeighborhoodAligner (alignment.rs, lib.rs:3040-3171): the graph (with its edge CDF) and
// the node embeddings stay on the device; runs align node ranges.
struct ep_aligner {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t n = 0;
  uint32_t dims = 0;
  uint64_t* indptr = nullptr;
  uint32_t* indices = nullptr;
  float* cdf = nullptr;
  float* emb = nullptr;
  float* out = nullptr;
  uint64_t out_rows = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};
  float last_ms = 0.f;  // duration of the last run's kernel (HIP events)
};

static void aligner_free(ep_aligner* a) {
  if (!a) return;
  if (a->stream) (void)hipSetDevice(a->device);
  void* ps[] = {a->indptr, a->indices, a->cdf, a->emb, a->out};
  for (void* q : ps)
    if (q) (void)hipFree(q);
  for (hipEvent_t e : a->ev)
    if (e) (void)hipEventDestroy(e);
  if (a->stream) (void)hipStreamDestroy(a->stream);
  delete a;
}

int ep_aligner_create(int device, uint64_t n_nodes, const uint64_t* indptr, const uint32_t* indices, const float* cdf,
                      uint32_t dims, const float* emb, ep_aligner** out) {
  if (!out || !indptr || !emb) return fail(EP_EINVAL, "NULL argument");
  *out = nullptr;
  if (n_nodes == 0) return fail(EP_EINVAL, "n_nodes must be > 0");
  if (!ep::row_dims_supported((int)dims)) return fail(EP_EINVAL, "embedding width not built for alignment");
  if (indptr[0] != 0) return fail(EP_EINVAL, "indptr[0] must be 0");
  for (uint64_t i = 0; i < n_nodes; ++i)
    if (indptr[i + 1] < indptr[i]) return fail(EP_EINVAL, "indptr must be non-decreasing");
  const uint64_t ne = indptr[n_nodes];
  if (ne && (!indices || !cdf)) return fail(EP_EINVAL, "NULL argument");
  for (uint64_t i = 0; i < ne; ++i) {
    if (indices[i] >= n_nodes) return fail(EP_EINVAL, "edge target out of range");
    if (!std::isfinite(cdf[i])) return fail(EP_EINVAL, "edge CDF values must be finite");
  }
  ep_aligner* a = new ep_aligner();
  a->device = device;
  a->n = n_nodes;
  a->dims = dims;
  auto bail = [&](int code, const std::string& msg) { aligner_free(a); return fail(code, msg); };
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&a->ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&a->ev[1]);
  if (e == hipSuccess) e = hipMalloc((void**)&a->indptr, (n_nodes + 1) * sizeof(uint64_t));
  if (e == hipSuccess) e = hipMalloc((void**)&a->indices, std::max<uint64_t>(ne, 1) * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc((void**)&a->cdf, std::max<uint64_t>(ne, 1) * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&a->emb, (size_t)n_nodes * dims * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(a->indptr, indptr, (n_nodes + 1) * sizeof(uint64_t), hipMemcpyHostToDevice);
  if (e == hipSuccess && ne) e = hipMemcpy(a->indices, indices, ne * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess && ne) e = hipMemcpy(a->cdf, cdf, ne * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(a->emb, emb, (size_t)n_nodes * dims * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) return bail(EP_EHIP, std::string("aligner setup: ") + hipGetErrorString(e));
  *out = a;
  return EP_OK;
}

int ep_aligner_run(ep_aligner* a, float alpha, uint32_t max_neighbors, uint64_t first, uint64_t count, float* out) {
  if (!a || !out) return fail(EP_EINVAL, "NULL argument");
  if (first > a->n || count > a->n - first) return fail(EP_EINVAL, "node range out of bounds");
  const int adaptive = std::isnan(alpha) ? 1 : 0;
  if (!adaptive && !std::isfinite(alpha)) return fail(EP_EINVAL, "alpha must be finite (NaN: adaptive)");
  if (count == 0) return EP_OK;
  EP_HIP(hipSetDevice(a->device));
  if (a->out_rows < count) {
    if (a->out) EP_HIP(hipFree(a->out));
    a->out = nullptr;
    a->out_rows = 0;
    EP_HIP(hipMalloc((void**)&a->out, (size_t)count * a->dims * sizeof(float)));
    a->out_rows = count;
  }
  EP_HIP(hipEventRecord(a->ev[0], a->stream));
  EP_HIP(ep::launch_align((int)a->dims, a->emb, a->indptr, a->indices, a->cdf, adaptive ? 0.f : alpha, adaptive,
                          max_neighbors, first, count, a->out, a->stream));
  EP_HIP(hipEventRecord(a->ev[1], a->stream));
  EP_HIP(hipMemcpyAsync(out, a->out, (size_t)count * a->dims * sizeof(float), hipMemcpyDeviceToHost, a->stream));
  EP_HIP(hipStreamSynchronize(a->stream));
  EP_HIP(hipEventElapsedTime(&a->last_ms, a->ev[0], a->ev[1]));
  return EP_OK;
}

float ep_aligner_last_kernel_ms(const ep_aligner* a) { return a ? a->last_ms : 0.f; }

void ep_aligner_destroy(ep_aligner* a) { aligner_free(a); }

int ep_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  if (!ctr || !key || !out) return fail(EP_EINVAL, "NULL argument");
  const ep::u32x4 r = ep::philox4x32(ep::u32x4{ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
  out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
  return EP_OK;
}

}  // extern "C"
