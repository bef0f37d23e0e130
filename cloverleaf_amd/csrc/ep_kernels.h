// Internal interface between the EP host runtime (ep_capi.hip) and the HIP kernels
// (ep_kernels.hip).  Not part of the public C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ep {

// A step processes a group of up to kMaxGroup reference batches that read the same
// parameters (the reference's concurrently running Hogwild batches, mod.rs:174 par_iter
// over chunks); their updates are applied one after the other in batch order.
constexpr uint32_t kMaxGroup = 8;
// Waves per 256-thread block of every kernel.
constexpr int kBlockWaves = 4;
// Sorted values: slot in the low bits, batch of the group above.
constexpr uint32_t kSlotBatchShift = 29;
constexpr uint32_t kSlotMask = (1u << kSlotBatchShift) - 1u;

// Device-resident per-step state.  Written by k_step_begin / k_scan / the scatter
// kernels; read by every kernel of the step, so one captured step graph can be
// replayed for every batch of a pass.
struct StepState {
  uint64_t next_step;      // LR-schedule step counter (1-based, like the reference's AtomicUsize)
  uint64_t stat_rows;      // cumulative feature-row reads (sum of R_e)
  uint64_t stat_nodes;     // cumulative node-id reads (sampled neighbours + negatives)
  uint64_t stat_unique;    // cumulative unique rows updated
  uint64_t stat_examples;  // cumulative anchors processed
  uint64_t stat_short;     // cumulative rows updated by the short-row kernel
  uint64_t stat_att_n2;    // attention: cumulative sum over blocks of items^2
  uint64_t stat_att_n;     // attention: cumulative sum over blocks of items
  uint32_t pass;           // 1-based pass
  uint32_t step_in_pass;   // batch index within the pass (0-based)
  uint32_t batch_n;        // anchors in the current batch
  uint32_t anchor_base;    // offset of the batch in the pass permutation
  uint32_t n_keys;         // gradient keys emitted this step (incl. sentinels)
  uint32_t n_unique;       // unique rows produced by the scatter this step
  uint32_t n_long;         // long segments deferred this step
  uint32_t n_pieces;       // pieces of long segments this step
  uint32_t n_touched;      // rows touched by the atomic path this step
  uint32_t n_short;        // short rows planned for k_seg_update this step
  uint32_t err;            // nonzero: key capacity exceeded this step
  uint32_t n_batches;      // reference batches in this step's group (<= kMaxGroup)
  float lr[8];             // learning rate of each batch of the group (kMaxGroup)
  float noise_sd[8];       // gradient-noise sd of each batch of the group
  uint64_t step0;          // step number of the group's first batch
  float bc1, bc2;          // Adam bias corrections 1 - beta^t, t = pass + 1
  uint32_t n_real;         // non-sentinel keys (the sorted key count)
  uint32_t reg_next;       // next plan region for k_seg_update (cleared by k_seg_plan)
  uint32_t dtot[4][2048];  // radix-sort digit totals, one row per sort pass (<= 11-bit digits)
};

// Static description of the step (fixed for the life of a captured graph).
struct StepConfig {
  // graph and features
  const uint64_t* indptr;
  const uint32_t* indices;
  const uint64_t* findptr;
  const uint32_t* fids;
  uint64_t n_nodes;
  uint32_t n_vocab;
  // Gradient key = perm(f) = f * key_a mod key_m (key_m = 2^sort_bits - 1 > every row
  // id, gcd(key_a, key_m) = 1, key_ainv the inverse of key_a mod key_m): the permutation
  // interleaves hot and cold rows in the sorted order, so 64-entry chunks of the sorted
  // keys hold similar numbers of rows.  sentinel = key_m (no real key reaches it).
  // The value sorted with a key is the slot of its adjoint row with the example's batch
  // (index in the step's group) in the top bits: (batch << kSlotBatchShift) | slot.  The
  // sort is stable and keys are emitted in example order, so within a row the values
  // rise, batch by batch.
  uint32_t sentinel;       // key for "no gradient", above every real key
  int sort_bits;
  uint32_t key_m, key_a, key_ainv, key_bits;  // key_m = 2^key_bits - 1
  uint32_t pbits, pmask;   // log2(group size), group size - 1
  // parameters
  void* table;             // float* (fp32) or uint16_t* (bf16)
  // Optimizer state: one record of `ostride` floats per row, [master |] m | v, so an
  // update touches one contiguous record (master only with a persistent FP32 master).
  float* master;           // record + 0 (persistent-master bf16 mode) else nullptr
  float* mom;              // record + 0 (fp32) / + D (bf16)
  float* var;              // record + D (fp32) / + 2D (bf16)
  uint32_t ostride;        // 2D (fp32) or 3D (bf16)
  int bf16;
  int dims;
  // batch
  const uint32_t* perm;    // anchors of the pass, in order
  const uint32_t* negpool; // negatives drawn uniformly from this pool
  uint32_t n_negpool;
  uint32_t hard_negs;      // random-walk hard negatives per example (min(hard_negatives, K))
  uint32_t n_batch_total;  // anchors in the pass
  uint32_t batch;          // anchors per step: group size x rbatch (buffer capacity)
  uint32_t rbatch;         // B, the reference batch size
  uint32_t S;              // neighbour cap (Fixed) ; 0 with all_nbrs
  int all_nbrs;            // 1: use every neighbour (Sample::All)
  // Weighted neighbours (model.rs:342-386).  With wavg every occurrence of neighbour j
  // in ~h(v) counts w_j + 1, w_j = scale * p_j * deg (p_j the edge's probability from the
  // CDF); its gradient goes through its own adjoint row nrow_base + nbr_base(b) + j.
  const float* cdf;        // per-edge CDF of each row (CumCSR) or nullptr
  float* nbr_w;            // [B*S] w_j of the sampled neighbours (Fixed)
  uint32_t* nbroff;        // [B] first neighbour row of each example (all_nbrs)
  int wavg;                // weighted_neighbor_averaging
  int wsample;             // weighted_neighbor_sampling
  uint32_t nrow_base;      // first per-neighbour adjoint row
  // max_features (utils.rs:193-216, model.rs:264-269): 0 All, 1 Fixed(mf_k), 2
  // Probability(mf_p); every node of an example keeps a subset of its (at most 64)
  // features, each kept occurrence counting weight + scale
  int mf_kind;
  uint32_t mf_k;
  float mf_p;
  uint32_t sel_cap;        // forward LDS stage: selection masks (ent_cap with max_features, else 0)
  int nbr_rows;            // ~h(v) with per-neighbour counts and adjoint rows (wavg or max_features)
  // max_nodes Probability(mn_p) (utils.rs:203-210): the sampled list of an anchor v and
  // the neighbours' weights live at v's CSR offset (anchors of a step are distinct)
  int pnbrs;
  float mn_p;
  uint32_t* nbr_sel;       // [E]
  float* nbr_selw;         // [E] (with nbr_rows)
  uint32_t K;              // negatives
  float gamma;
  int loss_kind;           // EP_LOSS_*: 0 margin, 1 contrastive, 2 starspace, 3 rank,
                           // 4 rankspace, 5 ppr (margin over random-walk positives)
  float loss_p2;           // contrastive negative margin / ppr restart probability
  int lw_kind;             // 0 none, 1 log, 2 exponential
  float lw_exp;
  int grad_mode;           // 0 first_leaf, 1 exact
  int train;               // 1: compute adjoints and emit keys
  int explicit_samples;    // 1: neighbour / negative lists are provided by the host
  uint64_t seed;
  // schedule
  float noise;             // > 0: gaussian gradient noise, sd = noise * noise_decay^step
  float noise_decay;
  float alpha;
  uint32_t total_updates;
  uint32_t steps_per_pass;
  // per-step buffers
  uint32_t* nbr_cnt;       // [B]
  uint32_t* nbr;           // [B*S]
  uint32_t* neg;           // [B*K]
  uint32_t* rcount;        // [B]
  // [B*ent_cap] per example, the feature list of each entry e < ent_cap (v, the sampled
  // neighbours, the negatives): its start in fids and its length, written by k_sample
  uint64_t* ex_fs;
  uint32_t* ex_fn;
  uint32_t* keyoff;        // [B+1]
  float* adj;              // [B*NC*D] FP32 adjoint rows (FP32 tables, attention gradient rows)
  // FP16 adjoint rows (BF16 tables, mean model): row r holds g / s in FP16 (round to
  // nearest) with s = max |g| of the row in ascale[r]; readers take s * value in FP32.
  // nullptr: the FP32 rows in adj.
  uint16_t* adjh;          // [B*NC*D]
  float* ascale;           // [B*NC]
  float* coef;             // [B*NC]
  uint8_t* elig;           // [B*NC]
  float* loss;             // [B]
  float* pass_loss;        // [n_batch_total] or nullptr
  float* xout;             // [B*NC*D] constructions (test hook) or nullptr
  uint32_t hash_entries;   // first-leaf LDS hash size (power of two)
  uint32_t ent_cap;        // forward LDS stage: max entries (nodes) per example, <= 256
  uint32_t ids_cap;        // forward LDS stage: max feature occurrences per example
  int fuse_emit;           // k_forward emits the keys of LDS-staged examples
  uint32_t fuse_hash;      // entries of k_forward's first-leaf hash (power of two that fits the
                           // construction region of its LDS)
  uint32_t fuse_max;       // first_leaf: k_forward emits examples of at most this many
                           // occurrences (5/8 of fuse_hash), k_emit the larger ones
  // keys
  uint32_t* keys[2];
  uint32_t* slots[2];
  uint32_t key_cap;
  uint32_t* ghist;         // [2048 * sort_grid]
  int sort_grid;
  // scatter (sort path)
  float* G;                // [key_cap*D] compact gradient rows (bounded by unique count)
  uint32_t* ufeat;         // [key_cap]
  // Plan regions: region r covers chunks [r*cpr, (r+1)*cpr) of 64 sorted entries (cpr =
  // ceil(chunks / n_regions), from the step's key count); k_seg_plan wave w owns the
  // n_regions / (unique_grid * NW) regions from w * that count and writes, without
  // atomics, each region's short rows (<= kSegShort entries) at srec[r*cpr*64 ...] with
  // count scount[r], one record {first entry, row id, run-start mask (bit i: entry i
  // starts a (row, batch) run), len | batch of run j << (6 + 3j)}, and its medium rows
  // (more entries, inside the 128-entry window, at most 2 per chunk) at mrec[r*cpr*2 ...]
  // with count mcount[r], {first entry, len}.  k_seg_update's wave w takes region w and,
  // when there are more regions than waves, further ones from the StepState queue
  // reg_next, so its blocks need not all be resident at once.
  uint4* srec;             // [srec_cap]
  uint32_t srec_cap;
  uint2* mrec;             // [mrec_cap]
  uint32_t mrec_cap;
  uint32_t* scount;        // [n_regions]
  uint32_t* mcount;        // [n_regions]
  uint32_t n_regions;      // a multiple of unique_grid * NW
  // Long rows (running past the window, so at least 65 entries): first sorted entry,
  // first piece, run count and per run (piece count | batch << 24) and its entry count.
  // Only runs of more than kPieceEntries entries are cut into pieces (a piece is a range
  // of one run: start entry and length; kPieceEntries entries, or more for hot runs so a
  // run has at most 64 pieces); the combine kernel sums the shorter runs straight from
  // their adjoint rows.  The pieces of a row follow each other run by run, so a step has
  // fewer than n_real / 16 pieces.
  uint32_t* long_st;       // [long_cap]
  uint32_t* long_len;      // runs of the row
  uint32_t* long_pbase;
  uint32_t* long_runs;     // [long_cap * kMaxGroup]
  uint32_t* long_rlen;     // [long_cap * kMaxGroup] entries of each run
  uint32_t* piece_st;      // [piece_cap]
  uint32_t* piece_info;    // [piece_cap] entries of the piece | batch << 29
  float* part;             // [piece_cap*D]
  uint32_t long_cap, piece_cap;
  // scatter (atomic path)
  float* Gd;               // [V*D]
  uint32_t* flag;          // [V]
  uint32_t* touched;       // [V]
  int scatter;             // 0 sort, 1 atomic
  int grad_out;            // test hook: write merged gradients (G/ufeat) instead of updating;
                           // ufeat = row | batch << 27
  int unique_grid;         // grid size for the persistent row kernels
  int seg_vpl;             // k_seg_adam values per lane (4, 8, 16; chosen by the tuner)
  int seg_unroll;          // k_seg_adam adjoint rows in flight per lane group (tuned)
  int fwd_x;               // k_forward rows-in-flight multiplier of the wide gathers: 1 or 2 (tuned)
  // Attention model (model = 1): rows are [Q_1..Q_H | K_1..K_H | V_1..V_H], dims =
  // att_heads * (2 att_dk + att_dm).  An example's attention blocks (h(v), one per
  // sampled neighbour, one per negative) start at blkoff[b]; bout/badj hold each
  // block's output and adjoint (att_dm values).  The gradient rows are written per
  // feature occurrence: adj[slot * dims] with slot = the occurrence's key position.
  int model;               // 0 mean, 1 attention
  int att_heads, att_dk, att_dm;
  uint32_t att_win;        // Sliding attention window_size (attention.rs:194-218); ~0u: Full
  // Random attention (attention.rs:221-250): max_features of the attention model as the
  // per-row column sample, 0 none, 1 Fixed(att_rk), 2 Probability(att_rp)
  int att_rkind;
  uint32_t att_rk;
  float att_rp;
  uint32_t* blkoff;        // [batch + 1]
  uint32_t* bitems;        // [blk_cap] items of each block (written by the forward)
  uint32_t* bnode;         // [blk_cap] node of each block        (written by k_att_map)
  uint32_t* bslot;         // [blk_cap] key position of its first feature occurrence
  uint32_t* bmeta;         // [blk_cap] example << 5 | construction
  float* bout;             // [blk_cap * att_dm]
  float* badj;             // [blk_cap * att_dm]
  uint32_t blk_cap;
};

// The neighbour list of example b (anchor row at s0): the CSR row itself (every
// neighbour), the Probability sample at the anchor's CSR offset, or the Fixed sample.
__device__ inline const uint32_t* nbr_list(const StepConfig& c, uint32_t b, uint64_t s0) {
  return c.all_nbrs ? c.indices + s0 : (c.pnbrs ? c.nbr_sel + s0 : c.nbr + (size_t)b * c.S);
}


// Launchers (all asynchronous on `stream`).
hipError_t launch_step_begin(const StepConfig& c, StepState* st, hipStream_t s);
// Same, with a fixed learning rate instead of the schedule (explicit test steps).
hipError_t launch_step_begin_lr(const StepConfig& c, StepState* st, float lr, hipStream_t s);
hipError_t launch_sample(const StepConfig& c, StepState* st, hipStream_t s);
hipError_t launch_scan(const StepConfig& c, StepState* st, hipStream_t s);
hipError_t launch_forward(const StepConfig& c, StepState* st, hipStream_t s);
hipError_t launch_emit(const StepConfig& c, StepState* st, hipStream_t s);
// Sort keys/slots; returns which buffer (0/1) holds the result in *out_buf.
hipError_t launch_sort(const StepConfig& c, StepState* st, hipStream_t s, int* out_buf);
// Segment reduce + Adam over the sorted keys (grad_out: write merged gradients instead).
// ev (nullable): events recorded after the plan, short-row, pieces and long-row kernels.
hipError_t launch_scatter_sorted(const StepConfig& c, StepState* st, int buf, hipStream_t s,
                                 hipEvent_t* ev = nullptr);
hipError_t launch_atomic_scatter(const StepConfig& c, StepState* st, hipStream_t s);
hipError_t launch_adam_touched(const StepConfig& c, StepState* st, hipStream_t s);
// One-off kernels.
hipError_t launch_init_table(const StepConfig& c, uint64_t seed, hipStream_t s);
hipError_t launch_fill_bf16(const float* src, uint16_t* dst, uint64_t n, uint32_t D, uint32_t stride, hipStream_t s);
hipError_t launch_embed_averaged(const StepConfig& c, const uint64_t* findptr, const uint32_t* fids,
                                 uint64_t n_nodes, float* out, hipStream_t s);
hipError_t launch_embed_weighted(const StepConfig& c, const uint64_t* findptr, const uint32_t* fids,
                                 const float* p_w, float alpha, uint64_t n_nodes, float* out, hipStream_t s);

bool dims_supported(int d);


// Row widths of the table-row kernels (init, update, averaged embedding): the mean
// model's widths and the attention row widths.
bool row_dims_supported(int d);
hipError_t launch_align(int dims, const float* emb, const uint64_t* indptr, const uint32_t* indices, const float* cdf,
                        float alpha, int adaptive, uint32_t max_nb, uint64_t first, uint64_t count, float* out,
                        hipStream_t s);
// Attention model kernels (ep_attention.hip): forward runs the block outputs and the
// loss/adjoints; backward writes the per-occurrence gradient rows (train steps).
bool att_shape_supported(int heads, int dk, int dm);
hipError_t launch_att_forward(const StepConfig& c, StepState* st, hipStream_t s);
hipError_t launch_att_backward(const StepConfig& c, StepState* st, hipStream_t s);
// AttentionAggregator inference (count 1 per listed feature): out[n_nodes * att_dm].
hipError_t launch_att_embed(const StepConfig& c, const uint64_t* findptr, const uint32_t* fids, uint64_t n_nodes,
                            float* out, hipStream_t s);
// Segments of the sorted keys longer than kSegShort entries are summed as pieces of
// kPieceEntries entries by k_pieces and combined in piece order by k_long_combine.
constexpr uint32_t kSegShort = 32;
constexpr uint32_t kPieceEntries = 32;
bool seg_vpl_supported(int d, int vpl);
// Whether k_seg_adam<D, VPL, U> is built (U in {8, 16} / (VPL/4)).
bool seg_variant_supported(int d, int vpl, int unroll);
// Forward variants: x = 1, and x = 2 (twice the rows in flight) where the wide gather runs.
bool fwd_variant_supported(int d, int bf16, int x);

}  // namespace ep
