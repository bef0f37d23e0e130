// Internal interface between the EP host runtime (ep_capi.hip) and the HIP kernels
// (ep_kernels.hip).  Not part of the public C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ep {

// Device-resident per-step state.  Written by k_step_begin / k_scan / the scatter
// kernels; read by every kernel of the step, so one captured step graph can be
// replayed for every batch of a pass.
struct StepState {
  uint64_t next_step;      // LR-schedule step counter (1-based, like the reference's AtomicUsize)
  uint64_t stat_rows;      // cumulative feature-row reads (sum of R_e)
  uint64_t stat_nodes;     // cumulative node-id reads (sampled neighbours + negatives)
  uint64_t stat_unique;    // cumulative unique rows updated
  uint64_t stat_examples;  // cumulative anchors processed
  uint32_t pass;           // 1-based pass
  uint32_t step_in_pass;   // batch index within the pass (0-based)
  uint32_t batch_n;        // anchors in the current batch
  uint32_t anchor_base;    // offset of the batch in the pass permutation
  uint32_t n_keys;         // gradient keys emitted this step (incl. sentinels)
  uint32_t n_unique;       // unique rows produced by the scatter this step
  uint32_t n_long;         // long segments deferred this step
  uint32_t n_pieces;       // pieces of long segments this step
  uint32_t n_touched;      // rows touched by the atomic path this step
  uint32_t err;            // nonzero: key capacity exceeded this step
  float lr;                // learning rate of this step
  float bc1, bc2;          // Adam bias corrections 1 - beta^t, t = pass + 1
  uint32_t n_real;         // non-sentinel keys (the sorted key count)
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
  uint32_t sentinel;       // key for "no gradient" (>= n_vocab, all ones in sort_bits)
  int sort_bits;
  // parameters
  void* table;             // float* (fp32) or uint16_t* (bf16)
  // Optimizer state: one record of `ostride` floats per row, [master |] m | v, so an
  // update touches one contiguous record (master only in bf16 mode).
  float* master;           // record + 0 (bf16 mode) else nullptr
  float* mom;              // record + 0 (fp32) / + D (bf16)
  float* var;              // record + D (fp32) / + 2D (bf16)
  uint32_t ostride;        // 2D (fp32) or 3D (bf16)
  int bf16;
  int dims;
  // batch
  const uint32_t* perm;    // anchors of the pass, in order
  const uint32_t* negpool; // negatives drawn uniformly from this pool
  uint32_t n_negpool;
  uint32_t n_batch_total;  // anchors in the pass
  uint32_t batch;          // B
  uint32_t S;              // neighbour cap (Fixed) ; 0 with all_nbrs
  int all_nbrs;            // 1: use every neighbour (Sample::All)
  uint32_t K;              // negatives
  float gamma;
  int lw_kind;             // 0 none, 1 log, 2 exponential
  float lw_exp;
  int grad_mode;           // 0 first_leaf, 1 exact
  int train;               // 1: compute adjoints and emit keys
  int explicit_samples;    // 1: neighbour / negative lists are provided by the host
  uint64_t seed;
  // schedule
  float alpha;
  uint32_t total_updates;
  uint32_t steps_per_pass;
  // per-step buffers
  uint32_t* nbr_cnt;       // [B]
  uint32_t* nbr;           // [B*S]
  uint32_t* neg;           // [B*K]
  uint32_t* rcount;        // [B]
  uint32_t* keyoff;        // [B+1]
  float* adj;              // [B*NC*D]
  float* coef;             // [B*NC]
  uint8_t* elig;           // [B*NC]
  float* loss;             // [B]
  float* pass_loss;        // [n_batch_total] or nullptr
  float* xout;             // [B*NC*D] constructions (test hook) or nullptr
  uint32_t hash_entries;   // first-leaf LDS hash size (power of two)
  uint32_t ent_cap;        // forward LDS stage: max entries (nodes) per example, <= 256
  uint32_t ids_cap;        // forward LDS stage: max feature occurrences per example
  int fuse_emit;           // k_forward emits the keys of LDS-staged examples
  // keys
  uint32_t* keys[2];
  uint32_t* slots[2];
  uint32_t key_cap;
  uint32_t* ghist;         // [2048 * sort_grid]
  int sort_grid;
  // scatter (sort path)
  float* G;                // [key_cap*D] compact gradient rows (bounded by unique count)
  uint32_t* ufeat;         // [key_cap]
  uint32_t* long_st;       // [long_cap]
  uint32_t* long_len;
  uint32_t* long_pbase;
  uint32_t* piece_seg;     // [piece_cap]
  float* part;             // [piece_cap*D]
  uint32_t long_cap, piece_cap;
  // scatter (atomic path)
  float* Gd;               // [V*D]
  uint32_t* flag;          // [V]
  uint32_t* touched;       // [V]
  int scatter;             // 0 sort, 1 atomic
  int grad_out;            // test hook: write merged gradients (G/ufeat) instead of updating
  int unique_grid;         // grid size for the persistent row kernels
  int seg_vpl;             // k_seg_adam values per lane (4, 8, 16; chosen by the tuner)
};

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
hipError_t launch_scatter_sorted(const StepConfig& c, StepState* st, int buf, hipStream_t s);
hipError_t launch_atomic_scatter(const StepConfig& c, StepState* st, hipStream_t s);
hipError_t launch_adam_touched(const StepConfig& c, StepState* st, hipStream_t s);
// One-off kernels.
hipError_t launch_init_table(const StepConfig& c, uint64_t seed, hipStream_t s);
hipError_t launch_fill_bf16(const float* src, uint16_t* dst, uint64_t n, uint32_t D, uint32_t stride, hipStream_t s);
hipError_t launch_embed_averaged(const StepConfig& c, const uint64_t* findptr, const uint32_t* fids,
                                 uint64_t n_nodes, float* out, hipStream_t s);

bool dims_supported(int d);
bool seg_vpl_supported(int d, int vpl);

}  // namespace ep
