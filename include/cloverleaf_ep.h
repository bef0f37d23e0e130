/*
 * cloverleaf_ep.h -- C ABI of the MI355X Embedding Propagation (EP) training step.
 *
 * This is the drop-in boundary for the path
 *     EmbeddingPropagator.learn_features(graph, features, feature_embeddings)
 * of the reference (src/lib.rs:1327-1370), which calls
 *     EmbeddingPropagation::learn(&G, &FeatureStore, Option<EmbeddingStore>, &Model)
 * (src/algos/ep/mod.rs:80-89 -> learn_feature_embeddings 92-287).
 *
 * Conventions
 *   - every call returns 0 on success or a negative EP_E* code; the message of the
 *     last failure on the calling thread is available from ep_last_error();
 *   - host buffers are caller-owned and copied in or out; device memory is owned by
 *     the ep_ctx;
 *   - one host thread per context; calls on one context are not re-entrant;
 *   - ids are uint32 (node ids < 2^32, feature ids < 2^27 - 1), CSR offsets uint64.
 */
#ifndef CLOVERLEAF_EP_H
#define CLOVERLEAF_EP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  EP_OK = 0,
  EP_EINVAL = -1,      /* bad argument / unsupported option                    */
  EP_ESTATE = -2,      /* call out of order (e.g. learn before set_graph)       */
  EP_EHIP = -3,        /* HIP runtime error                                     */
  EP_ENOMEM = -4,      /* device allocation failed                              */
  EP_ECAPACITY = -5,   /* a step exceeded a preallocated device capacity        */
  EP_EDUPLICATE = -6,  /* a file names a node twice (PyKeyError in the reference) */
};

/* Sample kinds of max_nodes / max_features (src/algos/utils.rs:166-191).       */
enum { EP_SAMPLE_ALL = 0, EP_SAMPLE_FIXED = 1, EP_SAMPLE_PROBABILITY = 2 };
/* LossWeighting (src/algos/ep/mod.rs:31-36).                                   */
enum { EP_LW_NONE = 0, EP_LW_LOG = 1, EP_LW_EXPONENTIAL = 2 };
/* Table precision.  FP32: FP32 table.  BF16: BF16 table; each touched row is taken
 * to FP32 for the update (FP32 master copy of the touched rows) and written back with
 * stochastic rounding.  BF16_MASTER: BF16 working table plus a persistent FP32 master
 * of every row (the update reads and writes the master, the BF16 copy is rounded
 * from it stochastically).  Adam moments are FP32 in every mode.                  */
enum { EP_PREC_FP32 = 0, EP_PREC_BF16 = 1, EP_PREC_BF16_MASTER = 2 };
/* Gradient extraction: the reference's first-leaf rule (mod.rs:347-352,366-367)
 * or the exact gradient (sum over every construction).                         */
enum { EP_GRAD_FIRST_LEAF = 0, EP_GRAD_EXACT = 1 };
/* Batch merge + update: sort + segment-reduce (deterministic) or float atomics. */
enum { EP_SCATTER_SORT = 0, EP_SCATTER_ATOMIC = 1 };

/* Mirrors EmbeddingPropagation (src/algos/ep/mod.rs:40-75) plus the model options
 * of EmbeddingPropagator::new (src/lib.rs:1205-1307).                           */
/* Losses (src/algos/ep/loss.rs:15-38).  PPR: margin loss against random-walk
 * positives (K walks with restart, loss.rs:191-226) instead of sampled neighbours. */
enum {
  EP_LOSS_MARGIN = 0,
  EP_LOSS_CONTRASTIVE = 1,
  EP_LOSS_STARSPACE = 2,
  EP_LOSS_RANK = 3,
  EP_LOSS_RANKSPACE = 4,
  EP_LOSS_PPR = 5
};

typedef struct ep_config {
  uint32_t dims;               /* d_model; mean model: 8,16,32,64,100,128,256;      */
                               /* attention: 8..256, a multiple of 8, with a table  */
                               /* row width heads*(2 d_k + dims) of 8..256, 384,    */
                               /* 512 or 768                                        */
  uint32_t batch_size;         /* default 50                                        */
  uint32_t passes;             /* default 100                                       */
  uint32_t negatives;          /* MarginLoss negatives, default 1                   */
  float alpha;                 /* learning rate, default 0.9                        */
  float gamma;                 /* MarginLoss margin, default 1.0                    */
  float valid_pct;             /* default 0.1                                       */
  int32_t max_nodes_kind;      /* EP_SAMPLE_ALL (default) / FIXED / PROBABILITY     */
  uint32_t max_nodes;          /* S when Fixed                                      */
  int32_t max_features_kind;   /* EP_SAMPLE_*: every node of an example keeps a    */
                               /* subset of its features (model.rs:264-269); Fixed  */
                               /* and Probability need <= 64 features per node and  */
                               /* the mean model                                    */
  uint32_t max_features;       /* k when Fixed                                      */
  int32_t loss_weighting;      /* EP_LW_*                                           */
  float loss_weight_exp;       /* exponent of EP_LW_EXPONENTIAL                     */
  uint32_t hard_negatives;     /* random-walk hard negatives per example, at most K */
                               /* used (RandomWalkHardSampler, node_sampler.rs:83-107)*/
  float noise;                 /* gradient noise sd at step 0, decaying to 1e-9 over */
                               /* the run (mod.rs:139-147, 238-247); 0 = off       */
  int32_t weighted_neighbor_sampling;   /* reservoir by edge weight (model.rs:373-376);*/
                                        /* changes the sample only together with     */
                                        /* weighted_neighbor_averaging               */
  int32_t weighted_neighbor_averaging;  /* ~h(v) counts scale*p_j*deg + 1 per        */
                                        /* occurrence of neighbour j (model.rs:357-  */
                                        /* 361); needs ep_set_edge_weights; mean     */
                                        /* model only                                */
  int32_t precision;           /* EP_PREC_*                                         */
  int32_t grad_mode;           /* EP_GRAD_*                                         */
  int32_t scatter;             /* EP_SCATTER_*                                      */
  int32_t use_graph;           /* 1: replay each step from a captured hipGraph      */
  uint64_t seed;               /* default 20222022 (src/lib.rs:90)                  */
  int32_t device;              /* HIP device ordinal                                */
  int32_t seg_vpl;             /* segment-Adam tile: values per lane (4, 8, 16) and */
  int32_t seg_unroll;          /* adjoint rows in flight ({8,16}/(vpl/4)); both 0 = */
                               /* autotuned over the first training steps           */
  uint32_t concurrent_batches; /* batches per step (1, 2, 4, 8): they read the same  */
                               /* parameters and update in batch order, as the      */
                               /* reference's concurrently running batches          */
                               /* (mod.rs:174); 1 = strictly sequential (default)   */
  uint32_t attention_dk;       /* 0: mean aggregation (AveragedFeatureModel);        */
                               /* d_k > 0: AttentionFeatureModel with Full attention  */
                               /* (lib.rs:1288-1299, `attention`); d_k in 4,8,16,32, */
                               /* at most 64 features per node                       */
  uint32_t attention_heads;    /* `attention_heads` (default 1); table rows are      */
                               /* heads * (2 d_k + dims) wide (model.rs:241-243)     */
  uint32_t context_window;     /* `context_window` k >= 1: Sliding attention with    */
                               /* window_size k (attention.rs:194-218); 0: Full      */
  int32_t loss_kind;           /* EP_LOSS_* (EPLoss, lib.rs:939-1110; loss.rs:15-38); */
                               /* gamma is the first parameter: margin / starspace   */
                               /* gamma, contrastive positive margin, rank tau,      */
                               /* rankspace gamma, ppr gamma                         */
  float loss_param2;           /* contrastive negative margin / ppr restart_p        */
  float max_features_p;        /* p in (0, 1) when max_features is Probability      */
  float max_nodes_p;           /* p in (0, 1) when max_nodes is Probability (mean   */
                               /* model)                                            */
  int32_t fp16_adjoints;       /* 1: with a BF16 table (mean model, sort scatter)   */
                               /* the per-construction adjoint rows the update      */
                               /* gathers are held in FP16 with a power-of-two scale */
                               /* per row (2^-11 relative to the row's largest      */
                               /* value); 0: FP32 adjoint rows (default)            */
  int32_t forward_variant;     /* forward gather rows in flight: 1 (x1), 2 (x2, for */
                               /* the widths the 16-byte gather serves); 0 = tuned  */
                               /* with seg_vpl/seg_unroll (default); every variant  */
                               /* gives identical results                           */
} ep_config;

typedef struct ep_ctx ep_ctx;

/* Progress callback, once per pass: pass (1-based), global step, mean train
 * error (mod.rs:256-265), mean validation loss (or NaN), current LR.            */
typedef void (*ep_progress_fn)(void* user, uint32_t pass, uint64_t step, float train_err, float valid_err,
                               float lr);

/* Per-run measurement of ep_bench_steps. */
typedef struct ep_bench_result {
  double ms_total;             /* wall time of the timed steps (HIP events)        */
  double ms_forward_avg;       /* average duration of the gather/loss kernel        */
  double ms_scatter_avg;       /* average duration of scatter+update (all kernels)  */
  uint64_t examples;           /* anchors processed in the timed steps              */
  uint64_t rows;               /* feature-row reads (sum of R_e) in the timed steps */
  uint64_t nodes;              /* sampled neighbour + negative ids read              */
  uint64_t unique_rows;        /* rows updated by Adam in the timed steps            */
  uint32_t steps;
  uint32_t forward_launches;
  int32_t seg_vpl;             /* segment-Adam variant in use (after tuning)        */
  int32_t seg_unroll;
  /* instrumented run: average duration per step of each phase, in ms:             */
  /* 0 step begin + sample + scan, 1 forward, 2 key emission, 3 radix sort,        */
  /* 4 row plan, 5 short-row update, 6 pieces + medium rows, 7 long rows           */
  double ms_phase[8];
  uint64_t short_rows;         /* rows updated by the short-row kernel (timed run)  */
  uint32_t last_long_rows;     /* long rows and their pieces in the last step        */
  uint32_t last_pieces;
  uint64_t att_items_sq;       /* attention model, timed steps: sum over attention    */
  uint64_t att_items;          /* blocks of items^2 and of items (FLOP model inputs)  */
  int32_t fwd_x;               /* forward variant in use: rows in flight x1 or x2     */
  int32_t reserved;
} ep_bench_result;

/* Fills the reference defaults (EmbeddingPropagator::new, src/lib.rs:1262-1274). */
int ep_config_default(ep_config* cfg);

/* Creates a context (device memory, stream).  Replaces the construction of
 * EmbeddingPropagation + AveragedFeatureModel (src/lib.rs:1262-1306).          */
int ep_create(const ep_config* cfg, ep_ctx** out);
void ep_destroy(ep_ctx* ctx);

/* Uploads the CumCSR adjacency (src/graph.rs:399-550; only offsets and columns
 * are used by the unweighted path).  indptr[n_nodes+1], indices[indptr[n]].    */
int ep_set_graph(ep_ctx* ctx, uint64_t n_nodes, const uint64_t* indptr, const uint32_t* indices);

/* Uploads the per-edge CDF of each row (CumCSR weights, graph.rs:414-422),
 * cdf[indptr[n]], after ep_set_graph.  Edge j of a row has probability
 * cdf[j] - cdf[j-1] (CDFtoP, graph.rs:630-661).  Needed for
 * weighted_neighbor_averaging.                                                  */
int ep_set_edge_weights(ep_ctx* ctx, const float* cdf);

/* Uploads the FeatureStore (src/feature_store.rs:19-25) as a CSR: node i has
 * features f_ids[f_indptr[i]..f_indptr[i+1]).  Every node must have >= 1 feature
 * (FeatureStore::fill_missing_nodes, feature_store.rs:83-89).                  */
int ep_set_features(ep_ctx* ctx, uint64_t n_nodes, uint32_t n_vocab, const uint64_t* f_indptr,
                    const uint32_t* f_ids);

/* Table rows are `dims` wide for the mean model and attention_heads * (2
 * attention_dk + dims) wide for the attention model ("row width" below).        */

/* Sets the feature table (n_vocab x row width, row-major FP32); rows == NULL draws the
 * random init of randomize_embedding_store (src/algos/ep/mod.rs:381-392) on the
 * device.  Resets the Adam state (optimizer.rs:34-38).                          */
int ep_set_table(ep_ctx* ctx, const float* rows);

/* Copies the FP32 table (FP32 master rows in BF16 mode) to rows[n_vocab*width]. */
int ep_get_table(ep_ctx* ctx, float* rows);
/* Copies the BF16 working table (BF16 mode only) to rows[n_vocab*width].        */
int ep_get_table_bf16(ep_ctx* ctx, uint16_t* rows);
/* Copies the Adam moments (optimizer.rs:29-30) to m, v [n_vocab*width].         */
int ep_get_adam(ep_ctx* ctx, float* m, float* v);

/* Sets the train / validation node ids.  ep_learn draws its own split
 * (mod.rs:119-122) when this has not been called.                               */
int ep_set_split(ep_ctx* ctx, const uint32_t* train, uint64_t n_train, const uint32_t* valid, uint64_t n_valid);

/* The training loop of learn_feature_embeddings (mod.rs:92-287): passes x
 * ceil(n_train/B) steps of sample -> forward -> backward -> merge -> Adam, with
 * the cosine LR schedule and per-pass validation.                               */
int ep_learn(ep_ctx* ctx, ep_progress_fn progress, void* user);

/* Test hook: the sampler of one batch (model.rs:354-377, node_sampler.rs:185-188)
 * for the given anchors at (pass, position base).  nbr_cnt[n], nbr[n*S],
 * neg[n*K].  Requires max_nodes Fixed.                                          */
int ep_sample_batch(ep_ctx* ctx, const uint32_t* anchors, uint32_t n, uint32_t pass, uint32_t* nbr_cnt,
                    uint32_t* nbr, uint32_t* neg);

/* Test hook: the neighbour weights w_j (model.rs:357-361) of the last
 * ep_sample_batch, w[n*S] (weighted_neighbor_averaging with max_nodes Fixed).   */
int ep_sample_weights(ep_ctx* ctx, uint32_t n, float* w);

/* Test hook: neighbour weights w[n*S] for the next ep_step_explicit calls
 * (needed with weighted_neighbor_averaging).                                    */
int ep_set_explicit_weights(ep_ctx* ctx, uint32_t n, const float* w);

/* Test hook: one training step on explicit samples (the body of the per-batch
 * closure, mod.rs:176-252).  anchors[n], nbr_cnt[n], nbr[n*S] (S = max_nodes),
 * neg[n*K].  lr and pass are given.  Outputs (each nullable):
 *   loss[n]                per-example weighted loss (0 when no gradient)
 *   cons[n*(K+2)*dims]     constructions h(v), ~h(v), h(u_1..K)
 *   n_unique, feat[], grad[]  merged batch gradient (mod.rs:223-233); feat and
 *                          grad must hold up to n_vocab entries / rows of the
 *                          row width
 * apply_update != 0 runs the Adam update (optimizer.rs:43-81).                  */
int ep_step_explicit(ep_ctx* ctx, const uint32_t* anchors, uint32_t n, const uint32_t* nbr_cnt,
                     const uint32_t* nbr, const uint32_t* neg, uint32_t pass, float lr, int apply_update,
                     float* loss, float* cons, uint32_t* n_unique, uint32_t* feat, float* grad);

/* No counterpart in the reference: for callers that cannot allocate device memory
 * (serving loops, harnesses that forbid hipMalloc).  ep_step_explicit with
 * apply_update = 1 and only the loss output, whose per-step arrays live in a 64 MiB
 * array the library holds on each GPU from its load until the process ends, instead of
 * buffers the context allocates: the call itself allocates nothing (no hipMalloc,
 * hipFree or allocating library routine).  Host arrays as in ep_step_explicit; loss[n]
 * nullable.  Needs the sort scatter, max_nodes Fixed (or the ppr loss) and a table set
 * beforehand (ep_set_table); a step needing more than 64 MiB of arrays returns
 * EP_ECAPACITY with the size in ep_last_error().  Synchronous.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (from several threads or
 * contexts) race.                                                                  */
int ep_step_explicit_noalloc(ep_ctx* ctx, const uint32_t* anchors, uint32_t n, const uint32_t* nbr_cnt,
                             const uint32_t* nbr, const uint32_t* neg, uint32_t pass, float lr, float* loss);

/* Measurement hook: runs `warmup` then `steps` training steps from the current
 * split with everything resident in HBM, timed with HIP events on the context
 * stream.  Per-kernel timings are taken in a separate instrumented run of the
 * same number of steps when instrument != 0.                                    */
int ep_bench_steps(ep_ctx* ctx, uint32_t warmup, uint32_t steps, int instrument, ep_bench_result* out);

/* NodeEmbedder(FeatureAggregator.Averaged).embed_feature_set (src/lib.rs:2092-2124,
 * aggregator.rs:34-47): out[n_nodes*width] = mean of each node's feature rows.  */
int ep_embed_averaged(ep_ctx* ctx, uint64_t n_nodes, const uint64_t* f_indptr, const uint32_t* f_ids, float* out);

/* NodeEmbedder(FeatureAggregator.Weighted(alpha, fs)).embed_feature_set
 * (src/lib.rs:2092-2124, 2019-2021; WeightedAggregator, aggregator.rs:82-116):
 * out[n_nodes*width] = sum_f w_f E_f / sum_f w_f over each node's listed features,
 * w_f = alpha / (alpha + p_w[f]), p_w[n_p] the unigram probabilities
 * (UnigramProbability, aggregator.rs:50-79); nodes without features get a zero row. */
int ep_embed_weighted(ep_ctx* ctx, uint64_t n_nodes, const uint64_t* f_indptr, const uint32_t* f_ids,
                      const float* p_w, uint32_t n_p, float alpha, float* out);

/* NodeEmbedder(FeatureAggregator.Attention(heads, d_k, None)).embed_feature_set
 * (src/lib.rs:2092-2124, 2028-2036; AttentionAggregator, aggregator.rs:124-154):
 * out[n_nodes*dims] = Full attention over each node's listed features, count 1
 * each, with the context's attention_heads / attention_dk; at most 64 features
 * per node; nodes without features get a zero row.                             */
int ep_embed_attention(ep_ctx* ctx, uint64_t n_nodes, const uint64_t* f_indptr, const uint32_t* f_ids, float* out);

/* NeighborhoodAligner (src/algos/alignment.rs:25-75; lib.rs:3040-3171 align and
 * align_to_disk).  Create uploads the graph (CumCSR offsets, columns and per-row
 * edge CDF) and the node embeddings emb[n_nodes*dims] (row i is graph node i);
 * run writes out[count*dims] for nodes [first, first+count):
 *   out_v = (1 - a) * sum_j p_j emb[u_j] + a * emb[v]
 * over the first min(deg, max_neighbors) edges (0 = all), p_j = CDFtoP, a = alpha,
 * or min(1 / ln(deg + 1), 1) when alpha is NaN (the reference's None).            */
typedef struct ep_aligner ep_aligner;
int ep_aligner_create(int device, uint64_t n_nodes, const uint64_t* indptr, const uint32_t* indices, const float* cdf,
                      uint32_t dims, const float* emb, ep_aligner** out);
int ep_aligner_run(ep_aligner* a, float alpha, uint32_t max_neighbors, uint64_t first, uint64_t count, float* out);
void ep_aligner_destroy(ep_aligner* a);
/* Measurement hook: duration (ms, HIP events) of the last run's alignment kernel.  */
float ep_aligner_last_kernel_ms(const ep_aligner* a);

/* Graph file loader (host): GraphReader::load (src/io.rs:270-323) with
 * CSR::construct_from_edges (graph.rs:70-142) and CumCSR::convert (graph.rs:414-422).
 * Lines "from_type\tfrom\tto_type\tto\tweight" (gzip for a ".gz" path); the first
 * skip_rows lines are skipped; undirected adds every edge reversed; weighted parses the
 * weight (else 1); deduplicate sorts by (from, to) and adds repeated edges' weights.
 * Node ids follow first appearance.  sizes, then copy: indptr[n+1], indices[E], the
 * per-row CDF cdf[E] and names = "type\tname\n" per node in id order.  Errors:
 * a negative EP_E* code, message in ep_graph_file_error() (the message of the last
 * failed call of either file loader on this thread).                       */
typedef struct ep_graph_file ep_graph_file;
int ep_graph_load_tsv(const char* path, int undirected, int weighted, int deduplicate, uint64_t skip_rows,
                      ep_graph_file** out);
int ep_graph_file_sizes(const ep_graph_file* g, uint64_t* n_nodes, uint64_t* n_entries, uint64_t* name_bytes);
int ep_graph_file_copy(const ep_graph_file* g, uint64_t* indptr, uint32_t* indices, float* cdf, char* names);
void ep_graph_file_free(ep_graph_file* g);
const char* ep_graph_file_error(void);

/* Feature file loader (host): FeatureSet.load_into (src/lib.rs:1579-1606) with
 * FeatureStore::set_features (src/feature_store.rs:36-48).  Lines
 * "node_type\tname\tf1 f2 ..." (gzip for a ".gz" path), exactly three tab-separated
 * fields, else an error naming the line; tokens split on Unicode whitespace
 * (str::split_whitespace).  Each token's (type, name) follows FeatureNamespace::parse
 * (lib.rs:1391-1403): EP_FNS_STATIC (ns_value, token), EP_FNS_NODE_TYPE (node_type,
 * token), EP_FNS_PREFIX split at the first ns_value (name from one byte past its start),
 * ("feat", token) when the prefix does not occur.  Lines of nodes missing from the node
 * table are ignored, tokens included; a later line of a node replaces an earlier one.
 * Name tables are a byte blob with offsets: entry k is (blob[off[2k], off[2k+1]),
 * blob[off[2k+1], off[2k+2])).  Known features keep their ids; new ones get ids
 * n_feats, n_feats + 1, ... in first-seen order.  sizes, then copy: the set nodes
 * (ascending), their lists as indptr[n_set + 1] / ids, the new feature names with
 * new_off[2 n_new + 1].  Errors: a negative EP_E* code, message in
 * ep_graph_file_error().                                                        */
#define EP_FNS_STATIC 0
#define EP_FNS_NODE_TYPE 1
#define EP_FNS_PREFIX 2
typedef struct ep_feature_file ep_feature_file;
int ep_features_load_tsv(const char* path, const char* node_blob, const uint64_t* node_off, uint64_t n_nodes,
                         const char* feat_blob, const uint64_t* feat_off, uint64_t n_feats, int ns_kind,
                         const char* ns_value, ep_feature_file** out);
int ep_feature_file_sizes(const ep_feature_file* f, uint64_t* n_set, uint64_t* n_ids, uint64_t* n_new,
                          uint64_t* new_name_bytes);
int ep_feature_file_copy(const ep_feature_file* f, uint32_t* nodes, uint64_t* indptr, uint32_t* ids,
                         char* new_blob, uint64_t* new_off);
void ep_feature_file_free(ep_feature_file* f);

/* Embedding file loader (host): NodeEmbeddings.load (src/lib.rs:2868-2894) with
 * EmbeddingReader::load (src/io.rs:135-200).  Lines "node_type\tname\t[v1,v2,...]"
 * (gzip for a ".gz" path); the first skip_rows lines are skipped, then, when filter is
 * set, lines whose text before the first tab is not one of the n_types types (blob
 * type_blob, entry k = [type_off[k], type_off[k+1])) are dropped.  A line is three tab
 * fields; its third loses its first and last byte and splits at ',' into values trimmed
 * of Unicode whitespace, each a whole f32 (fast_float::parse); otherwise "Error parsing
 * line: ..." (EP_EINVAL).  A repeated node is EP_EDUPLICATE ("found duplicate node at
 * i!"); a row of another width EP_EINVAL ("Embeddings have different sizes!").  sizes,
 * then copy: the table (rows x dims, file order) and the node names as a blob with
 * name_off[2 rows + 1] (entry k = (type, name)).  Messages in ep_graph_file_error().    */
typedef struct ep_embedding_file ep_embedding_file;
int ep_embeddings_load_tsv(const char* path, const char* type_blob, const uint64_t* type_off, uint64_t n_types,
                           int filter, uint64_t skip_rows, ep_embedding_file** out);
int ep_embedding_file_sizes(const ep_embedding_file* f, uint64_t* n_rows, uint64_t* dims, uint64_t* name_bytes);
int ep_embedding_file_copy(const ep_embedding_file* f, float* table, char* name_blob, uint64_t* name_off);
void ep_embedding_file_free(ep_embedding_file* f);

/* Embedding file writer (host): NodeEmbeddings.save (src/lib.rs:2828-2840) through
 * EmbeddingWriter::stream (src/io.rs:24-78) and open_file_for_writing (io.rs:212-222).
 * One line per row i: the bytes name_blob[name_off[i] .. name_off[i+1]) ("node_type\tname"),
 * a tab and "[v1,v2,...]" with each value as ryu::Buffer::format prints an f32: the
 * shortest digits that read back to the same value, "0.000001" down to 1e-6, "1e-7" and
 * "1e13" in exponent form, integers up to 13 digits as "123.0", "NaN", "inf".  gzip for a
 * ".gz" path at comp_level (negative: level 1, Compression::fast()).  Errors: a negative
 * EP_E* code, message in ep_graph_file_error().                                       */
int ep_embeddings_write_tsv(const char* path, int comp_level, uint64_t n_rows, uint64_t dims, const float* table,
                            const char* name_blob, const uint64_t* name_off);
/* Test hook: one value as the writer prints it into out (at least 24 bytes), terminated;
 * returns its length.                                                                   */
int ep_format_f32(float x, char* out);

/* Philox4x32-10 on the host (the device RNG), for known-answer tests.           */
int ep_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Thread-local message of the last failure ("" if none).                        */
const char* ep_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CLOVERLEAF_EP_H */
