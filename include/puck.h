/*
 * puck-mi355x C ABI: the drop-in boundary of the CKY parse path.
 *
 * Every entry point returns 0 on success and a negative status on failure; the message of the
 * last failure on the calling thread is returned by puck_last_error(). Host buffers belong to the
 * caller, device memory belongs to the library. A parser handle is internally serialised (the
 * reference's `synchronized` methods); distinct handles are independent and own one HIP stream
 * each.
 *
 * Reference interfaces replaced (Scala/JavaCL, citations relative to the reference checkout):
 *   CLParserData.make / readSequence      src/main/scala/puck/parser/CLParser.scala:1264-1305
 *   new CLParser(data, maxAllocSize)      src/main/scala/puck/parser/CLParser.scala:46-48
 *   CLParser.parse                        src/main/scala/puck/parser/CLParser.scala:50-53
 *   CLParser.partitions                   src/main/scala/puck/parser/CLParser.scala:55-58
 *   CLParser.insideOutside                src/main/scala/puck/parser/CLParser.scala:60-63
 *   ActualParser.updateMasks (coarse)     src/main/scala/puck/parser/CLParser.scala:208-229
 */
#ifndef PUCK_H_
#define PUCK_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct puck_grammar puck_grammar_t;
typedef struct puck_parser puck_parser_t;

/* RuleSemiring.scala:22-98: Viterbi (max, +), Real (+, x) with scaling, LogSum (log-add, +). */
enum puck_semiring { PUCK_VITERBI = 0, PUCK_REAL = 1, PUCK_LOGSUM = 2 };

/* Status codes. */
enum puck_status {
  PUCK_OK = 0,
  PUCK_ERR_ARG = -1,      /* bad argument */
  PUCK_ERR_IO = -2,       /* grammar file missing or malformed */
  PUCK_ERR_DEVICE = -3,   /* HIP runtime error */
  PUCK_ERR_NOMEM = -4,    /* allocation failure (host or device) */
  PUCK_ERR_INTERNAL = -5
};

/* Pre-order tree node of a Viterbi parse, 2n-1 per sentence (CLViterbi.scala:378 tree_t). */
typedef struct {
  int32_t top_sym; /* nonterminal id of the unary top, -1 when the node skips the unary layer */
  int32_t bot_sym; /* nonterminal id (width > 1) or terminal id (width == 1) */
  int32_t width;
  float score;
} puck_tree_node_t;

/* Pre-order node of an MBR (max-rule) parse with coarse labels (CLParser.scala:660-683). */
typedef struct {
  int32_t top_coarse; /* -1 when the unary layer is not used */
  int32_t bot_coarse;
  int32_t begin, end;
  float score;
} puck_mbr_node_t;

typedef struct {
  int32_t num_nonterms, num_terms, num_coarse, cell_size, mask_words, root;
  int32_t num_binary[4]; /* NN, NT, TN, TT */
  int32_t num_unary_n, num_unary_t, num_words, semiring;
  int32_t num_blocks[4]; /* CoarseParent blocks of the inside families */
  int32_t rules_dropped;
} puck_grammar_info_t;

typedef struct {
  double rule_kernel_ms;  /* summed device time of the rule-family kernels */
  double reduce_ms, other_ms, total_ms;
  double rule_apps;       /* rule applications executed by the rule kernels */
  int64_t rule_launches;
  double device_ms;       /* device time from input upload to the first result download */
} puck_stats_t;

/* Thread-local message of the last failure ("" if none). */
const char* puck_last_error(void);
int32_t puck_abi_version(void);

/* ---- grammars (CLParserData.make over SimpleGrammar.parseBerkeleyText(prefix, -12, ...)) ---- */
int32_t puck_grammar_load_text(const char* prefix, int32_t semiring, puck_grammar_t** out);
void puck_grammar_free(puck_grammar_t* g);
/* CLParser.reprojectGrammar + readHierarchy (CLParser.scala:1153-1208), the --reproject option of
 * CompileGrammar and CLParser.main (CompileGrammar.scala:57-59, CLParser.scala:1027-1029): new
 * handles for a coarse/fine pair whose coarse symbols are the coarse grammar's refined symbols;
 * each fine symbol projects to its ancestor at the coarse split level in the hierarchy, and fine
 * rules whose projection is not a coarse rule are scored fromLogSpace(-12). hierarchy_path NULL:
 * the fine prefix + ".hierarchy"; a level < 0: the last number in that grammar's prefix. The inputs
 * are not modified. */
int32_t puck_grammar_reproject(const puck_grammar_t* coarse, const puck_grammar_t* fine, const char* hierarchy_path,
                               int32_t coarse_level, int32_t fine_level, puck_grammar_t** coarse_out,
                               puck_grammar_t** fine_out);
/* CLParserData.writeSequence / readSequence (CLParser.scala:1249-1317) with ZipUtil
 * (ZipUtil.scala:13-98): the compiled layers in one zip ("grammar.grz") with the reference's entry
 * names (parserNames, parser-i/grammar, parser-i/structure, parser-i/semiring) and binary table
 * payloads of this library. load returns the number of layers and writes that many handles when
 * cap is large enough (otherwise it only returns the count). */
int32_t puck_grammar_save_grz(const puck_grammar_t* const* layers, int32_t n_layers, const char* path);
int32_t puck_grammar_load_grz(const char* path, puck_grammar_t** out, int32_t cap);
int32_t puck_grammar_info(const puck_grammar_t* g, puck_grammar_info_t* out);
/* Word id of `word` in the grammar lexicon, -1 if unknown. */
int32_t puck_grammar_word_id(const puck_grammar_t* g, const char* word);
/* Rule tables for cross-checks. family: 0..3 binary NN/NT/TN/TT (p,l,r), 4 unary N, 5 unary T
 * (p,c,-). Writes up to `cap` rows of 3 ints plus score w and kernel literal q. Returns the row
 * count (or a negative status). */
int32_t puck_grammar_rules(const puck_grammar_t* g, int32_t family, int32_t* idx, float* w, float* q,
                           int32_t cap);
/* Symbol table: coarse id and substate of nonterminal (terminal=0) or terminal (terminal=1) ids. */
int32_t puck_grammar_symbols(const puck_grammar_t* g, int32_t terminal, int32_t* coarse, int32_t* sub,
                             int32_t cap);
/* Name of coarse symbol `coarse` (.numstates line order), NUL-terminated into buf[cap]; returns its
 * length. */
int32_t puck_grammar_coarse_name(const puck_grammar_t* g, int32_t coarse, char* buf, int32_t cap);
/* CoarseParent blocks: which = 0 inside, 1 outside-left, 2 outside-right; returns #blocks and
 * writes block_of_coarse[num_coarse]. */
int32_t puck_grammar_blocks(const puck_grammar_t* g, int32_t which, int32_t family,
                            int32_t* block_of_coarse);
/* Lexicon of one word: writes (terminal id, ln P) pairs, returns count. */
int32_t puck_grammar_lexicon(const puck_grammar_t* g, int32_t word, int32_t* tags, double* logp,
                             int32_t cap);

/* ---- seeded inputs ---- */
/* Synthetic grammar with the wsj_6 shape; writes out_prefix.{numstates,binary,unary,lexicon,hierarchy}. */
int32_t puck_synth_wsj6(const char* data_dir, const char* out_prefix, uint64_t seed);
/* PCFG sentence sampling; mode 0 range filter, 1 log-normal lengths, 2 mixed long tail.
 * words_out needs capacity words_cap; offsets_out needs n+1 entries. */
int32_t puck_sample_sentences(const puck_grammar_t* g, uint64_t seed, int32_t n, int32_t mode,
                              int32_t min_len, int32_t max_len, int32_t* words_out, int64_t words_cap,
                              int32_t* offsets_out);

/* ---- parser (CLParser over coarse-to-fine layers; the last layer decides Viterbi vs MBR) ---- */
/* Flags of puck_parser_create. */
enum puck_parser_flags {
  PUCK_FLAG_ONE_ITEM_PER_LANE = 1, /* rule kernels compute one (cell, split) item per lane, not two */
  PUCK_FLAG_PERSISTENT = 2,        /* rule families of a span share one persistent work-queue launch */
  PUCK_FLAG_GRAPHS = 4,            /* parse calls capture each batch shape's device work into a HIP
                                      graph on its second sighting and replay it afterwards */
  PUCK_FLAG_NO_PLAN_AHEAD = 8,     /* accepted and ignored: every parser issues all of its device work,
                                      mask plans included, on its one HIP stream */
  PUCK_FLAG_OUTSIDE_FINAL = 16,    /* puck_parse_viterbi also runs the final layer's outside pass before
                                      the decode, as CLParser.insideOutside does
                                      (CLParser.scala:171-182); the trees do not change */
  PUCK_FLAG_DEVICE_SCHEDULE = 32,  /* each pass runs as one persistent launch that pulls the rule and
                                      reduce work of all its spans from a device queue, a span of a
                                      sentence starting once that sentence's shorter spans are done
                                      (per-block completion counters); would replace the span loop of
                                      CLParser.scala:284-291,331-342 on the device. DISABLED: rejected
                                      with PUCK_ERR_ARG (the launch did not finish on the GPU and no
                                      cause has been shown) */
  PUCK_FLAG_SLOT_PARTIALS = 64     /* Viterbi layers: the rule kernels write one partial row per
                                      (operand chunk, target) slot with plain stores, as Real and
                                      LogSum layers always do, instead of folding every slot of a
                                      target into one row with an integer atomic max (the default; the
                                      reference's Viterbi kernels accumulate with atomic_min on the
                                      score bits, SimpleGenRuleMultiply.java:326-350). Results are
                                      identical; the flag exists for comparison */
};
/* new CLParser(data, maxAllocSize) (CLParser.scala:46-48): mem_bytes bounds the device workspace,
 * as the reference's --mem option bounds WorkSpace.allocate (WorkSpace.scala:91-120). */
int32_t puck_parser_create(const puck_grammar_t* const* layers, int32_t n_layers, int32_t device,
                           int64_t mem_bytes, int32_t flags, puck_parser_t** out);
/* Launch shape of the rule kernels (the reference's equivalent knobs are the work-group shape and
 * rule-segment size fixed at kernel generation, JavaFriendlyGenRuleMultiply.scala:57-65): plain
 * launches cut every wave's rule stream into pieces while tiles x chunks x 2 x pieces <=
 * piece_target; persistent launches run queue_wgs_per_cu workgroups per CU. */
typedef struct {
  int32_t persistent;
  int32_t piece_target;     /* >= 1, default 512 */
  int32_t queue_wgs_per_cu; /* 1..8, default 1 */
  int32_t reduce_targets;   /* chart targets (waves) per reduce workgroup: 1, 2, 4, 8 or 16
                               (64-1,024 threads), default 4 */
} puck_launch_config_t;
int32_t puck_parser_configure(puck_parser_t* p, const puck_launch_config_t* cfg);
void puck_parser_free(puck_parser_t* p);
/* Sentences are word ids: sentence i is words[offsets[i] .. offsets[i+1]). */
int32_t puck_parse_viterbi(puck_parser_t* p, const int32_t* words, const int32_t* offsets, int32_t n,
                           puck_tree_node_t* nodes, int64_t nodes_cap);
int32_t puck_parse_mbr(puck_parser_t* p, const int32_t* words, const int32_t* offsets, int32_t n,
                       puck_mbr_node_t* nodes, int64_t nodes_cap);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): puck_parse_viterbi's result
 * with no device allocation. Every workspace buffer (charts, scales, masks, mask plans, the rule
 * kernels' partial rows and counters) is carved from a static per-device array of 256 MiB that the
 * library holds until the process ends; an input whose workspace would exceed it is rejected with
 * PUCK_ERR_ARG before any device work. Calls that overlap on one GPU share that array and race (the
 * handle's own lock serialises calls on one handle only). Viterbi final layers only. */
int32_t puck_parse_viterbi_noalloc(puck_parser_t* p, const int32_t* words, const int32_t* offsets, int32_t n,
                                   puck_tree_node_t* nodes, int64_t nodes_cap);
int32_t puck_partitions(puck_parser_t* p, const int32_t* words, const int32_t* offsets, int32_t n,
                        float* out);
int32_t puck_inside_outside(puck_parser_t* p, const int32_t* words, const int32_t* offsets, int32_t n);
/* Pruning masks of the last coarse layer: mask_words uint32 per chart cell, cells of sentence i at
 * [sum_{j<i} n_j(n_j+1), ...), bottom half then top half, span-diagonal-major. */
int32_t puck_masks(puck_parser_t* p, const int32_t* words, const int32_t* offsets, int32_t n,
                   uint32_t* out, int64_t cap);
/* Charts of layer `layer` after its passes on one batch (parity tests). Sizes are
 * cells*cell_size floats for the charts and cells for the scales; masks (cells*mask_words) are the
 * input masks of that layer when it has any. partition_scales has n entries. */
int32_t puck_debug_charts(puck_parser_t* p, const int32_t* words, const int32_t* offsets, int32_t n,
                          int32_t layer, int32_t with_outside, float* inside, float* outside,
                          float* in_scale, float* out_scale, uint32_t* masks,
                          double* partition_scales, int64_t chart_cap);
int32_t puck_parser_stats(const puck_parser_t* p, puck_stats_t* out);
/* Per rule-kernel launch of the last call: duration (ms), rule applications (under pruning masks:
 * the rules of the active blocks per planned item), and family id + 16 x layer index (families
 * 0-3 inside NN/NT/TN/TT, 4-7 outside-left, 8-11 outside-right, 12-15 unary). Returns the count. */
int64_t puck_parser_launches(const puck_parser_t* p, float* ms, double* apps, int32_t* family, int64_t cap);

/* ---- batched engine: many sentences of uneven length in one call ----
 * Replaces the batch loop of CLParser.parse (WorkSpace.getBatches, WorkSpace.scala:20-41, then
 * one batch at a time, CLParser.scala:145-206): sentences are sorted by length and cut into
 * buckets of bucket_cells chart cells; buckets run concurrently on `streams` HIP streams (one
 * workspace of mem_bytes / streams each), with persistent rule launches and, when `graphs` is
 * set, each bucket shape's device work captured into a HIP graph on its second sighting and
 * replayed after that. Results are identical to puck_parse_viterbi / puck_parse_mbr. */
typedef struct puck_batch puck_batch_t;
typedef struct {
  int32_t streams;        /* 1..8 */
  int64_t bucket_cells;   /* chart cells (sum of n(n+1)) per bucket, >= 1 */
  int32_t persistent;     /* persistent rule launches */
  int32_t graphs;         /* per-shape HIP graphs */
  int32_t piece_target;   /* see puck_launch_config_t */
  int32_t items_per_lane; /* 1 or 2 */
  int32_t outside_final;  /* Viterbi final layer: outside pass before the decode (insideOutside) */
  int32_t device_schedule; /* must be 0: see PUCK_FLAG_DEVICE_SCHEDULE (disabled) */
  int32_t reduce_targets; /* see puck_launch_config_t */
  int32_t slot_partials;  /* see PUCK_FLAG_SLOT_PARTIALS (0: fold, the default) */
} puck_batch_config_t;
typedef struct {
  double device_ms; /* first kernel of the first bucket to the last kernel of the last bucket */
  int64_t buckets, graph_captures, graph_replays;
} puck_batch_stats_t;
int32_t puck_batch_create(const puck_grammar_t* const* layers, int32_t n_layers, int32_t device, int64_t mem_bytes,
                          const puck_batch_config_t* cfg, puck_batch_t** out);
void puck_batch_free(puck_batch_t* b);
int32_t puck_batch_parse_viterbi(puck_batch_t* b, const int32_t* words, const int32_t* offsets, int32_t n,
                                 puck_tree_node_t* nodes, int64_t nodes_cap);
int32_t puck_batch_parse_mbr(puck_batch_t* b, const int32_t* words, const int32_t* offsets, int32_t n,
                             puck_mbr_node_t* nodes, int64_t nodes_cap);
int32_t puck_batch_stats(puck_batch_t* b, puck_batch_stats_t* out);

/* ---- serving (replaces BatchFunctionAnnotatorService, AnnotatorService.scala:29-78) ----
 * Continuous batching: up to max_inflight batches run concurrently (one HIP stream each); a free
 * slot admits queued requests in arrival order while the device memory their batch needs fits what
 * the slot holds plus the device's free memory. A slot parses its batch in two length classes, at
 * most 40 words first, and publishes each class's results when it is done; `batches` counts these
 * parse calls. flags: PUCK_FLAG_ONE_ITEM_PER_LANE and PUCK_FLAG_PERSISTENT as for
 * puck_parser_create. */
typedef struct puck_server puck_server_t;
typedef struct {
  int64_t completed, batches;
  double mean_batch, p50_ms, p90_ms, p99_ms, max_ms;
} puck_server_stats_t;
int32_t puck_server_create(const puck_grammar_t* const* layers, int32_t n_layers, int32_t device,
                           int64_t mem_bytes, int32_t max_inflight, int32_t max_batch, int32_t flags,
                           puck_server_t** out);
void puck_server_free(puck_server_t* s);
/* Enqueues one sentence (word ids); writes its ticket. Thread-safe. */
int32_t puck_server_submit(puck_server_t* s, const int32_t* words, int32_t len, int64_t* ticket);
/* Result of a ticket: returns 1 when done (nodes written: 2n-1 tree nodes, Viterbi final layer, or
 * MBR nodes for a Real final layer, cast to the matching struct), 0 when still pending (wait = 0),
 * negative on failure. */
int32_t puck_server_result(puck_server_t* s, int64_t ticket, int32_t wait, void* nodes, int64_t cap);
int32_t puck_server_flush(puck_server_t* s);
int32_t puck_server_stats(puck_server_t* s, int32_t reset, puck_server_stats_t* out);

#ifdef __cplusplus
}
#endif

#endif /* PUCK_H_ */
