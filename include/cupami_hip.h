/*
 * cupami_hip.h -- C ABI of libcupami_hip.so, the gfx950 support-counting path.
 *
 * Plain C types only.  Every function returns an int status (CPM_OK == 0,
 * negative on error); cpm_last_error() returns the calling thread's message
 * for the last failure.  Device memory belongs to the handle that allocated
 * it and is released by the matching *_destroy.  Host output buffers belong
 * to the caller.  A context is bound to one device and one HIP stream; use a
 * context from one host thread at a time.  Several contexts of one device
 * may mine the same store concurrently once its build has completed on its
 * own context (cpm_ctx_synchronize of that context); a store is destroyed
 * from the thread of the context that built it, when no search uses it.
 *
 * Items are addressed by COLUMN: the index of the item's label in string
 * order (the reference sorts labels as strings, cuApriori.py:397).  Rows are
 * transaction indexes 0..n_transactions-1.
 */
#ifndef CUPAMI_HIP_H
#define CUPAMI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes; the Python layer maps them to the reference's exception types */
#define CPM_OK              0
#define CPM_ERR_INVALID    -1   /* ValueError: bad argument */
#define CPM_ERR_RUNTIME    -2   /* RuntimeError: HIP call failed */
#define CPM_ERR_NOMEM      -3   /* MemoryError: device allocation failed */
#define CPM_ERR_NODEVICE   -4   /* RuntimeError: no usable gfx950 device */
#define CPM_ERR_LAYOUT     -5   /* input the device reader does not take: read it on the host */

typedef struct cpm_ctx cpm_ctx;
typedef struct cpm_bits cpm_bits;
typedef struct cpm_patterns cpm_patterns;
typedef struct cpm_fptree cpm_fptree;
typedef struct cpm_baskets cpm_baskets;

/* ABI version (major*10000 + minor*100 + patch). */
int cpm_version(void);
/* Message of the last failing call on this thread ("" if none). */
const char *cpm_last_error(void);
/* Number of visible HIP devices (0 when none); never fails. */
int cpm_device_count(void);

/* Replaces cuApriori._getCudaModule (cuApriori.py:458-471): binds a device,
 * creates its stream.  Fails with CPM_ERR_NODEVICE when none is present. */
int cpm_ctx_create(int device, cpm_ctx **out);
int cpm_ctx_destroy(cpm_ctx *ctx);
/* Block until all work queued on the context's stream is done. */
int cpm_ctx_synchronize(cpm_ctx *ctx);
/* The context's hipStream_t, as an opaque pointer. */
void *cpm_ctx_stream(cpm_ctx *ctx);

/* ---- vertical bitset store ------------------------------------------- */

/* Replaces _bitsets.build_bitsets (_bitsets.py:91-100): sets bit `rows[i]` of
 * item `cols[i]` for every pair (duplicates allowed).  The store holds one
 * row of ceil(n_transactions/64) uint64 words per item; bits past
 * n_transactions are zero.  Host pointers. */
int cpm_bits_build(cpm_ctx *ctx, const uint64_t *rows, const uint32_t *cols, uint64_t nnz,
                   uint64_t n_transactions, uint32_t n_items, cpm_bits **out);
/* Same, with rows/cols already in device memory (no host copies). */
int cpm_bits_build_device(cpm_ctx *ctx, const uint64_t *d_rows, const uint32_t *d_cols,
                          uint64_t nnz, uint64_t n_transactions, uint32_t n_items,
                          cpm_bits **out);
/* Many stores at once from device COO (one (rows, cols, nnz, n_transactions,
 * n_items) per database): _bitsets.build_bitsets (_bitsets.py:91-100) for a
 * batch of databases mined together.  The stores share one device
 * allocation and the batch is checked with two host synchronizations in
 * all.  out[i] receives store i; each is destroyed with cpm_bits_destroy.
 * On error no store is returned. */
int cpm_bits_build_device_batch(cpm_ctx *ctx, uint32_t n_db, const uint64_t *const *d_rows,
                                const uint32_t *const *d_cols, const uint64_t *nnz,
                                const uint64_t *n_transactions, const uint32_t *n_items, cpm_bits **out);
/* cpm_bits_build_device_batch with the pairs in host memory (each
 * database's rows/cols arrays, as cpm_bits_build takes them): the pairs are
 * copied into one device staging area first.  For a serving loop that
 * receives requests from the host. */
int cpm_bits_build_batch(cpm_ctx *ctx, uint32_t n_db, const uint64_t *const *rows, const uint32_t *const *cols,
                         const uint64_t *nnz, const uint64_t *n_transactions, const uint32_t *n_items,
                         cpm_bits **out);
/* Replaces the host side of the reference's BIN input (cuApriori.py:256-404):
 * the array checks of _readBinaryFile / _tryBinaryBulkRead
 * (cuApriori.py:273-276, 301-305), and the BIN branch of _getDevicePairs
 * (cuApriori.py:381-396) that finds the distinct ids, sorts them as strings
 * and maps every entry to its (row, column) for build_bitsets.  Input is a
 * basket file's arrays as stored (CUPAMI01, and the offsets and ids of
 * CUPAMIT1): offsets[n_baskets + 1] (basket i holds ids[offsets[i] ..
 * offsets[i+1])) and ids[n_ids], all ids <= universe.  Columns are the
 * distinct ids in decimal-string order ("10" < "2"); rows are the baskets in
 * file order, empty baskets left out unless keep_empty (the temporal readers
 * keep them, _temporal.py:125-139).  Repeated ids in a basket set one bit.
 * Fails with CPM_ERR_INVALID and the reference's message for offsets that
 * do not rise from 0 to n_ids ("Invalid BIN file: invalid transaction
 * offsets") and for ids above universe ("... item ID outside configured
 * universe"); then no store is returned.  cpm_bits_labels gives the id of
 * each column and cpm_bits_shape the row count.  Host pointers. */
int cpm_bits_build_csr(cpm_ctx *ctx, const uint64_t *offsets, uint64_t n_baskets, const uint32_t *ids,
                       uint64_t n_ids, uint32_t universe, int keep_empty, cpm_bits **out);
/* Same, with offsets/ids already in device memory. */
int cpm_bits_build_csr_device(cpm_ctx *ctx, const uint64_t *d_offsets, uint64_t n_baskets, const uint32_t *d_ids,
                              uint64_t n_ids, uint32_t universe, int keep_empty, cpm_bits **out);
/* Replaces the text side of the reference's input (_readDelimitedFile and
 * basket cleaning, cuApriori.py:428-450, and the GPU text reader it tries
 * first, cuApriori.py:223-247 with _gpu_io.py): delimited text (`text`,
 * host memory, n_bytes) parsed on the device into basket arrays kept on the
 * device.  Every line is a basket; the separator and '\n' end fields; fields
 * are stripped of ASCII whitespace; empty fields are dropped; every other
 * field must be a canonical decimal id (no sign or leading zero, at most
 * 2^32 - 1), its label being its own text.  Returns CPM_ERR_LAYOUT (no
 * baskets) for text the csv module would read differently (quotes, bytes
 * >= 0x80, NUL, a '\r' not followed by '\n' before the end) or with other
 * fields: the caller then reads it on the host. */
int cpm_baskets_parse_text(cpm_ctx *ctx, const char *text, uint64_t n_bytes, char sep, cpm_baskets **out);
/* Baskets (lines), ids, baskets with at least one id, largest id. */
int cpm_baskets_info(const cpm_baskets *b, uint64_t *n_baskets, uint64_t *n_ids, uint64_t *n_nonempty,
                     uint32_t *max_id);
/* offsets[n_baskets + 1] and ids[n_ids] to host memory (either may be NULL). */
int cpm_baskets_download(cpm_ctx *ctx, const cpm_baskets *b, uint64_t *offsets, uint32_t *ids);
/* cpm_bits_build_csr_device on the parsed baskets (universe: the largest id). */
int cpm_bits_build_baskets(cpm_ctx *ctx, const cpm_baskets *b, int keep_empty, cpm_bits **out);
int cpm_baskets_destroy(cpm_baskets *b);

/* The item id of each column of a store built by cpm_bits_build_csr*
 * (n_items values, host memory): the reference's `items` list
 * (cuApriori.py:384-385) as numbers. */
int cpm_bits_labels(cpm_ctx *ctx, const cpm_bits *bits, uint32_t *ids);
int cpm_bits_destroy(cpm_bits *bits);
/* words per row (ceil(N/64)), and the padded row stride in words */
int cpm_bits_shape(const cpm_bits *bits, uint64_t *n_transactions, uint32_t *n_items,
                   uint64_t *words, uint64_t *stride);
/* Copy the store to host: n_items rows of `words` uint64 each (no padding). */
int cpm_bits_download(cpm_ctx *ctx, const cpm_bits *bits, uint64_t *host);

/* ---- support operators ------------------------------------------------ */

/* Replaces _bitsets.count_supports (_bitsets.py:103-113): for each of the
 * n_cand candidates (width columns each, row-major in `cand`), the popcount
 * of the AND of its item rows.  Host pointers. */
int cpm_count_supports(cpm_ctx *ctx, const cpm_bits *bits, const uint32_t *cand,
                       uint64_t n_cand, uint32_t width, uint64_t *supports);

/* No counterpart in the reference: cpm_count_supports for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Candidates (`d_cand`, n_cand x
 * width u32 columns, row-major) and supports (`d_supports`, n_cand u64) are
 * DEVICE pointers; the task tables live in a fixed device array of the
 * library (about 17 MiB per device, held until the process ends), so the
 * call allocates nothing.  At most 65,536 candidates per call (the
 * reference's batch, cuEclat.py:185) and width <= 65; larger calls fail with
 * CPM_ERR_INVALID.  Synchronous, like cpm_count_supports: a column >= the
 * store's item count is reported as CPM_ERR_INVALID (supports then
 * unspecified).  Every call on one device shares that array: calls that
 * overlap on one device (several contexts or threads) race. */
int cpm_count_supports_noalloc(cpm_ctx *ctx, const cpm_bits *bits, const uint32_t *d_cand,
                               uint64_t n_cand, uint32_t width, uint64_t *d_supports);

/* Replaces _bitsets.intersect_extensions (_bitsets.py:116-125) with the keep
 * step cuEclat applies to its result (cuEclat.py:234-242) fused in:
 * child_j = prefix & ext_j for k extension rows of `words` uint64 each
 * (host arrays), supports[j] = popcount(child_j).  Children whose support is
 * >= keep_min are written, in extension order, to `kept_rows` (host,
 * up to k x words; NULL to count only) and their indexes to `kept_index`
 * (NULL allowed); *n_kept receives their count (NULL allowed).
 * keep_min = 0 keeps every child, as the reference returns them all. */
int cpm_intersect_extensions(cpm_ctx *ctx, const uint64_t *prefix, const uint64_t *ext,
                             uint32_t k, uint64_t words, uint64_t keep_min,
                             uint64_t *supports, uint64_t *kept_rows, uint32_t *kept_index,
                             uint32_t *n_kept);

/* Level-2 pair supports as a D^T D on the matrix cores (FP4 MFMA over 0/1
 * operands, exact: partial sums stay below 2^24 and are summed in int32):
 * gram[a*n + b] = |rows(cols[a]) & rows(cols[b])| for a <= b; entries below
 * the diagonal are left 0.  n <= 16384, n_transactions < 2^31.  Host output
 * of n*n uint32.  Replaces the level-2 count_supports over all pairs
 * (cuEclat.py:184-193). */
int cpm_pair_gram(cpm_ctx *ctx, const cpm_bits *bits, const uint32_t *cols, uint32_t n,
                  uint32_t *gram);

/* Replaces TemporalInput._evaluate's periodic_counts (_temporal.py:28-55,
 * 151-161): support and periodicity (largest gap of [0, ts of each
 * occurrence in row order, end]) of each candidate.  `timestamps` (host,
 * n_transactions int64, non-decreasing) is uploaded once per call. */
int cpm_periodic_eval(cpm_ctx *ctx, const cpm_bits *bits, const int64_t *timestamps,
                      int64_t end, const uint32_t *cand, uint64_t n_cand, uint32_t width,
                      uint64_t *supports, int64_t *periods);

/* ---- whole-search miners (device-resident search) --------------------- */

/* Replaces cuEclat.mine / cuApriori.mine's search (cuEclat.py:168-245,
 * cuApriori.py:501-535): every itemset whose support >= minsup, found by a
 * level-synchronous search whose candidate generation, pruning and support
 * counting all run on the device.  Level-2 supports come from
 * cpm_pair_gram when use_gram != 0. */
int cpm_eclat_mine(cpm_ctx *ctx, const cpm_bits *bits, uint64_t minsup, int use_gram,
                   cpm_patterns **out);

/* Many databases in one search: the levels of every database share each
 * launch (batched, persistent support kernel; one FP4 Gram launch for all
 * level-2 matrices).  bits[i] is mined at minsup[i]; the patterns of all
 * databases come back in one result, cpm_patterns_databases tells them
 * apart.  Results equal n_db separate cpm_eclat_mine calls. */
int cpm_eclat_mine_batch(cpm_ctx *ctx, const cpm_bits *const *bits, const uint64_t *minsup,
                         uint32_t n_db, int use_gram, cpm_patterns **out);

/* Search schedule of this context: entries per support work item (1..64),
 * words per tile (1024, the default, or 2048), replay of the level step as a hipGraph
 * (use_graph), and the subset test of candidates before counting. */
int cpm_set_search_options(cpm_ctx *ctx, uint32_t group, uint32_t tile_words, int use_graph,
                           int subset_prune);

/* Replaces cuPFEclat.mine (cuPFEclat.py:146-175): every itemset with
 * support >= minsup, support > 0 and periodicity <= maxper_floor (an integer
 * period p satisfies p <= m for a real m >= 0 iff p <= floor(m)).
 * timestamps: host, n_transactions non-decreasing int64. */
int cpm_pfeclat_mine(cpm_ctx *ctx, const cpm_bits *bits, const int64_t *timestamps, int64_t end,
                     uint64_t minsup, int64_t maxper_floor, cpm_patterns **out);

/* cpm_pfeclat_mine with the timestamps already in device memory (a caller
 * that keeps them resident, e.g. a serving loop mining one database
 * repeatedly); d_timestamps must stay valid until the call returns. */
int cpm_pfeclat_mine_device(cpm_ctx *ctx, const cpm_bits *bits, const int64_t *d_timestamps, int64_t end,
                            uint64_t minsup, int64_t maxper_floor, cpm_patterns **out);

/* ---- FP-tree (device arrays) -------------------------------------------- */

/* Replaces _fptree.unique_pairs (_fptree.py:97-105): distinct (row, col)
 * pairs in (row, col) order.  Host arrays; out_rows/out_cols hold up to nnz. */
int cpm_unique_pairs(cpm_ctx *ctx, const uint64_t *rows, const uint32_t *cols, uint64_t nnz,
                     uint64_t *out_rows, uint32_t *out_cols, uint64_t *out_n);

/* Replaces _fptree.build_tree (_fptree.py:108-177): weighted FP-tree of the
 * baskets given as (row, col) pairs with per-row weights (host arrays).
 * Node 0 is the root (parent -1, item -1); nodes of each depth follow,
 * ordered by (parent, item); baskets are ordered by (weighted support desc,
 * item id asc).  `supports` (item_count values, may be NULL) replaces the
 * support count and the infrequent-pair filter, as for conditional bases.
 * *out is set to NULL when no frequent pair remains (the reference's None). */
int cpm_fptree_build(cpm_ctx *ctx, const uint64_t *rows, const uint32_t *cols, uint64_t nnz,
                     const uint64_t *weights, uint64_t n_rows, uint32_t item_count,
                     uint64_t minsup, const uint64_t *supports, cpm_fptree **out);
/* Replaces _fptree.conditional_tree (_fptree.py:204-231); *out may be NULL. */
int cpm_fptree_conditional(cpm_ctx *ctx, cpm_fptree *tree, uint32_t item, uint64_t minsup,
                           cpm_fptree **out);
int cpm_fptree_info(const cpm_fptree *tree, uint64_t *n_nodes, uint32_t *depth,
                    int *single_path, uint32_t *header_len);
/* parents/items/counts: n_nodes each; header (items, supports): header_len
 * each, in ascending support order.  Any pointer may be NULL. */
int cpm_fptree_download(const cpm_fptree *tree, int64_t *parents, int64_t *items,
                        uint64_t *counts, uint32_t *header_items, uint64_t *header_supports);
int cpm_fptree_destroy(cpm_fptree *tree);

/* Replaces cuFPGrowth.mine (cuFPGrowth.py:161-213): unique pairs, unit
 * weights, tree build, then the depth-first search over conditional trees
 * (single-path trees enumerated directly).  minsup must be >= 1. */
int cpm_fpgrowth_mine(cpm_ctx *ctx, const uint64_t *rows, const uint32_t *cols, uint64_t nnz,
                      uint64_t n_transactions, uint32_t n_items, uint64_t minsup,
                      cpm_patterns **out);

/* ---- pattern results --------------------------------------------------- */

/* Patterns are a forest: pattern i = pattern parent[i] plus column item[i]
 * (parent == UINT32_MAX for singletons).  Parents always precede children. */
int cpm_patterns_count(const cpm_patterns *p, uint64_t *n);
/* Copy the forest to host arrays of length count (periods may be NULL; it is
 * filled only for periodic results). */
int cpm_patterns_download(const cpm_patterns *p, uint32_t *parent, uint32_t *item,
                          uint64_t *support, int64_t *period);
int cpm_patterns_destroy(cpm_patterns *p);
/* Database index of each pattern (0 for single-database results). */
int cpm_patterns_databases(const cpm_patterns *p, uint32_t *db);
/* The forest's own arrays (count elements each), valid until
 * cpm_patterns_destroy: period is NULL unless the result is periodic, db is
 * NULL when the result does not record databases, all are NULL for an empty
 * result.  No counterpart in the reference (it builds Python dicts); for
 * callers that read large results without copying them. */
int cpm_patterns_view(const cpm_patterns *p, const uint32_t **parent, const uint32_t **item,
                      const uint64_t **support, const int64_t **period, const uint32_t **db);
/* Replaces the key building of getPatterns / getPatternsAsDataFrame
 * (cuApriori.py:537-547, _temporal.py:170-176): the key of every pattern --
 * the labels of its columns in ascending column order joined by "\t" -- in
 * the result's order, separated by `sep`, into out[cap].  Labels are the
 * n_labels column labels as UTF-8, concatenated in `labels`, label i being
 * labels[label_off[i] .. label_off[i+1]).  *bytes receives the size of the
 * whole text; with out == NULL only the size is computed.  Host memory. */
int cpm_patterns_keys(const cpm_patterns *p, const char *labels, const uint64_t *label_off, uint32_t n_labels,
                      char sep, char *out, uint64_t cap, uint64_t *bytes);
/* Replaces save (cuApriori.py:548-552; periodic results _temporal.py:178-181):
 * writes "<key> : <support>\n" ("<key> : <support> : <period>\n" for a
 * periodic result) for every pattern to the file at `path`, in ascending
 * order of the label tuples (Python's sorted() over the reference's
 * _finalPatterns items).  Labels as for cpm_patterns_keys. */
int cpm_patterns_save(const cpm_patterns *p, const char *labels, const uint64_t *label_off, uint32_t n_labels,
                      const char *path);

/* ---- measurement -------------------------------------------------------- */

/* Totals since the last cpm_stats_reset on this context:
 * [0] support launches, [1] their kernel milliseconds (device clock stamps),
 * [2] algorithmic bitset bytes (8 * words per row read), [3] candidates
 * counted, [4] gram launches, [5] gram milliseconds (HIP events around each
 * launch), [6] gram ops (2 per multiply-add), [7] search levels.
 * Event times are read here: cpm_stats and cpm_stats_reset wait for the
 * context's stream when timed launches are outstanding. */
int cpm_stats(cpm_ctx *ctx, double *out8);
/* No counterpart in the reference (its BIN path has no device build to
 * time).  Totals since the last cpm_stats_reset of the store builds from
 * basket arrays: [0] builds, [1] ms of the tile build kernel, [2] its
 * algorithmic bytes (8 per row start + 4 per id read, 8 per store word
 * written), [3] ms from the first to the last launch of each build (the one
 * host readback of the build included).  Event-timed. */
int cpm_ingest_stats(cpm_ctx *ctx, double *out4);
int cpm_stats_reset(cpm_ctx *ctx);
/* Enable (1) / disable (0) per-launch event timing of the gram (default on;
 * it adds no synchronization). */
int cpm_set_timing(cpm_ctx *ctx, int on);

/* ---- measured roofs (no counterpart in the reference) --------------------
 * Device microbenchmarks that give the bounds the path's kernels are placed
 * against.  cpm_roof_stream: AND + popcount of the two halves of d_buf
 * (`bytes`, a multiple of 4096, device memory), `reps` passes, with
 * xcd_sliced != 0 each XCD reading only its eighth of each half; *ms is the
 * launch's event time, *bytes_read the bytes loaded.  d_sink: one device
 * u64 the kernel adds into.  cpm_roof_mfma: `iters` rounds of independent FP4
 * MFMAs on every SIMD, shape 16 (16x16x128) or 32 (32x32x64); *ops counts 2 per multiply-add; d_sink: one
 * float per block (8 per CU), never written in practice. */
int cpm_roof_stream(cpm_ctx *ctx, const void *d_buf, uint64_t bytes, uint32_t reps, int xcd_sliced,
                    unsigned long long *d_sink, double *ms, double *bytes_read);
int cpm_roof_mfma(cpm_ctx *ctx, int shape, uint32_t iters, float *d_sink, double *ms, double *ops);
/* cpm_roof_phases: the cost of a phase boundary inside one launch against a
 * launch boundary.  `blocks` workgroups of 256 lanes (all resident: the
 * launch is cooperative; more than fit is an error) run `rounds` (1..4096)
 * rounds of storing `dirty_bytes` each (a multiple of 16, at most 1 MiB) and
 * a grid-wide barrier (agent-scope release, arrival on a counter, a bounded
 * wait); *us_barrier is a round's time.  The same rounds as one launch each
 * in a hipGraph give *us_launch. */
int cpm_roof_phases(cpm_ctx *ctx, uint32_t blocks, uint32_t rounds, uint32_t dirty_bytes, double *us_barrier,
                    double *us_launch);

#ifdef __cplusplus
}
#endif
#endif
