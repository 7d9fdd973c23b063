/*
 * sparkalog_hip: MI355X (gfx950) native binary recursive operator pipeline.
 *
 * Part 1 is the drop-in operator ABI. Each sparkalog_hip_* operator replaces
 * the reference's extern "C" function of the same suffix in
 * crates/sparkalog-execution/native/column_ops.cu (declared for Rust at
 * crates/sparkalog-execution/src/lib.rs:28-169) with identical arguments and
 * conventions:
 *   - the return value is a hipError_t (0 = success);
 *   - all work is asynchronous on `stream` (a hipStream_t, or NULL);
 *   - nothing is allocated: buffers are caller-owned device memory, and
 *     scratch comes from `temporary`, sized by the matching
 *     *_temporary_bytes query;
 *   - an empty input writes a zero count on the stream and succeeds;
 *   - a null pointer with a non-zero length is hipErrorInvalidValue;
 *   - counts are written on the device (u32 for filter and anti-join, u64 for
 *     distinct, union and the join total);
 *   - the join count finds its output offsets in a single pass, waiting
 *     on earlier tiles with a bounded wait; if such a wait ever gives up (a
 *     scheduling failure of the device, never seen in practice), the total
 *     is written as all ones (UINT64_MAX) after its kernels, so the caller
 *     can tell it from a result. The filter, distinct, union and anti-join
 *     count first and write second, and never wait on another tile.
 * Scratch arrays the reference needed for its CUB pipeline (filter flags,
 * anti-join flags/selection, union packed copies, distinct packed/scratch)
 * keep their parameters; the kernels here use them only where noted.
 *
 * Part 2 is the device-resident relation store and semi-naive executor
 * (RelationStore / RecursiveExecutor of
 * crates/sparkalog-recursion/src/generic.rs:16-534), with FULL/DELTA/NEWT
 * kept in HBM between iterations.
 */
#ifndef SPARKALOG_HIP_H
#define SPARKALOG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int sparkalog_hip_error; /* hipError_t value */

/* ------------------------------------------------------------------ */
/* Part 1: operator ABI                                                */

/* column_ops.cu:231-246 (smoke test: data[i] += 1). */
sparkalog_hip_error sparkalog_hip_add_one_i32(int32_t* data, size_t len, void* stream);

/* column_ops.cu:248-250. */
sparkalog_hip_error sparkalog_hip_stream_synchronize(void* stream);

/* column_ops.cu:252-276. `flags` is unused (ballot compaction needs no flag
 * array) and may be NULL. */
sparkalog_hip_error sparkalog_hip_filter_u32_temporary_bytes(const uint32_t* flags, uint32_t* output,
                                                             uint32_t* count, size_t len,
                                                             size_t* temporary_bytes, void* stream);

/* column_ops.cu:278-318: output[0..*count) = ordered row ids i with
 * compare(input[i], operation, operand); operation 0..5 = Eq, Ne, Lt, Le,
 * Gt, Ge (crates/sparkalog-execution/src/lib.rs:1769-1778). */
sparkalog_hip_error sparkalog_hip_filter_u32(const uint32_t* input, size_t len, uint32_t operation,
                                             uint32_t operand, uint32_t* flags, uint32_t* output,
                                             uint32_t* count, void* temporary, size_t temporary_bytes,
                                             void* stream);

/* column_ops.cu:320-337: output[i] = i % modulus. */
sparkalog_hip_error sparkalog_hip_fill_mod_u32(uint32_t* output, size_t len, uint32_t modulus,
                                               void* stream);

/* Bench helper with no counterpart in the reference: a plain device copy of
 * `count` u64 words (both pointers 16-byte aligned). It measures the HBM copy
 * ceiling and is the known-size kernel that calibrates HBM byte counters. */
sparkalog_hip_error sparkalog_hip_copy_u64(const uint64_t* source, uint64_t* destination, size_t count,
                                           void* stream);

/* column_ops.cu:339-354. */
sparkalog_hip_error sparkalog_hip_join_u32_temporary_bytes(const uint64_t* counts, uint64_t* offsets,
                                                           size_t left_rows, size_t* temporary_bytes,
                                                           void* stream);

/* column_ops.cu:356-399: counts[l] = matches of left_keys[l] in the range
 * index (keys[unique_keys] sorted unique, starts[unique_keys+1]); offsets =
 * exclusive sum of counts; *total = sum. */
sparkalog_hip_error sparkalog_hip_join_u32_count(const uint32_t* left_keys, size_t left_rows,
                                                 const uint32_t* index_keys, const uint32_t* index_starts,
                                                 size_t unique_keys, uint64_t* counts, uint64_t* offsets,
                                                 uint64_t* total, void* temporary, size_t temporary_bytes,
                                                 void* stream);

/* No counterpart in the reference: sparkalog_hip_join_u32_count for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Same arguments and results
 * as sparkalog_hip_join_u32_count without `temporary`: its scratch (look-back
 * header and status words) is a static per-device array of
 * SPARKALOG_HIP_NO_ALLOC_SCRATCH_BYTES, so it allocates nothing (no hipMalloc,
 * hipFree or allocating library call) and issues no hipMemsetAsync. At most
 * SPARKALOG_HIP_NO_ALLOC_MAX_LEFT_ROWS left rows; larger inputs return
 * hipErrorInvalidValue before any device work. Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (other streams, other
 * threads) race and must be serialized by the caller. */
#define SPARKALOG_HIP_NO_ALLOC_SCRATCH_BYTES (1u << 20)
#define SPARKALOG_HIP_NO_ALLOC_MAX_LEFT_ROWS 268369920u /* (1 MiB - 256 B) / 8 B per 2,048-row tile */
sparkalog_hip_error sparkalog_hip_join_u32_count_no_alloc(const uint32_t* left_keys, size_t left_rows,
                                                          const uint32_t* index_keys, const uint32_t* index_starts,
                                                          size_t unique_keys, uint64_t* counts, uint64_t* offsets,
                                                          uint64_t* total, void* stream);

/* column_ops.cu:401-440: writes output rows offsets[l] .. offsets[l]+counts[l]
 * of left row l in index order; side 0 takes projection[l], side 1 takes
 * projection[index_rows[match]]. */
sparkalog_hip_error sparkalog_hip_join_u32_emit(const uint32_t* left_keys, size_t left_rows,
                                                const uint32_t* index_keys, const uint32_t* index_starts,
                                                const uint32_t* index_rows, size_t unique_keys,
                                                const uint32_t* projection0, uint32_t projection0_side,
                                                const uint32_t* projection1, uint32_t projection1_side,
                                                const uint64_t* offsets, uint32_t* output0, uint32_t* output1,
                                                void* stream);

/* column_ops.cu:442-484. */
sparkalog_hip_error sparkalog_hip_distinct_u32_temporary_bytes(uint64_t* packed, uint64_t* scratch,
                                                               uint64_t* unique_rows, size_t rows,
                                                               size_t* temporary_bytes, void* stream);

/* column_ops.cu:486-547: sorted unique (first, second) pairs. `packed` and
 * `scratch` (rows entries each) are the radix-sort ping-pong buffers. */
sparkalog_hip_error sparkalog_hip_distinct_u32(const uint32_t* first, const uint32_t* second, size_t rows,
                                               uint64_t* packed, uint64_t* scratch, uint64_t* unique_rows,
                                               uint32_t* output_first, uint32_t* output_second,
                                               void* temporary, size_t temporary_bytes, void* stream);

/* column_ops.cu:549-575. `flags` and `selected` are unused and may be NULL. */
sparkalog_hip_error sparkalog_hip_anti_join_u32_temporary_bytes(const uint32_t* flags, uint32_t* selected,
                                                                uint32_t* selected_rows, size_t left_rows,
                                                                size_t* temporary_bytes, void* stream);

/* column_ops.cu:577-647: left rows absent from the sorted right relation, in
 * left order; *selected_rows = count. */
sparkalog_hip_error sparkalog_hip_anti_join_u32(const uint32_t* left_first, const uint32_t* left_second,
                                                size_t left_rows, const uint32_t* right_first,
                                                const uint32_t* right_second, size_t right_rows,
                                                uint32_t* flags, uint32_t* selected, uint32_t* selected_rows,
                                                uint32_t* output_first, uint32_t* output_second,
                                                void* temporary, size_t temporary_bytes, void* stream);

/* column_ops.cu:649-689. left/right/merged/unique are unused and may be NULL. */
sparkalog_hip_error sparkalog_hip_union_u32_temporary_bytes(const uint64_t* left, size_t left_rows,
                                                            const uint64_t* right, size_t right_rows,
                                                            uint64_t* merged, uint64_t* unique,
                                                            uint64_t* unique_rows, size_t* temporary_bytes,
                                                            void* stream);

/* column_ops.cu:691-775: merge of two sorted relations with every adjacent
 * duplicate dropped; *unique_rows = count. The packed scratch arrays are
 * unused (the merge reads the columns directly) and may be NULL. */
sparkalog_hip_error sparkalog_hip_union_u32(const uint32_t* left_first, const uint32_t* left_second,
                                            size_t left_rows, const uint32_t* right_first,
                                            const uint32_t* right_second, size_t right_rows, uint64_t* left,
                                            uint64_t* right, uint64_t* merged, uint64_t* unique,
                                            uint64_t* unique_rows, uint32_t* output_first,
                                            uint32_t* output_second, void* temporary, size_t temporary_bytes,
                                            void* stream);

/* ---- device range index ----
 * U32RangeIndex::build_slice (crates/sparkalog-storage/src/lib.rs:357-393)
 * on the device: from `rows` values (the join key column of the right
 * relation), keys[K] sorted unique, starts[K + 1] (starts[K] = rows) and
 * index_rows[rows] = row ids sorted by (value, row), the layout
 * sparkalog_hip_join_u32_count / _emit take. *unique_keys (u64, device) = K;
 * keys and starts hold at least rows and rows + 1 entries. Asynchronous on
 * `stream`. */
sparkalog_hip_error sparkalog_hip_range_index_u32_temporary_bytes(size_t rows, size_t* temporary_bytes);
sparkalog_hip_error sparkalog_hip_range_index_u32(const uint32_t* values, size_t rows, uint32_t* keys,
                                                  uint32_t* starts, uint32_t* index_rows, uint64_t* unique_keys,
                                                  void* temporary, size_t temporary_bytes, void* stream);

/* ---- device text ingest ----
 * load_binary_u32 (crates/sparkalog-storage/src/io.rs:72-141) on the device:
 * a whitespace-separated text of two u32 fields per line, already in device
 * memory (`text` 16-byte aligned, `len` bytes), parsed into two device
 * columns in file order. Two phases, as the reference reads the file twice
 * (io.rs:75-83): sparkalog_hip_text_lines writes the line count (u64,
 * device), which bounds the rows and sizes the columns and the temporary of
 * sparkalog_hip_load_binary_u32; that call writes *rows (u64, device) and
 * *error (u64, device): SPARKALOG_HIP_LOAD_OK, or the reference's first
 * error (a Git LFS pointer line anywhere, io.rs:97-99; else the first line
 * with other than two fields, io.rs:120-125, or a field that is not a u32,
 * io.rs:126-127, first field before second). Lines are split at '\n' (a
 * trailing '\r' is whitespace), trimmed, and skipped when empty or starting
 * with '#' or '%' (io.rs:86-89); fields are split at ASCII whitespace and
 * parsed as Rust's str::parse::<u32> does (an optional '+', then decimal
 * digits, at most 4294967295). Differences: non-ASCII whitespace does not
 * separate fields, and bytes that are not UTF-8 are reported as a bad field
 * rather than an I/O error. Both calls are asynchronous on `stream`. */
#define SPARKALOG_HIP_LOAD_OK UINT64_MAX
#define SPARKALOG_HIP_LOAD_ERR_CAPACITY 1ull /* the text has more lines than `lines` */
#define SPARKALOG_HIP_LOAD_ERR_LFS_POINTER 0ull
#define SPARKALOG_HIP_LOAD_ERR_FIELD_COUNT ((1ull << 62) | 1ull)
#define SPARKALOG_HIP_LOAD_ERR_U32_FIRST ((1ull << 62) | 2ull)
#define SPARKALOG_HIP_LOAD_ERR_U32_SECOND ((1ull << 62) | 3ull)
/* An error code other than OK and CAPACITY is KIND | (line << 2), line
 * 1-based as in the reference's messages. */
#define SPARKALOG_HIP_LOAD_ERROR_KIND(code) ((code) & ((3ull << 62) | 3ull))
#define SPARKALOG_HIP_LOAD_ERROR_LINE(code) (((code) >> 2) & ((1ull << 60) - 1))
sparkalog_hip_error sparkalog_hip_text_lines_temporary_bytes(size_t len, size_t* temporary_bytes);
sparkalog_hip_error sparkalog_hip_text_lines(const char* text, size_t len, uint64_t* lines, void* temporary,
                                             size_t temporary_bytes, void* stream);
sparkalog_hip_error sparkalog_hip_load_binary_u32_temporary_bytes(size_t len, uint64_t lines,
                                                                  size_t* temporary_bytes);
/* `first` and `second` hold at least `lines` entries (the count from
 * sparkalog_hip_text_lines). */
sparkalog_hip_error sparkalog_hip_load_binary_u32(const char* text, size_t len, uint64_t lines, uint32_t* first,
                                                  uint32_t* second, uint64_t* rows, uint64_t* error,
                                                  void* temporary, size_t temporary_bytes, void* stream);

/* ------------------------------------------------------------------ */
/* Part 2: device-resident relation store and semi-naive executor      */

/* Status codes of the store/executor calls (GenericExecutionError and
 * RelationStoreError, crates/sparkalog-recursion/src/generic.rs:129-135 and
 * :544-559; execution::Error, crates/sparkalog-execution/src/lib.rs:171-188). */
enum sparkalog_hip_status {
    SPARKALOG_HIP_OK = 0,
    SPARKALOG_HIP_ERR_HIP = 1,               /* a HIP call failed: sparkalog_hip_last_hip_error() */
    SPARKALOG_HIP_ERR_TOO_MANY_ROWS = 2,     /* > u32::MAX rows */
    SPARKALOG_HIP_ERR_MISSING_COLUMN = 3,    /* plan column >= 2 */
    SPARKALOG_HIP_ERR_ITERATION_LIMIT = 4,
    SPARKALOG_HIP_ERR_INCOMPATIBLE_PLANS = 5,
    SPARKALOG_HIP_ERR_NOT_RECURSIVE = 6,
    SPARKALOG_HIP_ERR_MISSING_RELATION = 7,
    SPARKALOG_HIP_ERR_DUPLICATE_RELATION = 8,
    SPARKALOG_HIP_ERR_EMPTY_SCC = 9,
    SPARKALOG_HIP_ERR_DELTA_OUTSIDE_SCC = 10,
    SPARKALOG_HIP_ERR_TARGET_OUTSIDE_SCC = 11,
    SPARKALOG_HIP_ERR_MISSING_RULE_TARGET = 12,
    SPARKALOG_HIP_ERR_OUT_OF_MEMORY = 13,
    SPARKALOG_HIP_ERR_UNSUPPORTED_PLAN = 14, /* anti-join/union columns other than [0,1] */
    SPARKALOG_HIP_ERR_DEVICE = 15,           /* a kernel raised its device error flag */
    SPARKALOG_HIP_ERR_INVALID_ARGUMENT = 16
};

/* RecursiveRulePlan (crates/sparkalog-relational/src/lib.rs:74-82). */
typedef struct {
    uint32_t target, delta_input, right_input;
    uint32_t left_key, right_key;     /* BinaryEqualityJoin keys */
    uint32_t out_input[2];            /* JoinInput: 0 = Left, 1 = Right */
    uint32_t out_column[2];
    uint32_t distinct_columns[2];     /* BinaryDistinct */
    uint32_t anti_left[2], anti_right[2];   /* SortedBinaryAntiJoin */
    uint32_t union_left[2], union_right[2]; /* SortedBinaryUnion */
} sparkalog_hip_rule;

/* SccSummary (crates/sparkalog-recursion/src/generic.rs:536-542) plus
 * device timing and traffic of the run. */
typedef struct {
    uint64_t iterations;
    uint64_t total_new_rows;
    uint64_t delta_rows_left;   /* on ITERATION_LIMIT: DELTA rows still pending */
    double device_ms;           /* wall time of the run on the stream */
    double algorithmic_bytes;   /* sum over operators of their algorithmic bytes */
} sparkalog_hip_summary;

typedef struct sparkalog_hip_store sparkalog_hip_store;

/* A store bound to one device stream (NULL = the default stream). */
int sparkalog_hip_store_create(sparkalog_hip_store** out, void* stream);
void sparkalog_hip_store_destroy(sparkalog_hip_store* store);

/* RelationStore::insert_static (generic.rs:36-53). Columns are host memory
 * when on_device == 0 (staged through pinned memory) or device memory. */
int sparkalog_hip_store_insert_static(sparkalog_hip_store* store, uint32_t id, const uint32_t* c0,
                                      const uint32_t* c1, uint64_t rows, int on_device);

/* RelationStore::insert_recursive (generic.rs:55-79). FULL and DELTA must be
 * sorted unique pairs, as the reference requires of its callers. */
int sparkalog_hip_store_insert_recursive(sparkalog_hip_store* store, uint32_t id, const uint32_t* full0,
                                         const uint32_t* full1, uint64_t full_rows, const uint32_t* delta0,
                                         const uint32_t* delta1, uint64_t delta_rows, int on_device);

/* Forget every relation but keep the device buffers (and the executor's
 * operator workspaces) for reuse, so a repeated run allocates nothing. */
int sparkalog_hip_store_clear(sparkalog_hip_store* store);

/* Row count of FULL (which = 0) or DELTA (which = 1). */
int sparkalog_hip_store_rows(sparkalog_hip_store* store, uint32_t id, int which, uint64_t* rows);

/* Copy FULL / DELTA out as two columns of `capacity` rows (host memory when
 * to_device == 0, through pinned staging; device memory otherwise). */
int sparkalog_hip_store_read(sparkalog_hip_store* store, uint32_t id, int which, uint32_t* c0, uint32_t* c1,
                             uint64_t capacity, int to_device);

/* Order-independent digest (sum of mix64(((a<<32)|b) + 0x9E3779B97F4A7C15)),
 * strict-sortedness flag and row count of FULL, computed on the device. */
int sparkalog_hip_store_digest(sparkalog_hip_store* store, uint32_t id, uint64_t* digest, int* sorted_unique,
                               uint64_t* rows);

/* RecursiveExecutor::compile (generic.rs:407-470): validation only. */
int sparkalog_hip_executor_compile(sparkalog_hip_store* store, const uint32_t* relations, uint32_t n_relations,
                                   const sparkalog_hip_rule* rules, uint32_t n_rules);

/* RecursiveExecutor::compile + run (generic.rs:407-516) on the device. */
int sparkalog_hip_executor_run(sparkalog_hip_store* store, const uint32_t* relations, uint32_t n_relations,
                               const sparkalog_hip_rule* rules, uint32_t n_rules, uint64_t max_iterations,
                               sparkalog_hip_summary* summary);

/* Per-operator profiling: when enabled, every operator of a run is bracketed
 * by HIP events on the store's stream. op: 0 join count, 1 join emit,
 * 2 distinct, 3 anti-join (probe of FULL), 4 anti-join of few candidates
 * through FULL's first-column index, 5 union of rule contributions, 6 merge
 * of FULL's runs, 7 index build (sparkalog_hip_store_op_kernel_stats names
 * them). Reports launches, summed device milliseconds and summed algorithmic
 * bytes (8 bytes per pair read or written). */
int sparkalog_hip_store_set_profiling(sparkalog_hip_store* store, int enabled);
/* Executor tuning (no counterpart in the reference, whose union rewrites all
 * of FULL each iteration): FULL is kept as size-tiered sorted runs, and the
 * newest run is merged into the one before while that one holds fewer than
 * merge_ratio (>= 2, default 3) times its rows; segmented_distinct (default
 * 1) sorts candidates grouped by their first column per group in LDS instead
 * of a global radix sort; indexed_anti_join (default 0) gives FULL's runs a
 * first-column index (while first values are below 2^24) so the anti-join
 * reads only the rows of first values it holds; iterations with at most
 * sparse_probe_rows candidates (default 2^22) check them against FULL one
 * thread each through that index. Results do not depend on any of them. */
int sparkalog_hip_store_set_tuning(sparkalog_hip_store* store, uint32_t merge_ratio, int segmented_distinct,
                                   int indexed_anti_join, uint64_t sparse_probe_rows);
/* Named executor options (no counterpart in the reference); results do not
 * depend on any of them. Unknown names or values return
 * SPARKALOG_HIP_ERR_INVALID_ARGUMENT.
 *   "join_index": layout of the join's right-relation index (the reference's
 *     U32RangeIndex, crates/sparkalog-storage/src/lib.rs:345-423): 0 = chosen
 *     per relation (default: a dense start array over key values while the
 *     largest value is below 2^28 and the array is at most about 4 entries per
 *     row beyond 2^24; else the unique keys through a hash table, which
 *     indexes any u32 value), 1 = dense where the values allow it, 2 = hashed.
 *   "graph": 1 = replay each semi-naive iteration of a fixpoint with one rule
 *     over a static right relation (transitive closure's shape) from a
 *     captured HIP graph: FULL stays one sorted array, every count stays on
 *     the device, and the host synchronizes once per iteration (other plans
 *     run the regular executor); 0 = off (default).
 *   "graph_max_rows": FULL capacity (rows, 2^16..2^31, default 2^25) past
 *     which a graph-mode fixpoint continues in the regular executor, whose
 *     size-tiered FULL is not rewritten every iteration.
 *   "distinct_tile": nominal tile of the segmented distinct in keys (512..4096,
 *     a multiple of 64; default 3072): tiles are cut every distinct_tile keys
 *     and moved back to run starts, and each is sorted in LDS.
 *   "merge_ratio", "segmented_distinct", "indexed_anti_join",
 *     "sparse_probe_rows": the knobs of sparkalog_hip_store_set_tuning, one at
 *     a time (merge_ratio 2..64, the two switches 0/1). */
int sparkalog_hip_store_set_option(sparkalog_hip_store* store, const char* name, uint64_t value);
/* Graphs captured and graph replays of graph-mode iterations since the store
 * was created (no counterpart in the reference). */
int sparkalog_hip_store_graph_stats(sparkalog_hip_store* store, uint64_t* captures, uint64_t* replays);
/* Of those captures, how many updated an executable graph in place (the same
 * launches over other buffers or sizes) rather than instantiating a new one,
 * and the host milliseconds spent capturing and instantiating or updating
 * (no counterpart in the reference). */
int sparkalog_hip_store_graph_capture_stats(sparkalog_hip_store* store, uint64_t* updates, double* capture_ms);
int sparkalog_hip_store_op_stats(sparkalog_hip_store* store, int op, uint64_t* launches, double* ms,
                                 double* bytes);
/* The operator's name and main kernel, and that kernel's own launches and
 * time (events around it alone). `store` may be null to query names only;
 * returns SPARKALOG_HIP_ERR_INVALID_ARGUMENT past the last operator. */
int sparkalog_hip_store_op_kernel_stats(sparkalog_hip_store* store, int op, const char** name, const char** kernel,
                                        uint64_t* kernel_launches, double* kernel_ms);

/* ---- stepwise fixpoint (for callers that change relations between steps) ----
 * begin = RecursiveExecutor::compile; step = one loop trip of
 * RecursiveExecutor::run (generic.rs:481-512) unless every DELTA is empty,
 * returning the DELTA rows left; end = leave FULL as one array and report. */
int sparkalog_hip_session_begin(sparkalog_hip_store* store, const uint32_t* relations, uint32_t n_relations,
                                const sparkalog_hip_rule* rules, uint32_t n_rules);
int sparkalog_hip_session_step(sparkalog_hip_store* store, uint64_t* delta_rows);
int sparkalog_hip_session_end(sparkalog_hip_store* store, sparkalog_hip_summary* summary);

/* ---- many independent instances in one relation ----
 * Instance of a row = first column >> instance_bits; local pair = both
 * columns masked to instance_bits. No counterpart in the reference: these
 * serve the batched and served fixpoints. */

/* Record, per instance, 1 + the last iteration that produced new rows (0 =
 * none); applies to later session_begin / executor_run calls. */
int sparkalog_hip_set_instance_tracking(sparkalog_hip_store* store, uint32_t instance_bits, uint32_t instances);
int sparkalog_hip_instance_iterations(sparkalog_hip_store* store, uint32_t* last_new, uint32_t instances);

/* Per instance: DELTA rows after the last session_step, counted on the
 * device from that step's NEWT and read back with its counters, so a caller
 * learns which instances have converged without another synchronization
 * (host copy only; valid after a step of a session begun with tracking). */
int sparkalog_hip_instance_delta_rows(sparkalog_hip_store* store, uint64_t* rows, uint32_t instances);

/* Per instance: FULL rows and digest of its local pairs (the digest: the
 * sum over the pairs of splitmix64's finalizer applied to the local pair).
 * Right after a session that tracked these instances of this relation, and
 * while the relation is unchanged, the answer is the statistics the session
 * accumulated as it wrote each NEWT; otherwise FULL is read. */
int sparkalog_hip_store_instance_stats(sparkalog_hip_store* store, uint32_t id, uint32_t instance_bits,
                                       uint32_t instances, uint64_t* rows, uint64_t* digests);

/* Rows and digest of the FULL (which = 0) or DELTA (1) rows whose first
 * column is in [first_lo, first_hi), pairs taken relative to `base`. */
int sparkalog_hip_store_range_stats(sparkalog_hip_store* store, uint32_t id, int which, uint32_t first_lo,
                                    uint32_t first_hi, uint32_t base, uint64_t* rows, uint64_t* digest);

/* Row counts of FULL (which = 0) or DELTA (1) per range of the first
 * column: range j is [first_bounds[2j], first_bounds[2j+1]). */
int sparkalog_hip_store_range_rows(sparkalog_hip_store* store, uint32_t id, int which, const uint32_t* first_bounds,
                                   uint32_t ranges, uint64_t* rows);

/* Add sorted unique rows: to FULL (which = 0; for a recursive relation they
 * must be disjoint from FULL and become a run of their own) or DELTA (1). */
int sparkalog_hip_store_merge_rows(sparkalog_hip_store* store, uint32_t id, int which, const uint32_t* c0,
                                   const uint32_t* c1, uint64_t rows, int on_device);

/* Remove every row whose first column is in [first_lo, first_hi) from FULL
 * and DELTA. */
/* A non-recursive rule with stratified negation, target(x, y) :- left(x, y),
 * !right(x, y), over two stored relations (FULL of each): target becomes a
 * static relation holding left \ right in order (SortedBinaryAntiJoin,
 * crates/sparkalog-execution/src/lib.rs:672-710, as the general path's
 * negation against an earlier stratum, crates/sparkalog-datalog/src/
 * general.rs:522-536). Writes the row count to *rows (may be null). */
int sparkalog_hip_store_difference(sparkalog_hip_store* store, uint32_t target, uint32_t left, uint32_t right,
                                   uint64_t* rows);

int sparkalog_hip_store_remove_range(sparkalog_hip_store* store, uint32_t id, uint32_t first_lo, uint32_t first_hi);
/* Remove the rows of several ranges at once: first_bounds holds `ranges`
 * pairs [lo, hi), ascending and disjoint; each part of FULL and DELTA is
 * copied once. */
int sparkalog_hip_store_remove_ranges(sparkalog_hip_store* store, uint32_t id, const uint32_t* first_bounds,
                                      uint32_t ranges);

/* The hipError_t of the last failed HIP call in this thread (0 if none). */
int sparkalog_hip_last_hip_error(void);

/* Human-readable text of a status code. */
const char* sparkalog_hip_status_string(int status);

#ifdef __cplusplus
}
#endif

#endif
