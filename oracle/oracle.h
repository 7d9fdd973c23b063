/*
 * ORACLE - test infrastructure only.
 *
 * A plain-C CPU restatement of the reference's binary operator pipeline,
 * used as the checker for the HIP library. Only tests, the smoke check and
 * the CPU-baseline leg of the benchmark may load it; the product path never
 * does. Each function names the reference file:line it restates (paths are
 * relative to the reference repository root).
 *
 * Pinning (what fixes each function independently of its own formula) is
 * listed per function; see the tests under tests/test_oracle_*.py.
 */
#ifndef SPARKALOG_ORACLE_H
#define SPARKALOG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORACLE_OK = 0,
    ORACLE_TOO_MANY_ROWS = 1,        /* execution/src/lib.rs:325-331 */
    ORACLE_MISSING_COLUMN = 2,       /* execution/src/lib.rs:1258-1266 */
    ORACLE_INDEX_MISMATCH = 3,       /* execution/src/lib.rs:1293-1298 */
    ORACLE_ITERATION_LIMIT = 4,      /* recursion/src/generic.rs:483-488 */
    ORACLE_INCOMPATIBLE_PLANS = 5,   /* recursion/src/generic.rs:446-452 */
    ORACLE_NOT_RECURSIVE = 6,        /* recursion/src/generic.rs:420-423 */
    ORACLE_MISSING_RELATION = 7,     /* recursion/src/generic.rs:108-113 */
    ORACLE_DUPLICATE_RELATION = 8,   /* recursion/src/generic.rs:96-104 */
    ORACLE_EMPTY_SCC = 9,            /* recursion/src/generic.rs:411-413 */
    ORACLE_DELTA_OUTSIDE_SCC = 10,   /* recursion/src/generic.rs:428-430 */
    ORACLE_TARGET_OUTSIDE_SCC = 11,  /* recursion/src/generic.rs:431-433 */
    ORACLE_MISSING_RULE_TARGET = 12, /* recursion/src/generic.rs:443-445 */
    ORACLE_NO_MEMORY = 13
};

/* Sort packed u64 keys ascending (LSD radix; threads <= 1 runs serially). */
void oracle_sort_u64(uint64_t* keys, uint64_t n, int threads);

/* execution/src/lib.rs:398-413 (serial) and :416-434 (parallel). */
uint64_t oracle_distinct(const uint32_t* first, const uint32_t* second, uint64_t n,
                         uint32_t* out_first, uint32_t* out_second, int threads);

/* execution/src/lib.rs:672-710 (serial) and :602-669, :714-783 (parallel). */
uint64_t oracle_anti_join(const uint32_t* lf, const uint32_t* ls, uint64_t l,
                          const uint32_t* rf, const uint32_t* rs, uint64_t r,
                          uint32_t* of, uint32_t* os, int threads);

/* execution/src/lib.rs:934-979 (serial) and :981-1066 (parallel). */
uint64_t oracle_union(const uint32_t* lf, const uint32_t* ls, uint64_t l,
                      const uint32_t* rf, const uint32_t* rs, uint64_t r,
                      uint32_t* of, uint32_t* os, int threads);

/* storage/src/lib.rs:357-393. keys[K], starts[K+1], rows[n]; returns K. */
uint64_t oracle_range_index(const uint32_t* values, uint64_t n, uint32_t* keys,
                            uint32_t* starts, uint32_t* rows);

/* execution/src/lib.rs:1309-1337 (count_join_rows): the join's exact output
 * size, per-row counts in parallel chunks when threads > 1. */
uint64_t oracle_join_count(const uint32_t* left_keys, uint64_t l, const uint32_t* index_keys,
                           const uint32_t* index_starts, uint64_t k, int threads);

/*
 * execution/src/lib.rs:1340-1367 (serial) and :1370-1423 (parallel).
 * projection side 0 = left row, 1 = right row (execution/src/lib.rs:1272-1278).
 * Output order: left-row major, right rows in index order. Returns M.
 */
uint64_t oracle_join(const uint32_t* left_keys, uint64_t l, const uint32_t* index_keys,
                     const uint32_t* index_starts, const uint32_t* index_rows, uint64_t k,
                     const uint32_t* projection0, uint32_t side0,
                     const uint32_t* projection1, uint32_t side1,
                     uint32_t* out0, uint32_t* out1, int threads);

/* relational/src/lib.rs:155-164 + execution/src/lib.rs:1681-1702 / :1705-1767. */
uint64_t oracle_filter(const uint32_t* input, uint64_t n, uint32_t op, uint32_t operand,
                       uint32_t* out_rows, int threads);

/* ---- text loader: storage/src/io.rs:72-141 ---- */

enum {
    ORACLE_LOAD_OK = 0,
    ORACLE_LOAD_LFS_POINTER = 1,   /* io.rs:97-99 (checked over every line first) */
    ORACLE_LOAD_FIELD_COUNT = 2,   /* io.rs:120-125 */
    ORACLE_LOAD_INVALID_U32 = 3,   /* io.rs:136-141 */
    ORACLE_LOAD_CAPACITY = 4       /* more rows than `capacity` */
};

typedef struct {
    int kind;
    uint64_t line;         /* 1-based, as the reference reports it */
    uint32_t column;       /* 1 or 2 for ORACLE_LOAD_INVALID_U32 */
    uint64_t value_offset; /* the offending field, for ORACLE_LOAD_INVALID_U32 */
    uint64_t value_len;
} oracle_load_error;

/* load_binary_u32 over an in-memory text: lines split at '\n', trimmed of
 * ASCII whitespace, blank and '#'/'%' lines skipped (io.rs:86-89), two
 * whitespace-separated fields parsed as Rust's str::parse::<u32> (an optional
 * '+', decimal digits, no overflow). Returns the rows written to
 * first/second, or 0 with *error set. */
uint64_t oracle_load_binary_u32(const char* text, uint64_t len, uint32_t* first, uint32_t* second,
                                uint64_t capacity, oracle_load_error* error);

/* ---- semi-naive executor: recursion/src/generic.rs:16-534 ---- */

typedef struct {
    uint32_t target, delta_input, right_input;
    /* BinaryEqualityJoin (relational/src/lib.rs:46-50) */
    uint32_t left_key, right_key;
    uint32_t out_input[2]; /* 0 = Left, 1 = Right */
    uint32_t out_column[2];
    uint32_t distinct_columns[2];
    uint32_t anti_left[2], anti_right[2];
    uint32_t union_left[2], union_right[2];
} oracle_rule;

typedef struct {
    uint64_t iterations;
    uint64_t total_new_rows;
    uint64_t delta_rows_left; /* set on ORACLE_ITERATION_LIMIT */
} oracle_summary;

typedef struct oracle_store oracle_store;

oracle_store* oracle_store_new(void);
void oracle_store_free(oracle_store* store);
int oracle_store_insert_static(oracle_store* store, uint32_t id, const uint32_t* c0,
                               const uint32_t* c1, uint64_t n);
int oracle_store_insert_recursive(oracle_store* store, uint32_t id, const uint32_t* full0,
                                  const uint32_t* full1, uint64_t nfull, const uint32_t* delta0,
                                  const uint32_t* delta1, uint64_t ndelta);
/* Copy FULL (which = 0) or DELTA (which = 1) out; NULL buffers return the size only. */
uint64_t oracle_store_get(const oracle_store* store, uint32_t id, int which, uint32_t* c0,
                          uint32_t* c1);
/* RecursiveExecutor::compile + run (recursion/src/generic.rs:407-516). */
int oracle_run(oracle_store* store, const uint32_t* relations, uint32_t n_relations,
               const oracle_rule* rules, uint32_t n_rules, uint64_t max_iterations,
               int threads, oracle_summary* summary);

/* ---- brute force (independent of the semi-naive restatement) ---- */

typedef struct {
    uint64_t pairs;      /* |closure| */
    uint64_t digest;     /* sum of mix64(((x<<32)|z) + golden) */
    uint64_t max_dist;   /* longest shortest path, in edges */
} oracle_bfs_result;

/*
 * Transitive closure by breadth-first search from every source over the
 * sorted edge list; nodes are ids < n_nodes. per_source (optional, n_nodes)
 * receives each source's reach count. closure_first/second (optional) receive
 * the sorted closure when capacity >= pairs.
 */
int oracle_tc_bfs(const uint32_t* src, const uint32_t* dst, uint64_t edges, uint32_t n_nodes,
                  int threads, oracle_bfs_result* result, uint64_t* per_source,
                  uint32_t* closure_first, uint32_t* closure_second, uint64_t capacity);

/*
 * The general CPU path's semi-naive evaluation (datalog/src/general.rs:391-443
 * with the nested-loop clause evaluation of :478-566) for
 *   path(x,y) :- edge(x,y).  path(x,z) :- path(x,y), edge(y,z).
 *   nonlocal(x,z) :- path(x,z), !edge(x,z).
 * Tuple sets are kept sorted; each clause binding scans every tuple of the
 * joined relation as the reference does. Returns |path| and, via out params,
 * |nonlocal|, iterations and new rows; path/nonlocal copied out when non-NULL
 * (capacity = edges_cap rows each). Intended for inputs of a few thousand edges.
 */
uint64_t oracle_general_tc(const uint32_t* src, const uint32_t* dst, uint64_t edges,
                           uint64_t* nonlocal_rows, uint64_t* iterations, uint64_t* new_rows,
                           uint32_t* path_first, uint32_t* path_second,
                           uint32_t* nonlocal_first, uint32_t* nonlocal_second,
                           uint64_t capacity);

#ifdef __cplusplus
}
#endif

#endif
