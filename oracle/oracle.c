/*
 * ORACLE - test infrastructure only. Never linked into the product library.
 *
 * Plain-C restatement of the reference's CPU operators, range index and
 * semi-naive executor, plus brute-force transitive closure. Declarations and
 * reference citations: oracle.h. Paths are relative to the reference
 * repository root.
 *
 * Parallel variants reproduce Rayon's chunking, chunk = ceil(n / workers)
 * (execution/src/lib.rs:725-727, :1032-1033, :1384-1385); their results are
 * identical to the serial ones and only their timing differs.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define GOLDEN 0x9E3779B97F4A7C15ull

static uint64_t mix64(uint64_t z) {
    z += GOLDEN;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* execution/src/lib.rs:344-346 */
static uint64_t pack_tuple(uint32_t first, uint32_t second) {
    return ((uint64_t)first << 32) | (uint64_t)second;
}

/* ------------------------------------------------------------------ */
/* threads                                                             */

typedef void (*task_fn)(void* ctx, int worker, int workers);

typedef struct {
    task_fn fn;
    void* ctx;
    int worker, workers;
} task_arg;

static void* task_trampoline(void* p) {
    task_arg* a = (task_arg*)p;
    a->fn(a->ctx, a->worker, a->workers);
    return NULL;
}

static void run_workers(int workers, task_fn fn, void* ctx) {
    if (workers <= 1) {
        fn(ctx, 0, 1);
        return;
    }
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)workers);
    task_arg* args = (task_arg*)malloc(sizeof(task_arg) * (size_t)workers);
    for (int w = 0; w < workers; ++w) {
        args[w].fn = fn;
        args[w].ctx = ctx;
        args[w].worker = w;
        args[w].workers = workers;
        if (w > 0) pthread_create(&th[w], NULL, task_trampoline, &args[w]);
    }
    fn(ctx, 0, workers);
    for (int w = 1; w < workers; ++w) pthread_join(th[w], NULL);
    free(th);
    free(args);
}

static uint64_t div_ceil(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

/* ------------------------------------------------------------------ */
/* sort (LSD radix, 8-bit digits; stands in for sort_unstable /        */
/* par_sort_unstable, whose integer results are identical)             */

typedef struct {
    uint64_t* src;
    uint64_t* dst;
    uint64_t n;
    int shift;
    uint64_t* hist; /* workers x 256 */
} radix_ctx;

static void radix_hist(void* p, int w, int ws) {
    radix_ctx* c = (radix_ctx*)p;
    uint64_t chunk = div_ceil(c->n, (uint64_t)ws);
    uint64_t b = chunk * (uint64_t)w, e = b + chunk < c->n ? b + chunk : c->n;
    uint64_t* h = c->hist + (size_t)w * 256;
    memset(h, 0, 256 * sizeof(uint64_t));
    for (uint64_t i = b; i < e; ++i) h[(c->src[i] >> c->shift) & 0xFF]++;
}

static void radix_scatter(void* p, int w, int ws) {
    radix_ctx* c = (radix_ctx*)p;
    uint64_t chunk = div_ceil(c->n, (uint64_t)ws);
    uint64_t b = chunk * (uint64_t)w, e = b + chunk < c->n ? b + chunk : c->n;
    uint64_t* h = c->hist + (size_t)w * 256;
    for (uint64_t i = b; i < e; ++i) {
        uint64_t k = c->src[i];
        c->dst[h[(k >> c->shift) & 0xFF]++] = k;
    }
}

void oracle_sort_u64(uint64_t* keys, uint64_t n, int threads) {
    if (n < 2) return;
    int ws = threads < 1 ? 1 : threads;
    if ((uint64_t)ws > n / 65536 + 1) ws = (int)(n / 65536 + 1);
    uint64_t* tmp = (uint64_t*)malloc(n * sizeof(uint64_t));
    uint64_t* hist = (uint64_t*)malloc((size_t)ws * 256 * sizeof(uint64_t));
    uint64_t* src = keys;
    uint64_t* dst = tmp;
    uint64_t all_or = 0, all_and = ~0ull;
    for (uint64_t i = 0; i < n; ++i) {
        all_or |= keys[i];
        all_and &= keys[i];
    }
    for (int shift = 0; shift < 64; shift += 8) {
        if ((((all_or ^ all_and) >> shift) & 0xFF) == 0) continue; /* constant digit */
        radix_ctx c = {src, dst, n, shift, hist};
        run_workers(ws, radix_hist, &c);
        uint64_t sum = 0;
        for (int d = 0; d < 256; ++d)
            for (int w = 0; w < ws; ++w) {
                uint64_t v = hist[(size_t)w * 256 + d];
                hist[(size_t)w * 256 + d] = sum;
                sum += v;
            }
        run_workers(ws, radix_scatter, &c);
        uint64_t* t = src;
        src = dst;
        dst = t;
    }
    if (src != keys) memcpy(keys, src, n * sizeof(uint64_t));
    free(tmp);
    free(hist);
}

/* ------------------------------------------------------------------ */
/* distinct: execution/src/lib.rs:344-434                              */

/* execution/src/lib.rs:348-357 */
static uint64_t deduplicate_sorted(uint64_t* packed, uint64_t n) {
    uint64_t unique = 0;
    for (uint64_t read = 0; read < n; ++read) {
        if (unique == 0 || packed[read] != packed[unique - 1]) packed[unique++] = packed[read];
    }
    return unique;
}

typedef struct {
    const uint32_t *a, *b;
    uint64_t* packed;
    uint32_t *oa, *ob;
    uint64_t n;
} pack_ctx;

static void pack_worker(void* p, int w, int ws) {
    pack_ctx* c = (pack_ctx*)p;
    uint64_t chunk = div_ceil(c->n, (uint64_t)ws);
    uint64_t b = chunk * (uint64_t)w, e = b + chunk < c->n ? b + chunk : c->n;
    for (uint64_t i = b; i < e; ++i) c->packed[i] = pack_tuple(c->a[i], c->b[i]);
}

static void unpack_worker(void* p, int w, int ws) {
    pack_ctx* c = (pack_ctx*)p;
    uint64_t chunk = div_ceil(c->n, (uint64_t)ws);
    uint64_t b = chunk * (uint64_t)w, e = b + chunk < c->n ? b + chunk : c->n;
    for (uint64_t i = b; i < e; ++i) {
        c->oa[i] = (uint32_t)(c->packed[i] >> 32);
        c->ob[i] = (uint32_t)c->packed[i];
    }
}

static int clamp_workers(int threads, uint64_t n) {
    int ws = threads < 1 ? 1 : threads;
    if ((uint64_t)ws > n) ws = (int)(n == 0 ? 1 : n);
    return ws;
}

uint64_t oracle_distinct(const uint32_t* first, const uint32_t* second, uint64_t n,
                         uint32_t* out_first, uint32_t* out_second, int threads) {
    if (n == 0) return 0;
    uint64_t* packed = (uint64_t*)malloc(n * sizeof(uint64_t));
    int ws = clamp_workers(threads, n);
    pack_ctx c = {first, second, packed, out_first, out_second, n};
    run_workers(ws, pack_worker, &c);
    oracle_sort_u64(packed, n, ws);
    uint64_t unique = deduplicate_sorted(packed, n);
    c.n = unique;
    run_workers(clamp_workers(threads, unique), unpack_worker, &c);
    free(packed);
    return unique;
}

/* ------------------------------------------------------------------ */
/* anti-join: execution/src/lib.rs:579-783                             */

static int compare_pair(uint32_t a0, uint32_t a1, uint32_t b0, uint32_t b1) {
    if (a0 != b0) return a0 < b0 ? -1 : 1;
    if (a1 != b1) return a1 < b1 ? -1 : 1;
    return 0;
}

/* execution/src/lib.rs:588-600 */
static uint64_t lower_bound_pair(const uint32_t* rf, const uint32_t* rs, uint64_t r, uint32_t f,
                                 uint32_t s) {
    uint64_t low = 0, high = r;
    while (low < high) {
        uint64_t mid = low + (high - low) / 2;
        if (compare_pair(rf[mid], rs[mid], f, s) < 0)
            low = mid + 1;
        else
            high = mid;
    }
    return low;
}

/* Shared merge loop of execution/src/lib.rs:602-669 and :672-710. */
static uint64_t anti_join_range(const uint32_t* lf, const uint32_t* ls, uint64_t row_start,
                                uint64_t row_end, const uint32_t* rf, const uint32_t* rs,
                                uint64_t r, uint64_t right_row, uint32_t* of, uint32_t* os) {
    uint64_t selected = 0;
    for (uint64_t left_row = row_start; left_row < row_end; ++left_row) {
        while (right_row < r &&
               compare_pair(rf[right_row], rs[right_row], lf[left_row], ls[left_row]) < 0)
            ++right_row;
        int matched = right_row < r && rf[right_row] == lf[left_row] && rs[right_row] == ls[left_row];
        if (!matched) {
            if (of) {
                of[selected] = lf[left_row];
                os[selected] = ls[left_row];
            }
            ++selected;
        }
    }
    return selected;
}

typedef struct {
    const uint32_t *lf, *ls, *rf, *rs;
    uint64_t l, r, chunk_size, chunk_count;
    uint64_t* offsets;
    uint32_t *of, *os;
    int emit;
} anti_ctx;

static void anti_worker(void* p, int w, int ws) {
    anti_ctx* c = (anti_ctx*)p;
    for (uint64_t chunk = (uint64_t)w; chunk < c->chunk_count; chunk += (uint64_t)ws) {
        uint64_t b = chunk * c->chunk_size;
        uint64_t e = b + c->chunk_size < c->l ? b + c->chunk_size : c->l;
        if (b == e) continue;
        uint64_t start = lower_bound_pair(c->rf, c->rs, c->r, c->lf[b], c->ls[b]);
        if (!c->emit)
            c->offsets[chunk] = anti_join_range(c->lf, c->ls, b, e, c->rf, c->rs, c->r, start, NULL, NULL);
        else
            anti_join_range(c->lf, c->ls, b, e, c->rf, c->rs, c->r, start, c->of + c->offsets[chunk],
                            c->os + c->offsets[chunk]);
    }
}

uint64_t oracle_anti_join(const uint32_t* lf, const uint32_t* ls, uint64_t l, const uint32_t* rf,
                          const uint32_t* rs, uint64_t r, uint32_t* of, uint32_t* os, int threads) {
    if (l == 0) return 0;
    if (threads <= 1) return anti_join_range(lf, ls, 0, l, rf, rs, r, 0, of, os);
    int ws = clamp_workers(threads, l);
    uint64_t chunk_size = div_ceil(l, (uint64_t)ws);
    uint64_t chunk_count = div_ceil(l, chunk_size);
    uint64_t* offsets = (uint64_t*)malloc(chunk_count * sizeof(uint64_t));
    anti_ctx c = {lf, ls, rf, rs, l, r, chunk_size, chunk_count, offsets, of, os, 0};
    run_workers(ws, anti_worker, &c);
    uint64_t total = 0;
    for (uint64_t i = 0; i < chunk_count; ++i) {
        uint64_t v = offsets[i];
        offsets[i] = total;
        total += v;
    }
    c.emit = 1;
    run_workers(ws, anti_worker, &c);
    free(offsets);
    return total;
}

/* ------------------------------------------------------------------ */
/* union: execution/src/lib.rs:914-1066                                */

static uint64_t union_serial(const uint32_t* lf, const uint32_t* ls, uint64_t l, const uint32_t* rf,
                             const uint32_t* rs, uint64_t r, uint32_t* of, uint32_t* os) {
    uint64_t li = 0, ri = 0, written = 0;
    while (li < l || ri < r) {
        uint32_t t0, t1;
        if (ri == r || (li < l && compare_pair(lf[li], ls[li], rf[ri], rs[ri]) <= 0)) {
            t0 = lf[li];
            t1 = ls[li];
            ++li;
        } else {
            t0 = rf[ri];
            t1 = rs[ri];
            ++ri;
        }
        if (written == 0 || of[written - 1] != t0 || os[written - 1] != t1) {
            of[written] = t0;
            os[written] = t1;
            ++written;
        }
    }
    return written;
}

/* execution/src/lib.rs:981-997 */
static uint64_t merge_partition(const uint64_t* left, uint64_t l, const uint64_t* right, uint64_t r,
                                uint64_t pos) {
    uint64_t low = pos > r ? pos - r : 0;
    uint64_t high = pos < l ? pos : l;
    while (low < high) {
        uint64_t lp = low + (high - low) / 2;
        uint64_t rp = pos - lp;
        if (lp < l && rp > 0 && right[rp - 1] > left[lp])
            low = lp + 1;
        else
            high = lp;
    }
    return low;
}

typedef struct {
    const uint64_t *left, *right;
    uint64_t l, r, total, chunk;
    uint64_t* merged;
} merge_ctx;

/* execution/src/lib.rs:999-1012 */
static void merge_worker(void* p, int w, int ws) {
    merge_ctx* c = (merge_ctx*)p;
    uint64_t chunks = div_ceil(c->total, c->chunk);
    for (uint64_t k = (uint64_t)w; k < chunks; k += (uint64_t)ws) {
        uint64_t start = k * c->chunk;
        uint64_t end = start + c->chunk < c->total ? start + c->chunk : c->total;
        uint64_t li = merge_partition(c->left, c->l, c->right, c->r, start);
        uint64_t ri = start - li;
        for (uint64_t o = start; o < end; ++o) {
            if (ri == c->r || (li < c->l && c->left[li] <= c->right[ri]))
                c->merged[o] = c->left[li++];
            else
                c->merged[o] = c->right[ri++];
        }
    }
}

uint64_t oracle_union(const uint32_t* lf, const uint32_t* ls, uint64_t l, const uint32_t* rf,
                      const uint32_t* rs, uint64_t r, uint32_t* of, uint32_t* os, int threads) {
    uint64_t total = l + r;
    if (total == 0) return 0;
    if (threads <= 1) return union_serial(lf, ls, l, rf, rs, r, of, os);
    int ws = clamp_workers(threads, total);
    uint64_t* pl = (uint64_t*)malloc((l ? l : 1) * sizeof(uint64_t));
    uint64_t* pr = (uint64_t*)malloc((r ? r : 1) * sizeof(uint64_t));
    uint64_t* merged = (uint64_t*)malloc(total * sizeof(uint64_t));
    pack_ctx a = {lf, ls, pl, NULL, NULL, l};
    pack_ctx b = {rf, rs, pr, NULL, NULL, r};
    if (l) run_workers(clamp_workers(threads, l), pack_worker, &a);
    if (r) run_workers(clamp_workers(threads, r), pack_worker, &b);
    merge_ctx m = {pl, pr, l, r, total, div_ceil(total, (uint64_t)ws), merged};
    run_workers(ws, merge_worker, &m);
    uint64_t unique = deduplicate_sorted(merged, total);
    pack_ctx u = {NULL, NULL, merged, of, os, unique};
    run_workers(clamp_workers(threads, unique), unpack_worker, &u);
    free(pl);
    free(pr);
    free(merged);
    return unique;
}

/* ------------------------------------------------------------------ */
/* range index: storage/src/lib.rs:357-393                             */

uint64_t oracle_range_index(const uint32_t* values, uint64_t n, uint32_t* keys, uint32_t* starts,
                            uint32_t* rows) {
    if (n == 0) {
        if (starts) starts[0] = 0;
        return 0;
    }
    /* rows sorted by (value, row): pack and sort. */
    uint64_t* packed = (uint64_t*)malloc(n * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) packed[i] = pack_tuple(values[i], (uint32_t)i);
    oracle_sort_u64(packed, n, 1);
    uint64_t k = 0;
    for (uint64_t pos = 0; pos < n; ++pos) {
        uint32_t key = (uint32_t)(packed[pos] >> 32);
        if (rows) rows[pos] = (uint32_t)packed[pos];
        if (pos == 0 || (uint32_t)(packed[pos - 1] >> 32) != key) {
            if (keys) keys[k] = key;
            if (starts) starts[k] = (uint32_t)pos;
            ++k;
        }
    }
    if (starts) starts[k] = (uint32_t)n;
    free(packed);
    return k;
}

/* ------------------------------------------------------------------ */
/* join: execution/src/lib.rs:1206-1423                                */

/* IndexView::lookup, execution/src/lib.rs:1222-1227 */
static int64_t find_key(const uint32_t* keys, uint64_t k, uint32_t key) {
    uint64_t low = 0, high = k;
    while (low < high) {
        uint64_t mid = low + (high - low) / 2;
        if (keys[mid] < key)
            low = mid + 1;
        else
            high = mid;
    }
    return (low < k && keys[low] == key) ? (int64_t)low : -1;
}

/* count_join_rows, execution/src/lib.rs:1309-1337: per-row match counts
 * (in parallel chunks when threads > 1, as the par_iter of :1317-1320), then
 * a serial exclusive prefix into offsets (when non-null); returns the total. */
typedef struct {
    const uint32_t *left_keys, *keys, *starts;
    uint64_t l, k, chunk;
    uint64_t* counts;
} count_ctx;

static void join_count_worker(void* p, int w, int ws) {
    count_ctx* c = (count_ctx*)p;
    uint64_t chunks = div_ceil(c->l, c->chunk);
    for (uint64_t ch = (uint64_t)w; ch < chunks; ch += (uint64_t)ws) {
        uint64_t b = ch * c->chunk, e = b + c->chunk < c->l ? b + c->chunk : c->l;
        for (uint64_t i = b; i < e; ++i) {
            int64_t idx = find_key(c->keys, c->k, c->left_keys[i]);
            c->counts[i] = idx >= 0 ? (uint64_t)(c->starts[idx + 1] - c->starts[idx]) : 0;
        }
    }
}

static uint64_t count_join_rows(const uint32_t* left_keys, uint64_t l, const uint32_t* index_keys,
                                const uint32_t* index_starts, uint64_t k, uint64_t* offsets, int threads) {
    if (l == 0) return 0;
    uint64_t* counts = (uint64_t*)malloc(l * sizeof(uint64_t));
    int ws = clamp_workers(threads, l);
    count_ctx c = {left_keys, index_keys, index_starts, l, k, div_ceil(l, (uint64_t)ws), counts};
    run_workers(ws, join_count_worker, &c);
    uint64_t total = 0;
    for (uint64_t i = 0; i < l; ++i) {
        if (offsets) offsets[i] = total;
        total += counts[i];
    }
    free(counts);
    return total;
}

uint64_t oracle_join_count(const uint32_t* left_keys, uint64_t l, const uint32_t* index_keys,
                           const uint32_t* index_starts, uint64_t k, int threads) {
    return count_join_rows(left_keys, l, index_keys, index_starts, k, NULL, threads);
}

typedef struct {
    const uint32_t *left_keys, *keys, *starts, *rows, *p0, *p1;
    uint32_t side0, side1;
    uint64_t l, k, chunk;
    const uint64_t* offsets;
    uint32_t *out0, *out1;
} join_ctx;

static void join_emit_worker(void* p, int w, int ws) {
    join_ctx* c = (join_ctx*)p;
    uint64_t chunks = div_ceil(c->l, c->chunk);
    for (uint64_t ch = (uint64_t)w; ch < chunks; ch += (uint64_t)ws) {
        uint64_t b = ch * c->chunk, e = b + c->chunk < c->l ? b + c->chunk : c->l;
        for (uint64_t left_row = b; left_row < e; ++left_row) {
            int64_t idx = find_key(c->keys, c->k, c->left_keys[left_row]);
            if (idx < 0) continue;
            uint64_t out = c->offsets[left_row];
            for (uint32_t m = c->starts[idx]; m < c->starts[idx + 1]; ++m, ++out) {
                uint32_t right_row = c->rows[m];
                c->out0[out] = c->side0 == 0 ? c->p0[left_row] : c->p0[right_row];
                c->out1[out] = c->side1 == 0 ? c->p1[left_row] : c->p1[right_row];
            }
        }
    }
}

uint64_t oracle_join(const uint32_t* left_keys, uint64_t l, const uint32_t* index_keys,
                     const uint32_t* index_starts, const uint32_t* index_rows, uint64_t k,
                     const uint32_t* projection0, uint32_t side0, const uint32_t* projection1,
                     uint32_t side1, uint32_t* out0, uint32_t* out1, int threads) {
    if (l == 0) return 0;
    uint64_t* offsets = (uint64_t*)malloc(l * sizeof(uint64_t));
    uint64_t total = count_join_rows(left_keys, l, index_keys, index_starts, k, offsets, threads);
    int ws = clamp_workers(threads, l);
    join_ctx c = {left_keys, index_keys, index_starts, index_rows, projection0, projection1,
                  side0,     side1,      l,            k,          div_ceil(l, (uint64_t)ws),
                  offsets,   out0,       out1};
    run_workers(ws, join_emit_worker, &c);
    free(offsets);
    return total;
}

/* ------------------------------------------------------------------ */
/* filter: relational/src/lib.rs:155-164, execution/src/lib.rs:1681-1767 */

static int predicate_matches(uint32_t op, uint32_t operand, uint32_t value) {
    switch (op) {
        case 0: return value == operand;
        case 1: return value != operand;
        case 2: return value < operand;
        case 3: return value <= operand;
        case 4: return value > operand;
        case 5: return value >= operand;
        default: return 0;
    }
}

/* filter_cpu_parallel, execution/src/lib.rs:1705-1767: flags per chunk in
 * parallel (chunk = ceil(rows / workers)), per-chunk counts serially, then
 * each chunk writes its row ids at its offset in parallel. */
typedef struct {
    const uint32_t* input;
    uint32_t op, operand;
    uint64_t n, chunk;
    uint32_t* flags;
    const uint64_t* offsets;
    uint32_t* out_rows;
} filter_ctx;

static void filter_flag_worker(void* p, int w, int ws) {
    filter_ctx* c = (filter_ctx*)p;
    uint64_t chunks = div_ceil(c->n, c->chunk);
    for (uint64_t ch = (uint64_t)w; ch < chunks; ch += (uint64_t)ws) {
        uint64_t b = ch * c->chunk, e = b + c->chunk < c->n ? b + c->chunk : c->n;
        for (uint64_t i = b; i < e; ++i) c->flags[i] = (uint32_t)predicate_matches(c->op, c->operand, c->input[i]);
    }
}

static void filter_write_worker(void* p, int w, int ws) {
    filter_ctx* c = (filter_ctx*)p;
    uint64_t chunks = div_ceil(c->n, c->chunk);
    for (uint64_t ch = (uint64_t)w; ch < chunks; ch += (uint64_t)ws) {
        uint64_t b = ch * c->chunk, e = b + c->chunk < c->n ? b + c->chunk : c->n;
        uint64_t out = c->offsets[ch];
        for (uint64_t i = b; i < e; ++i)
            if (c->flags[i]) c->out_rows[out++] = (uint32_t)i;
    }
}

uint64_t oracle_filter(const uint32_t* input, uint64_t n, uint32_t op, uint32_t operand,
                       uint32_t* out_rows, int threads) {
    int ws = clamp_workers(threads, n);
    if (n == 0) return 0;
    if (ws <= 1) {
        /* filter_cpu_serial, execution/src/lib.rs:1681-1702 */
        uint64_t selected = 0;
        for (uint64_t row = 0; row < n; ++row)
            if (predicate_matches(op, operand, input[row])) out_rows[selected++] = (uint32_t)row;
        return selected;
    }
    uint64_t chunk = div_ceil(n, (uint64_t)ws), chunks = div_ceil(n, chunk);
    uint32_t* flags = (uint32_t*)malloc(n * sizeof(uint32_t));
    uint64_t* offsets = (uint64_t*)malloc(chunks * sizeof(uint64_t));
    filter_ctx c = {input, op, operand, n, chunk, flags, offsets, out_rows};
    run_workers(ws, filter_flag_worker, &c);
    uint64_t selected = 0;
    for (uint64_t ch = 0; ch < chunks; ++ch) {
        offsets[ch] = selected;
        uint64_t b = ch * chunk, e = b + chunk < n ? b + chunk : n;
        for (uint64_t i = b; i < e; ++i) selected += flags[i];
    }
    run_workers(ws, filter_write_worker, &c);
    free(flags);
    free(offsets);
    return selected;
}

/* ------------------------------------------------------------------ */
/* relation store + semi-naive executor: recursion/src/generic.rs       */

typedef struct {
    uint32_t* c[2];
    uint64_t len, cap;
} orel;

static int orel_reserve(orel* r, uint64_t n) {
    if (n <= r->cap) return 0;
    uint64_t cap = r->cap ? r->cap : 1;
    while (cap < n) cap *= 2;
    for (int i = 0; i < 2; ++i) {
        uint32_t* p = (uint32_t*)realloc(r->c[i], cap * sizeof(uint32_t));
        if (!p) return -1;
        r->c[i] = p;
    }
    r->cap = cap;
    return 0;
}

static void orel_free(orel* r) {
    free(r->c[0]);
    free(r->c[1]);
    memset(r, 0, sizeof(*r));
}

static int orel_assign(orel* r, const uint32_t* a, const uint32_t* b, uint64_t n) {
    if (orel_reserve(r, n)) return -1;
    if (n) {
        memcpy(r->c[0], a, n * sizeof(uint32_t));
        memcpy(r->c[1], b, n * sizeof(uint32_t));
    }
    r->len = n;
    return 0;
}

static void orel_swap(orel* a, orel* b) {
    orel t = *a;
    *a = *b;
    *b = t;
}

typedef struct {
    int present, recursive;
    orel full, delta, newt;
    uint64_t version;
} ostate;

struct oracle_store {
    ostate* rel;
    uint32_t n;
};

oracle_store* oracle_store_new(void) { return (oracle_store*)calloc(1, sizeof(oracle_store)); }

void oracle_store_free(oracle_store* s) {
    if (!s) return;
    for (uint32_t i = 0; i < s->n; ++i) {
        orel_free(&s->rel[i].full);
        orel_free(&s->rel[i].delta);
        orel_free(&s->rel[i].newt);
    }
    free(s->rel);
    free(s);
}

static ostate* store_slot(oracle_store* s, uint32_t id, int* err) {
    if (id >= s->n) {
        ostate* p = (ostate*)realloc(s->rel, sizeof(ostate) * (id + 1));
        if (!p) {
            *err = ORACLE_NO_MEMORY;
            return NULL;
        }
        memset(p + s->n, 0, sizeof(ostate) * (id + 1 - s->n));
        s->rel = p;
        s->n = id + 1;
    }
    if (s->rel[id].present) {
        *err = ORACLE_DUPLICATE_RELATION;
        return NULL;
    }
    return &s->rel[id];
}

/* RelationStore::insert_static, recursion/src/generic.rs:36-53 */
int oracle_store_insert_static(oracle_store* s, uint32_t id, const uint32_t* c0, const uint32_t* c1,
                               uint64_t n) {
    int err = 0;
    ostate* st = store_slot(s, id, &err);
    if (!st) return err;
    if (orel_assign(&st->full, c0, c1, n)) return ORACLE_NO_MEMORY;
    st->present = 1;
    st->recursive = 0;
    return ORACLE_OK;
}

/* RelationStore::insert_recursive, recursion/src/generic.rs:55-79 (no sorting) */
int oracle_store_insert_recursive(oracle_store* s, uint32_t id, const uint32_t* full0,
                                  const uint32_t* full1, uint64_t nfull, const uint32_t* delta0,
                                  const uint32_t* delta1, uint64_t ndelta) {
    int err = 0;
    ostate* st = store_slot(s, id, &err);
    if (!st) return err;
    if (orel_assign(&st->full, full0, full1, nfull) || orel_assign(&st->delta, delta0, delta1, ndelta))
        return ORACLE_NO_MEMORY;
    st->newt.len = 0;
    st->present = 1;
    st->recursive = 1;
    return ORACLE_OK;
}

uint64_t oracle_store_get(const oracle_store* s, uint32_t id, int which, uint32_t* c0, uint32_t* c1) {
    if (id >= s->n || !s->rel[id].present) return 0;
    const orel* r = which == 0 ? &s->rel[id].full : &s->rel[id].delta;
    if (c0 && r->len) memcpy(c0, r->c[0], r->len * sizeof(uint32_t));
    if (c1 && r->len) memcpy(c1, r->c[1], r->len * sizeof(uint32_t));
    return r->len;
}

static ostate* store_get(oracle_store* s, uint32_t id) {
    return (id < s->n && s->rel[id].present) ? &s->rel[id] : NULL;
}

typedef struct {
    uint32_t relation;
    uint64_t version;
    uint32_t column;
    int valid;
    uint32_t *keys, *starts, *rows;
    uint64_t k;
} cached_index;

typedef struct {
    const oracle_rule* plan;
    cached_index index;
    orel candidates, distinct;
} rule_rt;

typedef struct {
    uint32_t target;
    uint32_t* rule_indices;
    uint32_t n_rules;
    orel contributions, union_tmp, newt;
} target_rt;

/* Selected input columns (relation_column, execution/src/lib.rs:1258-1266). */
static int pick(const orel* r, uint32_t col, const uint32_t** out) {
    if (col > 1) return ORACLE_MISSING_COLUMN;
    *out = r->c[col];
    return ORACLE_OK;
}

/* RuleRuntime::evaluate, recursion/src/generic.rs:197-259 */
static int rule_evaluate(rule_rt* rt, oracle_store* s, int threads) {
    const oracle_rule* p = rt->plan;
    ostate* ds = store_get(s, p->delta_input);
    ostate* rs = store_get(s, p->right_input);
    if (!ds || !rs) return ORACLE_MISSING_RELATION;
    if (!ds->recursive) return ORACLE_NOT_RECURSIVE;
    const orel* delta = &ds->delta;
    const orel* right = &rs->full;
    const uint32_t* right_column;
    if (pick(right, p->right_key, &right_column)) return ORACLE_MISSING_COLUMN;
    if (!rt->index.valid || rt->index.relation != p->right_input || rt->index.version != rs->version ||
        rt->index.column != p->right_key) {
        free(rt->index.keys);
        free(rt->index.starts);
        free(rt->index.rows);
        uint64_t n = right->len;
        rt->index.keys = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
        rt->index.starts = (uint32_t*)malloc((n + 1) * sizeof(uint32_t));
        rt->index.rows = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
        rt->index.k = oracle_range_index(right_column, n, rt->index.keys, rt->index.starts, rt->index.rows);
        rt->index.relation = p->right_input;
        rt->index.version = rs->version;
        rt->index.column = p->right_key;
        rt->index.valid = 1;
    }
    const uint32_t* left_key;
    if (pick(delta, p->left_key, &left_key)) return ORACLE_MISSING_COLUMN;
    const uint32_t* proj[2];
    for (int i = 0; i < 2; ++i)
        if (pick(p->out_input[i] == 0 ? delta : right, p->out_column[i], &proj[i])) return ORACLE_MISSING_COLUMN;
    uint64_t m = oracle_join_count(left_key, delta->len, rt->index.keys, rt->index.starts, rt->index.k, threads);
    if (orel_reserve(&rt->candidates, m)) return ORACLE_NO_MEMORY;
    rt->candidates.len = oracle_join(left_key, delta->len, rt->index.keys, rt->index.starts,
                                     rt->index.rows, rt->index.k, proj[0], p->out_input[0], proj[1],
                                     p->out_input[1], rt->candidates.c[0], rt->candidates.c[1], threads);
    const uint32_t *d0, *d1;
    if (pick(&rt->candidates, p->distinct_columns[0], &d0) || pick(&rt->candidates, p->distinct_columns[1], &d1))
        return ORACLE_MISSING_COLUMN;
    if (orel_reserve(&rt->distinct, rt->candidates.len)) return ORACLE_NO_MEMORY;
    rt->distinct.len = oracle_distinct(d0, d1, rt->candidates.len, rt->distinct.c[0], rt->distinct.c[1], threads);
    return ORACLE_OK;
}

static int plan_union(const orel* l, const orel* r, const uint32_t lc[2], const uint32_t rc[2], orel* out,
                      int threads) {
    const uint32_t *l0, *l1, *r0, *r1;
    if (pick(l, lc[0], &l0) || pick(l, lc[1], &l1) || pick(r, rc[0], &r0) || pick(r, rc[1], &r1))
        return ORACLE_MISSING_COLUMN;
    if (orel_reserve(out, l->len + r->len)) return ORACLE_NO_MEMORY;
    out->len = oracle_union(l0, l1, l->len, r0, r1, r->len, out->c[0], out->c[1], threads);
    return ORACLE_OK;
}

static int plan_anti(const orel* l, const orel* r, const uint32_t lc[2], const uint32_t rc[2], orel* out,
                     int threads) {
    const uint32_t *l0, *l1, *r0, *r1;
    if (pick(l, lc[0], &l0) || pick(l, lc[1], &l1) || pick(r, rc[0], &r0) || pick(r, rc[1], &r1))
        return ORACLE_MISSING_COLUMN;
    if (orel_reserve(out, l->len)) return ORACLE_NO_MEMORY;
    out->len = oracle_anti_join(l0, l1, l->len, r0, r1, r->len, out->c[0], out->c[1], threads);
    return ORACLE_OK;
}

static int same_pair(const uint32_t a[2], const uint32_t b[2]) { return a[0] == b[0] && a[1] == b[1]; }

int oracle_run(oracle_store* s, const uint32_t* relations, uint32_t n_relations, const oracle_rule* rules,
               uint32_t n_rules, uint64_t max_iterations, int threads, oracle_summary* summary) {
    /* RecursiveExecutor::compile, recursion/src/generic.rs:407-470 */
    if (n_relations == 0) return ORACLE_EMPTY_SCC;
    for (uint32_t i = 0; i < n_relations; ++i) {
        for (uint32_t j = 0; j < i; ++j)
            if (relations[j] == relations[i]) return ORACLE_DUPLICATE_RELATION;
        ostate* st = store_get(s, relations[i]);
        if (!st) return ORACLE_MISSING_RELATION;
        if (!st->recursive) return ORACLE_NOT_RECURSIVE;
    }
    for (uint32_t r = 0; r < n_rules; ++r) {
        if (!store_get(s, rules[r].delta_input) || !store_get(s, rules[r].right_input))
            return ORACLE_MISSING_RELATION;
        int delta_in = 0, target_in = 0;
        for (uint32_t i = 0; i < n_relations; ++i) {
            delta_in |= relations[i] == rules[r].delta_input;
            target_in |= relations[i] == rules[r].target;
        }
        if (!delta_in) return ORACLE_DELTA_OUTSIDE_SCC;
        if (!target_in) return ORACLE_TARGET_OUTSIDE_SCC;
    }
    int rc = ORACLE_OK;
    target_rt* targets = (target_rt*)calloc(n_relations, sizeof(target_rt));
    rule_rt* rts = (rule_rt*)calloc(n_rules ? n_rules : 1, sizeof(rule_rt));
    for (uint32_t i = 0; i < n_relations && rc == ORACLE_OK; ++i) {
        targets[i].target = relations[i];
        targets[i].rule_indices = (uint32_t*)malloc(sizeof(uint32_t) * (n_rules ? n_rules : 1));
        for (uint32_t r = 0; r < n_rules; ++r)
            if (rules[r].target == relations[i]) targets[i].rule_indices[targets[i].n_rules++] = r;
        if (targets[i].n_rules == 0) {
            rc = ORACLE_MISSING_RULE_TARGET;
            break;
        }
        const oracle_rule* first = &rules[targets[i].rule_indices[0]];
        for (uint32_t k = 1; k < targets[i].n_rules; ++k) {
            const oracle_rule* o = &rules[targets[i].rule_indices[k]];
            if (!same_pair(o->anti_left, first->anti_left) || !same_pair(o->anti_right, first->anti_right) ||
                !same_pair(o->union_left, first->union_left) || !same_pair(o->union_right, first->union_right))
                rc = ORACLE_INCOMPATIBLE_PLANS;
        }
    }
    for (uint32_t r = 0; r < n_rules; ++r) rts[r].plan = &rules[r];

    /* RecursiveExecutor::run, recursion/src/generic.rs:472-516 */
    uint64_t iterations = 0, total_new = 0;
    while (rc == ORACLE_OK) {
        uint64_t delta_rows = 0;
        for (uint32_t i = 0; i < n_relations; ++i) delta_rows += store_get(s, relations[i])->delta.len;
        if (delta_rows == 0) break;
        if (iterations >= max_iterations) {
            if (summary) summary->delta_rows_left = delta_rows;
            rc = ORACLE_ITERATION_LIMIT;
            break;
        }
        for (uint32_t r = 0; r < n_rules && rc == ORACLE_OK; ++r) rc = rule_evaluate(&rts[r], s, threads);
        /* TargetRuntime::combine, recursion/src/generic.rs:299-358 */
        for (uint32_t i = 0; i < n_relations && rc == ORACLE_OK; ++i) {
            target_rt* t = &targets[i];
            const oracle_rule* plan0 = &rules[t->rule_indices[0]];
            for (uint32_t k = 0; k < t->n_rules && rc == ORACLE_OK; ++k) {
                rule_rt* rt = &rts[t->rule_indices[k]];
                if (k == 0) {
                    orel_swap(&rt->distinct, &t->contributions);
                    continue;
                }
                rc = plan_union(&t->contributions, &rt->distinct, plan0->union_left, plan0->union_right,
                                &t->union_tmp, threads);
                orel_swap(&t->union_tmp, &t->contributions);
            }
            if (rc != ORACLE_OK) break;
            ostate* target = store_get(s, t->target);
            rc = plan_anti(&t->contributions, &target->full, plan0->anti_left, plan0->anti_right, &t->newt,
                           threads);
        }
        /* TargetRuntime::apply, recursion/src/generic.rs:360-397 */
        for (uint32_t i = 0; i < n_relations && rc == ORACLE_OK; ++i) {
            target_rt* t = &targets[i];
            const oracle_rule* plan0 = &rules[t->rule_indices[0]];
            ostate* target = store_get(s, t->target);
            rc = plan_union(&target->full, &t->newt, plan0->union_left, plan0->union_right, &t->union_tmp,
                            threads);
            if (rc != ORACLE_OK) break;
            orel_swap(&t->newt, &target->newt);
            orel_swap(&t->union_tmp, &target->full);
            orel_swap(&target->delta, &target->newt);
            target->version += 1;
            total_new += target->delta.len;
        }
        if (rc == ORACLE_OK) ++iterations;
    }
    if (summary) {
        summary->iterations = iterations;
        summary->total_new_rows = total_new;
    }
    for (uint32_t i = 0; i < n_relations; ++i) {
        free(targets[i].rule_indices);
        orel_free(&targets[i].contributions);
        orel_free(&targets[i].union_tmp);
        orel_free(&targets[i].newt);
    }
    for (uint32_t r = 0; r < n_rules; ++r) {
        free(rts[r].index.keys);
        free(rts[r].index.starts);
        free(rts[r].index.rows);
        orel_free(&rts[r].candidates);
        orel_free(&rts[r].distinct);
    }
    free(targets);
    free(rts);
    return rc;
}

/* ------------------------------------------------------------------ */
/* brute-force closure by BFS                                          */

typedef struct {
    const uint64_t* offs;
    const uint32_t* dst;
    uint32_t n_nodes;
    uint64_t* per_source;
    uint64_t* src_offsets; /* when writing the closure */
    uint32_t *cf, *cs;
    pthread_mutex_t lock;
    uint32_t next_source;
    uint64_t pairs, digest, max_dist;
} bfs_ctx;

static int cmp_u32(const void* a, const void* b) {
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return x < y ? -1 : x > y;
}

static void bfs_worker(void* p, int w, int ws) {
    (void)w;
    (void)ws;
    bfs_ctx* c = (bfs_ctx*)p;
    uint32_t n = c->n_nodes;
    uint32_t* stamp = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
    uint32_t* queue = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    uint64_t pairs = 0, digest = 0, max_dist = 0;
    for (;;) {
        pthread_mutex_lock(&c->lock);
        uint32_t begin = c->next_source;
        uint32_t end = begin + 64 < n ? begin + 64 : n;
        c->next_source = end;
        pthread_mutex_unlock(&c->lock);
        if (begin >= n) break;
        for (uint32_t s = begin; s < end; ++s) {
            uint32_t mark = s + 1;
            uint64_t head = 0, tail = 0, dist = 0;
            for (uint64_t e = c->offs[s]; e < c->offs[s + 1]; ++e) {
                uint32_t v = c->dst[e];
                if (stamp[v] != mark) {
                    stamp[v] = mark;
                    queue[tail++] = v;
                }
            }
            /* level-synchronous BFS: level `dist` spans queue[head, level_end) */
            while (head < tail) {
                ++dist;
                uint64_t level_end = tail;
                for (; head < level_end; ++head) {
                    uint32_t u = queue[head];
                    for (uint64_t e = c->offs[u]; e < c->offs[u + 1]; ++e) {
                        uint32_t v = c->dst[e];
                        if (stamp[v] != mark) {
                            stamp[v] = mark;
                            queue[tail++] = v;
                        }
                    }
                }
            }
            if (dist > max_dist) max_dist = dist;
            pairs += tail;
            for (uint64_t i = 0; i < tail; ++i) digest += mix64(pack_tuple(s, queue[i]));
            if (c->per_source) c->per_source[s] = tail;
            if (c->cf) {
                qsort(queue, tail, sizeof(uint32_t), cmp_u32);
                uint64_t o = c->src_offsets[s];
                for (uint64_t i = 0; i < tail; ++i) {
                    c->cf[o + i] = s;
                    c->cs[o + i] = queue[i];
                }
            }
        }
    }
    pthread_mutex_lock(&c->lock);
    c->pairs += pairs;
    c->digest += digest;
    if (max_dist > c->max_dist) c->max_dist = max_dist;
    pthread_mutex_unlock(&c->lock);
    free(stamp);
    free(queue);
}

int oracle_tc_bfs(const uint32_t* src, const uint32_t* dst, uint64_t edges, uint32_t n_nodes, int threads,
                  oracle_bfs_result* result, uint64_t* per_source, uint32_t* closure_first,
                  uint32_t* closure_second, uint64_t capacity) {
    uint64_t* offs = (uint64_t*)calloc((size_t)n_nodes + 1, sizeof(uint64_t));
    if (!offs) return ORACLE_NO_MEMORY;
    for (uint64_t e = 0; e < edges; ++e) {
        if (src[e] >= n_nodes || dst[e] >= n_nodes) {
            free(offs);
            return ORACLE_MISSING_COLUMN;
        }
        offs[src[e] + 1]++;
    }
    for (uint32_t v = 0; v < n_nodes; ++v) offs[v + 1] += offs[v];
    /* edges arrive sorted by (src, dst); place dst by src in case they are not */
    uint32_t* adj = (uint32_t*)malloc((edges ? edges : 1) * sizeof(uint32_t));
    uint64_t* fill = (uint64_t*)malloc(((size_t)n_nodes + 1) * sizeof(uint64_t));
    memcpy(fill, offs, ((size_t)n_nodes + 1) * sizeof(uint64_t));
    for (uint64_t e = 0; e < edges; ++e) adj[fill[src[e]]++] = dst[e];
    free(fill);

    int ws = threads < 1 ? 1 : threads;
    bfs_ctx c;
    memset(&c, 0, sizeof(c));
    c.offs = offs;
    c.dst = adj;
    c.n_nodes = n_nodes;
    pthread_mutex_init(&c.lock, NULL);
    uint64_t* counts = per_source;
    int own_counts = 0;
    if (closure_first && !counts) {
        counts = (uint64_t*)malloc(((size_t)n_nodes + 1) * sizeof(uint64_t));
        own_counts = 1;
    }
    c.per_source = counts;
    run_workers(ws, bfs_worker, &c);
    if (closure_first && c.pairs <= capacity) {
        uint64_t* so = (uint64_t*)malloc(((size_t)n_nodes + 1) * sizeof(uint64_t));
        uint64_t acc = 0;
        for (uint32_t v = 0; v < n_nodes; ++v) {
            so[v] = acc;
            acc += counts[v];
        }
        bfs_ctx d = c;
        d.next_source = 0;
        d.pairs = d.digest = d.max_dist = 0;
        d.src_offsets = so;
        d.cf = closure_first;
        d.cs = closure_second;
        d.per_source = NULL;
        pthread_mutex_init(&d.lock, NULL);
        run_workers(ws, bfs_worker, &d);
        pthread_mutex_destroy(&d.lock);
        free(so);
    }
    if (result) {
        result->pairs = c.pairs;
        result->digest = c.digest;
        result->max_dist = c.max_dist;
    }
    pthread_mutex_destroy(&c.lock);
    if (own_counts) free(counts);
    free(offs);
    free(adj);
    return ORACLE_OK;
}

/* ------------------------------------------------------------------ */
/* general CPU path (datalog/src/general.rs:391-566) for TC + nonlocal */

typedef struct {
    uint64_t* v;
    uint64_t n, cap;
} u64vec;

static void vpush(u64vec* a, uint64_t x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? a->cap * 2 : 64;
        a->v = (uint64_t*)realloc(a->v, a->cap * sizeof(uint64_t));
    }
    a->v[a->n++] = x;
}

static void vsort_unique(u64vec* a) {
    oracle_sort_u64(a->v, a->n, 1);
    a->n = deduplicate_sorted(a->v, a->n);
}

static int vcontains(const u64vec* a, uint64_t x) {
    uint64_t lo = 0, hi = a->n;
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (a->v[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < a->n && a->v[lo] == x;
}

uint64_t oracle_general_tc(const uint32_t* src, const uint32_t* dst, uint64_t edges, uint64_t* nonlocal_rows,
                           uint64_t* iterations, uint64_t* new_rows, uint32_t* path_first,
                           uint32_t* path_second, uint32_t* nonlocal_first, uint32_t* nonlocal_second,
                           uint64_t capacity) {
    u64vec edge = {0}, full = {0}, delta = {0};
    for (uint64_t e = 0; e < edges; ++e) vpush(&edge, pack_tuple(src[e], dst[e]));
    vsort_unique(&edge); /* BTreeSet facts, general.rs:377-379 */
    /* seed clause path(x,y) :- edge(x,y): full = edge; delta = full (general.rs:396-405) */
    for (uint64_t i = 0; i < edge.n; ++i) vpush(&full, edge.v[i]);
    for (uint64_t i = 0; i < full.n; ++i) vpush(&delta, full.v[i]);
    uint64_t iters = 0, total_new = 0;
    while (delta.n) {
        /* path(x,z) :- path_delta(x,y), edge(y,z): one binding per delta tuple,
         * each scanning every edge tuple (general.rs:491-508). */
        u64vec next = {0};
        for (uint64_t b = 0; b < delta.n; ++b) {
            uint32_t x = (uint32_t)(delta.v[b] >> 32), y = (uint32_t)delta.v[b];
            for (uint64_t e = 0; e < edge.n; ++e) {
                if ((uint32_t)(edge.v[e] >> 32) != y) continue;
                uint64_t t = pack_tuple(x, (uint32_t)edge.v[e]);
                if (!vcontains(&full, t)) vpush(&next, t); /* general.rs:428-432 */
            }
        }
        vsort_unique(&next);
        total_new += next.n;
        /* full.extend(next) (general.rs:435-439) */
        u64vec merged = {0};
        for (uint64_t i = 0; i < full.n; ++i) vpush(&merged, full.v[i]);
        for (uint64_t i = 0; i < next.n; ++i) vpush(&merged, next.v[i]);
        vsort_unique(&merged);
        free(full.v);
        full = merged;
        free(delta.v);
        delta = next;
        ++iters;
    }
    /* nonlocal(x,z) :- path(x,z), !edge(x,z): negation scans every edge tuple
     * per binding (general.rs:522-536). */
    u64vec nonlocal = {0};
    for (uint64_t b = 0; b < full.n; ++b) {
        int found = 0;
        for (uint64_t e = 0; e < edge.n && !found; ++e) found = edge.v[e] == full.v[b];
        if (!found) vpush(&nonlocal, full.v[b]);
    }
    if (path_first && full.n <= capacity)
        for (uint64_t i = 0; i < full.n; ++i) {
            path_first[i] = (uint32_t)(full.v[i] >> 32);
            path_second[i] = (uint32_t)full.v[i];
        }
    if (nonlocal_first && nonlocal.n <= capacity)
        for (uint64_t i = 0; i < nonlocal.n; ++i) {
            nonlocal_first[i] = (uint32_t)(nonlocal.v[i] >> 32);
            nonlocal_second[i] = (uint32_t)nonlocal.v[i];
        }
    uint64_t path_rows = full.n;
    if (nonlocal_rows) *nonlocal_rows = nonlocal.n;
    if (iterations) *iterations = iters;
    if (new_rows) *new_rows = total_new;
    free(edge.v);
    free(full.v);
    free(delta.v);
    free(nonlocal.v);
    return path_rows;
}

/* ---- text loader (storage/src/io.rs:72-141) ---- */

static int ascii_ws(unsigned char c) { return c == ' ' || (c >= 9 && c <= 13); }

/* One line [s, e) after str::trim; 0 when it is not a data line (io.rs:86-89). */
static int data_line_span(const char* t, uint64_t* s, uint64_t* e) {
    while (*s < *e && ascii_ws((unsigned char)t[*s])) ++*s;
    while (*e > *s && ascii_ws((unsigned char)t[*e - 1])) --*e;
    return *s < *e && t[*s] != '#' && t[*s] != '%';
}

/* str::parse::<u32>: an optional '+', then at least one digit, no overflow. */
static int rust_parse_u32(const char* t, uint64_t s, uint64_t e, uint32_t* out) {
    uint64_t v = 0;
    if (s < e && t[s] == '+') ++s;
    if (s == e) return 0;
    for (; s < e; ++s) {
        if (t[s] < '0' || t[s] > '9') return 0;
        v = v * 10 + (uint64_t)(t[s] - '0');
        if (v > 0xFFFFFFFFull) return 0;
    }
    *out = (uint32_t)v;
    return 1;
}

uint64_t oracle_load_binary_u32(const char* text, uint64_t len, uint32_t* first, uint32_t* second,
                                uint64_t capacity, oracle_load_error* error) {
    static const char lfs[] = "version https://git-lfs.github.com/spec/v1";
    const uint64_t lfs_len = sizeof(lfs) - 1;
    memset(error, 0, sizeof(*error));
    /* count_rows (io.rs:91-103): a pointer line anywhere fails the load */
    uint64_t line = 0;
    for (uint64_t pos = 0; pos < len; ++line) {
        uint64_t end = pos;
        while (end < len && text[end] != '\n') ++end;
        uint64_t s = pos, e = end;
        if (data_line_span(text, &s, &e) && e - s == lfs_len && memcmp(text + s, lfs, lfs_len) == 0) {
            error->kind = ORACLE_LOAD_LFS_POINTER;
            error->line = line + 1;
            return 0;
        }
        pos = end + 1;
    }
    /* populate_columns (io.rs:105-131) */
    uint64_t rows = 0;
    line = 0;
    for (uint64_t pos = 0; pos < len; ++line) {
        uint64_t end = pos;
        while (end < len && text[end] != '\n') ++end;
        uint64_t s = pos, e = end;
        pos = end + 1;
        if (!data_line_span(text, &s, &e)) continue;
        uint64_t a_end = s;
        while (a_end < e && !ascii_ws((unsigned char)text[a_end])) ++a_end;
        uint64_t b = a_end;
        while (b < e && ascii_ws((unsigned char)text[b])) ++b;
        uint64_t b_end = b;
        while (b_end < e && !ascii_ws((unsigned char)text[b_end])) ++b_end;
        if (b == e || b_end != e) {
            error->kind = ORACLE_LOAD_FIELD_COUNT;
            error->line = line + 1;
            return 0;
        }
        uint32_t x, y;
        if (!rust_parse_u32(text, s, a_end, &x) || !rust_parse_u32(text, b, b_end, &y)) {
            const int col1 = !rust_parse_u32(text, s, a_end, &x);
            error->kind = ORACLE_LOAD_INVALID_U32;
            error->line = line + 1;
            error->column = col1 ? 1 : 2;
            error->value_offset = col1 ? s : b;
            error->value_len = col1 ? a_end - s : b_end - b;
            return 0;
        }
        if (rows >= capacity) {
            error->kind = ORACLE_LOAD_CAPACITY;
            return 0;
        }
        first[rows] = x;
        second[rows] = y;
        ++rows;
    }
    return rows;
}
