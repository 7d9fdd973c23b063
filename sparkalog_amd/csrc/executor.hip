// Device-resident relation store and semi-naive executor (part 2 of
// include/sparkalog_hip.h).
//
// Restates RelationStore / RuleRuntime / TargetRuntime / RecursiveExecutor of
// crates/sparkalog-recursion/src/generic.rs:16-534 with every relation kept in
// HBM between iterations:
//   * a relation is one packed u64 array (first<<32 | second), so the
//     lexicographic pair order is the integer order of the keys;
//   * the join's right input is indexed on the device: rows sorted by
//     (key, other) plus a dense CSR start array over key values, rebuilt only
//     when the right relation's version changes (generic.rs:217-229);
//   * row counts stay on the device inside an iteration; the host reads them
//     twice per iteration (join output sizes, then the new DELTA/FULL sizes),
//     which is also the termination test of generic.rs:518-525.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ops.hpp"
#include "prims.hpp"
#include "sparkalog_hip.h"

namespace spk {
namespace {

thread_local int g_last_hip_error = 0;

struct HipFailure {
    hipError_t code;
};

inline void check(hipError_t e) {
    if (e != hipSuccess) {
        g_last_hip_error = static_cast<int>(e);
        throw HipFailure{e};
    }
}

struct StatusError {
    int code;
};

// Stores keep released buffers pooled: a fresh multi-GB hipMalloc costs
// tens of milliseconds per GB here. When hipMalloc fails, g_relief frees
// pooled buffers of every store (smallest first) until `bytes` are free and
// the allocation is retried; it returns false when no pooled buffer is left.
bool (*g_relief)(size_t bytes) = nullptr;
// guards every store's pools and the store registry
inline std::mutex& pool_mutex() {
    static std::mutex m;
    return m;
}

// Held while a graph is captured, and around every call that synchronizes
// the whole device (a free does): synchronizing a stream that another
// thread is capturing would invalidate that capture.
inline std::recursive_mutex& device_mutex() {
    static std::recursive_mutex m;
    return m;
}

__host__ __device__ inline size_t tiles(size_t n, size_t t) { return (n + t - 1) / t; }

inline uint64_t bit_mask_for(uint64_t max_value) {
    uint64_t m = 0;
    while (m < max_value) m = (m << 1) | 1;
    return m;
}

// Device buffer with capacity; grow() keeps no contents.
template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), cap(o.cap) {
        o.ptr = nullptr;
        o.cap = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        std::swap(ptr, o.ptr);
        std::swap(cap, o.cap);
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (ptr) {
            std::lock_guard<std::recursive_mutex> lock(device_mutex());
            // kernels queued on any stream may still use the buffer
            (void)hipDeviceSynchronize();
            (void)hipFree(ptr);
        }
        ptr = nullptr;
        cap = 0;
    }
    void reserve(size_t n) {
        if (n <= cap && ptr) return;
        size_t c = cap ? cap : 1024;
        while (c < n) c += c / 2 + 1;
        release();
        std::lock_guard<std::recursive_mutex> lock(device_mutex());
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, c * sizeof(T));
        while (e != hipSuccess && g_relief && g_relief(c * sizeof(T))) {
            (void)hipGetLastError();
            e = hipMalloc(&p, c * sizeof(T));
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            g_last_hip_error = static_cast<int>(e);
            throw StatusError{SPARKALOG_HIP_ERR_OUT_OF_MEMORY};
        }
        ptr = static_cast<T*>(p);
        cap = c;
    }
};

struct DevRelation {
    DevBuf<uint64_t> keys;
    uint64_t rows = 0;
    // first-column index (xstart, n_x + 1 entries) of a FULL part, valid
    // while xidx_n is the current n_x and keys/rows are those it was built on
    DevBuf<uint32_t> xidx;
    uint32_t xidx_n = 0;
    const uint64_t* xidx_keys = nullptr;
    uint64_t xidx_rows = 0;
    // second column of a FULL part (u32 per row), valid while keys/rows are
    // those it was written for
    DevBuf<uint32_t> sec;
    // a run of FULL held only as first-column index + second column (keys
    // empty): inside a fixpoint, runs of a fused target are kept this way
    bool csr = false;
    const uint64_t* sec_keys = nullptr;
    uint64_t sec_rows = 0;
    void touched() {
        xidx_n = 0;
        sec_keys = nullptr;
    }
    bool xidx_valid(uint32_t n_x) const { return xidx_n == n_x && xidx_keys == keys.ptr && xidx_rows == rows; }
    void xidx_set(uint32_t n_x) {
        xidx_n = n_x;
        xidx_keys = keys.ptr;
        xidx_rows = rows;
    }
    bool sec_valid() const { return csr || (sec_keys != nullptr && sec_keys == keys.ptr && sec_rows == rows); }
    void sec_set() {
        sec_keys = keys.ptr;
        sec_rows = rows;
    }
};

struct RelationState {
    bool present = false;
    bool recursive = false;
    // FULL = full ∪ runs[0] ∪ runs[1] ∪ ...: sorted unique arrays, pairwise
    // disjoint, sizes decreasing. Inside a run new rows land in small runs so
    // that FULL is not rewritten every iteration; outside a run FULL is one
    // array (runs is empty).
    DevRelation full, delta, newt;
    std::vector<DevRelation> runs;
    uint64_t version = 0;
    uint32_t max_value = 0;  // max over both columns of every row ever stored
    // a step's run merges are issued by the next step, behind its join
    // count, so they overlap the host's wait for the join sizes
    bool settle_pending = false;
};

// Per-op timing (hipEvents) and algorithmic bytes when profiling is on: the
// whole operator (helper kernels included) and, separately, its main kernel.
struct OpStat {
    const char* name;
    const char* kernel;
    uint64_t launches = 0;
    double ms = 0.0;
    double bytes = 0.0;
    uint64_t kernel_launches = 0;
    double kernel_ms = 0.0;
};

enum OpKind {
    kOpJoinCount = 0, kOpJoinEmit, kOpDistinct, kOpAntiJoin, kOpProbeSparse, kOpUnion, kOpMergeRuns, kOpIndex, kOpCount
};
const char* const kOpNames[kOpCount] = {"join_count", "join_emit", "distinct", "anti_join", "anti_join_sparse",
                                        "union", "merge_runs", "index"};
// (the segmented distinct's events span its three kernels: tile sort, count
// scan and copy to the offsets)
const char* const kOpKernels[kOpCount] = {"dense_join_count_kernel", "dense_join_emit_kernel",
                                          "seg_tile_kernel+seg_scan_kernel+seg_write_kernel",
                                          "anti_join_kernel", "probe_sparse_kernel",
                                          "union_tile_kernel+group_scan_kernel+union_tile_kernel",
                                          "merge_disjoint_kernel", ""};
// Candidates up to this many (by default) are probed one thread each
// through the first-column index of FULL's runs (few candidates, long FULL
// ranges).
constexpr uint64_t kSparseProbeRows = 1ull << 22;

// Events of one profiled operator launch.
struct OpEvents {
    hipEvent_t begin = nullptr, end = nullptr;
    KernelTimer kernel;
    bool kernel_used = false;
    int op = 0;
    double bytes = 0.0;
};

// ------------------------------------------------------------------------
// executor kernels

__global__ void gather_index_keys_kernel(const uint64_t* rows, size_t n, uint32_t key_col, uint64_t* out) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t r = rows[i];
        const uint32_t c0 = static_cast<uint32_t>(r >> 32), c1 = static_cast<uint32_t>(r);
        out[i] = key_col == 0 ? r : pack_pair(c1, c0);
    }
}

// starts[v] = first sorted row with key >= v, for v in [0, n_keys].
__global__ void low_halves_kernel(const uint64_t* in, size_t n, uint32_t* out) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = static_cast<uint32_t>(in[i]);
}

__global__ void dense_starts_kernel(const uint64_t* sorted, size_t n, uint32_t* starts, size_t n_keys) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t v = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; v <= n_keys; v += stride) {
        const uint64_t target = static_cast<uint64_t>(v) << 32;
        size_t lo = 0, hi = n;
        while (lo < hi) {
            const size_t mid = (lo + hi) >> 1;
            if (sorted[mid] < target)
                lo = mid + 1;
            else
                hi = mid;
        }
        starts[v] = static_cast<uint32_t>(lo);
    }
}

__global__ void max_degree_kernel(const uint32_t* starts, size_t n_keys, uint32_t* out) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    uint32_t m = 0;
    for (size_t v = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n_keys; v += stride) {
        const uint32_t d = starts[v + 1] - starts[v];
        m = d > m ? d : m;
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const uint32_t o = __shfl_xor(m, off, kWave);
        m = o > m ? o : m;
    }
    if (lane_id() == 0 && m) atomicMax(out, m);
}

__device__ __forceinline__ uint32_t column_of(uint64_t key, uint32_t col) {
    return col == 0 ? static_cast<uint32_t>(key >> 32) : static_cast<uint32_t>(key);
}

// ---- hashed join index ----
// A right relation whose key values are sparse (any u32 value, or far more
// possible values than rows) is indexed by its unique keys instead of a dense
// start array over every value: rows sorted by (key, other) as for the dense
// index, starts[r] = the first row of the r-th unique key (K + 1 entries),
// and an open-addressing table of (key << 32 | r) slots, linear probing,
// Fibonacci hashing, at most half full (so every probe sequence meets an
// empty slot).
constexpr uint64_t kEmptySlot = ~0ull;

__host__ __device__ __forceinline__ uint64_t hash_slot(uint32_t key, uint32_t bits) {
    return (static_cast<uint64_t>(key) * 0x9E3779B97F4A7C15ull) >> (64 - bits);
}

// The join index a count kernel reads: dense (starts over key values
// [0, n_keys]) when table is null, else hashed (starts over unique keys).
struct KeyIndex {
    const uint32_t* starts;
    uint32_t n_keys;
    const uint64_t* table;
    uint32_t bits;
};

constexpr int kHeadItems = 8;  // consecutive sorted rows per thread
constexpr int kHeadTile = kBlock * kHeadItems;

// Unique-key heads of n >= 1 sorted index rows (key in the high half). With
// write == 0 each block stores its head count in tile_counts[block]; with
// write == 1 (tile_counts then holding their exclusive prefix) each head row
// goes to starts[tile_counts[block] + its rank].
__global__ __launch_bounds__(kBlock) void key_heads_kernel(const uint64_t* sorted, size_t n, uint64_t* tile_counts,
                                                           uint32_t* starts, int write) {
    __shared__ uint64_t lds[kWavesPerBlock];
    const size_t first = static_cast<size_t>(blockIdx.x) * kHeadTile + static_cast<size_t>(threadIdx.x) * kHeadItems;
    bool have_prev = first > 0 && first - 1 < n;
    uint32_t prev = have_prev ? static_cast<uint32_t>(sorted[first - 1] >> 32) : 0u;
    uint32_t heads = 0;
#pragma unroll
    for (int i = 0; i < kHeadItems; ++i) {
        const size_t idx = first + i;
        if (idx < n) {
            const uint32_t key = static_cast<uint32_t>(sorted[idx] >> 32);
            if (!have_prev || key != prev) heads |= 1u << i;
            prev = key;
            have_prev = true;
        }
    }
    uint64_t block_total = 0;
    const uint64_t excl = block_exclusive_sum<uint64_t>(static_cast<uint64_t>(__popc(heads)), &block_total, lds);
    if (!write) {
        if (threadIdx.x == 0) tile_counts[blockIdx.x] = block_total;
        return;
    }
    uint64_t pos = tile_counts[blockIdx.x] + excl;
#pragma unroll
    for (int i = 0; i < kHeadItems; ++i)
        if ((heads >> i) & 1u) starts[pos++] = static_cast<uint32_t>(first + i);
}

// starts[K] = n, K = *unique_keys (device).
__global__ void close_starts_kernel(uint32_t* starts, const uint64_t* unique_keys, size_t n) {
    starts[*unique_keys] = static_cast<uint32_t>(n);
}

// Each unique key r into the table (keys are distinct, so no slot is taken
// twice; the table has more slots than keys, so every loop ends).
__global__ void hash_insert_kernel(const uint64_t* sorted, const uint32_t* starts, uint32_t unique_keys,
                                   unsigned long long* table, uint32_t bits) {
    const uint64_t mask = (bits >= 64 ? ~0ull : (1ull << bits) - 1);
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t r = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < unique_keys; r += stride) {
        const uint32_t key = static_cast<uint32_t>(sorted[starts[r]] >> 32);
        const unsigned long long v = (static_cast<unsigned long long>(key) << 32) | static_cast<uint32_t>(r);
        uint64_t h = hash_slot(key, bits);
        while (atomicCAS(&table[h], static_cast<unsigned long long>(kEmptySlot), v) !=
               static_cast<unsigned long long>(kEmptySlot))
            h = (h + 1) & mask;
    }
}

// Join count over a dense CSR, in tiles of kCountTile delta rows: each
// row's first matching index row (for the emit), its exclusive offset within
// its tile (u64), and the tile's total in tile_sums. join_scan_kernel turns
// the totals into tile bases: the offset of row r is
// local[r] + base[r / kCountTile]. Tiles never wait for one another (a
// single-pass scan with look-back kept ~1,000 tiles waiting on aggregates).
// Layout: wave w owns tile rows [w*512, w*512+512), item i of lane j is row
// w*512 + i*64 + j, so loads are coalesced and the row-order prefix is one
// wave scan per item plus one block step.
constexpr int kCountItems = 8;
constexpr int kCountTile = kBlock * kCountItems;

// With n_dev, the row count is *n_dev (at most n, which then only sizes the
// launch) and tiles past it report a zero total.
template <typename O, bool Hashed>
__global__ __launch_bounds__(kBlock) void dense_join_count_kernel(const uint64_t* delta, size_t n, uint32_t key_col,
                                                                  KeyIndex ix, O* local, uint32_t* match,
                                                                  uint64_t* tile_sums, uint32_t* clear,
                                                                  uint32_t clear_words,
                                                                  const uint64_t* n_dev = nullptr) {
    const uint32_t* starts = ix.starts;
    // the step's status words, cleared by the first block (no kernel before
    // the count in a step reads them)
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < clear_words; i += kBlock) clear[i] = 0;
    constexpr int kWaveRows = kCountItems * kWave;
    __shared__ uint64_t wave_totals[kWavesPerBlock];
    const size_t base = static_cast<size_t>(blockIdx.x) * kCountTile;
    if (n_dev) n = *n_dev < n ? static_cast<size_t>(*n_dev) : n;
    if (base >= n) {  // block-uniform: a tile past the rows (never with n_dev null)
        if (threadIdx.x == 0) tile_sums[blockIdx.x] = 0;
        return;
    }
    const uint32_t lane = lane_id(), wave = wave_id();
    // loads in two batches (keys, then index entries), so that a wave has
    // all of them in flight at once instead of one dependent pair per item;
    // indexes are clamped (n >= 1, n_keys >= 1) so no load sits in a branch.
    // DELTA rows and the outputs stream (nontemporal loads and stores), so
    // that L2 keeps the randomly gathered starts[] (9.2 -> 8.0 ms a
    // fixpoint on the component graph)
    uint64_t dk[kCountItems];
#pragma unroll
    for (int i = 0; i < kCountItems; ++i) {
        const size_t row = base + wave * kWaveRows + i * kWave + lane;
        dk[i] = __builtin_nontemporal_load(&delta[row < n ? row : n - 1]);
    }
    uint32_t m0[kCountItems], m1[kCountItems];
    if constexpr (!Hashed) {
#pragma unroll
        for (int i = 0; i < kCountItems; ++i) {
            const size_t row = base + wave * kWaveRows + i * kWave + lane;
            const uint32_t key = column_of(dk[i], key_col);
            const bool hit = row < n && key < ix.n_keys;
            const uint32_t k = hit ? key : 0u;
            const uint32_t a = starts[k], b = starts[k + 1];
            m0[i] = hit ? a : 0u;
            m1[i] = hit ? b : 0u;
        }
    } else {
        // first probes of all items in flight together, then the few that
        // met another key walk on (the table is at most half full)
        const uint64_t mask = (1ull << ix.bits) - 1;
        uint64_t h[kCountItems], sl[kCountItems];
#pragma unroll
        for (int i = 0; i < kCountItems; ++i) {
            h[i] = hash_slot(column_of(dk[i], key_col), ix.bits);
            sl[i] = ix.table[h[i]];
        }
#pragma unroll
        for (int i = 0; i < kCountItems; ++i) {
            const uint32_t key = column_of(dk[i], key_col);
            while (sl[i] != kEmptySlot && static_cast<uint32_t>(sl[i] >> 32) != key) {
                h[i] = (h[i] + 1) & mask;
                sl[i] = ix.table[h[i]];
            }
        }
#pragma unroll
        for (int i = 0; i < kCountItems; ++i) {
            const size_t row = base + wave * kWaveRows + i * kWave + lane;
            const bool hit = row < n && sl[i] != kEmptySlot;
            const uint32_t r = hit ? static_cast<uint32_t>(sl[i]) : 0u;
            const uint32_t a = starts[r], b = starts[r + 1];
            m0[i] = hit ? a : 0u;
            m1[i] = hit ? b : 0u;
        }
    }
#pragma unroll
    for (int i = 0; i < kCountItems; ++i) {
        const size_t row = base + wave * kWaveRows + i * kWave + lane;
        if (row < n) __builtin_nontemporal_store(m0[i], &match[row]);
    }
    uint64_t c[kCountItems];
    uint64_t running = 0;  // wave-local prefix in row order
#pragma unroll
    for (int i = 0; i < kCountItems; ++i) {
        const uint64_t cnt = m1[i] - m0[i];
        const uint64_t incl = wave_inclusive_sum<uint64_t>(cnt);
        c[i] = running + incl - cnt;
        running += __shfl(incl, kWave - 1, kWave);
    }
    if (lane == 0) wave_totals[wave] = running;
    __syncthreads();
    uint64_t wave_base = 0, tile_sum = 0;
#pragma unroll
    for (int w = 0; w < kWavesPerBlock; ++w) {
        if (w < static_cast<int>(wave)) wave_base += wave_totals[w];
        tile_sum += wave_totals[w];
    }
#pragma unroll
    for (int i = 0; i < kCountItems; ++i) {
        const size_t row = base + wave * kWaveRows + i * kWave + lane;
        if (row < n) __builtin_nontemporal_store(static_cast<O>(wave_base + c[i]), &local[row]);
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tile_sum;
}

// Exclusive scan, in place, of the count's tile totals (one block, n >= 1),
// and the join total.
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void join_scan_kernel(uint64_t* sums, size_t n, uint64_t* total) {
    __shared__ uint64_t scan_lds[kScanThreads / kWave];
    block_scan_inplace<kScanThreads, 4>(sums, n, total, scan_lds);
}

// A delta row's first output: its offset within its count tile (u32 when
// every tile's total fits: the index's largest degree times kCountTile is
// below 2^32, else u64) plus the tile's base.
template <typename O>
__device__ __forceinline__ uint64_t join_offset(const O* local, const uint64_t* base, size_t row) {
    return static_cast<uint64_t>(local[row]) + base[row / kCountTile];
}

struct EmitSpec {
    uint32_t key_col;          // left key column
    uint32_t right_key;        // right key column
    uint32_t side[2];          // projection input of output column 0/1 (0 left, 1 right)
    uint32_t col[2];           // projection column
    uint32_t order[2];         // distinct columns: key = out[order0] << shift | out[order1]
    uint32_t shift;            // 32, or fewer bits to compress the sort key
};

constexpr int kEmitTile = 1024;   // output rows per tile
constexpr int kEmitStage = 1024;  // delta rows staged per tile

// row_start[t] = the delta row that contains output t*kEmitTile (one thread
// per tile; the last entry is n).
// With n_dev / total_dev the counts are read from the device (n, total and
// n_tiles then only bound them and the launch).
template <typename O>
__global__ void emit_partition_kernel(const O* local, const uint64_t* base, size_t n, size_t total,
                                      size_t* row_start, size_t n_tiles, const uint64_t* n_dev = nullptr,
                                      const uint64_t* total_dev = nullptr) {
    if (n_dev) n = *n_dev < n ? static_cast<size_t>(*n_dev) : n;
    if (total_dev) total = *total_dev < total ? static_cast<size_t>(*total_dev) : total;
    const size_t tiles_now = (total + kEmitTile - 1) / kEmitTile;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; t <= n_tiles; t += stride) {
        if (t > tiles_now) continue;
        if (t == tiles_now) {
            row_start[t] = n;
            continue;
        }
        const uint64_t o = static_cast<uint64_t>(t) * kEmitTile;
        // the last count tile whose first row starts at or before o (base[0]
        // is 0), then the last row of it that does: two short searches, the
        // first over a few thousand bases that stay in cache
        const size_t n_ct = (n + kCountTile - 1) / kCountTile;
        size_t lo = 0, hi = n_ct;
        while (lo < hi) {
            const size_t mid = (lo + hi) >> 1;
            if (base[mid] <= o)
                lo = mid + 1;
            else
                hi = mid;
        }
        const size_t ct = lo - 1;
        const uint64_t b = base[ct];
        // a count tile's first row has local offset 0
        size_t a = ct * kCountTile + 1, e = ct * kCountTile + kCountTile < n ? ct * kCountTile + kCountTile : n;
        while (a < e) {
            const size_t mid = (a + e) >> 1;
            if (b + static_cast<uint64_t>(local[mid]) <= o)
                a = mid + 1;
            else
                e = mid;
        }
        row_start[t] = a - 1;
    }
    (void)total;
}

// Output-balanced emit of the dense join, writing compressed sort keys. Each
// tile stages its delta rows in LDS, marks where each row's outputs begin,
// and a block max-scan gives every output position its row.
// match_base[row] = starts[key] of each delta row (written by the count),
// so the emit reads it in row order instead of gathering the index's starts.
// The delta rows, their first matches and the output stream (nontemporal
// loads and stores), so that L2 keeps the randomly gathered second column
// of the index (14.2 -> 13.3 ms a fixpoint on the component graph).
template <typename O>
__global__ __launch_bounds__(kBlock) void dense_join_emit_kernel(const uint64_t* delta, size_t n,
                                                                 const uint32_t* right_other, const O* local,
                                                                 const uint64_t* tile_base, const uint32_t* match_base,
                                                                 size_t total,
                                                                 const size_t* row_start, EmitSpec spec,
                                                                 uint64_t* out, const uint64_t* n_dev = nullptr,
                                                                 const uint64_t* total_dev = nullptr) {
    if (n_dev) n = *n_dev < n ? static_cast<size_t>(*n_dev) : n;
    if (total_dev) total = *total_dev < total ? static_cast<size_t>(*total_dev) : total;
    constexpr int kPer = kEmitTile / kBlock;  // 4 output positions per thread
    // 18 KB: 8 blocks a CU. A row's first output is kept relative to the
    // tile's first (u32, wrapping: differences are exact below 2^32, and a
    // row has fewer outputs than that); a position's row fits 16 bits.
    __shared__ uint32_t row_offset[kEmitStage];
    __shared__ uint64_t row_key[kEmitStage];
    __shared__ uint32_t row_match[kEmitStage];
    __shared__ uint16_t row_of[kEmitTile];
    static_assert(kEmitStage <= 65536, "a staged row index fits 16 bits");
    __shared__ uint32_t scan_lds[kWavesPerBlock];
    const uint64_t n_tiles = (total + kEmitTile - 1) / kEmitTile;
    for (uint64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const uint64_t o0 = t * kEmitTile;
        const uint64_t o1 = o0 + kEmitTile < total ? o0 + kEmitTile : total;
        const size_t l0 = row_start[t];
        const size_t l1 = row_start[t + 1] + 1 < n ? row_start[t + 1] + 1 : n;
        const size_t span = l1 - l0;
        const bool staged = span <= static_cast<size_t>(kEmitStage);
        if (staged) {
            for (uint32_t j = threadIdx.x; j < kEmitTile; j += kBlock) row_of[j] = 0;
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < span; i += kBlock) {
                const uint64_t dk = __builtin_nontemporal_load(&delta[l0 + i]);
                const uint64_t off = join_offset(local, tile_base, l0 + i);
                const uint32_t m0 = __builtin_nontemporal_load(&match_base[l0 + i]);
                const uint64_t next = l0 + i + 1 < n ? join_offset(local, tile_base, l0 + i + 1) : total;
                const uint64_t cnt = next - off;
                row_offset[i] = static_cast<uint32_t>(off - o0);
                row_key[i] = dk;
                row_match[i] = m0;
                // first output position of row i inside the tile (row 0 owns o0)
                if (i > 0 && cnt > 0 && off < o1) row_of[off - o0] = static_cast<uint16_t>(i);
            }
            __syncthreads();
            // inclusive max-scan of row_of: thread t owns positions [kPer*t, kPer*t + kPer)
            uint32_t local[kPer];
            uint32_t m = 0;
#pragma unroll
            for (int q = 0; q < kPer; ++q) {
                const uint32_t v = row_of[threadIdx.x * kPer + q];
                m = v > m ? v : m;
                local[q] = m;
            }
            // exclusive max over previous threads
            uint32_t incl = m;
#pragma unroll
            for (int off = 1; off < kWave; off <<= 1) {
                const uint32_t other = __shfl_up(incl, off, kWave);
                if (static_cast<int>(lane_id()) >= off) incl = other > incl ? other : incl;
            }
            if (lane_id() == kWave - 1) scan_lds[wave_id()] = incl;
            __syncthreads();
            uint32_t prior = __shfl_up(incl, 1, kWave);
            if (lane_id() == 0) prior = 0;
            for (int w = 0; w < static_cast<int>(wave_id()); ++w) prior = scan_lds[w] > prior ? scan_lds[w] : prior;
#pragma unroll
            for (int q = 0; q < kPer; ++q) {
                const uint32_t v = local[q] > prior ? local[q] : prior;
                row_of[threadIdx.x * kPer + q] = static_cast<uint16_t>(v);
            }
            __syncthreads();
        }
        for (uint64_t o = o0 + threadIdx.x; o < o1; o += kBlock) {
            uint64_t dk;
            uint32_t match, within;  // the row's first matching index row, and o's place among its outputs
            if (staged) {
                const uint32_t i = row_of[o - o0];
                dk = row_key[i];
                within = static_cast<uint32_t>(o - o0) - row_offset[i];
                match = row_match[i];
            } else {
                size_t lo = l0, hi = l1;
                while (lo < hi) {
                    const size_t mid = (lo + hi) >> 1;
                    if (join_offset(local, tile_base, mid) <= o)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                dk = delta[lo - 1];
                within = static_cast<uint32_t>(o - join_offset(local, tile_base, lo - 1));
                match = match_base[lo - 1];
            }
            const uint32_t r_other = right_other[match + within];
            // the right row is (join key, other); its join key equals the
            // left row's key column
            uint32_t v[2];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                if (spec.side[c] == 0)
                    v[c] = column_of(dk, spec.col[c]);
                else
                    v[c] = spec.col[c] == spec.right_key ? column_of(dk, spec.key_col) : r_other;
            }
            __builtin_nontemporal_store((static_cast<uint64_t>(v[spec.order[0]]) << spec.shift) | v[spec.order[1]], &out[o]);
        }
        __syncthreads();
    }
}

// Expand compressed keys (a << shift | b) to packed pairs (a << 32 | b).
__global__ void expand_keys_kernel(uint64_t* keys, const uint64_t* rows_dev, size_t bound, uint32_t shift) {
    const size_t n = rows_dev ? (*rows_dev < bound ? *rows_dev : bound) : bound;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    const uint64_t low_mask = (1ull << shift) - 1;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t k = keys[i];
        keys[i] = ((k >> shift) << 32) | (k & low_mask);
    }
}

__global__ void digest_kernel(const uint64_t* keys, size_t n, unsigned long long* digest, uint32_t* unsorted) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    unsigned long long acc = 0;
    uint32_t bad = 0;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t z = keys[i] + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        acc += z ^ (z >> 31);
        if (i > 0 && !(keys[i - 1] < keys[i])) bad = 1;
    }
    acc = wave_sum<unsigned long long>(acc);
    if (lane_id() == 0) atomicAdd(digest, acc);
    if (bad) atomicOr(unsorted, 1u);
}

__global__ void max_value_kernel(const uint32_t* a, const uint32_t* b, size_t n, uint32_t* out) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    uint32_t m = 0;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        m = a[i] > m ? a[i] : m;
        m = b[i] > m ? b[i] : m;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const uint32_t o = __shfl_xor(m, off, kWave);
        m = o > m ? o : m;
    }
    if (lane_id() == 0) atomicMax(out, m);
}

__global__ void pack_columns_kernel(const uint32_t* a, const uint32_t* b, size_t n, uint64_t* out) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = pack_pair(a[i], b[i]);
}

__global__ void unpack_columns_kernel(const uint64_t* keys, size_t n, uint32_t* a, uint32_t* b) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        a[i] = static_cast<uint32_t>(keys[i] >> 32);
        b[i] = static_cast<uint32_t>(keys[i]);
    }
}

__global__ void clear_words_kernel(uint32_t* p, size_t n) {
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

unsigned grid_for(size_t n) {
    size_t b = tiles(n, kBlock * 4);
    if (b == 0) b = 1;
    return static_cast<unsigned>(b < 16384 ? b : 16384);
}

// ------------------------------------------------------------------------
// host-side runtime

struct CachedIndex {
    bool valid = false;
    uint32_t relation = 0, column = 0;
    uint64_t version = 0;
    DevBuf<uint64_t> sorted;  // (key, other) rows sorted
    DevBuf<uint32_t> other;   // the other column of the sorted rows (the key is the join key)
    DevBuf<uint32_t> starts;  // dense: CSR over key values [0, n_keys]; hashed: over the n_keys unique keys
    uint32_t n_keys = 0;
    uint64_t rows = 0;
    uint32_t max_degree = 0;  // most rows of one key
    bool hashed = false;      // sparse key values: unique keys found through `table`
    DevBuf<uint64_t> table;   // (key << 32 | unique-key rank) slots, kEmptySlot when free
    uint32_t bits = 0;        // table size is 2^bits
    KeyIndex view() const { return KeyIndex{starts.ptr, n_keys, hashed ? table.ptr : nullptr, bits}; }
};

// Join index layout: chosen from the right relation's values (dense while
// its largest value is below 2^28 and the start array is at most about four
// entries per row beyond 2^24), or forced by sparkalog_hip_store_set_join_index.
enum JoinIndexMode { kIndexAuto = 0, kIndexDense = 1, kIndexHashed = 2 };

struct RuleRuntime {
    sparkalog_hip_rule plan{};
    CachedIndex index;
    DevBuf<uint64_t> tile_base; // join count: tile totals, then their exclusive prefix
    DevBuf<uint64_t> offsets;   // join offsets per delta row within its count tile (u32 when narrow)
    bool narrow = false;        // offsets held as u32
    DevBuf<uint32_t> match;     // first matching index row per delta row
    DevBuf<uint64_t> found;     // candidates already in FULL (one bit each)
    DevBuf<size_t> row_start;   // first delta row of each emit tile
    DevBuf<uint64_t> cand_a, cand_b, cand_c;
    DevBuf<unsigned char> seg_meta;  // header, tile counts and bounds of the segmented distinct
    DevBuf<uint64_t> distinct_out;
    uint64_t join_rows = 0;     // M (host, after the first sync)
    int distinct_slot = -1;     // profiling event of this iteration's distinct
    uint64_t delta_rows = 0;
    uint32_t shift = 32;
};

struct TargetRuntime {
    uint32_t target = 0;
    std::vector<uint32_t> rules;
    DevBuf<uint64_t> contributions, union_tmp;
    DevBuf<uint64_t> newt;
    DevBuf<uint64_t> full_union;
    DevBuf<uint64_t> new_run;      // NEWT's second copy, becomes a run of FULL
    DevBuf<uint32_t> new_sec;      // that run's second column (when FULL is probed by it)
    bool csr_run = false;          // this iteration's new run is kept as index + second column only
    DevBuf<uint32_t> new_xidx;     // its first-column index, written by the distinct
    uint64_t contrib_bound = 0;
    int anti_slot = -1;            // profiling event of this iteration's anti-join
    bool fused = false;            // this iteration's distinct wrote NEWT (candidates in FULL dropped first)
    bool instances_marked = false; // and marked the instances of NEWT's rows (instance tracking)
    std::vector<int> union_slots;  // profiling events of this iteration's contribution unions
};

// Graph-replayed iterations (sparkalog_hip_store_set_option "graph"): for a
// single rule over a static right relation (transitive closure's shape),
// FULL is one sorted array and every count lives on the device, so one
// iteration is a fixed sequence of launches over fixed buffers, captured
// once as a HIP graph and replayed; the host synchronizes once per
// iteration, to read the counts and the overflow flags. Device counter words:
enum GraphWord {
    kGwFull0 = 0, kGwFull1 = 1,    // FULL rows in buffer slot 0 / 1
    kGwDelta0 = 2, kGwDelta1 = 3,  // DELTA rows in buffer slot 0 / 1
    kGwJoin = 4,                   // join output rows (M)
    kGwJoinUsed = 5,               // M, or 0 when it overflows the candidate buffers
    kGwUnique = 6,                 // distinct candidates (U)
    kGwUniqueUsed = 7,             // U, or 0 after any overflow
    kGwFlags = 8,                  // bit 0: M > candidate capacity, bit 1: FULL + U > FULL capacity
    kGwError = 9,                  // look-back reports (low 32 bits)
    kGwDone = 10,                  // iterations completed
    kGwNext = 11,                  // the iteration's number + 1 (per-instance tracking)
    kGwWords = 16
};

struct GraphKey {
    const void* ptr[24];
    uint64_t val[12];
    bool operator==(const GraphKey& o) const {
        return std::memcmp(ptr, o.ptr, sizeof(ptr)) == 0 && std::memcmp(val, o.val, sizeof(val)) == 0;
    }
};

struct GraphRun {
    uint64_t cap_full = 0, cap_cand = 0, cap_delta = 0;
    DevBuf<uint64_t> full_next, delta_next;  // the buffers the next iteration writes
    DevBuf<uint64_t> cand_a, cand_b, uniq;
    DevBuf<uint64_t> offsets, tile_base;
    DevBuf<uint32_t> match;
    DevBuf<size_t> row_start;
    DevBuf<unsigned char> sort_temp, merge_temp;
    DevBuf<uint64_t> words;   // kGwWords device counters
    uint64_t* host = nullptr; // pinned mirror of `words`, one block per slot parity
    int slot = 0;             // which words slot holds the current FULL/DELTA rows
    // state seen after the last graph iteration: anything else means the
    // relations changed in between, and the device row counts are rewritten
    const void* seen_full = nullptr;
    const void* seen_delta = nullptr;
    uint64_t seen_full_rows = ~0ull, seen_delta_rows = ~0ull;
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    GraphKey key[2];
    // graphs are captured and launched on a stream of their own (the
    // store's stream may be the null stream, which cannot be captured),
    // ordered after the store's stream by an event
    hipStream_t stream = nullptr;
    hipEvent_t ready = nullptr;
    int next_evict = 0;
    uint64_t captures = 0, replays = 0;
    uint64_t updates = 0;     // captures applied to an existing executable graph
    double capture_ms = 0;    // host time spent capturing and instantiating
    void drop_graphs() {
        for (auto& e : exec)
            if (e) {
                (void)hipGraphExecDestroy(e);
                e = nullptr;
            }
    }
    ~GraphRun() {
        drop_graphs();
        if (host) (void)hipHostFree(host);
        if (ready) (void)hipEventDestroy(ready);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// Small device counters read back once per phase.
struct Counters {
    uint64_t join_total[64];
    uint64_t distinct_rows[64];
    uint64_t contrib_rows[64];
    uint64_t newt_rows[64];
    uint64_t full_rows[64];
    uint64_t merge_rows;         // output count of a run merge (equal to the summed sizes)
};

}  // namespace
}  // namespace spk

struct sparkalog_hip_store {
    hipStream_t stream = nullptr;
    std::vector<spk::RelationState> rel;
    spk::DevBuf<unsigned char> temp;
    spk::DevBuf<uint64_t> counters_dev;   // Counters on device
    spk::DevBuf<uint32_t> misc_dev;       // small counters: upload maxima, digests, the step's error word
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    // pinned host copies of the step's counters (read back once per phase)
    // and the event the host waits on for the join sizes
    spk::Counters* counters_host = nullptr;
    // [0]: device error flag, [1]: oversized-run rows of the distinct, [2]: index degree, [3]: a call's error word,
    // [4]: unique keys of a hashed index, [5]: output rows of a merge
    uint64_t* words_host = nullptr;
    hipEvent_t counters_ready = nullptr;
    bool profiling = false;
    spk::OpStat stats[spk::kOpCount];
    std::vector<spk::OpEvents> event_pool;
    size_t events_used = 0;
    uint64_t generation = 0;                    // bumped by every insert; versions derive from it
    std::vector<spk::DevBuf<uint64_t>> pool;    // released run buffers, reused by later merges
    uint32_t merge_ratio = 3;                   // merge two runs when the older is < ratio x the newer
    void* session = nullptr;                    // spk::Session of the running fixpoint
    bool segmented_distinct = true;             // sort grouped candidates per run in LDS

    spk::DevBuf<uint64_t> take(size_t n) {
        spk::DevBuf<uint64_t> b;
        {
            std::lock_guard<std::mutex> lock(spk::pool_mutex());
            size_t best = pool.size();
            for (size_t i = 0; i < pool.size(); ++i)
                if (pool[i].cap >= n && (best == pool.size() || pool[i].cap < pool[best].cap)) best = i;
            if (best < pool.size()) {
                b = std::move(pool[best]);
                pool.erase(pool.begin() + static_cast<long>(best));
            }
        }
        // a miss allocates; pooled buffers stay pooled (freeing them here, a
        // device-wide synchronization and hipFree each, would make the next
        // fixpoint allocate again in the middle of its iterations)
        if (!b.ptr) b.reserve(n ? n : 1);
        return b;
    }
    void give(spk::DevBuf<uint64_t>&& b) {
        if (!b.ptr) return;
        std::lock_guard<std::mutex> lock(spk::pool_mutex());
        pool.push_back(std::move(b));
    }
    // first-column index buffers (all n_x + 1 entries within a fixpoint)
    std::vector<spk::DevBuf<uint32_t>> xpool;
    spk::DevBuf<uint32_t> take_x(size_t n) {
        spk::DevBuf<uint32_t> b;
        {
            std::lock_guard<std::mutex> lock(spk::pool_mutex());
            for (size_t i = 0; i < xpool.size(); ++i)
                if (xpool[i].cap >= n) {
                    b = std::move(xpool[i]);
                    xpool.erase(xpool.begin() + static_cast<long>(i));
                    break;
                }
        }
        if (!b.ptr) b.reserve(n);
        return b;
    }
    void give_x(spk::DevBuf<uint32_t>&& b) {
        if (!b.ptr) return;
        std::lock_guard<std::mutex> lock(spk::pool_mutex());
        xpool.push_back(std::move(b));
    }
    // second-column buffers of FULL's parts (sized like runs: best fit)
    std::vector<spk::DevBuf<uint32_t>> secpool;
    spk::DevBuf<uint32_t> take_sec(size_t n) {
        spk::DevBuf<uint32_t> b;
        {
            std::lock_guard<std::mutex> lock(spk::pool_mutex());
            size_t best = secpool.size();
            for (size_t i = 0; i < secpool.size(); ++i)
                if (secpool[i].cap >= n && (best == secpool.size() || secpool[i].cap < secpool[best].cap)) best = i;
            if (best < secpool.size()) {
                b = std::move(secpool[best]);
                secpool.erase(secpool.begin() + static_cast<long>(best));
            }
        }
        if (!b.ptr) b.reserve(n ? n : 1);  // as take()
        return b;
    }
    void give_sec(spk::DevBuf<uint32_t>&& b) {
        if (!b.ptr) return;
        std::lock_guard<std::mutex> lock(spk::pool_mutex());
        secpool.push_back(std::move(b));
    }
    bool x_anti = false;  // anti-join FULL through its first-column index when it fits
    int join_index_mode = spk::kIndexAuto;
    int graph_mode = 0;   // replay captured iterations where the plan allows
    uint64_t graph_max_rows = 1ull << 25;  // larger FULLs leave graph mode
    spk::GraphRun graph;
    uint64_t sparse_probe_rows = spk::kSparseProbeRows;
    uint32_t seg_tile = spk::kSegTileDefault;  // nominal tile of the segmented distinct
    // small persistent device buffers for range probes and statistics (a
    // fresh allocation would synchronize the whole device when freed)
    spk::DevBuf<uint64_t> small_dev, probe_dev;
    uint64_t* small16() {
        small_dev.reserve(16);
        return small_dev.ptr;
    }
    uint64_t* probes(size_t n) {
        probe_dev.reserve(n ? n : 1);
        return probe_dev.ptr;
    }
    // operator workspaces kept across runs so repeated runs allocate nothing
    std::vector<spk::TargetRuntime> ws_targets;
    std::vector<spk::RuleRuntime> ws_rules;

    spk::Counters& host_counters() {
        if (!counters_host) {
            void* p = nullptr;
            std::lock_guard<std::recursive_mutex> lock(spk::device_mutex());
            spk::check(hipHostMalloc(&p, sizeof(spk::Counters) + 64, hipHostMallocDefault));
            counters_host = static_cast<spk::Counters*>(p);
            words_host = reinterpret_cast<uint64_t*>(static_cast<char*>(p) + sizeof(spk::Counters));
            spk::check(hipEventCreateWithFlags(&counters_ready, hipEventDisableTiming));
        }
        return *counters_host;
    }

    ~sparkalog_hip_store() {
        if (pinned) (void)hipHostFree(pinned);
        if (counters_host) (void)hipHostFree(counters_host);
        if (counters_ready) (void)hipEventDestroy(counters_ready);
        for (auto& e : event_pool) {
            (void)hipEventDestroy(e.begin);
            (void)hipEventDestroy(e.end);
            (void)hipEventDestroy(e.kernel.start);
            (void)hipEventDestroy(e.kernel.stop);
        }
    }

    void* staging(size_t bytes) {
        if (bytes > pinned_bytes) {
            std::lock_guard<std::recursive_mutex> lock(spk::device_mutex());
            if (pinned) (void)hipHostFree(pinned);
            pinned = nullptr;
            size_t b = pinned_bytes ? pinned_bytes : (1u << 20);
            while (b < bytes) b *= 2;
            spk::check(hipHostMalloc(&pinned, b, hipHostMallocDefault));
            pinned_bytes = b;
        }
        return pinned;
    }

    void* scratch(size_t bytes) {
        temp.reserve(bytes < 4096 ? 4096 : bytes);
        return temp.ptr;
    }

    // Profiling brackets around one operator.
    // Device word for the look-back reports of one call that synchronizes
    // before it returns: cleared on the stream, then routed to every launcher
    // of this thread by an ErrorSink; check_call_error reads it after that
    // synchronization.
    uint32_t* call_error_word() {
        misc_dev.reserve(256);
        uint32_t* w = misc_dev.ptr + 130;
        spk::check(spk::set_u32(w, 0, stream));
        return w;
    }
    void check_call_error(const uint32_t* w) {
        host_counters();
        uint32_t* h = reinterpret_cast<uint32_t*>(&words_host[3]);
        *h = 0;
        spk::check(hipMemcpyAsync(h, w, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        spk::check(hipStreamSynchronize(stream));
        if (*h) throw spk::StatusError{SPARKALOG_HIP_ERR_DEVICE};
    }

    int begin_op() {
        if (!profiling) return -1;
        if (events_used == event_pool.size()) {
            spk::OpEvents e;
            spk::check(hipEventCreate(&e.begin));
            spk::check(hipEventCreate(&e.end));
            spk::check(hipEventCreate(&e.kernel.start));
            spk::check(hipEventCreate(&e.kernel.stop));
            event_pool.push_back(e);
        }
        spk::OpEvents& e = event_pool[events_used];
        e.kernel_used = false;
        e.op = -1;
        e.bytes = 0.0;
        spk::check(hipEventRecord(e.begin, stream));
        return static_cast<int>(events_used++);
    }
    // Events around the operator's main kernel (null when not profiling).
    const spk::KernelTimer* kernel_timer(int slot) {
        if (slot < 0) return nullptr;
        event_pool[slot].kernel_used = true;
        return &event_pool[slot].kernel;
    }
    void end_op(int slot, int op, double bytes) {
        if (slot < 0) return;
        spk::check(hipEventRecord(event_pool[slot].end, stream));
        event_pool[slot].op = op;
        event_pool[slot].bytes = bytes;
    }
    // Bytes that depend on device-written counts, added after the sync.
    void add_op_bytes(int slot, double bytes) {
        if (slot >= 0 && static_cast<size_t>(slot) < events_used) event_pool[slot].bytes += bytes;
    }
    // Call after a stream synchronize.
    void collect_ops() {
        for (size_t i = 0; i < events_used; ++i) {
            const spk::OpEvents& e = event_pool[i];
            if (e.op < 0) continue;
            float ms = 0.f;
            spk::check(hipEventElapsedTime(&ms, e.begin, e.end));
            auto& st = stats[e.op];
            st.launches += 1;
            st.ms += ms;
            st.bytes += e.bytes;
            if (e.kernel_used) {
                float kms = 0.f;
                spk::check(hipEventElapsedTime(&kms, e.kernel.start, e.kernel.stop));
                st.kernel_launches += 1;
                st.kernel_ms += kms;
            }
        }
        events_used = 0;
    }
};

namespace spk {
namespace {

RelationState* find(sparkalog_hip_store* s, uint32_t id) {
    if (id >= s->rel.size() || !s->rel[id].present) return nullptr;
    return &s->rel[id];
}

RelationState* slot(sparkalog_hip_store* s, uint32_t id) {
    if (id >= s->rel.size()) s->rel.resize(id + 1);
    if (s->rel[id].present) throw StatusError{SPARKALOG_HIP_ERR_DUPLICATE_RELATION};
    // a fresh version space per insert keeps cached join indexes of an earlier
    // relation with the same id from matching
    s->rel[id].version = (++s->generation) << 32;
    return &s->rel[id];
}

// Upload two columns into a packed device relation; returns the max value.
uint32_t upload(sparkalog_hip_store* s, DevRelation& dst, const uint32_t* c0, const uint32_t* c1, uint64_t rows,
                int on_device) {
    if (rows > 0xFFFFFFFFull) throw StatusError{SPARKALOG_HIP_ERR_TOO_MANY_ROWS};
    dst.keys.reserve(rows ? rows : 1);
    dst.rows = rows;
    dst.touched();
    if (rows == 0) return 0;
    uint32_t max_value = 0;
    if (on_device) {
        pack_columns_kernel<<<grid_for(rows), kBlock, 0, s->stream>>>(c0, c1, rows, dst.keys.ptr);
        s->misc_dev.reserve(64);
        clear_words_kernel<<<1, 64, 0, s->stream>>>(s->misc_dev.ptr, 1);
        max_value_kernel<<<grid_for(rows), kBlock, 0, s->stream>>>(c0, c1, rows, s->misc_dev.ptr);
        check(hipGetLastError());
        check(hipMemcpyAsync(&max_value, s->misc_dev.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream));
        check(hipStreamSynchronize(s->stream));
        return max_value;
    }
    const size_t chunk = 1u << 22;  // rows per staging chunk (32 MiB)
    uint64_t* stage = static_cast<uint64_t*>(s->staging(chunk * sizeof(uint64_t)));
    for (uint64_t base = 0; base < rows; base += chunk) {
        const uint64_t n = rows - base < chunk ? rows - base : chunk;
        for (uint64_t i = 0; i < n; ++i) {
            const uint32_t a = c0[base + i], b = c1[base + i];
            max_value = a > max_value ? a : max_value;
            max_value = b > max_value ? b : max_value;
            stage[i] = (static_cast<uint64_t>(a) << 32) | b;
        }
        check(hipMemcpyAsync(dst.keys.ptr + base, stage, n * sizeof(uint64_t), hipMemcpyHostToDevice, s->stream));
        check(hipStreamSynchronize(s->stream));
    }
    return max_value;
}

void download(sparkalog_hip_store* s, const DevRelation& src, uint32_t* c0, uint32_t* c1, int to_device) {
    if (src.rows == 0) return;
    if (to_device) {
        unpack_columns_kernel<<<grid_for(src.rows), kBlock, 0, s->stream>>>(src.keys.ptr, src.rows, c0, c1);
        check(hipGetLastError());
        check(hipStreamSynchronize(s->stream));
        return;
    }
    const size_t chunk = 1u << 22;
    uint64_t* stage = static_cast<uint64_t*>(s->staging(chunk * sizeof(uint64_t)));
    for (uint64_t base = 0; base < src.rows; base += chunk) {
        const uint64_t n = src.rows - base < chunk ? src.rows - base : chunk;
        check(hipMemcpyAsync(stage, src.keys.ptr + base, n * sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
        check(hipStreamSynchronize(s->stream));
        for (uint64_t i = 0; i < n; ++i) {
            c0[base + i] = static_cast<uint32_t>(stage[i] >> 32);
            c1[base + i] = static_cast<uint32_t>(stage[i]);
        }
    }
}

bool identity(const uint32_t c[2]) { return c[0] == 0 && c[1] == 1; }

double settle_runs(sparkalog_hip_store* s, RelationState& ts, bool all);
void settle_all(sparkalog_hip_store* s, RelationState& st);

void build_index(sparkalog_hip_store* s, RuleRuntime& rt, RelationState& right, uint32_t relation) {
    const uint32_t col = rt.plan.right_key;
    CachedIndex& ix = rt.index;
    if (ix.valid && ix.relation == relation && ix.version == right.version && ix.column == col) return;
    // a recursive right input is indexed as one array
    settle_all(s, right);
    const int ev = s->begin_op();
    const uint64_t n = right.full.rows;
    // dense while the start array stays near the relation's own size; any
    // u32 value otherwise (storage/src/lib.rs:357-393 indexes every key)
    const uint64_t dense_entries = static_cast<uint64_t>(right.max_value) + 1;
    bool hashed = n > 0 && (right.max_value >= (1u << 28) || dense_entries > 4 * n + (1ull << 24));
    if (s->join_index_mode == kIndexDense && right.max_value < (1u << 28)) hashed = false;
    if (s->join_index_mode == kIndexHashed && n > 0) hashed = true;
    // look-back reports of the index sort (the step's own word is cleared
    // after the index is built)
    uint32_t* err = s->call_error_word();
    ErrorSink sink(err);
    ix.sorted.reserve(n ? n : 1);
    DevBuf<uint64_t> tmp = s->take(n ? n : 1);
    if (n) {
        gather_index_keys_kernel<<<grid_for(n), kBlock, 0, s->stream>>>(right.full.keys.ptr, n, col, ix.sorted.ptr);
        check(hipGetLastError());
        const uint64_t vm = bit_mask_for(right.max_value);
        const uint64_t known = (vm << 32) | vm;
        void* temp = s->scratch(distinct_temp_bytes(n));
        uint64_t* sorted = nullptr;
        PairInput in{ix.sorted.ptr, nullptr, nullptr, nullptr};
        check(radix_sort(in, n, ix.sorted.ptr, tmp.ptr, temp, distinct_temp_bytes(n), known, s->stream, &sorted));
        if (sorted != ix.sorted.ptr) std::swap(ix.sorted, tmp);
    }
    uint64_t n_keys = 0;
    s->host_counters();
    if (hashed) {
        // unique keys: head counts per tile, their exclusive scan (and K),
        // then each head's row into starts
        const size_t head_tiles = tiles(n, kHeadTile);
        DevBuf<uint64_t> counts = s->take(head_tiles + 1);
        uint64_t* k_dev = s->small16() + 4;
        key_heads_kernel<<<static_cast<unsigned>(head_tiles), kBlock, 0, s->stream>>>(ix.sorted.ptr, n, counts.ptr,
                                                                                     nullptr, 0);
        join_scan_kernel<<<1, kScanThreads, 0, s->stream>>>(counts.ptr, head_tiles, k_dev);
        check(hipGetLastError());
        uint64_t& k_host = s->words_host[4];
        k_host = 0;
        check(hipMemcpyAsync(&k_host, k_dev, sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
        check(hipStreamSynchronize(s->stream));
        n_keys = k_host;
        if (n_keys == 0 || n_keys > n) throw StatusError{SPARKALOG_HIP_ERR_DEVICE};
        ix.starts.reserve(static_cast<size_t>(n_keys) + 1);
        key_heads_kernel<<<static_cast<unsigned>(head_tiles), kBlock, 0, s->stream>>>(ix.sorted.ptr, n, counts.ptr,
                                                                                     ix.starts.ptr, 1);
        close_starts_kernel<<<1, 1, 0, s->stream>>>(ix.starts.ptr, k_dev, n);
        s->give(std::move(counts));
        // 2^bits >= 2K slots: at most half full
        uint32_t bits = 4;
        while ((1ull << bits) < 2 * n_keys) ++bits;
        ix.table.reserve(static_cast<size_t>(1) << bits);
        check(fill_u64(ix.table.ptr, static_cast<size_t>(1) << bits, kEmptySlot, s->stream));
        hash_insert_kernel<<<grid_for(n_keys), kBlock, 0, s->stream>>>(
            ix.sorted.ptr, ix.starts.ptr, static_cast<uint32_t>(n_keys),
            reinterpret_cast<unsigned long long*>(ix.table.ptr), bits);
        check(hipGetLastError());
        ix.bits = bits;
    } else {
        n_keys = n ? dense_entries : 1;  // an empty relation: one key with no rows
        ix.starts.reserve(static_cast<size_t>(n_keys) + 1);
        dense_starts_kernel<<<grid_for(static_cast<size_t>(n_keys) + 1), kBlock, 0, s->stream>>>(ix.sorted.ptr, n,
                                                                                              ix.starts.ptr, n_keys);
        ix.bits = 0;
    }
    ix.hashed = hashed;
    ix.other.reserve(n ? n : 1);
    if (n) low_halves_kernel<<<grid_for(n), kBlock, 0, s->stream>>>(ix.sorted.ptr, n, ix.other.ptr);
    uint32_t* degree_dev = reinterpret_cast<uint32_t*>(s->small16());
    clear_words_kernel<<<1, 64, 0, s->stream>>>(degree_dev, 1);
    max_degree_kernel<<<grid_for(n_keys), kBlock, 0, s->stream>>>(ix.starts.ptr, n_keys, degree_dev);
    uint32_t& degree_host = *reinterpret_cast<uint32_t*>(&s->words_host[2]);
    check(hipGetLastError());
    check(hipMemcpyAsync(&degree_host, degree_dev, sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream));
    s->check_call_error(err);  // synchronizes
    ix.max_degree = degree_host;
    ix.valid = true;
    ix.relation = relation;
    ix.version = right.version;
    ix.column = col;
    ix.n_keys = static_cast<uint32_t>(n_keys);
    ix.rows = n;
    s->give(std::move(tmp));
    s->end_op(ev, kOpIndex, 16.0 * n);
}

constexpr size_t kMaxRuns = 15;

__global__ void copy_keys_kernel(const uint64_t* in, const uint64_t* rows_dev, size_t bound, uint64_t* out,
                                 uint64_t* out_copy, uint64_t* out_rows) {
    const size_t n = *rows_dev < bound ? *rows_dev : bound;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    const size_t gtid = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    for (size_t i = gtid; i < n; i += stride) {
        out[i] = in[i];
        out_copy[i] = in[i];
    }
    if (gtid == 0) *out_rows = n;
}

hipError_t copy_keys(const uint64_t* in, const uint64_t* rows_dev, size_t bound, uint64_t* out, uint64_t* out_copy,
                     uint64_t* out_rows, hipStream_t stream) {
    copy_keys_kernel<<<grid_for(bound), kBlock, 0, stream>>>(in, rows_dev, bound, out, out_copy, out_rows);
    return hipGetLastError();
}

// Merge FULL's newest runs while the older run is smaller than merge_ratio
// times the newer (or always, when `all`), leaving FULL as one array when
// `all` is set. Inputs are disjoint sorted unique arrays, so a merge's output
// size is the sum of its inputs' sizes. Returns the algorithmic bytes moved.
// Build a FULL part's first-column index unless it is current.
void ensure_xidx(sparkalog_hip_store* s, DevRelation& r, uint32_t n_x) {
    if (r.xidx_valid(n_x)) return;
    if (r.csr) {
        // a part held as index + second column has no keys to build from:
        // its index grows with n_x (first values only grow within a store)
        const uint32_t old_n = r.xidx_n;
        if (n_x < old_n || r.rows > 0xFFFFFFFFull) throw StatusError{SPARKALOG_HIP_ERR_INVALID_ARGUMENT};
        DevBuf<uint32_t> grown = s->take_x(static_cast<size_t>(n_x) + 1);
        check(xindex_extend(r.xidx.ptr, old_n, grown.ptr, n_x, static_cast<uint32_t>(r.rows), s->stream));
        s->give_x(std::move(r.xidx));
        r.xidx = std::move(grown);
        r.xidx_set(n_x);
        return;
    }
    if (r.xidx.cap < static_cast<size_t>(n_x) + 1) {
        s->give_x(std::move(r.xidx));
        r.xidx = s->take_x(static_cast<size_t>(n_x) + 1);
    }
    check(xindex_build(r.keys.ptr, nullptr, r.rows, n_x, r.xidx.ptr, s->stream));
    r.xidx_set(n_x);
}

// Write a FULL part's second column unless it is current.
void ensure_sec(sparkalog_hip_store* s, DevRelation& r) {
    if (r.sec_valid()) return;
    if (r.sec.cap < r.rows + kSecPad) {
        s->give_sec(std::move(r.sec));
        r.sec = s->take_sec(r.rows + kSecPad);
    }
    check(second_column(r.keys.ptr, r.rows, r.sec.ptr, s->stream));
    r.sec_set();
}

// First-column index size for the store's current values: every stored row
// has both columns below it.
uint32_t store_n_x(const sparkalog_hip_store* s) {
    uint32_t vmax = 0;
    for (const auto& st : s->rel)
        if (st.present) vmax = st.max_value > vmax ? st.max_value : vmax;
    return vmax + 1;
}

// Give a FULL part held as index + second column its packed keys again.
void ensure_packed(sparkalog_hip_store* s, DevRelation& r) {
    if (!r.csr) return;
    const uint32_t nx = r.xidx_n;
    DevBuf<uint64_t> keys = s->take(r.rows ? r.rows : 1);
    if (r.rows) {
        const size_t tb = run_merge_temp_bytes(r.rows);
        void* temp = s->scratch(tb);
        check(run_merge(r.xidx.ptr, r.sec.ptr, r.rows, nullptr, nullptr, 0, nx, nullptr, keys.ptr, temp, tb,
                        s->stream));
    }
    s->give(std::move(r.keys));
    r.keys = std::move(keys);
    r.csr = false;
    r.touched();
    r.xidx_set(nx);
    r.sec_set();
}

void packed_parts(sparkalog_hip_store* s, RelationState& st) {
    ensure_packed(s, st.full);
    for (auto& r : st.runs) ensure_packed(s, r);
}

// Merge FULL's newest runs while the older run is smaller than merge_ratio
// times the newer (or always, when `all`, which leaves FULL as one packed
// array). Inputs are disjoint sorted unique arrays, so a merge's output size
// is the sum of its inputs' sizes. When either part is held as index +
// second column, the merge reads and writes only second columns (8 B a row
// instead of 20); the last merge of `all` also writes packed keys. Returns
// the algorithmic bytes moved.
double settle_runs(sparkalog_hip_store* s, RelationState& ts, bool all) {
    double bytes = 0.0;
    while (!ts.runs.empty()) {
        DevRelation& last = ts.runs.back();
        DevRelation& prev = ts.runs.size() >= 2 ? ts.runs[ts.runs.size() - 2] : ts.full;
        const bool merge = all || ts.runs.size() > kMaxRuns ||
                           prev.rows < static_cast<uint64_t>(s->merge_ratio) * last.rows;
        if (!merge) break;
        const uint64_t n = prev.rows + last.rows;
        if (prev.rows == 0) {
            const uint32_t last_nx = last.xidx_n && last.xidx_valid(last.xidx_n) ? last.xidx_n : 0;
            const bool last_sec = last.sec_valid(), last_csr = last.csr;
            std::swap(prev.keys, last.keys);
            std::swap(prev.xidx, last.xidx);
            std::swap(prev.sec, last.sec);
            prev.rows = last.rows;
            prev.touched();
            prev.csr = last_csr;
            last.csr = false;
            if (last_nx) prev.xidx_set(last_nx);
            if (last_sec && !last_csr) prev.sec_set();
        } else if (last.rows) {
            // an index covering every part: the store's current first values
            // (never below the index of a part held as index + second column)
            uint32_t nx = store_n_x(s);
            if (prev.csr && prev.xidx_n > nx) nx = prev.xidx_n;
            if (last.csr && last.xidx_n > nx) nx = last.xidx_n;
            // index entries are u32 row counts: a merged part must stay below 2^32 rows
            const bool cols = (prev.csr || last.csr) && nx && n <= 0xFFFFFFFFull;
            if (cols) {
                ensure_xidx(s, prev, nx);
                ensure_sec(s, prev);
                ensure_xidx(s, last, nx);
                ensure_sec(s, last);
                const bool packed_out = all && ts.runs.size() == 1;
                DevBuf<uint32_t> out_sec = s->take_sec(n + kSecPad);
                DevBuf<uint64_t> out_keys;
                if (packed_out) out_keys = s->take(n);
                const size_t tb = run_merge_temp_bytes(n);
                void* temp = s->scratch(tb);
                const int ev = s->begin_op();
                check(run_merge(prev.xidx.ptr, prev.sec.ptr, prev.rows, last.xidx.ptr, last.sec.ptr, last.rows, nx,
                                out_sec.ptr, packed_out ? out_keys.ptr : nullptr, temp, tb, s->stream,
                                s->kernel_timer(ev)));
                s->stats[kOpMergeRuns].kernel = "run_merge_kernel";
                const double mb = (packed_out ? 16.0 : 8.0) * static_cast<double>(n);
                s->end_op(ev, kOpMergeRuns, mb);
                check(xindex_add(prev.xidx.ptr, last.xidx.ptr, nx, s->stream));
                s->give(std::move(prev.keys));
                prev.keys = std::move(out_keys);
                s->give_sec(std::move(prev.sec));
                prev.sec = std::move(out_sec);
                prev.rows = n;
                prev.csr = !packed_out;
                prev.touched();
                prev.xidx_set(nx);
                if (packed_out) prev.sec_set();
                bytes += mb;
            } else {
                ensure_packed(s, prev);
                ensure_packed(s, last);
                // the merged run's index is the sum of the two indexes
                const uint32_t px = prev.xidx_n;
                const bool add = px && prev.xidx_valid(px) && last.xidx_valid(px) && n <= 0xFFFFFFFFull;
                // the merged run keeps a second column while its parts had one
                const bool sec = prev.sec_valid() && last.sec_valid();
                DevBuf<uint64_t> out = s->take(n);
                DevBuf<uint32_t> out_sec;
                if (sec) out_sec = s->take_sec(n + kSecPad);
                const size_t tb = merge_disjoint_temp_bytes(prev.rows, last.rows);
                void* temp = s->scratch(tb);
                const int ev = s->begin_op();
                check(merge_disjoint(prev.keys.ptr, prev.rows, last.keys.ptr, last.rows, out.ptr, temp, tb,
                                     s->stream, s->kernel_timer(ev), sec ? out_sec.ptr : nullptr));
                s->stats[kOpMergeRuns].kernel = "merge_disjoint_kernel";
                s->end_op(ev, kOpMergeRuns, (sec ? 20.0 : 16.0) * n);
                if (add) check(xindex_add(prev.xidx.ptr, last.xidx.ptr, px, s->stream));
                s->give(std::move(prev.keys));
                prev.keys = std::move(out);
                prev.rows = n;
                prev.touched();
                if (add) prev.xidx_set(px);
                if (sec) {
                    s->give_sec(std::move(prev.sec));
                    prev.sec = std::move(out_sec);
                    prev.sec_set();
                }
                bytes += (sec ? 20.0 : 16.0) * n;
            }
        }
        s->give(std::move(last.keys));
        s->give_x(std::move(last.xidx));
        s->give_sec(std::move(last.sec));
        last.csr = false;
        ts.runs.pop_back();
    }
    if (all) ensure_packed(s, ts.full);
    return bytes;
}

// FULL as one packed array (every run merged in).
void settle_all(sparkalog_hip_store* s, RelationState& st) {
    if (!st.runs.empty() || st.full.csr) settle_runs(s, st, true);
}


// ------------------------------------------------------------------------
// per-instance and per-range statistics, row merges and range removal (the
// batched and served fixpoints keep many independent instances in one
// relation, instance = first column >> instance bits)

// Rows and order-independent digest per instance over sorted keys. Each wave
// walks its own contiguous range of rows (instances are contiguous in sorted
// keys) and keeps the current instance's sums in registers, adding them to the
// instance's counters when the instance changes: an atomic per instance and
// wave instead of one per 64 rows, all of an instance's waves on one counter.
// A batch of 64 rows spanning two instances adds per row.
constexpr uint32_t kStatsBatch = 8;                             // row batches loaded at once
constexpr uint32_t kStatsWaveRows = 8 * kStatsBatch * kWave;    // rows per wave
__global__ __launch_bounds__(kBlock) void instance_stats_kernel(const uint64_t* keys, size_t n, uint32_t bits,
                                                                uint32_t instances, unsigned long long* rows,
                                                                unsigned long long* digests) {
    const size_t wave_g = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    const size_t r0 = wave_g * kStatsWaveRows;
    if (r0 >= n) return;  // wave-uniform
    const size_t r1 = r0 + kStatsWaveRows < n ? r0 + kStatsWaveRows : n;
    const uint32_t lane = lane_id();
    uint32_t cur = 0xFFFFFFFFu;     // instance being summed (wave-uniform)
    unsigned long long acc = 0;     // this lane's digest share of it
    unsigned long long acc_rows = 0;
    auto flush = [&]() {
        const unsigned long long sum = wave_sum<unsigned long long>(acc);
        if (lane == 0 && cur < instances && acc_rows) {
            atomicAdd(&digests[cur], sum);
            atomicAdd(&rows[cur], acc_rows);
        }
        acc = 0;
        acc_rows = 0;
    };
    for (size_t b = r0; b < r1; b += static_cast<size_t>(kStatsBatch) * kWave) {
        uint64_t kk[kStatsBatch];
#pragma unroll
        for (uint32_t j = 0; j < kStatsBatch; ++j) {
            const size_t i = b + j * kWave + lane;
            kk[j] = keys[i < r1 ? i : r1 - 1];
        }
#pragma unroll
        for (uint32_t j = 0; j < kStatsBatch; ++j) {
            const size_t i = b + j * kWave + lane;
            const bool valid = i < r1;
            const uint32_t x = static_cast<uint32_t>(kk[j] >> 32), z = static_cast<uint32_t>(kk[j]);
            const uint32_t inst = valid ? (bits >= 32 ? 0u : x >> bits) : 0xFFFFFFFFu;
            const unsigned long long c = valid ? instance_pair_hash(x, z, bits) : 0ull;
            const uint64_t valid_mask = ballot(valid);
            if (valid_mask == 0) continue;  // wave-uniform
            const uint32_t lead = __builtin_amdgcn_readfirstlane(inst);
            if (ballot(valid && inst == lead) == valid_mask) {
                if (lead != cur) {
                    flush();
                    cur = lead;
                }
                acc += c;
                acc_rows += static_cast<unsigned long long>(__popcll(valid_mask));
            } else {
                flush();
                cur = 0xFFFFFFFFu;
                if (valid && inst < instances) {
                    atomicAdd(&digests[inst], c);
                    atomicAdd(&rows[inst], 1ull);
                }
            }
        }
    }
    flush();
}

// lower bounds of a set of keys (one thread each)
__global__ void lower_bounds_kernel(const uint64_t* keys, size_t n, const uint64_t* probes, size_t count,
                                    uint64_t* out) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t j = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < count; j += stride) {
        size_t lo = 0, hi = n;
        while (lo < hi) {
            const size_t mid = (lo + hi) >> 1;
            if (keys[mid] < probes[j])
                lo = mid + 1;
            else
                hi = mid;
        }
        out[j] = lo;
    }
}

__global__ void range_stats_kernel(const uint64_t* keys, size_t begin, size_t end, uint32_t base,
                                   unsigned long long* rows_digest) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    unsigned long long acc = 0, cnt = 0;
    for (size_t i = begin + static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < end; i += stride) {
        const uint64_t k = keys[i];
        const uint64_t local = ((static_cast<uint64_t>(static_cast<uint32_t>(k >> 32) - base)) << 32) |
                               static_cast<uint32_t>(static_cast<uint32_t>(k) - base);
        uint64_t zz = local + 0x9E3779B97F4A7C15ull;
        zz = (zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9ull;
        zz = (zz ^ (zz >> 27)) * 0x94D049BB133111EBull;
        acc += zz ^ (zz >> 31);
        ++cnt;
    }
    acc = wave_sum<unsigned long long>(acc);
    cnt = wave_sum<unsigned long long>(cnt);
    if (lane_id() == 0 && cnt) {
        atomicAdd(&rows_digest[0], cnt);
        atomicAdd(&rows_digest[1], acc);
    }
}

// Positions [lo, hi) of the rows whose first column lies in [first_lo, first_hi).
std::pair<uint64_t, uint64_t> first_range(sparkalog_hip_store* s, const DevRelation& r, uint32_t first_lo,
                                          uint32_t first_hi) {
    if (r.rows == 0) return {0, 0};
    uint64_t probes[2] = {static_cast<uint64_t>(first_lo) << 32, static_cast<uint64_t>(first_hi) << 32};
    if (first_hi == 0) probes[1] = 0;
    uint64_t* dev = s->small16() + 8;  // [0, 8) is left to the callers
    check(hipMemcpyAsync(dev, probes, sizeof(probes), hipMemcpyHostToDevice, s->stream));
    lower_bounds_kernel<<<1, 64, 0, s->stream>>>(r.keys.ptr, r.rows, dev, 2, dev + 2);
    check(hipGetLastError());
    uint64_t out[2];
    check(hipMemcpyAsync(out, dev + 2, sizeof(out), hipMemcpyDeviceToHost, s->stream));
    check(hipStreamSynchronize(s->stream));
    return {out[0], out[1] > out[0] ? out[1] : out[0]};
}

void remove_rows(sparkalog_hip_store* s, DevRelation& r, uint64_t lo, uint64_t hi) {
    if (hi <= lo) return;
    const uint64_t n = r.rows - (hi - lo);
    DevBuf<uint64_t> out = s->take(n ? n : 1);
    if (lo) check(hipMemcpyAsync(out.ptr, r.keys.ptr, lo * sizeof(uint64_t), hipMemcpyDeviceToDevice, s->stream));
    if (r.rows > hi)
        check(hipMemcpyAsync(out.ptr + lo, r.keys.ptr + hi, (r.rows - hi) * sizeof(uint64_t),
                             hipMemcpyDeviceToDevice, s->stream));
    s->give(std::move(r.keys));
    r.keys = std::move(out);
    r.rows = n;
    r.touched();
}

// dst = dst ∪ keys (both sorted); the result size comes back from the device.
void union_into(sparkalog_hip_store* s, DevRelation& dst, const uint64_t* keys, uint64_t rows) {
    if (rows == 0) return;
    const uint64_t bound = dst.rows + rows;
    DevBuf<uint64_t> out = s->take(bound);
    const size_t tb = union_temp_bytes(dst.rows, rows);
    void* temp = s->scratch(tb);
    s->counters_dev.reserve(sizeof(Counters) / sizeof(uint64_t));
    Counters* cdev = reinterpret_cast<Counters*>(s->counters_dev.ptr);
    PairInput a{dst.keys.ptr, nullptr, nullptr, nullptr};
    PairInput b{keys, nullptr, nullptr, nullptr};
    uint32_t* err = s->call_error_word();
    {
        ErrorSink sink(err);
        check(sorted_union(a, dst.rows, b, rows, PairOutput{out.ptr, nullptr, nullptr}, &cdev->merge_rows, temp, tb,
                           s->stream));
    }
    s->host_counters();
    uint64_t& n = s->words_host[5];
    n = 0;
    check(hipMemcpyAsync(&n, &cdev->merge_rows, sizeof(n), hipMemcpyDeviceToHost, s->stream));
    s->check_call_error(err);
    s->give(std::move(dst.keys));
    dst.keys = std::move(out);
    dst.rows = n;
    dst.touched();
}

// RecursiveExecutor::compile (generic.rs:407-470): validation and runtime
// construction; returns a status code.
int compile_plan(sparkalog_hip_store* s, const uint32_t* relations, uint32_t n_relations,
                 const sparkalog_hip_rule* rules, uint32_t n_rules, std::vector<TargetRuntime>& targets,
                 std::vector<RuleRuntime>& rts) {
    if (n_rules > 64) return SPARKALOG_HIP_ERR_UNSUPPORTED_PLAN;
    if (n_relations == 0) return SPARKALOG_HIP_ERR_EMPTY_SCC;
    for (uint32_t i = 0; i < n_relations; ++i) {
        for (uint32_t j = 0; j < i; ++j)
            if (relations[j] == relations[i]) return SPARKALOG_HIP_ERR_DUPLICATE_RELATION;
        RelationState* st = find(s, relations[i]);
        if (!st) return SPARKALOG_HIP_ERR_MISSING_RELATION;
        if (!st->recursive) return SPARKALOG_HIP_ERR_NOT_RECURSIVE;
    }
    for (uint32_t r = 0; r < n_rules; ++r) {
        if (!find(s, rules[r].delta_input) || !find(s, rules[r].right_input)) return SPARKALOG_HIP_ERR_MISSING_RELATION;
        bool delta_in = false, target_in = false;
        for (uint32_t i = 0; i < n_relations; ++i) {
            delta_in |= relations[i] == rules[r].delta_input;
            target_in |= relations[i] == rules[r].target;
        }
        if (!delta_in) return SPARKALOG_HIP_ERR_DELTA_OUTSIDE_SCC;
        if (!target_in) return SPARKALOG_HIP_ERR_TARGET_OUTSIDE_SCC;
    }
    if (targets.size() != n_relations) {
        targets.clear();
        targets.resize(n_relations);
    }
    for (uint32_t i = 0; i < n_relations; ++i) {
        targets[i].target = relations[i];
        targets[i].rules.clear();
        for (uint32_t r = 0; r < n_rules; ++r)
            if (rules[r].target == relations[i]) targets[i].rules.push_back(r);
        if (targets[i].rules.empty()) return SPARKALOG_HIP_ERR_MISSING_RULE_TARGET;
        const sparkalog_hip_rule& f = rules[targets[i].rules[0]];
        for (uint32_t k : targets[i].rules) {
            const sparkalog_hip_rule& o = rules[k];
            if (std::memcmp(o.anti_left, f.anti_left, sizeof(f.anti_left)) ||
                std::memcmp(o.anti_right, f.anti_right, sizeof(f.anti_right)) ||
                std::memcmp(o.union_left, f.union_left, sizeof(f.union_left)) ||
                std::memcmp(o.union_right, f.union_right, sizeof(f.union_right)))
                return SPARKALOG_HIP_ERR_INCOMPATIBLE_PLANS;
        }
    }
    if (rts.size() != n_rules) {
        rts.clear();
        rts.resize(n_rules);
    }
    for (uint32_t r = 0; r < n_rules; ++r) {
        const sparkalog_hip_rule& p = rules[r];
        if (p.left_key > 1 || p.right_key > 1 || p.out_column[0] > 1 || p.out_column[1] > 1 ||
            p.distinct_columns[0] > 1 || p.distinct_columns[1] > 1 || p.out_input[0] > 1 || p.out_input[1] > 1)
            return SPARKALOG_HIP_ERR_MISSING_COLUMN;
        if (!identity(p.anti_left) || !identity(p.anti_right) || !identity(p.union_left) || !identity(p.union_right))
            return SPARKALOG_HIP_ERR_UNSUPPORTED_PLAN;
        rts[r].plan = p;
    }

    return SPARKALOG_HIP_OK;
}

}  // namespace
}  // namespace spk

using namespace spk;

extern "C" int sparkalog_hip_last_hip_error(void) { return g_last_hip_error; }

extern "C" const char* sparkalog_hip_status_string(int status) {
    switch (status) {
        case SPARKALOG_HIP_OK: return "ok";
        case SPARKALOG_HIP_ERR_HIP: return "HIP call failed";
        case SPARKALOG_HIP_ERR_TOO_MANY_ROWS: return "rows cannot be represented by u32 row IDs";
        case SPARKALOG_HIP_ERR_MISSING_COLUMN: return "plan names a missing column";
        case SPARKALOG_HIP_ERR_ITERATION_LIMIT: return "recursive SCC did not converge within the iteration limit";
        case SPARKALOG_HIP_ERR_INCOMPATIBLE_PLANS: return "clauses for one relation use incompatible set plans";
        case SPARKALOG_HIP_ERR_NOT_RECURSIVE: return "relation has no DELTA/NEWT state";
        case SPARKALOG_HIP_ERR_MISSING_RELATION: return "relation is not registered";
        case SPARKALOG_HIP_ERR_DUPLICATE_RELATION: return "relation is already registered";
        case SPARKALOG_HIP_ERR_EMPTY_SCC: return "recursive SCC has no relations";
        case SPARKALOG_HIP_ERR_DELTA_OUTSIDE_SCC: return "DELTA relation is outside its SCC";
        case SPARKALOG_HIP_ERR_TARGET_OUTSIDE_SCC: return "rule target is outside its SCC";
        case SPARKALOG_HIP_ERR_MISSING_RULE_TARGET: return "relation has no recursive rule";
        case SPARKALOG_HIP_ERR_OUT_OF_MEMORY: return "device allocation failed";
        case SPARKALOG_HIP_ERR_UNSUPPORTED_PLAN: return "plan shape not supported by the device executor";
        case SPARKALOG_HIP_ERR_DEVICE: return "a kernel raised its device error flag";
        case SPARKALOG_HIP_ERR_INVALID_ARGUMENT: return "invalid argument";
        default: return "unknown status";
    }
}

// Registry of live stores, for out-of-memory relief across stores (served
// lanes keep one store each, on their own threads).
static std::vector<sparkalog_hip_store*>& stores() {
    static std::vector<sparkalog_hip_store*> v;
    return v;
}

// g_relief: move pooled buffers of every store out of the pools, smallest
// first, until `bytes` are covered, then free them outside the lock (each
// free synchronizes the device first, so no queued kernel loses a buffer).
static bool relieve_pools(size_t bytes) {
    std::vector<spk::DevBuf<uint64_t>> free64;
    std::vector<spk::DevBuf<uint32_t>> free32;
    {
        std::lock_guard<std::mutex> lock(spk::pool_mutex());
        struct Pooled {
            size_t bytes;
            spk::DevBuf<uint64_t>* b64;
            spk::DevBuf<uint32_t>* b32;
        };
        std::vector<Pooled> all;
        for (auto* st : stores()) {
            for (auto& b : st->pool) all.push_back({b.cap * sizeof(uint64_t), &b, nullptr});
            for (auto& b : st->secpool) all.push_back({b.cap * sizeof(uint32_t), nullptr, &b});
            for (auto& b : st->xpool) all.push_back({b.cap * sizeof(uint32_t), nullptr, &b});
        }
        std::sort(all.begin(), all.end(), [](const Pooled& a, const Pooled& b) { return a.bytes < b.bytes; });
        size_t got = 0;
        for (auto& e : all) {
            if (got >= bytes) break;
            got += e.bytes;
            if (e.b64)
                free64.push_back(std::move(*e.b64));
            else
                free32.push_back(std::move(*e.b32));
        }
        auto empty64 = [](const spk::DevBuf<uint64_t>& x) { return x.ptr == nullptr; };
        auto empty32 = [](const spk::DevBuf<uint32_t>& x) { return x.ptr == nullptr; };
        for (auto* st : stores()) {
            st->pool.erase(std::remove_if(st->pool.begin(), st->pool.end(), empty64), st->pool.end());
            st->secpool.erase(std::remove_if(st->secpool.begin(), st->secpool.end(), empty32), st->secpool.end());
            st->xpool.erase(std::remove_if(st->xpool.begin(), st->xpool.end(), empty32), st->xpool.end());
        }
    }
    return !free64.empty() || !free32.empty();  // freed here, by the destructors
}

#define SPK_GUARD_BEGIN try {

#define SPK_GUARD_END                                    \
    }                                                    \
    catch (const StatusError& e) {                       \
        return e.code;                                   \
    }                                                    \
    catch (const HipFailure&) {                          \
        return SPARKALOG_HIP_ERR_HIP;                    \
    }                                                    \
    catch (...) {                                        \
        return SPARKALOG_HIP_ERR_OUT_OF_MEMORY;          \
    }

extern "C" int sparkalog_hip_store_create(sparkalog_hip_store** out, void* stream) {
    if (!out) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    SPK_GUARD_BEGIN
    auto* s = new sparkalog_hip_store();
    s->stream = static_cast<hipStream_t>(stream);
    {
        std::lock_guard<std::mutex> lock(spk::pool_mutex());
        stores().push_back(s);
        spk::g_relief = relieve_pools;
    }
    for (int i = 0; i < kOpCount; ++i) {
        s->stats[i].name = kOpNames[i];
        s->stats[i].kernel = kOpKernels[i];
    }
    *out = s;
    return SPARKALOG_HIP_OK;
    SPK_GUARD_END
}

namespace spk {
namespace {
void destroy_session(sparkalog_hip_store* s);
}  // namespace
}  // namespace spk

extern "C" void sparkalog_hip_store_destroy(sparkalog_hip_store* store) {
    if (!store) return;
    (void)hipStreamSynchronize(store->stream);
    destroy_session(store);
    {
        std::lock_guard<std::mutex> lock(spk::pool_mutex());
        auto& v = stores();
        v.erase(std::remove(v.begin(), v.end(), store), v.end());
    }
    delete store;
}

extern "C" int sparkalog_hip_store_insert_static(sparkalog_hip_store* s, uint32_t id, const uint32_t* c0,
                                                 const uint32_t* c1, uint64_t rows, int on_device) {
    if (!s || (rows && (!c0 || !c1))) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    SPK_GUARD_BEGIN
    RelationState* st = slot(s, id);
    st->max_value = upload(s, st->full, c0, c1, rows, on_device);
    st->present = true;
    st->recursive = false;
    return SPARKALOG_HIP_OK;
    SPK_GUARD_END
}

extern "C" int sparkalog_hip_store_difference(sparkalog_hip_store* s, uint32_t target, uint32_t left, uint32_t right,
                                              uint64_t* rows) {
    if (!s) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    if (!find(s, left) || !find(s, right)) return SPARKALOG_HIP_ERR_MISSING_RELATION;
    if (target == left || target == right) return SPARKALOG_HIP_ERR_DUPLICATE_RELATION;
    SPK_GUARD_BEGIN
    // the target's slot first: creating it may move every relation state
    RelationState* ts = slot(s, target);
    RelationState* ls = find(s, left);
    RelationState* rs = find(s, right);
    // both inputs as one sorted unique array each
    settle_all(s, *ls);
    settle_all(s, *rs);
    const uint64_t nl = ls->full.rows, nr = rs->full.rows;
    const uint32_t left_max = ls->max_value;
    ts->full.keys.reserve(nl ? nl : 1);
    ts->full.rows = 0;
    ts->full.touched();
    ts->present = true;
    ts->recursive = false;
    ts->max_value = left_max;
    s->counters_dev.reserve(sizeof(Counters) / sizeof(uint64_t));
    Counters* cdev = reinterpret_cast<Counters*>(s->counters_dev.ptr);
    AntiRight rights{};
    rights.count = 0;
    if (nr) {
        rights.in[0] = PairInput{rs->full.keys.ptr, nullptr, nullptr, nullptr};
        rights.rows[0] = nr;
        rights.count = 1;
    }
    const size_t tb = anti_join_temp_bytes(nl, rights.count);
    void* temp = s->scratch(tb);
    PairInput a{ls->full.keys.ptr, nullptr, nullptr, nullptr};
    PairOutput o{ts->full.keys.ptr, nullptr, nullptr};
    uint32_t* err = s->call_error_word();
    const int ev = s->begin_op();
    {
        ErrorSink sink(err);
        check(sorted_anti_join_multi(a, nl, rights, o, nullptr, &cdev->merge_rows, temp, tb, s->stream,
                                     s->kernel_timer(ev)));
    }
    s->end_op(ev, kOpAntiJoin, 8.0 * static_cast<double>(nl + nr));
    uint64_t n = 0;
    check(hipMemcpyAsync(&n, &cdev->merge_rows, sizeof(n), hipMemcpyDeviceToHost, s->stream));
    s->check_call_error(err);
    if (s->profiling) s->collect_ops();
    ts->full.rows = n;
    if (rows) *rows = n;
    return SPARKALOG_HIP_OK;
    SPK_GUARD_END
}

extern "C" int sparkalog_hip_store_insert_recursive(sparkalog_hip_store* s, uint32_t id, const uint32_t* full0,
                                                    const uint32_t* full1, uint64_t full_rows, const uint32_t* delta0,
                                                    const uint32_t* delta1, uint64_t delta_rows, int on_device) {
    if (!s || (full_rows && (!full0 || !full1)) || (delta_rows && (!delta0 || !delta1)))
        return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    SPK_GUARD_BEGIN
    RelationState* st = slot(s, id);
    const uint32_t m0 = upload(s, st->full, full0, full1, full_rows, on_device);
    const uint32_t m1 = upload(s, st->delta, delta0, delta1, delta_rows, on_device);
    st->max_value = m0 > m1 ? m0 : m1;
    st->newt.rows = 0;
    st->present = true;
    st->recursive = true;
    return SPARKALOG_HIP_OK;
    SPK_GUARD_END
}

extern "C" int sparkalog_hip_store_clear(sparkalog_hip_store* s) {
    if (!s) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    for (auto& st : s->rel) {
        st.present = false;
        st.recursive = false;
        st.full.rows = st.delta.rows = st.newt.rows = 0;
        st.full.csr = false;
        st.full.touched();
        for (auto& r : st.runs) {
            s->give(std::move(r.keys));
            s->give_x(std::move(r.xidx));
            s->give_sec(std::move(r.sec));
        }
        st.runs.clear();
    }
    return SPARKALOG_HIP_OK;
}

extern "C" int sparkalog_hip_store_rows(sparkalog_hip_store* s, uint32_t id, int which, uint64_t* rows) {
    if (!s || !rows) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    RelationState* st = find(s, id);
    if (!st) return SPARKALOG_HIP_ERR_MISSING_RELATION;
    if (which == 1 && !st->recursive) return SPARKALOG_HIP_ERR_NOT_RECURSIVE;
    uint64_t full_rows = st->full.rows;
    for (const auto& r : st->runs) full_rows += r.rows;
    *rows = which == 0 ? full_rows : st->delta.rows;
    return SPARKALOG_HIP_OK;
}

extern "C" int sparkalog_hip_store_read(sparkalog_hip_store* s, uint32_t id, int which, uint32_t* c0, uint32_t* c1,
                                        uint64_t capacity, int to_device) {
    if (!s) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    RelationState* st = find(s, id);
    if (!st) return SPARKALOG_HIP_ERR_MISSING_RELATION;
    if (which == 1 && !st->recursive) return SPARKALOG_HIP_ERR_NOT_RECURSIVE;
    SPK_GUARD_BEGIN
    if (which == 0) settle_all(s, *st);
    const DevRelation& r = which == 0 ? st->full : st->delta;
    if (r.rows > capacity || (r.rows && (!c0 || !c1))) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    download(s, r, c0, c1, to_device);
    return SPARKALOG_HIP_OK;
    SPK_GUARD_END
}

extern "C" int sparkalog_hip_store_digest(sparkalog_hip_store* s, uint32_t id, uint64_t* digest, int* sorted_unique,
                                          uint64_t* rows) {
    if (!s) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    RelationState* st = find(s, id);
    if (!st) return SPARKALOG_HIP_ERR_MISSING_RELATION;
    SPK_GUARD_BEGIN
    settle_all(s, *st);
    s->misc_dev.reserve(64);
    clear_words_kernel<<<1, 64, 0, s->stream>>>(s->misc_dev.ptr, 4);
    const uint64_t n = st->full.rows;
    if (n) {
        digest_kernel<<<grid_for(n), kBlock, 0, s->stream>>>(
            st->full.keys.ptr, n, reinterpret_cast<unsigned long long*>(s->misc_dev.ptr), s->misc_dev.ptr + 2);
    }
    check(hipGetLastError());
    uint32_t host[4];
    check(hipMemcpyAsync(host, s->misc_dev.ptr, sizeof(host), hipMemcpyDeviceToHost, s->stream));
    check(hipStreamSynchronize(s->stream));
    uint64_t d;
    std::memcpy(&d, host, sizeof(d));
    if (digest) *digest = d;
    if (sorted_unique) *sorted_unique = host[2] == 0;
    if (rows) *rows = n;
    return SPARKALOG_HIP_OK;
    SPK_GUARD_END
}

extern "C" int sparkalog_hip_store_set_tuning(sparkalog_hip_store* s, uint32_t merge_ratio, int segmented_distinct,
                                              int indexed_anti_join, uint64_t sparse_probe_rows) {
    if (!s || merge_ratio < 2) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    s->sparse_probe_rows = sparse_probe_rows;
    s->merge_ratio = merge_ratio;
    s->segmented_distinct = segmented_distinct != 0;
    s->x_anti = indexed_anti_join != 0;
    return SPARKALOG_HIP_OK;
}

extern "C" int sparkalog_hip_store_graph_stats(sparkalog_hip_store* s, uint64_t* captures, uint64_t* replays) {
    if (!s) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    if (captures) *captures = s->graph.captures;
    if (replays) *replays = s->graph.replays;
    return SPARKALOG_HIP_OK;
}

extern "C" int sparkalog_hip_store_graph_capture_stats(sparkalog_hip_store* s, uint64_t* updates, double* capture_ms) {
    if (!s) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    if (updates) *updates = s->graph.updates;
    if (capture_ms) *capture_ms = s->graph.capture_ms;
    return SPARKALOG_HIP_OK;
}

extern "C" int sparkalog_hip_store_set_option(sparkalog_hip_store* s, const char* name, uint64_t value) {
    if (!s || !name) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    const std::string key(name);
    if (key == "graph") {
        if (value > 1) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
        s->graph_mode = static_cast<int>(value);
        return SPARKALOG_HIP_OK;
    }
    if (key == "graph_max_rows") {
        if (value < (1ull << 16) || value > (1ull << 31)) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
        s->graph_max_rows = value;
        return SPARKALOG_HIP_OK;
    }
    if (key == "join_index") {
        if (value > 2) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
        s->join_index_mode = static_cast<int>(value);
        // cached indexes are rebuilt in the new layout
        for (auto& rt : s->ws_rules) rt.index.valid = false;
        return SPARKALOG_HIP_OK;
    }
    if (key == "distinct_tile") {
        if (value < 512 || value > 4096 || value % 64 != 0) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
        s->seg_tile = static_cast<uint32_t>(value);
        return SPARKALOG_HIP_OK;
    }
    // the set_tuning knobs one at a time
    if (key == "merge_ratio") {
        if (value < 2 || value > 64) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
        s->merge_ratio = static_cast<uint32_t>(value);
        return SPARKALOG_HIP_OK;
    }
    if (key == "sparse_probe_rows") {
        s->sparse_probe_rows = value;
        return SPARKALOG_HIP_OK;
    }
    if (key == "segmented_distinct" || key == "indexed_anti_join") {
        if (value > 1) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
        (key == "segmented_distinct" ? s->segmented_distinct : s->x_anti) = value != 0;
        return SPARKALOG_HIP_OK;
    }
    return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
}

extern "C" int sparkalog_hip_store_set_profiling(sparkalog_hip_store* s, int enabled) {
    if (!s) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    s->profiling = enabled != 0;
    for (auto& st : s->stats) {
        st.launches = 0;
        st.ms = 0;
        st.bytes = 0;
        st.kernel_launches = 0;
        st.kernel_ms = 0;
    }
    return SPARKALOG_HIP_OK;
}

extern "C" int sparkalog_hip_store_op_stats(sparkalog_hip_store* s, int op, uint64_t* launches, double* ms,
                                            double* bytes) {
    if (!s || op < 0 || op >= kOpCount) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    if (launches) *launches = s->stats[op].launches;
    if (ms) *ms = s->stats[op].ms;
    if (bytes) *bytes = s->stats[op].bytes;
    return SPARKALOG_HIP_OK;
}

extern "C" int sparkalog_hip_store_op_kernel_stats(sparkalog_hip_store* s, int op, const char** name,
                                                   const char** kernel, uint64_t* kernel_launches,
                                                   double* kernel_ms) {
    if (op < 0 || op >= kOpCount) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    if (name) *name = kOpNames[op];
    if (kernel) *kernel = s && s->stats[op].kernel ? s->stats[op].kernel : kOpKernels[op];
    if (kernel_launches) *kernel_launches = s ? s->stats[op].kernel_launches : 0;
    if (kernel_ms) *kernel_ms = s ? s->stats[op].kernel_ms : 0.0;
    return SPARKALOG_HIP_OK;
}

extern "C" int sparkalog_hip_executor_compile(sparkalog_hip_store* s, const uint32_t* relations,
                                              uint32_t n_relations, const sparkalog_hip_rule* rules,
                                              uint32_t n_rules) {
    if (!s || (n_relations && !relations) || (n_rules && !rules)) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    SPK_GUARD_BEGIN
    std::vector<TargetRuntime> targets;
    std::vector<RuleRuntime> rts;
    return compile_plan(s, relations, n_relations, rules, n_rules, targets, rts);
    SPK_GUARD_END
}

namespace spk {
namespace {

// State of one fixpoint between begin and end (the loop of generic.rs:472-516
// split into steps so that a caller may change relations between steps).
struct Session {
    bool active = false;
    bool graph = false;  // iterations replayed from captured graphs (graph_step)
    std::vector<uint32_t> relations;
    std::vector<sparkalog_hip_rule> rules;
    uint64_t iterations = 0, total_new = 0;
    double algo_bytes = 0.0;
    hipEvent_t t_begin = nullptr, t_end = nullptr;
    // per-instance tracking: instance of a row = first >> instance_bits
    uint32_t instance_bits = 0, instances = 0;
    DevBuf<uint32_t> last_new;  // per instance: 1 + last iteration with new rows
    // per instance: DELTA rows after the last step, counted on the device
    // from NEWT and read back with the step's counters (pinned)
    DevBuf<uint64_t> inst_delta;
    uint64_t* inst_delta_host = nullptr;
    uint32_t inst_delta_cap = 0;
    bool inst_delta_valid = false;
    // per-instance rows and digests of the one tracked relation's FULL,
    // accumulated from its initial FULL and every NEWT the distinct writes
    // (stat_live while every step's NEWT was counted); after the session
    // they answer instance_stats while the relation keeps stat_version
    DevBuf<uint64_t> stat_acc;
    bool stat_live = false, stat_valid = false;
    uint32_t stat_rel = 0, stat_bits = 0, stat_instances = 0;
    uint64_t stat_version = 0;
};

// Rows of each instance in sorted packed keys (count *rows_dev <= bound):
// thread j adds the rows whose first column >> bits == j to out[j].
__global__ void instance_rows_kernel(const uint64_t* keys, const uint64_t* rows_dev, size_t bound, uint32_t bits,
                                     uint32_t instances, unsigned long long* out) {
    const size_t n = *rows_dev < bound ? *rows_dev : bound;
    const auto lower = [&](uint64_t j) -> size_t {
        const uint64_t first = j << bits;
        if (first > 0xFFFFFFFFull) return n;
        const uint64_t target = first << 32;
        size_t lo = 0, hi = n;
        while (lo < hi) {
            const size_t mid = (lo + hi) >> 1;
            if (keys[mid] < target)
                lo = mid + 1;
            else
                hi = mid;
        }
        return lo;
    };
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t j = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < instances; j += stride) {
        const size_t c = lower(j + 1) - lower(j);
        if (c) atomicAdd(&out[j], static_cast<unsigned long long>(c));
    }
}

Session& session_of(sparkalog_hip_store* s) {
    if (!s->session) s->session = new Session();
    return *static_cast<Session*>(s->session);
}

void destroy_session(sparkalog_hip_store* s) {
    if (!s->session) return;
    Session* ss = static_cast<Session*>(s->session);
    if (ss->t_begin) (void)hipEventDestroy(ss->t_begin);
    if (ss->t_end) (void)hipEventDestroy(ss->t_end);
    if (ss->inst_delta_host) (void)hipHostFree(ss->inst_delta_host);
    delete ss;
    s->session = nullptr;
}

__global__ void track_instances_kernel(const uint64_t* keys, size_t n, uint32_t bits, uint32_t instances,
                                       uint32_t value, uint32_t* last_new, const uint64_t* n_dev = nullptr,
                                       const uint64_t* value_dev = nullptr) {
    if (n_dev) n = *n_dev < n ? static_cast<size_t>(*n_dev) : n;
    if (value_dev) value = static_cast<uint32_t>(*value_dev);
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t inst = static_cast<uint32_t>(keys[i] >> 32) >> bits;
        if (inst < instances && (i == 0 || (static_cast<uint32_t>(keys[i - 1] >> 32) >> bits) != inst))
            atomicMax(&last_new[inst], value);
    }
}

int session_begin(sparkalog_hip_store* s, const uint32_t* relations, uint32_t n_relations,
                  const sparkalog_hip_rule* rules, uint32_t n_rules) {
    Session& ss = session_of(s);
    const int compiled = compile_plan(s, relations, n_relations, rules, n_rules, s->ws_targets, s->ws_rules);
    if (compiled != SPARKALOG_HIP_OK) return compiled;
    ss.relations.assign(relations, relations + n_relations);
    ss.rules.assign(rules, rules + n_rules);
    // graph-replayed iterations serve one rule whose right input is static
    ss.graph = false;
    if (s->graph_mode && n_relations == 1 && n_rules == 1 && rules[0].target == relations[0] &&
        rules[0].delta_input == relations[0] && rules[0].right_input != relations[0]) {
        const RelationState* right = find(s, rules[0].right_input);
        ss.graph = right && !right->recursive;
    }
    s->graph.seen_full = s->graph.seen_delta = nullptr;  // device row counts rewritten at the first step
    // launch capacities follow this fixpoint's sizes (the buffers keep their
    // allocations): a small fixpoint after a large one launches small grids
    s->graph.cap_full = s->graph.cap_cand = s->graph.cap_delta = 0;
    ss.iterations = ss.total_new = 0;
    ss.algo_bytes = 0.0;
    s->events_used = 0;
    if (!ss.t_begin) {
        check(hipEventCreate(&ss.t_begin));
        check(hipEventCreate(&ss.t_end));
    }
    check(hipEventRecord(ss.t_begin, s->stream));
    s->counters_dev.reserve(sizeof(Counters) / sizeof(uint64_t));
    s->misc_dev.reserve(256);
    if (ss.instance_bits) {
        ss.last_new.reserve(ss.instances ? ss.instances : 1);
        clear_words_kernel<<<1, 1024, 0, s->stream>>>(ss.last_new.ptr, ss.instances);
        check(hipGetLastError());
        ss.inst_delta.reserve(ss.instances ? ss.instances : 1);
        if (ss.inst_delta_cap < ss.instances) {
            std::lock_guard<std::recursive_mutex> lock(device_mutex());
            if (ss.inst_delta_host) (void)hipHostFree(ss.inst_delta_host);
            ss.inst_delta_host = nullptr;
            void* p = nullptr;
            check(hipHostMalloc(&p, sizeof(uint64_t) * ss.instances, hipHostMallocDefault));
            ss.inst_delta_host = static_cast<uint64_t*>(p);
            ss.inst_delta_cap = ss.instances;
        }
    }
    ss.inst_delta_valid = false;
    ss.stat_live = ss.stat_valid = false;
    if (ss.instance_bits && ss.instances && n_relations == 1 && !ss.graph) {
        // the initial FULL's statistics, when all of it is packed keys
        RelationState& st = *find(s, relations[0]);
        bool packed = !st.full.csr;
        for (const auto& r : st.runs) packed = packed && !r.csr;
        if (packed) {
            // rows, then digests, each kInstanceStatSlots copies of every instance
            const size_t n_acc = 2 * static_cast<size_t>(kInstanceStatSlots) * ss.instances;
            ss.stat_acc.reserve(n_acc);
            unsigned long long* acc = reinterpret_cast<unsigned long long*>(ss.stat_acc.ptr);
            check(fill_u64(ss.stat_acc.ptr, n_acc, 0, s->stream));
            const auto add = [&](const DevRelation& part) {
                if (!part.rows) return;
                const size_t waves = (part.rows + kStatsWaveRows - 1) / kStatsWaveRows;
                const size_t blocks = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
                instance_stats_kernel<<<static_cast<unsigned>(blocks), kBlock, 0, s->stream>>>(
                    part.keys.ptr, part.rows, ss.instance_bits, ss.instances, acc,
                    acc + static_cast<size_t>(kInstanceStatSlots) * ss.instances);
            };
            add(st.full);
            for (const auto& r : st.runs) add(r);
            check(hipGetLastError());
            ss.stat_live = true;
        }
    }
    ss.active = true;
    return SPARKALOG_HIP_OK;
}

uint64_t pending_delta(sparkalog_hip_store* s) {
    uint64_t pending = 0;
    for (uint32_t r : session_of(s).relations) pending += find(s, r)->delta.rows;
    return pending;
}

void graph_step(sparkalog_hip_store* s);

// One semi-naive iteration: evaluate every rule, combine per target, apply
// per target (generic.rs:481-512).
void session_step(sparkalog_hip_store* s) {
    Session& ss = session_of(s);
    if (ss.graph) {
        graph_step(s);
        return;
    }
    std::vector<TargetRuntime>& targets = s->ws_targets;
    std::vector<RuleRuntime>& rts = s->ws_rules;
    const uint32_t n_rules = static_cast<uint32_t>(ss.rules.size());
    const uint32_t n_relations = static_cast<uint32_t>(ss.relations.size());
    const uint32_t* relations = ss.relations.data();
    Counters* cdev = reinterpret_cast<Counters*>(s->counters_dev.ptr);
    Counters& chost = s->host_counters();
    uint32_t* lb_error = s->misc_dev.ptr + 128;  // device error flag of this step's kernels
    // every look-back launcher of the step reports into it (the index build
    // below keeps a word of its own, since the step's word is cleared by the
    // first join count)
    ErrorSink step_sink(lb_error);
    uint64_t& iterations = ss.iterations;
    uint64_t& total_new = ss.total_new;
    double& algo_bytes = ss.algo_bytes;
    // the largest value of any stored relation bounds every key of this step
    uint32_t vmax = 0;
    for (const auto& st : s->rel)
        if (st.present) vmax = st.max_value > vmax ? st.max_value : vmax;
    const uint64_t vmask = bit_mask_for(vmax);
    uint32_t vbits = 0;
    while ((1ull << vbits) <= vmask && vbits < 32) ++vbits;
    if (vbits == 0) vbits = 1;
    const uint32_t shift = vbits <= 24 ? vbits : 32;  // compressed sort key width per column
    const uint64_t known_compressed = shift == 32 ? ((vmask << 32) | vmask) : ((vmask << shift) | vmask);
    // FULL's parts carry a first-column index (n_x + 1 u32 entries each)
    // while first values stay below 2^24
    const uint32_t n_x = vmax + 1;
    const bool use_x = s->x_anti && vmax < (1u << 24);

    // the step's status words are cleared by its first join count (or below
    // when no rule has DELTA rows)
    bool status_cleared = false;
    for (auto& t : targets) {
        t.fused = false;
        t.instances_marked = false;
        t.csr_run = false;
        t.anti_slot = -1;
    }
    // ---- RuleRuntime::evaluate, join count (generic.rs:197-241) ----
    for (uint32_t r = 0; r < n_rules; ++r) {
        RuleRuntime& rt = rts[r];
        RelationState& ds = *find(s, rt.plan.delta_input);
        RelationState& rs = *find(s, rt.plan.right_input);
        build_index(s, rt, rs, rt.plan.right_input);
        rt.delta_rows = ds.delta.rows;
        rt.shift = shift;
        rt.distinct_slot = -1;
        const uint64_t L = rt.delta_rows;
        if (L == 0) {
            check(set_u64(&cdev->join_total[r], 0, s->stream));
            continue;
        }
        rt.offsets.reserve(L);
        rt.match.reserve(L);
        const size_t n_tiles = tiles(L, kCountTile);
        rt.tile_base.reserve(n_tiles);
        const int ev = s->begin_op();
        // tile-local offsets as u32 while every count tile's total fits
        rt.narrow = static_cast<uint64_t>(rt.index.max_degree) * kCountTile < (1ull << 32);
        const uint32_t clear_words = status_cleared ? 0u : 256u;
        status_cleared = true;
        timer_start(s->kernel_timer(ev), s->stream);
        const KeyIndex kix = rt.index.view();
        uint32_t* local32 = reinterpret_cast<uint32_t*>(rt.offsets.ptr);
        const unsigned count_grid = static_cast<unsigned>(n_tiles);
        if (rt.narrow && !rt.index.hashed)
            dense_join_count_kernel<uint32_t, false><<<count_grid, kBlock, 0, s->stream>>>(
                ds.delta.keys.ptr, L, rt.plan.left_key, kix, local32, rt.match.ptr, rt.tile_base.ptr, s->misc_dev.ptr,
                clear_words);
        else if (rt.narrow)
            dense_join_count_kernel<uint32_t, true><<<count_grid, kBlock, 0, s->stream>>>(
                ds.delta.keys.ptr, L, rt.plan.left_key, kix, local32, rt.match.ptr, rt.tile_base.ptr, s->misc_dev.ptr,
                clear_words);
        else if (!rt.index.hashed)
            dense_join_count_kernel<uint64_t, false><<<count_grid, kBlock, 0, s->stream>>>(
                ds.delta.keys.ptr, L, rt.plan.left_key, kix, rt.offsets.ptr, rt.match.ptr, rt.tile_base.ptr,
                s->misc_dev.ptr, clear_words);
        else
            dense_join_count_kernel<uint64_t, true><<<count_grid, kBlock, 0, s->stream>>>(
                ds.delta.keys.ptr, L, rt.plan.left_key, kix, rt.offsets.ptr, rt.match.ptr, rt.tile_base.ptr,
                s->misc_dev.ptr, clear_words);
        timer_stop(s->kernel_timer(ev), s->stream);
        join_scan_kernel<<<1, kScanThreads, 0, s->stream>>>(rt.tile_base.ptr, n_tiles, &cdev->join_total[r]);
        check(hipGetLastError());
        s->end_op(ev, kOpJoinCount, 8.0 * L);
    }
    if (!status_cleared) clear_words_kernel<<<1, 256, 0, s->stream>>>(s->misc_dev.ptr, 256);
    check(hipMemcpyAsync(&chost, cdev, sizeof(Counters), hipMemcpyDeviceToHost, s->stream));
    check(hipEventRecord(s->counters_ready, s->stream));
    // the previous step's run merges: queued behind the join count (which
    // reads DELTA and the index, not FULL), they keep the device busy while
    // the host waits for the join sizes; the probe below sees FULL settled
    for (uint32_t i = 0; i < n_relations; ++i) {
        RelationState& ts = *find(s, relations[i]);
        if (ts.settle_pending) {
            ts.settle_pending = false;
            algo_bytes += settle_runs(s, ts, false);
        }
    }
    check(hipEventSynchronize(s->counters_ready));
    // ---- join emit + distinct (generic.rs:231-249) ----
    for (uint32_t r = 0; r < n_rules; ++r) {
        RuleRuntime& rt = rts[r];
        const uint64_t M = rt.delta_rows ? chost.join_total[r] : 0;
        rt.join_rows = M;
        if (M > 0xFFFFFFFFull) throw StatusError{SPARKALOG_HIP_ERR_TOO_MANY_ROWS};
        RelationState& ds = *find(s, rt.plan.delta_input);
        rt.cand_a.reserve(M ? M : 1);
        rt.cand_b.reserve(M ? M : 1);
        if (M == 0) {
            check(set_u64(&cdev->distinct_rows[r], 0, s->stream));
            continue;
        }
        EmitSpec spec;
        spec.key_col = rt.plan.left_key;
        spec.right_key = rt.plan.right_key;
        spec.side[0] = rt.plan.out_input[0];
        spec.side[1] = rt.plan.out_input[1];
        spec.col[0] = rt.plan.out_column[0];
        spec.col[1] = rt.plan.out_column[1];
        spec.order[0] = rt.plan.distinct_columns[0];
        spec.order[1] = rt.plan.distinct_columns[1];
        spec.shift = rt.shift;
        int ev = s->begin_op();
        const uint64_t emit_tiles = tiles(M, kEmitTile);
        const unsigned emit_grid = static_cast<unsigned>(emit_tiles < 65536 ? emit_tiles : 65536);
        rt.row_start.reserve(emit_tiles + 1);
        const uint32_t* local32 = reinterpret_cast<const uint32_t*>(rt.offsets.ptr);
        if (rt.narrow)
            emit_partition_kernel<uint32_t><<<grid_for(emit_tiles + 1), kBlock, 0, s->stream>>>(
                local32, rt.tile_base.ptr, rt.delta_rows, M, rt.row_start.ptr, emit_tiles);
        else
            emit_partition_kernel<uint64_t><<<grid_for(emit_tiles + 1), kBlock, 0, s->stream>>>(
                rt.offsets.ptr, rt.tile_base.ptr, rt.delta_rows, M, rt.row_start.ptr, emit_tiles);
        timer_start(s->kernel_timer(ev), s->stream);
        if (rt.narrow)
            dense_join_emit_kernel<uint32_t><<<emit_grid, kBlock, 0, s->stream>>>(
                ds.delta.keys.ptr, rt.delta_rows, rt.index.other.ptr, local32, rt.tile_base.ptr, rt.match.ptr, M,
                rt.row_start.ptr, spec, rt.cand_a.ptr);
        else
            dense_join_emit_kernel<uint64_t><<<emit_grid, kBlock, 0, s->stream>>>(
                ds.delta.keys.ptr, rt.delta_rows, rt.index.other.ptr, rt.offsets.ptr, rt.tile_base.ptr,
                rt.match.ptr, M, rt.row_start.ptr, spec, rt.cand_a.ptr);
        timer_stop(s->kernel_timer(ev), s->stream);
        check(hipGetLastError());
        s->end_op(ev, kOpJoinEmit, 8.0 * rt.index.rows + 8.0 * M);
        // The join writes candidates in DELTA order; when the distinct's major
        // column is DELTA's first column they are grouped by it already.
        const bool grouped = spec.side[spec.order[0]] == 0 && spec.col[spec.order[0]] == 0;
        // A target fed by this rule alone takes NEWT = distinct(candidates)
        // \ FULL as distinct(candidates \ FULL): candidates already in FULL
        // are found by probing FULL's runs and dropped before the sort, and
        // the distinct writes NEWT and FULL's new run itself.
        int fuse = -1;
        if (grouped && s->segmented_distinct && !use_x)
            for (size_t ti = 0; ti < targets.size(); ++ti)
                if (targets[ti].rules.size() == 1 && targets[ti].rules[0] == r) fuse = static_cast<int>(ti);
        const uint64_t* found = nullptr;
        if (fuse >= 0) {
            TargetRuntime& t = targets[static_cast<size_t>(fuse)];
            RelationState& ts = *find(s, t.target);
            if (ts.runs.size() >= kMaxRuns) settle_runs(s, ts, true);
            AntiRight rights{};
            rights.count = 0;
            double right_rows = 0;
            auto add_right = [&](const DevRelation& x) {
                if (!x.rows) return;
                rights.in[rights.count] = PairInput{x.keys.ptr, nullptr, nullptr, nullptr};
                rights.rows[rights.count] = x.rows;
                ++rights.count;
                right_rows += static_cast<double>(x.rows);
            };
            add_right(ts.full);
            for (const auto& x : ts.runs) add_right(x);
            bool column_probe = false, sparse_probe = false;
            if (rights.count) {
                rt.found.reserve(probe_found_words(M));
                bool parts_fit = ts.full.rows <= 0xFFFFFFFFull;
                for (const auto& x : ts.runs) parts_fit = parts_fit && x.rows <= 0xFFFFFFFFull;
                // FULL's parts are probed through their first-column indexes
                // while first values stay below 2^24
                const bool indexed = vmax < (1u << 24) && parts_fit;
                if (M <= s->sparse_probe_rows && indexed) {
                    XSecRights xr{};
                    xr.count = 0;
                    auto add_part = [&](DevRelation& x) {
                        if (!x.rows) return;
                        ensure_xidx(s, x, n_x);
                        ensure_sec(s, x);
                        xr.second[xr.count] = x.sec.ptr;
                        xr.xstart[xr.count] = x.xidx.ptr;
                        ++xr.count;
                    };
                    add_part(ts.full);
                    for (auto& x : ts.runs) add_part(x);
                    const int pev = s->begin_op();
                    check(probe_sparse_sec(rt.cand_a.ptr, M, rt.shift, xr, n_x, rt.found.ptr, s->stream,
                                           s->kernel_timer(pev)));
                    s->stats[kOpProbeSparse].kernel = "probe_sparse_sec_kernel";
                    // reads the candidates and, per candidate, a few rows of
                    // each run (not counted): not FULL as a whole
                    s->end_op(pev, kOpProbeSparse, 8.0 * M);
                    t.anti_slot = pev;
                    sparse_probe = true;
                } else if (indexed) {
                    // stream only the second columns of FULL's parts (4 bytes a row)
                    XSecRights xr{};
                    xr.count = 0;
                    auto add_part = [&](DevRelation& x) {
                        if (!x.rows) return;
                        ensure_xidx(s, x, n_x);
                        ensure_sec(s, x);
                        xr.second[xr.count] = x.sec.ptr;
                        xr.xstart[xr.count] = x.xidx.ptr;
                        ++xr.count;
                    };
                    add_part(ts.full);
                    for (auto& x : ts.runs) add_part(x);
                    const size_t xtb = probe_xsec_temp_bytes(M);
                    void* xtemp = s->scratch(xtb);
                    const int pev = s->begin_op();
                    check(probe_xsec(rt.cand_a.ptr, M, rt.shift, xr, n_x, rt.found.ptr, xtemp, xtb, s->stream,
                                     s->kernel_timer(pev)));
                    s->stats[kOpAntiJoin].kernel = "probe_xsec_kernel";
                    s->end_op(pev, kOpAntiJoin, 8.0 * M + 4.0 * right_rows);
                    t.anti_slot = pev;
                    column_probe = true;
                } else {
                    // packed keys of every part (a part held as index +
                    // second column is expanded first)
                    packed_parts(s, ts);
                    rights.count = 0;
                    right_rows = 0;
                    add_right(ts.full);
                    for (const auto& x : ts.runs) add_right(x);
                    const size_t ptb = probe_grouped_temp_bytes(M, rights.count);
                    void* ptemp = s->scratch(ptb);
                    const int pev = s->begin_op();
                    check(probe_grouped(rt.cand_a.ptr, nullptr, M, rt.shift, rights, rt.found.ptr, ptemp, ptb,
                                        s->stream, s->kernel_timer(pev)));
                    s->stats[kOpAntiJoin].kernel = "anti_join_kernel";
                    s->end_op(pev, kOpAntiJoin, 8.0 * M + 8.0 * right_rows);
                    t.anti_slot = pev;
                }
                found = rt.found.ptr;
            }
            algo_bytes += sparse_probe ? 8.0 * M : (column_probe ? 4.0 : 8.0) * right_rows + 8.0 * M;
            t.newt.reserve(M);
            // while FULL can be probed by columns the new run is kept as its
            // first-column index (built from NEWT, the same rows) and second
            // column only; else as packed keys (NEWT's second copy)
            t.csr_run = vmax < (1u << 24);
            if (t.csr_run) {
                t.new_sec = s->take_sec(M + kSecPad);
                t.new_xidx = s->take_x(static_cast<size_t>(n_x) + 1);
                check(xindex_clear(t.new_xidx.ptr, n_x, s->stream));
            } else {
                t.new_run = s->take(M);
            }
            t.fused = true;
            t.contrib_bound = M;
        }
        ev = s->begin_op();
        rt.distinct_slot = ev;
        if (grouped && s->segmented_distinct) {
            const size_t mb = segmented_distinct_meta_bytes(M, s->seg_tile);
            rt.seg_meta.reserve(mb);
            uint64_t& side_rows = s->words_host[1];
            side_rows = 0;
            check(segmented_distinct_prepare(rt.cand_a.ptr, M, rt.shift, rt.cand_b.ptr, rt.seg_meta.ptr, mb,
                                             &side_rows, s->stream, found, s->seg_tile));
            size_t tb = 0;
            void* temp = nullptr;
            if (side_rows) {
                rt.cand_c.reserve(side_rows);
                tb = distinct_temp_bytes(side_rows);
                temp = s->scratch(tb);
            }
            // a fused distinct writes NEWT and the new run instead
            if (fuse < 0) rt.distinct_out.reserve(M);
            PairOutput out{rt.distinct_out.ptr, nullptr, nullptr};
            uint64_t* count = &cdev->distinct_rows[r];
            if (fuse >= 0) {
                TargetRuntime& t = targets[static_cast<size_t>(fuse)];
                out = PairOutput{t.newt.ptr, nullptr, nullptr, t.csr_run ? nullptr : t.new_run.ptr, t.new_sec.ptr};
                count = &cdev->newt_rows[fuse];
            }
            uint32_t* xidx_out = nullptr;
            if (fuse >= 0 && targets[static_cast<size_t>(fuse)].csr_run)
                xidx_out = targets[static_cast<size_t>(fuse)].new_xidx.ptr;
            // NEWT written here is the next DELTA: its instances are marked
            // as it is written (instead of a pass over DELTA after the step)
            InstanceMarks marks;
            if (fuse >= 0 && ss.instance_bits) {
                marks = InstanceMarks{ss.last_new.ptr, ss.instance_bits, ss.instances,
                                      static_cast<uint32_t>(iterations + 1)};
                if (ss.stat_live) {
                    marks.rows_acc = reinterpret_cast<unsigned long long*>(ss.stat_acc.ptr);
                    marks.digest_acc = marks.rows_acc + static_cast<size_t>(kInstanceStatSlots) * ss.instances;
                }
                targets[static_cast<size_t>(fuse)].instances_marked = true;
            }
            check(segmented_distinct_finish(rt.cand_a.ptr, M, rt.shift, known_compressed, rt.cand_b.ptr,
                                            side_rows ? rt.cand_c.ptr : nullptr, side_rows, out, count,
                                            rt.seg_meta.ptr, mb, temp, tb, s->stream, s->kernel_timer(ev), found,
                                            xidx_out, s->seg_tile, marks));
            if (xidx_out) {
                // the new run's index: entries of first values without new rows
                const size_t ftb = xindex_fill_temp_bytes(n_x);
                void* ftemp = s->scratch(ftb);
                check(xindex_fill(xidx_out, n_x, count, ftemp, ftb, s->stream));
            }
        } else {
            const size_t tb = distinct_temp_bytes(M);
            void* temp = s->scratch(tb);
            PairInput in{rt.cand_a.ptr, nullptr, nullptr, nullptr};
            rt.distinct_out.reserve(M);
            PairOutput out{rt.distinct_out.ptr, nullptr, nullptr};
            check(distinct(in, M, rt.cand_a.ptr, rt.cand_b.ptr, out, &cdev->distinct_rows[r], temp, tb,
                           known_compressed, s->stream));
            if (rt.shift != 32)
                expand_keys_kernel<<<grid_for(M), kBlock, 0, s->stream>>>(rt.distinct_out.ptr,
                                                                         &cdev->distinct_rows[r], M, rt.shift);
        }
        check(hipGetLastError());
        s->end_op(ev, kOpDistinct, 8.0 * M);  // + 8U added after the sync
    }
    // ---- TargetRuntime::combine (generic.rs:299-358) ----
    for (auto& t : targets) {
        t.union_slots.clear();
        if (t.fused) continue;  // NEWT written by the distinct
        RelationState& ts = *find(s, t.target);
        const uint64_t* contrib = nullptr;
        const uint64_t* contrib_rows = nullptr;
        uint64_t bound = 0;
        for (size_t k = 0; k < t.rules.size(); ++k) {
            RuleRuntime& rt = rts[t.rules[k]];
            if (k == 0) {
                contrib = rt.distinct_out.ptr;
                contrib_rows = &cdev->distinct_rows[t.rules[k]];
                bound = rt.join_rows;
                continue;
            }
            const uint64_t nb = bound + rt.join_rows;
            t.union_tmp.reserve(nb ? nb : 1);
            const size_t tb = union_temp_bytes(bound, rt.join_rows);
            void* temp = s->scratch(tb);
            PairInput a{contrib, nullptr, nullptr, contrib_rows};
            PairInput b{rt.distinct_out.ptr, nullptr, nullptr, &cdev->distinct_rows[t.rules[k]]};
            PairOutput o{t.union_tmp.ptr, nullptr, nullptr};
            const int ev = s->begin_op();
            t.union_slots.push_back(ev);
            check(sorted_union(a, bound, b, rt.join_rows, o, &cdev->contrib_rows[t.rules[k]], temp, tb,
                               s->stream, s->kernel_timer(ev)));
            s->end_op(ev, kOpUnion, 0.0);  // bytes added after the sync
            std::swap(t.union_tmp, t.contributions);
            contrib = t.contributions.ptr;
            contrib_rows = &cdev->contrib_rows[t.rules[k]];
            bound = nb;
        }
        t.contrib_bound = bound;
        // NEWT = contributions \ FULL, as a chain of anti-joins over FULL's
        // runs (largest first). The last one also writes NEWT's copy that
        // becomes FULL's newest run.
        const uint32_t ti = static_cast<uint32_t>(&t - targets.data());
        t.newt.reserve(bound ? bound : 1);
        t.new_run = s->take(bound ? bound : 1);
        if (ts.runs.size() >= kMaxRuns) settle_runs(s, ts, true);
        packed_parts(s, ts);  // this anti-join reads FULL's packed keys
        AntiRight rights{};
        rights.count = 0;
        double right_rows = 0;
        auto add_right = [&](const DevRelation& r) {
            if (!r.rows) return;
            rights.in[rights.count] = PairInput{r.keys.ptr, nullptr, nullptr, nullptr};
            rights.rows[rights.count] = r.rows;
            ++rights.count;
            right_rows += static_cast<double>(r.rows);
        };
        add_right(ts.full);
        for (const auto& r : ts.runs) add_right(r);
        bool parts_fit = true;  // index entries are u32 row counts
        if (ts.full.rows > 0xFFFFFFFFull) parts_fit = false;
        for (const auto& r : ts.runs) parts_fit = parts_fit && r.rows <= 0xFFFFFFFFull;
        if (bound == 0) {
            check(set_u64(&cdev->newt_rows[ti], 0, s->stream));
        } else if (rights.count == 0) {
            check(copy_keys(contrib, contrib_rows, bound, t.newt.ptr, t.new_run.ptr, &cdev->newt_rows[ti],
                            s->stream));
        } else if (use_x && parts_fit) {
            XRights xr{};
            xr.count = 0;
            auto add_part = [&](DevRelation& r) {
                if (!r.rows) return;
                ensure_xidx(s, r, n_x);
                xr.keys[xr.count] = r.keys.ptr;
                xr.xstart[xr.count] = r.xidx.ptr;
                ++xr.count;
            };
            add_part(ts.full);
            for (auto& r : ts.runs) add_part(r);
            const size_t tb = anti_join_x_temp_bytes(bound);
            void* temp = s->scratch(tb);
            PairOutput o{t.newt.ptr, nullptr, nullptr, t.new_run.ptr};
            const int ev = s->begin_op();
            check(anti_join_x(contrib, contrib_rows, bound, xr, n_x, o, &cdev->newt_rows[ti], temp, tb, s->stream,
                              s->kernel_timer(ev)));
            s->stats[kOpAntiJoin].kernel = "anti_join_x_kernel";
            s->end_op(ev, kOpAntiJoin, 8.0 * right_rows);
            t.anti_slot = ev;
            algo_bytes += 8.0 * right_rows;
        } else {
            const size_t tb = anti_join_temp_bytes(bound, rights.count);
            void* temp = s->scratch(tb);
            PairInput a{contrib, nullptr, nullptr, contrib_rows};
            PairOutput o{t.newt.ptr, nullptr, nullptr, t.new_run.ptr};
            const int ev = s->begin_op();
            check(sorted_anti_join_multi(a, bound, rights, o, nullptr, &cdev->newt_rows[ti], temp, tb, s->stream,
                                         s->kernel_timer(ev)));
            s->stats[kOpAntiJoin].kernel = "anti_join_kernel";
            s->end_op(ev, kOpAntiJoin, 8.0 * right_rows);
            t.anti_slot = ev;
            algo_bytes += 8.0 * right_rows;
        }
    }
    // per instance: its NEWT rows (the next DELTA), read back with the counters
    if (ss.instance_bits && ss.instances) {
        clear_words_kernel<<<1, 1024, 0, s->stream>>>(reinterpret_cast<uint32_t*>(ss.inst_delta.ptr),
                                                      2 * static_cast<size_t>(ss.instances));
        for (uint32_t i = 0; i < n_relations; ++i) {
            TargetRuntime& t = targets[i];
            if (!t.newt.ptr) continue;
            instance_rows_kernel<<<grid_for(ss.instances), kBlock, 0, s->stream>>>(
                t.newt.ptr, &cdev->newt_rows[i], t.newt.cap, ss.instance_bits, ss.instances,
                reinterpret_cast<unsigned long long*>(ss.inst_delta.ptr));
        }
        check(hipGetLastError());
        check(hipMemcpyAsync(ss.inst_delta_host, ss.inst_delta.ptr, sizeof(uint64_t) * ss.instances,
                             hipMemcpyDeviceToHost, s->stream));
        ss.inst_delta_valid = true;
    }
    check(hipMemcpyAsync(&chost, cdev, sizeof(Counters), hipMemcpyDeviceToHost, s->stream));
    uint32_t* err_host = reinterpret_cast<uint32_t*>(&s->words_host[0]);
    *err_host = 0;
    check(hipMemcpyAsync(err_host, lb_error, sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream));
    check(hipStreamSynchronize(s->stream));
    if (*err_host) throw StatusError{SPARKALOG_HIP_ERR_DEVICE};
    // per-iteration algorithmic bytes of the operators run (pairs are 8 bytes)
    auto fused_target = [&](uint32_t r) -> int {
        for (size_t ti = 0; ti < targets.size(); ++ti)
            if (targets[ti].fused && targets[ti].rules[0] == r) return static_cast<int>(ti);
        return -1;
    };
    for (uint32_t r = 0; r < n_rules; ++r) {
        const double L = static_cast<double>(rts[r].delta_rows), M = static_cast<double>(rts[r].join_rows);
        const int fr = fused_target(r);
        // a fused distinct writes NEWT and FULL's new run: a packed copy
        // (8 B a row), or only its second column (4 B) when the run is kept
        // as index + second column
        const double U = !rts[r].join_rows ? 0.0
                         : fr >= 0 ? (targets[static_cast<size_t>(fr)].csr_run ? 1.5 : 2.0) *
                                         static_cast<double>(chost.newt_rows[fr])
                                   : static_cast<double>(chost.distinct_rows[r]);
        const double R = static_cast<double>(rts[r].index.rows);
        if (L > 0) algo_bytes += 8 * L + 8 * R + 8 * M + (M > 0 ? 8 * M + 8 * U : 0.0);
        s->add_op_bytes(rts[r].distinct_slot, 8 * U);
    }
    for (uint32_t i = 0; i < n_relations; ++i) {
        TargetRuntime& t = targets[i];
        if (t.fused) continue;  // counted with the rule's distinct and probe
        const uint64_t C = t.rules.size() > 1 ? chost.contrib_rows[t.rules.back()]
                                              : (t.contrib_bound ? chost.distinct_rows[t.rules[0]] : 0);
        // unions of the rules' contributions: 8 bytes per input and output row
        for (size_t k = 1; k < t.rules.size(); ++k) {
            const uint32_t rk = t.rules[k];
            const double a = static_cast<double>(k == 1 ? (rts[t.rules[0]].join_rows ? chost.distinct_rows[t.rules[0]] : 0)
                                                        : chost.contrib_rows[t.rules[k - 1]]);
            const double b = static_cast<double>(rts[rk].join_rows ? chost.distinct_rows[rk] : 0);
            const double ub = 8.0 * (a + b + static_cast<double>(chost.contrib_rows[rk]));
            if (k - 1 < t.union_slots.size()) s->add_op_bytes(t.union_slots[k - 1], ub);
            algo_bytes += ub;
        }
        // the anti-join reads the contributions and writes NEWT twice (DELTA
        // and FULL's new run)
        if (t.anti_slot >= 0 || t.contrib_bound) {
            const double extra = 8.0 * static_cast<double>(C) + 16.0 * static_cast<double>(chost.newt_rows[i]);
            s->add_op_bytes(t.anti_slot, extra);
            if (t.contrib_bound) algo_bytes += extra;
        }
    }
    // ---- TargetRuntime::apply (generic.rs:360-397) ----
    // FULL ∪ NEWT: NEWT is disjoint from every run of FULL, so it becomes
    // FULL's newest run; runs are merged (pure merges of disjoint sorted
    // arrays, output size known on the host) while the older run is less
    // than merge_ratio times the newer one.
    for (uint32_t i = 0; i < n_relations; ++i) {
        TargetRuntime& t = targets[i];
        RelationState& ts = *find(s, t.target);
        const uint64_t N = chost.newt_rows[i];
        std::swap(ts.newt.keys, t.newt);
        ts.newt.rows = N;
        if (N) {
            DevRelation run;
            run.rows = N;
            if (t.csr_run) {
                // first-column index and second column, both written with NEWT
                run.csr = true;
                run.sec = std::move(t.new_sec);
                run.xidx = std::move(t.new_xidx);
                run.xidx_set(n_x);
            } else {
                run.keys = std::move(t.new_run);
                const bool with_sec = t.new_sec.ptr != nullptr;
                if (with_sec) {
                    run.sec = std::move(t.new_sec);
                    run.sec_set();
                }
                // indexed while the anti-join uses the index (probing
                // sparsely or by columns), so merged runs get their index by
                // addition
                if ((use_x || with_sec || (vmax < (1u << 24) && N <= s->sparse_probe_rows)) && N <= 0xFFFFFFFFull)
                    ensure_xidx(s, run, n_x);
            }
            ts.runs.push_back(std::move(run));
        } else {
            s->give(std::move(t.new_run));
            s->give_sec(std::move(t.new_sec));
            s->give_x(std::move(t.new_xidx));
        }
        // merges wait for the next step unless FULL holds too many runs
        if (ts.runs.size() > kMaxRuns)
            algo_bytes += settle_runs(s, ts, false);
        else
            ts.settle_pending = !ts.runs.empty();
        std::swap(ts.delta, ts.newt);
        ts.version += 1;
        total_new += ts.delta.rows;
    }
    if (ss.instance_bits) {
        for (uint32_t i = 0; i < n_relations; ++i) {
            const RelationState& ts = *find(s, relations[i]);
            if (ts.delta.rows && !targets[i].instances_marked) {
                ss.stat_live = false;  // this NEWT was not counted
                track_instances_kernel<<<grid_for(ts.delta.rows), kBlock, 0, s->stream>>>(
                    ts.delta.keys.ptr, ts.delta.rows, ss.instance_bits, ss.instances,
                    static_cast<uint32_t>(iterations + 1), ss.last_new.ptr);
            }
        }
        check(hipGetLastError());
    }
    ++iterations;
}

// ---- graph-replayed iterations ----

__global__ void graph_begin_kernel(uint64_t* w) {
    w[kGwFlags] = 0;
    w[kGwError] = 0;
    w[kGwNext] = w[kGwDone] + 1;
}

// M beyond the candidate buffers: flag it and run the rest on no rows.
__global__ void graph_check_join_kernel(uint64_t* w, uint64_t cap) {
    const uint64_t m = w[kGwJoin];
    if (m > cap) {
        w[kGwFlags] |= 1;
        w[kGwJoinUsed] = 0;
    } else {
        w[kGwJoinUsed] = m;
    }
}

// FULL + U beyond the FULL buffers (or an earlier overflow): no rows merge.
__global__ void graph_check_full_kernel(uint64_t* w, int slot, uint64_t cap) {
    const uint64_t f = w[kGwFull0 + slot], u = w[kGwUnique];
    if (w[kGwFlags] & 1) {
        w[kGwUniqueUsed] = 0;
    } else if (f + u > cap) {
        w[kGwFlags] |= 2;
        w[kGwUniqueUsed] = 0;
    } else {
        w[kGwUniqueUsed] = u;
    }
}

__global__ void graph_end_kernel(uint64_t* w) {
    if (w[kGwFlags] == 0) w[kGwDone] = w[kGwNext];
}

__global__ void graph_set_rows_kernel(uint64_t* w, int slot, uint64_t full_rows, uint64_t delta_rows,
                                      uint64_t done) {
    w[kGwFull0 + slot] = full_rows;
    w[kGwDelta0 + slot] = delta_rows;
    w[kGwDone] = done;
}

inline uint64_t pow2_at_least(uint64_t n) {
    uint64_t c = 1;
    while (c < n) c <<= 1;
    return c;
}

// A buffer with at least `cap` entries holding the first `rows` of `b`.
void grow_keep(sparkalog_hip_store* s, DevBuf<uint64_t>& b, uint64_t cap, uint64_t rows) {
    if (b.ptr && b.cap >= cap) return;
    DevBuf<uint64_t> nb = s->take(cap);
    if (rows) check(hipMemcpyAsync(nb.ptr, b.ptr, rows * sizeof(uint64_t), hipMemcpyDeviceToDevice, s->stream));
    s->give(std::move(b));
    b = std::move(nb);
}

// One semi-naive iteration (generic.rs:481-512) for a single rule over a
// static right relation, from a captured HIP graph: join count, scan,
// output-balanced emit, global sort + unique of the candidates, and one merge
// giving FULL' = FULL ∪ U and NEWT = U \ FULL; counts stay on the device and
// the host synchronizes once, to read them. An overflow of the candidate or
// FULL buffers (flagged on the device, with no rows written) grows them, and
// the iteration runs again; nothing of the current FULL or DELTA changes
// until an iteration succeeds, since the graph writes the other buffers.
void graph_step(sparkalog_hip_store* s) {
    Session& ss = session_of(s);
    GraphRun& g = s->graph;
    RuleRuntime& rt = s->ws_rules[0];
    const sparkalog_hip_rule& plan = rt.plan;
    RelationState& ts = *find(s, ss.relations[0]);
    RelationState& rs = *find(s, plan.right_input);
    build_index(s, rt, rs, plan.right_input);
    if (!ts.runs.empty() || ts.full.csr) settle_all(s, ts);  // FULL as one packed array
    ts.settle_pending = false;
    uint32_t vmax = 0;
    for (const auto& st : s->rel)
        if (st.present) vmax = st.max_value > vmax ? st.max_value : vmax;
    const uint64_t vmask = bit_mask_for(vmax);
    uint32_t vbits = 0;
    while ((1ull << vbits) <= vmask && vbits < 32) ++vbits;
    if (vbits == 0) vbits = 1;
    const uint32_t shift = vbits <= 24 ? vbits : 32;
    const uint64_t known = shift == 32 ? ((vmask << 32) | vmask) : ((vmask << shift) | vmask);
    const uint64_t L = ts.delta.rows, F = ts.full.rows;
    if (F + 2 > 0xFFFFFFFFull) throw StatusError{SPARKALOG_HIP_ERR_TOO_MANY_ROWS};
    g.words.reserve(kGwWords);
    if (!g.host) {
        std::lock_guard<std::recursive_mutex> lock(device_mutex());
        void* p = nullptr;
        check(hipHostMalloc(&p, 2 * kGwWords * sizeof(uint64_t), hipHostMallocDefault));
        g.host = static_cast<uint64_t*>(p);
    }
    rt.narrow = static_cast<uint64_t>(rt.index.max_degree) * kCountTile < (1ull << 32);
    if (!g.stream) {
        std::lock_guard<std::recursive_mutex> lock(device_mutex());
        check(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
        check(hipEventCreateWithFlags(&g.ready, hipEventDisableTiming));
    }
    const hipStream_t gs = g.stream;
    for (int attempt = 0;; ++attempt) {
        if (attempt > 4) throw StatusError{SPARKALOG_HIP_ERR_DEVICE};
        // a FULL past graph_max_rows is rewritten every iteration at a cost the
        // regular executor's size-tiered runs avoid: the rest of the fixpoint
        // runs there (FULL and DELTA are single sorted arrays, as it expects)
        if (std::max(g.cap_full, pow2_at_least(2 * F + (1ull << 16))) > s->graph_max_rows) {
            ss.graph = false;
            session_step(s);
            return;
        }
        // capacities (powers of two, never shrinking within a store)
        g.cap_full = std::max(g.cap_full, pow2_at_least(2 * F + (1ull << 16)));
        g.cap_cand = std::max(g.cap_cand, pow2_at_least(4 * L + (1ull << 14)));
        if (g.cap_full > 0xFFFFFFFFull || g.cap_cand > 0xFFFFFFFFull) throw StatusError{SPARKALOG_HIP_ERR_TOO_MANY_ROWS};
        g.cap_delta = std::max(g.cap_cand, L ? L : 1);
        grow_keep(s, ts.full.keys, g.cap_full, F);
        grow_keep(s, ts.delta.keys, g.cap_delta, L);
        ts.full.touched();
        ts.delta.touched();
        g.full_next.reserve(g.cap_full);
        g.delta_next.reserve(g.cap_delta);
        g.cand_a.reserve(g.cap_cand);
        g.cand_b.reserve(g.cap_cand);
        g.uniq.reserve(g.cap_cand);
        g.offsets.reserve(g.cap_delta);
        g.match.reserve(g.cap_delta);
        const size_t count_tiles = tiles(g.cap_delta, kCountTile);
        g.tile_base.reserve(count_tiles);
        const size_t emit_tiles = tiles(g.cap_cand, kEmitTile);
        g.row_start.reserve(emit_tiles + 1);
        const size_t sort_bytes = distinct_temp_bytes(g.cap_cand);
        g.sort_temp.reserve(sort_bytes);
        const size_t merge_bytes = merge_new_temp_bytes(g.cap_full, g.cap_cand);
        g.merge_temp.reserve(merge_bytes);
        const bool tracking = ss.instance_bits && ss.instances;
        // rows the device does not know yet (first step, or relations changed)
        if (g.seen_full != ts.full.keys.ptr || g.seen_delta != ts.delta.keys.ptr || g.seen_full_rows != F ||
            g.seen_delta_rows != L) {
            graph_set_rows_kernel<<<1, 1, 0, s->stream>>>(g.words.ptr, g.slot, F, L, ss.iterations);
            check(hipGetLastError());
            g.seen_full = ts.full.keys.ptr;
            g.seen_delta = ts.delta.keys.ptr;
            g.seen_full_rows = F;
            g.seen_delta_rows = L;
        }
        GraphKey key{};
        const void* ptrs[] = {ts.full.keys.ptr, g.full_next.ptr, ts.delta.keys.ptr, g.delta_next.ptr, g.cand_a.ptr,
                              g.cand_b.ptr, g.uniq.ptr, g.offsets.ptr, g.tile_base.ptr, g.match.ptr, g.row_start.ptr,
                              g.sort_temp.ptr, g.merge_temp.ptr, g.words.ptr, rt.index.starts.ptr,
                              rt.index.table.ptr, rt.index.other.ptr, tracking ? ss.inst_delta.ptr : nullptr,
                              tracking ? ss.last_new.ptr : nullptr, tracking ? ss.inst_delta_host : nullptr};
        static_assert(sizeof(ptrs) <= sizeof(key.ptr), "graph key holds every pointer");
        std::memcpy(key.ptr, ptrs, sizeof(ptrs));
        const uint64_t vals[] = {g.cap_full, g.cap_cand, g.cap_delta, static_cast<uint64_t>(g.slot), shift,
                                 static_cast<uint64_t>(rt.narrow) | (static_cast<uint64_t>(rt.index.hashed) << 1),
                                 rt.index.bits | (static_cast<uint64_t>(rt.index.n_keys) << 8), known,
                                 tracking ? (static_cast<uint64_t>(ss.instance_bits) << 32 | ss.instances) : 0,
                                 static_cast<uint64_t>(plan.left_key) | (plan.right_key << 2) |
                                     (plan.out_input[0] << 4) | (plan.out_input[1] << 6) | (plan.out_column[0] << 8) |
                                     (plan.out_column[1] << 10) | (plan.distinct_columns[0] << 12) |
                                     (plan.distinct_columns[1] << 14)};
        static_assert(sizeof(vals) <= sizeof(key.val), "graph key holds every launch value");
        std::memcpy(key.val, vals, sizeof(vals));
        int which = -1;
        for (int k = 0; k < 2; ++k)
            if (g.exec[k] && g.key[k] == key) which = k;
        if (which < 0) {
            which = g.next_evict;
            g.next_evict ^= 1;
            const auto capture_t0 = std::chrono::steady_clock::now();
            // capture one iteration's launches (no allocation, no memset,
            // no synchronization inside)
            const int cur = g.slot, nxt = g.slot ^ 1;
            uint64_t* w = g.words.ptr;
            std::lock_guard<std::recursive_mutex> capture_lock(device_mutex());
            check(hipStreamBeginCapture(gs, hipStreamCaptureModeThreadLocal));
            try {
                ErrorSink sink(reinterpret_cast<uint32_t*>(&w[kGwError]));
                graph_begin_kernel<<<1, 1, 0, gs>>>(w);
                const KeyIndex kix = rt.index.view();
                const unsigned count_grid = static_cast<unsigned>(count_tiles);
                uint32_t* local32 = reinterpret_cast<uint32_t*>(g.offsets.ptr);
                const uint64_t* n_dev = &w[kGwDelta0 + cur];
                if (rt.narrow && !rt.index.hashed)
                    dense_join_count_kernel<uint32_t, false><<<count_grid, kBlock, 0, gs>>>(
                        ts.delta.keys.ptr, g.cap_delta, plan.left_key, kix, local32, g.match.ptr, g.tile_base.ptr,
                        nullptr, 0, n_dev);
                else if (rt.narrow)
                    dense_join_count_kernel<uint32_t, true><<<count_grid, kBlock, 0, gs>>>(
                        ts.delta.keys.ptr, g.cap_delta, plan.left_key, kix, local32, g.match.ptr, g.tile_base.ptr,
                        nullptr, 0, n_dev);
                else if (!rt.index.hashed)
                    dense_join_count_kernel<uint64_t, false><<<count_grid, kBlock, 0, gs>>>(
                        ts.delta.keys.ptr, g.cap_delta, plan.left_key, kix, g.offsets.ptr, g.match.ptr,
                        g.tile_base.ptr, nullptr, 0, n_dev);
                else
                    dense_join_count_kernel<uint64_t, true><<<count_grid, kBlock, 0, gs>>>(
                        ts.delta.keys.ptr, g.cap_delta, plan.left_key, kix, g.offsets.ptr, g.match.ptr,
                        g.tile_base.ptr, nullptr, 0, n_dev);
                join_scan_kernel<<<1, kScanThreads, 0, gs>>>(g.tile_base.ptr, count_tiles, &w[kGwJoin]);
                graph_check_join_kernel<<<1, 1, 0, gs>>>(w, g.cap_cand);
                EmitSpec spec;
                spec.key_col = plan.left_key;
                spec.right_key = plan.right_key;
                spec.side[0] = plan.out_input[0];
                spec.side[1] = plan.out_input[1];
                spec.col[0] = plan.out_column[0];
                spec.col[1] = plan.out_column[1];
                spec.order[0] = plan.distinct_columns[0];
                spec.order[1] = plan.distinct_columns[1];
                spec.shift = shift;
                const unsigned emit_grid = static_cast<unsigned>(emit_tiles < 65536 ? emit_tiles : 65536);
                const uint64_t* m_dev = &w[kGwJoinUsed];
                if (rt.narrow) {
                    emit_partition_kernel<uint32_t><<<grid_for(emit_tiles + 1), kBlock, 0, gs>>>(
                        local32, g.tile_base.ptr, g.cap_delta, g.cap_cand, g.row_start.ptr, emit_tiles, n_dev, m_dev);
                    dense_join_emit_kernel<uint32_t><<<emit_grid, kBlock, 0, gs>>>(
                        ts.delta.keys.ptr, g.cap_delta, rt.index.other.ptr, local32, g.tile_base.ptr, g.match.ptr,
                        g.cap_cand, g.row_start.ptr, spec, g.cand_a.ptr, n_dev, m_dev);
                } else {
                    emit_partition_kernel<uint64_t><<<grid_for(emit_tiles + 1), kBlock, 0, gs>>>(
                        g.offsets.ptr, g.tile_base.ptr, g.cap_delta, g.cap_cand, g.row_start.ptr, emit_tiles, n_dev,
                        m_dev);
                    dense_join_emit_kernel<uint64_t><<<emit_grid, kBlock, 0, gs>>>(
                        ts.delta.keys.ptr, g.cap_delta, rt.index.other.ptr, g.offsets.ptr, g.tile_base.ptr,
                        g.match.ptr, g.cap_cand, g.row_start.ptr, spec, g.cand_a.ptr, n_dev, m_dev);
                }
                PairInput in{g.cand_a.ptr, nullptr, nullptr, m_dev};
                PairOutput out{g.uniq.ptr, nullptr, nullptr};
                check(distinct(in, g.cap_cand, g.cand_a.ptr, g.cand_b.ptr, out, &w[kGwUnique], g.sort_temp.ptr,
                               sort_bytes, known, gs));
                if (shift != 32)
                    expand_keys_kernel<<<grid_for(g.cap_cand), kBlock, 0, gs>>>(g.uniq.ptr, &w[kGwUnique],
                                                                                       g.cap_cand, shift);
                graph_check_full_kernel<<<1, 1, 0, gs>>>(w, cur, g.cap_full);
                check(merge_new(ts.full.keys.ptr, &w[kGwFull0 + cur], g.cap_full, g.uniq.ptr, &w[kGwUniqueUsed],
                                g.cap_cand, g.full_next.ptr, &w[kGwFull0 + nxt], g.delta_next.ptr,
                                &w[kGwDelta0 + nxt], g.merge_temp.ptr, merge_bytes, gs));
                if (tracking) {
                    clear_words_kernel<<<1, 1024, 0, gs>>>(reinterpret_cast<uint32_t*>(ss.inst_delta.ptr),
                                                                  2 * static_cast<size_t>(ss.instances));
                    instance_rows_kernel<<<grid_for(ss.instances), kBlock, 0, gs>>>(
                        g.delta_next.ptr, &w[kGwDelta0 + nxt], g.cap_delta, ss.instance_bits, ss.instances,
                        reinterpret_cast<unsigned long long*>(ss.inst_delta.ptr));
                    track_instances_kernel<<<grid_for(g.cap_delta), kBlock, 0, gs>>>(
                        g.delta_next.ptr, g.cap_delta, ss.instance_bits, ss.instances, 0, ss.last_new.ptr,
                        &w[kGwDelta0 + nxt], &w[kGwNext]);
                    check(hipMemcpyAsync(ss.inst_delta_host, ss.inst_delta.ptr, sizeof(uint64_t) * ss.instances,
                                         hipMemcpyDeviceToHost, gs));
                }
                graph_end_kernel<<<1, 1, 0, gs>>>(w);
                check(hipGetLastError());
                check(hipMemcpyAsync(g.host + kGwWords * cur, w, kGwWords * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                     gs));
            } catch (...) {
                hipGraph_t broken = nullptr;
                (void)hipStreamEndCapture(gs, &broken);
                if (broken) (void)hipGraphDestroy(broken);
                (void)hipGetLastError();
                throw;
            }
            hipGraph_t graph = nullptr;
            check(hipStreamEndCapture(gs, &graph));
            // a graph of the same launches with other pointers or sizes (a new
            // layout, a new fixpoint) updates the slot's executable graph in
            // place; one that differs in its launches is instantiated anew
            bool updated = false;
            if (g.exec[which]) {
                hipGraphNode_t error_node = nullptr;
                hipGraphExecUpdateResult result = hipGraphExecUpdateError;
                const hipError_t ue = hipGraphExecUpdate(g.exec[which], graph, &error_node, &result);
                updated = ue == hipSuccess && result == hipGraphExecUpdateSuccess;
                if (!updated) {
                    (void)hipGetLastError();
                    (void)hipGraphExecDestroy(g.exec[which]);
                    g.exec[which] = nullptr;
                }
            }
            hipError_t inst = hipSuccess;
            if (!updated) inst = hipGraphInstantiate(&g.exec[which], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (inst != hipSuccess) g.exec[which] = nullptr;
            check(inst);
            g.key[which] = key;
            ++g.captures;
            if (updated) ++g.updates;
            g.capture_ms +=
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - capture_t0).count();
        } else {
            ++g.replays;
        }
        // after everything queued on the store's stream (uploads, the index,
        // the rows just set)
        check(hipEventRecord(g.ready, s->stream));
        check(hipStreamWaitEvent(gs, g.ready, 0));
        check(hipGraphLaunch(g.exec[which], gs));
        check(hipStreamSynchronize(gs));
        const uint64_t* h = g.host + kGwWords * g.slot;
        if (static_cast<uint32_t>(h[kGwError])) throw StatusError{SPARKALOG_HIP_ERR_DEVICE};
        const uint64_t flags = h[kGwFlags];
        if (flags) {
            // grow to what the device counted, and run the iteration again
            // (in the regular executor when FULL would pass graph_max_rows)
            if (flags & 1) g.cap_cand = std::max(g.cap_cand, pow2_at_least(h[kGwJoin] + h[kGwJoin] / 4 + 1));
            if (flags & 2) {
                const uint64_t need = pow2_at_least(F + h[kGwUnique] + (F + h[kGwUnique]) / 4 + 1);
                if (need > s->graph_max_rows) {
                    ss.graph = false;
                    session_step(s);
                    return;
                }
                g.cap_full = std::max(g.cap_full, need);
            }
            continue;
        }
        const int nxt = g.slot ^ 1;
        const uint64_t M = h[kGwJoin], U = h[kGwUnique];
        const uint64_t F2 = h[kGwFull0 + nxt], N = h[kGwDelta0 + nxt];
        if (F2 != F + N) throw StatusError{SPARKALOG_HIP_ERR_DEVICE};
        std::swap(ts.full.keys, g.full_next);
        std::swap(ts.delta.keys, g.delta_next);
        ts.full.rows = F2;
        ts.delta.rows = N;
        ts.newt.rows = 0;
        ts.full.touched();
        ts.delta.touched();
        ts.version += 1;
        g.slot = nxt;
        g.seen_full = ts.full.keys.ptr;
        g.seen_delta = ts.delta.keys.ptr;
        g.seen_full_rows = F2;
        g.seen_delta_rows = N;
        ss.total_new += N;
        ss.inst_delta_valid = tracking;
        const double R = static_cast<double>(rt.index.rows);
        ss.algo_bytes += 8.0 * L + 8.0 * R + 16.0 * M + 8.0 * U + 8.0 * (F + U) + 8.0 * (F2 + N);
        break;
    }
    ++ss.iterations;
}

int session_end(sparkalog_hip_store* s, sparkalog_hip_summary* summary, int rc) {
    Session& ss = session_of(s);
    // FULL leaves the run as one sorted array
    for (uint32_t r : ss.relations) ss.algo_bytes += settle_runs(s, *find(s, r), true);
    if (ss.stat_live && rc == SPARKALOG_HIP_OK && ss.relations.size() == 1) {
        ss.stat_valid = true;
        ss.stat_rel = ss.relations[0];
        ss.stat_version = find(s, ss.stat_rel)->version;
        ss.stat_bits = ss.instance_bits;
        ss.stat_instances = ss.instances;
    }
    ss.stat_live = false;
    check(hipEventRecord(ss.t_end, s->stream));
    check(hipEventSynchronize(ss.t_end));
    float ms = 0.f;
    check(hipEventElapsedTime(&ms, ss.t_begin, ss.t_end));
    if (s->profiling) s->collect_ops();
    if (summary) {
        summary->iterations = ss.iterations;
        summary->total_new_rows = ss.total_new;
        summary->device_ms = ms;
        summary->algorithmic_bytes = ss.algo_bytes;
        if (rc == SPARKALOG_HIP_OK) summary->delta_rows_left = 0;
    }
    ss.active = false;
    return rc;
}

}  // namespace
}  // namespace spk

extern "C" int sparkalog_hip_executor_run(sparkalog_hip_store* s, const uint32_t* relations, uint32_t n_relations,
                                          const sparkalog_hip_rule* rules, uint32_t n_rules, uint64_t max_iterations,
                                          sparkalog_hip_summary* summary) {
    if (!s || (n_relations && !relations) || (n_rules && !rules)) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    SPK_GUARD_BEGIN
    const int begun = session_begin(s, relations, n_relations, rules, n_rules);
    if (begun != SPARKALOG_HIP_OK) return begun;
    int rc = SPARKALOG_HIP_OK;
    for (;;) {
        const uint64_t pending = pending_delta(s);
        if (pending == 0) break;
        if (session_of(s).iterations >= max_iterations) {
            if (summary) summary->delta_rows_left = pending;
            rc = SPARKALOG_HIP_ERR_ITERATION_LIMIT;
            break;
        }
        session_step(s);
    }
    return session_end(s, summary, rc);
    SPK_GUARD_END
}

extern "C" int sparkalog_hip_session_begin(sparkalog_hip_store* s, const uint32_t* relations, uint32_t n_relations,
                                           const sparkalog_hip_rule* rules, uint32_t n_rules) {
    if (!s || (n_relations && !relations) || (n_rules && !rules)) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    SPK_GUARD_BEGIN
    return session_begin(s, relations, n_relations, rules, n_rules);
    SPK_GUARD_END
}

extern "C" int sparkalog_hip_session_step(sparkalog_hip_store* s, uint64_t* delta_rows) {
    if (!s || !session_of(s).active) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    SPK_GUARD_BEGIN
    if (pending_delta(s)) session_step(s);
    if (delta_rows) *delta_rows = pending_delta(s);
    return SPARKALOG_HIP_OK;
    SPK_GUARD_END
}

extern "C" int sparkalog_hip_session_end(sparkalog_hip_store* s, sparkalog_hip_summary* summary) {
    if (!s || !session_of(s).active) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    SPK_GUARD_BEGIN
    return session_end(s, summary, SPARKALOG_HIP_OK);
    SPK_GUARD_END
}

extern "C" int sparkalog_hip_store_instance_stats(sparkalog_hip_store* s, uint32_t id, uint32_t instance_bits,
                                                  uint32_t instances, uint64_t* rows, uint64_t* digests) {
    if (!s || (instances && (!rows || !digests))) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    RelationState* st = find(s, id);
    if (!st) return SPARKALOG_HIP_ERR_MISSING_RELATION;
    SPK_GUARD_BEGIN
    Session& ss = session_of(s);
    if (ss.stat_valid && !ss.active && id == ss.stat_rel && st->version == ss.stat_version &&
        instance_bits == ss.stat_bits && instances == ss.stat_instances) {
        // accumulated by the last session, and the relation is unchanged:
        // the sum of each instance's copies
        const size_t n_acc = 2 * static_cast<size_t>(kInstanceStatSlots) * instances;
        std::vector<uint64_t> copies(n_acc);
        if (n_acc)
            check(hipMemcpyAsync(copies.data(), ss.stat_acc.ptr, n_acc * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                 s->stream));
        check(hipStreamSynchronize(s->stream));
        for (uint32_t i = 0; i < instances; ++i) {
            uint64_t r = 0, d = 0;
            for (uint32_t j = 0; j < kInstanceStatSlots; ++j) {
                r += copies[static_cast<size_t>(j) * instances + i];
                d += copies[(static_cast<size_t>(kInstanceStatSlots) + j) * instances + i];
            }
            rows[i] = r;
            digests[i] = d;
        }
        return SPARKALOG_HIP_OK;
    }
    settle_all(s, *st);
    unsigned long long* acc =
        reinterpret_cast<unsigned long long*>(s->probes(2 * static_cast<size_t>(instances) + 2));
    check(fill_u64(reinterpret_cast<uint64_t*>(acc), 2 * static_cast<size_t>(instances), 0, s->stream));
    if (st->full.rows && instances) {
        const size_t waves = (st->full.rows + kStatsWaveRows - 1) / kStatsWaveRows;
        const size_t blocks = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
        instance_stats_kernel<<<static_cast<unsigned>(blocks), kBlock, 0, s->stream>>>(
            st->full.keys.ptr, st->full.rows, instance_bits, instances, acc, acc + instances);
    }
    check(hipGetLastError());
    if (instances) {
        check(hipMemcpyAsync(rows, acc, instances * sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
        check(hipMemcpyAsync(digests, acc + instances, instances * sizeof(uint64_t), hipMemcpyDeviceToHost,
                             s->stream));
    }
    check(hipStreamSynchronize(s->stream));
    return SPARKALOG_HIP_OK;
    SPK_GUARD_END
}

extern "C" int sparkalog_hip_store_range_stats(sparkalog_hip_store* s, uint32_t id, int which, uint32_t first_lo,
                                               uint32_t first_hi, uint32_t base, uint64_t* rows, uint64_t* digest) {
    if (!s) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    RelationState* st = find(s, id);
    if (!st) return SPARKALOG_HIP_ERR_MISSING_RELATION;
    if (which == 1 && !st->recursive) return SPARKALOG_HIP_ERR_NOT_RECURSIVE;
    SPK_GUARD_BEGIN
    std::vector<const DevRelation*> parts;
    if (which == 1) {
        parts.push_back(&st->delta);
    } else {
        packed_parts(s, *st);
        parts.push_back(&st->full);
        for (const auto& r : st->runs) parts.push_back(&r);
    }
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(s->small16());
    check(fill_u64(reinterpret_cast<uint64_t*>(acc), 2, 0, s->stream));
    for (const DevRelation* r : parts) {
        const auto range = first_range(s, *r, first_lo, first_hi);
        if (range.second > range.first)
            range_stats_kernel<<<grid_for(range.second - range.first), kBlock, 0, s->stream>>>(
                r->keys.ptr, range.first, range.second, base, acc);
    }
    check(hipGetLastError());
    unsigned long long host[2];
    check(hipMemcpyAsync(host, acc, sizeof(host), hipMemcpyDeviceToHost, s->stream));
    check(hipStreamSynchronize(s->stream));
    if (rows) *rows = host[0];
    if (digest) *digest = host[1];
    return SPARKALOG_HIP_OK;
    SPK_GUARD_END
}

extern "C" int sparkalog_hip_store_merge_rows(sparkalog_hip_store* s, uint32_t id, int which, const uint32_t* c0,
                                              const uint32_t* c1, uint64_t rows, int on_device) {
    if (!s || (rows && (!c0 || !c1))) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    RelationState* st = find(s, id);
    if (!st) return SPARKALOG_HIP_ERR_MISSING_RELATION;
    if (which == 1 && !st->recursive) return SPARKALOG_HIP_ERR_NOT_RECURSIVE;
    SPK_GUARD_BEGIN
    if (rows == 0) return SPARKALOG_HIP_OK;
    DevRelation incoming;
    const uint32_t mv = upload(s, incoming, c0, c1, rows, on_device);
    st->max_value = mv > st->max_value ? mv : st->max_value;
    if (which == 1) {
        union_into(s, st->delta, incoming.keys.ptr, incoming.rows);
    } else if (st->recursive) {
        // rows disjoint from FULL become a run of their own
        st->runs.push_back(std::move(incoming));
        st->version += 1;
    } else {
        ensure_packed(s, st->full);
        union_into(s, st->full, incoming.keys.ptr, incoming.rows);
        st->version += 1;
    }
    return SPARKALOG_HIP_OK;
    SPK_GUARD_END
}

extern "C" int sparkalog_hip_store_remove_ranges(sparkalog_hip_store* s, uint32_t id, const uint32_t* first_bounds,
                                                 uint32_t ranges) {
    if (!s || (ranges && !first_bounds)) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    for (uint32_t j = 0; j < ranges; ++j) {
        if (first_bounds[2 * j] > first_bounds[2 * j + 1]) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
        if (j && first_bounds[2 * j] < first_bounds[2 * j - 1]) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    }
    RelationState* st = find(s, id);
    if (!st) return SPARKALOG_HIP_ERR_MISSING_RELATION;
    SPK_GUARD_BEGIN
    if (ranges == 0) return SPARKALOG_HIP_OK;
    std::vector<uint64_t> probes(2 * static_cast<size_t>(ranges));
    for (uint32_t j = 0; j < 2 * ranges; ++j) probes[j] = static_cast<uint64_t>(first_bounds[j]) << 32;
    uint64_t* dev = s->probes(4 * static_cast<size_t>(ranges));
    check(hipMemcpyAsync(dev, probes.data(), probes.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s->stream));
    std::vector<uint64_t> pos(probes.size());
    // one pass per part: the rows outside every range are copied, in order,
    // into a new buffer
    auto strip = [&](DevRelation& r) {
        if (r.rows == 0) return;
        lower_bounds_kernel<<<grid_for(probes.size()), kBlock, 0, s->stream>>>(r.keys.ptr, r.rows, dev, probes.size(),
                                                                             dev + probes.size());
        check(hipGetLastError());
        check(hipMemcpyAsync(pos.data(), dev + probes.size(), pos.size() * sizeof(uint64_t), hipMemcpyDeviceToHost,
                             s->stream));
        check(hipStreamSynchronize(s->stream));
        uint64_t removed = 0;
        for (uint32_t j = 0; j < ranges; ++j) removed += pos[2 * j + 1] > pos[2 * j] ? pos[2 * j + 1] - pos[2 * j] : 0;
        if (removed == 0) return;
        const uint64_t n = r.rows - removed;
        DevBuf<uint64_t> out = s->take(n ? n : 1);
        uint64_t from = 0, to = 0;
        for (uint32_t j = 0; j <= ranges; ++j) {
            const uint64_t end = j < ranges ? pos[2 * j] : r.rows;
            if (end > from) {
                check(hipMemcpyAsync(out.ptr + to, r.keys.ptr + from, (end - from) * sizeof(uint64_t),
                                     hipMemcpyDeviceToDevice, s->stream));
                to += end - from;
            }
            if (j < ranges) from = pos[2 * j + 1] > from ? pos[2 * j + 1] : from;
        }
        s->give(std::move(r.keys));
        r.keys = std::move(out);
        r.rows = n;
        r.touched();
    };
    packed_parts(s, *st);
    strip(st->full);
    for (auto& r : st->runs) strip(r);
    if (st->recursive) strip(st->delta);
    st->version += 1;
    return SPARKALOG_HIP_OK;
    SPK_GUARD_END
}

extern "C" int sparkalog_hip_store_remove_range(sparkalog_hip_store* s, uint32_t id, uint32_t first_lo,
                                                uint32_t first_hi) {
    if (!s) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    RelationState* st = find(s, id);
    if (!st) return SPARKALOG_HIP_ERR_MISSING_RELATION;
    SPK_GUARD_BEGIN
    auto strip = [&](DevRelation& r) {
        const auto range = first_range(s, r, first_lo, first_hi);
        remove_rows(s, r, range.first, range.second);
    };
    packed_parts(s, *st);
    strip(st->full);
    for (auto& r : st->runs) strip(r);
    if (st->recursive) strip(st->delta);
    st->version += 1;
    return SPARKALOG_HIP_OK;
    SPK_GUARD_END
}

extern "C" int sparkalog_hip_store_range_rows(sparkalog_hip_store* s, uint32_t id, int which,
                                              const uint32_t* first_bounds, uint32_t ranges, uint64_t* rows) {
    if (!s || (ranges && (!first_bounds || !rows))) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    RelationState* st = find(s, id);
    if (!st) return SPARKALOG_HIP_ERR_MISSING_RELATION;
    if (which == 1 && !st->recursive) return SPARKALOG_HIP_ERR_NOT_RECURSIVE;
    SPK_GUARD_BEGIN
    for (uint32_t j = 0; j < ranges; ++j) rows[j] = 0;
    if (ranges == 0) return SPARKALOG_HIP_OK;
    std::vector<const DevRelation*> parts;
    if (which == 1) {
        parts.push_back(&st->delta);
    } else {
        packed_parts(s, *st);
        parts.push_back(&st->full);
        for (const auto& r : st->runs) parts.push_back(&r);
    }
    std::vector<uint64_t> probes(2 * static_cast<size_t>(ranges));
    for (uint32_t j = 0; j < 2 * ranges; ++j) probes[j] = static_cast<uint64_t>(first_bounds[j]) << 32;
    uint64_t* dev = s->probes(4 * static_cast<size_t>(ranges));
    check(hipMemcpyAsync(dev, probes.data(), probes.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s->stream));
    std::vector<uint64_t> pos(probes.size());
    for (const DevRelation* r : parts) {
        if (r->rows == 0) continue;
        lower_bounds_kernel<<<grid_for(probes.size()), kBlock, 0, s->stream>>>(r->keys.ptr, r->rows, dev,
                                                                             probes.size(), dev + probes.size());
        check(hipGetLastError());
        check(hipMemcpyAsync(pos.data(), dev + probes.size(), pos.size() * sizeof(uint64_t),
                             hipMemcpyDeviceToHost, s->stream));
        check(hipStreamSynchronize(s->stream));
        for (uint32_t j = 0; j < ranges; ++j)
            if (pos[2 * j + 1] > pos[2 * j]) rows[j] += pos[2 * j + 1] - pos[2 * j];
    }
    return SPARKALOG_HIP_OK;
    SPK_GUARD_END
}

extern "C" int sparkalog_hip_set_instance_tracking(sparkalog_hip_store* s, uint32_t instance_bits,
                                                   uint32_t instances) {
    if (!s || instance_bits >= 32) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    Session& ss = session_of(s);
    ss.instance_bits = instances ? instance_bits : 0;
    ss.instances = instances;
    return SPARKALOG_HIP_OK;
}

extern "C" int sparkalog_hip_instance_delta_rows(sparkalog_hip_store* s, uint64_t* rows, uint32_t instances) {
    if (!s || (instances && !rows)) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    Session& ss = session_of(s);
    if (!ss.instance_bits || instances > ss.instances || !ss.inst_delta_valid)
        return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    // written by the last session_step, which synchronized after the copy
    std::memcpy(rows, ss.inst_delta_host, sizeof(uint64_t) * instances);
    return SPARKALOG_HIP_OK;
}

extern "C" int sparkalog_hip_instance_iterations(sparkalog_hip_store* s, uint32_t* iterations, uint32_t instances) {
    if (!s || (instances && !iterations)) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    Session& ss = session_of(s);
    if (!ss.instance_bits || instances > ss.instances) return SPARKALOG_HIP_ERR_INVALID_ARGUMENT;
    SPK_GUARD_BEGIN
    if (instances)
        check(hipMemcpyAsync(iterations, ss.last_new.ptr, instances * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             s->stream));
    check(hipStreamSynchronize(s->stream));
    return SPARKALOG_HIP_OK;
    SPK_GUARD_END
}
