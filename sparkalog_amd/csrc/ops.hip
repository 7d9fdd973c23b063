// gfx950 kernels of the binary operator pipeline and their launchers.
//
// Design (all integer work, HBM-bound, no MFMA):
//  * distinct  = onesweep LSD radix sort (Adinets & Merrill 2022) over the
//    varying 8-bit digits of the packed key, with stable wave64 match-ranking
//    in LDS, followed by a two-pass unique compaction.
//  * union     = merge-path partitioned merge (Green, McColl & Bader 2012)
//    with fused duplicate drop, in two passes: tiles count their kept rows,
//    then merge again and write at their offsets.
//  * anti-join = left-tile partitioning with a per-tile right range found by
//    binary search; ranges that fit LDS are staged there, larger ones are
//    searched in place, so work is O(L log(R/L)) rather than O(L+R) when
//    R >> L; then a compaction of the kept rows at offsets from their tiles'
//    counts.
//  * join      = single-pass count + exclusive scan, then an output-balanced
//    emit (every block writes an equal slice of the output, whatever the key
//    skew).
//  * filter    = ballot compaction of matching row ids, in two passes.
// Where tiles of a compaction would wait on each other's counts (decoupled
// look-back, whose status polls cross XCDs on this chip), the filter, the
// union, the anti-join and the unique compaction count first and write
// second: a few bytes more read, no tile ever waits; the sort's digit passes
// count per chunk of tiles and scan the counts. Only the join count keeps a
// single pass with look-back.
#include "ops.hpp"
#include "prims.hpp"

namespace spk {

namespace {
thread_local uint32_t* g_error_sink = nullptr;
}  // namespace

uint32_t* error_sink() { return g_error_sink; }
ErrorSink::ErrorSink(uint32_t* word) : previous_(g_error_sink) { g_error_sink = word; }
ErrorSink::~ErrorSink() { g_error_sink = previous_; }

namespace {

constexpr int kSortItems = 16;
constexpr int kSortTile = kBlock * kSortItems;       // 4096 keys
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;              // 256 digits
constexpr int kPasses = 8;                           // 64-bit keys
constexpr int kCompactItems = 16;
constexpr int kCompactTile = kBlock * kCompactItems; // 4096 rows
constexpr int kTileGroup = 64;                      // tiles per group of a two-pass compaction
constexpr int kMergeItems = 8;
constexpr int kMergeTile = kBlock * kMergeItems;     // 2048 merged rows
constexpr int kAntiTile = 2048;                      // left rows per tile
constexpr int kAntiStage = 2048;                     // right rows staged in LDS per chunk
// Above this many right rows per 2048-row left tile, reading the range costs
// more HBM traffic than ~log2(span) probes per left row.
constexpr size_t kAntiStreamLimit = 1u << 18;
constexpr int kEmitTile = 2048;                      // output rows per tile
constexpr int kEmitStage = 2048;                     // left rows staged in LDS

__host__ __device__ inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
__host__ __device__ inline size_t tiles_of(size_t n, size_t t) { return (n + t - 1) / t; }

// Scratch header shared by every launcher (offset 0 of `temp`).
struct Header {
    uint32_t error;
    uint32_t counters[15];
};

// ------------------------------------------------------------------------
// generic helpers

__global__ void clear_u64_kernel(uint64_t* p, size_t n) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

__global__ void clear_header_kernel(Header* h) {
    if (threadIdx.x < 16) reinterpret_cast<uint32_t*>(h)[threadIdx.x] = 0;
}

hipError_t clear_u64(uint64_t* p, size_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const size_t blocks = tiles_of(n, kBlock * 4);
    const unsigned grid = static_cast<unsigned>(blocks < 8192 ? blocks : 8192);
    clear_u64_kernel<<<grid, kBlock, 0, stream>>>(p, n);
    return hipGetLastError();
}

__device__ __forceinline__ size_t actual_rows(const PairInput& in, size_t bound) {
    if (!in.rows_dev) return bound;
    const uint64_t n = *in.rows_dev;
    return n < bound ? n : bound;
}

__device__ __forceinline__ uint64_t load_key(const PairInput& in, size_t i) {
    return in.packed ? in.packed[i] : pack_pair(in.first[i], in.second[i]);
}

__device__ __forceinline__ void store_key(const PairOutput& out, size_t i, uint64_t k) {
    if (out.packed) out.packed[i] = k;
    if (out.packed_copy) out.packed_copy[i] = k;
    if (out.first) {
        out.first[i] = static_cast<uint32_t>(k >> 32);
        out.second[i] = static_cast<uint32_t>(k);
    }
}

// ------------------------------------------------------------------------
// radix sort
//
// LSD radix sort over the varying 8-bit digits of the packed keys, reduce
// then scan: each digit pass is an upsweep (every chunk of tiles counts its
// keys per digit), a scan of the counts in (digit, chunk) order, and a
// downsweep in which a block sorts its chunk's tiles one after another,
// stably inside each tile in LDS, and writes every digit's keys at the
// chunk's running offset for that digit. No tile waits for another: a
// onesweep pass (per-digit decoupled look-back) ran at a third of the copy
// ceiling here, its look-back chains crossing XCDs.

struct SortDesc {
    uint64_t rows;                       // actual row count of this sort
    uint64_t or_mask;
    uint64_t and_mask;
};

__device__ __forceinline__ uint32_t varying_digits(const SortDesc* d) {
    const uint64_t v = d->or_mask ^ d->and_mask;
    uint32_t m = 0;
#pragma unroll
    for (int p = 0; p < kPasses; ++p)
        if ((v >> (p * kRadixBits)) & (kRadix - 1)) m |= 1u << p;
    return m;
}

__global__ void sort_init_kernel(SortDesc* desc, uint64_t known_varying, PairInput in, size_t bound) {
    if (threadIdx.x == 0) {
        desc->rows = actual_rows(in, bound);
        desc->or_mask = known_varying;  // 0 when unknown: OR-reduced on device
        desc->and_mask = known_varying ? 0 : ~0ull;
    }
}

// With no varying digit every key is equal and no pass runs: the keys are
// copied to keys_a, where the sorted output of zero passes lies.
__global__ void sort_constant_copy_kernel(PairInput in, uint64_t* keys_a, const SortDesc* desc) {
    if (varying_digits(desc) != 0 || in.packed == keys_a) return;
    const size_t rows = desc->rows;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < rows; i += stride)
        keys_a[i] = load_key(in, i);
}

// A pass's source and destination: the first active pass reads the input
// (packed or two columns), later ones ping-pong keys_b -> keys_a -> keys_b.
struct PassIo {
    bool active;
    bool from_input;
    const uint64_t* src;
    uint64_t* dst;
};

__device__ __forceinline__ PassIo pass_io(const SortDesc* desc, int pass, uint64_t* keys_a, uint64_t* keys_b) {
    const uint32_t varying = varying_digits(desc);
    PassIo io{};
    io.active = (varying >> pass) & 1u;
    const uint32_t ordinal = __popc(varying & ((1u << pass) - 1u));
    io.from_input = ordinal == 0;
    io.src = (ordinal & 1u) ? keys_b : keys_a;
    io.dst = (ordinal & 1u) ? keys_a : keys_b;
    return io;
}

// Upsweep of a digit pass: block c counts the digits of the keys of chunk c
// (chunk_tiles tiles of kSortTile keys) into counts[d * chunks + c]. Whole
// tiles are read as 16-byte vectors (key order does not matter to a count)
// when the columns are 16-byte aligned. With `reduce` (the first pass of a
// sort whose varying digits are unknown) it also ORs and ANDs every key into
// the descriptor, and counts whether or not its digit varies.
__device__ __forceinline__ void count_digit(uint32_t* hist, uint64_t key, int shift, bool valid,
                                            uint64_t valid_mask) {
    const uint32_t digit = static_cast<uint32_t>(key >> shift) & (kRadix - 1);
    const uint32_t lead = __builtin_amdgcn_readfirstlane(digit);
    const uint64_t same = ballot(valid && digit == lead);
    if (same == valid_mask) {
        if (lane_id() == 0) atomicAdd(&hist[lead], static_cast<uint32_t>(__popcll(valid_mask)));
    } else if (valid) {
        atomicAdd(&hist[digit], 1u);
    }
}

__global__ __launch_bounds__(kBlock) void sort_upsweep_kernel(PairInput in, uint64_t* keys_a, uint64_t* keys_b,
                                                              SortDesc* desc, int pass, uint32_t chunk_tiles,
                                                              uint32_t* counts, int reduce) {
    __shared__ uint32_t hist[kRadix];
    __shared__ uint64_t red_or[kWavesPerBlock], red_and[kWavesPerBlock];
    const PassIo io = pass_io(desc, pass, keys_a, keys_b);
    if (!reduce && !io.active) return;
    const bool from_input = reduce || io.from_input;
    const size_t rows = desc->rows;
    const uint32_t chunks = gridDim.x;
    const uint32_t c = blockIdx.x;
    const int shift = pass * kRadixBits;
    hist[threadIdx.x] = 0;
    __syncthreads();
    const size_t c0 = static_cast<size_t>(c) * chunk_tiles * kSortTile;
    const size_t c1 = c0 + static_cast<size_t>(chunk_tiles) * kSortTile < rows
                          ? c0 + static_cast<size_t>(chunk_tiles) * kSortTile
                          : rows;
    const uint64_t* packed = from_input ? in.packed : io.src;
    const bool aligned = packed ? (reinterpret_cast<uintptr_t>(packed) & 15u) == 0
                                : ((reinterpret_cast<uintptr_t>(in.first) | reinterpret_cast<uintptr_t>(in.second)) &
                                   15u) == 0;
    uint64_t acc_or = 0, acc_and = ~0ull;
    for (size_t t0 = c0; t0 < c1; t0 += kSortTile) {
        uint64_t kv[kSortItems];
        const bool whole = c1 - t0 >= static_cast<size_t>(kSortTile) && aligned;  // block-uniform
        if (whole && packed) {
            // 2,048 vectors of two keys: thread t loads vectors t + 256 q
            const uint4* v = reinterpret_cast<const uint4*>(packed + t0);
#pragma unroll
            for (int q = 0; q < kSortItems / 2; ++q) {
                const uint4 x = v[q * kBlock + threadIdx.x];
                kv[2 * q] = (static_cast<uint64_t>(x.y) << 32) | x.x;
                kv[2 * q + 1] = (static_cast<uint64_t>(x.w) << 32) | x.z;
            }
        } else if (whole) {
            // 1,024 vectors of four values per column
            const uint4* f = reinterpret_cast<const uint4*>(in.first + t0);
            const uint4* g = reinterpret_cast<const uint4*>(in.second + t0);
            uint4 xf[kSortItems / 4], xg[kSortItems / 4];
#pragma unroll
            for (int q = 0; q < kSortItems / 4; ++q) {
                xf[q] = f[q * kBlock + threadIdx.x];
                xg[q] = g[q * kBlock + threadIdx.x];
            }
#pragma unroll
            for (int q = 0; q < kSortItems / 4; ++q) {
                kv[4 * q] = pack_pair(xf[q].x, xg[q].x);
                kv[4 * q + 1] = pack_pair(xf[q].y, xg[q].y);
                kv[4 * q + 2] = pack_pair(xf[q].z, xg[q].z);
                kv[4 * q + 3] = pack_pair(xf[q].w, xg[q].w);
            }
        } else {
#pragma unroll
            for (int i = 0; i < kSortItems; ++i) {
                const size_t idx = t0 + i * kBlock + threadIdx.x;
                const size_t at = idx < c1 ? idx : c1 - 1;
                kv[i] = from_input ? load_key(in, at) : io.src[at];
            }
        }
        const uint32_t valid_rows = whole ? kSortTile : static_cast<uint32_t>(c1 - t0);
#pragma unroll
        for (int i = 0; i < kSortItems; ++i) {
            // item i of every thread: which positions it holds does not
            // matter, only that the tile's first valid_rows keys are counted
            const bool valid = whole || i * kBlock + threadIdx.x < valid_rows;
            const uint64_t valid_mask = ballot(valid);
            if (valid_mask == 0) continue;
            if (valid) {
                acc_or |= kv[i];
                acc_and &= kv[i];
            }
            count_digit(hist, kv[i], shift, valid, valid_mask);
        }
    }
    __syncthreads();
    counts[static_cast<size_t>(threadIdx.x) * chunks + c] = hist[threadIdx.x];
    if (reduce) {
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) {
            acc_or |= __shfl_xor(acc_or, off, kWave);
            acc_and &= __shfl_xor(acc_and, off, kWave);
        }
        if (lane_id() == 0) {
            red_or[wave_id()] = acc_or;
            red_and[wave_id()] = acc_and;
        }
        __syncthreads();
        if (threadIdx.x == 0 && c0 < c1) {
            uint64_t o = 0, a = ~0ull;
            for (int w = 0; w < kWavesPerBlock; ++w) {
                o |= red_or[w];
                a &= red_and[w];
            }
            atomicOr(reinterpret_cast<unsigned long long*>(&desc->or_mask), static_cast<unsigned long long>(o));
            atomicAnd(reinterpret_cast<unsigned long long*>(&desc->and_mask), static_cast<unsigned long long>(a));
        }
    }
}

// Exclusive scan in place of each digit's chunk counts (block d, at most
// kBlock * 16 chunks) and the digit's total.
__global__ __launch_bounds__(kBlock) void sort_count_scan_kernel(const SortDesc* desc, int pass, uint32_t* counts,
                                                                 uint32_t chunks, uint64_t* digit_total) {
    __shared__ uint32_t scan_lds[kWavesPerBlock];
    if (!((varying_digits(desc) >> pass) & 1u)) return;
    constexpr int kItems = 16;
    uint32_t* row = counts + static_cast<size_t>(blockIdx.x) * chunks;
    uint32_t x[kItems];
    uint32_t mine = 0;
#pragma unroll
    for (int q = 0; q < kItems; ++q) {
        const uint32_t i = threadIdx.x * kItems + q;
        x[q] = i < chunks ? row[i] : 0u;
        mine += x[q];
    }
    uint32_t total;
    uint32_t run = block_exclusive_sum<uint32_t>(mine, &total, scan_lds);
#pragma unroll
    for (int q = 0; q < kItems; ++q) {
        const uint32_t i = threadIdx.x * kItems + q;
        if (i < chunks) row[i] = run;
        run += x[q];
    }
    if (threadIdx.x == 0) digit_total[blockIdx.x] = total;
}

// Downsweep of a digit pass: block c sorts the tiles of chunk c in order.
// Each tile is ranked stably by wave64 match masks (wave w owns keys
// [w*1024, w*1024+1024) of the tile; item i of lane j is key w*1024 + i*64 +
// j), placed in LDS in digit order and written out in runs of one digit at
// the chunk's running offset for that digit. The next tile's keys are
// loaded while the current one is ranked.
__global__ __launch_bounds__(kBlock) void sort_downsweep_kernel(PairInput in, uint64_t* keys_a, uint64_t* keys_b,
                                                                const SortDesc* desc, int pass,
                                                                uint32_t chunk_tiles, const uint32_t* counts,
                                                                const uint64_t* digit_total) {
    // 40 KB in all, so 4 blocks fit a CU: the scans' scratch lives in
    // tile_keys while it is idle (before the first tile, and while a tile's
    // keys are in registers)
    __shared__ uint64_t tile_keys[kSortTile];
    __shared__ uint32_t wave_count[kWavesPerBlock][kRadix];
    __shared__ uint64_t digit_base[kRadix];  // output position of the digit's tile run - its local start
    __shared__ uint64_t running[kRadix];     // next output position of the digit in this chunk
    uint32_t* const scan_lds = reinterpret_cast<uint32_t*>(tile_keys);
    uint64_t* const scan64 = tile_keys;
    const PassIo io = pass_io(desc, pass, keys_a, keys_b);
    if (!io.active) return;
    const size_t rows = desc->rows;
    const uint32_t chunks = gridDim.x;
    const uint32_t c = blockIdx.x;
    const int shift = pass * kRadixBits;
    const uint32_t lane = lane_id(), wave = wave_id();
    const uint32_t d = threadIdx.x;  // thread d owns digit d
    {
        uint64_t total;
        const uint64_t start = block_exclusive_sum<uint64_t>(digit_total[d], &total, scan64);
        running[d] = start + counts[static_cast<size_t>(d) * chunks + c];
    }
    const size_t c0 = static_cast<size_t>(c) * chunk_tiles * kSortTile;
    const size_t c1 = c0 + static_cast<size_t>(chunk_tiles) * kSortTile < rows
                          ? c0 + static_cast<size_t>(chunk_tiles) * kSortTile
                          : rows;
    if (c0 >= c1) return;  // block-uniform
    uint64_t keys[kSortItems];
    const auto load_tile = [&](size_t t0, uint64_t (&kv)[kSortItems]) {
#pragma unroll
        for (int i = 0; i < kSortItems; ++i) {
            const size_t idx = t0 + wave * (kSortItems * kWave) + i * kWave + lane;
            const size_t at = idx < c1 ? idx : c1 - 1;
            kv[i] = io.from_input ? load_key(in, at) : io.src[at];
        }
    };
    load_tile(c0, keys);
    for (size_t t0 = c0; t0 < c1; t0 += kSortTile) {
        uint64_t next[kSortItems];
        if (t0 + kSortTile < c1) load_tile(t0 + kSortTile, next);
        for (int i = threadIdx.x; i < kWavesPerBlock * kRadix; i += kBlock) (&wave_count[0][0])[i] = 0;
        __syncthreads();
        uint32_t ranks[kSortItems];
#pragma unroll
        for (int i = 0; i < kSortItems; ++i) {
            const size_t idx = t0 + wave * (kSortItems * kWave) + i * kWave + lane;
            const bool valid = idx < c1;
            const uint32_t digit = static_cast<uint32_t>(keys[i] >> shift) & (kRadix - 1);
            const uint64_t peers = match_digit<kRadixBits>(digit, valid);
            const uint32_t before = valid ? wave_count[wave][digit] : 0;
            const uint32_t below = lanes_below(peers);
            ranks[i] = before + below;
            if (valid && below == 0) wave_count[wave][digit] = before + static_cast<uint32_t>(__popcll(peers));
        }
        __syncthreads();
        // thread d: the tile's count of digit d, exclusive over waves
        uint32_t tile_count = 0;
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; ++w) {
            const uint32_t cnt = wave_count[w][d];
            wave_count[w][d] = tile_count;
            tile_count += cnt;
        }
        uint32_t tile_total;
        const uint32_t local_start = block_exclusive_sum<uint32_t>(tile_count, &tile_total, scan_lds);
        digit_base[d] = running[d] - local_start;
        running[d] += tile_count;
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; ++w) wave_count[w][d] += local_start;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kSortItems; ++i) {
            const size_t idx = t0 + wave * (kSortItems * kWave) + i * kWave + lane;
            if (idx < c1) {
                const uint32_t digit = static_cast<uint32_t>(keys[i] >> shift) & (kRadix - 1);
                tile_keys[wave_count[wave][digit] + ranks[i]] = keys[i];
            }
        }
        __syncthreads();
        // consecutive threads write consecutive positions of one digit's run
        const uint32_t valid_rows = static_cast<uint32_t>(c1 - t0 < static_cast<size_t>(kSortTile) ? c1 - t0
                                                                                                  : kSortTile);
        for (uint32_t i = threadIdx.x; i < valid_rows; i += kBlock) {
            const uint64_t k = tile_keys[i];
            const uint32_t digit = static_cast<uint32_t>(k >> shift) & (kRadix - 1);
            io.dst[digit_base[digit] + i] = k;
        }
        __syncthreads();  // tile_keys and digit_base are rewritten by the next tile
#pragma unroll
        for (int i = 0; i < kSortItems; ++i) keys[i] = next[i];
    }
}

// Unique compaction of sorted keys (in keys_a or keys_b as the sort
// descriptor says, else in keys_a), both passes of a two-pass compaction
// (the tile-group counts are defined with the union below): tiles of
// kCompactTile rows; wave w owns rows [w*1024, w*1024+1024), item i of lane
// j is row w*1024 + i*64 + j, so every load is coalesced and a row's
// predecessor is a shuffle away. A row is kept when it differs from the row
// before it. The count pass writes the tile's kept count and adds it to its
// group's sum; the write pass writes the kept rows at the tile's offset.
template <bool kWrite>
__global__ __launch_bounds__(kBlock) void unique_tile_kernel(const uint64_t* keys_a, const uint64_t* keys_b,
                                                             const SortDesc* sort_desc, size_t rows,
                                                             const uint64_t* rows_dev, uint32_t* counts,
                                                             uint64_t* group_sums, PairOutput out);

__global__ void set_u64_kernel(uint64_t* p, uint64_t v) { *p = v; }
__global__ void set_u32_kernel(uint32_t* p, uint32_t v) { *p = v; }

// After an operator's kernels (stream order: every tile has finished): a
// look-back report turns the output count into all ones.
__global__ void report_error_kernel(const uint32_t* error, uint32_t* count32, uint64_t* count64) {
    if (*error == 0) return;
    if (count32) *count32 = 0xFFFFFFFFu;
    if (count64) *count64 = ~0ull;
}

// The look-back error word of a launch: the thread's sink, else the launch's
// own header (reported into the count afterwards by report_error).
uint32_t* error_word(Header* header) {
    uint32_t* sink = error_sink();
    return sink ? sink : &header->error;
}

// Chunks of a digit pass: at most kSortChunks (the count scan's capacity),
// each a whole number of tiles.
constexpr size_t kSortChunks = static_cast<size_t>(kBlock) * 16;

struct SortLayout {
    size_t desc_off, counts_off, totals_off, unique_counts_off, unique_groups_off, total;
    size_t sort_tiles, chunk_tiles, chunks, unique_tiles, unique_groups;
};

SortLayout sort_layout(size_t rows) {
    SortLayout l{};
    l.sort_tiles = tiles_of(rows, kSortTile);
    l.chunk_tiles = l.sort_tiles > kSortChunks ? tiles_of(l.sort_tiles, kSortChunks) : 1;
    l.chunks = tiles_of(l.sort_tiles, l.chunk_tiles);
    l.unique_tiles = tiles_of(rows, kCompactTile);
    l.unique_groups = tiles_of(l.unique_tiles, kTileGroup);
    l.desc_off = 0;
    size_t off = align_up(sizeof(SortDesc), 256);
    l.counts_off = off;
    off += align_up(kRadix * (l.chunks ? l.chunks : 1) * sizeof(uint32_t), 256);
    l.totals_off = off;
    off += align_up(kRadix * sizeof(uint64_t), 256);
    l.unique_counts_off = off;
    off += align_up(l.unique_tiles * sizeof(uint32_t), 256);
    l.unique_groups_off = off;
    off += align_up((l.unique_groups ? l.unique_groups : 1) * sizeof(uint64_t), 256);
    l.total = off;
    return l;
}

unsigned grid_for(size_t n, size_t per_block, unsigned cap);

// rows >= 1; in.packed, when set, is not keys_b. The sorted keys end in
// keys_a or keys_b by the parity of the number of varying digits.
hipError_t run_sort(PairInput in, size_t rows, uint64_t* keys_a, uint64_t* keys_b, void* temp,
                    const SortLayout& l, uint64_t known_varying, hipStream_t stream) {
    char* base = static_cast<char*>(temp);
    SortDesc* desc = reinterpret_cast<SortDesc*>(base + l.desc_off);
    uint32_t* counts = reinterpret_cast<uint32_t*>(base + l.counts_off);
    uint64_t* totals = reinterpret_cast<uint64_t*>(base + l.totals_off);
    sort_init_kernel<<<1, 64, 0, stream>>>(desc, known_varying, in, rows);
    const unsigned chunks = static_cast<unsigned>(l.chunks);
    const uint32_t chunk_tiles = static_cast<uint32_t>(l.chunk_tiles);
    for (int p = 0; p < kPasses; ++p) {
        if (known_varying && !((known_varying >> (p * kRadixBits)) & (kRadix - 1))) continue;
        // unknown varying digits: pass 0's upsweep finds them
        const int reduce = !known_varying && p == 0;
        sort_upsweep_kernel<<<chunks, kBlock, 0, stream>>>(in, keys_a, keys_b, desc, p, chunk_tiles, counts, reduce);
        if (reduce)
            sort_constant_copy_kernel<<<grid_for(rows, kBlock * 4, 2048), kBlock, 0, stream>>>(in, keys_a, desc);
        sort_count_scan_kernel<<<kRadix, kBlock, 0, stream>>>(desc, p, counts, chunks, totals);
        sort_downsweep_kernel<<<chunks, kBlock, 0, stream>>>(in, keys_a, keys_b, desc, p, chunk_tiles, counts,
                                                              totals);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// merge-path union

// Merge-path split of diagonal `diag` (the number of A rows among the first
// `diag` merged rows; A goes first on ties), searched in [lo, hi], which
// must hold it.
__device__ __forceinline__ size_t merge_path_range(const PairInput& a, const PairInput& b, size_t diag, size_t lo,
                                                   size_t hi) {
    while (lo < hi) {
        const size_t mid = (lo + hi) >> 1;
        if (!(load_key(b, diag - 1 - mid) < load_key(a, mid)))
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t merge_path_lds(const uint64_t* a, uint32_t na, const uint64_t* b, uint32_t nb,
                                                   uint32_t diag) {
    uint32_t lo = diag > nb ? diag - nb : 0;
    uint32_t hi = diag < na ? diag : na;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (!(b[diag - 1 - mid] < a[mid]))
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// The serial merge of a thread's kMergeItems outputs from (ia, ib): both
// sides' heads stay in registers, so each output reads LDS once (the side it
// consumed). A goes first on ties. Outputs past n are 0.
__device__ __forceinline__ void merge_serial(const uint64_t* la, uint32_t na, const uint64_t* lb, uint32_t nb,
                                             uint32_t ia, uint32_t ib, uint32_t diag, uint32_t n,
                                             uint64_t (&v)[kMergeItems]) {
    uint64_t ha = ia < na ? la[ia] : 0, hb = ib < nb ? lb[ib] : 0;
#pragma unroll
    for (int i = 0; i < kMergeItems; ++i) {
        const bool take_a = ib >= nb || (ia < na && ha <= hb);
        v[i] = 0;
        if (diag + i < n) {
            if (take_a) {
                v[i] = ha;
                if (++ia < na) ha = la[ia];
            } else {
                v[i] = hb;
                if (++ib < nb) hb = lb[ib];
            }
        }
    }
}

// LDS index of merged row j in a blocked-to-striped transposition: thread t's
// rows [8t, 8t + 8) written unpadded would put lanes 64 bytes apart in a
// quarter of the banks; one u64 of padding per 32 rows spreads them.
__device__ __forceinline__ uint32_t merge_pad(uint32_t j) { return j + (j >> 5); }

// Merge-path partition of every tile boundary: part[t] = number of left rows
// among the first t*kMergeTile merged rows. One thread per boundary, so the
// searches of all tiles overlap instead of stalling each tile's start.
// Merge of two sorted arrays with no key in common (FULL's runs): merged row
// d goes to out[d], so every tile knows its output range and no tile waits
// for another. Tile t covers merged rows [t*kMergeTile, (t+1)*kMergeTile),
// whose split between the inputs comes from the partition pass.
__global__ __launch_bounds__(kBlock) void merge_disjoint_kernel(const uint64_t* a, size_t na, const uint64_t* b,
                                                                size_t nb, uint64_t* out, uint32_t* out_second,
                                                                const size_t* part) {
    __shared__ uint64_t lds[kMergeTile + kMergeTile / 32];  // + padding of the transposition (merge_pad)
    const size_t tile = blockIdx.x;
    const size_t total = na + nb;
    const size_t d0 = tile * kMergeTile;
    if (d0 >= total) return;
    const size_t d1 = d0 + kMergeTile < total ? d0 + kMergeTile : total;
    const size_t a0 = part[tile], a1 = part[tile + 1];
    const size_t b0 = d0 - a0, b1 = d1 - a1;
    const uint32_t n_a = static_cast<uint32_t>(a1 - a0), n_b = static_cast<uint32_t>(b1 - b0);
    uint64_t* la = lds;
    uint64_t* lb = lds + n_a;
    for (uint32_t i = threadIdx.x; i < n_a; i += kBlock) la[i] = a[a0 + i];
    for (uint32_t i = threadIdx.x; i < n_b; i += kBlock) lb[i] = b[b0 + i];
    __syncthreads();
    const uint32_t n = n_a + n_b;
    const uint32_t diag = threadIdx.x * kMergeItems < n ? threadIdx.x * kMergeItems : n;
    uint32_t ia = merge_path_lds(la, n_a, lb, n_b, diag);
    uint32_t ib = diag - ia;
    uint64_t v[kMergeItems];
    merge_serial(la, n_a, lb, n_b, ia, ib, diag, n, v);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kMergeItems; ++i)
        if (diag + i < n) lds[merge_pad(diag + i)] = v[i];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += kBlock) out[d0 + i] = lds[merge_pad(i)];
    if (out_second)
        for (uint32_t i = threadIdx.x; i < n; i += kBlock)
            out_second[d0 + i] = static_cast<uint32_t>(lds[merge_pad(i)]);
}

// Merge-path splits part[t] of the diagonals t*step (t < n_parts). Without
// `coarse` each split is searched over the whole input; with the splits of
// every group-th diagonal in `coarse` (a launch of this kernel with step *
// group), a split is searched only between those of the coarse diagonals
// around it: the split s of d satisfies i0 <= s <= i1 and D0 - i0 <= d - s
// <= D1 - i1, and the window's rows are few and close together (at 2^28
// rows a side, one level took 29 dependent random loads per split and 0.5 ms
// in all). Also zeroes `zero[0, n_zero)` (the tile-group sums of a two-pass
// union).
__global__ void merge_partition_kernel(PairInput left, size_t nl, PairInput right, size_t nr, size_t* part,
                                       size_t n_parts, size_t step, const size_t* coarse, size_t group,
                                       uint64_t* zero, size_t n_zero) {
    nl = actual_rows(left, nl);
    nr = actual_rows(right, nr);
    const size_t total = nl + nr;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n_zero; t += stride) zero[t] = 0;
    for (size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n_parts; t += stride) {
        const size_t d = t * step < total ? t * step : total;
        size_t lo = d > nr ? d - nr : 0;
        size_t hi = d < nl ? d : nl;
        if (coarse) {
            const size_t g = t / group;
            if (t % group == 0) {
                part[t] = coarse[g];
                continue;
            }
            const size_t D0 = g * group * step < total ? g * group * step : total;
            const size_t D1 = (g + 1) * group * step < total ? (g + 1) * group * step : total;
            const size_t i0 = coarse[g], i1 = coarse[g + 1];
            const size_t b0 = D0 - i0, b1 = D1 - i1;  // B rows before the coarse diagonals
            if (i0 > lo) lo = i0;
            if (i1 < hi) hi = i1;
            if (d > b1 && d - b1 > lo) lo = d - b1;
            if (d - b0 + i0 < hi) hi = d - b0 + i0;  // d >= D0 >= b0
        }
        part[t] = merge_path_range(left, right, d, lo, hi);
    }
}

// Diagonals of a partition searched directly; above this, through a coarse
// level of every kPartGroup-th diagonal.
constexpr size_t kPartGroup = 64;
constexpr size_t kPartDirect = 4 * kPartGroup;

size_t coarse_parts(size_t tiles) { return tiles_of(tiles, kPartGroup) + 1; }

// part[0, tiles] of a merge in kMergeTile-row tiles; `coarse` holds
// coarse_parts(tiles) entries.
void merge_partition(PairInput left, size_t nl, PairInput right, size_t nr, size_t* part, size_t tiles,
                     size_t* coarse, hipStream_t stream, uint64_t* zero = nullptr, size_t n_zero = 0) {
    const size_t n_parts = tiles + 1;
    if (n_parts <= kPartDirect) {
        merge_partition_kernel<<<grid_for(n_parts > n_zero ? n_parts : n_zero, kBlock, 4096), kBlock, 0, stream>>>(
            left, nl, right, nr, part, n_parts, kMergeTile, nullptr, 1, zero, n_zero);
        return;
    }
    const size_t n_coarse = coarse_parts(tiles);
    merge_partition_kernel<<<grid_for(n_coarse, 64, 4096), 64, 0, stream>>>(
        left, nl, right, nr, coarse, n_coarse, static_cast<size_t>(kMergeTile) * kPartGroup, nullptr, 1, zero, n_zero);
    merge_partition_kernel<<<grid_for(n_parts, kBlock, 4096), kBlock, 0, stream>>>(left, nl, right, nr, part, n_parts,
                                                                                  kMergeTile, coarse, kPartGroup,
                                                                                  nullptr, 0);
}

// Tile counts of the two-pass compactions (union, anti-join): tiles are
// grouped by kTileGroup; a group's sum is added by its tiles' count pass
// (one atomic per tile), a one-block scan turns the group sums into
// offsets, and a tile of the write pass adds the counts of the earlier
// tiles of its group (one wave load). No tile waits on another: on this
// chip a status poll crosses XCDs, and tiles waiting for predecessors in a
// decoupled look-back ran at a third of the copy ceiling.

__global__ __launch_bounds__(1024) void group_scan_kernel(uint64_t* group_sums, size_t n, uint64_t* total64,
                                                          uint32_t* total32) {
    __shared__ uint64_t scan_lds[1024 / kWave];
    __shared__ uint64_t total_slot;
    block_scan_inplace<1024, 4>(group_sums, n, &total_slot, scan_lds);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (total64) *total64 = total_slot;
        if (total32) *total32 = static_cast<uint32_t>(total_slot);
    }
}

// Output offset of `tile` from the scanned group sums and the counts of the
// earlier tiles of its group; call from one whole wave.
__device__ __forceinline__ uint64_t group_tile_offset(const uint64_t* group_prefix, const uint32_t* counts,
                                                      uint32_t tile) {
    const uint32_t g0 = tile & ~static_cast<uint32_t>(kTileGroup - 1);
    const uint32_t lane = lane_id();
    const uint32_t c = g0 + lane < tile ? counts[g0 + lane] : 0u;
    return group_prefix[tile / kTileGroup] + wave_sum(static_cast<uint64_t>(c));
}

template <bool kWrite>
__global__ __launch_bounds__(kBlock) void unique_tile_kernel(const uint64_t* keys_a, const uint64_t* keys_b,
                                                             const SortDesc* sort_desc, size_t rows,
                                                             const uint64_t* rows_dev, uint32_t* counts,
                                                             uint64_t* group_sums, PairOutput out) {
    constexpr int kItems = kCompactItems;      // 16
    constexpr int kWaveRows = kItems * kWave;  // 1024
    __shared__ uint32_t wave_tot[kWavesPerBlock];
    __shared__ uint64_t offset_slot;
    const uint64_t* keys = keys_a;
    if (sort_desc) {
        keys = (__popc(varying_digits(sort_desc)) & 1u) ? keys_b : keys_a;
        rows = sort_desc->rows;
    }
    if (rows_dev) rows = *rows_dev < rows ? *rows_dev : rows;
    const uint32_t tile = blockIdx.x;
    const size_t base = static_cast<size_t>(tile) * kCompactTile;
    if (base >= rows) {
        if (!kWrite && threadIdx.x == 0) counts[tile] = 0;
        return;
    }
    const uint32_t lane = lane_id(), wave = wave_id();
    const size_t w0 = base + static_cast<size_t>(wave) * kWaveRows;  // the wave's first row
    // every load is issued before any is used (rows clamped to the input)
    uint64_t k[kItems];
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const size_t r = w0 + i * kWave + lane;
        k[i] = keys[r < rows ? r : rows - 1];
    }
    uint64_t before0 = 0;  // the row before the wave's first (lane 0)
    if (lane == 0 && w0 > 0 && w0 <= rows) before0 = keys[w0 - 1];
    uint64_t offset = 0;
    if (kWrite && wave == 0) offset = group_tile_offset(group_sums, counts, tile);
    uint64_t masks[kItems];
    uint32_t mine = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const size_t r = w0 + i * kWave + lane;
        const uint64_t up = __shfl_up(k[i], 1, kWave);
        const uint64_t last_of_prev = i > 0 ? __shfl(k[i > 0 ? i - 1 : 0], kWave - 1, kWave) : before0;
        const uint64_t pred = lane > 0 ? up : last_of_prev;
        masks[i] = ballot(r < rows && (r == 0 || k[i] != pred));
        mine += static_cast<uint32_t>(__popcll(masks[i]));
    }
    if (lane == 0) wave_tot[wave] = mine;
    if (kWrite && threadIdx.x == 0) offset_slot = offset;
    __syncthreads();
    if (!kWrite) {
        if (threadIdx.x == 0) {
            uint32_t c = 0;
            for (int w = 0; w < kWavesPerBlock; ++w) c += wave_tot[w];
            counts[tile] = c;
            atomicAdd(reinterpret_cast<unsigned long long*>(&group_sums[tile / kTileGroup]),
                      static_cast<unsigned long long>(c));
        }
        return;
    }
    uint64_t pos = offset_slot;
    for (int w = 0; w < static_cast<int>(wave); ++w) pos += wave_tot[w];
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        if ((masks[i] >> lane) & 1ull) store_key(out, pos + lanes_below(masks[i]), k[i]);
        pos += static_cast<uint32_t>(__popcll(masks[i]));
    }
}

// Sorted union, both passes: tile t merges merged rows [t*kMergeTile,
// (t+1)*kMergeTile) (its split between the inputs from the partition pass)
// and keeps the rows that differ from their predecessor in merged order. The
// count pass writes the tile's kept count and adds it to its group's sum;
// the write pass merges again and writes the kept rows at the tile's offset.
template <bool kWrite>
__global__ __launch_bounds__(kBlock) void union_tile_kernel(PairInput left, size_t nl, PairInput right, size_t nr,
                                                            const size_t* part, uint32_t* counts,
                                                            uint64_t* group_sums, PairOutput out) {
    __shared__ uint64_t lds[kMergeTile + kMergeTile / 32 + 2];  // + padding of the compaction (merge_pad)
    __shared__ uint32_t scan_lds[kWavesPerBlock];
    __shared__ uint64_t wave_last[kWavesPerBlock];  // each wave's last merged row
    __shared__ uint64_t offset_slot;
    __shared__ uint64_t prev_slot;
    nl = actual_rows(left, nl);
    nr = actual_rows(right, nr);
    const uint32_t tile = blockIdx.x;
    const size_t total = nl + nr;
    const size_t d0 = static_cast<size_t>(tile) * kMergeTile;
    if (d0 >= total) {
        if (!kWrite && threadIdx.x == 0) counts[tile] = 0;
        return;
    }
    const size_t d1 = d0 + kMergeTile < total ? d0 + kMergeTile : total;
    const size_t a0 = part[tile], a1 = part[tile + 1];
    const size_t b0 = d0 - a0, b1 = d1 - a1;
    const uint32_t na = static_cast<uint32_t>(a1 - a0), nb = static_cast<uint32_t>(b1 - b0);
    const uint32_t n = na + nb;
    // every load of the tile is issued before any is used: position j of the
    // tile's A-then-B slice lands at lds[j] (positions past n reload a valid
    // row and are dropped); thread 0 also loads the row before the tile in
    // merged order (the larger of the last A and last B before it), and in
    // the write pass wave 0 reads the tile's offset
    uint64_t x[kMergeItems];
#pragma unroll
    for (int k = 0; k < kMergeItems; ++k) {
        const uint32_t j = threadIdx.x + k * kBlock;
        const uint32_t jj = j < n ? j : 0;
        x[k] = jj < na ? load_key(left, a0 + jj) : load_key(right, b0 + (jj - na));
    }
    uint64_t pa = 0, pb = 0;
    if (threadIdx.x == 0) {
        if (a0 > 0) pa = load_key(left, a0 - 1);
        if (b0 > 0) pb = load_key(right, b0 - 1);
    }
    uint64_t offset = 0;
    if (kWrite && wave_id() == 0) offset = group_tile_offset(group_sums, counts, tile);
#pragma unroll
    for (int k = 0; k < kMergeItems; ++k) {
        const uint32_t j = threadIdx.x + k * kBlock;
        if (j < n) lds[j] = x[k];
    }
    if (threadIdx.x == 0) {
        // flag of merged row 0: it differs from the row before the tile, if any
        prev_slot = a0 == 0 ? pb : (b0 == 0 ? pa : (pa > pb ? pa : pb));
        if (kWrite) offset_slot = offset;
    }
    __syncthreads();
    const bool have_prev = d0 > 0;
    const uint64_t* la = lds;
    const uint64_t* lb = lds + na;
    // each thread merges kMergeItems consecutive outputs
    const uint32_t diag = threadIdx.x * kMergeItems < n ? threadIdx.x * kMergeItems : n;
    const uint32_t ia = merge_path_lds(la, na, lb, nb, diag);
    uint64_t v[kMergeItems];
    merge_serial(la, na, lb, nb, ia, diag - ia, diag, n, v);
    // each row's predecessor in merged order: the thread's previous row, or
    // for its first row the previous thread's last (a shuffle inside the
    // wave, the previous wave's last row across waves)
    const uint64_t before = __shfl_up(v[kMergeItems - 1], 1, kWave);
    if (lane_id() == kWave - 1) wave_last[wave_id()] = v[kMergeItems - 1];
    __syncthreads();  // also: every LDS read of the merge is done
    const uint64_t first_prev = lane_id() > 0 ? before : (wave_id() > 0 ? wave_last[wave_id() - 1] : 0);
    uint32_t keep = 0;
#pragma unroll
    for (int i = 0; i < kMergeItems; ++i) {
        const uint32_t pos = diag + i;
        if (pos < n) {
            bool flag;
            if (pos == 0)
                flag = !have_prev || v[i] != prev_slot;
            else
                flag = v[i] != (i > 0 ? v[i - 1] : first_prev);
            keep |= static_cast<uint32_t>(flag) << i;
        }
    }
    uint32_t tile_total;
    const uint32_t local = block_exclusive_sum<uint32_t>(static_cast<uint32_t>(__popc(keep)), &tile_total, scan_lds);
    if (!kWrite) {
        if (threadIdx.x == 0) {
            counts[tile] = tile_total;
            atomicAdd(reinterpret_cast<unsigned long long*>(&group_sums[tile / kTileGroup]),
                      static_cast<unsigned long long>(tile_total));
        }
        return;
    }
    // compact in LDS (reads of lds[] above are complete after the scan's barriers)
    uint32_t pos = local;
#pragma unroll
    for (int i = 0; i < kMergeItems; ++i)
        if ((keep >> i) & 1u) lds[merge_pad(pos++)] = v[i];
    __syncthreads();
    const uint64_t base = offset_slot;
    for (uint32_t i = threadIdx.x; i < tile_total; i += kBlock) store_key(out, base + i, lds[merge_pad(i)]);
}

// ------------------------------------------------------------------------
// anti-join

__device__ __forceinline__ size_t lower_bound_key(const PairInput& r, size_t lo, size_t hi, uint64_t key) {
    while (lo < hi) {
        const size_t mid = (lo + hi) >> 1;
        if (load_key(r, mid) < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// For every left tile and every right input: the right rows that can match a
// sorted tile, [lower_bound(first key), upper_bound(last key)). One thread per
// (tile, right input).
// Left key i as a pair (first << 32 | second). With minor_bits != 0 the left
// input holds compressed keys (first << minor_bits | second) of candidate
// rows grouped by their first column.
__device__ __forceinline__ uint64_t left_pair(const PairInput& in, size_t i, uint32_t minor_bits) {
    const uint64_t k = load_key(in, i);
    if (minor_bits == 0) return k;
    return ((k >> minor_bits) << 32) | (k & ((1ull << minor_bits) - 1));
}

__global__ void anti_partition_kernel(PairInput left, size_t nl, AntiRight rights, size_t n_tiles,
                                      size_t* ranges, uint32_t minor_bits) {
    nl = actual_rows(left, nl);
    const size_t tiles = tiles_of(nl, kAntiTile) < n_tiles ? tiles_of(nl, kAntiTile) : n_tiles;
    const size_t jobs = tiles * rights.count;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t j = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < jobs; j += stride) {
        const size_t t = j / rights.count;
        const uint32_t r = static_cast<uint32_t>(j % rights.count);
        const PairInput& right = rights.in[r];
        const size_t nr = rights.rows[r];
        const size_t first = t * kAntiTile;
        const size_t last = (first + kAntiTile < nl ? first + kAntiTile : nl) - 1;
        uint64_t lo_key = left_pair(left, first, minor_bits), hi_key = left_pair(left, last, minor_bits);
        if (minor_bits) {  // grouped tile: every pair of its first and last first values
            lo_key &= ~0xFFFFFFFFull;
            hi_key |= 0xFFFFFFFFull;
        }
        const size_t lo = lower_bound_key(right, 0, nr, lo_key);
        size_t hi = lower_bound_key(right, lo, nr, hi_key);
        if (hi < nr && load_key(right, hi) == hi_key) ++hi;
        ranges[2 * j] = lo;
        ranges[2 * j + 1] = hi > lo ? hi : lo;
    }
}

// left \ (right_0 ∪ right_1 ∪ ...), first pass: which left rows some right
// input holds. Layout: wave w owns tile rows [w*512, w*512+512); item i of
// lane j is row w*512 + i*64 + j, so every load is coalesced and an item's
// result is one ballot word. Each right input's candidate range for the tile
// comes from the partition pass (or, for a tile that is not sorted, from a
// block-cooperative search); ranges are streamed through LDS once, or probed
// in place when huge. Tiles are independent: the time a tile takes depends
// on its ranges, and a single pass with look-back would keep finished tiles
// waiting for slower predecessors (measured: over a third of the time).
// With `tile_kept`, the tile also writes its count of rows no right input
// holds and adds it to its tile group's sum (the count pass of the two-pass
// compaction, anti_write_kernel).
__global__ __launch_bounds__(kBlock) void anti_join_kernel(PairInput left, size_t nl, AntiRight rights,
                                                           const size_t* ranges, uint64_t* found_words,
                                                           uint32_t minor_bits, uint32_t* tile_kept = nullptr,
                                                           uint64_t* group_sums = nullptr) {
    constexpr int kItems = kAntiTile / kBlock;  // 8
    constexpr int kWaveRows = kItems * kWave;   // 512
    __shared__ uint64_t stage[kAntiStage];
    __shared__ uint32_t wave_totals[kWavesPerBlock][kItems];
    __shared__ uint64_t red_min[kWavesPerBlock], red_max[kWavesPerBlock];
    nl = actual_rows(left, nl);
    const uint32_t tile = blockIdx.x;
    const size_t base = static_cast<size_t>(tile) * kAntiTile;
    if (base >= nl) {
        if (tile_kept && threadIdx.x == 0) tile_kept[tile] = 0;
        return;
    }
    const uint32_t lane = lane_id(), wave = wave_id();
    const size_t rows = nl - base < static_cast<size_t>(kAntiTile) ? nl - base : kAntiTile;
    const uint64_t first_key = left_pair(left, base, minor_bits), last_key = left_pair(left, base + rows - 1, minor_bits);
    uint64_t k[kItems];
    uint64_t kmin = ~0ull, kmax = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const size_t r = wave * kWaveRows + i * kWave + lane;
        k[i] = r < rows ? left_pair(left, base + r, minor_bits) : 0;
        if (r < rows) {
            kmin = k[i] < kmin ? k[i] : kmin;
            kmax = k[i] > kmax ? k[i] : kmax;
        }
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const uint64_t omin = __shfl_xor(kmin, off, kWave), omax = __shfl_xor(kmax, off, kWave);
        kmin = omin < kmin ? omin : kmin;
        kmax = omax > kmax ? omax : kmax;
    }
    if (lane == 0) {
        red_min[wave] = kmin;
        red_max[wave] = kmax;
    }
    __syncthreads();
    uint64_t lo_key = ~0ull, hi_key = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) {
        lo_key = red_min[w] < lo_key ? red_min[w] : lo_key;
        hi_key = red_max[w] > hi_key ? red_max[w] : hi_key;
    }
    // a sorted tile may use the precomputed ranges; any other order gets its
    // own (pure membership semantics for every input order)
    // (ranges of a grouped tile cover all of its first values, in any order)
    const bool sorted_tile = ranges != nullptr && (minor_bits != 0 || (lo_key == first_key && hi_key == last_key));
    uint32_t found = 0;  // bit i: item i matched
    for (uint32_t r = 0; r < rights.count; ++r) {
        const PairInput& right = rights.in[r];
        const size_t nr = rights.rows[r];
        size_t r0, r1;
        if (sorted_tile) {
            const size_t j = static_cast<size_t>(tile) * rights.count + r;
            r0 = ranges[2 * j];
            r1 = ranges[2 * j + 1];
        } else {
            r0 = block_search(0, nr, [&](size_t i) { return load_key(right, i) < lo_key; });
            r1 = block_search(r0, nr, [&](size_t i) { return load_key(right, i) <= hi_key; });
        }
        const size_t span = r1 - r0;
        if (span <= kAntiStreamLimit) {
            for (size_t c0 = r0; c0 < r1; c0 += kAntiStage) {
                const uint32_t n =
                    static_cast<uint32_t>(r1 - c0 < static_cast<size_t>(kAntiStage) ? r1 - c0 : kAntiStage);
                for (uint32_t i = threadIdx.x; i < n; i += kBlock) stage[i] = load_key(right, c0 + i);
                __syncthreads();
                const uint64_t cmin = stage[0], cmax = stage[n - 1];
#pragma unroll
                for (int i = 0; i < kItems; ++i) {
                    const size_t row = wave * kWaveRows + i * kWave + lane;
                    if (row < rows && !((found >> i) & 1u) && k[i] >= cmin && k[i] <= cmax) {
                        uint32_t lo = 0, hi = n;
                        while (lo < hi) {
                            const uint32_t mid = (lo + hi) >> 1;
                            if (stage[mid] < k[i])
                                lo = mid + 1;
                            else
                                hi = mid;
                        }
                        if (lo < n && stage[lo] == k[i]) found |= 1u << i;
                    }
                }
                __syncthreads();
            }
        } else {
#pragma unroll
            for (int i = 0; i < kItems; ++i) {
                const size_t row = wave * kWaveRows + i * kWave + lane;
                if (row < rows && !((found >> i) & 1u)) {
                    const size_t p = lower_bound_key(right, r0, r1, k[i]);
                    if (p < r1 && load_key(right, p) == k[i]) found |= 1u << i;
                }
            }
        }
    }
    // found rows as one 64-bit word per item: bit j of word
    // (base + wave*512 + i*64) / 64 is row base + wave*512 + i*64 + j
    uint32_t found_rows = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const uint64_t m = ballot(((found >> i) & 1u) != 0);
        const size_t row0 = base + wave * kWaveRows + i * kWave;
        if (lane == 0 && row0 < nl) found_words[row0 / kWave] = m;
        found_rows += static_cast<uint32_t>(__popcll(m));
    }
    if (tile_kept) {
        if (lane == 0) wave_totals[wave][0] = found_rows;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t f = 0;
            for (int w = 0; w < kWavesPerBlock; ++w) f += wave_totals[w][0];
            const uint32_t kept = static_cast<uint32_t>(rows) - f;
            tile_kept[tile] = kept;
            atomicAdd(reinterpret_cast<unsigned long long*>(&group_sums[tile / kTileGroup]),
                      static_cast<unsigned long long>(kept));
        }
    }
}

// Order-preserving compaction of the left rows whose found bit is clear, at
// the offset the tile counts give (the write pass of the two-pass
// compaction): the anti-join kernel's tiles and row layout, so a tile's
// found bits are eight words.
__global__ __launch_bounds__(kBlock) void anti_write_kernel(PairInput left, size_t nl, const uint64_t* found_words,
                                                            const uint32_t* tile_kept, const uint64_t* group_prefix,
                                                            PairOutput out) {
    constexpr int kItems = kAntiTile / kBlock;  // 8
    constexpr int kWaveRows = kItems * kWave;   // 512
    __shared__ uint32_t wave_tot[kWavesPerBlock];
    __shared__ uint64_t offset_slot;
    nl = actual_rows(left, nl);
    const uint32_t tile = blockIdx.x;
    const size_t base = static_cast<size_t>(tile) * kAntiTile;
    if (base >= nl) return;
    const uint32_t lane = lane_id(), wave = wave_id();
    const uint32_t rows = static_cast<uint32_t>(nl - base < static_cast<size_t>(kAntiTile) ? nl - base : kAntiTile);
    uint64_t k[kItems], words[kItems];
    // every load is issued before any is used (rows clamped to the tile)
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const uint32_t r = wave * kWaveRows + i * kWave + lane;
        k[i] = load_key(left, base + (r < rows ? r : rows - 1));
        const size_t row0 = base + wave * kWaveRows + i * kWave;  // multiple of 64
        words[i] = row0 < nl ? found_words[row0 / kWave] : ~0ull;
    }
    uint64_t offset = 0;
    if (wave == 0) offset = group_tile_offset(group_prefix, tile_kept, tile);
    uint64_t masks[kItems];
    uint32_t mine = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const uint32_t r = wave * kWaveRows + i * kWave + lane;
        masks[i] = ballot(r < rows && !((words[i] >> lane) & 1ull));
        mine += static_cast<uint32_t>(__popcll(masks[i]));
    }
    if (lane == 0) wave_tot[wave] = mine;
    if (threadIdx.x == 0) offset_slot = offset;
    __syncthreads();
    uint64_t pos = offset_slot;
    for (int w = 0; w < static_cast<int>(wave); ++w) pos += wave_tot[w];
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        if ((masks[i] >> lane) & 1ull) store_key(out, pos + lanes_below(masks[i]), k[i]);
        pos += static_cast<uint32_t>(__popcll(masks[i]));
    }
}

// ------------------------------------------------------------------------
// join (range-index layout: keys[K] sorted unique, starts[K+1], rows[R])

__device__ __forceinline__ int64_t find_index_key(const uint32_t* keys, size_t k, uint32_t key) {
    size_t lo = 0, hi = k;
    while (lo < hi) {
        const size_t mid = (lo + hi) >> 1;
        if (keys[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return (lo < k && keys[lo] == key) ? static_cast<int64_t>(lo) : -1;
}

// The same search through a sample of the key column held in LDS:
// s[j] = keys[min(j * stride, k - 1)] brackets the key's position within
// `stride` entries, a guess interpolated between the bracket's two keys is
// checked with a window of kKeyWindow entries (two loads), and only a key
// outside its window searches the whole bracket. With 2^23 uniform keys a
// search takes about 3 loads from global memory instead of 23.
constexpr int kKeySample = 2048;  // 8 KB of LDS
constexpr size_t kKeyWindow = 32;

__host__ __device__ __forceinline__ size_t key_sample_stride(size_t k) {
    return k <= static_cast<size_t>(kKeySample) ? 1 : (k + kKeySample - 1) / kKeySample;
}

// Every thread of the block calls it (k >= 1); ends with a barrier.
__device__ __forceinline__ void load_key_sample(const uint32_t* keys, size_t k, size_t stride, uint32_t* s) {
    for (uint32_t j = threadIdx.x; j < static_cast<uint32_t>(kKeySample); j += kBlock) {
        const size_t at = static_cast<size_t>(j) * stride;
        s[j] = keys[at < k ? at : k - 1];
    }
    __syncthreads();
}

__device__ __forceinline__ int64_t find_index_key_sampled(const uint32_t* keys, size_t k, uint32_t key,
                                                          const uint32_t* s, size_t stride) {
    // J = sample entries below the key
    uint32_t lo = 0, hi = kKeySample;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    const uint32_t J = lo;
    // the first position with keys[i] >= key lies in [a, b]; keys[b] >= key
    // unless b == k
    const size_t p0 = J == 0 ? 0 : ((J - 1) * stride < k - 1 ? (J - 1) * stride : k - 1);
    size_t a = J == 0 ? 0 : p0 + 1;
    size_t b = J == static_cast<uint32_t>(kKeySample) ? k : (J * stride < k - 1 ? J * stride : k - 1);
    if (a < b && J > 0 && J < static_cast<uint32_t>(kKeySample)) {
        const uint32_t k0 = s[J - 1], k1 = s[J];  // k0 < key <= k1, at p0 and b
        const size_t guess = p0 + static_cast<size_t>((static_cast<uint64_t>(key - k0) * (b - p0)) /
                                                       static_cast<uint64_t>(k1 - k0));
        const size_t wl = guess > a + kKeyWindow / 2 ? guess - kKeyWindow / 2 : a;
        const size_t wr = wl + kKeyWindow < b ? wl + kKeyWindow : b;
        const uint32_t left = keys[wl > a ? wl - 1 : a], right = keys[wr < b ? wr : b];
        if ((wl == a || left < key) && (wr == b || right >= key)) {
            a = wl;
            b = wr;
        }
    }
    while (a < b) {
        const size_t mid = (a + b) >> 1;
        if (keys[mid] < key)
            a = mid + 1;
        else
            b = mid;
    }
    return (a < k && keys[a] == key) ? static_cast<int64_t>(a) : -1;
}

__global__ __launch_bounds__(kBlock) void join_count_kernel(const uint32_t* left_keys, size_t nl,
                                                            const uint32_t* index_keys, const uint32_t* index_starts,
                                                            size_t k, uint64_t* counts, uint64_t* offsets,
                                                            uint64_t* total, uint64_t* status, Header* header,
                                                            uint32_t* error) {
    constexpr int kItems = 8;
    constexpr int kTile = kBlock * kItems;
    __shared__ uint64_t scan_lds[kWavesPerBlock];
    __shared__ uint64_t offset_slot;
    __shared__ uint32_t tile_slot;
    __shared__ uint32_t key_sample[kKeySample];
    const uint32_t tile = next_tile(&header->counters[0], &tile_slot);
    const size_t base = static_cast<size_t>(tile) * kTile;
    const size_t stride = key_sample_stride(k);
    if (k > 0) load_key_sample(index_keys, k, stride, key_sample);  // block-uniform
    // striped arrangement for coalescing: item i of thread t is row base + i*256 + t
    uint64_t c[kItems];
    uint64_t sum = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const size_t row = base + static_cast<size_t>(i) * kBlock + threadIdx.x;
        c[i] = 0;
        if (row < nl) {
            const int64_t idx = k > 0 ? find_index_key_sampled(index_keys, k, left_keys[row], key_sample, stride) : -1;
            if (idx >= 0) c[i] = index_starts[idx + 1] - index_starts[idx];
            counts[row] = c[i];
        }
    }
    // exclusive scan in row order: rows of item i precede rows of item i+1
    uint64_t running = 0;
    for (int i = 0; i < kItems; ++i) {
        uint64_t part_total;
        const uint64_t ex = block_exclusive_sum<uint64_t>(c[i], &part_total, scan_lds);
        c[i] = running + ex;
        running += part_total;
    }
    sum = running;
    if (wave_id() == 0) {
        const uint64_t excl = wave_lookback(status, tile, sum, 0, error);
        if (lane_id() == 0) offset_slot = excl;
    }
    __syncthreads();
    const uint64_t offset = offset_slot;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const size_t row = base + static_cast<size_t>(i) * kBlock + threadIdx.x;
        if (row < nl) offsets[row] = offset + c[i];
    }
    if (tile == tiles_of(nl, kTile) - 1 && threadIdx.x == 0) *total = offset + sum;
}

// Output-balanced emit: block b writes output rows [b*T, (b+1)*T). Runs over a
// fixed grid; the output size comes from the device (last offset + count).
__global__ __launch_bounds__(kBlock) void join_emit_kernel(const uint32_t* left_keys, size_t nl,
                                                           const uint32_t* index_keys, const uint32_t* index_starts,
                                                           const uint32_t* index_rows, size_t k,
                                                           const uint32_t* p0, uint32_t side0, const uint32_t* p1,
                                                           uint32_t side1, const uint64_t* offsets, uint32_t* out0,
                                                           uint32_t* out1) {
    __shared__ uint64_t row_offset[kEmitStage + 1];
    __shared__ uint32_t row_match[kEmitStage];
    __shared__ uint16_t row_list[kEmitStage];  // the tile's rows with outputs
    __shared__ uint32_t list_len;
    __shared__ size_t bounds[2];
    __shared__ uint64_t total_slot;
    __shared__ uint32_t key_sample[kKeySample];
    if (threadIdx.x == 0) {
        const int64_t idx = find_index_key(index_keys, k, left_keys[nl - 1]);
        const uint64_t last = idx >= 0 ? index_starts[idx + 1] - index_starts[idx] : 0;
        total_slot = offsets[nl - 1] + last;
    }
    __syncthreads();
    const uint64_t total = total_slot;
    const uint64_t tiles = (total + kEmitTile - 1) / kEmitTile;
    // a block with no output tile leaves before loading the key sample (the
    // test is block-uniform, so no barrier below is left half-reached)
    if (blockIdx.x >= tiles) return;
    const size_t stride = key_sample_stride(k);
    if (k > 0) load_key_sample(index_keys, k, stride, key_sample);  // block-uniform
    for (uint64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint64_t o0 = tile * kEmitTile;
        const uint64_t o1 = o0 + kEmitTile < total ? o0 + kEmitTile : total;
        // rows [l0, l1) cover outputs [o0, o1): l = upper_bound(offsets, o) - 1
        const size_t l0 = block_search(0, nl, [&](size_t i) { return offsets[i] <= o0; }) - 1;
        const size_t l1 = block_search(l0, nl, [&](size_t i) { return offsets[i] <= o1 - 1; });
        (void)bounds;
        const size_t span = l1 - l0;
        const bool staged = span <= static_cast<size_t>(kEmitStage);
        if (staged) {
            // only rows with outputs need their key searched (about a third of
            // uniform probe keys have none): the tile's offsets go to LDS, the
            // rows whose next offset is larger are listed (any order), and the
            // block's threads share the list's searches
            if (threadIdx.x == 0) list_len = 0;
            for (uint32_t i = threadIdx.x; i < span; i += kBlock) row_offset[i] = offsets[l0 + i];
            if (threadIdx.x == 0) row_offset[span] = l1 < nl ? offsets[l1] : total;
            __syncthreads();
            for (uint32_t c = 0; c < span; c += kBlock) {  // block-uniform trips
                const uint32_t i = c + threadIdx.x;
                const bool has = i < span && row_offset[i + 1] > row_offset[i];
                const uint64_t m = ballot(has);
                uint32_t at = 0;
                if (lane_id() == 0 && m) at = atomicAdd(&list_len, static_cast<uint32_t>(__popcll(m)));
                at = __shfl(at, 0, kWave);
                if (has) row_list[at + lanes_below(m)] = static_cast<uint16_t>(i);
            }
            __syncthreads();
            const uint32_t listed = list_len;
            for (uint32_t j = threadIdx.x; j < listed; j += kBlock) {
                const uint32_t i = row_list[j];
                const int64_t idx =
                    k > 0 ? find_index_key_sampled(index_keys, k, left_keys[l0 + i], key_sample, stride) : -1;
                row_match[i] = idx >= 0 ? index_starts[idx] : 0;
            }
        }
        __syncthreads();
        for (uint64_t o = o0 + threadIdx.x; o < o1; o += kBlock) {
            size_t row;
            uint64_t row_start;
            uint32_t match_start;
            if (staged) {
                uint32_t lo = 0, hi = static_cast<uint32_t>(span);
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (row_offset[mid] <= o)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                row = l0 + lo - 1;
                row_start = row_offset[lo - 1];
                match_start = row_match[lo - 1];
            } else {
                size_t lo = l0, hi = l1;
                while (lo < hi) {
                    const size_t mid = (lo + hi) >> 1;
                    if (offsets[mid] <= o)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                row = lo - 1;
                row_start = offsets[row];
                match_start = index_starts[find_index_key(index_keys, k, left_keys[row])];
            }
            const uint32_t right_row = index_rows[match_start + static_cast<uint32_t>(o - row_start)];
            out0[o] = side0 == 0 ? p0[row] : p0[right_row];
            out1[o] = side1 == 0 ? p1[row] : p1[right_row];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------
// filter

__device__ __forceinline__ bool compare_u32(uint32_t value, uint32_t op, uint32_t operand) {
    switch (op) {
        case 0: return value == operand;
        case 1: return value != operand;
        case 2: return value < operand;
        case 3: return value <= operand;
        case 4: return value > operand;
        case 5: return value >= operand;
        default: return false;
    }
}

// Filter, both passes of a two-pass compaction: tiles of kCompactTile rows;
// wave w owns rows [w*1024, w*1024+1024) of a tile, item i of lane j is row
// w*1024 + i*64 + j, and ballots rank the matches in row order. The count
// pass writes the tile's match count and adds it to its group's sum; the
// write pass writes the matching row ids at the tile's offset.
template <bool kWrite>
__global__ __launch_bounds__(kBlock) void filter_tile_kernel(const uint32_t* input, size_t n, uint32_t op,
                                                             uint32_t operand, uint32_t* out_rows, uint32_t* counts,
                                                             uint64_t* group_sums) {
    __shared__ uint32_t wave_tot[kWavesPerBlock];
    __shared__ uint64_t offset_slot;
    const uint32_t tile = blockIdx.x;
    const size_t base = static_cast<size_t>(tile) * kCompactTile;
    if (base >= n) {
        if (!kWrite && threadIdx.x == 0) counts[tile] = 0;
        return;
    }
    const uint32_t lane = lane_id(), wave = wave_id();
    const size_t w0 = base + static_cast<size_t>(wave) * (kCompactItems * kWave);
    // every load is issued before any is used (rows clamped to the input)
    uint32_t v[kCompactItems];
#pragma unroll
    for (int i = 0; i < kCompactItems; ++i) {
        const size_t row = w0 + i * kWave + lane;
        v[i] = input[row < n ? row : n - 1];
    }
    uint64_t offset = 0;
    if (kWrite && wave == 0) offset = group_tile_offset(group_sums, counts, tile);
    uint64_t masks[kCompactItems];
    uint32_t mine = 0;
#pragma unroll
    for (int i = 0; i < kCompactItems; ++i) {
        const size_t row = w0 + i * kWave + lane;
        masks[i] = ballot(row < n && compare_u32(v[i], op, operand));
        mine += static_cast<uint32_t>(__popcll(masks[i]));
    }
    if (lane == 0) wave_tot[wave] = mine;
    if (kWrite && threadIdx.x == 0) offset_slot = offset;
    __syncthreads();
    if (!kWrite) {
        if (threadIdx.x == 0) {
            uint32_t c = 0;
            for (int w = 0; w < kWavesPerBlock; ++w) c += wave_tot[w];
            counts[tile] = c;
            atomicAdd(reinterpret_cast<unsigned long long*>(&group_sums[tile / kTileGroup]),
                      static_cast<unsigned long long>(c));
        }
        return;
    }
    uint64_t pos = offset_slot;
    for (int w = 0; w < static_cast<int>(wave); ++w) pos += wave_tot[w];
#pragma unroll
    for (int i = 0; i < kCompactItems; ++i) {
        const size_t row = w0 + i * kWave + lane;
        if ((masks[i] >> lane) & 1ull) out_rows[pos + lanes_below(masks[i])] = static_cast<uint32_t>(row);
        pos += static_cast<uint32_t>(__popcll(masks[i]));
    }
}

// ------------------------------------------------------------------------
// small kernels

__global__ void add_one_kernel(int32_t* data, size_t len) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < len) data[i] += 1;
}

__global__ void fill_mod_kernel(uint32_t* out, size_t len, uint32_t modulus) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += stride)
        out[i] = static_cast<uint32_t>(i % modulus);
}

__global__ void fill_u64_kernel(uint64_t* out, size_t len, uint64_t value) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += stride) out[i] = value;
}

// Plain device copy: four 16-byte loads in flight per lane, then the stores.
// One 16-byte vector per thread, one contiguous 4 KiB chunk per block, with
// non-temporal loads and stores. On MI355X this measured 6.55 TB/s for a
// 1 GiB copy (82% of the 8 TB/s peak), against 4.5-5.1 TB/s for grid-stride
// loops over 1,024..65,536 blocks (with or without the hint) and 5.1 TB/s for
// hipMemcpyAsync: blocks dispatched in address order keep the active window
// of the copy compact.
typedef unsigned int copy_vec __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(kBlock) void copy_u64_kernel(const uint64_t* __restrict__ src, uint64_t* __restrict__ dst,
                                                          size_t count) {
    const copy_vec* s = reinterpret_cast<const copy_vec*>(src);
    copy_vec* d = reinterpret_cast<copy_vec*>(dst);
    const size_t n16 = count / 2;
    const size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n16) __builtin_nontemporal_store(__builtin_nontemporal_load(&s[i]), &d[i]);
    if ((count & 1) && i == 0) dst[count - 1] = src[count - 1];
}

unsigned grid_for(size_t n, size_t per_block, unsigned cap) {
    const size_t b = tiles_of(n, per_block);
    return static_cast<unsigned>(b == 0 ? 1 : (b < cap ? b : cap));
}

// Scratch of single-pass compaction kernels: header + one status word per tile.
size_t compact_temp(size_t tiles) { return kTempHeader + align_up(tiles * sizeof(uint64_t), 256); }

hipError_t prepare_compact(void* temp, size_t tiles, hipStream_t stream) {
    clear_header_kernel<<<1, 64, 0, stream>>>(static_cast<Header*>(temp));
    return clear_u64(reinterpret_cast<uint64_t*>(static_cast<char*>(temp) + kTempHeader), tiles, stream);
}

}  // namespace

// ========================================================================
// launchers

hipError_t report_error(const uint32_t* error, uint32_t* count32, uint64_t* count64, hipStream_t stream) {
    if (error_sink()) return hipSuccess;  // the sink's owner checks it
    report_error_kernel<<<1, 1, 0, stream>>>(error, count32, count64);
    return hipGetLastError();
}

size_t distinct_temp_bytes(size_t rows) {
    const SortLayout l = sort_layout(rows);
    return kTempHeader + l.total;
}

hipError_t radix_sort(PairInput in, size_t rows, uint64_t* keys_a, uint64_t* keys_b, void* temp,
                      size_t temp_bytes, uint64_t known_varying, hipStream_t stream, uint64_t** sorted_host) {
    const SortLayout l = sort_layout(rows);
    if (temp_bytes < kTempHeader + l.total) return hipErrorInvalidValue;
    if (rows == 0) return hipSuccess;
    if (in.packed && in.packed == keys_b) return hipErrorInvalidValue;
    char* sort_temp = static_cast<char*>(temp) + kTempHeader;
    hipError_t st = run_sort(in, rows, keys_a, keys_b, sort_temp, l, known_varying, stream);
    if (sorted_host && known_varying) {
        int active = 0;
        for (int p = 0; p < kPasses; ++p)
            if ((known_varying >> (p * kRadixBits)) & (kRadix - 1)) ++active;
        *sorted_host = (active & 1) ? keys_b : keys_a;
    }
    return st;
}

hipError_t distinct(PairInput in, size_t rows, uint64_t* keys_a, uint64_t* keys_b, PairOutput out,
                    uint64_t* unique_rows, void* temp, size_t temp_bytes, uint64_t known_varying,
                    hipStream_t stream) {
    if (rows == 0) {
        set_u64_kernel<<<1, 1, 0, stream>>>(unique_rows, 0);
        return hipGetLastError();
    }
    const SortLayout l = sort_layout(rows);
    if (temp_bytes < kTempHeader + l.total) return hipErrorInvalidValue;
    if (in.packed && in.packed == keys_b) return hipErrorInvalidValue;
    char* sort_temp = static_cast<char*>(temp) + kTempHeader;
    hipError_t st = run_sort(in, rows, keys_a, keys_b, sort_temp, l, known_varying, stream);
    if (st != hipSuccess) return st;
    uint32_t* ucounts = reinterpret_cast<uint32_t*>(sort_temp + l.unique_counts_off);
    uint64_t* ugroups = reinterpret_cast<uint64_t*>(sort_temp + l.unique_groups_off);
    st = clear_u64(ugroups, l.unique_groups, stream);
    if (st != hipSuccess) return st;
    const SortDesc* desc = reinterpret_cast<const SortDesc*>(sort_temp + l.desc_off);
    const unsigned grid = static_cast<unsigned>(l.unique_tiles);
    unique_tile_kernel<false><<<grid, kBlock, 0, stream>>>(keys_a, keys_b, desc, rows, nullptr, ucounts, ugroups, out);
    group_scan_kernel<<<1, 1024, 0, stream>>>(ugroups, l.unique_groups, unique_rows, nullptr);
    unique_tile_kernel<true><<<grid, kBlock, 0, stream>>>(keys_a, keys_b, desc, rows, nullptr, ucounts, ugroups, out);
    return hipGetLastError();
}

hipError_t unique_sorted(const uint64_t* keys, const uint64_t* rows_dev, size_t bound, PairOutput out,
                         uint64_t* unique_rows, void* temp, size_t temp_bytes, hipStream_t stream) {
    if (bound == 0) return set_u64(unique_rows, 0, stream);
    const SortLayout l = sort_layout(bound);
    if (temp_bytes < kTempHeader + l.total) return hipErrorInvalidValue;
    char* sort_temp = static_cast<char*>(temp) + kTempHeader;
    uint32_t* ucounts = reinterpret_cast<uint32_t*>(sort_temp + l.unique_counts_off);
    uint64_t* ugroups = reinterpret_cast<uint64_t*>(sort_temp + l.unique_groups_off);
    hipError_t st = clear_u64(ugroups, l.unique_groups, stream);
    if (st != hipSuccess) return st;
    const unsigned grid = static_cast<unsigned>(l.unique_tiles);
    unique_tile_kernel<false><<<grid, kBlock, 0, stream>>>(keys, keys, nullptr, bound, rows_dev, ucounts, ugroups, out);
    group_scan_kernel<<<1, 1024, 0, stream>>>(ugroups, l.unique_groups, unique_rows, nullptr);
    unique_tile_kernel<true><<<grid, kBlock, 0, stream>>>(keys, keys, nullptr, bound, rows_dev, ucounts, ugroups, out);
    return hipGetLastError();
}

// Scratch of the union: header, merge-path splits, per-tile kept counts and
// tile-group sums.
struct UnionLayout {
    size_t tiles, groups, part_off, coarse_off, counts_off, groups_off, total;
};

UnionLayout union_layout(size_t left_rows, size_t right_rows) {
    UnionLayout l{};
    l.tiles = tiles_of(left_rows + right_rows, kMergeTile);
    l.groups = tiles_of(l.tiles, kTileGroup);
    l.part_off = kTempHeader;
    l.coarse_off = l.part_off + align_up((l.tiles + 1) * sizeof(size_t), 256);
    l.counts_off = l.coarse_off + align_up(coarse_parts(l.tiles) * sizeof(size_t), 256);
    l.groups_off = l.counts_off + align_up(l.tiles * sizeof(uint32_t), 256);
    l.total = l.groups_off + align_up((l.groups ? l.groups : 1) * sizeof(uint64_t), 256);
    return l;
}

size_t union_temp_bytes(size_t left_rows, size_t right_rows) { return union_layout(left_rows, right_rows).total; }

hipError_t sorted_union(PairInput left, size_t left_rows, PairInput right, size_t right_rows, PairOutput out,
                        uint64_t* unique_rows, void* temp, size_t temp_bytes, hipStream_t stream,
                        const KernelTimer* timer) {
    const size_t total = left_rows + right_rows;
    if (total == 0) {
        set_u64_kernel<<<1, 1, 0, stream>>>(unique_rows, 0);
        return hipGetLastError();
    }
    const UnionLayout l = union_layout(left_rows, right_rows);
    if (temp_bytes < l.total) return hipErrorInvalidValue;
    if (l.tiles > 0xFFFFFFFFull) return hipErrorInvalidValue;
    char* base = static_cast<char*>(temp);
    size_t* part = reinterpret_cast<size_t*>(base + l.part_off);
    uint32_t* counts = reinterpret_cast<uint32_t*>(base + l.counts_off);
    uint64_t* groups = reinterpret_cast<uint64_t*>(base + l.groups_off);
    merge_partition(left, left_rows, right, right_rows, part, l.tiles, reinterpret_cast<size_t*>(base + l.coarse_off),
                    stream, groups, l.groups);
    timer_start(timer, stream);
    union_tile_kernel<false><<<static_cast<unsigned>(l.tiles), kBlock, 0, stream>>>(left, left_rows, right, right_rows,
                                                                                   part, counts, groups, out);
    group_scan_kernel<<<1, 1024, 0, stream>>>(groups, l.groups, unique_rows, nullptr);
    union_tile_kernel<true><<<static_cast<unsigned>(l.tiles), kBlock, 0, stream>>>(left, left_rows, right, right_rows,
                                                                                  part, counts, groups, out);
    timer_stop(timer, stream);
    return hipGetLastError();
}

size_t merge_disjoint_temp_bytes(size_t left_rows, size_t right_rows) {
    const size_t tiles = tiles_of(left_rows + right_rows, kMergeTile);
    return align_up((tiles + 1) * sizeof(size_t), 256) + align_up(coarse_parts(tiles) * sizeof(size_t), 256);
}

hipError_t merge_disjoint(const uint64_t* left, size_t left_rows, const uint64_t* right, size_t right_rows,
                          uint64_t* out, void* temp, size_t temp_bytes, hipStream_t stream, const KernelTimer* timer,
                          uint32_t* out_second) {
    const size_t total = left_rows + right_rows;
    if (total == 0) return hipSuccess;
    if (temp_bytes < merge_disjoint_temp_bytes(left_rows, right_rows)) return hipErrorInvalidValue;
    const size_t tiles = tiles_of(total, kMergeTile);
    size_t* part = static_cast<size_t*>(temp);
    const PairInput a{left, nullptr, nullptr, nullptr}, b{right, nullptr, nullptr, nullptr};
    merge_partition(a, left_rows, b, right_rows, part, tiles,
                    reinterpret_cast<size_t*>(static_cast<char*>(temp) + align_up((tiles + 1) * sizeof(size_t), 256)),
                    stream);
    timer_start(timer, stream);
    merge_disjoint_kernel<<<static_cast<unsigned>(tiles), kBlock, 0, stream>>>(left, left_rows, right, right_rows,
                                                                               out, out_second, part);
    timer_stop(timer, stream);
    return hipGetLastError();
}

// Scratch of the anti-join: header, right ranges per (tile, right input),
// found bits, per-tile kept counts and tile-group sums.
struct AntiLayout {
    size_t tiles, groups, ranges_off, found_off, kept_off, groups_off, total;
};

AntiLayout anti_layout(size_t left_rows, uint32_t right_inputs) {
    AntiLayout l{};
    l.tiles = tiles_of(left_rows, kAntiTile);
    l.groups = tiles_of(l.tiles, kTileGroup);
    l.ranges_off = kTempHeader;
    l.found_off = l.ranges_off + align_up(l.tiles * (right_inputs ? right_inputs : 1) * 2 * sizeof(size_t), 256);
    l.kept_off = l.found_off + align_up(l.tiles * (kAntiTile / kWave) * sizeof(uint64_t), 256);
    l.groups_off = l.kept_off + align_up(l.tiles * sizeof(uint32_t), 256);
    l.total = l.groups_off + align_up((l.groups ? l.groups : 1) * sizeof(uint64_t), 256);
    return l;
}

size_t anti_join_temp_bytes(size_t left_rows, uint32_t right_inputs) {
    return anti_layout(left_rows, right_inputs).total;
}

hipError_t sorted_anti_join_multi(PairInput left, size_t left_rows, const AntiRight& rights, PairOutput out,
                                  uint32_t* count32, uint64_t* count64, void* temp, size_t temp_bytes,
                                  hipStream_t stream, const KernelTimer* timer) {
    if (left_rows == 0) {
        if (count32) set_u32_kernel<<<1, 1, 0, stream>>>(count32, 0);
        if (count64) set_u64_kernel<<<1, 1, 0, stream>>>(count64, 0);
        return hipGetLastError();
    }
    if (rights.count > static_cast<uint32_t>(kMaxAntiRights)) return hipErrorInvalidValue;
    const AntiLayout l = anti_layout(left_rows, rights.count);
    if (temp_bytes < l.total) return hipErrorInvalidValue;
    if (l.tiles > 0xFFFFFFFFull) return hipErrorInvalidValue;
    char* base = static_cast<char*>(temp);
    size_t* ranges = reinterpret_cast<size_t*>(base + l.ranges_off);
    uint64_t* found_words = reinterpret_cast<uint64_t*>(base + l.found_off);
    uint32_t* kept = reinterpret_cast<uint32_t*>(base + l.kept_off);
    uint64_t* groups = reinterpret_cast<uint64_t*>(base + l.groups_off);
    hipError_t st = clear_u64(groups, l.groups, stream);
    if (st != hipSuccess) return st;
    if (rights.count) {
        anti_partition_kernel<<<grid_for(l.tiles * rights.count, kBlock, 4096), kBlock, 0, stream>>>(
            left, left_rows, rights, l.tiles, ranges, 0u);
    }
    timer_start(timer, stream);
    anti_join_kernel<<<static_cast<unsigned>(l.tiles), kBlock, 0, stream>>>(left, left_rows, rights, ranges,
                                                                           found_words, 0u, kept, groups);
    timer_stop(timer, stream);
    group_scan_kernel<<<1, 1024, 0, stream>>>(groups, l.groups, count64, count32);
    anti_write_kernel<<<static_cast<unsigned>(l.tiles), kBlock, 0, stream>>>(left, left_rows, found_words, kept,
                                                                            groups, out);
    return hipGetLastError();
}

size_t probe_grouped_temp_bytes(size_t rows, uint32_t right_inputs) {
    return align_up(tiles_of(rows, kAntiTile) * (right_inputs ? right_inputs : 1) * 2 * sizeof(size_t), 256);
}

size_t probe_found_words(size_t rows) { return tiles_of(rows, kAntiTile) * (kAntiTile / kWave); }

hipError_t probe_grouped(const uint64_t* keys, const uint64_t* rows_dev, size_t bound, uint32_t minor_bits,
                         const AntiRight& rights, uint64_t* found_words, void* temp, size_t temp_bytes,
                         hipStream_t stream, const KernelTimer* timer) {
    if (bound == 0) return hipSuccess;
    if (minor_bits == 0 || minor_bits > 32 || rights.count > static_cast<uint32_t>(kMaxAntiRights) ||
        temp_bytes < probe_grouped_temp_bytes(bound, rights.count))
        return hipErrorInvalidValue;
    const size_t tiles = tiles_of(bound, kAntiTile);
    const PairInput left{keys, nullptr, nullptr, rows_dev};
    size_t* ranges = static_cast<size_t*>(temp);
    if (rights.count) {
        anti_partition_kernel<<<grid_for(tiles * rights.count, kBlock, 4096), kBlock, 0, stream>>>(
            left, bound, rights, tiles, ranges, minor_bits);
    }
    timer_start(timer, stream);
    anti_join_kernel<<<static_cast<unsigned>(tiles), kBlock, 0, stream>>>(left, bound, rights, ranges, found_words,
                                                                         minor_bits);
    timer_stop(timer, stream);
    return hipGetLastError();
}

hipError_t sorted_anti_join(PairInput left, size_t left_rows, PairInput right, size_t right_rows, PairOutput out,
                            uint32_t* count32, uint64_t* count64, void* temp, size_t temp_bytes,
                            hipStream_t stream) {
    AntiRight rights{};
    rights.count = right_rows ? 1 : 0;
    rights.in[0] = right;
    rights.rows[0] = right_rows;
    return sorted_anti_join_multi(left, left_rows, rights, out, count32, count64, temp, temp_bytes, stream);
}

size_t join_temp_bytes(size_t left_rows) { return compact_temp(tiles_of(left_rows, kBlock * 8)); }

hipError_t join_count(const uint32_t* left_keys, size_t left_rows, const uint32_t* index_keys,
                      const uint32_t* index_starts, size_t unique_keys, uint64_t* counts, uint64_t* offsets,
                      uint64_t* total, void* temp, size_t temp_bytes, hipStream_t stream) {
    if (left_rows == 0) {
        set_u64_kernel<<<1, 1, 0, stream>>>(total, 0);
        return hipGetLastError();
    }
    const size_t tiles = tiles_of(left_rows, kBlock * 8);
    if (temp_bytes < compact_temp(tiles)) return hipErrorInvalidValue;
    hipError_t st = prepare_compact(temp, tiles, stream);
    if (st != hipSuccess) return st;
    uint32_t* error = error_word(static_cast<Header*>(temp));
    join_count_kernel<<<static_cast<unsigned>(tiles), kBlock, 0, stream>>>(
        left_keys, left_rows, index_keys, index_starts, unique_keys, counts, offsets, total,
        reinterpret_cast<uint64_t*>(static_cast<char*>(temp) + kTempHeader), static_cast<Header*>(temp), error);
    st = hipGetLastError();
    if (st != hipSuccess) return st;
    return report_error(error, nullptr, total, stream);
}

hipError_t join_emit(const uint32_t* left_keys, size_t left_rows, const uint32_t* index_keys,
                     const uint32_t* index_starts, const uint32_t* index_rows, size_t unique_keys,
                     const uint32_t* projection0, uint32_t side0, const uint32_t* projection1, uint32_t side1,
                     const uint64_t* offsets, uint32_t* out0, uint32_t* out1, hipStream_t stream) {
    if (left_rows == 0) return hipSuccess;
    join_emit_kernel<<<2048, kBlock, 0, stream>>>(left_keys, left_rows, index_keys, index_starts, index_rows,
                                                 unique_keys, projection0, side0, projection1, side1, offsets,
                                                 out0, out1);
    return hipGetLastError();
}

// Scratch of the filter: header, per-tile match counts and tile-group sums.
size_t filter_counts_off() { return kTempHeader; }
size_t filter_groups_off(size_t tiles) { return kTempHeader + align_up(tiles * sizeof(uint32_t), 256); }

size_t filter_temp_bytes(size_t rows) {
    const size_t tiles = tiles_of(rows, kCompactTile);
    return filter_groups_off(tiles) + align_up((tiles_of(tiles, kTileGroup) + 1) * sizeof(uint64_t), 256);
}

hipError_t filter(const uint32_t* input, size_t rows, uint32_t op, uint32_t operand, uint32_t* out_rows,
                  uint32_t* count, void* temp, size_t temp_bytes, hipStream_t stream) {
    if (rows == 0) {
        set_u32_kernel<<<1, 1, 0, stream>>>(count, 0);
        return hipGetLastError();
    }
    const size_t tiles = tiles_of(rows, kCompactTile);
    if (temp_bytes < filter_temp_bytes(rows) || tiles > 0xFFFFFFFFull) return hipErrorInvalidValue;
    char* base = static_cast<char*>(temp);
    uint32_t* counts = reinterpret_cast<uint32_t*>(base + filter_counts_off());
    uint64_t* groups = reinterpret_cast<uint64_t*>(base + filter_groups_off(tiles));
    const size_t n_groups = tiles_of(tiles, kTileGroup);
    hipError_t st = clear_u64(groups, n_groups, stream);
    if (st != hipSuccess) return st;
    const unsigned grid = static_cast<unsigned>(tiles);
    filter_tile_kernel<false><<<grid, kBlock, 0, stream>>>(input, rows, op, operand, out_rows, counts, groups);
    group_scan_kernel<<<1, 1024, 0, stream>>>(groups, n_groups, nullptr, count);
    filter_tile_kernel<true><<<grid, kBlock, 0, stream>>>(input, rows, op, operand, out_rows, counts, groups);
    return hipGetLastError();
}

hipError_t add_one_i32(int32_t* data, size_t len, hipStream_t stream) {
    if (len == 0) return hipSuccess;
    add_one_kernel<<<static_cast<unsigned>(tiles_of(len, kBlock)), kBlock, 0, stream>>>(data, len);
    return hipGetLastError();
}

hipError_t fill_mod_u32(uint32_t* out, size_t len, uint32_t modulus, hipStream_t stream) {
    if (len == 0) return hipSuccess;
    fill_mod_kernel<<<grid_for(len, kBlock * 4, 8192), kBlock, 0, stream>>>(out, len, modulus);
    return hipGetLastError();
}

hipError_t fill_u64(uint64_t* out, size_t len, uint64_t value, hipStream_t stream) {
    if (len == 0) return hipSuccess;
    fill_u64_kernel<<<grid_for(len, kBlock * 4, 8192), kBlock, 0, stream>>>(out, len, value);
    return hipGetLastError();
}

hipError_t copy_u64(const uint64_t* src, uint64_t* dst, size_t count, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const size_t blocks = tiles_of(count / 2 + 1, kBlock);
    if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    copy_u64_kernel<<<static_cast<unsigned>(blocks), kBlock, 0, stream>>>(src, dst, count);
    return hipGetLastError();
}

hipError_t set_u32(uint32_t* out, uint32_t value, hipStream_t stream) {
    set_u32_kernel<<<1, 1, 0, stream>>>(out, value);
    return hipGetLastError();
}

hipError_t set_u64(uint64_t* out, uint64_t value, hipStream_t stream) {
    set_u64_kernel<<<1, 1, 0, stream>>>(out, value);
    return hipGetLastError();
}

}  // namespace spk
