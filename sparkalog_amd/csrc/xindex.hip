// First-column index of sorted packed relations, and the anti-join that uses
// it.
//
// FULL is kept as a few sorted runs of packed pairs (first << 32 | second).
// For a run R and every first value x in [0, n_x], xstart[x] is the number
// of rows of R whose first column is below x, so the rows of R with first
// column x are R[xstart[x], xstart[x+1]). Merging two runs adds their
// indexes element-wise.
//
// The anti-join `left \ (R_0 ∪ R_1 ∪ ...)` (SortedBinaryAntiJoin,
// crates/sparkalog-execution/src/lib.rs:672-710, with FULL as right input)
// then reads, for a tile of left rows, only the segments of the first values
// present in the tile: no search, and none of the rows of first values
// without left rows. Each tile
//   1. loads its rows (sorted, unique) into LDS and lists its first-value
//      groups,
//   2. per run: looks up each group's segment, scans the segment lengths,
//      and streams the concatenated segments with four rows in flight per
//      lane, marking every left row found (binary search within its group),
//   3. compacts the unmarked rows in order at an offset from decoupled
//      look-back, writing them to both outputs (NEWT and FULL's new run).
#include "ops.hpp"
#include "prims.hpp"

namespace spk {
namespace {

constexpr int kXTile = 1024;                 // left rows per anti-join tile
constexpr int kXItems = kXTile / kBlock;     // 4 rows per thread
constexpr int kXWaveRows = kXItems * kWave;  // 256 rows per wave
constexpr int kXUnroll = 8;                  // segment rows in flight per lane
constexpr int kXHashBits = 11;               // hash set of a tile: 2048 slots (load <= 1/2)
constexpr int kXHashSlots = 1 << kXHashBits;
static_assert(kXHashSlots >= 2 * kXTile, "hash set load factor");
constexpr uint32_t kXPairCap = 1024;         // (run, group) segments looked up per batch
constexpr int kXPairItems = kXPairCap / kBlock;
constexpr int kCTile = 2048;                 // rows per compaction tile
constexpr int kCItems = kCTile / kBlock;
constexpr int kCWaveRows = kCItems * kWave;
constexpr uint32_t kXBuildSpan = 256;        // first values per index-build block

struct XHeader {
    uint32_t error;
    uint32_t counter;
};

__device__ __forceinline__ size_t rows_of(const uint64_t* rows_dev, size_t bound) {
    if (!rows_dev) return bound;
    const uint64_t n = *rows_dev;
    return n < bound ? n : bound;
}

// First row with first column >= x (rows sorted), by block-cooperative search.
__device__ __forceinline__ size_t block_first_at_least(const uint64_t* keys, size_t n, uint64_t x) {
    return block_search(0, n, [&](size_t i) { return (keys[i] >> 32) < x; });
}

// One block per kXBuildSpan first values: mark where each value's rows start,
// then xstart[x] = the smallest start at or after x (a suffix minimum).
__global__ __launch_bounds__(kBlock) void xindex_build_kernel(const uint64_t* keys, const uint64_t* rows_dev,
                                                              size_t bound, uint32_t n_x, uint32_t* xstart) {
    __shared__ uint32_t first_row[kXBuildSpan];
    __shared__ uint32_t scan_lds[kWavesPerBlock];
    const size_t n = rows_of(rows_dev, bound);
    const uint64_t xb = static_cast<uint64_t>(blockIdx.x) * kXBuildSpan;
    if (xb > n_x) return;  // block-uniform
    const uint64_t xe = xb + kXBuildSpan < static_cast<uint64_t>(n_x) + 1 ? xb + kXBuildSpan
                                                                          : static_cast<uint64_t>(n_x) + 1;
    const size_t r0 = block_first_at_least(keys, n, xb);
    const size_t r1 = block_first_at_least(keys, n, xe);
    for (uint32_t j = threadIdx.x; j < kXBuildSpan; j += kBlock) first_row[j] = static_cast<uint32_t>(r1);
    __syncthreads();
    // rows [r0, r1) have first values in [xb, xe); a row starts its value's
    // group when the row before it has another first value
    for (size_t c = r0; c < r1; c += static_cast<size_t>(kBlock) * kXUnroll) {
        uint64_t cur[kXUnroll], prev[kXUnroll];
#pragma unroll
        for (int u = 0; u < kXUnroll; ++u) {
            const size_t i = c + static_cast<size_t>(u) * kBlock + threadIdx.x;
            cur[u] = i < r1 ? keys[i] : 0;
            prev[u] = (i < r1 && i > r0) ? keys[i - 1] : 0;
        }
#pragma unroll
        for (int u = 0; u < kXUnroll; ++u) {
            const size_t i = c + static_cast<size_t>(u) * kBlock + threadIdx.x;
            if (i < r1 && (i == r0 || (prev[u] >> 32) != (cur[u] >> 32)))
                first_row[static_cast<uint32_t>((cur[u] >> 32) - xb)] = static_cast<uint32_t>(i);
        }
    }
    __syncthreads();
    // suffix minimum over the span: thread t owns 16 consecutive values
    constexpr int kPer = kXBuildSpan / kBlock;
    const uint32_t t = threadIdx.x;
    uint32_t v[kPer];
    uint32_t m = static_cast<uint32_t>(r1);
#pragma unroll
    for (int j = kPer - 1; j >= 0; --j) {
        const uint32_t f = first_row[t * kPer + j];
        m = f < m ? f : m;
        v[j] = m;
    }
    // minimum over later threads: inclusive scan of min over reversed order
    uint32_t later = m;
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t o = __shfl_down(later, off, kWave);
        if (lane_id() + off < static_cast<uint32_t>(kWave)) later = o < later ? o : later;
    }
    // `later` = min over lanes >= this lane of this wave
    if (lane_id() == 0) scan_lds[wave_id()] = later;
    __syncthreads();
    uint32_t after_wave = static_cast<uint32_t>(r1);
    for (int w = static_cast<int>(wave_id()) + 1; w < kWavesPerBlock; ++w)
        after_wave = scan_lds[w] < after_wave ? scan_lds[w] : after_wave;
    const uint32_t next_lane = __shfl_down(later, 1, kWave);
    uint32_t after = lane_id() + 1 < static_cast<uint32_t>(kWave) ? next_lane : static_cast<uint32_t>(r1);
    after = after_wave < after ? after_wave : after;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const uint64_t x = xb + t * kPer + j;
        if (x < xe) xstart[x] = v[j] < after ? v[j] : after;
    }
}

__global__ void xindex_add_kernel(uint32_t* into, const uint32_t* other, size_t n) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        into[i] += other[i];
}

__global__ void xanti_init_kernel(XHeader* h, uint64_t* status, size_t tiles) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    const size_t gtid = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gtid == 0) {
        h->error = 0;
        h->counter = 0;
    }
    for (size_t i = gtid; i < tiles; i += stride) status[i] = 0;
}

// Largest j in [j, last] with pre[j] <= t (pre non-decreasing, pre[j] <= t).
__device__ __forceinline__ uint32_t advance_pair(const uint64_t* pre, uint32_t j, uint32_t last, uint64_t t) {
    if (j + 1 > last || pre[j + 1] > t) return j;
    uint32_t lo = j + 1, hi = last;  // answer in [lo, hi]
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Slot of a packed pair in the tile's hash set (multiplicative mixing of
// both halves; top bits of the product).
__device__ __forceinline__ uint32_t pair_slot(uint64_t key) {
    const uint32_t a = static_cast<uint32_t>(key), b = static_cast<uint32_t>(key >> 32);
    const uint32_t h = (a ^ (b * 0x9E3779B9u)) * 0x85EBCA6Bu;
    return h >> (32 - kXHashBits);
}

// Marks, for every left row of a tile, whether FULL holds it: bit r of
// found_bits[tile * kXTile / 32 + r / 32]. Tiles are independent (their work
// varies with the segment sizes), so no tile waits for another.
__global__ __launch_bounds__(kBlock) void anti_join_x_kernel(const uint64_t* left, const uint64_t* left_rows,
                                                             size_t bound, XRights rights, uint32_t n_x,
                                                             uint32_t* found_bits) {
    __shared__ uint64_t tkeys[kXTile];
    __shared__ uint32_t table[kXHashSlots];  // tile row + 1, or 0 for an empty slot
    __shared__ uint32_t grp_x[kXTile];           // first value of each group of equal first values
    __shared__ uint64_t pair_pre[kXPairCap + 1];  // exclusive prefix of the batch's segment lengths
    __shared__ uint32_t pair_beg[kXPairCap];      // segment start row within its run
    __shared__ uint8_t pair_run[kXPairCap];       // the segment's run
    __shared__ const uint64_t* run_keys[kMaxAntiRights];
    __shared__ uint8_t found[kXTile];
    __shared__ uint32_t item_cnt[kWavesPerBlock][kXItems];
    __shared__ uint64_t scan_lds[kWavesPerBlock];
    const size_t nl = rows_of(left_rows, bound);
    const uint32_t tile = blockIdx.x;
    const size_t base = static_cast<size_t>(tile) * kXTile;
    if (base >= nl) return;
    const uint32_t lane = lane_id(), wave = wave_id();
    const uint32_t rows = static_cast<uint32_t>(nl - base < static_cast<size_t>(kXTile) ? nl - base : kXTile);
    uint64_t k[kXItems];
#pragma unroll
    for (int i = 0; i < kXItems; ++i) {
        const uint32_t r = wave * kXWaveRows + i * kWave + lane;
        k[i] = r < rows ? left[base + r] : ~0ull;
        if (r < rows) {
            tkeys[r] = k[i];
            found[r] = 0;
        }
    }
    for (uint32_t j = threadIdx.x; j < kXHashSlots; j += kBlock) table[j] = 0;
    if (threadIdx.x < rights.count) run_keys[threadIdx.x] = rights.keys[threadIdx.x];
    __syncthreads();
    // hash set of the tile's rows (unique keys: every insert takes a new slot)
    uint64_t starts[kXItems];
#pragma unroll
    for (int i = 0; i < kXItems; ++i) {
        const uint32_t r = wave * kXWaveRows + i * kWave + lane;
        if (r < rows) {
            uint32_t slot = pair_slot(k[i]);
            while (atomicCAS(&table[slot], 0u, r + 1) != 0u) slot = (slot + 1) & (kXHashSlots - 1);
        }
        // groups of equal first value, in tile order
        const bool st = r < rows && (r == 0 || (tkeys[r - 1] >> 32) != (k[i] >> 32));
        starts[i] = ballot(st);
        if (lane == 0) item_cnt[wave][i] = static_cast<uint32_t>(__popcll(starts[i]));
    }
    __syncthreads();
    uint32_t pos = 0, groups = 0;
    for (int w = 0; w < kWavesPerBlock; ++w)
        for (int i = 0; i < kXItems; ++i) {
            if (w < static_cast<int>(wave)) pos += item_cnt[w][i];
            groups += item_cnt[w][i];
        }
#pragma unroll
    for (int i = 0; i < kXItems; ++i) {
        if ((starts[i] >> lane) & 1ull) grp_x[pos + lanes_below(starts[i])] = static_cast<uint32_t>(k[i] >> 32);
        pos += static_cast<uint32_t>(__popcll(starts[i]));
    }
    __syncthreads();

    // (run, group) pairs in batches of whole runs: each batch looks up all
    // its segments at once, scans their lengths and streams them
    const uint32_t runs_per_batch = groups <= kXPairCap ? kXPairCap / groups : 1;
    for (uint32_t q0 = 0; q0 < rights.count; q0 += runs_per_batch) {
        const uint32_t q1 = q0 + runs_per_batch < rights.count ? q0 + runs_per_batch : rights.count;
        const uint32_t pairs = (q1 - q0) * groups;
        // a batch with more pairs than LDS slots (one run of > kXPairCap
        // groups) is cut into pieces of kXPairCap groups
        for (uint32_t p0 = 0; p0 < pairs; p0 += kXPairCap) {
            const uint32_t np = pairs - p0 < kXPairCap ? pairs - p0 : kXPairCap;
            uint32_t beg[kXPairItems], len[kXPairItems];
            uint64_t sum = 0;
#pragma unroll
            for (int g = 0; g < kXPairItems; ++g) {
                const uint32_t pi = threadIdx.x * kXPairItems + g;
                beg[g] = len[g] = 0;
                if (pi < np) {
                    const uint32_t pair = p0 + pi;
                    const uint32_t q = q0 + pair / groups, j = pair % groups;
                    const uint32_t x = grp_x[j];
                    if (x < n_x) {
                        const uint32_t* xs = rights.xstart[q];
                        beg[g] = xs[x];
                        len[g] = xs[x + 1] - beg[g];
                    }
                    pair_run[pi] = static_cast<uint8_t>(q);
                }
                sum += len[g];
            }
            uint64_t total = 0;
            uint64_t run = block_exclusive_sum<uint64_t>(sum, &total, scan_lds);
#pragma unroll
            for (int g = 0; g < kXPairItems; ++g) {
                const uint32_t pi = threadIdx.x * kXPairItems + g;
                if (pi < np) {
                    pair_pre[pi] = run;
                    pair_beg[pi] = beg[g];
                }
                run += len[g];
            }
            if (threadIdx.x == 0) pair_pre[np] = total;
            __syncthreads();
            // stream the concatenated segments, kXUnroll rows in flight per lane
            uint32_t pj = 0;
            const uint64_t* rk = run_keys[pair_run[0]];
            uint64_t row_off = pair_beg[0];  // row = row_off + t - pair_pre[pj]
            uint64_t pre = 0;
            for (uint64_t c = 0; c < total; c += kBlock * kXUnroll) {
                uint64_t row[kXUnroll];
#pragma unroll
                for (int u = 0; u < kXUnroll; ++u) {
                    const uint64_t t = c + u * kBlock + threadIdx.x;
                    row[u] = 0;
                    if (t < total) {
                        const uint32_t nj = advance_pair(pair_pre, pj, np - 1, t);
                        if (nj != pj) {
                            pj = nj;
                            rk = run_keys[pair_run[pj]];
                            row_off = pair_beg[pj];
                            pre = pair_pre[pj];
                        }
                        row[u] = rk[row_off + (t - pre)];
                    }
                }
#pragma unroll
                for (int u = 0; u < kXUnroll; ++u) {
                    const uint64_t t = c + u * kBlock + threadIdx.x;
                    if (t >= total) continue;
                    uint32_t slot = pair_slot(row[u]);
                    for (;;) {
                        const uint32_t v = table[slot];
                        if (v == 0) break;
                        if (tkeys[v - 1] == row[u]) {
                            found[v - 1] = 1;
                            break;
                        }
                        slot = (slot + 1) & (kXHashSlots - 1);
                    }
                }
            }
            __syncthreads();
        }
    }

    // found bits of the tile, one word per 32 rows
    if (threadIdx.x < kXTile / 32) {
        uint32_t w = 0;
        for (int b = 0; b < 32; ++b) {
            const uint32_t r = threadIdx.x * 32 + b;
            if (r < rows && found[r]) w |= 1u << b;
        }
        found_bits[static_cast<size_t>(tile) * (kXTile / 32) + threadIdx.x] = w;
    }
}

// Order-preserving compaction of the left rows whose found bit is clear,
// written to both outputs at offsets from decoupled look-back (uniform work
// per tile).
__global__ __launch_bounds__(kBlock) void xanti_compact_kernel(const uint64_t* left, const uint64_t* left_rows,
                                                               size_t bound, const uint32_t* found_bits,
                                                               PairOutput out, uint64_t* count64, uint64_t* status,
                                                               XHeader* h, uint32_t* error) {
    __shared__ uint32_t item_cnt[kWavesPerBlock][kCItems];
    __shared__ uint64_t offset_slot;
    __shared__ uint32_t tile_slot;
    const size_t nl = rows_of(left_rows, bound);
    const uint32_t tile = next_tile(&h->counter, &tile_slot);
    const size_t base = static_cast<size_t>(tile) * kCTile;
    if (nl == 0) {
        if (tile == 0 && threadIdx.x == 0) *count64 = 0;
        return;
    }
    if (base >= nl) return;
    const uint32_t lane = lane_id(), wave = wave_id();
    const uint32_t rows = static_cast<uint32_t>(nl - base < static_cast<size_t>(kCTile) ? nl - base : kCTile);
    uint64_t k[kCItems], keep[kCItems];
#pragma unroll
    for (int i = 0; i < kCItems; ++i) {
        const uint32_t r = wave * kCWaveRows + i * kWave + lane;
        const size_t g = base + r;
        bool alive = false;
        k[i] = 0;
        if (r < rows) {
            k[i] = left[g];
            alive = !((found_bits[g >> 5] >> (g & 31)) & 1u);
        }
        keep[i] = ballot(alive);
        if (lane == 0) item_cnt[wave][i] = static_cast<uint32_t>(__popcll(keep[i]));
    }
    __syncthreads();
    uint32_t wave_start = 0, tile_total = 0;
    for (int w = 0; w < kWavesPerBlock; ++w)
        for (int i = 0; i < kCItems; ++i) {
            if (w < static_cast<int>(wave)) wave_start += item_cnt[w][i];
            tile_total += item_cnt[w][i];
        }
    if (wave == 0) {
        const uint64_t excl = wave_lookback(status, tile, tile_total, 0, error);
        if (lane == 0) offset_slot = excl;
    }
    __syncthreads();
    const uint64_t offset = offset_slot;
    uint32_t p = wave_start;
#pragma unroll
    for (int i = 0; i < kCItems; ++i) {
        if ((keep[i] >> lane) & 1ull) {
            const size_t o = offset + p + lanes_below(keep[i]);
            if (out.packed) out.packed[o] = k[i];
            if (out.packed_copy) out.packed_copy[o] = k[i];
        }
        p += static_cast<uint32_t>(__popcll(keep[i]));
    }
    if (base + rows == nl && threadIdx.x == 0) *count64 = offset + tile_total;
}

// Sparse probe: one thread per compressed candidate (first << minor_bits
// | second); for every run, the index gives the candidate's segment (rows of
// its first value) and a branch-free binary search looks for the pair there.
// The searches of all runs advance in lockstep, so one step's loads are in
// flight together. Found candidates set their bit in found_words (one word
// per 64 candidates, written by the wave's first lane).
__global__ __launch_bounds__(kBlock) void probe_sparse_kernel(const uint64_t* keys, size_t rows, uint32_t minor_bits,
                                                              XRights rights, uint32_t n_x, uint64_t* found_words) {
    const size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
    const bool valid = i < rows;
    const uint64_t raw = valid ? keys[i] : 0;
    const uint32_t x = static_cast<uint32_t>(raw >> minor_bits);
    const uint64_t key = (static_cast<uint64_t>(x) << 32) | (raw & ((1ull << minor_bits) - 1));
    uint32_t lo[kMaxAntiRights], len[kMaxAntiRights];
    uint32_t longest = 0;
#pragma unroll
    for (int q = 0; q < kMaxAntiRights; ++q) {
        lo[q] = len[q] = 0;
        if (valid && q < static_cast<int>(rights.count) && x < n_x) {
            lo[q] = rights.xstart[q][x];
            len[q] = rights.xstart[q][x + 1] - lo[q];
            longest = len[q] > longest ? len[q] : longest;
        }
    }
    // pos[q] = rows of run q's segment below the key
    uint32_t pos[kMaxAntiRights];
#pragma unroll
    for (int q = 0; q < kMaxAntiRights; ++q) pos[q] = 0;
    uint32_t step = 1;
    while (step * 2 <= longest) step *= 2;
    for (; longest && step; step >>= 1) {
        uint64_t probe[kMaxAntiRights];
#pragma unroll
        for (int q = 0; q < kMaxAntiRights; ++q)
            probe[q] = pos[q] + step <= len[q] ? rights.keys[q][lo[q] + pos[q] + step - 1] : ~0ull;
#pragma unroll
        for (int q = 0; q < kMaxAntiRights; ++q)
            if (pos[q] + step <= len[q] && probe[q] < key) pos[q] += step;
    }
    bool hit = false;
#pragma unroll
    for (int q = 0; q < kMaxAntiRights; ++q)
        if (pos[q] < len[q] && rights.keys[q][lo[q] + pos[q]] == key) hit = true;
    const uint64_t m = ballot(valid && hit);
    const size_t row0 = i - lane_id();
    if (lane_id() == 0 && row0 < rows) found_words[row0 / kWave] = m;
}

// Candidates in FULL, by second columns. Layout as the grouped probe: wave w
// owns tile rows [w*512, w*512+512), item i of lane j is row w*512 + i*64 + j,
// so an item's result is one ballot word. Per run of FULL, the segment
// bounds [xstart[x], xstart[x+1]) of the tile's first values go to LDS: the
// index entries of the whole first-value range [x_lo, x_hi + 1] when it is
// at most kPSeg entries (one coalesced read), else the two entries of each
// group of equal first values (runs in row order). The tile's rows of that
// run, [xstart[x_lo], xstart[x_hi + 1]), stream through LDS in chunks of
// kPStage second values, and a candidate (x, y) searches y only in the part
// of its own segment that the chunk holds. A tile with a wide range and more
// than kPSeg / 2 groups (fewer than 4 candidates per first value), whose
// rows of some run exceed kPStreamLimit (few first values with huge
// segments), or whose runs hold more than kPRowsPerCandidate rows per
// candidate, is listed for probe_xsec_deferred_kernel instead, which searches
// every run in place (keeping that path out of this kernel keeps it at 64
// VGPRs).
constexpr int kPTile = 2048;                   // candidates per tile
constexpr int kPItems = kPTile / kBlock;       // 8 per thread
constexpr int kPWaveRows = kPItems * kWave;    // 512 per wave
constexpr int kPStage = 4096;                  // second values per LDS chunk (16 KB)
constexpr int kPStageVecs = kPStage / 4 / kBlock;  // 16-byte loads per thread and chunk
constexpr size_t kPStreamLimit = size_t{1} << 19;
constexpr size_t kPRowsPerCandidate = 256;
constexpr int kPSeg = 512;                     // LDS index entries of a tile's segment bounds (19.6 KB of LDS: 8 blocks a CU)

__global__ __launch_bounds__(kBlock) void probe_xsec_kernel(const uint64_t* keys, size_t rows, uint32_t minor_bits,
                                                            XSecRights rights, uint64_t* found_words,
                                                            uint32_t* deferred) {
    __shared__ __align__(16) uint32_t stage[kPStage];
    __shared__ uint32_t gx[kPSeg / 2];  // first value of each group (group mode)
    __shared__ uint32_t seg[kPSeg];     // segment bounds of the tile's first values in the current run
    __shared__ uint32_t red_lo[kWavesPerBlock], red_hi[kWavesPerBlock], wave_groups[kWavesPerBlock];
    __shared__ uint32_t run_lo[kMaxAntiRights], run_hi[kMaxAntiRights];  // the tile's rows of each run
    const size_t base = static_cast<size_t>(blockIdx.x) * kPTile;
    if (base >= rows) return;
    const uint32_t lane = lane_id(), wave = wave_id();
    const size_t n = rows - base < static_cast<size_t>(kPTile) ? rows - base : kPTile;
    const uint64_t minor_mask = (1ull << minor_bits) - 1;
    // g[i]: the item's first value, then k with its segment at seg[k], seg[k + 1]
    uint32_t g[kPItems], y[kPItems];
    uint32_t valid = 0;  // bit i: item i is a row of the tile
    uint32_t xmin = 0xFFFFFFFFu, xmax = 0;
    // key loads in pairs, each pair issued before it is used (rows clamped
    // to the tile, n >= 1; 32-bit offsets from the tile's base); larger
    // batches take more than 64 VGPRs and cost a block per CU
    constexpr int kHalf = 2;
    const uint64_t* tile_keys = keys + base;
    const uint32_t n_last = static_cast<uint32_t>(n - 1);
#pragma unroll
    for (int h = 0; h < kPItems; h += kHalf) {
        uint64_t kv[kHalf];
#pragma unroll
        for (int j = 0; j < kHalf; ++j) {
            const uint32_t r = wave * kPWaveRows + (h + j) * kWave + lane;
            kv[j] = tile_keys[r < n_last ? r : n_last];
        }
#pragma unroll
        for (int j = 0; j < kHalf; ++j) {
            const int i = h + j;
            const size_t r = wave * kPWaveRows + i * kWave + lane;
            g[i] = y[i] = 0;
            if (r < n) {
                const uint64_t k = kv[j];
                g[i] = static_cast<uint32_t>(k >> minor_bits);
                y[i] = static_cast<uint32_t>(k & minor_mask);
                valid |= 1u << i;
                xmin = g[i] < xmin ? g[i] : xmin;
                xmax = g[i] > xmax ? g[i] : xmax;
            }
        }
    }
    // group starts: a wave's first row, and rows whose first value differs
    // from the row before (a first value split over two waves gets two
    // groups with the same segment, which is harmless)
    uint32_t started = 0, starts_before = 0;
#pragma unroll
    for (int i = 0; i < kPItems; ++i) {
        const size_t r = wave * kPWaveRows + i * kWave + lane;
        uint32_t prev = __shfl_up(g[i], 1, kWave);
        if (i > 0) {
            const uint32_t last = __shfl(g[i > 0 ? i - 1 : 0], kWave - 1, kWave);
            if (lane == 0) prev = last;
        }
        const bool start = r < n && ((i == 0 && lane == 0) || prev != g[i]);
        starts_before += static_cast<uint32_t>(__popcll(ballot(start)));
        started |= (start ? 1u : 0u) << i;
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const uint32_t omin = __shfl_xor(xmin, off, kWave), omax = __shfl_xor(xmax, off, kWave);
        xmin = omin < xmin ? omin : xmin;
        xmax = omax > xmax ? omax : xmax;
    }
    if (lane == 0) {
        red_lo[wave] = xmin;
        red_hi[wave] = xmax;
        wave_groups[wave] = starts_before;
    }
    __syncthreads();
    uint32_t x_lo = 0xFFFFFFFFu, x_hi = 0, n_groups = 0, wave_off = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) {
        x_lo = red_lo[w] < x_lo ? red_lo[w] : x_lo;
        x_hi = red_hi[w] > x_hi ? red_hi[w] : x_hi;
        if (w < static_cast<int>(wave)) wave_off += wave_groups[w];
        n_groups += wave_groups[w];
    }
    // the tile's rows of every run, looked up at once (thread q: run q)
    if (threadIdx.x < rights.count) {
        run_lo[threadIdx.x] = rights.xstart[threadIdx.x][x_lo];
        run_hi[threadIdx.x] = rights.xstart[threadIdx.x][x_hi + 1];
    }
    __syncthreads();
    // dense: the range's index entries fit; else one pair per group (block-uniform)
    const uint32_t span_x = x_hi - x_lo + 2;
    const bool dense = span_x <= static_cast<uint32_t>(kPSeg);
    // every run's range must stream, and the bounds must fit; a tile whose
    // runs hold more than kPRowsPerCandidate rows per candidate searches in
    // place instead (a binary search per candidate reads fewer bytes than
    // streaming the rows) (block-uniform)
    bool stream = dense || n_groups <= static_cast<uint32_t>(kPSeg / 2);
    size_t tile_rows = 0;
    for (uint32_t q = 0; q < rights.count; ++q) {
        stream = stream && static_cast<size_t>(run_hi[q] - run_lo[q]) <= kPStreamLimit;
        tile_rows += run_hi[q] - run_lo[q];
    }
    stream = stream && tile_rows <= kPRowsPerCandidate * n;
    if (!stream) {
        // deferred[0] counts the listed tiles, deferred[1 ..] lists them
        if (threadIdx.x == 0) deferred[1 + atomicAdd(&deferred[0], 1u)] = blockIdx.x;
        return;
    }
    if (dense) {
#pragma unroll
        for (int i = 0; i < kPItems; ++i) g[i] -= x_lo;
    } else {
        // group of a row = group starts at or before it, minus one
        uint32_t before = wave_off;
#pragma unroll
        for (int i = 0; i < kPItems; ++i) {
            const bool start = (started >> i) & 1u;
            const uint64_t m = ballot(start);
            const uint32_t grp = before + lanes_below(m) + (start ? 1u : 0u) - 1u;
            before += static_cast<uint32_t>(__popcll(m));
            if (!((valid >> i) & 1u)) continue;
            if (start) gx[grp] = g[i];
            g[i] = 2 * grp;
        }
        __syncthreads();
    }
    uint32_t found = 0;  // bit i: item i is in some run
    for (uint32_t q = 0; q < rights.count; ++q) {
        const uint32_t* xs = rights.xstart[q];
        const uint32_t* sec = rights.second[q];
        if (dense) {
            for (uint32_t j = threadIdx.x; j < span_x; j += kBlock) seg[j] = xs[x_lo + j];
        } else {
            for (uint32_t j = threadIdx.x; j < n_groups; j += kBlock) {
                seg[2 * j] = xs[gx[j]];
                seg[2 * j + 1] = xs[gx[j] + 1];
            }
        }
        __syncthreads();
        const size_t r0 = run_lo[q], r1 = run_hi[q];
        {
            // chunks start at multiples of 4 rows and are read 16 bytes a
            // lane; the last vector may cover up to 3 rows past r1, inside
            // the buffer's kSecPad rows of padding
            for (size_t c0 = r0 & ~size_t{3}; c0 < r1; c0 += kPStage) {
                const size_t c1 = c0 + kPStage < r1 ? c0 + kPStage : r1;
                const uint32_t nvec = static_cast<uint32_t>((c1 - c0 + 3) / 4);
                const uint4* chunk = reinterpret_cast<const uint4*>(sec + c0);
                uint4* stage4 = reinterpret_cast<uint4*>(stage);
                // straight into LDS, every vector of the chunk in flight at
                // once (lanes past nvec repeat the last vector)
#pragma unroll
                for (int u = 0; u < kPStageVecs; ++u) {
                    const uint32_t j = u * kBlock + threadIdx.x;
                    load_lds16(chunk + (j < nvec ? j : nvec - 1), stage4 + u * kBlock + wave * kWave);
                }
                __syncthreads();
#pragma unroll
                for (int i = 0; i < kPItems; ++i) {
                    if (!(((valid & ~found) >> i) & 1u)) continue;
                    const uint32_t lo = seg[g[i]], hi = seg[g[i] + 1];
                    if (lo >= c1 || hi <= c0) continue;
                    uint32_t a = static_cast<uint32_t>((lo > c0 ? lo : c0) - c0);
                    const uint32_t e = static_cast<uint32_t>((hi < c1 ? hi : c1) - c0);
                    // the chunk's slice of the segment is ascending: search
                    // only the one slice whose values span y
                    if (y[i] < stage[a] || y[i] > stage[e - 1]) continue;
                    uint32_t b = e;
                    while (a < b) {
                        const uint32_t mid = (a + b) >> 1;
                        if (stage[mid] < y[i])
                            a = mid + 1;
                        else
                            b = mid;
                    }
                    if (a < e && stage[a] == y[i]) found |= 1u << i;
                }
                __syncthreads();
            }
        }
        __syncthreads();  // seg is rewritten for the next run
    }
    // bit j of word (base + wave*512 + i*64) / 64 is row base + wave*512 + i*64 + j
#pragma unroll
    for (int i = 0; i < kPItems; ++i) {
        const uint64_t m = ballot(((found >> i) & 1u) != 0);
        const size_t row0 = base + wave * kPWaveRows + i * kWave;
        if (lane == 0 && row0 < rows) found_words[row0 / kWave] = m;
    }
}

// Whether (x, y) is in some run: a lockstep binary search over every run's
// segment of first value x (one step of all runs in flight at a time).
__device__ __forceinline__ bool in_runs(uint32_t x, uint32_t y, const XSecRights& rights) {
    uint32_t lo[kMaxAntiRights], len[kMaxAntiRights];
    uint32_t longest = 0;
#pragma unroll
    for (int q = 0; q < kMaxAntiRights; ++q) {
        lo[q] = len[q] = 0;
        if (q < static_cast<int>(rights.count)) {
            lo[q] = rights.xstart[q][x];
            len[q] = rights.xstart[q][x + 1] - lo[q];
            longest = len[q] > longest ? len[q] : longest;
        }
    }
    // pos[q] = values of run q's segment below y
    uint32_t pos[kMaxAntiRights];
#pragma unroll
    for (int q = 0; q < kMaxAntiRights; ++q) pos[q] = 0;
    uint32_t step = 1;
    while (step * 2 <= longest) step *= 2;
    for (; longest && step; step >>= 1) {
        uint32_t probe[kMaxAntiRights];
#pragma unroll
        for (int q = 0; q < kMaxAntiRights; ++q)
            probe[q] = pos[q] + step <= len[q] ? rights.second[q][lo[q] + pos[q] + step - 1] : 0xFFFFFFFFu;
#pragma unroll
        for (int q = 0; q < kMaxAntiRights; ++q)
            if (pos[q] + step <= len[q] && probe[q] < y) pos[q] += step;
    }
    bool hit = false;
#pragma unroll
    for (int q = 0; q < kMaxAntiRights; ++q)
        if (pos[q] < len[q] && rights.second[q][lo[q] + pos[q]] == y) hit = true;
    return hit;
}

// The listed tiles of probe_xsec_kernel, one thread per candidate.
// Persistent blocks of kBlock candidates walk the list (kPTile / kBlock
// blocks per tile).
__global__ __launch_bounds__(kBlock) void probe_xsec_deferred_kernel(const uint64_t* keys, size_t rows,
                                                                     uint32_t minor_bits, XSecRights rights,
                                                                     uint64_t* found_words,
                                                                     const uint32_t* deferred) {
    constexpr uint32_t kPer = kPTile / kBlock;
    const uint64_t minor_mask = (1ull << minor_bits) - 1;
    const uint32_t listed = deferred[0];
    for (uint32_t w = blockIdx.x; w < listed * kPer; w += gridDim.x) {
        const size_t tile = deferred[1 + w / kPer];
        const size_t i = tile * kPTile + static_cast<size_t>(w % kPer) * kBlock + threadIdx.x;
        const bool valid = i < rows;
        const uint64_t k = valid ? keys[i] : 0;
        const bool hit = valid && in_runs(static_cast<uint32_t>(k >> minor_bits),
                                          static_cast<uint32_t>(k & minor_mask), rights);
        const uint64_t m = ballot(hit);
        const size_t row0 = i - lane_id();
        if (lane_id() == 0 && row0 < rows) found_words[row0 / kWave] = m;
    }
}

// probe_xsec for few candidates: one thread per candidate, every run
// searched in place.
__global__ __launch_bounds__(kBlock) void probe_sparse_sec_kernel(const uint64_t* keys, size_t rows,
                                                                  uint32_t minor_bits, XSecRights rights,
                                                                  uint64_t* found_words) {
    const uint64_t minor_mask = (1ull << minor_bits) - 1;
    const size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
    const bool valid = i < rows;
    const uint64_t k = valid ? keys[i] : 0;
    const bool hit = valid && in_runs(static_cast<uint32_t>(k >> minor_bits),
                                      static_cast<uint32_t>(k & minor_mask), rights);
    const uint64_t m = ballot(hit);
    const size_t row0 = i - lane_id();
    if (lane_id() == 0 && row0 < rows) found_words[row0 / kWave] = m;
}

// ---- merges of runs held as (first-column index, second column) ----
// C = A ∪ B for disjoint runs: C's index is xa + xb (xindex_add), and C's
// second column lists the merged rows in (first, second) order. Output tile
// t covers merged rows [t*kRTile, (t+1)*kRTile). Its first value bx[t] is
// the x whose merged segment [xa[x] + xb[x], xa[x+1] + xb[x+1]) holds row
// t*kRTile (written by the thread of that x), and its split pa[t] between A
// and B is a merge path inside the two segments of bx[t]. The tile rebuilds
// the keys (x << 32 | second) of its A and B rows in LDS, x from the index
// entries of its first-value range, and merges them as merge_disjoint does.
// With B empty it expands a run to packed keys.
constexpr int kRTile = 2048;                 // merged rows per tile
constexpr int kRItems = kRTile / kBlock;     // 8 per thread
constexpr uint32_t kRSpan = 256;             // index entries of a tile's first-value range kept in LDS (18 KB: 8 blocks a CU)

__global__ void run_bound_x_kernel(const uint32_t* xa, const uint32_t* xb, uint32_t n_x, uint32_t* bx) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t x = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; x < n_x; x += stride) {
        const uint64_t c0 = static_cast<uint64_t>(xa[x]) + (xb ? xb[x] : 0u);
        const uint64_t c1 = static_cast<uint64_t>(xa[x + 1]) + (xb ? xb[x + 1] : 0u);
        for (uint64_t t = (c0 + kRTile - 1) / kRTile; t * kRTile < c1; ++t) bx[t] = static_cast<uint32_t>(x);
    }
}

__global__ void run_partition_kernel(const uint32_t* xa, const uint32_t* sa, size_t na, const uint32_t* xb,
                                     const uint32_t* sb, const uint32_t* bx, size_t tiles, size_t* pa) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; t <= tiles; t += stride) {
        if (t == tiles) {
            pa[t] = na;
            continue;
        }
        const uint32_t x = bx[t];
        const uint64_t a0 = xa[x], b0 = xb ? xb[x] : 0u;
        const uint64_t la = xa[x + 1] - a0, lb = xb ? xb[x + 1] - b0 : 0u;
        const uint64_t j = static_cast<uint64_t>(t) * kRTile - a0 - b0;  // merged rows of x before the boundary
        uint64_t lo = j > lb ? j - lb : 0, hi = j < la ? j : la;
        while (lo < hi) {  // the rows of A among the first j of x's merged rows (no value in common)
            const uint64_t mid = (lo + hi) >> 1;
            if (!(sb[b0 + j - 1 - mid] < sa[a0 + mid]))
                lo = mid + 1;
            else
                hi = mid;
        }
        pa[t] = a0 + lo;
    }
}

__device__ __forceinline__ uint32_t merge_path_keys(const uint64_t* a, uint32_t na, const uint64_t* b, uint32_t nb,
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

__global__ __launch_bounds__(kBlock) void run_merge_kernel(const uint32_t* xa, const uint32_t* sa, size_t na,
                                                           const uint32_t* xb, const uint32_t* sb, size_t nb,
                                                           uint32_t n_x, const uint32_t* bx, const size_t* pa,
                                                           size_t tiles, uint32_t* out_second, uint64_t* out_packed) {
    // kRTile keys, then one u64 of padding per 32 for the merged rows'
    // transposition (see rpad)
    __shared__ uint64_t lds[kRTile + kRTile / 32];
    __shared__ uint32_t xa_l[kRSpan], xb_l[kRSpan];
    __shared__ uint32_t scan_max[kWavesPerBlock];
    const size_t t = blockIdx.x;
    const size_t total = na + nb;
    const size_t d0 = t * kRTile;
    if (d0 >= total) return;
    const size_t d1 = d0 + kRTile < total ? d0 + kRTile : total;
    const size_t a0 = pa[t], a1 = pa[t + 1];
    const size_t b0 = d0 - a0, b1 = d1 - a1;
    const uint32_t n_a = static_cast<uint32_t>(a1 - a0), n_b = static_cast<uint32_t>(b1 - b0);
    const uint32_t n = n_a + n_b;
    // the tile's second values (a's rows, then b's) are all loaded first:
    // position p of the tile is item p / kBlock of thread p % kBlock
    // (positions clamped to the tile, which holds >= 1 row)
    const uint32_t* sbb = sb ? sb : sa;
    uint32_t sv[kRTile / kBlock];
#pragma unroll
    for (int q = 0; q < kRTile / kBlock; ++q) {
        const uint32_t p = threadIdx.x + q * kBlock;
        const uint32_t pc = p < n ? p : n - 1;
        sv[q] = pc < n_a ? sa[a0 + pc] : sbb[b0 + (pc - n_a)];
    }
    // every row of the tile has its first value in [x_lo, x_hi]
    const uint32_t x_lo = bx[t];
    const uint32_t x_hi = t + 1 < tiles ? bx[t + 1] : n_x - 1;
    const uint32_t span = x_hi - x_lo + 2;  // entries x_lo .. x_hi + 1
    const bool staged = span <= kRSpan;     // block-uniform
    uint32_t xr[kRTile / kBlock];           // first value of each of the thread's rows
    if (staged) {
        // Each first value marks the tile position where its rows start, as
        // (part << 8 | x - x_lo) with part 1 for b's rows (so b's marks
        // dominate a's); a block max-scan then gives every position its
        // first value. The marks live in lds (u16) before the keys do.
        uint16_t* const mark = reinterpret_cast<uint16_t*>(lds);
        static_assert(kRTile * sizeof(uint16_t) == kBlock * sizeof(uint4), "one vector of marks per thread");
        static_assert(kRSpan <= 256, "x - x_lo fits a byte");
        for (uint32_t j = threadIdx.x; j < span; j += kBlock) {
            xa_l[j] = xa[x_lo + j];
            xb_l[j] = xb ? xb[x_lo + j] : 0u;
        }
        reinterpret_cast<uint4*>(lds)[threadIdx.x] = make_uint4(0u, 0u, 0u, 0u);
        __syncthreads();
        // x's rows start at position clamp(xs[x] - base) of their part; when
        // several first values start at one position (empty segments, or
        // segments begun before the tile) the largest marks it
        auto at = [](uint32_t v, size_t base, uint32_t len) -> uint32_t {
            return v <= base ? 0u : (v - base < len ? static_cast<uint32_t>(v - base) : len);
        };
        for (uint32_t j = threadIdx.x; j + 1 < span; j += kBlock) {
            const uint32_t pa0 = at(xa_l[j], a0, n_a), pa1 = at(xa_l[j + 1], a0, n_a);
            if (pa0 < pa1) mark[pa0] = static_cast<uint16_t>(j);
            if (xb) {
                const uint32_t pb0 = at(xb_l[j], b0, n_b), pb1 = at(xb_l[j + 1], b0, n_b);
                if (pb0 < pb1) mark[n_a + pb0] = static_cast<uint16_t>(0x100u | j);
            }
        }
        __syncthreads();
        // inclusive max-scan: thread t owns positions [8t, 8t + 8)
        uint4 w = reinterpret_cast<const uint4*>(lds)[threadIdx.x];
        uint32_t m[8] = {w.x & 0xFFFFu, w.x >> 16, w.y & 0xFFFFu, w.y >> 16,
                         w.z & 0xFFFFu, w.z >> 16, w.w & 0xFFFFu, w.w >> 16};
#pragma unroll
        for (int i = 1; i < 8; ++i) m[i] = m[i] > m[i - 1] ? m[i] : m[i - 1];
        uint32_t incl = m[7];
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t o = __shfl_up(incl, off, kWave);
            if (static_cast<int>(lane_id()) >= off) incl = o > incl ? o : incl;
        }
        if (lane_id() == kWave - 1) scan_max[wave_id()] = incl;
        __syncthreads();
        uint32_t prior = __shfl_up(incl, 1, kWave);
        if (lane_id() == 0) prior = 0;
        for (int wv = 0; wv < static_cast<int>(wave_id()); ++wv) prior = scan_max[wv] > prior ? scan_max[wv] : prior;
#pragma unroll
        for (int i = 0; i < 8; ++i) m[i] = m[i] > prior ? m[i] : prior;
        w = make_uint4(m[0] | (m[1] << 16), m[2] | (m[3] << 16), m[4] | (m[5] << 16), m[6] | (m[7] << 16));
        reinterpret_cast<uint4*>(lds)[threadIdx.x] = w;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kRTile / kBlock; ++q) {
            const uint32_t p = threadIdx.x + q * kBlock;
            xr[q] = x_lo + (mark[p < n ? p : 0] & 0xFFu);
        }
        __syncthreads();  // lds holds keys from here on
    } else {
        // a wide first-value range: binary search over the indexes in place
        auto first_of = [&](const uint32_t* xs, size_t row) -> uint32_t {
            uint32_t lo = 0, hi = x_hi - x_lo;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (xs[x_lo + mid] <= row)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            return x_lo + lo;
        };
#pragma unroll
        for (int q = 0; q < kRTile / kBlock; ++q) {
            const uint32_t p = threadIdx.x + q * kBlock;
            xr[q] = p < n_a ? first_of(xa, a0 + p) : (p < n ? first_of(xb, b0 + (p - n_a)) : 0u);
        }
    }
#pragma unroll
    for (int q = 0; q < kRTile / kBlock; ++q) {
        const uint32_t p = threadIdx.x + q * kBlock;
        if (p < n) lds[p] = (static_cast<uint64_t>(xr[q]) << 32) | sv[q];
    }
    __syncthreads();
    const uint32_t diag = threadIdx.x * kRItems < n ? threadIdx.x * kRItems : n;
    uint32_t ia = merge_path_keys(lds, n_a, lds + n_a, n_b, diag);
    uint32_t ib = diag - ia;
    // the heads of both sides stay in registers: one LDS read per merged row
    // (the side just consumed), not two
    constexpr uint64_t kPast = ~0ull;  // above every key (x < 2^24)
    uint64_t ha = ia < n_a ? lds[ia] : kPast;
    uint64_t hb = ib < n_b ? lds[n_a + ib] : kPast;
    uint64_t v[kRItems];
#pragma unroll
    for (int i = 0; i < kRItems; ++i) {
        const bool take_a = ha <= hb;
        v[i] = take_a ? ha : hb;
        if (diag + i < n) {
            if (take_a) {
                ++ia;
                ha = ia < n_a ? lds[ia] : kPast;
            } else {
                ++ib;
                hb = ib < n_b ? lds[n_a + ib] : kPast;
            }
        }
    }
    __syncthreads();
    // thread t's rows are [8t, 8t + 8): written unpadded, lanes 64 bytes
    // apart would meet in a quarter of the banks; one u64 of padding per 32
    // rows spreads them (rpad), and the coalesced stores read back through it
    auto rpad = [](uint32_t j) { return j + (j >> 5); };
#pragma unroll
    for (int i = 0; i < kRItems; ++i)
        if (diag + i < n) lds[rpad(diag + i)] = v[i];
    __syncthreads();
    // the merged run is read again only in later steps: streaming stores
    // (nontemporal) leave L2 to the index entries
    if (out_second)
        for (uint32_t i = threadIdx.x; i < n; i += kBlock)
            __builtin_nontemporal_store(static_cast<uint32_t>(lds[rpad(i)]), &out_second[d0 + i]);
    if (out_packed)
        for (uint32_t i = threadIdx.x; i < n; i += kBlock)
            __builtin_nontemporal_store(lds[rpad(i)], &out_packed[d0 + i]);
}

__global__ void xsec_init_kernel(uint32_t* deferred) { deferred[0] = 0; }

constexpr uint32_t kFillChunk = 4096;  // index entries per block of the suffix-minimum fill

__global__ void xindex_clear_kernel(uint32_t* xstart, size_t n) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        xstart[i] = 0xFFFFFFFFu;
}

// pass 1: entry n_x = the row count; each chunk's minimum. Loads are
// clamped into the index and all issued before any is used.
__global__ __launch_bounds__(kBlock) void xfill_chunk_min_kernel(uint32_t* xstart, uint32_t n_x,
                                                                 const uint64_t* rows_dev, uint32_t* chunk_min) {
    constexpr uint32_t kPer = kFillChunk / kBlock;  // entries per thread, strided by the block
    __shared__ uint32_t red[kWavesPerBlock];
    const size_t b0 = static_cast<size_t>(blockIdx.x) * kFillChunk;
    const size_t n = static_cast<size_t>(n_x) + 1;
    const uint32_t total = static_cast<uint32_t>(*rows_dev);
    if (b0 + kFillChunk >= n && threadIdx.x == 0) xstart[n_x] = total;
    uint32_t v[kPer];
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) {
        const size_t i = b0 + q * kBlock + threadIdx.x;
        v[q] = xstart[i < n_x ? i : n_x - 1];  // n_x >= 1
    }
    uint32_t m = 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) {
        const size_t i = b0 + q * kBlock + threadIdx.x;
        const uint32_t x = i < n_x ? v[q] : (i == n_x ? total : 0xFFFFFFFFu);
        m = x < m ? x : m;
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const uint32_t o = __shfl_xor(m, off, kWave);
        m = o < m ? o : m;
    }
    if (lane_id() == 0) red[wave_id()] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kWavesPerBlock; ++w) m = red[w] < m ? red[w] : m;
        chunk_min[blockIdx.x] = m < red[0] ? m : red[0];
    }
}

// pass 2: every entry = the minimum of itself and every entry after it.
// Thread t owns kPer consecutive entries; the minimum over later threads
// comes from a wave suffix scan and one value per later wave.
__global__ __launch_bounds__(kBlock) void xfill_apply_kernel(uint32_t* xstart, uint32_t n_x, const uint32_t* chunk_min,
                                                             uint32_t chunks) {
    constexpr uint32_t kPer = kFillChunk / kBlock;  // 16 consecutive entries per thread
    __shared__ uint32_t wave_min[kWavesPerBlock];
    __shared__ uint32_t carry_slot;
    const size_t b0 = static_cast<size_t>(blockIdx.x) * kFillChunk;
    const size_t n = static_cast<size_t>(n_x) + 1;
    const uint32_t lane = lane_id(), wave = wave_id();
    // carry: the minimum of every later chunk
    uint32_t c = 0xFFFFFFFFu;
    for (uint32_t j = blockIdx.x + 1 + threadIdx.x; j < chunks; j += kBlock) c = chunk_min[j] < c ? chunk_min[j] : c;
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const uint32_t o = __shfl_xor(c, off, kWave);
        c = o < c ? o : c;
    }
    if (threadIdx.x == 0) carry_slot = 0xFFFFFFFFu;
    __syncthreads();
    if (lane == 0) atomicMin(&carry_slot, c);
    // this thread's entries (loads clamped, all issued first), their suffix
    // minimum within the thread
    uint32_t v[kPer];
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) {
        const size_t i = b0 + threadIdx.x * kPer + q;
        v[q] = xstart[i < n ? i : n - 1];
    }
    uint32_t m = 0xFFFFFFFFu;
#pragma unroll
    for (int q = kPer - 1; q >= 0; --q) {
        const size_t i = b0 + threadIdx.x * kPer + q;
        const uint32_t x = i < n ? v[q] : 0xFFFFFFFFu;
        m = x < m ? x : m;
        v[q] = m;
    }
    // minimum over this and later lanes of the wave
    uint32_t incl = m;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t o = __shfl_down(incl, off, kWave);
        if (lane + off < static_cast<uint32_t>(kWave)) incl = o < incl ? o : incl;
    }
    if (lane == 0) wave_min[wave] = incl;
    uint32_t after = __shfl_down(incl, 1, kWave);  // later lanes of this wave
    if (lane == kWave - 1) after = 0xFFFFFFFFu;
    __syncthreads();
    for (int w = static_cast<int>(wave) + 1; w < kWavesPerBlock; ++w) after = wave_min[w] < after ? wave_min[w] : after;
    after = carry_slot < after ? carry_slot : after;
#pragma unroll
    for (int q = 0; q < static_cast<int>(kPer); ++q) {
        const size_t i = b0 + threadIdx.x * kPer + q;
        if (i < n) xstart[i] = v[q] < after ? v[q] : after;
    }
}

__global__ void xindex_extend_kernel(uint32_t* xstart, uint32_t old_n, uint32_t n_x, uint32_t rows) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t x = static_cast<size_t>(old_n) + 1 + static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         x <= n_x; x += stride)
        xstart[x] = rows;
}

__global__ void second_column_kernel(const uint64_t* keys, size_t rows, uint32_t* second) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < rows; i += stride)
        second[i] = static_cast<uint32_t>(keys[i]);
}

size_t xanti_tiles(size_t bound) { return (bound + kXTile - 1) / kXTile; }
size_t xcompact_tiles(size_t bound) { return (bound + kCTile - 1) / kCTile; }

}  // namespace

hipError_t xindex_build(const uint64_t* keys, const uint64_t* rows_dev, size_t bound, uint32_t n_x, uint32_t* xstart,
                        hipStream_t stream) {
    if (!xstart || (bound && !keys)) return hipErrorInvalidValue;
    const size_t blocks = (static_cast<size_t>(n_x) + 1 + kXBuildSpan - 1) / kXBuildSpan;
    xindex_build_kernel<<<static_cast<unsigned>(blocks), kBlock, 0, stream>>>(keys, rows_dev, bound, n_x, xstart);
    return hipGetLastError();
}

hipError_t xindex_add(uint32_t* into, const uint32_t* other, uint32_t n_x, hipStream_t stream) {
    if (!into || !other) return hipErrorInvalidValue;
    const size_t n = static_cast<size_t>(n_x) + 1;
    const size_t blocks = (n + kBlock * 4 - 1) / (kBlock * 4);
    xindex_add_kernel<<<static_cast<unsigned>(blocks < 8192 ? blocks : 8192), kBlock, 0, stream>>>(into, other, n);
    return hipGetLastError();
}

size_t anti_join_x_temp_bytes(size_t bound) {
    const size_t b = bound ? bound : 1;
    return 256 + (xcompact_tiles(b) * sizeof(uint64_t) + 255) / 256 * 256 +
           (xanti_tiles(b) * (kXTile / 32) * sizeof(uint32_t) + 255) / 256 * 256;
}

hipError_t probe_sparse(const uint64_t* keys, size_t rows, uint32_t minor_bits, const XRights& rights, uint32_t n_x,
                        uint64_t* found_words, hipStream_t stream, const KernelTimer* timer) {
    if (rows == 0) return hipSuccess;
    if (minor_bits == 0 || minor_bits > 32 || rights.count > static_cast<uint32_t>(kMaxAntiRights))
        return hipErrorInvalidValue;
    const size_t blocks = (rows + kBlock - 1) / kBlock;
    timer_start(timer, stream);
    probe_sparse_kernel<<<static_cast<unsigned>(blocks), kBlock, 0, stream>>>(keys, rows, minor_bits, rights, n_x,
                                                                             found_words);
    timer_stop(timer, stream);
    return hipGetLastError();
}

hipError_t xindex_extend(const uint32_t* from, uint32_t old_n, uint32_t* to, uint32_t n_x, uint32_t rows,
                         hipStream_t stream) {
    if (!from || !to || n_x < old_n) return hipErrorInvalidValue;
    hipError_t e = hipMemcpyAsync(to, from, (static_cast<size_t>(old_n) + 1) * sizeof(uint32_t),
                                  hipMemcpyDeviceToDevice, stream);
    if (e != hipSuccess || n_x == old_n) return e;
    const size_t n = static_cast<size_t>(n_x - old_n);
    const size_t blocks = (n + kBlock - 1) / kBlock;
    xindex_extend_kernel<<<static_cast<unsigned>(blocks < 4096 ? blocks : 4096), kBlock, 0, stream>>>(to, old_n, n_x,
                                                                                                   rows);
    return hipGetLastError();
}

hipError_t xindex_clear(uint32_t* xstart, uint32_t n_x, hipStream_t stream) {
    if (!xstart) return hipErrorInvalidValue;
    const size_t n = static_cast<size_t>(n_x) + 1;
    const size_t blocks = (n + kBlock - 1) / kBlock;
    xindex_clear_kernel<<<static_cast<unsigned>(blocks < 4096 ? blocks : 4096), kBlock, 0, stream>>>(xstart, n);
    return hipGetLastError();
}

size_t xindex_fill_temp_bytes(uint32_t n_x) {
    return ((static_cast<size_t>(n_x) + kFillChunk) / kFillChunk) * sizeof(uint32_t);
}

hipError_t xindex_fill(uint32_t* xstart, uint32_t n_x, const uint64_t* rows_dev, void* temp, size_t temp_bytes,
                       hipStream_t stream) {
    if (!xstart || !rows_dev || temp_bytes < xindex_fill_temp_bytes(n_x)) return hipErrorInvalidValue;
    const uint32_t chunks = static_cast<uint32_t>((static_cast<size_t>(n_x) + kFillChunk) / kFillChunk);
    uint32_t* chunk_min = static_cast<uint32_t*>(temp);
    xfill_chunk_min_kernel<<<chunks, kBlock, 0, stream>>>(xstart, n_x, rows_dev, chunk_min);
    xfill_apply_kernel<<<chunks, kBlock, 0, stream>>>(xstart, n_x, chunk_min, chunks);
    return hipGetLastError();
}

static bool xsec_rights_ok(const XSecRights& rights) {
    if (rights.count > static_cast<uint32_t>(kMaxAntiRights)) return false;
    for (uint32_t q = 0; q < rights.count; ++q)
        if (!rights.second[q] || !rights.xstart[q]) return false;
    return true;
}

size_t probe_xsec_temp_bytes(size_t rows) {
    return ((rows + kPTile - 1) / kPTile + 1) * sizeof(uint32_t);
}

hipError_t probe_xsec(const uint64_t* keys, size_t rows, uint32_t minor_bits, const XSecRights& rights,
                      uint32_t n_x, uint64_t* found_words, void* temp, size_t temp_bytes, hipStream_t stream,
                      const KernelTimer* timer) {
    if (rows == 0) return hipSuccess;
    if (minor_bits == 0 || minor_bits > 32 || !xsec_rights_ok(rights) || n_x == 0 || !keys || !found_words ||
        temp_bytes < probe_xsec_temp_bytes(rows))
        return hipErrorInvalidValue;
    const size_t tiles = (rows + kPTile - 1) / kPTile;
    uint32_t* deferred = static_cast<uint32_t*>(temp);
    xsec_init_kernel<<<1, 1, 0, stream>>>(deferred);
    timer_start(timer, stream);
    probe_xsec_kernel<<<static_cast<unsigned>(tiles), kBlock, 0, stream>>>(keys, rows, minor_bits, rights,
                                                                          found_words, deferred);
    timer_stop(timer, stream);
    const size_t per = kPTile / kBlock;
    const size_t grid = tiles * per < 2048 ? tiles * per : 2048;
    probe_xsec_deferred_kernel<<<static_cast<unsigned>(grid), kBlock, 0, stream>>>(keys, rows, minor_bits, rights,
                                                                                   found_words, deferred);
    return hipGetLastError();
}

hipError_t probe_sparse_sec(const uint64_t* keys, size_t rows, uint32_t minor_bits, const XSecRights& rights,
                            uint32_t n_x, uint64_t* found_words, hipStream_t stream, const KernelTimer* timer) {
    if (rows == 0) return hipSuccess;
    if (minor_bits == 0 || minor_bits > 32 || !xsec_rights_ok(rights) || n_x == 0 || !keys || !found_words)
        return hipErrorInvalidValue;
    timer_start(timer, stream);
    probe_sparse_sec_kernel<<<static_cast<unsigned>((rows + kBlock - 1) / kBlock), kBlock, 0, stream>>>(
        keys, rows, minor_bits, rights, found_words);
    timer_stop(timer, stream);
    return hipGetLastError();
}

size_t run_merge_temp_bytes(size_t rows) {
    const size_t tiles = (rows + kRTile - 1) / kRTile;
    return (tiles * sizeof(uint32_t) + 255) / 256 * 256 + (tiles + 1) * sizeof(size_t);
}

hipError_t run_merge(const uint32_t* xa, const uint32_t* sa, size_t na, const uint32_t* xb, const uint32_t* sb,
                     size_t nb, uint32_t n_x, uint32_t* out_second, uint64_t* out_packed, void* temp,
                     size_t temp_bytes, hipStream_t stream, const KernelTimer* timer) {
    const size_t total = na + nb;
    if (total == 0) return hipSuccess;
    if (n_x == 0 || !xa || (na && !sa) || (nb && (!xb || !sb)) || temp_bytes < run_merge_temp_bytes(total) ||
        (!out_second && !out_packed))
        return hipErrorInvalidValue;
    if (!nb) xb = sb = nullptr;
    const size_t tiles = (total + kRTile - 1) / kRTile;
    uint32_t* bx = static_cast<uint32_t*>(temp);
    size_t* pa = reinterpret_cast<size_t*>(static_cast<char*>(temp) + (tiles * sizeof(uint32_t) + 255) / 256 * 256);
    const size_t xblocks = (static_cast<size_t>(n_x) + kBlock - 1) / kBlock;
    run_bound_x_kernel<<<static_cast<unsigned>(xblocks < 8192 ? xblocks : 8192), kBlock, 0, stream>>>(xa, xb, n_x,
                                                                                                    bx);
    const size_t pblocks = (tiles + 1 + kBlock - 1) / kBlock;
    run_partition_kernel<<<static_cast<unsigned>(pblocks < 4096 ? pblocks : 4096), kBlock, 0, stream>>>(
        xa, sa, na, xb, sb, bx, tiles, pa);
    timer_start(timer, stream);
    run_merge_kernel<<<static_cast<unsigned>(tiles), kBlock, 0, stream>>>(xa, sa, na, xb, sb, nb, n_x, bx, pa, tiles,
                                                                          out_second, out_packed);
    timer_stop(timer, stream);
    return hipGetLastError();
}

hipError_t second_column(const uint64_t* keys, size_t rows, uint32_t* second, hipStream_t stream) {
    if (rows == 0) return hipSuccess;
    if (!keys || !second) return hipErrorInvalidValue;
    const size_t blocks = (rows + kBlock * 4 - 1) / (kBlock * 4);
    second_column_kernel<<<static_cast<unsigned>(blocks < 8192 ? blocks : 8192), kBlock, 0, stream>>>(keys, rows,
                                                                                                     second);
    return hipGetLastError();
}

hipError_t anti_join_x(const uint64_t* left, const uint64_t* left_rows, size_t bound, const XRights& rights,
                       uint32_t n_x, PairOutput out, uint64_t* count64, void* temp, size_t temp_bytes,
                       hipStream_t stream, const KernelTimer* timer) {
    if (bound == 0) return set_u64(count64, 0, stream);
    if (temp_bytes < anti_join_x_temp_bytes(bound) || rights.count > static_cast<uint32_t>(kMaxAntiRights) ||
        out.first != nullptr)
        return hipErrorInvalidValue;
    const size_t probe_tiles = xanti_tiles(bound), tiles = xcompact_tiles(bound);
    char* t = static_cast<char*>(temp);
    XHeader* h = reinterpret_cast<XHeader*>(t);
    uint64_t* status = reinterpret_cast<uint64_t*>(t + 256);
    uint32_t* found_bits = reinterpret_cast<uint32_t*>(t + 256 + (tiles * sizeof(uint64_t) + 255) / 256 * 256);
    const size_t init_blocks = tiles / kBlock + 1;
    xanti_init_kernel<<<static_cast<unsigned>(init_blocks < 4096 ? init_blocks : 4096), kBlock, 0, stream>>>(
        h, status, tiles);
    timer_start(timer, stream);
    anti_join_x_kernel<<<static_cast<unsigned>(probe_tiles), kBlock, 0, stream>>>(left, left_rows, bound, rights,
                                                                                  n_x, found_bits);
    timer_stop(timer, stream);
    uint32_t* error = error_sink() ? error_sink() : &h->error;
    xanti_compact_kernel<<<static_cast<unsigned>(tiles), kBlock, 0, stream>>>(left, left_rows, bound, found_bits, out,
                                                                             count64, status, h, error);
    const hipError_t st = hipGetLastError();
    if (st != hipSuccess) return st;
    return report_error(error, nullptr, count64, stream);
}

}  // namespace spk
