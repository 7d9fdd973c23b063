// Segmented distinct for keys whose major part is already non-decreasing.
//
// In the semi-naive loop the join emits candidates in DELTA order, and DELTA
// is sorted by its first column, so whenever the first output column is
// DELTA's first column the candidates arrive grouped by that column: only the
// runs of one major value need sorting. This replaces the multi-pass global
// radix sort (read + write of every key per 8-bit digit) by one read and one
// write per key:
//   1. partition: tile boundaries are moved back to the start of the run
//      that contains them (one thread per boundary), so no run crosses a tile;
//   2. a tile longer than kSegCap starts with a long run (one source's
//      candidates in one iteration): its keys set bits of a bitmap of their
//      minors in LDS, whose set bits are the run's sorted unique minors
//      (a run whose minors differ above their low 18 bits sends the tile,
//      gathered, through the global radix sort instead);
//   3. every tile (a long tile: its tail after the first run) sorts its keys
//      in LDS by (local run rank, minor) with LSD
//      passes over the bit range where its keys differ (a 12-bit digit, then
//      stable 9-bit digits: up to 30 varying bits take three passes), drops
//      duplicates, and writes the expanded pairs (major << 32 | minor) back
//      over its own input slice, with its unique count;
//   4. a one-block scan turns the counts into offsets, and a copy moves each
//      tile's rows to its offset. A single pass with decoupled look-back
//      kept tiles waiting for slower predecessors for about 38% of their
//      life; the copy costs 16 bytes a unique row instead.
#include "ops.hpp"
#include "prims.hpp"

namespace spk {
namespace {

constexpr int kSegCap = 4096;                 // largest tile sorted in LDS
constexpr int kSegBlock = 512;                // threads of a tile block (8 waves)
constexpr int kSegWaves = kSegBlock / kWave;
constexpr int kSegItems = kSegCap / kSegBlock;  // 8 keys per thread
constexpr int kSegDigitBits = 9;              // digit of a tile pass
constexpr int kSegRadix = 1 << kSegDigitBits;
constexpr int kSegFirstBits = 12;             // digit of the first (unstable) pass
constexpr int kSegFirstRadix = 1 << kSegFirstBits;
constexpr uint32_t kSegRankBits = 12;  // a tile holds fewer runs than kSegCap

// A tile longer than kSegCap holds one long run (its first: the run that
// contains the tile's nominal start) and a tail of shorter runs that lies
// within one nominal tile (the runs after the first start after the nominal
// start, and end before the next nominal boundary), so the tail is at most
// tile_rows - 1 <= kSegCap keys. The first run has one major: it is
// deduplicated through a bitmap of its minors' low bits in LDS when its
// minors agree above them (seg_run_kernel), and the tail is sorted like a
// small tile. A tile whose first run spans a wider range goes to the global
// sort whole. The bitmap covers 20 bits (128 KB of LDS, one block a CU) when
// minors have 19 or 20 bits, so no run is too wide; else 18 (32 KB, four
// blocks a CU).
constexpr int kRunItems = 8;  // keys a thread loads per chunk

__device__ __forceinline__ void store_pair(const PairOutput& out, size_t i, uint64_t k) {
    if (out.packed) out.packed[i] = k;
    if (out.packed_copy) out.packed_copy[i] = k;
    if (out.second_copy) out.second_copy[i] = static_cast<uint32_t>(k);
    if (out.first) {
        out.first[i] = static_cast<uint32_t>(k >> 32);
        out.second[i] = static_cast<uint32_t>(k);
    }
}

__host__ __device__ inline size_t seg_tiles(size_t n, uint32_t tile) { return (n + tile - 1) / tile; }
__host__ __device__ inline size_t seg_align(size_t v) { return (v + 255) / 256 * 256; }

struct SegHeader {
    uint32_t error;
    uint32_t counter;
    uint64_t side_rows;     // keys gathered for the global sort
    uint64_t side_unique;   // their unique count
    uint64_t big_tiles;     // tiles longer than kSegCap, listed for seg_run_kernel
    uint64_t deferred;      // tiles too wide for compressed u32 sort keys
};

struct SegLayout {
    size_t status_off, groups_off, bounds_off, list_off, side_off, run_off, total, tiles, groups;
};

// Tiles of a group: a tile adds its unique count to its group's sum, a
// one-block scan turns the group sums into offsets, and the copy of a tile
// adds the counts of the earlier tiles of its group (one wave load).
constexpr uint32_t kSegGroup = 64;

// Metadata of one segmented distinct: header, per-tile unique counts, group
// sums (then group offsets), tile bounds, the list of tiles longer than
// kSegCap (in finish, of the tiles deferred to u64 sort keys) and, for each
// long tile, a flag (1: its keys went to the global sort) and the length of
// its first run.
SegLayout seg_layout(size_t rows, uint32_t tile) {
    SegLayout l{};
    l.tiles = seg_tiles(rows, tile);
    l.groups = seg_tiles(l.tiles, kSegGroup);
    size_t off = seg_align(sizeof(SegHeader));
    l.status_off = off;
    off += seg_align(l.tiles * sizeof(uint64_t));
    l.groups_off = off;
    off += seg_align((l.groups ? l.groups : 1) * sizeof(uint64_t));
    l.bounds_off = off;
    off += seg_align((l.tiles + 1) * sizeof(size_t));
    l.list_off = off;
    off += seg_align((l.tiles ? l.tiles : 1) * sizeof(uint32_t));
    l.side_off = off;
    off += seg_align((l.tiles ? l.tiles : 1) * sizeof(uint8_t));
    l.run_off = off;
    off += seg_align((l.tiles ? l.tiles : 1) * sizeof(uint64_t));
    l.total = off;
    return l;
}

// bounds[t] = start of the run containing position t*tile (or rows);
// the first thread also resets the header and the grid the tile-group sums
// (the per-tile counts need no reset: every tile writes its own).
__global__ void seg_partition_kernel(const uint64_t* keys, size_t rows, uint32_t minor_bits, size_t* bounds,
                                     size_t tiles, SegHeader* h, uint32_t tile, uint64_t* group_sums,
                                     size_t groups) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t g = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < groups; g += stride)
        group_sums[g] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        h->error = 0;
        h->counter = 0;
        h->side_rows = 0;
        h->side_unique = 0;
        h->big_tiles = 0;
        h->deferred = 0;
    }
    for (size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; t <= tiles; t += stride) {
        const size_t p = t * tile;
        if (p >= rows) {
            bounds[t] = rows;
            continue;
        }
        const uint64_t major = keys[p] >> minor_bits;
        size_t lo = 0, hi = p;
        while (lo < hi) {
            const size_t mid = (lo + hi) >> 1;
            if ((keys[mid] >> minor_bits) < major)
                lo = mid + 1;
            else
                hi = mid;
        }
        bounds[t] = lo;
    }
}

__device__ __forceinline__ bool dropped(const uint64_t* found, size_t i) {
    return found != nullptr && ((found[i >> 6] >> (i & 63)) & 1ull);
}

// List the tiles longer than kSegCap (any order), one thread a tile.
__global__ void seg_list_kernel(const size_t* bounds, size_t tiles, SegHeader* h, uint32_t* list) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < tiles; t += stride)
        if (bounds[t + 1] - bounds[t] > static_cast<size_t>(kSegCap))
            list[atomicAdd(reinterpret_cast<unsigned long long*>(&h->big_tiles), 1ull)] = static_cast<uint32_t>(t);
}

// block_search (first index in [lo, hi) where `before` turns false, for a
// predicate true on a prefix) for the 512-thread tile kernel, reducing
// through `red` (kSegWaves u32 of LDS the caller has free) instead of
// __syncthreads_count, which brings LDS of its own.
template <typename Pred>
__device__ __forceinline__ size_t tile_search(size_t lo, size_t hi, Pred before, uint32_t* red) {
    while (lo < hi) {
        const size_t span = hi - lo;
        const size_t stride = span <= static_cast<size_t>(kSegBlock) ? 1 : (span + kSegBlock - 1) / kSegBlock;
        const size_t probe = lo + static_cast<size_t>(threadIdx.x) * stride;
        const uint64_t m = ballot(probe < hi && before(probe));
        if (lane_id() == 0) red[wave_id()] = static_cast<uint32_t>(__popcll(m));
        __syncthreads();
        uint32_t count = 0;
#pragma unroll
        for (int w = 0; w < kSegWaves; ++w) count += red[w];
        __syncthreads();
        if (count == 0) return lo;
        if (stride == 1) return lo + count;
        // answer in (probe[count - 1], probe[count]]
        const size_t last_true = lo + static_cast<size_t>(count - 1) * stride;
        const size_t next = lo + static_cast<size_t>(count) * stride;
        lo = last_true + 1;
        if (next < hi) hi = next;
    }
    return lo;
}

// One tile (at most kSegCap keys, or a long tile's tail) per call, with the
// block's LDS. The sort key is (local run rank << S | minor field): with
// kCompress the field is the bit range [mlb, mlb + mw) where the tile's
// minors differ (S = mw; the bits outside it are shared and put back when
// writing), else the whole minor (S = minor_bits). K is u32 when the rank
// (below kSegCap: 12 bits) and field fit; a compressing block whose field
// is wider than 20 bits lists its tile in `defer` for the u64 list kernel
// and returns.
template <typename K, bool kCompress>
__device__ __forceinline__ void seg_tile_body(uint32_t tile, uint64_t* keys, uint32_t minor_bits,
                                              const size_t* bounds, const uint64_t* side_sorted, uint64_t* counts,
                                              SegHeader* h, const uint64_t* found, uint64_t* group_sums,
                                              const uint8_t* side_tile, const uint64_t* run_rows, uint32_t* defer,
                                              uint64_t* seg_buf, uint16_t (*wave_count)[kSegRadix]) {
    // seg_buf: sort keys and run majors; before ranking, the same bytes hold
    // the tile's surviving raw keys (u64) when some keys are dropped
    K* const tile_keys = reinterpret_cast<K*>(seg_buf);
    uint32_t* const major_of = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(seg_buf) + kSegCap * sizeof(K));
    // wave_count: digit counts per wave (a wave ranks at most 512 keys: u16),
    // 40 KB in all for u32 sort keys, so 4 blocks fit a CU. The small arrays
    // alias space idle while they are used: per-wave totals and the OR/AND
    // of the minors and then of the keys live in wave_count outside the
    // passes, the scan's scratch in tile_keys during a pass's digit scan (the
    // keys are in registers then).
    uint32_t* const wave_tot = reinterpret_cast<uint32_t*>(&wave_count[0][0]);  // kept keys, run starts, unique keys
    K* const red_or = reinterpret_cast<K*>(&wave_count[0][0]) + 2 * kSegWaves;
    K* const red_and = red_or + kSegWaves;
    uint32_t* const scan_lds = reinterpret_cast<uint32_t*>(tile_keys);
    static_assert(sizeof(uint32_t) * kSegWaves + 2 * sizeof(K) * kSegWaves + 2 * sizeof(K) * kSegWaves <=
                      sizeof(uint16_t) * kSegWaves * kSegRadix, "aliases fit");
    static_assert(sizeof(uint16_t) * kSegFirstRadix == sizeof(uint16_t) * kSegWaves * kSegRadix,
                  "first-pass counts fit");
    static_assert(!kCompress || sizeof(K) == sizeof(uint32_t), "compressed keys are u32");
    const uint32_t lane = lane_id(), wave = wave_id();
    const size_t tb0 = bounds[tile], b1 = bounds[tile + 1];
    const bool long_tile = b1 - tb0 > static_cast<size_t>(kSegCap);
    const bool big = long_tile && side_tile[tile];  // its keys went to the global sort
    // a long tile's first run is deduplicated already (seg_run_kernel), its
    // unique rows at the tile's start: this block sorts the tail after it
    // (less than a nominal tile) and writes the tail's rows after them
    const uint64_t base_count = long_tile && !big ? counts[tile] : 0;
    const size_t b0 = long_tile && !big ? tb0 + run_rows[tile] : tb0;  // the first key sorted here
    const size_t out0 = tb0 + base_count;                              // where its unique rows go
    const size_t size = b1 - b0;
    const uint64_t minor_mask = (1ull << minor_bits) - 1;
    // each wave owns a contiguous share of the tile: items = ceil(size / 512)
    int items = static_cast<int>((size + kSegBlock - 1) / kSegBlock);
    uint32_t wave_len = static_cast<uint32_t>(items) * kWave;
    uint32_t n = static_cast<uint32_t>(size);

    K k[kSegItems];
    uint32_t count = 0;        // unique keys of this tile
    size_t side0 = 0;          // big tile: its slice of the side output
    uint64_t kept[kSegItems];  // ballot masks of the unique keys (wave-uniform)
    uint32_t wave_start = 0;   // unique keys in earlier waves
    // sort key = rank << key_shift | ((minor >> mlb) & field); the minor is
    // fixed | ((key & field) << mlb)
    uint32_t key_shift = minor_bits, mlb = 0;
    uint32_t field = static_cast<uint32_t>(minor_mask), fixed = 0;
    if (!big) {
        // load; k = (run starts at or before the position within this wave)
        // << minor_bits | minor; majors of run starts are staged in LDS (in
        // tile_keys, free until the first pass) at their wave-local rank
        uint64_t rawv[kSegItems];
        uint64_t prev_raw = 0;  // the key before this wave's share (lane 0 of item 0)
        // every load of the wave's share is issued before any is used
        // (positions clamped to the tile: no load sits in a branch of its own)
        const uint32_t last = n ? n - 1 : 0;
        uint32_t fw[kSegItems];  // the 32-bit half of the found bitmap holding each key's bit
        if (n > 0) {
#pragma unroll
            for (int i = 0; i < kSegItems; ++i) {
                if (i >= items) continue;
                const uint32_t p = wave * wave_len + i * kWave + lane;
                const size_t g = b0 + (p < n ? p : last);
                rawv[i] = keys[g];
                if (found != nullptr) fw[i] = reinterpret_cast<const uint32_t*>(found)[g >> 5];
            }
        }
        if (found == nullptr) {
#pragma unroll
            for (int i = 0; i < kSegItems; ++i) {
                if (i >= items) continue;
                const uint32_t p = wave * wave_len + i * kWave + lane;
                if (p >= n) rawv[i] = 0;
            }
            const uint32_t p0 = wave * wave_len;
            if (p0 > 0 && p0 < n) prev_raw = keys[b0 + p0 - 1];
        } else {
            // keep only the keys whose found bit is clear, in order; the
            // ballot masks stay in (scalar) registers, and only the waves'
            // totals go through LDS
            uint64_t keep[kSegItems];
            uint32_t mine = 0;
#pragma unroll
            for (int i = 0; i < kSegItems; ++i) {
                keep[i] = 0;
                if (i >= items) continue;
                const uint32_t p = wave * wave_len + i * kWave + lane;
                const bool valid = p < n;
                const size_t g = b0 + p;
                keep[i] = ballot(valid && !((fw[i] >> (g & 31)) & 1u));
                mine += static_cast<uint32_t>(__popcll(keep[i]));
            }
            if (lane == 0) wave_tot[wave] = mine;
            __syncthreads();
            uint32_t pos = 0, survivors = 0;
#pragma unroll
            for (int w = 0; w < kSegWaves; ++w) {
                const uint32_t c = wave_tot[w];
                if (w < static_cast<int>(wave)) pos += c;
                survivors += c;
            }
#pragma unroll
            for (int i = 0; i < kSegItems; ++i) {
                if (i >= items) continue;
                const uint64_t m = keep[i];
                if ((m >> lane) & 1ull) seg_buf[pos + lanes_below(m)] = rawv[i];
                pos += static_cast<uint32_t>(__popcll(m));
            }
            __syncthreads();
            n = survivors;
            items = static_cast<int>((n + kSegBlock - 1) / kSegBlock);
            wave_len = static_cast<uint32_t>(items) * kWave;
#pragma unroll
            for (int i = 0; i < kSegItems; ++i) {
                if (i >= items) continue;
                const uint32_t p = wave * wave_len + i * kWave + lane;
                rawv[i] = p < n ? seg_buf[p] : 0;
            }
            const uint32_t p0 = wave * wave_len;
            if (p0 > 0 && p0 < n) prev_raw = seg_buf[p0 - 1];
            __syncthreads();  // seg_buf now holds the sort keys and run majors
        }
        // with kCompress: the bit range where the kept minors differ, from
        // their OR and AND (through red_or / red_and, which the keys' own
        // reduction reuses only after the run starts' barrier below)
        if constexpr (kCompress) {
            uint32_t o = 0, a = 0xFFFFFFFFu;
#pragma unroll
            for (int i = 0; i < kSegItems; ++i) {
                if (i >= items) continue;
                const uint32_t p = wave * wave_len + i * kWave + lane;
                if (p < n) {
                    const uint32_t m = static_cast<uint32_t>(rawv[i]) & static_cast<uint32_t>(minor_mask);
                    o |= m;
                    a &= m;
                }
            }
#pragma unroll
            for (int off = kWave / 2; off > 0; off >>= 1) {
                o |= __shfl_xor(o, off, kWave);
                a &= __shfl_xor(a, off, kWave);
            }
            if (lane == 0) {
                red_or[wave] = o;
                red_and[wave] = a;
            }
            __syncthreads();
            uint32_t mo = 0, ma = 0xFFFFFFFFu;
#pragma unroll
            for (int w = 0; w < kSegWaves; ++w) {
                mo |= red_or[w];
                ma &= red_and[w];
            }
            const uint32_t vary = n ? (mo ^ ma) : 0u;
            mlb = vary ? static_cast<uint32_t>(__builtin_ctz(vary)) : 0u;
            const uint32_t mw = vary ? 32u - static_cast<uint32_t>(__builtin_clz(vary)) - mlb : 0u;
            if (mw + kSegRankBits > 32) {  // block-uniform: the u64 list kernel sorts this tile
                if (threadIdx.x == 0)
                    defer[atomicAdd(reinterpret_cast<unsigned long long*>(&h->deferred), 1ull)] = tile;
                return;
            }
            key_shift = mw;
            field = (1u << mw) - 1u;
            fixed = ma & ~(field << mlb);  // the minor bits every kept key shares
        }
        uint32_t* staged = reinterpret_cast<uint32_t*>(tile_keys) + wave * (kSegCap / kSegWaves);
        uint32_t starts_before = 0;  // run starts in earlier items of this wave
        uint32_t last_major = 0;     // major of lane 63 of the previous item
#pragma unroll
        for (int i = 0; i < kSegItems; ++i) {
            if (i >= items) continue;
            const uint32_t p = wave * wave_len + i * kWave + lane;
            const bool valid = p < n;
            const uint64_t raw = rawv[i];
            const uint32_t major = static_cast<uint32_t>(raw >> minor_bits);
            const uint32_t up = __shfl_up(major, 1, kWave);
            bool start = false;
            if (valid) {
                if (p == 0)
                    start = true;
                else if (lane > 0)
                    start = up != major;
                else if (i > 0)
                    start = last_major != major;
                else
                    start = static_cast<uint32_t>(prev_raw >> minor_bits) != major;
            }
            last_major = __shfl(major, kWave - 1, kWave);
            const uint64_t m = ballot(start);
            const uint32_t local = starts_before + lanes_below(m) + (start ? 1u : 0u);
            if (start) staged[local - 1] = major;
            k[i] = (static_cast<K>(local) << key_shift) | static_cast<K>((static_cast<uint32_t>(raw) >> mlb) & field);
            starts_before += static_cast<uint32_t>(__popcll(m));
        }
        if (lane == 0) wave_tot[wave] = starts_before;
        __syncthreads();
        uint32_t wave_base = 0;
        for (int w = 0; w < static_cast<int>(wave); ++w) wave_base += wave_tot[w];
        for (uint32_t j = lane; j < starts_before; j += kWave) major_of[wave_base + j] = staged[j];
        // rank = wave_base + local - 1 (unsigned wrap-around is exact)
        const K rank_add = static_cast<K>(static_cast<K>(static_cast<K>(wave_base) - static_cast<K>(1)) << key_shift);
        K or_acc = 0, and_acc = static_cast<K>(~static_cast<K>(0));
#pragma unroll
        for (int i = 0; i < kSegItems; ++i) {
            if (i >= items) continue;
            const uint32_t p = wave * wave_len + i * kWave + lane;
            if (p < n) {
                k[i] = static_cast<K>(k[i] + rank_add);
                or_acc |= k[i];
                and_acc &= k[i];
            }
        }
        // varying digits of the tile's (rank, minor) keys
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) {
            or_acc |= __shfl_xor(or_acc, off, kWave);
            and_acc &= __shfl_xor(and_acc, off, kWave);
        }
        if (lane == 0) {
            red_or[wave] = or_acc;
            red_and[wave] = and_acc;
        }
        __syncthreads();
        K vary = 0;
        {
            K o = 0, a = static_cast<K>(~static_cast<K>(0));
            for (int w = 0; w < kSegWaves; ++w) {
                o |= red_or[w];
                a &= red_and[w];
            }
            vary = n ? (o ^ a) : 0;
        }
        // LSD passes in LDS over the bit range [lb, lb + width) where the keys
        // differ (a last digit reaching past the range only adds bits that
        // are equal in every key of the tile). The first pass need not be
        // stable (the order it starts from carries no information): it ranks
        // with block-wide LDS atomics on a 12-bit digit, whose 4,096 block
        // counts take the same 8 KB as the later passes' per-wave counts; the
        // later passes are stable, on 9-bit digits. Up to 21 varying bits
        // take two passes, up to 30 three.
        int lb = 0, width = 0;
        if (vary) {
            lb = __builtin_ctzll(static_cast<uint64_t>(vary));
            width = 64 - __builtin_clzll(static_cast<uint64_t>(vary)) - lb;
        }
        const int passes =
            width == 0 ? 0
                       : 1 + (width > kSegFirstBits ? (width - kSegFirstBits + kSegDigitBits - 1) / kSegDigitBits : 0);
        __syncthreads();  // red_or / red_and (in wave_count) read by every thread
        uint32_t* const words = reinterpret_cast<uint32_t*>(&wave_count[0][0]);  // two u16 counts a word
        for (int pass = 0; pass < passes; ++pass) {  // block-uniform
            for (int i = threadIdx.x; i < kSegWaves * kSegRadix / 2; i += kSegBlock) words[i] = 0;
            __syncthreads();
            uint32_t ranks[kSegItems];
            if (pass == 0) {
                // rank within the digit over the whole block, from an LDS atomic
#pragma unroll
                for (int i = 0; i < kSegItems; ++i) {
                    if (i >= items) continue;
                    const uint32_t p = wave * wave_len + i * kWave + lane;
                    const uint32_t digit = static_cast<uint32_t>(k[i] >> lb) & (kSegFirstRadix - 1);
                    if (p < n) {
                        const uint32_t sh = 16u * (digit & 1u);
                        ranks[i] = (atomicAdd(&words[digit >> 1], 1u << sh) >> sh) & 0xFFFFu;
                    }
                }
                __syncthreads();
                // thread t owns digits [8t, 8t + 8): their counts become
                // exclusive starts in place (every count and start is at most
                // kSegCap, so the u16 halves never carry into each other)
                static_assert(kSegFirstRadix == 8 * kSegBlock, "eight digits per thread");
                uint4 q = reinterpret_cast<uint4*>(words)[threadIdx.x];
                const uint32_t pairs = q.x + q.y + q.z + q.w;
                uint32_t total_unused;
                uint32_t run = block_exclusive_sum<uint32_t, kSegWaves>((pairs & 0xFFFFu) + (pairs >> 16),
                                                                        &total_unused, scan_lds);
                const auto to_starts = [&run](uint32_t w) {
                    const uint32_t lo = w & 0xFFFFu;
                    const uint32_t out = run | ((run + lo) << 16);
                    run += lo + (w >> 16);
                    return out;
                };
                q.x = to_starts(q.x);
                q.y = to_starts(q.y);
                q.z = to_starts(q.z);
                q.w = to_starts(q.w);
                reinterpret_cast<uint4*>(words)[threadIdx.x] = q;
                __syncthreads();
#pragma unroll
                for (int i = 0; i < kSegItems; ++i) {
                    if (i >= items) continue;
                    const uint32_t p = wave * wave_len + i * kWave + lane;
                    if (p < n) {
                        const uint32_t digit = static_cast<uint32_t>(k[i] >> lb) & (kSegFirstRadix - 1);
                        const uint32_t at = (words[digit >> 1] >> (16u * (digit & 1u))) & 0xFFFFu;
                        tile_keys[at + ranks[i]] = k[i];
                    }
                }
            } else {
                const int shift = lb + kSegFirstBits + (pass - 1) * kSegDigitBits;
#pragma unroll
                for (int i = 0; i < kSegItems; ++i) {
                    if (i >= items) continue;
                    const uint32_t p = wave * wave_len + i * kWave + lane;
                    const bool valid = p < n;
                    const uint32_t digit = static_cast<uint32_t>(k[i] >> shift) & (kSegRadix - 1);
                    const uint64_t peers = match_digit<kSegDigitBits>(digit, valid);
                    const uint32_t before = wave_count[wave][digit];  // any lane: digit < kSegRadix
                    const uint32_t below = lanes_below(peers);
                    ranks[i] = before + below;
                    if (valid && below == 0)
                        wave_count[wave][digit] = static_cast<uint16_t>(before + static_cast<uint32_t>(__popcll(peers)));
                }
                __syncthreads();
                // digit d's start per wave: thread d owns digit d
                static_assert(kSegRadix == kSegBlock, "one digit per thread");
                const uint32_t d = threadIdx.x;
                uint32_t tc = 0;
                {
#pragma unroll
                    for (int w = 0; w < kSegWaves; ++w) {
                        const uint32_t c = wave_count[w][d];
                        wave_count[w][d] = static_cast<uint16_t>(tc);
                        tc += c;
                    }
                }
                uint32_t total_unused;
                const uint32_t local_start = block_exclusive_sum<uint32_t, kSegWaves>(tc, &total_unused, scan_lds);
                {
#pragma unroll
                    for (int w = 0; w < kSegWaves; ++w)
                        wave_count[w][d] = static_cast<uint16_t>(wave_count[w][d] + local_start);
                }
                __syncthreads();
#pragma unroll
                for (int i = 0; i < kSegItems; ++i) {
                    if (i >= items) continue;
                    const uint32_t p = wave * wave_len + i * kWave + lane;
                    if (p < n) {
                        const uint32_t digit = static_cast<uint32_t>(k[i] >> shift) & (kSegRadix - 1);
                        tile_keys[wave_count[wave][digit] + ranks[i]] = k[i];
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < kSegItems; ++i) {
                if (i >= items) continue;
                // p < items * 512 <= kSegCap for every lane; lanes past n
                // read keys nobody uses
                k[i] = tile_keys[wave * wave_len + i * kWave + lane];
            }
            __syncthreads();
        }
        // sorted keys by position; also in LDS for the predecessor of lane 0
#pragma unroll
        for (int i = 0; i < kSegItems; ++i) {
            if (i >= items) continue;
            const uint32_t p = wave * wave_len + i * kWave + lane;
            if (p < n) tile_keys[p] = k[i];
        }
        __syncthreads();
        uint32_t mine = 0;
#pragma unroll
        for (int i = 0; i < kSegItems; ++i) {
            kept[i] = 0;
            if (i >= items) continue;
            const uint32_t p = wave * wave_len + i * kWave + lane;
            const bool valid = p < n;
            const K prev = p > 0 && valid ? tile_keys[p - 1] : static_cast<K>(~k[i]);
            kept[i] = ballot(valid && (p == 0 || prev != k[i]));
            mine += static_cast<uint32_t>(__popcll(kept[i]));
        }
        if (lane == 0) wave_tot[wave] = mine;
        __syncthreads();
#pragma unroll
        for (int w = 0; w < kSegWaves; ++w) {
            const uint32_t c = wave_tot[w];
            if (w < static_cast<int>(wave)) wave_start += c;
            count += c;
        }
    } else {
        // oversized tile: its unique keys are a slice of the globally sorted side output
        const uint64_t first_major = keys[b0] >> minor_bits;
        const uint64_t last_major = keys[b1 - 1] >> minor_bits;
        const size_t n_side = h->side_unique;
        side0 = tile_search(
            0, n_side, [&](size_t i) { return (side_sorted[i] >> minor_bits) < first_major; }, wave_tot);
        const size_t side1 = tile_search(
            side0, n_side, [&](size_t i) { return (side_sorted[i] >> minor_bits) <= last_major; }, wave_tot);
        count = static_cast<uint32_t>(side1 - side0);
    }
    // the tile's unique rows go back over its own input slice (every load of
    // the slice is done: the write needs no other tile), expanded to pairs;
    // seg_write_kernel moves them to their offsets once every count is known
    if (threadIdx.x == 0) {
        counts[tile] = base_count + count;
        atomicAdd(reinterpret_cast<unsigned long long*>(&group_sums[tile / kSegGroup]),
                  static_cast<unsigned long long>(base_count + count));
    }
    if (!big) {
        uint32_t pos = wave_start;
#pragma unroll
        for (int i = 0; i < kSegItems; ++i) {
            if (i >= items) continue;
            const uint64_t kept_i = kept[i];
            if ((kept_i >> lane) & 1ull) {
                const uint32_t rank = static_cast<uint32_t>(k[i] >> key_shift);
                const uint32_t major = major_of[rank];
                keys[out0 + pos + lanes_below(kept_i)] =
                    (static_cast<uint64_t>(major) << 32) | (fixed | ((static_cast<uint32_t>(k[i]) & field) << mlb));
            }
            pos += static_cast<uint32_t>(__popcll(kept_i));
        }
    }
}

// One tile per block (blockIdx.x < tiles).
template <typename K, bool kCompress>
__global__ __launch_bounds__(kSegBlock) __attribute__((amdgpu_waves_per_eu(8))) void seg_tile_kernel(
    uint64_t* keys, size_t tiles, uint32_t minor_bits, const size_t* bounds, const uint64_t* side_sorted,
    uint64_t* counts, SegHeader* h, const uint64_t* found, uint64_t* group_sums, const uint8_t* side_tile,
    const uint64_t* run_rows, uint32_t* defer) {
    __shared__ uint64_t seg_buf[(kSegCap * (sizeof(K) + sizeof(uint32_t)) + 7) / 8];
    static_assert(sizeof(seg_buf) >= kSegCap * sizeof(uint64_t), "raw keys fit");
    __shared__ __attribute__((aligned(16))) uint16_t wave_count[kSegWaves][kSegRadix];
    // tiles are independent (no look-back), so dispatch order needs no ticket
    if (blockIdx.x >= tiles) return;
    seg_tile_body<K, kCompress>(blockIdx.x, keys, minor_bits, bounds, side_sorted, counts, h, found, group_sums,
                                side_tile, run_rows, defer, seg_buf, wave_count);
}

// The tiles a compressing launch deferred (u64 sort keys), one at a time per
// block of a persistent grid.
__global__ __launch_bounds__(kSegBlock) void seg_tile_list_kernel(uint64_t* keys, uint32_t minor_bits,
                                                                 const size_t* bounds, const uint64_t* side_sorted,
                                                                 uint64_t* counts, SegHeader* h,
                                                                 const uint64_t* found, uint64_t* group_sums,
                                                                 const uint8_t* side_tile, const uint64_t* run_rows,
                                                                 const uint32_t* list) {
    __shared__ uint64_t seg_buf[(kSegCap * (sizeof(uint64_t) + sizeof(uint32_t)) + 7) / 8];
    __shared__ __attribute__((aligned(16))) uint16_t wave_count[kSegWaves][kSegRadix];
    const uint64_t n_list = h->deferred;
    for (uint64_t j = blockIdx.x; j < n_list; j += gridDim.x) {
        seg_tile_body<uint64_t, false>(list[j], keys, minor_bits, bounds, side_sorted, counts, h, found, group_sums,
                                       side_tile, run_rows, nullptr, seg_buf, wave_count);
        __syncthreads();  // LDS is reused by the block's next tile
    }
}

// The first runs of the long tiles (listed by seg_list_kernel), one tile at
// a time per block of a persistent grid. A run has one major; its kept keys
// (found bit clear) set the bit of their minor's low kLowBits bits in a
// bitmap in LDS while the block checks that their minors agree above those
// bits. If they do, the bitmap's set bits, in order, are the run's sorted
// unique minors: they are written, expanded, over the start of the tile
// (every key of the tile is read by then), their count to counts[tile] and
// the run's length (kept or not) to run_rows[tile], for seg_tile_kernel to
// sort the tail after it. Otherwise every kept key of the tile is copied to
// the side buffer for the global sort and the tile is flagged in side_tile.
template <uint32_t kLowBits>
__global__ __launch_bounds__(kSegBlock) void seg_run_kernel(uint64_t* keys, uint32_t minor_bits, const size_t* bounds,
                                                           const uint32_t* list, SegHeader* h, uint64_t* side,
                                                           uint8_t* side_tile, uint64_t* counts, uint64_t* run_rows,
                                                           const uint64_t* found) {
    constexpr uint32_t kRunWords = (1u << kLowBits) / 32;
    __shared__ uint32_t bm[kRunWords];
    __shared__ uint32_t red_len[kSegWaves], red_or[kSegWaves], red_and[kSegWaves], red_min[kSegWaves],
        red_max[kSegWaves];
    __shared__ uint32_t scan_lds[kSegWaves];
    __shared__ uint64_t side_slot;
    constexpr uint32_t kChunk = kSegBlock * kRunItems;
    const uint64_t n_list = h->big_tiles;
    const uint32_t lane = lane_id(), wave = wave_id();
    const uint32_t minor_mask = static_cast<uint32_t>((1ull << minor_bits) - 1);
    const uint32_t low_mask = (1u << kLowBits) - 1;
    const uint32_t* const found32 = reinterpret_cast<const uint32_t*>(found);
    // the bitmap is cleared once here, and after each tile only between the
    // words of its least and greatest set bit
    if (blockIdx.x < n_list)
        for (uint32_t i = threadIdx.x; i < kRunWords; i += kSegBlock) bm[i] = 0;
    __syncthreads();
    for (uint64_t j = blockIdx.x; j < n_list; j += gridDim.x) {
        const uint32_t t = list[j];
        const size_t b0 = bounds[t];
        const size_t size = bounds[t + 1] - b0;  // > kSegCap
        const uint64_t major0 = keys[b0] >> minor_bits;
        // the run is a prefix of the tile: count its keys, set the kept
        // keys' bits, and gather the OR/AND of their high minor bits and the
        // least and greatest set bit; chunks of kChunk keys, every load of a
        // chunk issued before any is used
        uint32_t len = 0, hi_or = 0, hi_and = 0xFFFFFFFFu, lo_min = 0xFFFFFFFFu, lo_max = 0;
        for (size_t c = 0; c < size; c += kChunk) {
            uint64_t k[kRunItems];
            uint32_t fw[kRunItems];
#pragma unroll
            for (int i = 0; i < kRunItems; ++i) {
                const size_t p = c + static_cast<size_t>(i) * kSegBlock + threadIdx.x;
                const size_t g = b0 + (p < size ? p : size - 1);
                k[i] = keys[g];
                fw[i] = found32 != nullptr ? found32[g >> 5] : 0u;
            }
#pragma unroll
            for (int i = 0; i < kRunItems; ++i) {
                const size_t p = c + static_cast<size_t>(i) * kSegBlock + threadIdx.x;
                const bool in_run = p < size && (k[i] >> minor_bits) == major0;
                len += static_cast<uint32_t>(__popcll(ballot(in_run)));
                if (in_run && !((fw[i] >> ((b0 + p) & 31)) & 1u)) {
                    const uint32_t minor = static_cast<uint32_t>(k[i]) & minor_mask;
                    atomicOr(&bm[(minor & low_mask) >> 5], 1u << (minor & 31));
                    hi_or |= minor >> kLowBits;
                    hi_and &= minor >> kLowBits;
                    lo_min = min(lo_min, minor & low_mask);
                    lo_max = max(lo_max, minor & low_mask);
                }
            }
        }
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) {
            hi_or |= __shfl_xor(hi_or, off, kWave);
            hi_and &= __shfl_xor(hi_and, off, kWave);
            lo_min = min(lo_min, static_cast<uint32_t>(__shfl_xor(lo_min, off, kWave)));
            lo_max = max(lo_max, static_cast<uint32_t>(__shfl_xor(lo_max, off, kWave)));
        }
        if (lane == 0) {
            red_len[wave] = len;  // wave-uniform
            red_or[wave] = hi_or;
            red_and[wave] = hi_and;
            red_min[wave] = lo_min;
            red_max[wave] = lo_max;
        }
        __syncthreads();
        uint32_t run_len = 0, ho = 0, ha = 0xFFFFFFFFu, lo = 0xFFFFFFFFu, hi = 0;
#pragma unroll
        for (int w = 0; w < kSegWaves; ++w) {
            run_len += red_len[w];
            ho |= red_or[w];
            ha &= red_and[w];
            lo = min(lo, red_min[w]);
            hi = max(hi, red_max[w]);
        }
        if (ho & ~ha) {  // block-uniform: kept minors differ above their low kLowBits bits
            // the whole tile's kept keys to the side buffer (any order), a
            // chunk at a time: ballot compaction per item, one atomic a chunk
            for (size_t c = 0; c < size; c += kChunk) {
                uint64_t k[kRunItems], m[kRunItems];
                uint32_t mine = 0;
#pragma unroll
                for (int i = 0; i < kRunItems; ++i) {
                    const size_t p = c + static_cast<size_t>(i) * kSegBlock + threadIdx.x;
                    k[i] = keys[b0 + (p < size ? p : size - 1)];
                }
#pragma unroll
                for (int i = 0; i < kRunItems; ++i) {
                    const size_t p = c + static_cast<size_t>(i) * kSegBlock + threadIdx.x;
                    m[i] = ballot(p < size && !dropped(found, b0 + p));
                    mine += static_cast<uint32_t>(__popcll(m[i]));
                }
                if (lane == 0) red_len[wave] = mine;
                __syncthreads();
                uint32_t before = 0, total = 0;
#pragma unroll
                for (int w = 0; w < kSegWaves; ++w) {
                    if (w < static_cast<int>(wave)) before += red_len[w];
                    total += red_len[w];
                }
                if (threadIdx.x == 0)
                    side_slot = atomicAdd(reinterpret_cast<unsigned long long*>(&h->side_rows),
                                          static_cast<unsigned long long>(total));
                __syncthreads();
                uint64_t at = side_slot + before;
#pragma unroll
                for (int i = 0; i < kRunItems; ++i) {
                    if ((m[i] >> lane) & 1ull) side[at + lanes_below(m[i])] = k[i];
                    at += static_cast<uint32_t>(__popcll(m[i]));
                }
                __syncthreads();  // red_len and side_slot are rewritten by the next chunk
            }
            if (threadIdx.x == 0) side_tile[t] = 1;
        } else {
            // the set bits in order: of the words from the least set bit's
            // to the greatest's, thread i takes [i * per, i * per + per),
            // counts their bits, and after one scan writes them
            const uint64_t high = (static_cast<uint64_t>(major0) << 32) | (static_cast<uint64_t>(ho) << kLowBits);
            const uint32_t w_first = lo <= hi ? lo >> 5 : 0;
            const uint32_t words_used = lo <= hi ? (hi >> 5) + 1 : 0;  // one past the last word
            const uint32_t per = (words_used - w_first + kSegBlock - 1) / kSegBlock;
            const uint32_t w0 = w_first + threadIdx.x * per;
            uint32_t mine = 0;
            for (uint32_t q = 0; q < per; ++q)
                if (w0 + q < words_used) mine += static_cast<uint32_t>(__popc(bm[w0 + q]));
            uint32_t out;
            uint32_t at = block_exclusive_sum<uint32_t, kSegWaves>(mine, &out, scan_lds);
            for (uint32_t q = 0; q < per; ++q) {
                const uint32_t w = w0 + q;
                uint32_t bits = w < words_used ? bm[w] : 0u;
                while (bits) {
                    keys[b0 + at++] = high | ((w << 5) | static_cast<uint32_t>(__builtin_ctz(bits)));
                    bits &= bits - 1;
                }
            }
            if (threadIdx.x == 0) {
                side_tile[t] = 0;
                counts[t] = out;
                run_rows[t] = run_len;
            }
        }
        __syncthreads();  // every read of the bitmap is done
        if (lo <= hi)
            for (uint32_t w = (lo >> 5) + threadIdx.x; w <= (hi >> 5); w += kSegBlock) bm[w] = 0;
        __syncthreads();  // the bitmap and the reductions are reused by the block's next tile
    }
}

// Exclusive scan, in place, of the tile groups' unique counts (one block,
// n >= 1), and the total.
constexpr int kSegScanThreads = 1024;
__global__ __launch_bounds__(kSegScanThreads) void seg_scan_kernel(uint64_t* counts, size_t n, uint64_t* total) {
    __shared__ uint64_t scan_lds[kSegScanThreads / kWave];
    block_scan_inplace<kSegScanThreads, 4>(counts, n, total, scan_lds);
}

// Each tile's unique rows to their offset: NEWT (and the other outputs), and
// the output's first-column index at the first row of each first value. A
// tile's rows are its input slice (written by seg_tile_kernel), or for an
// oversized tile its slice of the globally sorted side output. Loads of a
// chunk are all issued before any is used (positions clamped to the tile).
constexpr int kSegWriteItems = 8;
__global__ __launch_bounds__(kBlock) void seg_write_kernel(const uint64_t* keys, const size_t* bounds,
                                                           const uint64_t* side_sorted, const uint64_t* counts,
                                                           const uint64_t* group_offsets, const SegHeader* h,
                                                           size_t tiles, uint32_t minor_bits, PairOutput out,
                                                           uint32_t* xidx_out, const uint8_t* side_tile,
                                                           InstanceMarks marks) {
    __shared__ uint64_t offset_slot;
    const uint64_t minor_mask = (1ull << minor_bits) - 1;
    const uint32_t lane = lane_id();
    // the instance this wave is summing, and this lane's share of its sums
    uint32_t stat_inst = 0xFFFFFFFFu;
    unsigned long long stat_digest = 0, stat_rows = 0;
    const size_t stat_slot = static_cast<size_t>(blockIdx.x % kInstanceStatSlots) * marks.instances;
    const auto flush_stats = [&]() {  // wave-uniform
        const unsigned long long d = wave_sum(stat_digest), r = wave_sum(stat_rows);
        if (lane == 0 && stat_inst < marks.instances && r) {
            atomicAdd(&marks.digest_acc[stat_slot + stat_inst], d);
            atomicAdd(&marks.rows_acc[stat_slot + stat_inst], r);
        }
        stat_digest = stat_rows = 0;
    };
    for (size_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint64_t c = counts[t];
        if (c == 0) continue;  // block-uniform
        // the tile's offset: its group's, plus the earlier tiles of the group
        if (wave_id() == 0) {
            const size_t g0 = t / kSegGroup * kSegGroup;
            const uint64_t before = g0 + lane < t ? counts[g0 + lane] : 0;
            const uint64_t o = group_offsets[t / kSegGroup] + wave_sum(before);
            if (lane == 0) offset_slot = o;
        }
        __syncthreads();
        const uint64_t o = offset_slot;
        const size_t b0 = bounds[t], b1 = bounds[t + 1];
        const bool big = b1 - b0 > static_cast<size_t>(kSegCap) && side_tile[t];
        const uint64_t* src = keys + b0;
        if (big) {
            // its unique keys are the side output's rows of its first values
            const uint64_t first_major = keys[b0] >> minor_bits;
            const size_t n_side = h->side_unique;
            const size_t side0 = block_search(
                0, n_side, [&](size_t i) { return (side_sorted[i] >> minor_bits) < first_major; });
            src = side_sorted + side0;
        }
        for (uint64_t c0 = 0; c0 < c; c0 += static_cast<uint64_t>(kBlock) * kSegWriteItems) {
            // a row's predecessor is the previous lane's row; lane 0 loads its own
            uint64_t v[kSegWriteItems], pv[kSegWriteItems];
#pragma unroll
            for (int q = 0; q < kSegWriteItems; ++q) {
                const uint64_t i = c0 + threadIdx.x + static_cast<uint64_t>(q) * kBlock;
                const uint64_t ic = i < c ? i : c - 1;
                v[q] = src[ic];
                pv[q] = lane == 0 ? src[ic > 0 ? ic - 1 : 0] : 0;
            }
#pragma unroll
            for (int q = 0; q < kSegWriteItems; ++q) {
                const uint64_t up = __shfl_up(v[q], 1, kWave);
                if (lane > 0) pv[q] = up;
            }
#pragma unroll
            for (int q = 0; q < kSegWriteItems; ++q) {
                const uint64_t i = c0 + threadIdx.x + static_cast<uint64_t>(q) * kBlock;
                if (i >= c) continue;
                uint64_t k = v[q], pk = pv[q];
                if (big) {
                    k = ((k >> minor_bits) << 32) | (k & minor_mask);
                    pk = ((pk >> minor_bits) << 32) | (pk & minor_mask);
                }
                store_pair(out, o + i, k);
                // the first output row of its first value starts that value's
                // segment of the output's first-column index
                if (xidx_out && (i == 0 || (pk >> 32) != (k >> 32)))
                    xidx_out[k >> 32] = static_cast<uint32_t>(o + i);
                // a tile's first row and each row whose instance differs from
                // its predecessor's mark their instance (rows are sorted)
                if (marks.last_new) {
                    const uint32_t inst = static_cast<uint32_t>(k >> 32) >> marks.bits;
                    if (inst < marks.instances &&
                        (i == 0 || (static_cast<uint32_t>(pk >> 32) >> marks.bits) != inst))
                        atomicMax(&marks.last_new[inst], marks.value);
                }
            }
            if (marks.rows_acc) {
                // per-instance rows and digest terms: a wave sums an item's
                // 64 rows while they belong to the instance it is summing
                // (valid lanes are a prefix of the wave), and adds the sums
                // when the instance changes or the tile ends
#pragma unroll
                for (int q = 0; q < kSegWriteItems; ++q) {
                    const uint64_t i = c0 + threadIdx.x + static_cast<uint64_t>(q) * kBlock;
                    const bool valid = i < c;
                    uint64_t k = v[q];
                    if (big) k = ((k >> minor_bits) << 32) | (k & minor_mask);
                    const uint32_t x = static_cast<uint32_t>(k >> 32);
                    const uint32_t inst = valid ? x >> marks.bits : 0xFFFFFFFFu;
                    if (ballot(valid) == 0) continue;  // wave-uniform
                    const uint32_t lead = __builtin_amdgcn_readfirstlane(inst);
                    const unsigned long long term =
                        valid ? instance_pair_hash(x, static_cast<uint32_t>(k), marks.bits) : 0ull;
                    if (ballot(valid && inst != lead) == 0) {
                        if (lead != stat_inst) {
                            flush_stats();
                            stat_inst = lead;
                        }
                        stat_digest += term;
                        stat_rows += valid ? 1ull : 0ull;
                    } else {
                        flush_stats();
                        stat_inst = 0xFFFFFFFFu;
                        if (valid && inst < marks.instances) {
                            atomicAdd(&marks.digest_acc[stat_slot + inst], term);
                            atomicAdd(&marks.rows_acc[stat_slot + inst], 1ull);
                        }
                    }
                }
            }
        }
        if (marks.rows_acc) {
            flush_stats();
            stat_inst = 0xFFFFFFFFu;
        }
        __syncthreads();  // offset_slot is rewritten for the block's next tile
    }
}

}  // namespace

size_t segmented_distinct_meta_bytes(size_t rows, uint32_t tile_rows) { return seg_layout(rows, tile_rows).total; }

hipError_t segmented_distinct_prepare(uint64_t* keys, size_t rows, uint32_t minor_bits, uint64_t* side,
                                      void* meta, size_t meta_bytes, uint64_t* side_rows, hipStream_t stream,
                                      const uint64_t* found, uint32_t tile_rows) {
    *side_rows = 0;
    if (rows == 0) return hipSuccess;
    if (tile_rows < 512 || tile_rows > static_cast<uint32_t>(kSegCap)) return hipErrorInvalidValue;
    const SegLayout l = seg_layout(rows, tile_rows);
    if (meta_bytes < l.total || minor_bits == 0 || minor_bits > 32) return hipErrorInvalidValue;
    char* base = static_cast<char*>(meta);
    SegHeader* h = reinterpret_cast<SegHeader*>(base);
    size_t* bounds = reinterpret_cast<size_t*>(base + l.bounds_off);
    const unsigned small_grid = static_cast<unsigned>(l.tiles / kBlock + 1 < 4096 ? l.tiles / kBlock + 1 : 4096);
    uint64_t* groups = reinterpret_cast<uint64_t*>(base + l.groups_off);
    seg_partition_kernel<<<small_grid, kBlock, 0, stream>>>(keys, rows, minor_bits, bounds, l.tiles, h, tile_rows,
                                                            groups, l.groups);
    uint32_t* list = reinterpret_cast<uint32_t*>(base + l.list_off);
    uint8_t* side_tile = reinterpret_cast<uint8_t*>(base + l.side_off);
    seg_list_kernel<<<small_grid, kBlock, 0, stream>>>(bounds, l.tiles, h, list);
    // long tiles: at most rows / (kSegCap + 1) of them
    if (rows > static_cast<size_t>(kSegCap)) {
        const size_t most = rows / (kSegCap + 1);
        uint64_t* counts = reinterpret_cast<uint64_t*>(base + l.status_off);
        uint64_t* run_rows = reinterpret_cast<uint64_t*>(base + l.run_off);
        if (minor_bits > 18 && minor_bits <= 20) {
            const unsigned grid = static_cast<unsigned>(most < 256 ? most : 256);  // one block a CU
            seg_run_kernel<20><<<grid, kSegBlock, 0, stream>>>(keys, minor_bits, bounds, list, h, side, side_tile,
                                                               counts, run_rows, found);
        } else {
            const unsigned grid = static_cast<unsigned>(most < 1024 ? most : 1024);  // four blocks a CU
            seg_run_kernel<18><<<grid, kSegBlock, 0, stream>>>(keys, minor_bits, bounds, list, h, side, side_tile,
                                                               counts, run_rows, found);
        }
    }
    hipError_t st = hipGetLastError();
    // minors of at most 20 bits fit the run bitmap whole: no key goes to
    // the global sort, and the host need not wait
    if (minor_bits <= 20) return st;
    // rows for the global sort (usually none): one 8-byte read sizes it
    if (st == hipSuccess) st = hipMemcpyAsync(side_rows, &h->side_rows, sizeof(uint64_t), hipMemcpyDeviceToHost, stream);
    if (st == hipSuccess) st = hipStreamSynchronize(stream);
    return st;
}

hipError_t segmented_distinct_finish(uint64_t* keys, size_t rows, uint32_t minor_bits, uint64_t side_varying,
                                     uint64_t* side_a, uint64_t* side_b, uint64_t side_rows, PairOutput out,
                                     uint64_t* unique_rows, void* meta, size_t meta_bytes, void* sort_temp,
                                     size_t sort_bytes, hipStream_t stream, const KernelTimer* timer,
                                     const uint64_t* found, uint32_t* xidx_out, uint32_t tile_rows,
                                     InstanceMarks marks) {
    if (rows == 0) return set_u64(unique_rows, 0, stream);
    if (tile_rows < 512 || tile_rows > static_cast<uint32_t>(kSegCap)) return hipErrorInvalidValue;
    const SegLayout l = seg_layout(rows, tile_rows);
    if (meta_bytes < l.total) return hipErrorInvalidValue;
    char* base = static_cast<char*>(meta);
    SegHeader* h = reinterpret_cast<SegHeader*>(base);
    uint64_t* status = reinterpret_cast<uint64_t*>(base + l.status_off);
    uint64_t* groups = reinterpret_cast<uint64_t*>(base + l.groups_off);
    size_t* bounds = reinterpret_cast<size_t*>(base + l.bounds_off);
    const uint8_t* side_tile = reinterpret_cast<const uint8_t*>(base + l.side_off);
    const uint64_t* run_rows = reinterpret_cast<const uint64_t*>(base + l.run_off);
    hipError_t st = hipSuccess;
    uint64_t* side_unique_out = side_a;
    if (side_rows) {
        if (!side_b || sort_bytes < distinct_temp_bytes(side_rows)) return hipErrorInvalidValue;
        PairInput side_in{side_a, nullptr, nullptr, nullptr};
        uint64_t* sorted = nullptr;
        st = radix_sort(side_in, side_rows, side_a, side_b, sort_temp, sort_bytes, side_varying, stream, &sorted);
        if (st != hipSuccess) return st;
        side_unique_out = sorted == side_a ? side_b : side_a;
        st = unique_sorted(sorted, nullptr, side_rows, PairOutput{side_unique_out, nullptr, nullptr}, &h->side_unique,
                           sort_temp, sort_bytes, stream);
        if (st != hipSuccess) return st;
    }
    timer_start(timer, stream);
    // rank bits: at most log2(kSegCap) = 12 (a run per key); wider minors
    // are compressed to the range where each tile's differ, and the tiles
    // where that is still too wide are sorted with u64 keys after
    uint32_t* defer = reinterpret_cast<uint32_t*>(base + l.list_off);  // the long-tile list is used up
    if (minor_bits + kSegRankBits <= 32) {
        seg_tile_kernel<uint32_t, false><<<static_cast<unsigned>(l.tiles), kSegBlock, 0, stream>>>(
            keys, l.tiles, minor_bits, bounds, side_unique_out, status, h, found, groups, side_tile, run_rows, defer);
    } else {
        seg_tile_kernel<uint32_t, true><<<static_cast<unsigned>(l.tiles), kSegBlock, 0, stream>>>(
            keys, l.tiles, minor_bits, bounds, side_unique_out, status, h, found, groups, side_tile, run_rows, defer);
        seg_tile_list_kernel<<<static_cast<unsigned>(l.tiles < 512 ? l.tiles : 512), kSegBlock, 0, stream>>>(
            keys, minor_bits, bounds, side_unique_out, status, h, found, groups, side_tile, run_rows, defer);
    }
    seg_scan_kernel<<<1, kSegScanThreads, 0, stream>>>(groups, l.groups, unique_rows);
    const unsigned write_grid = static_cast<unsigned>(l.tiles < 65536 ? l.tiles : 65536);
    seg_write_kernel<<<write_grid, kBlock, 0, stream>>>(keys, bounds, side_unique_out, status, groups, h, l.tiles,
                                                        minor_bits, out, xidx_out, side_tile, marks);
    timer_stop(timer, stream);
    return hipGetLastError();
}

}  // namespace spk
