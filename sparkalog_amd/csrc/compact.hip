// Kernels of the graph-replayed semi-naive iteration: every row count is read
// from device memory and every launch is sized by a capacity, so one
// iteration's launches can be captured once into a HIP graph and replayed
// while the buffers stay the same.
//
// merge_new: for FULL (F, sorted unique packed pairs) and the iteration's
// distinct candidates (U, sorted unique), one merge writes FULL' = F ∪ U and
// NEWT = U \ F (SortedBinaryAntiJoin followed by SortedBinaryUnion,
// crates/sparkalog-recursion/src/generic.rs:339-380, in one pass over F and U):
//   1. partition: the merge-path split of every tile boundary (ties put the F
//      row first, so a U row equal to an F row directly follows it);
//   2. count: per tile, its U rows already in F (duplicates) and the others;
//   3. a one-block scan of the per-tile (duplicates << 32 | new) counts;
//   4. write: every merged row to FULL' at its position less the duplicates
//      before it, and every new U row to NEWT.
#include "ops.hpp"
#include "prims.hpp"

namespace spk {
namespace {

constexpr int kMnTile = 2048;  // merged input rows per tile
constexpr int kMnPer = kMnTile / kBlock;

__host__ __device__ inline size_t mn_tiles(size_t n) { return (n + kMnTile - 1) / kMnTile; }
__host__ __device__ inline size_t mn_align(size_t v) { return (v + 255) / 256 * 256; }

__device__ __forceinline__ size_t dev_rows(const uint64_t* rows_dev, size_t cap) {
    const uint64_t n = *rows_dev;
    return n < cap ? static_cast<size_t>(n) : cap;
}

// part[t] = F rows among the first t * kMnTile merged rows, t in [0, tiles].
__global__ void mn_partition_kernel(const uint64_t* f, const uint64_t* f_rows, size_t f_cap, const uint64_t* u,
                                    const uint64_t* u_rows, size_t u_cap, size_t* part, size_t tiles) {
    const size_t nf = dev_rows(f_rows, f_cap), nu = dev_rows(u_rows, u_cap);
    const size_t total = nf + nu;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; t <= tiles; t += stride) {
        size_t diag = t * kMnTile;
        if (diag > total) diag = total;
        size_t lo = diag > nu ? diag - nu : 0, hi = diag < nf ? diag : nf;
        while (lo < hi) {
            const size_t mid = (lo + hi) >> 1;
            // F first on ties: F[mid] goes before U[diag - 1 - mid] when <=
            if (f[mid] <= u[diag - 1 - mid])
                lo = mid + 1;
            else
                hi = mid;
        }
        part[t] = lo;
    }
}

// The tile's rows merged in LDS: F rows [a0, a1) and U rows [b0, b1).
// Returns the tile's row count; merged[j] holds the key, from_u[j] whether it
// is a U row, and dup[j] whether it is a U row equal to the row before it.
__device__ __forceinline__ uint32_t mn_load_tile(const uint64_t* f, const uint64_t* u, size_t a0, size_t a1,
                                                 size_t b0, size_t b1, uint64_t* fa, uint64_t* ub,
                                                 uint64_t& before) {
    const uint32_t na = static_cast<uint32_t>(a1 - a0), nb = static_cast<uint32_t>(b1 - b0);
    for (uint32_t i = threadIdx.x; i < na; i += kBlock) fa[i] = f[a0 + i];
    for (uint32_t i = threadIdx.x; i < nb; i += kBlock) ub[i] = u[b0 + i];
    // the merged row before the tile: the larger of F[a0 - 1] and U[b0 - 1]
    // (F first on ties, so an equal pair puts the U row last)
    if (threadIdx.x == 0) {
        const bool hf = a0 > 0, hu = b0 > 0;
        before = 0;
        if (hf && hu) {
            const uint64_t x = f[a0 - 1], y = u[b0 - 1];
            before = x > y ? x : y;
        } else if (hf) {
            before = f[a0 - 1];
        } else if (hu) {
            before = u[b0 - 1];
        }
    }
    __syncthreads();
    return na + nb;
}

// Position j of the tile's merge: the number of its F rows before it.
__device__ __forceinline__ uint32_t mn_split(const uint64_t* fa, uint32_t na, const uint64_t* ub, uint32_t nb,
                                             uint32_t diag) {
    uint32_t lo = diag > nb ? diag - nb : 0, hi = diag < na ? diag : na;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (fa[mid] <= ub[diag - 1 - mid])
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// The merged row at tile position j (< n): its key and whether it is a new
// U row (a U row not equal to the merged row before it).
__device__ __forceinline__ void mn_row(const uint64_t* fa, uint32_t na, const uint64_t* ub, uint32_t nb, uint32_t j,
                                       bool have_before, uint64_t before, uint64_t& key, bool& is_u, bool& dup) {
    const uint32_t a = mn_split(fa, na, ub, nb, j);
    const uint32_t b = j - a;
    // the row at j is F[a] unless F is used up or U[b] sorts first
    is_u = a >= na || (b < nb && ub[b] < fa[a]);
    key = is_u ? ub[b] : fa[a];
    dup = false;
    if (is_u) {
        // its predecessor in the merge: the last row of F[0, a) or U[0, b)
        // (the F row when equal); an equal F row can only be F[a - 1]
        if (a > 0)
            dup = fa[a - 1] == key;
        else if (have_before)
            dup = before == key;
    }
}

__global__ __launch_bounds__(kBlock) void mn_count_kernel(const uint64_t* f, const uint64_t* f_rows, size_t f_cap,
                                                          const uint64_t* u, const uint64_t* u_rows, size_t u_cap,
                                                          const size_t* part, size_t tiles, uint64_t* tile_counts) {
    __shared__ uint64_t fa[kMnTile], ub[kMnTile];
    __shared__ uint64_t before;
    __shared__ uint32_t red[2][kWavesPerBlock];
    const size_t nf = dev_rows(f_rows, f_cap), nu = dev_rows(u_rows, u_cap);
    const size_t total = nf + nu;
    for (size_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const size_t d0 = t * kMnTile;
        if (d0 >= total) {  // block-uniform
            if (threadIdx.x == 0) tile_counts[t] = 0;
            continue;
        }
        const size_t d1 = d0 + kMnTile < total ? d0 + kMnTile : total;
        const size_t a0 = part[t], a1 = part[t + 1];
        const size_t b0 = d0 - a0, b1 = d1 - a1;
        const uint32_t n = mn_load_tile(f, u, a0, a1, b0, b1, fa, ub, before);
        const uint32_t na = static_cast<uint32_t>(a1 - a0), nb = static_cast<uint32_t>(b1 - b0);
        const bool have_before = d0 > 0;
        uint32_t dups = 0, news = 0;
        for (uint32_t j = threadIdx.x; j < n; j += kBlock) {
            uint64_t key;
            bool is_u, dup;
            mn_row(fa, na, ub, nb, j, have_before, before, key, is_u, dup);
            dups += dup ? 1u : 0u;
            news += (is_u && !dup) ? 1u : 0u;
        }
        dups = wave_sum(dups);
        news = wave_sum(news);
        if (lane_id() == 0) {
            red[0][wave_id()] = dups;
            red[1][wave_id()] = news;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t d = 0, w = 0;
            for (int i = 0; i < kWavesPerBlock; ++i) {
                d += red[0][i];
                w += red[1][i];
            }
            tile_counts[t] = (d << 32) | w;
        }
        __syncthreads();
    }
}

// Exclusive scan in place of the packed (duplicates << 32 | new) tile counts
// (neither sum reaches 2^32: F and U each hold fewer rows), and the totals:
// FULL' rows = F + U - duplicates, NEWT rows = new.
constexpr int kMnScanThreads = 1024;
__global__ __launch_bounds__(kMnScanThreads) void mn_scan_kernel(uint64_t* tile_counts, size_t tiles,
                                                                 const uint64_t* f_rows, size_t f_cap,
                                                                 const uint64_t* u_rows, size_t u_cap,
                                                                 uint64_t* full_rows, uint64_t* newt_rows) {
    __shared__ uint64_t scan_lds[kMnScanThreads / kWave];
    __shared__ uint64_t total;
    block_scan_inplace<kMnScanThreads, 4>(tile_counts, tiles, &total, scan_lds);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint64_t nf = dev_rows(f_rows, f_cap), nu = dev_rows(u_rows, u_cap);
        *full_rows = nf + nu - (total >> 32);
        *newt_rows = total & 0xFFFFFFFFull;
    }
}

__global__ __launch_bounds__(kBlock) void mn_write_kernel(const uint64_t* f, const uint64_t* f_rows, size_t f_cap,
                                                          const uint64_t* u, const uint64_t* u_rows, size_t u_cap,
                                                          const size_t* part, size_t tiles,
                                                          const uint64_t* tile_base, uint64_t* full_out,
                                                          uint64_t* newt_out) {
    __shared__ uint64_t fa[kMnTile], ub[kMnTile];
    __shared__ uint64_t before;
    __shared__ uint32_t scan_lds[2][kWavesPerBlock];
    const size_t nf = dev_rows(f_rows, f_cap), nu = dev_rows(u_rows, u_cap);
    const size_t total = nf + nu;
    const uint32_t lane = lane_id(), wave = wave_id();
    for (size_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const size_t d0 = t * kMnTile;
        if (d0 >= total) continue;  // block-uniform
        const size_t d1 = d0 + kMnTile < total ? d0 + kMnTile : total;
        const size_t a0 = part[t], a1 = part[t + 1];
        const size_t b0 = d0 - a0, b1 = d1 - a1;
        const uint32_t n = mn_load_tile(f, u, a0, a1, b0, b1, fa, ub, before);
        const uint32_t na = static_cast<uint32_t>(a1 - a0), nb = static_cast<uint32_t>(b1 - b0);
        const bool have_before = d0 > 0;
        const uint64_t base = tile_base[t];
        const uint64_t dup_base = base >> 32, new_base = base & 0xFFFFFFFFull;
        // thread i owns tile positions [i * kMnPer, i * kMnPer + kMnPer): a
        // block scan of its duplicates and new rows orders them
        uint64_t keys[kMnPer];
        uint32_t flags = 0;  // bit 2q: U row of position q, bit 2q+1: duplicate
        uint32_t my_dups = 0, my_news = 0;
#pragma unroll
        for (int q = 0; q < kMnPer; ++q) {
            const uint32_t j = threadIdx.x * kMnPer + q;
            keys[q] = 0;
            if (j < n) {
                bool is_u, dup;
                mn_row(fa, na, ub, nb, j, have_before, before, keys[q], is_u, dup);
                flags |= (is_u ? 1u : 0u) << (2 * q);
                flags |= (dup ? 1u : 0u) << (2 * q + 1);
                my_dups += dup ? 1u : 0u;
                my_news += (is_u && !dup) ? 1u : 0u;
            }
        }
        // block exclusive scans of both counts
        const uint32_t di = wave_inclusive_sum<uint32_t>(my_dups);
        const uint32_t ni = wave_inclusive_sum<uint32_t>(my_news);
        if (lane == kWave - 1) {
            scan_lds[0][wave] = di;
            scan_lds[1][wave] = ni;
        }
        __syncthreads();
        uint32_t dw = 0, nw = 0;
        for (uint32_t w = 0; w < wave; ++w) {
            dw += scan_lds[0][w];
            nw += scan_lds[1][w];
        }
        uint64_t dups_before = dup_base + dw + di - my_dups;
        uint64_t news_before = new_base + nw + ni - my_news;
#pragma unroll
        for (int q = 0; q < kMnPer; ++q) {
            const uint32_t j = threadIdx.x * kMnPer + q;
            if (j >= n) continue;
            const bool is_u = (flags >> (2 * q)) & 1u, dup = (flags >> (2 * q + 1)) & 1u;
            if (dup) {
                ++dups_before;
                continue;
            }
            full_out[d0 + j - dups_before] = keys[q];
            if (is_u) newt_out[news_before++] = keys[q];
        }
        __syncthreads();
    }
}

}  // namespace

size_t merge_new_temp_bytes(size_t full_cap, size_t cand_cap) {
    const size_t tiles = mn_tiles(full_cap + cand_cap);
    return mn_align((tiles + 1) * sizeof(size_t)) + mn_align((tiles + 1) * sizeof(uint64_t));
}

hipError_t merge_new(const uint64_t* full, const uint64_t* full_rows, size_t full_cap, const uint64_t* cand,
                     const uint64_t* cand_rows, size_t cand_cap, uint64_t* full_out, uint64_t* full_out_rows,
                     uint64_t* newt_out, uint64_t* newt_rows, void* temp, size_t temp_bytes, hipStream_t stream,
                     const KernelTimer* timer) {
    if (temp_bytes < merge_new_temp_bytes(full_cap, cand_cap)) return hipErrorInvalidValue;
    const size_t tiles = mn_tiles(full_cap + cand_cap) ? mn_tiles(full_cap + cand_cap) : 1;
    char* t = static_cast<char*>(temp);
    size_t* part = reinterpret_cast<size_t*>(t);
    uint64_t* counts = reinterpret_cast<uint64_t*>(t + mn_align((tiles + 1) * sizeof(size_t)));
    const unsigned pgrid = static_cast<unsigned>((tiles + 1 + kBlock - 1) / kBlock < 4096 ? (tiles + 1 + kBlock - 1) / kBlock
                                                                                         : 4096);
    const unsigned grid = static_cast<unsigned>(tiles < 65536 ? tiles : 65536);
    mn_partition_kernel<<<pgrid, kBlock, 0, stream>>>(full, full_rows, full_cap, cand, cand_rows, cand_cap, part, tiles);
    timer_start(timer, stream);
    mn_count_kernel<<<grid, kBlock, 0, stream>>>(full, full_rows, full_cap, cand, cand_rows, cand_cap, part, tiles,
                                                 counts);
    mn_scan_kernel<<<1, kMnScanThreads, 0, stream>>>(counts, tiles, full_rows, full_cap, cand_rows, cand_cap,
                                                     full_out_rows, newt_rows);
    mn_write_kernel<<<grid, kBlock, 0, stream>>>(full, full_rows, full_cap, cand, cand_rows, cand_cap, part, tiles,
                                                 counts, full_out, newt_out);
    timer_stop(timer, stream);
    return hipGetLastError();
}

}  // namespace spk
