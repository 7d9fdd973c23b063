// Device build of the join's range index: U32RangeIndex::build_slice
// (crates/sparkalog-storage/src/lib.rs:357-393) on the GPU.
//
// The reference sorts (value, row) pairs on the host and then walks them to
// record each distinct value and where its rows start. Here:
//   1. (value << 32 | row) is packed per row and sorted with the onesweep
//      radix sort over the bits that vary (all 32 value bits, the row bits);
//   2. a count pass marks the first row of each value run per 4,096-row
//      tile, a one-block scan gives each tile its first key slot, and a
//      write pass stores rows[i] (the sorted row ids), keys[j] (the run's
//      value) and starts[j] (the run's first position); starts[K] = n.
// The result is the reference's layout exactly: keys sorted unique,
// starts[K + 1], rows sorted by (value, row).
#include "ops.hpp"
#include "prims.hpp"
#include "sparkalog_hip.h"

namespace spk {
namespace {

constexpr int kIndexBlock = 256;
constexpr int kIndexItems = 16;
constexpr size_t kIndexTile = kIndexBlock * kIndexItems;  // 4,096 rows per tile
constexpr int kIndexScanThreads = 1024;

__host__ __device__ inline size_t index_align(size_t v) { return (v + 255) / 256 * 256; }
__host__ __device__ inline size_t index_tiles(size_t n) { return (n + kIndexTile - 1) / kIndexTile; }

__global__ void index_pack_kernel(const uint32_t* values, size_t n, uint64_t* packed) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        packed[i] = (static_cast<uint64_t>(values[i]) << 32) | static_cast<uint32_t>(i);
}

__device__ __forceinline__ bool run_start(const uint64_t* sorted, size_t i) {
    return i == 0 || (sorted[i] >> 32) != (sorted[i - 1] >> 32);
}

// Write = false: run starts per tile into tile_keys[t].
// Write = true: tile_keys holds exclusive offsets; writes rows, keys, starts.
template <bool Write>
__global__ __launch_bounds__(kIndexBlock) void index_runs_kernel(const uint64_t* sorted, size_t n,
                                                                 uint64_t* tile_keys, uint32_t* keys,
                                                                 uint32_t* starts, uint32_t* rows) {
    __shared__ uint32_t lds[kIndexBlock / kWave];
    const size_t t0 = static_cast<size_t>(blockIdx.x) * kIndexTile;
    uint64_t out = Write ? tile_keys[blockIdx.x] : 0;
    for (int q = 0; q < kIndexItems; ++q) {
        const size_t i = t0 + static_cast<size_t>(q) * kIndexBlock + threadIdx.x;
        const bool start = i < n && run_start(sorted, i);
        uint32_t total = 0;
        const uint32_t rank = block_exclusive_sum<uint32_t, kIndexBlock / kWave>(start ? 1u : 0u, &total, lds);
        if (Write && i < n) {
            const uint64_t k = sorted[i];
            rows[i] = static_cast<uint32_t>(k);
            if (start) {
                keys[out + rank] = static_cast<uint32_t>(k >> 32);
                starts[out + rank] = static_cast<uint32_t>(i);
            }
        }
        out += total;
    }
    if (!Write && threadIdx.x == 0) tile_keys[blockIdx.x] = out;
}

__global__ __launch_bounds__(kIndexScanThreads) void index_scan_kernel(uint64_t* tile_keys, size_t tiles,
                                                                       uint64_t* unique_keys) {
    __shared__ uint64_t lds[kIndexScanThreads / kWave];
    block_scan_inplace<kIndexScanThreads, 4>(tile_keys, tiles, unique_keys, lds);
}

// starts[K] = n (K read on the device).
__global__ void index_end_kernel(const uint64_t* unique_keys, size_t n, uint32_t* starts) {
    starts[*unique_keys] = static_cast<uint32_t>(n);
}


struct IndexLayout {
    size_t packed_a_off, packed_b_off, tiles_off, sort_off, total;
};

IndexLayout index_layout(size_t n) {
    IndexLayout l{};
    size_t off = 0;
    l.packed_a_off = off;
    off += index_align(n * sizeof(uint64_t));
    l.packed_b_off = off;
    off += index_align(n * sizeof(uint64_t));
    l.tiles_off = off;
    off += index_align((index_tiles(n) + 1) * sizeof(uint64_t));
    l.sort_off = off;
    off += index_align(distinct_temp_bytes(n));
    l.total = off;
    return l;
}

}  // namespace

size_t range_index_temp_bytes(size_t n) { return index_layout(n).total; }

hipError_t range_index(const uint32_t* values, size_t n, uint32_t* keys, uint32_t* starts, uint32_t* rows,
                       uint64_t* unique_keys, void* temp, size_t temp_bytes, hipStream_t stream) {
    if (n == 0) {
        hipError_t st = set_u64(unique_keys, 0, stream);
        if (st == hipSuccess) st = set_u32(starts, 0, stream);
        return st;
    }
    if (n > 0xFFFFFFFFull) return hipErrorInvalidValue;
    const IndexLayout l = index_layout(n);
    if (temp_bytes < l.total) return hipErrorInvalidValue;
    char* base = static_cast<char*>(temp);
    uint64_t* packed_a = reinterpret_cast<uint64_t*>(base + l.packed_a_off);
    uint64_t* packed_b = reinterpret_cast<uint64_t*>(base + l.packed_b_off);
    uint64_t* tile_keys = reinterpret_cast<uint64_t*>(base + l.tiles_off);
    void* sort_temp = base + l.sort_off;
    const size_t tiles = index_tiles(n);
    if (tiles > 0xFFFFFFFFull) return hipErrorInvalidValue;
    const size_t grid = (n + kIndexBlock * 4 - 1) / (kIndexBlock * 4);
    index_pack_kernel<<<static_cast<unsigned>(grid < 65536 ? grid : 65536), kIndexBlock, 0, stream>>>(values, n,
                                                                                                        packed_a);
    // every value bit may vary; the row bits are those of n - 1
    uint64_t row_mask = 0;
    while (row_mask < n - 1) row_mask = (row_mask << 1) | 1;
    const uint64_t varying = (0xFFFFFFFFull << 32) | (row_mask ? row_mask : 1);
    uint64_t* sorted = nullptr;
    const PairInput in{packed_a, nullptr, nullptr, nullptr};
    hipError_t st = radix_sort(in, n, packed_a, packed_b, sort_temp, l.total - l.sort_off, varying, stream, &sorted);
    if (st != hipSuccess) return st;
    index_runs_kernel<false><<<static_cast<unsigned>(tiles), kIndexBlock, 0, stream>>>(sorted, n, tile_keys, nullptr,
                                                                                        nullptr, nullptr);
    index_scan_kernel<<<1, kIndexScanThreads, 0, stream>>>(tile_keys, tiles, unique_keys);
    index_runs_kernel<true><<<static_cast<unsigned>(tiles), kIndexBlock, 0, stream>>>(sorted, n, tile_keys, keys,
                                                                                       starts, rows);
    index_end_kernel<<<1, 1, 0, stream>>>(unique_keys, n, starts);
    return hipGetLastError();
}

}  // namespace spk

extern "C" sparkalog_hip_error sparkalog_hip_range_index_u32_temporary_bytes(size_t rows, size_t* temporary_bytes) {
    if (!temporary_bytes) return hipErrorInvalidValue;
    *temporary_bytes = spk::range_index_temp_bytes(rows);
    return hipSuccess;
}

extern "C" sparkalog_hip_error sparkalog_hip_range_index_u32(const uint32_t* values, size_t rows, uint32_t* keys,
                                                             uint32_t* starts, uint32_t* index_rows,
                                                             uint64_t* unique_keys, void* temporary,
                                                             size_t temporary_bytes, void* stream) {
    if (!unique_keys || !starts || (rows && (!values || !keys || !index_rows || !temporary)))
        return hipErrorInvalidValue;
    return spk::range_index(values, rows, keys, starts, index_rows, unique_keys, temporary, temporary_bytes,
                            static_cast<hipStream_t>(stream));
}
