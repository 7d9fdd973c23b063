// Device text ingest: the whitespace-separated binary u32 loader
// (crates/sparkalog-storage/src/io.rs:72-141, load_binary_u32) on the GPU.
//
// The reference reads a file twice on the host: once to count rows (and to
// spot a Git LFS pointer, io.rs:91-103), once to parse each data line into
// two u32 columns (io.rs:105-131). Here the raw bytes are uploaded once and
// every step runs on the device, reading the text at HBM speed:
//   1. newline count per 16 KiB tile (four 16-byte loads per thread, an exact
//      SWAR count of '\n' bytes), and a one-block scan: the line count;
//   2. line starts: each tile is staged in LDS, every thread ranks the
//      newlines of its 64 bytes, and the start of each line is written;
//   3. a parse of 4,096 lines per block, one line per thread (adjacent
//      threads read adjacent lines): trim, skip blank and '#'/'%' lines,
//      split into fields on ASCII whitespace, parse u32 as Rust's
//      str::parse does (optional '+', decimal digits, no overflow); the
//      first pass counts data lines per tile and records the first error,
//      a scan gives each tile its output offset, the second pass writes the
//      columns in file order.
// Errors follow the reference's precedence: a Git LFS pointer line anywhere
// wins (the counting pass reads the whole file first); otherwise the first
// bad line, with a wrong field count before a bad first field before a bad
// second field. Each is one u64 code, kept as a device-wide minimum.
#include "ops.hpp"
#include "prims.hpp"
#include "sparkalog_hip.h"

namespace spk {
namespace {

constexpr int kTextBlock = 256;
constexpr int kTextVecs = 4;                                 // 16-byte vectors per thread
constexpr int kTextChunk = kTextVecs * 16;                   // 64 bytes per thread
constexpr size_t kTextTile = kTextBlock * kTextChunk;        // 16 KiB per block
constexpr int kLineItems = 16;
constexpr size_t kLineTile = kTextBlock * kLineItems;        // 4,096 lines per block
constexpr int kIngestScanThreads = 1024;

// io.rs:97-99
__constant__ char kLfsPointer[] = "version https://git-lfs.github.com/spec/v1";
constexpr size_t kLfsPointerLen = sizeof("version https://git-lfs.github.com/spec/v1") - 1;

struct IngestHeader {
    uint64_t newlines;  // '\n' bytes in the text
    uint64_t lines;     // lines as BufRead::lines counts them
};

__host__ __device__ inline size_t ingest_align(size_t v) { return (v + 255) / 256 * 256; }
__host__ __device__ inline size_t text_tiles(size_t len) { return (len + kTextTile - 1) / kTextTile; }
__host__ __device__ inline size_t line_tiles(uint64_t lines) { return (lines + kLineTile - 1) / kLineTile; }

// Bytes equal to '\n' in w (exact: the high bit of each byte of t is set
// only for a zero byte of x, with no carry between bytes).
__device__ __forceinline__ uint32_t newline_bytes(uint32_t w) {
    const uint32_t x = w ^ 0x0A0A0A0Au;
    const uint32_t t = ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu);
    return static_cast<uint32_t>(__popc(t));
}

__device__ __forceinline__ uint32_t newline_bytes(const uint4& v) {
    return newline_bytes(v.x) + newline_bytes(v.y) + newline_bytes(v.z) + newline_bytes(v.w);
}

// The 16-byte vector at byte offset `at` (a multiple of 16), zero-filled
// past `len` (a zero byte is neither '\n' nor whitespace nor a digit).
__device__ __forceinline__ uint4 load_text_vec(const uint8_t* text, size_t len, size_t at) {
    if (at + 16 <= len) return *reinterpret_cast<const uint4*>(text + at);
    uint32_t w[4] = {0, 0, 0, 0};
    for (size_t i = at; i < len && i < at + 16; ++i) w[(i - at) >> 2] |= static_cast<uint32_t>(text[i]) << (8 * ((i - at) & 3));
    return make_uint4(w[0], w[1], w[2], w[3]);
}

__global__ __launch_bounds__(kTextBlock) void text_count_kernel(const uint8_t* text, size_t len, uint64_t* tile_nl) {
    __shared__ uint32_t lds[kTextBlock / kWave];
    const size_t base = static_cast<size_t>(blockIdx.x) * kTextTile;
    uint32_t c = 0;
#pragma unroll
    for (int q = 0; q < kTextVecs; ++q) {
        const size_t at = base + (static_cast<size_t>(q) * kTextBlock + threadIdx.x) * 16;
        if (at < len) c += newline_bytes(load_text_vec(text, len, at));
    }
    uint32_t total = 0;
    (void)block_exclusive_sum<uint32_t, kTextBlock / kWave>(c, &total, lds);
    if (threadIdx.x == 0) tile_nl[blockIdx.x] = total;
}

// Exclusive scan of the tiles' newline counts, and the line count: a last
// line without a trailing '\n' is still a line (io.rs reads with
// BufRead::lines, which also yields no empty line after a final '\n').
__global__ __launch_bounds__(kIngestScanThreads) void text_scan_kernel(const uint8_t* text, size_t len,
                                                                       uint64_t* tile_nl, size_t tiles,
                                                                       IngestHeader* h, uint64_t* lines_out) {
    __shared__ uint64_t lds[kIngestScanThreads / kWave];
    block_scan_inplace<kIngestScanThreads, 4>(tile_nl, tiles, &h->newlines, lds);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint64_t nl = h->newlines;
        const uint64_t lines = nl + (len > 0 && text[len - 1] != '\n' ? 1 : 0);
        h->lines = lines;
        if (lines_out) *lines_out = lines;
    }
}

// line_start[k] = byte offset of line k, for k in [0, newlines]; line k <
// newlines ends at line_start[k + 1] - 1 (its '\n'). Entries past
// `capacity` are not written (the text then has more lines than counted).
__global__ __launch_bounds__(kTextBlock) void text_starts_kernel(const uint8_t* text, size_t len,
                                                                 const uint64_t* tile_nl, uint64_t* line_start,
                                                                 uint64_t capacity) {
    __shared__ uint4 tile[kTextTile / 16];
    __shared__ uint32_t lds[kTextBlock / kWave];
    const size_t base = static_cast<size_t>(blockIdx.x) * kTextTile;
#pragma unroll
    for (int q = 0; q < kTextVecs; ++q) {
        const int v = q * kTextBlock + static_cast<int>(threadIdx.x);
        const size_t at = base + static_cast<size_t>(v) * 16;
        tile[v] = at < len ? load_text_vec(text, len, at) : make_uint4(0, 0, 0, 0);
    }
    __syncthreads();
    // this thread's 64 bytes: vectors [threadIdx.x * 4, +4) of the tile
    uint4 mine[kTextVecs];
    uint32_t c = 0;
#pragma unroll
    for (int q = 0; q < kTextVecs; ++q) {
        mine[q] = tile[threadIdx.x * kTextVecs + q];
        c += newline_bytes(mine[q]);
    }
    uint32_t total = 0;
    uint64_t rank = tile_nl[blockIdx.x] + block_exclusive_sum<uint32_t, kTextBlock / kWave>(c, &total, lds);
    if (blockIdx.x == 0 && threadIdx.x == 0) line_start[0] = 0;
    if (c == 0) return;
    const size_t own = base + static_cast<size_t>(threadIdx.x) * kTextChunk;
#pragma unroll
    for (int q = 0; q < kTextVecs; ++q) {
        const uint32_t w[4] = {mine[q].x, mine[q].y, mine[q].z, mine[q].w};
#pragma unroll
        for (int d = 0; d < 4; ++d)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                if (((w[d] >> (8 * b)) & 0xFFu) == '\n') {
                    ++rank;
                    if (rank <= capacity) line_start[rank] = own + static_cast<size_t>(q * 16 + d * 4 + b) + 1;
                }
    }
}

__device__ __forceinline__ bool ascii_space(uint32_t c) { return c == ' ' || (c >= 9 && c <= 13); }

// str::parse::<u32> of text[s, e) (non-empty, no whitespace inside).
__device__ __forceinline__ bool parse_u32_field(const uint8_t* text, size_t s, size_t e, uint32_t* out) {
    if (s < e && text[s] == '+') ++s;
    if (s == e) return false;
    uint64_t v = 0;
    for (; s < e; ++s) {
        const uint32_t d = static_cast<uint32_t>(text[s]) - '0';
        if (d > 9) return false;
        v = v * 10 + d;
        if (v > 0xFFFFFFFFull) return false;
    }
    *out = static_cast<uint32_t>(v);
    return true;
}

enum LineKind : uint32_t { kSkip = 0, kData = 1 };

// Line k (0-based): kSkip, kData with (a, b), or an error code (> 1).
__device__ uint64_t parse_line(const uint8_t* text, size_t len, const uint64_t* line_start, uint64_t newlines,
                               uint64_t k, uint32_t* a, uint32_t* b) {
    size_t s = line_start[k];
    size_t e = k < newlines ? line_start[k + 1] - 1 : len;
    while (s < e && ascii_space(text[s])) ++s;  // str::trim (io.rs:86-89)
    while (e > s && ascii_space(text[e - 1])) --e;
    if (s == e || text[s] == '#' || text[s] == '%') return kSkip;
    const uint64_t line_no = k + 1;
    if (e - s == kLfsPointerLen) {
        bool same = true;
        for (size_t i = 0; i < kLfsPointerLen; ++i) same = same && text[s + i] == static_cast<uint8_t>(kLfsPointer[i]);
        if (same) return SPARKALOG_HIP_LOAD_ERR_LFS_POINTER | (line_no << 2);
    }
    // split_whitespace: the trimmed line starts and ends with a non-space
    size_t f1e = s;
    while (f1e < e && !ascii_space(text[f1e])) ++f1e;
    size_t f2s = f1e;
    while (f2s < e && ascii_space(text[f2s])) ++f2s;
    size_t f2e = f2s;
    while (f2e < e && !ascii_space(text[f2e])) ++f2e;
    if (f2s == e || f2e != e) return SPARKALOG_HIP_LOAD_ERR_FIELD_COUNT | (line_no << 2);
    if (!parse_u32_field(text, s, f1e, a)) return SPARKALOG_HIP_LOAD_ERR_U32_FIRST | (line_no << 2);
    if (!parse_u32_field(text, f2s, f2e, b)) return SPARKALOG_HIP_LOAD_ERR_U32_SECOND | (line_no << 2);
    return kData;
}

// Write = false: data lines per tile into tile_rows[t], errors into *error.
// Write = true: tile_rows holds exclusive offsets; rows go to first/second.
template <bool Write>
__global__ __launch_bounds__(kTextBlock) void text_parse_kernel(const uint8_t* text, size_t len,
                                                                const uint64_t* line_start, const IngestHeader* h,
                                                                uint64_t* tile_rows, uint32_t* first,
                                                                uint32_t* second, unsigned long long* error,
                                                                uint64_t capacity) {
    __shared__ uint32_t lds[kTextBlock / kWave];
    const uint64_t lines = h->lines, newlines = h->newlines;
    if (lines > capacity) return;  // block-uniform; reported by ingest_check_kernel
    const uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * kLineTile;
    uint64_t out = Write ? tile_rows[blockIdx.x] : 0;
    unsigned long long worst = ~0ull;
    for (int q = 0; q < kLineItems; ++q) {
        const uint64_t k = t0 + static_cast<uint64_t>(q) * kTextBlock + threadIdx.x;
        uint32_t a = 0, b = 0;
        uint64_t r = kSkip;
        if (k < lines) r = parse_line(text, len, line_start, newlines, k, &a, &b);
        if (r > kData) {
            worst = r < worst ? r : worst;
            r = kSkip;
        }
        uint32_t total = 0;
        const uint32_t rank = block_exclusive_sum<uint32_t, kTextBlock / kWave>(r == kData ? 1u : 0u, &total, lds);
        if (Write && r == kData) {
            first[out + rank] = a;
            second[out + rank] = b;
        }
        out += total;
    }
    if (!Write) {
        if (threadIdx.x == 0) tile_rows[blockIdx.x] = out;
        if (worst != ~0ull) atomicMin(error, worst);
    }
}

__global__ void ingest_init_kernel(unsigned long long* error) { *error = ~0ull; }

// Rows of the text (the scan's total) and the capacity check.
__global__ __launch_bounds__(kIngestScanThreads) void ingest_rows_kernel(uint64_t* tile_rows, size_t tiles,
                                                                         uint64_t* rows) {
    __shared__ uint64_t lds[kIngestScanThreads / kWave];
    block_scan_inplace<kIngestScanThreads, 4>(tile_rows, tiles, rows, lds);
}

__global__ void ingest_check_kernel(const IngestHeader* h, uint64_t capacity, unsigned long long* error) {
    if (h->lines > capacity) atomicMin(error, static_cast<unsigned long long>(SPARKALOG_HIP_LOAD_ERR_CAPACITY));
}

struct IngestLayout {
    size_t header_off, text_tiles_off, starts_off, line_tiles_off, total;
};

IngestLayout ingest_layout(size_t len, uint64_t lines) {
    IngestLayout l{};
    size_t off = 0;
    l.header_off = off;
    off += ingest_align(sizeof(IngestHeader));
    l.text_tiles_off = off;
    off += ingest_align((text_tiles(len) + 1) * sizeof(uint64_t));
    l.starts_off = off;
    off += ingest_align((lines + 1) * sizeof(uint64_t));
    l.line_tiles_off = off;
    off += ingest_align((line_tiles(lines) + 1) * sizeof(uint64_t));
    l.total = off;
    return l;
}

}  // namespace

size_t text_lines_temp_bytes(size_t len) { return ingest_layout(len, 0).starts_off; }

hipError_t text_lines(const uint8_t* text, size_t len, uint64_t* lines, void* temp, size_t temp_bytes,
                      hipStream_t stream) {
    if (len == 0) return set_u64(lines, 0, stream);
    const IngestLayout l = ingest_layout(len, 0);
    if (temp_bytes < l.starts_off) return hipErrorInvalidValue;
    char* base = static_cast<char*>(temp);
    IngestHeader* h = reinterpret_cast<IngestHeader*>(base + l.header_off);
    uint64_t* tile_nl = reinterpret_cast<uint64_t*>(base + l.text_tiles_off);
    const size_t tiles = text_tiles(len);
    text_count_kernel<<<static_cast<unsigned>(tiles), kTextBlock, 0, stream>>>(text, len, tile_nl);
    text_scan_kernel<<<1, kIngestScanThreads, 0, stream>>>(text, len, tile_nl, tiles, h, lines);
    return hipGetLastError();
}

size_t load_binary_u32_temp_bytes(size_t len, uint64_t lines) { return ingest_layout(len, lines).total; }

hipError_t load_binary_u32(const uint8_t* text, size_t len, uint64_t lines, uint32_t* first, uint32_t* second,
                           uint64_t* rows, uint64_t* error, void* temp, size_t temp_bytes, hipStream_t stream) {
    unsigned long long* err = reinterpret_cast<unsigned long long*>(error);
    ingest_init_kernel<<<1, 1, 0, stream>>>(err);
    if (len == 0) return set_u64(rows, 0, stream);
    const IngestLayout l = ingest_layout(len, lines);
    if (temp_bytes < l.total) return hipErrorInvalidValue;
    if (text_tiles(len) > 0xFFFFFFFFull || line_tiles(lines) > 0xFFFFFFFFull) return hipErrorInvalidValue;
    char* base = static_cast<char*>(temp);
    IngestHeader* h = reinterpret_cast<IngestHeader*>(base + l.header_off);
    uint64_t* tile_nl = reinterpret_cast<uint64_t*>(base + l.text_tiles_off);
    uint64_t* line_start = reinterpret_cast<uint64_t*>(base + l.starts_off);
    uint64_t* tile_rows = reinterpret_cast<uint64_t*>(base + l.line_tiles_off);
    const size_t tiles = text_tiles(len);
    text_count_kernel<<<static_cast<unsigned>(tiles), kTextBlock, 0, stream>>>(text, len, tile_nl);
    text_scan_kernel<<<1, kIngestScanThreads, 0, stream>>>(text, len, tile_nl, tiles, h, nullptr);
    // the line starts and the parse are sized for `lines`: a text with more
    // (not the one that was counted) stops here with a capacity error
    ingest_check_kernel<<<1, 1, 0, stream>>>(h, lines, err);
    if (lines == 0) return set_u64(rows, 0, stream);
    text_starts_kernel<<<static_cast<unsigned>(tiles), kTextBlock, 0, stream>>>(text, len, tile_nl, line_start,
                                                                                lines);
    const size_t ltiles = line_tiles(lines);
    text_parse_kernel<false><<<static_cast<unsigned>(ltiles), kTextBlock, 0, stream>>>(
        text, len, line_start, h, tile_rows, nullptr, nullptr, err, lines);
    ingest_rows_kernel<<<1, kIngestScanThreads, 0, stream>>>(tile_rows, ltiles, rows);
    text_parse_kernel<true><<<static_cast<unsigned>(ltiles), kTextBlock, 0, stream>>>(
        text, len, line_start, h, tile_rows, first, second, err, lines);
    return hipGetLastError();
}

}  // namespace spk

// ---- C ABI ----

extern "C" sparkalog_hip_error sparkalog_hip_text_lines_temporary_bytes(size_t len, size_t* temporary_bytes) {
    if (!temporary_bytes) return hipErrorInvalidValue;
    *temporary_bytes = spk::text_lines_temp_bytes(len);
    return hipSuccess;
}

extern "C" sparkalog_hip_error sparkalog_hip_text_lines(const char* text, size_t len, uint64_t* lines,
                                                        void* temporary, size_t temporary_bytes, void* stream) {
    if (!lines || (len && !text) || (len && !temporary)) return hipErrorInvalidValue;
    if (reinterpret_cast<uintptr_t>(text) % 16 != 0) return hipErrorInvalidValue;
    return spk::text_lines(reinterpret_cast<const uint8_t*>(text), len, lines, temporary, temporary_bytes,
                           static_cast<hipStream_t>(stream));
}

extern "C" sparkalog_hip_error sparkalog_hip_load_binary_u32_temporary_bytes(size_t len, uint64_t lines,
                                                                             size_t* temporary_bytes) {
    if (!temporary_bytes) return hipErrorInvalidValue;
    *temporary_bytes = spk::load_binary_u32_temp_bytes(len, lines);
    return hipSuccess;
}

extern "C" sparkalog_hip_error sparkalog_hip_load_binary_u32(const char* text, size_t len, uint64_t lines,
                                                             uint32_t* first, uint32_t* second, uint64_t* rows,
                                                             uint64_t* error, void* temporary,
                                                             size_t temporary_bytes, void* stream) {
    if (!rows || !error || (len && !text) || (len && !temporary) || (lines && (!first || !second)))
        return hipErrorInvalidValue;
    if (reinterpret_cast<uintptr_t>(text) % 16 != 0) return hipErrorInvalidValue;
    return spk::load_binary_u32(reinterpret_cast<const uint8_t*>(text), len, lines, first, second, rows, error,
                                temporary, temporary_bytes, static_cast<hipStream_t>(stream));
}
