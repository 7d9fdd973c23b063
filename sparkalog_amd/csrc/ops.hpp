// Host-side launchers of the gfx950 operator kernels (namespace spk). All
// launchers are asynchronous on `stream`, allocate nothing, and carve their
// scratch from a caller-provided temporary buffer whose size the matching
// *_temp_bytes function returns. Row counts are device-resident.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace spk {

// Key sources for the sort and set operators: packed u64 keys (first<<32 |
// second) or two u32 columns (structure of arrays).
// `rows_dev`, when non-null, holds the actual row count on the device; the
// host-side row count passed beside the input is then an upper bound used
// only to size the launch.
struct PairInput {
    const uint64_t* packed;   // used when non-null
    const uint32_t* first;
    const uint32_t* second;
    const uint64_t* rows_dev;
};

// Optional events recorded on the launch stream immediately before and after
// an operator's main kernel (the one that moves the operator's rows), so a
// caller can time that kernel alone; helper kernels fall outside.
struct KernelTimer {
    hipEvent_t start = nullptr;
    hipEvent_t stop = nullptr;
};

inline void timer_start(const KernelTimer* t, hipStream_t stream) {
    if (t && t->start) (void)hipEventRecord(t->start, stream);
}
inline void timer_stop(const KernelTimer* t, hipStream_t stream) {
    if (t && t->stop) (void)hipEventRecord(t->stop, stream);
}

struct PairOutput {
    uint64_t* packed;   // written when non-null
    uint32_t* first;    // written when non-null
    uint32_t* second;
    uint64_t* packed_copy = nullptr;  // second packed copy, written when non-null
    // second column of the rows (u32), written when non-null; only the
    // segmented distinct writes it
    uint32_t* second_copy = nullptr;
};

// ---- radix sort + unique (distinct) ----
size_t distinct_temp_bytes(size_t rows);
// Sort `rows` keys (from `in`) using keys_a / keys_b (rows entries each) as
// ping-pong buffers, drop adjacent duplicates, write the unique keys to `out`
// and the unique count to *unique_rows (u64, device). `known_varying` is a
// host-side superset of the key bits that differ between keys (0 = unknown:
// the kernels find it on the device).
hipError_t distinct(PairInput in, size_t rows, uint64_t* keys_a, uint64_t* keys_b,
                    PairOutput out, uint64_t* unique_rows, void* temp, size_t temp_bytes,
                    uint64_t known_varying, hipStream_t stream);
// Sort only (stable LSD); the sorted keys end in keys_a or keys_b: the
// returned pointer names which when known_varying != 0, else the device
// flag word at `temp` holds it and sorted_out (device u64*) is written.
hipError_t radix_sort(PairInput in, size_t rows, uint64_t* keys_a, uint64_t* keys_b,
                      void* temp, size_t temp_bytes, uint64_t known_varying, hipStream_t stream,
                      uint64_t** sorted_host);

// Right inputs of a multi-way anti-join: sorted relations (at most
// kMaxAntiRights), rows known on the host.
constexpr int kMaxAntiRights = 16;
struct AntiRight {
    PairInput in[kMaxAntiRights];
    uint64_t rows[kMaxAntiRights];
    uint32_t count;
};

// Unique compaction of sorted keys; the row count is *rows_dev (<= bound)
// when rows_dev is non-null. `temp` holds distinct_temp_bytes(bound).
hipError_t unique_sorted(const uint64_t* keys, const uint64_t* rows_dev, size_t bound, PairOutput out,
                         uint64_t* unique_rows, void* temp, size_t temp_bytes, hipStream_t stream);

// Distinct of compressed keys (major << minor_bits | minor) whose major part
// is non-decreasing along the input; writes sorted unique expanded pairs
// (major << 32 | minor). Two phases: prepare partitions the input into tiles
// aligned to runs, deduplicates the long first run of each tile over
// kSegCap keys through a bitmap of its minors, copies the keys of a long
// tile whose first run is too wide for the bitmap (minors of more than 20
// bits that differ above their low 18) to `side` and returns their row
// count on the host (with minor_bits <= 20 there are none: prepare returns
// 0 without waiting for the device); finish sorts
// those (side/side_b, sort_temp sized by distinct_temp_bytes(side_rows);
// unused when side_rows is 0) and every tile (a long tile's tail after its
// first run), overwriting `keys` (each tile's unique rows are staged over
// its input).
// `meta` (segmented_distinct_meta_bytes(rows)) must persist between phases.
// With `found` (probe_grouped's words over the same keys), keys whose found
// bit is set are dropped first: the output is then (keys \ right inputs),
// sorted and unique. With `xidx_out` (all entries 0xFFFFFFFF, from
// xindex_clear), finish also marks the first row of each first value in the
// output's first-column index (completed by xindex_fill).
// `tile_rows` is the nominal tile (512..4096 keys, 3,072 by default): tile
// boundaries sit every tile_rows keys before they move back to run starts.
constexpr uint32_t kSegTileDefault = 3072;
// Instance tracking for a batch of instances in one id space (instance =
// first value >> bits): finish raises last_new[instance] to `value` for every
// instance with an output row (at most `instances`; last_new null: none).
// With rows_acc / digest_acc, finish also adds each instance's output rows
// and the sum of instance_pair_hash over them (the per-instance statistics
// of the batched fixpoint, accumulated as NEWT is written), spread over
// kInstanceStatSlots copies (slot-major: copy j of instance i at
// j * instances + i; a block adds to copy blockIdx % slots) so that the
// atomics of many blocks do not all meet on one counter; the reader sums
// the copies.
constexpr uint32_t kInstanceStatSlots = 16;
struct InstanceMarks {
    uint32_t* last_new = nullptr;
    uint32_t bits = 0;
    uint32_t instances = 0;
    uint32_t value = 0;
    unsigned long long* rows_acc = nullptr;
    unsigned long long* digest_acc = nullptr;
};

// Digest term of a pair (x, z) of an instance: its local ids (the low `bits`
// bits of each) through the splitmix64 finalizer (Steele, Lea and Flood,
// 2014); an instance's digest is the sum over its pairs (order-independent).
__host__ __device__ inline uint64_t instance_pair_hash(uint32_t x, uint32_t z, uint32_t bits) {
    const uint32_t mask = bits >= 32 ? 0xFFFFFFFFu : ((1u << bits) - 1u);
    uint64_t zz = ((static_cast<uint64_t>(x & mask) << 32) | (z & mask)) + 0x9E3779B97F4A7C15ull;
    zz = (zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9ull;
    zz = (zz ^ (zz >> 27)) * 0x94D049BB133111EBull;
    return zz ^ (zz >> 31);
}
size_t segmented_distinct_meta_bytes(size_t rows, uint32_t tile_rows = kSegTileDefault);
hipError_t segmented_distinct_prepare(uint64_t* keys, size_t rows, uint32_t minor_bits, uint64_t* side,
                                      void* meta, size_t meta_bytes, uint64_t* side_rows, hipStream_t stream,
                                      const uint64_t* found = nullptr, uint32_t tile_rows = kSegTileDefault);
hipError_t segmented_distinct_finish(uint64_t* keys, size_t rows, uint32_t minor_bits, uint64_t side_varying,
                                     uint64_t* side_a, uint64_t* side_b, uint64_t side_rows, PairOutput out,
                                     uint64_t* unique_rows, void* meta, size_t meta_bytes, void* sort_temp,
                                     size_t sort_bytes, hipStream_t stream, const KernelTimer* timer = nullptr,
                                     const uint64_t* found = nullptr, uint32_t* xidx_out = nullptr,
                                     uint32_t tile_rows = kSegTileDefault, InstanceMarks marks = {});

// ---- sorted set operators ----
size_t union_temp_bytes(size_t left_rows, size_t right_rows);
hipError_t sorted_union(PairInput left, size_t left_rows, PairInput right, size_t right_rows,
                        PairOutput out, uint64_t* unique_rows, void* temp, size_t temp_bytes,
                        hipStream_t stream, const KernelTimer* timer = nullptr);

// Merge of two sorted packed arrays that share no key (FULL's runs):
// out[0, left_rows + right_rows) in order, and the second column of the
// merged rows in out_second when it is non-null.
size_t merge_disjoint_temp_bytes(size_t left_rows, size_t right_rows);
hipError_t merge_disjoint(const uint64_t* left, size_t left_rows, const uint64_t* right, size_t right_rows,
                          uint64_t* out, void* temp, size_t temp_bytes, hipStream_t stream,
                          const KernelTimer* timer = nullptr, uint32_t* out_second = nullptr);

size_t anti_join_temp_bytes(size_t left_rows, uint32_t right_inputs = 1);
// left \ right; `count32`/`count64` receive the output row count (either may
// be null).
hipError_t sorted_anti_join(PairInput left, size_t left_rows, PairInput right, size_t right_rows,
                            PairOutput out, uint32_t* count32, uint64_t* count64, void* temp,
                            size_t temp_bytes, hipStream_t stream);
// left \ (rights[0] ∪ rights[1] ∪ ...) in one pass over the left input.
hipError_t sorted_anti_join_multi(PairInput left, size_t left_rows, const AntiRight& rights, PairOutput out,
                                  uint32_t* count32, uint64_t* count64, void* temp, size_t temp_bytes,
                                  hipStream_t stream, const KernelTimer* timer = nullptr);

// Which candidate rows the right inputs hold. keys are compressed candidate
// rows (first << minor_bits | second, minor_bits in 1..32) grouped by their
// first column (count *rows_dev <= bound); bit j of found_words[i] is set
// when row 64*i + j is in some right input (probe_found_words(bound) words,
// written up to the last row).
size_t probe_grouped_temp_bytes(size_t rows, uint32_t right_inputs);
size_t probe_found_words(size_t rows);
hipError_t probe_grouped(const uint64_t* keys, const uint64_t* rows_dev, size_t bound, uint32_t minor_bits,
                         const AntiRight& rights, uint64_t* found_words, void* temp, size_t temp_bytes,
                         hipStream_t stream, const KernelTimer* timer = nullptr);

// ---- first-column index of sorted packed runs, and the anti-join using it ----
// xstart (n_x + 1 u32 entries) of a sorted packed relation with fewer than
// 2^32 rows whose first values are all below n_x: xstart[x] = rows with first
// column < x. The row count is *rows_dev (<= bound) when rows_dev is non-null.
hipError_t xindex_build(const uint64_t* keys, const uint64_t* rows_dev, size_t bound, uint32_t n_x, uint32_t* xstart,
                        hipStream_t stream);
// A first-column index from the entries the segmented distinct wrote
// (xidx_out: the output position of each first value's first row, every
// other entry 0xFFFFFFFF, set by xindex_clear): entry n_x becomes *rows_dev
// and every unset entry the next set one (a suffix minimum). temp holds
// xindex_fill_temp_bytes(n_x).
hipError_t xindex_clear(uint32_t* xstart, uint32_t n_x, hipStream_t stream);
size_t xindex_fill_temp_bytes(uint32_t n_x);
hipError_t xindex_fill(uint32_t* xstart, uint32_t n_x, const uint64_t* rows_dev, void* temp, size_t temp_bytes,
                       hipStream_t stream);
// to = from with n_x + 1 entries: a run's index for a larger n_x (its rows
// all have first values below old_n, so the new entries equal `rows`).
hipError_t xindex_extend(const uint32_t* from, uint32_t old_n, uint32_t* to, uint32_t n_x, uint32_t rows,
                         hipStream_t stream);
// into[x] += other[x]: the index of the merge of two disjoint runs.
hipError_t xindex_add(uint32_t* into, const uint32_t* other, uint32_t n_x, hipStream_t stream);

struct XRights {
    const uint64_t* keys[kMaxAntiRights];
    const uint32_t* xstart[kMaxAntiRights];
    uint32_t count;
};
// probe_grouped for few candidates: one thread per candidate, searching
// only its first value's segment of every run (rights' first-column indexes,
// first values below n_x); found_words as probe_grouped's.
hipError_t probe_sparse(const uint64_t* keys, size_t rows, uint32_t minor_bits, const XRights& rights, uint32_t n_x,
                        uint64_t* found_words, hipStream_t stream, const KernelTimer* timer = nullptr);
// FULL's runs by first-column index and second column only: the rows of run
// q with first value x have second values second[q][xstart[q][x],
// xstart[q][x+1]), ascending (4 bytes per row instead of 8).
struct XSecRights {
    const uint32_t* second[kMaxAntiRights];
    const uint32_t* xstart[kMaxAntiRights];
    uint32_t count;
};
// probe_grouped over FULL's second columns: for candidate rows grouped by
// their first value in ascending order (compressed keys, first value below
// n_x), sets bit i of found_words when candidate i is in some run. Each tile
// streams, per run, the second values of its first-value range through LDS.
// Every second column passed to probe_xsec is allocated with kSecPad rows
// past its last row (read, never used).
constexpr size_t kSecPad = 4;
size_t probe_xsec_temp_bytes(size_t rows);
hipError_t probe_xsec(const uint64_t* keys, size_t rows, uint32_t minor_bits, const XSecRights& rights,
                      uint32_t n_x, uint64_t* found_words, void* temp, size_t temp_bytes, hipStream_t stream,
                      const KernelTimer* timer = nullptr);
// probe_sparse over second columns (XSecRights), one thread per candidate.
hipError_t probe_sparse_sec(const uint64_t* keys, size_t rows, uint32_t minor_bits, const XSecRights& rights,
                            uint32_t n_x, uint64_t* found_words, hipStream_t stream,
                            const KernelTimer* timer = nullptr);
// Merge of two disjoint runs given by first-column index (n_x + 1 entries)
// and second column: writes the merged second column and/or the merged
// packed keys (either may be null, not both). The merged index is
// xa + xb (xindex_add). With nb == 0 it expands run A to packed keys.
size_t run_merge_temp_bytes(size_t rows);
hipError_t run_merge(const uint32_t* xa, const uint32_t* sa, size_t na, const uint32_t* xb, const uint32_t* sb,
                     size_t nb, uint32_t n_x, uint32_t* out_second, uint64_t* out_packed, void* temp,
                     size_t temp_bytes, hipStream_t stream, const KernelTimer* timer = nullptr);
// second[i] = low 32 bits of keys[i] (the second column of packed rows).
hipError_t second_column(const uint64_t* keys, size_t rows, uint32_t* second, hipStream_t stream);
// left \ (rights' union) for sorted unique packed left rows (count *left_rows
// <= bound) whose first values are below n_x; writes out.packed and
// out.packed_copy (out.first must be null) and the row count to *count64.
size_t anti_join_x_temp_bytes(size_t bound);
hipError_t anti_join_x(const uint64_t* left, const uint64_t* left_rows, size_t bound, const XRights& rights,
                       uint32_t n_x, PairOutput out, uint64_t* count64, void* temp, size_t temp_bytes,
                       hipStream_t stream, const KernelTimer* timer = nullptr);

// ---- one pass of anti-join and union for the graph-replayed iteration ----
// For FULL (sorted unique packed pairs, count *full_rows <= full_cap) and the
// distinct candidates (sorted unique, count *cand_rows <= cand_cap): writes
// FULL ∪ candidates to full_out (count *full_out_rows) and candidates \ FULL
// to newt_out (count *newt_rows), in order. Every count is read on the
// device and every launch is sized by the capacities, so the calls can be
// captured in a HIP graph; temp holds merge_new_temp_bytes(full_cap,
// cand_cap).
size_t merge_new_temp_bytes(size_t full_cap, size_t cand_cap);
hipError_t merge_new(const uint64_t* full, const uint64_t* full_rows, size_t full_cap, const uint64_t* cand,
                     const uint64_t* cand_rows, size_t cand_cap, uint64_t* full_out, uint64_t* full_out_rows,
                     uint64_t* newt_out, uint64_t* newt_rows, void* temp, size_t temp_bytes, hipStream_t stream,
                     const KernelTimer* timer = nullptr);

// ---- indexed join (U32RangeIndex layout) ----
size_t join_temp_bytes(size_t left_rows);
hipError_t join_count(const uint32_t* left_keys, size_t left_rows, const uint32_t* index_keys,
                      const uint32_t* index_starts, size_t unique_keys, uint64_t* counts,
                      uint64_t* offsets, uint64_t* total, void* temp, size_t temp_bytes,
                      hipStream_t stream);
hipError_t join_emit(const uint32_t* left_keys, size_t left_rows, const uint32_t* index_keys,
                     const uint32_t* index_starts, const uint32_t* index_rows, size_t unique_keys,
                     const uint32_t* projection0, uint32_t side0, const uint32_t* projection1,
                     uint32_t side1, const uint64_t* offsets, uint32_t* out0, uint32_t* out1,
                     hipStream_t stream);

// ---- filter / compaction ----
size_t filter_temp_bytes(size_t rows);
hipError_t filter(const uint32_t* input, size_t rows, uint32_t op, uint32_t operand,
                  uint32_t* out_rows, uint32_t* count, void* temp, size_t temp_bytes,
                  hipStream_t stream);

// ---- small kernels ----
hipError_t add_one_i32(int32_t* data, size_t len, hipStream_t stream);
hipError_t fill_mod_u32(uint32_t* out, size_t len, uint32_t modulus, hipStream_t stream);
hipError_t fill_u64(uint64_t* out, size_t len, uint64_t value, hipStream_t stream);
// dst[0..count) = src[0..count); both 16-byte aligned.
hipError_t copy_u64(const uint64_t* src, uint64_t* dst, size_t count, hipStream_t stream);
hipError_t set_u32(uint32_t* out, uint32_t value, hipStream_t stream);
hipError_t set_u64(uint64_t* out, uint64_t value, hipStream_t stream);

// Device error word of a temporary buffer (look-back failure or capacity
// overflow bits); every launcher reserves it at offset 0 of `temp`.
constexpr size_t kTempHeader = 256;

// Where look-back kernels report a bounded wait that gave up (kErrLookback).
// While an ErrorSink lives on a host thread, every launcher called on that
// thread passes the sink's device word to its look-back kernels, and the
// owner of the word checks it after its next synchronization. With no sink,
// a launcher reports into its temporary header and then overwrites the
// operator's output count with all ones (UINT32_MAX or UINT64_MAX), which is
// how the operator ABI makes such a failure visible to its caller.
uint32_t* error_sink();
class ErrorSink {
public:
    explicit ErrorSink(uint32_t* word);
    ~ErrorSink();
    ErrorSink(const ErrorSink&) = delete;
    ErrorSink& operator=(const ErrorSink&) = delete;

private:
    uint32_t* previous_;
};
// With no sink on this thread: after an operator's kernels (stream order),
// marks its output count all ones when `error` holds a report. A no-op under
// a sink, whose owner checks the word instead.
hipError_t report_error(const uint32_t* error, uint32_t* count32, uint64_t* count64, hipStream_t stream);

}  // namespace spk
