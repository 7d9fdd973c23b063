// extern "C" operator ABI (part 1 of include/sparkalog_hip.h). Argument
// validation and empty-input behaviour follow the reference's entry points
// in crates/sparkalog-execution/native/column_ops.cu:231-775; the work is
// done by the spk:: launchers in ops.hip.
#include <mutex>
#include <vector>

#include "ops.hpp"
#include "sparkalog_hip.h"

// Scratch of sparkalog_hip_join_u32_count_no_alloc: the join count's
// look-back header and one status word per 2,048 left rows. Each device
// holds its own copy until the process ends. Host code must never pass
// `scratch` itself to a kernel (it names a host-side stand-in); it uses the
// device address that hipGetSymbolAddress returned for the current device.
static __device__ __align__(256) unsigned char scratch[SPARKALOG_HIP_NO_ALLOC_SCRATCH_BYTES];

namespace {

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

// Empty inputs write their zero count with a one-thread kernel rather than
// hipMemsetAsync, so that callers may capture these calls in a HIP graph.
hipError_t zero_u32(uint32_t* p, void* stream) { return spk::set_u32(p, 0, as_stream(stream)); }
hipError_t zero_u64(uint64_t* p, void* stream) { return spk::set_u64(p, 0, as_stream(stream)); }

// Device address of `scratch` per device, looked up once per device (with
// that device current) on the first no-alloc call made there.
std::mutex g_scratch_mutex;
std::vector<void*> g_scratch_ptr;

hipError_t scratch_of_current_device(void** out) {
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (static_cast<size_t>(device) >= g_scratch_ptr.size()) g_scratch_ptr.resize(static_cast<size_t>(device) + 1, nullptr);
    if (g_scratch_ptr[device] == nullptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (p == nullptr) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace

extern "C" sparkalog_hip_error sparkalog_hip_add_one_i32(int32_t* data, size_t len, void* stream) {
    if (data == nullptr && len != 0) return hipErrorInvalidValue;
    return spk::add_one_i32(data, len, as_stream(stream));
}

extern "C" sparkalog_hip_error sparkalog_hip_stream_synchronize(void* stream) {
    return hipStreamSynchronize(as_stream(stream));
}

extern "C" sparkalog_hip_error sparkalog_hip_filter_u32_temporary_bytes(const uint32_t* flags, uint32_t* output,
                                                                        uint32_t* count, size_t len,
                                                                        size_t* temporary_bytes, void* stream) {
    (void)flags;
    (void)stream;
    if (temporary_bytes == nullptr || count == nullptr) return hipErrorInvalidValue;
    if (len != 0 && output == nullptr) return hipErrorInvalidValue;
    *temporary_bytes = spk::filter_temp_bytes(len);
    return hipSuccess;
}

extern "C" sparkalog_hip_error sparkalog_hip_filter_u32(const uint32_t* input, size_t len, uint32_t operation,
                                                        uint32_t operand, uint32_t* flags, uint32_t* output,
                                                        uint32_t* count, void* temporary, size_t temporary_bytes,
                                                        void* stream) {
    (void)flags;
    if (count == nullptr) return hipErrorInvalidValue;
    if (len == 0) return zero_u32(count, stream);
    if (input == nullptr || output == nullptr || temporary == nullptr) return hipErrorInvalidValue;
    return spk::filter(input, len, operation, operand, output, count, temporary, temporary_bytes,
                       as_stream(stream));
}

extern "C" sparkalog_hip_error sparkalog_hip_fill_mod_u32(uint32_t* output, size_t len, uint32_t modulus,
                                                          void* stream) {
    if (modulus == 0 || (output == nullptr && len != 0)) return hipErrorInvalidValue;
    return spk::fill_mod_u32(output, len, modulus, as_stream(stream));
}

extern "C" sparkalog_hip_error sparkalog_hip_copy_u64(const uint64_t* source, uint64_t* destination, size_t count,
                                                      void* stream) {
    if (count == 0) return hipSuccess;
    if (source == nullptr || destination == nullptr || (reinterpret_cast<uintptr_t>(source) & 15) ||
        (reinterpret_cast<uintptr_t>(destination) & 15))
        return hipErrorInvalidValue;
    return spk::copy_u64(source, destination, count, as_stream(stream));
}

extern "C" sparkalog_hip_error sparkalog_hip_join_u32_temporary_bytes(const uint64_t* counts, uint64_t* offsets,
                                                                      size_t left_rows, size_t* temporary_bytes,
                                                                      void* stream) {
    (void)stream;
    if (temporary_bytes == nullptr || (left_rows != 0 && (counts == nullptr || offsets == nullptr)))
        return hipErrorInvalidValue;
    *temporary_bytes = spk::join_temp_bytes(left_rows);
    return hipSuccess;
}

extern "C" sparkalog_hip_error sparkalog_hip_join_u32_count(const uint32_t* left_keys, size_t left_rows,
                                                            const uint32_t* index_keys, const uint32_t* index_starts,
                                                            size_t unique_keys, uint64_t* counts, uint64_t* offsets,
                                                            uint64_t* total, void* temporary, size_t temporary_bytes,
                                                            void* stream) {
    if (total == nullptr) return hipErrorInvalidValue;
    if (left_rows == 0) return zero_u64(total, stream);
    if (left_keys == nullptr || index_keys == nullptr || index_starts == nullptr || counts == nullptr ||
        offsets == nullptr || temporary == nullptr)
        return hipErrorInvalidValue;
    return spk::join_count(left_keys, left_rows, index_keys, index_starts, unique_keys, counts, offsets, total,
                           temporary, temporary_bytes, as_stream(stream));
}

extern "C" sparkalog_hip_error sparkalog_hip_join_u32_count_no_alloc(const uint32_t* left_keys, size_t left_rows,
                                                                     const uint32_t* index_keys,
                                                                     const uint32_t* index_starts, size_t unique_keys,
                                                                     uint64_t* counts, uint64_t* offsets,
                                                                     uint64_t* total, void* stream) {
    if (total == nullptr) return hipErrorInvalidValue;
    // sizes past the static scratch are refused before any device work
    if (left_rows > SPARKALOG_HIP_NO_ALLOC_MAX_LEFT_ROWS) return hipErrorInvalidValue;
    const size_t need = spk::join_temp_bytes(left_rows);
    if (need > SPARKALOG_HIP_NO_ALLOC_SCRATCH_BYTES) return hipErrorInvalidValue;
    if (left_rows == 0) return zero_u64(total, stream);
    if (left_keys == nullptr || index_keys == nullptr || index_starts == nullptr || counts == nullptr ||
        offsets == nullptr)
        return hipErrorInvalidValue;
    void* temp = nullptr;
    const hipError_t e = scratch_of_current_device(&temp);
    if (e != hipSuccess) return e;
    // join_count clears the header and the status words it uses with kernels
    // (no allocation, no hipMemsetAsync)
    return spk::join_count(left_keys, left_rows, index_keys, index_starts, unique_keys, counts, offsets, total, temp,
                           need, as_stream(stream));
}

extern "C" sparkalog_hip_error sparkalog_hip_join_u32_emit(const uint32_t* left_keys, size_t left_rows,
                                                           const uint32_t* index_keys, const uint32_t* index_starts,
                                                           const uint32_t* index_rows, size_t unique_keys,
                                                           const uint32_t* projection0, uint32_t projection0_side,
                                                           const uint32_t* projection1, uint32_t projection1_side,
                                                           const uint64_t* offsets, uint32_t* output0,
                                                           uint32_t* output1, void* stream) {
    if (left_rows == 0) return hipSuccess;
    if (left_keys == nullptr || index_keys == nullptr || index_starts == nullptr || index_rows == nullptr ||
        projection0 == nullptr || projection1 == nullptr || offsets == nullptr || output0 == nullptr ||
        output1 == nullptr)
        return hipErrorInvalidValue;
    return spk::join_emit(left_keys, left_rows, index_keys, index_starts, index_rows, unique_keys, projection0,
                          projection0_side, projection1, projection1_side, offsets, output0, output1,
                          as_stream(stream));
}

extern "C" sparkalog_hip_error sparkalog_hip_distinct_u32_temporary_bytes(uint64_t* packed, uint64_t* scratch,
                                                                          uint64_t* unique_rows, size_t rows,
                                                                          size_t* temporary_bytes, void* stream) {
    (void)stream;
    if (temporary_bytes == nullptr || unique_rows == nullptr) return hipErrorInvalidValue;
    if (rows != 0 && (packed == nullptr || scratch == nullptr)) return hipErrorInvalidValue;
    *temporary_bytes = spk::distinct_temp_bytes(rows);
    return hipSuccess;
}

extern "C" sparkalog_hip_error sparkalog_hip_distinct_u32(const uint32_t* first, const uint32_t* second, size_t rows,
                                                          uint64_t* packed, uint64_t* scratch, uint64_t* unique_rows,
                                                          uint32_t* output_first, uint32_t* output_second,
                                                          void* temporary, size_t temporary_bytes, void* stream) {
    if (unique_rows == nullptr) return hipErrorInvalidValue;
    if (rows == 0) return zero_u64(unique_rows, stream);
    if (first == nullptr || second == nullptr || packed == nullptr || scratch == nullptr ||
        output_first == nullptr || output_second == nullptr || temporary == nullptr)
        return hipErrorInvalidValue;
    spk::PairInput in{nullptr, first, second, nullptr};
    spk::PairOutput out{nullptr, output_first, output_second};
    return spk::distinct(in, rows, packed, scratch, out, unique_rows, temporary, temporary_bytes, 0,
                         as_stream(stream));
}

extern "C" sparkalog_hip_error sparkalog_hip_anti_join_u32_temporary_bytes(const uint32_t* flags, uint32_t* selected,
                                                                           uint32_t* selected_rows, size_t left_rows,
                                                                           size_t* temporary_bytes, void* stream) {
    (void)flags;
    (void)selected;
    (void)stream;
    if (temporary_bytes == nullptr || selected_rows == nullptr) return hipErrorInvalidValue;
    *temporary_bytes = spk::anti_join_temp_bytes(left_rows);
    return hipSuccess;
}

extern "C" sparkalog_hip_error sparkalog_hip_anti_join_u32(const uint32_t* left_first, const uint32_t* left_second,
                                                           size_t left_rows, const uint32_t* right_first,
                                                           const uint32_t* right_second, size_t right_rows,
                                                           uint32_t* flags, uint32_t* selected,
                                                           uint32_t* selected_rows, uint32_t* output_first,
                                                           uint32_t* output_second, void* temporary,
                                                           size_t temporary_bytes, void* stream) {
    (void)flags;
    (void)selected;
    if (selected_rows == nullptr) return hipErrorInvalidValue;
    if (left_rows == 0) return zero_u32(selected_rows, stream);
    if (left_first == nullptr || left_second == nullptr || output_first == nullptr || output_second == nullptr ||
        temporary == nullptr || (right_rows != 0 && (right_first == nullptr || right_second == nullptr)))
        return hipErrorInvalidValue;
    spk::PairInput l{nullptr, left_first, left_second, nullptr};
    spk::PairInput r{nullptr, right_first, right_second, nullptr};
    spk::PairOutput out{nullptr, output_first, output_second};
    return spk::sorted_anti_join(l, left_rows, r, right_rows, out, selected_rows, nullptr, temporary,
                                 temporary_bytes, as_stream(stream));
}

extern "C" sparkalog_hip_error sparkalog_hip_union_u32_temporary_bytes(const uint64_t* left, size_t left_rows,
                                                                       const uint64_t* right, size_t right_rows,
                                                                       uint64_t* merged, uint64_t* unique,
                                                                       uint64_t* unique_rows, size_t* temporary_bytes,
                                                                       void* stream) {
    (void)left;
    (void)right;
    (void)merged;
    (void)unique;
    (void)stream;
    if (temporary_bytes == nullptr || unique_rows == nullptr) return hipErrorInvalidValue;
    *temporary_bytes = spk::union_temp_bytes(left_rows, right_rows);
    return hipSuccess;
}

extern "C" sparkalog_hip_error sparkalog_hip_union_u32(const uint32_t* left_first, const uint32_t* left_second,
                                                       size_t left_rows, const uint32_t* right_first,
                                                       const uint32_t* right_second, size_t right_rows,
                                                       uint64_t* left, uint64_t* right, uint64_t* merged,
                                                       uint64_t* unique, uint64_t* unique_rows,
                                                       uint32_t* output_first, uint32_t* output_second,
                                                       void* temporary, size_t temporary_bytes, void* stream) {
    (void)left;
    (void)right;
    (void)merged;
    (void)unique;
    if (unique_rows == nullptr) return hipErrorInvalidValue;
    if (left_rows + right_rows == 0) return zero_u64(unique_rows, stream);
    if ((left_rows != 0 && (left_first == nullptr || left_second == nullptr)) ||
        (right_rows != 0 && (right_first == nullptr || right_second == nullptr)) || output_first == nullptr ||
        output_second == nullptr || temporary == nullptr)
        return hipErrorInvalidValue;
    spk::PairInput l{nullptr, left_first, left_second, nullptr};
    spk::PairInput r{nullptr, right_first, right_second, nullptr};
    spk::PairOutput out{nullptr, output_first, output_second};
    return spk::sorted_union(l, left_rows, r, right_rows, out, unique_rows, temporary, temporary_bytes,
                             as_stream(stream));
}
