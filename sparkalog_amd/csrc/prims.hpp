// Device primitives for the gfx950 operator kernels: wave64 ballots and
// scans, block scans over 256 threads (4 waves), and decoupled look-back
// (Merrill & Garland 2016) for single-pass tile prefixes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spk {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

// Look-back status word: flag in bits 63..62, value in bits 61..0.
constexpr uint64_t kFlagAggregate = 1ull << 62;
constexpr uint64_t kFlagPrefix = 2ull << 62;
constexpr uint64_t kValueMask = (1ull << 62) - 1;
// Bound on polls of one predecessor before a tile gives up and raises the
// device error flag; every wave therefore finishes even on a broken chain.
constexpr uint32_t kSpinLimit = 1u << 24;

// Device-side error bits, written into the scratch block of each launch.
constexpr uint32_t kErrLookback = 1u;
constexpr uint32_t kErrCapacity = 2u;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint32_t wave_id() { return threadIdx.x / kWave; }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Number of set bits of `mask` strictly below this lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Valid lanes whose `Bits`-bit digit equals this lane's (the stable rank of a
// lane within its digit is lanes_below of the result). One ballot per digit
// bit; sb is all ones where this lane's bit is set, so ~(ballot ^ sb) keeps
// the lanes that agree with it on that bit: one 32-bit bit op per half
// (v_bitop3) instead of a 64-bit select.
template <int Bits>
__device__ __forceinline__ uint64_t match_digit(uint32_t digit, bool valid) {
    const uint64_t v = ballot(valid);
    uint32_t lo = static_cast<uint32_t>(v), hi = static_cast<uint32_t>(v >> 32);
#pragma unroll
    for (int b = 0; b < Bits; ++b) {
        const uint32_t sb = static_cast<uint32_t>(__builtin_amdgcn_sbfe(static_cast<int32_t>(digit), b, 1));
        const uint64_t m = ballot(sb != 0);
        lo &= ~(static_cast<uint32_t>(m) ^ sb);
        hi &= ~(static_cast<uint32_t>(m >> 32) ^ sb);
    }
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

template <typename T>
__device__ __forceinline__ T wave_inclusive_sum(T v) {
#pragma unroll
    for (int offset = 1; offset < kWave; offset <<= 1) {
        T other = __shfl_up(v, offset, kWave);
        if (static_cast<int>(lane_id()) >= offset) v += other;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int offset = kWave / 2; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, kWave);
    return v;
}

// Block-wide exclusive sum of one value per thread over a block of kWaves
// waves. `lds` holds kWaves entries. Returns the exclusive prefix and the
// block total in *total.
template <typename T, int kWaves = kWavesPerBlock>
__device__ __forceinline__ T block_exclusive_sum(T v, T* total, T* lds) {
    const T inclusive = wave_inclusive_sum(v);
    if (lane_id() == kWave - 1) lds[wave_id()] = inclusive;
    __syncthreads();
    T wave_offset = 0, sum = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const T t = lds[w];
        if (w < static_cast<int>(wave_id())) wave_offset += t;
        sum += t;
    }
    *total = sum;
    __syncthreads();
    return wave_offset + inclusive - v;
}

// Exclusive scan in place of n >= 1 u64 values by one block of kThreads
// threads; *total = their sum. Chunks of kThreads * kItems consecutive values
// (thread t owns items [t * kItems, t * kItems + kItems) of a chunk) are each
// loaded at once (indexes clamped) and scanned with one block step, so a
// thread waits for memory once per chunk rather than once per value.
template <int kThreads, int kItems>
__device__ __forceinline__ void block_scan_inplace(uint64_t* v, size_t n, uint64_t* total, uint64_t* lds) {
    uint64_t carry = 0;
    for (size_t c0 = 0; c0 < n; c0 += static_cast<size_t>(kThreads) * kItems) {
        uint64_t x[kItems];
        uint64_t mine = 0;
#pragma unroll
        for (int q = 0; q < kItems; ++q) {
            const size_t idx = c0 + static_cast<size_t>(threadIdx.x) * kItems + q;
            const uint64_t val = v[idx < n ? idx : n - 1];
            x[q] = idx < n ? val : 0;
            mine += x[q];
        }
        uint64_t all = 0;
        uint64_t run = carry + block_exclusive_sum<uint64_t, kThreads / kWave>(mine, &all, lds);
#pragma unroll
        for (int q = 0; q < kItems; ++q) {
            const size_t idx = c0 + static_cast<size_t>(threadIdx.x) * kItems + q;
            if (idx < n) v[idx] = run;
            run += x[q];
        }
        carry += all;
    }
    if (threadIdx.x == 0) *total = carry;
}

// 16 bytes per lane from global memory straight into LDS (gfx950
// global_load_lds_dwordx4), with no VGPRs holding the data: lane i of the
// wave writes dst_wave[i], where dst_wave (wave-uniform) is the wave's slot
// of 64 vectors. The compiler waits for the load before the next barrier.
__device__ __forceinline__ void load_lds16(const uint4* src, uint4* dst_wave) {
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                     (__attribute__((address_space(3))) void*)dst_wave, 16, 0, 0);
}

__device__ __forceinline__ uint64_t load_status(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_status(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of `aggregate` over tiles 0..tile-1, computed by one
// whole wave (call from every lane of a single wave; all lanes return the
// same value). status[] must be zero for every tile at kernel start; `base`
// is added to tile 0's prefix. Look-back polls 64 predecessors per step.
__device__ inline uint64_t wave_lookback(uint64_t* status, uint32_t tile, uint64_t aggregate,
                                         uint64_t base, uint32_t* error) {
    const uint32_t lane = lane_id();
    if (tile == 0) {
        if (lane == 0) store_status(&status[0], kFlagPrefix | ((base + aggregate) & kValueMask));
        return base;
    }
    if (lane == 0) store_status(&status[tile], kFlagAggregate | (aggregate & kValueMask));
    uint64_t exclusive = 0;
    int64_t window_end = static_cast<int64_t>(tile) - 1;  // newest predecessor in the window
    for (;;) {
        const int64_t j = window_end - static_cast<int64_t>(lane);
        uint64_t word = 0;
        uint32_t spins = 0;
        bool failed = false;
        for (;;) {
            // Tiles before 0 read as a prefix of `base` so the window ends there.
            word = j >= 0 ? load_status(&status[j]) : (kFlagPrefix | base);
            const uint64_t empty = ballot((word >> 62) == 0);
            const uint64_t prefix = ballot((word >> 62) == 2);
            // Lanes up to and including the first prefix lane must be ready.
            const uint64_t needed = prefix ? ((prefix & (~prefix + 1)) << 1) - 1 : ~0ull;
            if ((empty & needed) == 0) break;
            if (++spins >= kSpinLimit) {
                failed = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (failed) {
            if (lane == 0) atomicOr(error, kErrLookback);
            return exclusive;
        }
        const uint64_t prefix = ballot((word >> 62) == 2);
        const uint32_t first_prefix = prefix ? static_cast<uint32_t>(__builtin_ctzll(prefix)) : 64u;
        const uint64_t contrib = lane <= first_prefix ? (word & kValueMask) : 0;
        exclusive += wave_sum(contrib);
        if (prefix) break;
        window_end -= kWave;
    }
    if (lane == 0) store_status(&status[tile], kFlagPrefix | ((exclusive + aggregate) & kValueMask));
    return exclusive;
}

// Dynamic tile index: tiles are numbered in the order blocks start, which
// guarantees every predecessor of a tile is already resident (forward
// progress for look-back).
__device__ __forceinline__ uint32_t next_tile(uint32_t* counter, uint32_t* lds_slot) {
    if (threadIdx.x == 0) *lds_slot = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t t = *lds_slot;
    __syncthreads();
    return t;
}

// Block-cooperative search: the first index i in [lo, hi) where the monotone
// predicate `before(i)` (true, ..., true, false, ..., false) is false, or hi.
// Every thread of the block must call it with the same arguments. Each step
// probes 256 evenly spaced indexes in parallel, so a range of 2^32 rows takes
// four dependent rounds instead of 32.
template <typename Pred>
__device__ __forceinline__ size_t block_search(size_t lo, size_t hi, Pred before) {
    while (lo < hi) {
        const size_t span = hi - lo;
        const size_t stride = span <= static_cast<size_t>(kBlock) ? 1 : (span + kBlock - 1) / kBlock;
        const size_t probe = lo + static_cast<size_t>(threadIdx.x) * stride;
        const bool in_range = probe < hi;
        const int count = __syncthreads_count(in_range && before(probe));
        if (count == 0) return lo;
        if (stride == 1) return lo + static_cast<size_t>(count);
        // answer in (probe[count-1], probe[count]]
        const size_t last_true = lo + static_cast<size_t>(count - 1) * stride;
        const size_t next = lo + static_cast<size_t>(count) * stride;
        lo = last_true + 1;
        if (next < hi) hi = next;
    }
    return lo;
}

__device__ __forceinline__ uint64_t pack_pair(uint32_t a, uint32_t b) {
    return (static_cast<uint64_t>(a) << 32) | b;
}

}  // namespace spk
