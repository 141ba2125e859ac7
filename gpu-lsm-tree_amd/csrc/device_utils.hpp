// Wave64 / workgroup helpers shared by the LSM and FliX kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gdict {

using u8 = std::uint8_t;
using u32 = std::uint32_t;
using u64 = std::uint64_t;

constexpr int kWave = 64;

__device__ __forceinline__ u32 lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ u64 lanemask_lt() {
  const u32 lane = lane_id();
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

__device__ __forceinline__ u64 ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ u32 popc64(u64 x) {
  return static_cast<u32>(__popcll(static_cast<unsigned long long>(x)));
}
// 1-based index of the lowest set bit, 0 when x == 0.
__device__ __forceinline__ u32 ffs64(u64 x) {
  return static_cast<u32>(__ffsll(static_cast<unsigned long long>(x)));
}

// Device-scope relaxed atomics used by the single-pass (look-back) scans.
__device__ __forceinline__ u64 load_relaxed(const u64 *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_relaxed(u64 *p, u64 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// LDS counters that the lanes of a wave update for each other (one lane
// writes, all read on the next item): relaxed workgroup-scope accesses, so
// the compiler neither caches nor forwards them.  The pointer must point
// into LDS; the address-space cast keeps them ds_read / ds_write (a
// volatile access through a generic pointer becomes a flat access that
// also waits for every outstanding global load).
using lds_u32 = __attribute__((address_space(3))) u32;
__device__ __forceinline__ u32 lds_load(u32 *p) {
  return __hip_atomic_load((lds_u32 *)(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_store(u32 *p, u32 v) {
  __hip_atomic_store((lds_u32 *)(p), v, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Wave-wide inclusive scan of a u32 through DPP-backed shuffles.
__device__ __forceinline__ u32 wave_inclusive_scan(u32 v) {
  const u32 lane = lane_id();
#pragma unroll
  for (int offset = 1; offset < kWave; offset <<= 1) {
    const u32 other = __shfl_up(v, offset, kWave);
    if (lane >= static_cast<u32>(offset)) v += other;
  }
  return v;
}

__device__ __forceinline__ u64 wave_sum_u64(u64 v) {
#pragma unroll
  for (int offset = kWave / 2; offset > 0; offset >>= 1)
    v += __shfl_xor(v, offset, kWave);
  return v;
}

__device__ __forceinline__ u32 wave_sum_u32(u32 v) {
#pragma unroll
  for (int offset = kWave / 2; offset > 0; offset >>= 1)
    v += __shfl_xor(v, offset, kWave);
  return v;
}

// Workgroup exclusive scan for blockDim.x == THREADS (a multiple of 64).
// `scratch` needs THREADS/64 + 1 words.  Returns the exclusive prefix; the
// total is left in scratch[THREADS/64].
template <int THREADS>
__device__ __forceinline__ u32 block_exclusive_scan(u32 v, u32 *scratch) {
  constexpr int kWaves = THREADS / kWave;
  const u32 lane = lane_id();
  const u32 wave = threadIdx.x / kWave;
  const u32 inclusive = wave_inclusive_scan(v);
  if (lane == kWave - 1) scratch[wave] = inclusive;
  __syncthreads();
  if (threadIdx.x == 0) {
    u32 run = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const u32 t = scratch[w];
      scratch[w] = run;
      run += t;
    }
    scratch[kWaves] = run;
  }
  __syncthreads();
  return scratch[wave] + inclusive - v;
}

// block_exclusive_scan for u64 values; `scratch` needs THREADS/64 + 1
// words and holds the total in scratch[THREADS/64] afterwards.
template <int THREADS>
__device__ __forceinline__ u64 block_exclusive_scan_u64(u64 v, u64 *scratch) {
  constexpr int kWaves = THREADS / kWave;
  const u32 lane = lane_id();
  const u32 wave = threadIdx.x / kWave;
  u64 inclusive = v;
#pragma unroll
  for (int offset = 1; offset < kWave; offset <<= 1) {
    const u64 other = __shfl_up(inclusive, offset, kWave);
    if (lane >= static_cast<u32>(offset)) inclusive += other;
  }
  if (lane == kWave - 1) scratch[wave] = inclusive;
  __syncthreads();
  if (threadIdx.x == 0) {
    u64 run = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const u64 t = scratch[w];
      scratch[w] = run;
      run += t;
    }
    scratch[kWaves] = run;
  }
  __syncthreads();
  return scratch[wave] + inclusive - v;
}

// --------------------------------------------------------------------------
// Single-pass chained scan ("decoupled look-back") over tiles that take
// tickets in launch order.  A status word packs
//   bits 63..62 : flag (0 = not ready, 1 = aggregate, 2 = inclusive prefix)
//   bits 61..0  : value
constexpr u64 kFlagShift = 62;
constexpr u64 kFlagAggregate = 1ull << kFlagShift;
constexpr u64 kFlagPrefix = 2ull << kFlagShift;
constexpr u64 kValueMask = (1ull << kFlagShift) - 1ull;

// Called by every thread of one wave (wave 0).  Publishes `aggregate` for
// `tile`, walks back over predecessors and returns the exclusive prefix.
// Tiles are processed in ticket order, so every predecessor is resident or
// finished and publishes its aggregate without waiting: the walk terminates.
__device__ __forceinline__ u64 wave_lookback(u64 *status, u32 tile,
                                             u64 aggregate) {
  const u32 lane = lane_id();
  if (tile == 0) {
    if (lane == 0) store_relaxed(&status[0], kFlagPrefix | aggregate);
    return 0;
  }
  if (lane == 0) store_relaxed(&status[tile], kFlagAggregate | aggregate);
  u64 exclusive = 0;
  long long window_end = static_cast<long long>(tile) - 1;  // inclusive
  while (true) {
    const long long t = window_end - static_cast<long long>(lane);
    u64 word = kFlagPrefix;  // lanes before tile 0 read as an empty prefix
    if (t >= 0) {
      do {
        word = load_relaxed(&status[t]);
      } while ((word >> kFlagShift) == 0);
    }
    const u64 prefix_lanes = ballot((word & kFlagPrefix) != 0);
    // first lane (closest predecessor) that holds an inclusive prefix
    const u32 stop = prefix_lanes ? ffs64(prefix_lanes) - 1
                                  : static_cast<u32>(kWave);
    const u64 contrib =
        (lane <= stop && t >= 0) ? (word & kValueMask) : 0ull;
    exclusive += wave_sum_u64(contrib);
    if (prefix_lanes) break;
    window_end -= kWave;
  }
  if (lane == 0)
    store_relaxed(&status[tile], kFlagPrefix | (exclusive + aggregate));
  return exclusive;
}

}  // namespace gdict
