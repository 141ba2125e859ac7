// Stable LSD radix sort of (u32 key, u32 payload) pairs for gfx950.
//
// One global-histogram launch computes the four 8-bit digit histograms in a
// single read of the keys; every digit pass is then ONE launch ("onesweep"):
// a 4096-key tile is ranked in registers with wave64 ballots (a multisplit:
// eight ballots give each lane the mask of lanes holding the same digit),
// tiles chain their per-digit offsets through a decoupled look-back, and the
// tile is re-ordered in LDS so the global stores leave as contiguous runs per
// digit.  The first pass writes payload = input position, which is what the
// callers need for "newest wins" (a stable sort keeps equal keys in input
// order, so the last of a run of equal keys is the newest update).
#pragma once

#include "device_utils.hpp"

namespace gdict {
namespace rs {

constexpr int kThreads = 256;
constexpr int kRadix = 256;
constexpr int kMaxPasses = 8;
constexpr int kWaves = kThreads / kWave;

template <class K>
constexpr int passes_for() {
  return static_cast<int>(sizeof(K));
}

// Digit histograms.  A block counts 16 keys per thread per round in LDS
// (the 16 loads are issued together), then adds its counts into one of
// kHistRows partial rows (row = block % kHistRows), so at most
// blocks / kHistRows atomics meet on one address; scan_kernel sums the rows.
// (Folding the scan into the histogram's last block through a done counter
// was measured slower on MI355X: every block's fence and the atomics of all
// blocks on one counter cost more than the extra launch.)
constexpr int kHistItems = 16;
constexpr int kHistRows = 32;

inline u32 hist_blocks_for(u64 n) {
  const u64 b = (n + kThreads * kHistItems - 1) / (kThreads * kHistItems);
  return static_cast<u32>(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

// rows[kHistRows][kPasses][kRadix] (zeroed) += counts of digits
// first .. first + kPasses - 1, digit p = (key >> 8p) & 0xff.
template <class K, int kPasses>
static __global__ __launch_bounds__(kThreads) void histogram_kernel(
    const K *__restrict__ keys, u32 n, u32 first, u32 *__restrict__ rows) {
  __shared__ u32 s_hist[kPasses][kRadix];
  for (int i = threadIdx.x; i < kPasses * kRadix; i += kThreads)
    (&s_hist[0][0])[i] = 0;
  __syncthreads();
  const u64 round = u64(kThreads) * kHistItems;
  for (u64 base = u64(blockIdx.x) * round; base < n;
       base += u64(gridDim.x) * round) {
    K k[kHistItems];
#pragma unroll
    for (int j = 0; j < kHistItems; ++j) {
      const u64 i = base + u64(j) * kThreads + threadIdx.x;
      k[j] = i < n ? keys[i] : K(0);
    }
#pragma unroll
    for (int j = 0; j < kHistItems; ++j) {
      if (base + u64(j) * kThreads + threadIdx.x < n) {
#pragma unroll
        for (int p = 0; p < kPasses; ++p)
          atomicAdd(&s_hist[p][static_cast<u32>(k[j] >> (8 * (first + p))) &
                               0xffu],
                    1u);
      }
    }
  }
  __syncthreads();
  u32 *row = rows + u64(blockIdx.x % kHistRows) * kPasses * kRadix;
  for (int i = threadIdx.x; i < kPasses * kRadix; i += kThreads) {
    const u32 c = (&s_hist[0][0])[i];
    if (c) atomicAdd(&row[i], c);
  }
}

// Sums the kHistRows partial rows of pass blockIdx.x and writes the
// exclusive scan over digits -> global digit offsets.
static __global__ __launch_bounds__(kRadix) void scan_kernel(
    const u32 *__restrict__ rows, u32 passes, u32 *__restrict__ offsets) {
  __shared__ u32 scratch[kRadix / kWave + 1];
  const u32 p = blockIdx.x;
  u32 c[kHistRows];
#pragma unroll
  for (int r = 0; r < kHistRows; ++r)
    c[r] = rows[(u64(r) * passes + p) * kRadix + threadIdx.x];
  u32 v = 0;
#pragma unroll
  for (int r = 0; r < kHistRows; ++r) v += c[r];
  offsets[p * kRadix + threadIdx.x] = block_exclusive_scan<kRadix>(v, scratch);
}

// Payload policies of a digit pass: payload(idx, valid, key, shift)
// returns the 32-bit payload of input element idx (and may rewrite key
// bits that no digit of the sort reads).
struct PayloadIn {  // the payload array of the previous pass
  const u32 *vals;
  template <class K>
  __device__ __forceinline__ u32 operator()(u64 idx, bool valid, K &,
                                            u32) const {
    return valid ? vals[idx] : 0u;
  }
};
struct PayloadPos {  // the input position
  template <class K>
  __device__ __forceinline__ u32 operator()(u64 idx, bool, K &, u32) const {
    return static_cast<u32>(idx);
  }
};

// One digit pass.  status has tiles*kRadix words and must be zero;
// *ticket must be zero.  kBlock threads (256 or 512) x kItems keys per
// tile; threads below 256 own one digit each for the per-digit bookkeeping.
template <class K, class Payload, int kItems, int kBlock>
__global__ __launch_bounds__(kBlock) void onesweep_kernel(
    const K *__restrict__ keys_in, Payload payload,
    K *__restrict__ keys_out, u32 *__restrict__ vals_out, u32 n,
    u32 shift_arg, const u32 *__restrict__ digit_offsets, u64 *status,
    u32 *ticket, u32 *__restrict__ dest_out, const u32 *shift_dev,
    const u32 *__restrict__ tile_prefix) {
  constexpr int kTile = kBlock * kItems;
  // the digit shift: an argument, or computed on the device by an earlier
  // launch
  const u32 shift = shift_dev ? *shift_dev : shift_arg;
  constexpr int kBlockWaves = kBlock / kWave;
  __shared__ u32 s_tile;
  __shared__ u32 s_wave_hist[kBlockWaves][kRadix];
  __shared__ u32 s_block_start[kRadix];
  __shared__ u32 s_global_base[kRadix];
  __shared__ u32 s_scan[kBlockWaves + 1];
  __shared__ K s_keys[kTile];
  __shared__ u32 s_vals[kTile];

  const u32 tid = threadIdx.x;
  const u32 lane = lane_id();
  const u32 wave = tid / kWave;
  // with precomputed per-tile digit prefixes (tile_prefix[d * tiles + t]:
  // keys of digit d in tiles before t; digit_offsets then holds the digit
  // totals) the tiles are independent: no ticket and no look-back
  if (tid == 0) s_tile = tile_prefix ? blockIdx.x : atomicAdd(ticket, 1u);
  for (int i = tid; i < kBlockWaves * kRadix; i += kBlock)
    (&s_wave_hist[0][0])[i] = 0;
  __syncthreads();
  const u32 tile = s_tile;
  const u64 base = u64(tile) * kTile;

  K key[kItems];
  u32 val[kItems], rank[kItems];
  const u64 wave_base = base + u64(wave) * (kWave * kItems);
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const u64 idx = wave_base + u64(j) * kWave + lane;
    const bool valid = idx < n;
    key[j] = valid ? keys_in[idx] : ~K(0);
    val[j] = payload(idx, valid, key[j], shift);
  }

  // Wave-local stable ranking (item-major, then lane order == input order).
  u32 *my_hist = s_wave_hist[wave];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const u64 idx = wave_base + u64(j) * kWave + lane;
    const bool valid = idx < n;
    const u32 d = static_cast<u32>(key[j] >> shift) & 0xffu;
    u64 match = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const u64 vote = ballot(bit);
      match &= bit ? vote : ~vote;
    }
    u32 prior = 0;
    if (valid) prior = lds_load(&my_hist[d]);
    rank[j] = prior + popc64(match & lanemask_lt());
    // the highest lane of each digit group advances the counter
    if (valid && (match >> lane) == 1ull)
      lds_store(&my_hist[d], prior + popc64(match));
  }
  __syncthreads();

  // Per digit: wave prefixes, tile total, global base via look-back.
  const u32 d = tid;
  u32 total = 0;
  u64 exclusive = 0;
  if (tid < kRadix) {
#pragma unroll
    for (int w = 0; w < kBlockWaves; ++w) {
      const u32 c = s_wave_hist[w][d];
      s_wave_hist[w][d] = total;
      total += c;
    }
    // look-back: every digit owns one column of the status array
    u64 *column = status;  // status[t * kRadix + d]
    if (tile_prefix) {
      exclusive = tile_prefix[u64(d) * gridDim.x + tile];
    } else if (tile == 0) {
      store_relaxed(&column[d], kFlagPrefix | total);
    } else {
      store_relaxed(&column[u64(tile) * kRadix + d], kFlagAggregate | total);
      long long t = static_cast<long long>(tile) - 1;
      while (true) {
        u64 word;
        do {
          word = load_relaxed(&column[u64(t) * kRadix + d]);
        } while ((word >> kFlagShift) == 0);
        exclusive += word & kValueMask;
        if (word & kFlagPrefix) break;
        --t;
      }
      store_relaxed(&column[u64(tile) * kRadix + d],
                    kFlagPrefix | (exclusive + total));
    }
  }
  const u32 block_start = block_exclusive_scan<kBlock>(total, s_scan);
  // the digit's start: given, or a scan of the digit totals
  u32 digit_start = 0;
  if (tile_prefix) {
    __shared__ u32 s_scan2[kBlockWaves + 1];
    digit_start = block_exclusive_scan<kBlock>(
        tid < kRadix ? digit_offsets[tid] : 0u, s_scan2);
  } else if (tid < kRadix) {
    digit_start = digit_offsets[d];
  }
  if (tid < kRadix) {
    s_block_start[d] = block_start;
    s_global_base[d] = digit_start + static_cast<u32>(exclusive) - block_start;
  }
  __syncthreads();

  // Scatter into LDS in digit order.
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const u64 idx = wave_base + u64(j) * kWave + lane;
    if (idx < n) {
      const u32 dj = static_cast<u32>(key[j] >> shift) & 0xffu;
      const u32 lp = s_block_start[dj] + s_wave_hist[wave][dj] + rank[j];
      s_keys[lp] = key[j];
      s_vals[lp] = val[j];
      // destination of input element idx, written in input order
      if (dest_out) dest_out[idx] = s_global_base[dj] + lp;
    }
  }
  __syncthreads();

  const u32 count =
      base >= n ? 0u : static_cast<u32>(n - base < u64(kTile) ? n - base
                                                              : u64(kTile));
  for (u32 i = tid; i < count; i += kBlock) {
    const K k = s_keys[i];
    const u32 pos = s_global_base[static_cast<u32>(k >> shift) & 0xffu] + i;
    keys_out[pos] = k;
    if (vals_out) vals_out[pos] = s_vals[i];
  }
}

// Clears a[0..na) and b[0..nb) in one launch.
static __global__ void zero_two_kernel(u32 *a, u64 na, u32 *b, u64 nb) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < na + nb;
       i += stride) {
    if (i < na) a[i] = 0;
    else b[i - na] = 0;
  }
}

static __global__ void zero_u32_kernel(u32 *p, u64 n) {
  const u64 stride = u64(gridDim.x) * blockDim.x;
  for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = 0;
}


// Tile = 4096 keys: 1024 threads x 4 keys below 2^22 keys (the 256-tile
// passes of 2^20-key batches run one tile per CU, so more waves per tile
// hide more latency), 256 x 16 above.  (2048-key tiles were measured
// slower on MI355X: more look-back steps.)
inline int block_for(u64 n) { return n < (u64(1) << 22) ? 1024 : 256; }

inline u32 tiles_for(u64 n) {
  const u64 tile = 4096;
  return static_cast<u32>((n + tile - 1) / tile);
}

// Workspace (u32 words) of a sort of n keys in `passes` digit passes, not
// counting the ping-pong buffers:
//   rows    [kHistRows][passes][256]     partial histograms
//   offsets [passes][256]                global digit offsets
//   tickets [64]                         tile tickets, one per pass
//   status  u64 [passes][tiles][256]     look-back words
// Everything but the offsets starts at zero (one clearing launch).
struct Control {
  u32 *rows, *offsets, *tickets;
  u64 *status;
};

inline u64 control_words(u64 n, int passes = 4) {
  return u64(kHistRows) * passes * kRadix + u64(passes) * kRadix + 64 +
         2ull * passes * u64(tiles_for(n)) * kRadix;
}

inline Control control_layout(u32 *control, int passes) {
  Control c;
  c.rows = control;
  c.offsets = control + u64(kHistRows) * passes * kRadix;
  c.tickets = c.offsets + passes * kRadix;
  c.status = reinterpret_cast<u64 *>(c.tickets + 64);  // even word offset
  return c;
}

// Clears the control words (at least control_words(n, passes); callers may
// extend the region with their own state) and an optional second region,
// in one launch.
inline void clear_control(u32 *control, u64 words, hipStream_t stream,
                          u32 *extra = nullptr, u64 extra_words = 0) {
  const u64 total = words + extra_words;
  zero_two_kernel<<<static_cast<u32>(std::min<u64>((total + 255) / 256, 4096)),
                    256, 0, stream>>>(control, words, extra, extra_words);
}

template <class K, class Payload>
inline void launch_pass(int block, u32 tiles, const K *src_k,
                        Payload payload, K *dst_k, u32 *dst_v, u32 n,
                        u32 shift, const u32 *offsets, u64 *status,
                        u32 *ticket, hipStream_t stream,
                        u32 *dest = nullptr,
                        const u32 *shift_dev = nullptr,
                        const u32 *tile_prefix = nullptr) {
  if (block == 1024)
    onesweep_kernel<K, Payload, 4, 1024><<<tiles, 1024, 0, stream>>>(
        src_k, payload, dst_k, dst_v, n, shift, offsets, status, ticket,
        dest, shift_dev, tile_prefix);
  else if (block == 512)
    onesweep_kernel<K, Payload, 8, 512><<<tiles, 512, 0, stream>>>(
        src_k, payload, dst_k, dst_v, n, shift, offsets, status, ticket,
        dest, shift_dev, tile_prefix);
  else
    onesweep_kernel<K, Payload, 16, 256><<<tiles, 256, 0, stream>>>(
        src_k, payload, dst_k, dst_v, n, shift, offsets, status, ticket,
        dest, shift_dev, tile_prefix);
}

// Sorts n (key, position) pairs stably.  keys_in is read only; the sorted
// keys and positions end in (keys_a, vals_a) (an even number of passes).
// keys_b/vals_b are scratch.  `control` holds control_words(n, passes) u32
// words.  All launches go to `stream`.
// clear_words (0 = control_words(n, passes)) words of `control` and
// extra[0..extra_words) are cleared by the first launch, so callers can
// place zero-initialised state of their own behind the sort's.
template <class K>
inline hipError_t sort_pairs(const K *keys_in, u32 n, K *keys_a, u32 *vals_a,
                             K *keys_b, u32 *vals_b, u32 *control,
                             hipStream_t stream, int block = 0,
                             u64 clear_words = 0, u32 *extra = nullptr,
                             u64 extra_words = 0) {
  constexpr int kPasses = passes_for<K>();
  if (n == 0) return hipSuccess;
  const int shape = block ? block : block_for(n);
  const u32 tiles = tiles_for(n);
  const Control c = control_layout(control, kPasses);
  clear_control(control, clear_words ? clear_words : control_words(n, kPasses),
                stream, extra, extra_words);
  histogram_kernel<K, kPasses><<<hist_blocks_for(n), kThreads, 0, stream>>>(
      keys_in, n, 0, c.rows);
  scan_kernel<<<kPasses, kRadix, 0, stream>>>(c.rows, kPasses, c.offsets);
  u32 *tickets = c.tickets;
  u64 *status = c.status;
  const u32 *offsets = c.offsets;
  const K *src_k = keys_in;
  const u32 *src_v = nullptr;
  K *dst_k = keys_b;
  u32 *dst_v = vals_b;
  for (int p = 0; p < kPasses; ++p) {
    u64 *pass_status = status + u64(p) * tiles * kRadix;
    if (p == 0) {
      launch_pass<K>(shape, tiles, src_k, PayloadPos{}, dst_k, dst_v, n, 0,
                     offsets, pass_status, tickets, stream);
    } else {
      launch_pass<K>(shape, tiles, src_k, PayloadIn{src_v}, dst_k, dst_v, n,
                     8u * p, offsets + p * kRadix, pass_status, tickets + p,
                     stream);
    }
    src_k = dst_k;
    src_v = dst_v;
    if (dst_k == keys_b) {
      dst_k = keys_a;
      dst_v = vals_a;
    } else {
      dst_k = keys_b;
      dst_v = vals_b;
    }
  }
  return hipGetLastError();
}

}  // namespace rs
}  // namespace gdict
