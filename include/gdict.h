/*
 * gdict: MI355X-native GPU dictionaries behind a plain C ABI.
 *
 * Two structures, both device resident, both driven from the host with
 * device pointers supplied by the caller (the caller owns every input and
 * output buffer; a handle owns its own device memory):
 *
 *   gd_lsm_*   an LSM dictionary over uint32 keys and uint32 values that
 *              replaces the GPULSMOpt class of the reference
 *              (GPULSMOpt.cuh:5599-7805) and the batch structs of
 *              gpu_dictionary_adapter.cuh:16-49;
 *   gd_flix_*  a coarse-granular updatable sorted index (bucket layer of
 *              fixed-size nodes chained into an allocation region) that
 *              replaces the FliX cg_rtx_index_updates class
 *              (FliX/impl_cg_rtx_index_updates.cuh:1422-2693, non-OptiX
 *              members), with 32- or 64-bit keys and values.
 *
 * Errors never cross the ABI as C++ exceptions: every call returns a
 * gd_status and stores a detail string retrievable with
 * gd_last_error_message().  All work of one handle is totally ordered: each
 * call makes its stream wait for the previous call's work and records a new
 * completion event (the reference's begin/end_operation,
 * GPULSMOpt.cuh:6852-6858), and a per-handle mutex serialises host calls
 * (GPULSMOpt.cuh:7685).  Calls are asynchronous with respect to the host
 * except where a function says otherwise.
 */
#ifndef GDICT_H
#define GDICT_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gd_status {
  GD_OK = 0,
  GD_INVALID_ARGUMENT = 1,   /* std::invalid_argument in the reference   */
  GD_OUT_OF_MEMORY = 2,      /* std::bad_alloc                           */
  GD_LENGTH = 3,             /* std::length_error                        */
  GD_PUBLICATION_FAILED = 4, /* std::runtime_error after a failed epoch   */
  GD_HIP_ERROR = 5,          /* std::runtime_error from CUDA_CHECK        */
  GD_CAPACITY = 6            /* allocation region / budget exhausted     */
} gd_status;

/* Publication status bits (GPULSMOpt.cuh:535-541). */
#define GD_PUBLICATION_SUCCESS 0u
#define GD_PUBLICATION_JOB_OVERFLOW (1u << 0)
#define GD_PUBLICATION_OUTPUT_OVERFLOW (1u << 3)
#define GD_PUBLICATION_JOB_TOO_LARGE (1u << 4)
#define GD_PUBLICATION_LEVEL_OVERFLOW (1u << 5)

/* Replaces DictionaryConfig (gpu_dictionary_adapter.cuh:16-23).
 * delete_order / base_delete_values are accepted for source compatibility;
 * the reference declares them and never reads them. */
typedef struct gd_config {
  uint64_t max_elements;
  uint64_t batch_capacity;
  uint64_t level_zero_capacity;
  uint32_t delete_order;
  uint32_t base_delete_values;
} gd_config;

typedef struct gd_lsm gd_lsm;

/* Detail string of the last failing call on `handle` (gd_lsm* or gd_flix*);
 * for a NULL handle, the last failing create call of this thread. */
const char *gd_last_error_message(const void *handle);

/* ------------------------------------------------------------------ LSM */

/* explicit GPULSMOpt(const DictionaryConfig&)  GPULSMOpt.cuh:5606 */
gd_status gd_lsm_create(const gd_config *config, gd_lsm **out);
/* ~GPULSMOpt()  GPULSMOpt.cuh:5723 (waits for outstanding work) */
gd_status gd_lsm_destroy(gd_lsm *h);
/* clear(stream)  GPULSMOpt.cuh:5731 */
gd_status gd_lsm_clear(gd_lsm *h, hipStream_t stream);
/* bulk_build(keys, values, count, stream)  GPULSMOpt.cuh:5738.  The last
 * occurrence of a key wins.  Synchronises the stream (as the reference does
 * for the unique count, GPULSMOpt.cuh:5776-5779). */
gd_status gd_lsm_bulk_build(gd_lsm *h, const uint32_t *keys,
                            const uint32_t *values, uint64_t count,
                            hipStream_t stream);
/* insert(const DeviceKeyValueBatch&, stream)  GPULSMOpt.cuh:5836 */
gd_status gd_lsm_insert(gd_lsm *h, const uint32_t *keys,
                        const uint32_t *values, uint64_t count,
                        hipStream_t stream);
/* erase(const DeviceKeyBatch&, stream)  GPULSMOpt.cuh:5843 */
gd_status gd_lsm_erase(gd_lsm *h, const uint32_t *keys, uint64_t count,
                       hipStream_t stream);
/* No reference counterpart: one batch mixing inserts and erases
 * (is_erase[i] != 0 makes entry i a tombstone); the age of an entry is its
 * position, exactly as if insert/erase had been one call. */
gd_status gd_lsm_apply_mixed(gd_lsm *h, const uint32_t *keys,
                             const uint32_t *values, const uint8_t *is_erase,
                             uint64_t count, hipStream_t stream);
/* lookup(const DeviceLookupBatch&, stream, quotients_grouped)
 * GPULSMOpt.cuh:5850.  out_found may be NULL: a miss then writes
 * 0xFFFFFFFF, otherwise value 0 and found 0.  quotients_grouped is a hint. */
gd_status gd_lsm_lookup(gd_lsm *h, const uint32_t *queries, uint64_t count,
                        uint32_t *out_values, uint8_t *out_found,
                        int quotients_grouped, hipStream_t stream);
/* No reference counterpart: gd_lsm_lookup that allocates nothing, for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and results as gd_lsm_lookup; the buffers of its grouped path
 * (query partition and grouped results) come from a library-owned device
 * array of 17 MiB per GPU, held until the process ends.  At most 2^20
 * queries per call (GD_INVALID_ARGUMENT above that, nothing launched).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (other streams, other handles) race, so the caller orders them. */
gd_status gd_lsm_lookup_noalloc(gd_lsm *h, const uint32_t *queries,
                                uint64_t count, uint32_t *out_values,
                                uint8_t *out_found, hipStream_t stream);
/* range(const DeviceRangeOutputBatch&, stream)  GPULSMOpt.cuh:5858:
 * out_sums[i] = sum of live values with lo[i] <= key <= hi[i], accumulated in
 * 64 bits and truncated to 32; lo > hi gives 0. */
gd_status gd_lsm_range_sum(gd_lsm *h, const uint32_t *lo, const uint32_t *hi,
                           uint64_t count, uint32_t *out_sums,
                           hipStream_t stream);
/* No GPULSMOpt counterpart (LSMu's count, FliX/impl_lsm_tree.cuh:1208):
 * number of live keys in [lo, hi], truncated to 32 bits. */
gd_status gd_lsm_range_count(gd_lsm *h, const uint32_t *lo,
                             const uint32_t *hi, uint64_t count,
                             uint32_t *out_counts, hipStream_t stream);
/* successor(const DeviceSuccessorBatch&, stream)  GPULSMOpt.cuh:5865:
 * smallest live key >= query, 0xFFFFFFFF when none. */
gd_status gd_lsm_successor(gd_lsm *h, const uint32_t *queries, uint64_t count,
                           uint32_t *out_keys, hipStream_t stream);
/* No reference counterpart: writes the live contents in ascending key order
 * (device buffers of `capacity` entries) and the number of live keys to
 * *out_count (host).  Synchronises the stream.  GD_LENGTH when capacity is
 * too small (out_count still receives the size). */
gd_status gd_lsm_dump_sorted(gd_lsm *h, uint32_t *out_keys,
                             uint32_t *out_values, uint64_t capacity,
                             uint64_t *out_count, hipStream_t stream);
/* No reference counterpart (LSMu's cleanup, FliX/impl_lsm_tree.cuh:1256):
 * merges every run into one, dropping tombstones and shadowed versions. */
gd_status gd_lsm_compact(gd_lsm *h, hipStream_t stream);
/* No reference counterpart (the reference has no on-disk format): writes
 * the live contents to `path` (header with key width, count and digests,
 * then the ascending keys and their values) and *out_count (nullable) =
 * live keys.  restore replaces the contents with a snapshot's (a bulk
 * build); GD_INVALID_ARGUMENT for a missing, foreign, truncated or
 * corrupted file or another key width, GD_OUT_OF_MEMORY beyond capacity.
 * Both synchronise the stream. */
gd_status gd_lsm_snapshot_save(gd_lsm *h, const char *path,
                               uint64_t *out_count, hipStream_t stream);
gd_status gd_lsm_snapshot_restore(gd_lsm *h, const char *path,
                                  uint64_t *out_count, hipStream_t stream);
/* canonical_carry_status()  GPULSMOpt.cuh:5872 (synchronous). */
gd_status gd_lsm_carry_status(gd_lsm *h, uint32_t *out_status);
/* gpu_resident_bytes()  GPULSMOpt.cuh:6341 */
gd_status gd_lsm_resident_bytes(gd_lsm *h, uint64_t *out_bytes);
/* No reference counterpart: the handle's device memory.  held = bytes its
 * pools hold from HIP (what gd_lsm_resident_bytes reports); cached = blocks
 * released by the structure and kept for reuse; live = held - cached; peak =
 * the highest held so far.  A pool keeps at most max(4 GiB, live / 2)
 * cached and returns the largest blocks above max(4 GiB, live / 4) to HIP
 * (hipFree waits for the device).  row_bytes_ub = the rows the runs hold
 * (an upper bound; tombstones included) times the row width. */
typedef struct gd_memory_stats {
  uint64_t held_bytes;
  uint64_t live_bytes;
  uint64_t cached_bytes;
  uint64_t peak_held_bytes;
  uint64_t trimmed_blocks;
  uint64_t row_bytes_ub;
} gd_memory_stats;
gd_status gd_lsm_memory(gd_lsm *h, gd_memory_stats *out);
/* No reference counterpart: returns every cached block to HIP (waits for the
 * handle's outstanding work). */
gd_status gd_lsm_trim(gd_lsm *h);
/* No reference counterpart: launch-shape knobs picked by the autotuner
 * (gpu-lsm-tree_amd/autotune.py).  None changes any result: they select
 * directory density, merge tile, sort block shape and when to group
 * lookups.  Zero fields keep the defaults. */
typedef struct gd_lsm_tuning {
  uint32_t dir_rows_log;           /* ~2^k rows per directory bucket (3)   */
  uint32_t merge_tile_log;         /* merge tile 2^k rows: 11 or 12 (12)   */
  uint32_t sort_block;             /* onesweep block: 256, 512 or 1024
                                      threads (by size)                    */
  uint32_t grouped_lookup_min_log; /* group lookups from 2^k queries (18);
                                      63 = never                           */
  uint32_t small_batch_max;        /* fused one-workgroup admission up to
                                      this many keys, <= 4096 (4096)       */
  uint32_t max_runs;               /* runs kept before the smallest adjacent
                                      pair merges, 2..47 (10)              */
  uint32_t segmented_admission;    /* 32-bit batches above 16384 keys: 1 =
                                      partition + per-segment LDS sort (1),
                                      2 = radix passes + finalize; 0 = keep */
  uint32_t merge_block;            /* merge workgroup: 256 or 512 threads
                                      per tile (512)                       */
  uint32_t merge_ways;             /* most runs one merge pass takes: 2 =
                                      pairwise only, up to 8 (8)           */
  uint32_t merge_fanout;           /* newest runs of one size that merge: 2
                                      = binary counter, up to merge_ways (2) */
} gd_lsm_tuning;

gd_status gd_lsm_set_tuning(gd_lsm *h, const gd_lsm_tuning *tuning);
gd_status gd_lsm_get_tuning(gd_lsm *h, gd_lsm_tuning *out);

/* No reference counterpart: opt-in per-kernel timing (HIP events on the
 * launch stream) of the merge, lookup and range kernels, with the merges'
 * row counts for their algorithmic bytes (rows_in + rows_out) * 8, and of
 * the admission of every update tile (all of its kernels, from the first
 * to the last). */
typedef struct gd_lsm_profile {
  uint64_t merge_launches;
  double merge_ms;
  uint64_t merge_rows_in;   /* source rows of all merges  */
  uint64_t merge_rows_out;  /* surviving rows             */
  uint64_t lookup_launches;
  double lookup_ms;
  uint64_t lookup_queries;
  uint64_t admission_batches;  /* admitted tiles          */
  double admission_ms;
  uint64_t admission_keys;     /* keys of those tiles     */
  uint64_t range_launches;     /* range_kernel launches   */
  double range_ms;
  uint64_t range_queries;
  uint64_t range_runs;         /* runs searched, summed over the launches */
} gd_lsm_profile;

gd_status gd_lsm_set_profiling(gd_lsm *h, int enable);
/* Synchronous. */
gd_status gd_lsm_read_profile(gd_lsm *h, gd_lsm_profile *out, int reset);

/* No reference counterpart: number of sorted runs currently searched by
 * reads and the rows they hold (tombstones and shadowed versions included).
 * Synchronous. */
gd_status gd_lsm_shape(gd_lsm *h, uint32_t *out_runs, uint64_t *out_rows_ub);

/* ---------------------------------------------------------- LSM, 64-bit */

/* No reference counterpart: the reference LSM is uint32-only
 * (FliX/impl_gpulsmopt.cuh:145-148, GPULSMOpt.cuh's 16/16-bit row split).
 * The same dictionary over uint64 keys and uint64 values, with the gd_lsm_*
 * semantics: newest update wins, tombstones mask older runs, a lookup miss
 * writes all-ones (value 0 + found 0 with out_found), range sums wrap
 * modulo 2^64, range counts are uint64, successor returns all-ones when no
 * live key is >= the query.  Capacity and tiling are those of gd_config
 * (at most 2^32 - 1 rows). */
typedef struct gd_lsm64 gd_lsm64;

gd_status gd_lsm64_create(const gd_config *config, gd_lsm64 **out);
gd_status gd_lsm64_destroy(gd_lsm64 *h);
gd_status gd_lsm64_clear(gd_lsm64 *h, hipStream_t stream);
gd_status gd_lsm64_bulk_build(gd_lsm64 *h, const uint64_t *keys,
                              const uint64_t *values, uint64_t count,
                              hipStream_t stream);
gd_status gd_lsm64_insert(gd_lsm64 *h, const uint64_t *keys,
                          const uint64_t *values, uint64_t count,
                          hipStream_t stream);
gd_status gd_lsm64_erase(gd_lsm64 *h, const uint64_t *keys, uint64_t count,
                         hipStream_t stream);
gd_status gd_lsm64_apply_mixed(gd_lsm64 *h, const uint64_t *keys,
                               const uint64_t *values,
                               const uint8_t *is_erase, uint64_t count,
                               hipStream_t stream);
gd_status gd_lsm64_lookup(gd_lsm64 *h, const uint64_t *queries,
                          uint64_t count, uint64_t *out_values,
                          uint8_t *out_found, hipStream_t stream);
gd_status gd_lsm64_range_sum(gd_lsm64 *h, const uint64_t *lo,
                             const uint64_t *hi, uint64_t count,
                             uint64_t *out_sums, hipStream_t stream);
gd_status gd_lsm64_range_count(gd_lsm64 *h, const uint64_t *lo,
                               const uint64_t *hi, uint64_t count,
                               uint64_t *out_counts, hipStream_t stream);
gd_status gd_lsm64_successor(gd_lsm64 *h, const uint64_t *queries,
                             uint64_t count, uint64_t *out_keys,
                             hipStream_t stream);
gd_status gd_lsm64_dump_sorted(gd_lsm64 *h, uint64_t *out_keys,
                               uint64_t *out_values, uint64_t capacity,
                               uint64_t *out_count, hipStream_t stream);
gd_status gd_lsm64_compact(gd_lsm64 *h, hipStream_t stream);
gd_status gd_lsm64_carry_status(gd_lsm64 *h, uint32_t *out_status);
gd_status gd_lsm64_resident_bytes(gd_lsm64 *h, uint64_t *out_bytes);
gd_status gd_lsm64_memory(gd_lsm64 *h, gd_memory_stats *out);
gd_status gd_lsm64_trim(gd_lsm64 *h);
gd_status gd_lsm64_set_tuning(gd_lsm64 *h, const gd_lsm_tuning *tuning);
gd_status gd_lsm64_get_tuning(gd_lsm64 *h, gd_lsm_tuning *out);
gd_status gd_lsm64_shape(gd_lsm64 *h, uint32_t *out_runs,
                         uint64_t *out_rows_ub);
gd_status gd_lsm64_snapshot_save(gd_lsm64 *h, const char *path,
                                 uint64_t *out_count, hipStream_t stream);
gd_status gd_lsm64_snapshot_restore(gd_lsm64 *h, const char *path,
                                    uint64_t *out_count, hipStream_t stream);

/* -------------------------------------------------------------- serving */

/* No reference counterpart: a continuous-batching engine over one LSM
 * dictionary.  Requests of any size are submitted from any thread and
 * admitted into steps (arrival order, HBM budget); a step applies all its
 * updates as one batch, then answers all its lookups and ranges with one
 * launch each, while the next step's updates overlap on a second stream.
 * Results equal a sequential replay in step order (updates before queries
 * within a step).  Request buffers are device memory owned by the caller:
 * inputs must be fully written before gd_serve_submit (the engine's streams
 * do not wait for the caller's streams), and every buffer must stay valid
 * until the request completes; outputs are valid once gd_serve_wait returns. */
typedef struct gd_server gd_server;

typedef enum gd_op {
  GD_OP_INSERT = 0,      /* a = keys, b = values                         */
  GD_OP_ERASE = 1,       /* a = keys                                     */
  GD_OP_LOOKUP = 2,      /* a = queries, out = values, out_found optional */
  GD_OP_RANGE_COUNT = 3, /* a = lo, b = hi, out = counts                 */
  GD_OP_RANGE_SUM = 4    /* a = lo, b = hi, out = sums                   */
} gd_op;

typedef struct gd_serve_config {
  gd_config lsm;               /* the dictionary                           */
  uint64_t hbm_budget_bytes;   /* dictionary + in-flight staging cap; 0 = none */
  uint64_t max_step_keys;      /* keys coalesced per step; 0 = 2^22 / 2^24 */
  uint32_t max_inflight_steps; /* pipelined steps; 0 = 2                   */
  uint32_t key_bytes;          /* 4 (0) or 8: uint32 or uint64 keys and
                                  values (the gd_lsm / gd_lsm64 dictionary) */
  const gd_lsm_tuning *tuning; /* nullable                                 */
  uint32_t step_launch;         /* 0 = every step kernel by kernel
                                  (default); 1 = a small step (updates
                                  that fit one fused admission, lookups
                                  below the grouped threshold, up to 64
                                  requests of each kind) is a descriptor
                                  copy, the fused admission reading the
                                  requests in place and one persistent
                                  query kernel; 2 = the same as one
                                  captured graph launch per step          */
} gd_serve_config;

typedef struct gd_serve_stats {
  uint64_t requests;           /* completed since the last reset           */
  uint64_t steps;
  uint64_t coalesced_updates;  /* update requests merged into steps        */
  uint64_t coalesced_queries;
  double latency_p50_us;       /* completion - submission                  */
  double latency_p99_us;
  double latency_p999_us;
  double latency_max_us;
  uint64_t peak_staged_bytes;
  uint64_t resident_bytes;     /* held by the dictionary and step pools,
                                  as of the last completed step           */
  uint64_t graph_steps;        /* small steps run fused (step_launch 1, 2) */
  uint64_t live_bytes;         /* resident_bytes minus cached free blocks */
  uint64_t peak_resident_bytes;/* highest resident_bytes seen (not reset) */
} gd_serve_stats;

gd_status gd_serve_create(const gd_serve_config *config, gd_server **out);
/* Drains every submitted request, then releases the engine. */
gd_status gd_serve_destroy(gd_server *h);
/* GD_CAPACITY when one request alone exceeds the HBM budget. */
gd_status gd_serve_submit(gd_server *h, gd_op op, const void *a,
                          const void *b, uint64_t count, void *out,
                          uint8_t *out_found, uint64_t *out_ticket);
/* Blocks until the request completes; returns its status (an update in a
 * step whose publication failed reports GD_PUBLICATION_FAILED); out_step
 * (nullable) receives the step that served it. */
gd_status gd_serve_wait(gd_server *h, uint64_t ticket, uint64_t *out_step);
/* Blocks until every submitted request has completed. */
gd_status gd_serve_drain(gd_server *h);
gd_status gd_serve_get_stats(gd_server *h, gd_serve_stats *out, int reset);

/* ----------------------------------------------------------------- FliX */

typedef struct gd_flix gd_flix;

typedef struct gd_flix_config {
  uint32_t key_bytes;        /* 4 or 8 (key32 / key64)                     */
  uint32_t value_bytes;      /* 4 (smallsize offsets, as the reference) or 8 */
  uint32_t node_slots;       /* pairs per node; 0 -> 32 (NODESIZE=5)       */
  uint32_t fill_percent;     /* initial fill; 0 -> 50                      */
  uint32_t rebuild_chain_nodes; /* rebuild policy: an insert that leaves a
                                   bucket chain longer than this many nodes
                                   rebuilds the bucket layer after it (the
                                   reference's periodic rebuild,
                                   FliX/benchmarks_updates.cuh:1919-1935);
                                   0 = only on demand                   */
} gd_flix_config;

/* Reserved keys: 0 (empty slot) and 1 (tombstone) (FliX/definitions.cuh:83-103,
 * FliX/definitions_updates.cuh:259).  Usable keys are [2, 2^bits - 4]. */

/* constructor + build_bucket_layer_only(keys, size, max_size, ...)
 * FliX/impl_cg_rtx_index_updates.cuh:1711.  Values are the input positions
 * (init_offsets); keys need not be sorted or unique (first occurrence wins).
 * max_size bounds the number of keys the index may hold; 0 -> 2*size. */
gd_status gd_flix_create(const gd_flix_config *config, gd_flix **out);
gd_status gd_flix_destroy(gd_flix *h);
gd_status gd_flix_build(gd_flix *h, const void *keys, uint64_t count,
                        uint64_t max_size, hipStream_t stream);
/* insert(sorted_keys, offsets, n, stream)  impl:2195.  Insert-if-absent;
 * input need not be sorted.  A batch is applied whole or not at all: when
 * the allocation region cannot hold its new nodes the index is first
 * rebuilt with a larger region; GD_CAPACITY (or GD_OUT_OF_MEMORY when the
 * larger region cannot be allocated) means no key of the batch was
 * inserted.  Synchronises the stream (the batch's new-key count and the
 * node plan are read back). */
gd_status gd_flix_insert(gd_flix *h, const void *keys, const void *values,
                         uint64_t count, hipStream_t stream);
/* remove(sorted_keys, n, stream)  impl:2253, for keys in any order (they
 * are sorted first); duplicates and absent keys are ignored. */
gd_status gd_flix_remove(gd_flix *h, const void *keys, uint64_t count,
                         hipStream_t stream);
/* remove(sorted_keys, n, stream)  impl:2253 with the reference's contract:
 * the keys are ascending (duplicates allowed) and are not sorted again.
 * Unsorted keys give undefined results, as in the reference. */
gd_status gd_flix_remove_sorted(gd_flix *h, const void *sorted_keys,
                                uint64_t count, hipStream_t stream);
/* lookups_ordered(sorted_keys, results, n, stream)  impl:2485: queries must
 * be ascending; results are in query order; miss -> all-ones value. */
gd_status gd_flix_lookup_sorted(gd_flix *h, const void *sorted_queries,
                                uint64_t count, void *out_values,
                                hipStream_t stream);
/* No reference counterpart: lookups in any order (sorted internally). */
gd_status gd_flix_lookup(gd_flix *h, const void *queries, uint64_t count,
                         void *out_values, hipStream_t stream);
/* lookups_successor(sorted_keys, out_keys, n, stream)  impl:2647: ceiling
 * key, key_type(0xFFFFFFFF) when none.  Queries must be ascending. */
gd_status gd_flix_successor_sorted(gd_flix *h, const void *sorted_queries,
                                   uint64_t count, void *out_keys,
                                   hipStream_t stream);
/* rebuild_bucket_layer_only(&ms, &bytes)  impl:1821 */
gd_status gd_flix_rebuild(gd_flix *h, hipStream_t stream);
/* No reference counterpart (range_lookup_sum is a TODO, impl:2640-2644):
 * pass 1 (out_keys == NULL) writes per-query counts to out_offsets[0..n)
 * and out_offsets[n] = total after an exclusive scan; pass 2 fills pairs of
 * query i at out_keys/out_values[out_offsets[i] ...]. */
gd_status gd_flix_range(gd_flix *h, const void *lo, const void *hi,
                        uint64_t count, uint64_t *out_offsets,
                        void *out_keys, void *out_values, hipStream_t stream);
/* compute_total_size(stream)  impl:1657 (synchronous) */
gd_status gd_flix_total_keys(gd_flix *h, uint64_t *out_keys,
                             hipStream_t stream);
/* gpu_resident_bytes()  impl:1600 */
gd_status gd_flix_resident_bytes(gd_flix *h, uint64_t *out_bytes);
/* No reference counterpart: as gd_lsm_memory; row_bytes_ub = the bucket
 * layer's node bytes in use. */
gd_status gd_flix_memory(gd_flix *h, gd_memory_stats *out);
/* No reference counterpart: as gd_lsm_trim. */
gd_status gd_flix_trim(gd_flix *h);

#ifdef __cplusplus
}
#endif

#endif /* GDICT_H */
