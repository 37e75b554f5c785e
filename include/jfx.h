/*
 * jfx.h — C ABI of libjfx.so, the MI355X (gfx950) sort-centric algorithms
 * layer behind jaxframes' sort_values / groupby().agg / merge / collect.
 *
 * The reference (jaxframes, pure Python on JAX) has no FFI: its "operator
 * API" is the module surface of src/jaxframes/distributed/parallel_algorithms.py.
 * Each entry point below names the reference function it replaces.
 *
 * Conventions
 *  - Every function returns 0 (JFX_OK) or a JFX_ERR_* code; jfx_last_error()
 *    returns a thread-local message for the last failure on this thread.
 *  - All data pointers are device pointers owned by the caller. The library
 *    never allocates device memory: scratch lives in a caller-provided
 *    workspace whose size comes from the matching *_workspace_size call
 *    (jfx_groupby_noalloc alone uses a library-owned static device array).
 *  - `stream` is a hipStream_t (NULL = default stream). Calls enqueue work on
 *    it. Calls whose output size is data dependent (they write the size to a
 *    host pointer) synchronise that stream once, as the reference does with
 *    int(np.asarray(...)) at parallel_algorithms.py:286, 712, 772.
 *  - Device-side safety: every cross-workgroup wait is bounded; a wait that
 *    runs out raises a device error word, reported by jfx_device_error().
 *  - Calls on different streams may overlap if they use different
 *    workspaces; calls sharing a workspace must be ordered.
 */
#ifndef JFX_H
#define JFX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define JFX_OK 0
#define JFX_ERR_INVALID 1      /* bad argument (null pointer, size, type) */
#define JFX_ERR_HIP 2          /* a HIP runtime call failed */
#define JFX_ERR_UNSUPPORTED 3  /* type/size combination not supported */
#define JFX_ERR_WORKSPACE 4    /* workspace too small */
#define JFX_ERR_DEVICE 5       /* a kernel raised the device error word */
#define JFX_ERR_CAPACITY 6     /* outputs too small for the groups the call may produce */

/* key / value type codes */
#define JFX_U32 0
#define JFX_I32 1
#define JFX_F32 2
#define JFX_U64 3
#define JFX_I64 4
#define JFX_F64 5

/* aggregation bit flags for jfx_groupby */
#define JFX_AGG_SUM 1
#define JFX_AGG_COUNT 2
#define JFX_AGG_MIN 4
#define JFX_AGG_MAX 8
#define JFX_AGG_MEAN 16

/* comparison codes for jfx_filter_rows */
#define JFX_CMP_GT 0
#define JFX_CMP_GE 1
#define JFX_CMP_LT 2
#define JFX_CMP_LE 3
#define JFX_CMP_EQ 4
#define JFX_CMP_NE 5

/* ---- library --------------------------------------------------------- */
const char* jfx_last_error(void);
int jfx_version(void);
/* Synchronises `stream`, reads and clears the device error word. */
int jfx_device_error(void* stream, uint32_t* out_code);

/* ---- radix sort ------------------------------------------------------- */
/* Workspace bytes for sorting n keys of key_type with pay_bytes (0, 4, 8)
 * of payload per key. */
int jfx_sort_workspace_size(int key_type, int pay_bytes, uint64_t n, uint64_t* out_bytes);

/* Stable LSD radix sort of keys with an optional payload.
 * Replaces ParallelRadixSort.sort (parallel_algorithms.py:463-581) and
 * parallel_sort (parallel_algorithms.py:1079-1114).
 *  - keys_in (n keys) is not modified; keys_out receives the sorted keys when
 *    want_keys != 0 (it is also used as a ping-pong buffer, so it must hold n
 *    keys and must not alias keys_in).
 *  - pay_bytes 0: no payload. 4 or 8: pay_in is permuted with the keys into
 *    pay_out; pay_bytes 4 with pay_in == NULL writes the uint32 row index
 *    (the reference's jnp.arange payload).
 *  - descending != 0 returns the reverse of the ascending order, so ties come
 *    out in reverse original order (parallel_algorithms.py:552-555).
 *  - digit_bits: 0 = autotuned, else 8 or 11.
 *  - Synchronises the stream once to read the digit plan. */
int jfx_sort(int key_type, const void* keys_in, void* keys_out, int want_keys,
             int pay_bytes, const void* pay_in, void* pay_out, uint64_t n,
             int descending, int digit_bits, void* workspace, uint64_t ws_bytes,
             void* stream);

/* ---- groupby (sort + head-flag segmented reduction) ------------------- */
/* Group outputs (out_keys, out_sum, out_count, out_min, out_max, out_mean)
 * are caller buffers with room for n groups; any of them may be NULL to skip
 * that aggregation. Types: out_keys has key_type; out_sum/min/max have
 * val_type; out_count is int64; out_mean is float32 for 4-byte value types
 * and float64 for 8-byte ones. Float sums are TwoSum-compensated in the value
 * type; integer sums wrap in the value type like the reference's
 * segment_sum. Groups come out in ascending key order. */

int jfx_segreduce_workspace_size(int val_type, uint64_t n, uint64_t* out_bytes);

/* Segmented reduction over already-sorted keys. Replaces the head-flag /
 * cumsum / segment_* steps of SortBasedGroupBy.groupby_aggregate
 * (parallel_algorithms.py:765-801) and _apply_aggregation_vectorized
 * (parallel_algorithms.py:626-660). Row i's value is vals[perm[i]] when perm
 * (uint32 row index, e.g. the sort's payload) is non-NULL, else vals[i].
 * When out_ngroups is non-NULL the stream is synchronised and G written. */
int jfx_segreduce(int key_type, const void* sorted_keys, int val_type, const void* vals,
                  const uint32_t* perm, uint64_t n, void* out_keys, void* out_sum,
                  int64_t* out_count, void* out_min, void* out_max, void* out_mean,
                  uint64_t* out_ngroups, void* workspace, uint64_t ws_bytes, void* stream);

int jfx_groupby_workspace_size(int key_type, int val_type, uint64_t n, uint64_t* out_bytes);

/* One-column groupby: radix sort of the keys carrying the values as payload,
 * then jfx_segreduce. Replaces groupby_aggregate (parallel_algorithms.py:
 * 1117-1148) for a single aggregated column. Synchronises once for the digit
 * plan and once to read G into *out_ngroups. Integer keys with at least 2^16
 * rows first try one pass (wave tables indexed by key mod 1024, retried once
 * at 2048 when the span seen allows): when max - min is below the table width
 * that pass gives the groups with one synchronisation, otherwise the
 * sort-based or small-range path runs as above (JFX_ONEPASS=0 skips the
 * one-pass try, JFX_ONEPASS=always makes it at any size). */
int jfx_groupby(int key_type, const void* keys, int val_type, const void* vals, uint64_t n,
                void* out_keys, void* out_sum, int64_t* out_count, void* out_min, void* out_max,
                void* out_mean, uint64_t* out_ngroups, int digit_bits, void* workspace,
                uint64_t ws_bytes, void* stream);

/* jfx_groupby with outputs of out_capacity entries (jfx_groupby: n). When
 * the groups may outnumber that -- on the one-pass path, more than the exact
 * group count; after the key statistics, more than min(n, max - min + 1) for
 * a small integer key range, else more than n -- nothing is written,
 * *out_ngroups is set to the capacity needed and JFX_ERR_CAPACITY is
 * returned; a caller that expects few groups passes small outputs and
 * retries with the capacity returned. */
int jfx_groupby_ex(int key_type, const void* keys, int val_type, const void* vals, uint64_t n,
                   void* out_keys, void* out_sum, int64_t* out_count, void* out_min,
                   void* out_max, void* out_mean, uint64_t* out_ngroups, uint64_t out_capacity,
                   int digit_bits, void* workspace, uint64_t ws_bytes, void* stream);

/* jfx_groupby_ex without a workspace, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). No counterpart in the reference. It runs only the
 * one-pass path of jfx_groupby_ex -- the same kernels with the same launch
 * shapes, so its outputs are bit-identical to jfx_groupby_ex's whenever that
 * call takes the one-pass path (n >= 2^16) -- and needs
 *  - integer keys (JFX_U32/I32/U64/I64), any value type, 1 <= n < 2^30;
 *  - a key span max - min below 2048 (below 1024 for float64 values with a
 *    min or max output); a wider span returns JFX_ERR_UNSUPPORTED with
 *    *out_ngroups = 0 and nothing written (use jfx_groupby_ex).
 * Scratch: a library-owned 27 MiB device array per GPU (held until the
 * process ends), shared by every call on that GPU, which keeps the last
 * call's contents: calls that overlap on one GPU (other streams or threads)
 * race and must be serialised by the caller. The call makes no hipMalloc /
 * hipFree / hipHostMalloc; it synchronises the stream once per one-pass
 * attempt (one or two) to read the result words with a stream-ordered
 * copy. Output and capacity semantics are jfx_groupby_ex's. */
int jfx_groupby_noalloc(int key_type, const void* keys, int val_type, const void* vals, uint64_t n,
                        void* out_keys, void* out_sum, int64_t* out_count, void* out_min,
                        void* out_max, void* out_mean, uint64_t* out_ngroups,
                        uint64_t out_capacity, void* stream);

/* ---- hashing and the local hash partition ----------------------------- */
/* Row hash of 1..8 key columns (device pointers, one per column, n rows
 * each). Replaces hash_multi_columns (parallel_algorithms.py:69-136): one key
 * -> fmix32(uint32(key)); K keys -> seed 0x1337, per key i
 * h ^= uint32(k_i)*(0x9E3779B1+i), h = rotl(h,13), h *= 0x85EBCA6B, then
 * fmix32. uint32(key) is the low 32 bits of an integer, the bits of a
 * float32, and the bits of a float64 rounded to float32. */
int jfx_hash_rows(const void* const* keys, const int* key_types, int nkeys, uint64_t n,
                  uint32_t* out_hashes, void* stream);

int jfx_hash_partition_workspace_size(int key_type, uint64_t n, uint32_t nbuckets,
                                      uint64_t* out_bytes);

/* Local hash partition into nbuckets (1..2048): bucket = hash % nbuckets with
 * hash = fmix32(uint32(key)), or hash = key when prehashed (key_type must be
 * JFX_U32, e.g. jfx_hash_rows output). Rows keep their input order inside a
 * bucket. Outputs (any of out_keys / out_rows / out_hashes may be NULL):
 * partitioned keys, the uint32 source row of each output row, out_offsets
 * (int64, nbuckets+1) and each input row's hash. Replaces the per-device step
 * of collective_hash_repartition (hash at :214, _bucketize_local_column at
 * :163-182) with a compact layout instead of the padded [P, local_n] one. */
int jfx_hash_partition(int key_type, const void* keys, uint64_t n, uint32_t nbuckets,
                       int prehashed, void* out_keys, uint32_t* out_rows, int64_t* out_offsets,
                       uint32_t* out_hashes, void* workspace, uint64_t ws_bytes, void* stream);

/* ---- sort-merge join -------------------------------------------------- */
/* Both sides are first sorted with jfx_sort carrying the uint32 row index
 * (ParallelSortMergeJoin.join step 1, parallel_algorithms.py:969-973). The
 * merge then emits every (left, right) pair with equal keys, in (key, left
 * sorted position, right sorted position) order — the order of
 * jnp.where(sorted_left[:, None] == sorted_right[None, :])
 * (parallel_algorithms.py:987-993) — without the L x R mask. NaN keys match
 * nothing (NaN != NaN); -0.0 matches +0.0. Each side must have < 2^32 rows. */
int jfx_join_workspace_size(uint64_t n_left, uint64_t* out_bytes);

/* Counts the pairs (per 2048-row left tile, scanned) and writes the total to
 * *out_pairs; synchronises the stream once. JFX_ERR_UNSUPPORTED when the
 * join has 2^32 pairs or more. The fills that follow write each tile's pairs
 * in chunks of at most 8192, one workgroup each, so a key hot on both sides
 * is spread over many workgroups. */
int jfx_join_count(int key_type, const void* sorted_left, uint64_t n_left, const void* sorted_right,
                   uint64_t n_right, uint64_t* out_pairs, void* workspace, uint64_t ws_bytes,
                   void* stream);

/* Writes the pairs counted by the preceding jfx_join_count on the same
 * workspace: out_left_rows / out_right_rows get left_rows[i] / right_rows[j]
 * of each pair (source row numbers), out_keys (optional) the key. Output
 * buffers hold *out_pairs entries. */
int jfx_join_fill(int key_type, const void* sorted_left, const uint32_t* left_rows, uint64_t n_left,
                  const void* sorted_right, const uint32_t* right_rows, uint64_t n_right,
                  uint32_t* out_left_rows, uint32_t* out_right_rows, void* out_keys,
                  void* workspace, uint64_t ws_bytes, void* stream);

/* jfx_join_fill with one payload column per side written by the fill: the
 * value gather of parallel_algorithms.py:1052-1073 fused into the pair
 * enumeration. out_left_values[k] = left_values[left_rows[i]] when
 * left_by_row, else left_values[i] (i the pair's sorted-left position: a
 * payload already in key order is read nearly sequentially); the right side
 * likewise. Payload elements are 4 or 8 bytes. Every output is optional;
 * left_rows / right_rows may be null when no output needs source rows (a
 * null left_rows reports sorted positions in out_left_rows). */
int jfx_join_fill_values(int key_type, const void* sorted_left, const uint32_t* left_rows,
                         uint64_t n_left, const void* sorted_right, const uint32_t* right_rows,
                         uint64_t n_right, int left_elem_bytes, const void* left_values,
                         int left_by_row, int right_elem_bytes, const void* right_values,
                         int right_by_row, uint32_t* out_left_rows, uint32_t* out_right_rows,
                         void* out_keys, void* out_left_values, void* out_right_values,
                         void* workspace, uint64_t ws_bytes, void* stream);

/* One side of jfx_join_fill_columns: what the fill writes for that side of
 * every pair. `rows` maps a sorted position to its source row (NULL: the
 * position itself); `out_rows` (optional) receives each pair's source row.
 * ncols payload columns are written to out_cols[j] (M entries, col_bytes[j]
 * = 4 or 8 bytes each), read either
 *  - from cols[j] at the source row (by_row = 1) or at the sorted position
 *    (by_row = 0), or
 *  - when record_bytes is 4 or 8, as the field at byte col_offset[j] of the
 *    record at the sorted position in `records`: a payload the sort carried
 *    in key order (jfx_pack_columns packs up to 8 bytes of columns into one
 *    record per row), read sequentially instead of gathered. */
#define JFX_JOIN_MAXCOLS 8
typedef struct jfx_join_side {
  const uint32_t* rows;
  uint32_t* out_rows;
  int32_t ncols;
  int32_t by_row;
  int32_t record_bytes;
  const void* records;
  const void* cols[JFX_JOIN_MAXCOLS];
  int32_t col_bytes[JFX_JOIN_MAXCOLS];
  int32_t col_offset[JFX_JOIN_MAXCOLS];
  void* out_cols[JFX_JOIN_MAXCOLS];
} jfx_join_side;

/* jfx_join_fill with any number of payload columns per side (up to 8), all
 * written by the fill: the pair enumeration of ParallelSortMergeJoin.join
 * (parallel_algorithms.py:987-994) with the payload gathers of :1052-1073
 * fused in, so no row columns need to be written for later gathers. left /
 * right may be NULL (nothing written for that side). Same workspace and
 * preceding jfx_join_count as jfx_join_fill. */
int jfx_join_fill_columns(int key_type, const void* sorted_left, uint64_t n_left,
                          const jfx_join_side* left, const void* sorted_right, uint64_t n_right,
                          const jfx_join_side* right, void* out_keys, void* workspace,
                          uint64_t ws_bytes, void* stream);

/* Packs ncols columns of n rows (4 or 8 bytes each, at most 8 bytes in all)
 * into one record per row, column j at the byte offset of the columns before
 * it; *out_record_bytes (optional) receives the record width. A join side's
 * payload packed this way rides through its radix sort as one 4- or 8-byte
 * payload (jfx_sort), so the fill reads it in key order: the column gather
 * of parallel_algorithms.py:1052-1073 moved into the sort. */
int jfx_pack_columns(int ncols, const void* const* cols, const int* col_bytes, uint64_t n,
                     void* out, int* out_record_bytes, void* stream);

/* Inner-join range pre-filter of one side, run before that side's sort. The
 * reference pairs only equal keys (parallel_algorithms.py:989-994), so a row
 * whose key lies outside [sorted_other[0], sorted_other[n_other - 1]] (the
 * other side sorted ascending) has no pair, and an inner join is unchanged
 * without it. Keeps the other rows in row order: their keys to out_keys, their
 * payload columns packed into one record per row (as jfx_pack_columns) to
 * out_records when ncols > 0, their row indices to out_rows when non-null,
 * and the kept count to *out_count (synchronises the stream once). Integer key
 * types only (JFX_ERR_UNSUPPORTED for floats); outputs hold n rows;
 * workspace of jfx_join_prefilter_workspace_size(n) bytes. Below 2^30 rows
 * one pass (a decoupled lookback over the tiles' kept counts) reads the keys
 * once; above it (or with JFX_PREFILTER=chunked) chunk counts, their scan and
 * a scatter read them twice. */
int jfx_join_prefilter_workspace_size(uint64_t n, uint64_t* out_bytes);
int jfx_join_prefilter(int key_type, const void* keys, uint64_t n, const void* sorted_other,
                       uint64_t n_other, int ncols, const void* const* cols, const int* col_bytes,
                       void* out_keys, void* out_records, uint32_t* out_rows, uint64_t* out_count,
                       void* workspace, uint64_t ws_bytes, void* stream);

/* For each sorted probe key, build_rows[first index of an equal key in
 * sorted_build], or 0xFFFFFFFF when none: the searchsorted + equality test
 * of the left / right / outer branches (parallel_algorithms.py:1014-1047). */
int jfx_join_first_match(int key_type, const void* sorted_probe, uint64_t n_probe,
                         const void* sorted_build, const uint32_t* build_rows, uint64_t n_build,
                         uint32_t* out_rows, void* stream);

/* dst[i] = src[rows[i]], or fill_bits when rows[i] == 0xFFFFFFFF, for 4- or
 * 8-byte elements: the payload gather of parallel_algorithms.py:1051-1073
 * (jnp.where(idx >= 0, col[idx], nan)). */
int jfx_gather(int elem_bytes, const void* src, const uint32_t* rows, uint64_t n, void* dst,
               uint64_t fill_bits, void* stream);

/* ---- batched groupby ---------------------------------------------------- */
/* Many independent single-key groupbys of uneven row counts ("frames"),
 * concatenated row-wise with frame_offsets (device int64[nframes+1], 0 .. n),
 * in one device pipeline with no host synchronisation (capturable in a HIP
 * graph): composite key (frame << key_bits) | (ordered(key) - ordered(key_min))
 * per row, one stable radix sort of the composites over exactly
 * ceil(log2 nframes) + key_bits bits, one head-flag segmented reduction.
 * Each frame's groups equal SortBasedGroupBy.groupby_aggregate
 * (parallel_algorithms.py:733-803) of that frame. Groups come out
 * frame-major, keys ascending; out_keys receives the composite keys (uint32
 * when the bit total is <= 32, else uint64); *out_ngroups (a DEVICE word) the
 * total group count. Keys must lie in [key_min, key_min + 2^key_bits)
 * (checked; a violation raises the device error word). Output buffers hold n
 * groups. No counterpart in the reference, which runs one groupby per frame. */
int jfx_batched_groupby_workspace_size(uint32_t total_bits, int val_type, uint64_t n,
                                       uint64_t* out_bytes);
/* sort_shape picks the onesweep tile shape of the composite sort (waves x
 * keys per lane, for 32-bit / 64-bit composites): 0 = 16 x 8 / 8 x 16 and 1 =
 * 8 x 16 / 16 x 8 (8192-key tiles), 2 = 8 x 8 / 4 x 16 and 3 = 4 x 16 / 8 x 8
 * (4096-key tiles); with digit_bits (8 or 11) it is what the autotuner of
 * jaxframes_amd.batched chooses per size class. */
int jfx_batched_groupby(int key_type, const void* keys, int val_type, const void* vals, uint64_t n,
                        const int64_t* frame_offsets, uint32_t nframes, uint64_t key_min,
                        uint32_t key_bits, int digit_bits, int sort_shape, void* out_keys,
                        void* out_sum, int64_t* out_count, void* out_min, void* out_max,
                        void* out_mean, uint32_t* out_ngroups, void* workspace, uint64_t ws_bytes,
                        void* stream);

/* Routed batched groupby: frames whose keys span fewer than
 * JFX_BATCH_NARROW_SPAN values skip the composite sort. Their rows are
 * reduced in one-pass LDS tables indexed by ordered(key) - ordered(frame
 * minimum), chunk by chunk, and the chunk tables merged per frame in chunk
 * order; only the other ("wide") frames' rows are composed, compacted, sorted
 * and reduced as in jfx_batched_groupby. Groups of both kinds are then
 * placed frame-major, as jfx_batched_groupby writes them (same outputs, same
 * composite keys, *out_ngroups a device word). Each frame's groups equal
 * SortBasedGroupBy.groupby_aggregate (parallel_algorithms.py:733-803) of that
 * frame; no counterpart in the reference.
 *
 * The split is fixed when the plan is made. jfx_batched_key_ranges reads the
 * keys once and writes the min and max ordered key image of each row chunk
 * (chunk c: rows [chunk_row0[c], chunk_row0[c] + chunk_rows[c]); device
 * arrays; out_minmax: device uint64[2 * nchunks]), from which the caller
 * builds a jfx_batched_routing of device arrays. A key outside its narrow
 * frame's measured range, or outside [key_min, key_min + 2^key_bits), raises
 * the device error word. float64 values are not routed (JFX_ERR_UNSUPPORTED). */
#define JFX_BATCH_NARROW_SPAN 1024
#define JFX_BATCH_CHUNK_ROWS 65536
int jfx_batched_key_ranges(int key_type, const void* keys, const uint64_t* chunk_row0,
                           const uint32_t* chunk_rows, uint32_t nchunks, uint64_t* out_minmax,
                           void* stream);

typedef struct jfx_batched_routing {
  uint32_t n_narrow;             /* narrow frames */
  uint32_t n_chunks;             /* row chunks of the narrow frames, frame by frame */
  uint32_t n_wide;               /* the other frames (empty frames among them) */
  uint32_t reserved;
  uint64_t n_wide_rows;          /* rows of the wide frames */
  uint64_t part_entries;         /* sum over chunks of their frame's span + 1 */
  uint64_t merged_entries;       /* sum over narrow frames of span + 1 */
  const uint32_t* narrow_frame;  /* [n_narrow] frame ids, ascending */
  const uint64_t* narrow_omin;   /* [n_narrow] ordered image of the frame's minimum key */
  const uint32_t* narrow_span;   /* [n_narrow] its max - min (ordered images), below 1024 */
  const uint32_t* narrow_chunk0; /* [n_narrow + 1] first chunk of each narrow frame */
  const uint64_t* narrow_mbase;  /* [n_narrow] first entry of its merged table */
  const uint64_t* chunk_row0;    /* [n_chunks] first row */
  const uint32_t* chunk_rows;    /* [n_chunks] rows, at most JFX_BATCH_CHUNK_ROWS */
  const uint32_t* chunk_narrow;  /* [n_chunks] narrow index of the chunk's frame */
  const uint64_t* chunk_part;    /* [n_chunks] first entry of its table; a frame's chunk
                                    tables follow each other, span + 1 entries apart */
  const uint32_t* wide_frame;    /* [n_wide] frame ids, ascending */
  const int64_t* wide_offsets;   /* [n_wide + 1] row offsets of the wide frames, compacted */
  const uint32_t* frame_narrow;  /* [nframes] narrow index of each frame, or 0xFFFFFFFF */
} jfx_batched_routing;

/* outputs: the outputs the call will write, bit 0 sum, 1 count, 2 min, 3 max,
 * 4 mean. */
int jfx_batched_routed_workspace_size(uint32_t total_bits, int val_type, uint32_t nframes,
                                      const jfx_batched_routing* route, int outputs,
                                      uint64_t* out_bytes);
int jfx_batched_groupby_routed(int key_type, const void* keys, int val_type, const void* vals,
                               uint64_t n, const int64_t* frame_offsets, uint32_t nframes,
                               uint64_t key_min, uint32_t key_bits, int digit_bits, int sort_shape,
                               const jfx_batched_routing* route, void* out_keys, void* out_sum,
                               int64_t* out_count, void* out_min, void* out_max, void* out_mean,
                               uint32_t* out_ngroups, void* workspace, uint64_t ws_bytes,
                               void* stream);

/* ---- indexed filter -> groupby (serving) ------------------------------- */
/* Per group of a frame indexed by (key, value) -- values ascending inside each
 * group by their order-preserving image, NaN last (nan_start[g] = first NaN
 * row of group g; ninf_end[g] = end of its leading -inf rows; pinf_start[g] =
 * its first +inf row) -- and prefix[] = float64 prefix sums of those values
 * restarted at every group (n + 1 entries, NaN and +-inf as 0: prefix[i + 1]
 * = sum of the values of row i's group up to row i, so a group's sums never
 * carry the rounding of other groups' values): count, sum and mean of the group's rows with
 * value <cmp> *value (JFX_CMP_*; IEEE semantics, a NaN value passes only !=
 * and makes its group's sum NaN), found by binary searches and prefix-sum
 * differences; out_mask[g] = 1 where the count is nonzero. The result of
 * filtering rows and then SortBasedGroupBy.groupby_aggregate
 * (parallel_algorithms.py:733-803), restricted to the groups that keep rows.
 * val_type: JFX_F32 or JFX_F64. No counterpart in the reference. */
int jfx_served_groups(int val_type, const void* vals, const double* prefix, const uint32_t* starts,
                      const uint32_t* nan_start, const uint32_t* ninf_end,
                      const uint32_t* pinf_start, uint32_t ngroups, int cmp, const void* value,
                      int64_t* out_count, void* out_sum, void* out_mean, uint8_t* out_mask,
                      void* stream);

/* ---- served join: a resident frame's key groups against a dimension table -- */
/* Inner join of a frame, given by its key groups (group_keys ascending in the
 * key order, group g's rows at key-order positions [group_starts[g],
 * group_starts[g+1])), with n_dim dimension rows: the pairs of
 * ParallelSortMergeJoin.join (parallel_algorithms.py:938-1075) in its (key,
 * left row, right row) order (pa:987-994), left rows taken in key order.
 * Integer keys (JFX_U32/I32/U64/I64). jfx_served_join_count maps the
 * dimension rows to groups, sorts them by group (fixed plan, no host read),
 * counts the pairs into *out_pairs and synchronises the stream once;
 * jfx_served_join_fill then writes, for every pair, the key, the frame's
 * payload (left_values in key order) and the dimension payload (dim_values
 * by row), on the same workspace. No counterpart in the reference. */
int jfx_served_join_workspace_size(uint64_t n_dim, uint32_t ngroups, uint64_t* out_bytes);
int jfx_served_join_count(int key_type, const void* group_keys, const uint32_t* group_starts,
                          uint32_t ngroups, const void* dim_keys, uint64_t n_dim,
                          uint64_t* out_pairs, void* workspace, uint64_t ws_bytes, void* stream);
int jfx_served_join_fill(int key_type, const void* group_keys, const uint32_t* group_starts,
                         uint32_t ngroups, int left_elem_bytes, const void* left_values,
                         uint64_t n_dim, int right_elem_bytes, const void* dim_values,
                         uint64_t n_pairs, void* out_keys, void* out_left, void* out_right,
                         void* workspace, uint64_t ws_bytes, void* stream);

/* One indexed filter -> groupby query of a batch: the index columns of
 * jfx_served_groups, the frame's group keys (in group order) and the
 * comparison. threshold is rounded to the value type. */
typedef struct jfx_served_query {
  const void* values;
  const double* prefix;
  const uint32_t* group_starts;
  const uint32_t* nan_start;
  const uint32_t* ninf_end;
  const uint32_t* pinf_start;
  const void* group_keys;
  uint32_t ngroups;
  int32_t cmp;
  double threshold;
} jfx_served_query;

int jfx_served_batch_workspace_size(uint32_t nqueries, uint64_t total_groups, uint64_t* out_bytes);

/* nqueries (<= 1024) queries of jfx_served_groups in one launch, with the
 * groups that keep rows compacted: query i's result is entries
 * [out_starts[i], out_starts[i+1]) of out_keys (key_bytes 4 or 8, copied
 * from group_keys), out_count, out_sum and out_mean (val_type JFX_F32 or
 * JFX_F64 for every query), groups ascending. Outputs hold sum(ngroups)
 * entries; out_starts is host memory of nqueries + 1 entries, written after
 * the call synchronises the stream once. Equal to jfx_served_groups followed
 * by a stable compaction of each query's groups. No counterpart in the
 * reference. */
int jfx_served_batch(int val_type, int key_bytes, const jfx_served_query* queries,
                     uint32_t nqueries, void* out_keys, int64_t* out_count, void* out_sum,
                     void* out_mean, uint64_t* out_starts, void* workspace, uint64_t ws_bytes,
                     void* stream);

/* ---- composite keys (multi-column sort / groupby / join) -------------- */
/* Packs 1..8 integer key columns (JFX_U32/I32/U64/I64, n rows each) into one
 * uint64 key per row whose order and equality are the lexicographic ones of
 * the columns: field j = ordered(k_j) - sub[j] (descending[j]: sub[j] -
 * ordered(k_j)), bits[j] wide, the last column least significant; ordered()
 * flips the sign bit of signed columns and sub[j] is the ordered minimum
 * (maximum when descending) of column j. Replaces the per-column stable
 * argsorts of multi_column_lexsort (parallel_algorithms.py:25-66) and the
 * OR of per-column head flags of groupby_aggregate_multi_column (:697-711)
 * by one sort / groupby of the composite. sum(bits) <= 64; descending may be
 * NULL (all ascending). */
int jfx_pack_keys(const void* const* keys, const int* key_types, int ncols, const uint64_t* sub,
                  const uint32_t* bits, const int* descending, uint64_t n, uint64_t* out,
                  void* stream);

/* Inverse of jfx_pack_keys with the same column description: writes every
 * column of each packed key to outs[j] (a column of key_types[j]). */
int jfx_unpack_keys(const uint64_t* packed, uint64_t n, const int* key_types, int ncols,
                    const uint64_t* sub, const uint32_t* bits, const int* descending,
                    void* const* outs, void* stream);

/* ---- stream compaction ------------------------------------------------ */
/* Workspace bytes for jfx_compact_rows / jfx_filter_rows (independent of n). */
int jfx_compact_workspace_size(uint64_t n, uint64_t* out_bytes);

/* Row positions, in row order, where mask[i] != 0 (n < 2^32): the
 * jnp.nonzero(valid_mask) step of compact_repartitioned_rows
 * (parallel_algorithms.py:277-293) and the frame layer's boolean row filter.
 * out_rows holds up to n entries; when out_count is non-NULL the stream is
 * synchronised and the number of positions written there. */
int jfx_compact_rows(const uint8_t* mask, uint64_t n, uint32_t* out_rows, uint64_t* out_count,
                     void* workspace, uint64_t ws_bytes, void* stream);

/* Row positions, in row order, where `col <cmp> *value` holds (cmp: JFX_CMP_*;
 * *value is a host value of col_type; IEEE comparisons, so NaN passes only
 * JFX_CMP_NE). The filter is evaluated on the fly, no mask is stored. */
int jfx_filter_rows(int col_type, const void* col, int cmp, const void* value, uint64_t n,
                    uint32_t* out_rows, uint64_t* out_count, void* workspace, uint64_t ws_bytes,
                    void* stream);

/* ---- HBM bandwidth probes ---------------------------------------------- */
/* No counterpart in the reference: the read and copy ceilings the kernels'
 * achieved bandwidth is compared with. mode 0 streams `bytes` of src with
 * 16-byte loads (`unroll` = 1, 2, 4 or 8 loads in flight per thread, a grid
 * of `blocks` workgroups of 256 threads, non-temporal loads when
 * nontemporal != 0) and stores one word per workgroup to dst (dst holds
 * `blocks` uint32); mode 1 copies src to dst (both `bytes`). Buffers and
 * bytes 16-byte aligned. Modes 2 and 3 are a radix pass's memory pattern
 * without its ranking: n = bytes / 12 (mode 2: n uint64 keys followed by n
 * uint32 payloads) or bytes / 8 (mode 3: keys only), n a multiple of 16384;
 * each 16384-key tile is read in order and written as runs of `unroll` keys
 * (a divisor of 16384), run d of tile t to [d*T*unroll + t*unroll, +unroll)
 * of the T = n / 16384 tiles; `blocks` is unused and nontemporal selects
 * non-temporal stores. Mode 4 writes `bytes` of dst with 16-byte stores and
 * no loads (the walk of mode 0; nontemporal selects non-temporal stores).
 * Asynchronous on `stream`. */
int jfx_bandwidth_probe(int mode, const void* src, void* dst, uint64_t bytes, uint32_t blocks,
                        uint32_t unroll, int nontemporal, void* stream);

/* ---- per-kernel timing ------------------------------------------------ */
/* No counterpart in the reference. When enabled, every kernel launch is
 * bracketed by HIP events on its own stream; totals per kernel name are
 * read back with jfx_prof_count / jfx_prof_get (which synchronise on the
 * pending events). */
int jfx_prof_enable(int on);
int jfx_prof_reset(void);
/* Time only launches whose kernel name starts with name_prefix ("" or NULL:
 * all), so a timed region can carry events around one kernel only. */
int jfx_prof_filter(const char* name_prefix);
/* Time every `every`-th launch that passes the filter (1, the default: all),
 * counted from the last jfx_prof_filter / jfx_prof_reset / jfx_prof_sample
 * call, so that a long timed region carries few events. */
int jfx_prof_sample(uint32_t every);
int jfx_prof_count(int* out_n);
int jfx_prof_get(int i, char* name_buf, uint64_t buf_len, double* total_ms, uint64_t* count);

#ifdef __cplusplus
}
#endif

#endif /* JFX_H */
