/*
 * wadjet_gpu.h — C-ABI of the MI355X hash-aggregate / hash-join library.
 *
 * Plain pointers and sizes only; no C++ or framework types cross this line.
 * Every entry point returns a wg_status; on failure wg_last_error() returns a
 * thread-local message (the Go adapter maps it to an `error`, the way
 * pipeline.go:38-75 turns operator failures into errors).
 *
 * Threading: a handle (wg_agg, wg_join, ...) is NOT thread-safe. The caller
 * serialises calls per handle, as HashAggregate does with h.mu
 * (internal/engine/exec/aggregate.go:1076); clones get their own handles.
 * Host pointers passed in are read only for the duration of the call and never
 * retained (the cgo rule; batch.go:113-131 ownership: a consumer may not keep a
 * pooled batch or its Sel past the call).
 *
 * Reference interfaces replaced (paths relative to the wadjet repository):
 *   Sink/SinkSource/MergeableSink      internal/engine/exec/operators.go:24-60
 *   HashAggregate                      internal/engine/exec/aggregate.go:98,1023-1220,3855,4202
 *   KernelFilter (selection vectors)   internal/engine/exec/filter.go:347-425
 *   HashJoin.Build / HashJoinProbe     internal/engine/exec/join.go:771-1162,2072-2139,2306-2515
 *   BoundedOutputOperator              internal/engine/exec/operators.go:86-90, join.go:2044-2060
 */
#ifndef WADJET_GPU_H
#define WADJET_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */

typedef enum wg_status {
  WG_OK = 0,
  WG_ERR_INVALID = 1,     /* bad argument / shape */
  WG_ERR_UNSUPPORTED = 2, /* shape the GPU path does not take: caller keeps the CPU operator */
  WG_ERR_DEVICE = 3,      /* HIP runtime error */
  WG_ERR_OOM = 4,         /* device allocation failed or footprint over budget */
  WG_ERR_STATE = 5,       /* call out of order (e.g. consume after finalize) */
  WG_ERR_NULL_KEY = 6     /* NULL in a multi-column group key (reference migrates to its generic path) */
} wg_status;

/* Thread-local text of the last failure on this thread ("" if none). */
const char* wg_last_error(void);
/* Library version string. */
const char* wg_version(void);

/* ------------------------------------------------------------------- types */

/* Column storage types. Semantics follow batch.Vector (internal/engine/batch/vector.go:302-465):
 *   WG_INT32   int32 storage, sign-extended: TypeInt32/Port/Date/Protocol
 *   WG_INT64   int64 storage: TypeInt64/Timestamp/Duration
 *   WG_FLOAT64 float64 storage
 *   WG_DECIMAL scaled int64 storage (device layout for DECIMAL(p<=18,s)); scale in wg_column.scale
 *   WG_DECIMAL128 batch.Int128 storage {int64 Hi; uint64 Lo} per row (decimal.go:11-14)
 *   WG_IPV4    int64 storage: TypeIPv4/MAC (keys, MIN/MAX; SUM/AVG emit NULL as in agg_scatter.go:519-578)
 *   WG_UINT32  compact IPv4: uint32 storage, zero-extended, otherwise as WG_IPV4
 *   WG_UINT16  compact Port: uint16 storage, zero-extended, otherwise as WG_INT32 */
typedef enum wg_type {
  WG_INT32 = 1,
  WG_INT64 = 2,
  WG_FLOAT64 = 3,
  WG_DECIMAL = 4,
  WG_DECIMAL128 = 5,
  WG_IPV4 = 6,
  WG_UINT32 = 7,
  WG_UINT16 = 8
} wg_type;

/* Aggregate functions, numbered as exec.AggFunc (aggregate.go:26-31).
 * COUNT(*) is WG_AGG_COUNT with input_col = -1 (aggColIdx < 0 in the reference). */
typedef enum wg_agg_func {
  WG_AGG_SUM = 0,
  WG_AGG_COUNT = 1,
  WG_AGG_MIN = 2,
  WG_AGG_MAX = 3,
  WG_AGG_AVG = 4
} wg_agg_func;

/* Comparison operators, numbered as kernel.CompareOp (kernel/compare.go:14-21). */
typedef enum wg_cmp { WG_EQ = 0, WG_NE = 1, WG_LT = 2, WG_LE = 3, WG_GT = 4, WG_GE = 5 } wg_cmp;

/* Where the column pointers of a call live. */
typedef enum wg_mem { WG_MEM_HOST = 0, WG_MEM_DEVICE = 1 } wg_mem;

/* One column of a batch: data (typed by `type`), optional validity bitmap
 * (batch.Bitmap words: bit i%64 of word i/64, 1 = valid; NULL = no nulls). */
typedef struct wg_column {
  const void* data;
  const uint64_t* validity;
  int32_t type;  /* wg_type */
  int32_t scale; /* DECIMAL scale, else 0 */
} wg_column;

/* One aggregate: exec.AggColumn{Func, InputCol} (aggregate.go:73-81).
 * input_col indexes the batch's columns (-1 = COUNT(*)).
 * input_col2 >= 0 makes the input the product input_col * input_col2 — the
 * aggPreProject product (planner/physical/plan.go:7900-7990) formed exactly:
 * INT x DECIMAL -> DECIMAL at the decimal's scale, INT x INT -> INT64. */
typedef struct wg_agg_spec {
  int32_t func;       /* wg_agg_func */
  int32_t input_col;  /* -1 for COUNT(*) */
  int32_t input_col2; /* -1 unless a product input */
  int32_t reserved;
} wg_agg_spec;

/* A fused predicate `col <op> constant` (KernelFilter with a typed kernel,
 * filter.go:373-425): rows that fail, or whose column is NULL, are dropped
 * exactly as the Sel vector it would have produced drops them. */
typedef struct wg_predicate {
  int32_t col; /* -1 = no predicate */
  int32_t op;  /* wg_cmp */
  int64_t i64; /* constant for int columns */
  double f64;  /* constant for float columns */
} wg_predicate;

/* ----------------------------------------------------------------- context */

typedef struct wg_ctx wg_ctx;

/* One context = one device + one HIP stream (default priority). */
wg_status wg_ctx_create(int device, wg_ctx** out);
wg_status wg_ctx_destroy(wg_ctx* ctx);
wg_status wg_ctx_sync(wg_ctx* ctx);
/* Device properties: compute units, HBM bytes, LDS bytes per workgroup. */
wg_status wg_ctx_info(wg_ctx* ctx, int32_t* cus, uint64_t* hbm_bytes, int32_t* lds_bytes);
/* Elapsed ms of the last timed region on the ctx stream (HIP events around
 * the library's own kernels: see wg_ctx_timer_*). */
wg_status wg_ctx_timer_start(wg_ctx* ctx);
wg_status wg_ctx_timer_stop(wg_ctx* ctx, float* ms);

/* Failure injection, for tests of the error paths: the nth device allocation
 * made through this context from now on (the library's own and wg_malloc's)
 * fails as out of memory, WG_ERR_OOM, and with sticky != 0 every allocation
 * after it fails too, as under real memory exhaustion. nth = 0 clears it. The
 * reference's tests force the same paths with a small memory.Tracker budget
 * (internal/memory) and FatalEvalPanic (pipeline.go:45-57). After any
 * WG_ERR_OOM a handle stays valid — every buffer it held is kept, and destroy
 * releases it; see wg_agg_consume for what the aggregate then holds. */
wg_status wg_ctx_inject_alloc_failure(wg_ctx* ctx, int64_t nth, int32_t sticky);

/* Device memory helpers so callers need no other GPU runtime. */
wg_status wg_malloc(wg_ctx* ctx, uint64_t bytes, void** out);
wg_status wg_free(wg_ctx* ctx, void* ptr);
wg_status wg_memcpy_h2d(wg_ctx* ctx, void* dst, const void* src, uint64_t bytes);
wg_status wg_memcpy_d2h(wg_ctx* ctx, void* dst, const void* src, uint64_t bytes);
wg_status wg_memset(wg_ctx* ctx, void* dst, int value, uint64_t bytes);
/* Device-to-device copy by a plain streaming kernel (16-byte loads and
 * stores). Both pointers and `bytes` must be multiples of 16. Also the
 * known-size kernel against which HBM traffic counters are calibrated. */
wg_status wg_memcpy_d2d(wg_ctx* ctx, void* dst, const void* src, uint64_t bytes);

/* ------------------------------------------------------------- aggregation */

typedef struct wg_agg wg_agg;

/* Create a grouped aggregation: GROUP BY key_cols[0..n_keys) over a batch
 * schema col_types[0..n_cols) (wg_type) with col_nullable[i] != 0 when the
 * column may carry a validity bitmap. Typed dispatch is resolved here, once,
 * like resolveIndices (aggregate.go:1296-1764): one int-class key -> 64-bit
 * key; 2-4 int-class keys -> packed composite key (packed_hash.go:386-472):
 * one 64-bit word when the columns fit it, else the reference's 128-bit
 * {lo, hi} layout (widest column first), dictionary-encoded on the device to
 * a 64-bit id per distinct key (a 128-bit key needs n_cols < 16).
 * group_hint sizes the table (InputRowHint/GroupNDVHint, aggregate.go:1464-1506);
 * the table grows when it is exceeded.
 * Keys wider than 16 bytes in total, string keys and other functions return
 * WG_ERR_UNSUPPORTED (the caller keeps the CPU operator). */
wg_status wg_agg_create(wg_ctx* ctx, int32_t n_cols, const int32_t* col_types,
                        const int32_t* col_scales, const uint8_t* col_nullable,
                        int32_t n_keys, const int32_t* key_cols, int32_t n_aggs,
                        const wg_agg_spec* aggs, uint64_t group_hint, wg_agg** out);
wg_status wg_agg_destroy(wg_agg* h);

/* Sink.Consume (operators.go:27; aggregate.go:1075). cols[n_cols] follow the
 * create-time schema; sel (ascending row indices, may be NULL) selects n_sel
 * of the n_rows rows; preds[0..n_preds) (n_preds <= 2) are fused filters that
 * must all pass — a chain of KernelFilters ANDed through one Sel
 * (filter.go:408-418).
 * mem = WG_MEM_HOST: the call copies what it needs before returning and
 * coalesces small batches into large device batches. mem = WG_MEM_DEVICE:
 * pointers are device-resident; the launch is enqueued on the ctx stream.
 * Out of device memory (WG_ERR_OOM): when every allocation the call needed
 * failed before its rows reached the aggregate (scratch, staging, capacity,
 * the partitioned passes' buffers), the aggregate holds exactly what it held
 * before and the call may be retried or the state handed over
 * (wg_agg_spill_run). When the table could not grow while the batch was
 * being applied, the message says the aggregate holds part of the call's
 * input; wg_agg_consume, _merge, _merge_run, _finalize and
 * wg_join_probe_agg then return WG_ERR_STATE until wg_agg_reset (the
 * adapter fails the query or re-runs it elsewhere, as pipeline.go:45-57
 * turns a failed operator into a query error). The same holds for
 * wg_agg_merge, wg_agg_merge_run, wg_agg_finalize and wg_join_probe_agg. */
wg_status wg_agg_consume(wg_agg* h, const wg_column* cols, uint64_t n_rows,
                         const uint32_t* sel, uint64_t n_sel, const wg_predicate* preds,
                         int32_t n_preds, int32_t mem);
/* MergeableSink.MergeSink (operators.go:56-60; aggregate.go:4682-4700):
 * fold src's partial state into dst (same create-time spec). src is left empty. */
wg_status wg_agg_merge(wg_agg* dst, wg_agg* src);
/* Sink.Finalize (aggregate.go:3855): flush staged rows, compact groups. */
wg_status wg_agg_finalize(wg_agg* h);
/* Number of groups after finalize. */
wg_status wg_agg_num_groups(wg_agg* h, uint64_t* out);
/* Source.Next payload (aggregate.go:4226-4524), group order unspecified.
 * Writes `count` groups starting at `offset` to host or device buffers:
 *   key_out[k]      : key column k in its input storage type (int32/int64/uint32/uint16)
 *   val_out[a]      : finalized value of aggregate a — float64 for SUM/AVG and
 *                     float MIN/MAX, int64 for COUNT and int MIN/MAX
 *                     (writeAccToColumn, aggregate_partial_spill.go:1855-1945)
 *   valid_out[a]    : 1 byte per group, 0 = NULL (SUM/AVG with Count==0, MIN/MAX without a value)
 * Any pointer may be NULL to skip that column. */
wg_status wg_agg_fetch(wg_agg* h, uint64_t offset, uint64_t count, void* const* key_out,
                       void* const* val_out, uint8_t* const* valid_out, int32_t mem);
/* Raw accumulator state per aggregate (kernel.Accumulator fields, types.go:21-40):
 * count[g] (int64), sum_lo/sum_hi[g] (int64 sum in sum_lo, or Int128 as lo/hi,
 * or float64 bits of SumF64 in sum_lo), minmax[g] (int64, or float64 bits).
 * Used for the exact parity check of Int128 sums. NULL pointers are skipped. */
wg_status wg_agg_fetch_state(wg_agg* h, int32_t agg, uint64_t offset, uint64_t count,
                             int64_t* count_out, uint64_t* sum_lo_out, int64_t* sum_hi_out,
                             int64_t* minmax_out, int32_t mem);
/* Validity of the group keys (1 byte per group, 0 = the NULL group of a
 * single-column key, aggregate_null_group_test.go:48-289). */
wg_status wg_agg_fetch_key_valid(wg_agg* h, uint64_t offset, uint64_t count, uint8_t* valid_out,
                                 int32_t mem);
/* The finalized groups as ONE partial-state spill run in the reference's
 * external-merge format ("WAGS\x01", aggregate_partial_spill.go:19-52: the run
 * drainStateToRuns writes, :1414-1470), groups ascending by their binary sort
 * key, so the CPU operator's k-way merge (finalizeViaPartialMerge) can combine
 * GPU partials with its own runs. group_names / group_type_ids (parquet.TypeID)
 * fill the header; NULL type ids derive them from the column types. Call with
 * out = NULL to learn *size; out is a host buffer of cap bytes. The aggregate
 * is not changed (wg_agg_reset afterwards drains it, as the reference's spill
 * resets its state). */
wg_status wg_agg_spill_run(wg_agg* h, const char* const* group_names, const uint16_t* group_type_ids,
                           void* out, uint64_t cap, uint64_t* size);
/* Merge one partial-state spill run (the same format, written by this library
 * or by the CPU operator's drainStateToRuns for the same GROUP BY and
 * aggregates) into a not yet finalized aggregate: the k-way merge's combine
 * rules (mergeFlatAccumRow, aggregate.go:5155-5228) applied on the device.
 * WG_ERR_INVALID when the run's aggregates or group columns differ. */
wg_status wg_agg_merge_run(wg_agg* h, const void* run, uint64_t size);
/* Drop all groups and staged rows; keep the spec and the table memory. Also
 * makes an aggregate usable again after a call left it holding part of a
 * failed input. */
wg_status wg_agg_reset(wg_agg* h);
/* Device bytes held by this aggregate (table + state + staging). */
wg_status wg_agg_footprint(wg_agg* h, uint64_t* bytes);
/* HashAggregate.StateBytes (aggregate.go:4050-4054, the bound
 * cappedPartialAgg flushes on, worker/shuffle_partial_agg.go:186-188): bytes
 * of live group state = groups so far x the bytes of one group's state
 * (slot words + exact Int128 bases). Rows staged from host batches are
 * aggregated first, so the figure covers every row consumed. */
wg_status wg_agg_state_bytes(wg_agg* h, uint64_t* bytes);
/* Tuning knobs (autotuner output): lds_slots = 0 disables the LDS table,
 * block = threads per workgroup (256, 512, 1024). */
wg_status wg_agg_set_tuning(wg_agg* h, int32_t lds_slots, int32_t block);
/* Counters from the last consume: rows that hit the LDS table, rows that went
 * to the global table (partitioned consume: rows written to the partitions),
 * LDS entries flushed, global table capacity. */
wg_status wg_agg_stats(wg_agg* h, uint64_t* lds_rows, uint64_t* global_rows,
                       uint64_t* flushed, uint64_t* capacity);

/* Device time (HIP events on the ctx stream) of the consume kernel launches
 * of the last consume call, and how many launches it took (1 unless the
 * table had to grow). */
wg_status wg_agg_kernel_time(wg_agg* h, float* ms, uint32_t* launches);

/* Device time (ms) of the last consume by phase: [0] direct consume kernel,
 * [1] partitioned consume pass 1 (heavy-hitter set, tiles sorted by level-1
 * partition, partition totals + scan), [2] pass 2 (records gathered into
 * contiguous final partitions), [3] partition aggregate. */
wg_status wg_agg_pass_times(wg_agg* h, float* ms4);
/* Consume strategy: 0 = pick by batch size and cardinality, 1 = direct
 * (shared table with LDS pre-aggregation), 2 = partitioned (two sorting
 * passes by hash partition, per-partition LDS aggregation); partition LDS table slots, workgroup size
 * and the largest partition chunk one workgroup takes. */
wg_status wg_agg_set_strategy(wg_agg* h, int32_t strategy, int32_t part_lds_slots, int32_t part_block,
                              uint64_t part_chunk_rows);
/* Heavy hitters of the partitioned consume (no reference counterpart; results
 * do not depend on it): up to hot_slots keys that a sample of the batch sees
 * most often are aggregated in LDS during the first sorting pass instead of
 * being partitioned. 0 disables; otherwise a power of two in [64, 2048]. */
wg_status wg_agg_set_heavy_hitters(wg_agg* h, int32_t hot_slots);

/* ------------------------------------------------------------------ filter */

/* KernelFilter.Execute (filter.go:373-425) on device-resident data: writes the
 * ascending indices of rows in [0,n_rows) (or of sel[0..n_sel) when sel != NULL)
 * whose column passes `col <op> constant` (NULL never passes) to out_sel and
 * their number to *n_out. Compaction is per 64-lane wavefront (ballot + mbcnt
 * prefix) with an order-preserving block scan. All pointers device-resident
 * except n_out (host). */
wg_status wg_filter(wg_ctx* ctx, const wg_column* col, uint64_t n_rows, const uint32_t* sel,
                    uint64_t n_sel, int32_t op, int64_t i64, double f64, uint32_t* out_sel,
                    uint64_t* n_out);

/* wg_filter without any allocation: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same arguments,
 * results and errors as wg_filter, for at most 2^26 (67 108 864) positions
 * per call (WG_ERR_INVALID beyond). Its scratch (about 8.1 MiB of tile counts,
 * offsets and ballot words) is one static device array per GPU, held until
 * the process ends and shared by every call on that GPU: calls that overlap
 * on one GPU (several threads or streams) race. */
wg_status wg_filter_noalloc(wg_ctx* ctx, const wg_column* col, uint64_t n_rows, const uint32_t* sel,
                            uint64_t n_sel, int32_t op, int64_t i64, double f64, uint32_t* out_sel,
                            uint64_t* n_out);

/* -------------------------------------------------------------------- join */

typedef struct wg_join wg_join;

/* HashJoin{InnerJoin, LeftKeys=[probe key], RightKeys=[build key]} (join.go:42-240)
 * with a single int-class key (tryEnableIntKey, join.go:520-546). The build
 * side schema is build_types[0..n_build_cols); build_key is its key column.
 * Columns listed in inline_cols are copied into the table slots so that a
 * fused probe reads them with the key (at most 4). build_hint sizes the table. */
wg_status wg_join_create(wg_ctx* ctx, int32_t n_build_cols, const int32_t* build_types,
                         const int32_t* build_scales, int32_t build_key, int32_t n_inline,
                         const int32_t* inline_cols, uint64_t build_hint, wg_join** out);
/* The same join on two int-class key columns (tryEnableDualIntKey with
 * dualIntHash / dualIntKeyFromVectors, join.go:592-661, 1621-1648; the chain
 * walk that verifies both components, inlineDualIntProbe :2848-3130): a build
 * row matches a probe row when both keys are equal as int64 (int32 columns
 * widened), any number of build rows per pair (chains), and a row with a NULL
 * in either key never matches. Each distinct (key0, key1) pair of the build
 * gets an exact 64-bit id from a device dictionary; the table keys on it. The
 * probe entry points then take the id column wg_join_dual_probe_key makes. */
wg_status wg_join_create_dual(wg_ctx* ctx, int32_t n_build_cols, const int32_t* build_types,
                              const int32_t* build_scales, int32_t build_key0, int32_t build_key1,
                              int32_t n_inline, const int32_t* inline_cols, uint64_t build_hint,
                              wg_join** out);
wg_status wg_join_destroy(wg_join* j);
/* HashJoin.Build per batch (join.go:1019-1065): rows with a NULL key are
 * skipped. Build columns are retained on the device for the gathers. Every
 * allocation happens before the call's rows reach the table: after
 * WG_ERR_OOM no row of the call was added, and the call may be repeated. */
wg_status wg_join_build_append(wg_join* j, const wg_column* cols, uint64_t n_rows, int32_t mem);
/* End of build (join.go:1126-1152). */
wg_status wg_join_build_finish(wg_join* j);
wg_status wg_join_build_rows(wg_join* j, uint64_t* rows);
/* The retained build column build_col as a device column (data and validity
 * owned by the join, valid until wg_join_destroy or the next build append):
 * the base of the build side's view columns (wg_view_flatten). */
wg_status wg_join_build_column(wg_join* j, int32_t build_col, wg_column* out);
/* Probe key of a two-column-key join (after build_finish): for every active
 * row (sel, or 0..n_rows) out_ids[row] = the id of its (key0, key1) pair, an id
 * no build row has when the build lacks the pair; out_valid (bitmap words,
 * (n_rows + 63) / 64) = key0 valid AND key1 valid. Pass
 * {out_ids, out_valid, WG_INT64} as the probe key of wg_join_probe,
 * wg_join_semi, wg_join_bloom_filter, or append it to the probe columns of
 * wg_join_probe_agg. Device buffers. */
wg_status wg_join_dual_probe_key(wg_join* j, const wg_column* probe_key0, const wg_column* probe_key1,
                                 uint64_t n_rows, const uint32_t* sel, uint64_t n_sel, int64_t* out_ids,
                                 uint64_t* out_valid);

/* Resumable probe cursor (HashJoinProbe.res, join.go:2670-2836). */
typedef struct wg_probe_cursor {
  uint64_t pos;   /* next probe position (index into sel or rows) */
  int64_t chain;  /* build row to resume a suspended chain at, -1 = none */
  int32_t done;   /* 1 when the input is fully probed */
  int32_t reserved;
} wg_probe_cursor;

/* Inner-join probe of one device-resident probe batch (inlineIntProbe,
 * join.go:2670-2814): emits at most out_cap (probe_row, build_row) pairs,
 * compacted per wavefront; resumes from *cursor and updates it. A call
 * examines a window of probe positions (twice its output room, at least 64K),
 * so it may return fewer than out_cap pairs, even none, before the input is
 * done: call again until cursor->done. Rows with a NULL key never match. Pair order is unspecified; the pair multiset is the
 * reference's. probe_key is a device column; out_* are device buffers;
 * *n_out is written on the host. */
wg_status wg_join_probe(wg_join* j, const wg_column* probe_key, uint64_t n_rows,
                        const uint32_t* sel, uint64_t n_sel, uint32_t* out_probe_rows,
                        uint32_t* out_build_rows, uint64_t out_cap, wg_probe_cursor* cursor,
                        uint64_t* n_out);
/* Semi / anti join probe of one device-resident probe batch
 * (executeSemiAntiJoin, join.go:3245-3310; semiProbeInt64 / antiProbeInt64,
 * join.go:4486-4650): writes to out_sel the probe positions (ascending, from
 * sel or 0..n_rows) whose key has a build row (anti = 0) or has none
 * (anti = 1); a NULL probe key never matches, so only the anti join emits it.
 * out_sel is a device buffer of at least n entries; *n_out on the host. */
wg_status wg_join_semi(wg_join* j, const wg_column* probe_key, uint64_t n_rows,
                       const uint32_t* sel, uint64_t n_sel, int32_t anti, uint32_t* out_sel,
                       uint64_t* n_out);
/* Bloom pre-filter of the probe side (HashJoin.buildBloom, join.go:1441-1483;
 * BloomFilterOp.Execute's int-key path, bloom_filter_op.go:39-219): build
 * sets two bits per distinct build key (bloomHashInt, join.go:1650-1654) in a
 * power-of-two array of ~10 bits per key, at least 8 words. The filter writes
 * to out_sel (device, n entries) the ascending positions of sel (or 0..n_rows)
 * whose non-NULL key may be in the build, *n_out on the host. The
 * 32-batch / 5 % adaptive bypass is the caller's (host operator state). */
wg_status wg_join_bloom_build(wg_join* j);
wg_status wg_join_bloom_filter(wg_join* j, const wg_column* probe_key, uint64_t n_rows,
                               const uint32_t* sel, uint64_t n_sel, uint32_t* out_sel,
                               uint64_t* n_out);
/* The filter's words (*words = count; copied to host `out` when cap >= count). */
wg_status wg_join_bloom_words(wg_join* j, uint64_t* words, uint64_t* out, uint64_t cap);
/* Gather a retained build column at build row indices (gatherBuildVector,
 * join.go:4024) into a device buffer of the column's storage type. */
wg_status wg_join_gather_build(wg_join* j, int32_t build_col, const uint32_t* build_rows,
                               uint64_t n, void* out);
/* Gather a device column at row indices (gatherVector, sort.go:611). */
wg_status wg_gather(wg_ctx* ctx, const wg_column* col, const uint32_t* rows, uint64_t n,
                    void* out);
/* Late materialization: the probe's (probe_row, build_row) pair arrays are the
 * indices of view columns over the probe batch and the retained build columns
 * (emitViewOutput, join.go:2517-2600; NewViewVector, batch/view.go:27-80),
 * kept on the device until a consumer needs values. This flattens one view
 * (Vector.Flatten, view.go:87-230): out[i] = base[indices[i]] and bit i of
 * out_valid ((n + 63) / 64 words, all written; may be NULL) set, unless row i
 * is NULL in the view's own bitmap own_valid (NULL = no own NULLs; the base
 * is then not read) or the base row is NULL — out[i] = 0 and the bit clear.
 * A view over a view composes (view.go:50-70) through this same call with
 * base = the inner view's indices as a WG_UINT32 column whose validity is the
 * inner view's own bitmap. Device pointers; WG_ERR_INVALID when an index is
 * >= base_rows. Synchronises. */
wg_status wg_view_flatten(wg_ctx* ctx, const wg_column* base, uint64_t base_rows, const uint32_t* indices,
                          const uint64_t* own_valid, uint64_t n, void* out, uint64_t* out_valid);

/* Fused probe -> aggregate: join the device-resident probe batch against the
 * build, and feed every matched (probe row, build row) pair to `agg` as one
 * joined row whose columns are the probe columns followed by the build
 * columns (joinOutputSchemaWithMapping, join.go:3898). The aggregate's
 * create-time schema must be exactly that concatenation. */
wg_status wg_join_probe_agg(wg_join* j, const wg_column* probe_cols, int32_t n_probe_cols,
                            int32_t probe_key, uint64_t n_rows, const uint32_t* sel,
                            uint64_t n_sel, wg_agg* agg);
/* Tuning of wg_join_probe_agg (no reference counterpart; results do not
 * depend on it). lean = 1 lets a query whose single group key and values come
 * from probe columns or the first two inline payload columns, none nullable,
 * take the streamlined kernel; lean = 0 keeps the general one. block: 512 or
 * 1024 threads; replicas: accumulator copies per LDS group slot (power of two
 * <= 64, 0 = as many as the LDS budget allows). */
wg_status wg_join_set_tuning(wg_join* j, int32_t lean, int32_t block, int32_t replicas);
/* Order-free digest of the join output multiset over (probe key, probe col q,
 * build col p1, build col p2): count plus two independent 64-bit sums of
 * per-row mixes, computed on the device from a pairs list. */
wg_status wg_join_digest(wg_join* j, const wg_column* probe_cols, const uint32_t* probe_rows,
                         const uint32_t* build_rows, uint64_t n, const int32_t* probe_sel_cols,
                         int32_t n_probe_sel, const int32_t* build_sel_cols, int32_t n_build_sel,
                         uint64_t digest_out[2]);

/* ---------------------------------------------------------------- parquet */

/* Parquet data pages decoded on the device (the reference's ColumnPageReader:
 * decodeDataPageV1 / V2, decodeValues, resolveDictForRows, the RLE /
 * bit-packing hybrid of rle.go — internal/storage/parquet/page_reader.go:
 * 434-700, reader.go:1020-1060). The host parses page headers, decompresses,
 * and scans the run headers of each hybrid stream (wg_pq_scan_runs); the
 * device expands runs, bit-unpacks, resolves dictionary indices and places
 * values and validity. */

/* Run headers of one RLE / bit-packing hybrid stream of `count` values at
 * bit width bit_width (host memory): runs[4 i..4 i+3] = {first value index,
 * value count, RLE value or bit offset of a bit-packed run from `data`,
 * 1 = bit-packed}. Call with runs = NULL to learn *n_runs; *used = bytes of
 * the stream consumed; *ones (may be NULL) = values equal to 1 among the
 * first `count` (the present rows of a definition-level stream).
 * WG_ERR_INVALID for a stream that ends early. */
wg_status wg_pq_scan_runs(const uint8_t* data, uint64_t len, int32_t bit_width, uint64_t count, uint32_t* runs,
                          uint64_t cap, uint64_t* n_runs, uint64_t* used, uint64_t* ones);

/* One hybrid stream on the device: its run table (from wg_pq_scan_runs) and
 * bytes; `data` 8-byte aligned with 16 readable bytes past the stream (the
 * decode reads whole 64-bit words). */
typedef struct wg_pq_runs {
  const uint32_t* runs;
  uint32_t n_runs;
  int32_t bit_width;
  const uint8_t* data;
} wg_pq_runs;

/* One decompressed data page of a flat (non-repeated) column. */
typedef struct wg_pq_page {
  uint64_t num_rows;      /* rows of the page, NULLs included */
  int32_t physical_type;  /* parquet Type: 1 INT32, 2 INT64, 5 DOUBLE */
  int32_t encoding;       /* parquet Encoding: 0 PLAIN, 2 PLAIN_DICTIONARY, 8 RLE_DICTIONARY */
  wg_pq_runs def;         /* definition levels (max level 1); n_runs = 0 for a required column or a page
                             without NULLs (every row then has a value) */
  const void* plain;      /* PLAIN: the non-NULL rows' values, packed */
  wg_pq_runs idx;         /* dictionary encodings: the non-NULL rows' indices */
  const void* dict;       /* the dictionary page's values (PLAIN, dict_n of them) */
  uint64_t dict_n;
} wg_pq_page;

/* Decode one page into out[row0 .. row0 + num_rows) (4- or 8-byte values by
 * physical type; NULL rows get 0) and, when out_valid != NULL, set the
 * validity bits of its present rows (bit (row0 + r) of the zero-initialised
 * bitmap words). Device pointers. WG_ERR_INVALID when a dictionary index is
 * out of range (ValidateDictIndices, reader.go:1020-1031) or a present row has
 * no index (a page whose index stream is empty must be all NULL).
 * Synchronises. */
wg_status wg_parquet_decode_page(wg_ctx* ctx, const wg_pq_page* page, void* out, uint64_t* out_valid,
                                 uint64_t row0);

/* Consecutive data pages of one column — a column chunk, or the chunks of
 * several row groups, each page naming its own chunk's dictionary — decoded
 * as one: page i lands at rows row0 + (rows of pages 0..i-1). The page array
 * is host memory, everything it points to device memory; all pages of one
 * physical type. Same results and errors as wg_parquet_decode_page on each
 * page in turn, in three launches and one synchronisation (the reference
 * decodes a chunk page by page, ColumnPageReader.ReadPage,
 * page_reader.go:434-520). */
wg_status wg_parquet_decode_pages(wg_ctx* ctx, const wg_pq_page* pages, uint64_t n_pages, void* out,
                                  uint64_t* out_valid, uint64_t row0);

/* One column chunk of a Parquet file, as the footer's ColumnMetaData gives it. */
typedef struct wg_pq_chunk {
  uint64_t offset;    /* file offset of the chunk's first page (its dictionary page when it has one) */
  uint64_t bytes;     /* total_compressed_size: every page header and body of the chunk */
  uint64_t num_rows;  /* rows of the chunk's row group */
  int32_t codec;      /* CompressionCodec: 0 UNCOMPRESSED, 1 SNAPPY, 2 GZIP, 6 ZSTD */
  int32_t reserved;   /* 0 */
} wg_pq_chunk;

/* The pages of one flat column's chunks, staged for wg_parquet_decode_pages. */
typedef struct wg_pq_staged wg_pq_staged;

/* Stage the chunks of one flat column of the file at `path` (several row
 * groups' chunks of the same column, in row order): on up to `threads` host
 * threads (0 = the usable cores, at most 16), each chunk is read, its Thrift
 * compact page headers parsed, its pages decompressed and the run headers of
 * each hybrid stream scanned (ColumnPageReader.ReadPage / decodeDataPageV1 /
 * decodeDataPageV2, internal/storage/parquet/page_reader.go:434-560;
 * Decompress, decompress.go:15-100); the streams go to device memory of ctx's
 * device through pinned staging slots, overlapping the next chunks' host work.
 * ctx = NULL stages into host memory instead (the page table then points at
 * host memory): the host half on its own. physical_type 1 INT32, 2 INT64 or
 * 5 DOUBLE; max_def 0 (required) or 1 (optional). Data pages v1 and v2,
 * PLAIN or dictionary encoded, with RLE definition levels. WG_ERR_UNSUPPORTED
 * for another codec, encoding or level encoding; WG_ERR_INVALID for a file
 * that cannot be read, a malformed page, or chunks whose data pages do not
 * hold their rows. Synchronises. */
wg_status wg_pq_stage(wg_ctx* ctx, const char* path, const wg_pq_chunk* chunks, uint64_t n_chunks,
                      int32_t physical_type, int32_t max_def, int32_t threads, wg_pq_staged** out);

/* Rows, page count, the stream bytes the decode reads (levels, values or
 * indices, dictionaries, run tables) and the staged memory's size. */
wg_status wg_pq_staged_info(const wg_pq_staged* s, uint64_t* rows, uint64_t* n_pages, uint64_t* stream_bytes,
                            uint64_t* staged_bytes);

/* Copy the page table (min(cap, n_pages) pages; host memory whose pointers
 * name the staged memory) into pages: the `pages` argument of
 * wg_parquet_decode_pages. */
wg_status wg_pq_staged_pages(const wg_pq_staged* s, wg_pq_page* pages, uint64_t cap);

/* Release the staged memory; NULL is a no-op. */
wg_status wg_pq_staged_free(wg_pq_staged* s);

/* One page body decompressed (Decompress, decompress.go:15-100): codec as in
 * wg_pq_chunk; dst_len must equal the page's uncompressed size.
 * WG_ERR_INVALID for corrupt input or a size mismatch. Host memory. */
wg_status wg_pq_decompress(int32_t codec, const uint8_t* src, uint64_t len, uint8_t* dst, uint64_t dst_len);

/* ----------------------------------------------------------------- shuffle */

/* Routing of a partitioned shuffle sink (partitionedShuffleSink.Consume,
 * internal/worker/partitioned_shuffle_sink.go:199-290): every active row goes
 * to partition FNV-1a-64(key bytes) % num_parts, as hashRowsIntoPartitions
 * computes it (:1039-1171: each key column's storage bytes in the reference's
 * vectors, little-endian — 4 for Int32/Port, 8 for Int64/IPv4/Float64 —,
 * 0xff for a NULL, one 0x00 byte for a decimal column; no key columns = all
 * rows to partition 0; num_parts 0 counts as 1), and the rows of each
 * partition keep their input order (perPartRows, :272-290).
 * Device columns; sel (device, ascending, < n_rows) selects the active rows,
 * NULL = all n_rows. Outputs: part (device, n active entries, or NULL) =
 * partition of the i-th active row (the `parts` array); rows (device, n
 * active entries) = the active rows' indices (b.Sel[i] or i) grouped by
 * partition; offsets (host, num_parts + 1 entries) = where each partition's
 * rows start in `rows`, then the total. num_parts <= 16384. Synchronises. */
wg_status wg_shuffle_partition(wg_ctx* ctx, const wg_column* cols, int32_t n_cols, const int32_t* key_cols,
                               int32_t n_keys, uint64_t n_rows, const uint32_t* sel, uint64_t n_sel,
                               int32_t num_parts, uint32_t* part, uint32_t* rows, uint64_t* offsets);

/* ------------------------------------------------------- batched execution */

/* Many independent GROUP BY queries in one persistent launch. Each query has
 * its own device-resident columns, spec and result buffers. The runner owns a
 * work queue of (query, tile) descriptors, a hipGraph of
 * {reset, persistent aggregate, finalize} and the autotuner's choices. */
typedef struct wg_batch wg_batch;

typedef struct wg_batch_query {
  wg_column key;       /* one int-class key column (device) */
  wg_column val;       /* value column for SUM (device); INT32, INT64 or DECIMAL */
  uint64_t n_rows;     /* rows [row_begin, row_begin + n_rows) of the columns */
  uint64_t est_groups; /* cardinality estimate used to size the table */
  uint64_t row_begin;
  /* optional window filter: rows with range_lo <= range_col < range_hi
   * (two ANDed KernelFilters); range_col.data == NULL disables it */
  wg_column range_col;
  int64_t range_lo, range_hi;
  /* optional inner join to a dimension (unique build keys): rows whose key has
   * no build row drop out; the group key becomes inline payload `dim_payload`
   * of the matched build row (fused probe -> GROUP BY build column) */
  const wg_join* dim;
  int32_t dim_payload;
  int32_t reserved;
} wg_batch_query;

wg_status wg_batch_create(wg_ctx* ctx, int32_t max_queries, uint64_t max_groups_total,
                          wg_batch** out);
wg_status wg_batch_destroy(wg_batch* b);
/* Run the queries: one persistent launch (graph-replayed when use_graph) or,
 * when plain != 0, one launch sequence per query. Results stay on the device
 * until wg_batch_result. Per query: GROUP BY key SUM(val), COUNT(*). */
wg_status wg_batch_run(wg_batch* b, const wg_batch_query* qs, int32_t n_queries,
                       int32_t use_graph, int32_t plain);
/* Autotune: time candidate tunings on these queries (device time of the
 * launch, HIP events) and keep the fastest: workgroup size, LDS threshold
 * (queries estimated at or below it aggregate in an LDS table per tile,
 * larger ones in their global table), rows per work item, and the partition
 * fan-out (a query above the threshold whose groups split into at most that
 * many key-hash partitions of at most the threshold each becomes one work
 * item per (tile, partition), each aggregating its partition in LDS), and the
 * LDS cache of the queries still above it. */
wg_status wg_batch_autotune(wg_batch* b, const wg_batch_query* qs, int32_t n_queries);
/* Per-query result: number of groups, then keys/sums(lo,hi as Int128)/counts. */
wg_status wg_batch_result(wg_batch* b, int32_t q, uint64_t* n_groups, int64_t* keys,
                          uint64_t* sum_lo, int64_t* sum_hi, int64_t* counts, uint64_t cap);
/* Chosen layout for query q in the last run: 1 = LDS table(s) holding all
 * its groups, 2 = global table behind an LDS cache of its frequent keys,
 * 0 = global table only; workgroup size; key partitions per tile (1 = no
 * fan-out). */
wg_status wg_batch_choice(wg_batch* b, int32_t q, int32_t* use_lds, int32_t* block, int32_t* fanout);
/* Current tuning (what wg_batch_autotune picked, or what was set): workgroup
 * size, LDS cardinality threshold (0 = never an LDS table), rows per work
 * item (an upper bound: a batch halves it, down to 8K rows, until it has four
 * work items per compute unit), largest partition fan-out (a power of two
 * <= 64; 1 = none), and
 * whether queries still above the threshold keep an LDS table as a cache of
 * their frequent keys (skewed keys: the hot keys' rows aggregate in LDS). */
wg_status wg_batch_tuning(wg_batch* b, int32_t* block, int32_t* lds_threshold, uint64_t* tile_rows,
                          int32_t* max_fanout, int32_t* lds_cache);
wg_status wg_batch_set_tuning(wg_batch* b, int32_t block, int32_t lds_threshold, uint64_t tile_rows,
                              int32_t max_fanout, int32_t lds_cache);

/* ---------------------------------------------------------------- serving */

/* A request loop over a device-resident table (continuous batching).
 * The reference serves queries through its worker (internal/worker/executor.go,
 * one pipeline per fragment); here requests are admitted by the HBM footprint
 * of their hash tables into `lanes` concurrent batches (each lane its own HIP
 * stream and persistent-launch runner): a request that arrives while other
 * batches run joins the next lane that frees up, so work overlaps on the
 * device. mode 1 runs the same requests one per launch, FIFO (the plain path). */
typedef struct wg_serve wg_serve;

typedef struct wg_serve_query {
  int32_t kind;        /* 0: GROUP BY key_col; 1: JOIN dim ON key_col GROUP BY dim payload */
  int32_t key_col;     /* table column: group key (kind 0) or join key (kind 1) */
  int32_t val_col;     /* table column under SUM */
  int32_t dim;         /* kind 1: id from wg_serve_add_dim */
  int32_t dim_payload; /* kind 1: inline payload index of the dimension */
  int32_t reserved;
  int64_t ts_lo, ts_hi; /* window [ts_lo, ts_hi) on the table's ts column */
  uint64_t est_groups;
} wg_serve_query;

typedef struct wg_serve_stats {
  uint64_t completed;
  uint64_t batches;       /* launches */
  uint64_t max_in_flight; /* most lanes busy at once */
  uint64_t peak_footprint;/* largest admitted hash-table bytes at once */
  double p50_ms, p99_ms, mean_ms, max_ms;
  double wall_s, qps;     /* first arrival to last completion; completed / wall_s */
  double device_busy_s;   /* union over lanes of each batch's device interval (first to last operation) */
  double lane_busy_s;     /* sum of those intervals (above device_busy_s while lanes overlap) */
  double host_submit_s;   /* engine thread: planning, uploading and launching batches */
  double host_retire_s;   /* engine thread: retiring finished batches, starting their result copies */
  double host_deliver_s;  /* engine thread: handing copied results to their requests */
} wg_serve_stats;

/* cols[n_cols] are device-resident columns of n_rows rows; ts_col is a
 * non-decreasing int64 column used for windows. hbm_budget bounds the bytes of
 * hash tables admitted at once (over all lanes). */
wg_status wg_serve_create(wg_ctx* ctx, const wg_column* cols, int32_t n_cols, uint64_t n_rows,
                          int32_t ts_col, int32_t lanes, uint64_t hbm_budget, int32_t max_batch,
                          int32_t mode, wg_serve** out);
wg_status wg_serve_destroy(wg_serve* s);
/* Register a built dimension join (unique keys); returns its id. */
wg_status wg_serve_add_dim(wg_serve* s, const wg_join* dim, int32_t* dim_id);
/* Thread-safe submit; the arrival time is taken here. */
wg_status wg_serve_submit(wg_serve* s, const wg_serve_query* q, uint64_t* id);
/* Block until request `id` completed; its group count and latency. */
wg_status wg_serve_wait(wg_serve* s, uint64_t id, uint64_t* n_groups, double* latency_ms);
/* Result rows of a completed request (keys, Int128 sums as lo/hi, counts). */
wg_status wg_serve_result(wg_serve* s, uint64_t id, int64_t* keys, uint64_t* sum_lo, int64_t* sum_hi,
                          int64_t* counts, uint64_t cap);
/* Open-loop load: submit qs[0..n) at Poisson arrival times of rate qps (an
 * arrival thread in the library), wait for all, write per-request latencies
 * (ms) and the request ids. */
wg_status wg_serve_replay(wg_serve* s, const wg_serve_query* qs, int32_t n, double qps, uint64_t seed,
                          double* latency_ms, uint64_t* ids);
wg_status wg_serve_stats_get(wg_serve* s, wg_serve_stats* out);
/* Forget completed requests and reset the statistics. */
wg_status wg_serve_reset(wg_serve* s);

#ifdef __cplusplus
}
#endif
#endif /* WADJET_GPU_H */
