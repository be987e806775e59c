/*
 * oracle.h — CPU restatement of wadjet's HashAggregate / HashJoin semantics.
 *
 * TEST INFRASTRUCTURE ONLY. This code is the checker for the HIP library: it
 * is called from tests/, from __graft_entry__.smoke() and from bench.py's CPU
 * baseline leg, never from the product path.
 *
 * Rows are processed in input order, with the reference's accumulation order,
 * so float64 sums equal the reference's serial result bit for bit.
 */
#ifndef WADJET_ORACLE_H
#define WADJET_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* column types: same numbering as wg_type in include/wadjet_gpu.h */
enum { OR_INT32 = 1, OR_INT64 = 2, OR_FLOAT64 = 3, OR_DECIMAL = 4, OR_DECIMAL128 = 5, OR_IPV4 = 6, OR_UINT32 = 7,
       OR_UINT16 = 8 };
enum { OR_SUM = 0, OR_COUNT = 1, OR_MIN = 2, OR_MAX = 3, OR_AVG = 4 };

typedef struct or_col {
  const void* data;
  const uint64_t* validity; /* batch.Bitmap words, NULL = no nulls */
  int32_t type;
  int32_t scale;
} or_col;

/* One group's kernel.Accumulator (internal/engine/exec/kernel/types.go:21-40). */
typedef struct or_acc {
  int64_t count;
  int64_t sum_i64;
  double sum_f64;
  uint64_t sum_lo;
  int64_t sum_hi;
  int64_t min_i64, max_i64;
  double min_f64, max_f64;
  int32_t has_min, has_max;
  int32_t is_float, is_decimal;
} or_acc;

/* Int128 (decimal.go) */
void or_i128_from_i64(int64_t v, uint64_t* lo, int64_t* hi);
void or_i128_add(uint64_t alo, int64_t ahi, uint64_t blo, int64_t bhi, uint64_t* lo, int64_t* hi);
void or_i128_neg(uint64_t lo, int64_t hi, uint64_t* olo, int64_t* ohi);
int or_i128_less(uint64_t alo, int64_t ahi, uint64_t blo, int64_t bhi);
double or_i128_to_f64(uint64_t lo, int64_t hi, int scale);
void or_i128_from_f64(double f, int scale, uint64_t* lo, int64_t* hi);
int or_i128_format(uint64_t lo, int64_t hi, int scale, char* buf, int cap);
int64_t or_parse_decimal(const char* s, int scale);

/* Grouped aggregation. Returns the number of groups (first-seen order) or -1.
 * keys_out[g*n_keys + k] = widened key value; key_null_out[g*n_keys + k].
 * acc_out[g*n_aggs + a]. max_groups bounds the outputs. */
int64_t or_hash_agg(const or_col* cols, int32_t n_cols, uint64_t n_rows, const uint32_t* sel, uint64_t n_sel,
                    int32_t pred_col, int32_t pred_op, int64_t pred_i64, double pred_f64, const int32_t* key_cols,
                    int32_t n_keys, const int32_t* agg_func, const int32_t* agg_col, int32_t n_aggs,
                    int64_t* keys_out, uint8_t* key_null_out, or_acc* acc_out, int64_t max_groups);

/* Finalized value (writeAccToColumn, aggregate_partial_spill.go:1855-1945):
 * returns 1 if valid; *f / *i hold the float64 or int64 output. */
int or_finalize(const or_acc* acc, int32_t func, int32_t in_type, int32_t scale, double* f, int64_t* i);

/* KernelFilter: ascending rows passing `col op constant` (NULL never passes). */
uint64_t or_filter(const or_col* col, uint64_t n_rows, const uint32_t* sel, uint64_t n_sel, int32_t op, int64_t ci,
                   double cf, uint32_t* out);

/* Inner hash join on one int-class key (HashJoin.Build + inlineIntProbe).
 * Emits (probe_row, build_row) pairs in the reference's order: probe order,
 * chain order LIFO. Returns the number of pairs (<= cap) or -1 on overflow. */
int64_t or_hash_join(const or_col* build_key, uint64_t n_build, const or_col* probe_key, uint64_t n_probe,
                     const uint32_t* probe_sel, uint64_t n_probe_sel, uint32_t* out_probe, uint32_t* out_build,
                     int64_t cap);
/* Semi / anti join probe (join.go:3245-3310, :4486-4650): probe positions in
 * input order whose key is (anti = 0) / is not (anti = 1) among the non-NULL
 * build keys; NULL probe keys never match. Returns the count written. */
int64_t or_semi_join(const or_col* build_key, uint64_t n_build, const or_col* probe_key, uint64_t n_probe,
                     const uint32_t* probe_sel, uint64_t n_probe_sel, int anti, uint32_t* out);

/* Order-free digest of a multiset of int64 tuples (rows of `width` values):
 * d[0] = sum mix(tuple; seed 1), d[1] = sum mix(tuple; seed 2), mod 2^64. */
void or_digest_rows(const int64_t* rows, uint64_t n, int32_t width, uint64_t d[2]);
uint64_t or_mix64(uint64_t k);

/* Multi-threaded partial aggregation + merge over `threads` threads, used as
 * the host CPU baseline (the morsel-parallel CloneSink/MergeSink shape). */
int64_t or_hash_agg_parallel(const or_col* cols, int32_t n_cols, uint64_t n_rows, int32_t pred_col, int32_t pred_op,
                             int64_t pred_i64, double pred_f64, const int32_t* key_cols, int32_t n_keys,
                             const int32_t* agg_func, const int32_t* agg_col, int32_t n_aggs, int32_t threads,
                             int64_t* keys_out, or_acc* acc_out, int64_t max_groups);

/* Inner join on an int key, then GROUP BY an int build column
 * SUM(probe int column x build DECIMAL column) exactly in Int128, COUNT(*),
 * with the probe rows split over `threads` threads (partials merged in range
 * order). Groups ascending; returns their count, -1 past max_groups. */
int64_t or_join_agg_parallel(const or_col* build_key, const or_col* build_group, const or_col* build_dec,
                             uint64_t n_build, const or_col* probe_key, const or_col* probe_mul, uint64_t n_probe,
                             int32_t threads, int64_t* keys_out, int64_t* count_out, uint64_t* sum_lo_out,
                             int64_t* sum_hi_out, int64_t max_groups);

#ifdef __cplusplus
}
#endif
#endif
