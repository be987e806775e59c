// wg_core.hpp — shared host/device pieces of the MI355X hash-aggregate /
// hash-join library: status handling, the context, column access, key
// packing and hashing, the accumulator-word program and the table layouts.
//
// Layout decisions (gfx950, 64-lane waves, 128-byte L2 lines, 160 KiB LDS/CU):
//  * Group table = open addressing, one array-of-structs slot per group:
//      word 0            packed 64-bit key (EMPTY sentinel = free slot)
//      words 1..W-1      accumulator words (counts, sums, min/max) of the group
//    Slot stride is a power of two (<= 128 B) so a slot never straddles an L2
//    line: a random row touches ONE line for probe + all updates, instead of
//    the 1 + n_fields lines of the reference's SoA arrays (agg_scatter.go:16-47).
//  * The key value equal to the EMPTY sentinel and the NULL key (single-column
//    keys) live in two side slots after the table, so every 64-bit key is
//    storable (the reference's intHashEmpty key is not, int_hash.go:42).
//  * Accumulator words are updated with fire-and-forget atomics (no return
//    value), so a row costs one probe load plus non-blocking atomics.
//  * DECIMAL sums are carry-save limbs: each row adds its low 32 bits
//    (unsigned) and its high 32 bits (signed) to two 64-bit words. Neither word
//    can overflow before 2^32 additions, so the Int128 value
//    limb_hi * 2^32 + limb_lo is exact; the host folds the limbs into an exact
//    Int128 base (side array) before that bound and on merges. This is
//    batch.Int128.Add (decimal.go:77-85) with the carry deferred.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "wg_host.hpp"

namespace wg {

constexpr int kMaxCols = 16;
constexpr int kMaxKeys = 4;
constexpr int kMaxOps = 24;
constexpr uint64_t kEmpty = 0x8000000000000001ull;  // == int64 -0x7FFFFFFFFFFFFFFF
constexpr int kSideSlots = 2;                        // [0] key==kEmpty, [1] NULL key

// ----------------------------------------------------------- column access

struct ColDesc {
  const void* data;
  const uint64_t* valid;
  int32_t type;
  int32_t scale;
};

__host__ __device__ inline int type_bytes(int t) {
  switch (t) {
    case WG_INT32: case WG_UINT32: return 4;
    case WG_UINT16: return 2;
    case WG_DECIMAL128: return 16;
    default: return 8;
  }
}
__host__ __device__ inline bool is_int_class(int t) {
  return t == WG_INT32 || t == WG_INT64 || t == WG_IPV4 || t == WG_UINT32 || t == WG_UINT16;
}
// Bits a group column occupies in the 64-bit packed key. int32-class columns
// take 32 bits as in packedKeyWidth (packed_hash.go:360-371); the compact
// uint32/uint16 forms hold values < 2^32 and also take 32 bits (zero-extended
// on unpack). The packed layout is table-internal: only the key VALUES are
// part of the result contract.
__host__ __device__ inline int packed_width(int t) {
  switch (t) {
    case WG_INT32: case WG_UINT16: case WG_UINT32: return 4;
    case WG_INT64: case WG_IPV4: return 8;
    default: return 0;
  }
}

__device__ __forceinline__ bool col_valid(const ColDesc& c, uint64_t row) {
  return c.valid == nullptr || ((c.valid[row >> 6] >> (row & 63)) & 1ull);
}

// Integer value widened to int64 exactly as the reference's consume loops
// widen it (aggregate.go:1910-1913: int32 sign-extends; IPv4 is int64 storage).
__device__ __forceinline__ int64_t load_int(const ColDesc& c, uint64_t row) {
  switch (c.type) {
    case WG_INT32: return (int64_t)((const int32_t*)c.data)[row];
    case WG_UINT32: return (int64_t)((const uint32_t*)c.data)[row];
    case WG_UINT16: return (int64_t)((const uint16_t*)c.data)[row];
    case WG_DECIMAL128: return (int64_t)((const uint64_t*)c.data)[2 * row + 1];
    default: return ((const int64_t*)c.data)[row];
  }
}
__device__ __forceinline__ double load_f64(const ColDesc& c, uint64_t row) {
  return ((const double*)c.data)[row];
}

// ------------------------------------------------------------------ hashing

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t k) {
  // murmur3 fmix64: a bijection on 64-bit words with full avalanche.
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

// ----------------------------------------------------------- key packing

// One group column's place in the 64-bit packed key; same placement rule as
// buildPackedLayout (packed_hash.go:384-425) restricted to <= 8 bytes total:
// an 8-byte column takes the whole word, 4-byte columns take the low then the
// high half. A single column is the whole word (widened int64).
struct KeyDesc {
  int32_t n;
  int32_t col[kMaxKeys];
  uint8_t shift[kMaxKeys];
  uint8_t narrow[kMaxKeys];  // 0 = whole word, 1 = 32 bits sign-extended, 2 = 32 bits zero-extended
  uint8_t word[kMaxKeys];    // keys wider than 64 bits: 0 = lo, 1 = hi word of the {lo, hi} pair
  int32_t key_type[kMaxKeys];
};

// --------------------------------------------------------- accumulator ops

enum OpKind : int32_t {
  OP_ROWCOUNT = 1,  // COUNT(*) / group row count: +1 per row
  OP_NNCOUNT = 2,   // non-null count of col (and col2 for a product)
  OP_SUMI = 3,      // int64 wrapping sum (SUM int32/int64, AVG int32)
  OP_SUMF = 4,      // float64 sum (SUM/AVG float64, AVG int64 as float64(v))
  OP_SUMD = 5,      // DECIMAL sum, 2 carry-save limbs (int64-storage decimal)
  OP_SUMD4 = 6,     // DECIMAL sum, 4 carry-save limbs (Int128 storage or products)
  OP_MIN = 7,       // ordered-u64 min (int biased, float ordered); NaN skipped
  OP_MAX = 8,
};

struct OpDesc {
  int32_t kind;
  int32_t col;   // input column (joined-row index for fused joins)
  int32_t col2;  // second factor for products, -1 otherwise
  int32_t word;  // first accumulator word in the slot
};

// Order-preserving map of a value to uint64 for MIN/MAX atomics.
__host__ __device__ __forceinline__ uint64_t enc_i64(int64_t v) { return (uint64_t)v ^ 0x8000000000000000ull; }
__host__ __device__ __forceinline__ int64_t dec_i64(uint64_t e) { return (int64_t)(e ^ 0x8000000000000000ull); }
__host__ __device__ __forceinline__ uint64_t enc_f64_bits(uint64_t b) {
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ __forceinline__ uint64_t dec_f64_bits(uint64_t e) {
  return (e >> 63) ? (e & 0x7FFFFFFFFFFFFFFFull) : ~e;
}
constexpr uint64_t kMinInit = ~0ull;  // above every encoded value
constexpr uint64_t kMaxInit = 0ull;   // below every encoded value

// --------------------------------------------------------------- predicate

struct PredDesc {
  int32_t col;  // -1 = none
  int32_t op;
  int64_t i64;
  double f64;
};

template <typename T>
__device__ __forceinline__ bool cmp_op(T a, T b, int op) {
  switch (op) {
    case WG_EQ: return a == b;
    case WG_NE: return a != b;
    case WG_LT: return a < b;
    case WG_LE: return a <= b;
    case WG_GT: return a > b;
    default: return a >= b;
  }
}

// Typed compare against the constant the way ResolveFilterKernel resolves it
// (kernel/compare.go:145-178): int32 columns compare as int32(constant).
__device__ __forceinline__ bool pred_pass(const ColDesc& c, uint64_t row, int op, int64_t ci, double cf) {
  if (!col_valid(c, row)) return false;
  switch (c.type) {
    case WG_FLOAT64: return cmp_op<double>(load_f64(c, row), cf, op);
    case WG_INT32: case WG_UINT16: return cmp_op<int32_t>((int32_t)load_int(c, row), (int32_t)ci, op);
    default: return cmp_op<int64_t>(load_int(c, row), ci, op);
  }
}

// ---------------------------------------------------------------- helpers

// Replace *p (room for *cap elements) by a buffer of `want` elements. The new
// buffer is allocated before the old one is freed, so a failed allocation
// leaves both *p and *cap as they were. The contents are not kept.
template <typename T, typename C>
wg_status regrow(wg_ctx* ctx, T** p, C* cap, uint64_t want, uint64_t elem_bytes, const char* what) {
  T* fresh = nullptr;
  WG_TRY(dev_alloc(ctx, (void**)&fresh, want * elem_bytes, nullptr, what));
  if (*p) (void)hipFree(*p);
  *p = fresh;
  *cap = (C)want;
  return WG_OK;
}

inline uint64_t next_pow2(uint64_t v) {
  uint64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

}  // namespace wg
