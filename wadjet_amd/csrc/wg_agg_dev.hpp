// wg_agg_dev.hpp — device side of the grouped aggregate: the per-row update
// program, the LDS pre-aggregation table, the global open-addressing table and
// the persistent consume kernel (plain rows, or joined rows from a fused probe).
//
// Reference semantics reproduced here (agg_scatter.go):
//   SUM(int32/int64) wrapping int64 + count          scatterSumInt      :178-215
//   SUM/AVG(float64), AVG(int64) as float64(v)       scatterSumFloat    :256-293, :525-530
//   SUM/AVG(DECIMAL) exact Int128                    scatterSumDecimal  :330-348
//   COUNT(col) non-null rows, COUNT(*) active rows   scatterCount(Star) :369-406
//   MIN/MAX int and float ("!has || v < min")        scatterMin/Max*    :408-510
#pragma once
#include "wg_core.hpp"

namespace wg {

struct AggProgram {
  int32_t n_ops;
  int32_t stride;  // words per slot (power of two), word 0 = key
  OpDesc ops[kMaxOps];
};

// Columns a row's values come from. For a plain aggregate every column is a
// batch column at `row`. For a fused join the first n_probe columns are probe
// columns (at the probe row) and the rest are build columns, read either from
// the payload words inlined in the join slot or from the retained build
// columns at the build row.
struct ColSet {
  int32_t n_cols;
  int32_t n_probe;
  ColDesc cols[kMaxCols];
  int32_t inline_word[kMaxCols];  // build columns: slot word holding the value, -1 = gather
};

struct RowRef {
  uint64_t prow;
  uint64_t brow;
  const uint64_t* bslot;  // join slot of the matched key (inline payload), or null
};

template <bool JOINED>
__device__ __forceinline__ bool rv_valid(const ColSet& cs, int c, const RowRef& r) {
  if (!JOINED || c < cs.n_probe) return col_valid(cs.cols[c], r.prow);
  return col_valid(cs.cols[c], r.brow);
}
template <bool JOINED>
__device__ __forceinline__ int64_t rv_int(const ColSet& cs, int c, const RowRef& r) {
  if (!JOINED || c < cs.n_probe) return load_int(cs.cols[c], r.prow);
  if (r.bslot != nullptr && cs.inline_word[c] >= 0) return (int64_t)r.bslot[cs.inline_word[c]];
  return load_int(cs.cols[c], r.brow);
}
template <bool JOINED>
__device__ __forceinline__ uint64_t rv_bits(const ColSet& cs, int c, const RowRef& r) {
  // raw 64-bit payload of a float64 column
  if (!JOINED || c < cs.n_probe) return ((const uint64_t*)cs.cols[c].data)[r.prow];
  if (r.bslot != nullptr && cs.inline_word[c] >= 0) return r.bslot[cs.inline_word[c]];
  return ((const uint64_t*)cs.cols[c].data)[r.brow];
}
// Int128 value of a DECIMAL128 column: {Hi, Lo} per row.
template <bool JOINED>
__device__ __forceinline__ void rv_i128(const ColSet& cs, int c, const RowRef& r, uint64_t& lo, int64_t& hi) {
  uint64_t row = (!JOINED || c < cs.n_probe) ? r.prow : r.brow;
  const uint64_t* d = (const uint64_t*)cs.cols[c].data;
  hi = (int64_t)d[2 * row];
  lo = d[2 * row + 1];
}

// ------------------------------------------------------------ atomics

// Scope of the target table: LDS (workgroup) or global (agent).
struct LdsT {};
struct GlbT {};

template <typename S> struct Scope;
template <> struct Scope<LdsT> { static constexpr int v = __HIP_MEMORY_SCOPE_WORKGROUP; };
template <> struct Scope<GlbT> { static constexpr int v = __HIP_MEMORY_SCOPE_AGENT; };

// 64-bit compare-and-swap returning the previous value.
template <typename S>
__device__ __forceinline__ uint64_t at_cas(uint64_t* p, uint64_t expected, uint64_t desired) {
  __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED, Scope<S>::v);
  return expected;
}
template <typename S>
__device__ __forceinline__ void at_add(uint64_t* p, uint64_t v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, Scope<S>::v);
}
template <typename S>
__device__ __forceinline__ void at_addf(uint64_t* p, double v) {
  __hip_atomic_fetch_add((double*)p, v, __ATOMIC_RELAXED, Scope<S>::v);
}
template <typename S>
__device__ __forceinline__ void at_min(uint64_t* p, uint64_t v) {
  __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, Scope<S>::v);
}
template <typename S>
__device__ __forceinline__ void at_max(uint64_t* p, uint64_t v) {
  __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, Scope<S>::v);
}

// Add a signed 128-bit value as four carry-save limbs (32-bit pieces; the top
// limb signed). Each limb word absorbs < 2^32 additions without overflow.
template <typename S>
__device__ __forceinline__ void add_limbs4(uint64_t* w, uint64_t lo, int64_t hi) {
  at_add<S>(w + 0, lo & 0xFFFFFFFFull);
  at_add<S>(w + 1, lo >> 32);
  at_add<S>(w + 2, (uint64_t)hi & 0xFFFFFFFFull);
  at_add<S>(w + 3, (uint64_t)(int64_t)(hi >> 32));
}
template <typename S>
__device__ __forceinline__ void add_limbs2(uint64_t* w, int64_t x) {
  at_add<S>(w + 0, (uint64_t)x & 0xFFFFFFFFull);
  at_add<S>(w + 1, (uint64_t)(x >> 32));  // arithmetic shift: signed high limb
}

// ------------------------------------------------------ row update program

// Apply one row to a group slot (LDS or global). `w` points at word 0 (key).
template <typename S, bool JOINED>
__device__ __forceinline__ void apply_row(uint64_t* w, const AggProgram& pg, const ColSet& cs,
                                          const RowRef& r) {
  for (int o = 0; o < pg.n_ops; ++o) {
    const OpDesc op = pg.ops[o];
    uint64_t* t = w + op.word;
    switch (op.kind) {
      case OP_ROWCOUNT:
        at_add<S>(t, 1);
        break;
      case OP_NNCOUNT:
        if (rv_valid<JOINED>(cs, op.col, r) && (op.col2 < 0 || rv_valid<JOINED>(cs, op.col2, r)))
          at_add<S>(t, 1);
        break;
      case OP_SUMI:
        if (rv_valid<JOINED>(cs, op.col, r) && (op.col2 < 0 || rv_valid<JOINED>(cs, op.col2, r))) {
          int64_t v = rv_int<JOINED>(cs, op.col, r);
          if (op.col2 >= 0) v = (int64_t)((uint64_t)v * (uint64_t)rv_int<JOINED>(cs, op.col2, r));
          at_add<S>(t, (uint64_t)v);
        }
        break;
      case OP_SUMF:
        if (rv_valid<JOINED>(cs, op.col, r)) {
          double v;
          if (cs.cols[op.col].type == WG_FLOAT64) v = __longlong_as_double((long long)rv_bits<JOINED>(cs, op.col, r));
          else v = (double)rv_int<JOINED>(cs, op.col, r);
          at_addf<S>(t, v);
        }
        break;
      case OP_SUMD:
        if (rv_valid<JOINED>(cs, op.col, r)) add_limbs2<S>(t, rv_int<JOINED>(cs, op.col, r));
        break;
      case OP_SUMD4:
        if (rv_valid<JOINED>(cs, op.col, r) && (op.col2 < 0 || rv_valid<JOINED>(cs, op.col2, r))) {
          uint64_t lo;
          int64_t hi;
          if (op.col2 >= 0) {
            // exact signed 64 x 64 -> 128 product of the two factors
            int64_t a = rv_int<JOINED>(cs, op.col, r);
            int64_t b = rv_int<JOINED>(cs, op.col2, r);
            lo = (uint64_t)a * (uint64_t)b;
            hi = (int64_t)__mul64hi((long long)a, (long long)b);
          } else if (cs.cols[op.col].type == WG_DECIMAL128) {
            rv_i128<JOINED>(cs, op.col, r, lo, hi);
          } else {
            int64_t a = rv_int<JOINED>(cs, op.col, r);
            lo = (uint64_t)a;
            hi = a < 0 ? -1 : 0;
          }
          add_limbs4<S>(t, lo, hi);
        }
        break;
      case OP_MIN:
      case OP_MAX:
        if (rv_valid<JOINED>(cs, op.col, r)) {
          uint64_t e;
          bool ok = true;
          if (cs.cols[op.col].type == WG_FLOAT64) {
            uint64_t b = rv_bits<JOINED>(cs, op.col, r);
            ok = (b & 0x7FFFFFFFFFFFFFFFull) <= 0x7FF0000000000000ull;  // NaN never wins
            e = enc_f64_bits(b);
          } else {
            e = enc_i64(rv_int<JOINED>(cs, op.col, r));
          }
          if (ok) {
            if (op.kind == OP_MIN) at_min<S>(t, e);
            else at_max<S>(t, e);
          }
        }
        break;
      default:
        break;
    }
  }
}

// Fold a partial slot (LDS entry or overflow record) into a global slot.
__device__ __forceinline__ void merge_partial(uint64_t* dst, const uint64_t* src, const AggProgram& pg) {
  for (int o = 0; o < pg.n_ops; ++o) {
    const OpDesc op = pg.ops[o];
    uint64_t* t = dst + op.word;
    const uint64_t* s = src + op.word;
    switch (op.kind) {
      case OP_ROWCOUNT: case OP_NNCOUNT: case OP_SUMI:
        if (s[0]) at_add<GlbT>(t, s[0]);
        break;
      case OP_SUMF:
        if (s[0]) at_addf<GlbT>(t, __longlong_as_double((long long)s[0]));
        break;
      case OP_SUMD:
        if (s[0]) at_add<GlbT>(t, s[0]);
        if (s[1]) at_add<GlbT>(t + 1, s[1]);
        break;
      case OP_SUMD4:
        for (int k = 0; k < 4; ++k)
          if (s[k]) at_add<GlbT>(t + k, s[k]);
        break;
      case OP_MIN:
        if (s[0] != kMinInit) at_min<GlbT>(t, s[0]);
        break;
      case OP_MAX:
        if (s[0] != kMaxInit) at_max<GlbT>(t, s[0]);
        break;
      default:
        break;
    }
  }
}

// Fold a partial slot into a global slot this workgroup owns exclusively
// (no other workgroup touches the key during the kernel): plain read-modify-write.
__device__ __forceinline__ void merge_partial_owned(uint64_t* dst, const uint64_t* src, const AggProgram& pg) {
  for (int o = 0; o < pg.n_ops; ++o) {
    const OpDesc op = pg.ops[o];
    uint64_t* t = dst + op.word;
    const uint64_t* s = src + op.word;
    switch (op.kind) {
      case OP_ROWCOUNT: case OP_NNCOUNT: case OP_SUMI: t[0] += s[0]; break;
      case OP_SUMF:
        t[0] = (uint64_t)__double_as_longlong(__longlong_as_double((long long)t[0]) +
                                              __longlong_as_double((long long)s[0]));
        break;
      case OP_SUMD: t[0] += s[0]; t[1] += s[1]; break;
      case OP_SUMD4: for (int k = 0; k < 4; ++k) t[k] += s[k]; break;
      case OP_MIN: if (s[0] < t[0]) t[0] = s[0]; break;
      case OP_MAX: if (s[0] > t[0]) t[0] = s[0]; break;
      default: break;
    }
  }
}

// Initial value of every word of an empty slot.
__host__ __device__ inline uint64_t slot_init_word(const AggProgram& pg, int word) {
  if (word == 0) return kEmpty;
  for (int o = 0; o < pg.n_ops; ++o) {
    if (pg.ops[o].word == word) {
      if (pg.ops[o].kind == OP_MIN) return kMinInit;
      if (pg.ops[o].kind == OP_MAX) return kMaxInit;
    }
  }
  return 0;
}

// ------------------------------------------------------------- the tables

struct GlobalTable {
  uint64_t* slots;      // cap * stride words, then kSideSlots side slots
  uint64_t mask;        // cap - 1
  uint64_t* counters;   // see Ctr
  uint64_t max_probe;   // a claim farther than this from the key's home slot overflows
};

enum Ctr : int {
  CTR_GROUPS = 0,    // slots claimed (summed once per workgroup)
  CTR_OVF_ROWS = 1,  // rows appended to the overflow row list
  CTR_OVF_PART = 2,  // partial records appended to the overflow partial list
  CTR_TILE = 3,      // next tile (dynamic scheduling)
  CTR_STOPPED = 4,   // a workgroup stopped for overflow headroom
  CTR_LDS_ROWS = 5,
  CTR_GLB_ROWS = 6,
  CTR_FLUSHED = 7,
  CTR_N = 8
};

// Linear probing stops after this many slots; a table that needs more is
// past its load budget and the host grows it (the claim reports overflow).
constexpr uint64_t kMaxProbe = 512;

__device__ __forceinline__ uint64_t key_hash(uint64_t key) { return mix64(key); }

// Find the slot of `key`, claiming a free one with a 64-bit CAS. Returns the
// slot pointer (claimed = 1 when this call took a free slot), or nullptr when
// no free slot lies within max_probe (the caller records the row or partial
// as overflow). Probe reads are workgroup-scope loads: a stale EMPTY is
// resolved by the CAS, and a slot never changes once it holds a key.
// No shared counter is touched here: callers sum `claimed` per workgroup.
__device__ __forceinline__ uint64_t* glb_find_or_claim(const GlobalTable& T, int stride, uint64_t key, uint64_t h,
                                                      uint32_t& claimed) {
  uint64_t idx = h & T.mask;
  const uint64_t lim = T.max_probe < T.mask + 1 ? T.max_probe : T.mask + 1;
  for (uint64_t p = 0; p < lim; ++p) {
    uint64_t* s = T.slots + idx * (uint64_t)stride;
    const uint64_t k = __hip_atomic_load(s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == key) return s;
    if (k == kEmpty) {
      const uint64_t prev = at_cas<GlbT>(s, kEmpty, key);
      if (prev == kEmpty) {
        ++claimed;
        return s;
      }
      if (prev == key) return s;
    }
    idx = (idx + 1) & T.mask;
  }
  return nullptr;
}

// Sum a per-thread count over the workgroup and add it to *dst with one
// atomic (a per-row atomic on one shared address serialises the whole grid).
__device__ __forceinline__ void block_add(uint64_t* dst, uint64_t v) {
  __shared__ unsigned long long s_sum;
  __syncthreads();
  if (threadIdx.x == 0) s_sum = 0;
  __syncthreads();
  if (v) atomicAdd(&s_sum, (unsigned long long)v);
  __syncthreads();
  if (threadIdx.x == 0 && s_sum) at_add<GlbT>(dst, s_sum);
}

// Per-workgroup LDS table: `lds_slots` slots of `stride` words.
constexpr int kLdsProbes = 8;

__device__ __forceinline__ uint64_t* lds_find_or_claim(uint64_t* lds, int lds_mask, int stride, uint64_t key,
                                                      uint64_t h, int shift = 40) {
  uint32_t idx = (uint32_t)(h >> shift) & (uint32_t)lds_mask;
  for (int p = 0; p < kLdsProbes; ++p) {
    uint64_t* s = lds + (uint64_t)idx * stride;
    uint64_t k = __hip_atomic_load(s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == key) return s;
    if (k == kEmpty) {
      uint64_t prev = at_cas<LdsT>(s, kEmpty, key);
      if (prev == kEmpty || prev == key) return s;
    }
    idx = (idx + 1) & (uint32_t)lds_mask;
  }
  return nullptr;
}

// ----------------------------------------------------------- key building

// Build the packed key of a (plain) row. Returns false for a NULL key; for a
// single-column key that is the NULL group (side slot 1), for multi-column
// keys the host rejected nullable key columns up front.
template <bool JOINED>
__device__ __forceinline__ bool build_key(const KeyDesc& K, const ColSet& cs, const RowRef& r, uint64_t& key) {
  if (K.n == 1) {
    if (!rv_valid<JOINED>(cs, K.col[0], r)) return false;
    key = (uint64_t)rv_int<JOINED>(cs, K.col[0], r);
    return true;
  }
  uint64_t k = 0;
  for (int i = 0; i < K.n; ++i) {
    uint64_t v = (uint64_t)rv_int<JOINED>(cs, K.col[i], r);
    if (K.narrow[i]) v = (v & 0xFFFFFFFFull) << K.shift[i];
    k |= v;
  }
  key = k;
  return true;
}

}  // namespace wg
