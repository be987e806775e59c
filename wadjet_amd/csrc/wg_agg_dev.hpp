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
  // ops are sorted by kind: those of kind k are ops[kind_end[k - 1], kind_end[k])
  int32_t kind_end[OP_MAX + 1];
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
  uint64_t pay2, pay3;    // with bslot: slot words 2 and 3, already loaded
};

// Inline payload word `w` of the matched join slot. Words 2 and 3 come from
// registers (scalar fields, not an array, so nothing is placed in scratch);
// `w` is uniform, so this is a scalar branch.
__device__ __forceinline__ uint64_t slot_payload(const RowRef& r, int w) {
  switch (w) {
    case 2: return r.pay2;
    case 3: return r.pay3;
    default: return r.bslot[w];
  }
}

template <bool JOINED>
__device__ __forceinline__ bool rv_valid(const ColSet& cs, int c, const RowRef& r) {
  if (!JOINED || c < cs.n_probe) return col_valid(cs.cols[c], r.prow);
  return col_valid(cs.cols[c], r.brow);
}
template <bool JOINED>
__device__ __forceinline__ int64_t rv_int(const ColSet& cs, int c, const RowRef& r) {
  if (!JOINED || c < cs.n_probe) return load_int(cs.cols[c], r.prow);
  if (r.bslot != nullptr && cs.inline_word[c] >= 0) return (int64_t)slot_payload(r, cs.inline_word[c]);
  return load_int(cs.cols[c], r.brow);
}
template <bool JOINED>
__device__ __forceinline__ uint64_t rv_bits(const ColSet& cs, int c, const RowRef& r) {
  // raw 64-bit payload of a float64 column
  if (!JOINED || c < cs.n_probe) return ((const uint64_t*)cs.cols[c].data)[r.prow];
  if (r.bslot != nullptr && cs.inline_word[c] >= 0) return slot_payload(r, cs.inline_word[c]);
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

// Exact Int128 value of a limb group (2 or 4 carry-save limbs).
__host__ __device__ inline void limbs_value(const uint64_t* l, int nl, uint64_t& lo, int64_t& hi) {
  // value = l0 + l1*2^32 (+ l2*2^64 + l3*2^96), limbs read as signed where
  // they may carry negative sums (the top limb), unsigned otherwise.
  unsigned __int128 acc = (unsigned __int128)l[0];
  if (nl == 2) {
    acc += (unsigned __int128)(__int128)(int64_t)l[1] << 32;
  } else {
    acc += (unsigned __int128)l[1] << 32;
    acc += (unsigned __int128)l[2] << 64;
    acc += (unsigned __int128)(__int128)(int64_t)l[3] << 96;
  }
  lo = (uint64_t)acc;
  hi = (int64_t)(uint64_t)(acc >> 64);
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
// `ws` is the distance between consecutive words of a slot: 1 in the global
// table (array of structs), the slot count in an LDS table (struct of arrays).
template <typename S>
__device__ __forceinline__ void add_limbs4(uint64_t* w, int ws, uint64_t lo, int64_t hi) {
  at_add<S>(w, lo & 0xFFFFFFFFull);
  at_add<S>(w + ws, lo >> 32);
  at_add<S>(w + 2 * ws, (uint64_t)hi & 0xFFFFFFFFull);
  at_add<S>(w + 3 * ws, (uint64_t)(int64_t)(hi >> 32));
}
template <typename S>
__device__ __forceinline__ void add_limbs2(uint64_t* w, int ws, int64_t x) {
  at_add<S>(w, (uint64_t)x & 0xFFFFFFFFull);
  at_add<S>(w + ws, (uint64_t)(x >> 32));  // arithmetic shift: signed high limb
}

// ------------------------------------------------------ row update program

// Apply one row to a group slot (LDS or global). `w` points at word 0 (key);
// word k is at w[k * ws].
template <typename S, bool JOINED>
__device__ __forceinline__ void apply_row(uint64_t* w, int ws, const AggProgram& pg, const ColSet& cs,
                                          const RowRef& r) {
  for (int o = 0; o < pg.n_ops; ++o) {
    const OpDesc op = pg.ops[o];
    uint64_t* t = w + op.word * ws;
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
        if (rv_valid<JOINED>(cs, op.col, r)) add_limbs2<S>(t, ws, rv_int<JOINED>(cs, op.col, r));
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
          add_limbs4<S>(t, ws, lo, hi);
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

// Fold a partial slot (LDS entry with word stride `sws`, or an overflow
// record with sws = 1) into a global slot (S = GlbT) or an LDS slot (LdsT).
template <typename S = GlbT>
__device__ __forceinline__ void merge_partial(uint64_t* dst, const uint64_t* src, int sws, const AggProgram& pg) {
  for (int o = 0; o < pg.n_ops; ++o) {
    const OpDesc op = pg.ops[o];
    uint64_t* t = dst + op.word;
    const uint64_t* s = src + op.word * sws;
    switch (op.kind) {
      case OP_ROWCOUNT: case OP_NNCOUNT: case OP_SUMI:
        if (s[0]) at_add<S>(t, s[0]);
        break;
      case OP_SUMF:
        if (s[0]) at_addf<S>(t, __longlong_as_double((long long)s[0]));
        break;
      case OP_SUMD:
        if (s[0]) at_add<S>(t, s[0]);
        if (s[sws]) at_add<S>(t + 1, s[sws]);
        break;
      case OP_SUMD4:
        for (int k = 0; k < 4; ++k)
          if (s[k * sws]) at_add<S>(t + k, s[k * sws]);
        break;
      case OP_MIN:
        if (s[0] != kMinInit) at_min<S>(t, s[0]);
        break;
      case OP_MAX:
        if (s[0] != kMaxInit) at_max<S>(t, s[0]);
        break;
      default:
        break;
    }
  }
}

// Fold a partial slot into a global slot this workgroup owns exclusively
// (no other workgroup touches the key during the kernel): plain read-modify-write.
__device__ __forceinline__ void merge_partial_owned(uint64_t* dst, const uint64_t* src, int sws,
                                                    const AggProgram& pg) {
  for (int o = 0; o < pg.n_ops; ++o) {
    const OpDesc op = pg.ops[o];
    uint64_t* t = dst + op.word;
    const uint64_t* s = src + op.word * sws;
    switch (op.kind) {
      case OP_ROWCOUNT: case OP_NNCOUNT: case OP_SUMI: t[0] += s[0]; break;
      case OP_SUMF:
        t[0] = (uint64_t)__double_as_longlong(__longlong_as_double((long long)t[0]) +
                                              __longlong_as_double((long long)s[0]));
        break;
      case OP_SUMD: t[0] += s[0]; t[1] += s[sws]; break;
      case OP_SUMD4: for (int k = 0; k < 4; ++k) t[k] += s[k * sws]; break;
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

// Fill an LDS table (struct of arrays: word k of slot i at lds[k * LW + i],
// LW >= L slots) with the empty-slot words.
__device__ __forceinline__ void lds_table_init(uint64_t* lds, int LW, const AggProgram& pg) {
  for (int w = 0; w < pg.stride; ++w) {
    const uint64_t v = slot_init_word(pg, w);
    for (int i = threadIdx.x; i < LW; i += blockDim.x) lds[(uint64_t)w * LW + i] = v;
  }
}

// An LDS table of L hashed slots is followed by kTrashSlots spare slots, one
// per lane: rows that do not aggregate in LDS still run the same (identity)
// LDS atomics on their lane's spare slot, so the common path has no per-row
// branch. Word stride of such a table: L + kTrashSlots.
constexpr int kTrashSlots = 64;
__host__ __device__ inline int lds_word_stride(int L) { return L + kTrashSlots; }
__host__ __device__ inline size_t lds_table_bytes(int L, int stride) {
  return (size_t)lds_word_stride(L) * stride * 8;
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
// device counter words per aggregate handle: the CTR_N counters, then the
// dense-output row count (outside the range the consume passes clear) and the
// partition aggregate's work-item cursor, so that one copy reads them all back
constexpr int kCtrDense = CTR_N;
constexpr int kCtrItem = CTR_N + 1;
constexpr int kAggCounterWords = CTR_N + 2;
// pinned words per aggregate handle: the counter words read back
constexpr int kAggPinnedWords = CTR_N + 8;
static_assert(kAggPinnedWords >= kAggCounterWords, "every counter word has a pinned word");

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
// Linear probing from slot `idx`, kProbeWindow slots per round trip: the key
// words of the window are loaded together, then the first slot that holds
// the key, or is free and is won by the CAS, is the answer (a CAS lost to
// another key moves on to the next slot of the window). At the table's load
// budget most probe runs end inside the first window, so a claim costs one
// load round trip and one CAS instead of one round trip per slot.
constexpr int kProbeWindow = 4;

__device__ __forceinline__ uint64_t* glb_probe_from(const GlobalTable& T, int stride, uint64_t key, uint64_t idx,
                                                   uint64_t first_p, bool& fresh) {
  const uint64_t lim = T.max_probe < T.mask + 1 ? T.max_probe : T.mask + 1;
  for (uint64_t p = first_p; p < lim; p += kProbeWindow) {
    uint64_t k[kProbeWindow];
#pragma unroll
    for (int j = 0; j < kProbeWindow; ++j)
      k[j] = __hip_atomic_load(T.slots + ((idx + j) & T.mask) * (uint64_t)stride, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    for (int j = 0; j < kProbeWindow && p + j < lim; ++j) {
      uint64_t* s = T.slots + ((idx + j) & T.mask) * (uint64_t)stride;
      if (k[j] == key) return s;
      if (k[j] == kEmpty) {
        const uint64_t prev = at_cas<GlbT>(s, kEmpty, key);
        if (prev == kEmpty) {
          fresh = true;
          return s;
        }
        if (prev == key) return s;
      }
    }
    idx = (idx + kProbeWindow) & T.mask;
  }
  return nullptr;
}

__device__ __forceinline__ uint64_t* glb_find_or_claim(const GlobalTable& T, int stride, uint64_t key, uint64_t h,
                                                      uint32_t& claimed) {
  bool fresh = false;
  uint64_t* s = glb_probe_from(T, stride, key, h & T.mask, 0, fresh);
  claimed += fresh ? 1u : 0u;
  return s;
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

// Per-workgroup LDS table of `lds_slots` slots, struct of arrays: word k of
// slot i is lds[k * lds_slots + i]. Random slots then spread over the LDS
// banks (a power-of-two slot stride would put every key in the same few).
constexpr int kLdsProbes = 16;  // a key beyond this run of probes aggregates in the global table

// Returns a pointer to the slot's key word (word k at ptr[k * lds_slots]).
__device__ __forceinline__ uint64_t* lds_find_or_claim(uint64_t* lds, int lds_mask, uint64_t key, uint64_t h,
                                                      int shift = 40, int max_probes = kLdsProbes) {
  uint32_t idx = (uint32_t)(h >> shift) & (uint32_t)lds_mask;
  for (int p = 0; p < max_probes; ++p) {
    uint64_t* s = lds + idx;
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

// The same claim with triangular probing (home, +1, +3, +6, ...: every slot of
// a power-of-two table once in lds_mask + 1 probes). Linear probing near 3/4
// load forms runs of occupied slots long enough to exhaust a probe window now
// and then; with triangular steps each probe lands in an independent place,
// so a window of max_probes slots is exhausted only when the table is full.
__device__ __forceinline__ uint64_t* lds_find_or_claim_tri(uint64_t* lds, int lds_mask, uint64_t key, uint64_t h,
                                                          int shift, int max_probes) {
  uint32_t idx = (uint32_t)(h >> shift) & (uint32_t)lds_mask;
  for (int p = 0; p < max_probes; ++p) {
    uint64_t* s = lds + idx;
    uint64_t k = __hip_atomic_load(s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == key) return s;
    if (k == kEmpty) {
      uint64_t prev = at_cas<LdsT>(s, kEmpty, key);
      if (prev == kEmpty || prev == key) return s;
    }
    idx = (idx + (uint32_t)p + 1) & (uint32_t)lds_mask;
  }
  return nullptr;
}

// ------------------------------------------------- batched row loading

constexpr int kMaxVals = 7;  // value columns a preloaded row carries
constexpr int kRowU = 4;     // rows per thread per step of the streaming loops (count pass)
constexpr int kRowU2 = 2;    // the same where more registers are live per row

// Raw 64-bit word of a column value: float64 bits, integers widened to int64.
__device__ __forceinline__ uint64_t raw_word(const ColDesc& c, uint64_t row) {
  if (c.type == WG_FLOAT64) return ((const uint64_t*)c.data)[row];
  return (uint64_t)load_int(c, row);
}

// A column value loaded with no branch on the column type. The type is
// uniform, but loads of different widths in the arms of a branch merge into
// one register with a copy, and the copy waits for the load right there: the
// loads of a batch of rows then run one at a time. Here every value takes the
// same three loads: the two 32-bit halves of a 4- or 8-byte element (the high
// half re-reads the low one for 4-byte types) and a 16-bit load for uint16;
// loads a type does not need read a fixed word in kSafeWord instead, so no
// access leaves the element or that word. value_of() assembles the value
// once every load of the batch is in flight.
__device__ const uint64_t kSafeWord = 0;

struct RawVal {
  uint32_t lo, hi, h16;
};

// Column data is device (global) memory: say so, so the selected pointers
// below still load with global (not flat) instructions.
typedef const __attribute__((address_space(1))) uint32_t* gptr_u32;
typedef const __attribute__((address_space(1))) uint16_t* gptr_u16;
typedef const __attribute__((address_space(1))) char* gptr_c;

// `active` is uniform; an inactive load (a column the query does not have)
// reads kSafeWord three times, so a batch's load sequence has no branches.
__device__ __forceinline__ RawVal load_raw(const ColDesc& c, uint64_t row, bool active = true) {
  const int t = c.type;
  const bool w2 = active && t == WG_UINT16;
  const bool w8 = active && !(t == WG_UINT16 || t == WG_INT32 || t == WG_UINT32);
  const uint64_t es = w2 ? 2 : w8 ? (t == WG_DECIMAL128 ? 16 : 8) : 4;
  const gptr_c e = (gptr_c)c.data + row * es + (t == WG_DECIMAL128 ? 8 : 0);  // DECIMAL128: the Lo word
  const gptr_u32 safe = (gptr_u32)&kSafeWord;
  const gptr_u32 plo = (w2 || !active) ? safe : (gptr_u32)e;
  const gptr_u32 phi = w8 ? plo + 1 : plo;
  const gptr_u16 p16 = w2 ? (gptr_u16)e : (gptr_u16)safe;
  return RawVal{*plo, *phi, *p16};
}

// A column resolved once per kernel for the streaming loops: the three loads
// of load_raw as base + row * stride each, with the type decisions taken here
// rather than per row (a load a type does not need reads kSafeWord with
// stride 0). Uniform values, so they live in scalar registers.
struct ColRes {
  const char* lo;
  const char* hi;
  const char* h16;
  uint32_t s_lo, s_hi, s16;
  const uint64_t* valid;
  int32_t type;
};

__device__ __forceinline__ ColRes resolve_col(const ColDesc& c, bool active) {
  const char* safe = (const char*)&kSafeWord;
  const char* d = (const char*)c.data;
  ColRes r{safe, safe, safe, 0, 0, 0, active ? c.valid : nullptr, c.type};
  if (!active) return r;
  switch (c.type) {
    case WG_UINT16: r.h16 = d; r.s16 = 2; break;
    case WG_INT32: case WG_UINT32: r.lo = r.hi = d; r.s_lo = r.s_hi = 4; break;
    case WG_DECIMAL128: r.lo = d + 8; r.hi = d + 12; r.s_lo = r.s_hi = 16; break;  // the Lo word
    default: r.lo = d; r.hi = d + 4; r.s_lo = r.s_hi = 8; break;
  }
  return r;
}

__device__ __forceinline__ RawVal load_res(const ColRes& c, uint64_t row) {
  return RawVal{*(gptr_u32)(c.lo + row * c.s_lo), *(gptr_u32)(c.hi + row * c.s_hi),
                *(gptr_u16)(c.h16 + row * c.s16)};
}
__device__ __forceinline__ uint64_t valid_res(const ColRes& c, uint64_t row) {
  return c.valid ? c.valid[row >> 6] : ~0ull;
}

// raw_word's form: float64 bits, integers widened to int64.
__device__ __forceinline__ uint64_t value_of(int type, const RawVal& r) {
  switch (type) {
    case WG_INT32: return (uint64_t)(int64_t)(int32_t)r.lo;
    case WG_UINT32: return r.lo;
    case WG_UINT16: return r.h16;
    default: return ((uint64_t)r.hi << 32) | r.lo;
  }
}

__device__ __forceinline__ uint64_t valid_word(const ColDesc& c, uint64_t row) {
  return c.valid ? c.valid[row >> 6] : ~0ull;
}

// pred_pass on a value already loaded (same typed comparison).
__device__ __forceinline__ bool pred_raw(const ColDesc& c, uint64_t raw, uint64_t vw, uint64_t row, int op,
                                         int64_t ci, double cf) {
  if (!((vw >> (row & 63)) & 1ull)) return false;
  switch (c.type) {
    case WG_FLOAT64: return cmp_op<double>(__longlong_as_double((long long)raw), cf, op);
    case WG_INT32: case WG_UINT16: return cmp_op<int32_t>((int32_t)(int64_t)raw, (int32_t)ci, op);
    default: return cmp_op<int64_t>((int64_t)raw, ci, op);
  }
}

// value_of for U rows of one column: the type switch once, outside the row loop.
template <int U>
__device__ __forceinline__ void values_of(int type, const RawVal (&r)[U], uint64_t (&out)[U]) {
  switch (type) {
    case WG_INT32:
#pragma unroll
      for (int u = 0; u < U; ++u) out[u] = (uint64_t)(int64_t)(int32_t)r[u].lo;
      break;
    case WG_UINT32:
#pragma unroll
      for (int u = 0; u < U; ++u) out[u] = r[u].lo;
      break;
    case WG_UINT16:
#pragma unroll
      for (int u = 0; u < U; ++u) out[u] = r[u].h16;
      break;
    default:
#pragma unroll
      for (int u = 0; u < U; ++u) out[u] = ((uint64_t)r[u].hi << 32) | r[u].lo;
      break;
  }
}

// pred_raw for U rows (bit u of the result): one dispatch on the column type
// and the operator, then a branch-free compare per row.
template <int U>
__device__ __forceinline__ uint32_t pred_mask(int type, const RawVal (&raw)[U], const uint64_t (&vw)[U],
                                              const uint64_t (&row)[U], int op, int64_t ci, double cf) {
  uint64_t v[U];
  values_of<U>(type, raw, v);
  uint32_t m = 0;
#pragma unroll
  for (int u = 0; u < U; ++u) m |= (uint32_t)((vw[u] >> (row[u] & 63)) & 1ull) << u;
  uint32_t pass = 0;
  if (type == WG_FLOAT64) {
    switch (op) {
#define WG_PRED_CASE(OPC, EXPR)                                                           \
  case OPC:                                                                               \
    _Pragma("unroll") for (int u = 0; u < U; ++u) {                                       \
      const double x = __longlong_as_double((long long)v[u]);                            \
      pass |= (uint32_t)(EXPR) << u;                                                      \
    }                                                                                     \
    break;
      WG_PRED_CASE(WG_EQ, x == cf)
      WG_PRED_CASE(WG_NE, x != cf)
      WG_PRED_CASE(WG_LT, x < cf)
      WG_PRED_CASE(WG_LE, x <= cf)
      WG_PRED_CASE(WG_GT, x > cf)
      default: WG_PRED_CASE(WG_GE, x >= cf)
#undef WG_PRED_CASE
    }
  } else if (type == WG_INT32 || type == WG_UINT16) {
    const int32_t k = (int32_t)ci;
    switch (op) {
#define WG_PRED_CASE(OPC, EXPR)                                                           \
  case OPC:                                                                               \
    _Pragma("unroll") for (int u = 0; u < U; ++u) {                                       \
      const int32_t x = (int32_t)(int64_t)v[u];                                           \
      pass |= (uint32_t)(EXPR) << u;                                                      \
    }                                                                                     \
    break;
      WG_PRED_CASE(WG_EQ, x == k)
      WG_PRED_CASE(WG_NE, x != k)
      WG_PRED_CASE(WG_LT, x < k)
      WG_PRED_CASE(WG_LE, x <= k)
      WG_PRED_CASE(WG_GT, x > k)
      default: WG_PRED_CASE(WG_GE, x >= k)
#undef WG_PRED_CASE
    }
  } else {
    switch (op) {
#define WG_PRED_CASE(OPC, EXPR)                                                           \
  case OPC:                                                                               \
    _Pragma("unroll") for (int u = 0; u < U; ++u) {                                       \
      const int64_t x = (int64_t)v[u];                                                    \
      pass |= (uint32_t)(EXPR) << u;                                                      \
    }                                                                                     \
    break;
      WG_PRED_CASE(WG_EQ, x == ci)
      WG_PRED_CASE(WG_NE, x != ci)
      WG_PRED_CASE(WG_LT, x < ci)
      WG_PRED_CASE(WG_LE, x <= ci)
      WG_PRED_CASE(WG_GT, x > ci)
      default: WG_PRED_CASE(WG_GE, x >= ci)
#undef WG_PRED_CASE
    }
  }
  return m & pass;
}

// Everything the streaming loops read per row, resolved once per kernel.
template <int NV>
struct RowSpec {
  ColRes pred[2];
  int32_t pop[2];
  int64_t pci[2];
  double pcf[2];
  bool pon[2];
  ColRes key[2];
  int32_t nkeys;
  int32_t narrow[2], shift[2];
  ColRes val[NV > 0 ? NV : 1];
  int32_t nvals;
  const uint32_t* sel;
  uint64_t row_base;
};

// PP provides cs, key (at most two packed columns), pred[2], sel, row_base,
// n_vals, val_col.
template <int NV, typename PP>
__device__ __forceinline__ RowSpec<NV> make_spec(const PP& P) {
  RowSpec<NV> S;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    S.pon[q] = P.pred[q].col >= 0;
    S.pred[q] = resolve_col(P.cs.cols[S.pon[q] ? P.pred[q].col : 0], S.pon[q]);
    S.pop[q] = P.pred[q].op;
    S.pci[q] = P.pred[q].i64;
    S.pcf[q] = P.pred[q].f64;
  }
  S.nkeys = P.key.n;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const bool on = k < P.key.n;
    S.key[k] = resolve_col(P.cs.cols[P.key.col[on ? k : 0]], on);
    S.narrow[k] = on ? P.key.narrow[k] : 0;
    S.shift[k] = on ? P.key.shift[k] : 0;
  }
  S.nvals = P.n_vals;
#pragma unroll
  for (int j = 0; j < (NV > 0 ? NV : 1); ++j) {
    const bool on = j < P.n_vals;
    S.val[j] = resolve_col(P.cs.cols[on ? P.val_col[j] : 0], on);
  }
  S.sel = P.sel;
  S.row_base = P.row_base;
  return S;
}

// Load and classify U rows per thread (positions i0 + u * step < end; the
// row is sel[i] or row_base + i). Every load of every row is issued before
// any value is used, so a wave keeps U rows of traffic in flight.
// kind: -1 filtered out or past the end, 0 keyed row, 1 sentinel-valued key,
// 2 NULL single-column key. NV value columns are loaded as raw words with
// their validity bits in vvalid (bit j).
template <int NV, int U>
__device__ __forceinline__ void load_rows(const RowSpec<NV>& S, uint64_t i0, uint64_t step, uint64_t end,
                                          uint64_t (&row)[U], uint64_t (&key)[U], int (&kind)[U],
                                          uint64_t (&val)[U][NV > 0 ? NV : 1], uint32_t (&vvalid)[U]) {
  bool in[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const uint64_t i = i0 + u * step;
    in[u] = i < end;
    const uint64_t ii = in[u] ? i : end - 1;
    row[u] = S.sel ? (uint64_t)S.sel[ii] : S.row_base + ii;
  }
  // loads: no use of a loaded value until every load of the batch is issued
  RawVal pr[2][U], kw[2][U], vr[U][NV > 0 ? NV : 1];
  uint64_t pv[2][U], kv[U], vv[U][NV > 0 ? NV : 1];
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int u = 0; u < U; ++u) {
      pr[q][u] = load_res(S.pred[q], row[u]);
      pv[q][u] = valid_res(S.pred[q], row[u]);
    }
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int u = 0; u < U; ++u) kw[k][u] = load_res(S.key[k], row[u]);
#pragma unroll
  for (int u = 0; u < U; ++u) kv[u] = valid_res(S.key[0], row[u]);
#pragma unroll
  for (int j = 0; j < NV; ++j)
#pragma unroll
    for (int u = 0; u < U; ++u) {
      vr[u][j] = load_res(S.val[j], row[u]);
      vv[u][j] = valid_res(S.val[j], row[u]);
    }
  // every load above is in flight; from here on the values are used, one
  // type / operator dispatch per column for all U rows (not one per row)
#pragma unroll
  for (int u = 0; u < U; ++u) {
    uint32_t vb = ~0u;
#pragma unroll
    for (int j = 0; j < NV; ++j) vb &= ~((uint32_t)!((vv[u][j] >> (row[u] & 63)) & 1ull) << j);
    vvalid[u] = vb;
  }
#pragma unroll
  for (int j = 0; j < NV; ++j)
    if (j < S.nvals) {
      uint64_t tmp[U];
      RawVal col[U];
#pragma unroll
      for (int u = 0; u < U; ++u) col[u] = vr[u][j];
      values_of<U>(S.val[j].type, col, tmp);
#pragma unroll
      for (int u = 0; u < U; ++u) val[u][j] = tmp[u];
    }
  uint32_t ok = 0;
#pragma unroll
  for (int u = 0; u < U; ++u) ok |= (uint32_t)in[u] << u;
#pragma unroll
  for (int q = 0; q < 2; ++q)
    if (S.pon[q]) ok &= pred_mask<U>(S.pred[q].type, pr[q], pv[q], row, S.pop[q], S.pci[q], S.pcf[q]);
  values_of<U>(S.key[0].type, kw[0], key);
  if (S.nkeys > 1) {
    uint64_t k1[U];
    values_of<U>(S.key[1].type, kw[1], k1);
    const int n0 = S.narrow[0], n1 = S.narrow[1], s0 = S.shift[0], s1 = S.shift[1];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t a = n0 ? (key[u] & 0xFFFFFFFFull) << s0 : key[u];
      const uint64_t b = n1 ? (k1[u] & 0xFFFFFFFFull) << s1 : k1[u];
      key[u] = a | b;
    }
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    // multi-column keys carry no validity (the host rejects them up front)
    const bool null_key = S.nkeys == 1 && !((kv[u] >> (row[u] & 63)) & 1ull);
    kind[u] = !((ok >> u) & 1u) ? -1 : null_key ? 2 : (key[u] == kEmpty ? 1 : 0);
  }
}

template <int NV>
__device__ __forceinline__ uint64_t pick(const uint64_t (&v)[NV], int c) {
  uint64_t r = 0;
#pragma unroll
  for (int j = 0; j < NV; ++j)
    if (j == c) r = v[j];
  return r;
}

// apply_row for a preloaded row: the op's columns index the value words v
// (float64 bits where vfloat bit j is set, else int64), vvalid bit j = value
// j is not NULL. Same semantics as apply_row for every column type except
// DECIMAL128 (16-byte values are not preloaded).
template <typename S, int NV>
__device__ __forceinline__ void apply_vals(uint64_t* w, int ws, const AggProgram& pg, const uint64_t (&v)[NV],
                                           uint32_t vfloat, uint32_t vvalid) {
  for (int o = 0; o < pg.n_ops; ++o) {
    const OpDesc op = pg.ops[o];
    uint64_t* t = w + op.word * ws;
    const bool ok1 = op.col >= 0 && ((vvalid >> op.col) & 1u);
    const bool ok2 = ok1 && (op.col2 < 0 || ((vvalid >> op.col2) & 1u));
    switch (op.kind) {
      case OP_ROWCOUNT:
        at_add<S>(t, 1);
        break;
      case OP_NNCOUNT:
        if (ok2) at_add<S>(t, 1);
        break;
      case OP_SUMI:
        if (ok2) {
          uint64_t x = pick(v, op.col);
          if (op.col2 >= 0) x *= pick(v, op.col2);
          at_add<S>(t, x);
        }
        break;
      case OP_SUMF:
        if (ok1) {
          const uint64_t b = pick(v, op.col);
          at_addf<S>(t, ((vfloat >> op.col) & 1u) ? __longlong_as_double((long long)b) : (double)(int64_t)b);
        }
        break;
      case OP_SUMD:
        if (ok1) add_limbs2<S>(t, ws, (int64_t)pick(v, op.col));
        break;
      case OP_SUMD4:
        if (ok2) {
          const int64_t a = (int64_t)pick(v, op.col);
          uint64_t lo;
          int64_t hi;
          if (op.col2 >= 0) {
            const int64_t b = (int64_t)pick(v, op.col2);
            lo = (uint64_t)a * (uint64_t)b;
            hi = (int64_t)__mul64hi((long long)a, (long long)b);
          } else {
            lo = (uint64_t)a;
            hi = a < 0 ? -1 : 0;
          }
          add_limbs4<S>(t, ws, lo, hi);
        }
        break;
      case OP_MIN: case OP_MAX:
        if (ok1) {
          const uint64_t b = pick(v, op.col);
          uint64_t e;
          if ((vfloat >> op.col) & 1u) {
            if ((b & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull) break;  // NaN never wins
            e = enc_f64_bits(b);
          } else {
            e = enc_i64((int64_t)b);
          }
          if (op.kind == OP_MIN) at_min<S>(t, e);
          else at_max<S>(t, e);
        }
        break;
      default:
        break;
    }
  }
}

// apply_vals on an LDS slot with no per-row branch: where the row does not
// apply (dead row, NULL input, NaN for MIN/MAX) the operand is the
// operation's identity (0, -0.0, the MIN/MAX initial word), so all lanes run
// the same atomics. Dead rows pass their lane's spare slot.
template <int NV>
__device__ __forceinline__ void apply_vals_lds(uint64_t* w, int ws, const AggProgram& pg, const uint64_t (&v)[NV],
                                               uint32_t vfloat, uint32_t vvalid, bool live) {
  for (int o = 0; o < pg.n_ops; ++o) {
    const OpDesc op = pg.ops[o];
    uint64_t* t = w + op.word * ws;
    const bool ok1 = live && op.col >= 0 && ((vvalid >> op.col) & 1u);
    const bool ok2 = ok1 && (op.col2 < 0 || ((vvalid >> op.col2) & 1u));
    switch (op.kind) {
      case OP_ROWCOUNT:
        at_add<LdsT>(t, live ? 1 : 0);
        break;
      case OP_NNCOUNT:
        at_add<LdsT>(t, ok2 ? 1 : 0);
        break;
      case OP_SUMI: {
        uint64_t x = pick(v, op.col);
        if (op.col2 >= 0) x *= pick(v, op.col2);
        at_add<LdsT>(t, ok2 ? x : 0);
        break;
      }
      case OP_SUMF: {
        const uint64_t b = pick(v, op.col);
        const double x = ((vfloat >> op.col) & 1u) ? __longlong_as_double((long long)b) : (double)(int64_t)b;
        at_addf<LdsT>(t, ok1 ? x : -0.0);
        break;
      }
      case OP_SUMD:
        add_limbs2<LdsT>(t, ws, ok1 ? (int64_t)pick(v, op.col) : 0);
        break;
      case OP_SUMD4: {
        const int64_t a = (int64_t)pick(v, op.col);
        uint64_t lo;
        int64_t hi;
        if (op.col2 >= 0) {
          const int64_t b = (int64_t)pick(v, op.col2);
          lo = (uint64_t)a * (uint64_t)b;
          hi = (int64_t)__mul64hi((long long)a, (long long)b);
        } else {
          lo = (uint64_t)a;
          hi = a < 0 ? -1 : 0;
        }
        add_limbs4<LdsT>(t, ws, ok2 ? lo : 0, ok2 ? hi : 0);
        break;
      }
      case OP_MIN: case OP_MAX: {
        const uint64_t b = pick(v, op.col);
        uint64_t e;
        bool ok = ok1;
        if ((vfloat >> op.col) & 1u) {
          ok = ok && (b & 0x7FFFFFFFFFFFFFFFull) <= 0x7FF0000000000000ull;  // NaN never wins
          e = enc_f64_bits(b);
        } else {
          e = enc_i64((int64_t)b);
        }
        if (op.kind == OP_MIN) at_min<LdsT>(t, ok ? e : kMinInit);
        else at_max<LdsT>(t, ok ? e : kMaxInit);
        break;
      }
      default:
        break;
    }
  }
}

// apply_vals_lds for U rows at once: each op is decoded once for the U rows
// (the op loop is the outer loop), so the per-op dispatch is shared.
template <int NV, int U>
__device__ __forceinline__ void apply_vals_lds_batch(uint64_t* const (&w)[U], int ws, const AggProgram& pg,
                                                     const uint64_t (&v)[U][NV], uint32_t vfloat,
                                                     const uint32_t (&vvalid)[U], uint32_t live) {
  // one loop per op kind over that kind's run of the (kind-sorted) program:
  // the kind is known at compile time in each loop body, so the updates of
  // one kind are straight-line code with no per-op dispatch
  int o = 0;
  for (; o < pg.kind_end[OP_ROWCOUNT]; ++o) {
    const int word = pg.ops[o].word;
#pragma unroll
    for (int u = 0; u < U; ++u) at_add<LdsT>(w[u] + word * ws, (live >> u) & 1u);
  }
  for (; o < pg.kind_end[OP_NNCOUNT]; ++o) {
    const OpDesc op = pg.ops[o];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool ok = ((live >> u) & 1u) && ((vvalid[u] >> op.col) & 1u) &&
                      (op.col2 < 0 || ((vvalid[u] >> op.col2) & 1u));
      at_add<LdsT>(w[u] + op.word * ws, ok ? 1 : 0);
    }
  }
  for (; o < pg.kind_end[OP_SUMI]; ++o) {
    const OpDesc op = pg.ops[o];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool ok = ((live >> u) & 1u) && ((vvalid[u] >> op.col) & 1u) &&
                      (op.col2 < 0 || ((vvalid[u] >> op.col2) & 1u));
      uint64_t x = pick(v[u], op.col);
      if (op.col2 >= 0) x *= pick(v[u], op.col2);
      at_add<LdsT>(w[u] + op.word * ws, ok ? x : 0);
    }
  }
  for (; o < pg.kind_end[OP_SUMF]; ++o) {
    const OpDesc op = pg.ops[o];
    // the column's type is the same in every lane: one branch per op, so a
    // float64 column pays for no int64 -> float64 conversion
    if ((vfloat >> op.col) & 1u) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool ok = ((live >> u) & 1u) && ((vvalid[u] >> op.col) & 1u);
        at_addf<LdsT>(w[u] + op.word * ws, ok ? __longlong_as_double((long long)pick(v[u], op.col)) : -0.0);
      }
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool ok = ((live >> u) & 1u) && ((vvalid[u] >> op.col) & 1u);
        at_addf<LdsT>(w[u] + op.word * ws, ok ? (double)(int64_t)pick(v[u], op.col) : -0.0);
      }
    }
  }
  for (; o < pg.kind_end[OP_SUMD]; ++o) {
    const OpDesc op = pg.ops[o];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool ok = ((live >> u) & 1u) && ((vvalid[u] >> op.col) & 1u);
      const int64_t x = ok ? (int64_t)pick(v[u], op.col) : 0;
      uint64_t* t = w[u] + op.word * ws;
      at_add<LdsT>(t, (uint64_t)x & 0xFFFFFFFFull);
      // the signed high limb is 0 for every value in [0, 2^32): a wave
      // whose values all are skips its atomic (adding 0 changes nothing)
      if (__any((x >> 32) != 0)) at_add<LdsT>(t + ws, (uint64_t)(x >> 32));
    }
  }
  for (; o < pg.kind_end[OP_SUMD4]; ++o) {
    const OpDesc op = pg.ops[o];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool ok = ((live >> u) & 1u) && ((vvalid[u] >> op.col) & 1u) &&
                      (op.col2 < 0 || ((vvalid[u] >> op.col2) & 1u));
      const int64_t a = (int64_t)pick(v[u], op.col);
      uint64_t lo;
      int64_t hi;
      if (op.col2 >= 0) {
        const int64_t b = (int64_t)pick(v[u], op.col2);
        lo = (uint64_t)a * (uint64_t)b;
        hi = (int64_t)__mul64hi((long long)a, (long long)b);
      } else {
        lo = (uint64_t)a;
        hi = a < 0 ? -1 : 0;
      }
      add_limbs4<LdsT>(w[u] + op.word * ws, ws, ok ? lo : 0, ok ? hi : 0);
    }
  }
  for (; o < pg.kind_end[OP_MAX]; ++o) {  // MIN then MAX
    const OpDesc op = pg.ops[o];
    const bool is_min = op.kind == OP_MIN;
    // encoded operands, one branch per op on the column's (lane-uniform) type
    uint64_t e[U];
    uint32_t ok = 0;
    if ((vfloat >> op.col) & 1u) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint64_t b = pick(v[u], op.col);
        const bool k = ((live >> u) & 1u) && ((vvalid[u] >> op.col) & 1u) &&
                       (b & 0x7FFFFFFFFFFFFFFFull) <= 0x7FF0000000000000ull;  // NaN never wins
        ok |= (uint32_t)k << u;
        e[u] = enc_f64_bits(b);
      }
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        ok |= (uint32_t)(((live >> u) & 1u) && ((vvalid[u] >> op.col) & 1u)) << u;
        e[u] = enc_i64((int64_t)pick(v[u], op.col));
      }
    }
    if (is_min) {
#pragma unroll
      for (int u = 0; u < U; ++u) at_min<LdsT>(w[u] + op.word * ws, ((ok >> u) & 1u) ? e[u] : kMinInit);
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) at_max<LdsT>(w[u] + op.word * ws, ((ok >> u) & 1u) ? e[u] : kMaxInit);
    }
  }
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
