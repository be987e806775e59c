// wg_agg.hip — grouped aggregation on MI355X: kernels and the wg_agg_* C-ABI.
//
// Reference: HashAggregate (internal/engine/exec/aggregate.go:98-434), its typed
// dispatch resolveIndices (:1296-1764), the int / packed consume loops
// (:1908-2146, :2326-2588), the SoA scatter kernels (agg_scatter.go), emission
// (:4226-4524, aggregate_partial_spill.go:1855-2021) and MergeSink (:4682-4700).
//
// Device pipeline of one consume call (all on the ctx stream):
//   agg_consume_kernel  persistent workgroups pull row tiles from an atomic
//                       tile counter; each row is filtered (fused predicate),
//                       keyed, and folded into the workgroup's LDS table when
//                       its key fits there, else into the global table; at
//                       the end the LDS table is flushed into the global one.
//   host read-back      groups / overflow counters (one sync per call).
//   grow + replay       only if the table ran past its load budget: rehash
//                       into a 2-4x table, merge the overflow partials, replay
//                       the overflow rows, continue with the untouched tiles.
#include <algorithm>
#include <cstdarg>
#include <vector>

#include "wg_agg_dev.hpp"
#include "wg_agg_host.hpp"

namespace wg {

// ===================================================================== kernels

struct ConsumeParams {
  ColSet cs;
  KeyDesc key;
  PredDesc pred[2];  // col = -1: unused
  AggProgram pg;
  GlobalTable T;
  const uint32_t* sel;  // null = rows [row_base, row_base + n)
  uint64_t row_base;
  uint64_t n;           // rows (or sel entries)
  uint64_t tile_rows;
  uint64_t n_tiles;
  uint32_t* ovf_rows;
  uint64_t ovf_rows_cap;
  uint64_t* ovf_part;
  uint64_t ovf_part_cap;
  int32_t lds_slots;  // power of two, 0 = no LDS table
  // preloaded values (NV > 0 kernels): pg's op columns index val_col
  int32_t n_vals;
  int32_t val_col[kMaxVals];
  uint32_t vfloat;  // bit j: value j is float64 bits
};

// NV > 0: every value a row needs is preloaded with its keys (kRowU2 rows per
// thread in flight) and applied with apply_vals; NV = 0: values are read by
// apply_row when the row is applied (DECIMAL128 inputs, more than three value
// columns).
template <bool USE_LDS, int NV>
__global__ void __launch_bounds__(1024) agg_consume_kernel(ConsumeParams P) {
  extern __shared__ uint64_t lds[];
  __shared__ int64_t s_tile;
  const int stride = P.pg.stride;
  const int L = USE_LDS ? P.lds_slots : 0;
  const int LW = USE_LDS ? lds_word_stride(L) : 0;  // L hashed slots, then one spare slot per lane
  uint64_t* const spare = lds + L + (threadIdx.x & 63);
  const RowSpec<NV> S = make_spec<NV>(P);

  if (USE_LDS) lds_table_init(lds, LW, P.pg);
  __syncthreads();

  uint32_t my_lds = 0, my_glb = 0, my_claims = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      int64_t t = -1;
      uint64_t o = __hip_atomic_load(P.T.counters + CTR_OVF_ROWS, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (o + (uint64_t)gridDim.x * P.tile_rows > P.ovf_rows_cap) {
        at_add<GlbT>(P.T.counters + CTR_STOPPED, 1);
      } else {
        uint64_t tt = __hip_atomic_fetch_add(P.T.counters + CTR_TILE, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (tt < P.n_tiles) t = (int64_t)tt;
      }
      s_tile = t;
    }
    __syncthreads();
    const int64_t tile = s_tile;
    __syncthreads();
    if (tile < 0) break;

    const uint64_t beg = (uint64_t)tile * P.tile_rows;
    const uint64_t end = min(beg + P.tile_rows, P.n);
    for (uint64_t base = beg; base < end; base += (uint64_t)blockDim.x * kRowU2) {
      uint64_t row[kRowU2], key[kRowU2], val[kRowU2][NV > 0 ? NV : 1];
      uint32_t vvalid[kRowU2];
      int kind[kRowU2];
      load_rows<NV, kRowU2>(S, base + threadIdx.x, blockDim.x, end, row, key, kind, val, vvalid);
      // LDS first: rows whose key is (or gets) a slot in the workgroup table
      uint64_t* ls[kRowU2];
#pragma unroll
      for (int u = 0; u < kRowU2; ++u) {
        ls[u] = nullptr;
        if (USE_LDS && kind[u] == 0) ls[u] = lds_find_or_claim(lds, L - 1, key[u], key_hash(key[u]));
      }
      if (USE_LDS && NV > 0) {
        // every lane runs the same LDS atomics (identity operands where the
        // row does not aggregate here): no per-row branch on the common path
        uint64_t* w[kRowU2];
        uint32_t live = 0;
#pragma unroll
        for (int u = 0; u < kRowU2; ++u) {
          const bool hit = ls[u] != nullptr;
          w[u] = hit ? ls[u] : spare;
          live |= (uint32_t)hit << u;
          my_lds += hit ? 1u : 0u;
        }
        apply_vals_lds_batch<(NV > 0 ? NV : 1), kRowU2>(w, LW, P.pg, val, P.vfloat, vvalid, live);
      }
      // the rest: side slots (NULL key, sentinel-valued key) and the global table
#pragma unroll
      for (int u = 0; u < kRowU2; ++u) {
        if (kind[u] < 0) continue;
        if (USE_LDS && ls[u] != nullptr) {
          if (NV == 0) {
            apply_row<LdsT, false>(ls[u], LW, P.pg, P.cs, RowRef{row[u], 0, nullptr});
            ++my_lds;
          }
          continue;
        }
        uint64_t* slot;
        if (kind[u] == 2) {
          slot = P.T.slots + (P.T.mask + 1 + 1) * (uint64_t)stride;  // NULL-key side slot
        } else if (kind[u] == 1) {
          slot = P.T.slots + (P.T.mask + 1) * (uint64_t)stride;  // sentinel-valued key side slot
        } else {
          slot = glb_find_or_claim(P.T, stride, key[u], key_hash(key[u]), my_claims);
          if (!slot) {
            uint64_t at = __hip_atomic_fetch_add(P.T.counters + CTR_OVF_ROWS, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            P.ovf_rows[at] = (uint32_t)row[u];
            continue;
          }
        }
        if (NV > 0) {
          uint64_t v[NV > 0 ? NV : 1];
#pragma unroll
          for (int j = 0; j < (NV > 0 ? NV : 1); ++j) v[j] = val[u][j];
          apply_vals<GlbT>(slot, 1, P.pg, v, P.vfloat, vvalid[u]);
        } else {
          apply_row<GlbT, false>(slot, 1, P.pg, P.cs, RowRef{row[u], 0, nullptr});
        }
        ++my_glb;
      }
    }
  }

  if (USE_LDS) {
    __syncthreads();
    uint32_t flushed = 0;
    for (int i = threadIdx.x; i < L; i += blockDim.x) {
      const uint64_t* ls = lds + i;
      const uint64_t key = ls[0];
      if (key == kEmpty) continue;
      ++flushed;
      uint64_t* slot = glb_find_or_claim(P.T, stride, key, key_hash(key), my_claims);
      if (slot != nullptr) {
        merge_partial(slot, ls, LW, P.pg);
      } else {
        uint64_t at = __hip_atomic_fetch_add(P.T.counters + CTR_OVF_PART, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint64_t* rec = P.ovf_part + at * (uint64_t)stride;
        for (int w = 0; w < stride; ++w) rec[w] = ls[(uint64_t)w * LW];
      }
    }
    block_add(P.T.counters + CTR_FLUSHED, flushed);
  }
  // one atomic per workgroup for each statistic and for the claimed groups
  block_add(P.T.counters + CTR_GROUPS, my_claims);
  block_add(P.T.counters + CTR_LDS_ROWS, my_lds);
  block_add(P.T.counters + CTR_GLB_ROWS, my_glb);
}

// Fill every slot (and the side slots) with the empty pattern: 16-byte
// stores of the precomputed slot words (the stride is a power of two >= 2).
__global__ void __launch_bounds__(256) agg_fill_kernel(uint64_t* slots, uint64_t n_words, AggProgram pg) {
  __shared__ uint64_t s_init[16];
  const int stride = pg.stride;
  if (threadIdx.x < (unsigned)stride) s_init[threadIdx.x] = slot_init_word(pg, threadIdx.x);
  __syncthreads();
  ulonglong2* d = (ulonglong2*)slots;
  const uint64_t n2 = n_words / 2;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n2; i += (uint64_t)gridDim.x * blockDim.x) {
    const int w = (int)((2 * i) & (uint64_t)(stride - 1));
    d[i] = make_ulonglong2(s_init[w], s_init[w + 1]);
  }
}

// Is the slot a live group? Table slots: key != EMPTY; side slots: rowcount > 0.
__device__ __forceinline__ bool slot_live(const uint64_t* s, bool side) {
  return side ? (s[1] != 0) : (s[0] != kEmpty);
}

// Move every live slot of `src` into `dst` (a larger, empty table). Keys are
// unique in src, so after the claim the words are copied with plain stores.
// base arrays (exact Int128 limb bases) move with their slot.
__global__ void agg_rehash_kernel(const uint64_t* src, uint64_t src_cap, const uint64_t* src_base,
                                  GlobalTable T, uint64_t* dst_base, int stride, int n_base_words) {
  const uint64_t total = src_cap + kSideSlots;
  uint32_t claims = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
    const bool side = i >= src_cap;
    const uint64_t* s = src + i * stride;
    if (!slot_live(s, side)) continue;
    uint64_t* d;
    uint64_t di;
    if (side) {
      di = (T.mask + 1) + (i - src_cap);
      d = T.slots + di * stride;
    } else {
      d = glb_find_or_claim(T, stride, s[0], key_hash(s[0]), claims);
      if (d == nullptr) {  // the new table is sized for all of src: a probe run this long is a bug
        at_add<GlbT>(T.counters + CTR_STOPPED, 1ull << 32);
        continue;
      }
      di = (uint64_t)(d - T.slots) / stride;
    }
    for (int w = 1; w < stride; ++w) d[w] = s[w];
    if (src_base != nullptr)
      for (int w = 0; w < n_base_words; ++w) dst_base[di * n_base_words + w] = src_base[i * n_base_words + w];
  }
  block_add(T.counters + CTR_GROUPS, claims);
}

// Merge overflow partial records (possibly several per key) with atomics.
__global__ void agg_merge_part_kernel(const uint64_t* recs, uint64_t n, GlobalTable T, AggProgram pg) {
  const int stride = pg.stride;
  uint32_t claims = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t* r = recs + i * stride;
    uint64_t* d = glb_find_or_claim(T, stride, r[0], key_hash(r[0]), claims);
    if (d == nullptr) {
      // the host grows the table before replaying partials, so this is a bug
      at_add<GlbT>(T.counters + CTR_STOPPED, 1ull << 32);
      continue;
    }
    merge_partial(d, r, 1, pg);
  }
  block_add(T.counters + CTR_GROUPS, claims);
}

// Fold each SUMD limb group into the exact Int128 base and zero the limbs
// (before the limbs could reach 2^32 additions).
__global__ void agg_normalize_kernel(uint64_t* slots, uint64_t total_slots, uint64_t* base, int stride,
                                     const int32_t* limb_word, const int32_t* limb_n, int n_sumd) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total_slots; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t* s = slots + i * stride;
    for (int d = 0; d < n_sumd; ++d) {
      uint64_t* l = s + limb_word[d];
      uint64_t lo;
      int64_t hi;
      limbs_value(l, limb_n[d], lo, hi);
      uint64_t* b = base + (i * n_sumd + d) * 2;
      unsigned __int128 v = ((unsigned __int128)b[1] << 64 | b[0]) + ((unsigned __int128)(uint64_t)hi << 64 | lo);
      b[0] = (uint64_t)v;
      b[1] = (uint64_t)(v >> 64);
      for (int k = 0; k < limb_n[d]; ++k) l[k] = 0;
    }
  }
}

// Merge every live slot of src (table + side slots, with its base) into dst.
// src keys are unique, so each dst slot receives at most one src slot: after
// the claim, words are combined with plain read-modify-writes.
__global__ void agg_merge_table_kernel(const uint64_t* src, uint64_t src_cap, const uint64_t* src_base,
                                       GlobalTable T, uint64_t* dst_base, AggProgram pg,
                                       const int32_t* limb_word, const int32_t* limb_n, int n_sumd) {
  const int stride = pg.stride;
  const uint64_t total = src_cap + kSideSlots;
  uint32_t claims = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
    const bool side = i >= src_cap;
    const uint64_t* s = src + i * stride;
    if (!slot_live(s, side)) continue;
    uint64_t* d;
    uint64_t di;
    if (side) {
      di = (T.mask + 1) + (i - src_cap);
      d = T.slots + di * stride;
    } else {
      d = glb_find_or_claim(T, stride, s[0], key_hash(s[0]), claims);
      if (d == nullptr) {
        at_add<GlbT>(T.counters + CTR_STOPPED, 1ull << 32);
        continue;
      }
      di = (uint64_t)(d - T.slots) / stride;
    }
    for (int o = 0; o < pg.n_ops; ++o) {
      const OpDesc op = pg.ops[o];
      uint64_t* t = d + op.word;
      const uint64_t* x = s + op.word;
      switch (op.kind) {
        case OP_ROWCOUNT: case OP_NNCOUNT: case OP_SUMI: t[0] += x[0]; break;
        case OP_SUMF: {
          double a = __longlong_as_double((long long)t[0]) + __longlong_as_double((long long)x[0]);
          t[0] = (uint64_t)__double_as_longlong(a);
          break;
        }
        case OP_MIN: t[0] = x[0] < t[0] ? x[0] : t[0]; break;
        case OP_MAX: t[0] = x[0] > t[0] ? x[0] : t[0]; break;
        default: break;  // SUMD limbs go through the base below
      }
    }
    for (int dd = 0; dd < n_sumd; ++dd) {
      uint64_t lo;
      int64_t hi;
      limbs_value(s + limb_word[dd], limb_n[dd], lo, hi);
      unsigned __int128 v = ((unsigned __int128)(uint64_t)hi << 64) | lo;
      if (src_base != nullptr) {
        const uint64_t* sb = src_base + (i * n_sumd + dd) * 2;
        v += ((unsigned __int128)sb[1] << 64) | sb[0];
      }
      uint64_t* b = dst_base + (di * n_sumd + dd) * 2;
      v += ((unsigned __int128)b[1] << 64) | b[0];
      b[0] = (uint64_t)v;
      b[1] = (uint64_t)(v >> 64);
    }
  }
  block_add(T.counters + CTR_GROUPS, claims);
}

// Compact live slots into a dense array (slot words + base words), order
// unspecified; out_null[pos] = 1 for the NULL-key group (side slot 1).
// Each workgroup owns a contiguous slot range: it counts its live slots,
// reserves its output range with ONE atomic, and places slots by a block
// prefix sum (wave ballot + per-wave totals).
constexpr int kCompactBlock = 256;

__global__ void __launch_bounds__(kCompactBlock) agg_compact_kernel(const uint64_t* slots, uint64_t cap,
                                                                   uint64_t first, const uint64_t* base, int stride,
                                                                   int n_base_words,
                                                                   uint64_t* out, uint64_t* out_base,
                                                                   uint8_t* out_null, uint64_t* out_n,
                                                                   uint64_t per_block) {
  __shared__ uint32_t wave_cnt[kCompactBlock / 64];
  __shared__ uint8_t live_lane[kCompactBlock];  // per wave: lanes of its live slots, in order
  __shared__ unsigned long long s_base;
  const uint64_t total = cap + kSideSlots;
  const uint64_t beg = first + (uint64_t)blockIdx.x * per_block;
  const uint64_t end = min(beg + per_block, total);
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  // pass 1: count this block's live slots
  uint32_t mine = 0;
  for (uint64_t i = beg + threadIdx.x; i < end; i += blockDim.x) mine += slot_live(slots + i * stride, i >= cap);
  __shared__ unsigned long long s_cnt;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  if (mine) atomicAdd(&s_cnt, (unsigned long long)mine);
  __syncthreads();
  if (threadIdx.x == 0) s_base = s_cnt ? atomicAdd((unsigned long long*)out_n, s_cnt) : 0;
  __syncthreads();
  // pass 2: place live slots in block order; a wave's live slots land in
  // consecutive output slots, which its lanes copy word by word (coalesced)
  const int sh = __ffs(stride) - 1;  // stride is a power of two
  uint64_t run = s_base;
  for (uint64_t i0 = beg; i0 < end; i0 += blockDim.x) {
    const uint64_t i = i0 + threadIdx.x;
    const bool side = i >= cap;
    const bool live = i < end && slot_live(slots + i * stride, side);
    const uint64_t b = __ballot(live);
    const uint32_t rank = (uint32_t)__popcll(b & ((1ull << lane) - 1));
    if (lane == 0) wave_cnt[wave] = (uint32_t)__popcll(b);
    if (live) live_lane[wave * 64 + rank] = (uint8_t)lane;
    __syncthreads();
    uint32_t before = 0, tot = 0;
    for (int w = 0; w < kCompactBlock / 64; ++w) {
      if (w < wave) before += wave_cnt[w];
      tot += wave_cnt[w];
    }
    const uint64_t pos0 = run + before;
    const uint32_t nl = (uint32_t)__popcll(b);
    const uint64_t wave_first = i0 + (uint64_t)wave * 64;
    for (uint32_t k = lane; k < (nl << sh); k += 64) {
      const uint32_t j = k >> sh, w = k & (uint32_t)(stride - 1);
      const uint64_t si = wave_first + live_lane[wave * 64 + j];
      // side slot 0 holds the key whose value is the sentinel
      out[(pos0 + j) * stride + w] = (w == 0 && si == cap) ? kEmpty : slots[si * stride + w];
    }
    if (live) {
      const uint64_t pos = pos0 + rank;
      out_null[pos] = (side && i == cap + 1) ? 1 : 0;
      if (out_base != nullptr)
        for (int w = 0; w < n_base_words; ++w) out_base[pos * n_base_words + w] = base ? base[i * n_base_words + w] : 0;
    }
    run += tot;
    __syncthreads();
  }
}

// Finalized output of one aggregate for groups [off, off+cnt) of the
// compacted array. Mirrors writeAccToColumn (aggregate_partial_spill.go:1855-1945)
// and Int128.ToFloat64 (decimal.go:39-50).
struct FinalDesc {
  int32_t func;
  int32_t count_word;  // word holding the NULL-deciding count
  int32_t val_kind;    // OpKind of the value word(s), 0 = none (always NULL)
  int32_t val_word;
  int32_t is_float_in; // MIN/MAX over float64
  int32_t sumd_index;  // index among SUMD ops (base array), -1
  int32_t scale;
  int32_t limbs;
};

__device__ __forceinline__ double i128_to_f64(uint64_t lo, int64_t hi, int scale) {
  // Int128.ToFloat64 (decimal.go:39-50), same operation order.
  double p10 = 1.0;
  for (int i = 0; i < scale; ++i) p10 *= 10.0;  // exact for scale <= 22 (math.Pow10 table)
  if ((hi == 0 && lo <= 0x7FFFFFFFFFFFFFFFull) || (hi == -1 && lo > 0x7FFFFFFFFFFFFFFFull))
    return (double)(int64_t)lo / p10;
  double f = (double)hi * 18446744073709551616.0 + (double)lo;
  return f / p10;
}

// Finalized value of one aggregate of a group: word(i) returns slot word i,
// b points at the group's exact Int128 bases (n_sumd pairs).
template <class Word>
__device__ __forceinline__ void decode_value(Word word, const uint64_t* b, const FinalDesc& fd, uint8_t& valid,
                                             double& vf, int64_t& vi) {
    const int64_t count = (int64_t)word(fd.count_word);
    switch (fd.func) {
      case WG_AGG_COUNT:
        vi = count;
        break;
      case WG_AGG_SUM:
      case WG_AGG_AVG: {
        if (count == 0 || fd.val_kind == 0) { valid = 0; break; }
        double sum;
        if (fd.val_kind == OP_SUMI) {
          vi = (int64_t)word(fd.val_word);
          sum = (double)vi;
        } else if (fd.val_kind == OP_SUMF) {
          sum = __longlong_as_double((long long)word(fd.val_word));
        } else {
          uint64_t l[4] = {word(fd.val_word), word(fd.val_word + 1), 0, 0};
          if (fd.limbs == 4) {
            l[2] = word(fd.val_word + 2);
            l[3] = word(fd.val_word + 3);
          }
          uint64_t lo;
          int64_t hi;
          limbs_value(l, fd.limbs, lo, hi);
          const uint64_t* bb = b + fd.sumd_index * 2;
          unsigned __int128 v = (((unsigned __int128)(uint64_t)hi << 64) | lo) + (((unsigned __int128)bb[1] << 64) | bb[0]);
          sum = i128_to_f64((uint64_t)v, (int64_t)(uint64_t)(v >> 64), fd.scale);
        }
        vf = fd.func == WG_AGG_AVG ? sum / (double)count : sum;
        break;
      }
      case WG_AGG_MIN:
      case WG_AGG_MAX: {
        if (count == 0 || fd.val_kind == 0) { valid = 0; break; }
        const uint64_t e = word(fd.val_word);
        const uint64_t init = fd.func == WG_AGG_MIN ? kMinInit : kMaxInit;
        if (fd.is_float_in) {
          if (e == init) vf = __longlong_as_double(0x7FF8000000000000ll);  // only NaN inputs
          else vf = __longlong_as_double((long long)dec_f64_bits(e));
        } else {
          vi = dec_i64(e);
        }
        break;
      }
    }
}

__global__ void agg_decode_kernel(const uint64_t* dense, const uint64_t* dense_base, uint64_t off, uint64_t cnt,
                                  int stride, int n_sumd, FinalDesc fd, double* out_f, int64_t* out_i,
                                  uint8_t* out_valid, int want_float) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < cnt; j += (uint64_t)gridDim.x * blockDim.x) {
    uint8_t valid = 1;
    double vf = 0.0;
    int64_t vi = 0;
    const uint64_t* row = dense + (off + j) * stride;
    decode_value([row](int i) { return row[i]; }, dense_base + (off + j) * n_sumd * 2, fd, valid, vf, vi);
    if (out_valid) out_valid[j] = valid;
    if (want_float) {
      if (out_f) out_f[j] = valid ? vf : 0.0;
    } else {
      if (out_i) out_i[j] = valid ? vi : 0;
    }
  }
}

// Unpack the key column k of groups [off, off+cnt) (writeIntKeyToColumn,
// aggregate_partial_spill.go:2006-2021; packedField.get, packed_hash.go:342-351).
__device__ __forceinline__ int64_t decode_key(uint64_t w, bool is_null, const KeyDesc& K, int k, const uint64_t* dict) {
  if (dict && !is_null) w = dict[2 * w + K.word[k]];  // 128-bit keys: the id's {lo, hi} word
  int64_t v;
  if (K.n == 1) v = (int64_t)w;
  else if (K.narrow[k] == 1) v = (int64_t)(int32_t)(uint32_t)(w >> K.shift[k]);
  else if (K.narrow[k] == 2) v = (int64_t)(uint32_t)(w >> K.shift[k]);
  else v = (int64_t)w;
  return is_null ? 0 : v;
}

__device__ __forceinline__ void store_key(void* out, uint64_t j, int type, int64_t v) {
  switch (type) {
    case WG_INT32: ((int32_t*)out)[j] = (int32_t)v; break;
    case WG_UINT32: ((uint32_t*)out)[j] = (uint32_t)v; break;
    case WG_UINT16: ((uint16_t*)out)[j] = (uint16_t)v; break;
    default: ((int64_t*)out)[j] = v; break;
  }
}

__global__ void agg_key_kernel(const uint64_t* dense, const uint8_t* dense_null, uint64_t off, uint64_t cnt,
                               int stride, KeyDesc K, int k, void* out, uint8_t* valid_out, const uint64_t* dict) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < cnt; j += (uint64_t)gridDim.x * blockDim.x) {
    const bool is_null = dense_null[off + j] != 0;
    const int64_t v = decode_key(dense[(off + j) * stride], is_null, K, k, dict);
    if (valid_out) valid_out[j] = is_null ? 0 : 1;
    if (out) store_key(out, j, K.key_type[k], v);
  }
}

// Every requested output column of groups [off, off+cnt) in one pass over the
// dense rows: each thread reads its group's slot words once and writes the
// keys and the finalized values (the Next drain of nextOwn,
// aggregate.go:4226-4524, into device columns).
struct EmitParams {
  const uint64_t* dense;
  const uint64_t* dense_base;
  const uint8_t* dense_null;
  const uint64_t* dict;
  uint64_t off, cnt;
  int32_t stride, n_sumd, n_keys, n_aggs;
  KeyDesc K;
  void* key_out[kMaxKeys];
  FinalDesc fd[kMaxOps];
  void* val_out[kMaxOps];
  uint8_t* valid_out[kMaxOps];
  uint32_t want_float;  // bit a: aggregate a emits float64
};

template <int S>
__global__ void __launch_bounds__(256) agg_emit_kernel(EmitParams E) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < E.cnt; j += (uint64_t)gridDim.x * blockDim.x) {
    // the group's slot words in registers, read once (S / 2 16-byte loads);
    // word i is picked by selects, so no word goes through scratch memory
    const uint64_t* s = E.dense + (E.off + j) * S;
    uint64_t w[S];
#pragma unroll
    for (int i = 0; i < S; i += 2) {
      const ulonglong2 x = *(const ulonglong2*)(s + i);
      w[i] = x.x;
      w[i + 1] = x.y;
    }
    auto word = [&w](int i) {
      uint64_t r = 0;
#pragma unroll
      for (int k = 0; k < S; ++k) r = i == k ? w[k] : r;
      return r;
    };
    const bool is_null = E.dense_null[E.off + j] != 0;
    for (int k = 0; k < E.n_keys; ++k)
      if (E.key_out[k]) store_key(E.key_out[k], j, E.K.key_type[k], decode_key(w[0], is_null, E.K, k, E.dict));
    const uint64_t* b = E.dense_base + (E.off + j) * E.n_sumd * 2;
    for (int a = 0; a < E.n_aggs; ++a) {
      uint8_t valid = 1;
      double vf = 0.0;
      int64_t vi = 0;
      decode_value(word, b, E.fd[a], valid, vf, vi);
      if (E.valid_out[a]) E.valid_out[a][j] = valid;
      if (!E.val_out[a]) continue;
      if ((E.want_float >> a) & 1u) ((double*)E.val_out[a])[j] = valid ? vf : 0.0;
      else ((int64_t*)E.val_out[a])[j] = valid ? vi : 0;
    }
  }
}

}  // namespace wg

// ================================================================ host side

namespace wg {

__global__ void zero_u64_kernel(uint64_t* p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = 0;
}

// Any NULL among rows [0,n) (or sel[0..n)) of a validity bitmap?
__global__ void any_null_kernel(const uint64_t* valid, const uint32_t* sel, uint64_t n, uint64_t* flag) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t row = sel ? sel[i] : i;
    if (!((valid[row >> 6] >> (row & 63)) & 1ull)) *flag = 1;
  }
}

int grid_for(uint64_t n, int block, int max_blocks) {
  uint64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > (uint64_t)max_blocks) g = max_blocks;
  return (int)g;
}

wg_status zero_dev(wg_ctx* ctx, uint64_t* p, uint64_t n) {
  hipLaunchKernelGGL(zero_u64_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, p, n);
  WG_HIP(hipGetLastError());
  return WG_OK;
}

// ---------------------------------------------------------- program compile

static bool nullable_in(const wg_agg* h, int c) { return c >= 0 && h->col_nullable[c]; }

static wg_status compile_program(wg_agg* h) {
  // word 0 = key, word 1 = ROWCOUNT (also the liveness mark of side slots)
  int next = 1;
  std::vector<OpDesc> ops;
  ops.push_back(OpDesc{OP_ROWCOUNT, -1, -1, next++});
  auto count_word_for = [&](int col, int col2) -> int {
    if (!nullable_in(h, col) && !nullable_in(h, col2)) return 1;
    for (auto& o : ops)
      if (o.kind == OP_NNCOUNT && o.col == col && o.col2 == col2) return o.word;
    ops.push_back(OpDesc{OP_NNCOUNT, col, col2, next++});
    return ops.back().word;
  };
  auto value_word_for = [&](int kind, int col, int col2, int words) -> int {
    for (auto& o : ops)
      if (o.kind == kind && o.col == col && o.col2 == col2) return o.word;
    ops.push_back(OpDesc{kind, col, col2, next});
    next += words;
    return ops.back().word;
  };
  h->finals.clear();
  h->sumd_word.clear();
  h->sumd_limbs.clear();
  for (const wg_agg_spec& a : h->aggs) {
    AggFinal f{};
    f.func = a.func;
    f.sumd_index = -1;
    f.val_kind = 0;
    WG_CHECK(a.func >= WG_AGG_SUM && a.func <= WG_AGG_AVG, WG_ERR_UNSUPPORTED, "aggregate function %d not on the GPU path", a.func);
    if (a.input_col < 0) {
      WG_CHECK(a.func == WG_AGG_COUNT, WG_ERR_INVALID, "only COUNT may omit its input column");
      f.count_word = 1;
      h->finals.push_back(f);
      continue;
    }
    WG_CHECK(a.input_col < h->n_cols && a.input_col2 < h->n_cols, WG_ERR_INVALID, "aggregate input column out of range");
    const int t = h->col_types[a.input_col];
    const int c2 = a.input_col2;
    int t_eff = t;
    int scale = h->col_scales[a.input_col];
    if (c2 >= 0) {
      const int t2 = h->col_types[c2];
      WG_CHECK(a.func == WG_AGG_SUM || a.func == WG_AGG_AVG, WG_ERR_UNSUPPORTED, "product inputs feed SUM/AVG only");
      const bool i1 = is_int_class(t) && t != WG_IPV4 && t != WG_UINT32;
      const bool i2 = is_int_class(t2) && t2 != WG_IPV4 && t2 != WG_UINT32;
      if (i1 && i2) t_eff = WG_INT64;
      else if (i1 && t2 == WG_DECIMAL) { t_eff = WG_DECIMAL128; scale = h->col_scales[c2]; }
      else if (t == WG_DECIMAL && i2) t_eff = WG_DECIMAL128;
      else WG_CHECK(false, WG_ERR_UNSUPPORTED, "product of types %d x %d not on the GPU path", t, t2);
    }
    f.in_type = t_eff;
    f.scale = scale;
    f.count_word = count_word_for(a.input_col, c2);
    switch (a.func) {
      case WG_AGG_COUNT:
        break;
      case WG_AGG_SUM:
      case WG_AGG_AVG:
        if (t_eff == WG_FLOAT64) {
          f.val_kind = OP_SUMF;
        } else if (t_eff == WG_INT64 && a.func == WG_AGG_AVG && c2 < 0) {
          f.val_kind = OP_SUMF;  // AVG(int64) accumulates float64 (agg_scatter.go:528-530)
        } else if (t_eff == WG_INT32 || t_eff == WG_UINT16 || t_eff == WG_INT64) {
          f.val_kind = OP_SUMI;
        } else if (t_eff == WG_DECIMAL) {
          f.val_kind = OP_SUMD;
          f.limbs = 2;
        } else if (t_eff == WG_DECIMAL128) {
          f.val_kind = OP_SUMD4;
          f.limbs = 4;
        } else {
          f.val_kind = 0;  // IPv4-class: no SUM kernel, emits NULL like the reference
        }
        break;
      case WG_AGG_MIN:
      case WG_AGG_MAX:
        if (t == WG_FLOAT64) {
          f.val_kind = a.func == WG_AGG_MIN ? OP_MIN : OP_MAX;
          f.is_float_in = 1;
        } else if (is_int_class(t)) {
          f.val_kind = a.func == WG_AGG_MIN ? OP_MIN : OP_MAX;
        } else {
          f.val_kind = 0;  // MIN/MAX(DECIMAL): no SoA kernel, emits NULL like the reference
        }
        break;
    }
    if (f.val_kind != 0) {
      const int words = f.val_kind == OP_SUMD ? 2 : f.val_kind == OP_SUMD4 ? 4 : 1;
      f.val_word = value_word_for(f.val_kind, a.input_col, f.val_kind == OP_MIN || f.val_kind == OP_MAX ? -1 : c2, words);
      if (f.val_kind == OP_SUMD || f.val_kind == OP_SUMD4) {
        int idx = -1;
        for (size_t i = 0; i < h->sumd_word.size(); ++i)
          if (h->sumd_word[i] == f.val_word) idx = (int)i;
        if (idx < 0) {
          idx = (int)h->sumd_word.size();
          h->sumd_word.push_back(f.val_word);
          h->sumd_limbs.push_back(words);
        }
        f.sumd_index = idx;
      }
    }
    h->finals.push_back(f);
  }
  WG_CHECK((int)ops.size() <= kMaxOps, WG_ERR_UNSUPPORTED, "too many accumulator ops (%d)", (int)ops.size());
  WG_CHECK(next <= 16, WG_ERR_UNSUPPORTED, "group state of %d words exceeds a 128-byte slot", next);
  h->n_words = next;
  // ops sorted by kind (ROWCOUNT, the smallest kind, stays first): the device
  // updates run one loop per kind over the kind's run of ops
  std::stable_sort(ops.begin(), ops.end(), [](const OpDesc& a, const OpDesc& b) { return a.kind < b.kind; });
  h->pg.n_ops = (int)ops.size();
  for (size_t i = 0; i < ops.size(); ++i) h->pg.ops[i] = ops[i];
  for (int k = 0; k <= OP_MAX; ++k) {
    int e = 0;
    for (const OpDesc& op : ops) e += op.kind <= k;
    h->pg.kind_end[k] = e;
  }
  h->pg.stride = (int)next_pow2((uint64_t)next);
  h->n_sumd = (int)h->sumd_word.size();
  return WG_OK;
}

static wg_status compile_key(wg_agg* h, int n_keys, const int32_t* key_cols) {
  WG_CHECK(n_keys >= 1 && n_keys <= kMaxKeys, WG_ERR_UNSUPPORTED, "GROUP BY of %d columns not on the GPU path", n_keys);
  KeyDesc& K = h->key;
  K = KeyDesc{};
  K.n = n_keys;
  int bits = 0;
  for (int i = 0; i < n_keys; ++i) {
    const int c = key_cols[i];
    WG_CHECK(c >= 0 && c < h->n_cols, WG_ERR_INVALID, "key column %d out of range", c);
    const int t = h->col_types[c];
    WG_CHECK(is_int_class(t), WG_ERR_UNSUPPORTED, "group key of type %d not on the GPU path", t);
    K.col[i] = c;
    K.key_type[i] = t;
    bits += packed_width(t) * 8;
  }
  if (n_keys == 1) return WG_OK;
  WG_CHECK(bits <= 128, WG_ERR_UNSUPPORTED, "composite key of %d bits: the GPU path packs <= 128", bits);
  if (bits > 64) {
    // {lo, hi} placement, widest first (buildPackedLayout, packed_hash.go:
    // 384-425): 8-byte columns take whole words, then 4-byte columns fill the
    // free half-words in (word, shift) order; the table keys on the pair's
    // dictionary id, held in a per-row column after the batch's own
    WG_CHECK(h->n_cols < kMaxCols, WG_ERR_UNSUPPORTED, "a 128-bit key needs a spare column slot (%d columns)",
             h->n_cols);
    KeyDesc W = K;
    int next_wide = 0;
    for (int i = 0; i < n_keys; ++i)
      if (packed_width(W.key_type[i]) == 8) {
        W.word[i] = (uint8_t)next_wide++;
        W.shift[i] = 0;
        W.narrow[i] = 0;
      }
    int slot = 0;
    for (int i = 0; i < n_keys; ++i) {
      if (packed_width(W.key_type[i]) != 4) continue;
      W.word[i] = (uint8_t)(next_wide + slot / 2);
      W.shift[i] = (uint8_t)(32 * (slot % 2));
      W.narrow[i] = W.key_type[i] == WG_INT32 ? 1 : 2;
      ++slot;
    }
    h->wide = true;
    h->wkey = W;
    K = KeyDesc{};
    K.n = 1;
    K.col[0] = h->n_cols;
    K.key_type[0] = WG_INT64;
    return WG_OK;
  }
  // 8-byte columns take the whole word (only possible alone); 4-byte columns
  // fill the low then the high half (buildPackedLayout, packed_hash.go:384-425)
  int shift = 0;
  for (int i = 0; i < n_keys; ++i) {
    const int t = K.key_type[i];
    K.narrow[i] = (t == WG_INT32) ? 1 : 2;
    K.shift[i] = (uint8_t)shift;
    shift += 32;
  }
  return WG_OK;
}

// ------------------------------------------------------------- table memory

uint64_t agg_footprint(const wg_agg* h) {
  uint64_t b = (h->cap + kSideSlots) * (uint64_t)h->pg.stride * 8;
  if (h->base) b += (h->cap + kSideSlots) * (uint64_t)h->n_sumd * 16;
  b += h->ovf_rows_cap * 4 + h->ovf_part_cap * (uint64_t)h->pg.stride * 8;
  b += h->dense_cap * ((uint64_t)h->pg.stride * 8 + (uint64_t)h->n_sumd * 16 + 1);
  for (void* p : h->stage.dev) (void)p;
  b += h->stage.cap_rows * 16 * (uint64_t)h->stage.dev.size();
  if (h->wide) b += kd_bytes(h->dict) + h->ids_cap * 8;
  return b;
}

static wg_status fill_table(wg_agg* h, uint64_t* slots, uint64_t cap) {
  const uint64_t words = (cap + kSideSlots) * (uint64_t)h->pg.stride;
  hipLaunchKernelGGL(agg_fill_kernel, dim3(grid_for(words, 256, 8192)), dim3(256), 0, h->ctx->stream, slots, words, h->pg);
  WG_HIP(hipGetLastError());
  return WG_OK;
}

// A filled table of cap slots, not yet attached to h.
static wg_status new_table(wg_agg* h, uint64_t cap, uint64_t** out) {
  uint64_t* s = nullptr;
  WG_ALLOC(h->ctx, &s, (cap + kSideSlots) * (uint64_t)h->pg.stride * 8, "a group table");
  const wg_status st = fill_table(h, s, cap);
  if (st != WG_OK) {
    (void)hipFree(s);
    return st;
  }
  *out = s;
  return WG_OK;
}

static wg_status alloc_table(wg_agg* h, uint64_t cap) {
  uint64_t* s = nullptr;
  WG_TRY(new_table(h, cap, &s));
  h->slots = s;
  h->cap = cap;
  return WG_OK;
}

wg_status ensure_base(wg_agg* h) {
  if (h->n_sumd == 0 || h->base) return WG_OK;
  const uint64_t words = (h->cap + kSideSlots) * (uint64_t)h->n_sumd * 2;
  WG_ALLOC(h->ctx, &h->base, words * 8, "exact Int128 bases");
  WG_TRY(zero_dev(h->ctx, h->base, words));
  return WG_OK;
}

GlobalTable table_of(wg_agg* h) {
  GlobalTable T;
  T.slots = h->slots;
  T.mask = h->cap - 1;
  T.counters = h->counters;
  T.max_probe = kMaxProbe;
  return T;
}

// Grow to a table of new_cap slots, moving every live slot (and its base).
// The new table and bases are allocated first: when that fails, h keeps the
// table it had.
wg_status grow_to(wg_agg* h, uint64_t new_cap) {
  if (new_cap <= h->cap) return WG_OK;
  uint64_t* fresh = nullptr;
  uint64_t* fresh_base = nullptr;
  WG_TRY(new_table(h, new_cap, &fresh));
  if (h->base) {
    const uint64_t words = (new_cap + kSideSlots) * (uint64_t)h->n_sumd * 2;
    wg_status st = dev_alloc(h->ctx, (void**)&fresh_base, words * 8, nullptr, "exact Int128 bases");
    if (st == WG_OK) st = zero_dev(h->ctx, fresh_base, words);
    if (st != WG_OK) {
      (void)hipFree(fresh);
      (void)hipFree(fresh_base);
      return st;
    }
  }
  uint64_t* old = h->slots;
  const uint64_t old_cap = h->cap;
  uint64_t* old_base = h->base;
  h->slots = fresh;
  h->cap = new_cap;
  h->base = fresh_base;
  // the rehash claims slots: reset the groups counter first
  WG_TRY(zero_dev(h->ctx, h->counters + CTR_GROUPS, 1));
  GlobalTable T = table_of(h);
  T.max_probe = new_cap;  // never refuse while moving into the larger table
  const uint64_t total = old_cap + kSideSlots;
  hipLaunchKernelGGL(agg_rehash_kernel, dim3(grid_for(total, 256, 8192)), dim3(256), 0, h->ctx->stream, old, old_cap,
                     old_base, T, h->base, h->pg.stride, h->n_sumd * 2);
  WG_HIP(hipGetLastError());
  WG_HIP(hipStreamSynchronize(h->ctx->stream));
  WG_HIP(hipFree(old));
  if (old_base) WG_HIP(hipFree(old_base));
  return WG_OK;
}

wg_status agg_ensure_capacity(wg_agg* h, uint64_t extra_groups) {
  // keep the table at <= 75 % load for what is known to be coming
  const uint64_t need = h->groups + extra_groups;
  if (need + need / 3 <= h->cap) return WG_OK;
  return grow_to(h, next_pow2(std::max<uint64_t>(need + need / 3, 1024)));
}

wg_status agg_ensure_dense(wg_agg* h, uint64_t rows, uint64_t keep) {
  if (h->dense_cap >= rows) return WG_OK;
  uint64_t* d = nullptr;
  uint64_t* db = nullptr;
  uint8_t* dn = nullptr;
  const uint64_t S = h->pg.stride, B = (uint64_t)h->n_sumd * 2;
  struct Release {  // a partial allocation is undone on every error path
    uint64_t** a;
    uint64_t** b;
    uint8_t** c;
    ~Release() {
      if (*a) (void)hipFree(*a);
      if (*b) (void)hipFree(*b);
      if (*c) (void)hipFree(*c);
    }
  } release{&d, &db, &dn};
  WG_ALLOC(h->ctx, &d, rows * S * 8, "the dense output");
  if (B) WG_ALLOC(h->ctx, &db, rows * B * 8, "the dense output");
  WG_ALLOC(h->ctx, &dn, rows, "the dense output");
  if (keep) {
    WG_HIP(hipMemcpyAsync(d, h->dense, keep * S * 8, hipMemcpyDeviceToDevice, h->ctx->stream));
    if (B) WG_HIP(hipMemcpyAsync(db, h->dense_base, keep * B * 8, hipMemcpyDeviceToDevice, h->ctx->stream));
    WG_HIP(hipMemcpyAsync(dn, h->dense_null, keep, hipMemcpyDeviceToDevice, h->ctx->stream));
  }
  WG_HIP(hipStreamSynchronize(h->ctx->stream));
  if (h->dense) WG_HIP(hipFree(h->dense));
  if (h->dense_base) WG_HIP(hipFree(h->dense_base));
  if (h->dense_null) WG_HIP(hipFree(h->dense_null));
  h->dense = d;
  h->dense_base = db;
  h->dense_null = dn;
  h->dense_cap = rows;
  d = db = nullptr;
  dn = nullptr;
  return WG_OK;
}

wg_status agg_settle(wg_agg* h) {
  if (h->n_dense == 0) return WG_OK;
  // dense rows carry zero exact bases: their limbs hold the whole sums
  const uint64_t n = h->n_dense;
  WG_TRY(agg_ensure_capacity(h, n));
  WG_TRY(agg_merge_partials(h, h->dense, n));
  h->n_dense = 0;
  h->dirty = true;
  return WG_OK;
}

wg_status normalize(wg_agg* h) {
  if (h->n_sumd == 0 || h->rows_since_norm == 0) {
    h->rows_since_norm = 0;
    return WG_OK;
  }
  WG_TRY(ensure_base(h));
  const uint64_t total = h->cap + kSideSlots;
  hipLaunchKernelGGL(agg_normalize_kernel, dim3(grid_for(total, 256, 8192)), dim3(256), 0, h->ctx->stream, h->slots,
                     total, h->base, h->pg.stride, (const int32_t*)h->d_sumd_word, (const int32_t*)h->d_sumd_limbs,
                     h->n_sumd);
  WG_HIP(hipGetLastError());
  h->base_valid = true;
  h->rows_since_norm = 0;
  return WG_OK;
}

wg_status agg_read_counters(wg_agg* h, uint64_t* c) {
  WG_HIP(hipMemcpyAsync(h->pinned, h->counters, CTR_N * 8, hipMemcpyDeviceToHost, h->ctx->stream));
  WG_HIP(hipStreamSynchronize(h->ctx->stream));
  memcpy(c, h->pinned, CTR_N * 8);
  return WG_OK;
}

wg_status ensure_overflow_buffers(wg_agg* h) { return ensure_overflow_for(h, h->grid, h->lds_slots); }

// Overflow lists for a launch of `grid` workgroups with `lds_slots`-slot LDS
// tables: row headroom for every in-flight tile, and room for every LDS entry
// a flush may spill (flushes write without a further bound check).
wg_status ensure_overflow_for(wg_agg* h, uint64_t grid, int32_t lds_slots) {
  const uint64_t rows_cap = std::max<uint64_t>(grid * h->tile_rows * 4, 1ull << 20);
  if (h->ovf_rows_cap < rows_cap) WG_TRY(regrow(h->ctx, &h->ovf_rows, &h->ovf_rows_cap, rows_cap, 4, "overflow rows"));
  const uint64_t part_cap = grid * (uint64_t)std::max(lds_slots, 1);
  if (h->ovf_part_cap < part_cap)
    WG_TRY(regrow(h->ctx, &h->ovf_part, &h->ovf_part_cap, part_cap, (uint64_t)h->pg.stride * 8, "overflow partials"));
  return WG_OK;
}

static void pick_launch(wg_agg* h) {
  // persistent grid: enough resident workgroups to fill every CU
  const int lds_bytes = (int)lds_table_bytes(h->lds_slots, h->pg.stride);
  int per_cu = 1;
  if (h->lds_slots == 0) per_cu = std::max(1, 2048 / h->block);
  else per_cu = std::max(1, std::min(160 * 1024 / std::max(lds_bytes, 1), 2048 / h->block));
  h->grid = h->ctx->cus * per_cu;
  h->tile_rows = (uint64_t)h->block * 16;
}

template <bool USE_LDS, int NV>
static wg_status launch_consume_t(int grid, int block, size_t lds, hipStream_t st, const ConsumeParams& P) {
  if (USE_LDS && lds > 64 * 1024)
    WG_HIP(hipFuncSetAttribute((const void*)agg_consume_kernel<USE_LDS, NV>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL((agg_consume_kernel<USE_LDS, NV>), dim3(grid), dim3(block), lds, st, P);
  return WG_OK;
}

static wg_status launch_consume(bool use_lds, int nv, int grid, int block, size_t lds, hipStream_t st,
                                const ConsumeParams& P) {
  switch (nv) {
    case 0: return use_lds ? launch_consume_t<true, 0>(grid, block, lds, st, P)
                           : launch_consume_t<false, 0>(grid, block, lds, st, P);
    case 1: return use_lds ? launch_consume_t<true, 1>(grid, block, lds, st, P)
                           : launch_consume_t<false, 1>(grid, block, lds, st, P);
    default: return use_lds ? launch_consume_t<true, 3>(grid, block, lds, st, P)
                            : launch_consume_t<false, 3>(grid, block, lds, st, P);
  }
}

// One launch over rows [row_base, row_base + n) or sel[0..n); reads back the
// counters (one stream sync).
static wg_status consume_once(wg_agg* h, const ColSet& cs, const uint32_t* sel, uint64_t row_base, uint64_t n,
                              const PredDesc* pred, uint64_t* c) {
  ConsumeParams P{};
  P.cs = cs;
  P.key = h->key;
  P.pred[0] = pred[0];
  P.pred[1] = pred[1];
  P.pg = h->pg;
  P.T = table_of(h);
  P.sel = sel;
  P.row_base = row_base;
  P.n = n;
  P.tile_rows = h->tile_rows;
  P.n_tiles = (n + h->tile_rows - 1) / h->tile_rows;
  P.ovf_rows = h->ovf_rows;
  P.ovf_rows_cap = h->ovf_rows_cap;
  P.ovf_part = h->ovf_part;
  P.ovf_part_cap = h->ovf_part_cap;
  P.lds_slots = h->lds_slots;
  WG_TRY(zero_dev(h->ctx, h->counters + CTR_OVF_ROWS, CTR_N - 1));  // keep CTR_GROUPS
  const int grid = (int)std::min<uint64_t>((uint64_t)h->grid, std::max<uint64_t>(P.n_tiles, 1));
  if (!h->ev0) {
    WG_HIP(hipEventCreate(&h->ev0));
    WG_HIP(hipEventCreate(&h->ev1));
  }
  // preload the values a row needs unless a column is DECIMAL128 or there
  // are more than kMaxVals of them
  std::vector<int> vcols;
  bool fast = true;
  AggProgram pg2 = h->pg;
  for (int o = 0; o < pg2.n_ops && fast; ++o)
    for (int32_t* c : {&pg2.ops[o].col, &pg2.ops[o].col2}) {
      if (*c < 0) continue;
      if (cs.cols[*c].type == WG_DECIMAL128) fast = false;
      int j = -1;
      for (size_t k = 0; k < vcols.size(); ++k)
        if (vcols[k] == *c) j = (int)k;
      if (j < 0) {
        j = (int)vcols.size();
        vcols.push_back(*c);
      }
      *c = j;
    }
  if (vcols.size() > 3) fast = false;  // more preloaded values would spill registers
  int nv = 0;
  if (fast) {
    P.pg = pg2;
    P.n_vals = (int32_t)vcols.size();
    for (size_t j = 0; j < vcols.size(); ++j) {
      P.val_col[j] = vcols[j];
      if (cs.cols[vcols[j]].type == WG_FLOAT64) P.vfloat |= 1u << j;
    }
    nv = vcols.size() <= 1 ? 1 : 3;
  }
  WG_HIP(hipEventRecord(h->ev0, h->ctx->stream));
  const size_t lds = h->lds_slots > 0 ? lds_table_bytes(h->lds_slots, h->pg.stride) : 0;
  h->touched = true;
  WG_TRY(launch_consume(h->lds_slots > 0, nv, grid, h->block, lds, h->ctx->stream, P));
  WG_HIP(hipGetLastError());
  WG_HIP(hipEventRecord(h->ev1, h->ctx->stream));
  h->dirty = true;
  WG_TRY(agg_read_counters(h, c));
  float ms = 0.f;
  WG_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->st_kernel_ms += ms;
  h->st_launches += 1;
  return WG_OK;
}

static wg_status consume_range(wg_agg* h, const ColSet& cs, const uint32_t* sel, uint64_t n, const PredDesc* pred,
                               int depth);

// Fold overflow partial records (possibly several per key) into the table
// with atomics; the table was grown to hold them.
// Merge a table of unique keys (cap slots + the side slots, exact Int128
// bases or null) into h: one claim per key, then plain combines.
wg_status agg_merge_slots(wg_agg* h, const uint64_t* src, uint64_t src_cap, const uint64_t* src_base,
                          uint64_t src_groups) {
  WG_TRY(agg_ensure_capacity(h, src_groups));
  WG_TRY(ensure_base(h));
  GlobalTable T = table_of(h);
  T.max_probe = h->cap;
  WG_TRY(zero_dev(h->ctx, h->counters + CTR_STOPPED, 1));
  const uint64_t total = src_cap + kSideSlots;
  h->touched = true;
  hipLaunchKernelGGL(agg_merge_table_kernel, dim3(grid_for(total, 256, 8192)), dim3(256), 0, h->ctx->stream, src,
                     src_cap, src_base, T, h->base, h->pg, (const int32_t*)h->d_sumd_word,
                     (const int32_t*)h->d_sumd_limbs, h->n_sumd);
  WG_HIP(hipGetLastError());
  uint64_t c[CTR_N];
  WG_TRY(agg_read_counters(h, c));
  WG_CHECK((c[CTR_STOPPED] >> 32) == 0, WG_ERR_DEVICE, "merge did not fit the destination table");
  h->groups = c[CTR_GROUPS];
  h->dirty = true;
  return WG_OK;
}

wg_status agg_merge_partials(wg_agg* h, const uint64_t* src, uint64_t n_part) {
  uint64_t* recs = nullptr;
  const uint64_t bytes = n_part * (uint64_t)h->pg.stride * 8;
  WG_ALLOC(h->ctx, &recs, bytes, "overflow records");  // out of the way of later launches
  WG_HIP(hipMemcpyAsync(recs, src, bytes, hipMemcpyDeviceToDevice, h->ctx->stream));
  GlobalTable T = table_of(h);
  T.max_probe = h->cap;
  h->touched = true;
  hipLaunchKernelGGL(agg_merge_part_kernel, dim3(grid_for(n_part, 256)), dim3(256), 0, h->ctx->stream, recs, n_part,
                     T, h->pg);
  WG_HIP(hipGetLastError());
  uint64_t c2[CTR_N];
  WG_TRY(agg_read_counters(h, c2));
  WG_HIP(hipFree(recs));
  WG_CHECK((c2[CTR_STOPPED] >> 32) == 0, WG_ERR_DEVICE, "overflow partials did not fit the grown table");
  h->groups = c2[CTR_GROUPS];
  return WG_OK;
}

// After a launch that overflowed: grow the table so that every group seen so
// far plus every pending row fits at <= 50 % load, fold the overflow partials
// in, then replay the overflow rows.
static wg_status resolve_overflow(wg_agg* h, const ColSet& cs, const PredDesc* pred, const uint64_t* c, int depth) {
  WG_CHECK(depth < 8, WG_ERR_DEVICE, "group table overflow did not converge");
  const uint64_t n_rows = c[CTR_OVF_ROWS], n_part = c[CTR_OVF_PART];
  h->groups = c[CTR_GROUPS];
  WG_TRY(grow_to(h, next_pow2(std::max<uint64_t>((h->groups + n_rows + n_part) * 2, h->cap * 2))));
  if (n_part) WG_TRY(agg_merge_partials(h, h->ovf_part, n_part));
  if (n_rows) {
    uint32_t* rows = nullptr;
    WG_ALLOC(h->ctx, &rows, n_rows * 4, "replayed rows");
    WG_HIP(hipMemcpyAsync(rows, h->ovf_rows, n_rows * 4, hipMemcpyDeviceToDevice, h->ctx->stream));
    wg_status st = consume_range(h, cs, rows, n_rows, pred, depth + 1);
    (void)hipStreamSynchronize(h->ctx->stream);
    (void)hipFree(rows);
    WG_TRY(st);
  }
  return WG_OK;
}

static wg_status consume_range(wg_agg* h, const ColSet& cs, const uint32_t* sel, uint64_t n, const PredDesc* pred,
                               int depth) {
  uint64_t pos = 0;
  while (pos < n) {
    uint64_t c[CTR_N];
    WG_TRY(consume_once(h, cs, sel ? sel + pos : nullptr, sel ? 0 : pos, n - pos, pred, c));
    const uint64_t n_tiles = (n - pos + h->tile_rows - 1) / h->tile_rows;
    const uint64_t tiles = std::min<uint64_t>(c[CTR_TILE], n_tiles);
    h->groups = c[CTR_GROUPS];
    h->st_lds += c[CTR_LDS_ROWS];
    h->st_glb += c[CTR_GLB_ROWS];
    h->st_flushed += c[CTR_FLUSHED];
    if (c[CTR_OVF_ROWS] || c[CTR_OVF_PART]) WG_TRY(resolve_overflow(h, cs, pred, c, depth));
    WG_CHECK(tiles > 0 || c[CTR_OVF_ROWS] > 0, WG_ERR_DEVICE, "consume launch made no progress");
    pos = std::min<uint64_t>(n, pos + tiles * h->tile_rows);
    if (pos < n) WG_TRY(agg_ensure_capacity(h, 0));
  }
  return WG_OK;
}

wg_status agg_consume_device(wg_agg* h, const ColSet& cs, const uint32_t* sel, uint64_t n, const PredDesc* pred) {
  WG_CHECK(!h->finalized, WG_ERR_STATE, "consume after finalize");
  if (n == 0) return WG_OK;
  WG_TRY(agg_settle(h));
  WG_CHECK(n < (1ull << 31), WG_ERR_INVALID, "batches of 2^31 rows or more must be split by the caller");
  if (h->rows_since_norm + n >= (1ull << 32) - 1) WG_TRY(normalize(h));
  WG_TRY(agg_ensure_capacity(h, 0));
  WG_TRY(ensure_overflow_buffers(h));
  h->st_lds = h->st_glb = h->st_flushed = 0;
  h->st_kernel_ms = 0.f;
  h->st_launches = 0;
  for (float& v : h->pass_ms) v = 0.f;
  // high-cardinality batches take the partitioned consume: rows of a key
  // meet in one workgroup instead of contending on shared slots
  const uint64_t est = std::max(h->group_hint, h->groups);
  const bool part = h->strategy == 2 ||
                    (h->strategy == 0 && n >= (1ull << 22) && est >= 8ull * (uint64_t)std::max(h->lds_slots, 256));
  if (part && part_eligible(h, cs)) {
    WG_TRY(agg_ensure_capacity(h, est));
    WG_TRY(agg_consume_partitioned(h, cs, sel, n, pred));
    h->st_kernel_ms = h->pass_ms[1] + h->pass_ms[2] + h->pass_ms[3];
    h->st_launches = 3;
  } else {
    WG_TRY(consume_range(h, cs, sel, n, pred, 0));
    h->pass_ms[0] = h->st_kernel_ms;
  }
  h->rows_since_norm += n;
  return WG_OK;
}

}  // namespace wg

// ============================================================ host staging

namespace wg {

static void stage_release(HostStage& S) {
  for (void* p : S.pinned) (void)hipHostFree(p);
  for (uint64_t* p : S.pvalid) (void)hipHostFree(p);
  for (void* p : S.dev) (void)hipFree(p);
  for (uint64_t* p : S.dvalid) (void)hipFree(p);
  S.pinned.clear();
  S.pvalid.clear();
  S.dev.clear();
  S.dvalid.clear();
  S.cap_rows = 0;
}

static wg_status stage_alloc(wg_agg* h, HostStage& S, uint64_t cap_rows) {
  S.pinned.assign(h->n_cols, nullptr);
  S.pvalid.assign(h->n_cols, nullptr);
  S.dev.assign(h->n_cols, nullptr);
  S.dvalid.assign(h->n_cols, nullptr);
  for (int c = 0; c < h->n_cols; ++c) {
    const uint64_t bytes = cap_rows * type_bytes(h->col_types[c]);
    WG_CHECK(hipHostMalloc(&S.pinned[c], bytes, hipHostMallocDefault) == hipSuccess, WG_ERR_OOM,
             "pinned staging of %llu bytes failed", (unsigned long long)bytes);
    WG_ALLOC(h->ctx, &S.dev[c], bytes, "host-batch staging");
    if (h->col_nullable[c]) {
      WG_CHECK(hipHostMalloc((void**)&S.pvalid[c], cap_rows / 8, hipHostMallocDefault) == hipSuccess, WG_ERR_OOM,
               "pinned staging of %llu bytes failed", (unsigned long long)(cap_rows / 8));
      WG_ALLOC(h->ctx, &S.dvalid[c], cap_rows / 8, "host-batch staging");
      memset(S.pvalid[c], 0, cap_rows / 8);
    }
  }
  S.cap_rows = cap_rows;
  return WG_OK;
}

// The staging buffers of host batches, all or none.
static wg_status stage_init(wg_agg* h) {
  HostStage& S = h->stage;
  if (S.cap_rows) return WG_OK;
  const wg_status st = stage_alloc(h, S, 1ull << 20);
  if (st != WG_OK) stage_release(S);
  return st;
}

wg_status agg_flush_stage(wg_agg* h) {
  HostStage& S = h->stage;
  if (S.rows == 0) return WG_OK;
  ColSet cs{};
  cs.n_cols = h->n_cols;
  cs.n_probe = h->n_cols;
  for (int c = 0; c < h->n_cols; ++c) {
    const uint64_t bytes = S.rows * type_bytes(h->col_types[c]);
    WG_HIP(hipMemcpyAsync(S.dev[c], S.pinned[c], bytes, hipMemcpyHostToDevice, h->ctx->stream));
    if (S.pvalid[c]) WG_HIP(hipMemcpyAsync(S.dvalid[c], S.pvalid[c], (S.rows + 63) / 64 * 8, hipMemcpyHostToDevice, h->ctx->stream));
    cs.cols[c] = ColDesc{S.dev[c], S.dvalid[c], h->col_types[c], h->col_scales[c]};
    cs.inline_word[c] = -1;
  }
  const PredDesc none[2] = {{-1, 0, 0, 0.0}, {-1, 0, 0, 0.0}};
  const uint64_t n = S.rows;
  if (h->wide) WG_TRY(wide_attach(h, cs, nullptr, n, n));
  // consume syncs on its counter read-back, after which the pinned buffers
  // are free; the staged rows stay staged until they have been consumed
  WG_TRY(agg_consume_device(h, cs, nullptr, n, none));
  S.rows = 0;
  for (int c = 0; c < h->n_cols; ++c)
    if (S.pvalid[c]) memset(S.pvalid[c], 0, (n + 63) / 64 * 8);
  return WG_OK;
}

static bool host_valid(const wg_column& c, uint64_t row) {
  return c.validity == nullptr || ((c.validity[row >> 6] >> (row & 63)) & 1ull);
}

template <typename T>
static bool cmp_host(T a, T b, int op) {
  switch (op) {
    case WG_EQ: return a == b;
    case WG_NE: return a != b;
    case WG_LT: return a < b;
    case WG_LE: return a <= b;
    case WG_GT: return a > b;
    default: return a >= b;
  }
}

static int64_t host_int(const wg_column& c, uint64_t row) {
  switch (c.type) {
    case WG_INT32: return ((const int32_t*)c.data)[row];
    case WG_UINT32: return ((const uint32_t*)c.data)[row];
    case WG_UINT16: return ((const uint16_t*)c.data)[row];
    default: return ((const int64_t*)c.data)[row];
  }
}

// Host batch -> staged rows: applies sel and the predicate (the KernelFilter
// the caller pushed down), copies the surviving rows and their validity bits.
static bool host_pred(const wg_column* cols, const wg_predicate& p, uint64_t row) {
  const wg_column& pc = cols[p.col];
  if (!host_valid(pc, row)) return false;
  if (pc.type == WG_FLOAT64) return cmp_host<double>(((const double*)pc.data)[row], p.f64, p.op);
  if (pc.type == WG_INT32 || pc.type == WG_UINT16) return cmp_host<int32_t>((int32_t)host_int(pc, row), (int32_t)p.i64, p.op);
  return cmp_host<int64_t>(host_int(pc, row), p.i64, p.op);
}

static wg_status consume_host(wg_agg* h, const wg_column* cols, uint64_t n_rows, const uint32_t* sel, uint64_t n_sel,
                              const wg_predicate* preds, int n_preds) {
  WG_TRY(stage_init(h));
  HostStage& S = h->stage;
  const uint64_t n = sel ? n_sel : n_rows;
  // multi-column keys: a NULL key is handed back before any state changes
  const KeyDesc& UK = h->wide ? h->wkey : h->key;
  if (UK.n > 1) {
    for (int k = 0; k < UK.n; ++k) {
      const wg_column& kc = cols[UK.col[k]];
      if (!kc.validity) continue;
      for (uint64_t i = 0; i < n; ++i)
        if (!host_valid(kc, sel ? sel[i] : i)) {
          set_error("NULL in a multi-column group key");
          return WG_ERR_NULL_KEY;
        }
    }
  }
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t row = sel ? sel[i] : i;
    bool pass = true;
    for (int k = 0; k < n_preds && pass; ++k) pass = host_pred(cols, preds[k], row);
    if (!pass) continue;
    const uint64_t d = S.rows;
    for (int c = 0; c < h->n_cols; ++c) {
      const int w = type_bytes(h->col_types[c]);
      memcpy((char*)S.pinned[c] + d * w, (const char*)cols[c].data + row * w, w);
      if (S.pvalid[c] && host_valid(cols[c], row)) S.pvalid[c][d >> 6] |= 1ull << (d & 63);
    }
    if (++S.rows == S.cap_rows) {
      const wg_status st = agg_flush_stage(h);
      if (st != WG_OK) {
        h->touched = true;  // part of this call's rows are staged or consumed
        return st;
      }
    }
  }
  return WG_OK;
}

wg_status agg_guard_begin(wg_agg* h) {
  WG_CHECK(!h->poisoned, WG_ERR_STATE,
           "the aggregate holds part of an input whose call failed; only wg_agg_reset or wg_agg_destroy apply");
  h->touched = false;
  return WG_OK;
}

wg_status agg_guard_end(wg_agg* h, wg_status st) {
  if (st != WG_OK && h->touched) {
    h->poisoned = true;
    h->dirty = true;  // the failed call may have left groups anywhere in the table: reset refills it all
    const std::string why = wg_last_error();
    set_error("%s (the aggregate now holds part of this call's input: only wg_agg_reset or wg_agg_destroy apply)",
              why.c_str());
  }
  h->touched = false;
  return st;
}

static wg_status free_dense(wg_agg* h) {
  if (h->dense) WG_HIP(hipFree(h->dense));
  if (h->dense_base) WG_HIP(hipFree(h->dense_base));
  if (h->dense_null) WG_HIP(hipFree(h->dense_null));
  h->dense = nullptr;
  h->dense_base = nullptr;
  h->dense_null = nullptr;
  h->dense_cap = 0;
  return WG_OK;
}

}  // namespace wg

// ================================================================= C-ABI

using namespace wg;

extern "C" {

wg_status wg_agg_create(wg_ctx* ctx, int32_t n_cols, const int32_t* col_types, const int32_t* col_scales,
                        const uint8_t* col_nullable, int32_t n_keys, const int32_t* key_cols, int32_t n_aggs,
                        const wg_agg_spec* aggs, uint64_t group_hint, wg_agg** out) {
  clear_error();
  WG_CHECK(ctx && out && col_types && key_cols && (n_aggs == 0 || aggs), WG_ERR_INVALID, "null argument");
  WG_CHECK(n_cols > 0 && n_cols <= kMaxCols, WG_ERR_UNSUPPORTED, "%d columns (max %d)", n_cols, kMaxCols);
  for (int c = 0; c < n_cols; ++c)
    WG_CHECK(col_types[c] >= WG_INT32 && col_types[c] <= WG_UINT16, WG_ERR_UNSUPPORTED, "column %d: type %d", c, col_types[c]);
  wg_agg* h = new wg_agg();
  h->ctx = ctx;
  h->n_cols = n_cols;
  for (int c = 0; c < n_cols; ++c) {
    h->col_types[c] = col_types[c];
    h->col_scales[c] = col_scales ? col_scales[c] : 0;
    h->col_nullable[c] = col_nullable ? col_nullable[c] : 0;
  }
  h->aggs.assign(aggs, aggs + n_aggs);
  wg_status st = compile_key(h, n_keys, key_cols);
  if (st == WG_OK) st = compile_program(h);
  if (st != WG_OK) {
    delete h;
    return st;
  }
  if (h->n_sumd) {
    st = dev_alloc(ctx, (void**)&h->d_sumd_word, h->n_sumd * 8, nullptr, "the DECIMAL sum layout");
    if (st != WG_OK) {
      wg_agg_destroy(h);
      return st;
    }
    h->d_sumd_limbs = h->d_sumd_word + h->n_sumd;
    std::vector<int32_t> tmp(h->sumd_word);
    tmp.insert(tmp.end(), h->sumd_limbs.begin(), h->sumd_limbs.end());
    WG_HIP(hipMemcpy(h->d_sumd_word, tmp.data(), h->n_sumd * 8, hipMemcpyHostToDevice));
  }
  st = dev_alloc(ctx, (void**)&h->counters, kAggCounterWords * 8, nullptr, "the aggregate's counters");
  if (st == WG_OK && hipHostMalloc((void**)&h->pinned, kAggPinnedWords * 8, hipHostMallocDefault) != hipSuccess) {
    set_error("pinned allocation of the counters failed");
    st = WG_ERR_OOM;
  }
  if (st != WG_OK) {
    wg_agg_destroy(h);
    return st;
  }
  WG_TRY(zero_dev(ctx, h->counters, kAggCounterWords));
  const uint64_t hint = std::max<uint64_t>(group_hint, 1024);
  h->group_hint = group_hint;
  st = alloc_table(h, next_pow2(hint + hint / 4));
  if (st == WG_OK && h->wide) st = wide_init(h);
  if (st != WG_OK) {
    wg_agg_destroy(h);
    return st;
  }
  pick_launch(h);
  WG_HIP(hipStreamSynchronize(ctx->stream));
  *out = h;
  return WG_OK;
}

wg_status wg_agg_destroy(wg_agg* h) {
  if (!h) return WG_OK;
  (void)hipStreamSynchronize(h->ctx->stream);
  (void)hipFree(h->slots);
  (void)hipFree(h->base);
  (void)hipFree(h->counters);
  (void)hipFree(h->ovf_rows);
  (void)hipFree(h->ovf_part);
  (void)hipFree(h->d_sumd_word);
  (void)free_dense(h);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  for (hipEvent_t e : h->pev)
    if (e) (void)hipEventDestroy(e);
  if (h->side) {
    (void)hipStreamSynchronize(h->side);
    (void)hipStreamDestroy(h->side);
  }
  (void)hipFree(h->part_cnt);
  (void)hipFree(h->part_scan);
  (void)hipFree(h->part_buf);
  (void)hipFree(h->part_items);
  (void)hipFree(h->hot_buf);
  if (h->part_host) (void)hipHostFree(h->part_host);
  if (h->pinned) (void)hipHostFree(h->pinned);
  stage_release(h->stage);
  if (h->wide) (void)wide_free(h);
  delete h;
  return WG_OK;
}

static wg_status consume_call(wg_agg* h, const wg_column* cols, uint64_t n_rows, const uint32_t* sel, uint64_t n_sel,
                              const wg_predicate* preds, int32_t n_preds, int32_t mem) {
  WG_CHECK(n_preds >= 0 && n_preds <= 2 && (n_preds == 0 || preds), WG_ERR_INVALID, "%d predicates (max 2)", n_preds);
  for (int c = 0; c < h->n_cols; ++c) {
    WG_CHECK(cols[c].type == h->col_types[c], WG_ERR_INVALID, "column %d: type %d, created as %d", c, cols[c].type, h->col_types[c]);
    WG_CHECK(cols[c].validity == nullptr || h->col_nullable[c], WG_ERR_INVALID, "column %d has a validity bitmap but was created non-nullable", c);
    WG_CHECK(cols[c].data != nullptr || (sel ? n_sel : n_rows) == 0, WG_ERR_INVALID, "column %d: null data", c);
  }
  PredDesc pd[2] = {{-1, 0, 0, 0.0}, {-1, 0, 0, 0.0}};
  int np = 0;
  for (int k = 0; k < n_preds; ++k) {
    const wg_predicate& p = preds[k];
    if (p.col < 0) continue;
    WG_CHECK(p.col < h->n_cols, WG_ERR_INVALID, "predicate column out of range");
    const int t = h->col_types[p.col];
    WG_CHECK(t != WG_DECIMAL && t != WG_DECIMAL128, WG_ERR_UNSUPPORTED, "no filter kernel for DECIMAL");
    WG_CHECK(p.op >= WG_EQ && p.op <= WG_GE, WG_ERR_INVALID, "bad compare op");
    pd[np++] = PredDesc{p.col, p.op, p.i64, p.f64};
  }
  if (mem == WG_MEM_HOST) {
    wg_predicate hp[2];
    for (int k = 0; k < np; ++k) hp[k] = wg_predicate{pd[k].col, pd[k].op, pd[k].i64, pd[k].f64};
    return consume_host(h, cols, n_rows, sel, n_sel, hp, np);
  }
  WG_CHECK(mem == WG_MEM_DEVICE, WG_ERR_INVALID, "bad mem kind %d", mem);
  const uint64_t n = sel ? n_sel : n_rows;
  if (n == 0) return WG_OK;
  const KeyDesc& UK = h->wide ? h->wkey : h->key;
  if (UK.n > 1) {
    for (int k = 0; k < UK.n; ++k) {
      const wg_column& kc = cols[UK.col[k]];
      if (!kc.validity) continue;
      uint64_t* flag = h->counters + CTR_STOPPED;
      WG_TRY(zero_dev(h->ctx, flag, 1));
      hipLaunchKernelGGL(any_null_kernel, dim3(grid_for(n, 256)), dim3(256), 0, h->ctx->stream, kc.validity, sel, n, flag);
      WG_HIP(hipGetLastError());
      uint64_t c[CTR_N];
      WG_TRY(agg_read_counters(h, c));
      if (c[CTR_STOPPED]) {
        set_error("NULL in a multi-column group key");
        return WG_ERR_NULL_KEY;
      }
    }
  }
  ColSet cs{};
  cs.n_cols = h->n_cols;
  cs.n_probe = h->n_cols;
  for (int c = 0; c < h->n_cols; ++c) {
    cs.cols[c] = ColDesc{cols[c].data, cols[c].validity, cols[c].type, cols[c].scale};
    cs.inline_word[c] = -1;
  }
  WG_TRY(agg_flush_stage(h));
  if (h->wide) WG_TRY(wide_attach(h, cs, sel, n, n_rows));
  return agg_consume_device(h, cs, sel, n, pd);
}

wg_status wg_agg_consume(wg_agg* h, const wg_column* cols, uint64_t n_rows, const uint32_t* sel, uint64_t n_sel,
                         const wg_predicate* preds, int32_t n_preds, int32_t mem) {
  clear_error();
  WG_CHECK(h && cols, WG_ERR_INVALID, "null argument");
  WG_CHECK(!h->finalized, WG_ERR_STATE, "consume after finalize");
  WG_TRY(agg_guard_begin(h));
  return agg_guard_end(h, consume_call(h, cols, n_rows, sel, n_sel, preds, n_preds, mem));
}

static wg_status merge_call(wg_agg* dst, wg_agg* src) {
  WG_CHECK(dst->pg.stride == src->pg.stride && dst->pg.n_ops == src->pg.n_ops && dst->ctx == src->ctx,
           WG_ERR_INVALID, "merge of aggregates with different specs");
  WG_CHECK(dst->wide == src->wide, WG_ERR_INVALID, "merge of aggregates with different keys");
  WG_TRY(agg_flush_stage(dst));
  WG_TRY(agg_flush_stage(src));
  WG_TRY(agg_settle(dst));
  WG_TRY(agg_settle(src));
  if (dst->wide) {
    // src's keys are ids of its own dictionary: rewrite them as dst's ids
    for (int k = 0; k < dst->wkey.n; ++k)
      WG_CHECK(dst->wkey.word[k] == src->wkey.word[k] && dst->wkey.shift[k] == src->wkey.shift[k] &&
                   dst->wkey.narrow[k] == src->wkey.narrow[k],
               WG_ERR_INVALID, "merge of aggregates with different key layouts");
    uint64_t* remapped = nullptr;
    WG_TRY(wide_remap_slots(dst, src, &remapped));
    const wg_status st = agg_merge_slots(dst, remapped, src->cap, src->base, src->groups);
    (void)hipFree(remapped);
    WG_TRY(st);
    return wg_agg_reset(src);
  }
  WG_TRY(agg_merge_slots(dst, src->slots, src->cap, src->base, src->groups));
  return wg_agg_reset(src);
}

wg_status wg_agg_merge(wg_agg* dst, wg_agg* src) {
  clear_error();
  WG_CHECK(dst && src && dst != src, WG_ERR_INVALID, "bad handles");
  WG_CHECK(!dst->finalized && !src->finalized, WG_ERR_STATE, "merge after finalize");
  WG_CHECK(!src->poisoned, WG_ERR_STATE, "merge of an aggregate that holds part of a failed call's input");
  WG_TRY(agg_guard_begin(dst));
  return agg_guard_end(dst, merge_call(dst, src));
}

static wg_status finalize_call(wg_agg* h) {
  WG_TRY(agg_flush_stage(h));
  // rows [0, n_dense) already hold the groups a partitioned consume emitted;
  // the table's live slots are compacted after them
  const uint64_t need = h->n_dense + h->groups + kSideSlots;
  WG_TRY(agg_ensure_dense(h, need, h->n_dense));
  uint64_t* out_n = h->counters + CTR_LDS_ROWS;  // reuse a counter word
  h->pinned[0] = h->n_dense;
  WG_HIP(hipMemcpyAsync(out_n, h->pinned, 8, hipMemcpyHostToDevice, h->ctx->stream));
  // an untouched table: only its side slots can be live
  const uint64_t first = h->dirty ? 0 : h->cap;
  const uint64_t total = h->cap + kSideSlots - first;
  const uint64_t per_block = std::max<uint64_t>(4096, (total + 8191) / 8192);
  const uint64_t blocks = (total + per_block - 1) / per_block;
  hipLaunchKernelGGL(agg_compact_kernel, dim3((unsigned)blocks), dim3(kCompactBlock), 0, h->ctx->stream, h->slots,
                     h->cap, first, (const uint64_t*)h->base, h->pg.stride, h->n_sumd * 2, h->dense, h->dense_base,
                     h->dense_null, out_n, per_block);
  WG_HIP(hipGetLastError());
  uint64_t c[CTR_N];
  WG_TRY(agg_read_counters(h, c));
  h->out_n = c[CTR_LDS_ROWS];
  h->finalized = true;
  return WG_OK;
}

wg_status wg_agg_finalize(wg_agg* h) {
  clear_error();
  WG_CHECK(h, WG_ERR_INVALID, "null handle");
  if (h->finalized) return WG_OK;
  WG_TRY(agg_guard_begin(h));
  return agg_guard_end(h, finalize_call(h));
}

wg_status wg_agg_num_groups(wg_agg* h, uint64_t* out) {
  WG_CHECK(h && out, WG_ERR_INVALID, "null argument");
  WG_CHECK(h->finalized, WG_ERR_STATE, "num_groups before finalize");
  *out = h->out_n;
  return WG_OK;
}

wg_status wg_agg_fetch(wg_agg* h, uint64_t offset, uint64_t count, void* const* key_out, void* const* val_out,
                       uint8_t* const* valid_out, int32_t mem) {
  clear_error();
  WG_CHECK(h, WG_ERR_INVALID, "null handle");
  WG_CHECK(h->finalized, WG_ERR_STATE, "fetch before finalize");
  WG_CHECK(offset + count <= h->out_n, WG_ERR_INVALID, "fetch [%llu,+%llu) of %llu groups",
           (unsigned long long)offset, (unsigned long long)count, (unsigned long long)h->out_n);
  if (count == 0) return WG_OK;
  const bool host = mem == WG_MEM_HOST;
  if (!host && (int)h->finals.size() <= kMaxOps) {
    // device columns: one pass over the groups writes every output
    const KeyDesc& K = h->wide ? h->wkey : h->key;
    EmitParams E{};
    E.dense = h->dense;
    E.dense_base = h->dense_base;
    E.dense_null = h->dense_null;
    E.dict = h->wide ? (const uint64_t*)h->dict.keys : nullptr;
    E.off = offset;
    E.cnt = count;
    E.stride = h->pg.stride;
    E.n_sumd = h->n_sumd;
    E.K = K;
    E.n_keys = K.n;
    for (int k = 0; k < K.n; ++k) E.key_out[k] = key_out ? key_out[k] : nullptr;
    E.n_aggs = (int)h->finals.size();
    for (int a = 0; a < E.n_aggs; ++a) {
      const AggFinal& f = h->finals[a];
      E.fd[a] = FinalDesc{f.func, f.count_word, f.val_kind, f.val_word, f.is_float_in, f.sumd_index, f.scale, f.limbs};
      E.val_out[a] = val_out ? val_out[a] : nullptr;
      E.valid_out[a] = valid_out ? valid_out[a] : nullptr;
      const bool fl = f.func == WG_AGG_SUM || f.func == WG_AGG_AVG ||
                      ((f.func == WG_AGG_MIN || f.func == WG_AGG_MAX) && f.is_float_in);
      E.want_float |= (uint32_t)fl << a;
    }
    const dim3 grid(grid_for(count, 256, 16384)), block(256);
    switch (h->pg.stride) {
      case 2: hipLaunchKernelGGL(agg_emit_kernel<2>, grid, block, 0, h->ctx->stream, E); break;
      case 4: hipLaunchKernelGGL(agg_emit_kernel<4>, grid, block, 0, h->ctx->stream, E); break;
      case 8: hipLaunchKernelGGL(agg_emit_kernel<8>, grid, block, 0, h->ctx->stream, E); break;
      default: hipLaunchKernelGGL(agg_emit_kernel<16>, grid, block, 0, h->ctx->stream, E); break;
    }
    WG_HIP(hipGetLastError());
    WG_HIP(hipStreamSynchronize(h->ctx->stream));
    return WG_OK;
  }
  void* tmp = nullptr;
  uint8_t* tmpv = nullptr;
  if (host) {
    WG_ALLOC(h->ctx, &tmp, count * 8, "fetch staging");
    WG_ALLOC(h->ctx, &tmpv, count, "fetch staging");
  }
  auto finish = [&](void* dst, const void* src, uint64_t bytes) -> wg_status {
    if (!host || !dst) return WG_OK;
    WG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->ctx->stream));
    WG_HIP(hipStreamSynchronize(h->ctx->stream));
    return WG_OK;
  };
  wg_status st = WG_OK;
  const int g = grid_for(count, 256);
  const KeyDesc& UK = h->wide ? h->wkey : h->key;
  for (int k = 0; k < UK.n && st == WG_OK; ++k) {
    void* kout = key_out ? key_out[k] : nullptr;
    if (!kout) continue;
    hipLaunchKernelGGL(agg_key_kernel, dim3(g), dim3(256), 0, h->ctx->stream, (const uint64_t*)h->dense,
                       (const uint8_t*)h->dense_null, offset, count, h->pg.stride, UK, k, host ? tmp : kout,
                       (uint8_t*)nullptr, h->wide ? (const uint64_t*)h->dict.keys : (const uint64_t*)nullptr);
    if (hipGetLastError() != hipSuccess) { set_error("agg_key_kernel launch failed"); st = WG_ERR_DEVICE; break; }
    st = finish(kout, tmp, count * type_bytes(UK.key_type[k]));
  }
  for (size_t a = 0; a < h->finals.size() && st == WG_OK; ++a) {
    void* vout = val_out ? val_out[a] : nullptr;
    uint8_t* vv = valid_out ? valid_out[a] : nullptr;
    if (!vout && !vv) continue;
    const AggFinal& f = h->finals[a];
    FinalDesc fd{f.func, f.count_word, f.val_kind, f.val_word, f.is_float_in, f.sumd_index, f.scale, f.limbs};
    const int want_float = (f.func == WG_AGG_SUM || f.func == WG_AGG_AVG || ((f.func == WG_AGG_MIN || f.func == WG_AGG_MAX) && f.is_float_in)) ? 1 : 0;
    hipLaunchKernelGGL(agg_decode_kernel, dim3(g), dim3(256), 0, h->ctx->stream, (const uint64_t*)h->dense,
                       (const uint64_t*)h->dense_base, offset, count, h->pg.stride, h->n_sumd, fd,
                       (double*)(host ? tmp : vout), (int64_t*)(host ? tmp : vout), host ? tmpv : vv, want_float);
    if (hipGetLastError() != hipSuccess) { set_error("agg_decode_kernel launch failed"); st = WG_ERR_DEVICE; break; }
    st = finish(vout, tmp, count * 8);
    if (st == WG_OK) st = finish(vv, tmpv, count);
  }
  if (st == WG_OK && key_out && h->key.n == 1 && valid_out == nullptr) {
    // nothing: key validity comes through wg_agg_fetch_key_valid
  }
  if (host) {
    (void)hipFree(tmp);
    (void)hipFree(tmpv);
  } else {
    WG_HIP(hipStreamSynchronize(h->ctx->stream));
  }
  return st;
}

wg_status wg_agg_fetch_state(wg_agg* h, int32_t agg, uint64_t offset, uint64_t count, int64_t* count_out,
                             uint64_t* sum_lo_out, int64_t* sum_hi_out, int64_t* minmax_out, int32_t mem) {
  clear_error();
  WG_CHECK(h, WG_ERR_INVALID, "null handle");
  WG_CHECK(h->finalized, WG_ERR_STATE, "fetch before finalize");
  WG_CHECK(agg >= 0 && agg < (int)h->finals.size(), WG_ERR_INVALID, "aggregate %d", agg);
  WG_CHECK(offset + count <= h->out_n, WG_ERR_INVALID, "range past the groups");
  if (count == 0) return WG_OK;
  // the raw state is decoded on the host from the dense slots
  const int S = h->pg.stride;
  std::vector<uint64_t> slots(count * S), base;
  WG_HIP(hipMemcpyAsync(slots.data(), h->dense + offset * S, count * S * 8, hipMemcpyDeviceToHost, h->ctx->stream));
  if (h->n_sumd) {
    base.resize(count * h->n_sumd * 2);
    WG_HIP(hipMemcpyAsync(base.data(), h->dense_base + offset * h->n_sumd * 2, base.size() * 8, hipMemcpyDeviceToHost, h->ctx->stream));
  }
  WG_HIP(hipStreamSynchronize(h->ctx->stream));
  const AggFinal& f = h->finals[agg];
  std::vector<int64_t> cnt(count), mm(count, 0), hi(count, 0);
  std::vector<uint64_t> lo(count, 0);
  for (uint64_t j = 0; j < count; ++j) {
    const uint64_t* s = slots.data() + j * S;
    cnt[j] = (int64_t)s[f.count_word];
    if (f.val_kind == OP_SUMI || f.val_kind == OP_SUMF) {
      lo[j] = s[f.val_word];
    } else if (f.val_kind == OP_SUMD || f.val_kind == OP_SUMD4) {
      uint64_t l;
      int64_t hh;
      limbs_value(s + f.val_word, f.limbs, l, hh);
      const uint64_t* b = base.data() + (j * h->n_sumd + f.sumd_index) * 2;
      unsigned __int128 v = (((unsigned __int128)(uint64_t)hh << 64) | l) + (((unsigned __int128)b[1] << 64) | b[0]);
      lo[j] = (uint64_t)v;
      hi[j] = (int64_t)(uint64_t)(v >> 64);
    } else if (f.val_kind == OP_MIN || f.val_kind == OP_MAX) {
      const uint64_t e = s[f.val_word];
      mm[j] = f.is_float_in ? (int64_t)dec_f64_bits(e) : dec_i64(e);
    }
  }
  auto put = [&](void* dst, const void* src, uint64_t bytes) -> wg_status {
    if (!dst) return WG_OK;
    if (mem == WG_MEM_HOST) memcpy(dst, src, bytes);
    else WG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return WG_OK;
  };
  WG_TRY(put(count_out, cnt.data(), count * 8));
  WG_TRY(put(sum_lo_out, lo.data(), count * 8));
  WG_TRY(put(sum_hi_out, hi.data(), count * 8));
  WG_TRY(put(minmax_out, mm.data(), count * 8));
  return WG_OK;
}

wg_status wg_agg_fetch_key_valid(wg_agg* h, uint64_t offset, uint64_t count, uint8_t* valid_out, int32_t mem) {
  clear_error();
  WG_CHECK(h && valid_out, WG_ERR_INVALID, "null argument");
  WG_CHECK(h->finalized, WG_ERR_STATE, "fetch before finalize");
  WG_CHECK(offset + count <= h->out_n, WG_ERR_INVALID, "range past the groups");
  if (count == 0) return WG_OK;
  std::vector<uint8_t> nul(count);
  WG_HIP(hipMemcpyAsync(nul.data(), h->dense_null + offset, count, hipMemcpyDeviceToHost, h->ctx->stream));
  WG_HIP(hipStreamSynchronize(h->ctx->stream));
  for (auto& v : nul) v = v ? 0 : 1;
  if (mem == WG_MEM_HOST) memcpy(valid_out, nul.data(), count);
  else WG_HIP(hipMemcpy(valid_out, nul.data(), count, hipMemcpyHostToDevice));
  return WG_OK;
}

wg_status wg_agg_reset(wg_agg* h) {
  clear_error();
  WG_CHECK(h, WG_ERR_INVALID, "null handle");
  h->stage.rows = 0;
  if (h->dirty) WG_TRY(fill_table(h, h->slots, h->cap));
  else if (h->side_dirty) WG_TRY(fill_table(h, h->slots + h->cap * (uint64_t)h->pg.stride, 0));
  if (h->base) WG_TRY(zero_dev(h->ctx, h->base, (h->cap + kSideSlots) * (uint64_t)h->n_sumd * 2));
  WG_TRY(zero_dev(h->ctx, h->counters, kAggCounterWords));
  h->dirty = false;
  h->side_dirty = false;
  h->n_dense = 0;
  h->groups = 0;
  h->rows_since_norm = 0;
  h->finalized = false;
  h->out_n = 0;
  if (h->wide) WG_TRY(wide_reset(h));
  h->poisoned = false;
  return WG_OK;
}

wg_status wg_agg_footprint(wg_agg* h, uint64_t* bytes) {
  WG_CHECK(h && bytes, WG_ERR_INVALID, "null argument");
  *bytes = agg_footprint(h);
  return WG_OK;
}

wg_status wg_agg_state_bytes(wg_agg* h, uint64_t* bytes) {
  clear_error();
  WG_CHECK(h && bytes, WG_ERR_INVALID, "null argument");
  if (!h->finalized) WG_TRY(agg_flush_stage(h));
  const uint64_t groups = h->finalized ? h->out_n : h->groups + h->n_dense;
  *bytes = groups * ((uint64_t)h->pg.stride * 8 + (uint64_t)h->n_sumd * 16);
  return WG_OK;
}

wg_status wg_agg_set_tuning(wg_agg* h, int32_t lds_slots, int32_t block) {
  clear_error();
  WG_CHECK(h, WG_ERR_INVALID, "null handle");
  WG_CHECK(block == 256 || block == 512 || block == 1024, WG_ERR_INVALID, "block %d", block);
  WG_CHECK(lds_slots == 0 || (lds_slots >= 64 && (lds_slots & (lds_slots - 1)) == 0), WG_ERR_INVALID,
           "lds_slots %d must be 0 or a power of two >= 64", lds_slots);
  WG_CHECK(lds_slots == 0 || lds_table_bytes(lds_slots, h->pg.stride) <= 160 * 1024 - 1024, WG_ERR_INVALID,
           "LDS table of %d slots x %d words exceeds the workgroup LDS", lds_slots, h->pg.stride);
  h->lds_slots = lds_slots;
  h->block = block;
  pick_launch(h);
  return WG_OK;
}

wg_status wg_agg_pass_times(wg_agg* h, float* ms4) {
  WG_CHECK(h && ms4, WG_ERR_INVALID, "null argument");
  for (int i = 0; i < 4; ++i) ms4[i] = h->pass_ms[i];
  return WG_OK;
}

wg_status wg_agg_set_strategy(wg_agg* h, int32_t strategy, int32_t part_lds_slots, int32_t part_block,
                              uint64_t part_chunk_rows) {
  clear_error();
  WG_CHECK(h, WG_ERR_INVALID, "null handle");
  WG_CHECK(strategy >= 0 && strategy <= 2, WG_ERR_INVALID, "strategy %d", strategy);
  WG_CHECK(part_block == 256 || part_block == 512 || part_block == 1024, WG_ERR_INVALID, "block %d", part_block);
  WG_CHECK(part_lds_slots >= 64 && (part_lds_slots & (part_lds_slots - 1)) == 0 &&
               lds_table_bytes(part_lds_slots, h->pg.stride) <= 150 * 1024,
           WG_ERR_INVALID, "partition LDS table of %d slots does not fit", part_lds_slots);
  WG_CHECK(part_chunk_rows >= 1024 && part_chunk_rows < (1ull << 31), WG_ERR_INVALID, "chunk rows");
  h->strategy = strategy;
  h->part_lds_slots = part_lds_slots;
  h->part_block = part_block;
  h->part_chunk_rows = part_chunk_rows;
  return WG_OK;
}

wg_status wg_agg_set_heavy_hitters(wg_agg* h, int32_t hot_slots) {
  clear_error();
  WG_CHECK(h, WG_ERR_INVALID, "null handle");
  WG_CHECK(hot_slots == 0 || (hot_slots >= 64 && hot_slots <= 2048 && (hot_slots & (hot_slots - 1)) == 0),
           WG_ERR_INVALID, "hot slots %d: 0 or a power of two in [64, 2048]", hot_slots);
  h->part_hot = hot_slots;
  return WG_OK;
}

wg_status wg_agg_kernel_time(wg_agg* h, float* ms, uint32_t* launches) {
  WG_CHECK(h && ms, WG_ERR_INVALID, "null argument");
  *ms = h->st_kernel_ms;
  if (launches) *launches = h->st_launches;
  return WG_OK;
}

wg_status wg_agg_stats(wg_agg* h, uint64_t* lds_rows, uint64_t* global_rows, uint64_t* flushed, uint64_t* capacity) {
  WG_CHECK(h, WG_ERR_INVALID, "null handle");
  if (lds_rows) *lds_rows = h->st_lds;
  if (global_rows) *global_rows = h->st_glb;
  if (flushed) *flushed = h->st_flushed;
  if (capacity) *capacity = h->cap;
  return WG_OK;
}

}  // extern "C"
