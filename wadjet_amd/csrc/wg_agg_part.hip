// wg_agg_part.hip — partitioned consume for high-cardinality GROUP BY.
//
// The direct consume folds every row into one shared table with device
// atomics; when a batch holds millions of groups with a skewed (Zipf) key,
// rows of warm keys queue up on the same slot and on the same L2/memory-side
// atomic unit. This path brings the rows of a key together first, so every
// key is aggregated by ONE workgroup in LDS. A key's final partition is the
// top pbits bits of its hash; pass 1 sorts by the top b1 of them (level 1),
// pass 2 by the rest (level 2):
//   1  part_l1_*kernel   each workgroup owns a contiguous row range and walks
//                        it in tiles: a tile's cold records (packed key, the
//                        value words the program reads) are ranked per
//                        level-1 partition in LDS and written as one block
//                        sorted by level-1 partition into the workgroup's own
//                        region; per-final-partition counts go to a histogram
//      part_tot_kernel   + exclusive scan: final partition boundaries
//   2  part_l2_kernel    per (level-1 partition, tile range): gathers its
//                        runs, ranks each record within its final partition
//                        and writes it into that partition's range, so every
//                        final partition is contiguous
//   3  part_agg_kernel   one workgroup per partition chunk aggregates it in
//                        an LDS table and writes its groups to the dense
//                        output (a partition it owns alone) or merges the
//                        table into the global table (atomics for the chunks
//                        of a large partition split over several workgroups)
// Both sorting passes write blocks, not scattered records: the writes of one
// tile or round land in a few kilobytes that L2 completes into whole lines,
// where one pass of scattered 32-byte records over thousands of partitions
// leaves every line partly written. Rows whose key is NULL or equals the
// table sentinel go straight to the side slots in pass 1. Results are the
// same accumulator words as the direct path.
//
// Heavy hitters: under a skewed key (Zipf) a few hundred keys carry half the
// rows. Before pass 1 a sample of the batch (kSampleTiles spread tiles) is
// counted and the keys seen most often (at most H of them) become the hot
// set. Pass 1 looks every row up in an LDS copy of the hot set: hot rows are
// not written out but aggregated in an LDS table of H slots; each workgroup
// writes its H partial slots out, and hot_merge_kernel folds them into the
// global table (or the dense output). Only the cold rows pay for the record
// traffic of passes 1-3.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "wg_agg_dev.hpp"
#include "wg_agg_host.hpp"

namespace wg {

constexpr int kAggU = 2;         // records per thread per step of the aggregate loop
constexpr int kL1Block = 512;       // pass-1 workgroup size, general loader
// and typed loads: 768 threads at up to 80 registers (amdgpu_waves_per_eu(6)
// below), so two workgroups with their LDS heavy-hitter tables keep six
// waves per SIMD resident
constexpr int kL1TypedBlock = 768;

struct PartParams {
  ColSet cs;
  KeyDesc key;
  PredDesc pred[2];
  AggProgram pg;  // for side-slot rows in pass 1 (original columns)
  GlobalTable T;
  const uint32_t* sel;
  uint64_t row_base;  // always 0: rows are sel[i] or i
  uint64_t n;
  uint64_t rows_per_wg;   // a multiple of the tile rows
  uint32_t tiles_per_wg;
  int32_t pbits;          // final partition = top pbits bits of the key hash
  int32_t b1;             // level-1 partition = top b1 of those (b1 <= 7)
  int32_t n_vals;
  uint32_t* hist;         // [G][2^pbits] cold rows per workgroup and final partition
  uint32_t* d1_off;       // [tiles][2^b1 + 1] per tile: run offsets, then its record count
  uint32_t* d1_base;      // [tiles] first position of each tile's block
  uint64_t* recs;         // pass-1 records: key word, value words, zero padding
  int32_t val_col[kMaxVals];
  // heavy hitters (hot_slots = 0: none)
  AggProgram pgv;            // pg with op columns remapped onto the value words
  uint32_t vfloat;           // bit j: value word j holds float64 bits
  const uint64_t* hot_keys;  // the hot set, *n_hot keys
  const uint32_t* n_hot;
  int32_t hot_slots;         // H: LDS accumulator slots of pass 1 (>= *n_hot)
  int32_t hot_bits;          // the LDS lookup of the hot set has 2^hot_bits entries
  uint64_t* hot_part;        // [G][H][stride] per-workgroup partial slots of pass 1
};

// The hot set in LDS: a bit filter, the lookup (2^hot_bits keys and their hot
// indices) and the hot accumulators (struct of arrays, H slots + one spare
// slot per lane).
// A bit filter over 16 hash bits answers most cold rows with one LDS read.
struct HotLds {
  uint32_t* bits;  // kHotFilterBits bits: set for the hash of every hot key
  uint64_t* key;
  uint32_t* idx;
  uint64_t* acc;
  uint32_t mask;
  int HW;  // word stride of acc
};

constexpr uint32_t kHotFilterBits = 1u << 16;

__host__ __device__ inline size_t hot_lookup_bytes(int hot_bits) {
  return kHotFilterBits / 8 + ((size_t)12 << hot_bits);
}

__device__ __forceinline__ HotLds hot_lds(uint32_t* base, uint32_t np, int hot_bits, int H) {
  HotLds L;
  L.bits = base + np;
  L.key = (uint64_t*)(L.bits + kHotFilterBits / 32);  // np >= 64: 8-byte aligned
  L.idx = (uint32_t*)(L.key + (1u << hot_bits));
  L.acc = (uint64_t*)(L.idx + (1u << hot_bits));  // 2^hot_bits >= 2: 8-byte aligned
  L.mask = (1u << hot_bits) - 1;
  L.HW = lds_word_stride(H);
  return L;
}

// Fill the LDS lookup with the hot keys (distinct, never the sentinel; the
// table is at least twice the key count, so every insertion ends).
__device__ __forceinline__ void hot_lookup_build(const HotLds& L, const uint64_t* keys, uint32_t n) {
  for (uint32_t i = threadIdx.x; i <= L.mask; i += blockDim.x) L.key[i] = kEmpty;
  for (uint32_t i = threadIdx.x; i < kHotFilterBits / 32; i += blockDim.x) L.bits[i] = 0;
  __syncthreads();
  for (uint32_t h = threadIdx.x; h < n; h += blockDim.x) {
    const uint64_t k = keys[h];
    const uint64_t hk = key_hash(k);
    atomicOr(&L.bits[(uint32_t)(hk & (kHotFilterBits - 1)) >> 5], 1u << (hk & 31));
    uint32_t i = (uint32_t)(hk >> 24) & L.mask;
    while (at_cas<LdsT>(L.key + i, kEmpty, k) != kEmpty) i = (i + 1) & L.mask;
    L.idx[i] = h;
  }
  __syncthreads();
}

// Hot index of a key (hk = key_hash(key)), or -1.
__device__ __forceinline__ int hot_find(const HotLds& L, uint64_t key, uint64_t hk) {
  if (!((L.bits[(uint32_t)(hk & (kHotFilterBits - 1)) >> 5] >> (hk & 31)) & 1u)) return -1;
  uint32_t i = (uint32_t)(hk >> 24) & L.mask;
  for (uint32_t p = 0; p <= L.mask; ++p) {
    const uint64_t k = L.key[i];
    if (k == key) return (int)L.idx[i];
    if (k == kEmpty) return -1;
    i = (i + 1) & L.mask;
  }
  return -1;
}

// ------------------------------------------------------------ pass 1
//
// Each workgroup owns a contiguous row range and walks it in tiles of
// blockDim * U rows. Per tile: hot rows aggregate in LDS, side-slot rows go to
// the table's side slots, and every cold row takes a rank within its level-1
// partition (the top b1 bits of its partition index) from an LDS counter. One
// wave scans the tile's counts; then each cold row writes its record at the
// tile's base + its level-1 partition's offset + its rank. A tile's records
// are one contiguous block sorted by level-1 partition, written right after
// the previous tile's block in the workgroup's own region (which starts at
// the position of its first row and can hold all of its rows): no count pass
// is needed, and the writes of one tile fall in a few kilobytes that L2
// assembles into whole lines before they reach HBM. The tile's offsets and
// base are its descriptor. The workgroup also counts its cold rows per final
// partition in an LDS histogram; pass 2's partition boundaries are the scan
// of those counts.

constexpr int kL1TU = 4;  // typed loads: tiles of kL1TypedBlock * 4 rows
constexpr int kL1GU = 2;  // general loader: tiles of kL1Block * 2 rows

// A record is RW words: the packed key, then the value words the program
// reads (zero words after them up to RW). Records are written and read in
// blocks, so they are packed back to back: an odd RW is moved as 8-byte
// words, an even one as 16-byte pairs.
template <int RW>
__device__ __forceinline__ void rec_store(uint64_t* p, const uint64_t (&w)[RW]) {
  if (RW % 2 == 0) {
#pragma unroll
    for (int j = 0; j < RW / 2; ++j) ((ulonglong2*)p)[j] = make_ulonglong2(w[2 * j], w[2 * j + 1]);
  } else {
#pragma unroll
    for (int j = 0; j < RW; ++j) p[j] = w[j];
  }
}

template <int RW>
__device__ __forceinline__ void rec_load(const uint64_t* p, uint64_t (&w)[RW]) {
  if (RW % 2 == 0) {
#pragma unroll
    for (int j = 0; j < RW / 2; ++j) {
      const ulonglong2 x = ((const ulonglong2*)p)[j];
      w[2 * j] = x.x;
      w[2 * j + 1] = x.y;
    }
  } else {
#pragma unroll
    for (int j = 0; j < RW; ++j) w[j] = p[j];
  }
}

// LDS of pass 1 (u32 words): cnt[3][np1] (the counters of tiles k % 3),
// hist[np], then the hot-set lookup and accumulators.
struct L1Lds {
  uint32_t* cnt;
  uint32_t* hist;
};

__host__ __device__ inline uint32_t l1_hot_offset(uint32_t np1, uint32_t np) { return (3 * np1 + np + 1) & ~1u; }

__device__ __forceinline__ L1Lds l1_lds(uint32_t* lds, uint32_t np1) { return L1Lds{lds, lds + 3 * np1}; }

// Zeroed counters and histogram, the hot accumulators at their initial values
// and the hot-set lookup; returns the hot layout.
__device__ __forceinline__ HotLds l1_setup(const PartParams& P, uint32_t* lds) {
  const uint32_t ho = l1_hot_offset(1u << P.b1, 1u << P.pbits);
  for (uint32_t i = threadIdx.x; i < ho; i += blockDim.x) lds[i] = 0;
  const HotLds HL = hot_lds(lds + ho, 0, P.hot_bits, P.hot_slots);
  const int H = P.hot_slots;
  if (H) {
    for (int w = 0; w < P.pg.stride; ++w) {
      const uint64_t v = slot_init_word(P.pg, w);
      for (int i = threadIdx.x; i < HL.HW; i += blockDim.x) HL.acc[(uint64_t)w * HL.HW + i] = v;
    }
    hot_lookup_build(HL, P.hot_keys, min(*P.n_hot, (uint32_t)H));
  }
  __syncthreads();
  return HL;
}

// One tile of U loaded rows per thread: hot rows into the LDS accumulators,
// side-slot rows into the table's side slots, cold rows ranked, scanned and
// written as the tile's block. `run` is the next free position of the
// workgroup's region (the same in every thread); k is the tile's index in the
// workgroup, `tile` its descriptor index. Every thread of the workgroup calls
// this once per tile: one barrier, after the ranks are taken. Each wave then
// scans the counts itself; the counters rotate over three buffers, so the
// buffer a tile zeroes was last read before this tile's barrier and is next
// counted into after the next tile's barrier.
template <int RW, int U>
__device__ __forceinline__ void l1_tile(const PartParams& P, const HotLds& HL, const L1Lds& S, uint32_t k,
                                        uint64_t tile, uint32_t& run, const uint64_t (&row)[U],
                                        const uint64_t (&key)[U], const int (&kind)[U],
                                        const uint64_t (&val)[U][RW - 1], const uint32_t (&vvalid)[U]) {
  constexpr int NV = RW - 1;
  const int H = P.hot_slots;
  const int stride = P.pg.stride;
  const int b2 = P.pbits - P.b1;
  const uint32_t np1 = 1u << P.b1;
  uint32_t* const cnt = S.cnt + (k % 3) * np1;
  uint64_t hk[U];
#pragma unroll
  for (int u = 0; u < U; ++u) hk[u] = key_hash(key[u]);
  uint32_t hot = 0;
  if (H) {
    // hot rows aggregate here: the same LDS atomics on every lane
    // (identity operands on the lane's spare slot for the others)
    uint64_t* const spare = HL.acc + H + (threadIdx.x & 63);
    uint64_t* w[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int hi = kind[u] == 0 ? hot_find(HL, key[u], hk[u]) : -1;
      w[u] = hi >= 0 ? HL.acc + hi : spare;
      hot |= (uint32_t)(hi >= 0) << u;
    }
    apply_vals_lds_batch<NV, U>(w, HL.HW, P.pgv, val, P.vfloat, vvalid, hot);
  }
  uint32_t part[U], rank[U], cold = 0;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    part[u] = (uint32_t)(hk[u] >> (64 - P.pbits));
    rank[u] = 0;
    if (kind[u] < 0 || ((hot >> u) & 1u)) continue;
    if (kind[u] != 0) {
      // side slots: [0] the sentinel-valued key, [1] the NULL key
      uint64_t* s = P.T.slots + (P.T.mask + (kind[u] == 1 ? 1 : 2)) * (uint64_t)stride;
      apply_row<GlbT, false>(s, 1, P.pg, P.cs, RowRef{row[u], 0, nullptr});
      continue;
    }
    cold |= 1u << u;
    atomicAdd(&S.hist[part[u]], 1u);
    rank[u] = atomicAdd(&cnt[part[u] >> b2], 1u);
  }
  __syncthreads();
  // exclusive scan of the np1 <= 128 counts, two per lane, in every wave
  const uint32_t l = threadIdx.x & 63, i0 = 2 * l, i1 = 2 * l + 1;
  const uint32_t c0 = i0 < np1 ? cnt[i0] : 0u, c1 = i1 < np1 ? cnt[i1] : 0u;
  uint32_t s = c0 + c1;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(s, d);
    if (l >= (uint32_t)d) s += y;
  }
  const uint32_t e0 = s - c0 - c1, e1 = e0 + c0;
  const uint32_t total = __shfl(s, 63);
  if (threadIdx.x < 64) {
    uint32_t* desc = P.d1_off + tile * (np1 + 1);
    uint32_t* next = S.cnt + ((k + 2) % 3) * np1;
    if (i0 < np1) {
      desc[i0] = e0;
      next[i0] = 0;
    }
    if (i1 < np1) {
      desc[i1] = e1;
      next[i1] = 0;
    }
    if (l == 0) {
      desc[np1] = total;
      P.d1_base[tile] = run;
    }
  }
  const uint64_t base = run;
  run += total;
  uint32_t off[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {  // every lane shuffles: the offset of the row's level-1 partition
    const uint32_t p1 = part[u] >> b2;
    const uint32_t v0 = __shfl(e0, (int)(p1 >> 1)), v1 = __shfl(e1, (int)(p1 >> 1));
    off[u] = (p1 & 1u) ? v1 : v0;
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (!((cold >> u) & 1u)) continue;
    uint64_t w[RW];
    w[0] = key[u];
#pragma unroll
    for (int j = 1; j < RW; ++j) w[j] = (j - 1 < P.n_vals) ? val[u][j - 1] : 0;
    rec_store<RW>(P.recs + (base + off[u] + rank[u]) * RW, w);
  }
}

// This workgroup's partial hot slots, dense: [H][stride], key in word 0.
__device__ __forceinline__ void scatter_flush_hot(const PartParams& P, const HotLds& HL) {
  const int H = P.hot_slots, stride = P.pg.stride;
  if (!H) return;
  __syncthreads();
  const uint32_t n_hot = min(*P.n_hot, (uint32_t)H);
  uint64_t* dst = P.hot_part + (uint64_t)blockIdx.x * H * stride;
  for (int i = threadIdx.x; i < H * stride; i += blockDim.x) {
    const int h = i / stride, w = i % stride;
    dst[i] = w ? HL.acc[(uint64_t)w * HL.HW + h] : ((uint32_t)h < n_hot ? P.hot_keys[h] : kEmpty);
  }
}

// End of pass 1: the workgroup's histogram row and its hot partials.
__device__ __forceinline__ void l1_finish(const PartParams& P, const L1Lds& S, const HotLds& HL) {
  const uint32_t np = 1u << P.pbits;
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) P.hist[(uint64_t)blockIdx.x * np + p] = S.hist[p];
  scatter_flush_hot(P, HL);
}

// Pass 1 through the general row loader (any key, predicate and value
// columns the partitioned consume takes). RW = record words: key + values
// (2, 4 or 8 words; zero pad words after the values).
template <int RW>
__global__ void __launch_bounds__(kL1Block) part_l1_kernel(PartParams P) {
  constexpr int NV = RW - 1, U = kL1GU;
  extern __shared__ uint32_t lds1[];
  const RowSpec<NV> S = make_spec<NV>(P);
  const HotLds HL = l1_setup(P, lds1);
  const L1Lds T = l1_lds(lds1, 1u << P.b1);
  const uint64_t beg = (uint64_t)blockIdx.x * P.rows_per_wg;
  const uint64_t end = min(beg + P.rows_per_wg, P.n);
  uint32_t run = (uint32_t)beg;  // the region of the workgroup starts at its first row's position
  const uint64_t t0 = (uint64_t)blockIdx.x * P.tiles_per_wg;
  for (uint32_t k = 0; k < P.tiles_per_wg; ++k) {
    uint64_t row[U], key[U], val[U][NV];
    uint32_t vvalid[U];
    int kind[U];
    load_rows<NV, U>(S, beg + (uint64_t)k * blockDim.x * U + threadIdx.x, blockDim.x, end, row, key, kind, val,
                     vvalid);
    l1_tile<RW, U>(P, HL, T, k, t0 + k, run, row, key, kind, val, vvalid);
  }
  l1_finish(P, T, HL);
}

template <int W>
__device__ __forceinline__ uint64_t load_w(const char* base, uint64_t row) {
  if (W == 4) return (uint64_t)*(const __attribute__((address_space(1))) uint32_t*)(base + row * 4);
  return *(const __attribute__((address_space(1))) uint64_t*)(base + row * 8);
}

template <int PW, int U>
__device__ __forceinline__ uint32_t pred_bits(const uint64_t (&v)[U], bool is_float, int op, int64_t ci, double cf) {
  uint32_t m = 0;
  if (PW == 0) return (1u << U) - 1;
  if (PW == 8 && is_float) {
#pragma unroll
    for (int u = 0; u < U; ++u) m |= (uint32_t)cmp_op<double>(__longlong_as_double((long long)v[u]), cf, op) << u;
  } else if (PW == 8) {
#pragma unroll
    for (int u = 0; u < U; ++u) m |= (uint32_t)cmp_op<int64_t>((int64_t)v[u], ci, op) << u;
  } else {
#pragma unroll
    for (int u = 0; u < U; ++u) m |= (uint32_t)cmp_op<int32_t>((int32_t)(uint32_t)v[u], (int32_t)ci, op) << u;
  }
  return m;
}

// Pass 1 for the common column shapes, with the type decisions taken at
// compile time: a key of one or two non-null 4- or 8-byte integer columns,
// at most one predicate on a non-null 4- or 8-byte column, and value columns
// that are non-null 8-byte words (int64, DECIMAL, float64, IPv4). W0 / W1 /
// PW are the byte widths of key column 0, key column 1 (0: none) and the
// predicate column (0: none); signedness, the operator and the key packing
// stay runtime values (uniform, resolved outside the row loop). Loads are
// typed by width, kL1TU rows in flight per thread, and a value column that is
// also the predicate column is read once. Classifies and writes exactly what
// part_l1_kernel does for the same rows.
template <int W0, int W1, int PW, int RW>
__global__ void __launch_bounds__(kL1TypedBlock) __attribute__((amdgpu_waves_per_eu(6)))
part_l1_typed_kernel(PartParams P) {
  constexpr int U = kL1TU, NV = RW - 1;
  extern __shared__ uint32_t lds1[];
  const HotLds HL = l1_setup(P, lds1);
  const L1Lds T = l1_lds(lds1, 1u << P.b1);
  const char* d0 = (const char*)P.cs.cols[P.key.col[0]].data;
  const char* d1 = (const char*)P.cs.cols[P.key.col[W1 ? 1 : 0]].data;
  const char* dp = (const char*)P.cs.cols[PW ? P.pred[0].col : 0].data;
  const char* dv[NV];
  bool vp[NV];  // value column j is the predicate column
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const bool on = j < P.n_vals;
    dv[j] = (const char*)P.cs.cols[on ? P.val_col[j] : 0].data;
    vp[j] = on && PW == 8 && P.val_col[j] == P.pred[0].col;
  }
  const bool sx0 = P.cs.cols[P.key.col[0]].type == WG_INT32;
  const bool pf = PW && P.cs.cols[P.pred[0].col].type == WG_FLOAT64;
  const int sh0 = P.key.shift[0], sh1 = P.key.shift[1];
  const int op = P.pred[0].op;
  const int64_t ci = P.pred[0].i64;
  const double cf = P.pred[0].f64;
  const uint64_t beg = (uint64_t)blockIdx.x * P.rows_per_wg;
  const uint64_t end = min(beg + P.rows_per_wg, P.n);
  uint32_t run = (uint32_t)beg;
  const uint64_t t0 = (uint64_t)blockIdx.x * P.tiles_per_wg;
  for (uint32_t k = 0; k < P.tiles_per_wg; ++k) {
    const uint64_t base = beg + (uint64_t)k * blockDim.x * U;
    uint64_t row[U], a[U], b[U], pv[U], key[U], val[U][NV];
    uint32_t in = 0, vvalid[U];
    int kind[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t i = base + (uint64_t)u * blockDim.x + threadIdx.x;
      in |= (uint32_t)(i < end) << u;
      const uint64_t ii = i < end ? i : end - 1;
      row[u] = P.sel ? (uint64_t)P.sel[ii] : ii;
      vvalid[u] = ~0u;
    }
    // every load of the tile in flight before any is used
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] = load_w<W0>(d0, row[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) b[u] = W1 ? load_w<W1 ? W1 : 4>(d1, row[u]) : 0;
#pragma unroll
    for (int u = 0; u < U; ++u) pv[u] = PW ? load_w<PW ? PW : 4>(dp, row[u]) : 0;
#pragma unroll
    for (int j = 0; j < NV; ++j)
      if (j < P.n_vals && !vp[j]) {
#pragma unroll
        for (int u = 0; u < U; ++u) val[u][j] = load_w<8>(dv[j], row[u]);
      }
#pragma unroll
    for (int j = 0; j < NV; ++j)
#pragma unroll
      for (int u = 0; u < U; ++u) val[u][j] = j >= P.n_vals ? 0 : vp[j] ? pv[u] : val[u][j];
    const uint32_t pass = in & pred_bits<PW, U>(pv, pf, op, ci, cf);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (W1) {  // two 4-byte columns packed into one word
        key[u] = ((a[u] & 0xFFFFFFFFull) << sh0) | ((b[u] & 0xFFFFFFFFull) << sh1);
      } else {
        key[u] = (W0 == 4 && sx0) ? (uint64_t)(int64_t)(int32_t)(uint32_t)a[u] : a[u];
      }
      kind[u] = !((pass >> u) & 1u) ? -1 : key[u] == kEmpty ? 1 : 0;
    }
    l1_tile<RW, U>(P, HL, T, k, t0 + k, run, row, key, kind, val, vvalid);
  }
  l1_finish(P, T, HL);
}

using L1Fn = void (*)(PartParams);

// The typed pass 1 for this batch's columns, or null (the general kernel).
static L1Fn typed_l1_for(const ColSet& cs, const KeyDesc& K, const PredDesc* pred, const std::vector<int>& vcols,
                         int rw) {
  auto width = [](int t) { return (t == WG_INT32 || t == WG_UINT32) ? 4 : (t == WG_INT64 || t == WG_IPV4) ? 8 : 0; };
  if (K.n > 2 || pred[1].col >= 0 || rw > 4) return nullptr;
  int w[2] = {0, 0};
  for (int k = 0; k < K.n; ++k) {
    const ColDesc& c = cs.cols[K.col[k]];
    if (c.valid) return nullptr;
    w[k] = width(c.type);
    if (!w[k]) return nullptr;
  }
  if (K.n == 2 && (w[0] != 4 || w[1] != 4)) return nullptr;
  int pw = 0;
  if (pred[0].col >= 0) {
    const ColDesc& c = cs.cols[pred[0].col];
    if (c.valid) return nullptr;
    pw = c.type == WG_FLOAT64 ? 8 : c.type == WG_INT32 ? 4 : (c.type == WG_INT64 || c.type == WG_IPV4) ? 8 : 0;
    if (!pw) return nullptr;
  }
  for (int c : vcols) {
    const ColDesc& d = cs.cols[c];
    if (d.valid || !(d.type == WG_INT64 || d.type == WG_DECIMAL || d.type == WG_FLOAT64 || d.type == WG_IPV4))
      return nullptr;
  }
#define WG_TL(A, B, C)                                                                 \
  if (w[0] == A && w[1] == B && pw == C)                                               \
    return rw == 2 ? part_l1_typed_kernel<A, B, C, 2>                                  \
                   : rw == 3 ? part_l1_typed_kernel<A, B, C, 3> : part_l1_typed_kernel<A, B, C, 4>;
  WG_TL(4, 4, 0) WG_TL(4, 4, 4) WG_TL(4, 4, 8)
  WG_TL(4, 0, 0) WG_TL(4, 0, 4) WG_TL(4, 0, 8)
  WG_TL(8, 0, 0) WG_TL(8, 0, 4) WG_TL(8, 0, 8)
#undef WG_TL
  return nullptr;
}

static L1Fn general_l1_for(int rw) {
  return rw == 2 ? part_l1_kernel<2> : rw == 3 ? part_l1_kernel<3> : rw == 4 ? part_l1_kernel<4> : part_l1_kernel<8>;
}

// Dense emission of finished groups. A partitioned consume that starts from an
// empty aggregate owns every key of a partition that one workgroup aggregates
// alone, and every hot key: no other item, pass or earlier consume holds them.
// Those groups go straight to the finalize output array in slot format (key
// word, accumulator words), with zero exact-Int128 bases and a valid key; the
// global table only receives what must still be merged there (split
// partitions, rows whose LDS claim failed, side slots).
struct DenseOut {
  uint64_t* rows;               // [cap][stride]
  uint64_t* base;               // [cap][base_words], or null
  uint8_t* null_key;            // [cap]
  unsigned long long* count;    // rows emitted so far
  uint64_t cap;
  int32_t base_words;
};

// Row `pos` of the dense output from slot words w0..: `word(k)` returns word k.
template <class F>
__device__ __forceinline__ void dense_put(const DenseOut& D, uint64_t pos, int stride, F word) {
  uint64_t* d = D.rows + pos * (uint64_t)stride;
  for (int w = 0; w < stride; w += 2) *(ulonglong2*)(d + w) = make_ulonglong2(word(w), word(w + 1));
  for (int w = 0; w < D.base_words; ++w) D.base[pos * (uint64_t)D.base_words + w] = 0;
  D.null_key[pos] = 0;
}

// ------------------------------------------------------------ hot set

constexpr uint64_t kSampleTiles = 128;     // spread tiles of the batch counted for the hot set
constexpr uint64_t kSampleTileRows = 1024;
constexpr int kSampleBits = 18;            // sample count table: 2^18 {key, count} slots
constexpr int kHotHistBins = 1024;
constexpr uint32_t kHotMinCount = 4;       // a hot key is seen at least this often in the sample

// Scratch of the hot-set selection: the sample table, the count histogram,
// the threshold and the key count, then the hot keys.
struct HotScratch {
  uint64_t* tab;
  uint32_t* hist;
  uint32_t* thr;  // [0] threshold count, [1] keys selected
  uint64_t* keys;
};

__global__ void hot_reset_kernel(HotScratch X) {
  const uint64_t n = 1ull << kSampleBits;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    X.tab[2 * i] = kEmpty;
    X.tab[2 * i + 1] = 0;
    if (i < kHotHistBins) X.hist[i] = 0;
    if (i < 2) X.thr[i] = 0;
  }
}

// Count the keys of kSampleTiles tiles of kSampleTileRows rows, spread over
// the batch (same filtering and key packing as the passes).
// Each workgroup counts its tile's keys in an LDS table first (one slot per
// distinct key of the tile), then adds the tile's counts to the sample
// table: one global claim and add per distinct key and tile, so the most
// frequent keys do not serialise on their sample slot.
constexpr uint32_t kSampleLdsSlots = 2 * (uint32_t)kSampleTileRows;

__global__ void __launch_bounds__(256) hot_sample_kernel(PartParams P, HotScratch X) {
  __shared__ uint64_t skey[kSampleLdsSlots];
  __shared__ uint32_t scnt[kSampleLdsSlots];
  const RowSpec<0> S = make_spec<0>(P);
  const uint64_t mask = (1ull << kSampleBits) - 1;
  const uint64_t gap = P.n / kSampleTiles;
  for (uint64_t t = blockIdx.x; t < kSampleTiles; t += gridDim.x) {
    for (uint32_t i = threadIdx.x; i < kSampleLdsSlots; i += blockDim.x) {
      skey[i] = kEmpty;
      scnt[i] = 0;
    }
    __syncthreads();
    const uint64_t beg = t * gap, end = min(beg + kSampleTileRows, P.n);
    for (uint64_t base = beg; base < end; base += (uint64_t)blockDim.x * kRowU) {
      uint64_t row[kRowU], key[kRowU], val[kRowU][1];
      uint32_t vvalid[kRowU];
      int kind[kRowU];
      load_rows<0, kRowU>(S, base + threadIdx.x, blockDim.x, end, row, key, kind, val, vvalid);
#pragma unroll
      for (int u = 0; u < kRowU; ++u) {
        if (kind[u] != 0) continue;
        // at most kSampleTileRows keys in a table twice that size: every probe ends
        uint32_t idx = (uint32_t)(key_hash(key[u]) >> 40) & (kSampleLdsSlots - 1);
        for (;;) {
          const uint64_t prev = at_cas<LdsT>(skey + idx, kEmpty, key[u]);
          if (prev == kEmpty || prev == key[u]) {
            atomicAdd(scnt + idx, 1u);
            break;
          }
          idx = (idx + 1) & (kSampleLdsSlots - 1);
        }
      }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kSampleLdsSlots; i += blockDim.x) {
      const uint64_t k = skey[i];
      if (k == kEmpty) continue;
      // at most kSampleTiles * kSampleTileRows keys in a table twice that size
      uint64_t idx = key_hash(k) & mask;
      for (uint64_t p = 0; p <= mask; ++p) {
        uint64_t* s = X.tab + 2 * idx;
        const uint64_t prev = at_cas<GlbT>(s, kEmpty, k);
        if (prev == kEmpty || prev == k) {
          at_add<GlbT>(s + 1, scnt[i]);
          break;
        }
        idx = (idx + 1) & mask;
      }
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) hot_hist_kernel(HotScratch X) {
  __shared__ uint32_t h[kHotHistBins];
  for (int i = threadIdx.x; i < kHotHistBins; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const uint64_t n = 1ull << kSampleBits;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    if (X.tab[2 * i] != kEmpty) {
      const uint64_t c = X.tab[2 * i + 1];
      atomicAdd(&h[c < (uint64_t)kHotHistBins ? (uint32_t)c : (uint32_t)kHotHistBins - 1], 1u);
    }
  __syncthreads();
  for (int i = threadIdx.x; i < kHotHistBins; i += blockDim.x)
    if (h[i]) atomicAdd(&X.hist[i], h[i]);
}

// The smallest count T >= kHotMinCount such that at most H keys were seen T
// times or more (kHotHistBins: none): a suffix sum over the histogram, then
// the first bin whose suffix fits.
__global__ void __launch_bounds__(kHotHistBins) hot_threshold_kernel(HotScratch X, uint32_t H) {
  __shared__ uint32_t sfx[kHotHistBins];
  __shared__ uint32_t s_t;
  const uint32_t c = threadIdx.x;
  sfx[c] = X.hist[c];
  if (c == 0) s_t = kHotHistBins;
  __syncthreads();
  for (uint32_t d = 1; d < kHotHistBins; d <<= 1) {
    const uint32_t add = c + d < kHotHistBins ? sfx[c + d] : 0;
    __syncthreads();
    sfx[c] += add;
    __syncthreads();
  }
  // keys seen c times or more: sfx[c], non-increasing in c
  if (c >= kHotMinCount && sfx[c] <= H) atomicMin(&s_t, c);
  __syncthreads();
  if (c == 0) X.thr[0] = s_t;
}

__global__ void __launch_bounds__(256) hot_collect_kernel(HotScratch X, uint32_t H) {
  const uint64_t n = 1ull << kSampleBits;
  const uint32_t T = X.thr[0];
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = X.tab[2 * i];
    if (k == kEmpty || X.tab[2 * i + 1] < T) continue;
    const uint32_t at = atomicAdd(&X.thr[1], 1u);
    if (at < H) X.keys[at] = k;  // at most H by the choice of T
  }
}

// Fold the per-workgroup hot partials of one hot key (block h) in LDS, then
// into the global table (overflow partial list if the table is past budget).
__global__ void __launch_bounds__(256) hot_merge_kernel(const uint64_t* part, uint32_t G, uint32_t H,
                                                       const uint32_t* n_hot, AggProgram pg, GlobalTable T,
                                                       uint64_t* ovf_part, DenseOut dense) {
  // one accumulator slot per wave (fewer lanes on each LDS word), folded
  // into the first at the end
  __shared__ uint64_t sw[4][64];
  uint64_t* const s = sw[0];
  const int stride = pg.stride;
  const uint32_t h = blockIdx.x;
  uint32_t claims = 0;
  if (h < min(*n_hot, H)) {  // uniform per block
    if (threadIdx.x < 4u * 64u) sw[threadIdx.x >> 6][threadIdx.x & 63] = slot_init_word(pg, threadIdx.x & 63);
    __syncthreads();
    uint64_t* const mine = sw[threadIdx.x >> 6];
    for (uint32_t g = threadIdx.x; g < G; g += blockDim.x) {
      const uint64_t* src = part + ((uint64_t)g * H + h) * stride;
      if (src[1]) merge_partial<LdsT>(mine, src, 1, pg);  // word 1: the slot's row count
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int q = 1; q < 4; ++q)
        if (sw[q][1]) merge_partial<LdsT>(s, sw[q], 1, pg);
    __syncthreads();
    if (threadIdx.x == 0 && s[1] && dense.rows) {
      // hot keys are never partitioned: this block holds all of the key's rows
      s[0] = part[(uint64_t)h * stride];
      const uint64_t pos = atomicAdd(dense.count, 1ull);  // < cap: the host reserved H rows
      dense_put(dense, pos, stride, [&](int w) { return w < stride ? s[w] : 0ull; });
    } else if (threadIdx.x == 0 && s[1]) {
      const uint64_t key = part[(uint64_t)h * stride];
      uint64_t* d = glb_find_or_claim(T, stride, key, key_hash(key), claims);
      if (d) {
        merge_partial(d, s, 1, pg);
      } else {
        const uint64_t at =
            __hip_atomic_fetch_add(T.counters + CTR_OVF_PART, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint64_t* rec = ovf_part + at * (uint64_t)stride;  // in bounds: the list holds far more than H partials
        s[0] = key;
        for (int w = 0; w < stride; ++w) rec[w] = s[w];
      }
    }
  }
  block_add(T.counters + CTR_GROUPS, claims);
}

// Partition totals: tot[p] += the cold rows of kTotRows workgroups in
// partition p (tot zeroed before; tot[np] stays 0, so the exclusive scan of
// np + 1 entries ends with the total).
constexpr uint32_t kTotRows = 16;

__global__ void __launch_bounds__(256) part_tot_kernel(const uint32_t* hist, uint32_t G, uint32_t np, uint32_t* tot) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const uint32_t g0 = blockIdx.y * kTotRows, g1 = min(g0 + kTotRows, G);
  uint32_t s = 0;
  for (uint32_t g = g0; g < g1; ++g) s += hist[(uint64_t)g * np + p];
  if (s) atomicAdd(tot + p, s);
}

// ------------------------------------------------------------ pass 2
//
// One workgroup per (level-1 partition p1, range of tpu pass-1 tiles). It
// gathers p1's runs of those tiles (run positions from the tile descriptors;
// a record's run is found by binary search over the runs' prefix sums in
// LDS) in rounds of kL2Block * kL2U records, ranks each record within its
// final partition (the level-2 bits below p1) with an LDS counter, reserves
// each final partition's share of the round with one atomic on that
// partition's fill cursor (which starts at its boundary), and writes the
// records there. Final partition p then holds all of its records in
// [bound[p], bound[p + 1]), each round's share of it as one contiguous run.
constexpr int kL2Block = 512;
constexpr int kL2U = 8;
constexpr uint32_t kL2MaxTiles = 2 * kL2Block;  // two tile descriptors per thread
constexpr uint32_t kMaxNp2 = 128;

struct L2Params {
  const uint64_t* in;       // pass-1 records
  uint64_t* out;            // records by final partition
  const uint32_t* d1_off;   // [nt1][np1 + 1] run offsets in each tile, then the tile's record count
  const uint32_t* d1_base;  // [nt1] first position of each tile
  uint32_t* fill;           // [np] next free position of each final partition
  uint64_t nt1;
  uint32_t tpu;
  int32_t b1, pbits;
};

template <int RW>
__global__ void __launch_bounds__(kL2Block) part_l2_kernel(L2Params P) {
  using Scan = hipcub::BlockScan<uint32_t, kL2Block>;
  __shared__ typename Scan::TempStorage scan_tmp;
  __shared__ uint32_t pre[kL2MaxTiles + 1];
  __shared__ uint32_t start[kL2MaxTiles];
  __shared__ uint32_t cnt[kMaxNp2], gbase[kMaxNp2];
  const uint32_t np1 = 1u << P.b1;
  const int b2 = P.pbits - P.b1;
  const uint32_t np2 = 1u << b2;
  const uint32_t p1 = blockIdx.x % np1;
  const uint64_t t0 = (uint64_t)(blockIdx.x / np1) * P.tpu;
  const uint32_t nt = (uint32_t)min<uint64_t>(P.tpu, P.nt1 - t0);
  uint32_t c[2], e[2], total;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const uint32_t i = 2 * threadIdx.x + j;
    c[j] = 0;
    if (i < nt) {
      const uint32_t* d = P.d1_off + (t0 + i) * (np1 + 1);
      const uint32_t a = d[p1];
      c[j] = d[p1 + 1] - a;  // d[np1] is the tile's record count
      start[i] = P.d1_base[t0 + i] + a;
    }
  }
  Scan(scan_tmp).ExclusiveSum(c, e, total);
#pragma unroll
  for (int j = 0; j < 2; ++j)
    if (2 * threadIdx.x + j < nt) pre[2 * threadIdx.x + j] = e[j];
  if (threadIdx.x == 0) pre[nt] = total;
  for (uint32_t q = threadIdx.x; q < np2; q += blockDim.x) cnt[q] = 0;
  __syncthreads();
  for (uint32_t r0 = 0; r0 < total; r0 += kL2Block * kL2U) {
    uint64_t rec[kL2U][RW];
    uint32_t q[kL2U], rank[kL2U], in = 0;
#pragma unroll
    for (int u = 0; u < kL2U; ++u) {
      const uint32_t i = r0 + u * kL2Block + threadIdx.x;
      if (i >= total) {
#pragma unroll
        for (int j = 0; j < RW; ++j) rec[u][j] = 0;
        continue;
      }
      in |= 1u << u;
      uint32_t lo = 0, hi = nt;  // pre[lo] <= i < pre[hi]
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= i) lo = mid;
        else hi = mid;
      }
      rec_load<RW>(P.in + (uint64_t)(start[lo] + (i - pre[lo])) * RW, rec[u]);
    }
#pragma unroll
    for (int u = 0; u < kL2U; ++u) {
      q[u] = (uint32_t)(key_hash(rec[u][0]) >> (64 - P.pbits)) & (np2 - 1);
      rank[u] = ((in >> u) & 1u) ? atomicAdd(&cnt[q[u]], 1u) : 0u;
    }
    __syncthreads();
    for (uint32_t x = threadIdx.x; x < np2; x += blockDim.x) {
      const uint32_t k = cnt[x];
      gbase[x] = k ? atomicAdd(P.fill + p1 * np2 + x, k) : 0u;
      cnt[x] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kL2U; ++u) {
      if (!((in >> u) & 1u)) continue;
      rec_store<RW>(P.out + (uint64_t)(gbase[q[u]] + rank[u]) * RW, rec[u]);
    }
  }
}

struct PartItem {
  uint32_t beg, end;  // positions in the partitioned records
  uint32_t split;     // 1: the partition is shared with other items
  uint32_t pad;
};

struct PartAggParams {
  AggProgram pg;  // ops remapped onto the record's value words
  GlobalTable T;
  const uint64_t* recs;
  const PartItem* items;
  uint32_t n_items;
  int32_t lds_slots;
  unsigned long long* item_ctr;
  const uint32_t* sel;  // replay: positions to aggregate, in chunks of sel_chunk (split items)
  uint64_t n_sel;
  uint64_t sel_chunk;
  uint64_t max_item_rows;  // headroom per in-flight item in the overflow row list
  uint32_t* ovf_rows;
  uint64_t ovf_rows_cap;
  uint64_t* ovf_part;
  uint64_t ovf_part_cap;
  uint32_t vfloat;  // bit j: value word j holds float64 bits (else an int64)
  DenseOut dense;   // rows == nullptr: merge every item into the global table
};

// Exclusive prefix sum of one u32 per thread over the workgroup (blockDim a
// multiple of 64, at most 1024) and its total; wsum is 16 words of shared
// memory. Every thread calls it (two barriers).
__device__ __forceinline__ uint32_t block_exclusive_sum(uint32_t v, uint32_t& total, uint32_t* wsum) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint32_t s = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(s, d);
    if (lane >= (uint32_t)d) s += y;
  }
  if (lane == 63) wsum[wave] = s;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (uint32_t w = 0; w < nw; ++w) {
    const uint32_t x = wsum[w];
    all += x;
    before += w < wave ? x : 0u;
  }
  total = all;
  __syncthreads();
  return before + s - v;
}

// Counters of one partition-aggregate launch: the per-launch words cleared,
// the work-item cursor set to the first item to take, the dense row count
// cleared when asked (once per consume, before anything writes the output).
__global__ void part_round_begin_kernel(uint64_t* ctr, uint64_t start_item, int zero_dense) {
  const int t = threadIdx.x;
  if (t >= CTR_OVF_ROWS && t < CTR_N) ctr[t] = 0;
  if (t == kCtrDense && zero_dense) ctr[t] = 0;
  if (t == kCtrItem) ctr[t] = start_item;
}

constexpr int kFlushU = 2;
// A partition's keys fill its LDS table to about 3/4 of L. A key that finds
// no LDS slot aggregates in the global table, and under dense emission one
// such key makes finalize scan (and reset refill) the whole table: the probe
// sequence is triangular, so 128 probes miss only in a table that is full.
constexpr int kPartLdsProbes = 128;

template <int RW>
__global__ void __launch_bounds__(1024) part_agg_kernel(PartAggParams P) {
  constexpr int NV = RW - 1;
  extern __shared__ uint64_t lds[];
  __shared__ int64_t s_item;
  __shared__ uint64_t s_init[16];
  __shared__ uint32_t s_wsum[16];
  const int stride = P.pg.stride;
  const int L = P.lds_slots;
  const int LW = lds_word_stride(L);  // L hashed slots, then one spare slot per lane
  uint64_t* const spare = lds + L + (threadIdx.x & 63);
  if (threadIdx.x < (unsigned)stride) s_init[threadIdx.x] = slot_init_word(P.pg, threadIdx.x);
  uint32_t claims = 0;
  const uint64_t n_work = P.sel ? (P.n_sel + P.sel_chunk - 1) / P.sel_chunk : P.n_items;
  for (;;) {
    if (threadIdx.x == 0) {
      int64_t t = -1;
      // take an item only while every in-flight item could still overflow entirely
      const uint64_t o = __hip_atomic_load(P.T.counters + CTR_OVF_ROWS, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint64_t op = __hip_atomic_load(P.T.counters + CTR_OVF_PART, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (o + (uint64_t)gridDim.x * P.max_item_rows > P.ovf_rows_cap ||
          op + (uint64_t)gridDim.x * P.lds_slots > P.ovf_part_cap) {
        at_add<GlbT>(P.T.counters + CTR_STOPPED, 1);
      } else {
        const unsigned long long w = atomicAdd(P.item_ctr, 1ull);
        if (w < n_work) t = (int64_t)w;
      }
      s_item = t;
    }
    __syncthreads();
    const int64_t item = s_item;
    if (item < 0) break;
    uint64_t beg, end;
    bool split;
    if (P.sel) {
      beg = (uint64_t)item * P.sel_chunk;
      end = min(beg + P.sel_chunk, P.n_sel);
      split = true;
    } else {
      const PartItem it = P.items[item];
      beg = it.beg;
      end = it.end;
      split = it.split != 0;
    }
    for (int w = 0; w < stride; ++w) {
      const uint64_t v = s_init[w];
      for (int i = threadIdx.x; i < LW; i += blockDim.x) lds[(uint64_t)w * LW + i] = v;
    }
    __syncthreads();
    for (uint64_t base = beg; base < end; base += (uint64_t)blockDim.x * kAggU) {
      uint64_t rec[kAggU][RW];
      uint64_t pos[kAggU];
      bool in[kAggU];
#pragma unroll
      for (int u = 0; u < kAggU; ++u) {
        const uint64_t i = base + (uint64_t)u * blockDim.x + threadIdx.x;
        in[u] = i < end;
        const uint64_t ii = in[u] ? i : beg;
        pos[u] = P.sel ? (uint64_t)P.sel[ii] : ii;
      }
#pragma unroll
      for (int u = 0; u < kAggU; ++u) {
        rec_load<RW>(P.recs + pos[u] * RW, rec[u]);
      }
      uint64_t* ls[kAggU];
#pragma unroll
      for (int u = 0; u < kAggU; ++u)
        ls[u] = in[u] ? lds_find_or_claim_tri(lds, L - 1, rec[u][0], key_hash(rec[u][0]), 16, kPartLdsProbes) : nullptr;
      // same LDS atomics on every lane (identity operands on the spare slot
      // where the row does not aggregate here)
      {
        uint64_t* w[kAggU];
        uint64_t v[kAggU][NV];
        uint32_t vvalid[kAggU];
        uint32_t live = 0;
#pragma unroll
        for (int u = 0; u < kAggU; ++u) {
          w[u] = ls[u] ? ls[u] : spare;
          live |= (uint32_t)(ls[u] != nullptr) << u;
          vvalid[u] = ~0u;
#pragma unroll
          for (int j = 0; j < NV; ++j) v[u][j] = rec[u][j + 1];
        }
        apply_vals_lds_batch<NV, kAggU>(w, LW, P.pg, v, P.vfloat, vvalid, live);
      }
#pragma unroll
      for (int u = 0; u < kAggU; ++u) {
        if (!in[u] || ls[u]) continue;
        const uint64_t key = rec[u][0];
        uint64_t v[NV];
#pragma unroll
        for (int j = 0; j < NV; ++j) v[j] = rec[u][j + 1];
        uint64_t* g = glb_find_or_claim(P.T, stride, key, key_hash(key), claims);
        if (!g) {
          const uint64_t at =
              __hip_atomic_fetch_add(P.T.counters + CTR_OVF_ROWS, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          P.ovf_rows[at] = (uint32_t)pos[u];  // in bounds: items are taken only with headroom
          continue;
        }
        apply_vals<GlbT, NV>(g, 1, P.pg, v, P.vfloat, ~0u);
      }
    }
    __syncthreads();
    if (P.dense.rows && !split) {
      // this workgroup holds every row of the item's keys: its LDS groups are
      // final, written to the dense output behind one reservation; a block
      // scan of the threads' live-slot counts numbers them (rows of adjacent
      // threads are adjacent)
      __shared__ unsigned long long s_at;
      uint32_t mine = 0;
      for (int i = threadIdx.x; i < L; i += blockDim.x) mine += lds[i] != kEmpty;
      uint32_t total;
      const uint32_t pre = block_exclusive_sum(mine, total, s_wsum);
      if (threadIdx.x == 0) s_at = total ? atomicAdd(P.dense.count, (unsigned long long)total) : 0ull;
      __syncthreads();
      uint64_t pos = s_at + pre;  // < cap: the host reserved L rows per item
      for (int i = threadIdx.x; i < L; i += blockDim.x) {
        if (lds[i] == kEmpty) continue;
        dense_put(P.dense, pos++, stride, [&](int w) { return w < stride ? lds[(uint64_t)w * LW + i] : 0ull; });
      }
      __syncthreads();
      continue;
    }
    // merge the LDS table into the global table; first probes of kFlushU
    // slots per thread are in flight together
    for (int i0 = threadIdx.x; i0 < L; i0 += blockDim.x * kFlushU) {
      uint64_t key[kFlushU], prev[kFlushU] = {}, home[kFlushU];
#pragma unroll
      for (int u = 0; u < kFlushU; ++u) {
        const int i = i0 + u * blockDim.x;
        key[u] = i < L ? lds[i] : kEmpty;
        home[u] = key_hash(key[u]) & P.T.mask;
      }
#pragma unroll
      for (int u = 0; u < kFlushU; ++u)
        if (key[u] != kEmpty) prev[u] = at_cas<GlbT>(P.T.slots + home[u] * (uint64_t)stride, kEmpty, key[u]);
#pragma unroll
      for (int u = 0; u < kFlushU; ++u) {
        if (key[u] == kEmpty) continue;
        const uint64_t* ls = lds + (i0 + u * blockDim.x);
        bool fresh = prev[u] == kEmpty;
        uint64_t* g = (fresh || prev[u] == key[u])
                          ? P.T.slots + home[u] * (uint64_t)stride
                          : glb_probe_from(P.T, stride, key[u], (home[u] + 1) & P.T.mask, 1, fresh);
        if (!g) {
          const uint64_t at =
              __hip_atomic_fetch_add(P.T.counters + CTR_OVF_PART, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          uint64_t* rec = P.ovf_part + at * (uint64_t)stride;  // in bounds: headroom as above
          for (int w = 0; w < stride; ++w) rec[w] = ls[(uint64_t)w * LW];
          continue;
        }
        if (fresh) ++claims;
        if (split) {
          merge_partial(g, ls, LW, P.pg);
        } else if (fresh) {
          // a slot this workgroup just claimed for a key no one else holds:
          // its words are still the initial values, store the partial
          g[1] = ls[LW];
          for (int w = 2; w < stride; w += 2)
            *(ulonglong2*)(g + w) = make_ulonglong2(ls[(uint64_t)w * LW], ls[(uint64_t)(w + 1) * LW]);
        } else {
          merge_partial_owned(g, ls, LW, P.pg);
        }
      }
    }
    __syncthreads();
  }
  block_add(P.T.counters + CTR_GROUPS, claims);
}

// ------------------------------------------------------------------ host

// Is the partitioned consume applicable to this aggregate and batch?
bool part_eligible(const wg_agg* h, const ColSet& cs) {
  if (h->key.n > 1)
    for (int k = 0; k < h->key.n; ++k)
      if (cs.cols[h->key.col[k]].valid) return false;
  for (int o = 0; o < h->pg.n_ops; ++o) {
    const OpDesc& op = h->pg.ops[o];
    for (int c : {op.col, op.col2}) {
      if (c < 0) continue;
      if (cs.cols[c].valid) return false;                // value NULLs are not carried
      if (cs.cols[c].type == WG_DECIMAL128) return false;  // 16-byte values are not carried
    }
  }
  return true;
}

static wg_status allow_lds(const void* kernel, size_t bytes) {
  if (bytes > 64 * 1024) WG_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return WG_OK;
}

static wg_status launch_part_agg(int rw, int grid, int block, size_t lds, hipStream_t st, const PartAggParams& Q) {
  switch (rw) {
    case 2:
      WG_TRY(allow_lds((const void*)part_agg_kernel<2>, lds));
      hipLaunchKernelGGL(part_agg_kernel<2>, dim3(grid), dim3(block), lds, st, Q);
      break;
    case 3:
      WG_TRY(allow_lds((const void*)part_agg_kernel<3>, lds));
      hipLaunchKernelGGL(part_agg_kernel<3>, dim3(grid), dim3(block), lds, st, Q);
      break;
    case 4:
      WG_TRY(allow_lds((const void*)part_agg_kernel<4>, lds));
      hipLaunchKernelGGL(part_agg_kernel<4>, dim3(grid), dim3(block), lds, st, Q);
      break;
    default:
      WG_TRY(allow_lds((const void*)part_agg_kernel<8>, lds));
      hipLaunchKernelGGL(part_agg_kernel<8>, dim3(grid), dim3(block), lds, st, Q);
      break;
  }
  WG_HIP(hipGetLastError());
  return WG_OK;
}

wg_status agg_consume_partitioned(wg_agg* h, const ColSet& cs, const uint32_t* sel, uint64_t n, const PredDesc* pred) {
  wg_ctx* ctx = h->ctx;
  // value columns the program reads, each carried as one 8-byte word
  std::vector<int> vcols;
  AggProgram pg2 = h->pg;
  for (int o = 0; o < pg2.n_ops; ++o) {
    for (int32_t* c : {&pg2.ops[o].col, &pg2.ops[o].col2}) {
      if (*c < 0) continue;
      int j = -1;
      for (size_t k = 0; k < vcols.size(); ++k)
        if (vcols[k] == *c) j = (int)k;
      if (j < 0) {
        j = (int)vcols.size();
        vcols.push_back(*c);
      }
      *c = j;
    }
  }
  WG_CHECK((int)vcols.size() <= kMaxVals, WG_ERR_UNSUPPORTED, "partitioned consume carries at most %d value columns",
           kMaxVals);
  WG_CHECK(h->key.n <= 2, WG_ERR_UNSUPPORTED, "partitioned consume takes keys of at most two columns");
  // records of 2, 3, 4 or 8 words (the key, the value words, zero words up to
  // the record size), packed back to back in the sorted blocks
  const int RW = vcols.size() < 2 ? 2 : vcols.size() == 2 ? 3 : vcols.size() == 3 ? 4 : 8;
  // an aggregate with no groups yet owns every key this consume sees: finished
  // groups go straight to the dense output instead of the table
  const bool dense = h->groups == 0 && h->n_dense == 0 && !h->dirty;
  // the partition table in LDS: the configured slots, halved until a table of
  // this program's slot stride fits
  int L = h->part_lds_slots;
  while (L > 64 && lds_table_bytes(L, h->pg.stride) > 150 * 1024) L >>= 1;
  WG_CHECK(lds_table_bytes(L, h->pg.stride) <= 150 * 1024, WG_ERR_UNSUPPORTED,
           "a partition table of this aggregate's slot stride does not fit in LDS");
  // partitions of at most about 3/4 L groups each: fewer, larger partitions
  // mean fewer, longer runs in the sorting passes, while triangular probing
  // in the LDS table still ends within its probe limit
  const uint64_t est = std::max<uint64_t>(std::max<uint64_t>(h->group_hint, h->groups), 1);
  int pbits = 6;
  while (pbits < 14 && (1ull << pbits) * (uint64_t)(3 * L / 4) < est) ++pbits;
  const uint32_t np = 1u << pbits;
  const int b1 = (pbits + 1) / 2;  // level 1 <= 2^7 partitions, level 2 the rest
  const uint32_t np1 = 1u << b1;
  // heavy hitters: batches large enough for a sample to find them
  // (as many of the requested slots as fit pass 1's LDS)
  int H = (h->part_hot > 0 && n >= (1ull << 20) && h->pg.stride <= 64) ? h->part_hot : 0;
  auto hot_bits_for = [](int hs) {
    int b = 0;
    while (hs && (1 << b) < 2 * hs) ++b;
    return b;
  };
  auto l1_lds_bytes = [&](int hs) {
    return (size_t)l1_hot_offset(np1, np) * 4 +
           (hs ? hot_lookup_bytes(hot_bits_for(hs)) + lds_table_bytes(hs, h->pg.stride) : 0);
  };
  while (H > 64 && l1_lds_bytes(H) > 150 * 1024) H >>= 1;
  if (l1_lds_bytes(H) > 150 * 1024) H = 0;
  const int hot_bits = hot_bits_for(H);
  const size_t l1_bytes = l1_lds_bytes(H);
  L1Fn l1_fn = typed_l1_for(cs, h->key, pred, vcols, RW);
  const int B1 = l1_fn ? kL1TypedBlock : kL1Block;
  const uint64_t T1 = (uint64_t)B1 * (l1_fn ? kL1TU : kL1GU);  // rows per pass-1 tile
  if (!l1_fn) l1_fn = general_l1_for(RW);
  WG_TRY(allow_lds((const void*)l1_fn, l1_bytes));
  // one resident round of pass-1 workgroups, each over a whole number of tiles
  int occ = 0;
  WG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, l1_fn, B1, l1_bytes));
  const int G = ctx->cus * std::max(1, occ);
  const uint64_t R = ((n + G - 1) / G + T1 - 1) / T1 * T1;
  const uint64_t tpw = R / T1, nt1 = (uint64_t)G * tpw;
  WG_CHECK((uint64_t)G * R < (1ull << 32), WG_ERR_UNSUPPORTED,
           "partitioned consume of %llu rows: record positions need more than 32 bits", (unsigned long long)n);
  // pass-2 units of tpu tiles: about two rounds of records if every row were cold
  const uint64_t T2 = (uint64_t)kL2Block * kL2U;
  const uint32_t tpu = (uint32_t)std::min<uint64_t>(kL2MaxTiles, std::max<uint64_t>(16, 2 * T2 * np1 / T1));
  const uint64_t units = (nt1 + tpu - 1) / tpu;

  PartParams P{};
  P.cs = cs;
  P.key = h->key;
  P.pred[0] = pred[0];
  P.pred[1] = pred[1];
  P.pg = h->pg;
  P.sel = sel;
  P.n = n;
  P.rows_per_wg = R;
  P.tiles_per_wg = (uint32_t)tpw;
  P.pbits = pbits;
  P.b1 = b1;
  P.n_vals = (int)vcols.size();
  for (size_t j = 0; j < vcols.size(); ++j) P.val_col[j] = vcols[j];
  P.pgv = pg2;
  for (size_t j = 0; j < vcols.size(); ++j)
    if (cs.cols[vcols[j]].type == WG_FLOAT64) P.vfloat |= 1u << j;

  // Everything the passes need is allocated here, before pass 1 changes any
  // state: hot-set scratch, histograms, descriptors, both record buffers,
  // work items, overflow lists and the dense output. (A failure to grow the
  // table while draining the items is the one allocation failure that can
  // come later.)
  HotScratch X{};
  if (H) {
    // scratch: sample table, histogram, threshold/count, keys, partial slots
    const uint64_t words = (2ull << kSampleBits) + kHotHistBins / 2 + 1 + (uint64_t)H +
                           (uint64_t)G * H * h->pg.stride;
    if (h->hot_buf_words < words) WG_TRY(regrow(ctx, &h->hot_buf, &h->hot_buf_words, words, 8, "heavy-hitter scratch"));
    X.tab = h->hot_buf;
    X.hist = (uint32_t*)(X.tab + (2ull << kSampleBits));
    X.thr = X.hist + kHotHistBins;
    X.keys = (uint64_t*)(X.thr + 2);
    P.hot_keys = X.keys;
    P.n_hot = X.thr + 1;
    P.hot_slots = H;
    P.hot_bits = hot_bits;
    P.hot_part = X.keys + H;
  }
  // u32 scratch: totals (8-byte aligned, zeroed as words), boundaries, fill
  // cursors, histogram rows, tile descriptors
  const uint64_t tot_words = ((uint64_t)np + 2) / 2 * 2;
  const uint64_t o_bound = tot_words, o_fill = o_bound + np + 1, o_hist = o_fill + np;
  const uint64_t o_d1 = o_hist + (uint64_t)G * np, o_base = o_d1 + nt1 * (np1 + 1);
  const uint64_t cnt_words = o_base + nt1;
  if (h->part_cnt_cap < cnt_words) WG_TRY(regrow(ctx, &h->part_cnt, &h->part_cnt_cap, cnt_words, 4, "partition counts"));
  size_t scan_bytes = 0;
  WG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, h->part_cnt, h->part_cnt, (int)(np + 1), ctx->stream));
  if (scan_bytes > h->part_scan_bytes)
    WG_TRY(regrow(ctx, &h->part_scan, &h->part_scan_bytes, scan_bytes, 1, "partition scan scratch"));
  uint32_t* tot = h->part_cnt;
  uint32_t* dbound = h->part_cnt + o_bound;
  uint32_t* fill = h->part_cnt + o_fill;
  P.hist = h->part_cnt + o_hist;
  P.d1_off = h->part_cnt + o_d1;
  P.d1_base = h->part_cnt + o_base;
  P.T = table_of(h);
  // records: pass 1's regions (room for every row of each workgroup), then
  // the partitioned records of pass 2 (room for every row)
  const uint64_t need = ((uint64_t)G * R + n) * RW;
  if (h->part_buf_words < need) WG_TRY(regrow(ctx, &h->part_buf, &h->part_buf_words, need, 8, "partition records"));
  P.recs = h->part_buf;
  uint64_t* recs2 = h->part_buf + (uint64_t)G * R * RW;
  // host side: boundaries, then the work items (largest possible count)
  const uint64_t chunk = h->part_chunk_rows;
  const uint64_t max_items = np + n / chunk + 1;
  const size_t host_bytes = (np + 1) * 4 + max_items * sizeof(PartItem) + 64;
  if (h->part_host_bytes < host_bytes) {
    if (h->side) WG_HIP(hipStreamSynchronize(h->side));  // no copy of an earlier call still reads it
    if (h->part_host) WG_HIP(hipHostFree(h->part_host));
    h->part_host = nullptr;
    h->part_host_bytes = 0;
    WG_HIP(hipHostMalloc(&h->part_host, host_bytes, hipHostMallocDefault));
    h->part_host_bytes = host_bytes;
  }
  uint32_t* bound = (uint32_t*)h->part_host;
  PartItem* hitems = (PartItem*)((char*)h->part_host + ((np + 1) * 4 + 15) / 16 * 16);
  const uint64_t ibytes = max_items * sizeof(PartItem);
  if (h->part_items_bytes < ibytes)
    WG_TRY(regrow(ctx, &h->part_items, &h->part_items_bytes, ibytes, 1, "partition work items"));
  const size_t lds_bytes = lds_table_bytes(L, h->pg.stride);
  const int per_cu_agg = std::max(1, std::min<int>((160 * 1024) / (int)lds_bytes, 2048 / h->part_block));
  const int grid_max = ctx->cus * per_cu_agg;
  // overflow lists with headroom for every in-flight item (see the kernel)
  const uint64_t need_rows = (uint64_t)grid_max * chunk * 2;
  if (h->ovf_rows_cap < need_rows) WG_TRY(regrow(ctx, &h->ovf_rows, &h->ovf_rows_cap, need_rows, 4, "overflow rows"));
  const uint64_t need_part = (uint64_t)grid_max * L * 2;
  if (h->ovf_part_cap < need_part)
    WG_TRY(regrow(ctx, &h->ovf_part, &h->ovf_part_cap, need_part, (uint64_t)h->pg.stride * 8, "overflow partials"));
  // dense output: at most L groups per partition a workgroup owns alone (and
  // never more groups than rows), and H hot keys
  if (dense)
    WG_TRY(agg_ensure_dense(h, (uint64_t)H + std::min<uint64_t>(n, (uint64_t)np * L) + kSideSlots, 0));
  if (!h->pev[0])
    for (hipEvent_t& e : h->pev) WG_HIP(hipEventCreate(&e));
  if (!h->side) WG_HIP(hipStreamCreate(&h->side));

  // 1: hot set, tiles sorted by level-1 partition (side-slot rows and hot
  // rows are folded here), partition totals and boundaries
  h->touched = true;
  if (dense) h->side_dirty = true;
  else h->dirty = true;
  WG_HIP(hipEventRecord(h->pev[0], ctx->stream));
  if (H) {
    hipLaunchKernelGGL(hot_reset_kernel, dim3(256), dim3(256), 0, ctx->stream, X);
    hipLaunchKernelGGL(hot_sample_kernel, dim3(kSampleTiles), dim3(256), 0, ctx->stream, P, X);
    hipLaunchKernelGGL(hot_hist_kernel, dim3(256), dim3(256), 0, ctx->stream, X);
    hipLaunchKernelGGL(hot_threshold_kernel, dim3(1), dim3(kHotHistBins), 0, ctx->stream, X, (uint32_t)H);
    hipLaunchKernelGGL(hot_collect_kernel, dim3(256), dim3(256), 0, ctx->stream, X, (uint32_t)H);
    WG_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(l1_fn, dim3(G), dim3(B1), l1_bytes, ctx->stream, P);
  WG_HIP(hipGetLastError());
  WG_TRY(zero_dev(ctx, (uint64_t*)tot, tot_words / 2));
  hipLaunchKernelGGL(part_tot_kernel, dim3((np + 255) / 256, (G + kTotRows - 1) / kTotRows), dim3(256), 0,
                     ctx->stream, (const uint32_t*)P.hist, (uint32_t)G, np, tot);
  WG_HIP(hipGetLastError());
  WG_HIP(hipcub::DeviceScan::ExclusiveSum(h->part_scan, scan_bytes, tot, dbound, (int)(np + 1), ctx->stream));
  WG_HIP(hipMemcpyAsync(fill, dbound, (uint64_t)np * 4, hipMemcpyDeviceToDevice, ctx->stream));
  WG_HIP(hipEventRecord(h->pev[1], ctx->stream));
  // the boundaries go to the host on the side stream: pass 2 starts at once
  WG_HIP(hipStreamWaitEvent(h->side, h->pev[1], 0));
  WG_HIP(hipMemcpyAsync(bound, dbound, (np + 1) * 4, hipMemcpyDeviceToHost, h->side));
  WG_HIP(hipEventRecord(h->pev[2], h->side));

  // 2: every final partition contiguous (queued before the host waits for
  // the boundaries, so it runs while the work items are built)
  L2Params Q2{};
  Q2.in = P.recs;
  Q2.out = recs2;
  Q2.d1_off = P.d1_off;
  Q2.d1_base = P.d1_base;
  Q2.fill = fill;
  Q2.nt1 = nt1;
  Q2.tpu = tpu;
  Q2.b1 = b1;
  Q2.pbits = pbits;
  switch (RW) {
    case 2: hipLaunchKernelGGL(part_l2_kernel<2>, dim3(np1 * units), dim3(kL2Block), 0, ctx->stream, Q2); break;
    case 3: hipLaunchKernelGGL(part_l2_kernel<3>, dim3(np1 * units), dim3(kL2Block), 0, ctx->stream, Q2); break;
    case 4: hipLaunchKernelGGL(part_l2_kernel<4>, dim3(np1 * units), dim3(kL2Block), 0, ctx->stream, Q2); break;
    default: hipLaunchKernelGGL(part_l2_kernel<8>, dim3(np1 * units), dim3(kL2Block), 0, ctx->stream, Q2); break;
  }
  WG_HIP(hipGetLastError());
  WG_HIP(hipEventRecord(h->pev[3], ctx->stream));
  WG_HIP(hipEventSynchronize(h->pev[2]));
  const uint64_t total = bound[np];
  h->st_glb = total;  // rows written to the partitions (the cold rows)

  // 3: work items, largest first; partitions above the chunk size are split
  uint64_t n_items = 0, dense_bound = (uint64_t)H;
  for (uint32_t p = 0; p < np; ++p) {
    const uint64_t b = bound[p], e = bound[p + 1];
    if (e <= b) continue;
    const bool split = e - b > chunk;
    if (!split) dense_bound += std::min<uint64_t>(e - b, (uint64_t)L);
    for (uint64_t x = b; x < e; x += chunk)
      hitems[n_items++] = PartItem{(uint32_t)x, (uint32_t)std::min<uint64_t>(e, x + chunk), split ? 1u : 0u, 0};
  }
  WG_CHECK(n_items <= max_items && total <= n, WG_ERR_DEVICE, "partition boundaries out of range");
  std::stable_sort(hitems, hitems + n_items,
                   [](const PartItem& a, const PartItem& b) { return a.end - a.beg > b.end - b.beg; });
  if (n_items) {  // uploaded on the side stream while pass 2 runs
    WG_HIP(hipMemcpyAsync(h->part_items, hitems, n_items * sizeof(PartItem), hipMemcpyHostToDevice, h->side));
    WG_HIP(hipEventRecord(h->pev[7], h->side));
    WG_HIP(hipStreamWaitEvent(ctx->stream, h->pev[7], 0));
  }

  PartAggParams Q{};
  Q.pg = pg2;
  Q.recs = recs2;
  Q.items = (const PartItem*)h->part_items;
  Q.n_items = (uint32_t)n_items;
  Q.lds_slots = L;
  Q.item_ctr = (unsigned long long*)(h->counters + kCtrItem);
  Q.vfloat = P.vfloat;
  DenseOut D{};
  if (dense) {
    D = DenseOut{h->dense, h->n_sumd ? h->dense_base : nullptr, h->dense_null,
                 (unsigned long long*)(h->counters + kCtrDense), dense_bound, h->n_sumd * 2};
  }
  Q.dense = D;
  Q.max_item_rows = chunk;
  Q.sel_chunk = chunk;
  // Items are drained in launches; a launch stops taking items when the
  // overflow lists run out of headroom (table past its load budget): the host
  // then grows the table, folds the partials in, replays the overflowed
  // positions and resumes with the items not yet taken.
  struct Pending { uint32_t* rows; uint64_t n; };
  struct Replays {  // replay lists still to aggregate, freed on every exit
    std::vector<Pending> v;
    ~Replays() {
      for (auto& r : v) (void)hipFree(r.rows);
    }
  } pending;
  std::vector<Pending>& replays = pending.v;
  uint64_t items_done = 0;
  bool first = true;
  // With dense emission the hot keys go to the dense output and never touch
  // the table: their merge is queued right behind the first aggregate launch,
  // and its results are read back with that launch's counters.
  bool hot_done = false, dense_read = false;
  // the dense row count is cleared by the first launch's counter set-up
  bool dense_clear = dense;
  auto begin_round = [&](uint64_t start_item) -> wg_status {
    hipLaunchKernelGGL(part_round_begin_kernel, dim3(1), dim3(64), 0, ctx->stream, h->counters, start_item,
                       dense_clear ? 1 : 0);
    WG_HIP(hipGetLastError());
    dense_clear = false;
    return WG_OK;
  };
  // every counter word to the pinned words in one copy (synchronises the stream)
  auto read_back = [&](uint64_t* c) -> wg_status {
    WG_HIP(hipMemcpyAsync(h->pinned, h->counters, kAggCounterWords * 8, hipMemcpyDeviceToHost, ctx->stream));
    WG_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(c, h->pinned, CTR_N * 8);
    return WG_OK;
  };
  for (int round = 0; round < 64; ++round) {
    const bool replay = !replays.empty();
    const uint32_t* rsel = replay ? replays.back().rows : nullptr;
    const uint64_t rn = replay ? replays.back().n : 0;
    const uint64_t total_work = replay ? (rn + chunk - 1) / chunk : n_items;
    if (!replay && items_done >= n_items) break;
    const uint64_t start_item = replay ? 0 : items_done;
    WG_TRY(begin_round(start_item));
    Q.T = table_of(h);
    Q.sel = rsel;
    Q.n_sel = rn;
    Q.ovf_rows = h->ovf_rows;
    Q.ovf_rows_cap = h->ovf_rows_cap;
    Q.ovf_part = h->ovf_part;
    Q.ovf_part_cap = h->ovf_part_cap;
    const uint64_t left = total_work - start_item;
    const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)grid_max, left));
    WG_HIP(hipEventRecord(h->pev[4], ctx->stream));
    if (left) WG_TRY(launch_part_agg(RW, grid, h->part_block, lds_bytes, ctx->stream, Q));
    WG_HIP(hipEventRecord(h->pev[5], ctx->stream));
    const bool hot_now = H && dense && !hot_done;
    if (hot_now) {
      hipLaunchKernelGGL(hot_merge_kernel, dim3(H), dim3(256), 0, ctx->stream, (const uint64_t*)P.hot_part,
                         (uint32_t)G, (uint32_t)H, P.n_hot, h->pg, table_of(h), h->ovf_part, D);
      WG_HIP(hipGetLastError());
      WG_HIP(hipEventRecord(h->pev[6], ctx->stream));
      hot_done = true;
    }
    dense_read = dense;
    uint64_t c[CTR_N];
    WG_TRY(read_back(c));
    const uint64_t taken = h->pinned[kCtrItem];
    float ms = 0.f;
    if (first) {
      WG_HIP(hipEventElapsedTime(&ms, h->pev[0], h->pev[1]));
      h->pass_ms[1] = ms;
      WG_HIP(hipEventElapsedTime(&ms, h->pev[1], h->pev[3]));
      h->pass_ms[2] = ms;
      first = false;
    }
    WG_HIP(hipEventElapsedTime(&ms, h->pev[4], h->pev[hot_now ? 6 : 5]));
    h->pass_ms[3] += ms;
    h->groups = c[CTR_GROUPS];
    const uint64_t done_now = std::min<uint64_t>(taken, total_work);
    if (replay) {
      if (done_now >= total_work) {
        WG_HIP(hipFree(replays.back().rows));
        replays.pop_back();
      } else {
        // keep the untaken tail of this replay list
        Pending& r = replays.back();
        const uint64_t skip = done_now * chunk;
        uint32_t* tail = nullptr;
        WG_ALLOC(h->ctx, &tail, (r.n - skip) * 4, "replayed rows");
        WG_HIP(hipMemcpyAsync(tail, r.rows + skip, (r.n - skip) * 4, hipMemcpyDeviceToDevice, ctx->stream));
        WG_HIP(hipStreamSynchronize(ctx->stream));
        WG_HIP(hipFree(r.rows));
        r.rows = tail;
        r.n -= skip;
      }
    } else {
      items_done = done_now;
    }
    const uint64_t nr = c[CTR_OVF_ROWS], npart = c[CTR_OVF_PART];
    if (nr || npart) {
      WG_TRY(grow_to(h, next_pow2(std::max<uint64_t>((h->groups + nr + npart) * 2, h->cap * 2))));
      if (npart) WG_TRY(agg_merge_partials(h, h->ovf_part, npart));
      if (nr) {
        uint32_t* rows = nullptr;
        WG_ALLOC(h->ctx, &rows, nr * 4, "replayed rows");
        WG_HIP(hipMemcpyAsync(rows, h->ovf_rows, nr * 4, hipMemcpyDeviceToDevice, ctx->stream));
        WG_HIP(hipStreamSynchronize(ctx->stream));
        replays.push_back(Pending{rows, nr});
      }
    }
  }
  WG_CHECK(replays.empty() && items_done >= n_items, WG_ERR_DEVICE, "partitioned consume did not converge");
  if (H && !hot_done) {
    // E: the hot keys' per-workgroup partials into the global table (or the
    // dense output when no aggregate launch ran)
    WG_TRY(begin_round(0));
    WG_HIP(hipEventRecord(h->pev[4], ctx->stream));
    hipLaunchKernelGGL(hot_merge_kernel, dim3(H), dim3(256), 0, ctx->stream, (const uint64_t*)P.hot_part, (uint32_t)G,
                       (uint32_t)H, P.n_hot, h->pg, table_of(h), h->ovf_part, D);
    WG_HIP(hipGetLastError());
    WG_HIP(hipEventRecord(h->pev[5], ctx->stream));
    uint64_t c[CTR_N];
    WG_TRY(read_back(c));
    dense_read = dense;
    float ms = 0.f;
    WG_HIP(hipEventElapsedTime(&ms, h->pev[4], h->pev[5]));
    h->pass_ms[3] += ms;
    h->groups = c[CTR_GROUPS];
    if (c[CTR_OVF_PART]) {
      WG_TRY(grow_to(h, next_pow2(std::max<uint64_t>((h->groups + c[CTR_OVF_PART]) * 2, h->cap * 2))));
      WG_TRY(agg_merge_partials(h, h->ovf_part, c[CTR_OVF_PART]));
    }
  }
  if (dense) {
    // no launch ran: nothing was written to the dense output
    h->n_dense = dense_read ? h->pinned[kCtrDense] : 0;
    WG_CHECK(h->n_dense <= D.cap, WG_ERR_DEVICE, "dense output overran its reservation");
    if (h->groups) h->dirty = true;  // split items, LDS misses or replays used the table
  }
  return WG_OK;
}

}  // namespace wg
