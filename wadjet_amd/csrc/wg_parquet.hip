// wg_parquet.hip — Parquet data pages decoded on the device.
//
// The reference reads Parquet with its own reader: per column chunk a
// ColumnPageReader decompresses each page, decodes the definition levels and
// the values — PLAIN, or dictionary indices in the RLE / bit-packing hybrid
// resolved against the dictionary page — into a Values vector
// (internal/storage/parquet/page_reader.go:434-700, rle.go, reader.go:
// 1020-1060). Here the host keeps the byte-level work that is sequential by
// nature — page headers, decompression, and the run headers of each hybrid
// stream (a few bytes per run of up to hundreds of values), on a pool of
// threads per column (wg_pq_stage, wg_pq_stage.cpp) — and the device does the
// per-value work:
//   count    present rows (definition level 1) of each 1024-row segment of
//            a nullable page, read from the level runs (RLE value, or bits
//            extracted from a bit-packed run);
//   scan     exclusive sum of the segment counts;
//   values   one wavefront per segment, 64 rows at a time: the ballot of the
//            present flags numbers the present rows, and each takes its
//            PLAIN value or the dictionary entry its hybrid-coded index names
//            (out-of-range indices are counted and reported, as
//            ValidateDictIndices does); NULL rows get 0 and a cleared
//            validity bit.
// Pages without NULLs skip the first two. Each stream's run cursor moves
// forward with the rows (one binary search per segment), and the pages of
// whole column chunks decode in one grid of 4096-row tiles
// (wg_parquet_decode_pages): at most three launches and one synchronisation
// however many pages there are.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <utility>
#include <vector>

#include "wg_core.hpp"

namespace wg {

// One run of a hybrid stream: first value index, value count, the RLE value
// or the bit offset of a bit-packed run, 1 when bit-packed.
struct PqRun {
  uint32_t start, count, payload, packed;
};

// From run r (whose start <= i) forward to the run holding i.
__device__ __forceinline__ uint32_t run_advance(const PqRun* runs, uint32_t n_runs, uint32_t r, uint64_t i) {
  while (r + 1 < n_runs && runs[r + 1].start <= i) ++r;
  return r;
}

// Value i of a stream from its run; false when i lies past the stream.
__device__ __forceinline__ bool run_value(const wg_pq_runs& S, const PqRun& R, uint64_t i, uint32_t* out) {
  if (i - R.start >= R.count) return false;
  if (!R.packed) {
    *out = R.payload;
    return true;
  }
  // values are packed least-significant bit first (parquet-format Encodings.md)
  const uint64_t bit = (uint64_t)R.payload + (i - R.start) * (uint64_t)S.bit_width;
  const uint64_t* w = (const uint64_t*)S.data;
  const uint64_t q = bit >> 6;
  const uint32_t s = (uint32_t)(bit & 63);
  const uint64_t v = s ? (w[q] >> s) | (w[q + 1] << (64 - s)) : w[q];
  *out = (uint32_t)(v & ((1ull << S.bit_width) - 1));
  return true;
}

// A page as the kernels see it: where its rows start in the batch and its
// first tile. Workgroup b decodes tile b (4096 rows) of the page whose tile
// range holds b, so the pages of whole column chunks decode in one grid; each
// of its four wavefronts owns one 1024-row segment of the tile.
constexpr uint32_t kPqSeg = 1024;  // rows per wavefront: a multiple of 64, at most 4096
constexpr uint32_t kPqSegsPerTile = 4;
constexpr uint32_t kPqTile = kPqSeg * kPqSegsPerTile;

struct PqPageDev {
  wg_pq_page p;
  uint64_t row_start;
  uint64_t tile0;
  uint64_t chunk0;  // first 64-row chunk of the page
};

// The page and the rows of the segment that wavefront `wave` of tile b owns.
struct PqSeg {
  wg_pq_page P;
  uint64_t row_start, tile0, chunk0, s0, s1;
};

__device__ __forceinline__ PqSeg segment_of(const PqPageDev* pages, const uint32_t* tile_page, uint32_t wave) {
  const PqPageDev& pg = pages[tile_page[blockIdx.x]];
  PqSeg g;
  g.P = pg.p;
  g.row_start = pg.row_start;
  g.tile0 = pg.tile0;
  g.chunk0 = pg.chunk0;
  g.s0 = (blockIdx.x - g.tile0) * (uint64_t)kPqTile + wave * (uint64_t)kPqSeg;
  g.s1 = min(g.s0 + kPqSeg, g.P.num_rows);
  return g;
}

// The last run whose first value index <= i, found by the whole wavefront 64
// candidates at a time (a run table of 4096 runs takes two rounds). Called
// with every lane active.
__device__ __forceinline__ uint32_t run_search_wave(const PqRun* runs, uint32_t n_runs, uint64_t i, int lane) {
  uint32_t lo = 0, len = n_runs;  // runs[lo].start <= i, and the answer lies in [lo, lo + len)
  while (len > 1) {
    const uint32_t step = (len + 63) / 64;
    const uint32_t idx = lo + step * (uint32_t)lane;
    const bool le = idx < lo + len && runs[idx].start <= i;
    const uint64_t b = __ballot(le) | 1ull;
    const uint32_t h = 63u - (uint32_t)__clzll(b);
    const uint32_t nlo = lo + step * h;
    len = min(step, lo + len - nlo);
    lo = nlo;
  }
  return lo;
}

// bw bits of a bit-packed stream at bit offset `bit`.
__device__ __forceinline__ uint64_t packed_bits(const uint8_t* data, uint64_t bit) {
  const uint64_t* w = (const uint64_t*)data;
  const uint64_t q = bit >> 6;
  const uint32_t s = (uint32_t)(bit & 63);
  return s ? (w[q] >> s) | (w[q + 1] << (64 - s)) : w[q];
}

// Present flags (definition level 1) of rows j .. min(j + 64, s1) - 1 of the
// page, bit l for row j + l, walking the runs that overlap the chunk from run
// r (which holds row j): one word per bit-packed run, nothing for an RLE run.
// *ok goes false when a row lies past the stream.
__device__ __forceinline__ uint64_t chunk_present(const wg_pq_runs& D, uint32_t r, uint64_t j, uint64_t s1,
                                                  bool* ok) {
  const PqRun* runs = (const PqRun*)D.runs;
  const uint64_t end = min(j + 64, s1);
  uint64_t m = 0;
  for (uint64_t pos = j; pos < end; ++r) {
    if (r >= D.n_runs) {
      *ok = false;
      break;
    }
    const PqRun R = runs[r];
    const uint64_t stop = min(end, (uint64_t)R.start + R.count);
    if (stop <= pos) {
      *ok = false;
      break;
    }
    const uint32_t n = (uint32_t)(stop - pos);
    const uint64_t nm = n >= 64 ? ~0ull : (1ull << n) - 1;
    const uint64_t bits = !R.packed ? (R.payload == 1u ? nm : 0ull)
                                    : packed_bits(D.data, (uint64_t)R.payload + (pos - R.start)) & nm;
    m |= bits << (pos - j);
    pos = stop;
  }
  return m;
}

// Up to 64 consecutive runs of a stream held by a wavefront, lane l holding
// run first + l, so the run of an index is found by a ballot instead of a
// load; reloaded when the indices move past it.
struct RunWindow {
  uint32_t first, n;                     // first run held, runs held
  uint32_t start, end, payload, packed;  // this lane's run
  uint64_t limit;                        // first index past the window
};

__device__ __forceinline__ void win_load(const wg_pq_runs& S, RunWindow& w, uint32_t first, int lane) {
  w.first = first;
  w.n = min(64u, S.n_runs - first);
  w.start = w.end = 0xFFFFFFFFu;
  w.payload = w.packed = 0;
  if ((uint32_t)lane < w.n) {
    const PqRun R = ((const PqRun*)S.runs)[first + lane];
    w.start = R.start;
    w.end = R.start + R.count;
    w.payload = R.payload;
    w.packed = R.packed;
  }
  w.limit = (uint32_t)__builtin_amdgcn_readlane((int)w.end, (int)w.n - 1);
}

constexpr uint32_t kPqChunks = kPqSeg / 64;  // 64-row chunks per segment
constexpr int kPqUnroll = 4;                 // chunks in flight per wavefront
constexpr uint32_t kPqLdsDict = 2048;        // dictionaries up to this many entries are staged in LDS
static_assert(kPqTile == 64 * 64, "the mask pass gives each lane of a wavefront one chunk of a tile");

__device__ __forceinline__ uint64_t low_bits(uint64_t n) { return n >= 64 ? ~0ull : (1ull << n) - 1; }

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
  const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src);
  const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}

// Present flags and present-row count of every 64-row chunk of the batch's
// pages: one wavefront per 4096-row tile, lane l taking chunk l. Each lane
// finds its chunk's first run by binary search between the runs that hold
// the tile's first and last rows (two wavefront-wide searches), then reads
// the runs over the chunk. Pages without definition levels get all rows.
__global__ void pq_mask_kernel(const PqPageDev* pages, const uint32_t* tile_page, uint64_t* masks, uint32_t* counts,
                               unsigned long long* bad) {
  const PqPageDev& pg = pages[tile_page[blockIdx.x]];
  const wg_pq_page P = pg.p;
  const uint64_t chunk0 = pg.chunk0;
  const int lane = __lane_id();
  const uint64_t t0 = (blockIdx.x - pg.tile0) * (uint64_t)kPqTile;
  const uint64_t t1 = min(t0 + kPqTile, P.num_rows);
  const uint64_t c = t0 / 64 + (uint32_t)lane;  // chunk within the page
  const uint64_t j = c * 64;
  uint64_t m = 0;
  bool ok = true;
  if (P.def.n_runs == 0) {
    if (j < P.num_rows) m = low_bits(P.num_rows - j);
  } else {
    const PqRun* runs = (const PqRun*)P.def.runs;
    const uint32_t lo = run_search_wave(runs, P.def.n_runs, t0, lane);
    const uint32_t hi = run_search_wave(runs, P.def.n_runs, t1 - 1, lane);
    if (j < P.num_rows) {
      uint32_t a = lo, b = hi + 1;  // runs[a].start <= j < runs[b].start
      while (b - a > 1) {
        const uint32_t mid = (a + b) >> 1;
        if (runs[mid].start <= j) a = mid;
        else b = mid;
      }
      m = chunk_present(P.def, a, j, P.num_rows, &ok);
    }
  }
  if (j < P.num_rows) {
    masks[chunk0 + c] = m;
    counts[chunk0 + c] = (uint32_t)__popcll(m);
  }
  if (__ballot(!ok) && lane == 0) atomicAdd(bad, 1ull);
}

// Each wavefront decodes its segment (kPqChunks chunks): lane c first holds
// chunk c's present flags and the page's present rows before it (from the
// mask pass and the scan of its counts; arithmetic for a page without
// NULLs), then kPqUnroll chunks at a time each present row takes its PLAIN
// value or the dictionary entry its index names, and lane c writes the
// validity word(s) its chunk covers — plain stores for words inside the
// segment, atomicOr for the two it shares with its neighbours. out and the
// validity bit index are offset by row0.
template <typename T>
__global__ void pq_values_kernel(const PqPageDev* pages, const uint32_t* tile_page, const uint64_t* masks,
                                 const uint32_t* prefix, T* out, uint64_t* out_valid, uint64_t row0,
                                 unsigned long long* bad) {
  const uint32_t wave = threadIdx.x >> 6;
  const int lane = __lane_id();
  const PqSeg g = segment_of(pages, tile_page, wave);
  const wg_pq_page& P = g.P;
  const uint64_t s0 = g.s0, s1 = g.s1;
  // a small dictionary is staged in LDS once for the workgroup (its four
  // wavefronts decode one page), so an index resolves without a global load
  __shared__ T sdict[kPqLdsDict];
  const bool lds_dict = P.encoding != 0 && P.dict != nullptr && P.dict_n <= kPqLdsDict;
  if (lds_dict) {
    for (uint32_t e = threadIdx.x; e < (uint32_t)P.dict_n; e += blockDim.x) sdict[e] = ((const T*)P.dict)[e];
    __syncthreads();
  }
  if (s0 >= s1) return;
  const bool nullable = P.def.n_runs != 0;
  const bool dict = P.encoding != 0;
  const bool has_idx = dict && P.idx.n_runs != 0;
  const uint32_t n_chunks = (uint32_t)((s1 - s0 + 63) / 64);
  const uint64_t my_j = s0 + 64ull * (uint32_t)lane;
  uint64_t my_mask = 0, my_before = 0;
  if ((uint32_t)lane < n_chunks) {
    if (nullable) {
      const uint64_t ci = g.chunk0 + my_j / 64;
      my_mask = masks[ci];
      my_before = (uint32_t)(prefix[ci] - prefix[g.chunk0]);
    } else {
      my_mask = low_bits(P.num_rows - my_j);
      my_before = my_j;
    }
  }
  RunWindow iw{};
  if (has_idx)
    win_load(P.idx, iw, run_search_wave((const PqRun*)P.idx.runs, P.idx.n_runs, shfl64(my_before, 0), lane), lane);
  const uint64_t lt = (1ull << lane) - 1;  // lanes below this one
  T* dst = out + row0 + g.row_start;
  uint32_t nbad = 0;
  for (uint32_t c0 = 0; c0 < n_chunks; c0 += kPqUnroll) {
    uint64_t m[kPqUnroll], kb[kPqUnroll];
    T v[kPqUnroll];
#pragma unroll
    for (int u = 0; u < kPqUnroll; ++u) {
      m[u] = shfl64(my_mask, (int)(c0 + u) & 63);
      kb[u] = shfl64(my_before, (int)(c0 + u) & 63);
      if (c0 + u >= n_chunks) m[u] = 0;
    }
#pragma unroll
    for (int u = 0; u < kPqUnroll; ++u) {
      v[u] = 0;
      const bool has = (m[u] >> lane) & 1;
      const uint64_t k = kb[u] + (uint64_t)__popcll(m[u] & lt);
      if (!dict) {
        if (has) v[u] = ((const T*)P.plain)[k];
      } else if (m[u]) {
        // the run of the chunk's first index, from the window
        if (has_idx && kb[u] >= iw.limit && iw.first + iw.n < P.idx.n_runs)
          win_load(P.idx, iw, run_search_wave((const PqRun*)P.idx.runs, P.idx.n_runs, kb[u], lane), lane);
        const uint64_t held = __ballot((uint32_t)lane < iw.n && iw.start <= kb[u]);
        const int h = held ? 63 - __clzll(held) : 0;
        const uint32_t hs = (uint32_t)__builtin_amdgcn_readlane((int)iw.start, h);
        const uint32_t he = (uint32_t)__builtin_amdgcn_readlane((int)iw.end, h);
        const uint32_t hp = (uint32_t)__builtin_amdgcn_readlane((int)iw.payload, h);
        const uint32_t hk = (uint32_t)__builtin_amdgcn_readlane((int)iw.packed, h);
        if (has) {
          uint32_t d = 0;
          bool got = false;
          if (!has_idx) {
          } else if (held && k >= hs && k < he) {
            got = true;
            if (!hk) d = hp;
            else d = (uint32_t)(packed_bits(P.idx.data, (uint64_t)hp + (k - hs) * (uint64_t)P.idx.bit_width) &
                                ((1ull << P.idx.bit_width) - 1));
          } else {
            // past the run of the chunk's first index: walk the runs from it
            const PqRun* runs = (const PqRun*)P.idx.runs;
            const uint32_t r0 = held ? iw.first + (uint32_t)h : 0u;
            got = run_value(P.idx, runs[run_advance(runs, P.idx.n_runs, r0, k)], k, &d);
          }
          if (got && d < P.dict_n) v[u] = lds_dict ? sdict[d] : ((const T*)P.dict)[d];
          else ++nbad;  // no index stream, an index past it, or one past the dictionary
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kPqUnroll; ++u) {
      const uint64_t j = s0 + 64ull * (c0 + u);
      if (j + lane < s1) dst[j + lane] = v[u];
    }
  }
  if (out_valid) {
    const uint64_t next = shfl64(my_mask, (lane + 1) & 63);
    if ((uint32_t)lane < n_chunks) {
      const uint64_t r = row0 + g.row_start + my_j;  // output row of the chunk's first row
      const uint32_t sh = (uint32_t)(r & 63);
      const bool full = my_j + 64 <= P.num_rows;
      if (sh == 0) {
        // the chunk is exactly one word
        if (full) out_valid[r >> 6] = my_mask;
        else if (my_mask) atomicOr((unsigned long long*)&out_valid[r >> 6], (unsigned long long)my_mask);
      } else {
        // word r >> 6 holds this chunk's first 64 - sh rows after rows of the
        // chunk before it (the lane below, or another segment for lane 0);
        // word (r >> 6) + 1 this chunk's last sh rows and the next chunk's first
        if (lane == 0 && my_mask << sh)
          atomicOr((unsigned long long*)&out_valid[r >> 6], (unsigned long long)(my_mask << sh));
        const bool next_full = (uint32_t)lane + 1 < n_chunks && my_j + 128 <= P.num_rows;
        const uint64_t w = (my_mask >> (64 - sh)) | ((uint32_t)lane + 1 < n_chunks ? next << sh : 0ull);
        if (next_full) out_valid[(r >> 6) + 1] = w;
        else if (w) atomicOr((unsigned long long*)&out_valid[(r >> 6) + 1], (unsigned long long)w);
      }
    }
  }
  if (nbad) atomicAdd(bad, (unsigned long long)nbad);
}

}  // namespace wg

using namespace wg;

extern "C" {

wg_status wg_parquet_decode_pages(wg_ctx* ctx, const wg_pq_page* pages, uint64_t n_pages, void* out,
                                  uint64_t* out_valid, uint64_t row0) {
  clear_error();
  WG_CHECK(ctx && out && (pages || n_pages == 0), WG_ERR_INVALID, "null argument");
  if (n_pages == 0) return WG_OK;
  const int32_t type = pages[0].physical_type;
  WG_CHECK(type == 1 || type == 2 || type == 5, WG_ERR_UNSUPPORTED, "parquet physical type %d not on the GPU path",
           type);
  // the page table with each page's first row and tile; empty pages dropped
  std::vector<PqPageDev> table;
  table.reserve(n_pages);
  std::vector<std::pair<size_t, uint64_t>> groups;  // (first table entry, rows) of each launch group
  bool any_nullable = false;
  uint64_t rows = 0, group_rows = 0, tiles = 0, chunks = 0;
  uint64_t max_chunks = 0;
  for (uint64_t i = 0; i < n_pages; ++i) {
    const wg_pq_page& P = pages[i];
    WG_CHECK(P.physical_type == type, WG_ERR_INVALID, "page %llu is of type %d, page 0 of %d", (unsigned long long)i,
             P.physical_type, type);
    WG_CHECK(P.encoding == 0 || P.encoding == 2 || P.encoding == 8, WG_ERR_UNSUPPORTED,
             "parquet encoding %d not on the GPU path", P.encoding);
    WG_CHECK(P.encoding == 0 ? P.plain != nullptr || P.num_rows == 0 : P.dict != nullptr || P.idx.n_runs == 0,
             WG_ERR_INVALID, "page %llu without its values", (unsigned long long)i);
    WG_CHECK(P.num_rows < (1ull << 30), WG_ERR_UNSUPPORTED, "page of %llu rows", (unsigned long long)P.num_rows);
    if (P.num_rows == 0) continue;
    // a launch group's rows stay below 2^30 (the scan counts in 32 bits)
    if (groups.empty() || group_rows + P.num_rows > (1ull << 30)) {
      groups.push_back({table.size(), 0});
      group_rows = 0;
      tiles = 0;
      chunks = 0;
    }
    table.push_back(PqPageDev{P, group_rows, tiles, chunks});
    any_nullable |= P.def.n_runs != 0;
    group_rows += P.num_rows;
    tiles += (P.num_rows + kPqTile - 1) / kPqTile;
    chunks += (P.num_rows + 63) / 64;
    groups.back().second = group_rows;
    max_chunks = std::max(max_chunks, chunks);
    rows += P.num_rows;
  }
  if (table.empty()) return WG_OK;
  // tile -> page (within its launch group), after the page table in one upload
  std::vector<uint64_t> tile_off(groups.size() + 1, 0);
  for (size_t g = 0; g < groups.size(); ++g) {
    const size_t last = g + 1 < groups.size() ? groups[g + 1].first : table.size();
    const PqPageDev& tail = table[last - 1];
    tile_off[g + 1] = tile_off[g] + tail.tile0 + (tail.p.num_rows + kPqTile - 1) / kPqTile;
  }
  const size_t table_bytes = table.size() * sizeof(PqPageDev);
  const size_t upload_bytes = table_bytes + tile_off.back() * 4;
  const size_t need = (upload_bytes + 7) / 8 * 8 + 8;  // + the error count read back
  std::lock_guard<std::mutex> upload_lock(ctx->upload_mu);  // held to the synchronising return
  if (ctx->upload_cap < need) {
    if (ctx->upload) WG_HIP(hipHostFree(ctx->upload));
    ctx->upload = nullptr;
    ctx->upload_cap = 0;
    WG_HIP(hipHostMalloc(&ctx->upload, need, hipHostMallocDefault));
    ctx->upload_cap = need;
  }
  uint8_t* upload = (uint8_t*)ctx->upload;
  uint64_t* nbad_host = (uint64_t*)(upload + (upload_bytes + 7) / 8 * 8);
  memcpy(upload, table.data(), table_bytes);
  uint32_t* map = (uint32_t*)(upload + table_bytes);
  for (size_t g = 0; g < groups.size(); ++g) {
    const size_t last = g + 1 < groups.size() ? groups[g + 1].first : table.size();
    for (size_t e = groups[g].first; e < last; ++e) {
      const uint64_t t0 = table[e].tile0, nt = (table[e].p.num_rows + kPqTile - 1) / kPqTile;
      for (uint64_t t = 0; t < nt; ++t) map[tile_off[g] + t0 + t] = (uint32_t)(e - groups[g].first);
    }
  }
  PqPageDev* dtable = nullptr;
  // per 64-row chunk: present flags, present-row counts, their exclusive sums
  uint64_t* masks = nullptr;
  unsigned long long* bad = nullptr;
  // scratch goes back to the stream's pool on every exit, errors included
  struct Release {
    hipStream_t st;
    void** p[3];
    ~Release() {
      for (void** q : p)
        if (*q) (void)hipFreeAsync(*q, st);
    }
  } release{ctx->stream, {(void**)&dtable, (void**)&masks, (void**)&bad}};
  uint32_t* counts = nullptr;
  uint32_t* prefix = nullptr;
  void* scan_tmp = nullptr;
  size_t scan_bytes = 0;
  WG_ALLOC_ASYNC(ctx, &bad, 8, ctx->stream, "an error count");
  WG_HIP(hipMemsetAsync(bad, 0, 8, ctx->stream));
  WG_ALLOC_ASYNC(ctx, &dtable, upload_bytes, ctx->stream, "the page table");
  WG_HIP(hipMemcpyAsync(dtable, upload, upload_bytes, hipMemcpyHostToDevice, ctx->stream));
  const uint32_t* dmap = (const uint32_t*)((const uint8_t*)dtable + table_bytes);
  if (any_nullable) {
    WG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                            (int)max_chunks, ctx->stream));
    WG_ALLOC_ASYNC(ctx, &masks, max_chunks * 16 + scan_bytes, ctx->stream, "decode scratch");
    counts = (uint32_t*)(masks + max_chunks);
    prefix = counts + max_chunks;
    scan_tmp = prefix + max_chunks;
  }
  uint64_t r0 = row0;
  for (size_t g = 0; g < groups.size(); ++g) {
    const size_t first = groups[g].first;
    const uint64_t grid = tile_off[g + 1] - tile_off[g];
    const PqPageDev* dt = dtable + first;
    const uint32_t* tp = dmap + tile_off[g];
    if (any_nullable) {
      const size_t last = g + 1 < groups.size() ? groups[g + 1].first : table.size();
      const uint64_t n_chunks = table[last - 1].chunk0 + (table[last - 1].p.num_rows + 63) / 64;
      hipLaunchKernelGGL(pq_mask_kernel, dim3((uint32_t)grid), dim3(64), 0, ctx->stream, dt, tp, masks, counts, bad);
      WG_HIP(hipGetLastError());
      WG_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, counts, prefix, (int)n_chunks, ctx->stream));
    }
    if (type == 1)
      hipLaunchKernelGGL(pq_values_kernel<uint32_t>, dim3((uint32_t)grid), dim3(256), 0, ctx->stream, dt, tp,
                         (const uint64_t*)masks, (const uint32_t*)prefix, (uint32_t*)out, out_valid, r0, bad);
    else
      hipLaunchKernelGGL(pq_values_kernel<uint64_t>, dim3((uint32_t)grid), dim3(256), 0, ctx->stream, dt, tp,
                         (const uint64_t*)masks, (const uint32_t*)prefix, (uint64_t*)out, out_valid, r0, bad);
    WG_HIP(hipGetLastError());
    r0 += groups[g].second;
  }
  WG_HIP(hipMemcpyAsync(nbad_host, bad, 8, hipMemcpyDeviceToHost, ctx->stream));
  WG_HIP(hipStreamSynchronize(ctx->stream));
  const uint64_t nbad = *nbad_host;
  WG_CHECK(nbad == 0, WG_ERR_INVALID, "%llu values with their dictionary index missing or out of range, or past "
           "a stream's end", (unsigned long long)nbad);
  return WG_OK;
}

wg_status wg_parquet_decode_page(wg_ctx* ctx, const wg_pq_page* page, void* out, uint64_t* out_valid, uint64_t row0) {
  clear_error();
  WG_CHECK(page, WG_ERR_INVALID, "null argument");
  return wg_parquet_decode_pages(ctx, page, 1, out, out_valid, row0);
}

}  // extern "C"
