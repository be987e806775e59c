// wg_parquet.hip — Parquet data pages decoded on the device.
//
// The reference reads Parquet with its own reader: per column chunk a
// ColumnPageReader decompresses each page, decodes the definition levels and
// the values — PLAIN, or dictionary indices in the RLE / bit-packing hybrid
// resolved against the dictionary page — into a Values vector
// (internal/storage/parquet/page_reader.go:434-700, rle.go, reader.go:
// 1020-1060). Here the host keeps the byte-level work that is sequential by
// nature — page headers, decompression, and the run headers of each hybrid
// stream (wg_pq_scan_runs, a few bytes per run of up to hundreds of values) —
// and the device does the per-value work:
//   levels   each row's definition level from its run (RLE value, or bits
//            extracted from a bit-packed run), 1 = a value is present;
//   scan     exclusive sum of the present flags = each row's value index;
//   values   each present row's value: PLAIN bytes at its value index, or the
//            dictionary entry its hybrid-coded index names (out-of-range
//            indices are counted and reported, as ValidateDictIndices does);
//            NULL rows get 0 and a cleared validity bit.
// Runs are found by binary search over the run table, so every value decodes
// independently, and the pages of whole column chunks decode in one grid of
// 4096-row tiles (wg_parquet_decode_pages): three launches and one
// synchronisation however many pages there are.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <utility>
#include <vector>

#include "wg_core.hpp"

namespace wg {

// One run of a hybrid stream: first value index, value count, the RLE value
// or the bit offset of a bit-packed run, 1 when bit-packed.
struct PqRun {
  uint32_t start, count, payload, packed;
};

__device__ __forceinline__ uint32_t hybrid_value(const wg_pq_runs& S, uint64_t i) {
  const PqRun* runs = (const PqRun*)S.runs;
  uint32_t lo = 0, hi = S.n_runs;  // the last run whose start <= i
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (runs[mid].start <= i) lo = mid;
    else hi = mid;
  }
  const PqRun r = runs[lo];
  if (!r.packed) return r.payload;
  // values are packed least-significant bit first (parquet-format Encodings.md)
  const uint64_t bit = (uint64_t)r.payload + (i - r.start) * (uint64_t)S.bit_width;
  const uint64_t* w = (const uint64_t*)S.data;
  const uint64_t q = bit >> 6;
  const uint32_t s = (uint32_t)(bit & 63);
  const uint64_t v = s ? (w[q] >> s) | (w[q + 1] << (64 - s)) : w[q];
  return (uint32_t)(v & ((1ull << S.bit_width) - 1));
}

// A page as the kernels see it: where its rows start in the batch and its
// first 4096-row tile. Workgroup b decodes tile b of the page whose tile range
// holds b, so every page of a column chunk (or of several) decodes in one grid.
constexpr uint32_t kPqTile = 4096;

struct PqPageDev {
  wg_pq_page p;
  uint64_t row_start;
  uint64_t tile0;
};

__device__ __forceinline__ const PqPageDev& page_of_tile(const PqPageDev* pages, uint32_t n_pages, uint64_t t) {
  uint32_t lo = 0, hi = n_pages;  // the last page whose first tile <= t (pages hold >= 1 row)
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pages[mid].tile0 <= t) lo = mid;
    else hi = mid;
  }
  return pages[lo];
}

__global__ void pq_levels_kernel(const PqPageDev* pages, uint32_t n_pages, uint32_t* present) {
  const PqPageDev& pg = page_of_tile(pages, n_pages, blockIdx.x);
  const uint64_t i0 = (blockIdx.x - pg.tile0) * (uint64_t)kPqTile;
  const uint64_t i1 = min(i0 + kPqTile, pg.p.num_rows);
  const bool nullable = pg.p.def.n_runs != 0;
  for (uint64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x)
    present[pg.row_start + i] = !nullable || hybrid_value(pg.p.def, i) == 1u ? 1u : 0u;
}

// out and the validity bit index are offset by row0; present / vidx are
// indexed by the row within the batch.
template <typename T>
__global__ void pq_values_kernel(const PqPageDev* pages, uint32_t n_pages, const uint32_t* present,
                                 const uint32_t* vidx, T* out, uint64_t* out_valid, uint64_t row0,
                                 unsigned long long* bad) {
  const PqPageDev& pg = page_of_tile(pages, n_pages, blockIdx.x);
  const wg_pq_page& P = pg.p;
  const uint64_t i0 = (blockIdx.x - pg.tile0) * (uint64_t)kPqTile;
  const uint64_t i1 = min(i0 + kPqTile, P.num_rows);
  const uint64_t base = pg.row_start;
  const bool nullable = P.def.n_runs != 0;
  const uint32_t v0 = nullable ? vidx[base] : 0u;  // values before the page
  const int lane = __lane_id();
  for (uint64_t j = i0; j < i1; j += blockDim.x) {
    const uint64_t i = j + threadIdx.x;
    const bool in = i < i1;
    const bool has = in && (!nullable || present[base + i]);
    T v = 0;
    if (has) {
      const uint64_t k = nullable ? (uint64_t)(vidx[base + i] - v0) : i;
      if (P.encoding == 0) {
        v = ((const T*)P.plain)[k];
      } else if (P.idx.n_runs == 0) {
        atomicAdd(bad, 1ull);  // a present row but no index stream
      } else {
        const uint32_t d = hybrid_value(P.idx, k);
        if (d < P.dict_n) v = ((const T*)P.dict)[d];
        else atomicAdd(bad, 1ull);
      }
    }
    if (in) out[row0 + base + i] = v;
    if (out_valid) {
      // the wavefront's 64 rows touch at most two bitmap words
      const uint64_t m = __ballot(has);
      if (lane == 0 && m) {
        const uint64_t g = row0 + base + j + (threadIdx.x & ~63u);
        const uint32_t s = (uint32_t)(g & 63);
        atomicOr((unsigned long long*)&out_valid[g >> 6], (unsigned long long)(m << s));
        if (s && (m >> (64 - s))) atomicOr((unsigned long long*)&out_valid[(g >> 6) + 1], (unsigned long long)(m >> (64 - s)));
      }
    }
  }
}

}  // namespace wg

using namespace wg;

extern "C" {

wg_status wg_pq_scan_runs(const uint8_t* data, uint64_t len, int32_t bit_width, uint64_t count, uint32_t* runs,
                          uint64_t cap, uint64_t* n_runs, uint64_t* used, uint64_t* ones) {
  clear_error();
  WG_CHECK(data || len == 0, WG_ERR_INVALID, "null data");
  WG_CHECK(n_runs && used, WG_ERR_INVALID, "null argument");
  WG_CHECK(bit_width >= 0 && bit_width <= 32, WG_ERR_INVALID, "bit width %d", bit_width);
  WG_CHECK(count < (1ull << 32) && len < (1ull << 29), WG_ERR_UNSUPPORTED, "a stream of %llu values / %llu bytes",
           (unsigned long long)count, (unsigned long long)len);
  const int vbytes = (bit_width + 7) / 8;
  uint64_t off = 0, got = 0, nr = 0, n1 = 0;
  // RLE / bit-packing hybrid (rle.go:50-120): varint header; header & 1 = 0:
  // RLE run of header >> 1 values, the value in ceil(bw / 8) bytes; else a
  // bit-packed run of (header >> 1) groups of 8 values, bw bytes per group
  while (got < count) {
    WG_CHECK(off < len, WG_ERR_INVALID, "hybrid stream ends after %llu of %llu values", (unsigned long long)got,
             (unsigned long long)count);
    uint64_t header = 0;
    int shift = 0;
    for (;;) {
      WG_CHECK(off < len && shift < 64, WG_ERR_INVALID, "bad run header varint");
      const uint8_t b = data[off++];
      header |= (uint64_t)(b & 0x7F) << shift;
      shift += 7;
      if (!(b & 0x80)) break;
    }
    uint64_t n;
    uint32_t payload = 0, packed = 0;
    if (header & 1) {
      const uint64_t groups = header >> 1;
      n = groups * 8;
      payload = (uint32_t)(off * 8);
      packed = 1;
      WG_CHECK(off + groups * (uint64_t)bit_width <= len, WG_ERR_INVALID, "bit-packed run past the stream");
      if (bit_width == 1) {  // ones among this run's values below `count`
        const uint64_t take = std::min<uint64_t>(n, count - got);
        for (uint64_t v = 0; v < take; ++v) n1 += (data[off + v / 8] >> (v % 8)) & 1u;
      }
      off += groups * (uint64_t)bit_width;
    } else {
      n = header >> 1;
      WG_CHECK(off + vbytes <= len, WG_ERR_INVALID, "RLE value past the stream");
      for (int i = 0; i < vbytes; ++i) payload |= (uint32_t)data[off + i] << (8 * i);
      off += vbytes;
      if (payload == 1) n1 += std::min<uint64_t>(n, count - got);
    }
    if (n == 0) continue;
    if (nr < cap && runs) {
      uint32_t* r = runs + 4 * nr;
      r[0] = (uint32_t)got;
      r[1] = (uint32_t)std::min<uint64_t>(n, count - got);
      r[2] = payload;
      r[3] = packed;
    }
    ++nr;
    got += n;  // a bit-packed run may pad its last group past `count`
  }
  *n_runs = nr;
  *used = off;
  if (ones) *ones = n1;
  if (nr > cap && runs) {
    set_error("%llu runs do not fit %llu", (unsigned long long)nr, (unsigned long long)cap);
    return WG_ERR_INVALID;
  }
  return WG_OK;
}

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
  uint64_t rows = 0, group_rows = 0, tiles = 0;
  uint64_t max_group = 0;
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
    }
    table.push_back(PqPageDev{P, group_rows, tiles});
    any_nullable |= P.def.n_runs != 0;
    group_rows += P.num_rows;
    tiles += (P.num_rows + kPqTile - 1) / kPqTile;
    groups.back().second = group_rows;
    max_group = std::max(max_group, group_rows);
    rows += P.num_rows;
  }
  if (table.empty()) return WG_OK;
  PqPageDev* dtable = nullptr;
  uint32_t* present = nullptr;
  uint32_t* vidx = nullptr;
  void* scan_tmp = nullptr;
  size_t scan_bytes = 0;
  unsigned long long* bad = nullptr;
  WG_HIP(hipMallocAsync((void**)&bad, 8, ctx->stream));
  WG_HIP(hipMemsetAsync(bad, 0, 8, ctx->stream));
  WG_HIP(hipMallocAsync((void**)&dtable, table.size() * sizeof(PqPageDev), ctx->stream));
  WG_HIP(hipMemcpyAsync(dtable, table.data(), table.size() * sizeof(PqPageDev), hipMemcpyHostToDevice, ctx->stream));
  if (any_nullable) {
    WG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                            (int)max_group, ctx->stream));
    WG_HIP(hipMallocAsync((void**)&present, max_group * 8 + scan_bytes, ctx->stream));
    vidx = present + max_group;
    scan_tmp = vidx + max_group;
  }
  uint64_t r0 = row0;
  for (size_t g = 0; g < groups.size(); ++g) {
    const size_t first = groups[g].first;
    const size_t last = g + 1 < groups.size() ? groups[g + 1].first : table.size();
    const uint64_t grows = groups[g].second;
    const PqPageDev& tail = table[last - 1];
    const uint64_t grid = tail.tile0 + (tail.p.num_rows + kPqTile - 1) / kPqTile;
    const uint32_t np = (uint32_t)(last - first);
    const PqPageDev* dt = dtable + first;
    if (any_nullable) {
      hipLaunchKernelGGL(pq_levels_kernel, dim3((uint32_t)grid), dim3(256), 0, ctx->stream, dt, np, present);
      WG_HIP(hipGetLastError());
      WG_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, present, vidx, (int)grows, ctx->stream));
    }
    if (type == 1)
      hipLaunchKernelGGL(pq_values_kernel<uint32_t>, dim3((uint32_t)grid), dim3(256), 0, ctx->stream, dt, np,
                         (const uint32_t*)present, (const uint32_t*)vidx, (uint32_t*)out, out_valid, r0, bad);
    else
      hipLaunchKernelGGL(pq_values_kernel<uint64_t>, dim3((uint32_t)grid), dim3(256), 0, ctx->stream, dt, np,
                         (const uint32_t*)present, (const uint32_t*)vidx, (uint64_t*)out, out_valid, r0, bad);
    WG_HIP(hipGetLastError());
    r0 += grows;
  }
  uint64_t nbad = 0;
  WG_HIP(hipMemcpyAsync(&nbad, bad, 8, hipMemcpyDeviceToHost, ctx->stream));
  if (present) WG_HIP(hipFreeAsync(present, ctx->stream));
  WG_HIP(hipFreeAsync(dtable, ctx->stream));
  WG_HIP(hipFreeAsync(bad, ctx->stream));
  WG_HIP(hipStreamSynchronize(ctx->stream));
  WG_CHECK(nbad == 0, WG_ERR_INVALID, "%llu dictionary indices missing or out of range", (unsigned long long)nbad);
  return WG_OK;
}

wg_status wg_parquet_decode_page(wg_ctx* ctx, const wg_pq_page* page, void* out, uint64_t* out_valid, uint64_t row0) {
  clear_error();
  WG_CHECK(page, WG_ERR_INVALID, "null argument");
  return wg_parquet_decode_pages(ctx, page, 1, out, out_valid, row0);
}

}  // extern "C"
