// wg_shuffle.hip — routing of a partitioned shuffle on the device.
//
// partitionedShuffleSink.Consume (internal/worker/partitioned_shuffle_sink.go:
// 199-290) assigns every active row of a batch to one of numParts partitions
// by an FNV-1a 64 hash of its key columns (hashRowsIntoPartitions, :1039-1171)
// and then appends each partition's rows in input order (perPartRows,
// :272-290). Here both steps run as two passes over per-wavefront row ranges:
//
//   count    each wavefront hashes its contiguous range of active rows, stores
//            the partition of every row and counts rows per partition in a
//            wavefront-private LDS histogram (no cross-wave ordering needed);
//   scan     exclusive sum over the (partition, wavefront) counts, partition-
//            major, so partition p's rows from wavefront w start at
//            offs[p * W + w] and the partitions come out contiguous;
//   scatter  each wavefront walks its range again 64 rows at a time; a row's
//            place among the step's rows of its partition comes from one
//            ballot per bit of the partition id (lanes in order), and
//            the wavefront's LDS cursor per partition advances by the step's
//            count, so every partition keeps input order — the stable
//            grouping the reference builds with appends. Up to 1024
//            partitions a workgroup first counting-sorts tiles of 8192 rows
//            in LDS and writes each partition's run of the tile together.
//
// Bytes per active row: key bytes read twice (count, then the stored 4-byte
// partition instead of the keys) + 4 B partition written + 4 B row id
// written (+ 4 B selection read per pass).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "wg_core.hpp"

namespace wg {

constexpr uint64_t kFnvOffset = 14695981039346656037ull;  // partitioned_shuffle_sink.go:1023
constexpr uint64_t kFnvPrime = 1099511628211ull;          // partitioned_shuffle_sink.go:1024
constexpr uint32_t kMaxShuffleParts = 16384;              // wavefront histograms of <= 64 KiB per workgroup
constexpr uint32_t kNoPart = 0xFFFFFFFFu;

struct ShufParams {
  ColDesc key[kMaxCols];
  int32_t n_keys;
  uint32_t np;
  uint32_t np_mask;    // np - 1 when np is a power of two, else 0
  uint32_t two32_mod;  // 2^32 mod np
  int32_t pbits;       // bits of a partition id: 2^pbits >= np
  uint32_t W;          // wavefronts of the grid
  const uint32_t* sel;
  uint64_t n;
  uint64_t rows_per_wave;  // multiple of 64
  uint32_t* part;
  uint32_t* counts;
  const uint32_t* offs;
  uint32_t* rows;
};

// FNV-1a over the W low bytes of v, least significant first (the bytes of
// binary.LittleEndian.PutUintN).
template <int W>
__device__ __forceinline__ uint64_t fnv_bytes(uint64_t h, uint64_t v) {
#pragma unroll
  for (int b = 0; b < W; ++b) h = (h ^ ((v >> (8 * b)) & 0xFFull)) * kFnvPrime;
  return h;
}

// The hashes of U rows (hashRowsIntoPartitions' type switch, :1057-1162),
// key column by key column so that the U rows' loads of a column are in
// flight together: each key column mixes its storage bytes in the
// reference's vectors — Int32/Port the 4 bytes of Int32Data, Int64/IPv4 the
// 8 bytes of Int64Data (a compact uint32 IPv4 zero-extended), Float64 its 8
// bits — or one 0xff byte when NULL; a decimal column takes the default case,
// one 0x00 byte, NULL or not.
template <int U>
__device__ __forceinline__ void shuffle_row_hashes(const ShufParams& P, const uint64_t (&row)[U], uint64_t (&h)[U]) {
#pragma unroll
  for (int u = 0; u < U; ++u) h[u] = kFnvOffset;
  for (int k = 0; k < P.n_keys; ++k) {
    const ColDesc c = P.key[k];
    if (c.type == WG_DECIMAL || c.type == WG_DECIMAL128) {
#pragma unroll
      for (int u = 0; u < U; ++u) h[u] = (h[u] ^ 0x00ull) * kFnvPrime;
      continue;
    }
    uint64_t v[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) ok[u] = col_valid(c, row[u]);
    const bool narrow = c.type == WG_INT32 || c.type == WG_UINT16;
    switch (c.type) {
      case WG_INT32:
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = (uint32_t)((const int32_t*)c.data)[row[u]];
        break;
      case WG_UINT16:
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = ((const uint16_t*)c.data)[row[u]];
        break;
      case WG_UINT32:
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = ((const uint32_t*)c.data)[row[u]];
        break;
      default:
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = ((const uint64_t*)c.data)[row[u]];
        break;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!ok[u]) h[u] = (h[u] ^ 0xFFull) * kFnvPrime;
      else h[u] = narrow ? fnv_bytes<4>(h[u], v[u]) : fnv_bytes<8>(h[u], v[u]);
    }
  }
}

// h % np with 32-bit remainders only: h = hi * 2^32 + lo, so
// h mod np = ((hi mod np) * (2^32 mod np) + lo mod np) mod np, every term
// below 2^31 for np <= 2^15.
__device__ __forceinline__ uint32_t part_of(const ShufParams& P, uint64_t h) {
  if (P.np_mask || P.np == 1) return (uint32_t)h & P.np_mask;
  const uint32_t np = P.np;
  const uint32_t r_hi = (uint32_t)(h >> 32) % np;
  const uint32_t r_lo = (uint32_t)h % np;
  return (r_hi * P.two32_mod + r_lo) % np;
}

constexpr int kShufU = 4;  // 64-row steps whose loads are issued together

__global__ void __launch_bounds__(256) shuffle_count_kernel(ShufParams P) {
  extern __shared__ uint32_t hist_all[];
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  uint32_t* hist = hist_all + (uint64_t)wave * P.np;
  for (uint32_t p = lane; p < P.np; p += 64) hist[p] = 0;
  const uint32_t gw = blockIdx.x * (blockDim.x >> 6) + wave;
  const uint64_t beg = (uint64_t)gw * P.rows_per_wave;
  const uint64_t end = min(beg + P.rows_per_wave, P.n);
  __builtin_amdgcn_wave_barrier();
  for (uint64_t base = beg; base < end; base += 64 * kShufU) {
    uint64_t row[kShufU], h[kShufU];
    bool in[kShufU];
#pragma unroll
    for (int u = 0; u < kShufU; ++u) {
      const uint64_t i = base + 64 * u + lane;
      in[u] = i < end;
      const uint64_t ii = in[u] ? i : beg;
      row[u] = P.sel ? (uint64_t)P.sel[ii] : ii;
    }
    shuffle_row_hashes<kShufU>(P, row, h);
#pragma unroll
    for (int u = 0; u < kShufU; ++u) {
      if (!in[u]) continue;
      const uint32_t p = P.n_keys ? part_of(P, h[u]) : 0;
      P.part[base + 64 * u + lane] = p;
      atomicAdd(&hist[p], 1u);
    }
  }
  __builtin_amdgcn_wave_barrier();
  if (gw < P.W)
    for (uint32_t p = lane; p < P.np; p += 64) P.counts[(uint64_t)p * P.W + gw] = hist[p];
}

__global__ void __launch_bounds__(256) shuffle_scatter_kernel(ShufParams P) {
  extern __shared__ uint32_t cur_all[];
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  uint32_t* cur = cur_all + (uint64_t)wave * P.np;
  const uint32_t gw = blockIdx.x * (blockDim.x >> 6) + wave;
  if (gw >= P.W) return;  // whole wavefronts only: no workgroup barrier follows
  for (uint32_t p = lane; p < P.np; p += 64) cur[p] = P.offs[(uint64_t)p * P.W + gw];
  const uint64_t beg = (uint64_t)gw * P.rows_per_wave;
  const uint64_t end = min(beg + P.rows_per_wave, P.n);
  const uint64_t lt = (1ull << lane) - 1;
  __builtin_amdgcn_wave_barrier();
  for (uint64_t base = beg; base < end; base += 64 * kShufU) {
    uint32_t part[kShufU], src[kShufU];
#pragma unroll
    for (int u = 0; u < kShufU; ++u) {
      const uint64_t i = base + 64 * u + lane;
      const bool in = i < end;
      part[u] = in ? P.part[i] : kNoPart;
      src[u] = in ? (P.sel ? P.sel[i] : (uint32_t)i) : 0;
    }
    // the steps in row order; within a step, the lanes holding the same
    // partition as this lane: one ballot per partition-id bit narrows the
    // wave to them (lanes whose id agrees on every bit), so a row's place
    // among the step's rows of its partition is the count of such lanes
    // below it
#pragma unroll
    for (int u = 0; u < kShufU; ++u) {
      const uint32_t p = part[u];
      const bool in = p != kNoPart;
      uint64_t same = __ballot(in);
      for (int b = 0; b < P.pbits; ++b) {
        const bool bit = (p >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        same &= bit ? bal : ~bal;
      }
      const uint32_t rank = (uint32_t)__popcll(same & lt);
      const uint32_t cnt = (uint32_t)__popcll(same);
      const bool last = rank + 1 == cnt;
      // every lane reads its cursor before the last lane of each group moves it
      const uint32_t b = in ? cur[p] : 0;
      __builtin_amdgcn_wave_barrier();
      if (in && last) cur[p] = b + cnt;
      if (in) P.rows[b + rank] = src[u];
      __builtin_amdgcn_wave_barrier();
    }
  }
}

// The same scatter for up to kTileMaxParts partitions, by workgroup tiles:
// a tile of kTileRows consecutive rows is ranked (each wavefront a quarter of
// it, with the ballot-per-bit ranks and a wavefront-private running count),
// counting-sorted by partition in LDS, and written out partition run by
// partition run, so the stores of a run are consecutive addresses (a few
// dozen rows per partition per tile) instead of one 4-byte store per row per
// partition front. The workgroup's range is its four wavefronts' ranges of
// the count pass, so its cursor in partition p starts at offs[p * W + 4 b].
constexpr uint32_t kTileRows = 8192;
constexpr uint32_t kTileMaxParts = 1024;
constexpr int kTileBlock = 256;
constexpr int kTileQuarter = kTileRows / 4;       // rows of one wavefront per tile
constexpr int kTileSteps = kTileQuarter / 64;     // 64-row steps of one wavefront per tile

__host__ __device__ inline size_t tile_lds_bytes(uint32_t np) { return (size_t)24 * np + 4 * kTileRows; }

// exclusive prefix sum of one value per thread over a 256-thread workgroup
__device__ __forceinline__ uint32_t block_exclusive_sum(uint32_t v, uint32_t* tmp) {
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  if (lane == 63) tmp[wave] = x;
  __syncthreads();
  uint32_t off = 0;
  for (int w = 0; w < wave; ++w) off += tmp[w];
  __syncthreads();
  return off + x - v;
}

__global__ void __launch_bounds__(kTileBlock) shuffle_scatter_tile_kernel(ShufParams P) {
  extern __shared__ uint32_t lds_t[];
  __shared__ uint32_t scan_tmp[kTileBlock / 64];
  const uint32_t np = P.np;
  uint32_t* cntw = lds_t;         // [4][np] running counts, then wavefront prefixes
  uint32_t* tstart = cntw + 4 * np;  // [np] where the tile's run of p starts
  uint32_t* cur = tstart + np;       // [np] this workgroup's cursor in p
  uint16_t* st_off = (uint16_t*)(cur + np);  // [kTileRows] row offset in the tile, partition-sorted
  uint16_t* st_p = st_off + kTileRows;       // [kTileRows] its partition
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  const uint64_t lt = (1ull << lane) - 1;
  for (uint32_t p = threadIdx.x; p < np; p += kTileBlock) cur[p] = P.offs[(uint64_t)p * P.W + 4ull * blockIdx.x];
  for (uint32_t i = threadIdx.x; i < 4 * np; i += kTileBlock) cntw[i] = 0;
  const uint64_t beg = (uint64_t)blockIdx.x * 4 * P.rows_per_wave;
  const uint64_t end = min(beg + 4 * P.rows_per_wave, P.n);
  uint32_t* myc = cntw + (uint64_t)wave * np;
  __syncthreads();
  for (uint64_t tb = beg; tb < end; tb += kTileRows) {
    const uint32_t trows = (uint32_t)min((uint64_t)kTileRows, end - tb);
    // a. rank this wavefront's quarter: (partition << 16) | rank in the quarter
    uint32_t pr[kTileSteps];
#pragma unroll
    for (int s = 0; s < kTileSteps; ++s) {
      const uint32_t off = wave * kTileQuarter + s * 64 + lane;
      pr[s] = off < trows ? P.part[tb + off] : kNoPart;
    }
#pragma unroll
    for (int s = 0; s < kTileSteps; ++s) {
      const uint32_t p = pr[s];
      const bool in = p != kNoPart;
      uint64_t same = __ballot(in);
      for (int b = 0; b < P.pbits; ++b) {
        const bool bit = (p >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        same &= bit ? bal : ~bal;
      }
      const uint32_t rank = (uint32_t)__popcll(same & lt);
      const uint32_t cnt = (uint32_t)__popcll(same);
      const uint32_t base = in ? myc[p] : 0;
      __builtin_amdgcn_wave_barrier();
      if (in && rank + 1 == cnt) myc[p] = base + cnt;
      __builtin_amdgcn_wave_barrier();
      pr[s] = in ? (p << 16) | (base + rank) : kNoPart;
    }
    __syncthreads();
    // b. per partition: wavefront prefixes, the tile's count and run start
    uint32_t tot[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t p = 4 * threadIdx.x + q;
      tot[q] = 0;
      if (p < np) {
        const uint32_t c0 = cntw[p], c1 = cntw[np + p], c2 = cntw[2 * np + p], c3 = cntw[3 * np + p];
        cntw[p] = 0;
        cntw[np + p] = c0;
        cntw[2 * np + p] = c0 + c1;
        cntw[3 * np + p] = c0 + c1 + c2;
        tot[q] = c0 + c1 + c2 + c3;
      }
    }
    const uint32_t ex = block_exclusive_sum(tot[0] + tot[1] + tot[2] + tot[3], scan_tmp);
    {
      uint32_t run = ex;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t p = 4 * threadIdx.x + q;
        if (p < np) tstart[p] = run;
        run += tot[q];
      }
    }
    __syncthreads();
    // c. counting-sort the tile's rows into LDS
#pragma unroll
    for (int s = 0; s < kTileSteps; ++s) {
      if (pr[s] == kNoPart) continue;
      const uint32_t p = pr[s] >> 16, r = pr[s] & 0xFFFFu;
      const uint32_t pos = tstart[p] + myc[p] + r;
      st_off[pos] = (uint16_t)(wave * kTileQuarter + s * 64 + lane);
      st_p[pos] = (uint16_t)p;
    }
    __syncthreads();
    // d. write the runs: consecutive positions of a partition are
    // consecutive rows of its output
    for (uint32_t j = threadIdx.x; j < trows; j += kTileBlock) {
      const uint32_t p = st_p[j];
      const uint64_t i = tb + st_off[j];
      P.rows[cur[p] + (j - tstart[p])] = P.sel ? P.sel[i] : (uint32_t)i;
    }
    __syncthreads();
    // e. advance the cursors, clear the running counts
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t p = 4 * threadIdx.x + q;
      if (p < np) {
        cur[p] += tot[q];
        cntw[p] = cntw[np + p] = cntw[2 * np + p] = cntw[3 * np + p] = 0;
      }
    }
    __syncthreads();
  }
}

// offsets[p] = offs[p * W] (the first wavefront's start in partition p),
// offsets[np] = total rows.
__global__ void shuffle_offsets_kernel(const uint32_t* offs, uint32_t np, uint32_t W, uint64_t total,
                                       uint64_t* out) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p <= np; p += gridDim.x * blockDim.x)
    out[p] = p < np ? (uint64_t)offs[(uint64_t)p * W] : total;
}

}  // namespace wg

using namespace wg;

extern "C" {

wg_status wg_shuffle_partition(wg_ctx* ctx, const wg_column* cols, int32_t n_cols, const int32_t* key_cols,
                               int32_t n_keys, uint64_t n_rows, const uint32_t* sel, uint64_t n_sel,
                               int32_t num_parts, uint32_t* part, uint32_t* rows, uint64_t* offsets) {
  clear_error();
  WG_CHECK(ctx && rows && offsets, WG_ERR_INVALID, "null argument");
  WG_CHECK(n_keys >= 0 && n_keys <= kMaxCols && (n_keys == 0 || (cols && key_cols)), WG_ERR_INVALID,
           "bad key columns (%d)", n_keys);
  WG_CHECK(num_parts >= 0, WG_ERR_INVALID, "negative partition count %d", num_parts);
  const uint32_t np = num_parts == 0 ? 1u : (uint32_t)num_parts;  // mod 0 routes everything to 0 (:1163-1166)
  WG_CHECK(np <= kMaxShuffleParts, WG_ERR_UNSUPPORTED, "%u partitions: the GPU routing takes at most %u", np,
           kMaxShuffleParts);
  const uint64_t n = sel ? n_sel : n_rows;
  WG_CHECK(n < (1ull << 32) && n_rows < (1ull << 32), WG_ERR_INVALID, "row ids are 32-bit");
  ShufParams P{};
  for (int k = 0; k < n_keys; ++k) {
    const int c = key_cols[k];
    WG_CHECK(c >= 0 && c < n_cols, WG_ERR_INVALID, "key column %d out of range", c);
    const wg_column& col = cols[c];
    WG_CHECK(col.type >= WG_INT32 && col.type <= WG_UINT16, WG_ERR_INVALID, "bad column type %d", col.type);
    WG_CHECK(col.data || col.type == WG_DECIMAL || col.type == WG_DECIMAL128 || n == 0, WG_ERR_INVALID,
             "key column %d has no data", c);
    P.key[k] = ColDesc{col.data, col.validity, col.type, col.scale};
  }
  P.n_keys = n_keys;
  P.np = np;
  P.np_mask = (np & (np - 1)) == 0 ? np - 1 : 0;
  P.two32_mod = (uint32_t)((1ull << 32) % np);
  while ((1u << P.pbits) < np) ++P.pbits;
  P.sel = sel;
  P.n = n;
  if (n == 0) {
    for (uint32_t p = 0; p <= np; ++p) offsets[p] = 0;
    return WG_OK;
  }
  // wavefronts: 32 per CU for large batches, >= 1024 rows each, and a
  // (partition x wavefront) count table of at most 2^25 entries
  const int wpb = np <= 4096 ? 4 : np <= 8192 ? 2 : 1;  // wavefronts per workgroup (LDS <= 64 KiB)
  uint64_t W = std::min<uint64_t>((uint64_t)ctx->cus * 32, (n + 1023) / 1024);
  W = std::min<uint64_t>(W, std::max<uint64_t>(1, (1ull << 25) / np));
  W = std::max<uint64_t>(W, 1);
  const uint64_t blocks = (W + wpb - 1) / wpb;
  W = blocks * wpb;
  P.W = (uint32_t)W;
  P.rows_per_wave = ((n + W - 1) / W + 63) / 64 * 64;
  const size_t lds = (size_t)wpb * np * 4;

  const uint64_t ncnt = (uint64_t)np * W;
  size_t scan_bytes = 0;
  WG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)ncnt,
                                          ctx->stream));
  const uint64_t part_bytes = part ? 0 : (n * 4 + 255) / 256 * 256;
  const uint64_t cnt_bytes = (ncnt * 4 + 255) / 256 * 256;
  const uint64_t off_bytes = ((uint64_t)(np + 1) * 8 + 255) / 256 * 256;
  char* scratch = nullptr;
  WG_ALLOC_ASYNC(ctx, &scratch, part_bytes + 2 * cnt_bytes + off_bytes + scan_bytes, ctx->stream, "shuffle scratch");
  P.part = part ? part : (uint32_t*)scratch;
  P.counts = (uint32_t*)(scratch + part_bytes);
  uint32_t* offs = (uint32_t*)(scratch + part_bytes + cnt_bytes);
  uint64_t* doff = (uint64_t*)(scratch + part_bytes + 2 * cnt_bytes);
  void* scan_tmp = scratch + part_bytes + 2 * cnt_bytes + off_bytes;
  P.offs = offs;
  P.rows = rows;
  hipLaunchKernelGGL(shuffle_count_kernel, dim3(blocks), dim3(64 * wpb), lds, ctx->stream, P);
  WG_HIP(hipGetLastError());
  WG_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, P.counts, offs, (int)ncnt, ctx->stream));
  if (np <= kTileMaxParts && wpb == 4) {
    hipLaunchKernelGGL(shuffle_scatter_tile_kernel, dim3(blocks), dim3(kTileBlock), tile_lds_bytes(np), ctx->stream,
                       P);
  } else {
    hipLaunchKernelGGL(shuffle_scatter_kernel, dim3(blocks), dim3(64 * wpb), lds, ctx->stream, P);
  }
  WG_HIP(hipGetLastError());
  hipLaunchKernelGGL(shuffle_offsets_kernel, dim3((np + 256) / 256), dim3(256), 0, ctx->stream, offs, np,
                     (uint32_t)W, n, doff);
  WG_HIP(hipGetLastError());
  WG_HIP(hipMemcpyAsync(offsets, doff, (uint64_t)(np + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
  WG_HIP(hipFreeAsync(scratch, ctx->stream));
  WG_HIP(hipStreamSynchronize(ctx->stream));
  return WG_OK;
}

}  // extern "C"
