// wg_select.hpp — ordered row selection: the positions of sel (or of 0..n)
// whose row satisfies a predicate, in ascending order, as the reference's
// selection vectors are built (filter.go:373-425, join.go:3245-3310).
// Two passes over tiles of kSelTile rows: the predicate once per row with its
// ballots kept and a count per tile, an exclusive scan of the counts, then
// each tile writes its rows at its offset, in order within the tile by the
// kept ballots + prefix counts.
#pragma once
#include <vector>

#include "wg_core.hpp"

namespace wg {

constexpr int kSelBlock = 256;
constexpr int kSelRowsPerThread = 16;
constexpr int kSelTile = kSelBlock * kSelRowsPerThread;

// Pred: a device functor bool(uint64_t row). Pass 1 (WRITE = false)
// evaluates it once per row, keeps each wavefront's ballot (one word per 64
// positions) and the tile's count; pass 2 reads the ballots back instead of
// evaluating the predicate again (a join probe's table lookup is not repeated).
template <bool WRITE, typename Pred>
__global__ void __launch_bounds__(kSelBlock) select_rows_kernel(Pred pred, uint64_t n, const uint32_t* sel,
                                                                uint64_t* masks, uint32_t* tile_counts,
                                                                const uint32_t* tile_offsets, uint32_t* out) {
  __shared__ uint32_t wave_tot[kSelBlock / 64];
  __shared__ uint32_t running;
  const uint64_t tile_beg = (uint64_t)blockIdx.x * kSelTile;
  const int lane = __lane_id();
  const int wave = threadIdx.x >> 6;
  if (!WRITE) {
    uint32_t cnt = 0;
    for (int r = 0; r < kSelRowsPerThread; ++r) {
      const uint64_t i = tile_beg + (uint64_t)r * kSelBlock + threadIdx.x;
      bool pass = false;
      if (i < n) pass = pred(sel ? (uint64_t)sel[i] : i);
      const uint64_t b = __ballot(pass);
      const uint64_t i0 = i - (uint64_t)lane;  // the wavefront's first position, a multiple of 64
      if (lane == 0 && i0 < n) masks[i0 >> 6] = b;
      cnt += (uint32_t)__popcll(b);
    }
    if (lane == 0) wave_tot[wave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t t = 0;
      for (int w = 0; w < kSelBlock / 64; ++w) t += wave_tot[w];
      tile_counts[blockIdx.x] = t;
    }
    return;
  }
  if (threadIdx.x == 0) running = tile_offsets[blockIdx.x];
  __syncthreads();
  // the tile in kSelRowsPerThread rounds of blockDim rows, in order
  for (int r = 0; r < kSelRowsPerThread; ++r) {
    const uint64_t i = tile_beg + (uint64_t)r * kSelBlock + threadIdx.x;
    const uint64_t i0 = i - (uint64_t)lane;
    const uint64_t b = i0 < n ? masks[i0 >> 6] : 0ull;
    const bool pass = (b >> lane) & 1ull;
    const uint32_t rank = (uint32_t)__popcll(b & ((1ull << lane) - 1));
    if (lane == 0) wave_tot[wave] = (uint32_t)__popcll(b);
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int w = 0; w < kSelBlock / 64; ++w) {
      if (w < wave) before += wave_tot[w];
      total += wave_tot[w];
    }
    if (pass) out[running + before + rank] = sel ? sel[i] : (uint32_t)i;
    __syncthreads();
    if (threadIdx.x == 0) running += total;
    __syncthreads();
  }
}

// Host side of the two passes on ctx's stream; *n_out on return.
template <typename Pred>
wg_status select_rows(wg_ctx* ctx, const Pred& pred, uint64_t n, const uint32_t* sel, uint32_t* out_sel,
                      uint64_t* n_out) {
  *n_out = 0;
  if (n == 0) return WG_OK;
  WG_CHECK(n < (1ull << 32), WG_ERR_INVALID, "selection vectors index < 2^32 rows");
  const uint64_t tiles = (n + kSelTile - 1) / kSelTile;
  const uint64_t words = (n + 63) / 64;
  uint32_t* counts = nullptr;  // tile counts, tile offsets, then the ballot words
  struct Release {  // the scratch goes back to the stream's pool on every exit
    hipStream_t st;
    uint32_t** p;
    ~Release() {
      if (*p) (void)hipFreeAsync(*p, st);
    }
  } release{ctx->stream, &counts};
  WG_ALLOC_ASYNC(ctx, &counts, tiles * 8 + words * 8 + 8, ctx->stream, "selection scratch");
  uint32_t* offsets = counts + tiles;
  uint64_t* masks = (uint64_t*)(((uintptr_t)(offsets + tiles) + 7) & ~(uintptr_t)7);
  hipLaunchKernelGGL((select_rows_kernel<false, Pred>), dim3(tiles), dim3(kSelBlock), 0, ctx->stream, pred, n, sel,
                     masks, counts, (const uint32_t*)nullptr, (uint32_t*)nullptr);
  WG_HIP(hipGetLastError());
  std::vector<uint32_t> h(tiles);
  WG_HIP(hipMemcpyAsync(h.data(), counts, tiles * 4, hipMemcpyDeviceToHost, ctx->stream));
  WG_HIP(hipStreamSynchronize(ctx->stream));
  uint64_t acc = 0;
  for (uint64_t t = 0; t < tiles; ++t) {
    const uint32_t v = h[t];
    h[t] = (uint32_t)acc;
    acc += v;
  }
  WG_HIP(hipMemcpyAsync(offsets, h.data(), tiles * 4, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL((select_rows_kernel<true, Pred>), dim3(tiles), dim3(kSelBlock), 0, ctx->stream, pred, n, sel,
                     masks, (uint32_t*)nullptr, (const uint32_t*)offsets, out_sel);
  WG_HIP(hipGetLastError());
  WG_HIP(hipFreeAsync(counts, ctx->stream));
  counts = nullptr;
  WG_HIP(hipStreamSynchronize(ctx->stream));
  *n_out = acc;
  return WG_OK;
}

}  // namespace wg
