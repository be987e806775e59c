// wg_view.hip — late-materialized (view) columns flattened on the device.
//
// With late materialization the join probe emits view columns instead of
// gathered copies (emitViewOutput, internal/engine/exec/join.go:2517-2600): a
// view is an owned base column, one base row index per row, and the view's
// own NULL bitmap (NewViewVector / NewViewVectorReuse,
// internal/engine/batch/view.go:27-80). The device probe's pair arrays
// already are those indices, so the probe output stays two index columns in
// HBM until a consumer needs values; flattening a view (Vector.Flatten,
// view.go:87-230) is this gather: row i takes base[indices[i]] unless it is
// NULL in the view's own bitmap (then the base is not read) or the base row
// is NULL; NULL rows hold 0. A view over a view composes (view.go:50-70) by
// flattening the inner view's index column, with the inner view's own NULLs,
// through the outer indices. Each wavefront writes whole validity words of
// its 64 rows, so the bitmap needs no atomics.
#include "wg_core.hpp"

namespace wg {

int grid_for(uint64_t n, int block, int max_blocks);

struct Word16 {
  uint64_t lo, hi;
};

// Each wavefront takes kViewU groups of 64 rows per step and issues their
// base reads together (the reads are random when the indices are).
constexpr int kViewU = 4;

template <typename W>
__global__ void view_flatten_kernel(const W* base, const uint64_t* base_valid, uint64_t base_rows, const uint32_t* idx,
                                    const uint64_t* own_valid, uint64_t n, W* out, uint64_t* out_valid,
                                    unsigned long long* bad) {
  const int lane = __lane_id();
  const uint64_t wave0 = (blockIdx.x * (uint64_t)blockDim.x + (threadIdx.x & ~63u)) * kViewU;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * kViewU;
  uint32_t nbad = 0;
  for (uint64_t g0 = wave0; g0 < n; g0 += stride) {
    uint32_t s[kViewU];
    bool take[kViewU], valid[kViewU];
#pragma unroll
    for (int u = 0; u < kViewU; ++u) {
      const uint64_t i = g0 + 64ull * u + lane;
      take[u] = false;
      valid[u] = false;
      s[u] = 0;
      if (i < n && (!own_valid || ((own_valid[i >> 6] >> (i & 63)) & 1ull))) {
        s[u] = idx[i];
        if (s[u] >= base_rows) ++nbad;
        else take[u] = true;
      }
    }
#pragma unroll
    for (int u = 0; u < kViewU; ++u)
      if (take[u]) valid[u] = !base_valid || ((base_valid[s[u] >> 6] >> (s[u] & 63)) & 1ull);
    W x[kViewU];
#pragma unroll
    for (int u = 0; u < kViewU; ++u) {
      x[u] = W{};
      if (valid[u]) x[u] = base[s[u]];
    }
#pragma unroll
    for (int u = 0; u < kViewU; ++u) {
      const uint64_t i = g0 + 64ull * u + lane;
      if (i < n) out[i] = x[u];
      const uint64_t m = __ballot(valid[u]);
      if (out_valid && lane == 0 && g0 + 64ull * u < n) out_valid[(g0 >> 6) + u] = m;
    }
  }
  if (nbad) atomicAdd(bad, (unsigned long long)nbad);
}

}  // namespace wg

using namespace wg;

extern "C" {

wg_status wg_view_flatten(wg_ctx* ctx, const wg_column* base, uint64_t base_rows, const uint32_t* indices,
                          const uint64_t* own_valid, uint64_t n, void* out, uint64_t* out_valid) {
  clear_error();
  WG_CHECK(ctx && base && (n == 0 || (indices && out)), WG_ERR_INVALID, "null argument");
  WG_CHECK(base->data || base_rows == 0, WG_ERR_INVALID, "null base column");
  if (n == 0) return WG_OK;
  unsigned long long* bad = nullptr;
  struct Release {  // the counter goes back to the stream's pool on every exit
    hipStream_t st;
    unsigned long long** p;
    ~Release() {
      if (*p) (void)hipFreeAsync(*p, st);
    }
  } release{ctx->stream, &bad};
  WG_ALLOC_ASYNC(ctx, &bad, 8, ctx->stream, "an error count");
  WG_HIP(hipMemsetAsync(bad, 0, 8, ctx->stream));
  const int grid = grid_for((n + kViewU - 1) / kViewU, 256, 16384);
  const uint64_t* bv = base->validity;
  switch (type_bytes(base->type)) {
    case 2:
      hipLaunchKernelGGL(view_flatten_kernel<uint16_t>, dim3(grid), dim3(256), 0, ctx->stream,
                         (const uint16_t*)base->data, bv, base_rows, indices, own_valid, n, (uint16_t*)out, out_valid,
                         bad);
      break;
    case 4:
      hipLaunchKernelGGL(view_flatten_kernel<uint32_t>, dim3(grid), dim3(256), 0, ctx->stream,
                         (const uint32_t*)base->data, bv, base_rows, indices, own_valid, n, (uint32_t*)out, out_valid,
                         bad);
      break;
    case 16:
      hipLaunchKernelGGL(view_flatten_kernel<Word16>, dim3(grid), dim3(256), 0, ctx->stream, (const Word16*)base->data,
                         bv, base_rows, indices, own_valid, n, (Word16*)out, out_valid, bad);
      break;
    default:
      hipLaunchKernelGGL(view_flatten_kernel<uint64_t>, dim3(grid), dim3(256), 0, ctx->stream,
                         (const uint64_t*)base->data, bv, base_rows, indices, own_valid, n, (uint64_t*)out, out_valid,
                         bad);
      break;
  }
  WG_HIP(hipGetLastError());
  uint64_t nbad = 0;
  WG_HIP(hipMemcpyAsync(&nbad, bad, 8, hipMemcpyDeviceToHost, ctx->stream));
  WG_HIP(hipStreamSynchronize(ctx->stream));
  WG_CHECK(nbad == 0, WG_ERR_INVALID, "%llu view indices past the base column's %llu rows", (unsigned long long)nbad,
           (unsigned long long)base_rows);
  return WG_OK;
}

}  // extern "C"
