// wg_ctx.hip — context, errors, device memory helpers and the selection-vector
// filter (KernelFilter.Execute, internal/engine/exec/filter.go:373-425).
#include <cstdarg>
#include <vector>

#include "wg_core.hpp"

namespace wg {

static thread_local char g_err[1024] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
void clear_error() { g_err[0] = 0; }

// --------------------------------------------------------------- the filter

// Order-preserving stream compaction of the rows that pass `col <op> const`.
// Each workgroup handles a contiguous tile: per wave, ballot + mbcnt give a
// lane's rank among passing lanes; a block scan over the wave totals gives
// the wave's offset; one atomic per tile would break ordering, so tiles take
// their output offset from a decoupled pass: pass 1 counts per tile, the host
// scans the (few) tile counts, pass 2 writes. Output stays ascending like the
// reference's Sel.
constexpr int kFilterBlock = 256;
constexpr int kFilterRowsPerThread = 16;
constexpr int kFilterTile = kFilterBlock * kFilterRowsPerThread;

__device__ __forceinline__ bool filter_pass(const ColDesc& c, uint64_t row, int op, int64_t ci, double cf) {
  return pred_pass(c, row, op, ci, cf);
}

// Grid-stride streaming copy of 16-byte words.
__global__ void __launch_bounds__(256) copy_words_kernel(ulonglong2* dst, const ulonglong2* src, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

template <bool WRITE>
__global__ void __launch_bounds__(kFilterBlock) filter_kernel(ColDesc col, uint64_t n, const uint32_t* sel, int op,
                                                              int64_t ci, double cf, uint32_t* tile_counts,
                                                              const uint32_t* tile_offsets, uint32_t* out) {
  __shared__ uint32_t wave_tot[kFilterBlock / 64];
  __shared__ uint32_t running;
  const uint64_t tile_beg = (uint64_t)blockIdx.x * kFilterTile;
  const int lane = __lane_id();
  const int wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) running = WRITE ? tile_offsets[blockIdx.x] : 0;
  __syncthreads();
  // process the tile in kFilterRowsPerThread rounds of blockDim rows, in order
  for (int r = 0; r < kFilterRowsPerThread; ++r) {
    const uint64_t i = tile_beg + (uint64_t)r * kFilterBlock + threadIdx.x;
    bool pass = false;
    uint64_t row = 0;
    if (i < n) {
      row = sel ? (uint64_t)sel[i] : i;
      pass = filter_pass(col, row, op, ci, cf);
    }
    const uint64_t b = __ballot(pass);
    const uint32_t rank = (uint32_t)__popcll(b & ((1ull << lane) - 1));
    if (lane == 0) wave_tot[wave] = (uint32_t)__popcll(b);
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int w = 0; w < kFilterBlock / 64; ++w) {
      if (w < wave) before += wave_tot[w];
      total += wave_tot[w];
    }
    if (WRITE && pass) out[running + before + rank] = (uint32_t)row;
    __syncthreads();
    if (threadIdx.x == 0) running += total;
    __syncthreads();
  }
  if (!WRITE && threadIdx.x == 0) tile_counts[blockIdx.x] = running;
}

}  // namespace wg

using namespace wg;

extern "C" {

const char* wg_last_error(void) { return g_err; }
const char* wg_version(void) { return "wadjet-gpu 0.1 (gfx950)"; }

wg_status wg_ctx_create(int device, wg_ctx** out) {
  clear_error();
  WG_CHECK(out != nullptr, WG_ERR_INVALID, "out is null");
  int n = 0;
  WG_HIP(hipGetDeviceCount(&n));
  WG_CHECK(device >= 0 && device < n, WG_ERR_INVALID, "device %d of %d", device, n);
  WG_HIP(hipSetDevice(device));
  wg_ctx* c = new wg_ctx();
  c->device = device;
  hipDeviceProp_t p;
  WG_HIP(hipGetDeviceProperties(&p, device));
  c->cus = p.multiProcessorCount;
  c->hbm = p.totalGlobalMem;
  c->lds_per_block = (int)p.sharedMemPerBlock;
  WG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  WG_HIP(hipEventCreate(&c->t0));
  WG_HIP(hipEventCreate(&c->t1));
  WG_HIP(hipHostMalloc((void**)&c->pinned, 4096, hipHostMallocDefault));
  *out = c;
  return WG_OK;
}

wg_status wg_ctx_destroy(wg_ctx* ctx) {
  if (!ctx) return WG_OK;
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipEventDestroy(ctx->t0);
  (void)hipEventDestroy(ctx->t1);
  (void)hipHostFree(ctx->pinned);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return WG_OK;
}

wg_status wg_ctx_sync(wg_ctx* ctx) {
  WG_CHECK(ctx, WG_ERR_INVALID, "ctx is null");
  WG_HIP(hipStreamSynchronize(ctx->stream));
  return WG_OK;
}

wg_status wg_ctx_info(wg_ctx* ctx, int32_t* cus, uint64_t* hbm_bytes, int32_t* lds_bytes) {
  WG_CHECK(ctx, WG_ERR_INVALID, "ctx is null");
  if (cus) *cus = ctx->cus;
  if (hbm_bytes) *hbm_bytes = ctx->hbm;
  if (lds_bytes) *lds_bytes = ctx->lds_per_block;
  return WG_OK;
}

wg_status wg_ctx_timer_start(wg_ctx* ctx) {
  WG_CHECK(ctx, WG_ERR_INVALID, "ctx is null");
  WG_HIP(hipEventRecord(ctx->t0, ctx->stream));
  return WG_OK;
}

wg_status wg_ctx_timer_stop(wg_ctx* ctx, float* ms) {
  WG_CHECK(ctx && ms, WG_ERR_INVALID, "null argument");
  WG_HIP(hipEventRecord(ctx->t1, ctx->stream));
  WG_HIP(hipEventSynchronize(ctx->t1));
  WG_HIP(hipEventElapsedTime(ms, ctx->t0, ctx->t1));
  return WG_OK;
}

wg_status wg_malloc(wg_ctx* ctx, uint64_t bytes, void** out) {
  WG_CHECK(ctx && out, WG_ERR_INVALID, "null argument");
  *out = nullptr;
  if (bytes == 0) return WG_OK;
  if (hipMalloc(out, bytes) != hipSuccess) {
    (void)hipGetLastError();
    set_error("hipMalloc(%llu) failed", (unsigned long long)bytes);
    return WG_ERR_OOM;
  }
  return WG_OK;
}

wg_status wg_free(wg_ctx* ctx, void* ptr) {
  WG_CHECK(ctx, WG_ERR_INVALID, "ctx is null");
  if (ptr) {
    WG_HIP(hipStreamSynchronize(ctx->stream));
    WG_HIP(hipFree(ptr));
  }
  return WG_OK;
}

wg_status wg_memcpy_h2d(wg_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  WG_CHECK(ctx, WG_ERR_INVALID, "ctx is null");
  if (bytes == 0) return WG_OK;
  WG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  WG_HIP(hipStreamSynchronize(ctx->stream));
  return WG_OK;
}

wg_status wg_memcpy_d2h(wg_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  WG_CHECK(ctx, WG_ERR_INVALID, "ctx is null");
  if (bytes == 0) return WG_OK;
  WG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  WG_HIP(hipStreamSynchronize(ctx->stream));
  return WG_OK;
}

wg_status wg_memcpy_d2d(wg_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  WG_CHECK(ctx, WG_ERR_INVALID, "ctx is null");
  WG_CHECK(((uintptr_t)dst | (uintptr_t)src | bytes) % 16 == 0, WG_ERR_INVALID,
           "wg_memcpy_d2d needs 16-byte aligned pointers and size");
  if (bytes == 0) return WG_OK;
  const uint64_t n = bytes / 16;
  const int grid = (int)std::min<uint64_t>((n + 255) / 256, (uint64_t)ctx->cus * 16);
  hipLaunchKernelGGL(copy_words_kernel, dim3(grid), dim3(256), 0, ctx->stream, (ulonglong2*)dst,
                     (const ulonglong2*)src, n);
  WG_HIP(hipGetLastError());
  return WG_OK;
}

wg_status wg_memset(wg_ctx* ctx, void* dst, int value, uint64_t bytes) {
  WG_CHECK(ctx, WG_ERR_INVALID, "ctx is null");
  if (bytes == 0) return WG_OK;
  WG_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
  return WG_OK;
}

wg_status wg_filter(wg_ctx* ctx, const wg_column* col, uint64_t n_rows, const uint32_t* sel, uint64_t n_sel,
                    int32_t op, int64_t i64, double f64, uint32_t* out_sel, uint64_t* n_out) {
  clear_error();
  WG_CHECK(ctx && col && out_sel && n_out, WG_ERR_INVALID, "null argument");
  WG_CHECK(op >= WG_EQ && op <= WG_GE, WG_ERR_INVALID, "bad compare op %d", op);
  WG_CHECK(col->type == WG_INT32 || col->type == WG_INT64 || col->type == WG_FLOAT64 || col->type == WG_UINT16 ||
               col->type == WG_UINT32 || col->type == WG_IPV4,
           WG_ERR_UNSUPPORTED, "no filter kernel for type %d", col->type);
  const uint64_t n = sel ? n_sel : n_rows;
  *n_out = 0;
  if (n == 0) return WG_OK;
  WG_CHECK(n < (1ull << 32), WG_ERR_INVALID, "selection vectors index < 2^32 rows");
  ColDesc c{col->data, col->validity, col->type, col->scale};
  const uint64_t tiles = (n + kFilterTile - 1) / kFilterTile;
  uint32_t* counts = nullptr;
  WG_HIP(hipMallocAsync((void**)&counts, tiles * 8, ctx->stream));
  uint32_t* offsets = counts + tiles;
  hipLaunchKernelGGL(filter_kernel<false>, dim3(tiles), dim3(kFilterBlock), 0, ctx->stream, c, n, sel, op, i64, f64,
                     counts, (const uint32_t*)nullptr, (uint32_t*)nullptr);
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
  hipLaunchKernelGGL(filter_kernel<true>, dim3(tiles), dim3(kFilterBlock), 0, ctx->stream, c, n, sel, op, i64, f64,
                     (uint32_t*)nullptr, (const uint32_t*)offsets, out_sel);
  WG_HIP(hipGetLastError());
  WG_HIP(hipFreeAsync(counts, ctx->stream));
  WG_HIP(hipStreamSynchronize(ctx->stream));
  *n_out = acc;
  return WG_OK;
}

}  // extern "C"
