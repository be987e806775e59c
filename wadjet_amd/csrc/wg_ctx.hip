// wg_ctx.hip — context, errors, device memory helpers and the selection-vector
// filter (KernelFilter.Execute, internal/engine/exec/filter.go:373-425).
#include <cstdarg>
#include <vector>

#include "wg_core.hpp"
#include "wg_select.hpp"

namespace wg {

static thread_local char g_err[1024] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
void clear_error() { g_err[0] = 0; }

wg_status dev_alloc(wg_ctx* ctx, void** p, size_t bytes, hipStream_t stream, const char* what) {
  *p = nullptr;
  if (ctx) {
    bool inject = ctx->failing.load();
    if (!inject) {
      const int64_t left = ctx->fail_at.load();
      if (left > 0 && ctx->fail_at.fetch_sub(1) == 1) {
        inject = true;
        if (ctx->fail_sticky.load()) ctx->failing.store(true);
      }
    }
    if (inject) {
      set_error("device allocation of %llu bytes for %s failed: out of memory (injected)", (unsigned long long)bytes,
                what);
      return WG_ERR_OOM;
    }
  }
  const hipError_t e = stream ? hipMallocAsync(p, bytes, stream) : hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    set_error("device allocation of %llu bytes for %s failed: %s", (unsigned long long)bytes, what,
              hipGetErrorString(e));
    return WG_ERR_OOM;
  }
  return WG_OK;
}

// --------------------------------------------------------------- the filter

// Order-preserving stream compaction of the rows that pass `col <op> const`.
// Each workgroup handles a contiguous tile: per wave, ballot + mbcnt give a
// lane's rank among passing lanes; a block scan over the wave totals gives
// the wave's offset; one atomic per tile would break ordering, so tiles take
// their output offset from a decoupled pass: pass 1 counts per tile, the host
// scans the (few) tile counts, pass 2 writes. Output stays ascending like the
// reference's Sel.
// Predicate of the filter kernel: the typed comparison of KernelFilter.
struct FilterPred {
  ColDesc col;
  int op;
  int64_t ci;
  double cf;
  __device__ bool operator()(uint64_t row) const { return pred_pass(col, row, op, ci, cf); }
};

// ------------------------------------------- the filter without allocation

// Scratch of wg_filter_noalloc, one copy per device for the life of the
// process: per kSelTile-row tile a count and an offset (u32), one ballot word
// per 64 rows, and the selected-row total. Host code never names this array
// in a device context: it uses the address hipGetSymbolAddress returns, taken
// once per context at wg_ctx_create with the context's device current.
constexpr uint64_t kNoAllocMaxRows = 1ull << 26;
constexpr uint64_t kNoAllocTiles = kNoAllocMaxRows / kSelTile;
constexpr uint64_t kNoAllocWords = kNoAllocTiles + kNoAllocMaxRows / 64 + 1;
__device__ __align__(256) uint64_t scratch[kNoAllocWords];

// Exclusive scan of the tile counts in one workgroup (a tile's offset and the
// total stay on the device): each thread sums a run of consecutive tiles, a
// scan over the 1024 run sums gives each run's start.
__global__ void __launch_bounds__(1024) tile_scan_kernel(const uint32_t* counts, uint32_t* offsets, uint64_t tiles,
                                                         uint64_t* total) {
  __shared__ uint64_t part[1024];
  const uint64_t per = (tiles + 1023) / 1024;
  const uint64_t b = min((uint64_t)threadIdx.x * per, tiles), e = min(b + per, tiles);
  uint64_t s = 0;
  for (uint64_t i = b; i < e; ++i) s += counts[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const uint64_t v = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint64_t run = part[threadIdx.x] - s;
  for (uint64_t i = b; i < e; ++i) {
    offsets[i] = (uint32_t)run;
    run += counts[i];
  }
  if (threadIdx.x == 1023) *total = part[1023];
}

// Grid-stride streaming copy of 16-byte words.
__global__ void __launch_bounds__(256) copy_words_kernel(ulonglong2* dst, const ulonglong2* src, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
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
  // this device's copy of the no-allocation filter's scratch
  WG_HIP(hipGetSymbolAddress((void**)&c->noalloc_scratch, HIP_SYMBOL(scratch)));
  *out = c;
  return WG_OK;
}

wg_status wg_ctx_destroy(wg_ctx* ctx) {
  if (!ctx) return WG_OK;
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipEventDestroy(ctx->t0);
  (void)hipEventDestroy(ctx->t1);
  if (ctx->upload) (void)hipHostFree(ctx->upload);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return WG_OK;
}

wg_status wg_ctx_inject_alloc_failure(wg_ctx* ctx, int64_t nth, int32_t sticky) {
  clear_error();
  WG_CHECK(ctx && nth >= 0, WG_ERR_INVALID, "bad argument");
  ctx->failing.store(false);
  ctx->fail_sticky.store(sticky != 0);
  ctx->fail_at.store(nth);
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
  return dev_alloc(ctx, out, bytes, nullptr, "wg_malloc");
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
  return select_rows(ctx, FilterPred{ColDesc{col->data, col->validity, col->type, col->scale}, op, i64, f64}, n, sel,
                     out_sel, n_out);
}

wg_status wg_filter_noalloc(wg_ctx* ctx, const wg_column* col, uint64_t n_rows, const uint32_t* sel, uint64_t n_sel,
                            int32_t op, int64_t i64, double f64, uint32_t* out_sel, uint64_t* n_out) {
  clear_error();
  WG_CHECK(ctx && col && out_sel && n_out, WG_ERR_INVALID, "null argument");
  WG_CHECK(op >= WG_EQ && op <= WG_GE, WG_ERR_INVALID, "bad compare op %d", op);
  WG_CHECK(col->type == WG_INT32 || col->type == WG_INT64 || col->type == WG_FLOAT64 || col->type == WG_UINT16 ||
               col->type == WG_UINT32 || col->type == WG_IPV4,
           WG_ERR_UNSUPPORTED, "no filter kernel for type %d", col->type);
  WG_CHECK(ctx->noalloc_scratch, WG_ERR_STATE, "the context has no filter scratch");
  const uint64_t n = sel ? n_sel : n_rows;
  WG_CHECK(n <= kNoAllocMaxRows, WG_ERR_INVALID,
           "wg_filter_noalloc takes at most %llu rows per call (%llu given): split the batch or use wg_filter",
           (unsigned long long)kNoAllocMaxRows, (unsigned long long)n);
  *n_out = 0;
  if (n == 0) return WG_OK;
  // every word read below is written first by this call: counts and ballots
  // by pass 1 (for every tile / 64-row word below n), offsets and the total
  // by the scan
  const uint64_t tiles = (n + kSelTile - 1) / kSelTile;
  uint32_t* counts = (uint32_t*)ctx->noalloc_scratch;
  uint32_t* offsets = counts + kNoAllocTiles;
  uint64_t* masks = ctx->noalloc_scratch + kNoAllocTiles;
  uint64_t* total = masks + kNoAllocMaxRows / 64;
  const FilterPred pred{ColDesc{col->data, col->validity, col->type, col->scale}, op, i64, f64};
  hipLaunchKernelGGL((select_rows_kernel<false, FilterPred>), dim3(tiles), dim3(kSelBlock), 0, ctx->stream, pred, n,
                     sel, masks, counts, (const uint32_t*)nullptr, (uint32_t*)nullptr);
  WG_HIP(hipGetLastError());
  hipLaunchKernelGGL(tile_scan_kernel, dim3(1), dim3(1024), 0, ctx->stream, (const uint32_t*)counts, offsets, tiles,
                     total);
  WG_HIP(hipGetLastError());
  hipLaunchKernelGGL((select_rows_kernel<true, FilterPred>), dim3(tiles), dim3(kSelBlock), 0, ctx->stream, pred, n,
                     sel, masks, (uint32_t*)nullptr, (const uint32_t*)offsets, out_sel);
  WG_HIP(hipGetLastError());
  WG_HIP(hipMemcpyAsync(n_out, total, 8, hipMemcpyDeviceToHost, ctx->stream));
  WG_HIP(hipStreamSynchronize(ctx->stream));
  return WG_OK;
}

}  // extern "C"
