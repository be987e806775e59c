// wg_host.hpp — the host-only part of the library's shared pieces: status
// handling and the context. Included by every translation unit, including
// the host-only C++ ones that are compiled without the device compiler.
#pragma once
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <mutex>

#include "../../include/wadjet_gpu.h"

namespace wg {

// The message wg_last_error returns, per calling thread.
void set_error(const char* fmt, ...);
void clear_error();

#define WG_HIP(call)                                                                 \
  do {                                                                               \
    hipError_t _e = (call);                                                          \
    if (_e != hipSuccess) {                                                          \
      ::wg::set_error("%s:%d: %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(_e)); \
      return WG_ERR_DEVICE;                                                          \
    }                                                                                \
  } while (0)

#define WG_CHECK(cond, code, ...)  \
  do {                             \
    if (!(cond)) {                 \
      ::wg::set_error(__VA_ARGS__); \
      return (code);               \
    }                              \
  } while (0)

#define WG_TRY(expr)                  \
  do {                                \
    wg_status _s = (expr);            \
    if (_s != WG_OK) return _s;       \
  } while (0)

// Every device allocation of the library: hipMalloc (stream = nullptr) or
// hipMallocAsync on `stream`, behind the context's injected failures
// (wg_ctx_inject_alloc_failure). *p is null after a failure, which is
// WG_ERR_OOM with the request in the message.
wg_status dev_alloc(wg_ctx* ctx, void** p, size_t bytes, hipStream_t stream, const char* what);

#define WG_ALLOC(ctx, pp, bytes, what) WG_TRY(::wg::dev_alloc((ctx), (void**)(pp), (bytes), nullptr, (what)))
#define WG_ALLOC_ASYNC(ctx, pp, bytes, stream, what) \
  WG_TRY(::wg::dev_alloc((ctx), (void**)(pp), (bytes), (stream), (what)))

}  // namespace wg

struct wg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  int cus = 0;
  int lds_per_block = 0;
  uint64_t hbm = 0;
  // Pinned staging for per-call tables and uploads (the Parquet page table,
  // the Parquet staging slots), grown on demand; a call holds upload_mu from
  // filling it until it has synchronised, so calls on one context from
  // several threads do not share it.
  void* upload = nullptr;
  uint64_t upload_cap = 0;
  std::mutex upload_mu;
  // device address of this device's no-allocation filter scratch
  uint64_t* noalloc_scratch = nullptr;
  // injected allocation failures: the fail_at-th allocation from the time of
  // injection fails (0 = none), and every later one too when sticky
  std::atomic<int64_t> fail_at{0};
  std::atomic<bool> fail_sticky{false};
  std::atomic<bool> failing{false};
};
