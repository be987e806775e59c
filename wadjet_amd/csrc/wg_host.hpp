// wg_host.hpp — the host-only part of the library's shared pieces: status
// handling and the context. Included by every translation unit, including
// the host-only C++ ones that are compiled without the device compiler.
#pragma once
#include <hip/hip_runtime_api.h>

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
};
