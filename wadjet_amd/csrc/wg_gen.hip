// wg_gen.hip — seeded synthetic column generators (device side).
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/wadjet_gen.h"
#include "wg_core.hpp"

namespace wg {

__host__ __device__ __forceinline__ uint64_t splitmix_at(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// murmur3 fmix32: a bijection of [0, 2^32)
__host__ __device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

__device__ __forceinline__ double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

__device__ __forceinline__ int64_t uniform_in(uint64_t r, int64_t lo, int64_t hi) {
  const uint64_t range = (uint64_t)(hi - lo) + 1;
  if (range == 0) return (int64_t)r;  // full 64-bit range
  return lo + (int64_t)__umul64hi(r, range);
}

// Keyed Feistel permutation of [0, 2^(2*half)), cycle-walked into [0, dom).
__device__ __forceinline__ uint64_t feistel_perm(uint64_t x, uint64_t dom, int half, uint64_t seed) {
  const uint64_t mask = (1ull << half) - 1;
  do {
    uint64_t l = x >> half, r = x & mask;
    for (int k = 0; k < 4; ++k) {
      const uint64_t f = mix64(r ^ (seed + 0x632BE59BD9B4E019ull * (k + 1))) & mask;
      const uint64_t nl = r;
      r = l ^ f;
      l = nl;
    }
    x = (l << half) | r;
  } while (x >= dom);
  return x;
}

__global__ void gen_kernel(int kind, void* out, uint64_t row0, uint64_t n, uint64_t seed, int64_t i0, int64_t i1,
                           double f0, double f1, const double* cdf, int half) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t row = row0 + j;
    const uint64_t r = splitmix_at(seed, row);
    switch (kind) {
      case WG_GEN_UNIFORM_I64: ((int64_t*)out)[j] = uniform_in(r, i0, i1); break;
      case WG_GEN_UNIFORM_I32: ((int32_t*)out)[j] = (int32_t)uniform_in(r, i0, i1); break;
      case WG_GEN_UNIFORM_U16: ((uint16_t*)out)[j] = (uint16_t)uniform_in(r, i0, i1); break;
      case WG_GEN_ZIPF_I32:
      case WG_GEN_ZIPF_U32: {
        const double u = u01(r);
        // smallest rank whose CDF exceeds u
        int64_t lo = 0, hi = i0 - 1;
        while (lo < hi) {
          const int64_t mid = (lo + hi) >> 1;
          if (cdf[mid] > u) hi = mid;
          else lo = mid + 1;
        }
        const uint64_t rank = (uint64_t)lo + 1;
        if (kind == WG_GEN_ZIPF_I32) ((int32_t*)out)[j] = (int32_t)rank;
        else ((uint32_t*)out)[j] = fmix32((uint32_t)rank ^ (uint32_t)seed);
        break;
      }
      case WG_GEN_NORMAL_F64:
      case WG_GEN_LOGNORMAL_I64: {
        uint64_t k = 0;
        double z;
        do {
          const uint64_t a = splitmix_at(seed ^ 0xA5A5A5A5A5A5A5A5ull, 2 * row + k);
          const uint64_t b = splitmix_at(seed ^ 0x5A5A5A5A5A5A5A5Aull, 2 * row + k);
          const double u1 = ((double)(a >> 11) + 0.5) * (1.0 / 9007199254740992.0);
          const double u2 = u01(b);
          z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
          k += 1000003;
        } while (z == 0.0);
        if (kind == WG_GEN_NORMAL_F64) ((double*)out)[j] = f0 + f1 * z;
        else ((int64_t*)out)[j] = (int64_t)llrint(exp(f0 + f1 * z));
        break;
      }
      case WG_GEN_PERM_KEY_I64:
        ((int64_t*)out)[j] = (int64_t)mix64(feistel_perm(row, (uint64_t)i0, half, seed));
        break;
      case WG_GEN_FK_I64: {
        const uint64_t r2 = splitmix_at(seed ^ 0x3C3C3C3C3C3C3C3Cull, row);
        const bool match = (int64_t)(r2 % 1000) < i1;
        const uint64_t u = __umul64hi(r, (uint64_t)i0);
        ((int64_t*)out)[j] = (int64_t)mix64(match ? u : (uint64_t)i0 + u);
        break;
      }
      case WG_GEN_SEQ_I64: ((int64_t*)out)[j] = i0 + (int64_t)row * i1; break;
      case WG_GEN_RANK_U32: ((uint32_t*)out)[j] = fmix32((uint32_t)(row + 1) ^ (uint32_t)seed); break;
      default: break;
    }
  }
}

// Zipf CDF tables, cached per (n, s) for the process lifetime.
static std::mutex g_cdf_mu;
static std::map<std::pair<int64_t, double>, double*> g_cdf;

static wg_status zipf_cdf(int64_t n, double s, double** out) {
  std::lock_guard<std::mutex> lk(g_cdf_mu);
  auto it = g_cdf.find({n, s});
  if (it != g_cdf.end()) {
    *out = it->second;
    return WG_OK;
  }
  std::vector<double> c(n);
  double acc = 0.0;
  for (int64_t r = 1; r <= n; ++r) {
    acc += std::pow((double)r, -s);
    c[r - 1] = acc;
  }
  for (auto& v : c) v /= acc;
  c[n - 1] = 2.0;  // every u < 1 lands at or before the last rank
  double* d = nullptr;
  WG_ALLOC(nullptr, &d, n * 8, "a Zipf CDF");
  WG_HIP(hipMemcpy(d, c.data(), n * 8, hipMemcpyHostToDevice));
  g_cdf[{n, s}] = d;
  *out = d;
  return WG_OK;
}

}  // namespace wg

using namespace wg;

extern "C" wg_status wg_gen(wg_ctx* ctx, int32_t kind, void* out, uint64_t row0, uint64_t n, uint64_t seed,
                            int64_t i0, int64_t i1, double f0, double f1) {
  clear_error();
  WG_CHECK(ctx && (out || n == 0), WG_ERR_INVALID, "null argument");
  WG_CHECK(kind >= WG_GEN_UNIFORM_I64 && kind <= WG_GEN_RANK_U32, WG_ERR_INVALID, "generator %d", kind);
  if (n == 0) return WG_OK;
  double* cdf = nullptr;
  int half = 1;
  if (kind == WG_GEN_ZIPF_I32 || kind == WG_GEN_ZIPF_U32) {
    WG_CHECK(i0 >= 1 && i0 <= (1ll << 30), WG_ERR_INVALID, "zipf over %lld values", (long long)i0);
    WG_TRY(zipf_cdf(i0, f0, &cdf));
  }
  if (kind == WG_GEN_PERM_KEY_I64 || kind == WG_GEN_FK_I64) {
    WG_CHECK(i0 >= 1 && i0 < (1ll << 40), WG_ERR_INVALID, "key domain %lld", (long long)i0);
  }
  if (kind == WG_GEN_PERM_KEY_I64) {
    WG_CHECK(row0 + n <= (uint64_t)i0, WG_ERR_INVALID, "permutation rows past the domain");
    int bits = 1;
    while ((1ll << bits) < i0) ++bits;
    half = (bits + 1) / 2;
  }
  const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(gen_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, kind, out, row0, n, seed, i0, i1,
                     f0, f1, (const double*)cdf, half);
  WG_HIP(hipGetLastError());
  WG_HIP(hipStreamSynchronize(ctx->stream));
  return WG_OK;
}
