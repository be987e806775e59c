// wg_key_dict.hip — device dictionary of {lo, hi} key pairs -> 64-bit ids.
//
// Keys the reference handles as two 64-bit words — the 128-bit packed GROUP BY
// key (buildPackedLayout / packedKeyAt, internal/engine/exec/packed_hash.go:
// 384-472) and the two-column join key (dualIntKeyFromVectors, join.go:
// 1621-1648) — are mapped here to a dense 64-bit id per distinct pair, so the
// group table and the join table keep their one-word keys and claims; the
// id -> {lo, hi} array turns ids back into key columns.
//
// Dictionary slot (4 words, 32-byte aligned, all ones when free):
//   fp      63-bit fingerprint of the pair (the claim word: CAS from free)
//   lo, hi  the pair
//   id      the pair's id (all ones until assigned)
// An encode is three kernels, so that no kernel reads a pair another lane of
// the same kernel is still writing (no fences on the hot path):
//   claim   every row probes from fp; a slot holding its fingerprint ends the
//           probe (the same pair, or — checked below — another pair with the
//           same fingerprint); a free slot is claimed by CAS and its pair
//           written;
//   assign  the new slots get ids (workgroup ranges of the table, one shared
//           counter atomic per workgroup) and the id -> pair array;
//   lookup  every row probes again with plain loads and takes the id of the
//           slot holding its exact pair.
// A pair the lookup does not find (its fingerprint was another pair's in the
// same claim pass) is claimed in another round, where the earlier claim's
// pair is visible and differs; a table past its probe budget grows first.
#include <algorithm>
#include <vector>

#include "wg_key_dict.hpp"

#include "wg_agg_host.hpp"

namespace wg {

constexpr uint64_t kDictFree = ~0ull;
constexpr int kDictMaxProbe = 256;
constexpr int kDictBlock = 256;           // threads per workgroup of the dictionary kernels
constexpr uint64_t kAssignRange = 16384;  // table slots per workgroup of the assign pass

struct WideDict {
  uint64_t* slots;
  uint64_t mask;
  uint64_t* keys;           // id -> {lo, hi}
  unsigned long long* ctr;  // [0] ids handed out, [1] a claim ran out of probes, [2] a lookup missed
};

__device__ __forceinline__ uint64_t wide_fp(uint64_t lo, uint64_t hi) {
  return mix64(lo ^ mix64(hi + 0x9E3779B97F4A7C15ull)) >> 1;  // below the free pattern
}

// Claim pass for one pair.
__device__ __forceinline__ void dict_claim(const WideDict& D, uint64_t lo, uint64_t hi) {
  const uint64_t fp = wide_fp(lo, hi);
  uint64_t i = fp & D.mask;
  for (int p = 0; p < kDictMaxProbe; ++p) {
    uint64_t* s = D.slots + i * 4;
    uint64_t st = *s;
    if (st == kDictFree) {
      uint64_t expected = kDictFree;
      if (__hip_atomic_compare_exchange_strong(s, &expected, fp, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT)) {
        s[1] = lo;
        s[2] = hi;
        return;
      }
      st = expected;
    }
    if (st == fp) {
      // the same pair, unless this slot's pair reads fully written (neither
      // word still the all-ones fill, so not half-written by a lane of this
      // pass) and differs: then it is another pair with this fingerprint
      const uint64_t l2 = s[1], h2 = s[2];
      const bool settled = l2 != kDictFree && h2 != kDictFree;
      if (!settled || (l2 == lo && h2 == hi)) return;
    }
    i = (i + 1) & D.mask;
  }
  D.ctr[1] = 1;  // out of probes: the table grows
}

// Lookup pass for one pair: its id, or kNoId (flagged when FLAG) when absent.
template <bool FLAG>
__device__ __forceinline__ uint64_t dict_lookup(const WideDict& D, uint64_t lo, uint64_t hi) {
  const uint64_t fp = wide_fp(lo, hi);
  uint64_t i = fp & D.mask;
  for (int p = 0; p < kDictMaxProbe; ++p) {
    const uint64_t* s = D.slots + i * 4;
    const ulonglong2 a = *(const ulonglong2*)s;
    if (a.x == kDictFree) break;
    if (a.x == fp) {
      const ulonglong2 b = *(const ulonglong2*)(s + 2);
      if (a.y == lo && b.x == hi) return b.y;
    }
    i = (i + 1) & D.mask;
  }
  if (FLAG) D.ctr[2] = 1;
  return kNoId;
}

// {lo, hi} of a row: each key column's widened value in its word and half
// (packedKeyAt, packed_hash.go:456-472).
__device__ __forceinline__ void wide_pack(const KeyDesc& K, const ColSet& cs, uint64_t row, uint64_t& lo,
                                          uint64_t& hi) {
  lo = hi = 0;
  for (int k = 0; k < K.n; ++k) {
    uint64_t v = (uint64_t)load_int(cs.cols[K.col[k]], row);
    if (K.narrow[k]) v = (v & 0xFFFFFFFFull) << K.shift[k];
    if (K.word[k]) hi |= v;
    else lo |= v;
  }
}

// Pair sources of an encode: a batch's rows, a list of pairs, or the live
// slots of another aggregate's table (whose keys are ids of its dictionary).
struct RowsSrc {
  ColSet cs;
  KeyDesc K;
  const uint32_t* sel;
  uint64_t* ids;  // per row
  bool skip_null;
  uint64_t row0;  // without sel: rows row0 + i
  __device__ bool get(uint64_t i, uint64_t& lo, uint64_t& hi) const {
    const uint64_t row = sel ? (uint64_t)sel[i] : row0 + i;
    if (skip_null)
      for (int k = 0; k < K.n; ++k)
        if (!col_valid(cs.cols[K.col[k]], row)) return false;
    wide_pack(K, cs, row, lo, hi);
    return true;
  }
  __device__ void put(uint64_t i, uint64_t id) const { ids[sel ? (uint64_t)sel[i] : row0 + i] = id; }
};

struct PairsSrc {
  const uint64_t* pairs;
  uint64_t* ids;
  __device__ bool get(uint64_t i, uint64_t& lo, uint64_t& hi) const {
    lo = pairs[2 * i];
    hi = pairs[2 * i + 1];
    return true;
  }
  __device__ void put(uint64_t i, uint64_t id) const { ids[i] = id; }
};

struct SlotsSrc {
  const uint64_t* slots;  // the other table
  int stride;
  const uint64_t* keys;  // its id -> pair array
  uint64_t* copy;        // a copy of its table, keys rewritten
  __device__ bool get(uint64_t j, uint64_t& lo, uint64_t& hi) const {
    const uint64_t w = slots[j * stride];
    if (w == kEmpty) return false;
    lo = keys[2 * w];
    hi = keys[2 * w + 1];
    return true;
  }
  __device__ void put(uint64_t j, uint64_t id) const { copy[j * stride] = id; }
};

template <typename Src>
__global__ void __launch_bounds__(kDictBlock) dict_claim_kernel(Src src, uint64_t n, WideDict D) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t lo, hi;
    if (src.get(i, lo, hi)) dict_claim(D, lo, hi);
  }
}

template <typename Src, bool FLAG>
__global__ void __launch_bounds__(kDictBlock) dict_lookup_kernel(Src src, uint64_t n, WideDict D) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t lo, hi;
    if (src.get(i, lo, hi)) src.put(i, dict_lookup<FLAG>(D, lo, hi));
  }
}

__device__ __forceinline__ bool needs_id(const uint64_t* s) { return s[0] != kDictFree && s[3] == kNoId; }

// Ids for the slots claimed since the last assign: each workgroup numbers
// the new slots of its kAssignRange-slot range in slot order from a base
// taken from the shared counter.
__global__ void __launch_bounds__(kDictBlock) dict_assign_kernel(WideDict D, uint64_t cap) {
  __shared__ uint32_t wave_tot[kDictBlock / 64];
  __shared__ unsigned long long base;
  constexpr uint64_t per = kAssignRange / kDictBlock;
  const uint64_t beg = blockIdx.x * kAssignRange + threadIdx.x * per;
  const uint64_t end = min(beg + per, cap);
  uint32_t mine = 0;
  for (uint64_t j = beg; j < end; ++j) mine += needs_id(D.slots + j * 4) ? 1u : 0u;
  // exclusive prefix of `mine` over the workgroup
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  uint32_t x = mine;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  if (lane == 63) wave_tot[wave] = x;
  __syncthreads();
  uint32_t off = 0, total = 0;
  for (int w = 0; w < kDictBlock / 64; ++w) {
    if (w < wave) off += wave_tot[w];
    total += wave_tot[w];
  }
  if (threadIdx.x == 0) base = total ? atomicAdd(D.ctr, (unsigned long long)total) : 0ull;
  __syncthreads();
  uint64_t id = base + off + x - mine;
  for (uint64_t j = beg; j < end; ++j) {
    uint64_t* s = D.slots + j * 4;
    if (!needs_id(s)) continue;
    s[3] = id;
    D.keys[2 * id] = s[1];
    D.keys[2 * id + 1] = s[2];
    ++id;
  }
}

// Move every claimed slot of an old table into an empty one (fingerprints,
// pairs and ids kept).
__global__ void dict_rehash_kernel(const uint64_t* old_slots, uint64_t old_cap, uint64_t* slots, uint64_t mask) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < old_cap; j += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t fp = old_slots[j * 4];
    if (fp == kDictFree) continue;
    uint64_t i = fp & mask;
    for (uint64_t p = 0; p <= mask; ++p) {
      uint64_t expected = kDictFree;
      if (__hip_atomic_compare_exchange_strong(slots + i * 4, &expected, fp, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT)) {
        slots[i * 4 + 1] = old_slots[j * 4 + 1];
        slots[i * 4 + 2] = old_slots[j * 4 + 2];
        slots[i * 4 + 3] = old_slots[j * 4 + 3];
        break;
      }
      i = (i + 1) & mask;
    }
  }
}

__global__ void fill_words_kernel(uint64_t* p, uint64_t n, uint64_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

static WideDict dict_of(const KeyDict& d) {
  return WideDict{d.slots, d.cap - 1, d.keys, (unsigned long long*)d.ctr};
}

static wg_status fill(wg_ctx* ctx, uint64_t* p, uint64_t n, uint64_t v) {
  if (n == 0) return WG_OK;
  hipLaunchKernelGGL(fill_words_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, ctx->stream, p, n, v);
  WG_HIP(hipGetLastError());
  return WG_OK;
}

// A table of `cap` slots and an id array of as many entries; the claimed
// slots of the old table move over with their ids.
static wg_status dict_alloc(wg_ctx* ctx, KeyDict& d, uint64_t cap) {
  uint64_t* slots = nullptr;
  uint64_t* keys = nullptr;
  WG_ALLOC(ctx, &slots, cap * 32, "a key dictionary");
  const wg_status st = dev_alloc(ctx, (void**)&keys, cap * 16, nullptr, "a key dictionary");
  if (st != WG_OK) {
    (void)hipFree(slots);
    return st;
  }
  WG_TRY(fill(ctx, slots, cap * 4, kDictFree));
  if (d.slots) {
    if (d.n) WG_HIP(hipMemcpyAsync(keys, d.keys, d.n * 16, hipMemcpyDeviceToDevice, ctx->stream));
    hipLaunchKernelGGL(dict_rehash_kernel, dim3(grid_for(d.cap, 256)), dim3(256), 0, ctx->stream,
                       (const uint64_t*)d.slots, d.cap, slots, cap - 1);
    WG_HIP(hipGetLastError());
  }
  WG_HIP(hipStreamSynchronize(ctx->stream));
  if (d.slots) WG_HIP(hipFree(d.slots));
  if (d.keys) WG_HIP(hipFree(d.keys));
  d.slots = slots;
  d.keys = keys;
  d.cap = cap;
  return WG_OK;
}

wg_status kd_init(wg_ctx* ctx, KeyDict& d, uint64_t hint) {
  WG_ALLOC(ctx, &d.ctr, 24, "the key dictionary's counters");
  WG_TRY(zero_dev(ctx, d.ctr, 3));
  d.n = 0;
  return dict_alloc(ctx, d, next_pow2(std::max<uint64_t>(2 * std::max<uint64_t>(hint, 1024), 64)));
}

void kd_free(KeyDict& d) {
  (void)hipFree(d.slots);
  (void)hipFree(d.keys);
  (void)hipFree(d.ctr);
  d = KeyDict{};
}

wg_status kd_reset(wg_ctx* ctx, KeyDict& d) {
  WG_TRY(fill(ctx, d.slots, d.cap * 4, kDictFree));
  WG_TRY(zero_dev(ctx, d.ctr, 3));
  d.n = 0;
  return WG_OK;
}

uint64_t kd_bytes(const KeyDict& d) { return d.cap * 48; }

// claim -> assign -> lookup over n items of src, growing the table when a
// claim ran out of probes or more than half of it is taken, and claiming
// again when a lookup missed (a fingerprint shared by two pairs).
template <typename Src>
static wg_status dict_encode(wg_ctx* ctx, KeyDict& d, const Src& src, uint64_t n) {
  if (n == 0) return WG_OK;
  const int grid = grid_for(n, kDictBlock, 8192);
  for (int round = 0; round < 12; ++round) {
    WG_TRY(zero_dev(ctx, d.ctr + 1, 2));
    hipLaunchKernelGGL(dict_claim_kernel<Src>, dim3(grid), dim3(kDictBlock), 0, ctx->stream, src, n, dict_of(d));
    WG_HIP(hipGetLastError());
    const uint64_t ranges = (d.cap + kAssignRange - 1) / kAssignRange;
    hipLaunchKernelGGL(dict_assign_kernel, dim3((unsigned)ranges), dim3(kDictBlock), 0, ctx->stream, dict_of(d),
                       d.cap);
    WG_HIP(hipGetLastError());
    uint64_t c[3];
    WG_HIP(hipMemcpyAsync(c, d.ctr, 24, hipMemcpyDeviceToHost, ctx->stream));
    WG_HIP(hipStreamSynchronize(ctx->stream));
    d.n = c[0];
    if (c[1] || 2 * c[0] > d.cap) {  // out of probes, or over half full: grow and claim again
      WG_TRY(dict_alloc(ctx, d, std::max<uint64_t>(2 * d.cap, next_pow2(4 * c[0]))));
      continue;
    }
    hipLaunchKernelGGL((dict_lookup_kernel<Src, true>), dim3(grid), dim3(kDictBlock), 0, ctx->stream, src, n,
                       dict_of(d));
    WG_HIP(hipGetLastError());
    WG_HIP(hipMemcpyAsync(c, d.ctr, 24, hipMemcpyDeviceToHost, ctx->stream));
    WG_HIP(hipStreamSynchronize(ctx->stream));
    if (!c[2]) return WG_OK;
  }
  set_error("key dictionary did not converge");
  return WG_ERR_DEVICE;
}

wg_status kd_encode_rows(wg_ctx* ctx, KeyDict& d, const ColSet& cs, const KeyDesc& K, const uint32_t* sel,
                         uint64_t n, bool skip_null, uint64_t* ids, uint64_t row0) {
  return dict_encode(ctx, d, RowsSrc{cs, K, sel, ids, skip_null, row0}, n);
}

wg_status kd_lookup_rows(wg_ctx* ctx, KeyDict& d, const ColSet& cs, const KeyDesc& K, const uint32_t* sel,
                         uint64_t n, bool skip_null, uint64_t* ids, uint64_t row0) {
  if (n == 0) return WG_OK;
  hipLaunchKernelGGL((dict_lookup_kernel<RowsSrc, false>), dim3(grid_for(n, kDictBlock, 8192)), dim3(kDictBlock), 0,
                     ctx->stream, RowsSrc{cs, K, sel, ids, skip_null, row0}, n, dict_of(d));
  WG_HIP(hipGetLastError());
  return WG_OK;
}

wg_status kd_encode_pairs(wg_ctx* ctx, KeyDict& d, const uint64_t* pairs, uint64_t n, uint64_t* ids) {
  return dict_encode(ctx, d, PairsSrc{pairs, ids}, n);
}

wg_status kd_remap_slots(wg_ctx* ctx, KeyDict& d, const uint64_t* slots, uint64_t cap, int stride,
                         const uint64_t* src_keys, uint64_t* copy) {
  return dict_encode(ctx, d, SlotsSrc{slots, stride, src_keys, copy}, cap);
}

wg_status kd_host_pairs(wg_ctx* ctx, const KeyDict& d, std::vector<uint64_t>* pairs) {
  pairs->resize(d.n * 2);
  if (d.n) {
    WG_HIP(hipMemcpyAsync(pairs->data(), d.keys, d.n * 16, hipMemcpyDeviceToHost, ctx->stream));
    WG_HIP(hipStreamSynchronize(ctx->stream));
  }
  return WG_OK;
}

}  // namespace wg
