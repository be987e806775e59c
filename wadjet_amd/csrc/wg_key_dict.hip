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
//   id      the pair's id
// A pair claimed in slot i of a table gets id base + i: the claiming lane
// knows it at once and writes the slot's pair and id and the id -> pair
// entry. A table that grows moves its slots, ids included, into one twice
// as large whose base is the old base + old cap, so ids are never reused.
// An encode is one claim pass, so that no kernel reads a pair another lane
// of the same kernel is still writing (no fences on the hot path):
//   claim   every row probes from fp; a free slot is claimed by CAS (the row
//           takes its id); a slot holding the row's fingerprint and a fully
//           written pair (neither word the all-ones fill) ends the probe when
//           the pair is the row's (the row takes the slot's id: the id word,
//           or base + slot while the claimer has not written it) and goes on
//           when it is another pair with the same fingerprint; a slot holding
//           the fingerprint and a pair not yet readable gives the row the
//           slot's id for now and puts the row on the verify list;
//   verify  the listed rows probe again with plain loads and take the id of
//           the slot holding their exact pair (every claim of the pass is
//           visible now).
// A listed pair the verify pass does not find (its fingerprint was another
// pair's) is claimed by a further claim pass, where the earlier claim's pair is
// visible and differs; a table past its probe budget or half full grows
// first, and the claim pass runs again over every row.
#include <algorithm>
#include <vector>

#include "wg_key_dict.hpp"

#include "wg_agg_host.hpp"

namespace wg {

constexpr uint64_t kDictFree = ~0ull;
constexpr int kDictMaxProbe = 256;
constexpr int kDictBlock = 256;           // threads per workgroup of the dictionary kernels
constexpr uint64_t kVerifyCap = 1 << 16;  // rows a claim pass can list for the verify pass

struct WideDict {
  uint64_t* slots;
  uint64_t mask;
  uint64_t base;            // id of slot i: base + i
  uint64_t* keys;           // id -> {lo, hi}
  unsigned long long* ctr;  // [0] claims, [1] a claim ran out of probes, [2] a lookup missed, [3] rows to verify
  uint64_t* vlist;
  uint64_t vcap;
};

__device__ __forceinline__ uint64_t wide_fp(uint64_t lo, uint64_t hi) {
  return mix64(lo ^ mix64(hi + 0x9E3779B97F4A7C15ull)) >> 1;  // below the free pattern
}

// Claim pass for one pair: 0 found (id exact), 1 claimed (id exact), 2 its
// fingerprint's slot does not show its pair yet (id to verify), 3 out of
// probes.
__device__ __forceinline__ int dict_claim(const WideDict& D, uint64_t lo, uint64_t hi, uint64_t& id) {
  const uint64_t fp = wide_fp(lo, hi);
  uint64_t i = fp & D.mask;
  for (int p = 0; p < kDictMaxProbe; ++p) {
    uint64_t* s = D.slots + i * 4;
    uint64_t st = *s;
    if (st == kDictFree) {
      uint64_t expected = kDictFree;
      if (__hip_atomic_compare_exchange_strong(s, &expected, fp, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT)) {
        id = D.base + i;
        *(ulonglong2*)(s + 2) = make_ulonglong2(hi, id);
        s[1] = lo;
        *(ulonglong2*)(D.keys + 2 * id) = make_ulonglong2(lo, hi);
        return 1;
      }
      st = expected;
    }
    if (st == fp) {
      // the id word is all ones only while this pass's claimer has not
      // written it, i.e. for a slot of this table claimed now: base + i
      const uint64_t l2 = s[1];
      const ulonglong2 b = *(const ulonglong2*)(s + 2);
      const bool settled = l2 != kDictFree && b.x != kDictFree;
      if (!settled || (l2 == lo && b.x == hi)) {
        id = b.y != kDictFree ? b.y : D.base + i;
        return settled ? 0 : 2;
      }
    }
    i = (i + 1) & D.mask;
  }
  return 3;
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

// Claim pass over n items: every item's id (exact, or listed to verify);
// claims counted per workgroup into ctr[0].
template <typename Src>
__global__ void __launch_bounds__(kDictBlock) dict_claim_kernel(Src src, uint64_t n, WideDict D) {
  __shared__ uint32_t claims;
  if (threadIdx.x == 0) claims = 0;
  __syncthreads();
  uint32_t mine = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t lo, hi, id = kNoId;
    if (!src.get(i, lo, hi)) continue;
    const int r = dict_claim(D, lo, hi, id);
    if (r == 3) {
      D.ctr[1] = 1;  // out of probes: the table grows and the pass runs again
      continue;
    }
    mine += r == 1;
    // rows to verify, one counter atomic per wave that has any (pairs with
    // an all-ones word never read as written, so a batch of such keys costs
    // one atomic per wave, not per row); past the list's end the host
    // verifies every item
    const uint64_t lm = __ballot(r == 2);  // the same in every active lane
    if (lm) {
      const int lead = __builtin_ctzll(lm);
      const uint32_t lane = __lane_id();
      unsigned long long at = 0;
      if (lane == (uint32_t)lead) at = atomicAdd(D.ctr + 3, (unsigned long long)__popcll(lm));
      at = __shfl(at, lead) + __popcll(lm & ((1ull << lane) - 1));
      if (r == 2 && at < D.vcap) D.vlist[at] = i;
    }
    src.put(i, id);
  }
  if (mine) atomicAdd(&claims, mine);
  __syncthreads();
  if (threadIdx.x == 0 && claims) atomicAdd(D.ctr, (unsigned long long)claims);
}

// Verify pass over the listed items (list) or over all n items (list null).
template <typename Src>
__global__ void __launch_bounds__(kDictBlock) dict_verify_kernel(Src src, const uint64_t* list, uint64_t n,
                                                                 WideDict D) {
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = list ? list[k] : k;
    uint64_t lo, hi;
    if (src.get(i, lo, hi)) src.put(i, dict_lookup<true>(D, lo, hi));
  }
}

template <typename Src, bool FLAG>
__global__ void __launch_bounds__(kDictBlock) dict_lookup_kernel(Src src, uint64_t n, WideDict D) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t lo, hi;
    if (src.get(i, lo, hi)) src.put(i, dict_lookup<FLAG>(D, lo, hi));
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
  return WideDict{d.slots, d.cap - 1, d.base, d.keys, (unsigned long long*)d.ctr, d.vlist, d.vcap};
}

static wg_status fill(wg_ctx* ctx, uint64_t* p, uint64_t n, uint64_t v) {
  if (n == 0) return WG_OK;
  hipLaunchKernelGGL(fill_words_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, ctx->stream, p, n, v);
  WG_HIP(hipGetLastError());
  return WG_OK;
}

// A table of `cap` slots whose ids start after the old table's, and an id
// array over the whole id space; the claimed slots of the old table move
// over with their ids.
static wg_status dict_alloc(wg_ctx* ctx, KeyDict& d, uint64_t cap) {
  const uint64_t base = d.slots ? d.base + d.cap : 0;
  uint64_t* slots = nullptr;
  uint64_t* keys = nullptr;
  WG_ALLOC(ctx, &slots, cap * 32, "a key dictionary");
  const wg_status st = dev_alloc(ctx, (void**)&keys, (base + cap) * 16, nullptr, "a key dictionary");
  if (st != WG_OK) {
    (void)hipFree(slots);
    return st;
  }
  WG_TRY(fill(ctx, slots, cap * 4, kDictFree));
  if (d.slots) {
    WG_HIP(hipMemcpyAsync(keys, d.keys, d.n * 16, hipMemcpyDeviceToDevice, ctx->stream));
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
  d.base = base;
  d.n = base + cap;
  return WG_OK;
}

wg_status kd_init(wg_ctx* ctx, KeyDict& d, uint64_t hint) {
  WG_ALLOC(ctx, &d.ctr, 32, "the key dictionary's counters");
  WG_TRY(zero_dev(ctx, d.ctr, 4));
  WG_ALLOC(ctx, &d.vlist, kVerifyCap * 8, "the key dictionary's verify list");
  d.vcap = kVerifyCap;
  d.claimed = 0;
  return dict_alloc(ctx, d, next_pow2(std::max<uint64_t>(2 * std::max<uint64_t>(hint, 1024), 64)));
}

void kd_free(KeyDict& d) {
  (void)hipFree(d.slots);
  (void)hipFree(d.keys);
  (void)hipFree(d.ctr);
  (void)hipFree(d.vlist);
  d = KeyDict{};
}

wg_status kd_reset(wg_ctx* ctx, KeyDict& d) {
  WG_TRY(fill(ctx, d.slots, d.cap * 4, kDictFree));
  WG_TRY(zero_dev(ctx, d.ctr, 4));
  d.claimed = 0;
  d.base = 0;  // every id is free again (the id array holds at least cap entries)
  d.n = d.cap;
  return WG_OK;
}

uint64_t kd_bytes(const KeyDict& d) { return d.cap * 32 + d.n * 16 + d.vcap * 8; }

// claim (-> verify the listed items) over n items of src, growing the table
// when a claim ran out of probes or more than half of it is taken, and
// claiming again when a verify missed (a fingerprint shared by two pairs).
template <typename Src>
static wg_status dict_encode(wg_ctx* ctx, KeyDict& d, const Src& src, uint64_t n) {
  if (n == 0) return WG_OK;
  const int grid = grid_for(n, kDictBlock, 8192);
  for (int round = 0; round < 12; ++round) {
    WG_TRY(zero_dev(ctx, d.ctr, 4));
    hipLaunchKernelGGL(dict_claim_kernel<Src>, dim3(grid), dim3(kDictBlock), 0, ctx->stream, src, n, dict_of(d));
    WG_HIP(hipGetLastError());
    uint64_t c[4];
    WG_HIP(hipMemcpyAsync(c, d.ctr, 32, hipMemcpyDeviceToHost, ctx->stream));
    WG_HIP(hipStreamSynchronize(ctx->stream));
    d.claimed += c[0];
    if (c[1] || 2 * d.claimed > d.cap) {  // out of probes, or over half full: grow and claim again
      WG_TRY(dict_alloc(ctx, d, std::max<uint64_t>(2 * d.cap, next_pow2(4 * d.claimed))));
      continue;
    }
    if (!c[3]) return WG_OK;
    const bool listed = c[3] <= d.vcap;
    const uint64_t m = listed ? c[3] : n;
    hipLaunchKernelGGL(dict_verify_kernel<Src>, dim3(grid_for(m, kDictBlock, 8192)), dim3(kDictBlock), 0,
                       ctx->stream, src, listed ? (const uint64_t*)d.vlist : nullptr, m, dict_of(d));
    WG_HIP(hipGetLastError());
    WG_HIP(hipMemcpyAsync(c, d.ctr, 32, hipMemcpyDeviceToHost, ctx->stream));
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
