// wg_agg_wide.hip — GROUP BY keys of 65 to 128 packed bits.
//
// The reference packs 2-4 int-class group columns into a 128-bit {lo, hi}
// key (buildPackedLayout / packedKeyAt, internal/engine/exec/packed_hash.go:
// 384-472) and probes a table of 24-byte entries (packedHashTable, :47-328).
// Here such a key is dictionary-encoded before it reaches the group table:
// a device hash table maps every distinct {lo, hi} pair to a dense 64-bit id
// (handed out in blocks per wavefront), the batch's rows get a per-row id
// column, and the aggregate runs on that id exactly as on a one-column int64
// key — direct or partitioned consume, heavy hitters, merge — so every
// aggregation kernel keeps its one-word claims. Finalize and the spill run
// turn ids back into the key columns through the id -> {lo, hi} array.
//
// Dictionary slot (4 words, 32-byte aligned, all ones when free): state, lo,
// hi, unused. state is kDictFree, kDictBusy while the claiming lane writes
// lo/hi, kDictDead for a claim that found no id left, else the id (published
// with release order after lo/hi). Ids are below every reserved state value.
#include <algorithm>
#include <vector>

#include "wg_agg_host.hpp"

namespace wg {

constexpr uint64_t kDictFree = ~0ull;
constexpr uint64_t kDictBusy = ~0ull - 1;
constexpr uint64_t kDictDead = ~0ull - 2;
constexpr uint64_t kNoId = ~0ull;
constexpr int kDictMaxProbe = 512;
constexpr uint64_t kDictMaxSpin = 1ull << 22;
constexpr uint64_t kIdBlock = 256;  // ids a wavefront takes from the shared counter at a time
constexpr int kDictBlock = 256;     // threads per workgroup of the dictionary kernels

struct WideDict {
  uint64_t* slots;
  uint64_t mask;
  uint64_t* keys;     // id -> {lo, hi}
  uint64_t keys_cap;  // ids that fit
  unsigned long long* ctr;  // [0] ids reserved in blocks, [1] rows left without an id, [2] ids exhausted
};

__device__ __forceinline__ uint64_t wide_hash(uint64_t lo, uint64_t hi) {
  return mix64(lo ^ mix64(hi + 0x9E3779B97F4A7C15ull));
}

// The pair's id, inserting it if new; kNoId (counted in ctr[1]) when the id
// space or the probe budget ran out — the host grows and runs again.
//
// Every word of a slot changes once: state Free -> Busy -> id (or Dead), lo
// and hi from the fill pattern (all ones) to the pair. A slot is first read
// with plain loads; the read settles it when the state is an id (or Dead) and
// neither pair word still shows the fill pattern, since a cached line may be
// older than the claim but never holds a half-written pair otherwise. Any
// other read (free, busy, an all-ones word) is settled with read-modify-write
// atomics, which act at the device's coherence point: a CAS to claim a free
// slot, fetch-ors of zero to read a published one, and for a busy slot a look
// again in the next round — never a spin inside the round, so a claiming
// lane of the same wavefront publishes first.
__device__ uint64_t dict_find_or_insert(const WideDict& D, uint64_t lo, uint64_t hi, unsigned long long* blk) {
  const uint64_t lanes_below = (1ull << __lane_id()) - 1;
  uint64_t i = wide_hash(lo, hi) & D.mask;
  int probes = 0;
  uint64_t spins = 0;
  bool settle = false;
  for (;;) {
    uint64_t* s = D.slots + i * 4;
    uint64_t st, l2, h2;
    if (!settle) {
      const ulonglong2 a = *(const ulonglong2*)s;
      const ulonglong2 b = *(const ulonglong2*)(s + 2);
      st = a.x;
      l2 = a.y;
      h2 = b.x;
      settle = !(st < kDictDead && l2 != kDictFree && h2 != kDictFree) && st != kDictDead;
    }
    bool claimed = false;
    if (settle) {
      settle = false;
      st = __hip_atomic_fetch_or(s, 0ull, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
      if (st == kDictFree && *(volatile const unsigned long long*)(D.ctr + 2)) break;  // a hint: ids ran out
      if (st == kDictFree) {
        uint64_t expected = kDictFree;
        claimed = __hip_atomic_compare_exchange_strong(s, &expected, kDictBusy, __ATOMIC_ACQUIRE, __ATOMIC_ACQUIRE,
                                                       __HIP_MEMORY_SCOPE_AGENT);
        if (!claimed) st = expected;
      }
      if (!claimed && st < kDictDead) {
        l2 = __hip_atomic_fetch_or(s + 1, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        h2 = __hip_atomic_fetch_or(s + 2, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    // this round's claimers take consecutive ids of the wavefront's block
    // (blk: the wavefront's next id and block end, in LDS); a new block
    // comes from the shared counter, kIdBlock >= 64 ids at a time
    const uint64_t m = __ballot(claimed);
    uint64_t nid = 0;
    if (m) {
      const int leader = __ffsll((long long)m) - 1;
      const unsigned long long cnt = (unsigned long long)__popcll(m);
      unsigned long long base = 0;
      if (__lane_id() == leader) {
        unsigned long long b = blk[0], e = blk[1];
        if (e - b < cnt) {
          b = atomicAdd(D.ctr, (unsigned long long)kIdBlock);
          e = b + kIdBlock;
        }
        base = b;
        blk[0] = b + cnt;
        blk[1] = e;
      }
      base = __shfl(base, leader);
      nid = base + (uint64_t)__popcll(m & lanes_below);
    }
    if (claimed) {
      if (nid < D.keys_cap) {
        (void)__hip_atomic_exchange(s + 1, lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        (void)__hip_atomic_exchange(s + 2, hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        D.keys[2 * nid] = lo;
        D.keys[2 * nid + 1] = hi;
        (void)__hip_atomic_exchange(s, nid, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        return nid;
      }
      (void)__hip_atomic_exchange(s, kDictDead, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(D.ctr + 2, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      atomicAdd(D.ctr + 1, 1ull);
      return kNoId;
    }
    if (st == kDictBusy) {  // being written: settle the same slot next round
      settle = true;
      if (++spins > kDictMaxSpin) break;
      continue;
    }
    if (st < kDictDead && l2 == lo && h2 == hi) return st;
    if (++probes >= kDictMaxProbe) break;
    i = (i + 1) & D.mask;
  }
  atomicAdd(D.ctr + 1, 1ull);
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

// The wavefront's id block in LDS, empty at the start of a kernel.
__device__ __forceinline__ unsigned long long* wave_block(unsigned long long (*blocks)[2]) {
  unsigned long long* blk = blocks[threadIdx.x >> 6];
  if (__lane_id() == 0) blk[0] = blk[1] = 0;
  __builtin_amdgcn_wave_barrier();
  return blk;
}

__global__ void __launch_bounds__(kDictBlock) wide_encode_kernel(ColSet cs, KeyDesc K, const uint32_t* sel,
                                                                 uint64_t n, WideDict D, uint64_t* ids) {
  __shared__ unsigned long long blocks[kDictBlock / 64][2];
  unsigned long long* blk = wave_block(blocks);
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t row = sel ? (uint64_t)sel[i] : i;
    uint64_t lo, hi;
    wide_pack(K, cs, row, lo, hi);
    ids[row] = dict_find_or_insert(D, lo, hi, blk);
  }
}

// Remap: the key word of every live slot of src's table (an id of src's
// dictionary, whose pair is src_keys[2 id]) becomes the pair's id in D,
// written into a copy of the table.
__global__ void __launch_bounds__(kDictBlock) wide_remap_kernel(const uint64_t* src_slots, uint64_t cap, int stride,
                                                                const uint64_t* src_keys, WideDict D,
                                                                uint64_t* copy) {
  __shared__ unsigned long long blocks[kDictBlock / 64][2];
  unsigned long long* blk = wave_block(blocks);
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < cap; j += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t w = src_slots[j * stride];
    if (w == kEmpty) continue;
    copy[j * stride] = dict_find_or_insert(D, src_keys[2 * w], src_keys[2 * w + 1], blk);
  }
}

// A run's pairs through the dictionary (ids[i] of pairs[2 i], pairs[2 i + 1]).
__global__ void __launch_bounds__(kDictBlock) wide_pairs_kernel(const uint64_t* pairs, uint64_t n, WideDict D,
                                                                uint64_t* ids) {
  __shared__ unsigned long long blocks[kDictBlock / 64][2];
  unsigned long long* blk = wave_block(blocks);
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    ids[i] = dict_find_or_insert(D, pairs[2 * i], pairs[2 * i + 1], blk);
}

// Move the published slots of an old table (ids below `keep`) into an empty
// one, ids kept.
__global__ void wide_rehash_kernel(const uint64_t* old_slots, uint64_t old_cap, uint64_t keep, uint64_t* slots,
                                   uint64_t mask) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < old_cap; j += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t id = old_slots[j * 4];
    if (id >= kDictDead || id >= keep) continue;
    const uint64_t lo = old_slots[j * 4 + 1], hi = old_slots[j * 4 + 2];
    uint64_t i = wide_hash(lo, hi) & mask;
    for (uint64_t p = 0; p <= mask; ++p) {
      uint64_t expected = kDictFree;
      if (__hip_atomic_compare_exchange_strong(slots + i * 4, &expected, kDictBusy, __ATOMIC_RELAXED,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        slots[i * 4 + 1] = lo;
        slots[i * 4 + 2] = hi;
        slots[i * 4] = id;
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

static WideDict dict_of(wg_agg* h) {
  return WideDict{h->dict, h->dict_cap - 1, h->dict_keys, h->dict_keys_cap, (unsigned long long*)h->dict_ctr};
}

static wg_status fill(wg_agg* h, uint64_t* p, uint64_t n, uint64_t v) {
  if (n == 0) return WG_OK;
  hipLaunchKernelGGL(fill_words_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, h->ctx->stream, p, n, v);
  WG_HIP(hipGetLastError());
  return WG_OK;
}

// Table of `cap` slots and room for cap/2 ids; the published pairs with ids
// below `keep` move over from the old table (ids are handed out in blocks,
// so [0, keep) may hold unused ids).
static wg_status dict_alloc(wg_agg* h, uint64_t cap, uint64_t keep) {
  uint64_t* slots = nullptr;
  uint64_t* keys = nullptr;
  if (hipMalloc(&slots, cap * 32) != hipSuccess || hipMalloc(&keys, cap / 2 * 16) != hipSuccess) {
    (void)hipGetLastError();
    if (slots) (void)hipFree(slots);
    set_error("hipMalloc for a %llu-slot key dictionary failed", (unsigned long long)cap);
    return WG_ERR_OOM;
  }
  WG_TRY(fill(h, slots, cap * 4, kDictFree));
  if (keep) {
    WG_HIP(hipMemcpyAsync(keys, h->dict_keys, keep * 16, hipMemcpyDeviceToDevice, h->ctx->stream));
    hipLaunchKernelGGL(wide_rehash_kernel, dim3(grid_for(h->dict_cap, 256)), dim3(256), 0, h->ctx->stream,
                       (const uint64_t*)h->dict, h->dict_cap, keep, slots, cap - 1);
    WG_HIP(hipGetLastError());
  }
  WG_HIP(hipStreamSynchronize(h->ctx->stream));
  if (h->dict) WG_HIP(hipFree(h->dict));
  if (h->dict_keys) WG_HIP(hipFree(h->dict_keys));
  h->dict = slots;
  h->dict_keys = keys;
  h->dict_cap = cap;
  h->dict_keys_cap = cap / 2;
  h->dict_n = keep;
  uint64_t c[3] = {keep, 0, 0};
  WG_HIP(hipMemcpy(h->dict_ctr, c, 24, hipMemcpyHostToDevice));
  return WG_OK;
}

wg_status wide_init(wg_agg* h) {
  WG_HIP(hipMalloc(&h->dict_ctr, 24));
  return dict_alloc(h, next_pow2(std::max<uint64_t>(2 * std::max<uint64_t>(h->group_hint, 1024), 64)), 0);
}

wg_status wide_free(wg_agg* h) {
  (void)hipFree(h->dict);
  (void)hipFree(h->dict_keys);
  (void)hipFree(h->dict_ctr);
  (void)hipFree(h->ids);
  h->dict = h->dict_keys = h->dict_ctr = h->ids = nullptr;
  return WG_OK;
}

wg_status wide_reset(wg_agg* h) {
  WG_TRY(fill(h, h->dict, h->dict_cap * 4, kDictFree));
  WG_TRY(zero_dev(h->ctx, h->dict_ctr, 3));
  h->dict_n = 0;
  return WG_OK;
}

// Run `launch` (an encode over the dictionary) until nothing overflows:
// after an overflow the dictionary grows (room for every row left without an
// id) and the encode runs again — it is idempotent, pairs already in keep
// their ids.
template <typename F>
static wg_status encode_until_fits(wg_agg* h, uint64_t n, F launch) {
  for (int round = 0; round < 16; ++round) {
    WG_TRY(launch(dict_of(h)));
    uint64_t c[3];
    WG_HIP(hipMemcpyAsync(c, h->dict_ctr, 24, hipMemcpyDeviceToHost, h->ctx->stream));
    WG_HIP(hipStreamSynchronize(h->ctx->stream));
    if (!c[1]) {
      h->dict_n = std::min<uint64_t>(c[0], h->dict_keys_cap);  // ids are below this
      return WG_OK;
    }
    // pairs with ids below keys_cap are published; later claims were dropped.
    // Room for every row left without an id (an upper bound of the new
    // pairs), at least twice the ids, at most a 2^23-id step beyond 4x.
    const uint64_t keep = std::min<uint64_t>(c[0], h->dict_keys_cap);
    uint64_t ids = next_pow2(keep + std::min<uint64_t>(c[1], n));
    ids = std::max<uint64_t>(ids, 2 * h->dict_keys_cap);
    ids = std::min<uint64_t>(ids, std::max<uint64_t>(4 * h->dict_keys_cap, 1ull << 23));
    WG_TRY(dict_alloc(h, 2 * ids, keep));
  }
  set_error("key dictionary did not converge");
  return WG_ERR_DEVICE;
}

wg_status wide_attach(wg_agg* h, ColSet& cs, const uint32_t* sel, uint64_t n, uint64_t n_rows) {
  if (h->ids_cap < std::max<uint64_t>(n_rows, 1)) {
    if (h->ids) WG_HIP(hipFree(h->ids));
    h->ids = nullptr;
    h->ids_cap = 0;
    WG_HIP(hipMalloc(&h->ids, std::max<uint64_t>(n_rows, 1) * 8));
    h->ids_cap = std::max<uint64_t>(n_rows, 1);
  }
  if (n) {
    const KeyDesc K = h->wkey;
    const ColSet c = cs;
    WG_TRY(encode_until_fits(h, n, [&](const WideDict& D) -> wg_status {
      hipLaunchKernelGGL(wide_encode_kernel, dim3(grid_for(n, kDictBlock, 8192)), dim3(kDictBlock), 0,
                         h->ctx->stream, c, K, sel, n, D, h->ids);
      WG_HIP(hipGetLastError());
      return WG_OK;
    }));
  }
  const int ic = h->n_cols;
  cs.cols[ic] = ColDesc{h->ids, nullptr, WG_INT64, 0};
  cs.inline_word[ic] = -1;
  cs.n_cols = ic + 1;
  cs.n_probe = ic + 1;
  return WG_OK;
}

wg_status wide_encode_pairs(wg_agg* h, const uint64_t* pairs, uint64_t n, uint64_t* ids_out) {
  if (n == 0) return WG_OK;
  return encode_until_fits(h, n, [&](const WideDict& D) -> wg_status {
    hipLaunchKernelGGL(wide_pairs_kernel, dim3(grid_for(n, kDictBlock, 8192)), dim3(kDictBlock), 0, h->ctx->stream,
                       pairs, n, D, ids_out);
    WG_HIP(hipGetLastError());
    return WG_OK;
  });
}

wg_status wide_dict_host(wg_agg* h, std::vector<uint64_t>* pairs) {
  pairs->resize(h->dict_n * 2);
  if (h->dict_n) {
    WG_HIP(hipMemcpyAsync(pairs->data(), h->dict_keys, h->dict_n * 16, hipMemcpyDeviceToHost, h->ctx->stream));
    WG_HIP(hipStreamSynchronize(h->ctx->stream));
  }
  return WG_OK;
}

// A copy of src's table whose keys are ids of dst's dictionary.
wg_status wide_remap_slots(wg_agg* dst, wg_agg* src, uint64_t** remapped) {
  *remapped = nullptr;
  uint64_t* copy = nullptr;
  const uint64_t words = (src->cap + kSideSlots) * (uint64_t)src->pg.stride;
  WG_HIP(hipMalloc(&copy, words * 8));
  wg_status st = WG_OK;
  if (hipMemcpyAsync(copy, src->slots, words * 8, hipMemcpyDeviceToDevice, dst->ctx->stream) != hipSuccess) {
    set_error("hipMemcpyAsync");
    st = WG_ERR_DEVICE;
  }
  if (st == WG_OK)
    st = encode_until_fits(dst, src->groups + 1, [&](const WideDict& D) -> wg_status {
      hipLaunchKernelGGL(wide_remap_kernel, dim3(grid_for(src->cap, kDictBlock, 8192)), dim3(kDictBlock), 0,
                         dst->ctx->stream, (const uint64_t*)src->slots, src->cap, src->pg.stride,
                         (const uint64_t*)src->dict_keys, D, copy);
      WG_HIP(hipGetLastError());
      return WG_OK;
    });
  if (st != WG_OK) {
    (void)hipFree(copy);
    return st;
  }
  *remapped = copy;
  return WG_OK;
}

}  // namespace wg
