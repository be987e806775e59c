// wg_join_dev.hpp — device view of a built join table (shared by the join
// probes and the batched / served runners). Slot layout: word 0 key, word 1
// {head build row (low 32 bits), build row count (high 32 bits)}, words 2..
// inline payload of the key's first build row.
#pragma once
#include "wg_core.hpp"

namespace wg {

constexpr int kMaxInline = 4;

struct JoinTable {
  uint64_t* slots;  // cap * stride words + 1 side slot (key == kEmpty)
  uint64_t mask;
  int32_t stride;
  const int32_t* next;
};

__device__ __forceinline__ int32_t slot_head(const uint64_t* s) { return (int32_t)(uint32_t)s[1]; }
__device__ __forceinline__ uint32_t slot_count(const uint64_t* s) { return (uint32_t)(s[1] >> 32); }

// Continue a probe for `key` at slot `idx` (after a miss at the slot before).
__device__ __forceinline__ const uint64_t* join_find_from(const JoinTable& J, uint64_t key, uint64_t idx) {
  for (uint64_t p = 0; p <= J.mask; ++p) {
    const uint64_t* s = J.slots + idx * (uint64_t)J.stride;
    const uint64_t k = s[0];
    if (k == key) return s;
    if (k == kEmpty) return nullptr;
    idx = (idx + 1) & J.mask;
  }
  return nullptr;
}

// The slot of `key`, or null when the key has no build row.
__device__ __forceinline__ const uint64_t* join_find(const JoinTable& J, uint64_t key) {
  if (key == kEmpty) {
    const uint64_t* s = J.slots + (J.mask + 1) * (uint64_t)J.stride;
    return slot_count(s) ? s : nullptr;
  }
  return join_find_from(J, key, mix64(key) & J.mask);
}

// The same probe, reading each slot's key and head/count words in one access
// (one round trip per slot instead of two); *w1 = the matching slot's second
// word. false when the key has no build row.
__device__ __forceinline__ bool join_find_w1(const JoinTable& J, uint64_t key, uint64_t* w1) {
  struct Pair {
    uint64_t k, w;
  } v;
  if (key == kEmpty) {
    __builtin_memcpy(&v, J.slots + (J.mask + 1) * (uint64_t)J.stride, 16);
    *w1 = v.w;
    return (v.w >> 32) != 0;
  }
  uint64_t idx = mix64(key) & J.mask;
  for (uint64_t p = 0; p <= J.mask; ++p) {
    __builtin_memcpy(&v, J.slots + idx * (uint64_t)J.stride, 16);
    if (v.k == key) {
      *w1 = v.w;
      return true;
    }
    if (v.k == kEmpty) return false;
    idx = (idx + 1) & J.mask;
  }
  return false;
}

}  // namespace wg
