// wg_key_dict.hpp — device dictionary of {lo, hi} key pairs -> 64-bit ids,
// shared by GROUP BY keys wider than 64 bits (wg_agg) and two-column join keys
// (wg_join). See wg_key_dict.hip for the slot layout and the passes.
#pragma once
#include <vector>

#include "wg_core.hpp"

namespace wg {

struct ColSet;

struct KeyDict {
  uint64_t* slots = nullptr;  // cap slots of 4 words {fingerprint, lo, hi, id}
  uint64_t cap = 0;
  uint64_t base = 0;          // id of a pair claimed in slot i of this table: base + i
  uint64_t* keys = nullptr;   // id -> {lo, hi}, n entries
  uint64_t n = 0;             // id space: base + cap (ids of earlier, smaller tables below base)
  uint64_t claimed = 0;       // taken slots of this table (as of the last read-back)
  uint64_t* ctr = nullptr;    // device: [0] claims of a pass, [1] a claim ran out of probes, [2] a lookup
                              // missed, [3] rows to verify
  uint64_t* vlist = nullptr;  // device: the rows a claim pass could not settle (vcap entries)
  uint64_t vcap = 0;
};

constexpr uint64_t kNoId = ~0ull;  // the id of a pair not in the dictionary

wg_status kd_init(wg_ctx* ctx, KeyDict& d, uint64_t hint);
void kd_free(KeyDict& d);
wg_status kd_reset(wg_ctx* ctx, KeyDict& d);
uint64_t kd_bytes(const KeyDict& d);
// ids[row] for the active rows (sel, or row0 .. row0 + n): the pair packed
// from the key columns K of cs (packedKeyAt's placement: K.word / K.shift /
// K.narrow), new pairs inserted. Rows with a NULL key column are skipped (their
// id left unwritten) when skip_null.
wg_status kd_encode_rows(wg_ctx* ctx, KeyDict& d, const ColSet& cs, const KeyDesc& K, const uint32_t* sel,
                         uint64_t n, bool skip_null, uint64_t* ids, uint64_t row0 = 0);
// The same without inserting: pairs not in the dictionary get kNoId.
wg_status kd_lookup_rows(wg_ctx* ctx, KeyDict& d, const ColSet& cs, const KeyDesc& K, const uint32_t* sel,
                         uint64_t n, bool skip_null, uint64_t* ids, uint64_t row0 = 0);
// ids[i] of pairs[2 i], pairs[2 i + 1] (device), new pairs inserted.
wg_status kd_encode_pairs(wg_ctx* ctx, KeyDict& d, const uint64_t* pairs, uint64_t n, uint64_t* ids);
// copy[j * stride] = this dictionary's id of the pair whose id in `src_keys`
// is slots[j * stride], for every slot j < cap whose word is not kEmpty.
wg_status kd_remap_slots(wg_ctx* ctx, KeyDict& d, const uint64_t* slots, uint64_t cap, int stride,
                         const uint64_t* src_keys, uint64_t* copy);
// The id -> pair array, to the host.
wg_status kd_host_pairs(wg_ctx* ctx, const KeyDict& d, std::vector<uint64_t>* pairs);

}  // namespace wg
