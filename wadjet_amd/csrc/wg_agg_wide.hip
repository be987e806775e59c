// wg_agg_wide.hip — GROUP BY keys of 65 to 128 packed bits.
//
// The reference packs 2-4 int-class group columns into a 128-bit {lo, hi}
// key (buildPackedLayout / packedKeyAt, internal/engine/exec/packed_hash.go:
// 384-472) and probes a table of 24-byte entries (packedHashTable, :47-328).
// Here such a key goes through the key dictionary (wg_key_dict.hip) before it
// reaches the group table: the batch's rows get a per-row id column and the
// aggregate runs on that id exactly as on a one-column int64 key — direct or
// partitioned consume, heavy hitters, merge — so every aggregation kernel
// keeps its one-word claims. Finalize and the spill run turn ids back into
// the key columns through the dictionary's id -> {lo, hi} array.
#include <algorithm>
#include <vector>

#include "wg_agg_host.hpp"

namespace wg {

wg_status wide_init(wg_agg* h) { return kd_init(h->ctx, h->dict, h->group_hint); }

wg_status wide_free(wg_agg* h) {
  kd_free(h->dict);
  (void)hipFree(h->ids);
  h->ids = nullptr;
  return WG_OK;
}

wg_status wide_reset(wg_agg* h) { return kd_reset(h->ctx, h->dict); }

wg_status wide_attach(wg_agg* h, ColSet& cs, const uint32_t* sel, uint64_t n, uint64_t n_rows) {
  if (h->ids_cap < std::max<uint64_t>(n_rows, 1)) {
    if (h->ids) WG_HIP(hipFree(h->ids));
    h->ids = nullptr;
    h->ids_cap = 0;
    WG_ALLOC(h->ctx, &h->ids, std::max<uint64_t>(n_rows, 1) * 8, "group-key ids");
    h->ids_cap = std::max<uint64_t>(n_rows, 1);
  }
  WG_TRY(kd_encode_rows(h->ctx, h->dict, cs, h->wkey, sel, n, false, h->ids));
  const int ic = h->n_cols;
  cs.cols[ic] = ColDesc{h->ids, nullptr, WG_INT64, 0};
  cs.inline_word[ic] = -1;
  cs.n_cols = ic + 1;
  cs.n_probe = ic + 1;
  return WG_OK;
}

wg_status wide_encode_pairs(wg_agg* h, const uint64_t* pairs, uint64_t n, uint64_t* ids_out) {
  return kd_encode_pairs(h->ctx, h->dict, pairs, n, ids_out);
}

wg_status wide_dict_host(wg_agg* h, std::vector<uint64_t>* pairs) { return kd_host_pairs(h->ctx, h->dict, pairs); }

// A copy of src's table whose keys are ids of dst's dictionary.
wg_status wide_remap_slots(wg_agg* dst, wg_agg* src, uint64_t** remapped) {
  *remapped = nullptr;
  uint64_t* copy = nullptr;
  const uint64_t words = (src->cap + kSideSlots) * (uint64_t)src->pg.stride;
  WG_ALLOC(dst->ctx, &copy, words * 8, "a remapped group table");
  wg_status st = WG_OK;
  if (hipMemcpyAsync(copy, src->slots, words * 8, hipMemcpyDeviceToDevice, dst->ctx->stream) != hipSuccess) {
    set_error("hipMemcpyAsync");
    st = WG_ERR_DEVICE;
  }
  if (st == WG_OK)
    st = kd_remap_slots(dst->ctx, dst->dict, src->slots, src->cap, src->pg.stride, src->dict.keys, copy);
  if (st != WG_OK) {
    (void)hipFree(copy);
    return st;
  }
  *remapped = copy;
  return WG_OK;
}

}  // namespace wg
