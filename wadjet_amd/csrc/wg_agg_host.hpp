// wg_agg_host.hpp — host-side state of a wg_agg handle (shared by the
// aggregate, the fused join probe and the batched runner).
#pragma once
#include <vector>

#include "wg_agg_dev.hpp"
#include "wg_key_dict.hpp"

namespace wg {

// Finalization recipe of one aggregate (which words decide NULL and value).
struct AggFinal {
  int32_t func;
  int32_t count_word;
  int32_t val_kind;  // OpKind or 0 (always NULL: MIN/MAX DECIMAL, SUM/AVG IPv4)
  int32_t val_word;
  int32_t is_float_in;
  int32_t sumd_index;
  int32_t scale;
  int32_t limbs;
  int32_t in_type;
};

// Staging of host batches (WG_MEM_HOST): rows are gathered through the
// selection vector into pinned buffers and shipped in large device batches.
struct HostStage {
  uint64_t cap_rows = 0;
  uint64_t rows = 0;
  std::vector<void*> pinned;     // per column
  std::vector<uint64_t*> pvalid;  // per nullable column (bitmap words)
  std::vector<void*> dev;         // per column
  std::vector<uint64_t*> dvalid;
  hipEvent_t copied = nullptr;    // last H2D of the stage finished
};

}  // namespace wg

struct wg_agg {
  wg_ctx* ctx = nullptr;
  // schema
  int32_t n_cols = 0;
  int32_t col_types[wg::kMaxCols] = {};
  int32_t col_scales[wg::kMaxCols] = {};
  uint8_t col_nullable[wg::kMaxCols] = {};
  // compiled program
  wg::KeyDesc key{};
  std::vector<wg_agg_spec> aggs;
  std::vector<wg::AggFinal> finals;
  wg::AggProgram pg{};
  int32_t n_words = 0;  // used words incl. key
  // exact Int128 bases of the SUMD limb groups
  int32_t n_sumd = 0;
  std::vector<int32_t> sumd_word, sumd_limbs;
  int32_t* d_sumd_word = nullptr;
  int32_t* d_sumd_limbs = nullptr;
  uint64_t* base = nullptr;  // (cap + side) * n_sumd * 2 words, lazily allocated
  bool base_valid = false;
  uint64_t rows_since_norm = 0;
  // global table
  uint64_t* slots = nullptr;
  uint64_t cap = 0;
  uint64_t groups = 0;  // as of the last read-back
  uint64_t* counters = nullptr;  // device CTR_N words
  // pinned read-back words of this handle (counters, work-item cursors): not
  // the context's, so handles on one context may run on different threads
  uint64_t* pinned = nullptr;
  uint32_t* ovf_rows = nullptr;
  uint64_t ovf_rows_cap = 0;
  uint64_t* ovf_part = nullptr;
  uint64_t ovf_part_cap = 0;
  bool dirty = false;       // table slots not in the freshly-filled state
  bool side_dirty = false;  // only the side slots written since the last fill
  // groups a partitioned consume into an empty aggregate wrote straight to
  // the dense output (rows [0, n_dense) of `dense`), not into the table
  uint64_t n_dense = 0;
  // partitioned consume (high-cardinality batches)
  int32_t strategy = 0;  // 0 auto, 1 direct, 2 partitioned
  uint64_t group_hint = 0;
  int32_t part_lds_slots = 2048;
  int32_t part_block = 1024;
  uint64_t part_chunk_rows = 1u << 16;
  uint32_t* part_cnt = nullptr;  // counts, offsets, boundaries (u32)
  uint64_t part_cnt_cap = 0;
  void* part_scan = nullptr;
  size_t part_scan_bytes = 0;
  uint64_t* part_buf = nullptr;  // partitioned key + value words
  uint64_t part_buf_words = 0;
  void* part_items = nullptr;
  uint64_t part_items_bytes = 0;
  void* part_host = nullptr;  // pinned: partition boundaries + work items
  uint64_t part_host_bytes = 0;
  int32_t part_hot = 512;          // heavy-hitter slots of the partitioned consume (0: none)
  uint64_t* hot_buf = nullptr;     // hot-set selection scratch + per-workgroup partial slots
  uint64_t hot_buf_words = 0;
  float pass_ms[4] = {0, 0, 0, 0};  // direct consume, count+scan, scatter, partition aggregate
  hipEvent_t pev[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  // the partitioned consume's copies of its boundaries and work items, on a
  // stream of their own so that pass 2 never waits on them
  hipStream_t side = nullptr;
  // tuning
  int32_t lds_slots = 1024;
  int32_t block = 512;
  int32_t grid = 0;
  uint64_t tile_rows = 0;
  // statistics of the last consume
  uint64_t st_lds = 0, st_glb = 0, st_flushed = 0;
  float st_kernel_ms = 0.f;  // consume-kernel time of the last consume (HIP events)
  uint32_t st_launches = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // failure handling: `touched` is set when a call starts changing the
  // aggregate's contents; a call that fails after that leaves it `poisoned`
  // (holding part of that call's input) and every later call but reset and
  // destroy is refused
  bool touched = false;
  bool poisoned = false;
  // finalize output
  bool finalized = false;
  uint64_t out_n = 0;
  uint64_t* dense = nullptr;
  uint64_t* dense_base = nullptr;
  uint8_t* dense_null = nullptr;
  uint64_t dense_cap = 0;
  // host staging
  wg::HostStage stage;
  // composite keys of 65-128 bits: the key columns pack into a {lo, hi} pair
  // (buildPackedLayout, packed_hash.go:384-425) and a device dictionary maps
  // each distinct pair to a 64-bit id; the id is the table's key (`key` then
  // names the per-row id column, index n_cols of the consumed ColSet)
  bool wide = false;
  wg::KeyDesc wkey{};      // the key columns' places in {lo, hi}
  wg::KeyDict dict;
  uint64_t* ids = nullptr;  // per-row ids of the batch being consumed
  uint64_t ids_cap = 0;
};

namespace wg {
// Internal entry points shared with the join and batch runners.
wg_status agg_consume_device(wg_agg* h, const ColSet& cs, const uint32_t* sel, uint64_t n,
                             const PredDesc* pred /* [2], col = -1 unused */);
wg_status agg_ensure_capacity(wg_agg* h, uint64_t extra_groups);
wg_status agg_read_counters(wg_agg* h, uint64_t* c);
wg_status agg_prepare_rows(wg_agg* h, uint64_t n);
wg_status agg_flush_stage(wg_agg* h);
uint64_t agg_footprint(const wg_agg* h);
int grid_for(uint64_t n, int block, int max_blocks = 4096);
wg_status zero_dev(wg_ctx* ctx, uint64_t* p, uint64_t n);
GlobalTable table_of(wg_agg* h);
wg_status grow_to(wg_agg* h, uint64_t new_cap);
wg_status normalize(wg_agg* h);
wg_status ensure_overflow_buffers(wg_agg* h);
wg_status ensure_overflow_for(wg_agg* h, uint64_t grid, int32_t lds_slots);
wg_status ensure_base(wg_agg* h);
wg_status agg_merge_partials(wg_agg* h, const uint64_t* recs, uint64_t n);
wg_status agg_merge_slots(wg_agg* h, const uint64_t* src, uint64_t src_cap, const uint64_t* src_base,
                          uint64_t src_groups);
bool part_eligible(const wg_agg* h, const ColSet& cs);
// keys wider than 64 bits (wg_agg_wide.hip)
wg_status wide_init(wg_agg* h);
wg_status wide_free(wg_agg* h);
wg_status wide_reset(wg_agg* h);
wg_status wide_attach(wg_agg* h, ColSet& cs, const uint32_t* sel, uint64_t n, uint64_t n_rows);
wg_status wide_encode_pairs(wg_agg* h, const uint64_t* pairs, uint64_t n, uint64_t* ids_out);
wg_status wide_dict_host(wg_agg* h, std::vector<uint64_t>* pairs);
wg_status wide_remap_slots(wg_agg* dst, wg_agg* src, uint64_t** remapped);
// Fold groups emitted to the dense output back into the table (before any
// consume, merge or merge-run that must see them there).
// Entry and exit of a call that changes the aggregate: refuse a poisoned one;
// poison it when the call failed after changing it.
wg_status agg_guard_begin(wg_agg* h);
wg_status agg_guard_end(wg_agg* h, wg_status st);
wg_status agg_settle(wg_agg* h);
// The dense output arrays with room for `rows` groups, keeping the first `keep`.
wg_status agg_ensure_dense(wg_agg* h, uint64_t rows, uint64_t keep);
wg_status agg_consume_partitioned(wg_agg* h, const ColSet& cs, const uint32_t* sel, uint64_t n, const PredDesc* pred);
}  // namespace wg
