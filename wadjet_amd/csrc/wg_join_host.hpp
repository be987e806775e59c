// wg_join_host.hpp — host-side state of a wg_join handle (shared with the
// batched / served runners that probe a built dimension table).
#pragma once
#include "wg_join_dev.hpp"
#include "wg_key_dict.hpp"

struct wg_join {
  wg_ctx* ctx = nullptr;
  int32_t n_cols = 0;
  int32_t types[wg::kMaxCols] = {};
  int32_t scales[wg::kMaxCols] = {};
  int32_t key_col = 0;
  // two-column keys (dualIntKeyFromVectors, join.go:1621-1648): each distinct
  // (key0, key1) pair has an id in `dict`; the retained column n_cols holds
  // the build rows' ids (NULL when either key is) and is the table's key
  bool dual = false;
  wg::KeyDesc dkey{};  // the two build key columns, lo = key0, hi = key1
  wg::KeyDict dict;
  int32_t n_ret = 0;  // retained columns: n_cols, plus the id column when dual
  int32_t n_inline = 0;
  int32_t inline_cols[wg::kMaxInline] = {};
  int32_t stride = 2;
  // retained build columns (device), capacity in rows
  void* col_data[wg::kMaxCols] = {};
  uint64_t* col_valid[wg::kMaxCols] = {};
  bool has_valid[wg::kMaxCols] = {};
  uint64_t rows = 0, row_cap = 0;
  int32_t* next = nullptr;
  // table
  uint64_t* slots = nullptr;
  uint64_t cap = 0;
  uint64_t* counter = nullptr;  // claimed slots + scratch words
  bool finished = false;
  bool claimed_unique = false;  // every non-NULL build key distinct (set at finish)
  uint64_t distinct = 0;        // distinct non-NULL build keys (set at finish)
  // Bloom pre-filter (wg_join_bloom_build): bloom_mask + 1 words of 64 bits
  uint64_t* bloom = nullptr;
  uint64_t bloom_mask = 0;
  // fused probe -> aggregate: streamlined kernel when the query has its shape
  int32_t lean = 1;           // 0: always the general kernel
  int32_t lean_block = 512;   // 512 or 1024 threads
  int32_t lean_replicas = 0;  // accumulator replicas (0: as many as the LDS budget allows)
  // probe scratch
  uint64_t* cnt = nullptr;
  uint64_t* off = nullptr;
  uint64_t scratch_rows = 0;
  void* cub_tmp = nullptr;
  size_t cub_bytes = 0;
};

