// wg_join_agg.hpp — the streamlined fused probe -> GROUP BY (parameters and
// launcher shared between the kernel file and the join's host code).
#pragma once
#include "wg_agg_dev.hpp"
#include "wg_join_dev.hpp"

namespace wg {

constexpr int kLeanMaxVals = 4;   // value words a joined row carries
constexpr int kLeanMaxReplicas = 64;

// A joined row in this shape reads its group key and values from probe
// columns or from the inline payload words 2..3 of the matched join slot, has
// no NULL group key or value, and aggregates into a small group table held in
// LDS with R replicas of every accumulator word (lane l updates replica
// l % R, so lanes of one wave that hit the same group do not serialise on one
// LDS address). Rows outside the shape (sentinel-valued keys, keys with
// several build rows, a group key equal to the sentinel) go to the overflow
// row list and are aggregated by the general kernel afterwards.
struct JoinLeanParams {
  JoinTable J;
  ColDesc pkey;                     // probe key column
  ColDesc gcol;                     // group key probe column (gword < 0)
  int32_t gword;                    // group key inline slot word (2 or 3), or -1
  int32_t n_vals;
  ColDesc vcol[kLeanMaxVals];       // value probe columns (vword[j] < 0)
  int32_t vword[kLeanMaxVals];      // value inline slot word (2 or 3), or -1
  uint32_t vfloat;                  // bit j: value word j holds float64 bits
  AggProgram pg;                    // op columns index the value words
  GlobalTable T;
  const uint32_t* sel;
  uint64_t row_base, n, tile_rows, n_tiles;
  uint32_t* ovf_rows;
  uint64_t ovf_rows_cap;
  uint64_t* ovf_part;
  int32_t L;  // LDS group slots (power of two)
  int32_t R;  // accumulator replicas (power of two, <= 64)
};

// LDS bytes of the replicated table: (L + 1 spare slot) x stride words x R.
inline size_t lean_lds_bytes(int L, int R, int stride) { return (size_t)(L + 1) * stride * R * 8; }

wg_status launch_join_agg_lean(const JoinLeanParams& P, int grid, int block, hipStream_t st);

}  // namespace wg
