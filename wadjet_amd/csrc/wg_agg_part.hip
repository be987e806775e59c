// wg_agg_part.hip — partitioned consume for high-cardinality GROUP BY.
//
// The direct consume folds every row into one shared table with device
// atomics; when a batch holds millions of groups with a skewed (Zipf) key,
// rows of warm keys queue up on the same slot and on the same L2/memory-side
// atomic unit. This path brings the rows of a key together first, so every
// key is aggregated by ONE workgroup in LDS:
//   A  part_count_kernel    each workgroup owns a contiguous row range, filters
//                           it, hashes the keys and counts rows per partition
//                           (top bits of the hash) in an LDS histogram
//   B  exclusive scan       over the [partition][workgroup] counts
//   C  part_scatter_kernel  the same workgroups re-read their rows and write
//                           (packed key, the value words the program reads)
//                           to their slice of each partition (LDS cursors)
//   D  part_agg_kernel      one workgroup per partition chunk aggregates it in
//                           an LDS table and merges the table into the global
//                           table: plain read-modify-writes for a partition it
//                           owns alone, atomics for the chunks of a large
//                           partition split over several workgroups
// Rows whose key is NULL or equals the table sentinel go straight to the side
// slots in pass C. Results are the same accumulator words as the direct path.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "wg_agg_dev.hpp"
#include "wg_agg_host.hpp"

namespace wg {

constexpr int kMaxVals = 8;

struct PartParams {
  ColSet cs;
  KeyDesc key;
  PredDesc pred[2];
  AggProgram pg;  // for side-slot rows in pass C (original columns)
  GlobalTable T;
  const uint32_t* sel;
  uint64_t n;
  uint64_t rows_per_wg;
  int32_t pbits;
  int32_t n_vals;
  uint32_t* counts;      // [P][G]
  const uint32_t* offs;  // [P][G] exclusive scan of counts
  uint64_t* pkey;
  uint64_t* pval[kMaxVals];
  int32_t val_col[kMaxVals];
};

__device__ __forceinline__ bool part_row(const PartParams& P, uint64_t row, uint64_t& key, int& kind) {
  // kind: 0 = partitioned, 1 = sentinel-valued key, 2 = NULL single key
  if (P.pred[0].col >= 0 && !pred_pass(P.cs.cols[P.pred[0].col], row, P.pred[0].op, P.pred[0].i64, P.pred[0].f64))
    return false;
  if (P.pred[1].col >= 0 && !pred_pass(P.cs.cols[P.pred[1].col], row, P.pred[1].op, P.pred[1].i64, P.pred[1].f64))
    return false;
  const RowRef r{row, 0, nullptr};
  if (!build_key<false>(P.key, P.cs, r, key)) {
    kind = 2;
    return true;
  }
  kind = key == kEmpty ? 1 : 0;
  return true;
}

__global__ void __launch_bounds__(256) part_count_kernel(PartParams P) {
  extern __shared__ uint32_t hist[];
  const uint32_t np = 1u << P.pbits;
  for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  const uint64_t beg = (uint64_t)blockIdx.x * P.rows_per_wg;
  const uint64_t end = min(beg + P.rows_per_wg, P.n);
  for (uint64_t i = beg + threadIdx.x; i < end; i += blockDim.x) {
    const uint64_t row = P.sel ? (uint64_t)P.sel[i] : i;
    uint64_t key;
    int kind;
    if (!part_row(P, row, key, kind) || kind != 0) continue;
    atomicAdd(&hist[key_hash(key) >> (64 - P.pbits)], 1u);
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) P.counts[(uint64_t)p * gridDim.x + blockIdx.x] = hist[p];
}

__device__ __forceinline__ uint64_t raw_word(const ColDesc& c, uint64_t row) {
  if (c.type == WG_FLOAT64) return ((const uint64_t*)c.data)[row];
  return (uint64_t)load_int(c, row);
}

__global__ void __launch_bounds__(256) part_scatter_kernel(PartParams P) {
  extern __shared__ uint32_t cur[];
  const uint32_t np = 1u << P.pbits;
  for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) cur[p] = P.offs[(uint64_t)p * gridDim.x + blockIdx.x];
  __syncthreads();
  const uint64_t beg = (uint64_t)blockIdx.x * P.rows_per_wg;
  const uint64_t end = min(beg + P.rows_per_wg, P.n);
  const int stride = P.pg.stride;
  for (uint64_t i = beg + threadIdx.x; i < end; i += blockDim.x) {
    const uint64_t row = P.sel ? (uint64_t)P.sel[i] : i;
    uint64_t key;
    int kind;
    if (!part_row(P, row, key, kind)) continue;
    if (kind != 0) {
      // side slots: [0] the sentinel-valued key, [1] the NULL key
      uint64_t* s = P.T.slots + (P.T.mask + (kind == 1 ? 1 : 2)) * (uint64_t)stride;
      apply_row<GlbT, false>(s, P.pg, P.cs, RowRef{row, 0, nullptr});
      continue;
    }
    const uint32_t pos = atomicAdd(&cur[key_hash(key) >> (64 - P.pbits)], 1u);
    P.pkey[pos] = key;
    for (int j = 0; j < P.n_vals; ++j) P.pval[j][pos] = raw_word(P.cs.cols[P.val_col[j]], row);
  }
}

struct PartItem {
  uint32_t beg, end;  // positions in the partitioned arrays
  uint32_t split;     // 1: the partition is shared with other items
  uint32_t pad;
};

struct PartAggParams {
  ColSet cs;        // value arrays (partitioned layout), one column per value word
  AggProgram pg;    // ops remapped onto cs
  GlobalTable T;
  const uint64_t* pkey;
  const PartItem* items;
  uint32_t n_items;
  int32_t lds_slots;
  unsigned long long* item_ctr;
  const uint32_t* sel;  // replay: positions to aggregate, in chunks of sel_chunk (split items)
  uint64_t n_sel;
  uint64_t sel_chunk;
  uint64_t max_item_rows;  // headroom per in-flight item in the overflow row list
  uint32_t* ovf_rows;
  uint64_t ovf_rows_cap;
  uint64_t* ovf_part;
  uint64_t ovf_part_cap;
};

__global__ void __launch_bounds__(1024) part_agg_kernel(PartAggParams P) {
  extern __shared__ uint64_t lds[];
  __shared__ int64_t s_item;
  const int stride = P.pg.stride;
  const int L = P.lds_slots;
  uint32_t claims = 0;
  const uint64_t n_work = P.sel ? (P.n_sel + P.sel_chunk - 1) / P.sel_chunk : P.n_items;
  for (;;) {
    if (threadIdx.x == 0) {
      int64_t t = -1;
      // take an item only while every in-flight item could still overflow entirely
      const uint64_t o = __hip_atomic_load(P.T.counters + CTR_OVF_ROWS, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint64_t op = __hip_atomic_load(P.T.counters + CTR_OVF_PART, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (o + (uint64_t)gridDim.x * P.max_item_rows > P.ovf_rows_cap ||
          op + (uint64_t)gridDim.x * P.lds_slots > P.ovf_part_cap) {
        at_add<GlbT>(P.T.counters + CTR_STOPPED, 1);
      } else {
        const unsigned long long w = atomicAdd(P.item_ctr, 1ull);
        if (w < n_work) t = (int64_t)w;
      }
      s_item = t;
    }
    __syncthreads();
    const int64_t item = s_item;
    if (item < 0) break;
    uint64_t beg, end;
    bool split;
    if (P.sel) {
      beg = (uint64_t)item * P.sel_chunk;
      end = min(beg + P.sel_chunk, P.n_sel);
      split = true;
    } else {
      const PartItem it = P.items[item];
      beg = it.beg;
      end = it.end;
      split = it.split != 0;
    }
    for (int i = threadIdx.x; i < L * stride; i += blockDim.x) lds[i] = slot_init_word(P.pg, i % stride);
    __syncthreads();
    for (uint64_t i = beg + threadIdx.x; i < end; i += blockDim.x) {
      const uint64_t pos = P.sel ? (uint64_t)P.sel[i] : i;
      const uint64_t key = P.pkey[pos];
      const uint64_t h = key_hash(key);
      const RowRef r{pos, 0, nullptr};
      uint64_t* ls = lds_find_or_claim(lds, L - 1, stride, key, h, 16);
      if (ls) {
        apply_row<LdsT, false>(ls, P.pg, P.cs, r);
        continue;
      }
      uint64_t* g = glb_find_or_claim(P.T, stride, key, h, claims);
      if (!g) {
        const uint64_t at = __hip_atomic_fetch_add(P.T.counters + CTR_OVF_ROWS, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        P.ovf_rows[at] = (uint32_t)pos;  // in bounds: items are taken only with headroom
        continue;
      }
      apply_row<GlbT, false>(g, P.pg, P.cs, r);
    }
    __syncthreads();
    // merge the LDS table into the global table
    for (int i = threadIdx.x; i < L; i += blockDim.x) {
      const uint64_t* ls = lds + (uint64_t)i * stride;
      const uint64_t key = ls[0];
      if (key == kEmpty) continue;
      uint64_t* g = glb_find_or_claim(P.T, stride, key, key_hash(key), claims);
      if (!g) {
        const uint64_t at = __hip_atomic_fetch_add(P.T.counters + CTR_OVF_PART, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint64_t* rec = P.ovf_part + at * (uint64_t)stride;  // in bounds: headroom as above
        for (int w = 0; w < stride; ++w) rec[w] = ls[w];
        continue;
      }
      if (split) merge_partial(g, ls, P.pg);
      else merge_partial_owned(g, ls, P.pg);
    }
    __syncthreads();
  }
  block_add(P.T.counters + CTR_GROUPS, claims);
}

// ------------------------------------------------------------------ host

// Is the partitioned consume applicable to this aggregate and batch?
bool part_eligible(const wg_agg* h, const ColSet& cs) {
  if (h->key.n > 1)
    for (int k = 0; k < h->key.n; ++k)
      if (cs.cols[h->key.col[k]].valid) return false;
  for (int o = 0; o < h->pg.n_ops; ++o) {
    const OpDesc& op = h->pg.ops[o];
    for (int c : {op.col, op.col2}) {
      if (c < 0) continue;
      if (cs.cols[c].valid) return false;                // value NULLs are not carried
      if (cs.cols[c].type == WG_DECIMAL128) return false;  // 16-byte values are not carried
    }
  }
  return true;
}

wg_status agg_consume_partitioned(wg_agg* h, const ColSet& cs, const uint32_t* sel, uint64_t n, const PredDesc* pred) {
  wg_ctx* ctx = h->ctx;
  // value columns the program reads, each carried as one 8-byte word
  std::vector<int> vcols;
  AggProgram pg2 = h->pg;
  for (int o = 0; o < pg2.n_ops; ++o) {
    for (int32_t* c : {&pg2.ops[o].col, &pg2.ops[o].col2}) {
      if (*c < 0) continue;
      int j = -1;
      for (size_t k = 0; k < vcols.size(); ++k)
        if (vcols[k] == *c) j = (int)k;
      if (j < 0) {
        j = (int)vcols.size();
        vcols.push_back(*c);
      }
      *c = j;
    }
  }
  WG_CHECK((int)vcols.size() <= kMaxVals, WG_ERR_UNSUPPORTED, "partitioned consume carries at most %d value columns",
           kMaxVals);
  const int L = h->part_lds_slots;
  const uint64_t est = std::max<uint64_t>(std::max<uint64_t>(h->group_hint, h->groups), 1);
  int pbits = 6;
  while (pbits < 14 && (1ull << pbits) * (uint64_t)(L / 3) < est) ++pbits;
  const uint32_t np = 1u << pbits;
  const int G = ctx->cus * 4;
  const uint64_t R = (n + G - 1) / G;

  PartParams P{};
  P.cs = cs;
  P.key = h->key;
  P.pred[0] = pred[0];
  P.pred[1] = pred[1];
  P.pg = h->pg;
  P.sel = sel;
  P.n = n;
  P.rows_per_wg = R;
  P.pbits = pbits;
  P.n_vals = (int)vcols.size();
  for (size_t j = 0; j < vcols.size(); ++j) P.val_col[j] = vcols[j];

  // scratch: counts + offsets
  const uint64_t ncnt = (uint64_t)np * G;
  if (h->part_cnt_cap < ncnt) {
    if (h->part_cnt) WG_HIP(hipFree(h->part_cnt));
    WG_HIP(hipMalloc(&h->part_cnt, ncnt * 8));
    h->part_cnt_cap = ncnt;
    size_t bytes = 0;
    WG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, h->part_cnt, h->part_cnt + ncnt, (int)ncnt, ctx->stream));
    if (bytes > h->part_scan_bytes) {
      if (h->part_scan) WG_HIP(hipFree(h->part_scan));
      WG_HIP(hipMalloc(&h->part_scan, bytes));
      h->part_scan_bytes = bytes;
    }
  }
  uint32_t* counts = h->part_cnt;
  uint32_t* offs = h->part_cnt + ncnt;
  P.counts = counts;
  P.offs = offs;
  P.T = table_of(h);

  if (!h->pev[0])
    for (hipEvent_t& e : h->pev) WG_HIP(hipEventCreate(&e));
  // A: count
  WG_HIP(hipEventRecord(h->pev[0], ctx->stream));
  hipLaunchKernelGGL(part_count_kernel, dim3(G), dim3(256), np * 4, ctx->stream, P);
  WG_HIP(hipGetLastError());
  size_t bytes = h->part_scan_bytes;
  WG_HIP(hipcub::DeviceScan::ExclusiveSum(h->part_scan, bytes, counts, offs, (int)ncnt, ctx->stream));
  WG_HIP(hipEventRecord(h->pev[1], ctx->stream));
  // partition boundaries (column 0 of each partition row) and the total
  std::vector<uint32_t> bound(np + 1);
  WG_HIP(hipMemcpy2DAsync(bound.data(), 4, offs, (size_t)G * 4, 4, np, hipMemcpyDeviceToHost, ctx->stream));
  uint32_t last_off = 0, last_cnt = 0;
  WG_HIP(hipMemcpyAsync(&last_off, offs + ncnt - 1, 4, hipMemcpyDeviceToHost, ctx->stream));
  WG_HIP(hipMemcpyAsync(&last_cnt, counts + ncnt - 1, 4, hipMemcpyDeviceToHost, ctx->stream));
  WG_HIP(hipStreamSynchronize(ctx->stream));
  const uint64_t total = (uint64_t)last_off + last_cnt;
  bound[np] = (uint32_t)total;

  // scratch: partitioned key + value arrays
  const uint64_t need = std::max<uint64_t>(total, 1) * (1 + vcols.size());
  if (h->part_buf_words < need) {
    if (h->part_buf) WG_HIP(hipFree(h->part_buf));
    WG_HIP(hipMalloc(&h->part_buf, need * 8));
    h->part_buf_words = need;
  }
  P.pkey = h->part_buf;
  for (size_t j = 0; j < vcols.size(); ++j) P.pval[j] = h->part_buf + (1 + j) * std::max<uint64_t>(total, 1);

  // C: scatter (side-slot rows are folded here)
  h->dirty = true;
  WG_HIP(hipEventRecord(h->pev[2], ctx->stream));
  hipLaunchKernelGGL(part_scatter_kernel, dim3(G), dim3(256), np * 4, ctx->stream, P);
  WG_HIP(hipGetLastError());
  WG_HIP(hipEventRecord(h->pev[3], ctx->stream));

  // D: work items, largest first; partitions above the chunk size are split
  const uint64_t chunk = h->part_chunk_rows;
  std::vector<PartItem> items;
  items.reserve(np + total / chunk + 1);
  for (uint32_t p = 0; p < np; ++p) {
    const uint64_t b = bound[p], e = bound[p + 1];
    if (e <= b) continue;
    const bool split = e - b > chunk;
    for (uint64_t x = b; x < e; x += chunk)
      items.push_back(PartItem{(uint32_t)x, (uint32_t)std::min<uint64_t>(e, x + chunk), split ? 1u : 0u, 0});
  }
  std::stable_sort(items.begin(), items.end(),
                   [](const PartItem& a, const PartItem& b) { return a.end - a.beg > b.end - b.beg; });
  const uint64_t ibytes = std::max<size_t>(items.size(), 1) * sizeof(PartItem) + 64;
  if (h->part_items_bytes < ibytes) {
    if (h->part_items) WG_HIP(hipFree(h->part_items));
    WG_HIP(hipMalloc(&h->part_items, ibytes));
    h->part_items_bytes = ibytes;
  }
  unsigned long long* item_ctr = (unsigned long long*)((char*)h->part_items + ibytes - 64);
  if (!items.empty())
    WG_HIP(hipMemcpyAsync(h->part_items, items.data(), items.size() * sizeof(PartItem), hipMemcpyHostToDevice, ctx->stream));

  PartAggParams Q{};
  Q.cs.n_cols = (int)vcols.size();
  Q.cs.n_probe = Q.cs.n_cols;
  for (size_t j = 0; j < vcols.size(); ++j) {
    const int t = cs.cols[vcols[j]].type;
    const int t2 = t == WG_FLOAT64 ? WG_FLOAT64 : (t == WG_DECIMAL ? WG_DECIMAL : WG_INT64);
    Q.cs.cols[j] = ColDesc{P.pval[j], nullptr, t2, cs.cols[vcols[j]].scale};
    Q.cs.inline_word[j] = -1;
  }
  Q.pg = pg2;
  Q.pkey = P.pkey;
  Q.items = (const PartItem*)h->part_items;
  Q.n_items = (uint32_t)items.size();
  Q.lds_slots = L;
  Q.item_ctr = item_ctr;
  const size_t lds_bytes = (size_t)L * h->pg.stride * 8;
  static bool attr_done = false;
  if (!attr_done) {
    WG_HIP(hipFuncSetAttribute((const void*)part_agg_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
    attr_done = true;
  }
  const int per_cu = std::max(1, std::min<int>((160 * 1024) / (int)lds_bytes, 2048 / h->part_block));
  const int grid_max = ctx->cus * per_cu;
  // overflow lists with headroom for every in-flight item (see the kernel)
  const uint64_t need_rows = (uint64_t)grid_max * chunk * 2;
  if (h->ovf_rows_cap < need_rows) {
    if (h->ovf_rows) WG_HIP(hipFree(h->ovf_rows));
    WG_HIP(hipMalloc(&h->ovf_rows, need_rows * 4));
    h->ovf_rows_cap = need_rows;
  }
  const uint64_t need_part = (uint64_t)grid_max * L * 2;
  if (h->ovf_part_cap < need_part) {
    if (h->ovf_part) WG_HIP(hipFree(h->ovf_part));
    WG_HIP(hipMalloc(&h->ovf_part, need_part * (uint64_t)h->pg.stride * 8));
    h->ovf_part_cap = need_part;
  }
  Q.max_item_rows = chunk;
  Q.sel_chunk = chunk;
  // Items are drained in launches; a launch stops taking items when the
  // overflow lists run out of headroom (table past its load budget): the host
  // then grows the table, folds the partials in, replays the overflowed
  // positions and resumes with the items not yet taken.
  struct Pending { uint32_t* rows; uint64_t n; };
  std::vector<Pending> replays;  // replay lists still to aggregate
  uint64_t items_done = 0;
  bool first = true;
  for (int round = 0; round < 64; ++round) {
    const bool replay = !replays.empty();
    const uint32_t* rsel = replay ? replays.back().rows : nullptr;
    const uint64_t rn = replay ? replays.back().n : 0;
    const uint64_t total_work = replay ? (rn + chunk - 1) / chunk : items.size();
    if (!replay && items_done >= items.size()) break;
    WG_TRY(zero_dev(ctx, h->counters + CTR_OVF_ROWS, CTR_N - 1));
    const uint64_t start_item = replay ? 0 : items_done;
    uint64_t* pin = ctx->pinned + 8;  // [8] start item, [9] items taken (counters use [0, 8))
    pin[0] = start_item;
    WG_HIP(hipMemcpyAsync(item_ctr, pin, 8, hipMemcpyHostToDevice, ctx->stream));
    Q.T = table_of(h);
    Q.sel = rsel;
    Q.n_sel = rn;
    Q.ovf_rows = h->ovf_rows;
    Q.ovf_rows_cap = h->ovf_rows_cap;
    Q.ovf_part = h->ovf_part;
    Q.ovf_part_cap = h->ovf_part_cap;
    const uint64_t left = total_work - start_item;
    const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)grid_max, left));
    WG_HIP(hipEventRecord(h->pev[4], ctx->stream));
    if (left) hipLaunchKernelGGL(part_agg_kernel, dim3(grid), dim3(h->part_block), lds_bytes, ctx->stream, Q);
    WG_HIP(hipGetLastError());
    WG_HIP(hipEventRecord(h->pev[5], ctx->stream));
    WG_HIP(hipMemcpyAsync(pin + 1, item_ctr, 8, hipMemcpyDeviceToHost, ctx->stream));
    uint64_t c[CTR_N];
    WG_TRY(agg_read_counters(h, c));  // synchronises the stream
    const uint64_t taken = pin[1];
    float ms = 0.f;
    if (first) {
      WG_HIP(hipEventElapsedTime(&ms, h->pev[0], h->pev[1]));
      h->pass_ms[1] = ms;
      WG_HIP(hipEventElapsedTime(&ms, h->pev[2], h->pev[3]));
      h->pass_ms[2] = ms;
      first = false;
    }
    WG_HIP(hipEventElapsedTime(&ms, h->pev[4], h->pev[5]));
    h->pass_ms[3] += ms;
    h->groups = c[CTR_GROUPS];
    const uint64_t done_now = std::min<uint64_t>(taken, total_work);
    if (replay) {
      if (done_now >= total_work) {
        WG_HIP(hipFree(replays.back().rows));
        replays.pop_back();
      } else {
        // keep the untaken tail of this replay list
        Pending& r = replays.back();
        const uint64_t skip = done_now * chunk;
        uint32_t* tail = nullptr;
        WG_HIP(hipMalloc(&tail, (r.n - skip) * 4));
        WG_HIP(hipMemcpyAsync(tail, r.rows + skip, (r.n - skip) * 4, hipMemcpyDeviceToDevice, ctx->stream));
        WG_HIP(hipStreamSynchronize(ctx->stream));
        WG_HIP(hipFree(r.rows));
        r.rows = tail;
        r.n -= skip;
      }
    } else {
      items_done = done_now;
    }
    const uint64_t nr = c[CTR_OVF_ROWS], npart = c[CTR_OVF_PART];
    if (nr || npart) {
      WG_TRY(grow_to(h, next_pow2(std::max<uint64_t>((h->groups + nr + npart) * 2, h->cap * 2))));
      if (npart) WG_TRY(agg_merge_partials(h, h->ovf_part, npart));
      if (nr) {
        uint32_t* rows = nullptr;
        WG_HIP(hipMalloc(&rows, nr * 4));
        WG_HIP(hipMemcpyAsync(rows, h->ovf_rows, nr * 4, hipMemcpyDeviceToDevice, ctx->stream));
        WG_HIP(hipStreamSynchronize(ctx->stream));
        replays.push_back(Pending{rows, nr});
      }
    }
  }
  for (auto& r : replays) (void)hipFree(r.rows);
  WG_CHECK(replays.empty() && items_done >= items.size(), WG_ERR_DEVICE, "partitioned consume did not converge");
  return WG_OK;
}

}  // namespace wg
