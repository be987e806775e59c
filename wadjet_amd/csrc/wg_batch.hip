// wg_batch.hip — many independent GROUP BY queries in ONE persistent launch.
//
// The batched form of the hash aggregate: the reference runs each query's
// HashAggregate in its own pipeline (internal/engine/exec/pipeline.go:81-690);
// on the GPU, small queries cannot fill 256 CUs on their own, so the runner
// packs every query's row tiles into one work queue that persistent
// workgroups drain:
//   * per query: its own open-addressing table in a shared arena, sized from
//     its cardinality estimate, and a layout choice made by the autotuner —
//     LDS table (per tile pre-aggregation, flushed with atomics) or direct
//     global-table updates;
//   * one hipGraph holds {arena fill, counter reset, persistent aggregate,
//     persistent compaction}; a run uploads the descriptors and replays it;
//   * a query whose table ran past its load budget is re-run alone through
//     the plain wg_agg path, so every result is exact whatever the estimate.
// Per query: GROUP BY key SUM(val), COUNT(*); SUM is exact Int128 (carry-save
// limbs), which is also the wrapping int64 sum for INT64 values (low word).
#include <algorithm>
#include <chrono>
#include <numeric>
#include <vector>

#include "wg_agg_dev.hpp"
#include "wg_agg_host.hpp"
#include "wg_join_dev.hpp"
#include "wg_join_host.hpp"

namespace wg {

constexpr int kBStride = 4;  // key, count, sum limb lo, sum limb hi

struct BQ {
  ColDesc key;
  ColDesc val;
  uint64_t* table;
  uint64_t mask;
  uint64_t max_probe;  // a claim farther than this from the key's home slot overflows
  uint32_t lds_slots;  // 0 = global table only
  int32_t dim_word;    // >= 0: group key = this word of the matched dimension slot
  uint32_t fan_bits;   // the query's keys split over 2^fan_bits work items per tile (top hash bits)
  uint32_t wrap;       // 1: integer values, whose SUM wraps at 64 bits: one sum word instead of two limbs
  ColDesc range;       // data == null: no window filter
  int64_t range_lo, range_hi;
  JoinTable dim;       // slots == null: no join
};

struct BWork {
  uint32_t q;
  uint32_t part;  // key partition of a fanned-out query (0 otherwise)
  uint64_t beg, end;
};

struct BCounters {
  unsigned long long groups;
  unsigned long long overflow;
  unsigned long long out_n;
  unsigned long long pad;
};

struct BatchParams {
  const BQ* qs;
  const BWork* work;
  const uint64_t* n_work;  // device scalar (graph-replay friendly)
  unsigned long long* work_ctr;
  BCounters* ctr;
};

__device__ __forceinline__ uint64_t bq_key(const BQ& q, uint64_t row) { return (uint64_t)load_int(q.key, row); }

__device__ __forceinline__ uint64_t* bq_slot(const BQ& q, uint64_t key) {
  if (key == kEmpty) return q.table + (q.mask + 1) * kBStride;  // side slot
  uint64_t idx = mix64(key) & q.mask;
  const uint64_t lim = q.max_probe < q.mask + 1 ? q.max_probe : q.mask + 1;
  for (uint64_t p = 0; p < lim; ++p) {
    uint64_t* s = q.table + idx * kBStride;
    const uint64_t k = __hip_atomic_load(s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == key) return s;
    if (k == kEmpty) {
      const uint64_t prev = at_cas<GlbT>(s, kEmpty, key);
      if (prev == kEmpty || prev == key) return s;
    }
    idx = (idx + 1) & q.mask;
  }
  return nullptr;
}

// Word k of the slot is s[k * ws]: ws = 1 in the arena (array of structs),
// the slot count in the LDS table (struct of arrays).
// An integer SUM wraps at 64 bits (scatterSumInt), so word 2 takes the value
// itself and word 3 stays 0; a DECIMAL SUM is exact: words 2 and 3 take the
// low 32 bits (unsigned) and the high 32 bits (signed) as carry-save limbs.
template <typename S>
__device__ __forceinline__ void bq_apply(uint64_t* s, uint32_t ws, int64_t v, bool wrap) {
  at_add<S>(s + ws, 1);
  if (wrap) {
    at_add<S>(s + 2 * ws, (uint64_t)v);
  } else {
    at_add<S>(s + 2 * ws, (uint64_t)v & 0xFFFFFFFFull);
    at_add<S>(s + 3 * ws, (uint64_t)(v >> 32));
  }
}

constexpr int kBU = 4;  // rows per thread per step, their loads all in flight together

// One work item = rows [beg, end) of one query, and with fan-out only the
// rows whose group key hashes to the item's partition: the LDS table then
// holds 1 / 2^fan_bits of the query's groups. Rows go to the LDS table when
// it has one (flushed into the query's table at the end of the item with
// one atomic per group word), else straight to the query's table.
__global__ void __launch_bounds__(1024) batch_agg_kernel(BatchParams P) {
  extern __shared__ uint64_t lds[];
  __shared__ int64_t s_item;
  const uint64_t n_work = *P.n_work;
  for (;;) {
    if (threadIdx.x == 0) {
      const unsigned long long w = atomicAdd(P.work_ctr, 1ull);
      s_item = w < n_work ? (int64_t)w : -1;
    }
    __syncthreads();
    const int64_t item = s_item;
    __syncthreads();
    if (item < 0) break;
    const BWork W = P.work[item];
    const BQ q = P.qs[W.q];
    BCounters* c = P.ctr + W.q;
    const uint32_t L = q.lds_slots;
    if (L) {
      for (uint32_t i = threadIdx.x; i < L * kBStride; i += blockDim.x) lds[i] = i < L ? kEmpty : 0;
      __syncthreads();
    }
    const ColRes kc = resolve_col(q.key, true), vc = resolve_col(q.val, true);
    const bool ranged = q.range.data != nullptr;
    const ColRes rc = resolve_col(q.range, ranged);
    const uint32_t fsh = 64 - q.fan_bits;
    for (uint64_t base = W.beg; base < W.end; base += (uint64_t)blockDim.x * kBU) {
      uint64_t row[kBU];
      bool in[kBU];
#pragma unroll
      for (int u = 0; u < kBU; ++u) {
        const uint64_t r = base + (uint64_t)u * blockDim.x + threadIdx.x;
        in[u] = r < W.end;
        row[u] = in[u] ? r : W.beg;
      }
      // every load of the step in flight before any value is used
      RawVal kr[kBU], vr[kBU], rr[kBU];
      uint64_t kv[kBU], vv[kBU], rv[kBU];
#pragma unroll
      for (int u = 0; u < kBU; ++u) {
        kr[u] = load_res(kc, row[u]);
        vr[u] = load_res(vc, row[u]);
        rr[u] = load_res(rc, row[u]);
        kv[u] = valid_res(kc, row[u]);
        vv[u] = valid_res(vc, row[u]);
        rv[u] = valid_res(rc, row[u]);
      }
      uint64_t key[kBU], val[kBU], tval[kBU];
      values_of<kBU>(kc.type, kr, key);
      values_of<kBU>(vc.type, vr, val);
      values_of<kBU>(rc.type, rr, tval);
#pragma unroll
      for (int u = 0; u < kBU; ++u) {
        const uint64_t bit = row[u] & 63;
        // NULL keys / values drop out, as do rows outside the window
        bool ok = in[u] && ((kv[u] >> bit) & 1ull) && ((vv[u] >> bit) & 1ull);
        if (ranged) ok = ok && ((rv[u] >> bit) & 1ull) && (int64_t)tval[u] >= q.range_lo && (int64_t)tval[u] < q.range_hi;
        if (!ok) continue;
        uint64_t k = key[u];
        if (q.dim.slots) {
          const uint64_t* js = join_find(q.dim, k);
          if (!js) continue;
          k = js[q.dim_word];
        }
        const uint64_t h = mix64(k);
        if (q.fan_bits && (uint32_t)(h >> fsh) != W.part) continue;  // another item's partition
        const int64_t v = (int64_t)val[u];
        if (L && k != kEmpty) {
          uint64_t* ls = lds_find_or_claim(lds, (int)L - 1, k, h);
          if (ls) {
            bq_apply<LdsT>(ls, L, v, q.wrap);
            continue;
          }
        }
        uint64_t* s = bq_slot(q, k);
        if (!s) {
          atomicAdd(&c->overflow, 1ull);  // rare: the query is re-run alone
          continue;
        }
        bq_apply<GlbT>(s, 1, v, q.wrap);
      }
    }
    if (L) {
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < L; i += blockDim.x) {
        const uint64_t* ls = lds + i;
        if (ls[0] == kEmpty) continue;
        uint64_t* s = bq_slot(q, ls[0]);
        if (!s) {
          atomicAdd(&c->overflow, 1ull);
          continue;
        }
        at_add<GlbT>(s + 1, ls[L]);
        if (ls[2 * L]) at_add<GlbT>(s + 2, ls[2 * L]);  // words that stayed 0 cost no atomic
        if (ls[3 * L]) at_add<GlbT>(s + 3, ls[3 * L]);
      }
      __syncthreads();
    }
  }
}

struct CWork {
  uint32_t q;
  uint32_t pad;
  uint64_t beg, end;  // slot range (the side slot is index mask + 1)
  uint64_t out_off;   // query's output region
};

struct CompactParams {
  const BQ* qs;
  const CWork* work;
  const uint64_t* n_work;
  unsigned long long* work_ctr;
  BCounters* ctr;
  int64_t* keys;
  uint64_t* sum_lo;
  int64_t* sum_hi;
  int64_t* counts;
};

__global__ void __launch_bounds__(256) batch_compact_kernel(CompactParams P) {
  __shared__ int64_t s_item;
  __shared__ uint32_t wave_cnt[4];
  __shared__ unsigned long long s_cnt, s_base;
  const uint64_t n_work = *P.n_work;
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  for (;;) {
    if (threadIdx.x == 0) {
      const unsigned long long w = atomicAdd(P.work_ctr, 1ull);
      s_item = w < n_work ? (int64_t)w : -1;
      s_cnt = 0;
    }
    __syncthreads();
    const int64_t item = s_item;
    if (item < 0) break;
    const CWork W = P.work[item];
    const BQ q = P.qs[W.q];
    auto live_at = [&](uint64_t i) -> bool {
      if (i >= W.end) return false;
      const uint64_t* s = q.table + i * kBStride;
      return i == q.mask + 1 ? s[1] != 0 : s[0] != kEmpty;
    };
    // count, reserve the item's output range with one atomic, then place
    uint32_t mine = 0;
    for (uint64_t i = W.beg + threadIdx.x; i < W.end; i += blockDim.x) mine += live_at(i);
    if (mine) atomicAdd(&s_cnt, (unsigned long long)mine);
    __syncthreads();
    if (threadIdx.x == 0) s_base = s_cnt ? atomicAdd(&P.ctr[W.q].out_n, s_cnt) : 0;
    __syncthreads();
    uint64_t run = W.out_off + s_base;
    for (uint64_t i0 = W.beg; i0 < W.end; i0 += blockDim.x) {
      const uint64_t i = i0 + threadIdx.x;
      const bool live = live_at(i);
      const uint64_t b = __ballot(live);
      if (lane == 0) wave_cnt[wave] = (uint32_t)__popcll(b);
      __syncthreads();
      uint32_t before = 0, tot = 0;
      for (int w = 0; w < 4; ++w) {
        if (w < wave) before += wave_cnt[w];
        tot += wave_cnt[w];
      }
      if (live) {
        const uint64_t pos = run + before + __popcll(b & ((1ull << lane) - 1));
        const uint64_t* s = q.table + i * kBStride;
        P.keys[pos] = (int64_t)s[0];
        const unsigned __int128 acc = (unsigned __int128)s[2] + ((unsigned __int128)(__int128)(int64_t)s[3] << 32);
        P.sum_lo[pos] = (uint64_t)acc;
        P.sum_hi[pos] = (int64_t)(uint64_t)(acc >> 64);
        P.counts[pos] = (int64_t)s[1];
      }
      run += tot;
      __syncthreads();
    }
  }
}

__global__ void batch_fill_kernel(uint64_t* arena, const uint64_t* n_words, BCounters* ctr, const uint64_t* n_q,
                                  unsigned long long* work_ctrs) {
  const uint64_t n = *n_words;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    arena[i] = (i % kBStride) == 0 ? kEmpty : 0;
  const uint64_t nq = *n_q;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nq; i += (uint64_t)gridDim.x * blockDim.x)
    ctr[i] = BCounters{0, 0, 0, 0};
  if (blockIdx.x == 0 && threadIdx.x < 2) work_ctrs[threadIdx.x] = 0;
}

// Every query's compacted results packed densely for one copy to the host:
// query i's n_i groups land at dst[i] as four runs of n_i words (keys, sum lo,
// sum hi, counts); an integer SUM keeps only its wrapped low word, so its hi
// word is the low word's sign (scatterSumInt).
struct PackItem {
  uint64_t src_off;  // the query's output region in the arena's outputs
  uint64_t n;
  uint64_t dst;      // word offset of its four runs in the packed buffer
  uint32_t wrap;     // 1: integer SUM
  uint32_t pad;
};

__global__ void batch_pack_kernel(const PackItem* items, uint32_t n_items, const int64_t* keys, const uint64_t* lo,
                                  const int64_t* hi, const int64_t* counts, uint64_t* out) {
  for (uint32_t q = blockIdx.y; q < n_items; q += gridDim.y) {
    const PackItem it = items[q];
    uint64_t* d = out + it.dst;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < it.n; i += (uint64_t)gridDim.x * blockDim.x) {
      const uint64_t l = lo[it.src_off + i];
      d[i] = (uint64_t)keys[it.src_off + i];
      d[it.n + i] = l;
      d[2 * it.n + i] = it.wrap ? ((int64_t)l < 0 ? ~0ull : 0ull) : (uint64_t)hi[it.src_off + i];
      d[3 * it.n + i] = (uint64_t)counts[it.src_off + i];
    }
  }
}

}  // namespace wg

// ================================================================ host side

namespace wg {

// One set of device buffers a batch is planned into: the table arena, the
// descriptors, the work queues and the outputs. A wg_batch owns one for the
// batched launch and one ("solo") for single-query launches.
struct BatchArena {
  uint64_t slots = 0;  // table slots the arena holds
  uint64_t* arena = nullptr;
  // the descriptors, one device block uploaded with one copy from a staging
  // area of the same layout: [scalars: 64 bytes][queries][compact items][work items]
  void* d_desc = nullptr;
  uint64_t q_cap = 0, cwork_cap = 0, work_cap = 0;
  uint64_t* d_scalars = nullptr;  // [0] n_work, [1] n_cwork, [2] arena words, [3] n_q
  BQ* d_qs = nullptr;
  CWork* d_cwork = nullptr;
  BWork* d_work = nullptr;
  unsigned long long* d_work_ctrs = nullptr;
  BCounters* d_ctr = nullptr;
  int64_t* d_keys = nullptr;
  uint64_t* d_sum_lo = nullptr;
  int64_t* d_sum_hi = nullptr;
  int64_t* d_counts = nullptr;
  void* pinned = nullptr;  // descriptor staging + counter read-back
  uint64_t pinned_bytes = 0;
  PackItem* pack_items = nullptr;  // pinned: the pack table of the last download
  PackItem* d_pack_items = nullptr;
  uint64_t pack_cap = 0;
  uint64_t* d_packed = nullptr;  // packed results of the last download
  uint64_t packed_cap = 0;
  // A captured graph holds the device pointers above as kernel arguments:
  // every reallocation bumps `gen`, and a graph of an older generation is
  // re-captured before it is replayed.
  uint64_t gen = 0;
  hipGraphExec_t graph = nullptr;
  int32_t graph_block = 0;
  uint32_t graph_lds = 0;
  uint64_t graph_gen = ~0ull;
};

struct Plan {
  std::vector<BQ> qs;
  std::vector<uint64_t> out_off;
  std::vector<int32_t> use_lds;
  std::vector<uint32_t> fanout;
  std::vector<BWork> work;
  std::vector<CWork> cwork;
  uint64_t tile_rows = 0;  // rows per work item this batch uses
  uint64_t arena_words = 0;
  uint32_t lds_bytes = 0;
};

}  // namespace wg

struct wg_batch {
  wg_ctx* ctx = nullptr;
  int32_t max_q = 0;
  wg::BatchArena main, solo;
  wg::Plan plan;
  // last run: per query group counts, whether its result sits in fb_*
  std::vector<uint64_t> n_groups;
  std::vector<uint8_t> fallback;
  std::vector<std::vector<int64_t>> fb_keys, fb_hi, fb_cnt;
  std::vector<std::vector<uint64_t>> fb_lo;
  std::vector<wg_batch_query> queries;
  // in-flight asynchronous run
  bool in_flight = false;
  hipEvent_t done = nullptr;
  // tuning: LDS for queries with est_groups <= lds_threshold, workgroup size
  int32_t block = 512;
  uint32_t lds_threshold = 1024;
  uint32_t lds_max_slots = 2048;
  uint64_t tile_rows = 16384;
  uint32_t max_fanout = 1;  // key partitions per tile a query above the threshold may use
  uint32_t lds_cache = 1;   // queries still above the threshold keep an LDS table as a cache of frequent keys
  bool tuned = false;
};

namespace wg {

static_assert(alignof(BQ) <= 8 && alignof(CWork) <= 8 && alignof(BWork) <= 8 && sizeof(BQ) % 8 == 0 &&
                  sizeof(CWork) % 8 == 0,
              "descriptor block layout");
static uint64_t desc_off_cwork(const BatchArena& A) { return 64 + A.q_cap * sizeof(BQ); }
static uint64_t desc_off_work(const BatchArena& A) { return desc_off_cwork(A) + A.cwork_cap * sizeof(CWork); }

// A descriptor block for nq queries, nw work items and nc compact items.
static wg_status ensure_desc(wg_ctx* ctx, BatchArena& A, uint64_t nq, uint64_t nw, uint64_t nc) {
  if (A.d_desc && nq <= A.q_cap && nw <= A.work_cap && nc <= A.cwork_cap) return WG_OK;
  BatchArena G = A;  // the capacities that grow double at least
  if (nq > G.q_cap) G.q_cap = std::max<uint64_t>(nq, 2 * G.q_cap);
  if (nw > G.work_cap) G.work_cap = std::max<uint64_t>(nw, 2 * G.work_cap);
  if (nc > G.cwork_cap) G.cwork_cap = std::max<uint64_t>(nc, 2 * G.cwork_cap);
  void* fresh = nullptr;
  WG_TRY(dev_alloc(ctx, &fresh, desc_off_work(G) + G.work_cap * sizeof(BWork), nullptr, "batch descriptors"));
  if (A.d_desc) (void)hipFree(A.d_desc);
  A.q_cap = G.q_cap;
  A.work_cap = G.work_cap;
  A.cwork_cap = G.cwork_cap;
  A.d_desc = fresh;
  A.d_scalars = (uint64_t*)fresh;
  A.d_qs = (BQ*)((char*)fresh + 64);
  A.d_cwork = (CWork*)((char*)fresh + desc_off_cwork(A));
  A.d_work = (BWork*)((char*)fresh + desc_off_work(A));
  ++A.gen;  // invalidates any graph that captured the old pointers
  return WG_OK;
}

static void arena_free(BatchArena& A);

static wg_status arena_alloc(wg_ctx* ctx, BatchArena& A, uint64_t slots, uint64_t max_q) {
  WG_ALLOC(ctx, &A.arena, (slots + 1) * kBStride * 8, "the batch arena");
  WG_ALLOC(ctx, &A.d_work_ctrs, 64, "batch counters");
  WG_ALLOC(ctx, &A.d_keys, slots * 8, "batch outputs");
  WG_ALLOC(ctx, &A.d_sum_lo, slots * 8, "batch outputs");
  WG_ALLOC(ctx, &A.d_sum_hi, slots * 8, "batch outputs");
  WG_ALLOC(ctx, &A.d_counts, slots * 8, "batch outputs");
  WG_TRY(ensure_desc(ctx, A, max_q, 1, 1));
  WG_ALLOC(ctx, &A.d_ctr, max_q * sizeof(BCounters), "batch counters");
  A.slots = slots;
  return WG_OK;
}

// An arena of `slots` table slots for up to max_q queries, all or nothing.
static wg_status arena_init(wg_ctx* ctx, BatchArena& A, uint64_t slots, uint64_t max_q) {
  const wg_status st = arena_alloc(ctx, A, slots, max_q);
  if (st != WG_OK) arena_free(A);
  return st;
}

static void arena_free(BatchArena& A) {
  if (A.graph) (void)hipGraphExecDestroy(A.graph);
  void* ps[] = {A.arena, A.d_desc, A.d_work_ctrs, A.d_ctr,
                A.d_keys, A.d_sum_lo, A.d_sum_hi, A.d_counts, A.d_pack_items, A.d_packed};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  if (A.pinned) (void)hipHostFree(A.pinned);
  if (A.pack_items) (void)hipHostFree(A.pack_items);
  A = BatchArena{};
}

constexpr uint64_t kLdsItemRows = 32768;

static uint64_t table_cap_for(uint64_t est) { return next_pow2(std::max<uint64_t>(std::max<uint64_t>(est, 8) * 2, 64)); }

static wg_status plan_into(const wg_batch* b, const wg_batch_query* qs, int32_t n, const BatchArena& A, Plan& P) {
  P.qs.assign(n, BQ{});
  P.out_off.assign(n, 0);
  P.use_lds.assign(n, 0);
  P.fanout.assign(n, 1);
  uint64_t slot_off = 0;
  uint32_t max_lds = 0;
  for (int i = 0; i < n; ++i) {
    const wg_batch_query& q = qs[i];
    WG_CHECK(is_int_class(q.key.type), WG_ERR_UNSUPPORTED, "query %d: key type %d", i, q.key.type);
    WG_CHECK(q.val.type == WG_INT64 || q.val.type == WG_DECIMAL || q.val.type == WG_INT32, WG_ERR_UNSUPPORTED,
             "query %d: value type %d", i, q.val.type);
    WG_CHECK(q.row_begin + q.n_rows < (1ull << 40), WG_ERR_INVALID, "query %d: row range", i);
    WG_CHECK(q.n_rows == 0 || (q.key.data && q.val.data), WG_ERR_INVALID, "query %d: null column", i);
    const uint64_t est = std::max<uint64_t>(q.est_groups, 8);
    const uint64_t cap = table_cap_for(est);
    BQ& d = P.qs[i];
    d.key = ColDesc{q.key.data, q.key.validity, q.key.type, 0};
    d.val = ColDesc{q.val.data, q.val.validity, q.val.type, q.val.scale};
    d.range = ColDesc{q.range_col.data, q.range_col.validity, q.range_col.type, 0};
    d.range_lo = q.range_lo;
    d.range_hi = q.range_hi;
    if (q.range_col.data) WG_CHECK(is_int_class(q.range_col.type), WG_ERR_UNSUPPORTED, "query %d: range column type", i);
    d.dim = JoinTable{nullptr, 0, 0, nullptr};
    d.dim_word = -1;
    if (q.dim) {
      WG_CHECK(q.dim->finished, WG_ERR_STATE, "query %d: dimension join not built", i);
      WG_CHECK(q.dim_payload >= 0 && q.dim_payload < q.dim->n_inline, WG_ERR_INVALID, "query %d: dim payload %d", i,
               q.dim_payload);
      WG_CHECK(q.dim->claimed_unique, WG_ERR_UNSUPPORTED, "query %d: dimension keys must be unique", i);
      d.dim = JoinTable{q.dim->slots, q.dim->cap - 1, q.dim->stride, q.dim->next};
      d.dim_word = 2 + q.dim_payload;
    }
    d.mask = cap - 1;
    d.max_probe = kMaxProbe;
    d.wrap = q.val.type != WG_DECIMAL ? 1u : 0u;
    P.out_off[i] = slot_off;  // output region = the query's slot range
    slot_off += cap + 1;
    // an LDS table for queries up to the threshold; above it, a query whose
    // groups split over at most max_fanout partitions of at most the
    // threshold each (and that is not a join: its group key needs the
    // dimension lookup first) fans out into that many items per tile
    uint32_t fan_bits = 0;
    if (b->lds_threshold > 0 && est > b->lds_threshold && !q.dim)
      while ((1u << fan_bits) < b->max_fanout && (est >> fan_bits) > b->lds_threshold) ++fan_bits;
    const uint64_t per_item = est >> fan_bits;
    const bool lds = b->lds_threshold > 0 && per_item <= b->lds_threshold;
    if (!lds) fan_bits = 0;
    d.fan_bits = fan_bits;
    P.use_lds[i] = lds ? 1 : b->lds_cache ? 2 : 0;
    P.fanout[i] = 1u << fan_bits;
    // an LDS table that cannot hold the query's groups still catches its
    // frequent keys (they claim slots first): rows of keys that find no slot
    // go to the query's table, the cached partials are flushed once per item
    d.lds_slots = lds ? (uint32_t)std::min<uint64_t>(table_cap_for(per_item), b->lds_max_slots)
                      : b->lds_cache ? b->lds_max_slots : 0;
    max_lds = std::max(max_lds, d.lds_slots);
  }
  WG_CHECK(slot_off <= A.slots, WG_ERR_OOM, "batch needs %llu table slots, the arena holds %llu",
           (unsigned long long)slot_off, (unsigned long long)A.slots);
  for (int i = 0; i < n; ++i) P.qs[i].table = A.arena + P.out_off[i] * kBStride;
  P.arena_words = slot_off * kBStride;
  P.lds_bytes = max_lds * kBStride * 8;
  // rows per work item: the tuned size, halved (down to 8K rows) until the
  // batch has at least four items per compute unit, so a batch of a few
  // queries still spreads over the whole device
  uint64_t rows_items = 0;
  for (int i = 0; i < n; ++i) rows_items += qs[i].n_rows * P.fanout[i];
  uint64_t tile = b->tile_rows;
  const uint64_t want_items = 4ull * (uint64_t)std::max(b->ctx->cus, 1);
  while (tile > 8192 && rows_items / tile < want_items) tile >>= 1;
  P.tile_rows = tile;
  // tiles, largest queries first so their tails do not trail the batch
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return qs[a].n_rows > qs[c].n_rows; });
  P.work.clear();
  for (int i : order) {
    const uint64_t r0 = qs[i].row_begin, r1 = qs[i].row_begin + qs[i].n_rows;
    // a query aggregated in an LDS table (no dimension lookup) flushes each
    // item's table into its global table with device atomics on the same
    // few slots: its items take at least kLdsItemRows rows, so the flushes
    // stay few even when the batch is too small to fill the device
    const uint64_t ti = P.use_lds[i] == 1 && !qs[i].dim ? std::max<uint64_t>(tile, kLdsItemRows) : tile;
    for (uint64_t r = r0; r < r1; r += ti)
      for (uint32_t f = 0; f < P.fanout[i]; ++f)  // a tile's partitions back to back: its rows stay in cache
        P.work.push_back(BWork{(uint32_t)i, f, r, std::min<uint64_t>(r1, r + ti)});
  }
  P.cwork.clear();
  for (int i = 0; i < n; ++i) {
    const uint64_t slots = P.qs[i].mask + 2;
    for (uint64_t s0 = 0; s0 < slots; s0 += 4096)
      P.cwork.push_back(CWork{(uint32_t)i, 0, s0, std::min<uint64_t>(slots, s0 + 4096), P.out_off[i]});
  }
  return WG_OK;
}

static wg_status upload(wg_ctx* ctx, BatchArena& A, const Plan& P) {
  const int n = (int)P.qs.size();
  WG_TRY(ensure_desc(ctx, A, std::max(n, 1), std::max<size_t>(P.work.size(), 1), std::max<size_t>(P.cwork.size(), 1)));
  // staging: the descriptor block's layout up to the last work item, then the counters' read-back
  const uint64_t desc_bytes = desc_off_work(A) + P.work.size() * sizeof(BWork);
  const uint64_t bytes = desc_bytes + 64 + n * sizeof(BCounters);
  if (bytes > A.pinned_bytes) {
    if (A.pinned) WG_HIP(hipHostFree(A.pinned));
    A.pinned = nullptr;
    A.pinned_bytes = 0;
    const uint64_t want = std::max<uint64_t>(bytes * 2, 1 << 20);
    WG_CHECK(hipHostMalloc(&A.pinned, want, hipHostMallocDefault) == hipSuccess, WG_ERR_OOM,
             "pinned staging of %llu bytes failed", (unsigned long long)want);
    A.pinned_bytes = want;
  }
  // the previous use of this arena has completed, so its staging is free
  char* p = (char*)A.pinned;
  const uint64_t sc[4] = {P.work.size(), P.cwork.size(), P.arena_words, (uint64_t)n};
  memcpy(p, sc, sizeof sc);
  memcpy(p + 64, P.qs.data(), n * sizeof(BQ));
  memcpy(p + desc_off_cwork(A), P.cwork.data(), P.cwork.size() * sizeof(CWork));
  memcpy(p + desc_off_work(A), P.work.data(), P.work.size() * sizeof(BWork));
  WG_HIP(hipMemcpyAsync(A.d_desc, p, desc_bytes, hipMemcpyHostToDevice, ctx->stream));
  return WG_OK;
}

static BCounters* counters_staging(BatchArena& A, const Plan& P) {
  char* p = (char*)A.pinned + (desc_off_work(A) + P.work.size() * sizeof(BWork) + 63) / 64 * 64;
  return (BCounters*)p;
}

static void enqueue(wg_ctx* ctx, BatchArena& A, int32_t block, uint32_t lds_bytes) {
  hipLaunchKernelGGL(batch_fill_kernel, dim3(ctx->cus * 4), dim3(256), 0, ctx->stream, A.arena,
                     (const uint64_t*)A.d_scalars + 2, A.d_ctr, (const uint64_t*)A.d_scalars + 3, A.d_work_ctrs);
  BatchParams P{A.d_qs, A.d_work, A.d_scalars, A.d_work_ctrs, A.d_ctr};
  const int per_cu = std::max(1, std::min<int>(lds_bytes ? (160 * 1024) / (int)lds_bytes : 8, 2048 / block));
  hipLaunchKernelGGL(batch_agg_kernel, dim3(ctx->cus * per_cu), dim3(block), lds_bytes, ctx->stream, P);
  CompactParams C{A.d_qs, A.d_cwork, A.d_scalars + 1, A.d_work_ctrs + 1, A.d_ctr,
                  A.d_keys, A.d_sum_lo, A.d_sum_hi, A.d_counts};
  hipLaunchKernelGGL(batch_compact_kernel, dim3(ctx->cus * 4), dim3(256), 0, ctx->stream, C);
}

static wg_status launch(wg_ctx* ctx, BatchArena& A, int32_t block, uint32_t lds_bytes, bool use_graph) {
  static bool attr_done = false;
  if (!attr_done) {
    WG_HIP(hipFuncSetAttribute((const void*)batch_agg_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
    attr_done = true;
  }
  if (use_graph) {
    // the graph is specialised on (workgroup size, LDS bytes) and holds the
    // descriptor pointers of its generation: re-capture when any changed
    if (!A.graph || A.graph_block != block || A.graph_lds != lds_bytes || A.graph_gen != A.gen) {
      if (A.graph) WG_HIP(hipGraphExecDestroy(A.graph));
      A.graph = nullptr;
      hipGraph_t g;
      WG_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      enqueue(ctx, A, block, lds_bytes);
      WG_HIP(hipStreamEndCapture(ctx->stream, &g));
      WG_HIP(hipGraphInstantiate(&A.graph, g, nullptr, nullptr, 0));
      WG_HIP(hipGraphDestroy(g));
      A.graph_block = block;
      A.graph_lds = lds_bytes;
      A.graph_gen = A.gen;
    }
    WG_HIP(hipGraphLaunch(A.graph, ctx->stream));
  } else {
    enqueue(ctx, A, block, lds_bytes);
    WG_HIP(hipGetLastError());
  }
  return WG_OK;
}

static void reset_results(wg_batch* b, int n) {
  b->n_groups.assign(n, 0);
  b->fallback.assign(n, 0);
  b->fb_keys.assign(n, {});
  b->fb_lo.assign(n, {});
  b->fb_hi.assign(n, {});
  b->fb_cnt.assign(n, {});
}

static void wrap_int_sums(const wg_batch_query& q, uint64_t n, const uint64_t* lo, int64_t* hi) {
  // INT64 / INT32 values: SUM wraps at 64 bits (scatterSumInt): keep the low word
  if (q.val.type != WG_DECIMAL)
    for (uint64_t g = 0; g < n; ++g) hi[g] = (int64_t)lo[g] < 0 ? -1 : 0;
}

// One query alone in its own launch (fill + aggregate + compact), synchronous,
// with the solo arena grown to fit `est` groups; results land in fb_*[i].
static wg_status run_solo(wg_batch* b, int i, const wg_batch_query& q0, uint64_t est) {
  wg_batch_query q = q0;
  q.est_groups = est;
  const uint64_t need = table_cap_for(est) + 1;
  if (b->solo.slots < need) {
    arena_free(b->solo);
    WG_TRY(arena_init(b->ctx, b->solo, std::max<uint64_t>(need, 1 << 16), 1));
  }
  Plan P;
  WG_TRY(plan_into(b, &q, 1, b->solo, P));
  WG_TRY(upload(b->ctx, b->solo, P));
  WG_TRY(launch(b->ctx, b->solo, b->block, P.lds_bytes, false));
  BCounters c;
  WG_HIP(hipMemcpyAsync(&c, b->solo.d_ctr, sizeof c, hipMemcpyDeviceToHost, b->ctx->stream));
  WG_HIP(hipStreamSynchronize(b->ctx->stream));
  if (c.overflow) return run_solo(b, i, q0, std::max<uint64_t>(c.groups * 4, est * 4));
  const uint64_t n = c.out_n;
  b->fb_keys[i].resize(n);
  b->fb_lo[i].resize(n);
  b->fb_hi[i].resize(n);
  b->fb_cnt[i].resize(n);
  if (n) {
    WG_HIP(hipMemcpyAsync(b->fb_keys[i].data(), b->solo.d_keys, n * 8, hipMemcpyDeviceToHost, b->ctx->stream));
    WG_HIP(hipMemcpyAsync(b->fb_lo[i].data(), b->solo.d_sum_lo, n * 8, hipMemcpyDeviceToHost, b->ctx->stream));
    WG_HIP(hipMemcpyAsync(b->fb_hi[i].data(), b->solo.d_sum_hi, n * 8, hipMemcpyDeviceToHost, b->ctx->stream));
    WG_HIP(hipMemcpyAsync(b->fb_cnt[i].data(), b->solo.d_counts, n * 8, hipMemcpyDeviceToHost, b->ctx->stream));
    WG_HIP(hipStreamSynchronize(b->ctx->stream));
    wrap_int_sums(q, n, b->fb_lo[i].data(), b->fb_hi[i].data());
  }
  b->n_groups[i] = n;
  b->fallback[i] = 1;
  return WG_OK;
}

}  // namespace wg

using namespace wg;

// ----------------------------------------------- internal asynchronous API
// (used by the serving engine: one wg_batch per lane, each on its own stream)

wg_status wg_batch_submit_async(wg_batch* b, const wg_batch_query* qs, int32_t n, int32_t use_graph) {
  WG_CHECK(!b->in_flight, WG_ERR_STATE, "a batch is already in flight on this runner");
  WG_CHECK(n > 0 && n <= b->max_q, WG_ERR_INVALID, "%d queries (max %d)", n, b->max_q);
  reset_results(b, n);
  b->queries.assign(qs, qs + n);
  WG_TRY(plan_into(b, qs, n, b->main, b->plan));
  WG_TRY(upload(b->ctx, b->main, b->plan));
  WG_TRY(launch(b->ctx, b->main, b->block, b->plan.lds_bytes, use_graph != 0));
  WG_HIP(hipMemcpyAsync(counters_staging(b->main, b->plan), b->main.d_ctr, n * sizeof(BCounters),
                        hipMemcpyDeviceToHost, b->ctx->stream));
  if (!b->done) WG_HIP(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
  WG_HIP(hipEventRecord(b->done, b->ctx->stream));
  b->in_flight = true;
  return WG_OK;
}

// 1 when the in-flight batch finished (results readable), 0 while running.
wg_status wg_batch_poll(wg_batch* b, int32_t* finished, int32_t wait) {
  *finished = 0;
  if (!b->in_flight) {
    *finished = 1;
    return WG_OK;
  }
  hipError_t e = wait ? hipEventSynchronize(b->done) : hipEventQuery(b->done);
  if (e == hipErrorNotReady) return WG_OK;
  b->in_flight = false;  // finished, or failed: either way nothing is in flight any more
  WG_HIP(e);
  const BCounters* c = counters_staging(b->main, b->plan);
  const int n = (int)b->queries.size();
  for (int i = 0; i < n; ++i) {
    b->n_groups[i] = c[i].out_n;
    // an underestimated table: redo this query alone, exactly
    if (c[i].overflow) WG_TRY(run_solo(b, i, b->queries[i], std::max<uint64_t>(c[i].groups * 4, b->queries[i].est_groups * 4)));
  }
  *finished = 1;
  return WG_OK;
}

// 1 when query q of the last batch was re-run alone (its results on the host).
int32_t wg_batch_rerun_alone(const wg_batch* b, int32_t q) {
  return q >= 0 && q < (int)b->fallback.size() ? b->fallback[q] : 0;
}

// After wg_batch_poll reported the batch finished: enqueue on the runner's
// stream one pack of every query's results and one copy of them into `host`
// (pinned, host_words words); query i's runs start at word offsets[i] (four
// runs of n_groups[i] words: keys, sum lo, sum hi, counts). Queries that were
// re-run alone (their results already on the host) get n = 0 here and are
// read with wg_batch_result. WG_ERR_OOM (nothing enqueued) when `host` is too
// small; *words = the words needed.
wg_status wg_batch_download_async(wg_batch* b, uint64_t* host, uint64_t host_words, uint64_t* offsets,
                                  uint64_t* words) {
  WG_CHECK(!b->in_flight, WG_ERR_STATE, "batch still running");
  BatchArena& A = b->main;
  const int n = (int)b->n_groups.size();
  uint64_t total = 0;
  for (int i = 0; i < n; ++i) {
    offsets[i] = total;
    if (!b->fallback[i]) total += 4 * b->n_groups[i];
  }
  *words = total;
  WG_CHECK(total <= host_words, WG_ERR_OOM, "results need %llu words, the buffer holds %llu",
           (unsigned long long)total, (unsigned long long)host_words);
  if (total == 0) return WG_OK;
  if (A.pack_cap < (uint64_t)n) {
    if (A.pack_items) WG_HIP(hipHostFree(A.pack_items));
    if (A.d_pack_items) WG_HIP(hipFree(A.d_pack_items));
    A.pack_items = nullptr;
    A.d_pack_items = nullptr;
    A.pack_cap = 0;
    WG_HIP(hipHostMalloc((void**)&A.pack_items, n * sizeof(PackItem), hipHostMallocDefault));
    WG_ALLOC(b->ctx, &A.d_pack_items, n * sizeof(PackItem), "batch result packing");
    A.pack_cap = n;
  }
  if (A.packed_cap < total) {
    if (A.d_packed) WG_HIP(hipFree(A.d_packed));
    A.d_packed = nullptr;
    A.packed_cap = 0;
    WG_ALLOC(b->ctx, &A.d_packed, total * 8, "batch result packing");
    A.packed_cap = total;
  }
  // the pack table of an earlier download is no longer read: downloads on this
  // stream run in order, and the caller waits for each before the next
  for (int i = 0; i < n; ++i)
    A.pack_items[i] = PackItem{b->plan.out_off[i], b->fallback[i] ? 0 : b->n_groups[i], offsets[i],
                               b->queries[i].val.type != WG_DECIMAL ? 1u : 0u, 0};
  WG_HIP(hipMemcpyAsync(A.d_pack_items, A.pack_items, n * sizeof(PackItem), hipMemcpyHostToDevice, b->ctx->stream));
  hipLaunchKernelGGL(batch_pack_kernel, dim3(64, std::min(n, 1024)), dim3(256), 0, b->ctx->stream,
                     (const PackItem*)A.d_pack_items, (uint32_t)n, (const int64_t*)A.d_keys, (const uint64_t*)A.d_sum_lo,
                     (const int64_t*)A.d_sum_hi, (const int64_t*)A.d_counts, A.d_packed);
  WG_HIP(hipGetLastError());
  WG_HIP(hipMemcpyAsync(host, A.d_packed, total * 8, hipMemcpyDeviceToHost, b->ctx->stream));
  return WG_OK;
}

extern "C" {

wg_status wg_batch_create(wg_ctx* ctx, int32_t max_queries, uint64_t max_groups_total, wg_batch** out) {
  clear_error();
  WG_CHECK(ctx && out, WG_ERR_INVALID, "null argument");
  WG_CHECK(max_queries > 0 && max_queries <= (1 << 20), WG_ERR_INVALID, "max_queries %d", max_queries);
  wg_batch* b = new wg_batch();
  b->ctx = ctx;
  b->max_q = max_queries;
  // each query's table is at most 4x its estimate (power of two of 2x), + side slot
  const uint64_t slots = std::max<uint64_t>(max_groups_total * 4 + (uint64_t)max_queries * 130, 1 << 16);
  wg_status st = arena_init(ctx, b->main, slots, max_queries);
  if (st != WG_OK) {
    wg_batch_destroy(b);
    return st;
  }
  *out = b;
  return WG_OK;
}

wg_status wg_batch_destroy(wg_batch* b) {
  if (!b) return WG_OK;
  (void)hipStreamSynchronize(b->ctx->stream);
  arena_free(b->main);
  arena_free(b->solo);
  if (b->done) (void)hipEventDestroy(b->done);
  delete b;
  return WG_OK;
}

wg_status wg_batch_run(wg_batch* b, const wg_batch_query* qs, int32_t n_queries, int32_t use_graph, int32_t plain) {
  clear_error();
  WG_CHECK(b && (qs || n_queries == 0), WG_ERR_INVALID, "null argument");
  WG_CHECK(n_queries >= 0 && n_queries <= b->max_q, WG_ERR_INVALID, "%d queries (max %d)", n_queries, b->max_q);
  WG_CHECK(!b->in_flight, WG_ERR_STATE, "a batch is in flight");
  if (n_queries == 0) {
    reset_results(b, 0);
    return WG_OK;
  }
  if (plain) {
    // one launch sequence per query, serialised: the baseline the batch beats
    reset_results(b, n_queries);
    b->queries.assign(qs, qs + n_queries);
    b->plan.use_lds.assign(n_queries, 0);
    b->plan.fanout.assign(n_queries, 1);
    for (int i = 0; i < n_queries; ++i) {
      WG_TRY(run_solo(b, i, qs[i], qs[i].est_groups));
      b->plan.use_lds[i] = b->lds_threshold > 0 && std::max<uint64_t>(qs[i].est_groups, 8) <= b->lds_threshold ? 1
                           : b->lds_cache ? 2 : 0;
    }
    return WG_OK;
  }
  WG_TRY(wg_batch_submit_async(b, qs, n_queries, use_graph));
  int32_t fin = 0;
  return wg_batch_poll(b, &fin, 1);
}

wg_status wg_batch_result(wg_batch* b, int32_t q, uint64_t* n_groups, int64_t* keys, uint64_t* sum_lo, int64_t* sum_hi,
                          int64_t* counts, uint64_t cap) {
  clear_error();
  WG_CHECK(b && n_groups, WG_ERR_INVALID, "null argument");
  WG_CHECK(!b->in_flight, WG_ERR_STATE, "results of a batch still in flight");
  WG_CHECK(q >= 0 && q < (int)b->n_groups.size(), WG_ERR_INVALID, "query %d", q);
  const uint64_t n = b->n_groups[q];
  *n_groups = n;
  if (cap < n || n == 0) return WG_OK;
  if (b->fallback[q]) {
    if (keys) memcpy(keys, b->fb_keys[q].data(), n * 8);
    if (sum_lo) memcpy(sum_lo, b->fb_lo[q].data(), n * 8);
    if (sum_hi) memcpy(sum_hi, b->fb_hi[q].data(), n * 8);
    if (counts) memcpy(counts, b->fb_cnt[q].data(), n * 8);
    return WG_OK;
  }
  const uint64_t off = b->plan.out_off[q];
  std::vector<uint64_t> lo_tmp;
  uint64_t* lo = sum_lo;
  if (!lo && sum_hi) {
    lo_tmp.resize(n);
    lo = lo_tmp.data();
  }
  if (keys) WG_HIP(hipMemcpyAsync(keys, b->main.d_keys + off, n * 8, hipMemcpyDeviceToHost, b->ctx->stream));
  if (lo) WG_HIP(hipMemcpyAsync(lo, b->main.d_sum_lo + off, n * 8, hipMemcpyDeviceToHost, b->ctx->stream));
  if (sum_hi) WG_HIP(hipMemcpyAsync(sum_hi, b->main.d_sum_hi + off, n * 8, hipMemcpyDeviceToHost, b->ctx->stream));
  if (counts) WG_HIP(hipMemcpyAsync(counts, b->main.d_counts + off, n * 8, hipMemcpyDeviceToHost, b->ctx->stream));
  WG_HIP(hipStreamSynchronize(b->ctx->stream));
  if (sum_hi) wrap_int_sums(b->queries[q], n, lo, sum_hi);
  return WG_OK;
}

wg_status wg_batch_choice(wg_batch* b, int32_t q, int32_t* use_lds, int32_t* block, int32_t* fanout) {
  WG_CHECK(b, WG_ERR_INVALID, "null handle");
  WG_CHECK(q >= 0 && q < (int)b->plan.use_lds.size(), WG_ERR_INVALID, "query %d", q);
  if (use_lds) *use_lds = b->plan.use_lds[q];
  if (block) *block = b->block;
  if (fanout) *fanout = q < (int)b->plan.fanout.size() ? (int32_t)b->plan.fanout[q] : 1;
  return WG_OK;
}

// Autotune: time (workgroup size x LDS threshold x tile size x partition
// fan-out) candidates on these queries with HIP events and keep the fastest.
wg_status wg_batch_autotune(wg_batch* b, const wg_batch_query* qs, int32_t n_queries) {
  clear_error();
  WG_CHECK(b && qs, WG_ERR_INVALID, "null argument");
  WG_CHECK(n_queries > 0 && n_queries <= b->max_q, WG_ERR_INVALID, "%d queries", n_queries);
  WG_CHECK(!b->in_flight, WG_ERR_STATE, "a batch is in flight");
  struct Cand { int32_t block; uint32_t thr; uint64_t tile; uint32_t fan; uint32_t cache; };
  std::vector<Cand> cands;
  for (int32_t blk : {256, 512, 1024})
    for (uint32_t thr : {0u, 256u, 1024u, 2048u})
      for (uint64_t tile : {8192ull, 32768ull, 65536ull})
        for (uint32_t fan : {1u, 4u, 16u, 64u})
          for (uint32_t cache : {0u, 1u})
            if (thr > 0 || (fan == 1 && cache == 0)) cands.push_back(Cand{blk, thr, tile, fan, cache});
  hipEvent_t e0, e1;
  WG_HIP(hipEventCreate(&e0));
  WG_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  Cand best_c = cands[0];
  Plan P;
  for (const Cand& c : cands) {
    b->block = c.block;
    b->lds_threshold = c.thr;
    b->tile_rows = c.tile;
    b->max_fanout = c.fan;
    b->lds_cache = c.cache;
    WG_TRY(plan_into(b, qs, n_queries, b->main, P));
    float t = 0.f;
    for (int rep = 0; rep < 3; ++rep) {
      WG_TRY(upload(b->ctx, b->main, P));
      WG_HIP(hipEventRecord(e0, b->ctx->stream));
      WG_TRY(launch(b->ctx, b->main, b->block, P.lds_bytes, false));
      WG_HIP(hipEventRecord(e1, b->ctx->stream));
      WG_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      WG_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (rep > 0) t += ms;  // the first rep warms caches and code objects
    }
    if (t < best) {
      best = t;
      best_c = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->block = best_c.block;
  b->lds_threshold = best_c.thr;
  b->tile_rows = best_c.tile;
  b->max_fanout = best_c.fan;
  b->lds_cache = best_c.cache;
  b->tuned = true;
  return WG_OK;
}

wg_status wg_batch_tuning(wg_batch* b, int32_t* block, int32_t* lds_threshold, uint64_t* tile_rows,
                          int32_t* max_fanout, int32_t* lds_cache) {
  WG_CHECK(b, WG_ERR_INVALID, "null handle");
  if (block) *block = b->block;
  if (lds_threshold) *lds_threshold = (int32_t)b->lds_threshold;
  if (tile_rows) *tile_rows = b->tile_rows;
  if (max_fanout) *max_fanout = (int32_t)b->max_fanout;
  if (lds_cache) *lds_cache = (int32_t)b->lds_cache;
  return WG_OK;
}

wg_status wg_batch_set_tuning(wg_batch* b, int32_t block, int32_t lds_threshold, uint64_t tile_rows,
                              int32_t max_fanout, int32_t lds_cache) {
  clear_error();
  WG_CHECK(b, WG_ERR_INVALID, "null handle");
  WG_CHECK(block == 256 || block == 512 || block == 1024, WG_ERR_INVALID, "block %d", block);
  WG_CHECK(lds_threshold >= 0 && lds_threshold <= 2048, WG_ERR_INVALID, "lds_threshold %d", lds_threshold);
  WG_CHECK(tile_rows >= 1024 && tile_rows <= (1u << 24), WG_ERR_INVALID, "tile_rows");
  WG_CHECK(max_fanout >= 1 && max_fanout <= 64 && (max_fanout & (max_fanout - 1)) == 0, WG_ERR_INVALID,
           "max_fanout %d (a power of two <= 64)", max_fanout);
  b->block = block;
  b->lds_threshold = (uint32_t)lds_threshold;
  b->tile_rows = tile_rows;
  b->max_fanout = (uint32_t)max_fanout;
  b->lds_cache = lds_cache ? 1u : 0u;
  return WG_OK;
}

}  // extern "C"
