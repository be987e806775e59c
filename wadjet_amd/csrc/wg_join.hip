// wg_join.hip — inner hash join on one int-class key: build, bounded
// materializing probe, fused probe -> grouped aggregate, gathers and the
// order-free output digest.
//
// Reference: HashJoin.Build flat path (internal/engine/exec/join.go:771-1162,
// arenaAppendInt :551-560), inlineIntProbe / resumeIntChain (:2670-2836),
// MaxProbeOutputRows (:2044-2060), gatherBuildVector (:4024), gatherVector
// (sort.go:611).
//
// Table slot (power-of-two words):
//   word 0   key (EMPTY sentinel = free)
//   word 1   low 32 bits: head build row of the key's chain (-1 = none)
//            high 32 bits: number of build rows with the key
//   word 2.. inline payload: build columns copied into the slot so a probe of
//            a unique key reads key and payload from one cache line
// Chains live in next[build_row] (newest first, like arenaNext). A slot's
// inline payload is the payload of its first build row; probes use it only
// when the key has exactly one build row and walk the chain otherwise.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "wg_agg_dev.hpp"
#include "wg_agg_host.hpp"
#include "wg_join_agg.hpp"
#include "wg_join_dev.hpp"
#include "wg_join_host.hpp"
#include "wg_select.hpp"

namespace wg {

__global__ void join_fill_kernel(uint64_t* slots, uint64_t n_words, int stride) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n_words; i += (uint64_t)gridDim.x * blockDim.x) {
    const int w = (int)(i % stride);
    slots[i] = w == 0 ? kEmpty : (w == 1 ? 0x00000000FFFFFFFFull : 0ull);
  }
}

struct BuildParams {
  ColDesc key;       // retained build key column (global build row index)
  ColDesc inl[kMaxInline];
  int32_t n_inline;
  uint64_t row0, n;  // build rows [row0, row0 + n)
  uint64_t* slots;
  uint64_t mask;
  int32_t stride;
  int32_t* next;
  uint64_t* claimed;  // slots claimed (device counter)
};

// HashJoin.Build inner loop (join.go:1051-1065): NULL keys are skipped; each
// row becomes the new head of its key's chain (arenaAppendInt, join.go:551-560).
__global__ void join_build_kernel(BuildParams B) {
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t iters = (B.n + step - 1) / step;
  uint32_t claims = 0, nonnull = 0;
  for (uint64_t it = 0; it < iters; ++it) {
    const uint64_t j = it * step + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t row = B.row0 + j;
    const bool live = j < B.n && col_valid(B.key, row);
    nonnull += live;
    if (j >= B.n) continue;
    if (!live) {
      B.next[row] = -1;
      continue;
    }
    const uint64_t key = (uint64_t)load_int(B.key, row);
    uint64_t* s = nullptr;
    if (key == kEmpty) {
      s = B.slots + (B.mask + 1) * (uint64_t)B.stride;
    } else {
      uint64_t idx = mix64(key) & B.mask;
      for (uint64_t p = 0; p <= B.mask; ++p) {
        uint64_t* c = B.slots + idx * (uint64_t)B.stride;
        uint64_t k = __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (k == kEmpty) {
          const uint64_t prev = at_cas<GlbT>(c, kEmpty, key);
          claims += prev == kEmpty;
          k = prev == kEmpty ? key : prev;
        }
        if (k == key) {
          s = c;
          break;
        }
        idx = (idx + 1) & B.mask;
      }
    }
    if (s == nullptr) continue;  // table full: the host sizes it so this never happens
    int32_t* head = (int32_t*)(s + 1);
    uint32_t* cnt = ((uint32_t*)(s + 1)) + 1;
    const int32_t old = __hip_atomic_exchange(head, (int32_t)row, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    B.next[row] = old;
    __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old == -1) {
      for (int i = 0; i < B.n_inline; ++i) s[2 + i] = (uint64_t)load_int(B.inl[i], row);
    }
  }
  // distinct keys and non-NULL rows (uniqueness check at finish), one atomic each per workgroup
  block_add(B.claimed, claims);
  block_add(B.claimed + 1, nonnull);
}

// Move a join table into a larger one (chains are per build row and stay).
__global__ void join_rehash_kernel(const uint64_t* src, uint64_t src_cap, uint64_t* dst, uint64_t dst_mask, int stride) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < src_cap; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t* s = src + i * stride;
    if (s[0] == kEmpty) continue;
    uint64_t idx = mix64(s[0]) & dst_mask;
    for (;;) {
      uint64_t* d = dst + idx * stride;
      if (at_cas<GlbT>(d, kEmpty, s[0]) == kEmpty) {
        for (int w = 1; w < stride; ++w) d[w] = s[w];
        break;
      }
      idx = (idx + 1) & dst_mask;
    }
  }
}

// ------------------------------------------------------- materializing probe

__device__ __forceinline__ const uint64_t* probe_slot(const JoinTable& J, const ColDesc& pk, uint64_t row) {
  if (!col_valid(pk, row)) return nullptr;  // NULL keys never match
  return join_find(J, (uint64_t)load_int(pk, row));
}

// The materializing probe works on tiles of kProbeTile probe positions, one
// workgroup each. Pass 1 looks every position up once and keeps its chain
// head and match count (count in the low word, head in the high word), and
// each tile's pair total; a scan over the tile totals gives every tile its
// output offset; pass 3 scans its tile's counts in LDS and writes the pairs
// from the kept heads, so the table is read once per probe row.
constexpr uint64_t kProbeTile = 4096;

__global__ void __launch_bounds__(256) join_count_kernel(JoinTable J, ColDesc pk, const uint32_t* sel, uint64_t pos0,
                                                         uint64_t n, uint64_t* hc, uint64_t* tile_cnt) {
  const uint64_t t0 = (uint64_t)blockIdx.x * kProbeTile;
  const uint64_t t1 = min(t0 + kProbeTile, n);
  uint64_t local = 0;
  constexpr int U = 4;  // positions whose home slots are read together
  struct Pair {
    uint64_t k, w;
  };
  for (uint64_t j0 = t0; j0 < t1; j0 += (uint64_t)blockDim.x * U) {
    uint64_t key[U], idx[U];
    bool live[U];
    Pair v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t j = j0 + (uint64_t)u * blockDim.x + threadIdx.x;
      live[u] = false;
      key[u] = 0;
      if (j < t1) {
        const uint64_t i = pos0 + j;
        const uint64_t row = sel ? sel[i] : i;
        live[u] = col_valid(pk, row);  // NULL keys never match
        if (live[u]) key[u] = (uint64_t)load_int(pk, row);
      }
      idx[u] = key[u] == kEmpty ? J.mask + 1 : (mix64(key[u]) & J.mask);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (live[u]) __builtin_memcpy(&v[u], J.slots + idx[u] * (uint64_t)J.stride, 16);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t j = j0 + (uint64_t)u * blockDim.x + threadIdx.x;
      if (j >= t1) continue;
      uint64_t w1 = 0;
      bool hit = false;
      if (live[u]) {
        if (key[u] == kEmpty) {
          hit = (v[u].w >> 32) != 0;  // the side slot of the sentinel key
          w1 = v[u].w;
        } else if (v[u].k == key[u]) {
          hit = true;
          w1 = v[u].w;
        } else if (v[u].k != kEmpty) {
          const uint64_t* s = join_find_from(J, key[u], (idx[u] + 1) & J.mask);  // past a collision
          hit = s != nullptr;
          if (hit) w1 = s[1];
        }
      }
      const uint32_t c = hit ? (uint32_t)(w1 >> 32) : 0u;
      const int32_t h = hit ? (int32_t)(uint32_t)w1 : -1;
      hc[j] = (uint64_t)c | ((uint64_t)(uint32_t)h << 32);
      local += c;
    }
  }
  __shared__ uint64_t part[4];
  for (int o = 32; o; o >>= 1) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)local, o);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(local >> 32), o);
    local += ((uint64_t)hi << 32) | lo;
  }
  if (__lane_id() == 0) part[threadIdx.x >> 6] = local;
  __syncthreads();
  if (threadIdx.x == 0) tile_cnt[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// First position whose pairs do not fit in `cap`. off is non-decreasing, so
// that position is the only one with off <= cap < off + cnt: one writer.
__global__ void join_cut_kernel(const uint64_t* off, const uint64_t* cnt, uint64_t n, uint64_t cap, uint64_t* cut) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x)
    if (off[j] <= cap && off[j] + cnt[j] > cap) *cut = j;
}

// Pass 3: tile b's pairs at its scanned offset, positions in order. Only
// positions whose pairs end at or before `room` write; the one position that
// straddles `room` records itself and its offset in cut[0..1].
__global__ void __launch_bounds__(256) join_emit_kernel(JoinTable J, const uint32_t* sel, uint64_t pos0, uint64_t n,
                                                        const uint64_t* hc, const uint64_t* tile_off, uint64_t room,
                                                        uint32_t* out_p, uint32_t* out_b, uint64_t* cut) {
  // 64-bit scan: 256 chains of up to 2^32 - 1 build rows each can pass 2^32
  using BlockScan = hipcub::BlockScan<uint64_t, 256>;
  __shared__ typename BlockScan::TempStorage tmp;
  const uint64_t t0 = (uint64_t)blockIdx.x * kProbeTile;
  const uint64_t t1 = min(t0 + kProbeTile, n);
  uint64_t base = tile_off[blockIdx.x];
  for (uint64_t j0 = t0; j0 < t1; j0 += blockDim.x) {
    const uint64_t j = j0 + threadIdx.x;
    const uint64_t v = j < t1 ? hc[j] : 0;
    const uint64_t c = (uint32_t)v;
    uint64_t pre = 0, total = 0;
    BlockScan(tmp).ExclusiveSum(c, pre, total);
    const uint64_t o = base + pre;
    if (c && o + c <= room) {
      const uint64_t row = sel ? sel[pos0 + j] : pos0 + j;
      int32_t b = (int32_t)(uint32_t)(v >> 32);
      for (uint64_t k = 0; k < c && b >= 0; ++k) {
        out_p[o + k] = (uint32_t)row;
        out_b[o + k] = (uint32_t)b;
        if (k + 1 < c) b = J.next[b];  // a chain's last link is not read
      }
    } else if (c && o <= room) {
      cut[0] = j;
      cut[1] = o;
    }
    base += total;
    __syncthreads();  // the scan storage is reused
  }
}

// One probe row's chain, from build row `start`, at most `limit` pairs;
// res[0] = pairs written, res[1] = build row to resume at (-1 = chain done).
__global__ void join_chain_kernel(const int32_t* next, int32_t start, uint32_t prow, uint64_t limit, uint32_t* out_p,
                                  uint32_t* out_b, int64_t* res) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  uint64_t k = 0;
  int32_t b = start;
  while (b >= 0 && k < limit) {
    out_p[k] = prow;
    out_b[k] = (uint32_t)b;
    ++k;
    b = next[b];
  }
  res[0] = (int64_t)k;
  res[1] = b;
}

__global__ void join_head_kernel(JoinTable J, ColDesc pk, uint64_t row, int64_t* res) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const uint64_t* s = probe_slot(J, pk, row);
  res[0] = s ? slot_head(s) : -1;
}

// ---------------------------------------------------------------- gathers

__global__ void gather_kernel(ColDesc c, const uint32_t* rows, uint64_t n, void* out) {
  const int w = type_bytes(c.type);
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = rows[j];
    switch (w) {
      case 2: ((uint16_t*)out)[j] = ((const uint16_t*)c.data)[r]; break;
      case 4: ((uint32_t*)out)[j] = ((const uint32_t*)c.data)[r]; break;
      case 16:
        ((uint64_t*)out)[2 * j] = ((const uint64_t*)c.data)[2 * r];
        ((uint64_t*)out)[2 * j + 1] = ((const uint64_t*)c.data)[2 * r + 1];
        break;
      default: ((uint64_t*)out)[j] = ((const uint64_t*)c.data)[r]; break;
    }
  }
}

// ----------------------------------------------------------------- digest

struct DigestParams {
  ColDesc pc[8];
  int32_t n_p;
  ColDesc bc[8];
  int32_t n_b;
};

__device__ __forceinline__ uint64_t raw64(const ColDesc& c, uint64_t row) {
  if (c.type == WG_FLOAT64) return ((const uint64_t*)c.data)[row];
  return (uint64_t)load_int(c, row);
}

// Same per-row mix as or_digest_rows (oracle/oracle.c): values in order
// (probe columns, then build columns), two independent folds, summed mod 2^64.
__global__ void join_digest_kernel(DigestParams D, const uint32_t* prow, const uint32_t* brow, uint64_t n,
                                   unsigned long long* out) {
  unsigned long long a = 0, b = 0;
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t x = 0x243F6A8885A308D3ull, y = 0x13198A2E03707344ull;
    for (int i = 0; i < D.n_p + D.n_b; ++i) {
      const uint64_t v = i < D.n_p ? raw64(D.pc[i], prow[j]) : raw64(D.bc[i - D.n_p], brow[j]);
      x = mix64(x ^ v);
      y = mix64(y + v * 0x9E3779B97F4A7C15ull);
    }
    a += x;
    b += y;
  }
  block_add((uint64_t*)out, a);
  block_add((uint64_t*)out + 1, b);
}

// ------------------------------------------------- fused probe -> aggregate

struct JoinAggParams {
  JoinTable J;
  ColDesc pkey;         // probe key column
  ColSet cs;            // joined row: probe columns, then build columns
  KeyDesc key;          // GROUP BY over joined-row columns
  AggProgram pg;
  GlobalTable T;
  const uint32_t* sel;
  uint64_t row_base;
  uint64_t n;
  uint64_t tile_rows;
  uint64_t n_tiles;
  uint32_t* ovf_rows;
  uint64_t ovf_rows_cap;
  uint64_t* ovf_part;
  int32_t lds_slots;
};

// The group slot of a joined row: LDS first, then the global table; null
// when the global table is over its budget (the probe row is replayed).
template <bool USE_LDS>
__device__ __forceinline__ uint64_t* group_slot(const JoinAggParams& P, uint64_t* lds, int L, const RowRef& r,
                                                bool& in_lds, uint32_t& claims) {
  uint64_t key;
  in_lds = false;
  if (!build_key<true>(P.key, P.cs, r, key)) return P.T.slots + (P.T.mask + 2) * (uint64_t)P.pg.stride;
  if (key == kEmpty) return P.T.slots + (P.T.mask + 1) * (uint64_t)P.pg.stride;
  const uint64_t h = key_hash(key);
  if (USE_LDS) {
    uint64_t* ls = lds_find_or_claim(lds, L - 1, key, h);
    if (ls) {
      in_lds = true;
      return ls;
    }
  }
  return glb_find_or_claim(P.T, P.pg.stride, key, h, claims);
}

// The probe row loop is software-pipelined: while row k is aggregated, the
// home slot of row k+1 and the key of row k+2 are already being loaded.
struct ProbeKey {
  uint64_t row, key, vw;
  bool live;
};
struct ProbeHome {
  ProbeKey k;
  uint64_t home;
  uint64_t w0, w1, w2, w3;  // the home slot: key, head|count, two payload words
};

__device__ __forceinline__ ProbeKey probe_key(const JoinAggParams& P, uint64_t i, uint64_t beg, uint64_t end) {
  ProbeKey r;
  r.live = i < end;
  const uint64_t ii = r.live ? i : beg;  // past the end: load a row in range, then ignore it
  r.row = P.sel ? (uint64_t)P.sel[ii] : P.row_base + ii;
  r.key = (uint64_t)load_int(P.pkey, r.row);
  r.vw = P.pkey.valid ? P.pkey.valid[r.row >> 6] : ~0ull;
  return r;
}

__device__ __forceinline__ ProbeHome probe_home(const JoinAggParams& P, const ProbeKey& k) {
  ProbeHome h;
  h.k = k;
  h.home = mix64(k.key) & P.J.mask;
  const ulonglong2* s = (const ulonglong2*)(P.J.slots + h.home * (uint64_t)P.J.stride);
  const ulonglong2 a = s[0];
  h.w0 = a.x;
  h.w1 = a.y;
  h.w2 = h.w3 = 0;
  if (P.J.stride >= 4) {
    const ulonglong2 b = s[1];
    h.w2 = b.x;
    h.w3 = b.y;
  }
  return h;
}

// Aggregate one probe row's matches (the home slot already loaded).
template <bool USE_LDS>
__device__ __forceinline__ void join_agg_row(const JoinAggParams& P, uint64_t* lds, int L, const ProbeHome& h,
                                             uint32_t& claims) {
  const ProbeKey& k = h.k;
  if (!k.live || !((k.vw >> (k.row & 63)) & 1ull)) return;  // NULL keys never match
  const uint64_t* js;
  bool at_home = false;
  if (k.key == kEmpty) {
    js = join_find(P.J, k.key);
  } else if (h.w0 == k.key) {
    js = P.J.slots + h.home * (uint64_t)P.J.stride;
    at_home = true;
  } else if (h.w0 == kEmpty) {
    return;
  } else {
    js = join_find_from(P.J, k.key, (h.home + 1) & P.J.mask);
  }
  if (!js) return;
  const uint32_t cnt = at_home ? (uint32_t)(h.w1 >> 32) : slot_count(js);
  bool in_lds;
  if (cnt == 1) {
    RowRef r{k.row, (uint64_t)(at_home ? (int32_t)(uint32_t)h.w1 : slot_head(js)), js, h.w2, h.w3};
    if (!at_home && P.J.stride >= 4) {
      r.pay2 = js[2];
      r.pay3 = js[3];
    }
    uint64_t* g = group_slot<USE_LDS>(P, lds, L, r, in_lds, claims);
    if (!g) {
      const uint64_t at = __hip_atomic_fetch_add(P.T.counters + CTR_OVF_ROWS, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      P.ovf_rows[at] = (uint32_t)k.row;
      return;
    }
    if (in_lds) apply_row<LdsT, true>(g, L, P.pg, P.cs, r);
    else apply_row<GlbT, true>(g, 1, P.pg, P.cs, r);
    return;
  }
  // several build rows: claim every group first, apply only if all fit
  bool ok = true;
  for (int32_t b = slot_head(js); b >= 0 && ok; b = P.J.next[b]) {
    const RowRef r{k.row, (uint64_t)b, nullptr};
    ok = group_slot<USE_LDS>(P, lds, L, r, in_lds, claims) != nullptr;
  }
  if (!ok) {
    const uint64_t at = __hip_atomic_fetch_add(P.T.counters + CTR_OVF_ROWS, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    P.ovf_rows[at] = (uint32_t)k.row;
    return;
  }
  for (int32_t b = slot_head(js); b >= 0; b = P.J.next[b]) {
    const RowRef r{k.row, (uint64_t)b, nullptr};
    uint64_t* g = group_slot<USE_LDS>(P, lds, L, r, in_lds, claims);
    if (in_lds) apply_row<LdsT, true>(g, L, P.pg, P.cs, r);
    else apply_row<GlbT, true>(g, 1, P.pg, P.cs, r);
  }
}

template <bool USE_LDS>
__global__ void __launch_bounds__(1024) join_agg_kernel(JoinAggParams P) {
  extern __shared__ uint64_t lds[];
  __shared__ int64_t s_tile;
  const int stride = P.pg.stride;
  const int L = USE_LDS ? P.lds_slots : 0;
  if (USE_LDS) lds_table_init(lds, L, P.pg);
  __syncthreads();
  uint32_t claims = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      int64_t t = -1;
      const uint64_t o = __hip_atomic_load(P.T.counters + CTR_OVF_ROWS, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (o + (uint64_t)gridDim.x * P.tile_rows > P.ovf_rows_cap) {
        at_add<GlbT>(P.T.counters + CTR_STOPPED, 1);
      } else {
        const uint64_t tt = __hip_atomic_fetch_add(P.T.counters + CTR_TILE, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (tt < P.n_tiles) t = (int64_t)tt;
      }
      s_tile = t;
    }
    __syncthreads();
    const int64_t tile = s_tile;
    __syncthreads();
    if (tile < 0) break;
    const uint64_t beg = (uint64_t)tile * P.tile_rows;
    const uint64_t end = min(beg + P.tile_rows, P.n);
    const uint64_t B = blockDim.x;
    uint64_t i = beg + threadIdx.x;
    ProbeKey k1 = probe_key(P, i + B, beg, end);
    ProbeHome h0 = probe_home(P, probe_key(P, i, beg, end));
    for (; i < end; i += B) {
      const ProbeKey k2 = probe_key(P, i + 2 * B, beg, end);
      const ProbeHome h1 = probe_home(P, k1);
      join_agg_row<USE_LDS>(P, lds, L, h0, claims);
      h0 = h1;
      k1 = k2;
    }
  }
  if (USE_LDS) {
    __syncthreads();
    for (int i = threadIdx.x; i < L; i += blockDim.x) {
      const uint64_t* ls = lds + i;
      const uint64_t key = ls[0];
      if (key == kEmpty) continue;
      uint64_t* slot = glb_find_or_claim(P.T, stride, key, key_hash(key), claims);
      if (slot != nullptr) {
        merge_partial(slot, ls, L, P.pg);
      } else {
        const uint64_t at = __hip_atomic_fetch_add(P.T.counters + CTR_OVF_PART, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint64_t* rec = P.ovf_part + at * (uint64_t)stride;
        for (int w = 0; w < stride; ++w) rec[w] = ls[(uint64_t)w * L];
      }
    }
  }
  block_add(P.T.counters + CTR_GROUPS, claims);
}

__global__ void agg_merge_part_kernel_j(const uint64_t* recs, uint64_t n, GlobalTable T, AggProgram pg) {
  const int stride = pg.stride;
  uint32_t claims = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t* r = recs + i * stride;
    uint64_t* d = glb_find_or_claim(T, stride, r[0], key_hash(r[0]), claims);
    if (d == nullptr) {
      at_add<GlbT>(T.counters + CTR_STOPPED, 1ull << 32);
      continue;
    }
    merge_partial(d, r, 1, pg);
  }
  block_add(T.counters + CTR_GROUPS, claims);
}

}  // namespace wg

// ================================================================ host side

namespace wg {

static wg_status join_alloc_table(wg_join* j, uint64_t cap) {
  uint64_t* s = nullptr;
  const uint64_t words = (cap + 1) * (uint64_t)j->stride;
  WG_ALLOC(j->ctx, &s, words * 8, "a join table");
  hipLaunchKernelGGL(join_fill_kernel, dim3(grid_for(words, 256, 8192)), dim3(256), 0, j->ctx->stream, s, words, j->stride);
  WG_HIP(hipGetLastError());
  if (j->slots) {
    hipLaunchKernelGGL(join_rehash_kernel, dim3(grid_for(j->cap, 256, 8192)), dim3(256), 0, j->ctx->stream,
                       (const uint64_t*)j->slots, j->cap, s, cap - 1, j->stride);
    WG_HIP(hipGetLastError());
    // the side slot (key == sentinel) moves as is
    WG_HIP(hipMemcpyAsync(s + cap * j->stride, j->slots + j->cap * j->stride, j->stride * 8, hipMemcpyDeviceToDevice,
                          j->ctx->stream));
    WG_HIP(hipStreamSynchronize(j->ctx->stream));
    WG_HIP(hipFree(j->slots));
  }
  j->slots = s;
  j->cap = cap;
  return WG_OK;
}

static wg_status join_reserve_rows(wg_join* j, uint64_t need) {
  if (need <= j->row_cap) return WG_OK;
  uint64_t cap = std::max<uint64_t>(need, j->row_cap * 2);
  cap = std::max<uint64_t>(cap, 1 << 16);
  for (int c = 0; c < j->n_ret; ++c) {
    void* d = nullptr;
    WG_ALLOC(j->ctx, &d, cap * type_bytes(j->types[c]), "a retained build column");
    if (j->col_data[c]) {
      WG_HIP(hipMemcpyAsync(d, j->col_data[c], j->rows * type_bytes(j->types[c]), hipMemcpyDeviceToDevice, j->ctx->stream));
      WG_HIP(hipStreamSynchronize(j->ctx->stream));
      WG_HIP(hipFree(j->col_data[c]));
    }
    j->col_data[c] = d;
    uint64_t* v = nullptr;
    WG_ALLOC(j->ctx, &v, (cap + 63) / 64 * 8, "a retained validity bitmap");
    WG_HIP(hipMemsetAsync(v, 0xFF, (cap + 63) / 64 * 8, j->ctx->stream));
    if (j->col_valid[c]) {
      WG_HIP(hipMemcpyAsync(v, j->col_valid[c], (j->rows + 63) / 64 * 8, hipMemcpyDeviceToDevice, j->ctx->stream));
      WG_HIP(hipStreamSynchronize(j->ctx->stream));
      WG_HIP(hipFree(j->col_valid[c]));
    }
    j->col_valid[c] = v;
  }
  int32_t* nx = nullptr;
  WG_ALLOC(j->ctx, &nx, cap * 4, "the chain array");
  if (j->next) {
    WG_HIP(hipMemcpyAsync(nx, j->next, j->rows * 4, hipMemcpyDeviceToDevice, j->ctx->stream));
    WG_HIP(hipStreamSynchronize(j->ctx->stream));
    WG_HIP(hipFree(j->next));
  }
  j->next = nx;
  j->row_cap = cap;
  return WG_OK;
}

// validity bits of appended rows land at row offset `at`: shift-copy on the host
static wg_status append_validity(wg_join* j, int c, const uint64_t* src, uint64_t n, bool src_device) {
  std::vector<uint64_t> words((n + 63) / 64);
  if (src_device) {
    WG_HIP(hipMemcpy(words.data(), src, words.size() * 8, hipMemcpyDeviceToHost));
  } else {
    memcpy(words.data(), src, words.size() * 8);
  }
  const uint64_t at = j->rows;
  const uint64_t first = at / 64, last = (at + n + 63) / 64;
  std::vector<uint64_t> dst(last - first);
  WG_HIP(hipMemcpy(dst.data(), j->col_valid[c] + first, dst.size() * 8, hipMemcpyDeviceToHost));
  for (uint64_t r = 0; r < n; ++r) {
    const uint64_t g = at + r;
    const bool v = (words[r >> 6] >> (r & 63)) & 1ull;
    uint64_t& w = dst[(g >> 6) - first];
    if (v) w |= 1ull << (g & 63);
    else w &= ~(1ull << (g & 63));
  }
  WG_HIP(hipMemcpy(j->col_valid[c] + first, dst.data(), dst.size() * 8, hipMemcpyHostToDevice));
  j->has_valid[c] = true;
  return WG_OK;
}

// Semi / anti probe predicate (executeSemiAntiJoin, join.go:3245-3310 and
// semiProbeInt64 / antiProbeInt64 :4486-4650): the key has a build row, or
// (anti) has none. A NULL probe key never matches, so the anti join emits it.
struct SemiPred {
  JoinTable J;
  ColDesc key;
  int anti;
  __device__ bool operator()(uint64_t row) const {
    if (!col_valid(key, row)) return anti != 0;
    const bool found = join_find(J, (uint64_t)load_int(key, row)) != nullptr;
    return anti ? !found : found;
  }
};

// ------------------------------------------------------ Bloom pre-filter
// HashJoin.buildBloom / bloomSet / bloomHashInt (join.go:1441-1483, :1650-1654)
// and bloomContains (bloom_filter_op.go:268-274): two bits per distinct build
// key in a power-of-two array of 64-bit words (~10 bits per key, at least 8
// words), both taken from one multiply-xorshift hash.
__host__ __device__ __forceinline__ uint64_t bloom_hash(int64_t key) {
  const uint64_t x = (uint64_t)key * 0x9E3779B97F4A7C15ull;
  return x ^ (x >> 32);
}

__global__ void bloom_build_kernel(const uint64_t* slots, uint64_t cap, int stride, uint64_t* bloom, uint64_t mask) {
  // table slots, then the side slot of the sentinel-valued key
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= cap; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t* s = slots + i * stride;
    const bool live = i < cap ? s[0] != kEmpty : (s[1] >> 32) != 0;
    if (!live) continue;
    const uint64_t h = bloom_hash((int64_t)(i < cap ? s[0] : kEmpty));
    atomicOr((unsigned long long*)&bloom[h & mask], 1ull << (h & 63));
    atomicOr((unsigned long long*)&bloom[(h >> 17) & mask], 1ull << ((h >> 6) & 63));
  }
}

struct BloomPred {
  const uint64_t* bloom;
  uint64_t mask;
  ColDesc key;
  __device__ bool operator()(uint64_t row) const {
    if (!col_valid(key, row)) return false;  // a NULL key never matches
    const uint64_t h = bloom_hash(load_int(key, row));
    return ((bloom[h & mask] >> (h & 63)) & 1ull) && ((bloom[(h >> 17) & mask] >> ((h >> 6) & 63)) & 1ull);
  }
};

static JoinTable jt_of(const wg_join* j) { return JoinTable{j->slots, j->cap - 1, j->stride, j->next}; }

static ColDesc retained_col(const wg_join* j, int c) {
  return ColDesc{j->col_data[c], j->has_valid[c] ? j->col_valid[c] : nullptr, j->types[c], j->scales[c]};
}

// Validity of two-column keys: a row's pair is NULL when either key is
// (dualIntKeyFromVectors, join.go:1630-1633); words [w0, w1) of the bitmaps.
__global__ void and_words_kernel(const uint64_t* a, const uint64_t* b, uint64_t w0, uint64_t w1, uint64_t* out) {
  for (uint64_t w = w0 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < w1; w += (uint64_t)gridDim.x * blockDim.x)
    out[w] = (a ? a[w] : ~0ull) & (b ? b[w] : ~0ull);
}

// The appended build rows [rows, rows + n) get their pairs' ids (new pairs
// enter the dictionary) in the retained id column, and its validity.
static wg_status dual_encode_build(wg_join* j, uint64_t n) {
  ColSet cs{};
  cs.n_cols = j->n_cols;
  cs.n_probe = j->n_cols;
  for (int c = 0; c < j->n_cols; ++c) {
    cs.cols[c] = retained_col(j, c);
    cs.inline_word[c] = -1;
  }
  const int ic = j->n_cols;
  const uint64_t w0 = j->rows / 64, w1 = (j->rows + n + 63) / 64;
  hipLaunchKernelGGL(and_words_kernel, dim3(grid_for(w1 - w0, 256)), dim3(256), 0, j->ctx->stream,
                     (const uint64_t*)j->col_valid[j->dkey.col[0]], (const uint64_t*)j->col_valid[j->dkey.col[1]], w0,
                     w1, j->col_valid[ic]);
  WG_HIP(hipGetLastError());
  j->has_valid[ic] = true;
  return kd_encode_rows(j->ctx, j->dict, cs, j->dkey, nullptr, n, true, (uint64_t*)j->col_data[ic], j->rows);
}

// Does the fused probe -> aggregate have the streamlined kernel's shape
// (wg_join_agg.hpp)? Fills the parts of Q that describe the query.
static bool lean_plan(const wg_join* j, const wg_agg* agg, const wg_column* probe_cols, int n_probe,
                      int probe_key, JoinLeanParams& Q) {
  if (!j->lean || agg->key.n != 1 || j->stride < 4) return false;
  // where a joined-row column comes from: a probe column or inline word 2/3
  auto source = [&](int c, ColDesc& cd, int32_t& word) -> bool {
    if (c < n_probe) {
      const wg_column& pc = probe_cols[c];
      if (pc.validity || pc.type == WG_DECIMAL128) return false;
      cd = ColDesc{pc.data, nullptr, pc.type, pc.scale};
      word = -1;
      return true;
    }
    const int b = c - n_probe;
    if (j->has_valid[b] || j->types[b] == WG_DECIMAL128) return false;
    for (int i = 0; i < j->n_inline && i < 2; ++i)
      if (j->inline_cols[i] == b) {
        cd = ColDesc{};
        word = 2 + i;
        return true;
      }
    return false;
  };
  const int gc = agg->key.col[0];
  if (!is_int_class(agg->col_types[gc]) || !source(gc, Q.gcol, Q.gword)) return false;
  Q.pg = agg->pg;
  std::vector<int> vcols;
  for (int o = 0; o < Q.pg.n_ops; ++o)
    for (int32_t* c : {&Q.pg.ops[o].col, &Q.pg.ops[o].col2}) {
      if (*c < 0) continue;
      int k = -1;
      for (size_t i = 0; i < vcols.size(); ++i)
        if (vcols[i] == *c) k = (int)i;
      if (k < 0) {
        k = (int)vcols.size();
        vcols.push_back(*c);
      }
      *c = k;
    }
  if ((int)vcols.size() > kLeanMaxVals) return false;
  Q.n_vals = (int32_t)vcols.size();
  Q.vfloat = 0;
  for (int k = 0; k < Q.n_vals; ++k) {
    if (!source(vcols[k], Q.vcol[k], Q.vword[k])) return false;
    if (agg->col_types[vcols[k]] == WG_FLOAT64) Q.vfloat |= 1u << k;
  }
  const wg_column& pk = probe_cols[probe_key];
  Q.pkey = ColDesc{pk.data, pk.validity, pk.type, pk.scale};
  Q.J = jt_of(j);
  return true;
}

static wg_status ensure_probe_scratch(wg_join* j, uint64_t n) {
  if (n <= j->scratch_rows) return WG_OK;
  const uint64_t tiles = (n + kProbeTile - 1) / kProbeTile;
  if (j->cnt) WG_HIP(hipFree(j->cnt));
  if (j->off) WG_HIP(hipFree(j->off));
  j->cnt = j->off = nullptr;
  j->scratch_rows = 0;
  WG_ALLOC(j->ctx, &j->cnt, n * 8, "probe scratch");                // per position: match count | chain head << 32
  WG_ALLOC(j->ctx, &j->off, (2 * tiles + 2) * 8, "probe scratch");  // per tile: pair count, then its offset
  size_t bytes = 0;
  WG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, j->off, j->off + tiles, (int)tiles, j->ctx->stream));
  if (bytes > j->cub_bytes) {
    if (j->cub_tmp) WG_HIP(hipFree(j->cub_tmp));
    j->cub_tmp = nullptr;
    j->cub_bytes = 0;
    WG_ALLOC(j->ctx, &j->cub_tmp, bytes, "scan scratch");
    j->cub_bytes = bytes;
  }
  j->scratch_rows = n;
  return WG_OK;
}

}  // namespace wg

using namespace wg;

extern "C" {

static wg_status join_create(wg_ctx* ctx, int32_t n_build_cols, const int32_t* build_types,
                             const int32_t* build_scales, int32_t build_key, int32_t build_key2, int32_t n_inline,
                             const int32_t* inline_cols, uint64_t build_hint, wg_join** out) {
  clear_error();
  WG_CHECK(ctx && out && build_types, WG_ERR_INVALID, "null argument");
  const bool dual = build_key2 >= 0;
  WG_CHECK(n_build_cols > 0 && n_build_cols + (dual ? 1 : 0) <= kMaxCols, WG_ERR_UNSUPPORTED, "%d build columns",
           n_build_cols);
  for (int key : {build_key, build_key2}) {
    if (key < 0 && key == build_key2) continue;
    WG_CHECK(key >= 0 && key < n_build_cols, WG_ERR_INVALID, "build key column %d", key);
    WG_CHECK(is_int_class(build_types[key]), WG_ERR_UNSUPPORTED, "join key of type %d not on the GPU path",
             build_types[key]);
  }
  WG_CHECK(n_inline >= 0 && n_inline <= kMaxInline, WG_ERR_INVALID, "%d inline columns (max %d)", n_inline, kMaxInline);
  wg_join* j = new wg_join();
  j->ctx = ctx;
  j->n_cols = n_build_cols;
  j->n_ret = n_build_cols;
  for (int c = 0; c < n_build_cols; ++c) {
    if (!(build_types[c] >= WG_INT32 && build_types[c] <= WG_UINT16)) {
      delete j;
      WG_CHECK(false, WG_ERR_UNSUPPORTED, "build column type %d", build_types[c]);
    }
    j->types[c] = build_types[c];
    j->scales[c] = build_scales ? build_scales[c] : 0;
  }
  j->key_col = build_key;
  if (dual) {
    // the table keys on the pair's id, a retained column after the batch's own
    j->dual = true;
    j->dkey.n = 2;
    j->dkey.col[0] = build_key;
    j->dkey.col[1] = build_key2;
    j->dkey.word[1] = 1;
    j->dkey.key_type[0] = build_types[build_key];
    j->dkey.key_type[1] = build_types[build_key2];
    j->key_col = n_build_cols;
    j->types[n_build_cols] = WG_INT64;
    j->n_ret = n_build_cols + 1;
  }
  j->n_inline = n_inline;
  for (int i = 0; i < n_inline; ++i) {
    const int c = inline_cols[i];
    if (!(c >= 0 && c < n_build_cols && j->types[c] != WG_DECIMAL128)) {
      delete j;
      WG_CHECK(false, WG_ERR_INVALID, "inline column %d", c);
    }
    j->inline_cols[i] = c;
  }
  j->stride = (int32_t)next_pow2(2 + n_inline);
  // 16 counter words: claims, digest, chain result, probe cut
  wg_status st = dev_alloc(ctx, (void**)&j->counter, 128, nullptr, "the join's counters");
  if (st == WG_OK) st = zero_dev(ctx, j->counter, 8);
  if (st == WG_OK) st = join_alloc_table(j, next_pow2(std::max<uint64_t>(build_hint, 1024) * 2));
  if (st == WG_OK && dual) st = kd_init(ctx, j->dict, build_hint);
  if (st != WG_OK) {
    wg_join_destroy(j);
    return st;
  }
  WG_HIP(hipStreamSynchronize(ctx->stream));
  *out = j;
  return WG_OK;
}

wg_status wg_join_create(wg_ctx* ctx, int32_t n_build_cols, const int32_t* build_types, const int32_t* build_scales,
                         int32_t build_key, int32_t n_inline, const int32_t* inline_cols, uint64_t build_hint,
                         wg_join** out) {
  return join_create(ctx, n_build_cols, build_types, build_scales, build_key, -1, n_inline, inline_cols, build_hint,
                     out);
}

wg_status wg_join_create_dual(wg_ctx* ctx, int32_t n_build_cols, const int32_t* build_types,
                              const int32_t* build_scales, int32_t build_key0, int32_t build_key1, int32_t n_inline,
                              const int32_t* inline_cols, uint64_t build_hint, wg_join** out) {
  WG_CHECK(build_key1 >= 0 && build_key1 != build_key0, WG_ERR_INVALID, "second build key column %d", build_key1);
  return join_create(ctx, n_build_cols, build_types, build_scales, build_key0, build_key1, n_inline, inline_cols,
                     build_hint, out);
}

wg_status wg_join_destroy(wg_join* j) {
  if (!j) return WG_OK;
  (void)hipStreamSynchronize(j->ctx->stream);
  for (int c = 0; c < j->n_ret; ++c) {
    (void)hipFree(j->col_data[c]);
    (void)hipFree(j->col_valid[c]);
  }
  if (j->dual) kd_free(j->dict);
  (void)hipFree(j->next);
  (void)hipFree(j->slots);
  (void)hipFree(j->counter);
  (void)hipFree(j->cnt);
  (void)hipFree(j->off);
  (void)hipFree(j->cub_tmp);
  (void)hipFree(j->bloom);
  delete j;
  return WG_OK;
}

wg_status wg_join_build_append(wg_join* j, const wg_column* cols, uint64_t n_rows, int32_t mem) {
  clear_error();
  WG_CHECK(j && cols, WG_ERR_INVALID, "null argument");
  WG_CHECK(!j->finished, WG_ERR_STATE, "build_append after build_finish");
  if (n_rows == 0) return WG_OK;
  WG_CHECK(j->rows + n_rows < (1ull << 31), WG_ERR_UNSUPPORTED, "build sides of 2^31 rows or more");
  for (int c = 0; c < j->n_cols; ++c)
    WG_CHECK(cols[c].type == j->types[c] && cols[c].data, WG_ERR_INVALID, "build column %d: type %d, expected %d", c,
             cols[c].type, j->types[c]);
  WG_TRY(join_reserve_rows(j, j->rows + n_rows));
  // keep the table at <= 50 % load
  if ((j->rows + n_rows) * 2 > j->cap) WG_TRY(join_alloc_table(j, next_pow2((j->rows + n_rows) * 2)));
  const hipMemcpyKind kind = mem == WG_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  for (int c = 0; c < j->n_cols; ++c) {
    const int w = type_bytes(j->types[c]);
    WG_HIP(hipMemcpyAsync((char*)j->col_data[c] + j->rows * w, cols[c].data, n_rows * w, kind, j->ctx->stream));
    if (cols[c].validity) WG_TRY(append_validity(j, c, cols[c].validity, n_rows, mem != WG_MEM_HOST));
  }
  if (j->dual) WG_TRY(dual_encode_build(j, n_rows));
  BuildParams B{};
  B.key = retained_col(j, j->key_col);
  for (int i = 0; i < j->n_inline; ++i) B.inl[i] = retained_col(j, j->inline_cols[i]);
  B.n_inline = j->n_inline;
  B.row0 = j->rows;
  B.n = n_rows;
  B.slots = j->slots;
  B.mask = j->cap - 1;
  B.stride = j->stride;
  B.next = j->next;
  B.claimed = j->counter;
  hipLaunchKernelGGL(join_build_kernel, dim3(grid_for(n_rows, 256, 16384)), dim3(256), 0, j->ctx->stream, B);
  WG_HIP(hipGetLastError());
  WG_HIP(hipStreamSynchronize(j->ctx->stream));
  j->rows += n_rows;
  return WG_OK;
}

wg_status wg_join_dual_probe_key(wg_join* j, const wg_column* probe_key0, const wg_column* probe_key1,
                                 uint64_t n_rows, const uint32_t* sel, uint64_t n_sel, int64_t* out_ids,
                                 uint64_t* out_valid) {
  clear_error();
  WG_CHECK(j && probe_key0 && probe_key1 && out_ids && out_valid, WG_ERR_INVALID, "null argument");
  WG_CHECK(j->dual, WG_ERR_STATE, "not a two-column-key join");
  WG_CHECK(j->finished, WG_ERR_STATE, "probe before build_finish");
  WG_CHECK(is_int_class(probe_key0->type) && is_int_class(probe_key1->type), WG_ERR_UNSUPPORTED,
           "probe key types %d, %d", probe_key0->type, probe_key1->type);
  const uint64_t n = sel ? n_sel : n_rows;
  const uint64_t words = (n_rows + 63) / 64;
  if (words) {
    hipLaunchKernelGGL(and_words_kernel, dim3(grid_for(words, 256)), dim3(256), 0, j->ctx->stream,
                       probe_key0->validity, probe_key1->validity, (uint64_t)0, words, out_valid);
    WG_HIP(hipGetLastError());
  }
  ColSet cs{};
  cs.n_cols = 2;
  cs.n_probe = 2;
  cs.cols[0] = ColDesc{probe_key0->data, probe_key0->validity, probe_key0->type, probe_key0->scale};
  cs.cols[1] = ColDesc{probe_key1->data, probe_key1->validity, probe_key1->type, probe_key1->scale};
  cs.inline_word[0] = cs.inline_word[1] = -1;
  KeyDesc K = j->dkey;
  K.col[0] = 0;
  K.col[1] = 1;
  WG_TRY(kd_lookup_rows(j->ctx, j->dict, cs, K, sel, n, true, (uint64_t*)out_ids, 0));
  WG_HIP(hipStreamSynchronize(j->ctx->stream));
  return WG_OK;
}

wg_status wg_join_build_finish(wg_join* j) {
  clear_error();
  WG_CHECK(j, WG_ERR_INVALID, "null handle");
  uint64_t c[2];
  WG_HIP(hipMemcpyAsync(c, j->counter, 16, hipMemcpyDeviceToHost, j->ctx->stream));
  uint64_t side = 0;
  WG_HIP(hipMemcpyAsync(&side, j->slots + j->cap * j->stride + 1, 8, hipMemcpyDeviceToHost, j->ctx->stream));
  WG_HIP(hipStreamSynchronize(j->ctx->stream));
  const uint64_t distinct = c[0] + ((side >> 32) ? 1 : 0);
  j->claimed_unique = distinct == c[1];
  j->distinct = distinct;
  j->finished = true;
  return WG_OK;
}

wg_status wg_join_set_tuning(wg_join* j, int32_t lean, int32_t block, int32_t replicas) {
  clear_error();
  WG_CHECK(j, WG_ERR_INVALID, "null handle");
  WG_CHECK(lean == 0 || lean == 1, WG_ERR_INVALID, "lean %d", lean);
  WG_CHECK(block == 512 || block == 1024, WG_ERR_INVALID, "block %d", block);
  WG_CHECK(replicas >= 0 && replicas <= kLeanMaxReplicas && (replicas & (replicas - 1)) == 0, WG_ERR_INVALID,
           "replicas %d", replicas);
  j->lean = lean;
  j->lean_block = block;
  j->lean_replicas = replicas;
  return WG_OK;
}

wg_status wg_join_build_rows(wg_join* j, uint64_t* rows) {
  WG_CHECK(j && rows, WG_ERR_INVALID, "null argument");
  *rows = j->rows;
  return WG_OK;
}

wg_status wg_join_build_column(wg_join* j, int32_t build_col, wg_column* out) {
  clear_error();
  WG_CHECK(j && out, WG_ERR_INVALID, "null argument");
  WG_CHECK(build_col >= 0 && build_col < j->n_cols, WG_ERR_INVALID, "build column %d", build_col);
  const ColDesc c = retained_col(j, build_col);
  *out = wg_column{c.data, c.valid, c.type, c.scale};
  return WG_OK;
}

wg_status wg_join_semi(wg_join* j, const wg_column* probe_key, uint64_t n_rows, const uint32_t* sel, uint64_t n_sel,
                       int32_t anti, uint32_t* out_sel, uint64_t* n_out) {
  clear_error();
  WG_CHECK(j && probe_key && out_sel && n_out, WG_ERR_INVALID, "null argument");
  WG_CHECK(j->finished, WG_ERR_STATE, "probe before build_finish");
  WG_CHECK(is_int_class(probe_key->type), WG_ERR_UNSUPPORTED, "probe key type %d", probe_key->type);
  const uint64_t n = sel ? n_sel : n_rows;
  const SemiPred pred{jt_of(j), ColDesc{probe_key->data, probe_key->validity, probe_key->type, probe_key->scale},
                      anti ? 1 : 0};
  return select_rows(j->ctx, pred, n, sel, out_sel, n_out);
}

wg_status wg_join_bloom_build(wg_join* j) {
  clear_error();
  WG_CHECK(j, WG_ERR_INVALID, "null handle");
  WG_CHECK(j->finished, WG_ERR_STATE, "bloom before build_finish");
  // nSlots: the power of two with nSlots * 64 >= 10 bits per key, at least 8
  uint64_t words = 1;
  while (words * 64 < j->distinct * 10) words *= 2;
  if (words < 8) words = 8;
  uint64_t cap_words = j->bloom ? j->bloom_mask + 1 : 0;
  if (cap_words != words) WG_TRY(regrow(j->ctx, &j->bloom, &cap_words, words, 8, "the Bloom filter"));
  WG_TRY(zero_dev(j->ctx, j->bloom, words));
  j->bloom_mask = words - 1;
  hipLaunchKernelGGL(bloom_build_kernel, dim3(grid_for(j->cap + 1, 256, 8192)), dim3(256), 0, j->ctx->stream,
                     (const uint64_t*)j->slots, j->cap, j->stride, j->bloom, j->bloom_mask);
  WG_HIP(hipGetLastError());
  WG_HIP(hipStreamSynchronize(j->ctx->stream));
  return WG_OK;
}

wg_status wg_join_bloom_filter(wg_join* j, const wg_column* probe_key, uint64_t n_rows, const uint32_t* sel,
                               uint64_t n_sel, uint32_t* out_sel, uint64_t* n_out) {
  clear_error();
  WG_CHECK(j && probe_key && out_sel && n_out, WG_ERR_INVALID, "null argument");
  WG_CHECK(j->bloom, WG_ERR_STATE, "bloom filter before wg_join_bloom_build");
  WG_CHECK(is_int_class(probe_key->type), WG_ERR_UNSUPPORTED, "probe key type %d", probe_key->type);
  const uint64_t n = sel ? n_sel : n_rows;
  const BloomPred pred{j->bloom, j->bloom_mask,
                       ColDesc{probe_key->data, probe_key->validity, probe_key->type, probe_key->scale}};
  return select_rows(j->ctx, pred, n, sel, out_sel, n_out);
}

wg_status wg_join_bloom_words(wg_join* j, uint64_t* words, uint64_t* out, uint64_t cap) {
  clear_error();
  WG_CHECK(j && words, WG_ERR_INVALID, "null argument");
  WG_CHECK(j->bloom, WG_ERR_STATE, "bloom filter before wg_join_bloom_build");
  *words = j->bloom_mask + 1;
  if (out && cap >= *words) WG_HIP(hipMemcpy(out, j->bloom, *words * 8, hipMemcpyDeviceToHost));
  return WG_OK;
}

wg_status wg_join_probe(wg_join* j, const wg_column* probe_key, uint64_t n_rows, const uint32_t* sel, uint64_t n_sel,
                        uint32_t* out_probe_rows, uint32_t* out_build_rows, uint64_t out_cap, wg_probe_cursor* cursor,
                        uint64_t* n_out) {
  clear_error();
  WG_CHECK(j && probe_key && cursor && n_out && out_probe_rows && out_build_rows, WG_ERR_INVALID, "null argument");
  WG_CHECK(j->finished, WG_ERR_STATE, "probe before build_finish");
  WG_CHECK(is_int_class(probe_key->type), WG_ERR_UNSUPPORTED, "probe key type %d", probe_key->type);
  WG_CHECK(out_cap > 0, WG_ERR_INVALID, "out_cap must be positive");
  const uint64_t n = sel ? n_sel : n_rows;
  *n_out = 0;
  if (cursor->done || cursor->pos >= n) {
    cursor->done = 1;
    return WG_OK;
  }
  WG_CHECK(n < (1ull << 31), WG_ERR_INVALID, "probe batches of 2^31 rows or more must be split");
  const JoinTable J = jt_of(j);
  const ColDesc pk{probe_key->data, probe_key->validity, probe_key->type, probe_key->scale};
  int64_t* res = (int64_t*)(j->counter + 4);
  uint64_t written = 0;
  // resume a suspended chain first (resumeIntChain, join.go:2819-2836)
  if (cursor->chain >= 0) {
    const uint64_t row = sel ? 0 : cursor->pos;
    uint32_t prow = (uint32_t)row;
    if (sel) WG_HIP(hipMemcpy(&prow, sel + cursor->pos, 4, hipMemcpyDeviceToHost));
    hipLaunchKernelGGL(join_chain_kernel, dim3(1), dim3(64), 0, j->ctx->stream, (const int32_t*)j->next,
                       (int32_t)cursor->chain, prow, out_cap, out_probe_rows, out_build_rows, res);
    WG_HIP(hipGetLastError());
    int64_t r[2];
    WG_HIP(hipMemcpyAsync(r, res, 16, hipMemcpyDeviceToHost, j->ctx->stream));
    WG_HIP(hipStreamSynchronize(j->ctx->stream));
    written = (uint64_t)r[0];
    cursor->chain = r[1];
    if (cursor->chain >= 0) {
      *n_out = written;
      return WG_OK;
    }
    cursor->pos += 1;
    if (cursor->pos >= n) {
      cursor->done = 1;
      *n_out = written;
      return WG_OK;
    }
  }
  const uint64_t pos = cursor->pos;
  const uint64_t room = out_cap - written;
  // a window of positions per call keeps a bounded probe of a large batch
  // linear: at most room pairs come out, and positions without a match
  // produce none, so twice the room (and at least 64K positions) is examined
  const uint64_t m = std::min<uint64_t>(n - pos, std::max<uint64_t>(2 * room, 1ull << 16));
  WG_TRY(ensure_probe_scratch(j, m));
  const uint64_t tiles = (m + kProbeTile - 1) / kProbeTile;
  uint64_t* hc = j->cnt;
  uint64_t* tile_cnt = j->off;
  uint64_t* tile_off = j->off + tiles;
  hipLaunchKernelGGL(join_count_kernel, dim3((uint32_t)tiles), dim3(256), 0, j->ctx->stream, J, pk, sel, pos, m, hc,
                     tile_cnt);
  WG_HIP(hipGetLastError());
  size_t bytes = j->cub_bytes;
  WG_HIP(hipcub::DeviceScan::ExclusiveSum(j->cub_tmp, bytes, tile_cnt, tile_off, (int)tiles, j->ctx->stream));
  uint64_t* cut_d = j->counter + 6;  // [0] first tile that does not fit, then the straddling position and offset
  const uint64_t init[3] = {~0ull, ~0ull, 0};
  WG_HIP(hipMemcpyAsync(cut_d, init, 24, hipMemcpyHostToDevice, j->ctx->stream));
  hipLaunchKernelGGL(join_cut_kernel, dim3(grid_for(tiles, 256, 4096)), dim3(256), 0, j->ctx->stream,
                     (const uint64_t*)tile_off, (const uint64_t*)tile_cnt, tiles, room, cut_d);
  WG_HIP(hipGetLastError());
  uint64_t cut_t = 0;
  WG_HIP(hipMemcpyAsync(&cut_t, cut_d, 8, hipMemcpyDeviceToHost, j->ctx->stream));
  WG_HIP(hipStreamSynchronize(j->ctx->stream));
  const uint64_t emit_tiles = cut_t == ~0ull ? tiles : cut_t + 1;
  hipLaunchKernelGGL(join_emit_kernel, dim3((uint32_t)emit_tiles), dim3(256), 0, j->ctx->stream, J, sel, pos, m,
                     (const uint64_t*)hc, (const uint64_t*)tile_off, room, out_probe_rows + written,
                     out_build_rows + written, cut_d + 1);
  WG_HIP(hipGetLastError());
  uint64_t fit = m, total = 0;
  if (cut_t == ~0ull) {
    uint64_t last[2] = {0, 0};
    WG_HIP(hipMemcpyAsync(&last[0], tile_off + tiles - 1, 8, hipMemcpyDeviceToHost, j->ctx->stream));
    WG_HIP(hipMemcpyAsync(&last[1], tile_cnt + tiles - 1, 8, hipMemcpyDeviceToHost, j->ctx->stream));
    WG_HIP(hipStreamSynchronize(j->ctx->stream));
    total = last[0] + last[1];
  } else {
    uint64_t c2[2] = {0, 0};
    WG_HIP(hipMemcpyAsync(c2, cut_d + 1, 16, hipMemcpyDeviceToHost, j->ctx->stream));
    WG_HIP(hipStreamSynchronize(j->ctx->stream));
    WG_CHECK(c2[0] < m && c2[0] / kProbeTile == cut_t, WG_ERR_DEVICE, "probe output cut not found in tile %llu",
             (unsigned long long)cut_t);
    fit = c2[0];
    total = c2[1];
  }
  written += total;
  if (fit == m) {
    cursor->pos = pos + m;
    if (cursor->pos >= n) cursor->done = 1;
  } else if (fit > 0) {
    cursor->pos = pos + fit;
  } else {
    // the first remaining row alone exceeds the room: split its chain
    uint32_t prow = (uint32_t)pos;
    if (sel) WG_HIP(hipMemcpy(&prow, sel + pos, 4, hipMemcpyDeviceToHost));
    hipLaunchKernelGGL(join_head_kernel, dim3(1), dim3(64), 0, j->ctx->stream, J, pk, (uint64_t)prow, res);
    int64_t head = -1;
    WG_HIP(hipMemcpyAsync(&head, res, 8, hipMemcpyDeviceToHost, j->ctx->stream));
    WG_HIP(hipStreamSynchronize(j->ctx->stream));
    hipLaunchKernelGGL(join_chain_kernel, dim3(1), dim3(64), 0, j->ctx->stream, (const int32_t*)j->next, (int32_t)head,
                       prow, room, out_probe_rows + written, out_build_rows + written, res);
    int64_t r[2];
    WG_HIP(hipMemcpyAsync(r, res, 16, hipMemcpyDeviceToHost, j->ctx->stream));
    WG_HIP(hipStreamSynchronize(j->ctx->stream));
    written += (uint64_t)r[0];
    cursor->pos = pos;
    cursor->chain = r[1];
    if (r[1] < 0) cursor->pos = pos + 1;
  }
  WG_HIP(hipStreamSynchronize(j->ctx->stream));
  *n_out = written;
  return WG_OK;
}

wg_status wg_join_gather_build(wg_join* j, int32_t build_col, const uint32_t* build_rows, uint64_t n, void* out) {
  clear_error();
  WG_CHECK(j && (n == 0 || (build_rows && out)), WG_ERR_INVALID, "null argument");
  WG_CHECK(build_col >= 0 && build_col < j->n_cols, WG_ERR_INVALID, "build column %d", build_col);
  if (n == 0) return WG_OK;
  hipLaunchKernelGGL(gather_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0, j->ctx->stream, retained_col(j, build_col),
                     build_rows, n, out);
  WG_HIP(hipGetLastError());
  WG_HIP(hipStreamSynchronize(j->ctx->stream));
  return WG_OK;
}

wg_status wg_gather(wg_ctx* ctx, const wg_column* col, const uint32_t* rows, uint64_t n, void* out) {
  clear_error();
  WG_CHECK(ctx && col && (n == 0 || (rows && out)), WG_ERR_INVALID, "null argument");
  if (n == 0) return WG_OK;
  ColDesc c{col->data, col->validity, col->type, col->scale};
  hipLaunchKernelGGL(gather_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0, ctx->stream, c, rows, n, out);
  WG_HIP(hipGetLastError());
  WG_HIP(hipStreamSynchronize(ctx->stream));
  return WG_OK;
}

wg_status wg_join_digest(wg_join* j, const wg_column* probe_cols, const uint32_t* probe_rows, const uint32_t* build_rows,
                         uint64_t n, const int32_t* probe_sel_cols, int32_t n_probe_sel, const int32_t* build_sel_cols,
                         int32_t n_build_sel, uint64_t digest_out[2]) {
  clear_error();
  WG_CHECK(j && digest_out, WG_ERR_INVALID, "null argument");
  WG_CHECK(n_probe_sel >= 0 && n_probe_sel <= 8 && n_build_sel >= 0 && n_build_sel <= 8, WG_ERR_INVALID, "at most 8 + 8 columns");
  DigestParams D{};
  D.n_p = n_probe_sel;
  D.n_b = n_build_sel;
  for (int i = 0; i < n_probe_sel; ++i) {
    const wg_column& c = probe_cols[probe_sel_cols[i]];
    WG_CHECK(c.type != WG_DECIMAL128, WG_ERR_UNSUPPORTED, "digest of DECIMAL128");
    D.pc[i] = ColDesc{c.data, c.validity, c.type, c.scale};
  }
  for (int i = 0; i < n_build_sel; ++i) {
    const int c = build_sel_cols[i];
    WG_CHECK(c >= 0 && c < j->n_cols && j->types[c] != WG_DECIMAL128, WG_ERR_INVALID, "digest build column %d", c);
    D.bc[i] = retained_col(j, c);
  }
  unsigned long long* acc = (unsigned long long*)(j->counter + 2);
  WG_TRY(zero_dev(j->ctx, (uint64_t*)acc, 2));
  if (n) {
    hipLaunchKernelGGL(join_digest_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, j->ctx->stream, D, probe_rows,
                       build_rows, n, acc);
    WG_HIP(hipGetLastError());
  }
  WG_HIP(hipMemcpyAsync(digest_out, acc, 16, hipMemcpyDeviceToHost, j->ctx->stream));
  WG_HIP(hipStreamSynchronize(j->ctx->stream));
  return WG_OK;
}

static wg_status probe_agg_call(wg_join* j, const wg_column* probe_cols, int32_t n_probe_cols, int32_t probe_key,
                                uint64_t n_rows, const uint32_t* sel, uint64_t n_sel, wg_agg* agg) {
  WG_CHECK(j->finished, WG_ERR_STATE, "probe before build_finish");
  WG_CHECK(!agg->finalized, WG_ERR_STATE, "aggregate already finalized");
  WG_CHECK(agg->ctx == j->ctx, WG_ERR_INVALID, "join and aggregate on different contexts");
  WG_CHECK(n_probe_cols > 0 && probe_key >= 0 && probe_key < n_probe_cols, WG_ERR_INVALID, "probe key column");
  WG_CHECK(agg->n_cols == n_probe_cols + j->n_cols && agg->n_cols <= kMaxCols, WG_ERR_INVALID,
           "aggregate schema must be the %d probe columns followed by the %d build columns", n_probe_cols, j->n_cols);
  for (int c = 0; c < n_probe_cols; ++c)
    WG_CHECK(probe_cols[c].type == agg->col_types[c], WG_ERR_INVALID, "probe column %d type", c);
  for (int c = 0; c < j->n_cols; ++c)
    WG_CHECK(j->types[c] == agg->col_types[n_probe_cols + c], WG_ERR_INVALID, "build column %d type", c);
  WG_CHECK(is_int_class(probe_cols[probe_key].type), WG_ERR_UNSUPPORTED, "probe key type");
  WG_CHECK(!agg->wide, WG_ERR_UNSUPPORTED, "a GROUP BY key wider than 64 bits after a join is not on the GPU path");
  if (agg->key.n > 1)
    for (int k = 0; k < agg->key.n; ++k)
      WG_CHECK(!agg->col_nullable[agg->key.col[k]], WG_ERR_UNSUPPORTED, "nullable multi-column key after a join");
  const uint64_t n = sel ? n_sel : n_rows;
  if (n == 0) return WG_OK;
  WG_CHECK(n < (1ull << 31), WG_ERR_INVALID, "probe batches of 2^31 rows or more must be split");
  WG_TRY(agg_flush_stage(agg));
  WG_TRY(agg_settle(agg));
  if (agg->rows_since_norm + n >= (1ull << 31)) WG_TRY(normalize(agg));
  WG_TRY(agg_ensure_capacity(agg, 0));
  WG_TRY(ensure_overflow_buffers(agg));

  JoinAggParams P{};
  P.J = jt_of(j);
  P.pkey = ColDesc{probe_cols[probe_key].data, probe_cols[probe_key].validity, probe_cols[probe_key].type,
                   probe_cols[probe_key].scale};
  P.cs.n_cols = agg->n_cols;
  P.cs.n_probe = n_probe_cols;
  for (int c = 0; c < n_probe_cols; ++c) {
    P.cs.cols[c] = ColDesc{probe_cols[c].data, probe_cols[c].validity, probe_cols[c].type, probe_cols[c].scale};
    P.cs.inline_word[c] = -1;
  }
  for (int c = 0; c < j->n_cols; ++c) {
    P.cs.cols[n_probe_cols + c] = retained_col(j, c);
    int w = -1;
    for (int i = 0; i < j->n_inline; ++i)
      if (j->inline_cols[i] == c && !j->has_valid[c]) w = 2 + i;
    P.cs.inline_word[n_probe_cols + c] = w;
  }
  P.key = agg->key;
  P.pg = agg->pg;
  // a join usually feeds a small GROUP BY: size the LDS table to the expected
  // groups (at most the aggregate's setting) so more workgroups fit per CU
  P.lds_slots = agg->lds_slots;
  if (P.lds_slots > 0 && agg->group_hint > 0)
    P.lds_slots = (int32_t)std::min<uint64_t>((uint64_t)P.lds_slots,
                                              next_pow2(std::max<uint64_t>(64, 2 * agg->group_hint)));
  P.tile_rows = agg->tile_rows;
  const int lds_bytes_j = P.lds_slots * P.pg.stride * 8;
  const int per_cu_j = P.lds_slots > 0
                           ? std::max(1, std::min(160 * 1024 / std::max(lds_bytes_j, 1), 2048 / agg->block))
                           : std::max(1, 2048 / agg->block);
  const int grid_j = agg->ctx->cus * per_cu_j;
  // streamlined kernel, for a GROUP BY declared small: group slots for twice
  // the expected groups, then as many accumulator replicas as fit the LDS of
  // the workgroups a CU holds
  JoinLeanParams Q{};
  const int per_cu_l = j->lean_block == 512 ? 2 : 1;
  const size_t lds_budget = (size_t)(160 * 1024 / per_cu_l) - 2048;
  Q.L = (int32_t)next_pow2(std::max<uint64_t>(32, 2 * agg->group_hint));
  const bool lean = agg->group_hint > 0 && lean_lds_bytes(Q.L, 1, agg->pg.stride) <= lds_budget &&
                    lean_plan(j, agg, probe_cols, n_probe_cols, probe_key, Q);
  int grid_l = 0;
  if (lean) {
    int R = j->lean_replicas ? j->lean_replicas : kLeanMaxReplicas;
    while (R > 1 && lean_lds_bytes(Q.L, R, Q.pg.stride) > lds_budget) R >>= 1;
    Q.R = R;
    Q.tile_rows = agg->tile_rows;
    grid_l = agg->ctx->cus * per_cu_l;
  }
  WG_TRY(ensure_overflow_for(agg, (uint64_t)std::max(grid_j, grid_l), std::max<int32_t>(P.lds_slots, lean ? Q.L : 0)));
  agg->st_kernel_ms = 0.f;
  agg->st_launches = 0;
  if (!agg->ev0) {
    WG_HIP(hipEventCreate(&agg->ev0));
    WG_HIP(hipEventCreate(&agg->ev1));
  }
  // work list: (sel, n) segments still to process; overflow rows become new segments
  struct Seg { const uint32_t* sel; uint64_t n; bool lean; };
  std::vector<Seg> work{{sel, n, lean}};
  struct Owned {  // replay lists, freed on every exit
    hipStream_t st;
    std::vector<uint32_t*> v;
    ~Owned() {
      if (!v.empty()) (void)hipStreamSynchronize(st);
      for (uint32_t* p : v) (void)hipFree(p);
    }
  } owned{agg->ctx->stream, {}};
  wg_status st = WG_OK;
  for (int guard = 0; !work.empty() && st == WG_OK && guard < 64; ++guard) {
    Seg seg = work.back();
    work.pop_back();
    uint64_t pos = 0;
    while (pos < seg.n && st == WG_OK) {
      P.T = table_of(agg);
      P.ovf_rows = agg->ovf_rows;
      P.ovf_rows_cap = agg->ovf_rows_cap;
      P.ovf_part = agg->ovf_part;
      P.sel = seg.sel ? seg.sel + pos : nullptr;
      P.row_base = seg.sel ? 0 : pos;
      P.n = seg.n - pos;
      P.n_tiles = (P.n + P.tile_rows - 1) / P.tile_rows;
      WG_TRY(zero_dev(agg->ctx, agg->counters + CTR_OVF_ROWS, CTR_N - 1));
      const int grid = (int)std::min<uint64_t>((uint64_t)(seg.lean ? grid_l : grid_j), std::max<uint64_t>(P.n_tiles, 1));
      WG_HIP(hipEventRecord(agg->ev0, agg->ctx->stream));
      agg->touched = true;
      if (seg.lean) {
        Q.T = P.T;
        Q.sel = P.sel;
        Q.row_base = P.row_base;
        Q.n = P.n;
        Q.n_tiles = P.n_tiles;
        Q.ovf_rows = P.ovf_rows;
        Q.ovf_rows_cap = P.ovf_rows_cap;
        Q.ovf_part = P.ovf_part;
        WG_TRY(launch_join_agg_lean(Q, grid, j->lean_block, agg->ctx->stream));
      } else if (P.lds_slots > 0) {
        static bool attr_done = false;
        if (!attr_done) {
          WG_HIP(hipFuncSetAttribute((const void*)join_agg_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     150 * 1024));
          attr_done = true;
        }
        hipLaunchKernelGGL(join_agg_kernel<true>, dim3(grid), dim3(agg->block),
                           (size_t)P.lds_slots * P.pg.stride * 8, agg->ctx->stream, P);
      } else {
        hipLaunchKernelGGL(join_agg_kernel<false>, dim3(grid), dim3(agg->block), 0, agg->ctx->stream, P);
      }
      WG_HIP(hipGetLastError());
      WG_HIP(hipEventRecord(agg->ev1, agg->ctx->stream));
      agg->dirty = true;
      uint64_t c[CTR_N];
      WG_TRY(agg_read_counters(agg, c));
      float ms = 0.f;
      WG_HIP(hipEventElapsedTime(&ms, agg->ev0, agg->ev1));
      agg->st_kernel_ms += ms;
      agg->st_launches += 1;
      agg->groups = c[CTR_GROUPS];
      const uint64_t tiles = std::min<uint64_t>(c[CTR_TILE], P.n_tiles);
      if (c[CTR_OVF_ROWS] || c[CTR_OVF_PART]) {
        const uint64_t nr = c[CTR_OVF_ROWS], np = c[CTR_OVF_PART];
        // rows the streamlined kernel leaves to the general one need no
        // room: grow only when a partial found the table past its budget
        if (!seg.lean || np)
          WG_TRY(grow_to(agg, next_pow2(std::max<uint64_t>((agg->groups + nr * 4 + np) * 2, agg->cap * 2))));
        if (np) {
          uint64_t* recs = nullptr;
          const uint64_t bytes = np * (uint64_t)agg->pg.stride * 8;
          WG_ALLOC(j->ctx, &recs, bytes, "overflow records");
          WG_HIP(hipMemcpyAsync(recs, agg->ovf_part, bytes, hipMemcpyDeviceToDevice, agg->ctx->stream));
          GlobalTable T = table_of(agg);
          T.max_probe = agg->cap;
          hipLaunchKernelGGL(agg_merge_part_kernel_j, dim3(grid_for(np, 256)), dim3(256), 0, agg->ctx->stream, recs, np,
                             T, agg->pg);
          WG_HIP(hipGetLastError());
          uint64_t c2[CTR_N];
          WG_TRY(agg_read_counters(agg, c2));
          WG_HIP(hipFree(recs));
          WG_CHECK((c2[CTR_STOPPED] >> 32) == 0, WG_ERR_DEVICE, "overflow partials did not fit the grown table");
          agg->groups = c2[CTR_GROUPS];
        }
        if (nr) {
          uint32_t* rows = nullptr;
          WG_ALLOC(j->ctx, &rows, nr * 4, "replayed rows");
          owned.v.push_back(rows);
          WG_HIP(hipMemcpyAsync(rows, agg->ovf_rows, nr * 4, hipMemcpyDeviceToDevice, agg->ctx->stream));
          WG_HIP(hipStreamSynchronize(agg->ctx->stream));
          work.push_back(Seg{rows, nr, false});  // the general kernel takes every row shape
        }
      }
      WG_CHECK(tiles > 0 || c[CTR_OVF_ROWS] > 0, WG_ERR_DEVICE, "join probe launch made no progress");
      pos = std::min<uint64_t>(seg.n, pos + tiles * P.tile_rows);
      if (pos < seg.n) WG_TRY(agg_ensure_capacity(agg, 0));
    }
  }
  WG_CHECK(work.empty(), WG_ERR_DEVICE, "join aggregate overflow did not converge");
  agg->rows_since_norm += n;
  return st;
}

wg_status wg_join_probe_agg(wg_join* j, const wg_column* probe_cols, int32_t n_probe_cols, int32_t probe_key,
                            uint64_t n_rows, const uint32_t* sel, uint64_t n_sel, wg_agg* agg) {
  clear_error();
  WG_CHECK(j && agg && probe_cols, WG_ERR_INVALID, "null argument");
  WG_TRY(agg_guard_begin(agg));
  return agg_guard_end(agg, probe_agg_call(j, probe_cols, n_probe_cols, probe_key, n_rows, sel, n_sel, agg));
}

}  // extern "C"
