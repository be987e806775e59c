// wg_join_agg.hip — streamlined fused probe -> GROUP BY.
//
// The join's usual consumer is a small GROUP BY over a build payload column
// (config: JOIN fk = key GROUP BY p1 SUM(qty * p2), 25 groups). The general
// fused kernel (wg_join.hip) walks one probe row at a time through the
// generic row program and aggregates into one LDS table per workgroup; with a
// few dozen groups the 64 lanes of a wave then hit the same few LDS words and
// their atomics serialise. This kernel keeps the semantics of
// HashJoinProbe.Execute -> HashAggregate.Consume (join.go:2306-2356,
// inlineIntProbe :2670-2814; aggregate.go:1908-2146 for the int group key)
// and changes only the schedule:
//   * every thread carries kLeanU probe rows; their probe-column loads and
//     then their home-slot loads (key, head|count, two inline payload words
//     in one 32-byte read) are all issued before any is used;
//   * the group table in LDS holds R replicas of every accumulator word;
//     lane l updates replica l % R, and the replicas are folded once at the
//     end of the kernel before the partials merge into the global table.
// Rows outside the shape (see wg_join_agg.hpp), and rows whose group finds
// the LDS table full, are appended to the overflow row list and aggregated
// afterwards by the general kernel, so the result is the same accumulator
// state the general path produces. The host takes this kernel only for a
// GROUP BY declared small (its group hint), so that list stays short.
#include "wg_join_agg.hpp"

namespace wg {

constexpr int kLeanU = 2;  // probe rows per thread per step (more spill registers)

// Find or claim `key` in an LDS key column where slot i's key word is
// lds[i * R]. Returns the slot index, or -1 past the probe limit.
__device__ __forceinline__ int lds_find_rep(uint64_t* lds, int mask, int R, uint64_t key, uint64_t h) {
  int idx = (int)((h >> 20) & (uint64_t)mask);
  for (int p = 0; p < kLdsProbes; ++p) {
    uint64_t* s = lds + (uint64_t)idx * R;
    const uint64_t k = __hip_atomic_load(s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == key) return idx;
    if (k == kEmpty) {
      const uint64_t prev = at_cas<LdsT>(s, kEmpty, key);
      if (prev == kEmpty || prev == key) return idx;
    }
    idx = (idx + 1) & mask;
  }
  return -1;
}

// Fold replicas 1..R-1 of one slot into replica 0 (word k of replica r at
// s[k * ws + r]). Counts, integer sums and carry-save limbs add (a limb word
// stays below its overflow bound: the folded additions are the rows the slot
// took); float sums add; MIN/MAX keep the extreme encoded word.
__device__ __forceinline__ void fold_replicas(uint64_t* s, int ws, int R, const AggProgram& pg) {
  for (int o = 0; o < pg.n_ops; ++o) {
    const OpDesc op = pg.ops[o];
    const int nw = op.kind == OP_SUMD ? 2 : op.kind == OP_SUMD4 ? 4 : 1;
    for (int k = 0; k < nw; ++k) {
      uint64_t* t = s + (uint64_t)(op.word + k) * ws;
      uint64_t a = t[0];
      for (int r = 1; r < R; ++r) {
        const uint64_t b = t[r];
        switch (op.kind) {
          case OP_SUMF:
            a = (uint64_t)__double_as_longlong(__longlong_as_double((long long)a) + __longlong_as_double((long long)b));
            break;
          case OP_MIN: a = b < a ? b : a; break;
          case OP_MAX: a = b > a ? b : a; break;
          default: a += b; break;
        }
      }
      t[0] = a;
    }
  }
}

__device__ __forceinline__ void load_slot4(const JoinTable& J, uint64_t idx, uint64_t& w0, uint64_t& w1, uint64_t& w2,
                                           uint64_t& w3) {
  const ulonglong2* s = (const ulonglong2*)(J.slots + idx * (uint64_t)J.stride);
  const ulonglong2 a = s[0], b = s[1];
  w0 = a.x;
  w1 = a.y;
  w2 = b.x;
  w3 = b.y;
}

template <int NV>
__global__ void __launch_bounds__(1024) join_agg_lean_kernel(JoinLeanParams P) {
  constexpr int U = kLeanU;
  extern __shared__ uint64_t lds[];
  __shared__ int64_t s_tile;
  const int L = P.L, R = P.R;
  const int ws = (L + 1) * R;  // distance between the words of a slot
  const int stride = P.pg.stride;
  for (int k = 0; k < stride; ++k) {
    const uint64_t v = slot_init_word(P.pg, k);
    for (int i = threadIdx.x; i < ws; i += blockDim.x) lds[(uint64_t)k * ws + i] = v;
  }
  __syncthreads();
  const uint32_t rep = threadIdx.x & (uint32_t)(R - 1);
  uint64_t* const spare = lds + (uint64_t)L * R + rep;  // identity updates of rows that do not aggregate here
  const ColRes pk = resolve_col(P.pkey, true);
  const ColRes gk = resolve_col(P.gcol, P.gword < 0);
  ColRes vc[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) vc[j] = resolve_col(P.vcol[j], j < P.n_vals && P.vword[j] < 0);
  uint32_t all_valid[U];
#pragma unroll
  for (int u = 0; u < U; ++u) all_valid[u] = ~0u;
  uint32_t claims = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      int64_t t = -1;
      // take a tile only while every in-flight tile could still overflow entirely
      const uint64_t o = __hip_atomic_load(P.T.counters + CTR_OVF_ROWS, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (o + (uint64_t)gridDim.x * P.tile_rows > P.ovf_rows_cap) {
        at_add<GlbT>(P.T.counters + CTR_STOPPED, 1);
      } else {
        const uint64_t tt =
            __hip_atomic_fetch_add(P.T.counters + CTR_TILE, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
    for (uint64_t base = beg; base < end; base += (uint64_t)blockDim.x * U) {
      uint64_t row[U];
      bool in[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint64_t i = base + (uint64_t)u * blockDim.x + threadIdx.x;
        in[u] = i < end;
        const uint64_t ii = in[u] ? i : beg;  // past the end: a row in range, ignored
        row[u] = P.sel ? (uint64_t)P.sel[ii] : P.row_base + ii;
      }
      // probe columns of all U rows in flight
      RawVal kr[U], gr[U], vr[NV][U];
      uint64_t kv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        kr[u] = load_res(pk, row[u]);
        kv[u] = valid_res(pk, row[u]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) gr[u] = load_res(gk, row[u]);
#pragma unroll
      for (int j = 0; j < NV; ++j)
#pragma unroll
        for (int u = 0; u < U; ++u) vr[j][u] = load_res(vc[j], row[u]);
      uint64_t key[U];
      values_of<U>(pk.type, kr, key);
      // home slots of all U rows in flight
      uint64_t idx[U], w0[U], w1[U], w2[U], w3[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        idx[u] = mix64(key[u]) & P.J.mask;
        load_slot4(P.J, idx[u], w0[u], w1[u], w2[u], w3[u]);
      }
      uint32_t live = 0, slow = 0;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool nn = in[u] && ((kv[u] >> (row[u] & 63)) & 1ull);  // a NULL probe key never matches
        if (nn && key[u] == kEmpty) slow |= 1u << u;                 // the sentinel-valued key: side slot
        else if (nn) live |= 1u << u;
      }
      // linear probes that did not end at the home slot continue (the
      // next slots are usually in the same or the next cache line)
      uint32_t pend = 0;
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (((live >> u) & 1u) && w0[u] != key[u] && w0[u] != kEmpty) pend |= 1u << u;
      for (uint64_t p = 0; __any(pend != 0) && p <= P.J.mask; ++p) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (!((pend >> u) & 1u)) continue;
          idx[u] = (idx[u] + 1) & P.J.mask;
          load_slot4(P.J, idx[u], w0[u], w1[u], w2[u], w3[u]);
          if (w0[u] == key[u] || w0[u] == kEmpty) pend &= ~(1u << u);
        }
      }
      uint32_t ok = 0;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (!((live >> u) & 1u) || w0[u] != key[u]) continue;
        if ((uint32_t)(w1[u] >> 32) == 1u) ok |= 1u << u;  // one build row: its payload is inline
        else slow |= 1u << u;                               // a chain: the general kernel walks it
      }
      // group key and value words of the joined rows
      uint64_t gkey[U], v[U][NV];
      if (P.gword >= 0) {
#pragma unroll
        for (int u = 0; u < U; ++u) gkey[u] = P.gword == 2 ? w2[u] : w3[u];
      } else {
        values_of<U>(gk.type, gr, gkey);
      }
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        if (j < P.n_vals && P.vword[j] >= 0) {
#pragma unroll
          for (int u = 0; u < U; ++u) v[u][j] = P.vword[j] == 2 ? w2[u] : w3[u];
        } else {
          uint64_t t[U];
          values_of<U>(vc[j].type, vr[j], t);
#pragma unroll
          for (int u = 0; u < U; ++u) v[u][j] = t[u];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (((ok >> u) & 1u) && gkey[u] == kEmpty) {
          ok &= ~(1u << u);
          slow |= 1u << u;
        }
      // LDS groups: every lane runs the same atomics (identity on the spare slot)
      int li[U];
      uint64_t* wp[U];
      uint32_t in_lds = 0;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        li[u] = ((ok >> u) & 1u) ? lds_find_rep(lds, L - 1, R, gkey[u], key_hash(gkey[u])) : -1;
        wp[u] = li[u] >= 0 ? lds + (uint64_t)li[u] * R + rep : spare;
        in_lds |= (uint32_t)(li[u] >= 0) << u;
      }
      apply_vals_lds_batch<NV, U>(wp, ws, P.pg, v, P.vfloat, all_valid, in_lds);
      slow |= ok & ~in_lds;  // more groups than the LDS table takes: the general kernel aggregates the row
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (!((slow >> u) & 1u)) continue;
        // in bounds: tiles are taken only with headroom for all their rows
        const uint64_t at =
            __hip_atomic_fetch_add(P.T.counters + CTR_OVF_ROWS, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        P.ovf_rows[at] = (uint32_t)row[u];
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < L; i += blockDim.x)
    if (lds[(uint64_t)i * R] != kEmpty) fold_replicas(lds + (uint64_t)i * R, ws, R, P.pg);
  __syncthreads();
  for (int i = threadIdx.x; i < L; i += blockDim.x) {
    const uint64_t* ls = lds + (uint64_t)i * R;
    const uint64_t key = ls[0];
    if (key == kEmpty) continue;
    uint64_t* slot = glb_find_or_claim(P.T, stride, key, key_hash(key), claims);
    if (slot != nullptr) {
      merge_partial(slot, ls, ws, P.pg);
    } else {
      const uint64_t at =
          __hip_atomic_fetch_add(P.T.counters + CTR_OVF_PART, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      uint64_t* rec = P.ovf_part + at * (uint64_t)stride;  // in bounds: L partials per workgroup are reserved
      for (int w = 0; w < stride; ++w) rec[w] = ls[(uint64_t)w * ws];
    }
  }
  block_add(P.T.counters + CTR_GROUPS, claims);
}

template <int NV>
static wg_status launch_nv(const JoinLeanParams& P, int grid, int block, size_t lds, hipStream_t st) {
  static bool attr_done = false;
  if (!attr_done) {
    WG_HIP(hipFuncSetAttribute((const void*)join_agg_lean_kernel<NV>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               150 * 1024));
    attr_done = true;
  }
  hipLaunchKernelGGL(join_agg_lean_kernel<NV>, dim3(grid), dim3(block), lds, st, P);
  WG_HIP(hipGetLastError());
  return WG_OK;
}

wg_status launch_join_agg_lean(const JoinLeanParams& P, int grid, int block, hipStream_t st) {
  WG_CHECK(P.n_vals >= 0 && P.n_vals <= kLeanMaxVals && P.J.stride >= 4 && P.L >= 2 && (P.L & (P.L - 1)) == 0 &&
               P.R >= 1 && P.R <= kLeanMaxReplicas && (P.R & (P.R - 1)) == 0 && (block == 512 || block == 1024),
           WG_ERR_INVALID, "streamlined join aggregate: shape out of range");
  const size_t lds = lean_lds_bytes(P.L, P.R, P.pg.stride);
  WG_CHECK(lds <= 150 * 1024, WG_ERR_INVALID, "streamlined join aggregate: %zu LDS bytes", lds);
  switch (P.n_vals) {
    case 0:
    case 1: return launch_nv<1>(P, grid, block, lds, st);
    case 2: return launch_nv<2>(P, grid, block, lds, st);
    case 3: return launch_nv<3>(P, grid, block, lds, st);
    default: return launch_nv<4>(P, grid, block, lds, st);
  }
}

}  // namespace wg
