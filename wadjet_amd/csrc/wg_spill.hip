// wg_spill.hip — the aggregate's groups as one partial-state spill run in
// wadjet's external-merge format, so a GPU aggregate can hand its partial
// state to the CPU operator's k-way merge (finalizeViaPartialMerge) or park it
// outside HBM.
//
// Format (internal/engine/exec/aggregate_partial_spill.go:19-52, little-endian):
//   "WAGS\x01" | u32 group cols | per col: u16 name len, name, u16 parquet.TypeID
//   | u32 aggs | per agg: u32 AggFunc, u8 isFloat, u8 isDecimal, i32 decScale
//   | u32 groups | per group, ascending by sort key (bytes.Compare):
//       u8 0x01, u32 sort key len, sort key, per group col a tagged value
//       (writeKeyValue :465-536), per agg the accumulator (emitAcc :302-351)
//   | u8 0x00
// The sort key of an int-class group is the decimal text of each column value
// joined by a 0 byte (appendIntModeSortKey / appendTypedIntKey,
// aggregate_partial_drain_cursor.go:421-452); a NULL group value is "<null>"
// (appendKeyValue, sort.go:992-995). The accumulator flags follow
// initFlatAccums (aggregate.go:6030-6119): SUM/AVG over float64 and AVG over
// int64 are float sums, SUM/AVG over DECIMAL are Int128 sums at the column's
// scale, MIN/MAX over DECIMAL have no scatter kernel and stay empty.
// Everything here is host code over the finalized dense state.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "wg_agg_host.hpp"

namespace wg {

namespace {

// parquet.TypeID (internal/storage/parquet/schema.go:12-35) of a group column
int parquet_type_of(int t) {
  switch (t) {
    case WG_INT32: return 1;   // TypeInt32
    case WG_UINT16: return 12; // TypePort
    case WG_IPV4: case WG_UINT32: return 8;  // TypeIPv4
    default: return 2;         // TypeInt64
  }
}

struct Out {
  std::vector<uint8_t> b;
  void u8(uint8_t v) { b.push_back(v); }
  void u16(uint16_t v) { for (int i = 0; i < 2; ++i) b.push_back((uint8_t)(v >> (8 * i))); }
  void u32(uint32_t v) { for (int i = 0; i < 4; ++i) b.push_back((uint8_t)(v >> (8 * i))); }
  void u64(uint64_t v) { for (int i = 0; i < 8; ++i) b.push_back((uint8_t)(v >> (8 * i))); }
  void bytes(const void* p, size_t n) { b.insert(b.end(), (const uint8_t*)p, (const uint8_t*)p + n); }
};

// The reference accumulator flags of aggregate a.
struct SpillSpec {
  int func;
  bool is_float, is_decimal;
  int32_t scale;
};

SpillSpec spec_of(const wg_agg* h, int a) {
  const AggFinal& f = h->finals[a];
  const wg_agg_spec& s = h->aggs[a];
  SpillSpec r{s.func, false, false, 0};
  if (s.input_col < 0 || s.func == WG_AGG_COUNT) return r;
  const int t = f.in_type;
  switch (s.func) {
    case WG_AGG_SUM:
    case WG_AGG_AVG:
      if (t == WG_FLOAT64) r.is_float = true;
      else if (t == WG_DECIMAL || t == WG_DECIMAL128) { r.is_decimal = true; r.scale = f.scale; }
      else if (t == WG_INT64 && s.func == WG_AGG_AVG) r.is_float = true;
      break;
    default:  // MIN / MAX
      if (t == WG_FLOAT64) r.is_float = true;
      else if (t == WG_DECIMAL || t == WG_DECIMAL128) r.is_decimal = true;
      break;
  }
  return r;
}

// Value of group column k of a dense slot's packed key (agg_key_kernel's rule).
int64_t key_value(const KeyDesc& K, uint64_t w, int k) {
  if (K.n == 1) return (int64_t)w;
  if (K.narrow[k] == 1) return (int64_t)(int32_t)(uint32_t)(w >> K.shift[k]);
  if (K.narrow[k] == 2) return (int64_t)(uint32_t)(w >> K.shift[k]);
  return (int64_t)w;
}

bool int32_tagged(int t) { return t == WG_INT32 || t == WG_UINT16; }

}  // namespace

}  // namespace wg

using namespace wg;

extern "C" {

wg_status wg_agg_spill_run(wg_agg* h, const char* const* group_names, const uint16_t* group_type_ids, void* out,
                           uint64_t cap, uint64_t* size) {
  clear_error();
  WG_CHECK(h && size, WG_ERR_INVALID, "null argument");
  WG_CHECK(h->finalized, WG_ERR_STATE, "spill run before finalize");
  const uint64_t n = h->out_n;
  WG_CHECK(n < (1ull << 32), WG_ERR_UNSUPPORTED, "a run holds fewer than 2^32 groups");
  const int S = h->pg.stride;
  std::vector<uint64_t> slots(n * S), base;
  std::vector<uint8_t> nulls(n);
  if (n) {
    WG_HIP(hipMemcpyAsync(slots.data(), h->dense, n * S * 8, hipMemcpyDeviceToHost, h->ctx->stream));
    WG_HIP(hipMemcpyAsync(nulls.data(), h->dense_null, n, hipMemcpyDeviceToHost, h->ctx->stream));
    if (h->n_sumd) {
      base.resize(n * h->n_sumd * 2);
      WG_HIP(hipMemcpyAsync(base.data(), h->dense_base, base.size() * 8, hipMemcpyDeviceToHost, h->ctx->stream));
    }
    WG_HIP(hipStreamSynchronize(h->ctx->stream));
  }
  const KeyDesc& K = h->wide ? h->wkey : h->key;
  std::vector<uint64_t> dict;  // 128-bit keys: id -> {lo, hi}
  if (h->wide) WG_TRY(wide_dict_host(h, &dict));
  // group column k of dense slot j
  auto kv = [&](uint64_t j, int k) -> int64_t {
    const uint64_t w = slots[j * S];
    return key_value(K, h->wide ? dict[2 * w + K.word[k]] : w, k);
  };
  const int na = (int)h->aggs.size();
  // sort keys, then the order of the groups
  std::vector<std::string> sk(n);
  for (uint64_t j = 0; j < n; ++j) {
    std::string& s = sk[j];
    for (int k = 0; k < K.n; ++k) {
      if (k) s.push_back('\0');
      if (nulls[j]) s += "<null>";
      else s += std::to_string(kv(j, k));
    }
  }
  std::vector<uint32_t> order(n);
  for (uint64_t j = 0; j < n; ++j) order[j] = (uint32_t)j;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return sk[a] < sk[b]; });

  Out o;
  o.bytes("WAGS\x01", 5);
  o.u32((uint32_t)K.n);
  for (int k = 0; k < K.n; ++k) {
    const char* nm = group_names && group_names[k] ? group_names[k] : "";
    const size_t len = strlen(nm);
    WG_CHECK(len < 65536, WG_ERR_INVALID, "group column name of %zu bytes", len);
    o.u16((uint16_t)len);
    o.bytes(nm, len);
    o.u16(group_type_ids ? group_type_ids[k] : (uint16_t)parquet_type_of(K.key_type[k]));
  }
  std::vector<SpillSpec> specs(na);
  o.u32((uint32_t)na);
  for (int a = 0; a < na; ++a) {
    specs[a] = spec_of(h, a);
    o.u32((uint32_t)specs[a].func);
    o.u8(specs[a].is_float ? 1 : 0);
    o.u8(specs[a].is_decimal ? 1 : 0);
    o.u32((uint32_t)specs[a].scale);
  }
  o.u32((uint32_t)n);
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t j = order[i];
    const uint64_t* s = slots.data() + (uint64_t)j * S;
    o.u8(0x01);
    o.u32((uint32_t)sk[j].size());
    o.bytes(sk[j].data(), sk[j].size());
    for (int k = 0; k < K.n; ++k) {
      if (nulls[j]) {
        o.u8(0);  // partialTagNull
      } else if (int32_tagged(K.key_type[k])) {
        o.u8(6);  // partialTagInt32
        o.u32((uint32_t)(int32_t)kv(j, k));
      } else {
        o.u8(3);  // partialTagInt64
        o.u64((uint64_t)kv(j, k));
      }
    }
    for (int a = 0; a < na; ++a) {
      const AggFinal& f = h->finals[a];
      const SpillSpec& sp = specs[a];
      // a sum with no scatter kernel (IPv4 inputs) never counted a row
      const int64_t cnt = (sp.func == WG_AGG_COUNT || f.val_kind) ? (int64_t)s[f.count_word] : 0;
      if (sp.func == WG_AGG_COUNT) {
        o.u64((uint64_t)cnt);
      } else if (sp.func == WG_AGG_SUM || sp.func == WG_AGG_AVG) {
        o.u64((uint64_t)cnt);
        if (sp.is_decimal) {
          uint64_t lo = 0;
          int64_t hi = 0;
          if (f.val_kind == OP_SUMD || f.val_kind == OP_SUMD4) {
            limbs_value(s + f.val_word, f.limbs, lo, hi);
            const uint64_t* b = base.data() + ((uint64_t)j * h->n_sumd + f.sumd_index) * 2;
            const unsigned __int128 v =
                (((unsigned __int128)(uint64_t)hi << 64) | lo) + (((unsigned __int128)b[1] << 64) | b[0]);
            lo = (uint64_t)v;
            hi = (int64_t)(uint64_t)(v >> 64);
          }
          o.u64((uint64_t)hi);  // writeInt128: Hi, then Lo
          o.u64(lo);
        } else {
          o.u64(f.val_kind ? s[f.val_word] : 0);  // SumF64 bits or SumI64
        }
      } else {  // MIN / MAX
        const uint64_t e = f.val_kind ? s[f.val_word] : 0;
        const uint64_t empty = sp.func == WG_AGG_MIN ? kMinInit : kMaxInit;
        const bool has = f.val_kind != 0 && !sp.is_decimal && e != empty;
        o.u8(has ? 1 : 0);
        if (has) o.u64(f.is_float_in ? dec_f64_bits(e) : (uint64_t)dec_i64(e));
      }
    }
  }
  o.u8(0x00);
  *size = o.b.size();
  if (out == nullptr) return WG_OK;
  WG_CHECK(cap >= o.b.size(), WG_ERR_INVALID, "run of %zu bytes does not fit %llu", o.b.size(),
           (unsigned long long)cap);
  memcpy(out, o.b.data(), o.b.size());
  return WG_OK;
}

static wg_status merge_run_call(wg_agg* h, const void* run, uint64_t size) {
  WG_TRY(agg_settle(h));
  const uint8_t* p = (const uint8_t*)run;
  const uint8_t* const end = p + size;
  bool ok = true;
  auto need = [&](uint64_t n) { ok = ok && (uint64_t)(end - p) >= n; return ok; };
  auto rd = [&](int bytes) -> uint64_t {
    uint64_t v = 0;
    if (!need(bytes)) return 0;
    for (int i = 0; i < bytes; ++i) v |= (uint64_t)p[i] << (8 * i);
    p += bytes;
    return v;
  };
  WG_CHECK(need(5) && memcmp(p, "WAGS\x01", 5) == 0, WG_ERR_INVALID, "not a partial-state spill run");
  p += 5;
  const KeyDesc& K = h->wide ? h->wkey : h->key;
  const int na = (int)h->aggs.size();
  WG_CHECK(rd(4) == (uint64_t)K.n && ok, WG_ERR_INVALID, "run has another number of group columns");
  for (int k = 0; k < K.n && ok; ++k) {
    const uint64_t len = rd(2);
    if (need(len)) p += len;
    (void)rd(2);
  }
  WG_CHECK(ok && rd(4) == (uint64_t)na && ok, WG_ERR_INVALID, "run has another number of aggregates");
  std::vector<SpillSpec> specs(na);
  for (int a = 0; a < na; ++a) {
    const SpillSpec sp = specs[a] = spec_of(h, a);
    const uint64_t func = rd(4), is_f = rd(1), is_d = rd(1), scale = rd(4);
    WG_CHECK(ok && func == (uint64_t)sp.func && (is_f != 0) == sp.is_float && (is_d != 0) == sp.is_decimal &&
                 (int32_t)scale == sp.scale,
             WG_ERR_INVALID, "aggregate %d of the run differs from this aggregate", a);
  }
  const uint64_t n = rd(4);
  WG_CHECK(ok, WG_ERR_INVALID, "truncated run header");
  // the run as a table of unique keys: n slots, then the two side slots
  const int S = h->pg.stride;
  const int nb = h->n_sumd * 2;
  std::vector<uint64_t> slots((n + kSideSlots) * S), base((n + kSideSlots) * (uint64_t)std::max(nb, 1));
  for (uint64_t j = 0; j < n + kSideSlots; ++j)
    for (int w = 0; w < S; ++w) slots[j * S + w] = slot_init_word(h->pg, w);
  for (uint64_t j = n; j < n + kSideSlots; ++j) slots[j * S + 1] = 0;  // a side slot is live once it has rows
  uint64_t g = 0;
  std::vector<uint64_t> pairs;  // 128-bit keys: {lo, hi} per group, in table order
  for (uint64_t i = 0; i < n; ++i) {
    WG_CHECK(rd(1) == 1 && ok, WG_ERR_INVALID, "bad row marker in group %llu", (unsigned long long)i);
    const uint64_t sl = rd(4);
    WG_CHECK(need(sl), WG_ERR_INVALID, "truncated sort key");
    p += sl;
    uint64_t key = 0, key_hi = 0;
    bool null_key = false;
    for (int k = 0; k < K.n; ++k) {
      const uint64_t tag = rd(1);
      int64_t v = 0;
      if (tag == 0) null_key = true;
      else if (tag == 6) v = (int64_t)(int32_t)(uint32_t)rd(4);
      else if (tag == 3) v = (int64_t)rd(8);
      else ok = false;
      WG_CHECK(ok, WG_ERR_INVALID, "group value tag %llu not on the GPU path", (unsigned long long)tag);
      uint64_t w = (uint64_t)v;
      if (K.n > 1 && K.narrow[k]) w = (w & 0xFFFFFFFFull) << K.shift[k];
      if (h->wide && K.word[k]) key_hi |= w;
      else key |= w;
    }
    if (h->wide) {  // the pair goes through the dictionary below
      pairs.push_back(key);
      pairs.push_back(key_hi);
    }
    WG_CHECK(!null_key || K.n == 1, WG_ERR_NULL_KEY, "NULL in a multi-column group key");
    // side slots: [n] the sentinel-valued key, [n + 1] the NULL key
    const uint64_t j = null_key ? n + 1 : (key == kEmpty && !h->wide) ? n : g++;
    uint64_t* s = slots.data() + j * S;
    s[0] = null_key || (key == kEmpty && !h->wide) ? kEmpty : key;
    uint64_t* b = base.data() + j * (uint64_t)std::max(nb, 1);
    int64_t rows = -1;  // the group's row count, from an aggregate counted on word 1
    for (int a = 0; a < na; ++a) {
      const AggFinal& f = h->finals[a];
      const SpillSpec& sp = specs[a];
      if (sp.func == WG_AGG_COUNT || sp.func == WG_AGG_SUM || sp.func == WG_AGG_AVG) {
        const int64_t cnt = (int64_t)rd(8);
        if (f.val_kind || sp.func == WG_AGG_COUNT) {
          s[f.count_word] = (uint64_t)cnt;
          if (f.count_word == 1) rows = cnt;
        }
        if (sp.func == WG_AGG_COUNT) continue;
        if (sp.is_decimal) {
          const uint64_t hi = rd(8), lo = rd(8);
          if (f.val_kind == OP_SUMD || f.val_kind == OP_SUMD4) {
            b[f.sumd_index * 2] = lo;
            b[f.sumd_index * 2 + 1] = hi;
          }
        } else {
          const uint64_t v = rd(8);
          if (f.val_kind == OP_SUMI || f.val_kind == OP_SUMF) s[f.val_word] = v;
        }
      } else {
        const uint64_t has = rd(1);
        if (has && sp.is_decimal) {
          // an Int128 MinDec / MaxDec (Hi, Lo: readAcc, aggregate_partial_spill.go:400-440); the GPU
          // path emits MIN/MAX(DECIMAL) as NULL like the flat accumulators, so it is read and dropped
          (void)rd(8);
          (void)rd(8);
        } else if (has) {
          const uint64_t v = rd(8);
          if (f.val_kind == OP_MIN || f.val_kind == OP_MAX) s[f.val_word] = f.is_float_in ? enc_f64_bits(v) : enc_i64((int64_t)v);
        }
      }
    }
    WG_CHECK(ok, WG_ERR_INVALID, "truncated group %llu", (unsigned long long)i);
    s[1] = (uint64_t)(rows >= 0 ? rows : 1);  // liveness when no aggregate carries the row count
  }
  WG_CHECK(rd(1) == 0 && ok && p == end, WG_ERR_INVALID, "run does not end after %llu groups", (unsigned long long)n);
  if (n == 0) return WG_OK;
  // the side-slot records move to just after the g table slots
  for (uint64_t k = 0; k < kSideSlots; ++k) {
    std::copy(slots.begin() + (n + k) * S, slots.begin() + (n + k + 1) * S, slots.begin() + (g + k) * S);
    if (nb) std::copy(base.begin() + (n + k) * nb, base.begin() + (n + k + 1) * nb, base.begin() + (g + k) * nb);
  }
  WG_TRY(agg_flush_stage(h));
  if (h->wide && g) {
    // the pairs through this aggregate's dictionary: the table keys on ids
    uint64_t* dp = nullptr;
    WG_ALLOC(h->ctx, &dp, g * 24, "spill-run groups");
    uint64_t* did = dp + 2 * g;
    std::vector<uint64_t> ids(g);
    wg_status st = hipMemcpy(dp, pairs.data(), g * 16, hipMemcpyHostToDevice) == hipSuccess ? WG_OK : WG_ERR_DEVICE;
    if (st == WG_OK) st = wide_encode_pairs(h, dp, g, did);
    if (st == WG_OK && hipMemcpy(ids.data(), did, g * 8, hipMemcpyDeviceToHost) != hipSuccess) st = WG_ERR_DEVICE;
    (void)hipFree(dp);
    if (st != WG_OK) {
      if (st == WG_ERR_DEVICE) set_error("dictionary encode of a run's keys failed");
      return st;
    }
    for (uint64_t j = 0; j < g; ++j) slots[j * S] = ids[j];
  }
  uint64_t* d_slots = nullptr;
  uint64_t* d_base = nullptr;
  WG_ALLOC(h->ctx, &d_slots, (g + kSideSlots) * S * 8, "spill-run groups");
  wg_status st = WG_OK;
  if (nb) st = dev_alloc(h->ctx, (void**)&d_base, (g + kSideSlots) * nb * 8, nullptr, "spill-run bases");
  if (st == WG_OK && hipMemcpy(d_slots, slots.data(), (g + kSideSlots) * S * 8, hipMemcpyHostToDevice) != hipSuccess)
    st = WG_ERR_DEVICE;
  if (st == WG_OK && nb && hipMemcpy(d_base, base.data(), (g + kSideSlots) * nb * 8, hipMemcpyHostToDevice) != hipSuccess)
    st = WG_ERR_DEVICE;
  if (st == WG_OK) st = agg_merge_slots(h, d_slots, g, d_base, g);
  (void)hipStreamSynchronize(h->ctx->stream);
  (void)hipFree(d_slots);
  if (d_base) (void)hipFree(d_base);
  if (st != WG_OK && !wg_last_error()[0]) set_error("merging a spill run failed");
  return st;
}

wg_status wg_agg_merge_run(wg_agg* h, const void* run, uint64_t size) {
  clear_error();
  WG_CHECK(h && (run || size == 0), WG_ERR_INVALID, "null argument");
  WG_CHECK(!h->finalized, WG_ERR_STATE, "merge after finalize");
  WG_TRY(agg_guard_begin(h));
  return agg_guard_end(h, merge_run_call(h, run, size));
}

}  // extern "C"
