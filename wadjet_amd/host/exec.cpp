// exec.cpp — the C++ operator layer over the wadjet_gpu C-ABI (see exec.hpp).
#include "exec.hpp"

#include <algorithm>
#include <cctype>
#include <cstring>
#include <thread>

namespace wadjet {

using batch::Column;
using batch::TypeID;
using batch::Vector;

static Status wrap(wg_status st, const char* what) {
  if (st == WG_OK) return Status::OK();
  return Status::Error(std::string(what) + ": " + wg_last_error(), st);
}
#define TRY(x)                       \
  do {                               \
    Status _s = (x);                 \
    if (!_s.ok()) return _s;         \
  } while (0)

// ------------------------------------------------------------- batch helpers

void batch::Bitmap::set_null(int i, int n) {
  const size_t need = (size_t)(std::max(n, i + 1) + 63) / 64;
  if (words.size() < need) words.resize(need, ~0ull);  // a vector that grew keeps its earlier bits
  words[i >> 6] &= ~(1ull << (i & 63));
}

LateMatCounters LateMat;

std::shared_ptr<Vector> Vector::make(TypeID t, int n, int scale) {
  auto v = std::make_shared<Vector>();
  v->type = t;
  v->len = n;
  v->scale = scale;
  switch (t) {
    case TypeID::Int32: case TypeID::Port: case TypeID::Date: v->i32.resize(n); break;
    case TypeID::Float64: v->f64.resize(n); break;
    case TypeID::Decimal: v->dec.resize(n); break;
    default: v->i64.resize(n); break;
  }
  return v;
}

int RecordBatch::column_index(const std::string& name) const {
  for (size_t i = 0; i < schema.size(); ++i)
    if (schema[i].name == name) return (int)i;
  // strip a table qualifier ("n1.n_name" -> "n_name"), as KernelFilter does
  const auto dot = name.find('.');
  if (dot != std::string::npos) return column_index(name.substr(dot + 1));
  return -1;
}

static int32_t wg_type_of(TypeID t) {
  switch (t) {
    case TypeID::Int32: case TypeID::Port: case TypeID::Date: return WG_INT32;
    case TypeID::Float64: return WG_FLOAT64;
    case TypeID::Decimal: return WG_DECIMAL128;
    case TypeID::IPv4: return WG_IPV4;
    default: return WG_INT64;
  }
}

static const void* data_of(const Vector& v) {
  switch (v.type) {
    case TypeID::Int32: case TypeID::Port: case TypeID::Date: return v.i32.data();
    case TypeID::Float64: return v.f64.data();
    case TypeID::Decimal: return v.dec.data();
    default: return v.i64.data();
  }
}

static int elem_bytes(TypeID t) {
  switch (t) {
    case TypeID::Int32: case TypeID::Port: case TypeID::Date: return 4;
    case TypeID::Decimal: return 16;
    default: return 8;
  }
}

static wg_column col_of(const Vector& v) {
  return wg_column{data_of(v), v.nulls.has_nulls() ? v.nulls.words.data() : nullptr, wg_type_of(v.type), v.scale};
}

void Vector::Flatten() {
  if (!base) return;
  const int n = (int)indices.size();
  auto flat = Vector::make(base->type, n, base->scale);
  const int w = elem_bytes(base->type);
  const char* s = (const char*)data_of(*base);
  char* d = (char*)data_of(*flat);
  for (int i = 0; i < n; ++i) {
    const uint32_t si = indices[i];
    if (!nulls.valid(i) || !base->nulls.valid((int)si)) {  // own NULL: the index is not read
      flat->nulls.set_null(i, n);
      continue;
    }
    memcpy(d + (size_t)i * w, s + (size_t)si * w, w);
  }
  type = flat->type;
  len = n;
  scale = flat->scale;
  nulls = std::move(flat->nulls);
  i32 = std::move(flat->i32);
  i64 = std::move(flat->i64);
  f64 = std::move(flat->f64);
  dec = std::move(flat->dec);
  base.reset();
  indices.clear();
}

std::shared_ptr<Vector> batch::NewViewVector(const std::shared_ptr<Vector>& base, std::vector<uint32_t> indices) {
  auto v = std::make_shared<Vector>();
  v->type = base->type;
  v->scale = base->scale;
  v->len = (int)indices.size();
  if (base->is_view()) {  // compose: fold the base view's indirection and own NULLs
    for (size_t i = 0; i < indices.size(); ++i) {
      const uint32_t x = indices[i];
      if (!base->nulls.valid((int)x)) v->nulls.set_null((int)i, v->len);
      indices[i] = base->indices[x];
    }
    v->base = base->base;
  } else {
    v->base = base;
  }
  v->indices = std::move(indices);
  return v;
}

bool RecordBatch::HasViews() const {
  for (auto& c : columns)
    if (c->is_view()) return true;
  return false;
}

void RecordBatch::FlattenViews() {
  for (auto& c : columns) c->Flatten();
}

// copy rows[idx] of src into a fresh vector
static std::shared_ptr<Vector> gather(const Vector& src, const uint32_t* idx, size_t n) {
  auto v = Vector::make(src.type, (int)n, src.scale);
  const int w = elem_bytes(src.type);
  const char* s = (const char*)data_of(src);
  char* d = (char*)data_of(*v);
  for (size_t i = 0; i < n; ++i) memcpy(d + i * w, s + (size_t)idx[i] * w, w);
  if (src.nulls.has_nulls())
    for (size_t i = 0; i < n; ++i)
      if (!src.nulls.valid(idx[i])) v->nulls.set_null((int)i, (int)n);
  return v;
}

// ------------------------------------------------------------- CPU operators

Status KernelFilter::Execute(Context&, const BatchPtr& in, BatchPtr* out) {
  const int ci = in->column_index(col_);
  if (ci < 0) return Status::Error("filter column \"" + col_ + "\" does not exist in the input schema");
  const Vector& v = *in->columns[ci];
  auto pass = [&](int r) -> bool {
    if (!v.nulls.valid(r)) return false;
    switch (v.type) {
      case TypeID::Float64: {
        const double a = v.f64[r], b = value_;
        switch (op_) {
          case CompareOp::Eq: return a == b;
          case CompareOp::Ne: return a != b;
          case CompareOp::Lt: return a < b;
          case CompareOp::Le: return a <= b;
          case CompareOp::Gt: return a > b;
          default: return a >= b;
        }
      }
      case TypeID::Int32: case TypeID::Port: case TypeID::Date: {
        const int32_t a = v.i32[r], b = (int32_t)(int64_t)value_;
        switch (op_) {
          case CompareOp::Eq: return a == b;
          case CompareOp::Ne: return a != b;
          case CompareOp::Lt: return a < b;
          case CompareOp::Le: return a <= b;
          case CompareOp::Gt: return a > b;
          default: return a >= b;
        }
      }
      case TypeID::Decimal: return false;
      default: {
        const int64_t a = v.i64[r], b = (int64_t)value_;
        switch (op_) {
          case CompareOp::Eq: return a == b;
          case CompareOp::Ne: return a != b;
          case CompareOp::Lt: return a < b;
          case CompareOp::Le: return a <= b;
          case CompareOp::Gt: return a > b;
          default: return a >= b;
        }
      }
    }
  };
  if (v.type == TypeID::Decimal) return Status::Error("no filter kernel for DECIMAL column \"" + col_ + "\"");
  auto o = std::make_shared<RecordBatch>(*in);  // shares the column vectors
  o->sel.clear();
  if (in->has_sel) {
    for (uint32_t r : in->sel)
      if (pass((int)r)) o->sel.push_back(r);
  } else {
    for (int r = 0; r < in->len; ++r)
      if (pass(r)) o->sel.push_back((uint32_t)r);
  }
  o->has_sel = true;
  *out = o->sel.empty() ? nullptr : o;
  return Status::OK();
}

Status BatchSource::Next(Context&, BatchPtr* out) {
  std::lock_guard<std::mutex> lk(mu_);
  *out = pos_ < batches_.size() ? batches_[pos_++] : nullptr;
  return Status::OK();
}

Status CollectSink::Consume(Context&, const BatchPtr& b) {
  std::vector<uint32_t> idx;
  if (b->has_sel) idx = b->sel;
  else {
    idx.resize(b->len);
    for (int i = 0; i < b->len; ++i) idx[i] = (uint32_t)i;
  }
  auto c = std::make_shared<RecordBatch>();
  c->schema = b->schema;
  c->len = (int)idx.size();
  for (auto& v : b->columns) c->columns.push_back(gather(*v, idx.data(), idx.size()));
  std::lock_guard<std::mutex> lk(mu_);
  batches.push_back(c);
  return Status::OK();
}

int CollectSink::rows() const {
  int n = 0;
  for (auto& b : batches) n += b->len;
  return n;
}

// --------------------------------------------------------------- pipeline

// push one batch through ops[i..] and into the sink, draining bounded output
// after every Execute before the next input (ChainDriver, pipeline.go:196-235)
static Status push(Context& ctx, std::vector<UnaryOperator*>& ops, size_t i, Sink* sink, const BatchPtr& b) {
  if (ctx.cancelled) return Status::Error("context canceled");
  if (i == ops.size()) {
    FlattenForConsumer(b.get(), sink);
    return sink->Consume(ctx, b);
  }
  FlattenForConsumer(b.get(), ops[i]);
  BatchPtr out;
  TRY(ops[i]->Execute(ctx, b, &out));
  if (out) TRY(push(ctx, ops, i + 1, sink, out));
  if (auto* bo = dynamic_cast<BoundedOutputOperator*>(ops[i])) {
    while (bo->HasPendingOutput()) {
      BatchPtr more;
      TRY(bo->NextOutput(ctx, &more));
      if (more) TRY(push(ctx, ops, i + 1, sink, more));
    }
  }
  return Status::OK();
}

Status Pipeline::Run(Context& ctx) {
  if (!source || !sink) return Status::Error("pipeline needs a source and a sink");
  TRY(source->Init(ctx));
  TRY(sink->Init(ctx));
  for (auto* op : ops) {
    TRY(op->Init(ctx));
    if (auto* bo = dynamic_cast<BoundedOutputOperator*>(op)) bo->EnableBoundedOutput();
  }
  auto* ms = dynamic_cast<MergeableSink*>(sink);
  bool cloneable = ms != nullptr;
  for (auto* op : ops) cloneable = cloneable && dynamic_cast<Cloneable*>(op) != nullptr;
  if (workers <= 1 || !cloneable) {
    for (;;) {
      BatchPtr b;
      TRY(source->Next(ctx, &b));
      if (!b) break;
      TRY(push(ctx, ops, 0, sink, b));
    }
    return sink->Finalize(ctx);
  }
  // parallel: per-worker op chains and sinks, merged into the primary
  std::vector<std::vector<std::unique_ptr<UnaryOperator>>> chains(workers);
  std::vector<std::unique_ptr<SinkSource>> sinks(workers);
  for (int w = 0; w < workers; ++w) {
    for (auto* op : ops) {
      chains[w].push_back(dynamic_cast<Cloneable*>(op)->Clone());
      TRY(chains[w].back()->Init(ctx));
      if (auto* bo = dynamic_cast<BoundedOutputOperator*>(chains[w].back().get())) bo->EnableBoundedOutput();
    }
    sinks[w] = ms->CloneSink();
    TRY(sinks[w]->Init(ctx));
  }
  std::vector<Status> st(workers);
  std::vector<std::thread> th;
  for (int w = 0; w < workers; ++w) {
    th.emplace_back([&, w] {
      std::vector<UnaryOperator*> chain;
      for (auto& o : chains[w]) chain.push_back(o.get());
      for (;;) {
        BatchPtr b;
        Status s = source->Next(ctx, &b);
        if (!s.ok()) { st[w] = s; return; }
        if (!b) return;
        s = push(ctx, chain, 0, sinks[w].get(), b);
        if (!s.ok()) { st[w] = s; return; }
      }
    });
  }
  for (auto& t : th) t.join();
  for (auto& s : st) TRY(s);
  for (int w = 0; w < workers; ++w) TRY(ms->MergeSink(sinks[w].get()));
  return sink->Finalize(ctx);
}

// ------------------------------------------------------- GpuHashAggregate

GpuHashAggregate::GpuHashAggregate(wg_ctx* ctx, std::vector<std::string> group_by, std::vector<AggColumn> aggs,
                                   uint64_t group_hint)
    : ctx_(ctx), group_by_(std::move(group_by)), aggs_(std::move(aggs)), hint_(group_hint) {}

GpuHashAggregate::~GpuHashAggregate() { (void)Close(); }

Status GpuHashAggregate::Close() {
  if (h_) wg_agg_destroy(h_);
  h_ = nullptr;
  return Status::OK();
}

Status GpuHashAggregate::AddPredicate(const KernelFilter& f) {
  if (pending_preds_.size() >= 2) return Status::Error("at most two pushed-down predicates");
  wg_predicate p{-1, (int32_t)f.op(), (int64_t)f.value(), f.value()};
  pending_preds_.push_back({f.column(), p});
  return Status::OK();
}

// resolveIndices (aggregate.go:1296-1764): map names to columns once, on the
// first batch, and create the typed device aggregate over the referenced
// columns only (keys first, then aggregate inputs).
Status GpuHashAggregate::resolve(const RecordBatch& b) {
  std::vector<int> used;
  auto slot = [&](const std::string& name) -> int {
    const int ci = b.column_index(name);
    if (ci < 0) return -1;
    for (size_t i = 0; i < used.size(); ++i)
      if (used[i] == ci) return (int)i;
    used.push_back(ci);
    return (int)used.size() - 1;
  };
  std::vector<int32_t> keys;
  for (auto& g : group_by_) {
    const int k = slot(g);
    if (k < 0) return Status::Error("GROUP BY column \"" + g + "\" not in the input schema");
    keys.push_back(k);
  }
  std::vector<wg_agg_spec> specs;
  for (auto& a : aggs_) {
    wg_agg_spec s{a.func, -1, -1, 0};
    if (!a.input_col.empty()) {
      s.input_col = slot(a.input_col);
      if (s.input_col < 0) return Status::Error("aggregate input \"" + a.input_col + "\" not in the input schema");
    }
    if (!a.input_col2.empty()) {
      s.input_col2 = slot(a.input_col2);
      if (s.input_col2 < 0) return Status::Error("aggregate input \"" + a.input_col2 + "\" not in the input schema");
    }
    specs.push_back(s);
  }
  preds_.clear();
  for (auto& pp : pending_preds_) {
    wg_predicate p = pp.second;
    p.col = slot(pp.first);
    if (p.col < 0) return Status::Error("filter column \"" + pp.first + "\" not in the input schema");
    preds_.push_back(p);
  }
  key_idx_.assign(used.begin(), used.end());
  in_schema_.clear();
  std::vector<int32_t> types, scales;
  std::vector<uint8_t> nullable;
  for (int ci : used) {
    in_schema_.push_back(b.schema[ci]);
    types.push_back(wg_type_of(b.schema[ci].type));
    scales.push_back(b.schema[ci].scale);
    nullable.push_back(b.schema[ci].nullable ? 1 : 0);
  }
  return wrap(wg_agg_create(ctx_, (int32_t)used.size(), types.data(), scales.data(), nullable.data(),
                            (int32_t)keys.size(), keys.data(), (int32_t)specs.size(), specs.data(), hint_, &h_),
              "wg_agg_create");
}

Status GpuHashAggregate::Consume(Context& ctx, const BatchPtr& b) {
  if (ctx.cancelled) return Status::Error("context canceled");
  std::lock_guard<std::mutex> lk(mu_);
  if (finalized_) return Status::Error("Consume after Finalize");
  FlattenForConsumer(b.get(), this);  // not view-aware: reads typed storage
  if (!h_) TRY(resolve(*b));
  std::vector<wg_column> cols;
  for (int ci : key_idx_) cols.push_back(col_of(*b->columns[ci]));
  return wrap(wg_agg_consume(h_, cols.data(), (uint64_t)b->len, b->has_sel ? b->sel.data() : nullptr,
                             b->has_sel ? b->sel.size() : 0, preds_.data(), (int32_t)preds_.size(), WG_MEM_HOST),
              "wg_agg_consume");
}

Status GpuHashAggregate::Finalize(Context&) {
  std::lock_guard<std::mutex> lk(mu_);
  if (finalized_) return Status::OK();
  if (h_) {  // no input: no groups
    TRY(wrap(wg_agg_finalize(h_), "wg_agg_finalize"));
    TRY(wrap(wg_agg_num_groups(h_, &groups_), "wg_agg_num_groups"));
  }
  finalized_ = true;  // a failed finalize can be called again
  return Status::OK();
}

std::vector<Column> GpuHashAggregate::OutputSchema() const {
  std::vector<Column> out;
  for (size_t k = 0; k < group_by_.size(); ++k) {
    Column c = in_schema_.empty() ? Column{group_by_[k], TypeID::Int64} : in_schema_[k];
    const auto dot = group_by_[k].find('.');
    c.name = dot == std::string::npos ? group_by_[k] : group_by_[k].substr(dot + 1);
    out.push_back(c);
  }
  for (auto& a : aggs_) {
    TypeID t = TypeID::Float64;  // SUM/AVG declared FLOAT64 (plan.go:9954-9973)
    if (a.func == WG_AGG_COUNT) t = TypeID::Int64;
    if ((a.func == WG_AGG_MIN || a.func == WG_AGG_MAX) && !in_schema_.empty()) {
      // minMaxOutputType (aggregate.go:4624-4642): int32 -> Int64, float -> Float64
      for (auto& c : in_schema_)
        if (c.name == a.input_col) t = c.type == TypeID::Float64 ? TypeID::Float64 : TypeID::Int64;
    }
    out.push_back(Column{a.output_col, t});
  }
  return out;
}

// Next: up to 2048 groups per batch (aggregate.go:4295-4303)
Status GpuHashAggregate::Next(Context&, BatchPtr* out) {
  *out = nullptr;
  if (!h_ || emitted_ >= groups_) return Status::OK();
  const uint64_t n = std::min<uint64_t>(groups_ - emitted_, batch::kDefaultBatchSize);
  auto b = std::make_shared<RecordBatch>();
  b->schema = OutputSchema();
  b->len = (int)n;
  std::vector<std::vector<int64_t>> raw_keys(group_by_.size(), std::vector<int64_t>(n));
  std::vector<void*> kp, vp;
  std::vector<uint8_t*> vv;
  std::vector<std::vector<uint8_t>> valid(aggs_.size(), std::vector<uint8_t>(n));
  std::vector<std::shared_ptr<Vector>> vals;
  for (size_t k = 0; k < group_by_.size(); ++k) {
    auto v = Vector::make(in_schema_[k].type, (int)n, in_schema_[k].scale);
    kp.push_back((void*)data_of(*v));
    b->columns.push_back(v);
  }
  const auto sch = b->schema;
  for (size_t a = 0; a < aggs_.size(); ++a) {
    auto v = Vector::make(sch[group_by_.size() + a].type, (int)n);
    vp.push_back((void*)data_of(*v));
    vv.push_back(valid[a].data());
    vals.push_back(v);
  }
  TRY(wrap(wg_agg_fetch(h_, emitted_, n, kp.data(), vp.data(), vv.data(), WG_MEM_HOST), "wg_agg_fetch"));
  std::vector<uint8_t> kvalid(n);
  TRY(wrap(wg_agg_fetch_key_valid(h_, emitted_, n, kvalid.data(), WG_MEM_HOST), "wg_agg_fetch_key_valid"));
  for (size_t k = 0; k < group_by_.size(); ++k)
    for (uint64_t i = 0; i < n; ++i)
      if (!kvalid[i]) b->columns[k]->nulls.set_null((int)i, (int)n);
  for (size_t a = 0; a < aggs_.size(); ++a) {
    for (uint64_t i = 0; i < n; ++i)
      if (!valid[a][i]) vals[a]->nulls.set_null((int)i, (int)n);
    b->columns.push_back(vals[a]);
  }
  emitted_ += n;
  *out = b;
  return Status::OK();
}

std::unique_ptr<SinkSource> GpuHashAggregate::CloneSink() {
  auto c = std::make_unique<GpuHashAggregate>(ctx_, group_by_, aggs_, hint_);
  c->pending_preds_ = pending_preds_;
  return c;
}

Status GpuHashAggregate::MergeSink(SinkSource* other) {
  auto* o = dynamic_cast<GpuHashAggregate*>(other);
  if (!o) return Status::Error("MergeSink: not a GpuHashAggregate");
  std::lock_guard<std::mutex> lk(mu_);
  if (!o->h_) return Status::OK();
  if (!h_) {  // adopt the clone's state (aggregate.go:4688-4697)
    std::swap(h_, o->h_);
    key_idx_ = o->key_idx_;
    in_schema_ = o->in_schema_;
    preds_ = o->preds_;
    return Status::OK();
  }
  return wrap(wg_agg_merge(h_, o->h_), "wg_agg_merge");
}

void PushDownFiltersIntoGpuAggregate(Pipeline& p) {
  auto* agg = dynamic_cast<GpuHashAggregate*>(p.sink);
  if (!agg) return;
  while (!p.ops.empty()) {
    auto* f = dynamic_cast<KernelFilter*>(p.ops.back());
    if (!f || !agg->AddPredicate(*f).ok()) break;
    p.ops.pop_back();
  }
}

// ------------------------------------------------------------ GpuHashJoin

GpuHashJoin::GpuHashJoin(wg_ctx* ctx, std::string probe_key, std::string build_key, uint64_t build_hint,
                         JoinType type)
    : GpuHashJoin(ctx, std::vector<std::string>{std::move(probe_key)}, std::vector<std::string>{std::move(build_key)},
                  build_hint, type) {}

GpuHashJoin::GpuHashJoin(wg_ctx* ctx, std::vector<std::string> probe_keys, std::vector<std::string> build_keys,
                         uint64_t build_hint, JoinType type)
    : ctx_(ctx), probe_keys_(std::move(probe_keys)), build_keys_(std::move(build_keys)), hint_(build_hint),
      type_(type) {}

GpuHashJoin::~GpuHashJoin() {
  if (j_) wg_join_destroy(j_);
}

Status GpuHashJoin::probe_key_idx(const RecordBatch& in, std::vector<int>* idx) const {
  idx->clear();
  for (auto& k : probe_keys_) {
    const int ki = in.column_index(k);
    if (ki < 0) return Status::Error("probe key \"" + k + "\" not in the probe schema");
    idx->push_back(ki);
  }
  return Status::OK();
}

Status GpuHashJoin::Build(Context& ctx, Source* build) {
  if (build_keys_.empty() || build_keys_.size() > 2 || build_keys_.size() != probe_keys_.size())
    return Status::Error("a GPU join takes one or two key columns per side");
  TRY(build->Init(ctx));
  for (;;) {
    if (ctx.cancelled) return Status::Error("context canceled");
    BatchPtr b;
    TRY(build->Next(ctx, &b));
    if (!b) break;
    FlattenForConsumer(b.get(), nullptr);  // the build retains typed storage (flattenSource, late_mat.go:55-68)
    // the build keeps every row of a batch (a Sel is applied by compacting first)
    BatchPtr c = b;
    if (b->has_sel) {
      c = std::make_shared<RecordBatch>();
      c->schema = b->schema;
      c->len = (int)b->sel.size();
      for (auto& v : b->columns) c->columns.push_back(gather(*v, b->sel.data(), b->sel.size()));
    }
    if (!j_) {
      std::vector<int> ki;
      for (auto& k : build_keys_) {
        ki.push_back(c->column_index(k));
        if (ki.back() < 0) return Status::Error("build key \"" + k + "\" not in the build schema");
      }
      build_schema_ = c->schema;
      std::vector<int32_t> types, scales;
      for (auto& col : build_schema_) {
        types.push_back(wg_type_of(col.type));
        scales.push_back(col.scale);
      }
      if (ki.size() == 1)
        TRY(wrap(wg_join_create(ctx_, (int32_t)types.size(), types.data(), scales.data(), ki[0], 0, nullptr, hint_, &j_),
                 "wg_join_create"));
      else
        TRY(wrap(wg_join_create_dual(ctx_, (int32_t)types.size(), types.data(), scales.data(), ki[0], ki[1], 0,
                                     nullptr, hint_, &j_),
                 "wg_join_create_dual"));
      for (auto& col : build_schema_) build_cols_.push_back(Vector::make(col.type, 0, col.scale));
    }
    std::vector<wg_column> cols;
    for (auto& v : c->columns) cols.push_back(col_of(*v));
    TRY(wrap(wg_join_build_append(j_, cols.data(), (uint64_t)c->len, WG_MEM_HOST), "wg_join_build_append"));
    // retain the build side on the host for output gathers (Detach, join.go:1014)
    for (size_t k = 0; k < c->columns.size(); ++k) {
      Vector& dst = *build_cols_[k];
      const Vector& src = *c->columns[k];
      const int old = dst.len;
      dst.len += src.len;
      dst.i32.insert(dst.i32.end(), src.i32.begin(), src.i32.end());
      dst.i64.insert(dst.i64.end(), src.i64.begin(), src.i64.end());
      dst.f64.insert(dst.f64.end(), src.f64.begin(), src.f64.end());
      dst.dec.insert(dst.dec.end(), src.dec.begin(), src.dec.end());
      if (src.nulls.has_nulls() || dst.nulls.has_nulls()) {
        dst.nulls.words.resize((dst.len + 63) / 64, ~0ull);
        for (int i = 0; i < src.len; ++i)
          if (!src.nulls.valid(i)) dst.nulls.set_null(old + i, dst.len);
      }
    }
    rows_ += (uint64_t)c->len;
  }
  if (!j_) return Status::OK();  // empty build side: every probe row is dropped
  return wrap(wg_join_build_finish(j_), "wg_join_build_finish");
}

std::unique_ptr<GpuHashJoinProbe> GpuHashJoin::Probe() { return std::make_unique<GpuHashJoinProbe>(this); }

// ------------------------------------------------------ DeviceProbeKey

void DeviceProbeKey::Release() {
  for (void*& p : buf_) {
    if (p) (void)wg_free(ctx_, p);
    p = nullptr;
  }
  if (sel_) (void)wg_free(ctx_, sel_);
  sel_ = nullptr;
  cap_ = 0;
}

Status DeviceProbeKey::Prepare(wg_ctx* ctx, wg_join* j, const std::vector<int>& key_idx, const RecordBatch& in,
                               wg_column* key, const uint32_t** sel) {
  const uint64_t n = (uint64_t)in.len;
  if (n > cap_ || ctx != ctx_) {
    Release();
    ctx_ = ctx;
    cap_ = std::max<uint64_t>(n, batch::kDefaultBatchSize);
    const uint64_t vw = (cap_ + 63) / 64 * 8;
    const uint64_t sizes[6] = {cap_ * 8, cap_ * 8, vw, vw, cap_ * 8, vw};
    for (int i = 0; i < 6; ++i) TRY(wrap(wg_malloc(ctx, sizes[i], &buf_[i]), "wg_malloc"));
    TRY(wrap(wg_malloc(ctx, cap_ * 4, (void**)&sel_), "wg_malloc"));
  }
  *sel = nullptr;
  if (in.has_sel) {
    if (!in.sel.empty()) TRY(wrap(wg_memcpy_h2d(ctx, sel_, in.sel.data(), in.sel.size() * 4), "upload sel"));
    *sel = sel_;
  }
  wg_column cols[2];
  for (size_t k = 0; k < key_idx.size(); ++k) {
    const Vector& v = *in.columns[key_idx[k]];
    TRY(wrap(wg_memcpy_h2d(ctx, buf_[k], data_of(v), n * elem_bytes(v.type)), "upload probe key"));
    cols[k] = col_of(v);
    cols[k].data = buf_[k];
    cols[k].validity = nullptr;
    if (v.nulls.has_nulls()) {
      TRY(wrap(wg_memcpy_h2d(ctx, buf_[2 + k], v.nulls.words.data(), v.nulls.words.size() * 8), "upload validity"));
      cols[k].validity = (const uint64_t*)buf_[2 + k];
    }
  }
  if (key_idx.size() == 1) {
    *key = cols[0];
    return Status::OK();
  }
  TRY(wrap(wg_join_dual_probe_key(j, &cols[0], &cols[1], n, *sel, in.has_sel ? in.sel.size() : 0,
                                  (int64_t*)buf_[4], (uint64_t*)buf_[5]),
           "wg_join_dual_probe_key"));
  *key = wg_column{buf_[4], (const uint64_t*)buf_[5], WG_INT64, 0};
  return Status::OK();
}

// ------------------------------------------------------ GpuHashJoinProbe

GpuHashJoinProbe::GpuHashJoinProbe(GpuHashJoin* j) : join_(j) {}

GpuHashJoinProbe::~GpuHashJoinProbe() { (void)Close(); }

Status GpuHashJoinProbe::Close() {
  dkey_.Release();
  for (uint32_t* p : {d_out_p_, d_out_b_})
    if (p) wg_free(join_->ctx_, p);
  d_out_p_ = d_out_b_ = nullptr;
  out_cap_ = 0;
  return Status::OK();
}

Status GpuHashJoinProbe::ensure_out(uint64_t rows) {
  if (rows <= out_cap_) return Status::OK();
  for (uint32_t* p : {d_out_p_, d_out_b_})
    if (p) wg_free(join_->ctx_, p);
  d_out_p_ = d_out_b_ = nullptr;
  out_cap_ = rows;
  TRY(wrap(wg_malloc(join_->ctx_, rows * 4, (void**)&d_out_p_), "wg_malloc"));
  return wrap(wg_malloc(join_->ctx_, rows * 4, (void**)&d_out_b_), "wg_malloc");
}

Status GpuHashJoinProbe::Execute(Context& ctx, const BatchPtr& in, BatchPtr* out) {
  *out = nullptr;
  pending_ = false;
  if (join_->type_ != JoinType::Inner) return semi_anti(in, out);
  if (!join_->j_) return Status::OK();  // empty build
  std::vector<int> ki;
  TRY(join_->probe_key_idx(*in, &ki));
  for (int k : ki)  // the keys go to the device as values; the other columns stay views
    if (in->columns[k]->is_view()) {
      in->FlattenColumn(k);
      LateMat.flattens.fetch_add(1);
    }
  if (!LateMaterialize) FlattenForConsumer(in.get(), nullptr);  // the eager gather reads typed storage
  {
    std::lock_guard<std::mutex> lock(join_->mu_);
    TRY(dkey_.Prepare(join_->ctx_, join_->j_, ki, *in, &key_, &dsel_));
  }
  TRY(ensure_out(kMaxProbeOutputRows));
  cur_ = in;
  cursor_ = wg_probe_cursor{0, -1, 0, 0};
  if (bounded_) {
    TRY(emit(out));
    pending_ = !cursor_.done;
    return Status::OK();
  }
  // unbounded: the whole input's output in one batch
  std::vector<BatchPtr> parts;
  while (!cursor_.done) {
    if (ctx.cancelled) return Status::Error("context canceled");
    BatchPtr part;
    TRY(emit(&part));
    if (part) parts.push_back(part);
  }
  if (parts.empty()) return Status::OK();
  if (parts.size() == 1) {
    *out = parts[0];
    return Status::OK();
  }
  auto all = std::make_shared<RecordBatch>();
  all->schema = parts[0]->schema;
  for (size_t ci = 0; ci < all->schema.size(); ++ci) {
    if (parts[0]->columns[ci]->is_view()) {
      // views of one base: one view with the parts' indices and own NULLs in turn
      auto v = std::make_shared<Vector>();
      v->type = parts[0]->columns[ci]->type;
      v->scale = parts[0]->columns[ci]->scale;
      v->base = parts[0]->columns[ci]->base;
      for (auto& p : parts) {
        const Vector& s = *p->columns[ci];
        const int old = (int)v->indices.size();
        v->indices.insert(v->indices.end(), s.indices.begin(), s.indices.end());
        v->len = (int)v->indices.size();
        if (s.nulls.has_nulls())
          for (int i = 0; i < s.len; ++i)
            if (!s.nulls.valid(i)) v->nulls.set_null(old + i, v->len);
      }
      all->columns.push_back(v);
      all->len = v->len;
      continue;
    }
    auto v = Vector::make(all->schema[ci].type, 0, all->schema[ci].scale);
    for (auto& p : parts) {
      const Vector& s = *p->columns[ci];
      const int old = v->len;
      v->len += s.len;
      v->i32.insert(v->i32.end(), s.i32.begin(), s.i32.end());
      v->i64.insert(v->i64.end(), s.i64.begin(), s.i64.end());
      v->f64.insert(v->f64.end(), s.f64.begin(), s.f64.end());
      v->dec.insert(v->dec.end(), s.dec.begin(), s.dec.end());
      if (s.nulls.has_nulls())
        for (int i = 0; i < s.len; ++i)
          if (!s.nulls.valid(i)) v->nulls.set_null(old + i, v->len);
    }
    all->columns.push_back(v);
  }
  all->len = all->columns.empty() ? 0 : all->columns[0]->len;
  *out = all;
  return Status::OK();
}

Status GpuHashJoinProbe::NextOutput(Context&, BatchPtr* out) {
  *out = nullptr;
  if (!pending_) return Status::OK();
  TRY(emit(out));
  pending_ = !cursor_.done;
  return Status::OK();
}

// Semi / anti: the input batch with a new Sel of the probe rows that have
// (semi) / have no (anti) build row; nullptr when no row remains
// (join.go:3468-3478). An empty build keeps no row (semi) / every row (anti).
Status GpuHashJoinProbe::semi_anti(const BatchPtr& in, BatchPtr* out) {
  const bool anti = join_->type_ == JoinType::Anti;
  const uint64_t n = in->has_sel ? in->sel.size() : (uint64_t)in->len;
  if (n == 0) return Status::OK();
  if (!join_->j_) {
    if (anti) *out = in;
    return Status::OK();
  }
  std::vector<int> ki;
  TRY(join_->probe_key_idx(*in, &ki));
  TRY(ensure_out(std::max<uint64_t>((uint64_t)in->len, kMaxProbeOutputRows)));
  uint64_t got = 0;
  std::unique_lock<std::mutex> lock(join_->mu_);
  TRY(dkey_.Prepare(join_->ctx_, join_->j_, ki, *in, &key_, &dsel_));
  TRY(wrap(wg_join_semi(join_->j_, &key_, (uint64_t)in->len, dsel_, in->has_sel ? in->sel.size() : 0, anti ? 1 : 0,
                        d_out_p_, &got),
           "wg_join_semi"));
  lock.unlock();
  if (got == 0) return Status::OK();
  auto o = std::make_shared<RecordBatch>(*in);  // columns shared, Sel replaced
  o->sel.resize(got);
  TRY(wrap(wg_memcpy_d2h(join_->ctx_, o->sel.data(), d_out_p_, got * 4), "download sel"));
  o->has_sel = true;
  *out = o;
  return Status::OK();
}

// One bounded chunk: probe on the device, then gather the probe columns and the
// retained build columns on the host (gatherVector / gatherBuildVector).
Status GpuHashJoinProbe::emit(BatchPtr* out) {
  *out = nullptr;
  const RecordBatch& in = *cur_;
  uint64_t got = 0;
  {
    std::lock_guard<std::mutex> lock(join_->mu_);
    TRY(wrap(wg_join_probe(join_->j_, &key_, (uint64_t)in.len, dsel_, in.has_sel ? in.sel.size() : 0, d_out_p_,
                           d_out_b_, kMaxProbeOutputRows, &cursor_, &got),
             "wg_join_probe"));
  }
  if (in.has_sel && in.sel.empty()) cursor_.done = 1;
  if (got == 0) return Status::OK();
  std::vector<uint32_t> pr(got), br(got);
  TRY(wrap(wg_memcpy_d2h(join_->ctx_, pr.data(), d_out_p_, got * 4), "download pairs"));
  TRY(wrap(wg_memcpy_d2h(join_->ctx_, br.data(), d_out_b_, got * 4), "download pairs"));
  auto o = std::make_shared<RecordBatch>();
  o->len = (int)got;
  // output = probe columns, then build columns (joinOutputSchemaWithMapping, join.go:3898);
  // late materialization: views at the pair indices (emitViewOutput, join.go:2517-2600)
  for (size_t ci = 0; ci < in.columns.size(); ++ci) {
    o->schema.push_back(in.schema[ci]);
    o->columns.push_back(LateMaterialize ? NewViewVector(in.columns[ci], pr) : gather(*in.columns[ci], pr.data(), got));
  }
  for (size_t ci = 0; ci < join_->build_cols_.size(); ++ci) {
    Column c = join_->build_schema_[ci];
    for (auto& pc : in.schema)
      if (pc.name == c.name) c.name = "build." + c.name;  // qualify duplicate names
    o->schema.push_back(c);
    o->columns.push_back(LateMaterialize ? NewViewVector(join_->build_cols_[ci], br)
                                         : gather(*join_->build_cols_[ci], br.data(), got));
  }
  if (LateMaterialize) {
    LateMat.batches_emitted.fetch_add(1);
    LateMat.view_columns.fetch_add((int64_t)o->columns.size());
  }
  *out = o;
  return Status::OK();
}

// --------------------------------------------------------- Bloom filter

GpuBloomFilterOp::GpuBloomFilterOp(GpuHashJoin* j) : join_(j) {}

GpuBloomFilterOp::~GpuBloomFilterOp() { (void)Close(); }

Status GpuBloomFilterOp::Close() {
  dkey_.Release();
  if (d_out_) (void)wg_free(join_->ctx_, d_out_);
  d_out_ = nullptr;
  cap_rows_ = 0;
  return Status::OK();
}

Status GpuBloomFilterOp::Execute(Context&, const BatchPtr& in, BatchPtr* out) {
  *out = in;
  const uint64_t active = in ? (in->has_sel ? in->sel.size() : (uint64_t)in->len) : 0;
  if (active == 0 || disabled_ || !join_->j_) return Status::OK();
  std::unique_lock<std::mutex> lock(join_->mu_);
  if (!join_->bloom_built_) {
    TRY(wrap(wg_join_bloom_build(join_->j_), "wg_join_bloom_build"));
    join_->bloom_built_ = true;
  }
  std::vector<int> ki;
  if (!join_->probe_key_idx(*in, &ki).ok()) return Status::OK();  // key column not in this batch: pass every row (:66-76)
  wg_ctx* c = join_->ctx_;
  if ((uint64_t)in->len > cap_rows_) {
    if (d_out_) (void)wg_free(c, d_out_);
    d_out_ = nullptr;
    cap_rows_ = std::max<uint64_t>((uint64_t)in->len, batch::kDefaultBatchSize);
    TRY(wrap(wg_malloc(c, cap_rows_ * 4, (void**)&d_out_), "wg_malloc"));
  }
  wg_column key{};
  const uint32_t* d_sel = nullptr;
  TRY(dkey_.Prepare(c, join_->j_, ki, *in, &key, &d_sel));
  uint64_t got = 0;
  TRY(wrap(wg_join_bloom_filter(join_->j_, &key, (uint64_t)in->len, d_sel, in->has_sel ? in->sel.size() : 0, d_out_,
                                &got),
           "wg_join_bloom_filter"));
  lock.unlock();
  if (got == 0) {
    *out = nullptr;  // every row rejected (:183-185); not counted toward the bypass
    return Status::OK();
  }
  total_checked_ += active;
  total_passed_ += got;
  if (++batches_seen_ >= 32 && total_checked_ > 0 && (total_checked_ - total_passed_) * 20 < total_checked_)
    disabled_ = true;  // under 5 % rejected
  auto o = std::make_shared<RecordBatch>(*in);  // columns shared, Sel replaced
  o->sel.resize(got);
  TRY(wrap(wg_memcpy_d2h(c, o->sel.data(), d_out_, got * 4), "download sel"));
  o->has_sel = true;
  *out = o;
  return Status::OK();
}

// ------------------------------------------------------ GpuHashAggregate

int64_t GpuHashAggregate::StateBytes() {
  std::lock_guard<std::mutex> lk(mu_);
  if (!h_) return 0;
  uint64_t b = 0;
  if (wg_agg_state_bytes(h_, &b) != WG_OK) return 0;
  return (int64_t)b;
}

// ------------------------------------------------------ GpuShuffleRouter

GpuShuffleRouter::GpuShuffleRouter(wg_ctx* ctx, std::vector<std::string> keys, int num_parts)
    : ctx_(ctx), keys_(std::move(keys)), num_parts_(num_parts), counts_(std::max(num_parts, 1), 0) {}

GpuShuffleRouter::~GpuShuffleRouter() {
  for (void* p : buf_) (void)wg_free(ctx_, p);
}

// grow-only device scratch slot i
void* GpuShuffleRouter::scratch(size_t i, uint64_t bytes) {
  if (buf_.size() <= i) {
    buf_.resize(i + 1, nullptr);
    cap_.resize(i + 1, 0);
  }
  if (cap_[i] < bytes) {
    if (buf_[i]) (void)wg_free(ctx_, buf_[i]);
    buf_[i] = nullptr;
    cap_[i] = 0;
    if (wg_malloc(ctx_, bytes, &buf_[i]) != WG_OK) return nullptr;
    cap_[i] = bytes;
  }
  return buf_[i];
}

Status GpuShuffleRouter::Route(Context& ctx, const BatchPtr& b, std::vector<std::vector<uint32_t>>* per_part_rows,
                               std::vector<int>* parts) {
  if (ctx.cancelled) return Status::Error("context canceled");
  if (!resolved_) {  // keyOnce: resolve key names against the first batch
    for (auto& k : keys_) {
      const int ci = b->column_index(k);
      if (ci < 0) return Status::Error("partitioned shuffle: key \"" + k + "\" not in schema");
      key_idx_.push_back(ci);
    }
    resolved_ = true;
  }
  for (int ci : key_idx_) b->FlattenColumn(ci);  // only the key columns are read
  const int np = std::max(num_parts_, 1);
  per_part_rows->assign(np, {});
  const uint64_t n = (uint64_t)b->active();
  if (parts) parts->assign(n, 0);
  if (n == 0) return Status::OK();
  // key columns and Sel to the device
  std::vector<wg_column> cols;
  size_t slot = 0;
  for (int ci : key_idx_) {
    const Vector& v = *b->columns[ci];
    const uint64_t bytes = (uint64_t)std::max(b->len, 1) * elem_bytes(v.type);
    void* d = scratch(slot++, bytes);
    if (!d) return Status::Error(std::string("wg_malloc: ") + wg_last_error());
    TRY(wrap(wg_memcpy_h2d(ctx_, d, data_of(v), (uint64_t)b->len * elem_bytes(v.type)), "upload key"));
    wg_column c = col_of(v);
    c.data = d;
    if (v.nulls.has_nulls()) {
      void* dv = scratch(slot++, v.nulls.words.size() * 8);
      if (!dv) return Status::Error(std::string("wg_malloc: ") + wg_last_error());
      TRY(wrap(wg_memcpy_h2d(ctx_, dv, v.nulls.words.data(), v.nulls.words.size() * 8), "upload validity"));
      c.validity = (const uint64_t*)dv;
    }
    cols.push_back(c);
  }
  const size_t s_sel = 2 * key_idx_.size(), s_part = s_sel + 1, s_rows = s_sel + 2;
  uint32_t* d_sel = nullptr;
  if (b->has_sel) {
    d_sel = (uint32_t*)scratch(s_sel, n * 4);
    if (!d_sel) return Status::Error(std::string("wg_malloc: ") + wg_last_error());
    TRY(wrap(wg_memcpy_h2d(ctx_, d_sel, b->sel.data(), n * 4), "upload sel"));
  }
  uint32_t* d_part = parts ? (uint32_t*)scratch(s_part, n * 4) : nullptr;
  uint32_t* d_rows = (uint32_t*)scratch(s_rows, n * 4);
  if (!d_rows || (parts && !d_part)) return Status::Error(std::string("wg_malloc: ") + wg_last_error());
  std::vector<int32_t> kidx(cols.size());
  for (size_t k = 0; k < cols.size(); ++k) kidx[k] = (int32_t)k;
  std::vector<uint64_t> offs(np + 1);
  TRY(wrap(wg_shuffle_partition(ctx_, cols.data(), (int32_t)cols.size(), kidx.data(), (int32_t)kidx.size(),
                                (uint64_t)b->len, d_sel, b->has_sel ? n : 0, num_parts_, d_part, d_rows, offs.data()),
           "wg_shuffle_partition"));
  std::vector<uint32_t> rows(n);
  TRY(wrap(wg_memcpy_d2h(ctx_, rows.data(), d_rows, n * 4), "download rows"));
  for (int p = 0; p < np; ++p) {
    (*per_part_rows)[p].assign(rows.begin() + offs[p], rows.begin() + offs[p + 1]);
    counts_[p] += (int64_t)(offs[p + 1] - offs[p]);
  }
  if (parts) {
    std::vector<uint32_t> pp(n);
    TRY(wrap(wg_memcpy_d2h(ctx_, pp.data(), d_part, n * 4), "download parts"));
    for (uint64_t i = 0; i < n; ++i) (*parts)[i] = (int)pp[i];
  }
  return Status::OK();
}

// ------------------------------------------------------ CappedPartialAgg

// parseAggFuncString (worker/executor_stage.go:251-300) for the functions the
// GPU aggregate carries; false for anything else (the operator then passes
// rows through, as it does for an unknown function).
static bool gpu_agg_func(const std::string& name, wg_agg_func* f) {
  std::string s;
  for (char c : name)
    if (c != ' ' && c != '\t' && c != '\n') s.push_back((char)std::tolower((unsigned char)c));
  if (s == "sum") *f = WG_AGG_SUM;
  else if (s == "count") *f = WG_AGG_COUNT;
  else if (s == "min") *f = WG_AGG_MIN;
  else if (s == "max") *f = WG_AGG_MAX;
  else if (s == "avg") *f = WG_AGG_AVG;
  else return false;
  return true;
}

CappedPartialAgg::CappedPartialAgg(wg_ctx* ctx, std::vector<std::string> keys, std::vector<AggSpec> specs,
                                   std::vector<std::string> partition_keys, int64_t cap_bytes)
    : ctx_(ctx), group_by_(std::move(keys)), partition_keys_(std::move(partition_keys)),
      cap_bytes_(cap_bytes <= 0 ? kDefaultCapBytes : cap_bytes) {
  bool unknown = false;
  for (auto& s : specs) {
    wg_agg_func f = WG_AGG_SUM;
    if (!gpu_agg_func(s.func, &f)) unknown = true;
    aggs_.push_back(AggColumn{f, s.input_col, s.output_col, ""});
  }
  resolved_ = disabled_ = unknown;  // newCappedPartialAggPartitioned (:85-113)
}

// resolveAgainst (:117-151)
void CappedPartialAgg::resolve_against(const RecordBatch& b) {
  resolved_ = true;
  auto present = [&](const std::string& name) {
    for (auto& c : b.schema)
      if (c.name == name) return true;
    return false;
  };
  for (auto& k : partition_keys_)
    if (!present(k)) {
      disabled_ = true;
      return;
    }
  std::vector<std::string> kept;
  for (auto& k : group_by_) {
    if (present(k)) kept.push_back(k);
    else ++dropped_keys_;
  }
  group_by_ = kept;
  std::vector<AggColumn> kept_aggs;
  for (auto& a : aggs_)
    if (present(a.input_col)) kept_aggs.push_back(a);
  aggs_ = kept_aggs;
  if (aggs_.empty() || group_by_.empty()) disabled_ = true;
}

Status CappedPartialAgg::Consume(Context& ctx, const BatchPtr& b, std::vector<BatchPtr>* flushed) {
  flushed->clear();
  FlattenForConsumer(b.get(), nullptr);
  if (!resolved_) resolve_against(*b);
  in_rows += b->active();
  if (disabled_) {  // raw passthrough
    out_rows += b->active();
    flushed->push_back(b);
    return Status::OK();
  }
  if (!agg_) agg_ = std::make_unique<GpuHashAggregate>(ctx_, group_by_, aggs_);
  TRY(agg_->Consume(ctx, b));
  if (agg_->StateBytes() < cap_bytes_) return Status::OK();
  return flush(ctx, flushed);
}

Status CappedPartialAgg::Drain(Context& ctx, std::vector<BatchPtr>* flushed) {
  flushed->clear();
  if (!agg_) return Status::OK();
  return flush(ctx, flushed);
}

// flush (:207-234): finalize the epoch, collect its batches, start afresh
Status CappedPartialAgg::flush(Context& ctx, std::vector<BatchPtr>* out) {
  TRY(agg_->Finalize(ctx));
  for (;;) {
    BatchPtr b;
    TRY(agg_->Next(ctx, &b));
    if (!b) break;
    out_rows += b->active();
    out->push_back(b);
  }
  TRY(agg_->Close());
  agg_.reset();
  ++flushes;
  return Status::OK();
}

}  // namespace wadjet
