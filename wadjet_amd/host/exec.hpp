// exec.hpp — C++ mirror of wadjet's operator interface for the GPU path.
//
// The reference's drop-in boundary is a set of Go interfaces in
// internal/engine/exec/operators.go:10-103 (Source, UnaryOperator, Sink,
// SinkSource, Cloneable, MergeableSink, BoundedOutputOperator) over
// batch.RecordBatch (internal/engine/batch/batch.go:9-28). This header restates
// them in C++ with the same names, argument meaning and error behaviour, plus
// the operators the GPU library backs:
//   GpuHashAggregate   SinkSource + MergeableSink      (aggregate.go:98, :1075, :3855, :4202, :4646-4700)
//   GpuHashJoin        Build(ctx, Source) / Probe()      (join.go:771, :1967)
//   GpuHashJoinProbe   UnaryOperator + BoundedOutput     (join.go:2072-2139, :2306-2600; view output
//                      + ViewAware                        late_mat.go, batch/view.go)
//   GpuBloomFilterOp   UnaryOperator + Cloneable         (bloom_filter_op.go:13-274)
//   GpuShuffleRouter   partitioned shuffle routing       (worker/partitioned_shuffle_sink.go:199-290)
//   CappedPartialAgg   exchange partial aggregation      (worker/shuffle_partial_agg.go:36-234)
// and the CPU pieces a pipeline needs around them: KernelFilter (filter.go:347-425),
// BatchSource, CollectSink and Pipeline{Source, Ops, Sink, Workers}
// (pipeline.go:31-690). Errors are returned as Status (Go's `error`).
#pragma once
#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/wadjet_gpu.h"

namespace wadjet {

// ---------------------------------------------------------------- status

struct Status {
  bool ok_ = true;
  std::string msg;
  int code = 0;  // the library's wg_status for failures it reported, else -1
  static Status OK() { return {}; }
  static Status Error(std::string m, int code = -1) { return Status{false, std::move(m), code}; }
  bool ok() const { return ok_; }
  // device memory ran out (WG_ERR_OOM): the adapter may fall back to the CPU operator
  bool out_of_memory() const { return !ok_ && code == WG_ERR_OOM; }
};

struct Context {
  bool cancelled = false;  // ctx.Err() != nil
};

namespace batch {

// TypeID subset on the GPU path (vector.go:302-465 storage classes).
enum class TypeID { Int32, Int64, Float64, Decimal, IPv4, Port, Date, Timestamp };

constexpr int kDefaultBatchSize = 2048;  // batch.go:6

struct Int128 {  // decimal.go:11-14 (Go field order: Hi, Lo)
  int64_t hi = 0;
  uint64_t lo = 0;
};

// Bitmap: bit i%64 of word i/64, 1 = valid (bitmap.go:7-26).
struct Bitmap {
  std::vector<uint64_t> words;  // empty = no nulls
  bool has_nulls() const { return !words.empty(); }
  bool valid(int i) const { return words.empty() || ((words[i >> 6] >> (i & 63)) & 1ull); }
  void set_null(int i, int n);
};

struct Column {  // parquet.Column subset: name, type, nullability, decimal scale
  std::string name;
  TypeID type;
  int scale = 0;
  bool nullable = true;
};

struct Vector {
  TypeID type = TypeID::Int64;
  int len = 0;
  int scale = 0;
  Bitmap nulls;
  std::vector<int32_t> i32;
  std::vector<int64_t> i64;
  std::vector<double> f64;
  std::vector<Int128> dec;  // DecimalData
  // A view (late materialization, view.go:10-26) owns no storage: row i is
  // row indices[i] of `base` (always an owned vector), NULL when the view's
  // own bitmap `nulls` or the base says so.
  std::shared_ptr<Vector> base;
  std::vector<uint32_t> indices;
  bool is_view() const { return base != nullptr; }
  // Vector.Flatten (view.go:87-230): gather through the indices into owned
  // storage in place (own-NULL rows do not read the base); no-op when owned.
  void Flatten();
  static std::shared_ptr<Vector> make(TypeID t, int n, int scale = 0);
};

// NewViewVector (view.go:27-80): a view of `base` at `indices`; over a view it
// composes (indices through the base view's, its own NULLs folded in) so the
// result's base is always owned.
std::shared_ptr<Vector> NewViewVector(const std::shared_ptr<Vector>& base, std::vector<uint32_t> indices);

struct RecordBatch {
  std::vector<std::shared_ptr<Vector>> columns;
  std::vector<Column> schema;
  int len = 0;
  std::vector<uint32_t> sel;  // selection vector; has_sel = false means all rows active
  bool has_sel = false;
  int column_index(const std::string& name) const;
  int active() const { return has_sel ? (int)sel.size() : len; }
  bool HasViews() const;   // view.go:259-267
  void FlattenViews();     // view.go:271-275
  void FlattenColumn(int i) { columns[i]->Flatten(); }
};

}  // namespace batch

using batch::RecordBatch;
using BatchPtr = std::shared_ptr<RecordBatch>;

// ------------------------------------------------------------ interfaces

struct Source {  // operators.go:11-15; Next returns nullptr at end of data
  virtual ~Source() = default;
  virtual Status Init(Context&) = 0;
  virtual Status Next(Context&, BatchPtr* out) = 0;
  virtual Status Close() = 0;
};

struct UnaryOperator {  // operators.go:18-22; nullptr output = fully filtered
  virtual ~UnaryOperator() = default;
  virtual Status Init(Context&) = 0;
  virtual Status Execute(Context&, const BatchPtr& in, BatchPtr* out) = 0;
  virtual Status Close() = 0;
};

struct Sink {  // operators.go:24-31: must handle concurrent Consume calls
  virtual ~Sink() = default;
  virtual Status Init(Context&) = 0;
  virtual Status Consume(Context&, const BatchPtr& b) = 0;
  virtual Status Finalize(Context&) = 0;
  virtual Status Close() = 0;
};

struct SinkSource : Sink, Source {  // operators.go:34-37
  Status Init(Context& c) override = 0;
  Status Close() override = 0;
};

struct Cloneable {  // operators.go:42-44
  virtual ~Cloneable() = default;
  virtual std::unique_ptr<UnaryOperator> Clone() = 0;
};

struct MergeableSink : SinkSource {  // operators.go:56-60
  virtual std::unique_ptr<SinkSource> CloneSink() = 0;
  virtual Status MergeSink(SinkSource* other) = 0;
};

// Late materialization (late_mat.go:11-68): process-wide counters that prove
// the feature engaged, the ViewAware opt-in, and the defensive flatten in
// front of every consumer that has not opted in.
struct LateMatCounters {
  std::atomic<int64_t> batches_emitted{0};  // probe output batches emitted as views
  std::atomic<int64_t> view_columns{0};     // view columns across those batches
  std::atomic<int64_t> flattens{0};         // FlattenViews calls made for consumers
};
extern LateMatCounters LateMat;

struct ViewAware {  // late_mat.go:32-34
  virtual ~ViewAware() = default;
  virtual bool AcceptsViews() const = 0;
};

// FlattenForConsumer (late_mat.go:43-53)
template <class T>
void FlattenForConsumer(RecordBatch* b, T* consumer) {
  if (!b || !b->HasViews()) return;
  if (auto* va = dynamic_cast<ViewAware*>(consumer); va && va->AcceptsViews()) return;
  b->FlattenViews();
  LateMat.flattens.fetch_add(1);
}
inline void FlattenForConsumer(RecordBatch* b, std::nullptr_t) {
  if (!b || !b->HasViews()) return;
  b->FlattenViews();
  LateMat.flattens.fetch_add(1);
}

struct BoundedOutputOperator {  // operators.go:86-90
  virtual ~BoundedOutputOperator() = default;
  virtual void EnableBoundedOutput() = 0;
  virtual bool HasPendingOutput() = 0;
  virtual Status NextOutput(Context&, BatchPtr* out) = 0;
};

// ------------------------------------------------------------ CPU pieces

enum class CompareOp { Eq, Ne, Lt, Le, Gt, Ge };  // kernel.CompareOp

// KernelFilter: typed compare of one column against a constant; compacts Sel
// in place when the input already has one (filter.go:373-425).
class KernelFilter : public UnaryOperator, public Cloneable {
 public:
  KernelFilter(std::string col, CompareOp op, double value) : col_(std::move(col)), op_(op), value_(value) {}
  Status Init(Context&) override { return Status::OK(); }
  Status Execute(Context&, const BatchPtr& in, BatchPtr* out) override;
  Status Close() override { return Status::OK(); }
  std::unique_ptr<UnaryOperator> Clone() override { return std::make_unique<KernelFilter>(col_, op_, value_); }
  const std::string& column() const { return col_; }
  CompareOp op() const { return op_; }
  double value() const { return value_; }

 private:
  std::string col_;
  CompareOp op_;
  double value_;
};

// A Source over prepared batches (NewSliceSource / BatchSource in tests).
class BatchSource : public Source {
 public:
  explicit BatchSource(std::vector<BatchPtr> batches) : batches_(std::move(batches)) {}
  Status Init(Context&) override { pos_ = 0; return Status::OK(); }
  Status Next(Context&, BatchPtr* out) override;
  Status Close() override { return Status::OK(); }

 private:
  std::vector<BatchPtr> batches_;
  size_t pos_ = 0;
  std::mutex mu_;
};

// Collects every row it receives (Sel applied).
class CollectSink : public Sink {
 public:
  Status Init(Context&) override { return Status::OK(); }
  Status Consume(Context&, const BatchPtr& b) override;
  Status Finalize(Context&) override { return Status::OK(); }
  Status Close() override { return Status::OK(); }
  std::vector<BatchPtr> batches;  // compacted copies
  int rows() const;

 private:
  std::mutex mu_;
};

// Pipeline{Source, Ops, Sink, Workers} (pipeline.go:31-36): serial when
// Workers <= 1; otherwise Workers threads pull batches, run cloned op chains
// and feed cloned sinks that are merged into the primary (MergeSink) before
// Finalize (pipeline.go:351-690, the non-partitioned shape).
struct Pipeline {
  Source* source = nullptr;
  std::vector<UnaryOperator*> ops;
  Sink* sink = nullptr;
  int workers = 0;
  Status Run(Context& ctx);
};

// Planner hook (the role of plan.go:7233-7240 with a WADJET_GPU switch): when
// a pipeline's trailing KernelFilters feed a GpuHashAggregate, push the
// predicates into the aggregate so the GPU evaluates them fused.
void PushDownFiltersIntoGpuAggregate(Pipeline& p);

// ------------------------------------------------------- GPU operators

struct AggColumn {  // exec.AggColumn (aggregate.go:73-81)
  wg_agg_func func;
  std::string input_col;  // "" = COUNT(*)
  std::string output_col;
  std::string input_col2;  // product input (aggPreProject's qty * price), optional
};

class GpuHashAggregate : public MergeableSink {
 public:
  GpuHashAggregate(wg_ctx* ctx, std::vector<std::string> group_by, std::vector<AggColumn> aggs,
                   uint64_t group_hint = 0);
  ~GpuHashAggregate() override;
  Status Init(Context&) override { return Status::OK(); }
  Status Consume(Context&, const BatchPtr& b) override;
  Status Finalize(Context&) override;
  Status Next(Context&, BatchPtr* out) override;
  Status Close() override;
  std::unique_ptr<SinkSource> CloneSink() override;
  Status MergeSink(SinkSource* other) override;
  // predicates pushed down from KernelFilters (at most 2)
  Status AddPredicate(const KernelFilter& f);
  std::vector<batch::Column> OutputSchema() const;
  wg_agg* handle() const { return h_; }
  // StateBytes (aggregate.go:4050-4054): bytes of live group state
  int64_t StateBytes();

 private:
  Status resolve(const RecordBatch& b);  // typed dispatch on the first batch (resolveIndices)
  wg_ctx* ctx_;
  std::vector<std::string> group_by_;
  std::vector<AggColumn> aggs_;
  uint64_t hint_;
  std::vector<std::pair<std::string, wg_predicate>> pending_preds_;
  wg_agg* h_ = nullptr;
  std::vector<batch::Column> in_schema_;
  std::vector<int32_t> key_idx_;
  std::vector<wg_predicate> preds_;
  uint64_t emitted_ = 0, groups_ = 0;
  bool finalized_ = false;
  std::mutex mu_;  // h.mu: Consume is serialised per handle
};

class GpuHashJoinProbe;

// JoinType subset on the GPU path (join.go JoinType): inner, semi, anti.
enum class JoinType { Inner, Semi, Anti };

// The probe key of a batch on the device: the key column (one-column keys) or
// the id column of its (key0, key1) pairs (two-column keys,
// wg_join_dual_probe_key), and the batch's Sel.
class DeviceProbeKey {
 public:
  ~DeviceProbeKey() { Release(); }
  // key columns key_idx of `in` (one or two); *key / *sel name device memory
  Status Prepare(wg_ctx* ctx, wg_join* j, const std::vector<int>& key_idx, const RecordBatch& in, wg_column* key,
                 const uint32_t** sel);
  void Release();

 private:
  wg_ctx* ctx_ = nullptr;
  void* buf_[6] = {};  // key0, key1, validity0, validity1, ids, id validity
  uint32_t* sel_ = nullptr;
  uint64_t cap_ = 0;
};

class GpuHashJoin {
 public:
  GpuHashJoin(wg_ctx* ctx, std::string probe_key, std::string build_key, uint64_t build_hint = 0,
              JoinType type = JoinType::Inner);
  // one or two key columns per side (NewHashJoin's LeftKeys / RightKeys, join.go:180-240;
  // two int-class keys take the dual-key path, join.go:592-661)
  GpuHashJoin(wg_ctx* ctx, std::vector<std::string> probe_keys, std::vector<std::string> build_keys,
              uint64_t build_hint = 0, JoinType type = JoinType::Inner);
  ~GpuHashJoin();
  Status Build(Context& ctx, Source* build);  // join.go:771
  std::unique_ptr<GpuHashJoinProbe> Probe();   // join.go:1967
  uint64_t BuildRows() const { return rows_; }
  wg_join* handle() const { return j_; }

 private:
  friend class GpuHashJoinProbe;
  Status probe_key_idx(const RecordBatch& in, std::vector<int>* idx) const;  // resolve names; error if absent
  wg_ctx* ctx_;
  std::vector<std::string> probe_keys_, build_keys_;
  uint64_t hint_;
  JoinType type_;
  wg_join* j_ = nullptr;
  std::vector<batch::Column> build_schema_;
  std::vector<std::shared_ptr<batch::Vector>> build_cols_;  // retained build side (host copy for gathers)
  uint64_t rows_ = 0;
  bool bloom_built_ = false;
  // Probe and Bloom clones run on pipeline worker threads over this one
  // handle, whose probe scratch and lazily built filter are per handle (the
  // C-ABI serialises nothing): every wg_join_* call they make holds this.
  std::mutex mu_;
  friend class GpuBloomFilterOp;
};

// MaxProbeOutputRows = 32768 (join.go:2044-2060)
constexpr uint64_t kMaxProbeOutputRows = 32768;

class GpuHashJoinProbe : public UnaryOperator, public BoundedOutputOperator, public Cloneable, public ViewAware {
 public:
  explicit GpuHashJoinProbe(GpuHashJoin* j);
  ~GpuHashJoinProbe() override;
  Status Init(Context&) override { return Status::OK(); }
  Status Execute(Context&, const BatchPtr& in, BatchPtr* out) override;
  Status Close() override;
  void EnableBoundedOutput() override { bounded_ = true; }
  bool HasPendingOutput() override { return pending_; }
  Status NextOutput(Context&, BatchPtr* out) override;
  std::unique_ptr<UnaryOperator> Clone() override {
    auto c = std::make_unique<GpuHashJoinProbe>(join_);
    c->LateMaterialize = LateMaterialize;
    return c;
  }
  // Reads view inputs (its key columns flattened, the rest composed), join.go:2234.
  bool AcceptsViews() const override { return true; }
  // HashJoinProbe.LateMaterialize (join.go:2436-2442): inner-join output as
  // view columns over the probe batch and the retained build columns at the
  // pair indices instead of gathered copies. Off by default.
  bool LateMaterialize = false;

 private:
  Status emit(BatchPtr* out);  // one bounded chunk of the current input
  Status semi_anti(const BatchPtr& in, BatchPtr* out);  // executeSemiAntiJoin, join.go:3245-3479
  Status ensure_out(uint64_t rows);
  GpuHashJoin* join_;
  bool bounded_ = false, pending_ = false;
  BatchPtr cur_;
  wg_probe_cursor cursor_{};
  DeviceProbeKey dkey_;
  wg_column key_{};
  const uint32_t* dsel_ = nullptr;
  uint32_t* d_out_p_ = nullptr;
  uint32_t* d_out_b_ = nullptr;
  uint64_t out_cap_ = 0;
};

// BloomFilterOp (bloom_filter_op.go:13-274): drops probe rows whose key cannot
// be in the build (device filter of the join's Bloom bits) before the probe.
// Clones share the read-only filter; each keeps its own adaptive state: after
// 32 batches a filter that rejected under 5 % of the rows it checked passes
// every later batch through untouched (:201-214).
class GpuBloomFilterOp : public UnaryOperator, public Cloneable {
 public:
  explicit GpuBloomFilterOp(GpuHashJoin* j);  // builds the join's filter on first use
  ~GpuBloomFilterOp() override;
  Status Init(Context&) override { return Status::OK(); }
  Status Execute(Context&, const BatchPtr& in, BatchPtr* out) override;
  Status Close() override;
  std::unique_ptr<UnaryOperator> Clone() override { return std::make_unique<GpuBloomFilterOp>(join_); }
  bool Disabled() const { return disabled_; }

 private:
  GpuHashJoin* join_;
  uint64_t total_checked_ = 0, total_passed_ = 0;
  int batches_seen_ = 0;
  bool disabled_ = false;
  DeviceProbeKey dkey_;
  uint32_t* d_out_ = nullptr;
  uint64_t cap_rows_ = 0;
};

// ------------------------------------------------------------ shuffle

// The routing half of partitionedShuffleSink (worker/partitioned_shuffle_sink.go:
// 153-290): key columns resolved by name on the first batch (keyOnce), every
// active row hashed to a partition on the device (hashRowsIntoPartitions,
// :1039-1171, via wg_shuffle_partition) and, per partition, the batch's rows in
// input order (perPartRows). Encoding and writing the partition files belongs
// to the storage layer and stays with the caller.
class GpuShuffleRouter {
 public:
  GpuShuffleRouter(wg_ctx* ctx, std::vector<std::string> keys, int num_parts);
  ~GpuShuffleRouter();
  // per_part_rows[p] = source rows (Sel values or positions) routed to p;
  // parts (optional) = partition of the i-th active row
  Status Route(Context&, const BatchPtr& b, std::vector<std::vector<uint32_t>>* per_part_rows,
               std::vector<int>* parts = nullptr);
  // rows routed to each partition so far (PartitionRowCounts, :1006-1019)
  const std::vector<int64_t>& PartitionRowCounts() const { return counts_; }
  int NumParts() const { return num_parts_; }

 private:
  void* scratch(size_t i, uint64_t bytes);
  wg_ctx* ctx_;
  std::vector<std::string> keys_;
  int num_parts_;
  std::vector<int> key_idx_;
  bool resolved_ = false;
  std::vector<int64_t> counts_;
  std::vector<void*> buf_;
  std::vector<uint64_t> cap_;
};

// distributed.AggSpec as the exchange carries it: the function by name.
struct AggSpec {
  std::string func;  // "sum", "min", "max", ... (parseAggFuncString)
  std::string input_col;
  std::string output_col;
};

// cappedPartialAgg (worker/shuffle_partial_agg.go:36-234): pre-combines a
// shuffle task's rows on the exchange's keys before they are partitioned.
// Name-preserving SUM/MIN/MAX specs; once the aggregate's state reaches
// cap_bytes the epoch is flushed (Finalize + drain of Next) and a fresh one
// starts, so memory stays bounded and flushed partials stay mergeable. On the
// first batch keys and specs are intersected with its schema; with nothing
// combinable left, an unknown function, or an absent partition key the
// operator passes batches through unchanged.
class CappedPartialAgg {
 public:
  static constexpr int64_t kDefaultCapBytes = 128ll << 20;  // defaultPartialAggCapBytes
  CappedPartialAgg(wg_ctx* ctx, std::vector<std::string> keys, std::vector<AggSpec> specs,
                   std::vector<std::string> partition_keys = {}, int64_t cap_bytes = 0);
  // consume: returns the batches flushed by this call (the input itself when disabled)
  Status Consume(Context&, const BatchPtr& b, std::vector<BatchPtr>* flushed);
  // drain: flushes the last epoch at end of stream
  Status Drain(Context&, std::vector<BatchPtr>* flushed);
  int64_t in_rows = 0, out_rows = 0, flushes = 0;
  bool Disabled() const { return disabled_; }
  int DroppedKeys() const { return dropped_keys_; }

 private:
  void resolve_against(const RecordBatch& b);
  Status flush(Context&, std::vector<BatchPtr>* out);
  wg_ctx* ctx_;
  std::vector<std::string> group_by_;
  std::vector<AggColumn> aggs_;
  std::vector<std::string> partition_keys_;
  int64_t cap_bytes_;
  std::unique_ptr<GpuHashAggregate> agg_;
  bool resolved_ = false, disabled_ = false;
  int dropped_keys_ = 0;
};

}  // namespace wadjet
