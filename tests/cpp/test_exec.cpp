// Operator-level tests of the C++ layer (wadjet_amd/host/exec.hpp), written
// like the reference's own exec tests. `cpu` runs the host-only pieces
// (KernelFilter, Pipeline, CollectSink); `gpu` runs GpuHashAggregate and
// GpuHashJoin on device 0.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>

#include "../../wadjet_amd/host/exec.hpp"

using namespace wadjet;
using batch::Column;
using batch::TypeID;
using batch::Vector;

static int g_fail = 0;
#define EXPECT(cond, ...)                                  \
  do {                                                     \
    if (!(cond)) {                                         \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__);     \
      std::printf(__VA_ARGS__);                            \
      std::printf("\n");                                   \
      ++g_fail;                                            \
    }                                                      \
  } while (0)
#define OK(st) EXPECT((st).ok(), "%s", (st).msg.c_str())

static BatchPtr make_batch(std::vector<Column> schema, int n) {
  auto b = std::make_shared<RecordBatch>();
  b->schema = schema;
  b->len = n;
  for (auto& c : schema) b->columns.push_back(Vector::make(c.type, n, c.scale));
  return b;
}

// ----------------------------------------------------------------- CPU tests

static void test_kernel_filter_and_pipeline() {
  // KernelFilter compacts Sel; a second filter compacts it again (AND chain)
  auto b = make_batch({{"k", TypeID::Int64}, {"f", TypeID::Float64}}, 10);
  for (int i = 0; i < 10; ++i) {
    b->columns[0]->i64[i] = i;
    b->columns[1]->f64[i] = i - 4.5;
  }
  b->columns[1]->nulls.set_null(7, 10);  // NULL never passes
  KernelFilter f1("f", CompareOp::Gt, -0.5), f2("k", CompareOp::Lt, 9);
  Context ctx;
  BatchPtr o1, o2;
  OK(f1.Execute(ctx, b, &o1));
  OK(f2.Execute(ctx, o1, &o2));
  // f = i - 4.5 > -0.5 keeps rows 5..9 except the NULL row 7; then k < 9
  EXPECT(o2 && o2->sel == std::vector<uint32_t>({5, 6, 8}), "sel mismatch");
  BatchPtr none;
  KernelFilter f3("f", CompareOp::Gt, 100.0);
  OK(f3.Execute(ctx, b, &none));
  EXPECT(!none, "fully filtered batch must be nullptr");
  KernelFilter bad("nope", CompareOp::Eq, 1);
  EXPECT(!bad.Execute(ctx, b, &none).ok(), "unknown filter column must error (filter.go:410)");

  // Pipeline serial and parallel through CollectSink
  std::vector<BatchPtr> src;
  for (int m = 0; m < 8; ++m) {
    auto x = make_batch({{"k", TypeID::Int64}, {"f", TypeID::Float64}}, 2048);
    for (int i = 0; i < 2048; ++i) {
      x->columns[0]->i64[i] = m * 2048 + i;
      x->columns[1]->f64[i] = (i % 10) - 4.5;
    }
    src.push_back(x);
  }
  for (int workers : {0, 4}) {
    BatchSource s(src);
    CollectSink sink;
    KernelFilter f("f", CompareOp::Gt, -0.5);
    Pipeline p{&s, {&f}, &sink, workers};
    OK(p.Run(ctx));
    int want = 0;
    for (int i = 0; i < 2048; ++i) want += (i % 10) - 4.5 > -0.5;
    EXPECT(sink.rows() == 8 * want, "pipeline rows %d (workers %d)", sink.rows(), workers);
  }
  // cancellation between batches (pipeline.go:268-270)
  {
    BatchSource s(src);
    CollectSink sink;
    Context cancelled;
    cancelled.cancelled = true;
    Pipeline p{&s, {}, &sink, 0};
    EXPECT(!p.Run(cancelled).ok(), "cancelled context must stop the pipeline");
  }
}

// ----------------------------------------------------------------- GPU tests

static void test_hash_aggregate_kat(wg_ctx* g) {
  // exec_test.go:39-96 with int keys: group 1 = "a", 2 = "b"; SUM(amount) a = 40, COUNT(amount) a = 2
  auto b = make_batch({{"group", TypeID::Int64}, {"amount", TypeID::Float64}}, 4);
  const int64_t gk[] = {1, 2, 1, 2};
  const double am[] = {10, 20, 30, 40};
  for (int i = 0; i < 4; ++i) {
    b->columns[0]->i64[i] = gk[i];
    b->columns[1]->f64[i] = am[i];
  }
  GpuHashAggregate agg(g, {"group"}, {{WG_AGG_SUM, "amount", "total"}, {WG_AGG_COUNT, "amount", "cnt"}});
  BatchSource src({b});
  Context ctx;
  Pipeline p{&src, {}, &agg, 0};
  OK(p.Run(ctx));
  BatchPtr out;
  OK(agg.Next(ctx, &out));
  EXPECT(out && out->len == 2, "expected 2 groups");
  bool seen = false;
  for (int i = 0; out && i < out->len; ++i)
    if (out->columns[0]->i64[i] == 1) {
      seen = true;
      EXPECT(out->columns[1]->f64[i] == 40.0, "total %f", out->columns[1]->f64[i]);
      EXPECT(out->columns[2]->i64[i] == 2, "cnt %lld", (long long)out->columns[2]->i64[i]);
    }
  EXPECT(seen, "group a missing");
  BatchPtr eof;
  OK(agg.Next(ctx, &eof));
  EXPECT(!eof, "Next after the last batch returns nullptr");
}

static void test_parallel_pushdown_equivalence(wg_ctx* g) {
  // Workers=4 with CloneSink/MergeSink and pushed-down filters == serial, CPU-filtered
  std::vector<BatchPtr> src;
  std::map<int64_t, std::pair<int64_t, int64_t>> want;  // k -> (sum v, count)
  for (int m = 0; m < 40; ++m) {
    auto x = make_batch({{"k", TypeID::Int32}, {"v", TypeID::Int64}, {"f", TypeID::Float64}}, 2048);
    for (int i = 0; i < 2048; ++i) {
      const int64_t r = m * 2048 + i;
      x->columns[0]->i32[i] = (int32_t)((r * 7919) % 1000);
      x->columns[1]->i64[i] = r * 3 - 1000;
      x->columns[2]->f64[i] = std::sin((double)r);
      if (x->columns[2]->f64[i] > -0.5) {
        auto& w = want[x->columns[0]->i32[i]];
        w.first += x->columns[1]->i64[i];
        w.second += 1;
      }
    }
    src.push_back(x);
  }
  for (int mode = 0; mode < 2; ++mode) {
    BatchSource s(src);
    GpuHashAggregate agg(g, {"k"}, {{WG_AGG_SUM, "v", "s"}, {WG_AGG_COUNT, "", "c"}, {WG_AGG_MIN, "k", "mn"}});
    KernelFilter f("f", CompareOp::Gt, -0.5);
    Pipeline p{&s, {&f}, &agg, mode == 0 ? 0 : 4};
    if (mode == 1) PushDownFiltersIntoGpuAggregate(p);
    Context ctx;
    OK(p.Run(ctx));
    int groups = 0;
    for (;;) {
      BatchPtr out;
      OK(agg.Next(ctx, &out));
      if (!out) break;
      EXPECT(out->schema[3].type == TypeID::Int64, "MIN(int32) emits Int64 (aggregate.go:4636)");
      for (int i = 0; i < out->len; ++i, ++groups) {
        const int64_t k = out->columns[0]->i32[i];
        EXPECT(want.count(k), "unexpected key %lld", (long long)k);
        EXPECT((int64_t)out->columns[1]->f64[i] == want[k].first, "sum of %lld", (long long)k);
        EXPECT(out->columns[2]->i64[i] == want[k].second, "count of %lld", (long long)k);
        EXPECT(out->columns[3]->i64[i] == k, "min of %lld", (long long)k);
      }
    }
    EXPECT(groups == (int)want.size(), "groups %d want %zu (mode %d)", groups, want.size(), mode);
  }
}

static void test_hash_join_kats(wg_ctx* g) {
  Context ctx;
  // join_test.go:10-56 (strings coded as ints): 3 rows, alice twice
  {
    auto build = make_batch({{"user_id", TypeID::Int64}, {"name", TypeID::Int64}}, 3);
    const int64_t bk[] = {1, 2, 3}, nm[] = {100, 200, 300};
    for (int i = 0; i < 3; ++i) {
      build->columns[0]->i64[i] = bk[i];
      build->columns[1]->i64[i] = nm[i];
    }
    auto probe = make_batch({{"user_id", TypeID::Int64}, {"amount", TypeID::Float64}}, 4);
    const int64_t pk[] = {1, 2, 1, 9};
    for (int i = 0; i < 4; ++i) {
      probe->columns[0]->i64[i] = pk[i];
      probe->columns[1]->f64[i] = 10.0 * i;
    }
    GpuHashJoin hj(g, "user_id", "user_id");
    BatchSource bs({build});
    OK(hj.Build(ctx, &bs));
    auto pr = hj.Probe();
    BatchSource ps({probe});
    CollectSink sink;
    Pipeline p{&ps, {pr.get()}, &sink, 0};
    OK(p.Run(ctx));
    EXPECT(sink.rows() == 3, "inner join rows %d, want 3", sink.rows());
    for (auto& ob : sink.batches)
      for (int i = 0; i < ob->len; ++i)
        EXPECT(ob->columns[0]->i64[i] == ob->columns[2]->i64[i], "joined keys differ");
  }
  // join_coverage_test.go:101-141: int key path, 1 row
  {
    auto build = make_batch({{"id", TypeID::Int64}}, 2);
    build->columns[0]->i64 = {1, 2};
    auto probe = make_batch({{"id", TypeID::Int64}, {"amount", TypeID::Float64}}, 2);
    probe->columns[0]->i64 = {1, 3};
    GpuHashJoin hj(g, "id", "id");
    BatchSource bs({build});
    OK(hj.Build(ctx, &bs));
    auto pr = hj.Probe();
    BatchSource ps({probe});
    CollectSink sink;
    Pipeline p{&ps, {pr.get()}, &sink, 0};
    OK(p.Run(ctx));
    EXPECT(sink.rows() == 1, "int key path rows %d, want 1", sink.rows());
  }
  // bounded output: one probe row fans out to 40000 build rows > MaxProbeOutputRows
  {
    auto build = make_batch({{"k", TypeID::Int32}, {"v", TypeID::Int64}}, 40000);
    for (int i = 0; i < 40000; ++i) {
      build->columns[0]->i32[i] = i < 39990 ? 7 : i;
      build->columns[1]->i64[i] = i;
    }
    auto probe = make_batch({{"k", TypeID::Int64}}, 3);
    probe->columns[0]->i64 = {7, 39995, 5};
    GpuHashJoin hj(g, "k", "k");
    BatchSource bs({build});
    OK(hj.Build(ctx, &bs));
    auto pr = hj.Probe();
    BatchSource ps({probe});
    CollectSink sink;
    Pipeline p{&ps, {pr.get()}, &sink, 0};
    OK(p.Run(ctx));
    EXPECT(sink.rows() == 39991, "fan-out rows %d, want 39991", sink.rows());
    size_t max_batch = 0;
    for (auto& ob : sink.batches) max_batch = std::max<size_t>(max_batch, (size_t)ob->len);
    EXPECT(max_batch <= kMaxProbeOutputRows, "probe output batch %zu exceeds the bound", max_batch);
    EXPECT(sink.batches.size() >= 2, "bounded output must span several NextOutput calls");
  }
}

// Probe clones on pipeline workers share one join handle (and, with a Bloom
// filter in front, its lazily built filter): Workers = 4 feeding a mergeable
// GpuHashAggregate must equal the serial run and the host-computed answer.
static void test_parallel_probe_into_aggregate(wg_ctx* g) {
  auto build = make_batch({{"k", TypeID::Int64}, {"p", TypeID::Int32}}, 30000);
  for (int i = 0; i < 30000; ++i) {
    build->columns[0]->i64[i] = (i % 20000) * 13;  // keys below 10000 * 13 have two build rows (chains)
    build->columns[1]->i32[i] = i % 37;
  }
  std::vector<BatchPtr> probes;
  std::map<int64_t, std::pair<int64_t, int64_t>> want;  // p -> (sum q, count)
  for (int m = 0; m < 48; ++m) {
    auto x = make_batch({{"fk", TypeID::Int64}, {"q", TypeID::Int64}}, 2048);
    for (int i = 0; i < 2048; ++i) {
      const int64_t r = m * 2048 + i;
      const int64_t key = ((r * 7919) % 26000) * 13;  // about 23 % miss the build
      x->columns[0]->i64[i] = key;
      x->columns[1]->i64[i] = r % 101;
      for (int b = 0; b < 30000; b += 20000)
        if (key / 13 < 20000 && (key / 13) + b < 30000) {
          auto& w = want[((key / 13) + b) % 37];
          w.first += r % 101;
          w.second += 1;
        }
    }
    probes.push_back(x);
  }
  for (int bloom = 0; bloom < 2; ++bloom)
    for (int workers : {0, 4}) {
      Context ctx;
      GpuHashJoin hj(g, "fk", "k");
      BatchSource bs({build});
      OK(hj.Build(ctx, &bs));
      auto pr = hj.Probe();
      GpuBloomFilterOp bf(&hj);
      BatchSource ps(probes);
      GpuHashAggregate agg(g, {"p"}, {{WG_AGG_SUM, "q", "s"}, {WG_AGG_COUNT, "", "c"}});
      std::vector<UnaryOperator*> ops;
      if (bloom) ops.push_back(&bf);
      ops.push_back(pr.get());
      Pipeline p{&ps, ops, &agg, workers};
      OK(p.Run(ctx));
      int groups = 0;
      for (;;) {
        BatchPtr out;
        OK(agg.Next(ctx, &out));
        if (!out) break;
        for (int i = 0; i < out->len; ++i, ++groups) {
          const int64_t k = out->columns[0]->i32[i];
          EXPECT(want.count(k), "unexpected group %lld", (long long)k);
          EXPECT((int64_t)out->columns[1]->f64[i] == want[k].first, "sum of %lld (workers %d, bloom %d)",
                 (long long)k, workers, bloom);
          EXPECT(out->columns[2]->i64[i] == want[k].second, "count of %lld (workers %d, bloom %d)", (long long)k,
                 workers, bloom);
        }
      }
      EXPECT(groups == (int)want.size(), "groups %d want %zu (workers %d, bloom %d)", groups, want.size(), workers,
             bloom);
    }
}

// join_coverage_test.go:272-351 (TestSemiJoin / TestAntiJoin) and :711-735
// (TestAntiJoinNoMatches), the string ids coded as ints; plus a NULL probe key.
static void test_semi_anti_kats(wg_ctx* g) {
  Context ctx;
  auto run = [&](JoinType type, std::vector<int64_t> bk, std::vector<int64_t> pk, std::vector<bool> pnull,
                 std::vector<int64_t>* kept) {
    GpuHashJoin hj(g, "id", "id", 0, type);
    std::vector<BatchPtr> bb;
    if (!bk.empty()) {
      auto build = make_batch({{"id", TypeID::Int64}}, (int)bk.size());
      build->columns[0]->i64 = bk;
      bb.push_back(build);
    }
    BatchSource bs(bb);
    OK(hj.Build(ctx, &bs));
    auto probe = make_batch({{"id", TypeID::Int64}, {"val", TypeID::Float64}}, (int)pk.size());
    probe->columns[0]->i64 = pk;
    for (size_t i = 0; i < pnull.size(); ++i)
      if (pnull[i]) probe->columns[0]->nulls.set_null((int)i, (int)pk.size());
    auto pr = hj.Probe();
    BatchSource ps({probe});
    CollectSink sink;
    Pipeline p{&ps, {pr.get()}, &sink, 0};
    OK(p.Run(ctx));
    kept->clear();
    for (auto& ob : sink.batches) {
      EXPECT(ob->columns.size() == 2, "semi/anti output must keep the probe schema only");
      if (ob->has_sel)
        for (uint32_t r : ob->sel) kept->push_back(ob->columns[0]->i64[r]);
      else
        for (int r = 0; r < ob->len; ++r) kept->push_back(ob->columns[0]->i64[r]);
    }
  };
  std::vector<int64_t> kept;
  run(JoinType::Semi, {1, 2}, {1, 2, 3}, {}, &kept);
  EXPECT(kept == std::vector<int64_t>({1, 2}), "semi join kept %zu rows, want 2", kept.size());
  run(JoinType::Anti, {1, 2}, {1, 2, 3}, {}, &kept);
  EXPECT(kept == std::vector<int64_t>({3}), "anti join kept %zu rows, want the row with id 3", kept.size());
  run(JoinType::Anti, {}, {1, 2}, {}, &kept);
  EXPECT(kept.size() == 2, "anti join with an empty build kept %zu rows, want 2", kept.size());
  run(JoinType::Semi, {1, 2}, {1, 0, 2}, {false, true, false}, &kept);
  EXPECT(kept.size() == 2, "semi join with a NULL probe key kept %zu rows, want 2", kept.size());
  run(JoinType::Anti, {1, 2}, {1, 0, 2}, {false, true, false}, &kept);
  EXPECT(kept.size() == 1, "anti join keeps the NULL probe key: kept %zu rows, want 1", kept.size());
}

// BloomFilterOp over a built join: never drops a build key, keeps Sel order,
// and bypasses itself after 32 batches that rejected under 5 % of their rows
// (bloom_filter_op.go:201-214), while a filter that rejects half keeps going.
static void test_bloom_filter_op(wg_ctx* g) {
  Context ctx;
  GpuHashJoin hj(g, "id", "id", 0, JoinType::Inner);
  auto build = make_batch({{"id", TypeID::Int64}}, 1000);
  for (int i = 0; i < 1000; ++i) build->columns[0]->i64[i] = (int64_t)i * 7;
  BatchSource bs({build});
  OK(hj.Build(ctx, &bs));
  auto probe_batch = [](int64_t base, bool members) {
    auto b = make_batch({{"id", TypeID::Int64}}, 2048);
    for (int i = 0; i < 2048; ++i)
      b->columns[0]->i64[i] = members ? ((base + i) % 1000) * 7 : (i % 2 ? ((base + i) % 1000) * 7 : 1000000 + base + i);
    return b;
  };
  GpuBloomFilterOp all_in(&hj);
  for (int bi = 0; bi < 40; ++bi) {
    auto in = probe_batch(bi * 2048, true);
    BatchPtr out;
    OK(all_in.Execute(ctx, in, &out));
    EXPECT(out != nullptr, "a batch of build keys was dropped");
    if (!out) continue;
    if (out.get() == in.get() && !out->has_sel) continue;  // bypassed
    EXPECT(out->has_sel && out->sel.size() == 2048, "a build key was filtered out (%zu kept)", out->sel.size());
  }
  EXPECT(all_in.Disabled(), "a filter that rejects nothing must bypass itself after 32 batches");
  GpuBloomFilterOp half(&hj);
  for (int bi = 0; bi < 40; ++bi) {
    auto in = probe_batch(bi * 2048, false);
    BatchPtr out;
    OK(half.Execute(ctx, in, &out));
    EXPECT(out && out->has_sel, "half-member batch: no selection");
    if (!out) continue;
    size_t members = 0;
    for (size_t i = 1; i < out->sel.size(); ++i) EXPECT(out->sel[i] > out->sel[i - 1], "selection out of order");
    for (uint32_t r : out->sel) members += (r % 2) == 1;
    EXPECT(members == 1024, "every build key of the batch is kept (%zu of 1024)", members);
  }
  EXPECT(!half.Disabled(), "a filter rejecting ~half the rows must stay on");
}

static void test_dual_key_join_kats(wg_ctx* g) {
  Context ctx;
  // TestHashJoin_DualIntKey (dual_int_test.go:11-75): 2 of 3 probe rows match
  auto build = make_batch({{"ps_partkey", TypeID::Int64}, {"ps_suppkey", TypeID::Int64}, {"ps_value", TypeID::Float64}}, 3);
  for (int i = 0; i < 3; ++i) {
    build->columns[0]->i64[i] = i + 1;
    build->columns[1]->i64[i] = 10 * (i + 1);
    build->columns[2]->f64[i] = 100.0 * (i + 1);
  }
  auto probe = make_batch({{"l_partkey", TypeID::Int64}, {"l_suppkey", TypeID::Int64}, {"l_qty", TypeID::Int64}}, 3);
  const int64_t pk[] = {1, 2, 99}, sk[] = {10, 20, 99};
  for (int i = 0; i < 3; ++i) {
    probe->columns[0]->i64[i] = pk[i];
    probe->columns[1]->i64[i] = sk[i];
    probe->columns[2]->i64[i] = 5 * (i + 1);
  }
  {
    GpuHashJoin hj(g, std::vector<std::string>{"l_partkey", "l_suppkey"},
                   std::vector<std::string>{"ps_partkey", "ps_suppkey"});
    BatchSource bs({build});
    OK(hj.Build(ctx, &bs));
    auto pr = hj.Probe();
    BatchSource ps({probe});
    CollectSink sink;
    Pipeline p{&ps, {pr.get()}, &sink, 0};
    OK(p.Run(ctx));
    EXPECT(sink.rows() == 2, "dual-key inner join rows %d, want 2", sink.rows());
    for (auto& ob : sink.batches)
      for (int i = 0; i < ob->len; ++i)
        EXPECT(ob->columns[0]->i64[i] == ob->columns[3]->i64[i] && ob->columns[1]->i64[i] == ob->columns[4]->i64[i],
               "probe keys differ from build keys");
  }
  // TestHashJoin_DualIntKey_SemiAnti (:77-140): semi keeps 2, anti 1
  auto build2 = make_batch({{"ps_partkey", TypeID::Int64}, {"ps_suppkey", TypeID::Int64}}, 2);
  for (int i = 0; i < 2; ++i) {
    build2->columns[0]->i64[i] = i + 1;
    build2->columns[1]->i64[i] = 10 * (i + 1);
  }
  auto probe2 = make_batch({{"l_partkey", TypeID::Int64}, {"l_suppkey", TypeID::Int64}}, 3);
  const int64_t pk2[] = {1, 3, 2}, sk2[] = {10, 30, 20};
  for (int i = 0; i < 3; ++i) {
    probe2->columns[0]->i64[i] = pk2[i];
    probe2->columns[1]->i64[i] = sk2[i];
  }
  for (JoinType t : {JoinType::Semi, JoinType::Anti}) {
    GpuHashJoin hj(g, std::vector<std::string>{"l_partkey", "l_suppkey"},
                   std::vector<std::string>{"ps_partkey", "ps_suppkey"}, 0, t);
    BatchSource bs({build2});
    OK(hj.Build(ctx, &bs));
    auto pr = hj.Probe();
    BatchPtr out;
    OK(pr->Execute(ctx, probe2, &out));
    const size_t want = t == JoinType::Semi ? 2 : 1;
    EXPECT(out && out->has_sel && out->sel.size() == want, "%s: %zu rows, want %zu",
           t == JoinType::Semi ? "semi" : "anti", out ? out->sel.size() : (size_t)0, want);
  }
}

// ------------------------------------------------------------ shuffle tests

static uint64_t fnv1a64(const unsigned char* p, size_t n, uint64_t h = 14695981039346656037ull) {
  for (size_t i = 0; i < n; ++i) h = (h ^ p[i]) * 1099511628211ull;
  return h;
}

// float64 sum of `val` per int64 `key` over flushed partial batches (the
// consumer's merge, as sumByIntKey in shuffle_partial_agg_test.go:50-60)
static std::map<int64_t, double> sum_by_key(const std::vector<BatchPtr>& bs, const char* key, const char* val) {
  std::map<int64_t, double> out;
  for (auto& b : bs) {
    const int k = b->column_index(key), v = b->column_index(val);
    for (int i = 0; i < b->len; ++i) out[b->columns[k]->i64[i]] += b->columns[v]->f64[i];
  }
  return out;
}

static BatchPtr kq_batch(std::vector<int64_t> k, std::vector<double> q) {
  auto b = make_batch({{"k", TypeID::Int64}, {"q", TypeID::Float64}}, (int)k.size());
  for (size_t i = 0; i < k.size(); ++i) {
    b->columns[0]->i64[i] = k[i];
    b->columns[1]->f64[i] = q[i];
  }
  return b;
}

static std::vector<BatchPtr> drain_all(CappedPartialAgg& p, std::vector<BatchPtr> in) {
  Context ctx;
  std::vector<BatchPtr> out, f;
  for (auto& b : in) {
    OK(p.Consume(ctx, b, &f));
    out.insert(out.end(), f.begin(), f.end());
  }
  OK(p.Drain(ctx, &f));
  out.insert(out.end(), f.begin(), f.end());
  return out;
}

static void test_capped_partial_agg(wg_ctx* g) {
  {  // TestCappedPartialAgg_SumMergesAndPreservesNames (shuffle_partial_agg_test.go:62-100)
    CappedPartialAgg p(g, {"k"}, {{"sum", "q", "q"}});
    auto rows = drain_all(p, {kq_batch({1, 1, 2}, {2, 3, 10}), kq_batch({1, 3}, {5, 7})});
    int n = 0;
    for (auto& b : rows) n += b->len;
    EXPECT(n == 3, "got %d partial rows, want 3 groups", n);
    for (auto& b : rows) EXPECT(b->column_index("k") >= 0 && b->column_index("q") >= 0, "names not preserved");
    auto s = sum_by_key(rows, "k", "q");
    EXPECT(s[1] == 10 && s[2] == 10 && s[3] == 7, "sums %g %g %g", s[1], s[2], s[3]);
    EXPECT(p.in_rows == 5 && p.out_rows == 3, "counters in=%lld out=%lld", (long long)p.in_rows,
           (long long)p.out_rows);
  }
  {  // TestCappedPartialAgg_CapFlushesEpochsAndStaysMergeable (:102-129): a 1-byte cap
    CappedPartialAgg p(g, {"k"}, {{"sum", "q", "q"}}, {}, 1);
    std::vector<BatchPtr> in;
    for (int i = 0; i < 4; ++i) in.push_back(kq_batch({1, 2}, {1, 2}));
    auto rows = drain_all(p, in);
    EXPECT(p.flushes >= 4, "flushes = %lld, want >= 4", (long long)p.flushes);
    auto s = sum_by_key(rows, "k", "q");
    EXPECT(s[1] == 4 && s[2] == 8, "merged sums %g %g", s[1], s[2]);
  }
  {  // TestCappedPartialAgg_MinMax (:131-158)
    auto b = make_batch({{"k", TypeID::Int64}, {"lo", TypeID::Float64}, {"hi", TypeID::Float64}}, 3);
    const double lo[] = {5, 2, 7}, hi[] = {5, 9, 1};
    for (int i = 0; i < 3; ++i) {
      b->columns[0]->i64[i] = 1;
      b->columns[1]->f64[i] = lo[i];
      b->columns[2]->f64[i] = hi[i];
    }
    CappedPartialAgg p(g, {"k"}, {{"min", "lo", "lo"}, {"max", "hi", "hi"}});
    auto rows = drain_all(p, {b});
    EXPECT(rows.size() == 1 && rows[0]->len == 1, "want one row");
    if (!rows.empty() && rows[0]->len == 1) {
      const auto& r = rows[0];
      EXPECT(r->columns[r->column_index("lo")]->f64[0] == 2.0 && r->columns[r->column_index("hi")]->f64[0] == 9.0,
             "min/max = %g/%g, want 2/9", r->columns[r->column_index("lo")]->f64[0],
             r->columns[r->column_index("hi")]->f64[0]);
    }
  }
  {  // TestCappedPartialAgg_SelectionVectorRespected (:160-176)
    auto b = kq_batch({1, 1, 2}, {100, 3, 4});
    b->sel = {1, 2};
    b->has_sel = true;
    CappedPartialAgg p(g, {"k"}, {{"sum", "q", "q"}});
    auto s = sum_by_key(drain_all(p, {b}), "k", "q");
    EXPECT(s[1] == 3 && s[2] == 4, "sums %g %g (deselected row must not count)", s[1], s[2]);
  }
  {  // resolveAgainst (shuffle_partial_agg.go:117-151): absent keys dropped,
     // absent partition key or nothing combinable -> raw passthrough
    CappedPartialAgg absent_part(g, {"k"}, {{"sum", "q", "q"}}, {"nope"});
    auto in = kq_batch({1, 1}, {1, 2});
    auto rows = drain_all(absent_part, {in});
    EXPECT(absent_part.Disabled() && rows.size() == 1 && rows[0].get() == in.get(), "absent partition key");
    CappedPartialAgg dropped(g, {"k", "ghost"}, {{"sum", "q", "q"}});
    auto s = sum_by_key(drain_all(dropped, {kq_batch({1, 1, 2}, {1, 2, 3})}), "k", "q");
    EXPECT(!dropped.Disabled() && dropped.DroppedKeys() == 1 && s[1] == 3 && s[2] == 3, "declared-but-absent key");
    CappedPartialAgg no_aggs(g, {"k"}, {{"sum", "ghost", "ghost"}});
    drain_all(no_aggs, {kq_batch({1}, {1})});
    EXPECT(no_aggs.Disabled(), "nothing combinable must pass through");
    CappedPartialAgg unknown(g, {"k"}, {{"median", "q", "q"}});
    auto raw = kq_batch({1, 2}, {1, 2});
    auto r2 = drain_all(unknown, {raw});
    EXPECT(unknown.Disabled() && r2.size() == 1 && r2[0].get() == raw.get(), "unknown function passes through");
  }
}

static void test_shuffle_router(wg_ctx* g) {
  Context ctx;
  {  // TestPartitionedShuffleSink_RoundTrip (partitioned_shuffle_sink_test.go:19-78)
    GpuShuffleRouter r(g, {"k"}, 4);
    auto b = make_batch({{"k", TypeID::Int64}}, 1000);
    for (int i = 0; i < 1000; ++i) b->columns[0]->i64[i] = i;
    std::vector<std::vector<uint32_t>> per;
    OK(r.Route(ctx, b, &per));
    int64_t total = 0;
    for (int p = 0; p < 4; ++p) {
      for (size_t j = 0; j < per[p].size(); ++j) {
        const int64_t k = b->columns[0]->i64[per[p][j]];
        unsigned char le[8];
        for (int x = 0; x < 8; ++x) le[x] = (unsigned char)((uint64_t)k >> (8 * x));
        EXPECT((int)(fnv1a64(le, 8) % 4) == p, "row k=%lld in partition %d", (long long)k, p);
        if (j) EXPECT(per[p][j] > per[p][j - 1], "partition %d out of input order", p);
      }
      total += (int64_t)per[p].size();
      EXPECT(r.PartitionRowCounts()[p] == (int64_t)per[p].size(), "row counts");
    }
    EXPECT(total == 1000, "total rows %lld, want 1000", (long long)total);
  }
  {  // multi-column keys with NULLs and a Sel: hashVectorValue's byte stream
    GpuShuffleRouter r(g, {"a", "p"}, 7);
    auto b = make_batch({{"a", TypeID::Int64}, {"p", TypeID::Port}, {"x", TypeID::Float64}}, 300);
    for (int i = 0; i < 300; ++i) {
      b->columns[0]->i64[i] = (int64_t)i * 2654435761ll - 99;
      b->columns[1]->i32[i] = (i * 37) % 65536;
    }
    for (int i = 0; i < 300; i += 11) b->columns[1]->nulls.set_null(i, 300);
    for (int i = 0; i < 300; i += 3) b->sel.push_back((uint32_t)i);
    b->has_sel = true;
    std::vector<std::vector<uint32_t>> per;
    std::vector<int> parts;
    OK(r.Route(ctx, b, &per, &parts));
    EXPECT(parts.size() == b->sel.size(), "one partition per active row");
    size_t seen = 0;
    for (int p = 0; p < 7; ++p) {
      seen += per[p].size();
      for (uint32_t row : per[p]) {
        unsigned char buf[12];
        size_t len = 0;
        const int64_t a = b->columns[0]->i64[row];
        for (int x = 0; x < 8; ++x) buf[len++] = (unsigned char)((uint64_t)a >> (8 * x));
        if (!b->columns[1]->nulls.valid((int)row)) {
          buf[len++] = 0xff;
        } else {
          const uint32_t v = (uint32_t)b->columns[1]->i32[row];
          for (int x = 0; x < 4; ++x) buf[len++] = (unsigned char)(v >> (8 * x));
        }
        EXPECT((int)(fnv1a64(buf, len) % 7) == p, "row %u in partition %d", row, p);
        EXPECT(row % 3 == 0, "row %u is not in the Sel", row);
        EXPECT(parts[row / 3] == p, "parts[%u] = %d, want %d", row / 3, parts[row / 3], p);
      }
    }
    EXPECT(seen == b->sel.size(), "rows lost");
  }
  {
    GpuShuffleRouter r(g, {"missing"}, 4);
    std::vector<std::vector<uint32_t>> per;
    EXPECT(!r.Route(ctx, kq_batch({1}, {1}), &per).ok(), "an absent key must error (:212-216)");
  }
}

// ------------------------------------------------- late materialization

static std::string cell(const Vector& v, int i) {
  if (!v.nulls.valid(i)) return "null";
  char buf[64];
  switch (v.type) {
    case TypeID::Int32: case TypeID::Port: case TypeID::Date: std::snprintf(buf, sizeof buf, "%d", v.i32[i]); break;
    case TypeID::Float64: std::snprintf(buf, sizeof buf, "%.17g", v.f64[i]); break;
    case TypeID::Decimal: std::snprintf(buf, sizeof buf, "%lld:%llu", (long long)v.dec[i].hi, (unsigned long long)v.dec[i].lo); break;
    default: std::snprintf(buf, sizeof buf, "%lld", (long long)v.i64[i]); break;
  }
  return buf;
}

// the rows of flattened batches as text, sorted (pair order is unspecified)
static std::vector<std::string> rows_of(const std::vector<BatchPtr>& bs) {
  std::vector<std::string> out;
  for (auto& b : bs)
    for (int i = 0; i < b->len; ++i) {
      std::string r;
      for (size_t c = 0; c < b->columns.size(); ++c) r += b->schema[c].name + "=" + cell(*b->columns[c], i) + " ";
      out.push_back(r);
    }
  std::sort(out.begin(), out.end());
  return out;
}

static std::vector<int64_t> i64_values(Vector& v) {
  std::vector<int64_t> out;
  for (int i = 0; i < v.len; ++i) out.push_back(v.nulls.valid(i) ? v.i64[i] : INT64_MIN);
  return out;
}

static void test_view_vectors() {
  const int64_t N = INT64_MIN;  // NULL marker in the expectations
  // TestViewFlattenAllFlatTypes, int64 case (view_test.go:53-93)
  auto base = Vector::make(TypeID::Int64, 4);
  base->i64 = {1, 0, -9000000000000000ll, 42};
  base->nulls.set_null(1, 4);
  auto v = NewViewVector(base, {3, 0, 0, 2, 1, 3});
  EXPECT(v->is_view() && v->len == 6, "a view of 6 rows");
  v->Flatten();
  EXPECT(!v->is_view() && i64_values(*v) == (std::vector<int64_t>{42, 1, 1, -9000000000000000ll, N, 42}),
         "flatten through a permuting, duplicating view");
  // TestViewOwnNullOverride (:155-169)
  auto ints = Vector::make(TypeID::Int64, 2);
  ints->i64 = {10, 20};
  auto iv = NewViewVector(ints, {1, 0});
  iv->nulls.set_null(0, 2);
  iv->Flatten();
  EXPECT(i64_values(*iv) == (std::vector<int64_t>{N, 10}), "own NULL overrides the index");
  // TestViewComposition (:171-183)
  auto b4 = Vector::make(TypeID::Int64, 4);
  b4->i64 = {0, 10, 20, 30};
  auto v1 = NewViewVector(b4, {3, 2, 1, 0});
  v1->nulls.set_null(2, 4);
  auto v2 = NewViewVector(v1, {2, 0, 1, 2});
  EXPECT(v2->base == b4, "composition collapses to the owned base");
  v2->Flatten();
  EXPECT(i64_values(*v2) == (std::vector<int64_t>{N, 30, 20, N}), "composed view values");
  // TestViewFlattenCopiesStorage (:199-211) and TestViewEmptyAndSingle (:213-219)
  auto f = Vector::make(TypeID::Float64, 2);
  f->f64 = {1.5, 2.5};
  auto fv = NewViewVector(f, {1, 0});
  fv->Flatten();
  f->f64[0] = -1;
  EXPECT(fv->f64[0] == 2.5 && fv->f64[1] == 1.5, "flatten copies, it does not alias");
  auto one = Vector::make(TypeID::Int64, 1);
  one->i64 = {5};
  auto e = NewViewVector(one, {});
  e->Flatten();
  EXPECT(e->len == 0, "empty view");
  auto s1 = NewViewVector(one, {0});
  s1->Flatten();
  EXPECT(i64_values(*s1) == std::vector<int64_t>{5}, "single-row view");
  // TestBatchHasViewsFlattenViews (:280-319)
  auto bt = make_batch({{"a", TypeID::Int64}, {"b", TypeID::Int64}}, 2);
  EXPECT(!bt->HasViews(), "owned batch has no views");
  bt->columns[1] = NewViewVector(b4, {0, 3});
  EXPECT(bt->HasViews(), "batch with a view column");
  bt->FlattenViews();
  EXPECT(!bt->HasViews() && bt->columns[1]->i64[1] == 30, "FlattenViews");
}

// An operator that re-emits its input as views (a stand-in for a late-
// materializing producer), for the pipeline's defensive flatten.
struct ViewEmitter : UnaryOperator {
  Status Init(Context&) override { return Status{}; }
  Status Close() override { return Status{}; }
  Status Execute(Context&, const BatchPtr& in, BatchPtr* out) override {
    auto o = std::make_shared<RecordBatch>();
    o->schema = in->schema;
    o->len = in->len;
    std::vector<uint32_t> idx(in->len);
    for (int i = 0; i < in->len; ++i) idx[i] = (uint32_t)(in->len - 1 - i);
    for (auto& c : in->columns) o->columns.push_back(NewViewVector(c, idx));
    *out = o;
    return Status{};
  }
};

static void test_pipeline_flattens_views() {
  Context ctx;
  auto b = make_batch({{"k", TypeID::Int64}}, 3);
  b->columns[0]->i64 = {1, 2, 3};
  BatchSource src({b});
  ViewEmitter op;
  CollectSink sink;
  Pipeline p{&src, {&op}, &sink, 0};
  const int64_t before = LateMat.flattens.load();
  OK(p.Run(ctx));
  EXPECT(LateMat.flattens.load() > before, "no flatten for a sink that is not view-aware");
  for (auto& ob : sink.batches) EXPECT(!ob->HasViews(), "the sink kept a view");
  EXPECT(sink.rows() == 3 && sink.batches[0]->columns[0]->i64[0] == 3, "reversed rows through the view");
}

// late_mat_test.go: orders (o_orderkey, o_comment) and lineitem (l_orderkey,
// l_price, l_ship); the reference's string columns are int64 codes here.
static BatchPtr late_build() {
  auto b = make_batch({{"o_orderkey", TypeID::Int64}, {"o_comment", TypeID::Int64}}, 3);
  b->columns[0]->i64 = {1, 2, 3};
  b->columns[1]->i64 = {101, 102, 100};
  return b;
}
static BatchPtr late_probe() {
  auto b = make_batch({{"l_orderkey", TypeID::Int64}, {"l_price", TypeID::Float64}, {"l_ship", TypeID::Int64}}, 4);
  b->columns[0]->i64 = {2, 9, 1, 2};  // 9: no match; the second 2: 1:N with row 0
  b->columns[1]->f64 = {20.5, 90.0, 10.0, 21.0};
  b->columns[2]->i64 = {7, 8, 0, 9};
  return b;
}

static std::vector<std::string> run_late_join(wg_ctx* g, bool late, JoinType t = JoinType::Inner,
                                              bool* had_views = nullptr) {
  Context ctx;
  GpuHashJoin hj(g, "l_orderkey", "o_orderkey", 0, t);
  BatchSource bs({late_build()});
  OK(hj.Build(ctx, &bs));
  auto pr = hj.Probe();
  pr->LateMaterialize = late;
  BatchPtr out;
  OK(pr->Execute(ctx, late_probe(), &out));
  if (!out) return {};
  if (had_views) *had_views = out->HasViews();
  out->FlattenViews();
  EXPECT(!out->HasViews(), "FlattenViews left views");
  return rows_of({out});
}

static void test_late_materialized_join(wg_ctx* g) {
  // TestLateMatInnerJoinParity (late_mat_test.go:76-92)
  const int64_t before = LateMat.batches_emitted.load();
  bool views = true;
  auto eager = run_late_join(g, false, JoinType::Inner, &views);
  EXPECT(!views, "eager output contains views");
  EXPECT(LateMat.batches_emitted.load() == before, "dormancy: the counter moved with the flag off");
  views = false;
  auto lazy = run_late_join(g, true, JoinType::Inner, &views);
  EXPECT(views, "late-materialized output has no view columns");
  EXPECT(LateMat.batches_emitted.load() > before, "engagement: the counter did not move with the flag on");
  EXPECT(eager == lazy && lazy.size() == 3, "inner join parity (%zu vs %zu rows)", eager.size(), lazy.size());
  // TestLateMatSemiAntiUnaffected (:119-149)
  for (JoinType t : {JoinType::Semi, JoinType::Anti}) {
    const int64_t b0 = LateMat.batches_emitted.load();
    views = false;
    run_late_join(g, true, t, &views);
    EXPECT(!views && LateMat.batches_emitted.load() == b0, "semi/anti must not emit views");
  }
  // TestLateMatPipelineFlattensForSink (:151-185)
  {
    Context ctx;
    GpuHashJoin hj(g, "l_orderkey", "o_orderkey");
    BatchSource bs({late_build()});
    OK(hj.Build(ctx, &bs));
    auto pr = hj.Probe();
    pr->LateMaterialize = true;
    BatchSource ps({late_probe()});
    CollectSink sink;
    Pipeline p{&ps, {pr.get()}, &sink, 0};
    const int64_t f0 = LateMat.flattens.load();
    OK(p.Run(ctx));
    EXPECT(LateMat.flattens.load() > f0, "the pipeline did not flatten for the sink");
    for (auto& ob : sink.batches) EXPECT(!ob->HasViews(), "the sink kept a batch with live views");
    EXPECT(sink.rows() == 3, "pipeline rows %d, want 3", sink.rows());
  }
  // TestLateMatChainComposition (:236-330): lineitem -> orders -> customer
  for (bool late : {false, true}) {
    Context ctx;
    auto line = make_batch({{"l_orderkey", TypeID::Int64}, {"l_price", TypeID::Float64}, {"l_comment", TypeID::Int64}}, 4);
    line->columns[0]->i64 = {1, 2, 1, 3};  // 3: no order
    line->columns[1]->f64 = {10.0, 20.0, 11.0, 30.0};
    line->columns[2]->i64 = {1, 2, 3, 4};
    auto orders = make_batch({{"o_orderkey", TypeID::Int64}, {"o_custkey", TypeID::Int64}}, 2);
    orders->columns[0]->i64 = {1, 2};
    orders->columns[1]->i64 = {100, 200};
    auto cust = make_batch({{"c_custkey", TypeID::Int64}, {"c_name", TypeID::Int64}}, 2);
    cust->columns[0]->i64 = {100, 200};
    cust->columns[1]->i64 = {7001, 7002};
    GpuHashJoin hj1(g, "l_orderkey", "o_orderkey"), hj2(g, "o_custkey", "c_custkey");
    BatchSource s1({orders}), s2({cust});
    OK(hj1.Build(ctx, &s1));
    OK(hj2.Build(ctx, &s2));
    auto p1 = hj1.Probe(), p2 = hj2.Probe();
    p1->LateMaterialize = p2->LateMaterialize = late;
    const int64_t f0 = LateMat.flattens.load();
    BatchPtr out1, out2;
    OK(p1->Execute(ctx, line, &out1));
    OK(p2->Execute(ctx, out1, &out2));
    if (!out1 || !out2) {
      EXPECT(false, "chain produced no output");
      continue;
    }
    if (late) {
      const int ck = out1->column_index("o_custkey"), pc = out1->column_index("l_price");
      EXPECT(!out1->columns[ck]->is_view(), "probe 2 did not flatten its key column");
      EXPECT(out1->columns[pc]->is_view(), "probe 2 flattened a pass-through column");
      EXPECT(LateMat.flattens.load() - f0 == 1, "flatten events %lld, want 1 (probe 2's key)",
             (long long)(LateMat.flattens.load() - f0));
      EXPECT(out1->columns[pc]->base == line->columns[1], "the pass-through view does not compose to the scan batch");
      const int pc2 = out2->column_index("l_price");
      EXPECT(out2->columns[pc2]->is_view() && out2->columns[pc2]->base == line->columns[1],
             "join 2's pass-through view does not compose to the scan batch");
    }
    out2->FlattenViews();
    static std::vector<std::string> eager_rows;
    auto rows = rows_of({out2});
    if (!late) eager_rows = rows;
    else EXPECT(rows == eager_rows && rows.size() == 3, "chain parity (%zu rows)", rows.size());
  }
}

static void test_allocation_failures_reach_the_pipeline(wg_ctx* g) {
  // The reference forces its error paths with a small memory budget and
  // FatalEvalPanic (pipeline.go:45-57). Here every device allocation a
  // pipeline run makes is in turn made to fail: Run returns the error with
  // its code (out of memory, so an adapter can fall back to the CPU
  // operator), serial and with 4 workers, and the next run on the same
  // context is exact.
  std::vector<BatchPtr> src;
  std::map<int64_t, std::pair<int64_t, int64_t>> want;  // k -> (sum v, count)
  for (int m = 0; m < 20; ++m) {
    auto x = make_batch({{"k", TypeID::Int64}, {"v", TypeID::Int64}}, 2048);
    for (int i = 0; i < 2048; ++i) {
      const int64_t r = m * 2048 + i;
      x->columns[0]->i64[i] = (r * 7919) % 5000;
      x->columns[1]->i64[i] = r * 3 - 1000;
      auto& w = want[x->columns[0]->i64[i]];
      w.first += x->columns[1]->i64[i];
      w.second += 1;
    }
    src.push_back(x);
  }
  for (int workers : {0, 4}) {
    int failures = 0;
    bool completed = false;
    for (int k = 1; k < 400 && !completed; ++k) {
      BatchSource s(src);
      GpuHashAggregate agg(g, {"k"}, {{WG_AGG_SUM, "v", "s"}, {WG_AGG_COUNT, "", "c"}}, 64);
      Pipeline p{&s, {}, &agg, workers};
      Context ctx;
      wg_ctx_inject_alloc_failure(g, k, 0);
      const Status st = p.Run(ctx);
      int groups = 0;
      Status st2;  // ok
      std::map<int64_t, std::pair<int64_t, int64_t>> got;
      if (st.ok()) {
        for (;;) {  // the output drain allocates too
          BatchPtr out;
          st2 = agg.Next(ctx, &out);
          if (!st2.ok() || !out) break;
          for (int i = 0; i < out->len; ++i, ++groups)
            got[out->columns[0]->i64[i]] = {(int64_t)out->columns[1]->f64[i], out->columns[2]->i64[i]};
        }
      }
      wg_ctx_inject_alloc_failure(g, 0, 0);
      const Status& bad = st.ok() ? st2 : st;
      if (bad.ok()) {
        completed = true;
        EXPECT(groups == (int)want.size(), "workers %d: %d groups, want %zu", workers, groups, want.size());
        for (auto& [key, sv] : want) {
          auto it = got.find(key);
          EXPECT(it != got.end() && it->second == sv, "workers %d: group %lld", workers, (long long)key);
        }
        continue;
      }
      ++failures;
      EXPECT(bad.out_of_memory(), "workers %d, allocation %d: code %d: %s", workers, k, bad.code, bad.msg.c_str());
      EXPECT(bad.msg.find("out of memory") != std::string::npos, "message: %s", bad.msg.c_str());
    }
    EXPECT(completed, "workers %d: the sequence never completed", workers);
    EXPECT(failures >= 3, "workers %d: only %d allocations failed", workers, failures);
  }
}

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "cpu";
  test_kernel_filter_and_pipeline();
  test_view_vectors();
  test_pipeline_flattens_views();
  if (mode == "gpu") {
    wg_ctx* g = nullptr;
    if (wg_ctx_create(0, &g) != WG_OK) {
      std::printf("FAIL: no device: %s\n", wg_last_error());
      return 2;
    }
    test_hash_aggregate_kat(g);
    test_parallel_pushdown_equivalence(g);
    test_hash_join_kats(g);
    test_parallel_probe_into_aggregate(g);
    test_semi_anti_kats(g);
    test_bloom_filter_op(g);
    test_dual_key_join_kats(g);
    test_capped_partial_agg(g);
    test_shuffle_router(g);
    test_late_materialized_join(g);
    test_allocation_failures_reach_the_pipeline(g);
    wg_ctx_destroy(g);
  }
  std::printf("%s: %d failure(s)\n", mode.c_str(), g_fail);
  return g_fail ? 1 : 0;
}
