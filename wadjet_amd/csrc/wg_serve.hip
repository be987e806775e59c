// wg_serve.hip — request loop with continuous batching over a device-resident
// table (the served form of the hash aggregate / join probe).
//
// One engine thread owns `lanes` batch runners, each on its own HIP stream.
// Requests queue with their arrival time; whenever a lane is idle the engine
// admits queued requests in FIFO order while their hash tables fit the free
// HBM budget (and the lane's batch size), and launches them as one persistent
// batch on that lane. Lanes run concurrently on the device, so requests that
// arrive while earlier batches run are admitted as soon as a lane and the
// memory free up. A finished lane's results are copied out, the requests are
// completed (latency = completion - arrival) and the lane takes the next batch.
// mode 1 (plain) is the same loop with one lane and one request per launch.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

#include "wg_core.hpp"
#include "wg_join_host.hpp"

// asynchronous batch API of wg_batch.hip
wg_status wg_batch_submit_async(wg_batch* b, const wg_batch_query* qs, int32_t n, int32_t use_graph);
wg_status wg_batch_poll(wg_batch* b, int32_t* finished, int32_t wait);
wg_status wg_batch_download_async(wg_batch* b, uint64_t* host, uint64_t host_words, uint64_t* offsets,
                                  uint64_t* words);
int32_t wg_batch_rerun_alone(const wg_batch* b, int32_t q);

namespace wg {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t id;
  wg_serve_query q;
  wg_batch_query bq;
  uint64_t footprint;
  Clock::time_point arrival, done_at;
  bool done = false;
  wg_status status = WG_OK;
  uint64_t n_groups = 0;
  std::vector<uint64_t> res;  // keys, sum low words, sum high words, counts: n_groups words each
};

// Results of a finished batch on their way to the host: one packed copy into
// one of the lane's two pinned buffers, the requests completed when its event
// fires. The lane itself is free again as soon as the copy is enqueued.
struct Download {
  std::vector<Request*> reqs;
  std::vector<uint64_t> off, n;  // word offset and group count of each request's runs
  hipEvent_t ev = nullptr;
  int buf = 0;
};

struct Lane {
  wg_ctx* ctx = nullptr;
  wg_batch* batch = nullptr;
  std::vector<Request*> reqs;
  uint64_t footprint = 0;
  bool busy = false;
  uint64_t* host[2] = {nullptr, nullptr};  // pinned result buffers
  uint64_t host_words[2] = {0, 0};
  int next_buf = 0;
  std::deque<Download> downloads;  // in flight, oldest first (at most two)
  std::vector<hipEvent_t> free_events;
  hipEvent_t t0 = nullptr, t1 = nullptr;  // timing: before the batch's first and after its last device operation
};

static uint64_t footprint_of(uint64_t est) {
  uint64_t cap = 64;
  while (cap < std::max<uint64_t>(est, 8) * 2) cap <<= 1;
  return (cap + 1) * 32 + (cap + 1) * 32;  // table slots + compacted outputs
}

}  // namespace wg

struct wg_serve {
  wg_ctx* ctx = nullptr;
  std::vector<wg_column> cols;
  uint64_t n_rows = 0;
  int32_t ts_col = -1;
  std::vector<int64_t> ts_index;  // ts at every kIndexStride-th row
  int32_t mode = 0;
  int32_t max_batch = 64;
  uint64_t budget = 0;
  std::vector<wg::Lane> lanes;
  std::vector<const wg_join*> dims;
  // queue and requests
  std::mutex mu;
  std::condition_variable cv_arrive, cv_done;
  std::deque<wg::Request*> queue;
  std::unordered_map<uint64_t, wg::Request*> reqs;
  uint64_t next_id = 1;
  bool stop = false;
  std::thread engine;
  // statistics
  uint64_t batches = 0, max_in_flight = 0, peak_fp = 0, in_use_fp = 0;
  hipEvent_t origin = nullptr;                  // device time zero of the intervals below
  std::vector<std::pair<double, double>> busy;  // per batch: device interval (ms from origin)
  double host_submit_s = 0, host_retire_s = 0, host_deliver_s = 0;
  wg::Clock::time_point first_arrival{}, last_done{};
  bool any = false;
  wg_status engine_error = WG_OK;
  // result buffers of released requests, by capacity in words: the engine
  // fills reused (already faulted-in) memory instead of fresh pages
  std::multimap<uint64_t, std::vector<uint64_t>> pool;
  uint64_t pool_words = 0;
};

namespace wg {

constexpr uint64_t kIndexStride = 4096;
constexpr uint64_t kPoolMaxWords = (8ull << 30) / 8;  // at most 8 GiB of pooled result buffers

// A pooled result buffer of at least `words` words and at most about four
// times that (empty if none): the caller holds s->mu.
static std::vector<uint64_t> take_buffer(wg_serve* s, uint64_t words) {
  auto it = s->pool.lower_bound(words);
  if (it == s->pool.end() || it->first > 4 * words + 16384) return {};
  std::vector<uint64_t> v = std::move(it->second);
  s->pool_words -= it->first;
  s->pool.erase(it);
  return v;
}

// Return a released request's result buffer to the pool: the caller holds s->mu.
static void give_buffer(wg_serve* s, std::vector<uint64_t>&& v) {
  const uint64_t c = v.capacity();
  if (c == 0 || s->pool_words + c > kPoolMaxWords) return;
  s->pool_words += c;
  s->pool.emplace(c, std::move(v));
}

// Row range [r0, r1) that covers ts window [lo, hi) at index granularity; the
// per-row window filter in the batch kernel makes it exact.
static void window_rows(const wg_serve* s, int64_t lo, int64_t hi, uint64_t& r0, uint64_t& r1) {
  const auto& ix = s->ts_index;
  // last indexed block whose first ts <= lo
  size_t b0 = (size_t)(std::upper_bound(ix.begin(), ix.end(), lo) - ix.begin());
  b0 = b0 ? b0 - 1 : 0;
  // first indexed block whose first ts >= hi: rows before it
  const size_t b1 = (size_t)(std::lower_bound(ix.begin(), ix.end(), hi) - ix.begin());
  r0 = std::min<uint64_t>(b0 * kIndexStride, s->n_rows);
  r1 = std::min<uint64_t>(b1 * kIndexStride, s->n_rows);
  if (r1 < r0) r1 = r0;
}

static wg_status to_batch_query(wg_serve* s, const wg_serve_query& q, wg_batch_query& bq) {
  WG_CHECK(q.key_col >= 0 && q.key_col < (int)s->cols.size() && q.val_col >= 0 && q.val_col < (int)s->cols.size(),
           WG_ERR_INVALID, "query columns out of range");
  bq = wg_batch_query{};
  bq.key = s->cols[q.key_col];
  bq.val = s->cols[q.val_col];
  bq.est_groups = std::max<uint64_t>(q.est_groups, 8);
  uint64_t r0 = 0, r1 = s->n_rows;
  if (s->ts_col >= 0) {
    window_rows(s, q.ts_lo, q.ts_hi, r0, r1);
    bq.range_col = s->cols[s->ts_col];
    bq.range_lo = q.ts_lo;
    bq.range_hi = q.ts_hi;
  }
  bq.row_begin = r0;
  bq.n_rows = r1 - r0;
  if (q.kind == 1) {
    WG_CHECK(q.dim >= 0 && q.dim < (int)s->dims.size(), WG_ERR_INVALID, "dimension %d", q.dim);
    bq.dim = s->dims[q.dim];
    bq.dim_payload = q.dim_payload;
  } else {
    WG_CHECK(q.kind == 0, WG_ERR_INVALID, "query kind %d", q.kind);
  }
  return WG_OK;
}

// Hand the results of the oldest download of L to its requests (the copy has
// completed) and complete them.
static void finish_download(wg_serve* s, Lane& L) {
  const auto t_enter = Clock::now();
  Download d = std::move(L.downloads.front());
  L.downloads.pop_front();
  const uint64_t* h = L.host[d.buf];
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (size_t i = 0; i < d.reqs.size(); ++i)
      if (d.reqs[i] && d.n[i]) d.reqs[i]->res = take_buffer(s, 4 * d.n[i]);
  }
  for (size_t i = 0; i < d.reqs.size(); ++i) {
    Request* r = d.reqs[i];
    const uint64_t n = d.n[i];
    if (r == nullptr) continue;  // completed already (re-run alone)
    // the download packs keys, sum words and counts of a request back to back
    const uint64_t* p = h + d.off[i];
    r->res.assign(p, p + 4 * n);
    r->n_groups = n;
  }
  L.free_events.push_back(d.ev);
  std::lock_guard<std::mutex> lk(s->mu);
  const auto now = Clock::now();
  for (Request* r : d.reqs) {
    if (r == nullptr) continue;
    r->done_at = now;
    r->done = true;
  }
  s->last_done = now;
  s->host_deliver_s += std::chrono::duration<double>(now - t_enter).count();
  s->cv_done.notify_all();
}

// Complete the requests of a lane whose batch finished with status st, or
// start their download; the lane is free for the next batch on return.
static void retire_lane(wg_serve* s, Lane& L, wg_status st) {
  const auto t_enter = Clock::now();
  float a_ms = 0, b_ms = 0;
  const bool timed = st == WG_OK && hipEventElapsedTime(&a_ms, s->origin, L.t0) == hipSuccess &&
                     hipEventElapsedTime(&b_ms, s->origin, L.t1) == hipSuccess;
  Download d;
  d.reqs = L.reqs;
  d.off.assign(L.reqs.size(), 0);
  d.n.assign(L.reqs.size(), 0);
  std::vector<Request*> now_done;
  if (st == WG_OK) {
    // requests whose query was re-run alone hold their results on the host already
    for (size_t i = 0; i < L.reqs.size(); ++i) {
      uint64_t n = 0;
      wg_status e = wg_batch_result(L.batch, (int32_t)i, &n, nullptr, nullptr, nullptr, nullptr, 0);
      d.n[i] = n;
      const bool alone = wg_batch_rerun_alone(L.batch, (int32_t)i) != 0;
      if (e == WG_OK && n && alone) {
        Request* r = L.reqs[i];
        {
          std::lock_guard<std::mutex> lk(s->mu);
          r->res = take_buffer(s, 4 * n);
        }
        r->res.resize(4 * n);
        uint64_t* q = r->res.data();
        e = wg_batch_result(L.batch, (int32_t)i, &n, (int64_t*)q, q + n, (int64_t*)(q + 2 * n), (int64_t*)(q + 3 * n),
                            n);
        r->n_groups = n;
      }
      if (e != WG_OK || alone) {
        L.reqs[i]->status = e;
        now_done.push_back(L.reqs[i]);
        d.reqs[i] = nullptr;
      }
    }
    // a pinned buffer free of undelivered results
    if (L.downloads.size() >= 2) {
      (void)hipEventSynchronize(L.downloads.front().ev);
      finish_download(s, L);
    }
    d.buf = L.next_buf;
    L.next_buf ^= 1;
    uint64_t words = 0;
    st = wg_batch_download_async(L.batch, L.host[d.buf], L.host_words[d.buf], d.off.data(), &words);
    if (st == WG_ERR_OOM) {  // grow this buffer (no download uses it now)
      (void)hipHostFree(L.host[d.buf]);
      L.host[d.buf] = nullptr;
      L.host_words[d.buf] = 0;
      const uint64_t want = std::max<uint64_t>(words + words / 2, 1 << 20);
      if (hipHostMalloc((void**)&L.host[d.buf], want * 8, hipHostMallocDefault) == hipSuccess) {
        L.host_words[d.buf] = want;
        st = wg_batch_download_async(L.batch, L.host[d.buf], L.host_words[d.buf], d.off.data(), &words);
      }
    }
    if (st == WG_OK) {
      if (L.free_events.empty()) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) st = WG_ERR_DEVICE;
        else L.free_events.push_back(e);
      }
    }
    if (st == WG_OK) {
      d.ev = L.free_events.back();
      L.free_events.pop_back();
      if (hipEventRecord(d.ev, L.ctx->stream) != hipSuccess) st = WG_ERR_DEVICE;
    }
  }
  if (st != WG_OK) {  // the batch or its download failed: every request fails
    for (Request*& r : d.reqs)
      if (r) {
        r->status = st;
        now_done.push_back(r);
        r = nullptr;
      }
  } else {
    L.downloads.push_back(std::move(d));
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (st != WG_OK) s->engine_error = st;
  const auto now = Clock::now();
  if (timed) s->busy.push_back({a_ms, b_ms});
  s->host_retire_s += std::chrono::duration<double>(now - t_enter).count();
  for (Request* r : now_done) {
    r->done_at = now;
    r->done = true;
  }
  s->in_use_fp -= L.footprint;  // the tables are packed away: the arena takes the next batch
  if (!now_done.empty()) s->last_done = now;
  L.reqs.clear();
  L.footprint = 0;
  L.busy = false;
  s->cv_done.notify_all();
}

static void engine_loop(wg_serve* s) {
  (void)hipSetDevice(s->ctx->device);
  for (;;) {
    bool progressed = false;
    // 1. retire finished lanes (their results start downloading), deliver
    //    downloads that completed
    for (Lane& L : s->lanes) {
      if (L.busy) {
        int32_t fin = 0;
        const wg_status st = wg_batch_poll(L.batch, &fin, 0);
        if (fin || st != WG_OK) {
          retire_lane(s, L, st);
          progressed = true;
        }
      }
      while (!L.downloads.empty() && hipEventQuery(L.downloads.front().ev) == hipSuccess) {
        finish_download(s, L);
        progressed = true;
      }
    }
    // 2. admit queued requests into idle lanes, FIFO, within the HBM budget
    for (Lane& L : s->lanes) {
      if (L.busy) continue;
      std::vector<Request*> take;
      {
        std::lock_guard<std::mutex> lk(s->mu);
        uint64_t fp = 0;
        while (!s->queue.empty() && (int)take.size() < s->max_batch) {
          Request* r = s->queue.front();
          const bool fits = s->in_use_fp + fp + r->footprint <= s->budget;
          if (!fits && !(take.empty() && s->in_use_fp == 0)) break;  // an oversized request runs alone
          take.push_back(r);
          fp += r->footprint;
          s->queue.pop_front();
        }
        if (take.empty()) continue;
        s->in_use_fp += fp;
        L.footprint = fp;
        s->peak_fp = std::max(s->peak_fp, s->in_use_fp);
      }
      std::vector<wg_batch_query> qs(take.size());
      for (size_t i = 0; i < take.size(); ++i) qs[i] = take[i]->bq;
      L.reqs = take;
      L.busy = true;
      const auto t_sub = Clock::now();
      wg_status st = hipEventRecord(L.t0, L.ctx->stream) == hipSuccess ? WG_OK : WG_ERR_DEVICE;
      if (st == WG_OK) st = wg_batch_submit_async(L.batch, qs.data(), (int32_t)qs.size(), 1);
      if (st == WG_OK && hipEventRecord(L.t1, L.ctx->stream) != hipSuccess) st = WG_ERR_DEVICE;
      {
        std::lock_guard<std::mutex> lk(s->mu);
        s->host_submit_s += std::chrono::duration<double>(Clock::now() - t_sub).count();
      }
      if (st != WG_OK) {
        for (Request* r : take) r->status = st;
        std::lock_guard<std::mutex> lk(s->mu);
        s->engine_error = st;
        for (Request* r : take) {
          r->done = true;
          r->done_at = Clock::now();
        }
        s->in_use_fp -= L.footprint;
        L.reqs.clear();
        L.footprint = 0;
        L.busy = false;
        s->cv_done.notify_all();
        continue;
      }
      ++s->batches;
      uint64_t busy = 0;
      for (Lane& M : s->lanes) busy += M.busy;
      s->max_in_flight = std::max(s->max_in_flight, busy);
      progressed = true;
    }
    // 3. sleep only when idle
    bool any_busy = false;
    for (Lane& L : s->lanes) any_busy |= L.busy || !L.downloads.empty();
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->stop && s->queue.empty() && !any_busy) return;
    if (!progressed) {
      if (any_busy) {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
      } else if (s->queue.empty()) {
        s->cv_arrive.wait_for(lk, std::chrono::milliseconds(5));
      }
    }
  }
}

}  // namespace wg

using namespace wg;

extern "C" {

wg_status wg_serve_create(wg_ctx* ctx, const wg_column* cols, int32_t n_cols, uint64_t n_rows, int32_t ts_col,
                          int32_t lanes, uint64_t hbm_budget, int32_t max_batch, int32_t mode, wg_serve** out) {
  clear_error();
  WG_CHECK(ctx && cols && out, WG_ERR_INVALID, "null argument");
  WG_CHECK(n_cols > 0 && n_cols <= kMaxCols, WG_ERR_INVALID, "%d columns", n_cols);
  WG_CHECK(ts_col < n_cols, WG_ERR_INVALID, "ts column %d", ts_col);
  WG_CHECK(mode == 0 || mode == 1, WG_ERR_INVALID, "mode %d", mode);
  if (mode == 1) {
    lanes = 1;
    max_batch = 1;
  }
  WG_CHECK(lanes >= 1 && lanes <= 8 && max_batch >= 1 && max_batch <= 4096, WG_ERR_INVALID, "lanes / batch size");
  wg_serve* s = new wg_serve();
  s->ctx = ctx;
  s->cols.assign(cols, cols + n_cols);
  s->n_rows = n_rows;
  s->ts_col = ts_col;
  s->mode = mode;
  s->max_batch = max_batch;
  s->budget = hbm_budget;
  if (ts_col >= 0) {
    WG_CHECK(cols[ts_col].type == WG_INT64, WG_ERR_UNSUPPORTED, "ts column must be INT64");
    const uint64_t m = (n_rows + kIndexStride - 1) / kIndexStride;
    s->ts_index.resize(m);
    if (m)
      WG_HIP(hipMemcpy2D(s->ts_index.data(), 8, cols[ts_col].data, kIndexStride * 8, 8, m, hipMemcpyDeviceToHost));
  }
  WG_HIP(hipEventCreate(&s->origin));
  WG_HIP(hipEventRecord(s->origin, ctx->stream));
  WG_HIP(hipEventSynchronize(s->origin));
  s->lanes.resize(lanes);
  for (Lane& L : s->lanes) {
    WG_TRY(wg_ctx_create(ctx->device, &L.ctx));
    WG_HIP(hipEventCreate(&L.t0));
    WG_HIP(hipEventCreate(&L.t1));
    // every lane's arena can hold the whole budget: the budget, not the arena, admits
    const uint64_t groups = std::max<uint64_t>(hbm_budget / 64 / 4, 1 << 14);
    WG_TRY(wg_batch_create(L.ctx, max_batch, groups, &L.batch));
    // the runner's tuning for request batches of window scans (1024-thread
    // workgroups, LDS tables up to 2048 groups and as a cache of frequent keys
    // above that, 128K-row work items, up to 16 key partitions): the fastest
    // of the candidates on GROUP BY port, GROUP BY address (Zipf keys) and
    // join-then-GROUP BY batches over the flow table
    WG_TRY(wg_batch_set_tuning(L.batch, 1024, 2048, 131072, 16, 1));
  }
  s->engine = std::thread(engine_loop, s);
  *out = s;
  return WG_OK;
}

wg_status wg_serve_destroy(wg_serve* s) {
  if (!s) return WG_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_arrive.notify_all();
  if (s->engine.joinable()) s->engine.join();
  for (Lane& L : s->lanes) {
    (void)wg_batch_destroy(L.batch);
    for (const Download& d : L.downloads) (void)hipEventDestroy(d.ev);
    for (hipEvent_t e : L.free_events) (void)hipEventDestroy(e);
    for (uint64_t* h : L.host)
      if (h) (void)hipHostFree(h);
    if (L.t0) (void)hipEventDestroy(L.t0);
    if (L.t1) (void)hipEventDestroy(L.t1);
    (void)wg_ctx_destroy(L.ctx);
  }
  for (auto& kv : s->reqs) delete kv.second;
  for (Request* r : s->queue) (void)r;
  if (s->origin) (void)hipEventDestroy(s->origin);
  delete s;
  return WG_OK;
}

wg_status wg_serve_add_dim(wg_serve* s, const wg_join* dim, int32_t* dim_id) {
  clear_error();
  WG_CHECK(s && dim && dim_id, WG_ERR_INVALID, "null argument");
  WG_CHECK(dim->finished && dim->claimed_unique, WG_ERR_UNSUPPORTED, "dimension must be built with unique keys");
  std::lock_guard<std::mutex> lk(s->mu);
  s->dims.push_back(dim);
  *dim_id = (int32_t)s->dims.size() - 1;
  return WG_OK;
}

wg_status wg_serve_submit(wg_serve* s, const wg_serve_query* q, uint64_t* id) {
  clear_error();
  WG_CHECK(s && q && id, WG_ERR_INVALID, "null argument");
  Request* r = new Request();
  r->arrival = Clock::now();
  r->q = *q;
  const wg_status st = to_batch_query(s, *q, r->bq);
  if (st != WG_OK) {
    delete r;
    return st;
  }
  r->footprint = footprint_of(r->bq.est_groups);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    WG_CHECK(!s->stop, WG_ERR_STATE, "server stopped");
    r->id = s->next_id++;
    if (!s->any) {
      s->first_arrival = r->arrival;
      s->any = true;
    }
    s->reqs[r->id] = r;
    s->queue.push_back(r);
    *id = r->id;
  }
  s->cv_arrive.notify_one();
  return WG_OK;
}

wg_status wg_serve_wait(wg_serve* s, uint64_t id, uint64_t* n_groups, double* latency_ms) {
  clear_error();
  WG_CHECK(s, WG_ERR_INVALID, "null handle");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  WG_CHECK(it != s->reqs.end(), WG_ERR_INVALID, "unknown request %llu", (unsigned long long)id);
  Request* r = it->second;
  s->cv_done.wait(lk, [&] { return r->done; });
  if (r->status != WG_OK) {
    set_error("request %llu failed with status %d", (unsigned long long)id, r->status);
    return r->status;
  }
  if (n_groups) *n_groups = r->n_groups;
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->done_at - r->arrival).count();
  return WG_OK;
}

wg_status wg_serve_result(wg_serve* s, uint64_t id, int64_t* keys, uint64_t* sum_lo, int64_t* sum_hi, int64_t* counts,
                          uint64_t cap) {
  clear_error();
  WG_CHECK(s, WG_ERR_INVALID, "null handle");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  WG_CHECK(it != s->reqs.end() && it->second->done, WG_ERR_STATE, "request %llu not completed", (unsigned long long)id);
  Request* r = it->second;
  const uint64_t n = r->n_groups;
  WG_CHECK(cap >= n, WG_ERR_INVALID, "result has %llu groups, capacity %llu", (unsigned long long)n, (unsigned long long)cap);
  const uint64_t* q = r->res.data();
  if (keys) memcpy(keys, q, n * 8);
  if (sum_lo) memcpy(sum_lo, q + n, n * 8);
  if (sum_hi) memcpy(sum_hi, q + 2 * n, n * 8);
  if (counts) memcpy(counts, q + 3 * n, n * 8);
  return WG_OK;
}

wg_status wg_serve_replay(wg_serve* s, const wg_serve_query* qs, int32_t n, double qps, uint64_t seed,
                          double* latency_ms, uint64_t* ids) {
  clear_error();
  WG_CHECK(s && qs && n > 0 && qps > 0, WG_ERR_INVALID, "bad argument");
  std::vector<uint64_t> id(n, 0);
  std::vector<wg_status> sub(n, WG_OK);
  // open loop: arrivals follow their own Poisson clock whatever the server does
  std::thread arrivals([&] {
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(qps);
    auto t = Clock::now();
    for (int i = 0; i < n; ++i) {
      t += std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(gap(rng)));
      std::this_thread::sleep_until(t);
      sub[i] = wg_serve_submit(s, &qs[i], &id[i]);
    }
  });
  arrivals.join();
  for (int i = 0; i < n; ++i) {
    WG_CHECK(sub[i] == WG_OK, sub[i], "submit %d failed", i);
    double ms = 0;
    WG_TRY(wg_serve_wait(s, id[i], nullptr, &ms));
    if (latency_ms) latency_ms[i] = ms;
    if (ids) ids[i] = id[i];
  }
  return WG_OK;
}

wg_status wg_serve_stats_get(wg_serve* s, wg_serve_stats* out) {
  clear_error();
  WG_CHECK(s && out, WG_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  std::vector<double> lat;
  for (auto& kv : s->reqs)
    if (kv.second->done) lat.push_back(std::chrono::duration<double, std::milli>(kv.second->done_at - kv.second->arrival).count());
  std::sort(lat.begin(), lat.end());
  *out = wg_serve_stats{};
  out->completed = lat.size();
  out->batches = s->batches;
  out->max_in_flight = s->max_in_flight;
  out->peak_footprint = s->peak_fp;
  if (!lat.empty()) {
    auto pct = [&](double p) { return lat[std::min<size_t>(lat.size() - 1, (size_t)(p * (lat.size() - 1) + 0.5))]; };
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
    out->max_ms = lat.back();
    double sum = 0;
    for (double v : lat) sum += v;
    out->mean_ms = sum / lat.size();
    out->wall_s = std::chrono::duration<double>(s->last_done - s->first_arrival).count();
    out->qps = out->wall_s > 0 ? lat.size() / out->wall_s : 0;
  }
  // device intervals: their union (the device busy with some batch) and their sum
  std::vector<std::pair<double, double>> iv = s->busy;
  std::sort(iv.begin(), iv.end());
  double uni = 0, sum = 0, cur_a = 0, cur_b = -1;
  for (const auto& [a, b] : iv) {
    sum += b - a;
    if (a > cur_b) {
      if (cur_b > cur_a) uni += cur_b - cur_a;
      cur_a = a;
      cur_b = b;
    } else {
      cur_b = std::max(cur_b, b);
    }
  }
  if (cur_b > cur_a) uni += cur_b - cur_a;
  out->device_busy_s = uni / 1000;
  out->lane_busy_s = sum / 1000;
  out->host_submit_s = s->host_submit_s;
  out->host_retire_s = s->host_retire_s;
  out->host_deliver_s = s->host_deliver_s;
  return WG_OK;
}

wg_status wg_serve_reset(wg_serve* s) {
  clear_error();
  WG_CHECK(s, WG_ERR_INVALID, "null handle");
  std::unique_lock<std::mutex> lk(s->mu);
  // wait for in-flight work to drain
  s->cv_done.wait(lk, [&] {
    if (!s->queue.empty()) return false;
    for (auto& kv : s->reqs)
      if (!kv.second->done) return false;
    return true;
  });
  for (auto& kv : s->reqs) {
    give_buffer(s, std::move(kv.second->res));
    delete kv.second;
  }
  s->reqs.clear();
  s->batches = 0;
  s->max_in_flight = 0;
  s->peak_fp = 0;
  s->any = false;
  s->busy.clear();
  s->host_submit_s = s->host_retire_s = s->host_deliver_s = 0;
  return WG_OK;
}

}  // extern "C"
