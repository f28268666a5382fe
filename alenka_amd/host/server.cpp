// Served form of the operators: a request loop over device-resident tables with continuous
// batching.  Requests are filter + GROUP BY queries (aln_batch_query) against a registered
// table, or star pipelines (aln_star_args).  Each of `streams` worker slots owns an aln_ctx
// on its own HIP stream and a device region; a slot that comes free takes the oldest waiting
// request and, in batched mode, every other waiting query on the same table that fits the
// batch and the memory budget, and answers them with one shared scan - or, for filter +
// GROUP BY requests, runs the oldest alone through the operators when that has been
// measured to take less time per request on the table (a single request always does).
// Requests that arrive while slots are busy are admitted as slots and budget free up, so
// batches overlap on the device.  In plain mode every query runs alone through the
// per-statement operators
// (aln_scan_filter -> aln_eval -> aln_groupby), the way the reference runs a script
// statement by statement (emit_filter operators.cu:2306, emit_select 1801-2084).
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../csrc/aln_internal.h"

using namespace aln;
using Clock = std::chrono::steady_clock;

namespace {

struct Table {
  std::vector<aln_col> cols;
  uint32_t seg_rows = 0;
  uint64_t nrows = 0;
  std::vector<int64_t> cmin, cmax;  // per column, from the zone maps (or a device scan)
  std::vector<std::vector<aln_seg>> hsegs;  // per column, host copies of the segment headers
  std::vector<const aln_seg *> hsp;         // their data pointers
  // milliseconds per filter + GROUP BY request on this table, run alone through the
  // operators and inside a shared-scan batch (running means under the server lock; 0 = not
  // measured yet), and the batch decisions taken since batching was last tried
  double plain_ms = 0, batch_ms = 0, batch1_ms = 0;  // batch1: a shared-scan run of one request
  uint32_t since_batch = 0;
};

struct Request {
  uint64_t id = 0;
  int kind = 0;  // 0 filter + GROUP BY, 1 star pipeline
  int table = -1;
  aln_batch_query q{};
  uint64_t groups = 0;  // dense group range of the query
  aln_star_args star{};
  std::vector<aln_col> scols[3];
  std::vector<aln_range> sranges[3];
  std::vector<char> smodes[3];
  size_t mem = 0;  // admission estimate (device bytes held while running)
  Clock::time_point t_submit, t_start, t_done;
  bool done = false;
  int status = ALN_OK;
  std::string error;
  std::vector<int64_t> res[3];  // keys, sums, counts
  std::vector<aln_star_row> top;
  aln_star_stats stats{};
  uint64_t batch = 0;
  int batch_size = 0;
};

struct Slot {
  aln_ctx *ctx = nullptr;
  hipStream_t stream = nullptr;
  char *region = nullptr;
  size_t region_size = 0;
  std::thread th;
};

}  // namespace

struct aln_server {
  int device = 0;
  aln_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request *> pending;
  std::unordered_map<uint64_t, Request *> reqs;
  std::deque<Table> tables;  // stable addresses: workers read entries without the lock
  std::vector<Slot *> slots;
  aln_ctx *admin = nullptr;  // table registration, on its own stream, under the lock
  hipStream_t admin_stream = nullptr;
  uint64_t next_id = 1, next_batch = 1;
  size_t mem_in_use = 0;
  int running = 0;
  bool stop = false;
  aln_server_totals st{};
  // star requests by fact table (first column's words, segments): milliseconds per request
  // run alone and inside shared-scan batches (running means under the lock; 0 = not measured
  // yet), and the batching decisions taken since batching was last tried
  struct StarTimes {
    double one_ms = 0, batch_ms = 0;
    uint32_t since_batch = 0;
  };
  std::map<std::pair<const void *, uint32_t>, StarTimes> star_times;
};

namespace {

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Device bytes of a filter + GROUP BY run through the operators: row ids, two evaluated
// columns, group-by workspace and outputs.
size_t plain_mem(const Table &t, uint64_t g) { return (size_t)(4 + 8 + 8 + 48) * t.nrows + 24 * g + (16u << 20); }

// Device bytes a request holds while it runs (what admission charges against the budget):
// plain, the operators' bytes; batched, dense group tables + compaction + outputs (a
// batched-mode request run through the operators is charged their bytes when it starts).
size_t request_mem(const aln_server *s, const Request &r) {
  if (r.kind == 1) {
    // bitmaps over both key ranges, per-block counts, group tables and candidates
    const aln_star_args &a = r.star;
    uint64_t mrows = 0, drows = 0;
    for (const aln_col &c : r.scols[1]) mrows = std::max<uint64_t>(mrows, (uint64_t)c.nsegs * a.mid.seg_rows);
    for (const aln_col &c : r.scols[0]) drows = std::max<uint64_t>(drows, (uint64_t)c.nsegs * a.dim.seg_rows);
    return (size_t)(48 * mrows + 2 * drows) + (64u << 20);
  }
  const Table &t = s->tables[r.table];
  const uint64_t g = r.groups ? r.groups : 1;
  return s->cfg.plain ? plain_mem(t, g) : align256(12 * g) + align256(36 * g) + 4096;
}

// Batched mode, filter + GROUP BY: whether the requests taken together run through one
// shared scan (true) or each alone through the operators (caller holds the lock).
// The requests run batched unless the measured per-request time of such runs on this
// table (batches, or single requests through the batch kernel, measured apart) is not
// below that of plain runs, and then every 8th time they still run batched so the
// estimate follows the load.  (A single request gains no shared scan, but the batch
// kernel aggregates without the operators' row ids and evaluated columns.)
constexpr uint32_t BATCH_RETRY = 8;

bool use_batch(aln_server *s, const std::vector<Request *> &rs) {
  Table &t = s->tables[rs[0]->table];
  const uint64_t g = rs[0]->groups ? rs[0]->groups : 1;
  if (s->mem_in_use - rs[0]->mem + plain_mem(t, g) > s->cfg.budget_bytes) return true;  // no room
  const double bm = rs.size() < 2 ? t.batch1_ms : t.batch_ms;
  if (bm > 0 && t.plain_ms <= 0) return false;  // measure the plain form too
  if (bm <= 0 || bm < t.plain_ms) {
    t.since_batch = 0;
    return true;
  }
  if (++t.since_batch >= BATCH_RETRY) {
    t.since_batch = 0;
    return true;
  }
  return false;
}

// Star requests taken together: one shared-scan batch (aln_star_agg_batch) or the oldest
// alone (caller holds the lock) - the same measured choice as for filter + GROUP BY, single
// runs measured first (the single pipeline avoids the batch's per-query bitmaps and tables,
// and was faster per request on every table size measured).
aln_server::StarTimes &star_times_of(aln_server *s, const Request &r) {
  if (r.scols[2].empty()) return s->star_times[{nullptr, 0u}];
  const aln_col &c = r.scols[2][0];
  return s->star_times[{(const void *)c.words, c.nsegs}];
}

bool use_star_batch(aln_server *s, const std::vector<Request *> &rs) {
  if (rs.size() < 2) return false;
  aln_server::StarTimes &t = star_times_of(s, *rs[0]);
  if (t.one_ms <= 0) return false;  // single runs are measured first
  if (t.batch_ms <= 0 || t.batch_ms < t.one_ms) {
    t.since_batch = 0;
    return true;
  }
  if (++t.since_batch >= BATCH_RETRY) {
    t.since_batch = 0;
    return true;
  }
  return false;
}

void note_time(double *mean, double ms) { *mean = *mean <= 0 ? ms : 0.8 * *mean + 0.2 * ms; }

void finish(aln_server *s, Request *r, int status, const char *err) {
  r->status = status;
  if (status != ALN_OK && err) r->error = err;
  r->t_done = Clock::now();
  r->done = true;
}

int ensure_region(Slot *sl, size_t need) {
  if (sl->region_size >= need) return ALN_OK;
  if (sl->region) {
    (void)hipStreamSynchronize(sl->stream);
    (void)hipFree(sl->region);
    sl->region = nullptr;
    sl->region_size = 0;
  }
  const size_t sz = std::max(need, (size_t)64 << 20);
  if (hipMalloc((void **)&sl->region, sz) != hipSuccess) {
    (void)hipGetLastError();
    sl->region = nullptr;
    return set_error(ALN_E_NOMEM, "aln_server: device region of %zu bytes", sz);
  }
  sl->region_size = sz;
  return ALN_OK;
}

// One filter + GROUP BY through the per-statement operators.  The key program groups by
// floor(key / div): keys are first shifted by a multiple of div at or below the column
// minimum, so the truncating division of the operators equals the floor.
int run_plain(aln_server *s, Slot *sl, Request *r) {
  const Table &t = s->tables[r->table];
  const aln_batch_query &q = r->q;
  const int64_t div = q.key_div;
  const int64_t kmin = t.cmin[q.key_col];
  const int64_t base = floor_div(kmin, div) * div;
  const size_t n = t.nrows;
  ALN_TRY(ensure_region(sl, align256(4 * n) + 2 * align256(8 * n) + 3 * align256(8 * (r->groups + 1))));
  char *p = sl->region;
  uint32_t *rows = (uint32_t *)p;
  p += align256(4 * n);
  int64_t *key = (int64_t *)p;
  p += align256(8 * n);
  int64_t *rev = (int64_t *)p;
  p += align256(8 * n);
  const uint64_t cap = r->groups + 1;
  int64_t *ok = (int64_t *)p;
  p += align256(8 * cap);
  int64_t *os = (int64_t *)p;
  p += align256(8 * cap);
  int64_t *oc = (int64_t *)p;
  const int32_t pc = q.pred_col;
  const aln_insn pred[7] = {{ALN_I_COL, pc, 1}, {ALN_I_CONST, 0, q.lo}, {ALN_I_CMP, 6, 0}, {ALN_I_COL, pc, 1},
                            {ALN_I_CONST, 0, q.hi},  {ALN_I_CMP, 5, 0},   {ALN_I_AND, 0, 0}};
  const aln_insn kp[5] = {{ALN_I_COL, q.key_col, 1}, {ALN_I_CONST, 0, base}, {ALN_I_SUB, 0, 0},
                          {ALN_I_CONST, 0, div}, {ALN_I_DIV, 0, 0}};
  const aln_insn vp[7] = {{ALN_I_CONST, 0, q.a0}, {ALN_I_COL, q.x_col, q.a1}, {ALN_I_ADD, 0, 0},
                          {ALN_I_CONST, 0, q.b0}, {ALN_I_COL, q.y_col, q.b1}, {ALN_I_ADD, 0, 0},
                          {ALN_I_MUL, 0, 0}};
  const int nc = (int)t.cols.size();
  uint64_t cnt = 0;
  ALN_TRY(aln_scan_filter(sl->ctx, t.cols.data(), nc, t.seg_rows, pred, 7, nullptr, rows, n, &cnt));
  ALN_TRY(aln_eval(sl->ctx, t.cols.data(), nc, t.seg_rows, kp, 5, rows, cnt, key));
  ALN_TRY(aln_eval(sl->ctx, t.cols.data(), nc, t.seg_rows, vp, 7, rows, cnt, rev));
  const int64_t *keys[1] = {key};
  const int64_t *vals[2] = {rev, rev};
  const int32_t aggs[2] = {ALN_AGG_SUM, ALN_AGG_COUNT};
  int64_t *okeys[1] = {ok};
  int64_t *ovals[2] = {os, oc};
  uint64_t g = 0;
  ALN_TRY(aln_groupby(sl->ctx, keys, 1, vals, aggs, 2, cnt, ALN_ORDER_KEYS, 100, okeys, ovals, cap, &g));
  for (auto &v : r->res) v.resize(g);
  if (g) {
    ALN_TRY(copy_to_host(sl->ctx, r->res[0].data(), ok, 8 * g));
    ALN_TRY(copy_to_host(sl->ctx, r->res[1].data(), os, 8 * g));
    ALN_TRY(copy_to_host(sl->ctx, r->res[2].data(), oc, 8 * g));
  }
  const int64_t shift = base / div;
  for (auto &k : r->res[0]) k += shift;
  return ALN_OK;
}

// The queries of one batch (same table) through one shared scan.
int run_batched(aln_server *s, Slot *sl, std::vector<Request *> &rs) {
  const Table &t = s->tables[rs[0]->table];
  std::vector<aln_batch_query> qs;
  std::vector<uint64_t> groups;
  for (Request *r : rs) {
    qs.push_back(r->q);
    groups.push_back(r->groups);
  }
  const size_t need = batch_pool_bytes(groups.data(), (int)qs.size());
  ALN_TRY(ensure_region(sl, need));
  aln_batch *b = nullptr;
  ALN_TRY(batch_create(sl->ctx, t.cols.data(), (int)t.cols.size(), t.seg_rows, qs.data(), (int)qs.size(),
                       &s->cfg.tune, sl->region, sl->region_size, &b, t.hsp.data(), t.cmin.data(), t.cmax.data()));
  int rc = aln_batch_run(b);
  for (size_t i = 0; rc == ALN_OK && i < rs.size(); ++i) rc = batch_fetch(b, (int)i, rs[i]->res);
  aln_batch_destroy(b);
  return rc;
}

int run_star(Slot *sl, Request *r) {
  r->top.resize(std::max<uint32_t>(1, r->star.limit));
  uint32_t ntop = 0;
  ALN_TRY(aln_star_agg(sl->ctx, &r->star, r->top.data(), &ntop, nullptr, 0, &r->stats));
  r->top.resize(ntop);
  return ALN_OK;
}

// Star requests over the same tables share one scan of each table; when the batched form
// does not apply (hash join needed, duplicate keys) they run one by one.
int run_star_batch(Slot *sl, std::vector<Request *> &rs) {
  if (rs.size() == 1) return run_star(sl, rs[0]);
  std::vector<aln_star_args> args;
  std::vector<aln_star_row *> tops;
  std::vector<uint32_t> ntop(rs.size());
  std::vector<aln_star_stats> st(rs.size());
  for (Request *r : rs) {
    r->top.resize(std::max<uint32_t>(1, r->star.limit));
    args.push_back(r->star);
    tops.push_back(r->top.data());
  }
  const int rc = aln_star_agg_batch(sl->ctx, args.data(), (int)rs.size(), tops.data(), ntop.data(), st.data());
  if (rc == ALN_E_UNSUPPORTED) {
    for (Request *r : rs) ALN_TRY(run_star(sl, r));
    return ALN_OK;
  }
  ALN_TRY(rc);
  for (size_t i = 0; i < rs.size(); ++i) {
    rs[i]->top.resize(ntop[i]);
    rs[i]->stats = st[i];
  }
  return ALN_OK;
}

bool same_scan(const aln_table_scan &x, const aln_table_scan &y) {
  if (x.ncols != y.ncols || x.seg_rows != y.seg_rows || x.seg_modes || y.seg_modes) return false;
  for (int i = 0; i < x.ncols; ++i)
    if (x.cols[i].words != y.cols[i].words || x.cols[i].segs != y.cols[i].segs || x.cols[i].nsegs != y.cols[i].nsegs)
      return false;
  return true;
}

// Two star requests can share scans: same tables and column slots, no group output.
bool star_compatible(const Request &a, const Request &b) {
  const aln_star_args &x = a.star, &y = b.star;
  return same_scan(x.dim, y.dim) && same_scan(x.mid, y.mid) && same_scan(x.fact, y.fact) && x.d_key == y.d_key &&
         x.m_key == y.m_key && x.m_fk == y.m_fk && x.m_g1 == y.m_g1 && x.m_g2 == y.m_g2 && x.f_fk == y.f_fk &&
         x.f_x == y.f_x && x.f_y == y.f_y && !(x.flags & 3) && !(y.flags & 3);
}

// Takes the next batch off the queue (caller holds the lock).  Returns an empty vector
// when the oldest request must wait for budget.
std::vector<Request *> take_batch(aln_server *s) {
  std::vector<Request *> out;
  while (!s->pending.empty()) {
    Request *h = s->pending.front();
    if (h->mem > s->cfg.budget_bytes) {  // can never be admitted
      s->pending.pop_front();
      char msg[160];
      snprintf(msg, sizeof(msg), "aln_server: request %llu needs %zu bytes, budget %llu",
               (unsigned long long)h->id, h->mem, (unsigned long long)s->cfg.budget_bytes);
      finish(s, h, ALN_E_NOMEM, msg);
      s->cv_done.notify_all();
      continue;
    }
    if (s->mem_in_use + h->mem > s->cfg.budget_bytes) return out;
    break;
  }
  if (s->pending.empty()) return out;
  Request *h = s->pending.front();
  s->pending.pop_front();
  out.push_back(h);
  size_t mem = h->mem;
  const int cap = h->kind == 0 ? (int)ALN_BATCH_MAX : (int)ALN_STAR_BATCH_MAX;
  const int maxb = s->cfg.plain ? 1 : std::max(1, std::min(s->cfg.max_batch, cap));
  if (maxb > 1) {
    for (auto it = s->pending.begin(); it != s->pending.end() && (int)out.size() < maxb;) {
      Request *r = *it;
      const bool fits = r->kind == h->kind &&
                        (h->kind == 0 ? r->table == h->table : star_compatible(*h, *r));
      if (fits && s->mem_in_use + mem + r->mem <= s->cfg.budget_bytes) {
        mem += r->mem;
        out.push_back(r);
        it = s->pending.erase(it);
      } else {
        ++it;
      }
    }
  }
  s->mem_in_use += mem;
  return out;
}

void slot_main(aln_server *s, Slot *sl) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_work.wait(lk, [&] {
      return (s->stop && s->pending.empty()) ||
             (!s->pending.empty() &&
              (s->pending.front()->mem > s->cfg.budget_bytes ||
               s->mem_in_use + s->pending.front()->mem <= s->cfg.budget_bytes));
    });
    if (s->pending.empty()) {
      if (s->stop) break;
      continue;
    }
    std::vector<Request *> rs = take_batch(s);
    if (rs.empty()) continue;
    const bool batched = !s->cfg.plain && (rs[0]->kind == 0 ? use_batch(s, rs) : use_star_batch(s, rs));
    if (rs[0]->kind == 1 && rs.size() > 1 && !batched) {
      // the oldest alone; the others back to the front of the queue, in order
      for (size_t i = rs.size(); i-- > 1;) {
        s->mem_in_use -= rs[i]->mem;
        s->pending.push_front(rs[i]);
      }
      rs.resize(1);
      s->cv_work.notify_all();
    }
    if (rs[0]->kind == 0 && !s->cfg.plain && !batched) {
      // run the oldest alone through the operators, charged their bytes; the others go back
      // to the front of the queue, in order, for the next free slot
      for (size_t i = rs.size(); i-- > 1;) {
        s->mem_in_use -= rs[i]->mem;
        s->pending.push_front(rs[i]);
      }
      rs.resize(1);
      const size_t pm = plain_mem(s->tables[rs[0]->table], rs[0]->groups ? rs[0]->groups : 1);
      s->mem_in_use += pm - rs[0]->mem;
      rs[0]->mem = pm;
      s->cv_work.notify_all();
    }
    const uint64_t bid = s->next_batch++;
    ++s->running;
    s->st.max_concurrent = std::max<uint64_t>(s->st.max_concurrent, (uint64_t)s->running);
    const Clock::time_point t0 = Clock::now();
    for (Request *r : rs) {
      r->t_start = t0;
      r->batch = bid;
      r->batch_size = (int)rs.size();
    }
    lk.unlock();
    int rc = ALN_OK;
    if (rs[0]->kind == 1)
      rc = run_star_batch(sl, rs);
    else if (batched)
      rc = run_batched(s, sl, rs);
    else
      for (size_t i = 0; rc == ALN_OK && i < rs.size(); ++i) rc = run_plain(s, sl, rs[i]);
    const char *err = rc == ALN_OK ? nullptr : aln_last_error();
    const double per = ms_between(t0, Clock::now()) / (double)rs.size();
    lk.lock();
    if (rs[0]->kind == 0 && rc == ALN_OK) {
      Table &tb = s->tables[rs[0]->table];
      note_time(!batched ? &tb.plain_ms : rs.size() < 2 ? &tb.batch1_ms : &tb.batch_ms, per);
    } else if (rc == ALN_OK && !s->cfg.plain) {
      aln_server::StarTimes &tt = star_times_of(s, *rs[0]);
      note_time(batched ? &tt.batch_ms : &tt.one_ms, per);
    }
    --s->running;
    size_t mem = 0;
    for (Request *r : rs) {
      mem += r->mem;
      finish(s, r, rc, err);
    }
    s->mem_in_use -= mem;
    s->st.batches += 1;
    s->st.completed += rs.size();
    if (rs.size() > 1 && (batched || rs[0]->kind == 1)) s->st.batched_queries += rs.size();
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
}

int add_request(aln_server *s, Request *r, uint64_t *h_id) {
  std::lock_guard<std::mutex> g(s->mu);
  if (s->stop) {
    delete r;
    return set_error(ALN_E_ARG, "aln_server: stopped");
  }
  r->id = s->next_id++;
  r->mem = request_mem(s, *r);
  r->t_submit = Clock::now();
  s->reqs[r->id] = r;
  s->pending.push_back(r);
  s->st.submitted += 1;
  *h_id = r->id;
  s->cv_work.notify_one();
  return ALN_OK;
}

Request *find_done(aln_server *s, uint64_t id, std::unique_lock<std::mutex> &lk, int *rc) {
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) {
    *rc = set_error(ALN_E_ARG, "aln_server: no request %llu", (unsigned long long)id);
    return nullptr;
  }
  Request *r = it->second;
  s->cv_done.wait(lk, [&] { return r->done; });
  *rc = ALN_OK;
  return r;
}

}  // namespace

extern "C" {

int aln_server_create(int device, const aln_server_config *cfg, aln_server **out) {
  ALN_REQUIRE(cfg && out, "aln_server_create: bad arguments");
  ALN_REQUIRE(cfg->streams >= 1 && cfg->streams <= 8, "aln_server_create: %d streams (1..8)", cfg->streams);
  ALN_REQUIRE(cfg->budget_bytes > 0, "aln_server_create: zero memory budget");
  int ndev = 0;
  ALN_HIP(hipGetDeviceCount(&ndev));
  ALN_REQUIRE(device >= 0 && device < ndev, "aln_server_create: device %d of %d", device, ndev);
  ALN_HIP(hipSetDevice(device));
  aln_server *s = new aln_server();
  s->device = device;
  s->cfg = *cfg;
  if (s->cfg.max_batch <= 0) s->cfg.max_batch = ALN_BATCH_MAX;
  // one persistent batch block per CU by default: a block's LDS (80 KB) then leaves room on
  // each CU for the other workers' kernels (burst of the served mix: 785 against 603 q/s with
  // two blocks per CU, which fill the LDS)
  if (s->cfg.tune.blocks_per_cu <= 0) s->cfg.tune.blocks_per_cu = 1;
  if (hipStreamCreateWithFlags(&s->admin_stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return set_error(ALN_E_HIP, "aln_server_create: stream creation failed");
  }
  {
    const int rc = aln_create(device, s->admin_stream, &s->admin);
    if (rc != ALN_OK) {
      aln_server_destroy(s);
      return rc;
    }
  }
  for (int i = 0; i < cfg->streams; ++i) {
    Slot *sl = new Slot();
    int rc = ALN_OK;
    if (hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking) != hipSuccess)
      rc = set_error(ALN_E_HIP, "aln_server_create: stream creation failed");
    if (rc == ALN_OK) rc = aln_create(device, sl->stream, &sl->ctx);
    if (rc != ALN_OK) {
      if (sl->stream) (void)hipStreamDestroy(sl->stream);
      delete sl;
      aln_server_destroy(s);
      return rc;
    }
    s->slots.push_back(sl);
  }
  for (Slot *sl : s->slots) sl->th = std::thread(slot_main, s, sl);
  *out = s;
  return ALN_OK;
}

int aln_server_add_table(aln_server *s, const aln_col *cols, int ncols, uint32_t seg_rows, int32_t *h_table) {
  ALN_REQUIRE(s && cols && h_table && seg_rows > 0, "aln_server_add_table: bad arguments");
  ALN_REQUIRE(ncols >= 1 && ncols <= ALN_MAX_COLS, "aln_server_add_table: %d columns", ncols);
  Table t;
  t.cols.assign(cols, cols + ncols);
  t.seg_rows = seg_rows;
  const uint32_t nsegs = cols[0].nsegs;
  for (int i = 0; i < ncols; ++i)
    ALN_REQUIRE(cols[i].nsegs == nsegs, "aln_server_add_table: columns differ in segments");
  // zone maps give each column's range (a plain 64-bit column is scanned once)
  std::lock_guard<std::mutex> g(s->mu);
  aln_ctx *c = s->admin;
  t.cmin.resize(ncols);
  t.cmax.resize(ncols);
  t.hsegs.resize(ncols);
  for (int i = 0; i < ncols; ++i) {
    std::vector<aln_seg> &hs = t.hsegs[i];
    hs.resize(nsegs);
    if (nsegs) ALN_TRY(copy_to_host(c, hs.data(), cols[i].segs, sizeof(aln_seg) * nsegs));
    int64_t mn = INT64_MAX, mx = INT64_MIN;
    uint64_t rows = 0;
    bool open = false;
    for (uint32_t k = 0; k < nsegs; ++k) {
      if (!hs[k].nrows) continue;
      mn = std::min(mn, hs[k].vmin);
      mx = std::max(mx, hs[k].vmax);
      open = open || (hs[k].bits == 64 && hs[k].vmin == INT64_MIN && hs[k].vmax == INT64_MAX);
      rows += hs[k].nrows;
    }
    if (open) ALN_TRY(minmax_i64(c, (const int64_t *)cols[i].words, rows, &mn, &mx));
    t.cmin[i] = mn;
    t.cmax[i] = mx;
    if (i == 0) t.nrows = rows;
  }
  ALN_REQUIRE(t.nrows < (1ull << 32), "aln_server_add_table: %llu rows", (unsigned long long)t.nrows);
  for (auto &hs : t.hsegs) t.hsp.push_back(hs.data());
  s->tables.push_back(std::move(t));
  *h_table = (int32_t)s->tables.size() - 1;
  return ALN_OK;
}

int aln_server_submit(aln_server *s, int32_t table, const aln_batch_query *q, uint64_t *h_id) {
  ALN_REQUIRE(s && q && h_id, "aln_server_submit: bad arguments");
  int64_t mn, mx;
  {
    std::lock_guard<std::mutex> g(s->mu);
    ALN_REQUIRE(table >= 0 && table < (int32_t)s->tables.size(), "aln_server_submit: table %d", table);
    const Table &t = s->tables[table];
    const int nc = (int)t.cols.size();
    ALN_REQUIRE(q->pred_col >= 0 && q->pred_col < nc && q->key_col >= 0 && q->key_col < nc && q->x_col >= 0 &&
                    q->x_col < nc && q->y_col >= 0 && q->y_col < nc,
                "aln_server_submit: column slot out of range");
    ALN_REQUIRE(q->key_div >= 1, "aln_server_submit: key_div %lld", (long long)q->key_div);
    mn = t.cmin[q->key_col];
    mx = t.cmax[q->key_col];
  }
  Request *r = new Request();
  r->kind = 0;
  r->table = table;
  r->q = *q;
  r->groups = dense_groups(mn, mx, q->key_div);
  if (r->groups > (1ull << 31)) {
    delete r;
    return set_error(ALN_E_UNSUPPORTED, "aln_server_submit: more than 2^31 dense groups");
  }
  return add_request(s, r, h_id);
}

int aln_server_submit_star(aln_server *s, const aln_star_args *a, uint64_t *h_id) {
  ALN_REQUIRE(s && a && h_id, "aln_server_submit_star: bad arguments");
  Request *r = new Request();
  r->kind = 1;
  r->star = *a;
  r->star.flags &= ~1u;  // top rows only
  const aln_table_scan *src[3] = {&a->dim, &a->mid, &a->fact};
  aln_table_scan *dst[3] = {&r->star.dim, &r->star.mid, &r->star.fact};
  for (int i = 0; i < 3; ++i) {
    const aln_table_scan &t = *src[i];
    if (!t.cols || t.ncols < 1 || t.ncols > ALN_MAX_COLS || t.nranges < 0 || t.nranges > ALN_MAX_RANGES ||
        (t.nranges && !t.ranges)) {
      delete r;
      return set_error(ALN_E_ARG, "aln_server_submit_star: table scan %d", i);
    }
    r->scols[i].assign(t.cols, t.cols + t.ncols);
    r->sranges[i].assign(t.ranges, t.ranges + t.nranges);
    if (t.seg_modes) r->smodes[i].assign(t.seg_modes, t.seg_modes + t.cols[0].nsegs);
    dst[i]->cols = r->scols[i].data();
    dst[i]->ranges = r->sranges[i].empty() ? nullptr : r->sranges[i].data();
    dst[i]->seg_modes = t.seg_modes ? r->smodes[i].data() : nullptr;
  }
  return add_request(s, r, h_id);
}

int aln_server_wait(aln_server *s, uint64_t id, aln_server_result *h_res) {
  ALN_REQUIRE(s && h_res, "aln_server_wait: bad arguments");
  std::unique_lock<std::mutex> lk(s->mu);
  int rc;
  Request *r = find_done(s, id, lk, &rc);
  if (!r) return rc;
  memset(h_res, 0, sizeof(*h_res));
  h_res->status = r->status;
  h_res->kind = r->kind;
  h_res->batch = r->batch;
  h_res->batch_size = r->batch_size;
  h_res->ngroups = r->kind == 0 ? r->res[0].size() : r->top.size();
  const bool started = r->batch != 0;  // requests failed at admission never started
  h_res->queue_ms = ms_between(r->t_submit, started ? r->t_start : r->t_done);
  h_res->run_ms = started ? ms_between(r->t_start, r->t_done) : 0.0;
  h_res->latency_ms = ms_between(r->t_submit, r->t_done);
  if (r->status != ALN_OK) return set_error(r->status, "%s", r->error.c_str());
  return ALN_OK;
}

int aln_server_fetch(aln_server *s, uint64_t id, int64_t *h_keys, int64_t *h_sums, int64_t *h_counts,
                     uint64_t cap) {
  ALN_REQUIRE(s, "aln_server_fetch: bad arguments");
  std::unique_lock<std::mutex> lk(s->mu);
  int rc;
  Request *r = find_done(s, id, lk, &rc);
  if (!r) return rc;
  ALN_REQUIRE(r->kind == 0, "aln_server_fetch: request %llu is a star pipeline", (unsigned long long)id);
  if (r->status != ALN_OK) return set_error(r->status, "%s", r->error.c_str());
  const uint64_t g = r->res[0].size();
  if (g > cap)
    return set_error(ALN_E_CAPACITY, "aln_server_fetch: %llu groups, capacity %llu", (unsigned long long)g,
                     (unsigned long long)cap);
  ALN_REQUIRE(g == 0 || (h_keys && h_sums && h_counts), "aln_server_fetch: no output arrays");
  if (g) {
    memcpy(h_keys, r->res[0].data(), 8 * g);
    memcpy(h_sums, r->res[1].data(), 8 * g);
    memcpy(h_counts, r->res[2].data(), 8 * g);
  }
  return ALN_OK;
}

int aln_server_fetch_star(aln_server *s, uint64_t id, aln_star_row *h_top, uint32_t cap, uint32_t *h_ntop) {
  ALN_REQUIRE(s && h_ntop, "aln_server_fetch_star: bad arguments");
  std::unique_lock<std::mutex> lk(s->mu);
  int rc;
  Request *r = find_done(s, id, lk, &rc);
  if (!r) return rc;
  ALN_REQUIRE(r->kind == 1, "aln_server_fetch_star: request %llu is a query", (unsigned long long)id);
  if (r->status != ALN_OK) return set_error(r->status, "%s", r->error.c_str());
  const uint32_t n = (uint32_t)r->top.size();
  *h_ntop = n;
  if (n > cap) return set_error(ALN_E_CAPACITY, "aln_server_fetch_star: %u rows, capacity %u", n, cap);
  if (n) memcpy(h_top, r->top.data(), sizeof(aln_star_row) * n);
  return ALN_OK;
}

int aln_server_release(aln_server *s, uint64_t id) {
  ALN_REQUIRE(s, "aln_server_release: NULL server");
  std::unique_lock<std::mutex> lk(s->mu);
  int rc;
  Request *r = find_done(s, id, lk, &rc);
  if (!r) return rc;
  s->reqs.erase(id);
  delete r;
  return ALN_OK;
}

int aln_server_stats(aln_server *s, aln_server_totals *h) {
  ALN_REQUIRE(s && h, "aln_server_stats: bad arguments");
  std::lock_guard<std::mutex> g(s->mu);
  *h = s->st;
  h->pending = s->pending.size();
  h->mem_in_use = s->mem_in_use;
  return ALN_OK;
}

void aln_server_destroy(aln_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (Slot *sl : s->slots)
    if (sl->th.joinable()) sl->th.join();
  for (Slot *sl : s->slots) {
    if (sl->ctx) aln_destroy(sl->ctx);
    if (sl->region) (void)hipFree(sl->region);
    if (sl->stream) (void)hipStreamDestroy(sl->stream);
    delete sl;
  }
  if (s->admin) aln_destroy(s->admin);
  if (s->admin_stream) (void)hipStreamDestroy(s->admin_stream);
  for (auto &kv : s->reqs) delete kv.second;
  delete s;
}

}  // extern "C"
