// Serving loop: search requests (sequence, model) admitted as they arrive and
// run as continuous batches over several streams of one device.
//
// The reference has no server: cmsearch/cmscan run each (model, sequence)
// search to completion, splitting a sequence into the (chunk x strand) tasks
// of cm_search.rs:1774-1808 and filtering each with f3_filter_sequence
// (cm_pipeline.rs:1863). A request here is that filter stage for one
// sequence against one resident p7 model, both strands; its result is what
// ixg_window_tasks + ixg_f3_filter_tasks give for the sequence alone
// (windows per task, F1F3Acct per task), bit for bit.
//
// A search request (ixg_server_submit_search) is a whole search of one
// sequence against one resident model bundle (p7, CM, CP9s, band maps) with
// the caller's options (thresholds and E-value scale fixed by the caller, as
// cmsearch -Z fixes them): a batch of such requests for the same bundle and
// options is one ixg_search over a multi-sequence target of their
// sequences, and each request's hits are the batch's hits on its sequence,
// in the same order: what ixg_search gives for the sequence alone.
//
// Scheduling. Requests queue in arrival order. Each worker owns a context
// (its own stream and grow-only scratch) and, when idle, takes the oldest
// queued request together with every later queued request for the same model
// and options, up to max_batch residues, as one batch: their sequences and
// reverse complements go to the device in one copy and run through one
// cascade (one F1 sweep over all their tasks, one F2/F3/F3b launch over all
// surviving windows). Workers run concurrently, so batches of different
// models overlap on the device and one batch's host glue overlaps another's
// kernels. Admission is by residues in flight: the head of the queue waits
// until the residues already on the device plus its own fit max_inflight
// (a request larger than that runs when nothing else is in flight); later
// requests do not overtake it. Each worker context gets mem_budget/nstreams
// bytes of device memory; a batch that exceeds it fails with IXG_ENOMEM.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include "ctx.h"

using namespace ixg;

namespace ixg {
int cascade_views(ixg_ctx *c, const ixg_p7 *p, const uint8_t *d_seq, const int64_t *base, const int64_t *len,
                  int64_t nt, const ixg_f3opts &o, std::vector<ixg_window> &out, std::vector<int64_t> &first,
                  std::vector<ixg_acct> &acct);
}  // namespace ixg

namespace {

using Clock = std::chrono::steady_clock;

struct Req {
  uint64_t ticket = 0;
  int kind = 0;  // 0: filter stage, 1: whole search
  const ixg_p7 *p7 = nullptr;
  ixg_f3opts o{};
  ixg_search_model m{};
  ixg_searchopts so{};
  std::vector<ixg_searchhit> hits;
  int64_t L = 0;
  std::vector<uint8_t> seq;  // the L residue codes
  int state = 0;             // 0 queued, 1 running, 2 done
  int rc = IXG_OK;
  std::string err;
  std::vector<ixg_task> tasks;
  std::vector<ixg_window> wins;  // task index local to the request
  std::vector<ixg_acct> acct;
  int64_t batch_size = 0;
  Clock::time_point t_sub, t_run, t_done;
};

bool codes_valid(const uint8_t *p, int64_t n) {
  uint8_t acc = 0;
  for (int64_t i = 0; i < n; i++) acc |= (uint8_t)(p[i] >= IXG_KP);
  if (acc) g_err = "digitized residue code >= 18 in the sequence";
  return !acc;
}

bool same_search(const Req &a, const Req &b) {  // same bundle and options (compared bytewise)
  return !memcmp(&a.m, &b.m, sizeof(a.m)) && !memcmp(&a.so, &b.so, sizeof(a.so));
}

bool same_opts(const ixg_f3opts &a, const ixg_f3opts &b) {
  return a.f1 == b.f1 && a.f1b == b.f1b && a.f2 == b.f2 && a.f2b == b.f2b && a.f3 == b.f3 && a.f3b == b.f3b &&
         a.cmw == b.cmw && a.maxw == b.maxw && a.opts == b.opts;
}

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

struct ixg_server {
  int device = 0;
  std::vector<ixg_ctx *> ctx;
  std::vector<std::thread> th;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req *> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
  int64_t max_inflight = 0, max_batch = 0, inflight = 0;
  uint64_t next = 1;
  bool stop = false;
  int64_t stats[4] = {0, 0, 0, 0};  // batches, requests, largest batch (requests), residues

  // Takes the next batch off the queue (lock held); empty if the head does not fit yet.
  std::vector<Req *> take(int64_t &res) {
    std::vector<Req *> b;
    res = 0;
    if (queue.empty()) return b;
    Req *head = queue.front();
    if (inflight > 0 && inflight + head->L > max_inflight) return b;
    b.push_back(head);
    res = head->L;
    for (auto it = queue.begin() + 1; it != queue.end(); ++it) {
      Req *q = *it;
      if (q->kind != head->kind) continue;
      if (head->kind == 0 && (q->p7 != head->p7 || !same_opts(q->o, head->o))) continue;
      if (head->kind == 1 && !same_search(*q, *head)) continue;
      if (res + q->L > max_batch || inflight + res + q->L > max_inflight) continue;
      b.push_back(q);
      res += q->L;
    }
    for (Req *q : b) queue.erase(std::find(queue.begin(), queue.end(), q));
    return b;
  }

  // A batch of search requests: one ixg_search over a target holding their
  // sequences; hits go back to their request (sequence index -> 0).
  static int run_search_batch(ixg_ctx *c, std::vector<Req *> &b) {
    std::vector<const uint8_t *> ptr(b.size());
    std::vector<std::vector<uint8_t>> pad(b.size());
    std::vector<int64_t> len(b.size());
    for (size_t r = 0; r < b.size(); r++) {
      pad[r].assign(b[r]->L + 2, 255);
      if (b[r]->L) memcpy(pad[r].data() + 1, b[r]->seq.data(), b[r]->L);
      ptr[r] = pad[r].data();
      len[r] = b[r]->L;
      b[r]->hits.clear();
    }
    ixg_target *t = nullptr;
    RCHK(ixg_target_create_multi(c, ptr.data(), len.data(), (int64_t)b.size(), &t));
    const Req &h = *b[0];
    std::vector<ixg_searchhit> hits(256);
    int64_t n = 0;
    int rc = ixg_search(c, h.m.p7, h.m.cm, h.m.cp9, h.m.map, h.m.trcp9, h.m.trmap, t, &h.so, hits.data(),
                        (int64_t)hits.size(), &n, nullptr);
    if (rc == IXG_ERANGE) {
      hits.resize(n);
      rc = ixg_search(c, h.m.p7, h.m.cm, h.m.cp9, h.m.map, h.m.trcp9, h.m.trmap, t, &h.so, hits.data(), n, &n,
                      nullptr);
    }
    const std::string err = g_err;
    ixg_target_free(c, t);
    g_err = err;
    RCHK(rc);
    hits.resize(n);
    for (ixg_searchhit x : hits) {
      Req &q = *b[x.seq];
      x.seq = 0;
      q.hits.push_back(x);
    }
    return IXG_OK;
  }

  static int run_batch(ixg_ctx *c, std::vector<Req *> &b) {
    if (b[0]->kind == 1) return run_search_batch(c, b);
    // layout per request: [255, seq, 255][255, revcomp(seq), 255]
    int64_t tot = 0;
    std::vector<int64_t> off(b.size());
    for (size_t r = 0; r < b.size(); r++) off[r] = tot, tot += 2 * (b[r]->L + 2);
    uint8_t *h, *d;
    RCHK(pinned(c, P_SERVE, (size_t)tot, (void **)&h));
    RCHK(scratch(c, S_SERVE_SEQ, (size_t)tot, (void **)&d));
    std::vector<int64_t> base, len;
    for (size_t r = 0; r < b.size(); r++) {
      Req &q = *b[r];
      const int64_t L = q.L;
      uint8_t *f = h + off[r], *rc = f + L + 2;
      f[0] = f[L + 1] = rc[0] = rc[L + 1] = 255;
      if (L) memcpy(f + 1, q.seq.data(), L);
      for (int64_t j = 1; j <= L; j++) {  // reverse complement (cm_search.rs:1822-1827)
        const uint8_t x = f[L + 1 - j];
        rc[j] = x < 4 ? (uint8_t)(3 - x) : x;
      }
      const int64_t nt = ixg_window_tasks(L, q.o.maxw, 1, 1, nullptr, 0);
      q.tasks.resize(nt);
      ixg_window_tasks(L, q.o.maxw, 1, 1, q.tasks.data(), nt);
      for (auto &t : q.tasks) {
        base.push_back(off[r] + (t.strand == 0 ? t.start - 1 : (L + 2) + (L - t.end)));
        len.push_back(t.end - t.start + 1);
      }
    }
    for (Req *q : b) q->wins.clear(), q->acct.clear();
    if (base.empty()) return IXG_OK;  // only empty sequences: no tasks
    HIPCHK(hipMemcpyAsync(d, h, (size_t)tot, hipMemcpyHostToDevice, c->stream));
    std::vector<ixg_window> wins;
    std::vector<int64_t> first;
    std::vector<ixg_acct> acct;
    RCHK(cascade_views(c, b[0]->p7, d, base.data(), len.data(), (int64_t)base.size(), b[0]->o, wins, first, acct));
    size_t t = 0;
    for (size_t r = 0; r < b.size(); r++) {
      Req &q = *b[r];
      const size_t t0 = t, nt = q.tasks.size();
      q.wins.assign(wins.begin() + first[t0], wins.begin() + first[t0 + nt]);
      for (auto &w : q.wins) w.task -= (int32_t)t0;
      q.acct.assign(acct.begin() + t0, acct.begin() + t0 + nt);
      t += nt;
    }
    return IXG_OK;
  }

  void worker(int w) {
    ixg_ctx *c = ctx[w];
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      int64_t res = 0;
      std::vector<Req *> b;
      cv_work.wait(lk, [&] {
        if (stop && queue.empty()) return true;
        b = take(res);
        return !b.empty();
      });
      if (b.empty()) return;  // stopping with nothing queued
      inflight += res;
      const auto now = Clock::now();
      for (Req *q : b) q->state = 1, q->t_run = now, q->batch_size = (int64_t)b.size();
      lk.unlock();
      g_err.clear();
      const int rc = run_batch(c, b);
      const std::string err = g_err;
      lk.lock();
      inflight -= res;
      const auto done = Clock::now();
      for (Req *q : b) q->state = 2, q->rc = rc, q->err = err, q->t_done = done;
      stats[0]++;
      stats[1] += (int64_t)b.size();
      stats[2] = std::max<int64_t>(stats[2], (int64_t)b.size());
      stats[3] += 2 * res;
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }
};

extern "C" {

int ixg_server_create(ixg_ctx *owner, int nstreams, size_t mem_budget, int64_t max_inflight, int64_t max_batch,
                      ixg_server **out) {
  if (!owner || !out || nstreams < 1 || nstreams > 16 || max_inflight < 1 || max_batch < 1) return IXG_EINVAL;
  auto s = std::make_unique<ixg_server>();
  s->device = owner->device;
  s->max_inflight = max_inflight;
  s->max_batch = max_batch;
  for (int i = 0; i < nstreams; i++) {
    ixg_ctx *c = nullptr;
    int rc = ixg_init(owner->device, mem_budget / (size_t)nstreams, &c);
    if (rc != IXG_OK) {
      for (auto *x : s->ctx) ixg_shutdown(x);
      return rc;
    }
    c->autotune = 0;  // serving keeps the fixed launch shapes: no tuning runs inside a request's latency
    s->ctx.push_back(c);
    // staging sized up front for a full batch, so serving does not free and
    // reallocate pinned memory (a device-wide synchronising call) while other
    // workers' copies are in flight
    void *h;
    int r1 = pinned(c, P_SERVE, (size_t)(2 * max_batch) + ((size_t)1 << 20), &h);
    int r2 = r1 == IXG_OK ? pinned(c, P_UP, (size_t)4 << 20, &h) : r1;
    int r3 = r2 == IXG_OK ? pinned(c, P_DOWN, (size_t)4 << 20, &h) : r2;
    if (r3 != IXG_OK) {
      for (auto *x : s->ctx) ixg_shutdown(x);
      return r3;
    }
  }
  ixg_server *raw = s.get();
  for (int i = 0; i < nstreams; i++) s->th.emplace_back([raw, i] { raw->worker(i); });
  *out = s.release();
  return IXG_OK;
}

int ixg_server_submit(ixg_server *s, const ixg_p7 *p7, const uint8_t *dsq, int64_t L, const ixg_f3opts *o,
                      uint64_t *ticket) {
  if (!s || !p7 || !dsq || L < 0 || !o || !ticket || o->maxw < 1 || o->cmw < 1) return IXG_EINVAL;
  if (!codes_valid(dsq + 1, L)) return IXG_EINVAL;
  auto q = std::make_unique<Req>();
  q->p7 = p7;
  q->o = *o;
  q->L = L;
  q->seq.assign(dsq + 1, dsq + 1 + L);
  q->t_sub = Clock::now();
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return IXG_EINVAL;
  q->ticket = s->next++;
  *ticket = q->ticket;
  s->queue.push_back(q.get());
  s->reqs.emplace(q->ticket, std::move(q));
  s->cv_work.notify_one();
  return IXG_OK;
}

int ixg_server_submit_search(ixg_server *s, const ixg_search_model *m, const uint8_t *dsq, int64_t L,
                             const ixg_searchopts *o, uint64_t *ticket) {
  if (!s || !m || !m->p7 || !m->cm || !m->cp9 || !m->map || !dsq || L < 0 || !o || !ticket || o->f3.maxw < 1 ||
      o->f3.cmw < 1 || (o->trunc_mode == 1 && (!m->trcp9[0] || !m->trcp9[1] || !m->trcp9[2] || !m->trmap[0] ||
                                               !m->trmap[1] || !m->trmap[2]))) {
    g_err = "ixg_server_submit_search: bad arguments";
    return IXG_EINVAL;
  }
  if (!codes_valid(dsq + 1, L)) return IXG_EINVAL;
  auto q = std::make_unique<Req>();
  q->kind = 1;
  q->m = *m;
  q->so = *o;
  q->L = L;
  q->seq.assign(dsq + 1, dsq + 1 + L);
  q->t_sub = Clock::now();
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return IXG_EINVAL;
  q->ticket = s->next++;
  *ticket = q->ticket;
  s->queue.push_back(q.get());
  s->reqs.emplace(q->ticket, std::move(q));
  s->cv_work.notify_one();
  return IXG_OK;
}

int ixg_server_search_result(ixg_server *s, uint64_t ticket, ixg_searchhit *hits, int64_t cap, int64_t *n,
                             double *times_ms) {
  if (!s || !n || (cap > 0 && !hits)) return IXG_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end() || it->second->kind != 1) {
    g_err = "ixg_server_search_result: no such search ticket";
    return IXG_EINVAL;
  }
  Req *q = it->second.get();
  s->cv_done.wait(lk, [&] { return q->state == 2; });
  if (q->rc != IXG_OK) {
    g_err = q->err;
    const int rc = q->rc;
    s->reqs.erase(it);
    return rc;
  }
  *n = (int64_t)q->hits.size();
  if (*n > cap) return IXG_ERANGE;  // ticket kept: ask again with room
  if (*n) memcpy(hits, q->hits.data(), q->hits.size() * sizeof(ixg_searchhit));
  if (times_ms) {
    times_ms[0] = ms_between(q->t_sub, q->t_run);
    times_ms[1] = ms_between(q->t_run, q->t_done);
    times_ms[2] = ms_between(q->t_sub, q->t_done);
    times_ms[3] = (double)q->batch_size;
  }
  s->reqs.erase(it);
  return IXG_OK;
}

int ixg_server_wait(ixg_server *s, uint64_t ticket, int64_t timeout_us, int *done) {
  if (!s || !done) return IXG_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return IXG_EINVAL;
  Req *q = it->second.get();
  if (timeout_us < 0)
    s->cv_done.wait(lk, [&] { return q->state == 2; });
  else if (timeout_us > 0)
    s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), [&] { return q->state == 2; });
  *done = q->state == 2;
  return IXG_OK;
}

int ixg_server_result(ixg_server *s, uint64_t ticket, ixg_task *tasks, int64_t tcap, int64_t *ntasks,
                      ixg_window *wins, int64_t wcap, int64_t *nwins, ixg_acct *acct, double *times_ms) {
  if (!s || !ntasks || !nwins) return IXG_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end() || it->second->kind != 0) return IXG_EINVAL;
  Req *q = it->second.get();
  s->cv_done.wait(lk, [&] { return q->state == 2; });
  if (q->rc != IXG_OK) {
    g_err = q->err;
    const int rc = q->rc;
    s->reqs.erase(it);
    return rc;
  }
  *ntasks = (int64_t)q->tasks.size();
  *nwins = (int64_t)q->wins.size();
  if (*ntasks > tcap || *nwins > wcap) return IXG_ERANGE;  // ticket kept: ask again with room
  if (*ntasks && tasks) memcpy(tasks, q->tasks.data(), q->tasks.size() * sizeof(ixg_task));
  if (*nwins && wins) memcpy(wins, q->wins.data(), q->wins.size() * sizeof(ixg_window));
  if (*ntasks && acct) memcpy(acct, q->acct.data(), q->acct.size() * sizeof(ixg_acct));
  if (times_ms) {
    times_ms[0] = ms_between(q->t_sub, q->t_run);
    times_ms[1] = ms_between(q->t_run, q->t_done);
    times_ms[2] = ms_between(q->t_sub, q->t_done);
    times_ms[3] = (double)q->batch_size;
  }
  s->reqs.erase(it);
  return IXG_OK;
}

int ixg_server_stats(ixg_server *s, int64_t *stats4) {
  if (!s || !stats4) return IXG_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  for (int i = 0; i < 4; i++) stats4[i] = s->stats[i];
  return IXG_OK;
}

int ixg_server_free(ixg_server *s) {
  if (!s) return IXG_EINVAL;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &t : s->th) t.join();
  for (auto *c : s->ctx) ixg_shutdown(c);
  delete s;
  return IXG_OK;
}

}  // extern "C"
