// librbg: the served form.  A request loop on one device: assemblies arrive
// as sessions (rbg_server_submit), later contig batches extend them
// (rbg_server_extend), and a worker thread serves whatever is pending in
// steps.  A step is one batched build (rbg_build.cpp) over the requests
// admitted to it, so requests that arrive while others are on the device
// share the next step's launches (continuous batching).  Admission keeps the
// device memory of resident sessions plus the step's working set within the
// HBM budget.  A session's sketches, depth rows, int8 rows and candidate
// lists stay resident after its build, so an extension uploads and sketches
// only its new contigs and evaluates only the pairs that involve them.
// Steps are pipelined over two workspaces: the worker uploads and builds a
// step and saves its sessions' state (device copies on its own stream, so a
// session's next request can join the very next step), then a finisher
// thread copies the step's graphs to the host on a second stream while the
// worker goes on with the next step.  Steps finish in order.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "rbg_build.h"

using namespace rbg;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t id = 0, session = 0;
  bool extend = false;
  rbg_assembly_in in{};  // caller-owned buffers, valid until the request is done
  int state = 0;         // 0 pending, 1 running, 2 done
  rbg_status status = RBG_OK;
  std::string err;
  rbg_graph_result result{};
  rbg_request_info info{};
};

struct Session {
  uint64_t id = 0;
  uint32_t S = 0;
  uint32_t n_accepted = 0;  // contigs after every accepted request
  uint32_t queued = 0;      // requests not yet done
  bool released = false;
  bool failed = false;
  std::string fail_msg;
  Resident res;
};

// Resident bytes per row of a session (planes, padded depth rows, int8 rows,
// list offsets and an allowance of 2K list entries per row); a session whose
// pairs are counted densely also keeps its n x n counts (resident_matrix_bytes).
size_t resident_matrix_bytes(uint32_t n, uint32_t S, const rbg_graph_params& P) {
  return dense_counts(regime_for(P.regime, n), S, P.min_edge_weight) ? (size_t)n * n * 2 : 0;
}
size_t resident_row_bytes(uint32_t S, const rbg_graph_params& P) {
  const size_t W = words_for(P.m);
  size_t b = 8 * W * 4 + 8 + (size_t)2 * P.max_edges * 8;
  if (S > 1) b += 2 * 16 * nc_for(S) * 4 + kpad_for(S);
  return b;
}

void set_err(char* dst, size_t cap, const std::string& s) {
  std::strncpy(dst, s.c_str(), cap - 1);
  dst[cap - 1] = 0;
}

}  // namespace

struct rbg_server {
  rbg_ctx* ctx = nullptr;   // the worker's stream: uploads and builds
  rbg_ctx* ctx2 = nullptr;  // the finisher's stream: results and saved state
  rbg_server_config cfg{};
  std::thread worker, finisher;
  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_ws, cv_jobs;
  std::deque<Request*> pending;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  std::unordered_map<uint64_t, std::unique_ptr<Session>> sessions;
  uint64_t next_req = 1, next_session = 1;
  bool stop = false;
  rbg_asm ws[2];                // step workspaces (buffers reused across steps)
  hipEvent_t built[2] = {};     // a workspace's build is complete
  bool ws_busy[2] = {false, false};
  struct Job {                  // a built step for the finisher
    std::vector<Request*> step;
    int w = 0;
    rbg_status st = RBG_OK;
    std::string msg;
    std::vector<rbg_status> saved;  // per request: its state was kept
    std::vector<std::string> save_msg;
    std::vector<uint32_t> n;        // per request: the session's contigs as of it
    double t_start = 0, t_up = 0, t_built = 0, t_save = 0;
  };
  std::deque<Job> jobs;
  bool worker_done = false;
  Clock::time_point t0 = Clock::now();
  rbg_server_stats stats{};
  std::string err;
  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

  size_t resident_bytes() const {  // under mu
    size_t s = 0;
    for (auto& kv : sessions) s += kv.second->res.device_bytes();
    return s;
  }
  size_t busy_ws_bytes() const {  // under mu
    return (ws_busy[0] ? ws[0].device_bytes() : 0) + (ws_busy[1] ? ws[1].device_bytes() : 0);
  }
  // Requests of the next step, FIFO, at most one per session, within the budget.
  void select(std::vector<Request*>& step);
  void build_step(std::vector<Request*>& step, int w);
  void finish_step(Job& job);
  void loop();
  void finish_loop();
  void finish(Request* r, rbg_status st, const std::string& msg);
  void drop_session_if_idle(Session* s);
};

void rbg_server::finish(Request* r, rbg_status st, const std::string& msg) {
  r->state = 2;
  r->status = st;
  r->err = msg;
  r->info.t_done = now();
  r->info.latency_ms = 1e3 * (r->info.t_done - r->info.t_submit);
  r->info.status = st;
  set_err(r->info.error, sizeof(r->info.error), msg);
  auto it = sessions.find(r->session);
  if (it != sessions.end()) {
    Session* s = it->second.get();
    if (s->queued) --s->queued;
    if (st != RBG_OK) {
      s->failed = true;
      s->fail_msg = msg;
    }
    drop_session_if_idle(s);
  }
}

void rbg_server::drop_session_if_idle(Session* s) {
  if (s->released && s->queued == 0) sessions.erase(s->id);
}

void rbg_server::select(std::vector<Request*>& step) {
  const rbg_graph_params& P = cfg.params;
  std::unordered_set<uint64_t> seen;  // sessions with an earlier pending request
  std::vector<AsmInput> ins;
  size_t growth = 0;
  uint64_t rows = 0;
  const size_t resident = resident_bytes() + busy_ws_bytes();
  for (auto it = pending.begin(); it != pending.end();) {
    Request* r = *it;
    if (step.size() >= cfg.max_step_requests) break;
    if (!seen.insert(r->session).second) {  // keep each session's requests in order
      ++it;
      continue;
    }
    Session* s = sessions.at(r->session).get();
    if (s->failed) {
      it = pending.erase(it);
      finish(r, RBG_ERR_ARG, "session failed earlier: " + s->fail_msg);
      continue;
    }
    const uint32_t n = (r->extend ? s->res.n : 0) + r->in.n;
    if (rows + n > cfg.max_step_rows && !step.empty()) {
      ++it;
      continue;
    }
    AsmInput x;
    x.in = r->in;
    x.src = r->extend ? &s->res : nullptr;
    ins.push_back(x);
    // growth of the session's state (its buffers grow with 25% headroom)
    const size_t after = ((size_t)n * resident_row_bytes(r->in.S, P) + resident_matrix_bytes(n, r->in.S, P)) * 5 / 4;
    const size_t g = after - std::min<size_t>(s->res.device_bytes(), after);
    const size_t need = resident + growth + g + estimate_build_bytes(ins.data(), (uint32_t)ins.size(), &P);
    if (need > cfg.hbm_budget) {
      ins.pop_back();
      if (step.empty() && stats.running == 0) {
        // alone it does not fit: only a release by the client could make room
        const size_t alone = resident + g + estimate_build_bytes(&x, 1, &P);
        if (alone > cfg.hbm_budget) {
          it = pending.erase(it);
          finish(r, RBG_ERR_BUDGET,
                 "request needs ~" + std::to_string(alone >> 20) + " MiB with " +
                     std::to_string(resident >> 20) + " MiB resident; budget " +
                     std::to_string(cfg.hbm_budget >> 20) + " MiB");
          continue;
        }
      }
      break;  // admission stays in arrival order: later requests wait behind this one
    }
    growth += g;
    rows += n;
    r->state = 1;
    r->info.t_admit = now();
    step.push_back(r);
    it = pending.erase(it);
  }
}

void rbg_server::build_step(std::vector<Request*>& step, int w) {
  const rbg_graph_params& P = cfg.params;
  const uint32_t A = (uint32_t)step.size();
  std::vector<AsmInput> ins(A);
  size_t others;
  {
    std::lock_guard<std::mutex> lk(mu);
    for (uint32_t k = 0; k < A; ++k) {
      Session* se = sessions.at(step[k]->session).get();
      ins[k].in = step[k]->in;
      ins[k].src = step[k]->extend ? &se->res : nullptr;
    }
    others = resident_bytes() + (ws_busy[1 - w] ? ws[1 - w].device_bytes() : 0);
  }
  // The workspace holds the largest step so far: give it back when the
  // budget needs the room.
  if (others + ws[w].device_bytes() > cfg.hbm_budget) ws[w].release_buffers();
  Job job;
  job.step = step;
  job.w = w;
  job.t_start = now();
  rbg_status st = batch_upload(ctx, &ws[w], ins.data(), A, P.min_cv);
  job.t_up = now();
  if (st == RBG_OK) st = run_build(ctx, &ws[w], &P, false, false, nullptr);
  if (st == RBG_OK && hipEventRecord(built[w], ctx->stream) != hipSuccess) st = ctx->fail(RBG_ERR_HIP, "event");
  job.t_built = now();
  job.st = st;
  if (st != RBG_OK) job.msg = ctx->err;
  // the sessions' state, on this stream: the next step that extends a session
  // reads it after these copies
  job.saved.assign(A, st);
  job.save_msg.assign(A, std::string());
  job.n.assign(A, 0);
  if (st == RBG_OK) {
    std::lock_guard<std::mutex> lk(mu);  // resident sizes are read under mu
    for (uint32_t k = 0; k < A; ++k) {
      Session* se = sessions.at(step[k]->session).get();
      job.saved[k] = save_resident(ctx, &ws[w], k, &P, &se->res);
      if (job.saved[k] != RBG_OK) job.save_msg[k] = ctx->err;
      job.n[k] = ws[w].asms[k].n;
    }
  }
  job.t_save = now();
  {
    std::lock_guard<std::mutex> lk(mu);
    // A failed build or save leaves a session without (or with partly
    // refitted) device state: mark it failed now, before the next select()
    // could admit one of its queued requests against that state.
    for (uint32_t k = 0; k < A; ++k) {
      const rbg_status sk = st != RBG_OK ? st : job.saved[k];
      if (sk == RBG_OK) continue;
      auto it = sessions.find(step[k]->session);
      if (it == sessions.end()) continue;
      it->second->failed = true;
      it->second->fail_msg = st != RBG_OK ? (job.msg.empty() ? "build failed" : job.msg) : job.save_msg[k];
    }
    jobs.push_back(std::move(job));
  }
  cv_jobs.notify_one();
}

void rbg_server::finish_step(Job& job) {
  const rbg_graph_params& P = cfg.params;
  const uint32_t A = (uint32_t)job.step.size();
  rbg_asm& b = ws[job.w];
  rbg_status st = job.st;
  if (st == RBG_OK && hipEventSynchronize(built[job.w]) != hipSuccess) st = RBG_ERR_HIP;
  std::vector<rbg_status> rs(A, st);
  std::vector<std::string> msg(A, st == RBG_OK ? std::string() : (job.msg.empty() ? "build failed" : job.msg));
  double t_res = 0, t_save = 0;
  if (st == RBG_OK)
    for (uint32_t k = 0; k < A; ++k) {
      Request* r = job.step[k];
      const AsmHost& h = b.asms[k];
      const bool incr = h.incr || h.dense_incr;
      r->info.incremental = incr ? 1 : 0;
      r->info.pass_pairs = incr ? (uint64_t)h.n * (h.n - 1) / 2 - (uint64_t)h.n_old * (h.n_old - 1) / 2
                                : (uint64_t)h.n * (h.n - 1) / 2;
      const double t0s = now();
      rs[k] = assembly_result(ctx2, &b, k, &P, false, &r->result, nullptr);
      t_res += now() - t0s;
      if (rs[k] != RBG_OK) msg[k] = ctx2->err;
      if (rs[k] == RBG_OK && job.saved[k] != RBG_OK) {
        rs[k] = job.saved[k];
        msg[k] = job.save_msg[k];
      }
      if (rs[k] != RBG_OK) rbg_graph_result_free(&r->result);
    }
  t_save = job.t_save - job.t_built;
  // the copies read the workspace: done before it takes the next step
  (void)hipStreamSynchronize(ctx2->stream);
  const double t_end = now();
  std::lock_guard<std::mutex> lk(mu);
  stats.steps += 1;
  stats.busy_ms += 1e3 * ((job.t_save - job.t_start) + t_res);
  stats.upload_ms += 1e3 * (job.t_up - job.t_start);
  stats.build_ms += 1e3 * (job.t_built - job.t_up);
  stats.result_ms += 1e3 * t_res;
  stats.save_ms += 1e3 * t_save;
  (void)t_end;
  for (uint32_t k = 0; k < A; ++k) {
    Request* r = job.step[k];
    r->info.step = stats.steps;
    r->info.step_requests = A;
    if (job.n.size() == A && job.n[k]) r->info.n = job.n[k];  // later steps may have grown the session
    if (rs[k] == RBG_OK) {
      stats.requests += 1;
      stats.graph_pairs += (uint64_t)r->info.n * (r->info.n - 1) / 2;
      stats.pass_pairs += r->info.pass_pairs;
    }
    finish(r, rs[k], msg[k]);
  }
  stats.running -= std::min<uint32_t>(stats.running, A);
  ws_busy[job.w] = false;
  cv_ws.notify_all();
  cv_work.notify_all();
  cv_done.notify_all();
}

void rbg_server::loop() {
  (void)hipSetDevice(ctx->device);
  for (;;) {
    std::vector<Request*> step;
    int w = 0;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] { return stop || !pending.empty(); });
      if (!stop) cv_ws.wait(lk, [&] { return stop || !ws_busy[0] || !ws_busy[1]; });
      if (stop) {
        while (!pending.empty()) {
          Request* r = pending.front();
          pending.pop_front();
          finish(r, RBG_ERR_ARG, "server stopped");
        }
        worker_done = true;
        cv_jobs.notify_all();
        cv_done.notify_all();
        return;
      }
      w = ws_busy[0] ? 1 : 0;
      select(step);
      if (step.empty()) {
        cv_done.notify_all();  // requests may have been failed by admission
        // wait for a finished step (its sessions rejoin) or for new requests
        if (!pending.empty()) cv_work.wait_for(lk, std::chrono::milliseconds(5));
        continue;
      }
      ws_busy[w] = true;
      stats.running += (uint32_t)step.size();
    }
    build_step(step, w);
  }
}

void rbg_server::finish_loop() {
  (void)hipSetDevice(ctx->device);
  for (;;) {
    Job job;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_jobs.wait(lk, [&] { return !jobs.empty() || worker_done; });
      if (jobs.empty()) return;
      job = std::move(jobs.front());
      jobs.pop_front();
    }
    finish_step(job);
  }
}

extern "C" {

void rbg_server_default_config(rbg_server_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->hbm_budget = 0;
  c->max_step_requests = 64;
  c->max_step_rows = 1u << 22;
  rbg_default_params(&c->params);
  c->schedule = rbg_schedule{16, 64};
}

rbg_status rbg_server_start(int device, const rbg_server_config* cfg, rbg_server** out) {
  if (!cfg || !out) return RBG_ERR_ARG;
  *out = nullptr;
  const rbg_graph_params& P = cfg->params;
  if (P.m < 2 || P.m > 512 || P.max_edges == 0 || P.max_edges > 1024) return RBG_ERR_UNSUPPORTED;
  if (P.semantics) return RBG_ERR_UNSUPPORTED;  // non-default reference semantics (rbg_env_semantics)
  if (!(P.near_tol >= 0.0 && P.near_tol <= 1.0)) return RBG_ERR_ARG;
  if (cfg->max_step_requests == 0 || cfg->max_step_requests > 65535 || cfg->max_step_rows == 0 ||
      cfg->max_step_rows >= (1u << 24))
    return RBG_ERR_ARG;
  const rbg_schedule& sc = cfg->schedule;
  if (sc.gate_tiles_per_item < 1 || sc.gate_tiles_per_item > 64 || sc.dense_rows_per_item < 1 ||
      sc.dense_rows_per_item > 4096)
    return RBG_ERR_ARG;
  auto s = std::make_unique<rbg_server>();
  RBG_TRY(rbg_open(device, &s->ctx));
  if (rbg_open(device, &s->ctx2) != RBG_OK) {
    rbg_close(s->ctx);
    return RBG_ERR_HIP;
  }
  for (auto& e : s->built)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      rbg_close(s->ctx2);
      rbg_close(s->ctx);
      return RBG_ERR_HIP;
    }
  s->cfg = *cfg;
  s->cfg.params.flags = 0;
  if (s->cfg.hbm_budget == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
      rbg_close(s->ctx2);
      rbg_close(s->ctx);
      return RBG_ERR_HIP;
    }
    s->cfg.hbm_budget = free_b / 10 * 9;
  }
  s->ws[0].sched = s->ws[1].sched = s->cfg.schedule;
  s->stats.budget = s->cfg.hbm_budget;
  // the sketch tables of the configuration, before any request is timed
  rbg_ctx::Tables* t;
  const rbg_status st = get_tables(s->ctx, P.m, P.pmh_seed, &t);
  if (st != RBG_OK) {
    rbg_close(s->ctx2);
    rbg_close(s->ctx);
    return st;
  }
  rbg_server* raw = s.release();
  raw->worker = std::thread([raw] { raw->loop(); });
  raw->finisher = std::thread([raw] { raw->finish_loop(); });
  *out = raw;
  return RBG_OK;
}

void rbg_server_stop(rbg_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_ws.notify_all();
  if (s->worker.joinable()) s->worker.join();  // then the finisher drains the built steps
  if (s->finisher.joinable()) s->finisher.join();
  (void)hipSetDevice(s->ctx->device);
  for (auto& kv : s->reqs) rbg_graph_result_free(&kv.second->result);
  s->reqs.clear();
  s->sessions.clear();
  for (auto& w : s->ws) w.release_buffers();
  for (auto& e : s->built)
    if (e) (void)hipEventDestroy(e);
  rbg_close(s->ctx2);
  rbg_close(s->ctx);
  delete s;
}

const char* rbg_server_last_error(const rbg_server* s) {
  static thread_local std::string copy;
  if (!s) return "no server";
  std::lock_guard<std::mutex> lk(const_cast<rbg_server*>(s)->mu);
  copy = s->err;
  return copy.c_str();
}

static rbg_status check_input(rbg_server* s, const rbg_assembly_in* a) {
  if (!a || a->n == 0 || !a->seqs || !a->offsets || (a->S > 0 && !a->depth)) {
    s->err = "empty or missing input";
    return RBG_ERR_ARG;
  }
  if (a->S > 64) {
    s->err = "graph builds support at most 64 depth samples";
    return RBG_ERR_UNSUPPORTED;
  }
  return RBG_OK;
}

rbg_status rbg_server_submit(rbg_server* s, const rbg_assembly_in* a, uint64_t* session,
                             uint64_t* request) {
  if (!s || !session || !request) return RBG_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  RBG_TRY(check_input(s, a));
  if (a->n >= s->cfg.max_step_rows) {
    s->err = "assembly larger than a step";
    return RBG_ERR_UNSUPPORTED;
  }
  if (regime_for(s->cfg.params.regime, a->n) == 2 && a->n < 2500) {
    s->err = "fused regime needs n >= 2500";
    return RBG_ERR_ARG;
  }
  auto ses = std::make_unique<Session>();
  ses->id = s->next_session++;
  ses->S = a->S;
  ses->n_accepted = a->n;
  ses->queued = 1;
  auto r = std::make_unique<Request>();
  r->id = s->next_req++;
  r->session = ses->id;
  r->in = *a;
  r->info.session = ses->id;
  r->info.n = a->n;
  r->info.n_new = a->n;
  r->info.t_submit = s->now();
  *session = ses->id;
  *request = r->id;
  s->pending.push_back(r.get());
  s->reqs[r->id] = std::move(r);
  s->sessions[ses->id] = std::move(ses);
  s->cv_work.notify_one();
  return RBG_OK;
}

rbg_status rbg_server_extend(rbg_server* s, uint64_t session, const rbg_assembly_in* more,
                             uint64_t* request) {
  if (!s || !request) return RBG_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  RBG_TRY(check_input(s, more));
  auto it = s->sessions.find(session);
  if (it == s->sessions.end() || it->second->released) {
    s->err = "unknown or released session";
    return RBG_ERR_ARG;
  }
  Session* ses = it->second.get();
  if (more->S != ses->S) {
    s->err = "added contigs need the session's sample count";
    return RBG_ERR_ARG;
  }
  if ((uint64_t)ses->n_accepted + more->n >= s->cfg.max_step_rows) {
    s->err = "session larger than a step";
    return RBG_ERR_UNSUPPORTED;
  }
  ses->n_accepted += more->n;
  ses->queued += 1;
  auto r = std::make_unique<Request>();
  r->id = s->next_req++;
  r->session = session;
  r->extend = true;
  r->in = *more;
  r->info.session = session;
  r->info.n = ses->n_accepted;
  r->info.n_new = more->n;
  r->info.t_submit = s->now();
  *request = r->id;
  s->pending.push_back(r.get());
  s->reqs[r->id] = std::move(r);
  s->cv_work.notify_one();
  return RBG_OK;
}

rbg_status rbg_server_wait(rbg_server* s, uint64_t request, double timeout_s, rbg_graph_result* out,
                           rbg_request_info* info) {
  if (!s || !out) return RBG_ERR_ARG;
  std::memset(out, 0, sizeof(*out));
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(request);
  if (it == s->reqs.end()) {
    s->err = "unknown request";
    return RBG_ERR_ARG;
  }
  Request* r = it->second.get();
  auto done = [&] { return r->state == 2; };
  if (timeout_s < 0) {
    s->cv_done.wait(lk, done);
  } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
    s->err = "request not done within the timeout";
    return RBG_ERR_TIMEOUT;
  }
  *out = r->result;
  r->result = rbg_graph_result{};
  if (info) *info = r->info;
  const rbg_status st = r->status;
  if (st != RBG_OK) s->err = r->err;
  s->reqs.erase(it);
  return st;
}

rbg_status rbg_server_release(rbg_server* s, uint64_t session) {
  if (!s) return RBG_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->sessions.find(session);
  if (it == s->sessions.end() || it->second->released) {
    s->err = "unknown or released session";
    return RBG_ERR_ARG;
  }
  it->second->released = true;
  s->drop_session_if_idle(it->second.get());  // else when its last request is done
  s->cv_work.notify_one();
  return RBG_OK;
}

rbg_status rbg_server_get_stats(rbg_server* s, rbg_server_stats* out) {
  if (!s || !out) return RBG_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  out->sessions = s->sessions.size();
  out->pending = s->pending.size();
  out->resident_bytes = s->resident_bytes();
  out->workspace_bytes = s->ws[0].device_bytes() + s->ws[1].device_bytes();
  return RBG_OK;
}

}  // extern "C"
