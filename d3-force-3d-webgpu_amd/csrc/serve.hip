// serve.hip -- the request loop behind fg_serve_* (include/d3fgpu.h).
//
// A worker thread drives one batch (fg_batch_*): admit queued jobs in arrival order while
// they fit the budget, queue a round of ticks, read the job table, read back and release
// the jobs that finished, repeat.  Submission and waiting happen on the caller's threads
// under one mutex; the batch itself is touched by the worker only.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/d3fgpu.h"

namespace {

using Clock = std::chrono::steady_clock;

// A job copied at submission: the descriptor's arrays live here until admission.
struct Request {
  uint64_t id = 0;
  uint32_t status = FG_SERVE_QUEUED;
  uint64_t batch_id = 0;
  fg_job_desc desc{};
  std::vector<double> pos, vel, fix;
  std::vector<uint8_t> mask;
  std::vector<fg_force_desc> forces;
  std::vector<std::vector<double>> farrays;
  std::vector<std::vector<uint32_t>> flinks;
  std::vector<double> out_pos, out_vel;
  uint32_t ticks = 0;
  double t_submit = 0, t_admit = 0, t_finish = 0;
  std::string error;
};

std::vector<double> copy_d(const double* p, size_t n) {
  return p ? std::vector<double>(p, p + n) : std::vector<double>();
}

void copy_job(const fg_job_desc* j, Request& r) {
  r.desc = *j;
  const size_t n = j->n;
  r.pos = copy_d(j->pos, 3 * n);
  r.vel = copy_d(j->vel, 3 * n);
  r.fix = copy_d(j->fix, 3 * n);
  if (j->fixmask) r.mask.assign(j->fixmask, j->fixmask + n);
  r.forces.assign(j->forces, j->forces + j->nforces);
  for (fg_force_desc& f : r.forces) {
    auto keep = [&](const double*& p, size_t count) {
      if (!p) return;
      r.farrays.push_back(copy_d(p, count));
      p = r.farrays.back().data();
    };
    const size_t m = f.count;
    if (f.links) {
      r.flinks.emplace_back(f.links, f.links + 2 * m);
      f.links = r.flinks.back().data();
    }
    const bool per_link = f.type == FG_FORCE_LINK;
    keep(f.strengths, per_link ? m : n);
    keep(f.radii, n);
    keep(f.distances, m);
    keep(f.bias, m);
    keep(f.targets, n);
  }
  r.desc.pos = r.pos.empty() ? nullptr : r.pos.data();
  r.desc.vel = r.vel.empty() ? nullptr : r.vel.data();
  r.desc.fix = r.fix.empty() ? nullptr : r.fix.data();
  r.desc.fixmask = r.mask.empty() ? nullptr : r.mask.data();
  r.desc.forces = r.forces.empty() ? nullptr : r.forces.data();
}

void drop_inputs(Request& r) {
  std::vector<double>().swap(r.pos);
  std::vector<double>().swap(r.vel);
  std::vector<double>().swap(r.fix);
  std::vector<uint8_t>().swap(r.mask);
  std::vector<fg_force_desc>().swap(r.forces);
  std::vector<std::vector<double>>().swap(r.farrays);
  std::vector<std::vector<uint32_t>>().swap(r.flinks);
  r.desc.pos = r.desc.vel = r.desc.fix = nullptr;
  r.desc.fixmask = nullptr;
  r.desc.forces = nullptr;
  r.desc.nforces = 0;
}

}  // namespace

struct fg_server {
  int device = 0;
  fg_batch* batch = nullptr;
  fg_serve_config cfg{};
  std::string tune_cache;
  Clock::time_point t0 = Clock::now();
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::map<uint64_t, std::unique_ptr<Request>> all;  // every job not yet collected
  std::map<uint64_t, Request*> running;             // batch id -> request
  uint64_t next_id = 1;
  bool stop = false;
  std::thread worker;
  std::string err;
  // statistics
  uint64_t submitted = 0, admitted = 0, finished = 0, failed = 0, rounds = 0, ticks = 0, peak_bytes = 0;
  uint32_t peak_running = 0;
  std::vector<double> latencies;
  double first_submit = -1, last_finish = 0;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void worker_loop(fg_server* s) {
  std::vector<fg_job_info> info(4096);
  while (true) {
    std::vector<Request*> admit_now;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty() || !s->running.empty(); });
      if (s->stop) return;
    }
    // admission in arrival order; the head waits for memory rather than being overtaken.
    // Only this thread takes from the queue, so its head stays put while the job uploads
    // outside the lock (submitters are never held up by an upload).
    while (true) {
      Request* r = nullptr;
      bool busy = false;
      {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop || s->queue.empty()) break;
        r = s->queue.front();
        busy = !s->running.empty();
      }
      uint64_t bid = 0;
      const int rc = fg_batch_submit(s->batch, &r->desc, &bid);
      std::lock_guard<std::mutex> lk(s->mu);
      if (rc == FG_ERR_NOMEM && busy) break;
      s->queue.pop_front();
      if (rc != FG_OK) {
        r->status = FG_SERVE_FAILED;
        r->error = fg_batch_last_error(s->batch);
        r->t_finish = s->now();
        s->failed++;
        drop_inputs(*r);
        s->cv_done.notify_all();
        continue;
      }
      r->status = FG_SERVE_RUNNING;
      r->batch_id = bid;
      r->t_admit = s->now();
      s->running[bid] = r;
      s->admitted++;
      admit_now.push_back(r);
    }
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->peak_running = std::max(s->peak_running, (uint32_t)s->running.size());
      if (s->running.empty()) continue;
    }
    for (Request* r : admit_now) drop_inputs(*r);
    if (!admit_now.empty() && s->cfg.autotune)
      (void)fg_batch_autotune(s->batch, s->tune_cache.empty() ? nullptr : s->tune_cache.c_str(), nullptr);
    const uint32_t k = s->cfg.ticks_per_round ? s->cfg.ticks_per_round : 10;
    fg_batch_stats bst{};
    if (fg_batch_step(s->batch, k) != FG_OK || fg_batch_get_stats(s->batch, &bst) != FG_OK) {
      std::lock_guard<std::mutex> lk(s->mu);
      s->err = fg_batch_last_error(s->batch);
      // the failed jobs leave the batch, so their memory no longer counts against the budget
      for (auto& kv : s->running) {
        (void)fg_batch_release(s->batch, kv.first);
        kv.second->status = FG_SERVE_FAILED;
        kv.second->error = s->err;
        kv.second->t_finish = s->now();
        s->failed++;
      }
      s->running.clear();
      s->cv_done.notify_all();
      continue;
    }
    uint32_t count = 0;
    (void)fg_batch_jobs(s->batch, info.data(), (uint32_t)info.size(), &count);
    std::vector<std::pair<uint64_t, uint32_t>> done;
    for (uint32_t q = 0; q < count && q < info.size(); ++q)
      if (info[q].status == FG_JOB_DONE) done.push_back({info[q].id, info[q].ticks});
    std::vector<std::pair<Request*, uint32_t>> collected;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      for (auto& d : done) {
        auto it = s->running.find(d.first);
        if (it != s->running.end()) collected.push_back({it->second, d.second});
      }
    }
    for (auto& c : collected) {
      Request* r = c.first;
      r->out_pos.assign(3 * (size_t)r->desc.n, 0.0);
      r->out_vel.assign(3 * (size_t)r->desc.n, 0.0);
      int rc = fg_batch_read(s->batch, r->batch_id, r->out_pos.data(), r->out_vel.data());
      (void)fg_batch_release(s->batch, r->batch_id);
      std::lock_guard<std::mutex> lk(s->mu);
      s->running.erase(r->batch_id);
      r->ticks = c.second;
      r->t_finish = s->now();
      if (rc == FG_OK) {
        r->status = FG_SERVE_FINISHED;
        s->finished++;
        s->latencies.push_back(r->t_finish - r->t_submit);
        s->last_finish = r->t_finish;
      } else {
        r->status = FG_SERVE_FAILED;
        r->error = fg_batch_last_error(s->batch);
        s->failed++;
      }
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->rounds++;
    s->ticks += k;
    s->peak_bytes = std::max(s->peak_bytes, (uint64_t)bst.bytes);
    if (!collected.empty()) s->cv_done.notify_all();
  }
}

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t i = (size_t)floor(pos);
  const double f = pos - (double)i;
  return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

}  // namespace

extern "C" {

int fg_serve_create(fg_server** out, int device, const fg_serve_config* cfg) {
  if (!out) return FG_ERR_INVALID;
  *out = nullptr;
  fg_server* s = new (std::nothrow) fg_server();
  if (!s) return FG_ERR_NOMEM;
  s->device = device;
  if (cfg) s->cfg = *cfg;
  if (cfg && cfg->tune_cache) s->tune_cache = cfg->tune_cache;
  s->cfg.tune_cache = nullptr;
  int rc = fg_batch_create(&s->batch, device, s->cfg.hbm_budget);
  if (rc != FG_OK) {
    delete s;
    return rc;
  }
  try {
    s->worker = std::thread(worker_loop, s);
  } catch (...) {
    fg_batch_destroy(s->batch);
    delete s;
    return FG_ERR_NOMEM;
  }
  *out = s;
  return FG_OK;
}

int fg_serve_destroy(fg_server* s) {
  if (!s) return FG_ERR_INVALID;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_done.notify_all();
  if (s->worker.joinable()) s->worker.join();
  fg_batch_destroy(s->batch);
  delete s;
  return FG_OK;
}

const char* fg_serve_last_error(const fg_server* s) { return s ? s->err.c_str() : "null server"; }

int fg_serve_submit(fg_server* s, const fg_job_desc* job, uint64_t* id) {
  if (!s || !job || !id) return FG_ERR_INVALID;
  uint64_t bytes = 0;
  int rc = fg_batch_job_bytes(job, &bytes);
  if (rc != FG_OK) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->err = "the job was refused (see fg_batch_job_bytes)";
    return rc;
  }
  std::unique_ptr<Request> r(new (std::nothrow) Request());
  if (!r) return FG_ERR_NOMEM;
  try {
    copy_job(job, *r);
  } catch (...) {
    return FG_ERR_NOMEM;
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    r->t_submit = s->now();
    if (s->first_submit < 0) s->first_submit = r->t_submit;
    *id = r->id;
    s->queue.push_back(r.get());
    s->all[r->id] = std::move(r);
    s->submitted++;
  }
  s->cv_work.notify_all();
  return FG_OK;
}

int fg_serve_wait(fg_server* s, uint64_t id, double timeout_s, double* pos, double* vel,
                  fg_serve_result* result) {
  if (!s) return FG_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) {
    s->err = "no such job";
    return FG_ERR_INVALID;
  }
  Request* r = it->second.get();
  auto ready = [&] { return r->status == FG_SERVE_FINISHED || r->status == FG_SERVE_FAILED || s->stop; };
  if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s > 0 ? timeout_s : 0.0), ready)) {
    s->err = "timed out waiting for the job";
    return FG_ERR_STATE;
  }
  if (result) {
    result->id = r->id;
    result->status = r->status;
    result->n = r->desc.n;
    result->ndim = r->desc.ndim;
    result->ticks = r->ticks;
    result->submitted_s = r->t_submit;
    result->admitted_s = r->t_admit;
    result->finished_s = r->t_finish;
  }
  if (r->status != FG_SERVE_FINISHED) {
    s->err = r->error.empty() ? "the server stopped" : r->error;
    // a job still queued or running when the server stops may be in the worker's hands (an
    // upload outside the lock): the server frees it once the worker has ended
    if (r->status == FG_SERVE_FAILED) s->all.erase(it);
    return FG_ERR_STATE;
  }
  if (pos) memcpy(pos, r->out_pos.data(), sizeof(double) * r->out_pos.size());
  if (vel) memcpy(vel, r->out_vel.data(), sizeof(double) * r->out_vel.size());
  s->all.erase(it);
  return FG_OK;
}

int fg_serve_drain(fg_server* s, double timeout_s) {
  if (!s) return FG_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto idle = [&] { return (s->queue.empty() && s->running.empty()) || s->stop; };
  if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s > 0 ? timeout_s : 0.0), idle)) {
    s->err = "timed out waiting for the queue to drain";
    return FG_ERR_STATE;
  }
  return FG_OK;
}

int fg_serve_get_stats(fg_server* s, fg_serve_stats* out) {
  if (!s || !out) return FG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  fg_serve_stats st{};
  st.submitted = s->submitted;
  st.admitted = s->admitted;
  st.finished = s->finished;
  st.failed = s->failed;
  st.queued = (uint32_t)s->queue.size();
  st.running = (uint32_t)s->running.size();
  st.peak_running = s->peak_running;
  st.p50_latency_s = percentile(s->latencies, 0.50);
  st.p99_latency_s = percentile(s->latencies, 0.99);
  double sum = 0, mx = 0;
  for (double v : s->latencies) sum += v, mx = std::max(mx, v);
  st.mean_latency_s = s->latencies.empty() ? 0.0 : sum / (double)s->latencies.size();
  st.max_latency_s = mx;
  const double span = s->last_finish - (s->first_submit < 0 ? 0.0 : s->first_submit);
  st.layouts_per_s = span > 0 ? (double)s->finished / span : 0.0;
  st.rounds = s->rounds;
  st.ticks = s->ticks;
  st.peak_bytes = s->peak_bytes;
  fg_batch_stats b{};
  st.budget = s->cfg.hbm_budget;
  *out = st;
  (void)b;
  return FG_OK;
}

}  // extern "C"
