// Serving loop above the engine's C ABI: jobs (deck + method + walkers + seed) arrive at any
// time, are admitted in arrival order while their walker state fits the HBM budget and fewer
// than max_active run, and then run concurrently, each on its own engine (own HIP stream) and
// host thread.  When a job finishes its memory is returned and the scheduler admits the next.
//
// A job's numbers depend only on its deck, seed and method (every random number is keyed by
// the job's seed, walker, step and electron), so a job gives the same block records whether it
// runs alone or next to others.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/qwalk_amd.h"

namespace {

struct Job {
  long long id = 0;
  std::string deck;
  qw_job_desc desc{};
  std::vector<double> eblock;
  double stats[QW_JOB_STATS] = {0};
  qw_status status = QW_OK;
  std::string err;
  bool done = false;
  qw_engine* eng = nullptr;  // created at admission
  unsigned long long bytes = 0;
};

double now_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct qw_server {
  qw_server_opts opts{};
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  std::mutex mu;
  std::condition_variable cv;
  std::deque<long long> pending;
  std::map<long long, std::unique_ptr<Job>> jobs;
  std::vector<std::thread> workers;
  unsigned long long used = 0;
  int active = 0;
  long long next_id = 1;
  bool stopping = false;
  std::thread sched;

  void finish(Job& j, qw_status st, const std::string& err) {
    std::lock_guard<std::mutex> g(mu);
    j.status = st;
    j.err = err;
    j.stats[2] = now_since(t0);
    j.stats[3] = j.stats[2] - j.stats[0];
    j.done = true;
    used -= j.bytes;
    j.bytes = 0;
    active--;
    cv.notify_all();
  }

  // Runs one admitted job to completion on the calling (worker) thread.
  void run_job(Job* j) {
    qw_engine* e = j->eng;
    const qw_job_desc& d = j->desc;
    int info[6];
    qw_engine_info(e, info);
    const int ne = info[0] + info[1], nw = d.nwalkers;
    std::vector<double> pos((size_t)nw * ne * 3);
    qw_status st = qw_walkers_random(e, nw, d.seed, 3.0, pos.data());
    if (st == QW_OK) st = qw_walkers_set(e, nw, pos.data(), nullptr);
    if (st == QW_OK && d.equil_steps > 0) {
      qw_method_desc eq = d.method;
      eq.method = 0;
      eq.nconfig = nw;
      eq.nblock = 1;
      eq.nstep = d.equil_steps;
      eq.ndecorr = 1;
      eq.timestep = d.method.method == 1 ? (d.method.timestep > 0.05 ? d.method.timestep : 0.05) : d.method.timestep;
      double tmp[QW_BLOCK_FIELDS];
      st = qw_run(e, &eq, tmp, nullptr);
    }
    double rs[4] = {0, 0, 0, 0};
    if (st == QW_OK) {
      qw_method_desc m = d.method;
      m.nconfig = nw;
      j->eblock.assign((size_t)(m.nblock > 0 ? m.nblock : 1) * QW_BLOCK_FIELDS, 0.0);
      st = qw_run(e, &m, j->eblock.data(), rs);
    }
    std::string err = st == QW_OK ? std::string() : std::string(qw_last_error(e));
    j->stats[4] = rs[1];
    j->stats[5] = rs[2];
    j->stats[6] = rs[3];
    qw_engine_destroy(e);
    j->eng = nullptr;
    finish(*j, st, err);
  }

  // Admission in arrival order: the head job is admitted when its walker bytes fit.
  void schedule() {
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      cv.wait(lk, [&] { return stopping || !pending.empty(); });
      if (pending.empty() && stopping) return;
      Job* j = jobs[pending.front()].get();
      if (!j->eng) {
        lk.unlock();
        qw_engine_opts eo{j->desc.precision, opts.device, j->desc.seed, 0};
        qw_engine* e = nullptr;
        qw_status st = qw_engine_create(j->deck.c_str(), &eo, &e);
        unsigned long long b = 0;
        if (st == QW_OK) st = qw_engine_walker_bytes(e, j->desc.nwalkers, &b);
        lk.lock();
        if (st != QW_OK) {
          j->status = st;
          j->err = qw_last_error(e);
          if (e) qw_engine_destroy(e);
          j->stats[2] = now_since(t0);
          j->stats[3] = j->stats[2] - j->stats[0];
          j->done = true;
          pending.pop_front();
          cv.notify_all();
          continue;
        }
        j->eng = e;
        j->stats[7] = (double)b;
        if (b > opts.hbm_budget) {  // can never fit: fail it rather than block the queue
          lk.unlock();
          qw_engine_destroy(e);
          lk.lock();
          j->eng = nullptr;
          j->status = QW_ERR_NOMEM;
          j->err = "job needs more device memory than the server's budget";
          j->stats[2] = now_since(t0);
          j->stats[3] = j->stats[2] - j->stats[0];
          j->done = true;
          pending.pop_front();
          cv.notify_all();
          continue;
        }
      }
      const unsigned long long b = (unsigned long long)j->stats[7];
      if (used + b > opts.hbm_budget || active >= opts.max_active) {
        cv.wait(lk);  // a finishing job (or a stop) wakes us
        continue;
      }
      used += b;
      j->bytes = b;
      active++;
      j->stats[1] = now_since(t0);
      pending.pop_front();
      workers.emplace_back([this, j] { run_job(j); });
    }
  }
};

extern "C" qw_status qw_server_create(const qw_server_opts* opts, qw_server** out) {
  if (!out) return QW_ERR_ARG;
  *out = nullptr;
  qw_server* s = new (std::nothrow) qw_server;
  if (!s) return QW_ERR_NOMEM;
  if (opts) s->opts = *opts;
  if (s->opts.hbm_budget == 0) s->opts.hbm_budget = 64ull << 30;
  if (s->opts.max_active <= 0) s->opts.max_active = 8;
  s->sched = std::thread([s] { s->schedule(); });
  *out = s;
  return QW_OK;
}

extern "C" qw_status qw_server_submit(qw_server* s, const char* deck_path, const qw_job_desc* job, long long* job_id) {
  if (!s || !deck_path || !job || !job_id || job->nwalkers <= 0 || job->method.nblock <= 0) return QW_ERR_ARG;
  std::unique_ptr<Job> j(new Job);
  j->deck = deck_path;
  j->desc = *job;
  std::lock_guard<std::mutex> g(s->mu);
  j->id = s->next_id++;
  j->stats[0] = now_since(s->t0);
  *job_id = j->id;
  s->pending.push_back(j->id);
  s->jobs[j->id] = std::move(j);
  s->cv.notify_all();
  return QW_OK;
}

extern "C" qw_status qw_server_wait(qw_server* s, long long job_id, double* eblock, double* stats) {
  if (!s) return QW_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return QW_ERR_ARG;
  Job* j = it->second.get();
  s->cv.wait(lk, [&] { return j->done; });
  if (eblock && !j->eblock.empty()) std::memcpy(eblock, j->eblock.data(), j->eblock.size() * sizeof(double));
  if (stats) std::memcpy(stats, j->stats, sizeof(j->stats));
  return j->status;
}

extern "C" const char* qw_server_error(qw_server* s, long long job_id) {
  if (!s) return "null server";
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->jobs.find(job_id);
  return it == s->jobs.end() ? "unknown job" : it->second->err.c_str();
}

extern "C" void qw_server_destroy(qw_server* s) {
  if (!s) return;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv.wait(lk, [&] { return s->pending.empty() && s->active == 0; });
    s->stopping = true;
    s->cv.notify_all();
  }
  s->sched.join();
  for (auto& t : s->workers) t.join();
  delete s;
}
