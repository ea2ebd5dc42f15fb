// Serving loop above the engine's C ABI: jobs (deck + method + walkers + seed) arrive at any
// time and are admitted in arrival order while their walker state fits the HBM budget and
// fewer than max_active run.  An admitted job's engine is made (deck read, tables and walkers
// on the device) by the admission thread; the job then joins the batch of its precision, and
// the stepping thread advances every batch member by one step per qw_batch_step, in launches
// shared by all of them (continuous batching: jobs join and leave between steps, never
// waiting for the others).  Jobs a batch cannot hold (BF16 precisions, several determinants,
// T-moves, more than 128 electrons of a spin in mixed precision) run on their own engine and
// host thread as before.  When a job finishes its memory is returned and the next job is
// admitted.
//
// A job's numbers depend only on its deck, seed and method (every random number is keyed by
// the job's seed, walker, step and electron, and a batch member's results are bitwise those
// of its engine alone), so a job gives the same block records alone or next to others.
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
  int member = 0;  // batch member id while it runs in a batch
  std::chrono::steady_clock::time_point t_run;
};

double now_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// The runs of a job: a VMC equilibration from the random walkers (when asked), then its method.
std::vector<qw_method_desc> job_runs(const qw_job_desc& d) {
  std::vector<qw_method_desc> r;
  if (d.equil_steps > 0) {
    qw_method_desc eq = d.method;
    eq.method = 0;
    eq.nconfig = d.nwalkers;
    eq.nblock = 1;
    eq.nstep = d.equil_steps;
    eq.ndecorr = 1;
    eq.timestep = d.method.method == 1 ? (d.method.timestep > 0.05 ? d.method.timestep : 0.05) : d.method.timestep;
    r.push_back(eq);
  }
  qw_method_desc m = d.method;
  m.nconfig = d.nwalkers;
  r.push_back(m);
  return r;
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
  std::thread sched, stepper;
  // jobs whose engines are ready to join a batch, handed from the admission thread
  std::deque<Job*> ready;
  std::map<int, qw_batch*> batches;  // by precision (stepping thread only)
  std::map<int, std::map<int, Job*>> members;  // precision -> member id -> job

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

  // Walkers from the random guess of the job's seed; a walker whose determinant is exactly
  // zero (an electron beyond every basis function's cutoff) is drawn again from the next
  // seed.  false (and the job failed) on an error.
  bool place_walkers(Job* j) {
    qw_engine* e = j->eng;
    int info[6];
    qw_engine_info(e, info);
    const int ne = info[0] + info[1], nw = j->desc.nwalkers;
    const size_t row = (size_t)ne * 3;
    std::vector<double> pos(nw * row), alt(nw * row), ld((size_t)nw * 2);
    qw_status st = qw_walkers_random(e, nw, j->desc.seed, 3.0, pos.data());
    for (int t = 1; st == QW_OK && t <= 20; t++) {
      st = qw_walkers_set(e, nw, pos.data(), nullptr);
      if (st == QW_OK) st = qw_walkers_state(e, nw, ld.data(), nullptr, nullptr, nullptr);
      if (st != QW_OK) break;
      std::vector<int> bad;
      for (int w = 0; w < nw; w++)
        if (!(ld[2 * w] > -1e200 && ld[2 * w + 1] > -1e200)) bad.push_back(w);
      if (bad.empty()) break;
      st = qw_walkers_random(e, nw, j->desc.seed + 7919ull * t, 3.0, alt.data());
      for (int w : bad) std::copy(alt.begin() + w * row, alt.begin() + (w + 1) * row, pos.begin() + w * row);
    }
    if (st != QW_OK) {
      std::string err = qw_last_error(e);
      qw_engine_destroy(e);
      j->eng = nullptr;
      finish(*j, st, err);
      return false;
    }
    return true;
  }

  // A job that no batch holds: its own engine and host thread, qw_run for each run.
  void run_alone(Job* j) {
    qw_engine* e = j->eng;
    qw_status st = QW_OK;
    double rs[4] = {0, 0, 0, 0};
    const auto runs = job_runs(j->desc);
    for (size_t r = 0; r < runs.size() && st == QW_OK; r++) {
      std::vector<double> eb((size_t)runs[r].nblock * QW_BLOCK_FIELDS, 0.0);
      st = qw_run(e, &runs[r], eb.data(), r + 1 == runs.size() ? rs : nullptr);
      if (r + 1 == runs.size()) j->eblock = eb;
    }
    std::string err = st == QW_OK ? std::string() : std::string(qw_last_error(e));
    j->stats[4] = rs[1];
    j->stats[5] = rs[2];
    j->stats[6] = rs[3];
    qw_engine_destroy(e);
    j->eng = nullptr;
    finish(*j, st, err);
  }

  // Stepping thread: new jobs join their precision's batch, every batch takes one step, and
  // finished members leave with their results.
  void step_loop() {
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      size_t live = 0;
      for (auto& p : members) live += p.second.size();
      cv.wait(lk, [&] { return !ready.empty() || live > 0 || (stopping && pending.empty()); });
      std::deque<Job*> take;
      take.swap(ready);
      if (take.empty() && live == 0 && stopping) return;
      lk.unlock();
      for (Job* j : take) join(j);
      for (auto& p : batches) {
        if (members[p.first].empty()) continue;
        int fin[64], nf = 0;
        qw_status st = qw_batch_step(p.second, fin, &nf);
        if (st != QW_OK) {
          // a failed step fails every job of that batch
          std::string err = qw_batch_error(p.second);
          for (auto& m : members[p.first]) {
            Job* j = m.second;
            qw_engine_destroy(j->eng);
            j->eng = nullptr;
            finish(*j, st, err);
          }
          members[p.first].clear();
          qw_batch_destroy(p.second);
          p.second = nullptr;
          continue;
        }
        for (int k = 0; k < nf; k++) {
          Job* j = members[p.first][fin[k]];
          members[p.first].erase(fin[k]);
          j->eblock.assign((size_t)j->desc.method.nblock * QW_BLOCK_FIELDS, 0.0);
          double rs[4] = {0, 0, 0, 0};
          qw_status ls = qw_batch_leave(p.second, fin[k], j->eblock.data(), rs);
          j->stats[4] = rs[1];
          j->stats[5] = now_since(t0) - j->stats[1];
          j->stats[6] = j->stats[5] > 0 ? rs[1] / j->stats[5] : 0.0;
          qw_engine_destroy(j->eng);
          j->eng = nullptr;
          finish(*j, ls, ls == QW_OK ? std::string() : std::string(qw_batch_error(p.second)));
        }
      }
      for (auto it = batches.begin(); it != batches.end();)
        if (!it->second) it = batches.erase(it);
        else ++it;
      lk.lock();
    }
  }

  // Joins the batch of the job's precision, or runs the job alone when no batch holds it.
  void join(Job* j) {
    const int prec = j->desc.precision;
    qw_batch* b = nullptr;
    auto it = batches.find(prec);
    if (it != batches.end()) {
      b = it->second;
    } else if (qw_batch_open(opts.device, prec, 0, &b) == QW_OK) {
      batches[prec] = b;
    }
    if (b) {
      const auto runs = job_runs(j->desc);
      int mid = 0;
      if (qw_batch_join(b, j->eng, (int)runs.size(), runs.data(), &mid) == QW_OK) {
        members[prec][mid] = j;
        j->member = mid;
        return;
      }
    }
    std::lock_guard<std::mutex> g(mu);
    workers.emplace_back([this, j] { run_alone(j); });
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
      lk.unlock();
      const bool ok = place_walkers(j);
      lk.lock();
      if (ok) {
        ready.push_back(j);
        cv.notify_all();
      }
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
  if (s->opts.max_active <= 0) s->opts.max_active = 32;
  if (s->opts.max_active > 64) s->opts.max_active = 64;  // members of one batch
  s->sched = std::thread([s] { s->schedule(); });
  s->stepper = std::thread([s] { s->step_loop(); });
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
  s->stepper.join();
  for (auto& t : s->workers) t.join();
  for (auto& p : s->batches)
    if (p.second) qw_batch_destroy(p.second);
  delete s;
}
