// Serving loop above the engine's C ABI: jobs (deck + method + walkers + seed) arrive at any
// time and are admitted in arrival order while their walker state fits the HBM budget and
// fewer than max_active run.  A deck is parsed once per server (a template engine per deck
// file and precision, checked against the file's size and modification time); an admitted
// job's engine is cloned from it and its walkers placed by a pool of set-up threads, so the
// set-up of several jobs overlaps and never holds up admission or stepping.  The job then
// joins the batch of its precision in its size lane (at most 64 electrons of a spin, or
// more), and the lane's stepping thread advances every batch member by one step per
// qw_batch_step, in launches shared by all of them (continuous batching: jobs join and leave
// between steps, never waiting for the others; the two lanes step independently on their own
// streams, so small jobs are not held to the step time of large ones).  Jobs a batch cannot hold
// (BF16 precisions, several determinants, T-moves, more than 128 electrons of a spin in mixed
// precision) run on their own engine and host thread.  A finished job's engine is released
// by the set-up pool; its memory returns to the budget when it is freed, and the next job is
// admitted.
//
// A job's numbers depend only on its deck, seed and method (every random number is keyed by
// the job's seed, walker, step and electron, a cloned engine is the engine the deck gives,
// and a batch member's results are bitwise those of its engine alone), so a job gives the
// same block records alone or next to others.
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <functional>
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
  qw_engine* eng = nullptr;   // cloned from the template after admission
  qw_engine* tmpl = nullptr;  // the deck's template engine (owned by the server)
  unsigned long long bytes = 0;
  int member = 0;  // batch member id while it runs in a batch
  std::chrono::steady_clock::time_point t_run;
};

constexpr int kSetupThreads = 4;  // engine clones and walker placement running at once

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
  std::thread sched;
  // Two stepping threads, one for jobs with at most 64 electrons of a spin and one for larger
  // ones: each advances its own batches (one per precision) on their own streams, so a small
  // job's step never waits for a large job's, and the device runs both sets of launches.
  struct Lane {
    std::deque<Job*> ready;                      // set-up done, waiting to join (under mu)
    std::map<int, qw_batch*> batches;            // by precision (this lane's thread only)
    std::map<int, std::map<int, Job*>> members;  // precision -> member id -> job
    std::thread th;
  };
  Lane lanes[2];

  // template engines by (deck path, precision), with the file's size and modification time
  struct Tmpl {
    qw_engine* e = nullptr;
    long long size = -1, mtime = -1;
  };
  std::map<std::pair<std::string, int>, Tmpl> templates;  // admission thread (and destroy)
  // set-up pool: engine clones + walker placement, engine releases
  std::mutex tmu;
  std::condition_variable tcv;
  std::deque<std::function<void()>> tasks;
  std::vector<std::thread> pool;
  bool pool_stop = false;

  void submit_task(std::function<void()> f) {
    std::lock_guard<std::mutex> g(tmu);
    tasks.push_back(std::move(f));
    tcv.notify_one();
  }
  void pool_loop() {
    std::unique_lock<std::mutex> lk(tmu);
    while (true) {
      tcv.wait(lk, [&] { return pool_stop || !tasks.empty(); });
      if (tasks.empty()) return;
      std::function<void()> f = std::move(tasks.front());
      tasks.pop_front();
      lk.unlock();
      f();
      lk.lock();
    }
  }

  // The job's result is final (waiters wake); its memory is still held until release().
  void complete(Job& j, qw_status st, const std::string& err) {
    std::lock_guard<std::mutex> g(mu);
    j.status = st;
    j.err = err;
    j.stats[2] = now_since(t0);
    j.stats[3] = j.stats[2] - j.stats[0];
    j.done = true;
    cv.notify_all();
  }
  // The job's engine has been freed: its bytes and slot go back to admission.
  void release(Job& j) {
    std::lock_guard<std::mutex> g(mu);
    used -= j.bytes;
    j.bytes = 0;
    active--;
    cv.notify_all();
  }
  void finish(Job& j, qw_status st, const std::string& err) {
    complete(j, st, err);
    release(j);
  }
  // Completes the job now and frees its engine on the set-up pool (freeing device memory
  // synchronises the device, which the stepping thread should not wait for).
  void finish_later(Job* j, qw_status st, const std::string& err) {
    qw_engine* e = j->eng;
    j->eng = nullptr;
    complete(*j, st, err);
    submit_task([this, j, e] {
      qw_engine_destroy(e);
      release(*j);
    });
  }

  // The deck's template engine for this precision (parsed on first use, and again when the
  // deck file changed); the error message in *err on failure.
  qw_status get_template(const Job& j, qw_engine** out, std::string* err) {
    struct stat st {};
    const long long size = stat(j.deck.c_str(), &st) == 0 ? (long long)st.st_size : -1;
    const long long mtime = size >= 0 ? (long long)st.st_mtime : -1;
    Tmpl& t = templates[{j.deck, j.desc.precision}];
    if (t.e && (t.size != size || t.mtime != mtime)) {
      qw_engine_destroy(t.e);
      t.e = nullptr;
    }
    if (!t.e) {
      qw_engine_opts eo{j.desc.precision, opts.device, 0, 0};
      qw_status s = qw_engine_create(j.deck.c_str(), &eo, &t.e);
      if (s != QW_OK) {
        *err = qw_last_error(nullptr);
        t.e = nullptr;
        return s;
      }
      t.size = size;
      t.mtime = mtime;
    }
    *out = t.e;
    return QW_OK;
  }

  // Set-up of an admitted job on the pool: its engine, then its walkers; then it is ready.
  void setup_job(Job* j) {
    const double ts = now_since(t0);
    qw_engine_opts eo{j->desc.precision, opts.device, j->desc.seed, 0};
    qw_engine* e = nullptr;
    qw_status st = qw_engine_clone(j->tmpl, &eo, &e);
    if (st != QW_OK) {
      finish(*j, st, qw_last_error(nullptr));
      return;
    }
    j->eng = e;
    int info[6] = {0, 0, 0, 0, 0, 0};
    qw_engine_info(e, info);
    const int lane = std::max(info[0], info[1]) > 64 ? 1 : 0;
    const bool ok = place_walkers(j);
    j->stats[8] += now_since(t0) - ts;
    if (!ok) return;
    std::lock_guard<std::mutex> g(mu);
    lanes[lane].ready.push_back(j);
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

  // Stepping thread of lane L: new jobs join their precision's batch, every batch takes one
  // step, and finished members leave with their results.
  void step_loop(Lane& L) {
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      size_t live = 0;
      for (auto& p : L.members) live += p.second.size();
      cv.wait(lk, [&] { return !L.ready.empty() || live > 0 || (stopping && pending.empty()); });
      std::deque<Job*> take;
      take.swap(L.ready);
      if (take.empty() && live == 0 && stopping) return;
      lk.unlock();
      for (Job* j : take) join(L, j);
      for (auto& p : L.batches) {
        auto& mem = L.members[p.first];
        if (mem.empty()) continue;
        int fin[64], nf = 0;
        qw_status st = qw_batch_step(p.second, fin, &nf);
        if (st != QW_OK) {
          // a failed step fails every job of that batch
          std::string err = qw_batch_error(p.second);
          for (auto& m : mem) finish_later(m.second, st, err);
          mem.clear();
          qw_batch_destroy(p.second);
          p.second = nullptr;
          continue;
        }
        for (int k = 0; k < nf; k++) {
          Job* j = mem[fin[k]];
          mem.erase(fin[k]);
          j->eblock.assign((size_t)j->desc.method.nblock * QW_BLOCK_FIELDS, 0.0);
          double rs[4] = {0, 0, 0, 0};
          qw_status ls = qw_batch_leave(p.second, fin[k], j->eblock.data(), rs);
          j->stats[4] = rs[1];
          j->stats[5] = now_since(t0) - j->stats[1];
          j->stats[6] = j->stats[5] > 0 ? rs[1] / j->stats[5] : 0.0;
          finish_later(j, ls, ls == QW_OK ? std::string() : std::string(qw_batch_error(p.second)));
        }
      }
      for (auto it = L.batches.begin(); it != L.batches.end();)
        if (!it->second) it = L.batches.erase(it);
        else ++it;
      lk.lock();
    }
  }

  // Joins the lane's batch of the job's precision, or runs the job alone when no batch holds it.
  void join(Lane& L, Job* j) {
    const int prec = j->desc.precision;
    qw_batch* b = nullptr;
    auto it = L.batches.find(prec);
    if (it != L.batches.end()) {
      b = it->second;
    } else if (qw_batch_open(opts.device, prec, 0, &b) == QW_OK) {
      L.batches[prec] = b;
    }
    if (b) {
      const auto runs = job_runs(j->desc);
      int mid = 0;
      if (qw_batch_join(b, j->eng, (int)runs.size(), runs.data(), &mid) == QW_OK) {
        L.members[prec][mid] = j;
        j->member = mid;
        j->stats[9] = 1;
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
      if (!j->tmpl) {
        lk.unlock();
        const double ts = now_since(t0);
        qw_engine* t = nullptr;
        std::string err;
        qw_status st = get_template(*j, &t, &err);
        unsigned long long b = 0;
        if (st == QW_OK) st = qw_engine_walker_bytes(t, j->desc.nwalkers, &b);
        lk.lock();
        j->stats[8] = now_since(t0) - ts;
        if (st != QW_OK || b > opts.hbm_budget) {
          // a deck that cannot be read, or a job that can never fit: fail it rather than
          // block the queue
          j->status = st != QW_OK ? st : QW_ERR_NOMEM;
          j->err = st != QW_OK ? err : "job needs more device memory than the server's budget";
          j->stats[7] = (double)b;
          j->stats[2] = now_since(t0);
          j->stats[3] = j->stats[2] - j->stats[0];
          j->done = true;
          pending.pop_front();
          cv.notify_all();
          continue;
        }
        j->tmpl = t;
        j->stats[7] = (double)b;
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
      submit_task([this, j] { setup_job(j); });
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
  for (int i = 0; i < kSetupThreads; i++) s->pool.emplace_back([s] { s->pool_loop(); });
  s->sched = std::thread([s] { s->schedule(); });
  for (auto& L : s->lanes) L.th = std::thread([s, &L] { s->step_loop(L); });
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
  for (auto& L : s->lanes) L.th.join();
  for (auto& t : s->workers) t.join();
  {
    std::lock_guard<std::mutex> g(s->tmu);
    s->pool_stop = true;
    s->tcv.notify_all();
  }
  for (auto& t : s->pool) t.join();
  for (auto& L : s->lanes)
    for (auto& p : L.batches)
      if (p.second) qw_batch_destroy(p.second);
  for (auto& t : s->templates)
    if (t.second.e) qw_engine_destroy(t.second.e);
  delete s;
}
