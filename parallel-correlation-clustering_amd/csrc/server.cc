// server.cc -- continuous-batching request loop over the batch path.
//
// Requests are host copies of graphs.  Worker threads (each drives its own
// batch handles, hence its own HIP stream) take requests from the queue in
// arrival order and admit them into the next batch while the estimated HBM
// footprint of that batch fits the budget left by the batches other workers
// have on the device.  A request too large for the whole budget fails with
// RESOURCE_EXHAUSTED; one that only has to wait for memory stays queued.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "pcc_amd.h"
#include "status.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t id = 0;
  int64_t n = 0, m = 0;
  std::vector<int64_t> off;
  std::vector<uint32_t> nbr;
  std::vector<float> w;
  double resolution = 0;
  double est_bytes = 0;
  Clock::time_point submitted;
  // Result.
  bool done = false;
  int status = PCC_OK;
  std::string error;
  std::vector<uint32_t> ids;
  double objective = 0;
  double latency = 0;
};

// HBM needed to cluster a graph: CSR of every level (refinement keeps them),
// cluster state, round workspace, and the contraction's sort buffers.
double EstimateBytes(int64_t n, int64_t m) {
  return 48.0 * (double)m + 160.0 * (double)n + (1 << 20);
}

}  // namespace

struct pcc_server {
  pcc_config cfg;
  int objective = PCC_OBJECTIVE_CORRELATION;
  pcc_server_options opt;
  int device = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::map<uint64_t, std::shared_ptr<Request>> requests;
  uint64_t next_id = 1;
  double reserved = 0;
  int64_t graphs_in_flight = 0;
  bool stop = false;
  std::vector<std::thread> workers;
  // Stats.
  int64_t completed = 0, failed = 0, batches = 0, max_in_flight = 0, batch_graphs = 0;
  std::vector<double> latencies;
  bool have_first = false;
  Clock::time_point first_submit, last_done;

  void Worker();
};

void pcc_server::Worker() {
  (void)hipSetDevice(device);
  while (true) {
    std::vector<std::shared_ptr<Request>> batch;
    double est = 0;
    {
      std::unique_lock<std::mutex> lock(mu);
      cv_work.wait(lock, [&] { return stop || !queue.empty(); });
      if (stop && queue.empty()) return;
      if (opt.batch_window_seconds > 0 && (int64_t)queue.size() < opt.max_batch_graphs) {
        lock.unlock();
        std::this_thread::sleep_for(std::chrono::duration<double>(opt.batch_window_seconds));
        lock.lock();
        if (queue.empty()) continue;
      }
      // Requests that can never fit fail now.
      while (!queue.empty() && queue.front()->est_bytes > opt.hbm_budget_bytes) {
        auto r = queue.front();
        queue.pop_front();
        r->done = true;
        r->status = PCC_RESOURCE_EXHAUSTED;
        r->error = "request needs more HBM than the server's budget";
        ++failed;
        cv_done.notify_all();
      }
      if (queue.empty()) continue;
      // Admit in arrival order while the batch fits what other batches leave.
      int64_t edges = 0;
      while (!queue.empty() && (int64_t)batch.size() < opt.max_batch_graphs) {
        auto& r = queue.front();
        if (reserved + est + r->est_bytes > opt.hbm_budget_bytes) break;
        if (!batch.empty() && edges + r->m > opt.max_batch_edges) break;
        est += r->est_bytes;
        edges += r->m;
        batch.push_back(r);
        queue.pop_front();
      }
      if (batch.empty()) {
        // Wait for memory held by other workers' batches.
        cv_done.wait(lock, [&] {
          return stop || queue.empty() || reserved + queue.front()->est_bytes <= opt.hbm_budget_bytes;
        });
        continue;
      }
      reserved += est;
      graphs_in_flight += (int64_t)batch.size();
      max_in_flight = std::max(max_in_flight, graphs_in_flight);
      if (!queue.empty()) cv_work.notify_one();
    }
    // Build and run the batch outside the lock.
    const int64_t G = (int64_t)batch.size();
    std::vector<int64_t> counts((size_t)G), off(1, 0);
    std::vector<uint32_t> nbr;
    std::vector<float> w;
    std::vector<double> res((size_t)G);
    int64_t run = 0;
    for (int64_t i = 0; i < G; ++i) {
      const Request& r = *batch[(size_t)i];
      counts[(size_t)i] = r.n;
      res[(size_t)i] = r.resolution;
      for (int64_t v = 0; v < r.n; ++v) off.push_back(run + r.off[(size_t)v + 1]);
      run += r.m;
      nbr.insert(nbr.end(), r.nbr.begin(), r.nbr.end());
      w.insert(w.end(), r.w.begin(), r.w.end());
    }
    const int64_t total_n = (int64_t)off.size() - 1;
    std::vector<uint32_t> ids((size_t)std::max<int64_t>(total_n, 1));
    std::vector<double> obj((size_t)G);
    pcc_graph* b = nullptr;
    int rc = pcc_batch_from_csr(G, counts.data(), off.data(), nbr.data(), w.data(), &b);
    std::string err;
    if (rc == PCC_OK) {
      rc = pcc_batch_cluster(b, &cfg, objective, res.data(), ids.data(), obj.data(), nullptr, nullptr);
      if (rc != PCC_OK) err = pcc_last_error();
      pcc_graph_destroy(b);
    } else {
      err = pcc_last_error();
    }
    const auto now = Clock::now();
    {
      std::lock_guard<std::mutex> lock(mu);
      int64_t pos = 0;
      for (int64_t i = 0; i < G; ++i) {
        Request& r = *batch[(size_t)i];
        r.status = rc;
        r.error = err;
        if (rc == PCC_OK) {
          r.ids.assign(ids.begin() + pos, ids.begin() + pos + r.n);
          r.objective = obj[(size_t)i];
        }
        pos += r.n;
        r.latency = std::chrono::duration<double>(now - r.submitted).count();
        r.done = true;
        r.off.clear();
        r.off.shrink_to_fit();
        r.nbr.clear();
        r.nbr.shrink_to_fit();
        r.w.clear();
        r.w.shrink_to_fit();
        if (rc == PCC_OK) {
          ++completed;
          latencies.push_back(r.latency);
        } else {
          ++failed;
        }
      }
      last_done = now;
      ++batches;
      batch_graphs += G;
      reserved -= est;
      graphs_in_flight -= G;
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }
}

extern "C" {

void pcc_server_options_default(pcc_server_options* opt) {
  std::memset(opt, 0, sizeof(*opt));
  opt->workers = 2;
  opt->max_batch_graphs = 512;
  opt->max_batch_edges = (int64_t)1 << 31;
}

int pcc_server_create(const pcc_config* cfg, int objective, const pcc_server_options* opt,
                      pcc_server** out) {
  pcc::ClearError();
  if (!cfg || !out) return pcc::SetError(PCC_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  auto s = std::make_unique<pcc_server>();
  s->cfg = *cfg;
  s->objective = objective;
  if (opt)
    s->opt = *opt;
  else
    pcc_server_options_default(&s->opt);
  if (s->opt.workers < 1) s->opt.workers = 1;
  if (s->opt.max_batch_graphs < 1) s->opt.max_batch_graphs = 1;
  // Subclustering runs on one graph at a time: each request is its own batch.
  if (cfg->subclustering_method != 0) s->opt.max_batch_graphs = 1;
  if (s->opt.max_batch_edges < 1) s->opt.max_batch_edges = (int64_t)1 << 31;
  if (hipGetDevice(&s->device) != hipSuccess)
    return pcc::SetError(PCC_INTERNAL, "no HIP device");
  if (s->opt.hbm_budget_bytes <= 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return pcc::SetError(PCC_INTERNAL, "hipMemGetInfo failed");
    s->opt.hbm_budget_bytes = 0.9 * (double)fr;
  }
  for (int i = 0; i < s->opt.workers; ++i) s->workers.emplace_back([p = s.get()] { p->Worker(); });
  *out = s.release();
  return PCC_OK;
}

int pcc_server_submit(pcc_server* s, int64_t n, int64_t m, const int64_t* offsets,
                      const uint32_t* nbrs, const float* weights, double resolution,
                      uint64_t* out_request_id) {
  pcc::ClearError();
  if (!s || !offsets || !out_request_id || (m > 0 && !nbrs))
    return pcc::SetError(PCC_INVALID_ARGUMENT, "null argument");
  if (n < 1 || m < 0 || offsets[0] != 0 || offsets[n] != m)
    return pcc::SetError(PCC_INVALID_ARGUMENT, "bad CSR shape");
  for (int64_t p = 0; p < m; ++p)
    if ((int64_t)nbrs[p] >= n) return pcc::SetError(PCC_INVALID_ARGUMENT, "neighbour id out of range");
  auto r = std::make_shared<Request>();
  r->n = n;
  r->m = m;
  r->off.assign(offsets, offsets + n + 1);
  r->nbr.assign(nbrs, nbrs + m);
  if (weights)
    r->w.assign(weights, weights + m);
  else
    r->w.assign((size_t)m, 1.0f);
  r->resolution = resolution;
  r->est_bytes = EstimateBytes(n, m);
  r->submitted = Clock::now();
  {
    std::lock_guard<std::mutex> lock(s->mu);
    r->id = s->next_id++;
    if (!s->have_first) {
      s->have_first = true;
      s->first_submit = r->submitted;
    }
    s->requests[r->id] = r;
    s->queue.push_back(r);
    *out_request_id = r->id;
  }
  s->cv_work.notify_one();
  return PCC_OK;
}

int pcc_server_wait(pcc_server* s, uint64_t request_id, uint32_t* out_cluster_ids,
                    double* out_objective, double* out_latency_seconds) {
  pcc::ClearError();
  if (!s) return pcc::SetError(PCC_INVALID_ARGUMENT, "null server");
  std::shared_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lock(s->mu);
    auto it = s->requests.find(request_id);
    if (it == s->requests.end()) return pcc::SetError(PCC_NOT_FOUND, "unknown request id");
    // The first waiter claims the request: later waits on the same id get
    // NOT_FOUND instead of sharing (and erasing) the entry.
    r = it->second;
    s->requests.erase(it);
    s->cv_done.wait(lock, [&] { return r->done; });
  }
  if (out_latency_seconds) *out_latency_seconds = r->latency;
  if (r->status != PCC_OK) return pcc::SetError(r->status, r->error);
  if (out_cluster_ids) std::copy(r->ids.begin(), r->ids.end(), out_cluster_ids);
  if (out_objective) *out_objective = r->objective;
  return PCC_OK;
}

int pcc_server_get_stats(pcc_server* s, pcc_server_stats* out) {
  if (!s || !out) return pcc::SetError(PCC_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lock(s->mu);
  std::memset(out, 0, sizeof(*out));
  out->completed = s->completed;
  out->failed = s->failed;
  out->batches = s->batches;
  out->max_graphs_in_flight = s->max_in_flight;
  out->mean_batch_graphs = s->batches ? (double)s->batch_graphs / (double)s->batches : 0.0;
  std::vector<double> l = s->latencies;
  if (!l.empty()) {
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) {
      const size_t i = (size_t)std::min<double>((double)l.size() - 1, std::ceil(q * (double)l.size()) - 1);
      return l[i];
    };
    out->p50_latency_seconds = pct(0.50);
    out->p99_latency_seconds = pct(0.99);
    const double span = std::chrono::duration<double>(s->last_done - s->first_submit).count();
    out->graphs_per_second = span > 0 ? (double)s->completed / span : 0.0;
  }
  return PCC_OK;
}

void pcc_server_destroy(pcc_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_done.notify_all();
  for (auto& t : s->workers) t.join();
  delete s;
}

}  // extern "C"
