// driver.cc -- host orchestration of the clustering path and the C ABI.
//
// Restates ParallelCorrelationClusterer::RefineClusters_subroutine
// (clustering/clusterers/parallel-correlation-clusterer.cc:476-629),
// IterateBestMoves (:340-365), BestMovesForVertexSubset (:115-338) and the
// modularity wrapper (clustering/clusterers/parallel-modularity-clusterer.cc:68-111)
// over device-resident CSR and cluster state.  One HIP stream per graph
// handle; the host synchronizes once per best-move round to read the next
// frontier's size (IterateBestMoves' loop condition) and per-bin counts.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "kernels.h"
#include "pcc_amd.h"
#include "status.h"

namespace pcc {

// ------------------------------------------------------------- allocator --

namespace {
size_t RoundSize(size_t b) {
  size_t r = 256;
  while (r < b) r <<= 1;
  return r;
}
}  // namespace

DeviceAllocator::~DeviceAllocator() { ReleaseAll(); }

void* DeviceAllocator::Alloc(size_t bytes) {
  const size_t sz = RoundSize(bytes);
  auto it = free_.find(sz);
  void* p = nullptr;
  if (it != free_.end() && !it->second.empty()) {
    p = it->second.back();
    it->second.pop_back();
  } else {
    if (hipMalloc(&p, sz) != hipSuccess) {
      (void)hipGetLastError();
      // Return cached blocks to the device and try once more.
      for (auto& kv : free_)
        for (void* q : kv.second) (void)hipFree(q);
      free_.clear();
      if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        throw AllocError{bytes};
      }
    }
  }
  live_[p] = sz;
  in_use_ += sz;
  return p;
}

void DeviceAllocator::Free(void* p) {
  auto it = live_.find(p);
  if (it == live_.end()) return;
  in_use_ -= it->second;
  free_[it->second].push_back(p);
  live_.erase(it);
}

void DeviceAllocator::ReleaseAll() {
  for (auto& kv : free_)
    for (void* q : kv.second) (void)hipFree(q);
  free_.clear();
  for (auto& kv : live_) (void)hipFree(kv.first);
  live_.clear();
  in_use_ = 0;
}

}  // namespace pcc

using namespace pcc;

struct pcc_graph {
  int device = 0;
  hipStream_t stream = nullptr;
  DeviceAllocator alloc;
  // Import staging (GbbsGraph, clustering/gbbs-graph.cc:32-76).
  std::mutex mu;
  std::vector<std::vector<std::pair<uint32_t, float>>> adj;
  bool finished = false;
  DevCsr csr;
  // Options.
  bool timing = false;
  bool force_ordered = false;
  BinThresholds th;
  ~pcc_graph() {
    csr = DevCsr();
    alloc.ReleaseAll();
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int Guard(const char* where, const std::function<int()>& body) {
  try {
    return body();
  } catch (const HipError& e) {
    return SetError(PCC_INTERNAL, std::string(where) + ": HIP error " + hipGetErrorString(e.err) +
                                      " at " + e.file + ":" + std::to_string(e.line) + " (" +
                                      e.what + ")");
  } catch (const AllocError& e) {
    return SetError(PCC_RESOURCE_EXHAUSTED,
                    std::string(where) + ": device allocation of " + std::to_string(e.bytes) +
                        " bytes failed (free HBM exhausted)");
  } catch (const std::bad_alloc&) {
    return SetError(PCC_RESOURCE_EXHAUSTED, std::string(where) + ": host allocation failed");
  }
}

int NewHandle(pcc_graph** out) {
  auto h = std::make_unique<pcc_graph>();
  PCC_HIP_CHECK(hipGetDevice(&h->device));
  PCC_HIP_CHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  *out = h.release();
  return PCC_OK;
}

__global__ void k_validate(int64_t n, int64_t m, const int64_t* off, const uint32_t* nbr,
                           unsigned long long* bad) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x)
    if ((int64_t)nbr[i] >= n) atomicAdd(bad, 1ull);
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    if (off[v + 1] < off[v]) atomicAdd(bad, 1ull);
}

__global__ void k_fill_ones(float* w, int64_t m) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x)
    w[i] = 1.0f;
}

int UploadCsr(pcc_graph* h, int64_t n, int64_t m, const int64_t* off, const uint32_t* nbrs,
              const float* w) {
  if (n < 0 || m < 0 || n >= (int64_t)0xFFFFFFFFll)
    return SetError(PCC_INVALID_ARGUMENT, "vertex count must lie in [0, 2^32 - 1)");
  if (!off || (m > 0 && !nbrs)) return SetError(PCC_INVALID_ARGUMENT, "null CSR array");
  if (off[0] != 0 || off[n] != m)
    return SetError(PCC_INVALID_ARGUMENT, "offsets must start at 0 and end at m");
  hipStream_t s = h->stream;
  h->csr.n = n;
  h->csr.m = m;
  h->csr.off.Reset(&h->alloc, n + 1);
  h->csr.nbr.Reset(&h->alloc, m > 0 ? m : 1);
  h->csr.w.Reset(&h->alloc, m > 0 ? m : 1);
  PCC_HIP_CHECK(hipMemcpyAsync(h->csr.off.get(), off, sizeof(int64_t) * (size_t)(n + 1),
                               hipMemcpyHostToDevice, s));
  if (m > 0) {
    PCC_HIP_CHECK(hipMemcpyAsync(h->csr.nbr.get(), nbrs, sizeof(uint32_t) * (size_t)m,
                                 hipMemcpyHostToDevice, s));
    if (w) {
      PCC_HIP_CHECK(hipMemcpyAsync(h->csr.w.get(), w, sizeof(float) * (size_t)m,
                                   hipMemcpyHostToDevice, s));
    } else {
      hipLaunchKernelGGL(k_fill_ones, dim3(1024), dim3(256), 0, s, h->csr.w.get(), m);
    }
  }
  DevBuf<unsigned long long> bad(&h->alloc, 1);
  PCC_HIP_CHECK(hipMemsetAsync(bad.get(), 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL(k_validate, dim3(2048), dim3(256), 0, s, n, m, h->csr.off.get(),
                     h->csr.nbr.get(), bad.get());
  unsigned long long nbad = 0;
  PCC_HIP_CHECK(hipMemcpyAsync(&nbad, bad.get(), sizeof(nbad), hipMemcpyDeviceToHost, s));
  PCC_HIP_CHECK(hipStreamSynchronize(s));
  if (nbad) {
    h->csr = DevCsr();
    return SetError(PCC_INVALID_ARGUMENT,
                    "CSR has out-of-range neighbour ids or decreasing offsets");
  }
  h->finished = true;
  return PCC_OK;
}

// ------------------------------------------------------------ the engine --

struct Level {
  DevCsr own;    // contracted graphs (level > 0)
  CsrView g;     // view of the graph of this level
  DevBuf<uint32_t> ids, sizes;
  DevBuf<double> nw, K;
  StateView State() const {
    StateView s;
    s.ids = ids.get();
    s.sizes = sizes.get();
    s.nw = nw.get();
    s.K = K.get();
    return s;
  }
};

struct Timer {
  hipEvent_t a = nullptr, b = nullptr;
};

class Engine {
 public:
  Engine(pcc_graph* h, const pcc_config& cfg, pcc_stats* stats)
      : h_(h), cfg_(cfg), s_(h->stream), al_(&h->alloc), stats_(stats) {}

  ~Engine() {
    for (auto& t : pending_) {
      (void)hipEventDestroy(t.first.a);
      (void)hipEventDestroy(t.first.b);
    }
  }

  // IterateBestMoves (PCC.cc:340-365).  Returns "moved".
  bool IterateBestMoves(const CsrView& g, const StateView& st, int64_t inner) {
    const int64_t n = g.n;
    Workspace ws(al_, n, s_);
    // Bin capacities and max degree of this level.
    DevBuf<unsigned long long> caps(al_, kBins);
    PCC_HIP_CHECK(hipMemsetAsync(caps.get(), 0, sizeof(unsigned long long) * kBins, s_));
    LaunchBinCapacity(g, h_->th, caps.get(), s_);
    unsigned long long hcaps[kBins];
    PCC_HIP_CHECK(hipMemcpyAsync(hcaps, caps.get(), sizeof(hcaps), hipMemcpyDeviceToHost, s_));
    PCC_HIP_CHECK(hipStreamSynchronize(s_));
    BinLayout layout;
    int64_t run = 0;
    for (int b = 0; b < kBins; ++b) {
      layout.base[b] = run;
      run += (int64_t)hcaps[b];
    }
    HashScratch hs;
    DevBuf<uint32_t> hkeys;
    DevBuf<double> hvals;
    const bool need_hash = hcaps[5] + hcaps[6] + hcaps[7] > 0;
    if (need_hash) {
      // Regions sized for the largest hashed degree on this level.
      const int64_t maxdeg = MaxDegree(g);
      uint32_t cap = 64;
      while (cap < (uint64_t)(2 * maxdeg + 2)) cap <<= 1;
      hs.cap = cap;
      hs.regions = cap >= (1u << 22) ? 16 : (cap >= (1u << 20) ? 32 : 64);
      hkeys.Reset(al_, (int64_t)hs.cap * hs.regions);
      hvals.Reset(al_, (int64_t)hs.cap * hs.regions);
      hs.keys = hkeys.get();
      hs.vals = hvals.get();
    }
    MoveParams mp;
    mp.res = cfg_.resolution;
    mp.offset = cfg_.edge_weight_offset;
    mp.n = (uint32_t)n;
    mp.force_ordered = h_->force_ordered ? 1 : 0;

    // Round 0 frontier: every vertex (PCC.cc:346-352).
    int64_t counts[kBins], te[kBins];
    ResetCounters(ws.rc.get(), true);
    LaunchBinFrontier(g, nullptr, h_->th, layout, ws.list.get(), ws.rc.get(), s_);
    ReadCounters(ws.rc.get(), counts, te, nullptr);

    bool moved = false, local_moved = true;
    int64_t it = 0;
    for (it = 0; it < inner && local_moved; ++it) {
      int64_t fsize = 0;
      RoundStep(g, st, mp, layout, hs, ws, counts, te, it, &fsize);
      local_moved = fsize != 0;
      moved |= local_moved;
    }
    return moved;
  }

  int64_t MaxDegree(const CsrView& g);

  struct Workspace {
    DevBuf<uint32_t> list, fb_list, moves;
    DevBuf<uint8_t> moved, flags, next, new_flag, frontier;
    DevBuf<RoundCounters> rc;
    Workspace(DeviceAllocator* al, int64_t n, hipStream_t s) {
      const int64_t nn = n > 0 ? n : 1;
      list.Reset(al, nn);
      fb_list.Reset(al, nn);
      moves.Reset(al, nn);
      moved.Reset(al, nn);
      flags.Reset(al, nn);
      next.Reset(al, nn);
      new_flag.Reset(al, nn);
      rc.Reset(al, 1);
      // new_flag is kept all-zero between rounds by the new-cluster assignment.
      PCC_HIP_CHECK(hipMemsetAsync(new_flag.get(), 0, (size_t)nn, s));
      PCC_HIP_CHECK(hipMemsetAsync(rc.get(), 0, sizeof(RoundCounters), s));
    }
  };

  void ResetCounters(RoundCounters* rc, bool all) {
    if (all) {
      PCC_HIP_CHECK(hipMemsetAsync(rc, 0, sizeof(RoundCounters), s_));
    } else {
      PCC_HIP_CHECK(hipMemsetAsync(rc, 0, offsetof(RoundCounters, fb_count), s_));
    }
  }

  void ReadCounters(RoundCounters* rc, int64_t counts[kBins], int64_t te[kBins],
                    RoundCounters* full) {
    RoundCounters hrc;
    PCC_HIP_CHECK(hipMemcpyAsync(&hrc, rc, sizeof(hrc), hipMemcpyDeviceToHost, s_));
    PCC_HIP_CHECK(hipStreamSynchronize(s_));
    for (int b = 0; b < kBins; ++b) {
      counts[b] = (int64_t)hrc.cnt[b];
      te[b] = (int64_t)hrc.te[b];
    }
    if (full) *full = hrc;
    CollectTimers();
  }

  Timer StartTimer() {
    Timer t;
    if (!h_->timing) return t;
    PCC_HIP_CHECK(hipEventCreate(&t.a));
    PCC_HIP_CHECK(hipEventCreate(&t.b));
    PCC_HIP_CHECK(hipEventRecord(t.a, s_));
    return t;
  }
  void StopTimer(Timer t, int kind) {
    if (!t.a) return;
    PCC_HIP_CHECK(hipEventRecord(t.b, s_));
    pending_.emplace_back(t, kind);
  }
  // kind: 0..7 best-move bin, 8 ordered fallback, 9 apply+frontier, 10 contraction
  void CollectTimers() {
    for (auto& p : pending_) {
      float ms = 0;
      PCC_HIP_CHECK(hipEventSynchronize(p.first.b));
      PCC_HIP_CHECK(hipEventElapsedTime(&ms, p.first.a, p.first.b));
      kind_seconds_[p.second] += ms * 1e-3;
      kind_launches_[p.second] += 1;
      (void)hipEventDestroy(p.first.a);
      (void)hipEventDestroy(p.first.b);
    }
    pending_.clear();
  }

  // One BestMovesForVertexSubset round (PCC.cc:115-338).
  void RoundStep(const CsrView& g, const StateView& st, const MoveParams& mp,
                 const BinLayout& layout, const HashScratch& hs, Workspace& ws,
                 int64_t counts[kBins], int64_t te[kBins], int64_t it, int64_t* fsize) {
    const int64_t n = g.n;
    const bool async = cfg_.async != 0;
    rounds_ += 1;
    int64_t F = 0, TE = 0;
    for (int b = 0; b < kBins; ++b) {
      F += counts[b];
      TE += te[b];
      bin_edges_[b] += te[b];
      bin_vertices_[b] += counts[b];
    }
    edges_scanned_ += TE;
    frontier_vertices_ += F;
    bestmove_bytes_ += 12.0 * (double)TE + 40.0 * (double)F;

    MoveOut out;
    out.moves = ws.moves.get();
    out.moved = ws.moved.get();
    out.moved_clusters = ws.flags.get();
    if (async) PCC_HIP_CHECK(hipMemsetAsync(ws.flags.get(), 0, (size_t)n, s_));
    PCC_HIP_CHECK(hipMemsetAsync(&ws.rc.get()->fb_count, 0, 2 * sizeof(unsigned long long), s_));
    const uint64_t salt =
        (async && cfg_.permute) ? (cfg_.permute_seed * 0x9E3779B97F4A7C15ull + (uint64_t)(rounds_) + 1) : 0;
    Timer all = StartTimer();
    for (int b = 0; b < kBins; ++b) {
      if (counts[b] == 0) continue;
      Timer tb = StartTimer();
      LaunchBestMoveBin(b, counts[b], async, g, st, mp, ws.list.get() + layout.base[b], out,
                        ws.fb_list.get(), ws.rc.get(), hs, salt, s_);
      StopTimer(tb, b);
    }
    if (!async && counts[5] + counts[6] + counts[7] > 0) {
      Timer tf = StartTimer();
      LaunchBestMoveOrdered(g, st, mp, ws.fb_list.get(), ws.rc.get(), out, hs, s_);
      StopTimer(tf, 8);
    }
    StopTimer(all, 11);

    Timer ta = StartTimer();
    if (!async) {
      // MoveNodesToCluster (PCCI.cc:256-369).
      PCC_HIP_CHECK(hipMemsetAsync(ws.flags.get(), 0, (size_t)n, s_));
      LaunchApply(counts, layout, ws.list.get(), st, (uint32_t)n, ws.moves.get(), ws.flags.get(),
                  ws.new_flag.get(), ws.rc.get(), s_);
      RunNewClusterAssign(al_, st, (uint32_t)n, ws.new_flag.get(), ws.flags.get(), ws.rc.get(), s_);
    }
    if (cfg_.move_method == PCC_ALL_MOVE) {  // PCC.cc:299-307
      StopTimer(ta, 9);
      *fsize = n;
      RoundCounters hrc;
      int64_t c2[kBins], t2[kBins];
      ReadCounters(ws.rc.get(), c2, t2, &hrc);
      fallback_ += (int64_t)hrc.fb_count;
      return;  // counts / te unchanged: the frontier stays all vertices
    }
    // Next frontier = neighbours of the seeds (PCC.cc:309-331).
    PCC_HIP_CHECK(hipMemsetAsync(ws.next.get(), 0, (size_t)n, s_));
    if (cfg_.move_method == PCC_NBHR_CLUSTER_MOVE) {
      LaunchExpandClusters(g, st.ids, ws.flags.get(), ws.next.get(), s_);
    } else {
      LaunchExpandMoved(counts, layout, ws.list.get(), g, ws.moved.get(), ws.next.get(), s_);
    }
    ResetCounters(ws.rc.get(), false);
    LaunchBinFrontier(g, ws.next.get(), h_->th, layout, ws.list.get(), ws.rc.get(), s_);
    StopTimer(ta, 9);
    RoundCounters hrc;
    ReadCounters(ws.rc.get(), counts, te, &hrc);
    fallback_ += (int64_t)hrc.fb_count;
    int64_t f = 0;
    for (int b = 0; b < kBins; ++b) f += counts[b];
    *fsize = f;
  }

  // RefineClusters_subroutine (PCC.cc:476-629).  nw0: device node weights of
  // the input graph (taken over); init: device initial ids or null.
  int Run(DevBuf<double> nw0, const uint32_t* init, uint32_t* host_out) {
    const CsrView g0 = h_->csr.View();
    const int64_t n = g0.n;
    int64_t num_iterations = 0, num_inner = 0;
    if (cfg_.clustering_moves_method == PCC_LOUVAIN) {  // PCC.cc:493-500
      num_iterations = cfg_.louvain_num_iterations > 0 ? cfg_.louvain_num_iterations : 10;
      num_inner = cfg_.louvain_num_inner_iterations > 0 ? cfg_.louvain_num_inner_iterations : 10;
    } else if (cfg_.clustering_moves_method == PCC_DEFAULT_CLUSTER_MOVES) {
      return SetError(PCC_UNIMPLEMENTED,
                      "DEFAULT_CLUSTER_MOVES (whole-cluster moves) is not on this path; use LOUVAIN");
    } else {
      return SetError(PCC_UNIMPLEMENTED,
                      "Correlation clustering moves must be DEFAULT_CLUSTER_MOVES or LOUVAIN.");
    }
    if (cfg_.subclustering_method != 0)
      return SetError(PCC_UNIMPLEMENTED, "subclustering_method other than NONE is not on this path");

    hipEvent_t e0, e1;
    PCC_HIP_CHECK(hipEventCreate(&e0));
    PCC_HIP_CHECK(hipEventCreate(&e1));
    PCC_HIP_CHECK(hipEventRecord(e0, s_));
    const auto w0 = std::chrono::steady_clock::now();

    std::vector<std::unique_ptr<Level>> levels;
    levels.emplace_back(new Level());
    {
      Level& L = *levels.back();
      L.g = g0;
      L.nw = std::move(nw0);
      L.ids.Reset(al_, n > 0 ? n : 1);
      L.sizes.Reset(al_, n > 0 ? n : 1);
      L.K.Reset(al_, n > 0 ? n : 1);
      if (init)
        LaunchSetClustering(n, init, L.State(), s_);
      else
        LaunchSingletons(n, L.State(), s_);
    }
    DevBuf<uint32_t> cluster_ids(al_, n > 0 ? n : 1), flat(al_, n > 0 ? n : 1);
    IotaU32(cluster_ids.get(), n);
    int64_t iter = 0;
    for (iter = 0; iter < num_iterations; ++iter) {
      Level& L = *levels[(size_t)iter];
      const bool moved = IterateBestMoves(L.g, L.State(), num_inner);
      if (!moved) {  // PCC.cc:543-547
        iter--;
        break;
      }
      if (!cfg_.refine) {  // PCC.cc:550
        LaunchFlatten(n, cluster_ids.get(), L.ids.get(), flat.get(), s_);
        std::swap(cluster_ids, flat);
      }
      if (iter == num_iterations - 1) break;  // PCC.cc:556
      // CompressGraph (PCC.cc:569-571) and singletons with aggregated weights
      // (PCC.cc:583-585).
      Timer tc = StartTimer();
      auto next = std::make_unique<Level>();
      CompressGraph(al_, L.g, L.ids.get(), L.nw.get(), &next->own, &next->nw, s_);
      next->g = next->own.View();
      const int64_t n2 = next->g.n;
      next->ids.Reset(al_, n2);
      next->sizes.Reset(al_, n2);
      next->K.Reset(al_, n2);
      LaunchSingletons(n2, next->State(), s_);
      StopTimer(tc, 10);
      if (!cfg_.refine && iter > 0) {
        // Only the current level is needed without refinement.
        L.own = DevCsr();
        L.ids.Release();
        L.sizes.Release();
        L.K.Release();
        L.nw.Release();
      }
      levels.push_back(std::move(next));
    }
    if (cfg_.refine && iter > 0) {  // PCC.cc:589-607
      for (int64_t i = iter - 1; i >= 0; --i) {
        Level& L = *levels[(size_t)i];
        const int64_t ni = L.g.n;
        DevBuf<uint32_t> fl(al_, ni > 0 ? ni : 1);
        LaunchFlatten(ni, L.ids.get(), levels[(size_t)i + 1]->ids.get(), fl.get(), s_);
        ResetClusteringFromIds(al_, ni, fl.get(), L.State(), s_);
        IterateBestMoves(L.g, L.State(), num_inner);
      }
      PCC_HIP_CHECK(hipMemcpyAsync(cluster_ids.get(), levels[0]->ids.get(),
                                   sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, s_));
    }
    PCC_HIP_CHECK(hipEventRecord(e1, s_));
    PCC_HIP_CHECK(hipEventSynchronize(e1));
    const auto w1 = std::chrono::steady_clock::now();
    float ms = 0;
    PCC_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    CollectTimers();
    if (n > 0)
      PCC_HIP_CHECK(hipMemcpyAsync(host_out, cluster_ids.get(), sizeof(uint32_t) * (size_t)n,
                                   hipMemcpyDeviceToHost, s_));
    PCC_HIP_CHECK(hipStreamSynchronize(s_));
    if (stats_) {
      stats_->levels = iter + 1;
      stats_->rounds = rounds_;
      stats_->edges_scanned = edges_scanned_;
      stats_->frontier_vertices = frontier_vertices_;
      stats_->fallback_vertices = fallback_;
      stats_->cluster_seconds = std::max((double)ms * 1e-3,
                                         0.0 * std::chrono::duration<double>(w1 - w0).count());
      stats_->bestmove_bytes = bestmove_bytes_;
      stats_->bestmove_seconds = kind_seconds_[11];
      stats_->bestmove_launches = 0;
      for (int b = 0; b <= 8; ++b) stats_->bestmove_launches += kind_launches_[b];
      stats_->apply_seconds = kind_seconds_[9];
      stats_->contract_seconds = kind_seconds_[10];
      for (int b = 0; b < kBins; ++b) {
        stats_->bin_seconds[b] = kind_seconds_[b];
        stats_->bin_edges[b] = bin_edges_[b];
        stats_->bin_vertices[b] = bin_vertices_[b];
        stats_->bin_launches[b] = kind_launches_[b];
      }
      stats_->ordered_seconds = kind_seconds_[8];
    }
    return PCC_OK;
  }

  void IotaU32(uint32_t* p, int64_t n);

 private:
  pcc_graph* h_;
  pcc_config cfg_;
  hipStream_t s_;
  DeviceAllocator* al_;
  pcc_stats* stats_;
  std::vector<std::pair<Timer, int>> pending_;
  double kind_seconds_[16] = {0};
  int64_t kind_launches_[16] = {0};
  int64_t bin_edges_[kBins] = {0};
  int64_t bin_vertices_[kBins] = {0};
  int64_t rounds_ = 0, edges_scanned_ = 0, frontier_vertices_ = 0, fallback_ = 0;
  double bestmove_bytes_ = 0;
};

__global__ void k_iota(uint32_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (uint32_t)i;
}

__global__ void k_max_degree(int64_t n, const int64_t* off, unsigned long long* out) {
  unsigned long long m = 0;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    m = max(m, (unsigned long long)(off[v + 1] - off[v]));
  for (int s = 32; s > 0; s >>= 1) m = max(m, (unsigned long long)__shfl_xor(m, s));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

void Engine::IotaU32(uint32_t* p, int64_t n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_iota, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 8192)), dim3(256),
                     0, s_, p, n);
  PCC_HIP_CHECK(hipGetLastError());
}

int64_t Engine::MaxDegree(const CsrView& g) {
  DevBuf<unsigned long long> d(al_, 1);
  PCC_HIP_CHECK(hipMemsetAsync(d.get(), 0, sizeof(unsigned long long), s_));
  if (g.n > 0)
    hipLaunchKernelGGL(k_max_degree, dim3((unsigned)std::min<int64_t>((g.n + 255) / 256, 4096)),
                       dim3(256), 0, s_, g.n, g.off, d.get());
  unsigned long long hd = 0;
  PCC_HIP_CHECK(hipMemcpyAsync(&hd, d.get(), sizeof(hd), hipMemcpyDeviceToHost, s_));
  PCC_HIP_CHECK(hipStreamSynchronize(s_));
  return (int64_t)hd;
}

int CheckGraph(pcc_graph* g) {
  if (!g) return SetError(PCC_INVALID_ARGUMENT, "null graph");
  if (!g->finished) return SetError(PCC_INVALID_ARGUMENT, "graph import not finished");
  PCC_HIP_CHECK(hipSetDevice(g->device));
  return PCC_OK;
}

// Node weights of a Cluster / RefineClusters call on device.
DevBuf<double> NodeWeights(pcc_graph* h, int objective, const double* host_nw) {
  const int64_t n = h->csr.n;
  DevBuf<double> nw(&h->alloc, n > 0 ? n : 1);
  if (host_nw) {
    if (n > 0)
      PCC_HIP_CHECK(hipMemcpyAsync(nw.get(), host_nw, sizeof(double) * (size_t)n,
                                   hipMemcpyHostToDevice, h->stream));
  } else if (objective == PCC_OBJECTIVE_MODULARITY) {
    LaunchDegrees(h->csr.View(), nw.get(), h->stream);  // PGU.cc:141-153
  } else {
    std::vector<double> ones((size_t)(n > 0 ? n : 1), 1.0);
    PCC_HIP_CHECK(hipMemcpyAsync(nw.get(), ones.data(), sizeof(double) * ones.size(),
                                 hipMemcpyHostToDevice, h->stream));
    PCC_HIP_CHECK(hipStreamSynchronize(h->stream));
  }
  return nw;
}

int ClusterImpl(pcc_graph* g, const pcc_config* cfg_in, int objective, const double* host_nw,
                const uint32_t* host_init, uint32_t* out_ids, pcc_stats* stats) {
  PCC_RETURN_IF_ERROR(CheckGraph(g));
  if (!cfg_in || !out_ids) return SetError(PCC_INVALID_ARGUMENT, "null argument");
  if (objective != PCC_OBJECTIVE_CORRELATION && objective != PCC_OBJECTIVE_MODULARITY)
    return SetError(PCC_INVALID_ARGUMENT, "unknown objective");
  pcc_config cfg = *cfg_in;
  const double gamma = cfg.resolution;
  const int64_t n = g->csr.n;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (objective == PCC_OBJECTIVE_MODULARITY) {
    // ComputeModularityConfig + PMC.cc:76-79: k_v = deg(v), res' = gamma / sum deg.
    cfg.resolution = gamma / (double)g->csr.m;
    cfg.edge_weight_offset = 0;
  }
  DevBuf<uint32_t> init;
  if (host_init) {
    for (int64_t v = 0; v < n; ++v)
      if (host_init[v] >= (uint64_t)n)
        return SetError(PCC_INVALID_ARGUMENT, "initial cluster index out of range [0, n)");
    init.Reset(&g->alloc, n > 0 ? n : 1);
    if (n > 0)
      PCC_HIP_CHECK(hipMemcpyAsync(init.get(), host_init, sizeof(uint32_t) * (size_t)n,
                                   hipMemcpyHostToDevice, g->stream));
  }
  DevBuf<double> nw = NodeWeights(g, objective, host_nw);
  DevBuf<double> nw_keep(&g->alloc, n > 0 ? n : 1);
  PCC_HIP_CHECK(hipMemcpyAsync(nw_keep.get(), nw.get(), sizeof(double) * (size_t)(n > 0 ? n : 1),
                               hipMemcpyDeviceToDevice, g->stream));
  {
    Engine e(g, cfg, stats);
    PCC_RETURN_IF_ERROR(e.Run(std::move(nw), host_init ? init.get() : nullptr, out_ids));
  }
  if (stats) {
    // Final objective on the input graph (PCC.cc:619-624) and modularity
    // (PMC.cc:91-93); outside the timed span like the reference.
    DevBuf<uint32_t> ids(&g->alloc, n > 0 ? n : 1);
    DevBuf<double> res(&g->alloc, 2);
    if (n > 0)
      PCC_HIP_CHECK(hipMemcpyAsync(ids.get(), out_ids, sizeof(uint32_t) * (size_t)n,
                                   hipMemcpyHostToDevice, g->stream));
    ComputeObjective(&g->alloc, g->csr.View(), ids.get(), nw_keep.get(), cfg.resolution,
                     cfg.edge_weight_offset, res.get(), g->stream);
    if (objective == PCC_OBJECTIVE_MODULARITY)
      ComputeModularity(&g->alloc, g->csr.View(), ids.get(), gamma, res.get() + 1, g->stream);
    double hres[2] = {0, 0};
    PCC_HIP_CHECK(hipMemcpyAsync(hres, res.get(), sizeof(hres), hipMemcpyDeviceToHost, g->stream));
    PCC_HIP_CHECK(hipStreamSynchronize(g->stream));
    stats->objective = hres[0];
    stats->modularity = objective == PCC_OBJECTIVE_MODULARITY ? hres[1] : 0.0;
    std::vector<uint32_t> sorted(out_ids, out_ids + n);
    std::sort(sorted.begin(), sorted.end());
    stats->num_clusters = (int64_t)(std::unique(sorted.begin(), sorted.end()) - sorted.begin());
  }
  return PCC_OK;
}

// Uploads a host ClusteringHelper state.
struct DevState {
  DevBuf<uint32_t> ids, sizes;
  DevBuf<double> nw, K;
  DevState(pcc_graph* h, const uint32_t* ids_h, const uint32_t* sizes_h, const double* nw_h,
           const double* K_h) {
    const int64_t n = h->csr.n > 0 ? h->csr.n : 1;
    DeviceAllocator* a = &h->alloc;
    ids.Reset(a, n);
    sizes.Reset(a, n);
    nw.Reset(a, n);
    K.Reset(a, n);
    const size_t nn = (size_t)h->csr.n;
    if (nn) {
      PCC_HIP_CHECK(hipMemcpyAsync(ids.get(), ids_h, 4 * nn, hipMemcpyHostToDevice, h->stream));
      PCC_HIP_CHECK(hipMemcpyAsync(sizes.get(), sizes_h, 4 * nn, hipMemcpyHostToDevice, h->stream));
      PCC_HIP_CHECK(hipMemcpyAsync(nw.get(), nw_h, 8 * nn, hipMemcpyHostToDevice, h->stream));
      PCC_HIP_CHECK(hipMemcpyAsync(K.get(), K_h, 8 * nn, hipMemcpyHostToDevice, h->stream));
    }
  }
  StateView View() const {
    StateView s;
    s.ids = ids.get();
    s.sizes = sizes.get();
    s.nw = nw.get();
    s.K = K.get();
    return s;
  }
};

int CheckState(pcc_graph* g, const uint32_t* ids, const uint32_t* sizes) {
  const int64_t n = g->csr.n;
  for (int64_t v = 0; v < n; ++v)
    if (ids[v] >= (uint64_t)n) return SetError(PCC_INVALID_ARGUMENT, "cluster id out of range");
  (void)sizes;
  return PCC_OK;
}

}  // namespace

namespace pcc {
// Used by the one-phase entry points below.
class EngineAccess;
}  // namespace pcc

extern "C" {

int pcc_graph_new(pcc_graph** out) {
  ClearError();
  if (!out) return SetError(PCC_INVALID_ARGUMENT, "null out");
  return Guard("pcc_graph_new", [&] { return NewHandle(out); });
}

int pcc_graph_import(pcc_graph* g, uint32_t id, double node_weight, const uint32_t* nbrs,
                     const double* weights, uint64_t degree) {
  (void)node_weight;  // GbbsGraph ignores node weights (gbbs-graph.h:34)
  if (!g) return SetError(PCC_INVALID_ARGUMENT, "null graph");
  if (id == 0xFFFFFFFFu) return SetError(PCC_INVALID_ARGUMENT, "node id out of range");
  if (degree > 0 && !nbrs) return SetError(PCC_INVALID_ARGUMENT, "null neighbour list");
  std::vector<std::pair<uint32_t, float>> edges((size_t)degree);
  for (uint64_t i = 0; i < degree; ++i)
    edges[(size_t)i] = {nbrs[i], weights ? (float)weights[i] : 1.0f};
  std::lock_guard<std::mutex> lock(g->mu);
  if (g->finished) return SetError(PCC_INVALID_ARGUMENT, "Import after FinishImport");
  if (g->adj.size() <= id) g->adj.resize((size_t)id + 1);
  g->adj[id] = std::move(edges);
  return PCC_OK;
}

int pcc_graph_finish_import(pcc_graph* g) {
  ClearError();
  if (!g) return SetError(PCC_INVALID_ARGUMENT, "null graph");
  return Guard("pcc_graph_finish_import", [&]() -> int {
    std::lock_guard<std::mutex> lock(g->mu);
    if (g->finished) return SetError(PCC_INVALID_ARGUMENT, "FinishImport called twice");
    // n = max(imported id, max neighbour id) + 1 (gbbs-graph.cc:55-68).
    int64_t n = (int64_t)g->adj.size();
    for (const auto& a : g->adj)
      for (const auto& e : a) n = std::max<int64_t>(n, (int64_t)e.first + 1);
    std::vector<int64_t> off((size_t)n + 1, 0);
    for (int64_t v = 0; v < (int64_t)g->adj.size(); ++v)
      off[(size_t)v + 1] = (int64_t)g->adj[(size_t)v].size();
    for (int64_t v = 0; v < n; ++v) off[(size_t)v + 1] += off[(size_t)v];
    const int64_t m = off[(size_t)n];
    std::vector<uint32_t> nb((size_t)std::max<int64_t>(m, 1));
    std::vector<float> w((size_t)std::max<int64_t>(m, 1));
    for (int64_t v = 0; v < (int64_t)g->adj.size(); ++v) {
      int64_t p = off[(size_t)v];
      for (const auto& e : g->adj[(size_t)v]) {
        nb[(size_t)p] = e.first;
        w[(size_t)p] = e.second;
        ++p;
      }
    }
    g->adj.clear();
    PCC_HIP_CHECK(hipSetDevice(g->device));
    return UploadCsr(g, n, m, off.data(), nb.data(), w.data());
  });
}

int pcc_graph_from_csr(int64_t n, int64_t m, const int64_t* offsets, const uint32_t* nbrs,
                       const float* weights, pcc_graph** out) {
  ClearError();
  if (!out) return SetError(PCC_INVALID_ARGUMENT, "null out");
  *out = nullptr;
  return Guard("pcc_graph_from_csr", [&]() -> int {
    pcc_graph* h = nullptr;
    PCC_RETURN_IF_ERROR(NewHandle(&h));
    const int rc = UploadCsr(h, n, m, offsets, nbrs, weights);
    if (rc != PCC_OK) {
      delete h;
      return rc;
    }
    *out = h;
    return PCC_OK;
  });
}

int64_t pcc_graph_num_nodes(const pcc_graph* g) { return g ? g->csr.n : -1; }
int64_t pcc_graph_num_edges(const pcc_graph* g) { return g ? g->csr.m : -1; }

int pcc_graph_copy_csr(const pcc_graph* gc, int64_t* offsets, uint32_t* nbrs, float* weights) {
  ClearError();
  pcc_graph* g = const_cast<pcc_graph*>(gc);
  return Guard("pcc_graph_copy_csr", [&]() -> int {
    PCC_RETURN_IF_ERROR(CheckGraph(g));
    const int64_t n = g->csr.n, m = g->csr.m;
    if (offsets)
      PCC_HIP_CHECK(hipMemcpyAsync(offsets, g->csr.off.get(), 8 * (size_t)(n + 1),
                                   hipMemcpyDeviceToHost, g->stream));
    if (nbrs && m)
      PCC_HIP_CHECK(hipMemcpyAsync(nbrs, g->csr.nbr.get(), 4 * (size_t)m, hipMemcpyDeviceToHost,
                                   g->stream));
    if (weights && m)
      PCC_HIP_CHECK(hipMemcpyAsync(weights, g->csr.w.get(), 4 * (size_t)m, hipMemcpyDeviceToHost,
                                   g->stream));
    PCC_HIP_CHECK(hipStreamSynchronize(g->stream));
    return PCC_OK;
  });
}

void pcc_graph_destroy(pcc_graph* g) {
  if (!g) return;
  (void)hipSetDevice(g->device);
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  delete g;
}

int pcc_set_option(pcc_graph* g, int key, int64_t value) {
  ClearError();
  if (!g) return SetError(PCC_INVALID_ARGUMENT, "null graph");
  switch (key) {
    case PCC_OPT_TIMING:
      g->timing = value != 0;
      return PCC_OK;
    case PCC_OPT_FORCE_ORDERED:
      g->force_ordered = value != 0;
      return PCC_OK;
    case PCC_OPT_WAVE_MAX_DEG:
      if (value < 64 || value > 512)
        return SetError(PCC_INVALID_ARGUMENT, "wave bin upper degree must lie in [64, 512]");
      g->th.wave_max = value;
      return PCC_OK;
    case PCC_OPT_BLOCK_MAX_DEG:
      if (value < g->th.wave_max || value > 4096)
        return SetError(PCC_INVALID_ARGUMENT, "block bin upper degree must lie in [wave_max, 4096]");
      g->th.block_max = value;
      return PCC_OK;
    default:
      return SetError(PCC_INVALID_ARGUMENT, "unknown option");
  }
}

int pcc_cluster(pcc_graph* g, const pcc_config* cfg, int objective, uint32_t* out_cluster_ids,
                pcc_stats* stats) {
  ClearError();
  return Guard("pcc_cluster", [&] {
    return ClusterImpl(g, cfg, objective, nullptr, nullptr, out_cluster_ids, stats);
  });
}

int pcc_refine_clusters(pcc_graph* g, const pcc_config* cfg, int objective,
                        const double* node_weights, uint32_t* inout_cluster_ids,
                        pcc_stats* stats) {
  ClearError();
  if (!inout_cluster_ids) return SetError(PCC_INVALID_ARGUMENT, "null cluster ids");
  return Guard("pcc_refine_clusters", [&]() -> int {
    std::vector<uint32_t> init(inout_cluster_ids,
                               inout_cluster_ids + (g ? std::max<int64_t>(g->csr.n, 0) : 0));
    return ClusterImpl(g, cfg, objective, node_weights, init.data(), inout_cluster_ids, stats);
  });
}

int pcc_compute_objective(pcc_graph* g, const pcc_config* cfg, const uint32_t* cluster_ids,
                          const double* node_weights, double* out_objective) {
  ClearError();
  return Guard("pcc_compute_objective", [&]() -> int {
    PCC_RETURN_IF_ERROR(CheckGraph(g));
    if (!cfg || !cluster_ids || !out_objective) return SetError(PCC_INVALID_ARGUMENT, "null argument");
    const int64_t n = g->csr.n;
    for (int64_t v = 0; v < n; ++v)
      if (cluster_ids[v] >= (uint64_t)n) return SetError(PCC_INVALID_ARGUMENT, "cluster id out of range");
    DevBuf<uint32_t> ids(&g->alloc, n > 0 ? n : 1);
    if (n) PCC_HIP_CHECK(hipMemcpyAsync(ids.get(), cluster_ids, 4 * (size_t)n, hipMemcpyHostToDevice, g->stream));
    DevBuf<double> nw = NodeWeights(g, PCC_OBJECTIVE_CORRELATION, node_weights);
    DevBuf<double> res(&g->alloc, 1);
    ComputeObjective(&g->alloc, g->csr.View(), ids.get(), nw.get(), cfg->resolution,
                     cfg->edge_weight_offset, res.get(), g->stream);
    PCC_HIP_CHECK(hipMemcpyAsync(out_objective, res.get(), 8, hipMemcpyDeviceToHost, g->stream));
    PCC_HIP_CHECK(hipStreamSynchronize(g->stream));
    return PCC_OK;
  });
}

int pcc_compute_modularity(pcc_graph* g, double resolution, const uint32_t* cluster_ids,
                           double* out_modularity) {
  ClearError();
  return Guard("pcc_compute_modularity", [&]() -> int {
    PCC_RETURN_IF_ERROR(CheckGraph(g));
    if (!cluster_ids || !out_modularity) return SetError(PCC_INVALID_ARGUMENT, "null argument");
    const int64_t n = g->csr.n;
    for (int64_t v = 0; v < n; ++v)
      if (cluster_ids[v] >= (uint64_t)n) return SetError(PCC_INVALID_ARGUMENT, "cluster id out of range");
    DevBuf<uint32_t> ids(&g->alloc, n > 0 ? n : 1);
    if (n) PCC_HIP_CHECK(hipMemcpyAsync(ids.get(), cluster_ids, 4 * (size_t)n, hipMemcpyHostToDevice, g->stream));
    DevBuf<double> res(&g->alloc, 1);
    ComputeModularity(&g->alloc, g->csr.View(), ids.get(), resolution, res.get(), g->stream);
    PCC_HIP_CHECK(hipMemcpyAsync(out_modularity, res.get(), 8, hipMemcpyDeviceToHost, g->stream));
    PCC_HIP_CHECK(hipStreamSynchronize(g->stream));
    return PCC_OK;
  });
}

int pcc_best_moves(pcc_graph* g, const pcc_config* cfg, const uint32_t* cluster_ids,
                   const uint32_t* cluster_sizes, const double* node_weights,
                   const double* cluster_weights, const uint8_t* frontier, uint32_t* out_moves,
                   uint8_t* out_moved, uint32_t* out_target, double* out_delta) {
  ClearError();
  return Guard("pcc_best_moves", [&]() -> int {
    PCC_RETURN_IF_ERROR(CheckGraph(g));
    if (!cfg || !cluster_ids || !cluster_sizes || !node_weights || !cluster_weights ||
        !out_moves || !out_moved)
      return SetError(PCC_INVALID_ARGUMENT, "null argument");
    PCC_RETURN_IF_ERROR(CheckState(g, cluster_ids, cluster_sizes));
    const int64_t n = g->csr.n;
    if (n == 0) return PCC_OK;
    hipStream_t s = g->stream;
    DevState st(g, cluster_ids, cluster_sizes, node_weights, cluster_weights);
    const CsrView cg = g->csr.View();
    Engine::Workspace ws(&g->alloc, n, g->stream);
    DevBuf<uint8_t> fr;
    if (frontier) {
      fr.Reset(&g->alloc, n);
      PCC_HIP_CHECK(hipMemcpyAsync(fr.get(), frontier, (size_t)n, hipMemcpyHostToDevice, s));
    }
    DevBuf<uint32_t> tgt(&g->alloc, n);
    DevBuf<double> dlt(&g->alloc, n);
    // Entries outside the frontier keep "no move".
    PCC_HIP_CHECK(hipMemsetAsync(ws.moves.get(), 0xFF, 4 * (size_t)n, s));
    PCC_HIP_CHECK(hipMemsetAsync(ws.moved.get(), 0, (size_t)n, s));
    PCC_HIP_CHECK(hipMemsetAsync(tgt.get(), 0xFF, 4 * (size_t)n, s));
    PCC_HIP_CHECK(hipMemsetAsync(dlt.get(), 0, 8 * (size_t)n, s));
    DevBuf<unsigned long long> caps(&g->alloc, kBins);
    PCC_HIP_CHECK(hipMemsetAsync(caps.get(), 0, 8 * kBins, s));
    LaunchBinCapacity(cg, g->th, caps.get(), s);
    unsigned long long hcaps[kBins];
    PCC_HIP_CHECK(hipMemcpyAsync(hcaps, caps.get(), sizeof(hcaps), hipMemcpyDeviceToHost, s));
    PCC_HIP_CHECK(hipMemsetAsync(ws.rc.get(), 0, sizeof(RoundCounters), s));
    PCC_HIP_CHECK(hipStreamSynchronize(s));
    BinLayout layout;
    int64_t run = 0;
    for (int b = 0; b < kBins; ++b) {
      layout.base[b] = run;
      run += (int64_t)hcaps[b];
    }
    LaunchBinFrontier(cg, frontier ? fr.get() : nullptr, g->th, layout, ws.list.get(),
                      ws.rc.get(), s);
    RoundCounters hrc;
    PCC_HIP_CHECK(hipMemcpyAsync(&hrc, ws.rc.get(), sizeof(hrc), hipMemcpyDeviceToHost, s));
    PCC_HIP_CHECK(hipStreamSynchronize(s));
    // Hash scratch sized for the largest degree.
    int64_t maxdeg = 0;
    {
      std::vector<int64_t> off((size_t)n + 1);
      PCC_HIP_CHECK(hipMemcpy(off.data(), cg.off, 8 * (size_t)(n + 1), hipMemcpyDeviceToHost));
      for (int64_t v = 0; v < n; ++v) maxdeg = std::max(maxdeg, off[(size_t)v + 1] - off[(size_t)v]);
    }
    HashScratch hs;
    uint32_t cap = 64;
    while (cap < (uint64_t)(2 * maxdeg + 2)) cap <<= 1;
    hs.cap = cap;
    hs.regions = 16;
    DevBuf<uint32_t> hk(&g->alloc, (int64_t)cap * hs.regions);
    DevBuf<double> hv(&g->alloc, (int64_t)cap * hs.regions);
    hs.keys = hk.get();
    hs.vals = hv.get();
    MoveParams mp;
    mp.res = cfg->resolution;
    mp.offset = cfg->edge_weight_offset;
    mp.n = (uint32_t)n;
    mp.force_ordered = g->force_ordered ? 1 : 0;
    MoveOut out;
    out.moves = ws.moves.get();
    out.moved = ws.moved.get();
    out.moved_clusters = ws.flags.get();
    out.dbg_target = tgt.get();
    out.dbg_delta = dlt.get();
    const StateView sv = st.View();
    for (int b = 0; b < kBins; ++b)
      LaunchBestMoveBin(b, (int64_t)hrc.cnt[b], false, cg, sv, mp, ws.list.get() + layout.base[b],
                        out, ws.fb_list.get(), ws.rc.get(), hs, 0, s);
    LaunchBestMoveOrdered(cg, sv, mp, ws.fb_list.get(), ws.rc.get(), out, hs, s);
    PCC_HIP_CHECK(hipMemcpyAsync(out_moves, ws.moves.get(), 4 * (size_t)n, hipMemcpyDeviceToHost, s));
    PCC_HIP_CHECK(hipMemcpyAsync(out_moved, ws.moved.get(), (size_t)n, hipMemcpyDeviceToHost, s));
    if (out_target)
      PCC_HIP_CHECK(hipMemcpyAsync(out_target, tgt.get(), 4 * (size_t)n, hipMemcpyDeviceToHost, s));
    if (out_delta)
      PCC_HIP_CHECK(hipMemcpyAsync(out_delta, dlt.get(), 8 * (size_t)n, hipMemcpyDeviceToHost, s));
    PCC_HIP_CHECK(hipStreamSynchronize(s));
    return PCC_OK;
  });
}

int pcc_compress(pcc_graph* g, const uint32_t* cluster_ids, const double* node_weights,
                 pcc_graph** out_graph, double* out_node_weights) {
  ClearError();
  if (!out_graph) return SetError(PCC_INVALID_ARGUMENT, "null out");
  *out_graph = nullptr;
  return Guard("pcc_compress", [&]() -> int {
    PCC_RETURN_IF_ERROR(CheckGraph(g));
    if (!cluster_ids || !node_weights || !out_node_weights)
      return SetError(PCC_INVALID_ARGUMENT, "null argument");
    const int64_t n = g->csr.n;
    for (int64_t v = 0; v < n; ++v)
      if (cluster_ids[v] >= (uint64_t)n) return SetError(PCC_INVALID_ARGUMENT, "cluster id out of range");
    if (n == 0) return SetError(PCC_INVALID_ARGUMENT, "empty graph");
    hipStream_t s = g->stream;
    DevBuf<uint32_t> ids(&g->alloc, n);
    DevBuf<double> nw(&g->alloc, n);
    PCC_HIP_CHECK(hipMemcpyAsync(ids.get(), cluster_ids, 4 * (size_t)n, hipMemcpyHostToDevice, s));
    PCC_HIP_CHECK(hipMemcpyAsync(nw.get(), node_weights, 8 * (size_t)n, hipMemcpyHostToDevice, s));
    DevCsr c;
    DevBuf<double> nw2;
    CompressGraph(&g->alloc, g->csr.View(), ids.get(), nw.get(), &c, &nw2, s);
    // Hand the result to a new handle (host round trip keeps ownership simple).
    std::vector<int64_t> off((size_t)c.n + 1);
    std::vector<uint32_t> nb((size_t)std::max<int64_t>(c.m, 1));
    std::vector<float> w((size_t)std::max<int64_t>(c.m, 1));
    PCC_HIP_CHECK(hipMemcpyAsync(off.data(), c.off.get(), 8 * off.size(), hipMemcpyDeviceToHost, s));
    if (c.m) {
      PCC_HIP_CHECK(hipMemcpyAsync(nb.data(), c.nbr.get(), 4 * (size_t)c.m, hipMemcpyDeviceToHost, s));
      PCC_HIP_CHECK(hipMemcpyAsync(w.data(), c.w.get(), 4 * (size_t)c.m, hipMemcpyDeviceToHost, s));
    }
    PCC_HIP_CHECK(hipMemcpyAsync(out_node_weights, nw2.get(), 8 * (size_t)c.n, hipMemcpyDeviceToHost, s));
    PCC_HIP_CHECK(hipStreamSynchronize(s));
    pcc_graph* h = nullptr;
    PCC_RETURN_IF_ERROR(NewHandle(&h));
    const int rc = UploadCsr(h, c.n, c.m, off.data(), nb.data(), w.data());
    if (rc != PCC_OK) {
      delete h;
      return rc;
    }
    *out_graph = h;
    return PCC_OK;
  });
}

int pcc_round(pcc_graph* g, const pcc_config* cfg, uint32_t* cluster_ids, uint32_t* cluster_sizes,
              const double* node_weights, double* cluster_weights, uint8_t* frontier,
              int64_t* out_frontier_size) {
  ClearError();
  return Guard("pcc_round", [&]() -> int {
    PCC_RETURN_IF_ERROR(CheckGraph(g));
    if (!cfg || !cluster_ids || !cluster_sizes || !node_weights || !cluster_weights || !frontier)
      return SetError(PCC_INVALID_ARGUMENT, "null argument");
    PCC_RETURN_IF_ERROR(CheckState(g, cluster_ids, cluster_sizes));
    const int64_t n = g->csr.n;
    if (n == 0) {
      if (out_frontier_size) *out_frontier_size = 0;
      return PCC_OK;
    }
    hipStream_t s = g->stream;
    DevState st(g, cluster_ids, cluster_sizes, node_weights, cluster_weights);
    const CsrView cg = g->csr.View();
    Engine e(g, *cfg, nullptr);
    Engine::Workspace ws(&g->alloc, n, g->stream);
    DevBuf<unsigned long long> caps(&g->alloc, kBins);
    PCC_HIP_CHECK(hipMemsetAsync(caps.get(), 0, 8 * kBins, s));
    LaunchBinCapacity(cg, g->th, caps.get(), s);
    unsigned long long hcaps[kBins];
    PCC_HIP_CHECK(hipMemcpyAsync(hcaps, caps.get(), sizeof(hcaps), hipMemcpyDeviceToHost, s));
    PCC_HIP_CHECK(hipStreamSynchronize(s));
    BinLayout layout;
    int64_t run = 0;
    for (int b = 0; b < kBins; ++b) {
      layout.base[b] = run;
      run += (int64_t)hcaps[b];
    }
    HashScratch hs;
    const int64_t maxdeg = e.MaxDegree(cg);
    uint32_t cap = 64;
    while (cap < (uint64_t)(2 * maxdeg + 2)) cap <<= 1;
    hs.cap = cap;
    hs.regions = 16;
    DevBuf<uint32_t> hk(&g->alloc, (int64_t)cap * hs.regions);
    DevBuf<double> hv(&g->alloc, (int64_t)cap * hs.regions);
    hs.keys = hk.get();
    hs.vals = hv.get();
    MoveParams mp;
    mp.res = cfg->resolution;
    mp.offset = cfg->edge_weight_offset;
    mp.n = (uint32_t)n;
    mp.force_ordered = g->force_ordered ? 1 : 0;
    DevBuf<uint8_t> fr(&g->alloc, n);
    PCC_HIP_CHECK(hipMemcpyAsync(fr.get(), frontier, (size_t)n, hipMemcpyHostToDevice, s));
    e.ResetCounters(ws.rc.get(), true);
    LaunchBinFrontier(cg, fr.get(), g->th, layout, ws.list.get(), ws.rc.get(), s);
    int64_t counts[kBins], te[kBins];
    e.ReadCounters(ws.rc.get(), counts, te, nullptr);
    int64_t fsize = 0;
    e.RoundStep(cg, st.View(), mp, layout, hs, ws, counts, te, 0, &fsize);
    if (cfg->move_method == PCC_ALL_MOVE) {
      std::memset(frontier, 1, (size_t)n);
    } else {
      PCC_HIP_CHECK(hipMemcpyAsync(frontier, ws.next.get(), (size_t)n, hipMemcpyDeviceToHost, s));
    }
    const size_t nn = (size_t)n;
    PCC_HIP_CHECK(hipMemcpyAsync(cluster_ids, st.ids.get(), 4 * nn, hipMemcpyDeviceToHost, s));
    PCC_HIP_CHECK(hipMemcpyAsync(cluster_sizes, st.sizes.get(), 4 * nn, hipMemcpyDeviceToHost, s));
    PCC_HIP_CHECK(hipMemcpyAsync(cluster_weights, st.K.get(), 8 * nn, hipMemcpyDeviceToHost, s));
    PCC_HIP_CHECK(hipStreamSynchronize(s));
    if (out_frontier_size) *out_frontier_size = fsize;
    return PCC_OK;
  });
}

}  // extern "C"
