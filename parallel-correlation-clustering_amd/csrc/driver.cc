// driver.cc -- the C ABI (include/pcc_amd.h) over the engine.
//
// Each entry point names the reference interface it replaces in
// include/pcc_amd.h.  Errors come back as absl status codes with a
// thread-local message; nothing throws across the ABI.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "engine.h"
#include "kernels.h"
#include "pcc_amd.h"
#include "status.h"

using namespace pcc;

namespace {

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
    PCC_HIP_CHECK(hipStreamSynchronize(h->stream));
  } else if (objective == PCC_OBJECTIVE_MODULARITY) {
    LaunchDegrees(h->csr.View(), nw.get(), h->stream);  // PGU.cc:141-153
  } else {
    LaunchFillF64(nw.get(), n > 0 ? n : 1, 1.0, h->stream);
  }
  return nw;
}

// Shared body of pcc_cluster / pcc_refine_clusters / pcc_batch_cluster.
// host_init: per-vertex initial cluster indices local to each graph.
// out_ids: final ids local to each graph.  objectives / modularities: per
// graph (may be null).
int ClusterImpl(pcc_graph* g, const pcc_config* cfg_in, int objective, const double* host_nw,
                const double* resolutions, const uint32_t* host_init, uint32_t* out_ids,
                double* objectives, double* modularities, pcc_stats* stats);

// Subclustering numbers subclusters and lays out levels per graph, so a batch
// with it is clustered graph by graph: each graph on a handle of its own with
// the batch's options, through the single-graph path.  Statistics are summed
// (levels: the maximum; objective and modularity: the first graph's, as for
// batches).
int ClusterEachGraph(pcc_graph* g, const pcc_config* cfg_in, int objective, const double* host_nw,
                     const double* resolutions, const uint32_t* host_init, uint32_t* out_ids,
                     double* objectives, double* modularities, pcc_stats* stats) {
  const int G = g->num_graphs();
  const std::vector<int64_t>& base = g->gbase;
  const int64_t n = g->csr.n, m = g->csr.m;
  std::vector<int64_t> off((size_t)n + 1);
  std::vector<uint32_t> nbr((size_t)std::max<int64_t>(m, 1));
  std::vector<float> w((size_t)std::max<int64_t>(m, 1));
  PCC_HIP_CHECK(hipMemcpyAsync(off.data(), g->csr.off.get(), 8 * (size_t)(n + 1),
                               hipMemcpyDeviceToHost, g->stream));
  if (m > 0) {
    PCC_HIP_CHECK(hipMemcpyAsync(nbr.data(), g->csr.nbr.get(), 4 * (size_t)m,
                                 hipMemcpyDeviceToHost, g->stream));
    PCC_HIP_CHECK(hipMemcpyAsync(w.data(), g->csr.w.get(), 4 * (size_t)m, hipMemcpyDeviceToHost,
                                 g->stream));
  }
  PCC_HIP_CHECK(hipStreamSynchronize(g->stream));
  if (stats) std::memset(stats, 0, sizeof(*stats));
  for (int o = 0; o < G; ++o) {
    const int64_t v0 = base[(size_t)o], v1 = base[(size_t)o + 1];
    const int64_t e0 = off[(size_t)v0], e1 = off[(size_t)v1];
    std::vector<int64_t> lo((size_t)(v1 - v0) + 1);
    for (int64_t i = v0; i <= v1; ++i) lo[(size_t)(i - v0)] = off[(size_t)i] - e0;
    std::vector<uint32_t> ln((size_t)std::max<int64_t>(e1 - e0, 1));
    for (int64_t k = e0; k < e1; ++k) ln[(size_t)(k - e0)] = nbr[(size_t)k] - (uint32_t)v0;
    pcc_graph* h = nullptr;
    PCC_RETURN_IF_ERROR(NewHandle(&h));
    std::unique_ptr<pcc_graph> hold(h);
    h->timing = g->timing;
    h->force_ordered = g->force_ordered;
    h->async_subrounds = g->async_subrounds;
    h->async_singleton_rule = g->async_singleton_rule;
    h->cluster_subrounds = g->cluster_subrounds;
    h->hub_part_edges = g->hub_part_edges;
    h->pull_fraction = g->pull_fraction;
    h->device_rounds = g->device_rounds;
    h->contract_method = g->contract_method;
    h->contract_chunk = g->contract_chunk;
    h->th = g->th;
    PCC_RETURN_IF_ERROR(UploadCsr(h, v1 - v0, e1 - e0, lo.data(), ln.data(), w.data() + e0));
    pcc_config c = *cfg_in;
    if (resolutions) c.resolution = resolutions[o];
    pcc_stats so;
    double ob = 0.0, mo = 0.0;
    PCC_RETURN_IF_ERROR(ClusterImpl(h, &c, objective, host_nw ? host_nw + v0 : nullptr, nullptr,
                                    host_init ? host_init + v0 : nullptr, out_ids + v0, &ob, &mo,
                                    &so));
    if (objectives) objectives[o] = ob;
    if (modularities) modularities[o] = mo;
    if (!stats) continue;
    if (o == 0) {
      stats->objective = so.objective;
      stats->modularity = so.modularity;
    }
    stats->levels = std::max(stats->levels, so.levels);
    stats->rounds += so.rounds;
    stats->edges_scanned += so.edges_scanned;
    stats->frontier_vertices += so.frontier_vertices;
    stats->num_clusters += so.num_clusters;
    stats->fallback_vertices += so.fallback_vertices;
    stats->cluster_seconds += so.cluster_seconds;
    stats->bestmove_seconds += so.bestmove_seconds;
    stats->bestmove_bytes += so.bestmove_bytes;
    stats->contract_seconds += so.contract_seconds;
    stats->apply_seconds += so.apply_seconds;
    stats->bestmove_launches += so.bestmove_launches;
    stats->ordered_seconds += so.ordered_seconds;
    for (int b = 0; b < 8; ++b) {
      stats->bin_seconds[b] += so.bin_seconds[b];
      stats->bin_edges[b] += so.bin_edges[b];
      stats->bin_vertices[b] += so.bin_vertices[b];
      stats->bin_launches[b] += so.bin_launches[b];
    }
    stats->cluster_move_seconds += so.cluster_move_seconds;
    stats->cluster_moves += so.cluster_moves;
  }
  return PCC_OK;
}

int ClusterImpl(pcc_graph* g, const pcc_config* cfg_in, int objective, const double* host_nw,
                const double* resolutions, const uint32_t* host_init, uint32_t* out_ids,
                double* objectives, double* modularities, pcc_stats* stats) {
  PCC_RETURN_IF_ERROR(CheckGraph(g));
  if (cfg_in && cfg_in->subclustering_method != 0 && g->num_graphs() > 1)
    return ClusterEachGraph(g, cfg_in, objective, host_nw, resolutions, host_init, out_ids,
                            objectives, modularities, stats);
  if (!cfg_in || (!out_ids && g->csr.n > 0)) return SetError(PCC_INVALID_ARGUMENT, "null argument");
  if (objective != PCC_OBJECTIVE_CORRELATION && objective != PCC_OBJECTIVE_MODULARITY)
    return SetError(PCC_INVALID_ARGUMENT, "unknown objective");
  pcc_config cfg = *cfg_in;
  const int G = g->num_graphs();
  const int64_t n = g->csr.n;
  const std::vector<int64_t>& base = g->gbase;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  std::vector<double> gamma((size_t)G, cfg.resolution);
  if (resolutions) gamma.assign(resolutions, resolutions + G);
  std::vector<double> res = gamma;
  if (objective == PCC_OBJECTIVE_MODULARITY) {
    // ComputeModularityConfig + PMC.cc:76-79: k_v = deg(v), res' = gamma / sum deg.
    const std::vector<int64_t>& off_at = g->goff;
    for (int o = 0; o < G; ++o)
      res[(size_t)o] = gamma[(size_t)o] / (double)(off_at[(size_t)o + 1] - off_at[(size_t)o]);
    cfg.edge_weight_offset = 0;
  }
  cfg.resolution = res[0];
  DevBuf<uint32_t> init;
  if (host_init) {
    std::vector<uint32_t> gl((size_t)n);
    for (int o = 0; o < G; ++o)
      for (int64_t v = base[(size_t)o]; v < base[(size_t)o + 1]; ++v) {
        if ((int64_t)host_init[v] >= base[(size_t)o + 1] - base[(size_t)o])
          return SetError(PCC_INVALID_ARGUMENT, "initial cluster index out of range [0, n)");
        gl[(size_t)v] = (uint32_t)(base[(size_t)o] + host_init[v]);
      }
    init.Reset(&g->alloc, n > 0 ? n : 1);
    if (n > 0)
      PCC_HIP_CHECK(hipMemcpyAsync(init.get(), gl.data(), sizeof(uint32_t) * (size_t)n,
                                   hipMemcpyHostToDevice, g->stream));
    PCC_HIP_CHECK(hipStreamSynchronize(g->stream));
  }
  DevBuf<double> nw = NodeWeights(g, objective, host_nw);
  DevBuf<double> nw_keep(&g->alloc, n > 0 ? n : 1);
  PCC_HIP_CHECK(hipMemcpyAsync(nw_keep.get(), nw.get(), sizeof(double) * (size_t)(n > 0 ? n : 1),
                               hipMemcpyDeviceToDevice, g->stream));
  DevBuf<uint32_t> ids;  // final ids as vertex ids of the whole batch
  {
    Engine e(g, cfg, stats);
    PCC_RETURN_IF_ERROR(e.Run(std::move(nw), G > 1 ? res.data() : nullptr,
                              host_init ? init.get() : nullptr, out_ids, nullptr, &ids));
  }
  if (stats || objectives || modularities) {
    // Final objective on the input graph (PCC.cc:619-624) and modularity
    // (PMC.cc:91-93), per graph; outside the timed span like the reference.
    if (n == 0) ids.Reset(&g->alloc, 1);
    DevBuf<int64_t> dbase(&g->alloc, (int64_t)base.size());
    DevBuf<double> dres(&g->alloc, G), dgam(&g->alloc, G);
    PCC_HIP_CHECK(hipMemcpyAsync(dbase.get(), base.data(), 8 * base.size(), hipMemcpyHostToDevice, g->stream));
    PCC_HIP_CHECK(hipMemcpyAsync(dres.get(), res.data(), 8 * (size_t)G, hipMemcpyHostToDevice, g->stream));
    PCC_HIP_CHECK(hipMemcpyAsync(dgam.get(), gamma.data(), 8 * (size_t)G, hipMemcpyHostToDevice, g->stream));
    DevBuf<double> obj(&g->alloc, G), mod(&g->alloc, G);
    ComputeObjective(&g->alloc, g->csr.View(), ids.get(), nw_keep.get(), G, dbase.get(), dres.get(),
                     cfg.edge_weight_offset, obj.get(), g->stream);
    if (objective == PCC_OBJECTIVE_MODULARITY)
      ComputeModularity(&g->alloc, g->csr.View(), ids.get(), G, dbase.get(), dgam.get(), mod.get(),
                        g->stream);
    std::vector<double> hobj((size_t)G), hmod((size_t)G, 0.0);
    PCC_HIP_CHECK(hipMemcpyAsync(hobj.data(), obj.get(), 8 * (size_t)G, hipMemcpyDeviceToHost, g->stream));
    if (objective == PCC_OBJECTIVE_MODULARITY)
      PCC_HIP_CHECK(hipMemcpyAsync(hmod.data(), mod.get(), 8 * (size_t)G, hipMemcpyDeviceToHost, g->stream));
    PCC_HIP_CHECK(hipStreamSynchronize(g->stream));
    if (objectives) std::copy(hobj.begin(), hobj.end(), objectives);
    if (modularities) std::copy(hmod.begin(), hmod.end(), modularities);
    if (stats) {
      stats->objective = hobj[0];
      stats->modularity = hmod[0];
      // Graphs own disjoint id ranges, so the distinct ids of the batch are
      // the sum of every graph's cluster count.
      stats->num_clusters = CountDistinctIds(&g->alloc, n, ids.get(), g->stream);
    }
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

int CheckIds(pcc_graph* g, const uint32_t* ids) {
  const int64_t n = g->csr.n;
  for (int64_t v = 0; v < n; ++v)
    if (ids[v] >= (uint64_t)n) return SetError(PCC_INVALID_ARGUMENT, "cluster id out of range");
  return PCC_OK;
}

int CheckSingle(pcc_graph* g) {
  if (g->num_graphs() != 1)
    return SetError(PCC_INVALID_ARGUMENT, "this entry point takes a single graph, not a batch");
  return PCC_OK;
}

}  // namespace

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

int pcc_autotune(pcc_graph* g, const pcc_config* cfg, int objective, const double* resolutions,
                 int max_trials, pcc_tune_result* out) {
  ClearError();
  return Guard("pcc_autotune", [&]() -> int {
    PCC_RETURN_IF_ERROR(CheckGraph(g));
    if (!cfg || !out) return SetError(PCC_INVALID_ARGUMENT, "null argument");
    struct Setting {
      int64_t wave, block, part;
    };
    const int64_t n = g->csr.n;
    std::vector<uint32_t> ids((size_t)std::max<int64_t>(n, 1));
    const bool timing0 = g->timing;
    g->timing = false;  // per-launch events would distort the comparison
    const int G = g->num_graphs();
    int trials = 0;
    const int budget = max_trials > 0 ? max_trials : 16;
    // Device span of one Cluster call with setting x.
    auto run = [&](const Setting& x, double* sec) -> int {
      g->th.wave_max = x.wave;
      g->th.block_max = x.block;
      g->hub_part_edges = x.part;
      pcc_stats st;
      ++trials;
      PCC_RETURN_IF_ERROR(ClusterImpl(g, cfg, objective, nullptr, G > 1 ? resolutions : nullptr,
                                      nullptr, ids.data(), nullptr, nullptr, &st));
      *sec = st.cluster_seconds;
      return PCC_OK;
    };
    Setting best{g->th.wave_max, g->th.block_max, g->hub_part_edges};
    double best_s = 0;
    int rc = run(best, &best_s);
    if (rc != PCC_OK) {
      g->timing = timing0;
      return rc;
    }
    out->initial_seconds = best_s;
    // One coordinate at a time: workgroup bin bound, wave bin bound, hub
    // partition size.
    const int64_t blocks[] = {1024, 2048, 4096};
    const int64_t waves[] = {128, 256, 512};
    const int64_t parts[] = {1024, 2048, 4096};
    for (int coord = 0; coord < 3 && rc == PCC_OK; ++coord) {
      const int64_t* cand = coord == 0 ? blocks : coord == 1 ? waves : parts;
      const Setting base = best;
      for (int i = 0; i < 3 && trials < budget; ++i) {
        Setting x = base;
        if (coord == 0) x.block = cand[i];
        if (coord == 1) x.wave = cand[i];
        if (coord == 2) x.part = cand[i];
        if (x.wave > x.block) continue;
        if (x.wave == base.wave && x.block == base.block && x.part == base.part) continue;
        double sec = 0;
        rc = run(x, &sec);
        if (rc != PCC_OK) break;
        if (sec < best_s) {
          best_s = sec;
          best = x;
        }
      }
    }
    g->timing = timing0;
    g->th.wave_max = best.wave;
    g->th.block_max = best.block;
    g->hub_part_edges = best.part;
    if (rc != PCC_OK) return rc;
    out->wave_max_deg = best.wave;
    out->block_max_deg = best.block;
    out->hub_part_edges = best.part;
    out->trials = trials;
    out->best_seconds = best_s;
    return PCC_OK;
  });
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
    case PCC_OPT_ASYNC_SUBROUNDS:
      if (value < 0 || value > 1024)
        return SetError(PCC_INVALID_ARGUMENT, "async sub-rounds must lie in [0, 1024]");
      g->async_subrounds = (int)value;
      return PCC_OK;
    case PCC_OPT_ASYNC_SINGLETON_RULE:
      g->async_singleton_rule = value != 0;
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
    case PCC_OPT_HUB_PART_EDGES:
      if (value < 16 || value > 4096)
        return SetError(PCC_INVALID_ARGUMENT, "hub partition size must lie in [16, 4096]");
      g->hub_part_edges = value;
      return PCC_OK;
    case PCC_OPT_HASH_WAVES:
      if (value < 0 || value > 64)
        return SetError(PCC_INVALID_ARGUMENT, "hash-bin waves must lie in [0, 64]");
      g->th.hash_waves = value;
      return PCC_OK;
    case PCC_OPT_CLUSTER_SUBROUNDS:
      if (value < 1 || value > 1024)
        return SetError(PCC_INVALID_ARGUMENT, "cluster sub-rounds must lie in [1, 1024]");
      g->cluster_subrounds = (int)value;
      return PCC_OK;
    case PCC_OPT_DEVICE_ROUNDS:
      if (value < 0 || value > 2)
        return SetError(PCC_INVALID_ARGUMENT, "device rounds must be 0, 1 or 2");
      g->device_rounds = (int)value;
      return PCC_OK;
    case PCC_OPT_CONTRACT_METHOD:
      if (value < 0 || value > 1)
        return SetError(PCC_INVALID_ARGUMENT, "contraction method must be 0 or 1");
      g->contract_method = (int)value;
      return PCC_OK;
    case PCC_OPT_CONTRACT_CHUNK:
      if (value < 1 || value > kContractChunkTerms)
        return SetError(PCC_INVALID_ARGUMENT, "contraction chunk must lie in [1, 2^31]");
      g->contract_chunk = value;
      return PCC_OK;
    case PCC_OPT_PULL_FRACTION:
      if (value < 0 || value > 1024)
        return SetError(PCC_INVALID_ARGUMENT, "pull fraction must lie in [0, 1024]");
      g->pull_fraction = (int)value;
      return PCC_OK;
    default:
      return SetError(PCC_INVALID_ARGUMENT, "unknown option");
  }
}

int pcc_cluster(pcc_graph* g, const pcc_config* cfg, int objective, uint32_t* out_cluster_ids,
                pcc_stats* stats) {
  ClearError();
  return Guard("pcc_cluster", [&]() -> int {
    PCC_RETURN_IF_ERROR(CheckGraph(g));
    PCC_RETURN_IF_ERROR(CheckSingle(g));
    return ClusterImpl(g, cfg, objective, nullptr, nullptr, nullptr, out_cluster_ids, nullptr,
                       nullptr, stats);
  });
}

int pcc_refine_clusters(pcc_graph* g, const pcc_config* cfg, int objective,
                        const double* node_weights, uint32_t* inout_cluster_ids,
                        pcc_stats* stats) {
  ClearError();
  if (!inout_cluster_ids) return SetError(PCC_INVALID_ARGUMENT, "null cluster ids");
  return Guard("pcc_refine_clusters", [&]() -> int {
    PCC_RETURN_IF_ERROR(CheckGraph(g));
    PCC_RETURN_IF_ERROR(CheckSingle(g));
    std::vector<uint32_t> init(inout_cluster_ids, inout_cluster_ids + std::max<int64_t>(g->csr.n, 0));
    return ClusterImpl(g, cfg, objective, node_weights, nullptr, init.data(), inout_cluster_ids,
                       nullptr, nullptr, stats);
  });
}

int pcc_compute_objective(pcc_graph* g, const pcc_config* cfg, const uint32_t* cluster_ids,
                          const double* node_weights, double* out_objective) {
  ClearError();
  return Guard("pcc_compute_objective", [&]() -> int {
    PCC_RETURN_IF_ERROR(CheckGraph(g));
    PCC_RETURN_IF_ERROR(CheckSingle(g));
    if (!cfg || !cluster_ids || !out_objective) return SetError(PCC_INVALID_ARGUMENT, "null argument");
    PCC_RETURN_IF_ERROR(CheckIds(g, cluster_ids));
    const int64_t n = g->csr.n;
    DevBuf<uint32_t> ids(&g->alloc, n > 0 ? n : 1);
    if (n) PCC_HIP_CHECK(hipMemcpyAsync(ids.get(), cluster_ids, 4 * (size_t)n, hipMemcpyHostToDevice, g->stream));
    DevBuf<double> nw = NodeWeights(g, PCC_OBJECTIVE_CORRELATION, node_weights);
    DevBuf<double> res(&g->alloc, 1), dres(&g->alloc, 1);
    DevBuf<int64_t> dbase(&g->alloc, 2);
    const int64_t hb[2] = {0, n};
    PCC_HIP_CHECK(hipMemcpyAsync(dbase.get(), hb, 16, hipMemcpyHostToDevice, g->stream));
    PCC_HIP_CHECK(hipMemcpyAsync(dres.get(), &cfg->resolution, 8, hipMemcpyHostToDevice, g->stream));
    ComputeObjective(&g->alloc, g->csr.View(), ids.get(), nw.get(), 1, dbase.get(), dres.get(),
                     cfg->edge_weight_offset, res.get(), g->stream);
    PCC_HIP_CHECK(hipMemcpyAsync(out_objective, res.get(), 8, hipMemcpyDeviceToHost, g->stream));
    PCC_HIP_CHECK(hipStreamSynchronize(g->stream));
    return PCC_OK;
  });
}

int pcc_compute_objective_noalloc(pcc_graph* g, const pcc_config* cfg,
                                  const uint32_t* cluster_ids_dev, const double* node_weights_dev,
                                  double* out_objective_dev, void* stream) {
  ClearError();
  return Guard("pcc_compute_objective_noalloc", [&]() -> int {
    PCC_RETURN_IF_ERROR(CheckGraph(g));
    PCC_RETURN_IF_ERROR(CheckSingle(g));
    const int64_t n = g->csr.n, m = g->csr.m;
    if (!cfg || !out_objective_dev || (n > 0 && !cluster_ids_dev))
      return SetError(PCC_INVALID_ARGUMENT, "null argument");
    if (n > kNoAllocMaxVertices || m > kNoAllocMaxEdges)
      return SetError(PCC_INVALID_ARGUMENT,
                      "pcc_compute_objective_noalloc takes at most 2^21 vertices and 2^30 "
                      "directed edges (its work memory is a fixed per-device array)");
    void* work = NoAllocScratch(g->device);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g->stream;
    ComputeObjectiveNoAlloc(work, g->csr.View(), cluster_ids_dev, node_weights_dev,
                            cfg->resolution, cfg->edge_weight_offset, out_objective_dev, s);
    return PCC_OK;
  });
}

int pcc_compute_modularity(pcc_graph* g, double resolution, const uint32_t* cluster_ids,
                           double* out_modularity) {
  ClearError();
  return Guard("pcc_compute_modularity", [&]() -> int {
    PCC_RETURN_IF_ERROR(CheckGraph(g));
    PCC_RETURN_IF_ERROR(CheckSingle(g));
    if (!cluster_ids || !out_modularity) return SetError(PCC_INVALID_ARGUMENT, "null argument");
    PCC_RETURN_IF_ERROR(CheckIds(g, cluster_ids));
    const int64_t n = g->csr.n;
    DevBuf<uint32_t> ids(&g->alloc, n > 0 ? n : 1);
    if (n) PCC_HIP_CHECK(hipMemcpyAsync(ids.get(), cluster_ids, 4 * (size_t)n, hipMemcpyHostToDevice, g->stream));
    DevBuf<double> res(&g->alloc, 1), dgam(&g->alloc, 1);
    DevBuf<int64_t> dbase(&g->alloc, 2);
    const int64_t hb[2] = {0, n};
    PCC_HIP_CHECK(hipMemcpyAsync(dbase.get(), hb, 16, hipMemcpyHostToDevice, g->stream));
    PCC_HIP_CHECK(hipMemcpyAsync(dgam.get(), &resolution, 8, hipMemcpyHostToDevice, g->stream));
    ComputeModularity(&g->alloc, g->csr.View(), ids.get(), 1, dbase.get(), dgam.get(), res.get(),
                      g->stream);
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
    PCC_RETURN_IF_ERROR(CheckSingle(g));
    if (!cfg || !cluster_ids || !cluster_sizes || !node_weights || !cluster_weights ||
        !out_moves || !out_moved)
      return SetError(PCC_INVALID_ARGUMENT, "null argument");
    PCC_RETURN_IF_ERROR(CheckIds(g, cluster_ids));
    PCC_RETURN_IF_ERROR(CheckMovesConfig(g, *cfg));
    const int64_t n = g->csr.n;
    if (n == 0) return PCC_OK;
    hipStream_t s = g->stream;
    DevState st(g, cluster_ids, cluster_sizes, node_weights, cluster_weights);
    Engine e(g, *cfg, nullptr);
    Engine::LevelCtx c;
    Layout lay;
    e.PrepareLevel(g->csr.View(), st.View(), lay, &c);
    Engine::Workspace ws(&g->alloc, n, s);
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
    int64_t counts[kBins], te[kBins];
    e.BinFrontier(c, ws, frontier ? fr.get() : nullptr, counts, te);
    MoveOut out;
    out.moves = ws.moves.get();
    out.moved = ws.moved.get();
    out.moved_clusters = ws.flags.get();
    out.dbg_target = tgt.get();
    out.dbg_delta = dlt.get();
    const uint32_t* lists[kBins];
    for (int b = 0; b < kBins; ++b) lists[b] = ws.list.get() + c.layout.base[b];
    e.LaunchMoves(c, ws, c.mp, false, lists, counts, te, out, 0, true);
    LaunchBestMoveOrdered(c.g, c.st, c.mp, ws.fb_list.get(), ws.rc.get(), out, c.hs, s);
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
    PCC_RETURN_IF_ERROR(CheckSingle(g));
    if (!cluster_ids || !node_weights || !out_node_weights)
      return SetError(PCC_INVALID_ARGUMENT, "null argument");
    PCC_RETURN_IF_ERROR(CheckIds(g, cluster_ids));
    const int64_t n = g->csr.n;
    if (n == 0) return SetError(PCC_INVALID_ARGUMENT, "empty graph");
    hipStream_t s = g->stream;
    DevBuf<uint32_t> ids(&g->alloc, n);
    DevBuf<double> nw(&g->alloc, n);
    PCC_HIP_CHECK(hipMemcpyAsync(ids.get(), cluster_ids, 4 * (size_t)n, hipMemcpyHostToDevice, s));
    PCC_HIP_CHECK(hipMemcpyAsync(nw.get(), node_weights, 8 * (size_t)n, hipMemcpyHostToDevice, s));
    const uint32_t max_id = *std::max_element(cluster_ids, cluster_ids + n);
    DevCsr c;
    DevBuf<double> nw2;
    CompressGraph(&g->alloc, g->csr.View(), ids.get(), nw.get(), nullptr, nullptr, nullptr,
                  (int64_t)max_id + 1, &c, &nw2, s, g->contract_method,
                  g->contract_chunk);
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
    PCC_RETURN_IF_ERROR(CheckSingle(g));
    if (!cfg || !cluster_ids || !cluster_sizes || !node_weights || !cluster_weights || !frontier)
      return SetError(PCC_INVALID_ARGUMENT, "null argument");
    PCC_RETURN_IF_ERROR(CheckIds(g, cluster_ids));
    PCC_RETURN_IF_ERROR(CheckMovesConfig(g, *cfg));
    const int64_t n = g->csr.n;
    if (n == 0) {
      if (out_frontier_size) *out_frontier_size = 0;
      return PCC_OK;
    }
    hipStream_t s = g->stream;
    DevState st(g, cluster_ids, cluster_sizes, node_weights, cluster_weights);
    const bool exact = NodeWeightsExact(&g->alloc, st.nw.get(), n, s);
    if (!exact && cfg->clustering_moves_method == PCC_DEFAULT_CLUSTER_MOVES)
      return SetError(PCC_INVALID_ARGUMENT,
                      "DEFAULT_CLUSTER_MOVES needs integer node weights whose magnitudes sum "
                      "below 2^53");
    Engine e(g, *cfg, nullptr);
    e.SetExactWeights(exact);
    Engine::LevelCtx c;
    Layout lay;
    e.PrepareLevel(g->csr.View(), st.View(), lay, &c);
    Engine::Workspace ws(&g->alloc, n, s);
    DevBuf<uint8_t> fr(&g->alloc, n), gmoved(&g->alloc, 1);
    PCC_HIP_CHECK(hipMemcpyAsync(fr.get(), frontier, (size_t)n, hipMemcpyHostToDevice, s));
    int64_t counts[kBins], te[kBins];
    e.BinFrontier(c, ws, fr.get(), counts, te);
    const int64_t fsize = e.RoundStep(c, ws, nullptr, 1, gmoved.get(), counts, te);
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
    if (out_frontier_size) *out_frontier_size = cfg->move_method == PCC_ALL_MOVE ? n : fsize;
    return PCC_OK;
  });
}

int pcc_batch_from_csr(int64_t num_graphs, const int64_t* vertex_counts, const int64_t* offsets,
                       const uint32_t* nbrs, const float* weights, pcc_graph** out) {
  ClearError();
  if (!out) return SetError(PCC_INVALID_ARGUMENT, "null out");
  *out = nullptr;
  if (num_graphs < 1 || !vertex_counts || !offsets)
    return SetError(PCC_INVALID_ARGUMENT, "a batch needs at least one graph");
  return Guard("pcc_batch_from_csr", [&]() -> int {
    std::vector<int64_t> base((size_t)num_graphs + 1, 0);
    for (int64_t gi = 0; gi < num_graphs; ++gi) {
      if (vertex_counts[gi] < 1)
        return SetError(PCC_INVALID_ARGUMENT, "every graph of a batch needs at least one vertex");
      base[(size_t)gi + 1] = base[(size_t)gi] + vertex_counts[gi];
    }
    const int64_t n = base.back();
    if (n >= (int64_t)0xFFFFFFF0ll)
      return SetError(PCC_INVALID_ARGUMENT, "vertex count must lie in [0, 2^32 - 16)");
    const int64_t m = offsets[n];
    if (offsets[0] != 0 || m < 0)
      return SetError(PCC_INVALID_ARGUMENT, "offsets must start at 0 and end at m >= 0");
    if (m > 0 && !nbrs) return SetError(PCC_INVALID_ARGUMENT, "null CSR array");
    // The graph boundaries must be non-decreasing within [0, m] before the
    // rows are read (UploadCsr checks every offset afterwards).
    for (int64_t gi = 0; gi < num_graphs; ++gi) {
      const int64_t a = offsets[base[(size_t)gi]], b = offsets[base[(size_t)gi + 1]];
      if (a < 0 || b < a || b > m)
        return SetError(PCC_INVALID_ARGUMENT, "offsets must be non-decreasing within [0, m]");
    }
    // Local neighbour ids -> ids of the disjoint union.
    std::vector<uint32_t> gl((size_t)std::max<int64_t>(m, 1));
    for (int64_t gi = 0; gi < num_graphs; ++gi) {
      const int64_t b = base[(size_t)gi], e = base[(size_t)gi + 1];
      for (int64_t p = offsets[b]; p < offsets[e]; ++p) {
        if ((int64_t)nbrs[p] >= e - b)
          return SetError(PCC_INVALID_ARGUMENT, "neighbour id outside its graph");
        gl[(size_t)p] = (uint32_t)(b + nbrs[p]);
      }
    }
    pcc_graph* h = nullptr;
    PCC_RETURN_IF_ERROR(NewHandle(&h));
    h->gbase = base;
    int rc = UploadCsr(h, n, m, offsets, gl.data(), weights);
    if (rc == PCC_OK && num_graphs > 1) {
      std::vector<uint32_t> vg((size_t)n);
      for (int64_t gi = 0; gi < num_graphs; ++gi)
        std::fill(vg.begin() + base[(size_t)gi], vg.begin() + base[(size_t)gi + 1], (uint32_t)gi);
      h->vgraph.Reset(&h->alloc, n);
      PCC_HIP_CHECK(hipMemcpy(h->vgraph.get(), vg.data(), 4 * (size_t)n, hipMemcpyHostToDevice));
    }
    if (rc != PCC_OK) {
      delete h;
      return rc;
    }
    *out = h;
    return PCC_OK;
  });
}

int64_t pcc_batch_num_graphs(const pcc_graph* g) { return g ? g->num_graphs() : -1; }

int pcc_batch_cluster(pcc_graph* g, const pcc_config* cfg, int objective, const double* resolutions,
                      uint32_t* out_cluster_ids, double* out_objectives, double* out_modularities,
                      pcc_stats* stats) {
  ClearError();
  return Guard("pcc_batch_cluster", [&] {
    return ClusterImpl(g, cfg, objective, nullptr, resolutions, nullptr, out_cluster_ids,
                       out_objectives, out_modularities, stats);
  });
}

}  // extern "C"
