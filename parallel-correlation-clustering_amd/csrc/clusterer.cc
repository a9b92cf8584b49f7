// clusterer.cc -- the C++ facade (see clusterer.h) implemented over the C ABI.

#include "clusterer.h"

#include <algorithm>
#include <mutex>
#include <numeric>

namespace pcc {

Status Status::FromCode(int code) {
  if (code == PCC_OK) return OkStatus();
  return Status(code, pcc_last_error());
}

std::string Status::ToString() const {
  if (ok()) return "OK";
  const char* name = "UNKNOWN";
  switch (code_) {
    case PCC_INVALID_ARGUMENT: name = "INVALID_ARGUMENT"; break;
    case PCC_NOT_FOUND: name = "NOT_FOUND"; break;
    case PCC_RESOURCE_EXHAUSTED: name = "RESOURCE_EXHAUSTED"; break;
    case PCC_UNIMPLEMENTED: name = "UNIMPLEMENTED"; break;
    case PCC_INTERNAL: name = "INTERNAL"; break;
    default: break;
  }
  return std::string(name) + ": " + message_;
}

StatusOr<ClustererConfig> ClustererConfig::ParseFromText(const std::string& text) {
  ClustererConfig c;
  const int rc = pcc_config_parse_text(text.c_str(), &c.correlation_clusterer_config);
  if (rc != PCC_OK) return Status::FromCode(rc);
  return c;
}

namespace in_memory {

namespace {
std::mutex g_create_mu;
}  // namespace

DeviceGraph::~DeviceGraph() { pcc_graph_destroy(g_); }

Status DeviceGraph::EnsureHandle() {
  std::lock_guard<std::mutex> lock(g_create_mu);
  if (g_) return OkStatus();
  return Status::FromCode(pcc_graph_new(&g_));
}

Status DeviceGraph::Import(AdjacencyList adjacency_list) {
  Status s = EnsureHandle();
  if (!s.ok()) return s;
  std::vector<uint32_t> nbrs(adjacency_list.outgoing_edges.size());
  std::vector<double> w(nbrs.size());
  for (size_t i = 0; i < nbrs.size(); ++i) {
    nbrs[i] = adjacency_list.outgoing_edges[i].first;
    w[i] = adjacency_list.outgoing_edges[i].second;
  }
  return Status::FromCode(pcc_graph_import(g_, adjacency_list.id, adjacency_list.weight,
                                           nbrs.data(), w.data(), nbrs.size()));
}

Status DeviceGraph::FinishImport() {
  Status s = EnsureHandle();
  if (!s.ok()) return s;
  return Status::FromCode(pcc_graph_finish_import(g_));
}

Status DeviceGraph::SetCsr(int64_t n, int64_t m, const int64_t* offsets, const uint32_t* nbrs,
                           const float* weights) {
  pcc_graph* h = nullptr;
  const int rc = pcc_graph_from_csr(n, m, offsets, nbrs, weights, &h);
  if (rc != PCC_OK) return Status::FromCode(rc);
  pcc_graph_destroy(g_);
  g_ = h;
  return OkStatus();
}

InMemoryClusterer::Clustering OutputIndicesById(const std::vector<uint32_t>& ids) {
  std::vector<uint32_t> order(ids.size());
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return ids[a] < ids[b]; });
  InMemoryClusterer::Clustering out;
  for (size_t i = 0; i < order.size(); ++i) {
    if (i == 0 || ids[order[i]] != ids[order[i - 1]]) out.emplace_back();
    out.back().push_back(order[i]);
  }
  return out;
}

StatusOr<InMemoryClusterer::Clustering> ParallelCorrelationClusterer::Cluster(
    const ClustererConfig& config) const {
  if (!graph_.handle()) return Status(PCC_INVALID_ARGUMENT, "graph not imported");
  std::vector<uint32_t> ids((size_t)std::max<int64_t>(graph_.num_nodes(), 1));
  const int rc = pcc_cluster(graph_.handle(), &config.correlation_clusterer_config, objective(),
                             ids.data(), &stats_);
  if (rc != PCC_OK) return Status::FromCode(rc);
  ids.resize((size_t)graph_.num_nodes());
  return OutputIndicesById(ids);
}

std::string InMemoryClusterer::StringId(NodeId id) const {
  if (node_id_map_ == nullptr) return std::to_string(id);
  if ((size_t)id < node_id_map_->size()) return (*node_id_map_)[id];
  return "missing-id-" + std::to_string(id);
}

Status ParallelCorrelationClusterer::RefineClusters(const ClustererConfig& config,
                                                    Clustering* clustering) const {
  return RefineClusters(config, clustering, {});
}

Status ParallelCorrelationClusterer::RefineClusters(const ClustererConfig& config,
                                                    Clustering* clustering,
                                                    const std::vector<double>& node_weights,
                                                    double resolution) const {
  (void)resolution;
  if (!graph_.handle()) return Status(PCC_INVALID_ARGUMENT, "graph not imported");
  if (!clustering) return Status(PCC_INVALID_ARGUMENT, "null clustering");
  const int64_t n = graph_.num_nodes();
  if (!node_weights.empty() && (int64_t)node_weights.size() != n)
    return Status(PCC_INVALID_ARGUMENT, "node_weights must have one entry per node");
  // The initial clustering must cover [0, n) exactly once
  // (parallel-correlation-clusterer.h:53-56).
  std::vector<uint32_t> ids((size_t)std::max<int64_t>(n, 1), PCC_NO_MOVE);
  for (size_t i = 0; i < clustering->size(); ++i)
    for (uint32_t v : (*clustering)[i]) {
      if ((int64_t)v >= n)
        return Status(PCC_INVALID_ARGUMENT,
                      "clustering names node " + StringId(v) + ", which is not in the graph");
      if (ids[v] != PCC_NO_MOVE)
        return Status(PCC_INVALID_ARGUMENT,
                      "clustering holds node " + StringId(v) + " more than once");
      ids[v] = (uint32_t)i;
    }
  for (int64_t v = 0; v < n; ++v)
    if (ids[(size_t)v] == PCC_NO_MOVE)
      return Status(PCC_INVALID_ARGUMENT,
                    "clustering misses node " + StringId((NodeId)v));
  const int rc = pcc_refine_clusters(graph_.handle(), &config.correlation_clusterer_config,
                                     objective(), node_weights.empty() ? nullptr : node_weights.data(),
                                     ids.data(), &stats_);
  if (rc != PCC_OK) return Status::FromCode(rc);
  ids.resize((size_t)n);
  *clustering = OutputIndicesById(ids);
  return OkStatus();
}

}  // namespace in_memory
}  // namespace pcc
