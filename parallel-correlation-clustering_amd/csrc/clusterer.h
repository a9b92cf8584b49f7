// clusterer.h -- C++ facade over the C ABI with the reference's interface.
//
// Mirrors research_graph::in_memory::InMemoryClusterer
// (clustering/in-memory-clusterer.h:34-121), GbbsGraph
// (clustering/gbbs-graph.h:39-56), ParallelCorrelationClusterer
// (clustering/clusterers/parallel-correlation-clusterer.h:44-95) and
// ParallelModularityClusterer
// (clustering/clusterers/parallel-modularity-clusterer.h:45-56).  absl::Status
// and absl::StatusOr are replaced by the small pcc::Status / pcc::StatusOr
// below with the same codes; the config proto by ClustererConfig, which wraps
// pcc_config and parses the same text format.
#ifndef PCC_AMD_CLUSTERER_H_
#define PCC_AMD_CLUSTERER_H_

#include <stdint.h>

#include <string>
#include <utility>
#include <vector>

#include "pcc_amd.h"

namespace pcc {

class Status {
 public:
  Status() = default;
  Status(int code, std::string message) : code_(code), message_(std::move(message)) {}
  static Status FromCode(int code);  // reads pcc_last_error() when code != 0
  bool ok() const { return code_ == PCC_OK; }
  int code() const { return code_; }
  const std::string& message() const { return message_; }
  std::string ToString() const;

 private:
  int code_ = PCC_OK;
  std::string message_;
};

inline Status OkStatus() { return Status(); }

template <typename T>
class StatusOr {
 public:
  StatusOr(Status s) : status_(std::move(s)) {}  // NOLINT: implicit like absl
  StatusOr(T v) : value_(std::move(v)) {}        // NOLINT
  bool ok() const { return status_.ok(); }
  const Status& status() const { return status_; }
  T& value() { return value_; }
  const T& value() const { return value_; }
  T& operator*() { return value_; }
  T* operator->() { return &value_; }
  const T* operator->() const { return &value_; }

 private:
  Status status_;
  T value_{};
};

// ClustererConfig holding a CorrelationClustererConfig (clustering/config.proto:188-196).
struct ClustererConfig {
  pcc_config correlation_clusterer_config;
  ClustererConfig() { pcc_config_default(&correlation_clusterer_config); }
  // google::protobuf::TextFormat::ParseFromString equivalent.
  static StatusOr<ClustererConfig> ParseFromText(const std::string& text);
};

namespace in_memory {

class InMemoryClusterer {
 public:
  class Graph {
   public:
    using NodeId = uint32_t;
    struct AdjacencyList {
      NodeId id = static_cast<NodeId>(-1);
      double weight = 1;
      std::vector<std::pair<NodeId, double>> outgoing_edges;
    };
    virtual ~Graph() = default;
    // Safe to call concurrently (in-memory-clusterer.h:60-65).
    virtual Status Import(AdjacencyList adjacency_list) = 0;
    virtual Status FinishImport() = 0;
  };

  using NodeId = Graph::NodeId;
  using AdjacencyList = Graph::AdjacencyList;
  using Clustering = std::vector<std::vector<NodeId>>;

  virtual ~InMemoryClusterer() = default;
  virtual Graph* MutableGraph() = 0;
  virtual StatusOr<Clustering> Cluster(const ClustererConfig& config) const = 0;
  virtual Status RefineClusters(const ClustererConfig& config, Clustering* clustering) const {
    (void)config;
    (void)clustering;
    return OkStatus();
  }

  // String ids of the NodeIds, used in error messages
  // (in-memory-clusterer.h:101-108).  Not owned: the vector must outlive
  // every call on this object.
  void set_node_id_map(const std::vector<std::string>* node_id_map) {
    node_id_map_ = node_id_map;
  }

 protected:
  // The map's entry for id, "missing-id-<id>" for an id outside the map, or
  // the decimal id when no map was set (in-memory-clusterer.cc:27-35).
  std::string StringId(NodeId id) const;

 private:
  const std::vector<std::string>* node_id_map_ = nullptr;
};

// GbbsGraph stand-in: imports adjacency lists and keeps the CSR in HBM.
class DeviceGraph : public InMemoryClusterer::Graph {
 public:
  DeviceGraph() = default;
  ~DeviceGraph() override;
  DeviceGraph(const DeviceGraph&) = delete;
  DeviceGraph& operator=(const DeviceGraph&) = delete;

  Status Import(AdjacencyList adjacency_list) override;
  Status FinishImport() override;
  // The CLI path: the whole CSR at once (cluster-in-memory_main.cc:306).
  Status SetCsr(int64_t n, int64_t m, const int64_t* offsets, const uint32_t* nbrs,
                const float* weights);
  pcc_graph* handle() const { return g_; }
  int64_t num_nodes() const { return g_ ? pcc_graph_num_nodes(g_) : 0; }

 private:
  Status EnsureHandle();
  pcc_graph* g_ = nullptr;
};

// Groups vertices by id, ids ascending, members ascending
// (OutputIndicesById, parallel/parallel-sequence-ops.h:183-209).
InMemoryClusterer::Clustering OutputIndicesById(const std::vector<uint32_t>& ids);

class ParallelCorrelationClusterer : public InMemoryClusterer {
 public:
  Graph* MutableGraph() override { return &graph_; }
  StatusOr<Clustering> Cluster(const ClustererConfig& config) const override;
  Status RefineClusters(const ClustererConfig& config, Clustering* clustering) const override;
  // RefineClusters with explicit node weights (parallel-correlation-clusterer.h:58-91).
  // resolution is accepted and not used: the reference passes it on as
  // original_resolution, which it never reads
  // (parallel-correlation-clusterer.cc:407-412, 631-639).
  Status RefineClusters(const ClustererConfig& config, Clustering* clustering,
                        const std::vector<double>& node_weights, double resolution = 0) const;
  const pcc_stats& last_stats() const { return stats_; }
  DeviceGraph& device_graph() { return graph_; }

 protected:
  virtual int objective() const { return PCC_OBJECTIVE_CORRELATION; }
  DeviceGraph graph_;
  mutable pcc_stats stats_{};
};

class ParallelModularityClusterer : public ParallelCorrelationClusterer {
 protected:
  int objective() const override { return PCC_OBJECTIVE_MODULARITY; }
};

}  // namespace in_memory
}  // namespace pcc

#endif  // PCC_AMD_CLUSTERER_H_
