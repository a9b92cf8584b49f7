// cluster_in_memory_main.cc -- command-line runner with the reference CLI's
// flags and output (clustering/cluster-in-memory_main.cc:39-64, 252-327):
//
//   cluster-in-memory --clusterer_name=ParallelCorrelationClusterer
//       --clusterer_config='correlation_clusterer_config { resolution: 0.01 }'
//       --input_graph=<GBBS adjacency file> [--float_weighted]
//       [--output_clustering=<file>] [--input_communities=<file>]
//
// The graph must be a symmetric GBBS (Weighted)AdjacencyGraph.  Clustering
// runs on the current HIP device.

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "clusterer.h"
#include "pcc_amd.h"

using pcc::Status;
using pcc::in_memory::InMemoryClusterer;

namespace {

struct Flags {
  std::string clusterer_name, clusterer_config, input_graph, output_clustering, input_communities;
  bool is_symmetric_graph = true;
  bool float_weighted = false;
};

bool ParseBoolFlag(const std::string& v) { return v.empty() || v == "true" || v == "1"; }

Status ParseFlags(int argc, char** argv, Flags* f) {
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a.rfind("--", 0) != 0) return Status(PCC_INVALID_ARGUMENT, "unexpected argument " + a);
    a = a.substr(2);
    std::string name = a, value;
    bool has_value = false;
    const size_t eq = a.find('=');
    if (eq != std::string::npos) {
      name = a.substr(0, eq);
      value = a.substr(eq + 1);
      has_value = true;
    }
    auto take = [&]() -> std::string {
      if (has_value) return value;
      if (i + 1 < argc) return argv[++i];
      return "";
    };
    if (name == "clusterer_name") f->clusterer_name = take();
    else if (name == "clusterer_config") f->clusterer_config = take();
    else if (name == "input_graph") f->input_graph = take();
    else if (name == "output_clustering") f->output_clustering = take();
    else if (name == "input_communities") f->input_communities = take();
    else if (name == "float_weighted") f->float_weighted = ParseBoolFlag(value);
    else if (name == "nofloat_weighted") f->float_weighted = false;
    else if (name == "is_symmetric_graph") f->is_symmetric_graph = ParseBoolFlag(value);
    else return Status(PCC_INVALID_ARGUMENT, "unknown flag --" + name);
  }
  return pcc::OkStatus();
}

void PrintTime(std::chrono::steady_clock::time_point begin, std::chrono::steady_clock::time_point end,
               const std::string& what) {
  std::cout << what << " Time: "
            << std::chrono::duration_cast<std::chrono::microseconds>(end - begin).count() / 1000000.0
            << std::endl;
}

// WriteClustering, cluster-in-memory_main.cc:132-145.
Status WriteClustering(const std::string& filename, const InMemoryClusterer::Clustering& c) {
  std::ofstream file(filename);
  if (!file.is_open()) return Status(PCC_NOT_FOUND, "Unable to open file.");
  for (const auto& cluster : c) {
    for (auto node_id : cluster) file << node_id << "\t";
    file << std::endl;
  }
  return pcc::OkStatus();
}

// ReadCommunities + CompareCommunities, cluster-in-memory_main.cc:147-225.
void CompareCommunities(const std::string& filename, InMemoryClusterer::Clustering clustering) {
  std::vector<std::vector<uint32_t>> communities;
  std::ifstream infile(filename);
  std::string line;
  while (infile.is_open() && std::getline(infile, line)) {
    std::vector<uint32_t> row;
    std::stringstream ss(line);
    std::string item;
    while (std::getline(ss, item, '\t'))
      if (!item.empty()) row.push_back((uint32_t)std::stoul(item));
    std::sort(row.begin(), row.end());
    communities.push_back(row);
  }
  for (auto& c : clustering) std::sort(c.begin(), c.end());
  double avg_precision = 0, avg_recall = 0;
  for (const auto& community : communities) {
    size_t max_intersect = 0, max_idx = 0;
    std::vector<uint32_t> inter(community.size());
    for (size_t i = 0; i < clustering.size(); ++i) {
      auto it = std::set_intersection(clustering[i].begin(), clustering[i].end(), community.begin(),
                                      community.end(), inter.begin());
      const size_t sz = (size_t)(it - inter.begin());
      if (sz > max_intersect) {
        max_intersect = sz;
        max_idx = i;
      }
    }
    avg_precision += clustering.empty() ? 0.0 : (double)max_intersect / (double)clustering[max_idx].size();
    avg_recall += community.empty() ? 0.0 : (double)max_intersect / (double)community.size();
  }
  avg_precision /= (double)communities.size();
  avg_recall /= (double)communities.size();
  std::cout << "Avg precision: " << std::setprecision(17) << avg_precision << std::endl;
  std::cout << "Avg recall: " << std::setprecision(17) << avg_recall << std::endl;
  size_t min_size = 0xFFFFFFFFu, max_size = 0;
  double avg_size = 0;
  for (const auto& c : clustering) {
    min_size = std::min(min_size, c.size());
    max_size = std::max(max_size, c.size());
    avg_size += (double)c.size();
  }
  avg_size /= (double)clustering.size();
  std::cout << "Num communities: " << clustering.size() << std::endl;
  std::cout << "Min size: " << min_size << std::endl;
  std::cout << "Max size: " << max_size << std::endl;
  std::cout << "Avg size: " << avg_size << std::endl;
}

Status Main(int argc, char** argv) {
  Flags f;
  Status s = ParseFlags(argc, argv, &f);
  if (!s.ok()) return s;
  auto cfg = pcc::ClustererConfig::ParseFromText(f.clusterer_config);
  if (!cfg.ok())
    return Status(PCC_INVALID_ARGUMENT,
                  "Cannot parse --clusterer_config as a text-format "
                  "research_graph.in_memory.ClustererConfig proto: " + f.clusterer_config);
  std::unique_ptr<pcc::in_memory::ParallelCorrelationClusterer> clusterer;
  if (f.clusterer_name == "ParallelCorrelationClusterer") {
    clusterer.reset(new pcc::in_memory::ParallelCorrelationClusterer);
  } else if (f.clusterer_name == "ParallelModularityClusterer") {
    clusterer.reset(new pcc::in_memory::ParallelModularityClusterer);
  } else {
    return Status(PCC_UNIMPLEMENTED,
                  "ParallelCorrelationClusterer and ParallelModularityClusterer are the only "
                  "supported clusterers.");
  }
  const auto begin_read = std::chrono::steady_clock::now();
  pcc_host_csr csr{};
  int rc = pcc_read_adjacency_graph(f.input_graph.c_str(), f.float_weighted ? 1 : 0, &csr);
  if (rc != PCC_OK) return Status::FromCode(rc);
  s = clusterer->device_graph().SetCsr(csr.n, csr.m, csr.offsets, csr.nbrs, csr.weights);
  pcc_host_csr_free(&csr);
  if (!s.ok()) return s;
  PrintTime(begin_read, std::chrono::steady_clock::now(), "Read graph");

  auto clustering = clusterer->Cluster(*cfg);
  if (!clustering.ok()) return clustering.status();
  const pcc_stats& st = clusterer->last_stats();
  std::cout << "Cluster Time: " << st.cluster_seconds << std::endl;
  std::cout << "Objective: " << std::setprecision(17) << st.objective << std::endl;
  std::cout << "Number of Clusters: " << clustering->size() << std::endl;
  if (f.clusterer_name == "ParallelModularityClusterer")
    std::cout << "Modularity: " << std::setprecision(17) << st.modularity << std::endl;
  if (!f.input_communities.empty()) CompareCommunities(f.input_communities, *clustering);
  if (!f.output_clustering.empty()) {
    s = WriteClustering(f.output_clustering, *clustering);
    if (!s.ok()) return s;
  }
  return pcc::OkStatus();
}

}  // namespace

int main(int argc, char** argv) {
  const Status status = Main(argc, argv);
  if (!status.ok()) {
    std::cerr << status.ToString() << std::endl;
    return EXIT_FAILURE;
  }
  return 0;
}
