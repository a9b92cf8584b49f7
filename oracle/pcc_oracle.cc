// pcc_oracle.cc -- CPU restatement of the ParallelCorrelationClusterer /
// ParallelModularityClusterer hot path.
//
// TEST INFRASTRUCTURE ONLY: the parity checker for the HIP path.  Nothing in
// the product library links, loads or calls this file.
//
// Reference files restated (paths relative to the reference repository root):
//   clustering/clusterers/parallel-correlation-clusterer.cc        (PCC.cc)
//   clustering/clusterers/parallel-correlation-clusterer-internal.cc (PCCI.cc)
//   clustering/clusterers/correlation-clusterer-util.h             (UTIL.h)
//   clustering/clusterers/parallel-modularity-clusterer.cc         (PMC.cc)
//   parallel/parallel-graph-utils.cc                                (PGU.cc)
//   parallel/parallel-sequence-ops.h                                (PSO.h)
//   clustering/clusterers/correlation-clusterer.cc, correlation-clusterer-
//   internal.cc, modularity-clusterer.cc: the sequential clusterer (SEQ-CC /
//   SEQ-MOD, oracle_config.sequential), the paper's own CPU baseline
//
// Floating-point policy (pinned identically in the HIP kernels):
//   * EdgeSum::NetWeight (UTIL.h:36-39) computes weight_ - res * x.  The
//     reference is built by GCC with -O2 -march=native and GCC's default
//     -ffp-contract=fast for C++, which fuses that expression into one
//     fnmadd; it is written here as std::fma(-res, x, weight_).  Everything
//     else is compiled with -ffp-contract=off.
//   * Block-structured sums follow pbbslib's reduce (GBBS d385adac,
//     pbbslib/sequence_ops.h): blocks of max(1024, 4*ceil(sqrt(n))) elements,
//     each folded left starting from its first element, then the block sums
//     reduced the same way.  This blocking is restated from the published
//     pbbslib source, which is not vendored in the reference: parity with the
//     reference for sums longer than 1024 terms is unpinned.
//   * Per-cluster sums of hub vertices (degree > 1000) go through
//     sparse_additive_map atomics in the reference (PCCI.cc:518-536), whose
//     order is nondeterministic there.  The oracle uses adjacency order, one
//     of the reference's possible interleavings.
//
// Subclustering (PCC.cc:175-249, 436-474; PCCI.cc:784-1120):
//   * TRIANGLE labels a component by its smallest member's position; the
//     reference's workefficient_cc::CC labels come from a randomized
//     decomposition, so only the partition is pinned, not the ids.
//   * Between LOUVAIN levels the oracle composes the levels through the
//     subcluster ids (the next level's vertices).  The reference composes
//     cluster ids there (PCC.cc:550, 594-595) and so indexes the next level by
//     the wrong vertex; its results on that step cannot pin this restatement,
//     which is parity unpinned against the reference and pinned only by the
//     composition test (tests/test_oracle_subclustering.py).
//   * Async subcluster moves are not restated.

#include "pcc_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
#endif

namespace {

constexpr uint32_t kNone = ORACLE_NO_MOVE;
constexpr int64_t kHubDegree = 1000;  // PCCI.cc:456

int g_threads = 1;
int64_t g_edges_scanned = 0;  // sum over rounds of sum_{v in frontier} deg(v)

// Wall seconds per phase of the whole-pipeline runs: 0 best-move map, 1 sync
// apply, 2 frontier expansion, 3 contraction, 4 refinement relabel.
double g_phase[5] = {0, 0, 0, 0, 0};
double Now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}
struct PhaseTimer {
  int k;
  double t0;
  explicit PhaseTimer(int kk) : k(kk), t0(Now()) {}
  ~PhaseTimer() { g_phase[k] += Now() - t0; }
};

// Stable sort on g_threads threads (libstdc++'s parallel multiway mergesort).
// A stable sort has one possible output, so the thread count never changes a
// result.
template <class It, class Cmp>
void StableSort(It b, It e, Cmp cmp) {
#ifdef _OPENMP
  if (g_threads > 1 && e - b > 65536) {
    __gnu_parallel::stable_sort(b, e, cmp,
                                __gnu_parallel::multiway_mergesort_tag((unsigned)g_threads));
    return;
  }
#endif
  std::stable_sort(b, e, cmp);
}

// pbbslib reduce: block size and left folds (see the header comment).
template <class T, class F>
T PbbsReduce(const T* a, size_t n, T identity, F f) {
  if (n == 0) return identity;
  size_t bs = std::max<size_t>(1024, 4 * (size_t)std::ceil(std::sqrt((double)n)));
  size_t l = (n + bs - 1) / bs;
  if (l <= 1) {
    T r = a[0];
    for (size_t i = 1; i < n; ++i) r = f(r, a[i]);
    return r;
  }
  std::vector<T> sums(l);
  for (size_t b = 0; b < l; ++b) {
    size_t s = b * bs, e = std::min(n, s + bs);
    T r = a[s];
    for (size_t i = s + 1; i < e; ++i) r = f(r, a[i]);
    sums[b] = r;
  }
  return PbbsReduce(sums.data(), l, identity, f);
}

// EdgeSum::NetWeight, UTIL.h:36-39, fused as the reference binary computes it.
inline double NetWeight(double weight, double sum_prod_node_weights, double res) {
  return std::fma(-res, sum_prod_node_weights, weight);
}

// Per-thread open-addressing map ClusterId -> double (absl::flat_hash_map /
// sparse_additive_map stand-in; iteration order never matters because the
// best-move selection is a total order, PCCI.cc:493-503).
struct ClusterMap {
  std::vector<uint32_t> keys;
  std::vector<double> vals;
  std::vector<uint32_t> used;
  uint32_t mask = 0;
  void Reset(size_t expected) {
    size_t cap = 16;
    while (cap < 2 * expected + 2) cap <<= 1;
    if (keys.size() < cap) {
      keys.assign(cap, kNone);
      vals.assign(cap, 0.0);
    } else {
      for (uint32_t s : used) keys[s] = kNone, vals[s] = 0.0;
      cap = keys.size();
    }
    used.clear();
    mask = (uint32_t)(cap - 1);
  }
  double& At(uint32_t key) {
    uint32_t h = (key * 0x9E3779B1u) & mask;
    while (true) {
      if (keys[h] == key) return vals[h];
      if (keys[h] == kNone) {
        keys[h] = key;
        vals[h] = 0.0;
        used.push_back(h);
        return vals[h];
      }
      h = (h + 1) & mask;
    }
  }
};

thread_local ClusterMap t_map;

// ClusteringHelper::EfficientBestMove, PCCI.cc:448-583.  flat: every degree
// takes the adjacency-order path, as SeqClusteringHelper::EfficientBestMove
// (correlation-clusterer-internal.cc:116-174) does.
// cands (may be null) receives every candidate (cluster, delta) the
// reference evaluates, in evaluation order, and change the "new cluster"
// value.
struct Candidates {
  std::vector<std::pair<uint32_t, double>> list;
  double change = 0.0;
};

void BestMove(const oracle_graph& g, const oracle_state& s, double res,
              double offset, uint32_t v, uint32_t* out_t, double* out_d, bool flat = false,
              Candidates* cands = nullptr) {
  const int64_t b = g.offsets[v], e = g.offsets[(int64_t)v + 1];
  const int64_t deg = e - b;
  const uint32_t n = (uint32_t)g.n;
  const uint32_t node_cluster = s.cluster_ids[v];
  const double k = s.node_weights[v];
  ClusterMap& after = t_map;
  after.Reset((size_t)deg);

  if (flat || deg <= kHubDegree) {
    // PCCI.cc:457-511
    double together = 0.0;  // class_1_currently_together (EdgeSum starts at 0)
    double moving_nodes_weight = 0;
    moving_nodes_weight += k;
    for (int64_t p = b; p < e; ++p) {
      double weight = (double)g.weights[p];
      weight -= offset;
      const uint32_t u = g.nbrs[p];
      const uint32_t neighbor_cluster = s.cluster_ids[u];
      if (v != u) {
        if (node_cluster == neighbor_cluster) together += weight;
        after.At(neighbor_cluster) += weight;
      }
    }
    double change = 0;
    double max_edges = k * (s.cluster_weights[node_cluster] - k);
    change -= NetWeight(together, max_edges, res);
    bool has = false;
    uint32_t best = 0;
    double best_delta = change;
    if (cands) cands->change = change;
    for (uint32_t slot : after.used) {
      const uint32_t c = after.keys[slot];
      max_edges = moving_nodes_weight * s.cluster_weights[c];
      const double overall = change + NetWeight(after.vals[slot], max_edges, res);
      if (cands) cands->list.emplace_back(c, overall);
      if (overall > best_delta || (overall == best_delta && has && c < best)) {
        best = c;
        best_delta = overall;
        has = true;
      }
    }
    *out_t = has ? best : n;
    *out_d = best_delta;
    return;
  }

  // Hub path, PCCI.cc:516-582.
  std::vector<double> curr_together_seq((size_t)deg, 0.0);
  double moving_nodes_weight = 0;
  moving_nodes_weight += k;
  for (int64_t j = 0; j < deg; ++j) {
    float weight = g.weights[b + j];
    weight -= offset;  // float arithmetic, PCCI.cc:529
    const uint32_t u = g.nbrs[b + j];
    const uint32_t neighbor_cluster = s.cluster_ids[u];
    if (v != u) {
      if (node_cluster == neighbor_cluster) curr_together_seq[(size_t)j] = weight;
      after.At(neighbor_cluster) += (double)weight;
    }
  }
  const double max_edges = k * (s.cluster_weights[node_cluster] - k);
  const double curr_together = NetWeight(
      PbbsReduce(curr_together_seq.data(), curr_together_seq.size(), 0.0,
                 [](double x, double y) { return x + y; }),
      max_edges, res);
  double change = 0;
  change -= curr_together;
  // best_move_seq + max reduce with identity (0, 0.0), PCCI.cc:551-575.
  std::vector<std::pair<uint32_t, double>> seq;
  seq.reserve(after.used.size());
  for (uint32_t slot : after.used) {
    const uint32_t c = after.keys[slot];
    double data = after.vals[slot];
    const double me = moving_nodes_weight * s.cluster_weights[c];
    data = NetWeight(data, me, res);  // data -= res * max_edges
    seq.emplace_back(c, change + data);
  }
  if (cands) {
    cands->change = change;
    cands->list = seq;
  }
  auto f_max = [](const std::pair<uint32_t, double>& a,
                  const std::pair<uint32_t, double>& bb) {
    if (a.second == bb.second) return a.first < bb.first ? a : bb;
    if (a.second > bb.second) return a;
    return bb;
  };
  const auto max_move = PbbsReduce(seq.data(), seq.size(),
                                   std::make_pair(uint32_t{0}, 0.0), f_max);
  if (change >= max_move.second) {  // PCCI.cc:577-581
    *out_t = n;
    *out_d = change;
    return;
  }
  *out_t = max_move.first;
  *out_d = max_move.second;
}

// Sync decision for one vertex, PCC.cc:140-156.
inline void SyncDecision(const oracle_graph& g, const oracle_state& s,
                         const oracle_config& cfg, uint32_t v, uint32_t* move,
                         uint8_t* moved, uint32_t* tgt, double* dlt) {
  uint32_t t;
  double d;
  BestMove(g, s, cfg.resolution, cfg.edge_weight_offset, v, &t, &d);
  if (tgt) *tgt = t;
  if (dlt) *dlt = d;
  const uint32_t cur = s.cluster_ids[v];
  if (t < (uint32_t)g.n && s.cluster_sizes[t] == 1 && s.cluster_sizes[cur] == 1 &&
      cur >= t) {
    t = cur;
    d = 0;
  }
  if (d > 0) {
    *move = t;
    *moved = 1;
  } else {
    *move = kNone;
    *moved = 0;
  }
}

// ClusteringHelper::AsyncMove(node), PCCI.cc:371-403 (sequential execution).
bool AsyncMove(const oracle_graph& g, oracle_state& s, const oracle_config& cfg,
               uint32_t v) {
  uint32_t t;
  double d;
  BestMove(g, s, cfg.resolution, cfg.edge_weight_offset, v, &t, &d);
  const uint32_t cur = s.cluster_ids[v];
  if (d <= 0) return false;
  const uint32_t n = (uint32_t)g.n;
  s.cluster_sizes[cur] -= 1;
  s.cluster_weights[cur] += -1 * s.node_weights[v];
  if (t != n) {
    s.cluster_sizes[t] += 1;
    s.cluster_weights[t] += s.node_weights[v];
    s.cluster_ids[v] = t;
    return true;
  }
  size_t i = v;
  while (true) {
    if (s.cluster_sizes[i] == 0) {
      s.cluster_sizes[i] = 1;
      s.cluster_weights[i] += s.node_weights[v];
      s.cluster_ids[v] = (uint32_t)i;
      return true;
    }
    i = (i + 1) % n;
  }
}

// ClusteringHelper::BestMove(graph, moving_nodes), PCCI.cc:644-707, with
// BestMoveFromStats, UTIL.h:70-113.  members is the moving set in the order
// the reference visits it (ascending vertex ids: the groups come from
// OutputIndicesById, PSO.h:183-209); flag is an all-zero scratch of n bytes.
void BestMoveSet(const oracle_graph& g, const oracle_state& s, const oracle_config& cfg,
                 const std::vector<uint32_t>& members, std::vector<uint8_t>& flag,
                 uint32_t* out_t, double* out_d) {
  const double offset = cfg.edge_weight_offset, res = cfg.resolution;
  for (uint32_t u : members) flag[u] = 1;
  // cluster_moving_weights (PCCI.cc:657): a handful of entries.
  std::vector<std::pair<uint32_t, double>> cmw;
  auto cmw_at = [&](uint32_t c) -> double& {
    for (auto& e : cmw)
      if (e.first == c) return e.second;
    cmw.emplace_back(c, 0.0);
    return cmw.back().second;
  };
  int64_t total_deg = 0;
  for (uint32_t u : members) total_deg += g.offsets[u + 1] - g.offsets[u];
  ClusterMap& after = t_map;  // class_1_together_after
  after.Reset((size_t)total_deg);
  double class2 = 0.0, class1_together = 0.0, moving_w = 0.0;
  for (uint32_t node : members) {  // PCCI.cc:665-688
    const uint32_t nc = s.cluster_ids[node];
    cmw_at(nc) += s.node_weights[node];
    moving_w += s.node_weights[node];
    for (int64_t p = g.offsets[node]; p < g.offsets[(int64_t)node + 1]; ++p) {
      const uint32_t nb = g.nbrs[p];
      // The lambda takes the weight as float: the offset is subtracted in
      // double and the result rounded back to float (PCCI.cc:668-669).
      const float w = (float)((double)g.weights[p] - offset);
      const uint32_t ncl = s.cluster_ids[nb];
      if (flag[nb]) {
        if (nc != ncl) class2 += (double)w;
      } else {
        if (nc == ncl) class1_together += (double)w;
        after.At(ncl) += (double)w;
      }
    }
  }
  for (uint32_t u : members) flag[u] = 0;
  class2 /= 2;  // RemoveDoubleCounting, UTIL.h:32
  // BestMoveFromStats, UTIL.h:79-111.  `max_edges += a * b` is one fused
  // multiply-add under GCC's default contraction, as NetWeight is.
  double change = 0.0;
  double max_edges = moving_w * moving_w / 2;
  for (const auto& e : cmw) max_edges -= e.second * e.second / 2;
  change += NetWeight(class2, max_edges, res);
  max_edges = 0.0;
  for (const auto& e : cmw)
    max_edges = std::fma(e.second, s.cluster_weights[e.first] - e.second, max_edges);
  change -= NetWeight(class1_together, max_edges, res);
  bool have = false;
  uint32_t best_c = 0;
  double best_d = change;
  for (uint32_t slot : after.used) {
    const uint32_t c = after.keys[slot];
    double moving_in_c = 0.0;
    for (const auto& e : cmw)
      if (e.first == c) moving_in_c = e.second;
    const double me = moving_w * (s.cluster_weights[c] - moving_in_c);
    const double overall = change + NetWeight(after.vals[slot], me, res);
    if (overall > best_d || (overall == best_d && have && c < best_c)) {
      best_c = c;
      best_d = overall;
      have = true;
    }
  }
  *out_t = have ? best_c : (uint32_t)g.n;
  *out_d = best_d;
}

// ClusteringHelper::AsyncMove(graph, moving_nodes), PCCI.cc:405-445
// (sequential execution).
bool AsyncMoveSet(const oracle_graph& g, oracle_state& s, const oracle_config& cfg,
                  const std::vector<uint32_t>& members, std::vector<uint8_t>& flag) {
  if (members.empty()) return false;
  if (members.size() == 1) return AsyncMove(g, s, cfg, members[0]);
  uint32_t t;
  double d;
  BestMoveSet(g, s, cfg, members, flag, &t, &d);
  if (d <= 0) return false;
  const uint32_t n = (uint32_t)g.n;
  const uint32_t cur = s.cluster_ids[members[0]];
  const uint32_t cnt = (uint32_t)members.size();
  s.cluster_sizes[cur] -= cnt;
  // The moved weight is a float reduce of the node weights (PCCI.cc:416-419).
  std::vector<float> fw(members.size());
  for (size_t i = 0; i < members.size(); ++i) fw[i] = (float)s.node_weights[members[i]];
  const float total = PbbsReduce(fw.data(), fw.size(), 0.0f, [](float a, float b) { return a + b; });
  s.cluster_weights[cur] += (double)(-1 * total);
  if (t != n) {
    s.cluster_sizes[t] += cnt;
    s.cluster_weights[t] += (double)total;
    for (uint32_t u : members) s.cluster_ids[u] = t;
    return true;
  }
  size_t i = cur;
  while (true) {
    if (s.cluster_sizes[i] == 0) {
      s.cluster_sizes[i] = cnt;
      s.cluster_weights[i] += (double)total;
      for (uint32_t u : members) s.cluster_ids[u] = (uint32_t)i;
      return true;
    }
    i = (i + 1) % n;
  }
}

// Sync set moves of the given vertex groups, members ascending (PCC.cc:220-241
// for subclusters, 263-285 for clusters): every group decides on the same
// state, moves only to a larger id than the cluster of its first member, iff
// delta > 0, and the moves are applied as MoveNodesToCluster.  moved_clusters
// is or-ed with the clusters modified.
void GroupMovesSync(const oracle_graph& g, oracle_state& s, const oracle_config& cfg,
                    const std::vector<std::vector<uint32_t>>& groups,
                    std::vector<uint8_t>& moved_clusters) {
  const int64_t n = g.n;
  std::vector<uint8_t> flag((size_t)n, 0);
  std::vector<uint8_t> additional((size_t)n, 0);
  std::vector<uint32_t> moves((size_t)n, kNone);
  for (const auto& grp : groups) {
    if (grp.empty()) continue;
    uint32_t t;
    double d;
    BestMoveSet(g, s, cfg, grp, flag, &t, &d);
    const uint32_t cur = s.cluster_ids[grp.front()];
    if (t < (uint32_t)n && cur >= t) {
      t = cur;
      d = 0;
    }
    if (d > 0)
      for (uint32_t u : grp) moves[u] = t;
  }
  oracle_apply_moves(n, &s, moves.data(), additional.data());
  for (int64_t i = 0; i < n; ++i) moved_clusters[(size_t)i] |= additional[(size_t)i];
}

// Vertex groups by key (key[v] < n), members ascending (OutputIndicesById).
std::vector<std::vector<uint32_t>> GroupsOf(int64_t n, const uint32_t* key) {
  std::vector<std::vector<uint32_t>> groups((size_t)n);
  for (int64_t v = 0; v < n; ++v) groups[key[v]].push_back((uint32_t)v);
  return groups;
}

int64_t SubclusterIds(const oracle_graph& g, const uint32_t* ids, int method,
                      std::vector<uint32_t>* sub, std::vector<uint32_t>* parent, double offset,
                      double res);

// Sync DEFAULT_CLUSTER_MOVES with subclustering, PCC.cc:175-297: the clusters
// after the vertex phase are split into subclusters, the subclusters move as
// sets, then the clusters as they were before the subcluster moves (which may
// span several clusters by now) move as sets.
void SubclusterMovesPhase(const oracle_graph& g, oracle_state& s, const oracle_config& cfg,
                          std::vector<uint8_t>& moved_clusters) {
  const int64_t n = g.n;
  const auto clusters = GroupsOf(n, s.cluster_ids);
  std::vector<uint32_t> sub, parent;
  SubclusterIds(g, s.cluster_ids, cfg.subclustering_method, &sub, &parent,
                cfg.edge_weight_offset, cfg.resolution);
  GroupMovesSync(g, s, cfg, GroupsOf(n, sub.data()), moved_clusters);
  GroupMovesSync(g, s, cfg, clusters, moved_clusters);
}

// The whole-cluster phase of a DEFAULT_CLUSTER_MOVES round, PCC.cc:167-296
// (subclustering NONE): every non-empty cluster, as grouped after the vertex
// phase, tries to move as a whole.  moved_clusters is or-ed with the clusters
// it modifies.
void ClusterMovesPhase(const oracle_graph& g, oracle_state& s, const oracle_config& cfg,
                       std::vector<uint8_t>& moved_clusters) {
  const int64_t n = g.n;
  const std::vector<std::vector<uint32_t>> groups = GroupsOf(n, s.cluster_ids);
  if (!cfg.async) {
    GroupMovesSync(g, s, cfg, groups, moved_clusters);  // PCC.cc:263-285
    return;
  }
  std::vector<uint8_t> flag((size_t)n, 0);
  std::vector<uint8_t> additional((size_t)n, 0);
  for (int64_t c = 0; c < n; ++c) {  // PCC.cc:254-262
    const auto& grp = groups[(size_t)c];
    if (grp.empty()) continue;
    // The reference marks the entries indexed by the members' vertex ids.
    if (AsyncMoveSet(g, s, cfg, grp, flag))
      for (uint32_t u : grp) additional[u] = 1;
  }
  for (int64_t i = 0; i < n; ++i) moved_clusters[(size_t)i] |= additional[(size_t)i];
}

uint64_t SplitMix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Next frontier = neighbours of the seed set (gbbs::edgeMap with an
// always-true map, PCC.cc:46-55, 329-331).
int64_t ExpandFrontier(const oracle_graph& g, const std::vector<uint8_t>& seed,
                       uint8_t* frontier) {
  std::memset(frontier, 0, (size_t)g.n);
  // Every write stores 1, so the set does not depend on the thread count.
#pragma omp parallel for schedule(dynamic, 1024) num_threads(g_threads)
  for (int64_t v = 0; v < g.n; ++v) {
    if (!seed[(size_t)v]) continue;
    for (int64_t p = g.offsets[v]; p < g.offsets[v + 1]; ++p)
      __atomic_store_n(&frontier[g.nbrs[p]], (uint8_t)1, __ATOMIC_RELAXED);
  }
  int64_t cnt = 0;
  for (int64_t v = 0; v < g.n; ++v) cnt += frontier[v];
  return cnt;
}

// BestMovesForVertexSubset of the sequential clusterer,
// correlation-clusterer.cc:96-152: the listed vertices move one after another
// in ascending order with SeqClusteringHelper::AsyncMove
// (correlation-clusterer-internal.cc:87-114), whose new cluster is the first
// empty id counted from 0; a mover marks its destination cluster.
int64_t SeqRound(const oracle_graph& g, oracle_state& s, const oracle_config& cfg,
                 uint8_t* frontier) {
  const int64_t n = g.n;
  std::vector<uint8_t> moved_vertex((size_t)n, 0), moved_clusters((size_t)n, 0);
  for (int64_t v = 0; v < n; ++v) {
    if (!frontier[v]) continue;
    g_edges_scanned += g.offsets[v + 1] - g.offsets[v];
    uint32_t t;
    double d;
    BestMove(g, s, cfg.resolution, cfg.edge_weight_offset, (uint32_t)v, &t, &d, true);
    if (d <= 0) continue;
    const uint32_t cur = s.cluster_ids[v];
    const double k = s.node_weights[v];
    s.cluster_sizes[cur] -= 1;
    s.cluster_weights[cur] -= k;
    if (t != (uint32_t)n) {
      s.cluster_sizes[t] += 1;
      s.cluster_weights[t] += k;
      s.cluster_ids[v] = t;
    } else {
      size_t i = 0;
      while (s.cluster_sizes[i] != 0) i = (i + 1) % (size_t)n;
      s.cluster_sizes[i] = 1;
      s.cluster_weights[i] += k;
      s.cluster_ids[v] = (uint32_t)i;
    }
    moved_vertex[(size_t)v] = 1;
    moved_clusters[s.cluster_ids[v]] = 1;
  }
  if (cfg.move_method == 1) {  // ALL_MOVE, correlation-clusterer.cc:111-118
    std::memset(frontier, 1, (size_t)n);
    return n;
  }
  std::vector<uint8_t> seed((size_t)n);
  for (int64_t v = 0; v < n; ++v)  // correlation-clusterer.cc:125-134
    seed[(size_t)v] = cfg.move_method == 0 ? moved_clusters[s.cluster_ids[v]] : moved_vertex[v];
  return ExpandFrontier(g, seed, frontier);
}

struct OwnedCsr {
  int64_t n = 0, m = 0;
  std::vector<int64_t> offsets;
  std::vector<uint32_t> nbrs;
  std::vector<float> weights;
  std::vector<double> node_weights;
  oracle_graph View() const {
    return oracle_graph{n, m, offsets.data(), nbrs.data(), weights.data()};
  }
};

// CompressGraph, PCCI.cc:709-782, with ComputeInterClusterEdgesSort,
// PGU.cc:156-231, RetrieveInterClusterEdges, PGU.cc:70-122, GetOffsets,
// PGU.cc:33-62.
// seq: SeqCompressGraph (correlation-clusterer-internal.cc:178-279) instead:
// each run is a float left fold from 0 in CSR order (all_edges[cu][cv] =
// aggregate(element, w), :195-198) and node weights are summed from 0 in
// vertex order (:273-275).  The sequential clusterer lists a compressed
// vertex's neighbours in hash-map order (:223-228), which nothing fixes; this
// restatement lists them by ascending id.
void Compress(const oracle_graph& g, const uint32_t* ids, const double* nw,
              OwnedCsr* out, bool seq = false) {
  PhaseTimer timer(3);
  uint32_t max_id = 0;
  for (int64_t i = 0; i < g.n; ++i) max_id = std::max(max_id, ids[i]);
  const int64_t n2 = 1 + (int64_t)max_id;  // PCCI.cc:713-717
  struct E {
    uint32_t cu, cv;
    float w;
  };
  // PGU.cc:93-103: keep (u, v) iff v <= u or the clusters differ, in CSR
  // order (counted per vertex, then written at each vertex's prefix).
  auto keep = [&](int64_t u, int64_t p) {
    const uint32_t v = g.nbrs[p];
    return (int64_t)v <= u || ids[v] != ids[u];
  };
  std::vector<int64_t> kept_at((size_t)g.n + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024) num_threads(g_threads)
  for (int64_t u = 0; u < g.n; ++u) {
    int64_t c = 0;
    for (int64_t p = g.offsets[u]; p < g.offsets[u + 1]; ++p) c += keep(u, p) ? 1 : 0;
    kept_at[(size_t)u + 1] = c;
  }
  for (int64_t u = 0; u < g.n; ++u) kept_at[(size_t)u + 1] += kept_at[(size_t)u];
  std::vector<E> edges((size_t)kept_at[(size_t)g.n]);
#pragma omp parallel for schedule(dynamic, 1024) num_threads(g_threads)
  for (int64_t u = 0; u < g.n; ++u) {
    int64_t q = kept_at[(size_t)u];
    for (int64_t p = g.offsets[u]; p < g.offsets[u + 1]; ++p)
      if (keep(u, p)) edges[(size_t)q++] = E{ids[u], ids[g.nbrs[p]], g.weights[p]};
  }
  StableSort(edges.begin(), edges.end(), [](const E& a, const E& b) {
    return a.cu != b.cu ? a.cu < b.cu : a.cv < b.cv;
  });
  out->n = n2;
  out->offsets.assign((size_t)n2 + 1, 0);
  out->nbrs.clear();
  out->weights.clear();
  std::vector<float> seg;
  std::vector<int64_t> counts((size_t)n2, 0);
  for (size_t i = 0; i < edges.size();) {
    size_t j = i;
    seg.clear();
    while (j < edges.size() && edges[j].cu == edges[i].cu && edges[j].cv == edges[i].cv) {
      seg.push_back(edges[j].w);
      ++j;
    }
    // PGU.cc:206-217: float aggregate through a (double, double) -> double lambda.
    float w = 0.0f;
    if (seq) {
      for (float x : seg) w = (float)((double)w + (double)x);
    } else {
      w = PbbsReduce(seg.data(), seg.size(), 0.0f, [](float a, float b) {
        return (float)((double)a + (double)b);
      });
    }
    out->nbrs.push_back(edges[i].cv);
    out->weights.push_back(w);
    counts[edges[i].cu] += 1;
    i = j;
  }
  out->m = (int64_t)out->nbrs.size();
  for (int64_t c = 0; c < n2; ++c) out->offsets[c + 1] = out->offsets[c] + counts[c];
  // Node weights, PCCI.cc:733-777: stable sort of (id, weight) by id, then a
  // pbbs reduce per group.
  std::vector<uint32_t> order((size_t)g.n);
  std::iota(order.begin(), order.end(), 0u);
  StableSort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return ids[a] < ids[b]; });
  out->node_weights.assign((size_t)n2, 0.0);
  std::vector<double> grp;
  for (size_t i = 0; i < order.size();) {
    size_t j = i;
    grp.clear();
    while (j < order.size() && ids[order[j]] == ids[order[i]]) {
      grp.push_back(nw[order[j]]);
      ++j;
    }
    if (seq) {
      double acc = 0.0;
      for (double x : grp) acc += x;
      out->node_weights[ids[order[i]]] = acc;
    } else {
      out->node_weights[ids[order[i]]] =
          PbbsReduce(grp.data(), grp.size(), 0.0, [](double a, double b) { return a + b; });
    }
    i = j;
  }
}

// ClusteringHelper::SetClustering from per-vertex ids whose groups are
// numbered by ascending id (PCCI.cc:60-77 applied to OutputIndicesById's
// output, PSO.h:183-209).  ids are replaced by the group rank.
void ResetFromIds(int64_t n, oracle_state& s, const std::vector<uint32_t>& flat) {
  std::vector<uint32_t> rank;
  std::vector<uint32_t> sorted(flat);
  std::sort(sorted.begin(), sorted.end());
  sorted.erase(std::unique(sorted.begin(), sorted.end()), sorted.end());
  for (int64_t i = 0; i < n; ++i) {  // ResetClustering, PCCI.cc:50-58
    s.cluster_weights[i] = 0;
    s.cluster_sizes[i] = 0;
  }
  for (int64_t j = 0; j < n; ++j) {  // ascending member order within a group
    const uint32_t r =
        (uint32_t)(std::lower_bound(sorted.begin(), sorted.end(), flat[j]) - sorted.begin());
    s.cluster_ids[j] = r;
    s.cluster_sizes[r] += 1;
    s.cluster_weights[r] += s.node_weights[j];
  }
}

// ---------------------------------------------------------- subclustering --

// CountTriangles, PCCI.cc:786-799, for one edge (u, v): GBBS's vertex
// intersect counts the common entries of the two adjacency lists by a merge
// (it assumes ascending lists), divided by the sum of the two degrees.
double TriangleScore(const oracle_graph& g, uint32_t u, uint32_t v) {
  int64_t i = g.offsets[u], j = g.offsets[v];
  const int64_t ie = g.offsets[(int64_t)u + 1], je = g.offsets[(int64_t)v + 1];
  size_t common = 0;
  while (i < ie && j < je) {
    if (g.nbrs[i] == g.nbrs[j]) {
      ++i;
      ++j;
      ++common;
    } else if (g.nbrs[i] < g.nbrs[j]) {
      ++i;
    } else {
      ++j;
    }
  }
  return (double)common / (double)((ie - g.offsets[u]) + (je - g.offsets[v]));
}

// The numbering of CompressSubclusters (PCC.cc:436-474) with
// ComputeSubcluster (PCCI.cc:1101-1120).  Clusters are taken in ascending id
// order with their members ascending (OutputIndicesById,
// parallel-sequence-ops.h:184-209).  Within cluster r, member i gets label
// root(i) from a union-find over the cluster's members:
//   method 2 CONNECTIVITY (PCCI.cc:846-895): every edge between two members;
//   method 1 TRIANGLE (PCCI.cc:1015-1083): edges with
//     TriangleScore > 0.25 between two members.
// GBBS's SimpleUnionAsyncStruct links the larger root under the smaller one,
// so root(i) is the smallest member index of i's component.  (TRIANGLE's
// workefficient_cc::CC labels come from a randomized decomposition; the
// smallest index is used for them too.)  A cluster spans size (CONNECTIVITY)
// or 1 + its largest label (TRIANGLE) ids; the spans are laid out in cluster
// order, sub[v] = span start + label, and every id j of cluster r's span
// starts the next level in cluster r (PCC.cc:457-462, PCCI.cc:60-77).
// Returns the sum of the spans; parent[j] is filled for j < min(sum, n).
// ComputeSubclusterConnectivitySlow, PCCI.cc:897-1013, for one cluster cl
// (members ascending), restated step by step: members are visited in order;
// a member still alone, with at least |cl| - 1 edges leaving the cluster,
// looks at the current groups (OutputIndicesById of tmp: ascending value,
// members ascending) other than its own; a group is valid if its edges
// leaving the cluster number at least |group| (|cl| - |group|) and the
// member's edges into it at least offset * |group| + resolution * |group|
// (one fused multiply-add under GCC's contraction); the member joins the
// last valid group with the largest surplus.  fast marks the cluster (1) and
// the group in hand (2), as the reference's fast_intersect; it is all zero on
// entry and on exit.  Returns the labels tmp (indexed like cl).
std::vector<uint32_t> SlowConnectivityLabels(const oracle_graph& g, const std::vector<uint32_t>& cl,
                                             double offset, double res,
                                             std::vector<uint8_t>& fast) {
  const size_t size = cl.size();
  std::vector<uint32_t> tmp(size);
  std::iota(tmp.begin(), tmp.end(), 0u);
  std::vector<uint8_t> singletons(size, 0);
  for (uint32_t v : cl) fast[v] = 1;
  auto edges_to = [&](uint32_t v, uint8_t mark) {
    int count = 0;
    for (int64_t p = g.offsets[v]; p < g.offsets[(int64_t)v + 1]; ++p)
      if (fast[g.nbrs[p]] == mark) ++count;
    return count;
  };
  for (size_t i = 0; i < size; ++i) {
    if (singletons[i] != 0) continue;
    const uint32_t v = cl[i];
    const int curr_num_edges = edges_to(v, 0);
    if ((size_t)curr_num_edges < size - 1) continue;
    std::vector<std::vector<uint32_t>> groups(size);
    for (size_t j = 0; j < size; ++j) groups[tmp[j]].push_back((uint32_t)j);
    groups.erase(std::remove_if(groups.begin(), groups.end(),
                                [](const std::vector<uint32_t>& x) { return x.empty(); }),
                 groups.end());
    std::vector<uint8_t> invalid(groups.size(), 0);
    size_t num_valid = groups.size();
    for (size_t j = 0; j < groups.size(); ++j) {
      const auto& curr = groups[j];
      if (curr[0] == i) {
        invalid[j] = 1;
        --num_valid;
        continue;
      }
      for (uint32_t k : curr) fast[cl[k]] = 2;
      int num_edges = 0;
      for (uint32_t k : curr) num_edges += edges_to(cl[k], 0);
      for (uint32_t k : curr) fast[cl[k]] = 1;
      if ((size_t)num_edges < curr.size() * (size - curr.size())) {
        invalid[j] = 1;
        --num_valid;
      }
    }
    if (num_valid == 0) continue;
    double max_max = 0;
    size_t idx = 0;
    for (size_t j = 0; j < groups.size(); ++j) {
      if (invalid[j]) continue;
      const auto& curr = groups[j];
      for (uint32_t k : curr) fast[cl[k]] = 2;
      const double weight = (double)edges_to(v, 2);
      const double cs = (double)curr.size();
      const double off = std::fma(offset, cs, res * cs);
      for (uint32_t k : curr) fast[cl[k]] = 1;
      if (weight < off) {
        invalid[j] = 1;
        --num_valid;
      } else if (weight - off >= max_max) {
        max_max = weight - off;
        idx = j;
      }
    }
    if (num_valid == 0) continue;
    const auto& curr = groups[idx];
    tmp[i] = tmp[curr[0]];
    singletons[i] = 1;
    singletons[curr[0]] = 1;
  }
  for (uint32_t v : cl) fast[v] = 0;
  return tmp;
}

int64_t SubclusterIds(const oracle_graph& g, const uint32_t* ids, int method,
                      std::vector<uint32_t>* sub, std::vector<uint32_t>* parent, double offset,
                      double res) {
  const int64_t n = g.n;
  std::vector<uint8_t> fast(method == 3 ? (size_t)n : 0, 0);
  std::vector<std::vector<uint32_t>> members((size_t)n);
  for (int64_t v = 0; v < n; ++v) members[ids[v]].push_back((uint32_t)v);
  std::vector<int64_t> index((size_t)n, 0);
  sub->assign((size_t)n, 0);
  parent->assign((size_t)n, 0);
  std::vector<uint32_t> uf;
  auto find = [&](uint32_t x) {
    while (uf[x] != x) x = uf[x];
    return x;
  };
  int64_t next = 0, rank = 0;
  for (int64_t c = 0; c < n; ++c) {
    const auto& cl = members[(size_t)c];
    if (cl.empty()) continue;
    const uint32_t size = (uint32_t)cl.size();
    if (method == 3) {
      const std::vector<uint32_t> tmp = SlowConnectivityLabels(g, cl, offset, res, fast);
      uint32_t max_label = 0;
      for (uint32_t i = 0; i < size; ++i) {
        (*sub)[cl[i]] = (uint32_t)(next + tmp[i]);
        max_label = std::max(max_label, tmp[i]);
      }
      const int64_t span = (int64_t)max_label + 1;  // PCCI.cc:1006-1012
      for (int64_t j = next; j < next + span && j < n; ++j) (*parent)[(size_t)j] = (uint32_t)rank;
      next += span;
      ++rank;
      continue;
    }
    for (uint32_t i = 0; i < size; ++i) index[cl[i]] = i;
    uf.resize(size);
    std::iota(uf.begin(), uf.end(), 0u);
    for (uint32_t i = 0; i < size; ++i) {
      const uint32_t u = cl[i];
      for (int64_t p = g.offsets[u]; p < g.offsets[(int64_t)u + 1]; ++p) {
        const uint32_t v = g.nbrs[p];
        if (ids[v] != (uint32_t)c) continue;
        if (method == 1 && !(TriangleScore(g, u, v) > 0.25)) continue;
        uint32_t a = find(i), b = find((uint32_t)index[v]);
        if (a == b) continue;
        if (a < b) std::swap(a, b);
        uf[a] = b;
      }
    }
    uint32_t max_label = 0;
    for (uint32_t i = 0; i < size; ++i) {
      const uint32_t r = find(i);
      (*sub)[cl[i]] = (uint32_t)(next + r);
      max_label = std::max(max_label, r);
    }
    const int64_t span = method == 2 ? (int64_t)size : (int64_t)max_label + 1;
    for (int64_t j = next; j < next + span && j < n; ++j) (*parent)[(size_t)j] = (uint32_t)rank;
    next += span;
    ++rank;
  }
  return next;
}

struct Level {
  OwnedCsr csr;  // unused for level 0 (the input graph)
  // With subclustering: the next level's vertex of every vertex of this one.
  std::vector<uint32_t> map;
  std::vector<uint32_t> ids, sizes;
  std::vector<double> nw, K;
  oracle_state State() {
    return oracle_state{ids.data(), sizes.data(), nw.data(), K.data()};
  }
};

}  // namespace

extern "C" {

void oracle_set_threads(int threads) {
  g_threads = threads < 1 ? 1 : threads;
}
int oracle_get_threads(void) { return g_threads; }

int64_t oracle_best_move_candidates(const oracle_graph* g, const oracle_state* s,
                                    const oracle_config* cfg, uint32_t v, uint32_t* clusters,
                                    double* deltas, int64_t cap, double* out_change) {
  Candidates c;
  uint32_t t;
  double d;
  BestMove(*g, *s, cfg->resolution, cfg->edge_weight_offset, v, &t, &d, false, &c);
  for (int64_t i = 0; i < (int64_t)c.list.size() && i < cap; ++i) {
    clusters[i] = c.list[(size_t)i].first;
    deltas[i] = c.list[(size_t)i].second;
  }
  if (out_change) *out_change = c.change;
  return (int64_t)c.list.size();
}

void oracle_phase_seconds(double out[5], int reset) {
  for (int i = 0; i < 5; ++i) {
    out[i] = g_phase[i];
    if (reset) g_phase[i] = 0;
  }
}

void oracle_best_move(const oracle_graph* g, const oracle_state* s,
                      const oracle_config* cfg, uint32_t v, uint32_t* out_target,
                      double* out_delta) {
  BestMove(*g, *s, cfg->resolution, cfg->edge_weight_offset, v, out_target, out_delta);
}

void oracle_best_move_set(const oracle_graph* g, const oracle_state* s,
                          const oracle_config* cfg, const uint32_t* members, int64_t count,
                          uint32_t* out_target, double* out_delta) {
  std::vector<uint32_t> set(members, members + count);
  std::vector<uint8_t> flag((size_t)g->n, 0);
  BestMoveSet(*g, *s, *cfg, set, flag, out_target, out_delta);
}

void oracle_sync_decisions(const oracle_graph* g, const oracle_state* s,
                           const oracle_config* cfg, const uint8_t* frontier,
                           uint32_t* out_moves, uint8_t* out_moved,
                           uint32_t* out_target, double* out_delta) {
  const int64_t n = g->n;
#pragma omp parallel for schedule(dynamic, 512) num_threads(g_threads)
  for (int64_t v = 0; v < n; ++v) {
    if (frontier && !frontier[v]) {
      out_moves[v] = kNone;
      out_moved[v] = 0;
      if (out_target) out_target[v] = kNone;
      if (out_delta) out_delta[v] = 0.0;
      continue;
    }
    SyncDecision(*g, *s, *cfg, (uint32_t)v, &out_moves[v], &out_moved[v],
                 out_target ? &out_target[v] : nullptr, out_delta ? &out_delta[v] : nullptr);
  }
}

void oracle_apply_moves(int64_t n, oracle_state* s, const uint32_t* moves,
                        uint8_t* out_modified) {
  // ClusteringHelper::MoveNodesToCluster(moves), PCCI.cc:256-369.
  std::memset(out_modified, 0, (size_t)n);
  std::vector<uint32_t> moving;
  for (int64_t v = 0; v < n; ++v)
    if (moves[v] != kNone && moves[v] != s->cluster_ids[v]) moving.push_back((uint32_t)v);
  if (moving.empty()) return;
  std::vector<uint32_t> by_old(moving);
  StableSort(by_old.begin(), by_old.end(), [&](uint32_t a, uint32_t b) {
    return s->cluster_ids[a] < s->cluster_ids[b];
  });
  for (size_t i = 0; i < by_old.size();) {  // PCCI.cc:294-303
    size_t j = i;
    const uint32_t prev = s->cluster_ids[by_old[i]];
    while (j < by_old.size() && s->cluster_ids[by_old[j]] == prev) ++j;
    s->cluster_sizes[prev] -= (uint32_t)(j - i);
    out_modified[prev] = 1;
    for (size_t q = i; q < j; ++q) s->cluster_weights[prev] -= s->node_weights[by_old[q]];
    i = j;
  }
  std::vector<uint32_t> by_new(moving);
  StableSort(by_new.begin(), by_new.end(),
             [&](uint32_t a, uint32_t b) { return moves[a] < moves[b]; });
  size_t new_start = by_new.size();
  for (size_t i = 0; i < by_new.size();) {  // PCCI.cc:325-337
    size_t j = i;
    const uint32_t move_id = moves[by_new[i]];
    while (j < by_new.size() && moves[by_new[j]] == move_id) ++j;
    if (move_id != (uint32_t)n) {
      s->cluster_sizes[move_id] += (uint32_t)(j - i);
      out_modified[move_id] = 1;
      for (size_t q = i; q < j; ++q) {
        s->cluster_ids[by_new[q]] = move_id;
        s->cluster_weights[move_id] += s->node_weights[by_new[q]];
      }
    } else {
      new_start = i;
    }
    i = j;
  }
  if (moves[by_new.back()] == (uint32_t)n) {  // PCCI.cc:340-366
    std::vector<uint32_t> zero;
    for (int64_t id = 0; id < n; ++id)
      if (s->cluster_sizes[id] == 0) zero.push_back((uint32_t)id);
    for (size_t i = new_start; i < by_new.size(); ++i) {
      const uint32_t c = zero[i - new_start];
      const uint32_t v = by_new[i];
      s->cluster_ids[v] = c;
      s->cluster_sizes[c] = 1;
      out_modified[c] = 1;
      s->cluster_weights[c] = s->node_weights[v];
    }
  }
}

int64_t oracle_round(const oracle_graph* g, oracle_state* s,
                     const oracle_config* cfg, uint8_t* frontier,
                     uint64_t round_seed) {
  // BestMovesForVertexSubset, PCC.cc:115-338.
  if (cfg->sequential) return SeqRound(*g, *s, *cfg, frontier);
  const int64_t n = g->n;
  for (int64_t v = 0; v < n; ++v)
    if (frontier[v]) g_edges_scanned += g->offsets[v + 1] - g->offsets[v];
  std::vector<uint8_t> moved_vertex((size_t)n, 0);
  std::vector<uint8_t> moved_clusters((size_t)n, 0);
  if (cfg->async) {
    std::vector<uint32_t> order;
    for (int64_t v = 0; v < n; ++v)
      if (frontier[v]) order.push_back((uint32_t)v);
    if (cfg->permute) {  // vertexMapPermute, PCC.cc:69-109 (own seeded shuffle)
      uint64_t st = SplitMix64(cfg->permute_seed ^ SplitMix64(round_seed));
      for (size_t i = order.size(); i > 1; --i) {
        st = SplitMix64(st);
        std::swap(order[i - 1], order[(size_t)(st % i)]);
      }
    }
    for (uint32_t v : order) {  // PCC.cc:133-138
      if (AsyncMove(*g, *s, *cfg, v)) {
        moved_clusters[s->cluster_ids[v]] = 1;
        moved_vertex[v] = 1;
      }
    }
  } else {
    std::vector<uint32_t> moves((size_t)n);
    {
      PhaseTimer t(0);
      oracle_sync_decisions(g, s, cfg, frontier, moves.data(), moved_vertex.data(), nullptr,
                            nullptr);
    }
    PhaseTimer t(1);
    oracle_apply_moves(n, s, moves.data(), moved_clusters.data());  // PCC.cc:163-165
  }
  if (cfg->clustering_moves_method == 0) {  // DEFAULT_CLUSTER_MOVES, PCC.cc:167-297
    // Subcluster moves are restated for sync rounds only.
    if (cfg->subclustering_method >= 1 && cfg->subclustering_method <= 3 && !cfg->async)
      SubclusterMovesPhase(*g, *s, *cfg, moved_clusters);
    else
      ClusterMovesPhase(*g, *s, *cfg, moved_clusters);
  }
  if (cfg->move_method == 1) {  // ALL_MOVE, PCC.cc:299-307
    std::memset(frontier, 1, (size_t)n);
    return n;
  }
  PhaseTimer t(2);
  std::vector<uint8_t> seed((size_t)n);
  for (int64_t v = 0; v < n; ++v)  // PCC.cc:312-324
    seed[(size_t)v] = cfg->move_method == 0 ? moved_clusters[s->cluster_ids[v]] : moved_vertex[v];
  return ExpandFrontier(*g, seed, frontier);
}

int oracle_iterate_best_moves(const oracle_graph* g, oracle_state* s,
                              const oracle_config* cfg, int64_t inner, uint64_t seed,
                              int64_t* out_rounds) {
  // IterateBestMoves, PCC.cc:340-365.
  std::vector<uint8_t> frontier((size_t)g->n, 1);
  bool moved = false, local_moved = true;
  int64_t it = 0;
  // all_iter (sequential clusterer): until nothing moves,
  // correlation-clusterer.cc:173-183.
  const bool until_done = cfg->sequential && cfg->all_iter;
  for (it = 0; (until_done || it < inner) && local_moved; ++it) {
    const int64_t sz = oracle_round(g, s, cfg, frontier.data(), seed * 1315423911ull + (uint64_t)it);
    local_moved = sz != 0;
    moved |= local_moved;
  }
  if (out_rounds) *out_rounds += it;
  return moved ? 1 : 0;
}

oracle_csr* oracle_compress(const oracle_graph* g, const uint32_t* cluster_ids,
                            const double* node_weights) {
  auto* c = new OwnedCsr();
  Compress(*g, cluster_ids, node_weights, c);
  return reinterpret_cast<oracle_csr*>(c);
}

void oracle_csr_sizes(const oracle_csr* c, int64_t* n, int64_t* m) {
  const auto* o = reinterpret_cast<const OwnedCsr*>(c);
  *n = o->n;
  *m = o->m;
}

void oracle_csr_copy(const oracle_csr* c, int64_t* offsets, uint32_t* nbrs, float* weights,
                     double* node_weights) {
  const auto* o = reinterpret_cast<const OwnedCsr*>(c);
  if (offsets) std::copy(o->offsets.begin(), o->offsets.end(), offsets);
  if (nbrs) std::copy(o->nbrs.begin(), o->nbrs.end(), nbrs);
  if (weights) std::copy(o->weights.begin(), o->weights.end(), weights);
  if (node_weights) std::copy(o->node_weights.begin(), o->node_weights.end(), node_weights);
}

void oracle_csr_free(oracle_csr* c) { delete reinterpret_cast<OwnedCsr*>(c); }

double oracle_objective(const oracle_graph* g, const uint32_t* ids, const double* nw,
                        double resolution, double offset) {
  // ClusteringHelper::ComputeObjective, PCCI.cc:118-149.
  const int64_t n = g->n;
  std::vector<double> K((size_t)n, 0.0);
  for (int64_t v = 0; v < n; ++v) K[ids[v]] += nw[v];
  std::vector<double> shifted((size_t)n), tmp;
  for (int64_t i = 0; i < n; ++i) {
    tmp.clear();
    for (int64_t p = g->offsets[i]; p < g->offsets[i + 1]; ++p) {
      const float weight = g->weights[p];
      tmp.push_back(ids[i] == ids[g->nbrs[p]] ? (weight - offset) / 2 : 0.0);
    }
    shifted[(size_t)i] =
        PbbsReduce(tmp.data(), tmp.size(), 0.0, [](double a, double b) { return a + b; });
  }
  auto add = [](double a, double b) { return a + b; };
  double objective = PbbsReduce(shifted.data(), shifted.size(), 0.0, add);
  std::vector<double> rs((size_t)n);
  for (int64_t i = 0; i < n; ++i) rs[(size_t)i] = nw[i] * (K[ids[i]] - nw[i]);
  objective -= resolution * PbbsReduce(rs.data(), rs.size(), 0.0, add) / 2;
  return objective;
}

double oracle_modularity(const oracle_graph* g, const uint32_t* ids, double resolution) {
  // ComputeModularity, PMC.cc:36-66.  Clusters are visited in ascending id
  // order, as OutputIndicesById lists them.
  double total_edge_weight = 0;
  double modularity = 0;
  for (int64_t i = 0; i < g->n; ++i) {
    for (int64_t p = g->offsets[i]; p < g->offsets[i + 1]; ++p) {
      total_edge_weight++;
      if (ids[i] == ids[g->nbrs[p]]) modularity++;
    }
  }
  std::vector<std::pair<uint32_t, uint32_t>> order((size_t)g->n);
  for (int64_t i = 0; i < g->n; ++i) order[(size_t)i] = {ids[i], (uint32_t)i};
  std::sort(order.begin(), order.end());
  for (size_t i = 0; i < order.size();) {
    size_t j = i;
    double degree = 0;
    while (j < order.size() && order[j].first == order[i].first) {
      const uint32_t v = order[j].second;
      degree += (double)(g->offsets[v + 1] - g->offsets[v]);
      ++j;
    }
    modularity -= (resolution * degree * degree) / (total_edge_weight);
    i = j;
  }
  modularity = modularity / (total_edge_weight);
  return modularity;
}

int64_t oracle_subclusters(const oracle_graph* g, const uint32_t* ids, int method, uint32_t* sub,
                           uint32_t* parent, double offset, double res) {
  if (method < 1 || method > 3) return -1;
  std::vector<uint32_t> s, p;
  const int64_t total = SubclusterIds(*g, ids, method, &s, &p, offset, res);
  std::copy(s.begin(), s.end(), sub);
  std::copy(p.begin(), p.end(), parent);
  return total;
}

double oracle_triangle_score(const oracle_graph* g, uint32_t u, uint32_t v) {
  return TriangleScore(*g, u, v);
}

void oracle_flatten(int64_t n, const uint32_t* ids, const uint32_t* compressed, uint32_t* out) {
  for (int64_t i = 0; i < n; ++i)
    out[i] = ids[i] == kNone ? kNone : compressed[ids[i]];  // PGU.cc:237-241
}

void oracle_rank_relabel(int64_t n, const uint32_t* ids, uint32_t* out) {
  std::vector<uint32_t> sorted(ids, ids + n);
  std::sort(sorted.begin(), sorted.end());
  sorted.erase(std::unique(sorted.begin(), sorted.end()), sorted.end());
  for (int64_t i = 0; i < n; ++i)
    out[i] = (uint32_t)(std::lower_bound(sorted.begin(), sorted.end(), ids[i]) - sorted.begin());
}

int oracle_cluster(const oracle_graph* g, const oracle_config* cfg_in, int modularity,
                   const double* node_weights, const uint32_t* initial_ids, uint32_t* out_ids,
                   oracle_result* out) {
  oracle_config cfg = *cfg_in;
  const int64_t n = g->n;
  std::vector<double> nw0((size_t)n, 1.0);
  if (node_weights) {
    std::copy(node_weights, node_weights + n, nw0.begin());
  } else if (modularity) {
    // ComputeModularityConfig, PGU.cc:141-153.
    for (int64_t v = 0; v < n; ++v) nw0[(size_t)v] = (double)(g->offsets[v + 1] - g->offsets[v]);
  }
  const double gamma = cfg.resolution;
  if (modularity) {  // PMC.cc:74-79
    const double total = PbbsReduce(nw0.data(), nw0.size(), 0.0,
                                    [](double a, double b) { return a + b; });
    cfg.resolution = gamma / total;
    cfg.edge_weight_offset = 0;
  }
  // RefineClusters_subroutine, PCC.cc:476-629.
  const auto t0 = std::chrono::steady_clock::now();
  int64_t num_iterations = 0, num_inner = 0;
  if (cfg.clustering_moves_method == 1) {  // PCC.cc:493-500
    num_iterations = cfg.louvain_num_iterations > 0 ? cfg.louvain_num_iterations : 10;
    num_inner = cfg.louvain_num_inner_iterations > 0 ? cfg.louvain_num_inner_iterations : 10;
  } else if (cfg.clustering_moves_method == 0) {  // DEFAULT_CLUSTER_MOVES, PCC.cc:488-492
    num_iterations = 1;
    num_inner = cfg.num_iterations > 0 ? cfg.num_iterations : 20;
  } else {
    return 12;  // PCC.cc:501-505
  }
  if (cfg.sequential && cfg.all_iter) num_iterations = INT64_MAX;  // correlation-clusterer.cc:266-267
  const bool subclust = cfg.subclustering_method != 0 && !cfg.sequential;
  if (subclust && (cfg.subclustering_method > 3 || (cfg.clustering_moves_method == 0 && cfg.async)))
    return 12;  // async subcluster moves are not restated
  int64_t rounds = 0;
  g_edges_scanned = 0;
  std::vector<Level> levels;
  levels.reserve((size_t)std::min<int64_t>(num_iterations, 64) + 1);
  levels.emplace_back();
  {
    Level& L = levels.back();
    L.nw = nw0;
    L.ids.assign((size_t)n, 0);
    L.sizes.assign((size_t)n, 0);
    L.K.assign((size_t)n, 0.0);
    if (initial_ids) {  // SetClustering with a non-empty clustering, PCCI.cc:68-75
      for (int64_t j = 0; j < n; ++j) {
        L.ids[(size_t)j] = initial_ids[j];
        L.sizes[initial_ids[j]] += 1;
        L.K[initial_ids[j]] += L.nw[(size_t)j];
      }
    } else {
      for (int64_t j = 0; j < n; ++j) {
        L.ids[(size_t)j] = (uint32_t)j;
        L.sizes[(size_t)j] = 1;
        L.K[(size_t)j] = 0 + L.nw[(size_t)j];
      }
    }
  }
  std::vector<uint32_t> cluster_ids((size_t)n);
  std::iota(cluster_ids.begin(), cluster_ids.end(), 0u);
  auto graph_of = [&](size_t lvl) { return lvl == 0 ? *g : levels[lvl].csr.View(); };
  int64_t iter = 0;
  for (iter = 0; iter < num_iterations; ++iter) {
    oracle_graph cg = graph_of((size_t)iter);
    oracle_state st = levels[(size_t)iter].State();
    const int moved = oracle_iterate_best_moves(&cg, &st, &cfg, num_inner,
                                                cfg.permute_seed + (uint64_t)iter, &rounds);
    if (!moved) {  // PCC.cc:543-547
      // With subclustering the level started from its parent clustering,
      // which is then the result.
      if (subclust && !cfg.refine && iter > 0) {
        std::vector<uint32_t> flat((size_t)n);
        oracle_flatten(n, cluster_ids.data(), levels[(size_t)iter].ids.data(), flat.data());
        cluster_ids.swap(flat);
      }
      iter--;
      break;
    }
    const bool last = iter == num_iterations - 1;
    if (!cfg.refine && (!subclust || last)) {  // PCC.cc:550
      std::vector<uint32_t> flat((size_t)n);
      oracle_flatten(n, cluster_ids.data(), levels[(size_t)iter].ids.data(), flat.data());
      cluster_ids.swap(flat);
    }
    if (last) break;          // PCC.cc:556
    levels.emplace_back();    // CompressGraph, PCC.cc:569-571
    Level& cur = levels[(size_t)iter];
    Level& nxt = levels.back();
    if (subclust) {
      // CompressSubclusters (PCC.cc:564-567): the subclusters become the next
      // level's vertices and start in their parent cluster.  The next level's
      // vertex of a vertex is its subcluster id, which is what the flatten
      // steps compose (the reference composes cluster ids there, PCC.cc:550
      // and 594-595, which index the next level by the wrong vertex).
      std::vector<uint32_t> parent;
      SubclusterIds(cg, cur.ids.data(), cfg.subclustering_method, &cur.map, &parent,
                    cfg.edge_weight_offset, cfg.resolution);
      Compress(cg, cur.map.data(), cur.nw.data(), &nxt.csr);
      const int64_t n2 = nxt.csr.n;
      nxt.nw = nxt.csr.node_weights;
      nxt.ids.resize((size_t)n2);
      nxt.sizes.assign((size_t)n2, 0);
      nxt.K.assign((size_t)n2, 0.0);
      for (int64_t j = 0; j < n2; ++j) {  // SetClustering, PCCI.cc:68-75
        nxt.ids[(size_t)j] = parent[(size_t)j];
        nxt.sizes[parent[(size_t)j]] += 1;
        nxt.K[parent[(size_t)j]] += nxt.nw[(size_t)j];
      }
      if (!cfg.refine) {
        std::vector<uint32_t> flat((size_t)n);
        oracle_flatten(n, cluster_ids.data(), cur.map.data(), flat.data());
        cluster_ids.swap(flat);
      }
    } else {
      Compress(cg, cur.ids.data(), cur.nw.data(), &nxt.csr, cfg.sequential != 0);
      const int64_t n2 = nxt.csr.n;
      nxt.nw = nxt.csr.node_weights;  // singletons with aggregated weights, PCC.cc:583-585
      nxt.ids.resize((size_t)n2);
      nxt.sizes.assign((size_t)n2, 1);
      nxt.K.resize((size_t)n2);
      for (int64_t i = 0; i < n2; ++i) {
        nxt.ids[(size_t)i] = (uint32_t)i;
        nxt.K[(size_t)i] = nxt.nw[(size_t)i];
      }
    }
    if (!cfg.refine) {
      // Without refinement only the current level is needed.
      cur.ids.clear();
      cur.ids.shrink_to_fit();
    }
  }
  if (cfg.refine && iter > 0) {  // PCC.cc:589-607
    for (int64_t i = iter - 1; i >= 0; --i) {
      oracle_graph cg = graph_of((size_t)i);
      Level& L = levels[(size_t)i];
      const int64_t ni = cg.n;
      std::vector<uint32_t> flat((size_t)ni);
      oracle_flatten(ni, subclust ? L.map.data() : L.ids.data(), levels[(size_t)i + 1].ids.data(),
                     flat.data());
      oracle_state st = L.State();
      {
        PhaseTimer t(4);
        ResetFromIds(ni, st, flat);
      }
      oracle_iterate_best_moves(&cg, &st, &cfg, num_inner, cfg.permute_seed + 7919ull * (uint64_t)(i + 1),
                                &rounds);
    }
    cluster_ids = levels[0].ids;
  }
  const auto t1 = std::chrono::steady_clock::now();
  std::copy(cluster_ids.begin(), cluster_ids.end(), out_ids);
  if (out) {
    out->seconds = std::chrono::duration<double>(t1 - t0).count();
    out->levels = iter + 1;
    out->rounds = rounds;
    out->edges_scanned = g_edges_scanned;
    std::vector<uint32_t> ranks((size_t)n);
    oracle_rank_relabel(n, cluster_ids.data(), ranks.data());
    uint32_t k = 0;
    for (int64_t v = 0; v < n; ++v) k = std::max(k, ranks[(size_t)v] + 1);
    out->num_clusters = n ? k : 0;
    out->objective = oracle_objective(g, ranks.data(), nw0.data(), cfg.resolution,
                                      cfg.edge_weight_offset);
    out->modularity = modularity ? oracle_modularity(g, ranks.data(), gamma) : 0.0;
  }
  return 0;
}

}  // extern "C"
