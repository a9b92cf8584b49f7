// engine.h -- the graph handle and the multilevel Louvain engine.
//
// One engine drives one graph or a batch of independent graphs.  A batch is
// laid out as one disjoint union: graph g owns the vertex range
// [base[g], base[g + 1]) of every level, cluster ids are vertex ids of the
// level, and per-graph quantities (resolution, level progress, termination)
// are kept per graph, so each graph follows exactly the control flow of its
// own ParallelCorrelationClusterer run.
#ifndef PCC_AMD_ENGINE_H_
#define PCC_AMD_ENGINE_H_

#include <hip/hip_runtime.h>

#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "kernels.h"
#include "pcc_amd.h"
#include "status.h"

struct pcc_graph {
  int device = 0;
  hipStream_t stream = nullptr;
  pcc::DeviceAllocator alloc;
  // Import staging (GbbsGraph, clustering/gbbs-graph.cc:32-76).
  std::mutex mu;
  std::vector<std::vector<std::pair<uint32_t, float>>> adj;
  bool finished = false;
  pcc::DevCsr csr;
  // Batch layout: graph g owns vertices [gbase[g], gbase[g + 1]); one graph
  // has gbase = {0, n} and no vgraph.
  std::vector<int64_t> gbase;
  std::vector<int64_t> goff;  // CSR offset at every gbase entry (edges before graph g)
  pcc::DevBuf<uint32_t> vgraph;
  // Options.
  bool timing = false;
  bool force_ordered = false;
  int async_subrounds = 0;  // 0: automatic (see Engine::RoundStep)
  int async_singleton_rule = 0;
  int cluster_subrounds = (int)pcc::kClusterSubRounds;  // async whole-cluster moves
  // Sync LOUVAIN rounds driven from device (PCC_OPT_DEVICE_ROUNDS): 0 host
  // read-back per round (measured fastest, the default), 1 graph replay,
  // 2 device counts with plain launches.
  int device_rounds = 0;
  int64_t hub_part_edges = pcc::kHubPartEdges;
  int contract_method = 0;  // PCC_OPT_CONTRACT_METHOD
  int64_t contract_chunk = pcc::kContractChunkTerms;  // PCC_OPT_CONTRACT_CHUNK
  // Pull expansion when the seed volume exceeds pull_fraction / 1024 of the
  // graph's (edges + vertices); 1024 disables it.
  int pull_fraction = 512;
  // Edge multiset symmetric: -1 unknown, 0 no, 1 yes (set on first use).
  int symmetric = -1;
  pcc::BinThresholds th;
  int num_graphs() const { return gbase.empty() ? 1 : (int)gbase.size() - 1; }
  ~pcc_graph() {
    vgraph.Release();
    csr = pcc::DevCsr();
    alloc.ReleaseAll();
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace pcc {

int Guard(const char* where, const std::function<int()>& body);
int NewHandle(pcc_graph** out);
// Uploads a CSR (neighbour ids global) and validates it on device.
int UploadCsr(pcc_graph* h, int64_t n, int64_t m, const int64_t* off, const uint32_t* nbrs,
              const float* w);

// Per-level layout of a batch.
struct Layout {
  int G = 1;                   // graphs present at this level
  std::vector<int> orig;       // original graph index of each present graph
  std::vector<int64_t> base;   // G + 1 vertex bases
  DevBuf<int64_t> dbase;       // device copy of base (batches only)
  DevBuf<uint32_t> vgraph;     // present-graph index of each vertex (batches only)
  DevBuf<double> gres;         // resolution of each present graph (batches only)
  bool batch = false;
};

struct Level {
  DevCsr own;   // contracted graphs (level > 0)
  CsrView g;    // this level's graph
  // With subclustering: the next level's vertex (subcluster id) of every
  // vertex of this one, kept for the refinement flatten.
  DevBuf<uint32_t> map;
  DevBuf<uint32_t> ids, sizes;
  DevBuf<double> nw, K;
  Layout lay;
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
  Engine(pcc_graph* h, const pcc_config& cfg, pcc_stats* stats);
  ~Engine();

  // RefineClusters_subroutine (parallel-correlation-clusterer.cc:476-629) for
  // every graph of the handle.  nw0: level-0 node weights (taken over);
  // gres: per-graph resolutions (null: cfg.resolution for all); init: device
  // initial cluster ids (global vertex ids of the level) or null.
  // host_out receives each vertex's final cluster id local to its graph;
  // out_iter (may be null) the final level index of each graph; dev_global
  // (may be null) the final ids on device as vertex ids of the whole batch.
  int Run(DevBuf<double> nw0, const double* gres, const uint32_t* init, uint32_t* host_out,
          std::vector<int64_t>* out_iter, DevBuf<uint32_t>* dev_global = nullptr);

  struct Workspace {
    DevBuf<uint32_t> list, list2, fb_list, moves;
    DevBuf<uint32_t> sub_cnt, sub_cur;  // async sub-round segments, kBins * 1024
    DevBuf<unsigned long long> sub_te;  // bin-7 degree sum per sub-round, 1024
    DevBuf<uint8_t> moved, flags, next, new_flag;
    DevBuf<RoundCounters> rc;
    // Device-driven rounds: the current round's list counts, the statistics
    // of the call's rounds, and the new-cluster assignment's temporaries.
    DevBuf<DeviceCounts> dc;
    DevBuf<DeviceRoundStats> dstats;
    NewAssignWork na;
    // Clears new_flag, rc and (optionally) also_bytes bytes at also_zero in
    // one launch.
    Workspace(DeviceAllocator* al, int64_t n, hipStream_t s, uint8_t* also_zero = nullptr,
              int64_t also_bytes = 0);
  };

  // One level's best-move machinery, shared by Run and the one-phase C ABI.
  struct LevelCtx {
    CsrView g;
    StateView st;
    MoveParams mp;
    BinLayout layout;
    int64_t caps[kBins];     // vertices of each bin on this level
    int64_t caps_te[kBins];  // edges of each bin on this level
    DevBuf<int32_t> gscale;  // MoveParams::gscale
    // Ordered fallback scratch.
    HashScratch hs;
    DevBuf<uint32_t> hkeys;
    DevBuf<double> hvals;
    // Bin-7 scratch, sized for every bin-7 vertex of the level.
    HubScratch hb;
    DevBuf<int64_t> hb_part, hb_ent;
    DevBuf<uint32_t> hb_flag, hb_pc1, hb_item_of, hb_bcur, hb_pfine, hb_cbuf;
    DevBuf<uint2> hb_bkt;
    DevBuf<HubPart> hb_hpart;
    DevBuf<double> hb_tval, hb_px1, hb_px2;
    // Every bin-7 vertex with its partition prefix (the NBHR_CLUSTER_MOVE
    // expansion and the pull expansion).
    DevBuf<uint32_t> hubs, hubs_item_of;
    DevBuf<int64_t> hubs_part, hubs_ent;
    int64_t hub_count = 0, hub_te = 0;
    // Pull expansion allowed (symmetric graph, a neighbourhood move method);
    // vol receives the seed volume of each round.
    bool pull = false;
    DevBuf<unsigned long long> vol;
    HubList Hubs() const {
      HubList h;
      h.list = hubs.get();
      h.count = hub_count;
      h.te = hub_te;
      h.part_off = hubs_part.get();
      h.item_of = hubs_item_of.get();
      return h;
    }
  };
  void PrepareLevel(const CsrView& g, const StateView& st, const Layout& lay, LevelCtx* ctx);
  // Best-move kernels of every bin; lists[b] holds counts[b] vertices of bin
  // b, te[7] bounds the degree sum of bin 7's list.  scan_hubs: refresh the
  // bin-7 prefix sums first.
  void LaunchMoves(LevelCtx& c, Workspace& ws, const MoveParams& mp, bool async,
                   const uint32_t* const lists[kBins], const int64_t counts[kBins],
                   const int64_t te[kBins], const MoveOut& out, uint64_t salt, bool scan_hubs,
                   const RoundResetJob* reset = nullptr);
  // IterateBestMoves (parallel-correlation-clusterer.cc:340-365) over the
  // graphs selected by sel (present-graph flags; null = all).  Returns the
  // per-present-graph "moved" flags.
  std::vector<uint8_t> IterateBestMoves(const Level& L, const std::vector<uint8_t>* sel,
                                        int64_t inner);
  // One BestMovesForVertexSubset round; counts / te describe the listed
  // frontier and receive the next one.  Returns the next frontier size.
  int64_t RoundStep(LevelCtx& c, Workspace& ws, const uint32_t* vgraph, int G, uint8_t* gmoved,
                    int64_t counts[kBins], int64_t te[kBins]);
  // The same round with every count on device (DeviceCounts): launches sized
  // by the level's bin capacities, no host read-back, no allocation, so the
  // rounds of a call run back to back and one round can be captured as a HIP
  // graph and replayed.  Sync LOUVAIN rounds without subclustering and with
  // exact node weights (DeviceRounds()).
  void RoundStepDevice(LevelCtx& c, Workspace& ws, const uint32_t* vgraph, uint8_t* gmoved);
  bool DeviceRounds() const;
  void BinFrontier(LevelCtx& c, Workspace& ws, const uint8_t* frontier, int64_t counts[kBins],
                   int64_t te[kBins]);
  // Node weights on which FP64 atomics are exact (see NodeWeightsExact);
  // otherwise the sync apply and the cluster-weight rebuilds take the ordered
  // passes.  Run sets it from the level-0 weights.
  void SetExactWeights(bool exact) { exact_nw_ = exact; }

 private:
  void ResetCounters(RoundCounters* rc, bool all);
  void ReadCounters(RoundCounters* rc, int64_t counts[kBins], int64_t te[kBins],
                    RoundCounters* full);
  Timer StartTimer();
  void StopTimer(Timer t, int kind);
  void CollectTimers();

  pcc_graph* h_;
  pcc_config cfg_;
  hipStream_t s_;
  DeviceAllocator* al_;
  pcc_stats* stats_;
  std::vector<std::pair<Timer, int>> pending_;
  double kind_seconds_[16] = {0};
  int64_t kind_launches_[16] = {0};
  int64_t bin_edges_[kStatGroups] = {0};
  int64_t bin_vertices_[kStatGroups] = {0};
  int64_t rounds_ = 0, edges_scanned_ = 0, frontier_vertices_ = 0, fallback_ = 0;
  // Fallback and repair counts of host-driven rounds, summed on device and
  // read once at the end of Run.
  DevBuf<unsigned long long> fb_acc_;
  int64_t cluster_moves_ = 0;
  double bestmove_bytes_ = 0;
  bool exact_nw_ = true;
  // With refinement every level's graph is visited twice; its best-move
  // context (bins, thresholds, hub lists and scratch) depends only on the
  // graph and its node weights, so the second visit reuses it.
  std::vector<std::pair<const Level*, std::unique_ptr<LevelCtx>>> ctx_cache_;
};

// The config checks every entry point that runs best moves shares
// (RefineClusters_subroutine, parallel-correlation-clusterer.cc:487-500, and
// the proto's enums, config.proto:115-160): clustering_moves_method LOUVAIN or
// DEFAULT_CLUSTER_MOVES (UNIMPLEMENTED otherwise, as the reference), a known
// move_method and subclustering_method (INVALID_ARGUMENT), no async subcluster
// moves inside DEFAULT_CLUSTER_MOVES rounds and no subclustering of batches
// (UNIMPLEMENTED), strictly ascending rows for TRIANGLE (INVALID_ARGUMENT).
int CheckMovesConfig(pcc_graph* h, const pcc_config& cfg);

// Node weights on which FP64 atomics are exact: every weight an integer and
// the sum of their magnitudes below 2^53, so every partial sum of cluster
// weights is exactly representable and its order cannot change a bit.
// Reads back one flag (synchronizes the stream).
bool NodeWeightsExact(DeviceAllocator* al, const double* nw, int64_t n, hipStream_t s);

void LaunchIota(uint32_t* p, int64_t n, hipStream_t s);
void LaunchFillF64(double* p, int64_t n, double x, hipStream_t s);
// Number of distinct values among ids[0..n), each below n.  Synchronizes.
int64_t CountDistinctIds(DeviceAllocator* al, int64_t n, const uint32_t* ids, hipStream_t s);

}  // namespace pcc

#endif  // PCC_AMD_ENGINE_H_
