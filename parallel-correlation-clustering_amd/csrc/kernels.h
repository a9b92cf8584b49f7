// kernels.h -- host-side launchers of the gfx950 kernels and the small device
// allocator they draw temporaries from.  Everything runs on one HIP stream
// per graph handle.
#ifndef PCC_AMD_KERNELS_H_
#define PCC_AMD_KERNELS_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace pcc {

// ------------------------------------------------------------- allocator --

// Caching device allocator: freed blocks go to per-size free lists and are
// reused without hipFree, so a clustering call never synchronizes on memory
// management after its first level.  When the device is out of memory, the
// cached free blocks of this allocator and then of every other allocator in
// the process (each graph handle has one) are returned to the device before
// the allocation fails.  Thread-safe.
class DeviceAllocator {
 public:
  DeviceAllocator();
  ~DeviceAllocator();
  DeviceAllocator(const DeviceAllocator&) = delete;
  DeviceAllocator& operator=(const DeviceAllocator&) = delete;
  void* Alloc(size_t bytes);  // throws std::bad_alloc-like AllocError
  void Free(void* p);
  void ReleaseAll();
  // Returns the cached free blocks to the device; the bytes released.
  size_t ReleaseFree();
  size_t BytesInUse() const;

 private:
  mutable std::mutex mu_;
  std::map<size_t, std::vector<void*>> free_;
  std::map<void*, size_t> live_;
  size_t in_use_ = 0;
};

struct AllocError {
  size_t bytes;
};

// Owning typed device buffer tied to an allocator.
template <typename T>
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(DeviceAllocator* a, int64_t n) { Reset(a, n); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : a_(o.a_), p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      Release();
      a_ = o.a_;
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  ~DevBuf() { Release(); }
  void Reset(DeviceAllocator* a, int64_t n) {
    Release();
    a_ = a;
    n_ = n;
    p_ = n > 0 ? static_cast<T*>(a->Alloc(sizeof(T) * (size_t)n)) : nullptr;
  }
  void Release() {
    if (p_ && a_) a_->Free(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  int64_t size() const { return n_; }

 private:
  DeviceAllocator* a_ = nullptr;
  T* p_ = nullptr;
  int64_t n_ = 0;
};

#define PCC_HIP_CHECK(expr)                                                   \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) throw ::pcc::HipError{_e, #expr, __FILE__, __LINE__}; \
  } while (0)

struct HipError {
  hipError_t err;
  const char* what;
  const char* file;
  int line;
};

// ------------------------------------------------------------- structures --

struct CsrView {
  int64_t n = 0;
  int64_t m = 0;
  const int64_t* off = nullptr;
  const uint32_t* nbr = nullptr;
  const float* w = nullptr;
};

// ClusteringHelper SoA (parallel-correlation-clusterer-internal.h:158-164).
struct StateView {
  uint32_t* ids = nullptr;
  uint32_t* sizes = nullptr;
  const double* nw = nullptr;
  double* K = nullptr;
};

// Degree bins of the best-move kernels: one frontier list and one kernel
// shape per bin.
//   0..4   deg <= 4, 8, 16, 32, 64 -> 4..64 lanes per vertex, ordered shuffles
//   5..7   deg <= 127, 255, wave_max         -> one wave, LDS hash
//   8..10  deg <= 1023, 2047, block_max      -> one workgroup, LDS hash
//   11     larger -> several workgroups per vertex, each an LDS hash over one
//          hash partition of the neighbour clusters
// Each hash table size gets its own list, so a launch only visits vertices of
// its range.  Statistics and timers report 8 groups: bins 0..4, the wave
// hash bins, the workgroup hash bins, the partitioned bin.
constexpr int kBins = 12;
constexpr int kWaveBin0 = 5;   // first one-wave hash bin
constexpr int kBlockBin0 = 8;  // first one-workgroup hash bin
constexpr int kHubBin = 11;
constexpr int kStatGroups = 8;
__host__ __device__ inline int StatGroupOf(int b) {
  return b < kWaveBin0 ? b : (b < kBlockBin0 ? 5 : (b < kHubBin ? 6 : 7));
}
struct BinThresholds {
  int64_t wave_max = 512;
  int64_t block_max = 4096;
  // Workgroups per hash-bin launch, as a multiple of the workgroups that fit
  // on the chip at once (0: one per listed vertex, up to 8192).  Fewer
  // workgroups each run the software pipeline over several vertices (one
  // chip-full measured fastest on the RMAT scale-22 workload: 37.1 vs 34.3
  // best-move GTEPS with one workgroup per vertex).
  int64_t hash_waves = 1;
};

// Device-resident per-round counters.
struct RoundCounters {
  unsigned long long cnt[kBins];
  unsigned long long te[kBins];
  unsigned long long fb_count;      // vertices listed for the ordered kernel
  unsigned long long new_count;
  unsigned long long repair_count;  // hash tables with order-repaired slots
  unsigned long long moved_count;
  // fb_count + repair_count of earlier rounds, folded in by the round reset
  // (statistics read once per IterateBestMoves call).
  unsigned long long fb_total;
};

struct BinLayout {
  int64_t base[kBins];
};

// Device-driven rounds: the counts of the current round's lists (the
// previous round's binning, moved here when the round starts) live on device,
// so a round's launches are sized by the level's capacities, every kernel
// reads its list's count, and nothing is read back between rounds.
struct DeviceCounts {
  unsigned long long cnt[kBins];
  unsigned long long te[kBins];
  // 1 when the frontier's volume can exceed the pull threshold (a pull
  // expansion may then be chosen; NBHR_MOVE must then clear every mover flag).
  unsigned int pull_possible;
};

// Statistics of device-driven rounds, accumulated on device and read once per
// IterateBestMoves call.  A round counts when its frontier is non-empty, or
// when it is the call's first (IterateBestMoves, PCC.cc:346-363).
struct DeviceRoundStats {
  unsigned long long rounds, edges, verts;
  unsigned long long bin_edges[kStatGroups], bin_verts[kStatGroups];
  unsigned int first;
};

// Parameters of the best-move kernels.  Cluster ids are global vertex ids of
// the level (a batch of graphs occupies consecutive vertex ranges), so ties on
// the smaller id and the singleton rule compare ids within one graph exactly
// as the reference compares them.
struct MoveParams {
  double res = 0;           // resolution when every graph shares one
  double offset = 0;        // edge_weight_offset
  uint32_t new_id = 0;      // "form a new cluster" target (n for one graph)
  int force_ordered = 0;
  // Batches (null for a single graph): graph of each vertex, per-graph
  // resolution and vertex ranges [gbase[g], gbase[g + 1]).
  const uint32_t* vgraph = nullptr;
  const double* gres = nullptr;
  const int64_t* gbase = nullptr;
  uint32_t n = 0;           // vertices of the level
  // Async sub-rounds: only vertices with SubRoundOf(v, sub_count) == sub_index
  // are processed by a launch (sub_count <= 1: all).
  uint32_t sub_count = 1;
  uint32_t sub_index = 0;
  // Async: apply the sync singleton rule before moving (0/1).
  int async_singleton_rule = 0;
  // Async moves to a new cluster: ascending empty ids of the round and, per
  // graph, the next / end index into them (null: probe only).
  const uint32_t* free_ids = nullptr;
  unsigned long long* free_cur = nullptr;
  const unsigned long long* free_end = nullptr;
  // Per-vertex scale of the order-free test of the hashed bins (deg > 64):
  // PackScale(threshold, magnitude exponent), see SlotInexact.
  const int32_t* gscale = nullptr;
  // 1 when every resolution and node weight is >= 0: then a candidate c with
  // S_c <= 0 has delta <= change + S_c <= change and can never be chosen over
  // staying, so its cluster weight K_c is not looked up.
  int prune = 0;
};

struct MoveOut {
  uint32_t* moves = nullptr;      // sync: target or kNoneId
  uint8_t* moved = nullptr;       // moved_vertex
  uint8_t* moved_clusters = nullptr;  // async: destination clusters
  uint32_t* dbg_target = nullptr;  // EfficientBestMove's return (may be null)
  double* dbg_delta = nullptr;
};

// Default edges per partition of a bin-7 vertex.
constexpr int64_t kHubPartEdges = 2048;

// Bin-7 scratch.  part_off / ent_off: exclusive prefix sums over the bin's
// list of the partition counts ceil(deg / kHubPartEdges) and of the degrees;
// hflag (bit 0: own-cluster sum found in tval, bit 1: inexact or table
// overflow, bit 2: a candidate exists) is zero between launches.
// One partition item of a bin-7 list, written by the gather pass so that the
// partition pass reads its vertex, its adjacency range and the vertex state
// it needs with one load instead of a chain of dependent ones (no vertex
// moves between the two passes).
struct alignas(16) HubPart {
  int64_t start = 0;   // offset of v's adjacency
  int64_t deg = 0;     // degree of v
  uint32_t v = 0;      // the vertex
  uint32_t item = 0;   // its list item
  uint32_t part = 0;   // this partition's index among v's
  uint32_t parts = 0;  // v's partition count
  int32_t scale = 0;   // gscale[v] (0 without order-test scales)
  uint32_t cv = 0;     // ids[v]
  double kv = 0.0;     // nw[v]
  double res = 0.0;    // resolution of v's graph
  uint32_t pad[2] = {0, 0};
};


struct HubScratch {
  int64_t part_edges = kHubPartEdges;
  int64_t* part_off = nullptr;
  uint32_t* item_of = nullptr;  // list item of each partition
  HubPart* hpart = nullptr;     // record of each partition
  int64_t* ent_off = nullptr;
  uint32_t* hflag = nullptr;
  double* tval = nullptr;
  // Per partition item t, over its candidate clusters c with
  // x_c = NetWeight(S_c, k_v K_c, res): the largest x (px1), the smallest
  // cluster with that value (pc1), and the largest x below it (px2).
  double* px1 = nullptr;
  uint32_t* pc1 = nullptr;
  double* px2 = nullptr;
  // The same clusters bucketed by partition: item t's (cluster, weight bits)
  // pairs at bkt[t * bcap ...], bcur[t] of them; bcur[t] > bcap means the
  // bucket overflowed and the partition gathers and filters the whole
  // adjacency instead.  The partition pass resets bcur[t] to 0.
  uint2* bkt = nullptr;
  uint32_t* bcur = nullptr;
  int64_t bcap = 0;
  // pfine[t] != 0: partition item t received a fine value (sync rounds; see
  // kNoFineThr).  Without one its totals are exact in every order.  The
  // partition pass resets it with bcur[t].
  uint32_t* pfine = nullptr;
  // Async rounds: the cluster of every neighbour of list item i in adjacency
  // order at cbuf[ent_off[i] ...] (kNoneId for a self-loop), written by the
  // gather pass, so that a partition whose bucket overflowed (a heavy
  // cluster, frequent on unit weights) streams it instead of gathering the
  // whole row again.  Null in sync rounds.
  uint32_t* cbuf = nullptr;
};

// Bucket capacity per partition item: a partition's share of a vertex's
// edges is at most part_edges when clusters hash evenly; heavy clusters can
// overflow it.
inline int64_t HubBucketCap(int64_t part_edges) { return part_edges + part_edges / 2 + 64; }
// The larger capacity async rounds take when memory allows.
inline int64_t HubBucketCapAsync(int64_t part_edges) { return 8 * part_edges + 64; }

// Scratch for the ordered fallback path.
struct HashScratch {
  uint32_t* keys = nullptr;
  double* vals = nullptr;
  uint32_t cap = 0;   // entries per workgroup region (power of 2)
  int regions = 0;    // number of workgroup regions
};

// ---------------------------------------------------------------- launchers --

// Bins the frontier (frontier == nullptr: every vertex) into list regions
// list[layout.base[b] + i], counting into rc->cnt / rc->te.
void LaunchBinFrontier(const CsrView& g, const uint8_t* frontier, const BinThresholds& th,
                       const BinLayout& layout, uint32_t* list, RoundCounters* rc,
                       hipStream_t s);

// Per-level statistics in one pass (caps zeroed by the caller, kLevelStats
// entries): caps[b] vertices and caps[kBins + b] edges of bin b (capacity of
// each list region), caps[2 kBins] the largest degree, caps[2 kBins + 1] = 1
// if a node weight nw[v] (nw may be null) or a resolution gres[g] (gres may
// be null, G entries) is negative or NaN.
constexpr int kLevelStats = 2 * kBins + 2;
void LaunchBinCapacity(const CsrView& g, const BinThresholds& th, const double* nw,
                       const double* gres, int64_t G, unsigned long long* caps, hipStream_t s);

// Async sub-rounds: cnt[b * S + s] (zeroed by the caller) counts the listed
// vertices of bin b in sub-round s; with cur[b * S + s] set to the segment
// starts within each bin region, LaunchSubScatter copies the list into out
// grouped by (bin, sub-round).  te7[s] (zeroed by the caller) receives the
// degree sum of sub-round s's bin-7 vertices.  S <= 1024.
void LaunchSubCount(const CsrView& g, const uint32_t* list, const BinLayout& layout,
                    const int64_t counts[kBins], uint32_t S, uint32_t* cnt,
                    unsigned long long* te7, hipStream_t s);
void LaunchSubScatter(const uint32_t* list, const BinLayout& layout, const int64_t counts[kBins],
                      uint32_t S, uint32_t* cur, uint32_t* out, hipStream_t s);

// gscale[v] for every vertex with deg > 64 (see MoveParams::gscale).
// max_deg (-1: unknown) bounds the degrees; without rows longer than
// kLongRow the long-row passes are skipped.
void LaunchGranThreshold(DeviceAllocator* al, const CsrView& g, double offset, int32_t* gscale,
                         hipStream_t s, int64_t max_deg = -1);

// Bin-7 prefix sums (HubScratch::part_off / ent_off) of a list, with
// partitions of part_edges edges, and the list item of every partition
// (item_of[part_off[i] + p] = i).
void LaunchHubScan(const CsrView& g, const uint32_t* list, int64_t count, int64_t part_edges,
                   int64_t* part_off, int64_t* ent_off, uint32_t* item_of, hipStream_t s,
                   const unsigned long long* dcount = nullptr);

// Every vertex of bin 7 (deg > th.block_max) appended to list; *count zeroed
// by the caller.
void LaunchCollectHubs(const CsrView& g, const BinThresholds& th, uint32_t* list,
                       unsigned long long* count, hipStream_t s);

// One best-move kernel per bin 0..6.  count = number of listed vertices in
// the bin.
// dc (may be null): device-driven round; count is then the bin's capacity
// on the level (the launch bound) and the kernel reads dc->cnt[bin].
void LaunchBestMoveBin(int bin, int64_t count, bool async, const BinThresholds& th,
                       const CsrView& g, const StateView& st, const MoveParams& mp,
                       const uint32_t* list, const MoveOut& out, uint32_t* fb_list,
                       RoundCounters* rc, uint64_t permute_salt, hipStream_t s,
                       const DeviceCounts* dc = nullptr);

// Bins 0..10 of a round (count / list per bin, as LaunchBestMoveBin).  Sync
// rounds run bins 0-4 as one launch and, with at most 1024 hashed vertices
// and host-known counts, bins 5-10 as one launch of the largest table
// needed; async rounds launch bin by bin.  Results equal the per-bin launches.
// reset (may be null): the round's clears (LaunchRoundReset without a third
// array or accumulator), done by extra workgroups of the bins-0-4 launch in
// sync rounds; it must not clear anything the best-move kernels write.
struct RoundResetJob {
  int64_t n;
  uint8_t* a;
  uint8_t* b;
  RoundCounters* rc;
  int counters;
};
void LaunchBestMoveBins(const int64_t counts[kHubBin], const uint32_t* const lists[kHubBin],
                        bool async, const BinThresholds& th, const CsrView& g,
                        const StateView& st, const MoveParams& mp, const MoveOut& out,
                        uint32_t* fb_list, RoundCounters* rc, uint64_t permute_salt,
                        hipStream_t s, const DeviceCounts* dc = nullptr,
                        const RoundResetJob* reset = nullptr);

// Bin 7: partition pass and selection pass over a list whose HubScratch
// prefix sums are current; te = sum of the listed degrees.
void LaunchBestMoveHub(int64_t count, int64_t te, bool async, const CsrView& g,
                       const StateView& st, const MoveParams& mp, const uint32_t* list,
                       const HubScratch& hb, const MoveOut& out, uint32_t* fb_list,
                       RoundCounters* rc, hipStream_t s, const DeviceCounts* dc = nullptr);

// Ordered (sequential-order-exact) best moves for the fallback list.
void LaunchBestMoveOrdered(const CsrView& g, const StateView& st, const MoveParams& mp,
                           const uint32_t* fb_list, const RoundCounters* rc,
                           const MoveOut& out, const HashScratch& hs, hipStream_t s);

// Sync apply, MoveNodesToCluster (parallel-correlation-clusterer-internal.cc:256-369).
void LaunchApply(const int64_t counts[kBins], const BinLayout& layout, const uint32_t* list,
                 const StateView& st, uint32_t new_id, const uint32_t* moves, uint8_t* modified,
                 uint8_t* new_flag, RoundCounters* rc, hipStream_t s,
                 const DeviceCounts* dc = nullptr);

// The same apply with the cluster-weight arithmetic in the reference's order
// (losses of each old cluster, then gains of each new one, members in
// ascending vertex order), for node weights on which FP64 atomics are not
// exact.  Synchronizes the stream once (the mover count).
void LaunchApplyOrdered(DeviceAllocator* al, const int64_t counts[kBins], const BinLayout& layout,
                        const uint32_t* list, const StateView& st, uint32_t n, uint32_t new_id,
                        const uint32_t* moves, uint8_t* modified, uint8_t* new_flag,
                        RoundCounters* rc, hipStream_t s);

// K[c] = left fold of nw over the members of c in ascending vertex order,
// from 0 (SetClustering, parallel-correlation-clusterer-internal.cc:60-77);
// sizes are left as they are.
void OrderedClusterWeights(DeviceAllocator* al, int64_t n, const StateView& st, hipStream_t s);

// Empty-id reuse for vertices forming new clusters, per graph of a batch
// (vgraph / gbase null for one graph); gated on rc->new_count.  work (may be
// null: allocated per call) holds its n-sized temporaries, so a caller that
// keeps one per level runs the assignment with no allocation.
// Up to four byte ranges cleared by one launch (a kernel, so the clear can
// sit in a captured graph).
struct ZeroSpans {
  uint8_t* p[4];
  int64_t bytes[4];
  int n;
};
void LaunchZeroSpans(const ZeroSpans& z, hipStream_t s);

// *acc += rc->fb_total + rc->fb_count + rc->repair_count, on the stream.
void LaunchAddFallbacks(const RoundCounters* rc, unsigned long long* acc, hipStream_t s);

struct NewAssignWork {
  DevBuf<uint32_t> zrank, zero_list, nrank, scan;
  DevBuf<uint8_t> z;
  NewAssignWork(DeviceAllocator* al, int64_t n);
};
void RunNewClusterAssign(DeviceAllocator* al, const StateView& st, uint32_t n,
                         const uint32_t* vgraph, const int64_t* gbase, uint8_t* new_flag,
                         uint8_t* modified, RoundCounters* rc, hipStream_t s,
                         const NewAssignWork* work = nullptr);

// Sync set moves of arbitrary vertex groups (grp[v] < n; one graph): every
// group with members gets BestMove(graph, members)
// (parallel-correlation-clusterer-internal.cc:644-707), the sync rule against
// the cluster of its smallest member (parallel-correlation-clusterer.cc:224-233,
// 270-279) and moves iff delta > 0; the moves are applied as
// MoveNodesToCluster, or-ing modified clusters into modified.  Returns the
// number of groups that move.  Synchronizes the stream.
int64_t RunGroupMoves(DeviceAllocator* al, const CsrView& g, const StateView& st,
                      const MoveParams& mp, const uint32_t* grp, uint8_t* modified,
                      uint8_t* new_flag, RoundCounters* rc, hipStream_t s);

// The whole-cluster phase of a DEFAULT_CLUSTER_MOVES round
// (parallel-correlation-clusterer.cc:167-297) over the clusters of the graphs
// whose frontier list (counts / layout / list) is non-empty.  Sync: the moves
// are applied as MoveNodesToCluster, or-ing modified clusters into modified.
// Async: sub_rounds sub-rounds over a hash of the cluster id (mixed with
// salt); modified is marked at the members' vertex ids as the reference does.
// Returns the number of cluster moves.  Synchronizes the stream.
int64_t RunClusterMoves(DeviceAllocator* al, const CsrView& g, const StateView& st,
                        const MoveParams& mp, int G, const int64_t counts[kBins],
                        const BinLayout& layout, const uint32_t* list, bool async,
                        uint32_t sub_rounds, uint32_t salt, uint8_t* modified, uint8_t* new_flag,
                        RoundCounters* rc, hipStream_t s);
// Async sub-rounds of the whole-cluster phase.
constexpr uint32_t kClusterSubRounds = 16;

// Empty cluster ids at the start of an async round (MoveParams::free_*).
struct FreeIds {
  DevBuf<uint32_t> ids;
  DevBuf<unsigned long long> cur, end;
};
void PrepareFreeIds(DeviceAllocator* al, const StateView& st, uint32_t n, int G,
                    const int64_t* gbase, FreeIds* f, hipStream_t s);

// Clears of one round in a single launch: byte arrays a, b, c (each may be
// null) of n entries, the counters selected by `counters` (bit 0: cnt / te,
// bit 1: fb_count / new_count / repair_count, after adding fb_count and
// repair_count to fb_total) and *acc (may be null).  Byte
// arrays must start 16-byte aligned (device allocations are).
// c_gate (may be null): c is cleared only when *c_gate != 0.
void LaunchRoundReset(int64_t n, uint8_t* a, uint8_t* b, uint8_t* c, RoundCounters* rc,
                      int counters, unsigned long long* acc, hipStream_t s,
                      const unsigned int* c_gate = nullptr);

// Start of a device-driven round: dc <- the binned counts of rc (then zeroed
// there unless keep, the ALL_MOVE frontier that never changes), the round's
// statistics into stats, and dc->pull_possible from the frontier's volume
// against pull_thr (always 1 unless nbhr_move: the NBHR_MOVE seeds are a
// subset of the frontier).
void LaunchRoundBegin(RoundCounters* rc, DeviceCounts* dc, DeviceRoundStats* stats, bool keep,
                      bool nbhr_move, double pull_thr, hipStream_t s);
// Zero statistics with the next round marked as a call's first.
void LaunchRoundStatsInit(DeviceRoundStats* stats, hipStream_t s);

// flags[graph of v] = 1 for every v with next[v] != 0.
void LaunchGraphFlags(int64_t n, const uint8_t* next, const uint32_t* vgraph, uint8_t* flags,
                      hipStream_t s);

// Direction of the next-frontier expansion (the sparse / dense switch of
// gbbs's edgeMap): with vol null or *vol <= thr the push kernels mark the
// neighbours of every seed; otherwise the pull kernels test every vertex for a
// seed neighbour.  Both give the same frontier on a symmetric graph, so the
// pull side is only enabled for graphs checked to be symmetric.
struct ExpandDir {
  const unsigned long long* vol = nullptr;
  unsigned long long thr = 0;
  __device__ __forceinline__ bool Pull() const { return vol != nullptr && *vol > thr; }
};

// Vertices with deg > block_max of a level graph, with partitions of
// part_edges edges (prefix part_off, owner item_of; see LaunchHubScan).
struct HubList {
  const uint32_t* list = nullptr;
  int64_t count = 0, te = 0;
  const int64_t* part_off = nullptr;
  const uint32_t* item_of = nullptr;
};

// Next frontier = neighbours of the seeds.  Bin 7 of the list is walked in
// the partitions of hub_part_off (its HubScratch::part_off, partitions of
// part_edges edges); te[7] = sum of its degrees.
void LaunchExpandMoved(const int64_t counts[kBins], const int64_t te[kBins],
                       const BinLayout& layout, const uint32_t* list,
                       const int64_t* hub_part_off, const uint32_t* hub_item_of,
                       int64_t part_edges, const CsrView& g, const uint8_t* moved,
                       const ExpandDir& dir, uint8_t* next, hipStream_t s,
                       const DeviceCounts* dc = nullptr);
// Neighbours of every vertex whose cluster is modified.  Vertices with
// deg > block_max are taken from hubs instead of the vertex sweep.
void LaunchExpandClusters(const CsrView& g, const uint32_t* ids, const uint8_t* modified,
                          int64_t block_max, const HubList& hubs, int64_t part_edges,
                          const ExpandDir& dir, uint8_t* next, hipStream_t s);
// vol += sum of (deg + 1) over the seeds: modified[ids[v]] if ids is non-null,
// else moved[v] (which must then hold 0 for every vertex that did not move).
void LaunchSeedVolume(const CsrView& g, const uint32_t* ids, const uint8_t* modified,
                      const uint8_t* moved, unsigned long long* vol, hipStream_t s,
                      const unsigned int* gate = nullptr);
// Pull side of the expansion (no-op unless dir selects it); hubs must list
// every vertex with deg > block_max.
void LaunchPullFrontier(const CsrView& g, const uint32_t* ids, const uint8_t* modified,
                        const uint8_t* moved, int64_t block_max, const HubList& hubs,
                        int64_t part_edges, const ExpandDir& dir, uint8_t* next, hipStream_t s);
// Symmetry fingerprint: true when the multiset of directed edges (u, v)
// equals that of (v, u), decided by comparing sums of 64-bit hashes of the
// pairs (a false "symmetric" needs a hash-sum collision, probability ~2^-64).
bool GraphIsSymmetric(DeviceAllocator* al, const CsrView& g, hipStream_t s);

// ------------------------------------------------------------- primitives --

// Copies bytes from device memory to host memory after the work queued on s
// (synchronizes s), through a pinned staging buffer of the calling thread.
void ReadBack(void* host, const void* dev, size_t bytes, hipStream_t s);

// Exclusive scan; total (device pointer, may be null) receives the sum.  If
// gate is non-null and *gate == 0 the kernels return immediately.
// work (may be null): ScanWorkElems(n) preallocated elements, so the scan
// allocates nothing.
int64_t ScanWorkElems(int64_t n);
void ExclusiveScanU8toU32(DeviceAllocator* al, const uint8_t* in, uint32_t* out, int64_t n,
                          uint32_t* total, hipStream_t s,
                          const unsigned long long* gate = nullptr, uint32_t* work = nullptr);
void ExclusiveScanU32toI64(DeviceAllocator* al, const uint32_t* in, int64_t* out, int64_t n,
                           int64_t* total, hipStream_t s);
void ExclusiveScanU32toU32(DeviceAllocator* al, const uint32_t* in, uint32_t* out, int64_t n,
                           uint32_t* total, hipStream_t s);

void ReduceMaxU32(DeviceAllocator* al, const uint32_t* in, int64_t n, uint32_t* out,
                  hipStream_t s);

// Stable LSD radix sort of (key, value) pairs on key bits [begin_bit, end_bit).
void RadixSortPairsF32(DeviceAllocator* al, uint64_t* keys, float* vals, int64_t n, int begin_bit,
                       int end_bit, hipStream_t s);
void RadixSortPairsF64(DeviceAllocator* al, uint64_t* keys, double* vals, int64_t n,
                       int begin_bit, int end_bit, hipStream_t s);
void RadixSortPairsU32(DeviceAllocator* al, uint64_t* keys, uint32_t* vals, int64_t n,
                       int begin_bit, int end_bit, hipStream_t s);

// Segment boundaries of sorted keys: returns the number of runs of equal keys
// and fills start[0..nseg] (start[nseg] = n).  Synchronizes the stream.
int64_t Segments(DeviceAllocator* al, const uint64_t* keys, int64_t n, DevBuf<int64_t>* start,
                 hipStream_t s);

// Sums of the runs [start[i], start[i + 1]) of vals, i < nseg, in pbbslib's
// reduce order (left fold up to 1024 terms, blocked beyond), into out[i].
void SegmentSumsF32(DeviceAllocator* al, const float* vals, const int64_t* start, int64_t nseg,
                    int64_t nvals, float* out, hipStream_t s);
void SegmentSumsF64(DeviceAllocator* al, const double* vals, const int64_t* start, int64_t nseg,
                    int64_t nvals, double* out, hipStream_t s);

// ------------------------------------------------------------ contraction --

struct DevCsr {
  int64_t n = 0;
  int64_t m = 0;
  DevBuf<int64_t> off;
  DevBuf<uint32_t> nbr;
  DevBuf<float> w;
  CsrView View() const {
    CsrView v;
    v.n = n;
    v.m = m;
    v.off = off.get();
    v.nbr = nbr.get();
    v.w = w.get();
    return v;
  }
};

// Per-graph maximum cluster id (gmax[graph] = max ids[v]); gmax zeroed by the
// caller.  vgraph null: one graph.
void LaunchGraphMaxIds(int64_t n, const uint32_t* ids, const uint32_t* vgraph, uint32_t* gmax,
                       hipStream_t s);

// CompressGraph (parallel-correlation-clusterer-internal.cc:709-782) of every
// graph g of the level with newbase[g] >= 0: cluster id c of graph g becomes
// vertex newbase[g] + (c - gbase[g]) of the contracted level, whose size is n2.
// Vertices of graphs with newbase < 0 are dropped.  vgraph / gbase / newbase
// null: one graph, identity numbering (n2 = 1 + max id).
// method 0: cluster-grouped (k_group.hip); 1: one LSD radix sort of every
// kept edge by (c(u), c(v)) (k_contract.hip).  Both give the same graph.
// chunk_terms bounds the terms one radix sort of the grouped path takes.
constexpr int64_t kContractChunkTerms = (int64_t)1 << 31;
void CompressGraph(DeviceAllocator* al, const CsrView& g, const uint32_t* ids, const double* nw,
                   const uint32_t* vgraph, const int64_t* gbase, const int64_t* newbase,
                   int64_t n2, DevCsr* out, DevBuf<double>* out_nw, hipStream_t s,
                   int method = 0, int64_t chunk_terms = kContractChunkTerms);
void CompressGraphSorted(DeviceAllocator* al, const CsrView& g, const uint32_t* ids,
                         const double* nw, const uint32_t* vgraph, const int64_t* gbase,
                         const int64_t* newbase, int64_t n2, DevCsr* out,
                         DevBuf<double>* out_nw, hipStream_t s);
// Returns false (touching nothing) for inputs it does not take: n or n2 of
// 2^32 - 1 or more, or keys of its radix path wider than 64 bits.
bool CompressGraphGrouped(DeviceAllocator* al, const CsrView& g, const uint32_t* ids,
                          const double* nw, const uint32_t* vgraph, const int64_t* gbase,
                          const int64_t* newbase, int64_t n2, DevCsr* out,
                          DevBuf<double>* out_nw, hipStream_t s,
                          int64_t chunk_terms = kContractChunkTerms);

// The subcluster numbering of CompressSubclusters
// (parallel-correlation-clusterer.cc:436-474) of one graph's clustering ids:
// method 1 TRIANGLE, 2 CONNECTIVITY, 3 SLOW_CONNECTIVITY (which reads offset
// and res: the edge weight offset and resolution).  Fills sub (n entries: the next level's
// vertex of every vertex) and parent (n2 entries: the cluster every next-level
// vertex starts in) and returns n2 = 1 + max sub, or -1 when method 1 meets an
// adjacency row that is not strictly ascending.  Synchronizes the stream.
// check_rows false skips the row-order check (the caller made it already);
// half: the graph is known to be symmetric, so each edge is visited once.
bool RowsAscending(DeviceAllocator* al, const CsrView& g, hipStream_t s);
int64_t SubclusterIds(DeviceAllocator* al, const CsrView& g, const uint32_t* ids, int method,
                      DevBuf<uint32_t>* sub, DevBuf<uint32_t>* parent, double offset,
                      double res, hipStream_t s, bool check_rows = true, bool half = false);

// FlattenClustering (parallel/parallel-graph-utils.cc:233-243).
void LaunchFlatten(int64_t n, const uint32_t* ids, const uint32_t* compressed, uint32_t* out,
                   hipStream_t s);

// Flatten of the level-0 assignment of batch graphs: for vertex v of original
// graph o = vgraph0[v] with sel[o] != 0, ids[v] = next_ids[ids[v] + shift[o]]
// (next_ids null: only the shift is applied).
void LaunchFlattenShift(int64_t n, uint32_t* ids, const uint32_t* vgraph0, const uint8_t* sel,
                        const int64_t* shift, const uint32_t* next_ids, hipStream_t s);

// Compact rank relabel + ResetClustering/SetClustering
// (parallel-correlation-clusterer.cc:594-600): ids := rank of id among the
// distinct ids; sizes/K rebuilt.
void ResetClusteringFromIds(DeviceAllocator* al, int64_t n, const uint32_t* flat,
                            const StateView& st, hipStream_t s);

// Batch form of the refinement step of one level: for every vertex v of a
// graph p with sel[p] != 0 (vgraph / gbase / nbase describe the level and the
// next level), flat = next_ids[nbase[p] + ids[v] - gbase[p]] is relabelled to
// gbase[p] + (rank of flat among the graph's distinct flat ids) and that
// graph's sizes / cluster weights are rebuilt.
void ResetClusteringBatch(DeviceAllocator* al, int64_t n, int64_t n_next, const uint32_t* next_ids,
                          const uint32_t* vgraph, const int64_t* gbase, const int64_t* nbase,
                          const uint8_t* sel, const StateView& st, hipStream_t s);

// Singleton state (ids = i, sizes = 1, K = nw).
void LaunchSingletons(int64_t n, const StateView& st, hipStream_t s);

// SetClustering from per-vertex cluster indices (initial clustering).
void LaunchSetClustering(int64_t n, const uint32_t* init, const StateView& st, hipStream_t s);

// Node weights = degrees (ComputeModularityConfig,
// parallel/parallel-graph-utils.cc:141-153).
void LaunchDegrees(const CsrView& g, double* nw, hipStream_t s);

// Fixed-order double sum.
void SumF64(DeviceAllocator* al, const double* in, int64_t n, double* out, hipStream_t s);

// ComputeObjective (parallel-correlation-clusterer-internal.cc:118-149) of
// each of the G graphs [gbase[g], gbase[g + 1]) with resolution res[g]; ids
// are global cluster ids.  out_dev[g] receives graph g's objective.
void ComputeObjective(DeviceAllocator* al, const CsrView& g, const uint32_t* ids,
                      const double* nw, int G, const int64_t* gbase, const double* res,
                      double offset, double* out_dev, hipStream_t s);

// The same objective for one graph with no device allocation: every work
// array lives in a fixed per-device array (NoAllocScratch(device), with that
// device current), so graphs are limited to kNoAllocMaxVertices vertices and
// kNoAllocMaxEdges directed edges.  nw null means unit weights.  Cluster ids
// outside [0, n) give NaN.  Asynchronous on s; calls that overlap on one
// device share the array and race.
constexpr int64_t kNoAllocMaxVertices = (int64_t)1 << 21;
constexpr int64_t kNoAllocMaxEdges = (int64_t)1 << 30;
void* NoAllocScratch(int device);
void ComputeObjectiveNoAlloc(void* scratch_dev, const CsrView& g, const uint32_t* ids,
                             const double* nw, double res, double offset, double* out_dev,
                             hipStream_t s);

// ComputeModularity (parallel-modularity-clusterer.cc:36-66) of each graph
// with resolution gamma[g].
void ComputeModularity(DeviceAllocator* al, const CsrView& g, const uint32_t* ids, int G,
                       const int64_t* gbase, const double* gamma, double* out_dev, hipStream_t s);

}  // namespace pcc

#endif  // PCC_AMD_KERNELS_H_
