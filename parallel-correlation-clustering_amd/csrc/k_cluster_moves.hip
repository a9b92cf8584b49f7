// k_cluster_moves.hip -- the whole-cluster phase of a DEFAULT_CLUSTER_MOVES
// round (clustering/clusterers/parallel-correlation-clusterer.cc:167-297,
// subclustering NONE): after the vertex moves, every cluster tries to move as
// a whole, with ClusteringHelper::BestMove(graph, moving_nodes)
// (parallel-correlation-clusterer-internal.cc:644-707) and BestMoveFromStats
// (correlation-clusterer-util.h:70-113).
//
// The moving set of cluster c is all of c, so of BestMoveFromStats' terms only
// these survive (with W = the summed node weights of c's members, S_cd = the
// summed weights of the edges from c's members into cluster d, each weight
// taken as float(w - offset) as the reference's float lambda parameter does):
//   change  = 0 - NetWeight(0, W * (K_c - W))         (class 1 together: none)
//   delta_d = change + NetWeight(S_cd, W * (K_d - 0))
// and the best candidate is the largest delta_d > change, smallest d on ties,
// else "new cluster" with delta = change.  S_cd is a left fold in double over
// c's members in ascending vertex order and their adjacency order -- exactly
// the order of the directed edges (u -> v), c(u) = c != c(v), in CSR order.
// So the edges are keyed (c(u), c(v)), stably radix-sorted (CSR order breaks
// ties) and each run of equal keys is folded left in double: by an unordered
// sum when that is exact in every order, else by an ordered wave fold.
//
// Sync (PCC.cc:263-285): a cluster moves only to a larger id, iff delta > 0;
// the moves are applied as MoveNodesToCluster, so new-cluster movers take
// empty ids one vertex each.  Async (PCC.cc:254-262, AsyncMove(set),
// PCCI.cc:405-445): clusters are processed in sub-rounds over a hash of the
// cluster id; within a sub-round every cluster decides on the same snapshot,
// each target admits only its best incoming cluster and a cluster does not
// follow a target that moves away (a mutual pair merges into the larger id),
// sizes and weights move by the float-rounded set weight,
// and the reference's moved_clusters marking at the members' vertex ids is
// kept.  W is accumulated with FP64 atomics: exact for integer node weights
// (unit weights, degrees and their sums), as the cluster weights K are.

#include "device_common.h"
#include "kernels.h"

namespace pcc {

namespace {

constexpr int kCmItems = 16;
constexpr int kCmThreads = 256;
constexpr int64_t kCmChunk = (int64_t)kCmItems * kCmThreads;
constexpr int64_t kCmShortRun = 64;

int CmGrid(int64_t items, int per_block, int max_blocks) {
  int64_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (int)b;
}

int CmBits(uint64_t maxval) {
  int b = 0;
  while (b < 63 && (maxval >> b) != 0) ++b;
  return b < 1 ? 1 : b;
}

// Which clusters take part: those of graphs active in this round and, for
// async sub-rounds, of the current sub-round.
struct CmFilter {
  const uint32_t* vgraph = nullptr;  // null: one graph
  const uint8_t* gact = nullptr;     // per present graph (null: all active)
  uint32_t sub_count = 1, sub_index = 0, salt = 0;
  __device__ __forceinline__ bool Cluster(uint32_t c) const {
    if (gact && !gact[vgraph ? vgraph[c] : 0]) return false;
    return sub_count <= 1 || SubRoundOf(c ^ salt, sub_count) == sub_index;
  }
};

__device__ __forceinline__ double ResOf(const MoveParams& mp, uint32_t c) {
  return mp.gres ? mp.gres[mp.vgraph[c]] : mp.res;
}

// change = 0 - NetWeight(class_1_currently_together = 0, W * (K_c - W))
// (UTIL.h:79-93 with one moving cluster; the class-2 term is exactly 0).
__device__ __forceinline__ double ChangeOf(double W, double Kc, double res) {
  const double me = __fma_rn(W, Kc - W, 0.0);
  return 0.0 - NetWeight(0.0, me, res);
}

// Order-preserving map of doubles to unsigned integers (0 is below every
// finite value and serves as "no candidate").
__device__ __forceinline__ unsigned long long OrdBits(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double FromOrd(unsigned long long o) {
  const unsigned long long b = (o >> 63) ? (o & 0x7FFFFFFFFFFFFFFFull) : ~o;
  return __longlong_as_double((long long)b);
}

// Active graphs of the round: gact[graph of v] = 1 for every listed vertex.
struct CmBins {
  int64_t base[kBins];
  int64_t count[kBins];
};
__global__ __launch_bounds__(256) void k_cm_active(CmBins bins, const uint32_t* __restrict__ list,
                                                   const uint32_t* __restrict__ vgraph,
                                                   uint8_t* __restrict__ gact) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int b = 0; b < kBins; ++b)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < bins.count[b]; i += stride)
      gact[vgraph[list[bins.base[b] + i]]] = 1;
}

// Directed edges u -> v with c(u) taking part and c(v) != c(u), chunk by
// chunk (kCmChunk consecutive CSR edges per workgroup, kCmItems per thread
// in order), so the emitted edges come out in CSR order.
__global__ __launch_bounds__(kCmThreads) void k_cm_count(CsrView g, const uint32_t* __restrict__ ids,
                                                         CmFilter f, int64_t nchunks,
                                                         uint32_t* __restrict__ chunk_cnt) {
  __shared__ uint32_t sw[kCmThreads / kWave];
  for (int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const int64_t c0 = ch * kCmChunk;
    const int64_t len = min(kCmChunk, g.m - c0);
    uint32_t c = 0;
    const int64_t q0 = (int64_t)threadIdx.x * kCmItems;
    if (q0 < len) {
      const int64_t q1 = min(q0 + kCmItems, len);
      int64_t u = -1, uend = 0;
      uint32_t cu = 0;
      bool take = false;
      for (int64_t q = q0; q < q1; ++q) {
        const int64_t p = c0 + q;
        if (p >= uend) {
          u = VertexOfEdge(g.off, u < 0 ? 0 : u + 1, g.n, p);
          uend = g.off[u + 1];
          cu = ids[u];
          take = f.Cluster(cu);
        }
        if (take && ids[g.nbr[p]] != cu) ++c;
      }
    }
    uint32_t total = 0;
    (void)BlockExclusiveScan(c, sw, &total);
    if (threadIdx.x == 0) chunk_cnt[ch] = total;
  }
}

__global__ __launch_bounds__(kCmThreads) void k_cm_emit(CsrView g, const uint32_t* __restrict__ ids,
                                                        CmFilter f, int64_t nchunks,
                                                        const int64_t* __restrict__ chunk_off,
                                                        int vbits, double offset,
                                                        uint64_t* __restrict__ keys,
                                                        float* __restrict__ vals) {
  __shared__ uint32_t sw[kCmThreads / kWave];
  for (int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const int64_t c0 = ch * kCmChunk;
    const int64_t len = min(kCmChunk, g.m - c0);
    uint64_t key[kCmItems];
    float val[kCmItems];
    uint32_t keep = 0;
    uint32_t c = 0;
    const int64_t q0 = (int64_t)threadIdx.x * kCmItems;
    if (q0 < len) {
      int64_t u = -1, uend = 0;
      uint32_t cu = 0;
      bool take = false;
#pragma unroll
      for (int j = 0; j < kCmItems; ++j) {
        key[j] = 0;
        val[j] = 0.0f;
        const int64_t q = q0 + j;
        if (q >= len) continue;
        const int64_t p = c0 + q;
        if (p >= uend) {
          u = VertexOfEdge(g.off, u < 0 ? 0 : u + 1, g.n, p);
          uend = g.off[u + 1];
          cu = ids[u];
          take = f.Cluster(cu);
        }
        if (!take) continue;
        const uint32_t cv = ids[g.nbr[p]];
        if (cv == cu) continue;
        key[j] = ((uint64_t)cu << vbits) | (uint64_t)cv;
        // The reference's lambda parameter is a float: w - offset is computed
        // in double and rounded back to float
        // (parallel-correlation-clusterer-internal.cc:668-669).
        val[j] = (float)((double)g.w[p] - offset);
        keep |= 1u << j;
        ++c;
      }
    }
    uint32_t total = 0;
    const uint32_t at = BlockExclusiveScan(c, sw, &total);
    const int64_t base = chunk_off[ch] + at;
    int64_t k = 0;
#pragma unroll
    for (int j = 0; j < kCmItems; ++j) {
      if (keep & (1u << j)) {
        keys[base + k] = key[j];
        vals[base + k] = val[j];
        ++k;
      }
    }
  }
}

// EdgeSum::Add sums of each run (a left fold in double starting from 0.0).
// Runs of at most kCmShortRun terms are folded by one thread; longer ones are
// queued for k_cm_long_sums.
__global__ __launch_bounds__(256) void k_cm_short_sums(const float* __restrict__ vals,
                                                       const int64_t* __restrict__ start,
                                                       int64_t nseg, double* __restrict__ out,
                                                       uint32_t* __restrict__ longq,
                                                       unsigned long long* __restrict__ nlong) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nseg;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = start[r], e = start[r + 1];
    if (e - s > kCmShortRun) {
      longq[atomicAdd(nlong, 1ull)] = (uint32_t)r;
      continue;
    }
    double acc = 0.0;
    for (int64_t i = s; i < e; ++i) acc += (double)vals[i];
    out[r] = acc;
  }
}

// One wave per queued run.  Every term is a multiple of 2^g; if the sum of
// magnitudes stays below 2^(g + 52) every partial sum is exact, so any order
// gives the left fold's result; otherwise the run is folded in order.
__global__ __launch_bounds__(256) void k_cm_long_sums(const float* __restrict__ vals,
                                                      const int64_t* __restrict__ start,
                                                      const uint32_t* __restrict__ longq,
                                                      int64_t nlong, double* __restrict__ out) {
  const int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (q >= nlong) return;
  const uint32_t r = longq[q];
  const int64_t s = start[r], len = start[r + 1] - s;
  const float* src = vals + s;
  const int lane = LaneId();
  double sum = 0.0, sabs = 0.0;
  int gmin = 1 << 20;
  for (int64_t i = lane; i < len; i += kWave) {
    const double x = (double)src[i];
    sum += x;
    sabs += fabs(x);
    gmin = min(gmin, GranExp(x));
  }
  sum = WaveSum(sum);
  sabs = WaveSum(sabs);
  gmin = WaveMinInt(gmin);
  double r_out;
  if (sabs == 0.0) {
    r_out = 0.0;  // 0.0 + (+-0) + ... = +0
  } else if (sabs < INFINITY && (gmin >= (1 << 19) || sabs < ldexp(1.0, gmin + 52))) {
    r_out = sum;
  } else {
    double acc = 0.0;
    for (int64_t c0 = 0; c0 < len; c0 += kWave) {
      const int64_t nk = min((int64_t)kWave, len - c0);
      const double x = lane < nk ? (double)src[c0 + lane] : 0.0;
      for (int j = 0; j < nk; ++j) acc += __shfl(x, j);
    }
    r_out = acc;
  }
  if (lane == 0) out[r] = r_out;
}

// W_c and member counts of the clusters taking part.
__global__ __launch_bounds__(256) void k_cm_weights(int64_t n, StateView st, CmFilter f,
                                                    double* __restrict__ wsum,
                                                    uint32_t* __restrict__ csum) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t c = st.ids[v];
    if (!f.Cluster(c)) continue;
    atomicAdd(&wsum[c], st.nw[v]);
    atomicAdd(&csum[c], 1u);
  }
}

// delta of every run (c, d) and the per-cluster maximum.
__global__ __launch_bounds__(256) void k_cm_delta(const uint64_t* __restrict__ keys,
                                                  const int64_t* __restrict__ start,
                                                  const double* __restrict__ sums, int64_t nseg,
                                                  int vbits, StateView st, MoveParams mp,
                                                  const double* __restrict__ wsum,
                                                  double* __restrict__ dlt,
                                                  unsigned long long* __restrict__ cmax) {
  const uint64_t vmask = (1ull << vbits) - 1ull;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nseg;
       r += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[start[r]];
    const uint32_t c = (uint32_t)(k >> vbits), d = (uint32_t)(k & vmask);
    const double res = ResOf(mp, c);
    const double W = wsum[c];
    const double change = ChangeOf(W, st.K[c], res);
    // max_edges = W * (K_d - moving weight in d), which is 0 (UTIL.h:104-105).
    const double x = change + NetWeight(sums[r], W * (st.K[d] - 0.0), res);
    dlt[r] = x;
    atomicMax(&cmax[c], OrdBits(x));
  }
}

// Smallest d among the runs of c at the maximum.
__global__ __launch_bounds__(256) void k_cm_pick(const uint64_t* __restrict__ keys,
                                                 const int64_t* __restrict__ start, int64_t nseg,
                                                 int vbits, const double* __restrict__ dlt,
                                                 const unsigned long long* __restrict__ cmax,
                                                 uint32_t* __restrict__ ctgt) {
  const uint64_t vmask = (1ull << vbits) - 1ull;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nseg;
       r += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[start[r]];
    const uint32_t c = (uint32_t)(k >> vbits);
    if (dlt[r] == FromOrd(cmax[c])) atomicMin(&ctgt[c], (uint32_t)(k & vmask));
  }
}

// Per cluster: the best move (BestMoveFromStats' selection), then the sync
// rule (move only to a larger id, PCC.cc:268-276) or none (async), and
// delta > 0.  cmove[c] = target, new_id for "new cluster", or kNoneId.
__global__ __launch_bounds__(256) void k_cm_decide(int64_t n, StateView st, MoveParams mp,
                                                   CmFilter f, int async,
                                                   const double* __restrict__ wsum,
                                                   const unsigned long long* __restrict__ cmax,
                                                   const uint32_t* __restrict__ ctgt,
                                                   uint32_t* __restrict__ cmove,
                                                   double* __restrict__ cdelta,
                                                   unsigned long long* __restrict__ nmoves) {
  unsigned long long moved = 0;
  for (int64_t ci = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; ci < n;
       ci += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t c = (uint32_t)ci;
    uint32_t out = kNoneId;
    if (st.sizes[c] != 0 && f.Cluster(c)) {
      const double change = ChangeOf(wsum[c], st.K[c], ResOf(mp, c));
      uint32_t t = mp.new_id;
      double d = change;
      if (cmax[c] != 0ull) {
        const double mx = FromOrd(cmax[c]);
        if (mx > change) {
          t = ctgt[c];
          d = mx;
        }
      }
      if (!async && t != mp.new_id && c >= t) d = 0;
      if (d > 0) {
        out = t;
        ++moved;
      }
      if (cdelta) cdelta[c] = d;
    }
    cmove[c] = out;
  }
  for (int s = 32; s > 0; s >>= 1) moved += __shfl_xor(moved, s);
  if (LaneId() == 0 && moved) atomicAdd(nmoves, moved);
}

// Sync apply over every vertex: MoveNodesToCluster with moves[v] =
// cmove[c(v)] (same update rules as the vertex phase's apply).
// grp non-null: moves are per group, cmove[grp[v]].
__global__ __launch_bounds__(256) void k_cm_apply_sync(int64_t n, StateView st, uint32_t new_id,
                                                       const uint32_t* __restrict__ cmove,
                                                       const uint32_t* __restrict__ grp,
                                                       uint8_t* __restrict__ modified,
                                                       uint8_t* __restrict__ new_flag,
                                                       RoundCounters* __restrict__ rc) {
  unsigned long long new_movers = 0;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t cur = st.ids[v];
    const uint32_t t = cmove[grp ? grp[v] : cur];
    if (t == kNoneId || t == cur) continue;
    const double k = st.nw[v];
    atomicSub(&st.sizes[cur], 1u);
    atomicAdd(&st.K[cur], -k);
    modified[cur] = 1;
    if (t != new_id) {
      atomicAdd(&st.sizes[t], 1u);
      atomicAdd(&st.K[t], k);
      modified[t] = 1;
      st.ids[v] = t;
    } else {
      new_flag[v] = 1;
      ++new_movers;
    }
  }
  for (int s = 32; s > 0; s >>= 1) new_movers += __shfl_xor(new_movers, s);
  if (LaneId() == 0 && new_movers) atomicAdd(&rc->new_count, new_movers);
}

// Async conflicts within one sub-round, whose clusters all decided on the same
// snapshot.  A sequential pass would let the second cluster heading for t see
// the first one's arrival, so each target admits only its best incoming
// cluster (largest delta, then smallest id): tbest[t] = OrdBits max, tsel[t]
// = smallest id at that maximum.
__global__ __launch_bounds__(256) void k_cm_target_max(int64_t n, uint32_t new_id,
                                                       const uint32_t* __restrict__ cmove,
                                                       const double* __restrict__ cdelta,
                                                       unsigned long long* __restrict__ tbest) {
  for (int64_t ci = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; ci < n;
       ci += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t t = cmove[ci];
    if (t != kNoneId && t != new_id) atomicMax(&tbest[t], OrdBits(cdelta[ci]));
  }
}
__global__ __launch_bounds__(256) void k_cm_target_sel(int64_t n, uint32_t new_id,
                                                       const uint32_t* __restrict__ cmove,
                                                       const double* __restrict__ cdelta,
                                                       const unsigned long long* __restrict__ tbest,
                                                       uint32_t* __restrict__ tsel) {
  for (int64_t ci = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; ci < n;
       ci += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t t = cmove[ci];
    if (t != kNoneId && t != new_id && cdelta[ci] == FromOrd(tbest[t]))
      atomicMin(&tsel[t], (uint32_t)ci);
  }
}
// A cluster admitted by its target still waits when the target itself moves
// away in this sub-round, except in a mutual pair, where the smaller id moves.
__global__ __launch_bounds__(256) void k_cm_yield(int64_t n, uint32_t new_id,
                                                  const uint32_t* __restrict__ cmove,
                                                  const uint32_t* __restrict__ tsel,
                                                  uint32_t* __restrict__ cmove2) {
  for (int64_t ci = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; ci < n;
       ci += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t c = (uint32_t)ci;
    uint32_t t = cmove[c];
    if (t != kNoneId && t != new_id) {
      // Admitted by t, and t stays (or t is heading for c and c < t).
      const bool admitted = tsel[t] == c;
      const uint32_t tt = cmove[t];
      const bool t_leaves = tt != kNoneId && tt != new_id && tsel[tt] == t;
      if (!admitted || (t_leaves && !(tt == c && c < t))) t = kNoneId;
    }
    cmove2[c] = t;
  }
}

// Async cluster state: the set weight is a float reduce in the reference
// (PCCI.cc:416-419); W rounded to float stands for it.  "New cluster" for a
// whole cluster reclaims the cluster's own, now empty, id: no change.
__global__ __launch_bounds__(256) void k_cm_async_update(int64_t n, StateView st, uint32_t new_id,
                                                         const uint32_t* __restrict__ cmove,
                                                         const double* __restrict__ wsum,
                                                         const uint32_t* __restrict__ csum) {
  for (int64_t ci = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; ci < n;
       ci += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t c = (uint32_t)ci;
    const uint32_t t = cmove[c];
    if (t == kNoneId || t == new_id) continue;
    const double wf = (double)(float)wsum[c];
    const uint32_t cnt = csum[c];
    atomicSub(&st.sizes[c], cnt);
    atomicAdd(&st.K[c], -wf);
    atomicAdd(&st.sizes[t], cnt);
    atomicAdd(&st.K[t], wf);
  }
}

// Async member update: ids follow their cluster's move, and moved_clusters
// is marked at the members' vertex ids, as the reference does
// (parallel-correlation-clusterer.cc:257-259).
__global__ __launch_bounds__(256) void k_cm_async_members(int64_t n, uint32_t* __restrict__ ids,
                                                          uint32_t new_id,
                                                          const uint32_t* __restrict__ cmove,
                                                          uint8_t* __restrict__ marks) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t t = cmove[ids[v]];
    if (t == kNoneId) continue;
    marks[v] = 1;
    if (t != new_id) ids[v] = t;
  }
}

// ----------------------------------------------------------- group moves --
// BestMove(graph, moving_nodes) for arbitrary vertex groups (subclusters, or
// clusters as they were before the subcluster moves, which may span several
// clusters by now).  For group S with moving weight W and W_c = its weight
// in cluster c (PCCI.cc:655-688, UTIL.h:79-111):
//   change  = NetWeight(class2 / 2, W^2/2 - sum_c W_c^2/2)
//             - NetWeight(class1_together, sum_c W_c (K_c - W_c))
//   delta_d = change + NetWeight(S_d, W (K_d - W_d))
// class2 = edges inside S between different clusters, class1_together =
// edges leaving S inside their own cluster, S_d = edges leaving S into d.
// Edge runs are keyed (group, code): code d < n for S_d, n for
// class1_together, n + 1 for class2; each is a left fold in CSR order, i.e.
// over the members in ascending order, as the reference's member loop.
// W, W_c and the max_edges sums are exact for integer node weights, as the
// cluster weights are.

__device__ __forceinline__ uint32_t GmRecords(uint32_t gu, uint32_t gv, uint32_t cu, uint32_t cv) {
  if (gu == gv) return cu != cv ? 1u : 0u;
  return cu == cv ? 2u : 1u;
}

__global__ __launch_bounds__(256) void k_gm_vert(int64_t n, StateView st,
                                                 const uint32_t* __restrict__ grp, int vbits,
                                                 double* __restrict__ wsum,
                                                 uint32_t* __restrict__ front,
                                                 uint64_t* __restrict__ vkeys,
                                                 double* __restrict__ vvals) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t gv = grp[v];
    atomicAdd(&wsum[gv], st.nw[v]);
    atomicMin(&front[gv], (uint32_t)v);
    vkeys[v] = ((uint64_t)gv << vbits) | (uint64_t)st.ids[v];
    vvals[v] = st.nw[v];
  }
}

// W_{g,c} of every (group, cluster) run, and the group sums of W_c^2 / 2 and
// W_c (K_c - W_c).
__global__ __launch_bounds__(256) void k_gm_cmw(const uint64_t* __restrict__ vkeys,
                                                const double* __restrict__ vvals,
                                                const int64_t* __restrict__ vstart, int64_t nv,
                                                int vbits, StateView st,
                                                uint64_t* __restrict__ ukeys,
                                                double* __restrict__ cw, double* __restrict__ hsq,
                                                double* __restrict__ fm) {
  const uint64_t vmask = (1ull << vbits) - 1ull;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nv;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s0 = vstart[r], s1 = vstart[r + 1];
    double acc = 0.0;
    for (int64_t i = s0; i < s1; ++i) acc += vvals[i];
    const uint64_t k = vkeys[s0];
    const uint32_t gg = (uint32_t)(k >> vbits), c = (uint32_t)(k & vmask);
    ukeys[r] = k;
    cw[r] = acc;
    atomicAdd(&hsq[gg], acc * acc / 2);
    atomicAdd(&fm[gg], acc * (st.K[c] - acc));
  }
}

// One pass over kCmItems consecutive CSR edges of a thread: the number of
// records (count) or the records themselves (emit, at out).
template <bool kEmit>
__device__ __forceinline__ uint32_t GmEdges(const CsrView& g, const uint32_t* __restrict__ ids,
                                            const uint32_t* __restrict__ grp, int64_t c0,
                                            int64_t q0, int64_t len, int cbits, double offset,
                                            uint64_t* __restrict__ keys,
                                            float* __restrict__ vals, int64_t out) {
  uint32_t c = 0;
  if (q0 >= len) return 0;
  const int64_t q1 = min(q0 + (int64_t)kCmItems, len);
  const uint64_t n = (uint64_t)g.n;
  int64_t u = -1, uend = 0;
  uint32_t gu = 0, cu = 0;
  for (int64_t q = q0; q < q1; ++q) {
    const int64_t p = c0 + q;
    if (p >= uend) {
      u = VertexOfEdge(g.off, u < 0 ? 0 : u + 1, g.n, p);
      uend = g.off[u + 1];
      gu = grp[u];
      cu = ids[u];
    }
    const uint32_t v = g.nbr[p];
    const uint32_t gv = grp[v], cv = ids[v];
    if (!kEmit) {
      c += GmRecords(gu, gv, cu, cv);
      continue;
    }
    // The reference's lambda parameter is a float: w - offset rounded back
    // to float (PCCI.cc:668-669).
    const float w = (float)((double)g.w[p] - offset);
    const uint64_t gk = (uint64_t)gu << cbits;
    if (gu == gv) {
      if (cu != cv) {
        keys[out + c] = gk | (n + 1);
        vals[out + c] = w;
        ++c;
      }
    } else {
      keys[out + c] = gk | (uint64_t)cv;
      vals[out + c] = w;
      ++c;
      if (cu == cv) {
        keys[out + c] = gk | n;
        vals[out + c] = w;
        ++c;
      }
    }
  }
  return c;
}

__global__ __launch_bounds__(kCmThreads) void k_gm_count(CsrView g, const uint32_t* __restrict__ ids,
                                                         const uint32_t* __restrict__ grp,
                                                         int64_t nchunks,
                                                         uint32_t* __restrict__ chunk_cnt) {
  __shared__ uint32_t sw[kCmThreads / kWave];
  for (int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const int64_t c0 = ch * kCmChunk;
    const int64_t len = min(kCmChunk, g.m - c0);
    const uint32_t c = GmEdges<false>(g, ids, grp, c0, (int64_t)threadIdx.x * kCmItems, len, 0,
                                      0.0, nullptr, nullptr, 0);
    uint32_t total = 0;
    (void)BlockExclusiveScan(c, sw, &total);
    if (threadIdx.x == 0) chunk_cnt[ch] = total;
  }
}

__global__ __launch_bounds__(kCmThreads) void k_gm_emit(CsrView g, const uint32_t* __restrict__ ids,
                                                        const uint32_t* __restrict__ grp,
                                                        int64_t nchunks,
                                                        const int64_t* __restrict__ chunk_off,
                                                        int cbits, double offset,
                                                        uint64_t* __restrict__ keys,
                                                        float* __restrict__ vals) {
  __shared__ uint32_t sw[kCmThreads / kWave];
  for (int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const int64_t c0 = ch * kCmChunk;
    const int64_t len = min(kCmChunk, g.m - c0);
    const int64_t q0 = (int64_t)threadIdx.x * kCmItems;
    const uint32_t c = GmEdges<false>(g, ids, grp, c0, q0, len, 0, 0.0, nullptr, nullptr, 0);
    uint32_t total = 0;
    const uint32_t at = BlockExclusiveScan(c, sw, &total);
    (void)GmEdges<true>(g, ids, grp, c0, q0, len, cbits, offset, keys, vals,
                        chunk_off[ch] + (int64_t)at);
  }
}

// class1_together and class2 sums of every group.
__global__ __launch_bounds__(256) void k_gm_special(const uint64_t* __restrict__ keys,
                                                    const int64_t* __restrict__ start,
                                                    const double* __restrict__ sums, int64_t nseg,
                                                    int cbits, uint64_t n,
                                                    double* __restrict__ c1,
                                                    double* __restrict__ c2) {
  const uint64_t cmask = (1ull << cbits) - 1ull;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nseg;
       r += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[start[r]];
    const uint64_t code = k & cmask;
    if (code == n) c1[k >> cbits] = sums[r];
    if (code == n + 1) c2[k >> cbits] = sums[r];
  }
}

__global__ __launch_bounds__(256) void k_gm_change(int64_t n, MoveParams mp,
                                                   const uint32_t* __restrict__ front,
                                                   const double* __restrict__ wsum,
                                                   const double* __restrict__ hsq,
                                                   const double* __restrict__ fm,
                                                   const double* __restrict__ c1,
                                                   const double* __restrict__ c2,
                                                   double* __restrict__ chg) {
  for (int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; gi < n;
       gi += (int64_t)gridDim.x * blockDim.x) {
    if (front[gi] == kNoneId) continue;
    const double W = wsum[gi];
    const double me2 = W * W / 2 - hsq[gi];
    double change = 0.0;
    change += NetWeight(c2[gi] / 2, me2, mp.res);  // RemoveDoubleCounting, UTIL.h:32
    change -= NetWeight(c1[gi], fm[gi], mp.res);
    chg[gi] = change;
  }
}

// W_{g,d} by binary search over the sorted (group, cluster) keys; 0 if S has
// no member in d.
__device__ __forceinline__ double MovingIn(const uint64_t* __restrict__ ukeys,
                                           const double* __restrict__ cw, int64_t nv,
                                           uint64_t key) {
  int64_t lo = 0, hi = nv;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (ukeys[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < nv && ukeys[lo] == key ? cw[lo] : 0.0;
}

__global__ __launch_bounds__(256) void k_gm_delta(const uint64_t* __restrict__ keys,
                                                  const int64_t* __restrict__ start,
                                                  const double* __restrict__ sums, int64_t nseg,
                                                  int cbits, int vbits, uint64_t n, StateView st,
                                                  MoveParams mp, const double* __restrict__ wsum,
                                                  const double* __restrict__ chg,
                                                  const uint64_t* __restrict__ ukeys,
                                                  const double* __restrict__ cw, int64_t nv,
                                                  double* __restrict__ dlt,
                                                  unsigned long long* __restrict__ gmax) {
  const uint64_t cmask = (1ull << cbits) - 1ull;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nseg;
       r += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[start[r]];
    const uint64_t d = k & cmask;
    if (d >= n) continue;
    const uint32_t gg = (uint32_t)(k >> cbits);
    const double W = wsum[gg];
    const double wd = MovingIn(ukeys, cw, nv, ((uint64_t)gg << vbits) | d);
    const double x = chg[gg] + NetWeight(sums[r], W * (st.K[d] - wd), mp.res);
    dlt[r] = x;
    atomicMax(&gmax[gg], OrdBits(x));
  }
}

__global__ __launch_bounds__(256) void k_gm_pick(const uint64_t* __restrict__ keys,
                                                 const int64_t* __restrict__ start, int64_t nseg,
                                                 int cbits, uint64_t n,
                                                 const double* __restrict__ dlt,
                                                 const unsigned long long* __restrict__ gmax,
                                                 uint32_t* __restrict__ gtgt) {
  const uint64_t cmask = (1ull << cbits) - 1ull;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nseg;
       r += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[start[r]];
    const uint64_t d = k & cmask;
    if (d >= n) continue;
    const uint32_t gg = (uint32_t)(k >> cbits);
    if (dlt[r] == FromOrd(gmax[gg])) atomicMin(&gtgt[gg], (uint32_t)d);
  }
}

// Best move of every group, then the sync rule against the cluster of the
// group's first member (PCC.cc:224-233, 270-279) and delta > 0.
__global__ __launch_bounds__(256) void k_gm_decide(int64_t n, StateView st, MoveParams mp,
                                                   const uint32_t* __restrict__ front,
                                                   const double* __restrict__ chg,
                                                   const unsigned long long* __restrict__ gmax,
                                                   const uint32_t* __restrict__ gtgt,
                                                   uint32_t* __restrict__ gmove,
                                                   unsigned long long* __restrict__ nmoves) {
  unsigned long long moved = 0;
  for (int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; gi < n;
       gi += (int64_t)gridDim.x * blockDim.x) {
    uint32_t out = kNoneId;
    if (front[gi] != kNoneId) {
      uint32_t t = mp.new_id;
      double d = chg[gi];
      if (gmax[gi] != 0ull) {
        const double mx = FromOrd(gmax[gi]);
        if (mx > d) {
          t = gtgt[gi];
          d = mx;
        }
      }
      if (t != mp.new_id && st.ids[front[gi]] >= t) d = 0;
      if (d > 0) {
        out = t;
        ++moved;
      }
    }
    gmove[gi] = out;
  }
  for (int s = 32; s > 0; s >>= 1) moved += __shfl_xor(moved, s);
  if (LaneId() == 0 && moved) atomicAdd(nmoves, moved);
}

}  // namespace

int64_t RunGroupMoves(DeviceAllocator* al, const CsrView& g, const StateView& st,
                      const MoveParams& mp, const uint32_t* grp, uint8_t* modified,
                      uint8_t* new_flag, RoundCounters* rc, hipStream_t s) {
  const int64_t n = g.n;
  if (n == 0) return 0;
  const int vbits = CmBits((uint64_t)(n > 1 ? n - 1 : 1));
  const int cbits = CmBits((uint64_t)n + 1);
  const int grid_n = CmGrid(n, 256, 8192);
  DevBuf<double> wsum(al, n), hsq(al, n), fm(al, n), c1(al, n), c2(al, n), chg(al, n);
  DevBuf<uint32_t> front(al, n), gtgt(al, n), gmove(al, n);
  DevBuf<unsigned long long> gmax(al, n), nmoves(al, 1);
  for (double* p : {wsum.get(), hsq.get(), fm.get(), c1.get(), c2.get()})
    PCC_HIP_CHECK(hipMemsetAsync(p, 0, sizeof(double) * (size_t)n, s));
  PCC_HIP_CHECK(hipMemsetAsync(front.get(), 0xFF, sizeof(uint32_t) * (size_t)n, s));
  PCC_HIP_CHECK(hipMemsetAsync(gtgt.get(), 0xFF, sizeof(uint32_t) * (size_t)n, s));
  PCC_HIP_CHECK(hipMemsetAsync(gmax.get(), 0, sizeof(unsigned long long) * (size_t)n, s));
  PCC_HIP_CHECK(hipMemsetAsync(nmoves.get(), 0, sizeof(unsigned long long), s));
  // Moving weight per (group, cluster).
  DevBuf<uint64_t> vkeys(al, n);
  DevBuf<double> vvals(al, n);
  hipLaunchKernelGGL(k_gm_vert, dim3(grid_n), dim3(256), 0, s, n, st, grp, vbits, wsum.get(),
                     front.get(), vkeys.get(), vvals.get());
  PCC_HIP_CHECK(hipGetLastError());
  RadixSortPairsF64(al, vkeys.get(), vvals.get(), n, 0, 2 * vbits, s);
  DevBuf<int64_t> vstart;
  const int64_t nv = Segments(al, vkeys.get(), n, &vstart, s);
  DevBuf<uint64_t> ukeys(al, nv);
  DevBuf<double> cw(al, nv);
  hipLaunchKernelGGL(k_gm_cmw, dim3(CmGrid(nv, 256, 8192)), dim3(256), 0, s, vkeys.get(),
                     vvals.get(), vstart.get(), nv, vbits, st, ukeys.get(), cw.get(), hsq.get(),
                     fm.get());
  PCC_HIP_CHECK(hipGetLastError());
  vkeys.Release();
  vvals.Release();
  // Edge runs keyed (group, code).
  const int64_t nchunks = (g.m + kCmChunk - 1) / kCmChunk;
  const int cgrid = (int)std::min<int64_t>(nchunks > 0 ? nchunks : 1, 1 << 16);
  int64_t me = 0;
  DevBuf<int64_t> coff(al, nchunks + 1);
  if (nchunks > 0) {
    DevBuf<uint32_t> ccnt(al, nchunks);
    hipLaunchKernelGGL(k_gm_count, dim3(cgrid), dim3(kCmThreads), 0, s, g, st.ids, grp, nchunks,
                       ccnt.get());
    PCC_HIP_CHECK(hipGetLastError());
    ExclusiveScanU32toI64(al, ccnt.get(), coff.get(), nchunks, coff.get() + nchunks, s);
    PCC_HIP_CHECK(hipMemcpyAsync(&me, coff.get() + nchunks, sizeof(int64_t),
                                 hipMemcpyDeviceToHost, s));
    PCC_HIP_CHECK(hipStreamSynchronize(s));
  }
  DevBuf<uint64_t> keys;
  DevBuf<int64_t> start;
  DevBuf<double> sums, dlt;
  int64_t nseg = 0;
  if (me > 0) {
    keys.Reset(al, me);
    DevBuf<float> vals(al, me);
    hipLaunchKernelGGL(k_gm_emit, dim3(cgrid), dim3(kCmThreads), 0, s, g, st.ids, grp, nchunks,
                       coff.get(), cbits, mp.offset, keys.get(), vals.get());
    PCC_HIP_CHECK(hipGetLastError());
    RadixSortPairsF32(al, keys.get(), vals.get(), me, 0, vbits + cbits, s);
    nseg = Segments(al, keys.get(), me, &start, s);
    sums.Reset(al, nseg);
    dlt.Reset(al, nseg);
    DevBuf<uint32_t> longq(al, nseg);
    DevBuf<unsigned long long> nlong(al, 1);
    PCC_HIP_CHECK(hipMemsetAsync(nlong.get(), 0, sizeof(unsigned long long), s));
    const int grid_r = CmGrid(nseg, 256, 8192);
    hipLaunchKernelGGL(k_cm_short_sums, dim3(grid_r), dim3(256), 0, s, vals.get(), start.get(),
                       nseg, sums.get(), longq.get(), nlong.get());
    PCC_HIP_CHECK(hipGetLastError());
    if (me > kCmShortRun) {
      unsigned long long hl = 0;
      PCC_HIP_CHECK(hipMemcpyAsync(&hl, nlong.get(), sizeof(hl), hipMemcpyDeviceToHost, s));
      PCC_HIP_CHECK(hipStreamSynchronize(s));
      if (hl > 0) {
        constexpr int kWavesPerBlock = 256 / kWave;
        hipLaunchKernelGGL(k_cm_long_sums,
                           dim3((unsigned)((hl + kWavesPerBlock - 1) / kWavesPerBlock)),
                           dim3(256), 0, s, vals.get(), start.get(), longq.get(), (int64_t)hl,
                           sums.get());
        PCC_HIP_CHECK(hipGetLastError());
      }
    }
    hipLaunchKernelGGL(k_gm_special, dim3(grid_r), dim3(256), 0, s, keys.get(), start.get(),
                       sums.get(), nseg, cbits, (uint64_t)n, c1.get(), c2.get());
    PCC_HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_gm_change, dim3(grid_n), dim3(256), 0, s, n, mp, front.get(), wsum.get(),
                     hsq.get(), fm.get(), c1.get(), c2.get(), chg.get());
  PCC_HIP_CHECK(hipGetLastError());
  if (nseg > 0) {
    const int grid_r = CmGrid(nseg, 256, 8192);
    hipLaunchKernelGGL(k_gm_delta, dim3(grid_r), dim3(256), 0, s, keys.get(), start.get(),
                       sums.get(), nseg, cbits, vbits, (uint64_t)n, st, mp, wsum.get(), chg.get(),
                       ukeys.get(), cw.get(), nv, dlt.get(), gmax.get());
    hipLaunchKernelGGL(k_gm_pick, dim3(grid_r), dim3(256), 0, s, keys.get(), start.get(), nseg,
                       cbits, (uint64_t)n, dlt.get(), gmax.get(), gtgt.get());
    PCC_HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_gm_decide, dim3(grid_n), dim3(256), 0, s, n, st, mp, front.get(),
                     chg.get(), gmax.get(), gtgt.get(), gmove.get(), nmoves.get());
  PCC_HIP_CHECK(hipGetLastError());
  // MoveNodesToCluster(moves) (PCC.cc:241, 285).
  PCC_HIP_CHECK(hipMemsetAsync(&rc->new_count, 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL(k_cm_apply_sync, dim3(grid_n), dim3(256), 0, s, n, st, mp.new_id,
                     gmove.get(), grp, modified, new_flag, rc);
  PCC_HIP_CHECK(hipGetLastError());
  RunNewClusterAssign(al, st, (uint32_t)n, mp.vgraph, mp.gbase, new_flag, modified, rc, s);
  unsigned long long hm = 0;
  PCC_HIP_CHECK(hipMemcpyAsync(&hm, nmoves.get(), sizeof(hm), hipMemcpyDeviceToHost, s));
  PCC_HIP_CHECK(hipStreamSynchronize(s));
  return (int64_t)hm;
}

int64_t RunClusterMoves(DeviceAllocator* al, const CsrView& g, const StateView& st,
                        const MoveParams& mp, int G, const int64_t counts[kBins],
                        const BinLayout& layout, const uint32_t* list, bool async,
                        uint32_t sub_rounds, uint32_t salt, uint8_t* modified, uint8_t* new_flag,
                        RoundCounters* rc, hipStream_t s) {
  const int64_t n = g.n;
  if (n == 0) return 0;
  const int vbits = CmBits((uint64_t)(n > 1 ? n - 1 : 1));
  const int grid_n = CmGrid(n, 256, 8192);
  CmFilter f;
  DevBuf<uint8_t> gact;
  if (mp.vgraph) {
    // Only graphs whose round runs (a non-empty frontier) move clusters.
    gact.Reset(al, G);
    PCC_HIP_CHECK(hipMemsetAsync(gact.get(), 0, (size_t)G, s));
    CmBins b;
    int64_t total = 0;
    for (int i = 0; i < kBins; ++i) {
      b.base[i] = layout.base[i];
      b.count[i] = counts[i];
      total += counts[i];
    }
    if (total > 0)
      hipLaunchKernelGGL(k_cm_active, dim3(CmGrid(total, 256, 4096)), dim3(256), 0, s, b, list,
                         mp.vgraph, gact.get());
    f.vgraph = mp.vgraph;
    f.gact = gact.get();
  }
  const uint32_t S = async ? (sub_rounds < 1 ? 1 : sub_rounds) : 1;
  f.sub_count = S;
  f.salt = salt;
  DevBuf<double> wsum(al, n);
  DevBuf<uint32_t> csum(al, n), ctgt(al, n), cmove(al, n), cmove2, tsel;
  DevBuf<unsigned long long> cmax(al, n), nmoves(al, 1), tbest;
  DevBuf<double> cdelta;
  if (async) {
    cmove2.Reset(al, n);
    tsel.Reset(al, n);
    tbest.Reset(al, n);
    cdelta.Reset(al, n);
  }
  PCC_HIP_CHECK(hipMemsetAsync(nmoves.get(), 0, sizeof(unsigned long long), s));
  const int64_t nchunks = (g.m + kCmChunk - 1) / kCmChunk;
  DevBuf<uint32_t> ccnt(al, nchunks > 0 ? nchunks : 1);
  DevBuf<int64_t> coff(al, nchunks + 1);
  const int cgrid = (int)std::min<int64_t>(nchunks > 0 ? nchunks : 1, 1 << 16);
  for (uint32_t sub = 0; sub < S; ++sub) {
    f.sub_index = sub;
    PCC_HIP_CHECK(hipMemsetAsync(wsum.get(), 0, sizeof(double) * (size_t)n, s));
    PCC_HIP_CHECK(hipMemsetAsync(csum.get(), 0, sizeof(uint32_t) * (size_t)n, s));
    PCC_HIP_CHECK(hipMemsetAsync(cmax.get(), 0, sizeof(unsigned long long) * (size_t)n, s));
    PCC_HIP_CHECK(hipMemsetAsync(ctgt.get(), 0xFF, sizeof(uint32_t) * (size_t)n, s));
    hipLaunchKernelGGL(k_cm_weights, dim3(grid_n), dim3(256), 0, s, n, st, f, wsum.get(),
                       csum.get());
    PCC_HIP_CHECK(hipGetLastError());
    // Inter-cluster edges of the clusters taking part, keyed (c(u), c(v)).
    int64_t me = 0;
    if (nchunks > 0) {
      hipLaunchKernelGGL(k_cm_count, dim3(cgrid), dim3(kCmThreads), 0, s, g, st.ids, f, nchunks,
                         ccnt.get());
      PCC_HIP_CHECK(hipGetLastError());
      ExclusiveScanU32toI64(al, ccnt.get(), coff.get(), nchunks, coff.get() + nchunks, s);
      PCC_HIP_CHECK(hipMemcpyAsync(&me, coff.get() + nchunks, sizeof(int64_t),
                                   hipMemcpyDeviceToHost, s));
      PCC_HIP_CHECK(hipStreamSynchronize(s));
    }
    if (me > 0) {
      DevBuf<uint64_t> keys(al, me);
      DevBuf<float> vals(al, me);
      hipLaunchKernelGGL(k_cm_emit, dim3(cgrid), dim3(kCmThreads), 0, s, g, st.ids, f, nchunks,
                         coff.get(), vbits, mp.offset, keys.get(), vals.get());
      PCC_HIP_CHECK(hipGetLastError());
      RadixSortPairsF32(al, keys.get(), vals.get(), me, 0, 2 * vbits, s);
      DevBuf<int64_t> start;
      const int64_t nseg = Segments(al, keys.get(), me, &start, s);
      DevBuf<double> sums(al, nseg), dlt(al, nseg);
      DevBuf<uint32_t> longq(al, nseg);
      DevBuf<unsigned long long> nlong(al, 1);
      PCC_HIP_CHECK(hipMemsetAsync(nlong.get(), 0, sizeof(unsigned long long), s));
      const int grid_r = CmGrid(nseg, 256, 8192);
      hipLaunchKernelGGL(k_cm_short_sums, dim3(grid_r), dim3(256), 0, s, vals.get(), start.get(),
                         nseg, sums.get(), longq.get(), nlong.get());
      PCC_HIP_CHECK(hipGetLastError());
      if (me > kCmShortRun) {
        unsigned long long hl = 0;
        PCC_HIP_CHECK(hipMemcpyAsync(&hl, nlong.get(), sizeof(hl), hipMemcpyDeviceToHost, s));
        PCC_HIP_CHECK(hipStreamSynchronize(s));
        if (hl > 0) {
          constexpr int kWavesPerBlock = 256 / kWave;
          hipLaunchKernelGGL(k_cm_long_sums,
                             dim3((unsigned)((hl + kWavesPerBlock - 1) / kWavesPerBlock)),
                             dim3(256), 0, s, vals.get(), start.get(), longq.get(), (int64_t)hl,
                             sums.get());
          PCC_HIP_CHECK(hipGetLastError());
        }
      }
      hipLaunchKernelGGL(k_cm_delta, dim3(grid_r), dim3(256), 0, s, keys.get(), start.get(),
                         sums.get(), nseg, vbits, st, mp, wsum.get(), dlt.get(), cmax.get());
      hipLaunchKernelGGL(k_cm_pick, dim3(grid_r), dim3(256), 0, s, keys.get(), start.get(), nseg,
                         vbits, dlt.get(), cmax.get(), ctgt.get());
      PCC_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_cm_decide, dim3(grid_n), dim3(256), 0, s, n, st, mp, f, (int)async,
                       wsum.get(), cmax.get(), ctgt.get(), cmove.get(), cdelta.get(),
                       nmoves.get());
    PCC_HIP_CHECK(hipGetLastError());
    if (!async) {
      // MoveNodesToCluster(moves) (PCC.cc:285): new-cluster movers are
      // counted afresh for the empty-id assignment.
      PCC_HIP_CHECK(hipMemsetAsync(&rc->new_count, 0, sizeof(unsigned long long), s));
      hipLaunchKernelGGL(k_cm_apply_sync, dim3(grid_n), dim3(256), 0, s, n, st, mp.new_id,
                         cmove.get(), nullptr, modified, new_flag, rc);
      PCC_HIP_CHECK(hipGetLastError());
      RunNewClusterAssign(al, st, (uint32_t)n, mp.vgraph, mp.gbase, new_flag, modified, rc, s);
    } else {
      PCC_HIP_CHECK(hipMemsetAsync(tbest.get(), 0, sizeof(unsigned long long) * (size_t)n, s));
      PCC_HIP_CHECK(hipMemsetAsync(tsel.get(), 0xFF, sizeof(uint32_t) * (size_t)n, s));
      hipLaunchKernelGGL(k_cm_target_max, dim3(grid_n), dim3(256), 0, s, n, mp.new_id,
                         cmove.get(), cdelta.get(), tbest.get());
      hipLaunchKernelGGL(k_cm_target_sel, dim3(grid_n), dim3(256), 0, s, n, mp.new_id,
                         cmove.get(), cdelta.get(), tbest.get(), tsel.get());
      hipLaunchKernelGGL(k_cm_yield, dim3(grid_n), dim3(256), 0, s, n, mp.new_id, cmove.get(),
                         tsel.get(), cmove2.get());
      hipLaunchKernelGGL(k_cm_async_update, dim3(grid_n), dim3(256), 0, s, n, st, mp.new_id,
                         cmove2.get(), wsum.get(), csum.get());
      hipLaunchKernelGGL(k_cm_async_members, dim3(grid_n), dim3(256), 0, s, n, st.ids, mp.new_id,
                         cmove2.get(), modified);
      PCC_HIP_CHECK(hipGetLastError());
    }
  }
  unsigned long long hm = 0;
  PCC_HIP_CHECK(hipMemcpyAsync(&hm, nmoves.get(), sizeof(hm), hipMemcpyDeviceToHost, s));
  PCC_HIP_CHECK(hipStreamSynchronize(s));
  return (int64_t)hm;
}

}  // namespace pcc
