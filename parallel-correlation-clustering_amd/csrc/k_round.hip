// k_round.hip -- the rest of one best-move round: the sync apply
// (ClusteringHelper::MoveNodesToCluster,
// clustering/clusterers/parallel-correlation-clusterer-internal.cc:256-369),
// the empty-id reuse for vertices forming new clusters (same function,
// :339-366) and the next-frontier expansion
// (clustering/clusterers/parallel-correlation-clusterer.cc:299-337).
//
// Cluster sizes and weights are updated with device atomics.  The reference
// subtracts/adds node weights group by group; with integer node weights (unit
// weights, degrees, and their aggregates on contracted levels) every partial
// sum is an exact integer, so the atomic order cannot change a result.

#include <algorithm>
#include <memory>

#include "device_common.h"
#include "kernels.h"

namespace pcc {

namespace {

struct Bins {
  int64_t base[kBins];
  int64_t count[kBins];
};

// Device-driven rounds: the launch sized bins by their capacities; the bins
// a launch covers (count > 0) take their current counts from dcnt.  The
// kernel argument itself is never written (a written copy indexed by a
// variable would live in scratch memory).
__device__ __forceinline__ int64_t BinCount(const Bins& bins, const unsigned long long* dcnt,
                                            int b) {
  return (dcnt && bins.count[b] > 0) ? (int64_t)dcnt[b] : bins.count[b];
}
__device__ __forceinline__ int64_t BinTotal(const Bins& bins, const unsigned long long* dcnt) {
  int64_t total = 0;
#pragma unroll
  for (int b = 0; b < kBins; ++b) total += BinCount(bins, dcnt, b);
  return total;
}

__global__ __launch_bounds__(256) void k_apply(Bins bins, const uint32_t* __restrict__ list,
                                               StateView st, uint32_t new_id,
                                               const uint32_t* __restrict__ moves,
                                               uint8_t* __restrict__ modified,
                                               uint8_t* __restrict__ new_flag,
                                               RoundCounters* __restrict__ rc,
                                               const unsigned long long* __restrict__ dcnt) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  unsigned long long new_movers = 0;
  for (int b = 0; b < kBins; ++b) {
    const int64_t cb = BinCount(bins, dcnt, b);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cb; i += stride) {
      const uint32_t v = list[bins.base[b] + i];
      const uint32_t t = moves[v];
      if (t == kNoneId) continue;
      const uint32_t cur = st.ids[v];
      if (t == cur) continue;
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
  }
  // One global atomic per workgroup: atomics on one address from every wave
  // of a large grid serialize.
  __shared__ unsigned long long s_new[256 / kWave];
  for (int s = 32; s > 0; s >>= 1) new_movers += __shfl_xor(new_movers, s);
  if (LaneId() == 0) s_new[threadIdx.x / kWave] = new_movers;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += s_new[w];
    if (t) atomicAdd(&rc->new_count, t);
  }
}

__global__ __launch_bounds__(256) void k_zero_flags(uint32_t n, const uint32_t* __restrict__ sizes,
                                                    uint8_t* __restrict__ z,
                                                    const unsigned long long* __restrict__ gate) {
  if (*gate == 0) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    z[i] = sizes[i] == 0 ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_zero_compact(uint32_t n, const uint8_t* __restrict__ z,
                                                      const uint32_t* __restrict__ zrank,
                                                      uint32_t* __restrict__ zero_list,
                                                      const unsigned long long* __restrict__ gate) {
  if (*gate == 0) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if (z[i]) zero_list[zrank[i]] = (uint32_t)i;
}

// i-th new-cluster mover of a graph in ascending vertex order takes the i-th
// empty id of that graph in ascending order (PCCI.cc:346-365).  Exclusive
// ranks over the whole level minus the rank at the graph's first vertex give
// ranks within the graph.
__global__ __launch_bounds__(256) void k_new_assign(uint32_t n, StateView st,
                                                    const uint32_t* __restrict__ vgraph,
                                                    const int64_t* __restrict__ gbase,
                                                    uint8_t* __restrict__ new_flag,
                                                    const uint32_t* __restrict__ nrank,
                                                    const uint32_t* __restrict__ zrank,
                                                    const uint32_t* __restrict__ zero_list,
                                                    uint8_t* __restrict__ modified,
                                                    const unsigned long long* __restrict__ gate) {
  if (*gate == 0) return;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    if (!new_flag[v]) continue;
    int64_t base = 0, end = n;
    if (vgraph) {
      const uint32_t g = vgraph[v];
      base = gbase[g];
      end = gbase[g + 1];
    }
    const uint32_t r = nrank[v] - nrank[base];
    const uint32_t c = zero_list[zrank[base] + r];
    new_flag[v] = 0;
    // Sizes summing to the graph's vertex count guarantee enough empty ids;
    // inconsistent caller state must not turn into an out-of-range write.
    if ((int64_t)c < base || (int64_t)c >= end) continue;
    st.ids[v] = c;
    st.sizes[c] = 1;
    modified[c] = 1;
    st.K[c] = st.nw[v];
  }
}

// Levels of at most kSmallAssignMax vertices run the whole empty-id
// assignment in one workgroup (one launch instead of a dozen; the same
// result): thread t owns the contiguous ids [t * per, (t + 1) * per).
constexpr int kSmallAssignThreads = 1024;
constexpr uint32_t kSmallAssignMax = 1u << 18;
__global__ __launch_bounds__(kSmallAssignThreads) void k_new_assign_small(
    uint32_t n, StateView st, const uint32_t* __restrict__ vgraph,
    const int64_t* __restrict__ gbase, uint8_t* __restrict__ new_flag,
    uint32_t* __restrict__ zero_list, uint32_t* __restrict__ zrank, uint32_t* __restrict__ nrank,
    uint8_t* __restrict__ modified, const unsigned long long* __restrict__ gate) {
  if (*gate == 0) return;
  __shared__ uint32_t sw[kSmallAssignThreads / kWave];
  const uint32_t per = (n + blockDim.x - 1) / blockDim.x;
  const uint32_t lo = min(n, threadIdx.x * per), hi = min(n, lo + per);
  uint32_t z = 0, m = 0;
  for (uint32_t i = lo; i < hi; ++i) {
    z += st.sizes[i] == 0 ? 1u : 0u;
    m += new_flag[i] ? 1u : 0u;
  }
  uint32_t zt = 0, mt = 0;
  uint32_t zr = BlockExclusiveScan(z, sw, &zt);
  uint32_t mr = BlockExclusiveScan(m, sw, &mt);
  for (uint32_t i = lo; i < hi; ++i) {
    zrank[i] = zr;
    nrank[i] = mr;
    if (st.sizes[i] == 0) zero_list[zr++] = i;
    if (new_flag[i]) ++mr;
  }
  __threadfence_block();
  __syncthreads();
  for (uint32_t v = lo; v < hi; ++v) {
    if (!new_flag[v]) continue;
    int64_t base = 0, end = n;
    if (vgraph) {
      const uint32_t g = vgraph[v];
      base = gbase[g];
      end = gbase[g + 1];
    }
    const uint32_t idx = zrank[base] + (nrank[v] - nrank[base]);
    new_flag[v] = 0;
    if (idx >= zt) continue;  // inconsistent caller state: no write out of range
    const uint32_t c = zero_list[idx];
    if ((int64_t)c < base || (int64_t)c >= end) continue;
    st.ids[v] = c;
    st.sizes[c] = 1;
    modified[c] = 1;
    st.K[c] = st.nw[v];
  }
}

// Per-round clears in one launch: up to three byte arrays of n entries, the
// round counters selected by `counters` (bit 0: cnt / te, bit 1: fb_count /
// new_count / repair_count, folded into fb_total first) and one 64-bit
// accumulator.
__global__ __launch_bounds__(256) void k_round_reset(int64_t n, uint8_t* __restrict__ a,
                                                     uint8_t* __restrict__ b,
                                                     uint8_t* __restrict__ c,
                                                     RoundCounters* __restrict__ rc, int counters,
                                                     unsigned long long* __restrict__ acc,
                                                     const unsigned int* __restrict__ c_gate) {
  if (c_gate && *c_gate == 0) c = nullptr;
  RoundResetSpan((int64_t)blockIdx.x * blockDim.x + threadIdx.x, (int64_t)gridDim.x * blockDim.x, n,
                 a, b, c, rc, counters, acc, kBins);
}

// 16 lanes per seed vertex mark its neighbours.
// Marks the neighbours of the selected vertices.  Each thread owns one
// vertex; degrees up to 16 are walked by the thread itself, larger ones by
// the whole wave in turn.
__device__ __forceinline__ void MarkNeighbours(bool sel, int64_t st, int64_t deg,
                                               const uint32_t* __restrict__ nbr,
                                               uint8_t* __restrict__ next) {
  if (sel && deg <= 16)
    for (int64_t p = 0; p < deg; ++p) next[nbr[st + p]] = 1;
  uint64_t big = __ballot(sel && deg > 16);
  while (big) {
    const int l = __ffsll((unsigned long long)big) - 1;
    big &= big - 1;
    const int64_t vs = __shfl(st, l);
    const int64_t vd = __shfl(deg, l);
    for (int64_t p = LaneId(); p < vd; p += kWave) next[nbr[vs + p]] = 1;
  }
}

__device__ __forceinline__ int64_t WaveUniform(int64_t x) {
  const uint64_t u = (uint64_t)x;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Moved vertices of bins 5 and 6: one wave per vertex, four per workgroup,
// 64 neighbours per step (two steps' loads in flight).  The vertex index is
// uniform over the wave, so the bin search runs on the scalar unit; with
// device counts (kDev) the loaded counts are made wave-uniform explicitly so
// that it stays there.
template <bool kDev>
__global__ __launch_bounds__(256) void k_expand_block(Bins bins, int64_t total,
                                                      const uint32_t* __restrict__ list,
                                                      const int64_t* __restrict__ off,
                                                      const uint32_t* __restrict__ nbr,
                                                      const uint8_t* __restrict__ moved,
                                                      ExpandDir dir, uint8_t* __restrict__ next,
                                                      const unsigned long long* __restrict__ dcnt) {
  if (dir.Pull()) return;
  if (kDev) total = WaveUniform(BinTotal(bins, dcnt));
  const int lane = LaneId();
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / kWave);
  for (int64_t t = WaveUniform(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave); t < total;
       t += nwaves) {
    int64_t r = t;
    int b = 0;
    if (kDev) {
      while (r >= WaveUniform(BinCount(bins, dcnt, b))) r -= WaveUniform(BinCount(bins, dcnt, b++));
    } else {
      while (r >= bins.count[b]) r -= bins.count[b++];
    }
    const uint32_t v = list[bins.base[b] + r];
    if (!moved[v]) continue;
    const int64_t st = off[v], en = off[v + 1];
    int64_t p = st + lane;
    for (; p + kWave < en; p += 2 * kWave) {
      const uint32_t a = nbr[p], c = nbr[p + kWave];
      next[a] = 1;
      next[c] = 1;
    }
    if (p < en) next[nbr[p]] = 1;
  }
}

// Moved vertices of bins 0..4 (bins 5-6: k_expand_block, bin 7: k_expand_hub).
__global__ __launch_bounds__(256) void k_expand_moved(Bins bins, int64_t total,
                                                      const uint32_t* __restrict__ list,
                                                      const int64_t* __restrict__ off,
                                                      const uint32_t* __restrict__ nbr,
                                                      const uint8_t* __restrict__ moved,
                                                      ExpandDir dir, uint8_t* __restrict__ next,
                                                      const unsigned long long* __restrict__ dcnt) {
  if (dir.Pull()) return;
  if (dcnt) total = BinTotal(bins, dcnt);
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < total;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = base + threadIdx.x;
    bool sel = false;
    int64_t st = 0, deg = 0;
    if (t < total) {
      int64_t r = t;
      int b = 0;
      while (r >= BinCount(bins, dcnt, b)) r -= BinCount(bins, dcnt, b++);
      const uint32_t v = list[bins.base[b] + r];
      if (moved[v]) {
        sel = true;
        st = off[v];
        deg = off[v + 1] - st;
      }
    }
    MarkNeighbours(sel, st, deg, nbr, next);
  }
}

// Vertices with deg <= block_max whose cluster is modified.
__global__ __launch_bounds__(256) void k_expand_clusters(int64_t n, const int64_t* __restrict__ off,
                                                         const uint32_t* __restrict__ nbr,
                                                         const uint32_t* __restrict__ ids,
                                                         const uint8_t* __restrict__ modified,
                                                         int64_t block_max, ExpandDir dir,
                                                         uint8_t* __restrict__ next) {
  if (dir.Pull()) return;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = base + threadIdx.x;
    bool sel = false;
    int64_t st = 0, deg = 0;
    if (v < n) {
      st = off[v];
      deg = off[v + 1] - st;
      sel = deg <= block_max && modified[ids[v]] != 0;
    }
    MarkNeighbours(sel, st, deg, nbr, next);
  }
}

// Bin-7 vertices: one workgroup per partition (part_off prefix of the list)
// marks that partition's share of the adjacency.  The seed test is moved[v]
// (ids null) or modified[ids[v]].
__global__ __launch_bounds__(256) void k_expand_hub(const uint32_t* __restrict__ list, int64_t count,
                                                    const int64_t* __restrict__ part_off,
                                                    const uint32_t* __restrict__ item_of,
                                                    const int64_t* __restrict__ off,
                                                    const uint32_t* __restrict__ nbr,
                                                    const uint8_t* __restrict__ moved,
                                                    const uint32_t* __restrict__ ids,
                                                    const uint8_t* __restrict__ modified,
                                                    ExpandDir dir, uint8_t* __restrict__ next,
                                                    const unsigned long long* __restrict__ dcount) {
  if (dir.Pull()) return;
  if (dcount) count = (int64_t)*dcount;
  const int64_t t = XcdChunkItem(blockIdx.x);
  if (t >= part_off[count]) return;
  const int64_t i = item_of[t];
  const uint32_t v = list[i];
  if (ids ? modified[ids[v]] == 0 : moved[v] == 0) return;
  const int64_t parts = part_off[i + 1] - part_off[i];
  const int64_t part = t - part_off[i];
  const int64_t st = off[v], deg = off[v + 1] - st;
  const int64_t per = (deg + parts - 1) / parts;
  const int64_t e1 = min(deg, (part + 1) * per);
  for (int64_t p = part * per + threadIdx.x; p < e1; p += blockDim.x) next[nbr[st + p]] = 1;
}

// Dense (pull) expansion, used when the seeds' edge volume is large: v
// joins the next frontier iff some neighbour is a seed.  On a symmetric graph
// this is the same set the push kernels produce (u is a neighbour of v iff v
// is one of u), reached with an early exit per vertex instead of a write per
// seed edge.  The seed test is modified[ids[u]] (ids non-null) or moved[u].
__device__ __forceinline__ bool IsSeed(uint32_t u, const uint32_t* __restrict__ ids,
                                       const uint8_t* __restrict__ modified,
                                       const uint8_t* __restrict__ moved) {
  return ids ? modified[ids[u]] != 0 : moved[u] != 0;
}

// vol += sum over seeds of (deg + 1), the edgeMap density measure.
__global__ __launch_bounds__(256) void k_seed_volume(int64_t n, const int64_t* __restrict__ off,
                                                     const uint32_t* __restrict__ ids,
                                                     const uint8_t* __restrict__ modified,
                                                     const uint8_t* __restrict__ moved,
                                                     unsigned long long* __restrict__ vol,
                                                     const unsigned int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ unsigned long long sm[256 / kWave];
  unsigned long long x = 0;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    if (IsSeed((uint32_t)v, ids, modified, moved)) x += (unsigned long long)(off[v + 1] - off[v]) + 1ull;
  for (int sh = 32; sh > 0; sh >>= 1) x += __shfl_xor(x, sh);
  if (LaneId() == 0) sm[threadIdx.x / kWave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int i = 0; i < 256 / kWave; ++i) t += sm[i];
    if (t) atomicAdd(vol, t);
  }
}

// Vertices with deg <= block_max: degrees up to 16 by their own thread, larger
// ones by the whole wave, 64 neighbours per step.
__global__ __launch_bounds__(256) void k_pull_vertices(int64_t n, const int64_t* __restrict__ off,
                                                       const uint32_t* __restrict__ nbr,
                                                       const uint32_t* __restrict__ ids,
                                                       const uint8_t* __restrict__ modified,
                                                       const uint8_t* __restrict__ moved,
                                                       int64_t block_max, ExpandDir dir,
                                                       uint8_t* __restrict__ next) {
  if (!dir.Pull()) return;
  const int lane = LaneId();
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = base + threadIdx.x;
    int64_t st = 0, deg = 0;
    if (v < n) {
      st = off[v];
      deg = off[v + 1] - st;
      if (deg <= 16) {
        for (int64_t p = 0; p < deg; ++p)
          if (IsSeed(nbr[st + p], ids, modified, moved)) {
            next[v] = 1;
            break;
          }
      }
    }
    uint64_t big = __ballot(v < n && deg > 16 && deg <= block_max);
    const int64_t wave_v0 = base + (threadIdx.x & ~(kWave - 1));
    while (big) {
      const int l = __ffsll((unsigned long long)big) - 1;
      big &= big - 1;
      const int64_t vs = __shfl(st, l), vd = __shfl(deg, l);
      for (int64_t p0 = 0; p0 < vd; p0 += kWave) {
        const int64_t p = p0 + lane;
        const bool hit = p < vd && IsSeed(nbr[vs + p], ids, modified, moved);
        if (__ballot(hit)) {
          if (lane == 0) next[wave_v0 + l] = 1;
          break;
        }
      }
    }
  }
}

// Vertices with deg > block_max, one workgroup per partition of the hub list
// (part_off / item_of as in k_expand_hub); a partition stops at its first hit
// or when another partition has already marked the hub.
__global__ __launch_bounds__(256) void k_pull_hub(const uint32_t* __restrict__ hubs, int64_t count,
                                                  const int64_t* __restrict__ part_off,
                                                  const uint32_t* __restrict__ item_of,
                                                  const int64_t* __restrict__ off,
                                                  const uint32_t* __restrict__ nbr,
                                                  const uint32_t* __restrict__ ids,
                                                  const uint8_t* __restrict__ modified,
                                                  const uint8_t* __restrict__ moved,
                                                  ExpandDir dir, uint8_t* __restrict__ next) {
  if (!dir.Pull()) return;
  const int64_t t = XcdChunkItem(blockIdx.x);
  if (t >= part_off[count]) return;
  const int64_t i = item_of[t];
  const uint32_t v = hubs[i];
  const int64_t parts = part_off[i + 1] - part_off[i];
  const int64_t part = t - part_off[i];
  const int64_t st = off[v], deg = off[v + 1] - st;
  const int64_t per = (deg + parts - 1) / parts;
  const int64_t e0 = part * per, e1 = min(deg, e0 + per);
  for (int64_t p0 = e0; p0 < e1; p0 += blockDim.x) {
    const int64_t p = p0 + threadIdx.x;
    const bool hit = p < e1 && IsSeed(nbr[st + p], ids, modified, moved);
    const bool marked = threadIdx.x == 0 && *(volatile const uint8_t*)&next[v] != 0;
    // Uniform exit: every thread takes the block-wide result.
    if (__syncthreads_or(hit || marked)) {
      if (threadIdx.x == 0) next[v] = 1;
      return;
    }
  }
}

int Grid(int64_t items, int per_block, int max_blocks) {
  int64_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (int)b;
}

Bins MakeBins(const int64_t counts[kBins], const BinLayout& layout, int64_t* total) {
  Bins b;
  *total = 0;
  for (int i = 0; i < kBins; ++i) {
    b.base[i] = layout.base[i];
    b.count[i] = counts[i];
    *total += counts[i];
  }
  return b;
}

}  // namespace

void LaunchApply(const int64_t counts[kBins], const BinLayout& layout, const uint32_t* list,
                 const StateView& st, uint32_t new_id, const uint32_t* moves, uint8_t* modified,
                 uint8_t* new_flag, RoundCounters* rc, hipStream_t s, const DeviceCounts* dc) {
  int64_t total;
  const Bins b = MakeBins(counts, layout, &total);
  if (total == 0) return;
  hipLaunchKernelGGL(k_apply, dim3(Grid(total, 256, 4096)), dim3(256), 0, s, b, list, st, new_id,
                     moves, modified, new_flag, rc, dc ? dc->cnt : (const unsigned long long*)nullptr);
  PCC_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------- ordered weight updates --
// Node weights that are not integers (or whose sums reach 2^53) make FP64
// atomics order-dependent.  These passes update the cluster weights in the
// reference's order instead: movers are sorted stably by cluster (keys
// cluster << vb | vertex keep ascending vertex order within a cluster), and
// one thread folds each cluster's run left to right.

namespace {

int BitsFor(uint64_t x) {  // bits to hold every value in [0, x]
  int b = 1;
  while (b < 64 && (x >> b) != 0) ++b;
  return b;
}

__global__ __launch_bounds__(256) void k_mark_movers(Bins bins, const uint32_t* __restrict__ list,
                                                     const uint32_t* __restrict__ ids,
                                                     const uint32_t* __restrict__ moves,
                                                     uint8_t* __restrict__ flag) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int b = 0; b < kBins; ++b)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < bins.count[b];
         i += stride) {
      const uint32_t v = list[bins.base[b] + i];
      const uint32_t t = moves[v];
      flag[v] = (t != kNoneId && t != ids[v]) ? 1 : 0;
    }
}

// Movers in ascending vertex order, keyed by old and by new cluster.
__global__ __launch_bounds__(256) void k_emit_movers(int64_t n, const uint8_t* __restrict__ flag,
                                                     const uint32_t* __restrict__ rank,
                                                     StateView st, const uint32_t* __restrict__ moves,
                                                     int vb, uint64_t* __restrict__ kold,
                                                     double* __restrict__ wold,
                                                     uint64_t* __restrict__ knew,
                                                     double* __restrict__ wnew) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    if (!flag[v]) continue;
    const uint32_t i = rank[v];
    kold[i] = ((uint64_t)st.ids[v] << vb) | (uint64_t)v;
    knew[i] = ((uint64_t)moves[v] << vb) | (uint64_t)v;
    wold[i] = st.nw[v];
    wnew[i] = st.nw[v];
  }
}

// Every vertex keyed by its cluster (OrderedClusterWeights).
__global__ __launch_bounds__(256) void k_emit_members(int64_t n, StateView st, int vb,
                                                      uint64_t* __restrict__ keys,
                                                      double* __restrict__ w) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    keys[v] = ((uint64_t)st.ids[v] << vb) | (uint64_t)v;
    w[v] = st.nw[v];
    st.K[v] = 0.0;
  }
}

// mode 0: a run of old cluster c loses its members (sizes -= count,
// K[c] -= w one after another, PCCI.cc:283-292); mode 1: a run of new cluster
// c != new_id gains them (sizes += count, ids = c, K[c] += w, PCCI.cc:309-320),
// a run of new_id flags its members for the empty-id assignment; mode 2: K[c]
// = 0 + w + w + ... over all members (SetClustering, PCCI.cc:68-75).
__global__ __launch_bounds__(256) void k_fold_runs(int64_t cnt, const uint64_t* __restrict__ keys,
                                                   const double* __restrict__ w, int vb, int mode,
                                                   StateView st, uint32_t new_id,
                                                   uint8_t* __restrict__ modified,
                                                   uint8_t* __restrict__ new_flag,
                                                   RoundCounters* __restrict__ rc) {
  const uint64_t vmask = vb >= 64 ? ~0ull : ((1ull << vb) - 1ull);
  unsigned long long news = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t c = (uint32_t)(keys[i] >> vb);
    if (mode == 1 && c == new_id) {
      new_flag[keys[i] & vmask] = 1;
      ++news;
      continue;
    }
    if (i > 0 && (uint32_t)(keys[i - 1] >> vb) == c) continue;  // not the run's head
    int64_t j = i;
    if (mode == 0) {
      double k = st.K[c];
      for (; j < cnt && (uint32_t)(keys[j] >> vb) == c; ++j) k -= w[j];
      st.K[c] = k;
      st.sizes[c] -= (uint32_t)(j - i);
      modified[c] = 1;
    } else if (mode == 1) {
      double k = st.K[c];
      for (; j < cnt && (uint32_t)(keys[j] >> vb) == c; ++j) {
        st.ids[keys[j] & vmask] = c;
        k += w[j];
      }
      st.K[c] = k;
      st.sizes[c] += (uint32_t)(j - i);
      modified[c] = 1;
    } else {
      double k = 0.0;
      for (; j < cnt && (uint32_t)(keys[j] >> vb) == c; ++j) k += w[j];
      st.K[c] = k;
    }
  }
  if (mode == 1) {
    for (int s = 32; s > 0; s >>= 1) news += __shfl_xor(news, s);
    if (LaneId() == 0 && news) atomicAdd(&rc->new_count, news);
  }
}

}  // namespace

void LaunchApplyOrdered(DeviceAllocator* al, const int64_t counts[kBins], const BinLayout& layout,
                        const uint32_t* list, const StateView& st, uint32_t n, uint32_t new_id,
                        const uint32_t* moves, uint8_t* modified, uint8_t* new_flag,
                        RoundCounters* rc, hipStream_t s) {
  int64_t total;
  const Bins b = MakeBins(counts, layout, &total);
  if (total == 0 || n == 0) return;
  DevBuf<uint8_t> flag(al, n);
  DevBuf<uint32_t> rank(al, n), dcount(al, 1);
  PCC_HIP_CHECK(hipMemsetAsync(flag.get(), 0, n, s));
  hipLaunchKernelGGL(k_mark_movers, dim3(Grid(total, 256, 4096)), dim3(256), 0, s, b, list, st.ids,
                     moves, flag.get());
  PCC_HIP_CHECK(hipGetLastError());
  ExclusiveScanU8toU32(al, flag.get(), rank.get(), n, dcount.get(), s);
  uint32_t movers = 0;
  PCC_HIP_CHECK(hipMemcpyAsync(&movers, dcount.get(), 4, hipMemcpyDeviceToHost, s));
  PCC_HIP_CHECK(hipStreamSynchronize(s));
  if (movers == 0) return;
  const int vb = BitsFor((uint64_t)n - 1);
  const int cb = BitsFor((uint64_t)new_id > (uint64_t)n ? (uint64_t)new_id : (uint64_t)n);
  DevBuf<uint64_t> kold(al, movers), knew(al, movers);
  DevBuf<double> wold(al, movers), wnew(al, movers);
  hipLaunchKernelGGL(k_emit_movers, dim3(Grid(n, 256, 4096)), dim3(256), 0, s, (int64_t)n,
                     flag.get(), rank.get(), st, moves, vb, kold.get(), wold.get(), knew.get(),
                     wnew.get());
  PCC_HIP_CHECK(hipGetLastError());
  RadixSortPairsF64(al, kold.get(), wold.get(), movers, vb, vb + cb, s);
  RadixSortPairsF64(al, knew.get(), wnew.get(), movers, vb, vb + cb, s);
  const int grid = Grid(movers, 256, 4096);
  // Every loss is applied before any gain, as the reference's two passes.
  hipLaunchKernelGGL(k_fold_runs, dim3(grid), dim3(256), 0, s, (int64_t)movers, kold.get(),
                     wold.get(), vb, 0, st, new_id, modified, new_flag, rc);
  hipLaunchKernelGGL(k_fold_runs, dim3(grid), dim3(256), 0, s, (int64_t)movers, knew.get(),
                     wnew.get(), vb, 1, st, new_id, modified, new_flag, rc);
  PCC_HIP_CHECK(hipGetLastError());
}

void OrderedClusterWeights(DeviceAllocator* al, int64_t n, const StateView& st, hipStream_t s) {
  if (n <= 0) return;
  const int vb = BitsFor((uint64_t)n - 1);
  const int cb = BitsFor((uint64_t)n);
  DevBuf<uint64_t> keys(al, n);
  DevBuf<double> w(al, n);
  hipLaunchKernelGGL(k_emit_members, dim3(Grid(n, 256, 4096)), dim3(256), 0, s, n, st, vb,
                     keys.get(), w.get());
  PCC_HIP_CHECK(hipGetLastError());
  RadixSortPairsF64(al, keys.get(), w.get(), n, vb, vb + cb, s);
  hipLaunchKernelGGL(k_fold_runs, dim3(Grid(n, 256, 4096)), dim3(256), 0, s, n, keys.get(),
                     w.get(), vb, 2, st, 0u, (uint8_t*)nullptr, (uint8_t*)nullptr,
                     (RoundCounters*)nullptr);
  PCC_HIP_CHECK(hipGetLastError());
}

__global__ __launch_bounds__(256) void k_graph_flags(int64_t n, const uint8_t* __restrict__ next,
                                                     const uint32_t* __restrict__ vgraph,
                                                     uint8_t* __restrict__ flags) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    if (next[v]) flags[vgraph ? vgraph[v] : 0] = 1;
}

void LaunchRoundReset(int64_t n, uint8_t* a, uint8_t* b, uint8_t* c, RoundCounters* rc,
                      int counters, unsigned long long* acc, hipStream_t s,
                      const unsigned int* c_gate) {
  const int64_t work = (a || b || c) ? (n + 15) / 16 : 1;
  hipLaunchKernelGGL(k_round_reset, dim3(Grid(work, 256, 2048)), dim3(256), 0, s, n, a, b, c, rc,
                     counters, acc, c_gate);
  PCC_HIP_CHECK(hipGetLastError());
}

// One wave: the round's list counts move from rc (where the binning
// accumulated them) to dc, and the round's statistics are added up.
__global__ __launch_bounds__(64) void k_round_begin(RoundCounters* __restrict__ rc,
                                                    DeviceCounts* __restrict__ dc,
                                                    DeviceRoundStats* __restrict__ st, int keep,
                                                    int nbhr_move, double pull_thr) {
  const int b = threadIdx.x;
  unsigned long long c = 0, t = 0;
  if (b < kBins) {
    c = rc->cnt[b];
    t = rc->te[b];
    dc->cnt[b] = c;
    dc->te[b] = t;
    if (!keep) {
      rc->cnt[b] = 0;
      rc->te[b] = 0;
    }
    const int grp = StatGroupOf(b);
    if (c) atomicAdd(&st->bin_verts[grp], c);
    if (t) atomicAdd(&st->bin_edges[grp], t);
  }
  unsigned long long F = c, TE = t;
  for (int sh = 32; sh > 0; sh >>= 1) {
    F += __shfl_xor(F, sh);
    TE += __shfl_xor(TE, sh);
  }
  if (b == 0) {
    st->edges += TE;
    st->verts += F;
    if (st->first || F > 0) st->rounds += 1;
    st->first = 0;
    dc->pull_possible = (!nbhr_move || (double)(TE + F) > pull_thr) ? 1u : 0u;
  }
}

__global__ void k_round_stats_init(DeviceRoundStats* __restrict__ st) {
  if (threadIdx.x == 0) {
    st->rounds = st->edges = st->verts = 0;
    for (int k = 0; k < kStatGroups; ++k) st->bin_edges[k] = st->bin_verts[k] = 0;
    st->first = 1;
  }
}

void LaunchRoundStatsInit(DeviceRoundStats* stats, hipStream_t s) {
  hipLaunchKernelGGL(k_round_stats_init, dim3(1), dim3(64), 0, s, stats);
  PCC_HIP_CHECK(hipGetLastError());
}

void LaunchRoundBegin(RoundCounters* rc, DeviceCounts* dc, DeviceRoundStats* stats, bool keep,
                      bool nbhr_move, double pull_thr, hipStream_t s) {
  hipLaunchKernelGGL(k_round_begin, dim3(1), dim3(64), 0, s, rc, dc, stats, keep ? 1 : 0,
                     nbhr_move ? 1 : 0, pull_thr);
  PCC_HIP_CHECK(hipGetLastError());
}

void LaunchGraphFlags(int64_t n, const uint8_t* next, const uint32_t* vgraph, uint8_t* flags,
                      hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_graph_flags, dim3(Grid(n, 256, 8192)), dim3(256), 0, s, n, next, vgraph,
                     flags);
  PCC_HIP_CHECK(hipGetLastError());
}

__global__ __launch_bounds__(256) void k_fill_u32_gated(uint32_t* __restrict__ p, int64_t n,
                                                        uint32_t x,
                                                        const unsigned long long* __restrict__ gate) {
  if (*gate == 0) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = x;
}

// acc += the fallback and repair counts of a workspace's rounds (one thread).
__global__ void k_add_fallbacks(const RoundCounters* __restrict__ rc,
                                unsigned long long* __restrict__ acc) {
  *acc += rc->fb_total + rc->fb_count + rc->repair_count;
}

void LaunchAddFallbacks(const RoundCounters* rc, unsigned long long* acc, hipStream_t s) {
  hipLaunchKernelGGL(k_add_fallbacks, dim3(1), dim3(1), 0, s, rc, acc);
  PCC_HIP_CHECK(hipGetLastError());
}

// Word stores over each range's aligned middle, byte stores at its ends.
__global__ __launch_bounds__(256) void k_zero_spans(ZeroSpans z) {
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int k = 0; k < z.n; ++k) {
    uint8_t* p = z.p[k];
    const int64_t nb = z.bytes[k];
    const int64_t head = std::min<int64_t>(nb, (int64_t)((4 - ((uintptr_t)p & 3)) & 3));
    const int64_t words = (nb - head) / 4;
    uint32_t* w = reinterpret_cast<uint32_t*>(p + head);
    for (int64_t i = t0; i < words; i += step) w[i] = 0u;
    const int64_t tail0 = head + 4 * words;
    for (int64_t i = t0; i < head; i += step) p[i] = 0;
    for (int64_t i = tail0 + t0; i < nb; i += step) p[i] = 0;
  }
}

void LaunchZeroSpans(const ZeroSpans& z, hipStream_t s) {
  int64_t most = 0;
  for (int k = 0; k < z.n; ++k) most = std::max(most, z.bytes[k]);
  if (most <= 0) return;
  const int64_t blocks = std::min<int64_t>(4096, (most / 4 + 255) / 256 + 1);
  hipLaunchKernelGGL(k_zero_spans, dim3((unsigned)blocks), dim3(256), 0, s, z);
  PCC_HIP_CHECK(hipGetLastError());
}

NewAssignWork::NewAssignWork(DeviceAllocator* al, int64_t n) {
  const int64_t nn = n > 0 ? n : 1;
  zrank.Reset(al, nn);
  zero_list.Reset(al, nn);
  nrank.Reset(al, nn);
  if (n > (int64_t)kSmallAssignMax) {
    z.Reset(al, nn);
    scan.Reset(al, std::max<int64_t>(1, ScanWorkElems(n)));
  }
}

void RunNewClusterAssign(DeviceAllocator* al, const StateView& st, uint32_t n,
                         const uint32_t* vgraph, const int64_t* gbase, uint8_t* new_flag,
                         uint8_t* modified, RoundCounters* rc, hipStream_t s,
                         const NewAssignWork* work) {
  const unsigned long long* gate = &rc->new_count;
  if (n == 0) return;
  std::unique_ptr<NewAssignWork> own;
  if (!work) {
    own.reset(new NewAssignWork(al, n));
    work = own.get();
  }
  if (n <= kSmallAssignMax) {
    hipLaunchKernelGGL(k_new_assign_small, dim3(1), dim3(kSmallAssignThreads), 0, s, n, st,
                       vgraph, gbase, new_flag, work->zero_list.get(), work->zrank.get(),
                       work->nrank.get(), modified, gate);
    PCC_HIP_CHECK(hipGetLastError());
    return;
  }
  const int grid = Grid(n, 256, 4096);
  hipLaunchKernelGGL(k_fill_u32_gated, dim3(grid), dim3(256), 0, s, work->zero_list.get(),
                     (int64_t)n, 0xFFFFFFFFu, gate);
  hipLaunchKernelGGL(k_zero_flags, dim3(grid), dim3(256), 0, s, n, st.sizes, work->z.get(), gate);
  ExclusiveScanU8toU32(al, work->z.get(), work->zrank.get(), n, nullptr, s, gate,
                       work->scan.get());
  hipLaunchKernelGGL(k_zero_compact, dim3(grid), dim3(256), 0, s, n, work->z.get(),
                     work->zrank.get(), work->zero_list.get(), gate);
  ExclusiveScanU8toU32(al, new_flag, work->nrank.get(), n, nullptr, s, gate, work->scan.get());
  hipLaunchKernelGGL(k_new_assign, dim3(grid), dim3(256), 0, s, n, st, vgraph, gbase, new_flag,
                     work->nrank.get(), work->zrank.get(), work->zero_list.get(), modified, gate);
  PCC_HIP_CHECK(hipGetLastError());
}

// Ranges of each graph's ids in the ascending empty-id list: free_cur[g] /
// free_end[g] = first index with id >= gbase[g] / gbase[g + 1].
__global__ void k_free_ranges(int G, const int64_t* __restrict__ gbase, uint32_t n,
                              const uint8_t* __restrict__ z, const uint32_t* __restrict__ zrank,
                              const uint32_t* __restrict__ zero_list,
                              unsigned long long* __restrict__ free_cur,
                              unsigned long long* __restrict__ free_end) {
  const uint32_t total = n ? zrank[n - 1] + z[n - 1] : 0;
  for (int gi = blockIdx.x * blockDim.x + threadIdx.x; gi <= G; gi += gridDim.x * blockDim.x) {
    const int64_t key = gi == G ? (int64_t)n : (gbase ? gbase[gi] : (gi == 0 ? 0 : (int64_t)n));
    uint32_t lo = 0, hi = total;  // first index with zero_list[idx] >= key
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if ((int64_t)zero_list[mid] < key)
        lo = mid + 1;
      else
        hi = mid;
    }
    if (gi < G) free_cur[gi] = lo;
    if (gi > 0) free_end[gi - 1] = lo;
  }
}

void PrepareFreeIds(DeviceAllocator* al, const StateView& st, uint32_t n, int G,
                    const int64_t* gbase, FreeIds* f, hipStream_t s) {
  f->ids.Reset(al, n > 0 ? n : 1);
  f->cur.Reset(al, G);
  f->end.Reset(al, G);
  if (n == 0) {
    PCC_HIP_CHECK(hipMemsetAsync(f->cur.get(), 0, sizeof(unsigned long long) * (size_t)G, s));
    PCC_HIP_CHECK(hipMemsetAsync(f->end.get(), 0, sizeof(unsigned long long) * (size_t)G, s));
    return;
  }
  DevBuf<uint8_t> z(al, n);
  DevBuf<uint32_t> zrank(al, n);
  DevBuf<unsigned long long> one(al, 1);
  const unsigned long long kOne = 1;
  PCC_HIP_CHECK(hipMemcpyAsync(one.get(), &kOne, sizeof(kOne), hipMemcpyHostToDevice, s));
  const int grid = Grid(n, 256, 4096);
  hipLaunchKernelGGL(k_zero_flags, dim3(grid), dim3(256), 0, s, n, st.sizes, z.get(), one.get());
  ExclusiveScanU8toU32(al, z.get(), zrank.get(), n, nullptr, s);
  hipLaunchKernelGGL(k_zero_compact, dim3(grid), dim3(256), 0, s, n, z.get(), zrank.get(),
                     f->ids.get(), one.get());
  hipLaunchKernelGGL(k_free_ranges, dim3((G + 256) / 256), dim3(256), 0, s, G, gbase, n, z.get(),
                     zrank.get(), f->ids.get(), f->cur.get(), f->end.get());
  PCC_HIP_CHECK(hipGetLastError());
  // z, zrank and one go back to the caching allocator; the stream orders
  // their reuse after these kernels.
  PCC_HIP_CHECK(hipStreamSynchronize(s));
}

void LaunchExpandMoved(const int64_t counts[kBins], const int64_t te[kBins],
                       const BinLayout& layout, const uint32_t* list,
                       const int64_t* hub_part_off, const uint32_t* hub_item_of,
                       int64_t part_edges, const CsrView& g, const uint8_t* moved,
                       const ExpandDir& dir, uint8_t* next, hipStream_t s, const DeviceCounts* dc) {
  const unsigned long long* dcnt = dc ? dc->cnt : nullptr;
  int64_t total;
  const Bins all = MakeBins(counts, layout, &total);
  Bins lo = all, mid = all;
  int64_t nlo = 0, nmid = 0;
  for (int b = 0; b < kBins; ++b) {
    const bool is_lo = b < kWaveBin0, is_mid = b >= kWaveBin0 && b < kHubBin;
    lo.count[b] = is_lo ? all.count[b] : 0;
    mid.count[b] = is_mid ? all.count[b] : 0;
    nlo += lo.count[b];
    nmid += mid.count[b];
  }
  if (nlo > 0) {
    hipLaunchKernelGGL(k_expand_moved, dim3(Grid(nlo, 256, dc ? 4096 : 8192)), dim3(256), 0, s, lo,
                       nlo, list, g.off, g.nbr, moved, dir, next, dcnt);
    PCC_HIP_CHECK(hipGetLastError());
  }
  if (nmid > 0) {
    if (dcnt)
      hipLaunchKernelGGL(k_expand_block<true>, dim3(Grid(nmid, 4, 4096)), dim3(256), 0, s, mid,
                         nmid, list, g.off, g.nbr, moved, dir, next, dcnt);
    else
      hipLaunchKernelGGL(k_expand_block<false>, dim3(Grid(nmid, 4, 4096)), dim3(256), 0, s, mid,
                         nmid, list, g.off, g.nbr, moved, dir, next, dcnt);
    PCC_HIP_CHECK(hipGetLastError());
  }
  const int64_t c7 = counts[kBins - 1];
  if (c7 > 0) {
    int64_t items = c7 + te[kBins - 1] / part_edges + 1;
    items = (items + kXcdGrid - 1) / kXcdGrid * kXcdGrid;
    hipLaunchKernelGGL(k_expand_hub, dim3((unsigned)items), dim3(256), 0, s,
                       list + layout.base[kBins - 1], c7, hub_part_off, hub_item_of, g.off, g.nbr,
                       moved, (const uint32_t*)nullptr, (const uint8_t*)nullptr, dir, next,
                       dc ? &dc->cnt[kBins - 1] : (const unsigned long long*)nullptr);
    PCC_HIP_CHECK(hipGetLastError());
  }
}

void LaunchExpandClusters(const CsrView& g, const uint32_t* ids, const uint8_t* modified,
                          int64_t block_max, const HubList& hubs, int64_t part_edges,
                          const ExpandDir& dir, uint8_t* next, hipStream_t s) {
  if (g.n == 0) return;
  hipLaunchKernelGGL(k_expand_clusters, dim3(Grid(g.n, 256, 4096)), dim3(256), 0, s, g.n, g.off,
                     g.nbr, ids, modified, block_max, dir, next);
  PCC_HIP_CHECK(hipGetLastError());
  if (hubs.count > 0) {
    int64_t items = hubs.count + hubs.te / part_edges + 1;
    items = (items + kXcdGrid - 1) / kXcdGrid * kXcdGrid;
    hipLaunchKernelGGL(k_expand_hub, dim3((unsigned)items), dim3(256), 0, s, hubs.list,
                       hubs.count, hubs.part_off, hubs.item_of, g.off, g.nbr,
                       (const uint8_t*)nullptr, ids, modified, dir, next,
                       (const unsigned long long*)nullptr);
    PCC_HIP_CHECK(hipGetLastError());
  }
}

void LaunchSeedVolume(const CsrView& g, const uint32_t* ids, const uint8_t* modified,
                      const uint8_t* moved, unsigned long long* vol, hipStream_t s,
                      const unsigned int* gate) {
  if (g.n == 0) return;
  // 512 workgroups: one atomic each on the single accumulator.
  hipLaunchKernelGGL(k_seed_volume, dim3(Grid(g.n, 256, 512)), dim3(256), 0, s, g.n, g.off, ids,
                     modified, moved, vol, gate);
  PCC_HIP_CHECK(hipGetLastError());
}

void LaunchPullFrontier(const CsrView& g, const uint32_t* ids, const uint8_t* modified,
                        const uint8_t* moved, int64_t block_max, const HubList& hubs,
                        int64_t part_edges, const ExpandDir& dir, uint8_t* next, hipStream_t s) {
  if (g.n == 0) return;
  hipLaunchKernelGGL(k_pull_vertices, dim3(Grid(g.n, 256, 4096)), dim3(256), 0, s, g.n, g.off,
                     g.nbr, ids, modified, moved, block_max, dir, next);
  PCC_HIP_CHECK(hipGetLastError());
  if (hubs.count > 0) {
    int64_t items = hubs.count + hubs.te / part_edges + 1;
    items = (items + kXcdGrid - 1) / kXcdGrid * kXcdGrid;
    hipLaunchKernelGGL(k_pull_hub, dim3((unsigned)items), dim3(256), 0, s, hubs.list, hubs.count,
                       hubs.part_off, hubs.item_of, g.off, g.nbr, ids, modified, moved, dir, next);
    PCC_HIP_CHECK(hipGetLastError());
  }
}

}  // namespace pcc
