// k_group.hip -- cluster-grouped contraction for gfx950.
//
// Builds the same contracted graph as the all-edge sort (CompressGraphSorted
// in k_contract.hip, which restates CompressGraph,
// clustering/clusterers/parallel-correlation-clusterer-internal.cc:709-782,
// with ComputeInterClusterEdgesSort, parallel/parallel-graph-utils.cc:70-231),
// bit for bit, one cluster at a time:
//   1. vertices are stably sorted by their contracted vertex (new id: the
//      cluster id plus the graph's shift in a batch), so each cluster's members
//      form a run in ascending vertex order, and their rows laid end to end
//      are the cluster's terms in CSR order.  This "padded" layout gives
//      cluster c the positions [rowpos[vstart[c]], rowpos[vstart[c + 1])) and
//      its capacity (the members' degree sum);
//   2. the kept terms of a cluster (v <= u or c(v) != c(u), PGU.cc:93-103) are
//      ordered by (c(v), position) -- the order of the reference's stable sort
//      by (c(u), c(v)) restricted to one c(u) -- and every run of equal c(v) is
//      summed in float in pbbslib's reduce order (left fold up to 1024 terms,
//      blocks of max(1024, 4 ceil(sqrt len)) otherwise), which yields the
//      cluster's contracted row, written at its padded position.  Clusters
//      are classed by capacity:
//        <= 64:   G lanes of a wave per cluster (G = 4 ... 64), a bitonic
//                 network over shuffles;
//        <= 512:  one wave per cluster, 64 E terms in registers (E = 2, 4, 8),
//                 a bitonic network over shuffles and register swaps;
//        <= 8192: W = 2 ... 16 waves per cluster, 8 terms per lane, the
//                 same network with its longest stages through LDS;
//        larger:  those clusters' terms only go through the LSD radix sort of
//                 (rank, c(v)) keys and the segmented sums;
//   3. the rows are copied, class by class, behind the exclusive scan of
//      their lengths.
// Node weights are summed per cluster in ascending vertex order over the same
// vertex sort.  Compared with sorting every kept edge by a 44-bit key (six
// global passes of 12-byte pairs at scale 22), most terms are read once from
// the CSR, sorted on chip and written once as contracted edges.

#include <algorithm>
#include <cmath>
#include <vector>

#include "device_common.h"
#include "kernels.h"

namespace pcc {

namespace {

int GvGrid(int64_t items, int per_block, int max_blocks) {
  int64_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (int)b;
}

int GvBitsFor(uint64_t maxval) {
  int b = 0;
  while (b < 63 && (maxval >> b) != 0) ++b;
  return b < 1 ? 1 : b;
}

constexpr int kGvThreads = 256;
constexpr int kGvWaves = kGvThreads / kWave;
// Capacity classes: G = 4, 8, 16, 32, 64 lanes; one wave with E = 2, 4, 8
// terms per lane; 2, 4, 8, 16 waves with 8 terms per lane; the radix path.
constexpr int kGvClasses = 13;
constexpr int64_t kGvSortedMax = 8192;  // larger clusters take the radix path
constexpr int kGvBig = kGvClasses - 1;

__device__ __forceinline__ int GvClassOf(int64_t cap) {
  if (cap <= 4) return 0;
  if (cap <= 8) return 1;
  if (cap <= 16) return 2;
  if (cap <= 32) return 3;
  if (cap <= 64) return 4;
  if (cap <= 128) return 5;
  if (cap <= 256) return 6;
  if (cap <= 512) return 7;
  if (cap <= 1024) return 8;
  if (cap <= 2048) return 9;
  if (cap <= 4096) return 10;
  if (cap <= kGvSortedMax) return 11;
  return kGvBig;
}

// Where the members of a contracted vertex and their rows are.  Member i
// (position in the vertex sort) is vertex perm[i], whose row starts at
// moff[i] in the CSR and at rowpos[i] in the padded layout; mcu[i] is its
// cluster id.
struct GroupIndex {
  const uint32_t* perm;
  const uint32_t* vstart;  // n2 + 1 entries: first member of each new vertex
  const int64_t* rowpos;   // n + 1 entries
  const int64_t* moff;
  const uint32_t* mcu;
};

// A cluster to contract: new id, member range [vs, ve), capacity; its padded
// start rowpos[vs] is kept in a parallel array.
struct GvItems {
  const uint4* item;  // (c, vs, ve, cap)
  const int64_t* rp0;
};

// Batch renumbering (cluster c of graph p -> newbase[p] + c - gbase[p]).
struct GvMap {
  const uint32_t* vgraph;
  const int64_t* gbase;
  const int64_t* newbase;
  __device__ __forceinline__ int64_t Shift(int64_t v) const {
    if (!vgraph) return 0;
    const uint32_t p = vgraph[v];
    return newbase[p] < 0 ? INT64_MIN : newbase[p] - gbase[p];
  }
};

// Sort keys of the vertices: new id, or n2 for vertices of graphs that leave
// the batch, which sort last.  Also clears the n2 member counts and new node
// weights.
__global__ __launch_bounds__(256) void k_gv_keys(int64_t n, const uint32_t* __restrict__ ids,
                                                 GvMap gm, int64_t n2, uint64_t* __restrict__ vkey,
                                                 uint32_t* __restrict__ vval,
                                                 uint32_t* __restrict__ cnt,
                                                 double* __restrict__ nw2) {
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = t0; v < n; v += step) {
    const int64_t shift = gm.Shift(v);
    vkey[v] = shift == INT64_MIN ? (uint64_t)n2 : (uint64_t)((int64_t)ids[v] + shift);
    vval[v] = (uint32_t)v;
  }
  for (int64_t c = t0; c < n2; c += step) {
    cnt[c] = 0u;
    nw2[c] = 0.0;
  }
}

// Member counts per new id from the runs of the sorted keys (cnt zeroed).
__global__ __launch_bounds__(256) void k_gv_counts(const uint64_t* __restrict__ vkey,
                                                   const int64_t* __restrict__ start, int64_t nseg,
                                                   int64_t n2, uint32_t* __restrict__ cnt) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nseg;
       s += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = vkey[start[s]];
    if ((int64_t)k < n2) cnt[k] = (uint32_t)(start[s + 1] - start[s]);
  }
}

// Per member: row length (0 for dropped vertices), CSR row start, cluster id
// and node weight, all in member order.
__global__ __launch_bounds__(256) void k_gv_members(int64_t n, const int64_t* __restrict__ off,
                                                    const uint32_t* __restrict__ ids,
                                                    const uint64_t* __restrict__ vkey,
                                                    const uint32_t* __restrict__ perm, int64_t n2,
                                                    const double* __restrict__ nw,
                                                    uint32_t* __restrict__ deg,
                                                    int64_t* __restrict__ moff,
                                                    uint32_t* __restrict__ mcu,
                                                    double* __restrict__ nws) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t u = perm[i];
    const int64_t o = off[u];
    deg[i] = (int64_t)vkey[i] < n2 ? (uint32_t)(off[u + 1] - o) : 0u;
    moff[i] = o;
    mcu[i] = ids[u];
    nws[i] = nw[u];
  }
}

// Class counts per workgroup of kGvThreads consecutive new vertices, class
// major: bcnt[k * nblk + b].  Empty new vertices get row length 0 here.
__global__ __launch_bounds__(kGvThreads) void k_gv_class_count(int64_t n2, GroupIndex gi,
                                                               int64_t nblk,
                                                               uint32_t* __restrict__ bcnt,
                                                               uint32_t* __restrict__ len) {
  __shared__ uint32_t wc[kGvWaves][kGvClasses];
  const int lane = LaneId(), w = threadIdx.x / kWave;
  for (int64_t b = blockIdx.x; b < nblk; b += gridDim.x) {
    const int64_t c = b * kGvThreads + threadIdx.x;
    int cls = -1;
    if (c < n2) {
      const int64_t cap = gi.rowpos[gi.vstart[c + 1]] - gi.rowpos[gi.vstart[c]];
      if (cap == 0)
        len[c] = 0;
      else
        cls = GvClassOf(cap);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kGvClasses; ++k) {
      const uint64_t mask = __ballot(cls == k);
      if (lane == 0) wc[w][k] = (uint32_t)__popcll(mask);
    }
    __syncthreads();
    if (threadIdx.x < kGvClasses) {
      uint32_t s = 0;
      for (int ww = 0; ww < kGvWaves; ++ww) s += wc[ww][threadIdx.x];
      bcnt[(int64_t)threadIdx.x * nblk + b] = s;
    }
  }
}

// Places every nonempty new vertex in its class's list, in ascending id order
// within the class (base: exclusive scan of bcnt).
__global__ __launch_bounds__(kGvThreads) void k_gv_class_place(int64_t n2, GroupIndex gi,
                                                               int64_t nblk,
                                                               const uint32_t* __restrict__ base,
                                                               uint4* __restrict__ item,
                                                               int64_t* __restrict__ rp0) {
  __shared__ uint32_t wc[kGvWaves][kGvClasses];
  const int lane = LaneId(), w = threadIdx.x / kWave;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (int64_t b = blockIdx.x; b < nblk; b += gridDim.x) {
    const int64_t c = b * kGvThreads + threadIdx.x;
    int cls = -1;
    uint32_t vs = 0, ve = 0;
    int64_t r0 = 0, cap = 0;
    if (c < n2) {
      vs = gi.vstart[c];
      ve = gi.vstart[c + 1];
      r0 = gi.rowpos[vs];
      cap = gi.rowpos[ve] - r0;
      if (cap > 0) cls = GvClassOf(cap);
    }
    uint32_t rank = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kGvClasses; ++k) {
      const uint64_t mask = __ballot(cls == k);
      if (cls == k) rank = (uint32_t)__popcll(mask & lt);
      if (lane == 0) wc[w][k] = (uint32_t)__popcll(mask);
    }
    __syncthreads();
    if (cls >= 0) {
      uint32_t pos = base[(int64_t)cls * nblk + b] + rank;
      for (int ww = 0; ww < w; ++ww) pos += wc[ww][cls];
      item[pos] = make_uint4((uint32_t)c, vs, ve, (uint32_t)cap);
      rp0[pos] = r0;
    }
  }
}

// Class boundaries in the item list: bounds[k] = first item of class k,
// bounds[kGvClasses] = number of items.
__global__ void k_gv_class_bounds(const uint32_t* __restrict__ base, int64_t nblk,
                                  const uint32_t* __restrict__ total,
                                  unsigned long long* __restrict__ bounds) {
  const int k = threadIdx.x;
  if (k < kGvClasses) bounds[k] = base[(int64_t)k * nblk];
  if (k == kGvClasses) bounds[k] = *total;
}

// Capacities <= G: G lanes per cluster, 64 / G clusters per wave.  Lane j
// takes the cluster's j-th term in CSR order (the row holding it is found
// through a per-wave LDS map of row starts), the group sorts its keys
// (c(v) << 6 | j; dropped terms ~0) with a bitonic network over shuffles, and
// the first lane of each run of equal c(v) folds the run left in term order.
template <int G>
__global__ __launch_bounds__(kGvThreads) void k_gv_small(CsrView g,
                                                         const uint32_t* __restrict__ ids,
                                                         GroupIndex gi, GvItems L, int64_t count,
                                                         uint32_t* __restrict__ nbr2p,
                                                         float* __restrict__ w2p,
                                                         uint32_t* __restrict__ len) {
  constexpr int PER = kWave / G;
  constexpr uint64_t kLow = G == 64 ? ~0ull : ((1ull << G) - 1ull);
  __shared__ uint8_t own[kGvThreads];
  const int lane = LaneId();
  const int gl = lane & (G - 1);
  const int grp = lane / G;
  const int gfirst = threadIdx.x - gl;  // LDS slot of the group's lane 0
  const int64_t wave = ((int64_t)blockIdx.x * kGvThreads + threadIdx.x) / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * kGvWaves;
  for (int64_t base = wave * PER; base < count; base += nwaves * PER) {  // uniform per wave
    const int64_t it = base + grp;
    const bool live = it < count;
    uint32_t c = 0, vs = 0, ve = 0;
    int64_t rp0 = 0;
    if (live) {
      const uint4 itm = L.item[it];
      c = itm.x;
      vs = itm.y;
      ve = itm.z;
      rp0 = L.rp0[it];
    }
    // The source (u, p) of term gl, member chunk by member chunk (one chunk
    // unless the cluster has more than G members, some of them isolated).
    uint32_t ue = 0, cue = 0;
    int64_t pe = -1;
    for (uint32_t ms = 0; __any(live && vs + ms < ve); ms += G) {
      const uint32_t i = vs + ms + gl;
      const bool mem = live && i < ve;
      uint32_t mu = 0, mcu = 0, mr = 0, md = 0;
      int64_t mo = 0;
      if (mem) {
        mu = gi.perm[i];
        mcu = gi.mcu[i];
        mo = gi.moff[i];
        const int64_t r0 = gi.rowpos[i];
        md = (uint32_t)(gi.rowpos[i + 1] - r0);
        mr = (uint32_t)(r0 - rp0);
      }
      own[threadIdx.x] = 0xFF;
      __builtin_amdgcn_wave_barrier();
      if (mem && md > 0) own[gfirst + (int)mr] = (uint8_t)gl;  // mr < G for a nonempty row
      __builtin_amdgcn_wave_barrier();
      const uint8_t o = own[threadIdx.x];
      const uint64_t starts = (__ballot(o != 0xFF) >> (grp * G)) & kLow;
      const uint64_t below = starts & (~0ull >> (63 - gl));
      const int s0 = below ? 63 - __clzll(below) : gl;
      const int l = below ? (int)own[gfirst + s0] : gl;
      const int src = grp * G + l;
      const uint32_t su = __shfl(mu, src);
      const uint32_t scu = __shfl(mcu, src);
      const int64_t so = __shfl(mo, src);
      const uint32_t sd = __shfl(md, src);
      __builtin_amdgcn_wave_barrier();
      if (live && below && (uint32_t)gl < (uint32_t)s0 + sd) {
        ue = su;
        cue = scu;
        pe = so + (gl - s0);
      }
    }
    uint64_t key = ~0ull;
    float x = 0.0f;
    if (pe >= 0) {
      const uint32_t v = g.nbr[pe];
      x = g.w[pe];
      const uint32_t cvo = ids[v];
      if (v <= ue || cvo != cue) key = ((uint64_t)(cvo + (c - cue)) << 6) | (uint64_t)gl;
    }
    // Bitonic sort of the group's keys (ascending); partners lie in the group.
#pragma unroll
    for (int k = 2; k <= G; k <<= 1) {
#pragma unroll
      for (int d = k >> 1; d > 0; d >>= 1) {
        const uint64_t ok = __shfl_xor(key, d);
        const float ox = __shfl_xor(x, d);
        const bool up = (gl & k) == 0;
        const bool lower = (gl & d) == 0;
        const bool take = (lower == up) ? (ok < key) : (ok > key);
        if (take) {
          key = ok;
          x = ox;
        }
      }
    }
    const bool valid = key != ~0ull;
    const uint32_t cv = (uint32_t)(key >> 6);
    const uint64_t pk = __shfl_up(key, 1);
    const bool head = valid && (gl == 0 || (uint32_t)(pk >> 6) != cv);
    const uint64_t hm = (__ballot(head) >> (grp * G)) & kLow;
    const uint64_t vm = (__ballot(valid) >> (grp * G)) & kLow;
    const int nvalid = __popcll(vm);
    const uint64_t later = hm & ~(~0ull >> (63 - gl));  // heads after gl
    const int end = later ? __ffsll((unsigned long long)later) - 1 : nvalid;
    const int runlen = head ? end - gl : 0;
    const int maxrun = WaveMaxInt(runlen);
    float acc = x;
    for (int t = 1; t < maxrun; ++t) {
      const float y = __shfl(x, lane + t < kWave ? lane + t : lane);
      if (t < runlen) acc += y;
    }
    if (live && head) {
      const int64_t dst = rp0 + __popcll(hm & ~(~0ull << gl));
      nbr2p[dst] = cv;
      w2p[dst] = acc;
    }
    if (live && gl == 0) len[c] = (uint32_t)__popcll(hm);
  }
}

// Left fold of x[0, L) (L >= 1); the loads go out ahead of the adds, which
// stay in order.
__device__ __forceinline__ float GvFold(const float* x, uint32_t L) {
  float acc = x[0];
#pragma unroll 8
  for (uint32_t j = 1; j < L; ++j) acc += x[j];
  return acc;
}

// pbbslib's reduce order over x[0, L) (L <= kGvSortedMax: at most a handful of
// blocks of 1024, whose sums are folded left).
__device__ __forceinline__ float GvPbbsSum(const float* x, uint32_t L) {
  if (L <= 1024) return GvFold(x, L);
  uint32_t bs = 4u * (uint32_t)ceil(sqrt((double)L));
  if (bs < 1024) bs = 1024;
  float acc = GvFold(x, bs < L ? bs : L);
  for (uint32_t b = bs; b < L; b += bs) acc += GvFold(x + b, (L - b) < bs ? (L - b) : bs);
  return acc;
}

// Bitonic sorting network (ascending) over N = 64 E W keys: element
// i = 64 E w + 64 e + lane is key[e] of lane `lane` in wave w.  Stage (K, D)
// compares i with i ^ D; the lower index keeps the smaller key where bit K of
// i is clear, the larger one where it is set.  Partners at distance D < 64
// are exchanged by shuffles, D < 64 E by register swaps, longer ones (W > 1)
// through LDS (sk holds N keys).  The stages are template arguments, so every
// loop below unrolls and the key array stays in registers.
template <int E, int W, int K, int D>
__device__ __forceinline__ void GvNetStage(uint64_t (&key)[E], int lane, int w, uint64_t* sk) {
  constexpr int NW = kWave * E;
  if constexpr (D >= NW) {
    __syncthreads();  // the previous readers of sk are done
#pragma unroll
    for (int e = 0; e < E; ++e) sk[w * NW + e * kWave + lane] = key[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = w * NW + e * kWave + lane;
      const uint64_t ok = sk[i ^ D];
      const bool up = (i & K) == 0;
      const bool lower = (i & D) == 0;
      const bool take = (lower == up) ? (ok < key[e]) : (ok > key[e]);
      key[e] = take ? ok : key[e];
    }
  } else if constexpr (D >= kWave) {
    constexpr int de = D / kWave;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      if ((e & de) == 0) {
        const bool up = ((w * NW + e * kWave + lane) & K) == 0;
        const uint64_t k1 = key[e], k2 = key[e | de];
        const bool sw = (k1 > k2) == up;  // equal keys: swapping changes nothing
        key[e] = sw ? k2 : k1;
        key[e | de] = sw ? k1 : k2;
      }
    }
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const uint64_t ok = __shfl_xor(key[e], D);
      const bool up = ((w * NW + e * kWave + lane) & K) == 0;
      const bool lower = (lane & D) == 0;
      const bool take = (lower == up) ? (ok < key[e]) : (ok > key[e]);
      key[e] = take ? ok : key[e];
    }
  }
  if constexpr (D > 1) GvNetStage<E, W, K, D / 2>(key, lane, w, sk);
}

template <int E, int W, int K = 2>
__device__ __forceinline__ void GvNet(uint64_t (&key)[E], int lane, int w, uint64_t* sk) {
  GvNetStage<E, W, K, K / 2>(key, lane, w, sk);
  if constexpr (K < kWave * E * W) GvNet<E, W, K * 2>(key, lane, w, sk);
}

// Capacities in (64 E / 2, 64 E]: one wave per cluster; term q = 64 e + lane
// sits in register e of its lane.  Keys c(v) << 10 | q are sorted by a bitonic
// network (partners within a register column by shuffles, across columns by
// swaps), staged in LDS, and each run's first lane folds the run's values
// (kept in term order in LDS) left.
template <int E>
__global__ __launch_bounds__(kGvThreads) void k_gv_wave(CsrView g, const uint32_t* __restrict__ ids,
                                                        GroupIndex gi, GvItems L, int64_t count,
                                                        uint32_t* __restrict__ nbr2p,
                                                        float* __restrict__ w2p,
                                                        uint32_t* __restrict__ len) {
  constexpr int N = kWave * E;
  constexpr uint64_t kQ = 1023;  // q < N <= 1024
  __shared__ uint64_t sk[kGvWaves][N];
  __shared__ float sv[kGvWaves][N];
  __shared__ int64_t smo[kGvWaves][kWave];
  __shared__ uint32_t smu[kGvWaves][kWave], smcu[kGvWaves][kWave], smr[kGvWaves][kWave + 1];
  const int lane = LaneId();
  const int w = threadIdx.x / kWave;
  const int64_t wave = ((int64_t)blockIdx.x * kGvThreads + threadIdx.x) / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * kGvWaves;
  for (int64_t it = wave; it < count; it += nwaves) {  // uniform per wave
    const uint4 itm = L.item[it];
    const int64_t rp0 = L.rp0[it];
    const uint32_t c = itm.x, vs = itm.y, ve = itm.z;
    uint64_t key[E];
#pragma unroll
    for (int e = 0; e < E; ++e) key[e] = ~0ull;
    for (uint32_t ms = vs; ms < ve; ms += kWave) {
      const uint32_t nm = min((uint32_t)kWave, ve - ms);
      if ((uint32_t)lane < nm) {
        const uint32_t i = ms + lane;
        smu[w][lane] = gi.perm[i];
        smcu[w][lane] = gi.mcu[i];
        smo[w][lane] = gi.moff[i];
        smr[w][lane] = (uint32_t)(gi.rowpos[i] - rp0);
      }
      if (lane == 0) smr[w][nm] = (uint32_t)(gi.rowpos[ms + nm] - rp0);
      __builtin_amdgcn_wave_barrier();
      const uint32_t lo = smr[w][0], hi = smr[w][nm];
      int64_t p[E];
      int a[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const uint32_t q = (uint32_t)e * kWave + (uint32_t)lane;
        p[e] = -1;
        a[e] = 0;
        if (q >= lo && q < hi) {
          int x = 0, y = (int)nm - 1;  // the last member whose row starts at or before q
          while (x < y) {
            const int mid = (x + y + 1) >> 1;
            if (smr[w][mid] <= q)
              x = mid;
            else
              y = mid - 1;
          }
          a[e] = x;
          p[e] = smo[w][x] + (int64_t)(q - smr[w][x]);
        }
      }
      uint32_t v[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        v[e] = p[e] >= 0 ? g.nbr[p[e]] : 0u;
        if (p[e] >= 0) sv[w][e * kWave + lane] = g.w[p[e]];
      }
#pragma unroll
      for (int e = 0; e < E; ++e) {
        if (p[e] < 0) continue;
        const uint32_t cvo = ids[v[e]];
        const uint32_t cuo = smcu[w][a[e]];
        if (v[e] <= smu[w][a[e]] || cvo != cuo)
          key[e] = ((uint64_t)(cvo + (c - cuo)) << 10) | (uint64_t)(e * kWave + lane);
      }
      __builtin_amdgcn_wave_barrier();
    }
    GvNet<E, 1>(key, lane, 0, nullptr);
#pragma unroll
    for (int e = 0; e < E; ++e) sk[w][e * kWave + lane] = key[e];
    __builtin_amdgcn_wave_barrier();
    // Run heads per register column (sorted position i = 64 e + lane).  Each
    // head's c(v) goes out at once; then the sorted keys' space holds the
    // values in sorted order (first half) and the head positions (second
    // half), and the runs are folded one per lane, round robin.
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint64_t hb[E];
    int nvalid = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = e * kWave + lane;
      const uint64_t k = key[e];
      const bool valid = k != ~0ull;
      hb[e] = __ballot(valid && (i == 0 || (sk[w][i - 1] >> 10) != (k >> 10)));
      nvalid += __popcll(__ballot(valid));
    }
    __builtin_amdgcn_wave_barrier();
    float* svs = reinterpret_cast<float*>(&sk[w][0]);
    uint16_t* hpos = reinterpret_cast<uint16_t*>(&sk[w][N / 2]);
    uint32_t heads = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = e * kWave + lane;
      if (key[e] != ~0ull) svs[i] = sv[w][key[e] & kQ];
      if ((hb[e] >> lane) & 1ull) {
        const uint32_t r = heads + (uint32_t)__popcll(hb[e] & lt);
        hpos[r] = (uint16_t)i;
        nbr2p[rp0 + r] = (uint32_t)(key[e] >> 10);
      }
      heads += (uint32_t)__popcll(hb[e]);
    }
    if (lane == 0) hpos[heads] = (uint16_t)nvalid;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t h = lane; h < heads; h += kWave) {
      const uint32_t i = hpos[h];
      w2p[rp0 + h] = GvPbbsSum(svs + i, (uint32_t)hpos[h + 1] - i);
    }
    if (lane == 0) len[c] = heads;
    __builtin_amdgcn_wave_barrier();
  }
}

// Capacities in (64 E W / 2, 64 E W] for W = 2 ... 16 waves: one workgroup of
// W waves per cluster; term q = 64 E w + 64 e + lane sits in register e of
// lane `lane` of wave w.  The bitonic network over N = 64 E W keys c(v) << QB
// | q runs its stages with partner distance below 64 over shuffles, below
// 64 E by register swaps and only the longest ones (between waves) through
// LDS.  Run heads are ranked across the waves, and each run's first lane
// folds it in pbbslib's order.
template <int E, int W>
__global__ __launch_bounds__(kWave* W) void k_gv_multi(CsrView g, const uint32_t* __restrict__ ids,
                                                      GroupIndex gi, GvItems L, int64_t count,
                                                      uint32_t* __restrict__ nbr2p,
                                                      float* __restrict__ w2p,
                                                      uint32_t* __restrict__ len) {
  constexpr int T = kWave * W;
  constexpr int NW = kWave * E;  // keys per wave
  constexpr int N = NW * W;
  constexpr int QB = N <= 2048 ? 11 : (N <= 4096 ? 12 : 13);
  constexpr uint64_t kQ = (1ull << QB) - 1ull;
  static_assert(N <= (1 << QB), "term index field too narrow");
  __shared__ uint64_t sk[N];
  __shared__ float sv[N];
  __shared__ int64_t smo[T];
  __shared__ uint32_t smu[T], smcu[T], smr[T + 1];
  __shared__ uint32_t wvalid[W], wheads[W];
  const int lane = LaneId();
  const int w = threadIdx.x / kWave;
  const uint32_t t = threadIdx.x;
  for (int64_t it = blockIdx.x; it < count; it += gridDim.x) {
    const uint4 itm = L.item[it];
    const int64_t rp0 = L.rp0[it];
    const uint32_t c = itm.x, vs = itm.y, ve = itm.z;
    uint64_t key[E];
#pragma unroll
    for (int e = 0; e < E; ++e) key[e] = ~0ull;
    for (uint32_t ms = vs; ms < ve; ms += T) {
      const uint32_t nm = min((uint32_t)T, ve - ms);
      __syncthreads();  // the previous chunk / cluster is done with the staging
      if (t < nm) {
        const uint32_t i = ms + t;
        smu[t] = gi.perm[i];
        smcu[t] = gi.mcu[i];
        smo[t] = gi.moff[i];
        smr[t] = (uint32_t)(gi.rowpos[i] - rp0);
      }
      if (t == 0) smr[nm] = (uint32_t)(gi.rowpos[ms + nm] - rp0);
      __syncthreads();
      const uint32_t lo = smr[0], hi = smr[nm];
      int64_t p[E];
      int a[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const uint32_t q = (uint32_t)(w * NW + e * kWave + lane);
        p[e] = -1;
        a[e] = 0;
        if (q >= lo && q < hi) {
          int x = 0, y = (int)nm - 1;  // the last member whose row starts at or before q
          while (x < y) {
            const int mid = (x + y + 1) >> 1;
            if (smr[mid] <= q)
              x = mid;
            else
              y = mid - 1;
          }
          a[e] = x;
          p[e] = smo[x] + (int64_t)(q - smr[x]);
        }
      }
      uint32_t v[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        v[e] = p[e] >= 0 ? g.nbr[p[e]] : 0u;
        if (p[e] >= 0) sv[w * NW + e * kWave + lane] = g.w[p[e]];
      }
#pragma unroll
      for (int e = 0; e < E; ++e) {
        if (p[e] < 0) continue;
        const uint32_t cvo = ids[v[e]];
        const uint32_t cuo = smcu[a[e]];
        if (v[e] <= smu[a[e]] || cvo != cuo)
          key[e] = ((uint64_t)(cvo + (c - cuo)) << QB) | (uint64_t)(w * NW + e * kWave + lane);
      }
    }
    GvNet<E, W>(key, lane, w, sk);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) sk[w * NW + e * kWave + lane] = key[e];
    __syncthreads();
    // Run heads per register column (sorted position i = 64 E w + 64 e +
    // lane), ranked across the waves.  Each head's c(v) goes out at once;
    // then the sorted keys' space holds the values in sorted order (first
    // half) and the head positions (second half), and the runs are folded one
    // per thread, round robin.
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint64_t hb[E];
    int nvalid = 0, nheads = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = w * NW + e * kWave + lane;
      const uint64_t k = key[e];
      const bool valid = k != ~0ull;
      hb[e] = __ballot(valid && (i == 0 || (sk[i - 1] >> QB) != (k >> QB)));
      nvalid += __popcll(__ballot(valid));
      nheads += __popcll(hb[e]);
    }
    if (lane == 0) {
      wvalid[w] = (uint32_t)nvalid;
      wheads[w] = (uint32_t)nheads;
    }
    __syncthreads();  // every read of the sorted keys is done
    uint32_t total = 0, hbefore = 0, htotal = 0;
    for (int ww = 0; ww < W; ++ww) {
      total += wvalid[ww];
      htotal += wheads[ww];
      if (ww < w) hbefore += wheads[ww];
    }
    float* svs = reinterpret_cast<float*>(&sk[0]);
    uint16_t* hpos = reinterpret_cast<uint16_t*>(&sk[N / 2]);
    uint32_t heads = hbefore;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = w * NW + e * kWave + lane;
      if (key[e] != ~0ull) svs[i] = sv[key[e] & kQ];
      if ((hb[e] >> lane) & 1ull) {
        const uint32_t r = heads + (uint32_t)__popcll(hb[e] & lt);
        hpos[r] = (uint16_t)i;
        nbr2p[rp0 + r] = (uint32_t)(key[e] >> QB);
      }
      heads += (uint32_t)__popcll(hb[e]);
    }
    if (t == 0) hpos[htotal] = (uint16_t)total;
    __syncthreads();
    for (uint32_t h = t; h < htotal; h += T) {
      const uint32_t i = hpos[h];
      w2p[rp0 + h] = GvPbbsSum(svs + i, (uint32_t)hpos[h + 1] - i);
    }
    if (t == 0) len[c] = htotal;
  }
}

// Radix path, 1: the terms of a chunk of nr listed (large) clusters in CSR
// order, keyed (rank << (vbits + 1)) | c(v), dropped terms | 1 << vbits (last
// in their cluster).  Term q of the chunk's concatenation lies in cluster rank
// a with boff[a] - boff[0] <= q < boff[a + 1] - boff[0].
__global__ __launch_bounds__(256) void k_gv_big_emit(CsrView g, const uint32_t* __restrict__ ids,
                                                     GroupIndex gi, GvItems L, int64_t nr,
                                                     const int64_t* __restrict__ boff, int vbits,
                                                     uint64_t* __restrict__ keys,
                                                     float* __restrict__ vals) {
  const int64_t b0 = boff[0], total = boff[nr] - b0;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * blockDim.x) {
    int64_t a = 0, b = nr - 1;
    while (a < b) {
      const int64_t mid = (a + b + 1) >> 1;
      if (boff[mid] - b0 <= q)
        a = mid;
      else
        b = mid - 1;
    }
    const uint4 itm = L.item[a];
    const uint32_t c = itm.x;
    const int64_t P = L.rp0[a] + (q + b0 - boff[a]);
    int64_t x = itm.y, y = (int64_t)itm.z - 1;  // the last member whose row starts at or before P
    while (x < y) {
      const int64_t mid = (x + y + 1) >> 1;
      if (gi.rowpos[mid] <= P)
        x = mid;
      else
        y = mid - 1;
    }
    const uint32_t u = gi.perm[x];
    const int64_t p = gi.moff[x] + (P - gi.rowpos[x]);
    const uint32_t v = g.nbr[p];
    const uint32_t cuo = gi.mcu[x], cvo = ids[v];
    const bool kept = v <= u || cvo != cuo;
    const uint64_t low = kept ? (uint64_t)(cvo + (c - cuo)) : (1ull << vbits);
    keys[q] = ((uint64_t)a << (vbits + 1)) | low;
    vals[q] = g.w[p];
  }
}

// Radix path, 2: the first run of every cluster rank.
__global__ __launch_bounds__(256) void k_gv_big_first(int64_t nseg, const uint64_t* __restrict__ keys,
                                                      const int64_t* __restrict__ start, int vbits,
                                                      int64_t* __restrict__ first) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nseg;
       s += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t r = keys[start[s]] >> (vbits + 1);
    if (s == 0 || (keys[start[s - 1]] >> (vbits + 1)) != r) first[r] = s;
  }
}

// Radix path, 3: run s of rank r is entry s - first[r] of the cluster's row.
__global__ __launch_bounds__(256) void k_gv_big_rows(int64_t nseg, const uint64_t* __restrict__ keys,
                                                     const int64_t* __restrict__ start,
                                                     const float* __restrict__ sums,
                                                     const int64_t* __restrict__ first, GvItems L,
                                                     int vbits, uint32_t* __restrict__ nbr2p,
                                                     float* __restrict__ w2p,
                                                     uint32_t* __restrict__ len) {
  const uint64_t lowmask = (2ull << vbits) - 1ull;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nseg;
       s += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[start[s]];
    const uint64_t r = k >> (vbits + 1);
    const uint64_t low = k & lowmask;
    const bool kept = low != (1ull << vbits);
    const int64_t pos = s - first[r];
    if (kept) {
      const int64_t dst = L.rp0[r] + pos;
      nbr2p[dst] = (uint32_t)low;
      w2p[dst] = sums[s];
    }
    const bool last = s == nseg - 1 || (keys[start[s + 1]] >> (vbits + 1)) != r;
    if (last) len[L.item[r].x] = (uint32_t)(pos + (kept ? 1 : 0));
  }
}

__global__ __launch_bounds__(256) void k_gv_big_caps(int64_t nbig, GvItems L,
                                                     uint32_t* __restrict__ bcap) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nbig;
       r += (int64_t)gridDim.x * blockDim.x)
    bcap[r] = L.item[r].w;
}

// Rows from their padded positions to the compacted CSR, class by class: G
// consecutive threads per cluster (G divides the workgroup size), each
// copying every G-th entry of the row.
template <int G>
__global__ __launch_bounds__(kGvThreads) void k_gv_copy(GvItems L, int64_t count,
                                                        const uint32_t* __restrict__ len,
                                                        const int64_t* __restrict__ off2,
                                                        const uint32_t* __restrict__ nbr2p,
                                                        const float* __restrict__ w2p,
                                                        uint32_t* __restrict__ nbr,
                                                        float* __restrict__ w) {
  const int gl = threadIdx.x % G;
  const int64_t groups = (int64_t)gridDim.x * (kGvThreads / G);
  for (int64_t it = ((int64_t)blockIdx.x * kGvThreads + threadIdx.x) / G; it < count;
       it += groups) {
    const uint32_t c = L.item[it].x;
    const int64_t from = L.rp0[it], to = off2[c];
    const uint32_t l = len[c];
    for (uint32_t j = gl; j < l; j += G) {
      nbr[to + j] = nbr2p[from + j];
      w[to + j] = w2p[from + j];
    }
  }
}

__global__ __launch_bounds__(256) void k_gv_node_weights(const uint64_t* __restrict__ vkey,
                                                         const int64_t* __restrict__ start,
                                                         const double* __restrict__ sums,
                                                         int64_t nseg, int64_t n2,
                                                         double* __restrict__ nw2) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nseg;
       s += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = vkey[start[s]];
    if ((int64_t)k < n2) nw2[k] = sums[s];
  }
}

// Up to this many items in classes 0-4 (or 5-11, or all classes for the
// copies) go through one launch.
constexpr int64_t kGvFewItems = 256;

template <int G>
void LaunchSmall(const CsrView& g, const uint32_t* ids, const GroupIndex& gi, const GvItems& L,
                 int64_t count, uint32_t* nbr2p, float* w2p, uint32_t* len, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_gv_small<G>, dim3(GvGrid(count, kGvThreads / G, 65536)), dim3(kGvThreads),
                     0, s, g, ids, gi, L, count, nbr2p, w2p, len);
}

template <int G>
void LaunchCopy(const GvItems& L, int64_t count, const uint32_t* len, const int64_t* off2,
                const uint32_t* nbr2p, const float* w2p, uint32_t* nbr, float* w, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_gv_copy<G>, dim3(GvGrid(count, kGvThreads / G, 65536)), dim3(kGvThreads), 0,
                     s, L, count, len, off2, nbr2p, w2p, nbr, w);
}

template <int W>
void LaunchMulti(const CsrView& g, const uint32_t* ids, const GroupIndex& gi, const GvItems& L,
                 int64_t count, uint32_t* nbr2p, float* w2p, uint32_t* len, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL((k_gv_multi<8, W>), dim3(GvGrid(count, 1, 16384)), dim3(kWave * W), 0, s, g,
                     ids, gi, L, count, nbr2p, w2p, len);
}

template <int E>
void LaunchWave(const CsrView& g, const uint32_t* ids, const GroupIndex& gi, const GvItems& L,
                int64_t count, uint32_t* nbr2p, float* w2p, uint32_t* len, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_gv_wave<E>, dim3(GvGrid(count, kGvWaves, 65536)), dim3(kGvThreads), 0, s,
                     g, ids, gi, L, count, nbr2p, w2p, len);
}

}  // namespace

bool CompressGraphGrouped(DeviceAllocator* al, const CsrView& g, const uint32_t* ids,
                          const double* nw, const uint32_t* vgraph, const int64_t* gbase,
                          const int64_t* newbase, int64_t n2, DevCsr* out, DevBuf<double>* out_nw,
                          hipStream_t s, int64_t chunk_terms) {
  const int64_t n = g.n;
  const int vbits = GvBitsFor((uint64_t)(n2 > 1 ? n2 - 1 : 1));
  if (n <= 0 || n2 <= 0 || n >= (int64_t)0xFFFFFFFFll || n2 >= (int64_t)0xFFFFFFFFll)
    return false;  // the all-edge path handles these
  const GvMap gm{vgraph, gbase, newbase};
  const int grid_n = GvGrid(n, 256, 8192);

  // 1. Members of every new vertex, in ascending vertex order.
  DevBuf<uint64_t> vkey(al, n);
  DevBuf<uint32_t> perm(al, n), vstart(al, n2 + 1);
  DevBuf<int64_t> nstart;  // runs of equal new id in the vertex order
  int64_t ng = 0;
  {
    DevBuf<uint32_t> cnt(al, n2);
    out_nw->Reset(al, n2);
    hipLaunchKernelGGL(k_gv_keys, dim3(grid_n), dim3(256), 0, s, n, ids, gm, n2, vkey.get(),
                       perm.get(), cnt.get(), out_nw->get());
    PCC_HIP_CHECK(hipGetLastError());
    RadixSortPairsU32(al, vkey.get(), perm.get(), n, 0, GvBitsFor((uint64_t)n2), s);
    ng = Segments(al, vkey.get(), n, &nstart, s);
    if (ng > 0)
      hipLaunchKernelGGL(k_gv_counts, dim3(GvGrid(ng, 256, 8192)), dim3(256), 0, s, vkey.get(),
                         nstart.get(), ng, n2, cnt.get());
    PCC_HIP_CHECK(hipGetLastError());
    ExclusiveScanU32toU32(al, cnt.get(), vstart.get(), n2, vstart.get() + n2, s);
  }
  DevBuf<int64_t> rowpos(al, n + 1), moff(al, n);
  DevBuf<uint32_t> mcu(al, n);
  DevBuf<double> nws(al, n);
  {
    DevBuf<uint32_t> deg(al, n);
    hipLaunchKernelGGL(k_gv_members, dim3(grid_n), dim3(256), 0, s, n, g.off, ids, vkey.get(),
                       perm.get(), n2, nw, deg.get(), moff.get(), mcu.get(), nws.get());
    PCC_HIP_CHECK(hipGetLastError());
    ExclusiveScanU32toI64(al, deg.get(), rowpos.get(), n, rowpos.get() + n, s);
  }
  const GroupIndex gi{perm.get(), vstart.get(), rowpos.get(), moff.get(), mcu.get()};

  // 2. Capacity classes: item lists in ascending id order within each class.
  DevBuf<uint32_t> len(al, n2);
  const int64_t nblk = (n2 + kGvThreads - 1) / kGvThreads;
  DevBuf<uint32_t> bcnt(al, kGvClasses * nblk), bbase(al, kGvClasses * nblk + 1);
  hipLaunchKernelGGL(k_gv_class_count, dim3(GvGrid(nblk, 1, 65536)), dim3(kGvThreads), 0, s, n2,
                     gi, nblk, bcnt.get(), len.get());
  PCC_HIP_CHECK(hipGetLastError());
  ExclusiveScanU32toU32(al, bcnt.get(), bbase.get(), kGvClasses * nblk,
                        bbase.get() + kGvClasses * nblk, s);
  DevBuf<unsigned long long> dbounds(al, kGvClasses + 1);
  hipLaunchKernelGGL(k_gv_class_bounds, dim3(1), dim3(64), 0, s, bbase.get(), nblk,
                     bbase.get() + kGvClasses * nblk, dbounds.get());
  PCC_HIP_CHECK(hipGetLastError());
  unsigned long long hb[kGvClasses + 1] = {};
  ReadBack(hb, dbounds.get(), sizeof(hb), s);
  const int64_t nitems = (int64_t)hb[kGvClasses];
  const int64_t nbig = (int64_t)(hb[kGvClasses] - hb[kGvBig]);
  const int rbits = GvBitsFor((uint64_t)(nbig > 1 ? nbig - 1 : 1));
  if (nbig > 0 && rbits + vbits + 1 > 64) return false;
  DevBuf<uint4> items(al, nitems > 0 ? nitems : 1);
  DevBuf<int64_t> irp0(al, nitems > 0 ? nitems : 1);
  hipLaunchKernelGGL(k_gv_class_place, dim3(GvGrid(nblk, 1, 65536)), dim3(kGvThreads), 0, s, n2,
                     gi, nblk, bbase.get(), items.get(), irp0.get());
  PCC_HIP_CHECK(hipGetLastError());
  bcnt.Release();
  bbase.Release();
  auto cls = [&](int k) { return GvItems{items.get() + hb[k], irp0.get() + hb[k]}; };
  auto cnt = [&](int k) { return (int64_t)(hb[k + 1] - hb[k]); };

  // 3. Contracted rows at the clusters' padded positions.
  const int64_t mpad = std::max<int64_t>(g.m, 1);
  DevBuf<uint32_t> nbr2p(al, mpad);
  DevBuf<float> w2p(al, mpad);
  uint32_t* np = nbr2p.get();
  float* wp = w2p.get();
  // Each class kernel sorts any capacity up to its own, and the sorted order
  // (hence the row) does not depend on the network size, so a few clusters
  // spread over several classes take one launch of the largest kernel they
  // need: on small graphs the class launches are short and dependent.
  const int64_t nsmall = (int64_t)(hb[5] - hb[0]), nmid = (int64_t)(hb[kGvBig] - hb[5]);
  if (nsmall <= kGvFewItems) {
    LaunchSmall<64>(g, ids, gi, cls(0), nsmall, np, wp, len.get(), s);
  } else {
    LaunchSmall<4>(g, ids, gi, cls(0), cnt(0), np, wp, len.get(), s);
    LaunchSmall<8>(g, ids, gi, cls(1), cnt(1), np, wp, len.get(), s);
    LaunchSmall<16>(g, ids, gi, cls(2), cnt(2), np, wp, len.get(), s);
    LaunchSmall<32>(g, ids, gi, cls(3), cnt(3), np, wp, len.get(), s);
    LaunchSmall<64>(g, ids, gi, cls(4), cnt(4), np, wp, len.get(), s);
  }
  if (nmid <= kGvFewItems) {
    int top = 5;
    for (int k = 5; k < kGvBig; ++k)
      if (cnt(k) > 0) top = k;
    switch (top) {
      case 5: LaunchWave<2>(g, ids, gi, cls(5), nmid, np, wp, len.get(), s); break;
      case 6: LaunchWave<4>(g, ids, gi, cls(5), nmid, np, wp, len.get(), s); break;
      case 7: LaunchWave<8>(g, ids, gi, cls(5), nmid, np, wp, len.get(), s); break;
      case 8: LaunchMulti<2>(g, ids, gi, cls(5), nmid, np, wp, len.get(), s); break;
      case 9: LaunchMulti<4>(g, ids, gi, cls(5), nmid, np, wp, len.get(), s); break;
      case 10: LaunchMulti<8>(g, ids, gi, cls(5), nmid, np, wp, len.get(), s); break;
      default: LaunchMulti<16>(g, ids, gi, cls(5), nmid, np, wp, len.get(), s); break;
    }
  } else {
    LaunchWave<2>(g, ids, gi, cls(5), cnt(5), np, wp, len.get(), s);
    LaunchWave<4>(g, ids, gi, cls(6), cnt(6), np, wp, len.get(), s);
    LaunchWave<8>(g, ids, gi, cls(7), cnt(7), np, wp, len.get(), s);
    LaunchMulti<2>(g, ids, gi, cls(8), cnt(8), np, wp, len.get(), s);
    LaunchMulti<4>(g, ids, gi, cls(9), cnt(9), np, wp, len.get(), s);
    LaunchMulti<8>(g, ids, gi, cls(10), cnt(10), np, wp, len.get(), s);
    LaunchMulti<16>(g, ids, gi, cls(11), cnt(11), np, wp, len.get(), s);
  }
  PCC_HIP_CHECK(hipGetLastError());
  if (nbig > 0) {
    // Large clusters through the radix sort, in chunks of consecutive ranks
    // of at most chunk_terms terms (one cluster per chunk when it alone is
    // larger), which bounds the sort's temporaries and its key count.
    const GvItems B = cls(kGvBig);
    DevBuf<uint32_t> bcap(al, nbig);
    DevBuf<int64_t> boff(al, nbig + 1);
    hipLaunchKernelGGL(k_gv_big_caps, dim3(GvGrid(nbig, 256, 1024)), dim3(256), 0, s, nbig, B,
                       bcap.get());
    PCC_HIP_CHECK(hipGetLastError());
    ExclusiveScanU32toI64(al, bcap.get(), boff.get(), nbig, boff.get() + nbig, s);
    std::vector<int64_t> hoff((size_t)nbig + 1);
    ReadBack(hoff.data(), boff.get(), sizeof(int64_t) * (size_t)(nbig + 1), s);
    const int64_t chunk = std::max<int64_t>(chunk_terms, 1);
    for (int64_t r0 = 0; r0 < nbig;) {
      int64_t r1 = r0 + 1;
      while (r1 < nbig && hoff[(size_t)r1 + 1] - hoff[(size_t)r0] <= chunk) ++r1;
      const int64_t nr = r1 - r0, mb = hoff[(size_t)r1] - hoff[(size_t)r0];
      const int cbits = GvBitsFor((uint64_t)(nr > 1 ? nr - 1 : 1)) + vbits + 1;
      const GvItems C{B.item + r0, B.rp0 + r0};
      DevBuf<uint64_t> keys(al, mb);
      DevBuf<float> vals(al, mb);
      hipLaunchKernelGGL(k_gv_big_emit, dim3(GvGrid(mb, 256, 16384)), dim3(256), 0, s, g, ids, gi,
                         C, nr, boff.get() + r0, vbits, keys.get(), vals.get());
      PCC_HIP_CHECK(hipGetLastError());
      RadixSortPairsF32(al, keys.get(), vals.get(), mb, 0, cbits, s);
      DevBuf<int64_t> start;
      const int64_t nseg = Segments(al, keys.get(), mb, &start, s);
      DevBuf<float> sums(al, nseg);
      SegmentSumsF32(al, vals.get(), start.get(), nseg, mb, sums.get(), s);
      DevBuf<int64_t> first(al, nr);
      hipLaunchKernelGGL(k_gv_big_first, dim3(GvGrid(nseg, 256, 8192)), dim3(256), 0, s, nseg,
                         keys.get(), start.get(), vbits, first.get());
      hipLaunchKernelGGL(k_gv_big_rows, dim3(GvGrid(nseg, 256, 8192)), dim3(256), 0, s, nseg,
                         keys.get(), start.get(), sums.get(), first.get(), C, vbits, np, wp,
                         len.get());
      PCC_HIP_CHECK(hipGetLastError());
      r0 = r1;
    }
  }
  // 4. Compaction.
  out->n = n2;
  out->off.Reset(al, n2 + 1);
  ExclusiveScanU32toI64(al, len.get(), out->off.get(), n2, out->off.get() + n2, s);
  int64_t m2 = -1;
  ReadBack(&m2, out->off.get() + n2, sizeof(int64_t), s);
  if (m2 < 0 || m2 > g.m)
    throw HipError{hipErrorUnknown, "CompressGraphGrouped: contracted rows exceed the terms",
                   __FILE__, __LINE__};
  out->m = m2;
  out->nbr.Reset(al, m2 > 0 ? m2 : 1);
  out->w.Reset(al, m2 > 0 ? m2 : 1);
  if (m2 > 0) {
    const int64_t* o2 = out->off.get();
    uint32_t* onb = out->nbr.get();
    float* ow = out->w.get();
    if (nitems <= kGvFewItems) {
      LaunchCopy<64>(cls(0), nitems, len.get(), o2, np, wp, onb, ow, s);
    } else {
      LaunchCopy<4>(cls(0), cnt(0), len.get(), o2, np, wp, onb, ow, s);
      LaunchCopy<8>(cls(1), cnt(1), len.get(), o2, np, wp, onb, ow, s);
      LaunchCopy<16>(cls(2), cnt(2), len.get(), o2, np, wp, onb, ow, s);
      LaunchCopy<32>(cls(3), cnt(3), len.get(), o2, np, wp, onb, ow, s);
      LaunchCopy<64>(cls(4), hb[9] - hb[4], len.get(), o2, np, wp, onb, ow, s);
      LaunchCopy<256>(cls(9), hb[kGvClasses] - hb[9], len.get(), o2, np, wp, onb, ow, s);
    }
  }
  PCC_HIP_CHECK(hipGetLastError());
  nbr2p.Release();
  w2p.Release();
  len.Release();
  items.Release();
  irp0.Release();

  // 5. Node weights: members' weights in ascending vertex order, summed per
  // new vertex in pbbslib's order.
  DevBuf<double> gsum(al, ng > 0 ? ng : 1);
  SegmentSumsF64(al, nws.get(), nstart.get(), ng, n, gsum.get(), s);
  if (ng > 0) {
    hipLaunchKernelGGL(k_gv_node_weights, dim3(GvGrid(ng, 256, 8192)), dim3(256), 0, s,
                       vkey.get(), nstart.get(), gsum.get(), ng, n2, out_nw->get());
    PCC_HIP_CHECK(hipGetLastError());
  }
  return true;
}

}  // namespace pcc
