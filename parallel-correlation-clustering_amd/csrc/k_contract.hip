// k_contract.hip -- cluster contraction, flatten / relabel, state set-up and
// the objective and modularity reductions.
//
// Contraction restates CompressGraph
// (clustering/clusterers/parallel-correlation-clusterer-internal.cc:709-782)
// with ComputeInterClusterEdgesSort, RetrieveInterClusterEdges and GetOffsets
// (parallel/parallel-graph-utils.cc:33-62, 70-122, 156-231):
//   * directed edge u->v is kept iff v <= u or c(u) != c(v), so an
//     intra-cluster edge survives once, as a self-loop of its cluster;
//   * kept edges are stably sorted by (c(u), c(v)) (CSR order breaks ties),
//     here an LSD radix sort of (c(u) << b | c(v)) keys;
//   * each run of equal keys is summed in float in pbbslib's reduce order
//     (left fold for runs up to 1024, blocked for longer runs);
//   * node weights of a cluster are summed in ascending vertex order the same
//     way; n' = 1 + max cluster id, so unused ids become isolated vertices.

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "device_common.h"
#include "kernels.h"

namespace pcc {

namespace {

int Grid(int64_t items, int per_block, int max_blocks) {
  int64_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (int)b;
}

int BitsFor(uint64_t maxval) {
  int b = 0;
  while (b < 63 && (maxval >> b) != 0) ++b;
  return b < 1 ? 1 : b;
}

// Renumbering of a batch level: cluster c of graph p -> newbase[p] + c - gbase[p].
struct GraphMap {
  const uint32_t* vgraph;
  const int64_t* gbase;
  const int64_t* newbase;
  __device__ __forceinline__ int64_t Shift(int64_t v) const {
    if (!vgraph) return 0;
    const uint32_t p = vgraph[v];
    return newbase[p] < 0 ? INT64_MIN : newbase[p] - gbase[p];
  }
};

// Kept edges (PGU.cc:93-103) are found chunk by chunk: a workgroup stages
// kEdgeChunk consecutive CSR edges in LDS and each thread walks kEdgeItems of
// them in order, tracking the source vertex, so hub vertices are spread over
// many workgroups and the kept edges come out in CSR order.
constexpr int kEdgeItems = 16;
constexpr int kEdgeThreads = 256;
constexpr int64_t kEdgeChunk = (int64_t)kEdgeItems * kEdgeThreads;

struct EdgeWalk {
  int64_t u = -1, uend = 0;
  uint32_t cu = 0;
  int64_t shift = 0;
  // Advances to the source of edge p.
  __device__ __forceinline__ void To(const CsrView& g, const uint32_t* __restrict__ ids,
                                     const GraphMap& gm, int64_t p) {
    if (p < uend) return;
    u = u < 0 ? VertexOfEdge(g.off, 0, g.n, p) : VertexOfEdge(g.off, u + 1, g.n, p);
    uend = g.off[u + 1];
    cu = ids[u];
    shift = gm.Shift(u);
  }
};

// Symmetry fingerprint: sum over directed edges of H(u, v) - H(v, u), H a
// 64-bit mix of the pair; zero for a symmetric edge multiset.
__device__ __forceinline__ uint64_t PairHash(uint64_t a, uint64_t b) {
  uint64_t z = ((a << 32) | b) + 0x9E3779B97F4A7C15ull;  // splitmix64 finaliser
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(kEdgeThreads) void k_sym_fingerprint(CsrView g, int64_t nchunks,
                                                                  unsigned long long* __restrict__ acc) {
  unsigned long long x = 0;
  for (int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const int64_t c0 = ch * kEdgeChunk;
    const int64_t len = min(kEdgeChunk, g.m - c0);
    const int64_t q0 = (int64_t)threadIdx.x * kEdgeItems;
    if (q0 >= len) continue;
    const int64_t q1 = min(q0 + kEdgeItems, len);
    int64_t u = VertexOfEdge(g.off, 0, g.n, c0 + q0);
    int64_t uend = g.off[u + 1];
    for (int64_t q = q0; q < q1; ++q) {
      const int64_t p = c0 + q;
      if (p >= uend) {
        u = VertexOfEdge(g.off, u + 1, g.n, p);
        uend = g.off[u + 1];
      }
      const uint64_t v = g.nbr[p];
      x += PairHash((uint64_t)u, v) - PairHash(v, (uint64_t)u);
    }
  }
  for (int sh = 32; sh > 0; sh >>= 1) x += __shfl_xor(x, sh);
  if (LaneId() == 0 && x) atomicAdd(acc, x);
}

__global__ __launch_bounds__(kEdgeThreads) void k_count_kept(CsrView g,
                                                             const uint32_t* __restrict__ ids,
                                                             GraphMap gm, int64_t nchunks,
                                                             uint32_t* __restrict__ chunk_cnt) {
  __shared__ uint32_t snbr[kEdgeChunk];
  __shared__ uint32_t sw[kEdgeThreads / kWave];
  for (int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const int64_t c0 = ch * kEdgeChunk;
    const int64_t len = min(kEdgeChunk, g.m - c0);
    __syncthreads();
    for (int64_t k = threadIdx.x; k < len; k += kEdgeThreads) snbr[k] = g.nbr[c0 + k];
    __syncthreads();
    uint32_t c = 0;
    const int64_t q0 = (int64_t)threadIdx.x * kEdgeItems;
    if (q0 < len) {
      EdgeWalk wk;
      const int64_t q1 = min(q0 + kEdgeItems, len);
      for (int64_t q = q0; q < q1; ++q) {
        wk.To(g, ids, gm, c0 + q);
        if (wk.shift == INT64_MIN) continue;
        const uint32_t v = snbr[q];
        if ((int64_t)v <= wk.u || ids[v] != wk.cu) ++c;
      }
    }
    uint32_t total = 0;
    (void)BlockExclusiveScan(c, sw, &total);
    if (threadIdx.x == 0) chunk_cnt[ch] = total;
  }
}

// Writes the kept edges as (key, weight) in CSR order; chunk ch starts at
// chunk_off[ch].
__global__ __launch_bounds__(kEdgeThreads) void k_emit_kept(CsrView g,
                                                            const uint32_t* __restrict__ ids,
                                                            GraphMap gm, int64_t nchunks,
                                                            const int64_t* __restrict__ chunk_off,
                                                            int vbits, uint64_t* __restrict__ keys,
                                                            float* __restrict__ vals) {
  __shared__ uint32_t snbr[kEdgeChunk];
  __shared__ float sw8[kEdgeChunk];
  __shared__ uint32_t sw[kEdgeThreads / kWave];
  for (int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const int64_t c0 = ch * kEdgeChunk;
    const int64_t len = min(kEdgeChunk, g.m - c0);
    __syncthreads();
    for (int64_t k = threadIdx.x; k < len; k += kEdgeThreads) {
      snbr[k] = g.nbr[c0 + k];
      sw8[k] = g.w[c0 + k];
    }
    __syncthreads();
    uint64_t key[kEdgeItems];
    uint32_t keep = 0;  // bit j: edge q0 + j kept
    uint32_t c = 0;
    const int64_t q0 = (int64_t)threadIdx.x * kEdgeItems;
    if (q0 < len) {
      EdgeWalk wk;
#pragma unroll
      for (int j = 0; j < kEdgeItems; ++j) {
        const int64_t q = q0 + j;
        key[j] = 0;
        if (q >= len) continue;
        wk.To(g, ids, gm, c0 + q);
        if (wk.shift == INT64_MIN) continue;
        const uint32_t v = snbr[q];
        const uint32_t cv = ids[v];
        if ((int64_t)v <= wk.u || cv != wk.cu) {
          key[j] = ((uint64_t)(wk.cu + wk.shift) << vbits) | (uint64_t)(cv + wk.shift);
          keep |= 1u << j;
          ++c;
        }
      }
    }
    uint32_t total = 0;
    uint32_t at = BlockExclusiveScan(c, sw, &total);
    const int64_t base = chunk_off[ch] + at;
    int64_t k = 0;
#pragma unroll
    for (int j = 0; j < kEdgeItems; ++j) {
      if (keep & (1u << j)) {
        keys[base + k] = key[j];
        vals[base + k] = sw8[q0 + j];
        ++k;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_heads(const uint64_t* __restrict__ keys, int64_t n,
                                               uint8_t* __restrict__ head) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_seg_starts(const uint8_t* __restrict__ head,
                                                    const uint32_t* __restrict__ rank, int64_t n,
                                                    int64_t* __restrict__ start) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if (head[i]) start[rank[i]] = i;
  if (blockIdx.x == 0 && threadIdx.x == 0) start[0] = 0;
}

__global__ void k_set_i64(int64_t* p, int64_t v) { *p = v; }

// Segment sums, pbbslib reduce order: a run of len terms is folded left when
// len <= 1024; otherwise it is cut into blocks of bs = max(1024, 4 ceil(sqrt
// len)) terms, each block is folded left, and the block sums are reduced the
// same way.  Runs of at most 1024 terms are folded by one thread here; longer
// runs are queued for the blocked kernels below.
template <typename V>
__global__ __launch_bounds__(256) void k_seg_short(const V* __restrict__ vals,
                                                   const int64_t* __restrict__ start, int64_t nseg,
                                                   V* __restrict__ out, uint32_t* __restrict__ longq,
                                                   unsigned long long* __restrict__ nlong) {
  for (int64_t sgi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; sgi < nseg;
       sgi += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = start[sgi], e = start[sgi + 1];
    if (e - s > 1024) {
      const unsigned long long q = atomicAdd(nlong, 1ull);
      longq[q] = (uint32_t)sgi;
      continue;
    }
    V r = vals[s];
    for (int64_t i = s + 1; i < e; ++i) r += vals[i];
    out[sgi] = r;
  }
}

__device__ __forceinline__ int64_t PbbsBlockSize(int64_t len) {
  const int64_t bs = (int64_t)ceil(sqrt((double)len)) * 4;
  return bs < 1024 ? 1024 : bs;
}

// Number of level-1 blocks of each queued run.
__global__ __launch_bounds__(256) void k_seg_nblocks(const int64_t* __restrict__ start,
                                                     const uint32_t* __restrict__ longq,
                                                     int64_t nlong, uint32_t* __restrict__ nblk) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nlong;
       q += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t sgi = longq[q];
    const int64_t len = start[sgi + 1] - start[sgi];
    const int64_t bs = PbbsBlockSize(len);
    nblk[q] = (uint32_t)((len + bs - 1) / bs);
  }
}

// Ordered left fold of src[0, len) (len >= 1) by one wave; every lane
// returns the result.
template <typename V>
__device__ __forceinline__ V WaveFoldLeft(const V* __restrict__ src, int64_t len) {
  const int lane = LaneId();
  V acc = src[0];
  for (int64_t c0 = 1; c0 < len; c0 += kWave) {
    const int64_t nk = min((int64_t)kWave, len - c0);
    const V x = lane < nk ? src[c0 + lane] : (V)0;
    for (int j = 0; j < nk; ++j) acc += __shfl(x, j);
  }
  return acc;
}

// One wave per level-1 block (item w of the concatenated block lists, boff =
// exclusive prefix of the block counts): the block's sum in the reference's
// order.  An unordered sum is used when it is exact in every order -- every
// term a multiple of 2^g and the sum of magnitudes below 2^(g + p - 1), p the
// significand width of V -- and an ordered fold otherwise (also for all-zero
// blocks, whose result sign depends on the order, and non-finite terms).
template <typename V>
__global__ __launch_bounds__(256) void k_seg_blocks(const V* __restrict__ vals,
                                                    const int64_t* __restrict__ start,
                                                    const uint32_t* __restrict__ longq,
                                                    const int64_t* __restrict__ boff, int64_t nlong,
                                                    V* __restrict__ part) {
  constexpr int kP = sizeof(V) == 4 ? 24 : 53;
  const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (w >= boff[nlong]) return;
  int64_t lo = 0, hi = nlong - 1;  // largest q with boff[q] <= w
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (boff[mid] <= w) lo = mid; else hi = mid - 1;
  }
  const uint32_t sgi = longq[lo];
  const int64_t s = start[sgi], len = start[sgi + 1] - s;
  const int64_t bs = PbbsBlockSize(len);
  const int64_t k = w - boff[lo];
  const int64_t b0 = k * bs, b1 = min(len, b0 + bs);
  const V* src = vals + s + b0;
  const int64_t bl = b1 - b0;
  double sum = 0.0, sabs = 0.0;
  int gmin = 1 << 20;
  for (int64_t i = LaneId(); i < bl; i += kWave) {
    const double x = (double)src[i];
    sum += x;
    sabs += fabs(x);
    gmin = min(gmin, GranExp(x));
  }
  sum = WaveSum(sum);
  sabs = WaveSum(sabs);
  gmin = WaveMinInt(gmin);
  const bool exact = sabs > 0.0 && sabs < INFINITY &&
                     (gmin >= (1 << 19) || sabs < ldexp(1.0, gmin + kP - 1));
  const V r = exact ? (V)sum : WaveFoldLeft(src, bl);
  if (LaneId() == 0) part[w] = r;
}

// One wave per queued run: reduces its level-1 block sums in the reference's
// order (blocks of max(1024, 4 ceil(sqrt l)) folded left, repeated), in place.
template <typename V>
__global__ __launch_bounds__(256) void k_seg_top(const uint32_t* __restrict__ longq,
                                                 const int64_t* __restrict__ boff, int64_t nlong,
                                                 V* __restrict__ part, V* __restrict__ out) {
  const int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (q >= nlong) return;
  V* p = part + boff[q];
  int64_t len = boff[q + 1] - boff[q];
  while (len > 1) {
    const int64_t bs = PbbsBlockSize(len);
    const int64_t l = (len + bs - 1) / bs;
    if (l <= 1) break;
    // Block k's sum lands at index k <= k * bs, after the block was read.
    for (int64_t k = 0; k < l; ++k) {
      const int64_t b0 = k * bs;
      const V r = WaveFoldLeft(p + b0, min(bs, len - b0));
      if (LaneId() == 0) p[k] = r;
      __threadfence();  // the next pass reads p[k] from every lane
    }
    len = l;
  }
  const V r = WaveFoldLeft(p, len);
  if (LaneId() == 0) out[longq[q]] = r;
}

__global__ __launch_bounds__(256) void k_emit_compressed(const uint64_t* __restrict__ keys,
                                                         const int64_t* __restrict__ start,
                                                         int64_t nseg, int vbits,
                                                         uint32_t* __restrict__ nbr,
                                                         uint32_t* __restrict__ cnt) {
  // Segments are sorted by source, so a wave's sources form runs; the first
  // lane of each run adds the run length.
  const uint64_t vmask = (1ull << vbits) - 1ull;
  const int lane = LaneId();
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < nseg;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = base + threadIdx.x;
    const bool in = s < nseg;
    uint64_t src = ~0ull;
    if (in) {
      const uint64_t k = keys[start[s]];
      nbr[s] = (uint32_t)(k & vmask);
      src = k >> vbits;
    }
    const uint64_t prev = __shfl_up(src, 1);
    const bool head = in && (lane == 0 || prev != src);
    const uint64_t heads = __ballot(head);
    const uint64_t valid = __ballot(in);  // a prefix of the wave's lanes
    if (head) {
      const uint64_t later = lane == 63 ? 0ull : (heads & (~0ull << (lane + 1)));
      const int end = later ? __ffsll((unsigned long long)later) - 1 : 64 - __clzll(valid);
      atomicAdd(&cnt[src], (uint32_t)(end - lane));
    }
  }
}

__global__ __launch_bounds__(256) void k_node_keys(int64_t n, const uint32_t* __restrict__ ids,
                                                   GraphMap gm, int64_t n2,
                                                   const double* __restrict__ nw,
                                                   uint64_t* __restrict__ keys,
                                                   double* __restrict__ vals) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t shift = gm.Shift(v);
    // Vertices of graphs that leave the batch sort last under key n2.
    keys[v] = shift == INT64_MIN ? (uint64_t)n2 : (uint64_t)(ids[v] + shift);
    vals[v] = nw[v];
  }
}

__global__ __launch_bounds__(256) void k_scatter_node_weights(const uint64_t* __restrict__ keys,
                                                              const int64_t* __restrict__ start,
                                                              const double* __restrict__ sums,
                                                              int64_t nseg, int64_t n2,
                                                              double* __restrict__ nw2) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nseg;
       s += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[start[s]];
    if ((int64_t)k < n2) nw2[k] = sums[s];
  }
}

__global__ __launch_bounds__(256) void k_graph_max(int64_t n, const uint32_t* __restrict__ ids,
                                                   const uint32_t* __restrict__ vgraph,
                                                   uint32_t* __restrict__ gmax) {
  if (vgraph == nullptr) {  // one graph: a block maximum, one atomic per block
    __shared__ uint32_t smax;
    if (threadIdx.x == 0) smax = 0;
    __syncthreads();
    uint32_t mx = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x)
      mx = max(mx, ids[v]);
    for (int s = 32; s > 0; s >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, s));
    if (LaneId() == 0) atomicMax(&smax, mx);
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(&gmax[0], smax);
    return;
  }
  // Batches: one atomic per wave when the wave's vertices share a graph,
  // otherwise one per vertex.
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = base + threadIdx.x;
    const bool in = v < n;
    const uint32_t x = in ? ids[v] : 0u;
    const uint32_t gi = (in && vgraph) ? vgraph[v] : 0u;
    const uint32_t g0 = __shfl(gi, 0);
    if (__all(!in || gi == g0)) {
      uint32_t mx = x;
      for (int s = 32; s > 0; s >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, s));
      if (LaneId() == 0) atomicMax(&gmax[g0], mx);
    } else if (in) {
      atomicMax(&gmax[gi], x);
    }
  }
}

__global__ __launch_bounds__(256) void k_flatten_shift(int64_t n, uint32_t* __restrict__ ids,
                                                       const uint32_t* __restrict__ vgraph0,
                                                       const uint8_t* __restrict__ sel,
                                                       const int64_t* __restrict__ shift,
                                                       const uint32_t* __restrict__ next_ids) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t o = vgraph0 ? vgraph0[v] : 0u;
    if (sel && !sel[o]) continue;
    const int64_t x = (int64_t)ids[v] + (shift ? shift[o] : 0);
    ids[v] = next_ids ? next_ids[x] : (uint32_t)x;
  }
}

// Batch refinement relabel (see ResetClusteringBatch).
__global__ __launch_bounds__(256) void k_batch_flat(int64_t n, const uint32_t* __restrict__ ids,
                                                    const uint32_t* __restrict__ next_ids,
                                                    const uint32_t* __restrict__ vgraph,
                                                    const int64_t* __restrict__ gbase,
                                                    const int64_t* __restrict__ nbase,
                                                    const uint8_t* __restrict__ sel,
                                                    uint32_t* __restrict__ flat,
                                                    uint8_t* __restrict__ present) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t p = vgraph ? vgraph[v] : 0u;
    if (sel && !sel[p]) continue;
    const int64_t gb = gbase ? gbase[p] : 0, nb = nbase ? nbase[p] : 0;
    const uint32_t y = next_ids[nb + (int64_t)ids[v] - gb];
    flat[v] = y;
    present[y] = 1;
  }
}

__global__ __launch_bounds__(256) void k_batch_clear(int64_t n, const uint32_t* __restrict__ vgraph,
                                                     const uint8_t* __restrict__ sel, StateView st) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (sel && !sel[vgraph ? vgraph[i] : 0u]) continue;
    st.sizes[i] = 0;
    st.K[i] = 0.0;
  }
}

__global__ __launch_bounds__(256) void k_batch_assign(int64_t n, const uint32_t* __restrict__ flat,
                                                      const uint32_t* __restrict__ prank,
                                                      const uint32_t* __restrict__ vgraph,
                                                      const int64_t* __restrict__ gbase,
                                                      const int64_t* __restrict__ nbase,
                                                      const uint8_t* __restrict__ sel, StateView st) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t p = vgraph ? vgraph[v] : 0u;
    if (sel && !sel[p]) continue;
    const int64_t gb = gbase ? gbase[p] : 0, nb = nbase ? nbase[p] : 0;
    const uint32_t r = (uint32_t)(gb + (int64_t)(prank[flat[v]] - prank[nb]));
    st.ids[v] = r;
    atomicAdd(&st.sizes[r], 1u);
    atomicAdd(&st.K[r], st.nw[v]);
  }
}

__global__ __launch_bounds__(256) void k_flatten(int64_t n, const uint32_t* __restrict__ ids,
                                                 const uint32_t* __restrict__ comp,
                                                 uint32_t* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = ids[i] == kNoneId ? kNoneId : comp[ids[i]];
}

__global__ __launch_bounds__(256) void k_mark_present(int64_t n, const uint32_t* __restrict__ ids,
                                                      uint8_t* __restrict__ present) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    present[ids[i]] = 1;
}

__global__ __launch_bounds__(256) void k_clear_state(int64_t n, StateView st) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    st.sizes[i] = 0;
    st.K[i] = 0.0;
  }
}

__global__ __launch_bounds__(256) void k_assign_ranked(int64_t n, const uint32_t* __restrict__ flat,
                                                       const uint32_t* __restrict__ rank,
                                                       StateView st) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t r = rank[flat[v]];
    st.ids[v] = r;
    atomicAdd(&st.sizes[r], 1u);
    atomicAdd(&st.K[r], st.nw[v]);
  }
}

__global__ __launch_bounds__(256) void k_assign_init(int64_t n, const uint32_t* __restrict__ init,
                                                     StateView st) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t r = init[v];
    st.ids[v] = r;
    atomicAdd(&st.sizes[r], 1u);
    atomicAdd(&st.K[r], st.nw[v]);
  }
}

__global__ __launch_bounds__(256) void k_singletons(int64_t n, StateView st) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    st.ids[i] = (uint32_t)i;
    st.sizes[i] = 1;
    st.K[i] = 0 + st.nw[i];
  }
}

__global__ __launch_bounds__(256) void k_degrees(CsrView g, double* __restrict__ nw) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n;
       v += (int64_t)gridDim.x * blockDim.x)
    nw[v] = (double)(g.off[v + 1] - g.off[v]);
}

// Adds x into a[key] with one atomic per distinct key of the wave (large
// clusters otherwise serialize on a single address).  All lanes call it; act
// marks the lanes that contribute.
__device__ __forceinline__ void WaveKeyedAdd(bool act, uint32_t key, double x, double* a) {
  uint64_t todo = __ballot(act);
  while (todo) {
    const int leader = __ffsll((unsigned long long)todo) - 1;
    const uint32_t k = __shfl(key, leader);
    const bool mine = act && key == k;
    const uint64_t grp = __ballot(mine);
    double y = mine ? x : 0.0;
    for (int sh = 32; sh > 0; sh >>= 1) y += __shfl_xor(y, sh);
    if (LaneId() == leader) atomicAdd(&a[k], y);
    todo &= ~grp;
  }
}

// Node weight of v; nw null means unit weights.
__device__ __forceinline__ double NwAt(const double* __restrict__ nw, int64_t v) {
  return nw ? nw[v] : 1.0;
}

// gate (may be null): when *gate != 0 the kernel does nothing (the cluster
// ids were found out of range by the allocation-free entry point).
__global__ __launch_bounds__(256) void k_cluster_weights(int64_t n, const uint32_t* __restrict__ ids,
                                                         const double* __restrict__ nw,
                                                         double* __restrict__ K,
                                                         const unsigned int* __restrict__ gate) {
  if (gate && *gate) return;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = base + threadIdx.x;
    const bool in = v < n;
    WaveKeyedAdd(in, in ? ids[v] : 0u, in ? NwAt(nw, v) : 0.0, K);
  }
}

// Per-vertex terms of ComputeObjective (PCCI.cc:124-145): the intra-cluster
// half weights folded in pbbs order, and k_v (K_c(v) - k_v).
// Per-vertex intra-cluster edge weight (halved) in pbbslib's reduce order
// (PCCI.cc:128-138): a left fold for degrees up to one reduce block of
// max(1024, 4 ceil(sqrt(deg))) terms, block sums folded left otherwise.
// Degrees up to kObjThreadDeg are folded by the vertex's own thread; larger
// ones by the whole wave, which loads kObjChunk consecutive terms at once and
// folds them in order through shuffles.
__device__ __forceinline__ double ObjTerm(const CsrView& g, const uint32_t* __restrict__ ids,
                                          uint32_t cv, double offset, int64_t p) {
  const float w = g.w[p];
  return ids[g.nbr[p]] == cv ? (w - offset) / 2 : 0.0;
}

constexpr int kObjChunk = 4 * kWave;
// Graphs with more vertices than this queue their mid-size rows.
constexpr int64_t kObjQueueN = 65536;
// Rows up to this degree are folded by their vertex's own thread.
constexpr int64_t kObjThreadDeg = 16;

// Ordered left fold of terms [s0, en) of vertex (vb, vc) onto acc (have: acc
// holds a value); wave-uniform.
__device__ __forceinline__ double ObjFold(const CsrView& g, const uint32_t* __restrict__ ids,
                                          uint32_t vc, double offset, int64_t vb, int64_t s0,
                                          int64_t en, double acc, bool have) {
  const int lane = LaneId();
  for (int64_t c0 = s0; c0 < en; c0 += kObjChunk) {
    double x[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t p = c0 + k * kWave + lane;
      x[k] = p < en ? ObjTerm(g, ids, vc, offset, vb + p) : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t nk = min((int64_t)kWave, en - c0 - k * kWave);
      for (int j = 0; j < nk; ++j) {
        const double y = __shfl(x[k], j);
        acc = have ? acc + y : y;
        have = true;
      }
    }
  }
  return acc;
}

// The intra-cluster half weight of one vertex with more than kObjThreadDeg
// and at most kLongRow edges (row [vb, vb + vd), cluster vc), by one wave.
__device__ __forceinline__ double WaveObjRow(const CsrView& g, const uint32_t* __restrict__ ids,
                                             uint32_t vc, double offset, int64_t vb, int64_t vd) {
  const int lane = LaneId();
  // First an any-order sum with the order-free test: at most two nonzero
  // terms, or all terms multiples of 2^gmin with sum of magnitudes below
  // 2^(gmin + 52), make every partial sum exact, so the fold order of the
  // reference cannot change the result.
  double sum = 0.0, sabs = 0.0;
  int gmin = 1 << 20, nz = 0;
  for (int64_t c0 = 0; c0 < vd; c0 += kObjChunk) {
    double x[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t p = c0 + k * kWave + lane;
      x[k] = p < vd ? ObjTerm(g, ids, vc, offset, vb + p) : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      sum += x[k];
      sabs += fabs(x[k]);
      gmin = min(gmin, GranExp(x[k]));
      nz += x[k] != 0.0;
    }
  }
  sum = WaveSum(sum);
  sabs = WaveSum(sabs);
  gmin = WaveMinInt(gmin);
  for (int sh = 32; sh > 0; sh >>= 1) nz += __shfl_xor(nz, sh);
  double tot = sum;
  if (nz > 2 && !(gmin >= (1 << 19) || sabs < ldexp(1.0, gmin + 52))) {
    int64_t bs = (int64_t)ceil(sqrt((double)vd)) * 4;
    if (bs < 1024) bs = 1024;
    bool have = false;
    for (int64_t s0 = 0; s0 < vd; s0 += bs) {
      const double blk = ObjFold(g, ids, vc, offset, vb, s0, min(s0 + bs, vd), 0.0, false);
      tot = have ? tot + blk : blk;
      have = true;
    }
  }
  return tot;
}

__global__ __launch_bounds__(256) void k_objective_terms(CsrView g, const uint32_t* __restrict__ ids,
                                                         const double* __restrict__ nw,
                                                         const double* __restrict__ K,
                                                         double offset, double* __restrict__ e,
                                                         double* __restrict__ r,
                                                         uint32_t* __restrict__ longq,
                                                         unsigned int* __restrict__ nlong,
                                                         const unsigned int* __restrict__ gate,
                                                         uint32_t* __restrict__ midq,
                                                         unsigned int* __restrict__ nmid) {
  if (gate && *gate) return;
  const int lane = LaneId();
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < g.n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = base + threadIdx.x;
    const bool in = v < g.n;
    uint32_t cv = 0;
    int64_t b = 0, deg = 0;
    if (in) {
      cv = ids[v];
      b = g.off[v];
      deg = g.off[v + 1] - b;
      const double k = NwAt(nw, v);
      r[v] = k * (K[cv] - k);
      if (deg > kLongRow) longq[atomicAdd(nlong, 1u)] = (uint32_t)v;  // k_objective_long
      if (deg <= kObjThreadDeg) {
        // The left fold (the reference's order for rows of at most 1024
        // terms), eight terms' loads in flight at a time.
        double tot = 0.0;
        for (int64_t p0 = 0; p0 < deg; p0 += 8) {
          uint32_t u[8];
          float w[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const bool ok = p0 + k < deg;
            u[k] = ok ? g.nbr[b + p0 + k] : 0u;
            w[k] = ok ? g.w[b + p0 + k] : 0.0f;
          }
          uint32_t c[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) c[k] = p0 + k < deg ? ids[u[k]] : ~cv;
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            if (p0 + k >= deg) break;
            const double x = c[k] == cv ? ((double)w[k] - offset) / 2 : 0.0;
            tot = p0 + k == 0 ? x : tot + x;
          }
        }
        e[v] = tot;
      }
    }
    uint64_t big = __ballot(in && deg > kObjThreadDeg && deg <= kLongRow);
    if (midq) {  // queued for k_objective_mid
      if (big) {
        const int leader = __ffsll((unsigned long long)big) - 1;
        unsigned int pos = 0;
        if (lane == leader) pos = atomicAdd(nmid, (unsigned int)__popcll(big));
        pos = __shfl(pos, leader);
        const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        if ((big >> lane) & 1ull) midq[pos + __popcll(big & lt)] = (uint32_t)v;
      }
      continue;
    }
    const int64_t wave_v0 = base + (threadIdx.x & ~(kWave - 1));
    while (big) {
      const int l = __ffsll((unsigned long long)big) - 1;
      big &= big - 1;
      const double tot = WaveObjRow(g, ids, __shfl(cv, l), offset, __shfl(b, l), __shfl(deg, l));
      if (lane == 0) e[wave_v0 + l] = tot;
    }
  }
}

// The rows k_objective_terms queued (more than kObjThreadDeg and at most
// kLongRow edges), one wave each over all waves of the grid, so that rows of
// similar id (and, on RMAT-like graphs, similar degree) do not pile up on a
// few waves.
__global__ __launch_bounds__(256) void k_objective_mid(CsrView g, const uint32_t* __restrict__ ids,
                                                       double offset,
                                                       const uint32_t* __restrict__ midq,
                                                       const unsigned int* __restrict__ nmid,
                                                       double* __restrict__ e) {
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / kWave);
  const int64_t count = (int64_t)*nmid;
  for (int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; q < count; q += nwaves) {
    const uint32_t v = midq[q];
    const int64_t vb = g.off[v];
    const double tot = WaveObjRow(g, ids, ids[v], offset, vb, g.off[v + 1] - vb);
    if (LaneId() == 0) e[v] = tot;
  }
}

// Rows longer than kLongRow: the any-order sum and the order-free test's
// inputs, chunk by chunk over the whole grid (acc: sum, sum of magnitudes;
// gmin; nonzero count).
__global__ __launch_bounds__(256) void k_objective_long(CsrView g, const uint32_t* __restrict__ ids,
                                                        double offset,
                                                        const uint32_t* __restrict__ longq,
                                                        const unsigned int* __restrict__ nlong,
                                                        double* __restrict__ acc,
                                                        int* __restrict__ gmin,
                                                        unsigned long long* __restrict__ nz) {
  __shared__ double sm[256 / kWave];
  __shared__ int si[256 / kWave];
  __shared__ unsigned long long su[256 / kWave];
  ForLongChunks(longq, *nlong, g.off, [&](uint32_t q, uint32_t v, int64_t e0, int64_t e1) {
    const uint32_t cv = ids[v];
    double sum = 0.0, sabs = 0.0;
    int gm = 1 << 20;
    unsigned long long z = 0;
    for (int64_t p = e0 + threadIdx.x; p < e1; p += blockDim.x) {
      const double x = ObjTerm(g, ids, cv, offset, p);
      sum += x;
      sabs += fabs(x);
      gm = min(gm, GranExp(x));
      z += x != 0.0;
    }
    sum = BlockReduce(sum, sm, [](double a, double b) { return a + b; });
    sabs = BlockReduce(sabs, sm, [](double a, double b) { return a + b; });
    gm = BlockReduce(gm, si, [](int a, int b) { return a < b ? a : b; });
    z = BlockReduce(z, su, [](unsigned long long a, unsigned long long b) { return a + b; });
    if (threadIdx.x == 0) {
      atomicAdd(&acc[2 * q], sum);
      atomicAdd(&acc[2 * q + 1], sabs);
      atomicMin(&gmin[q], gm);
      atomicAdd(&nz[q], z);
    }
  });
}

// One wave per long row: the any-order sum when it is exact in every order
// (at most two nonzero terms, or all terms multiples of 2^gmin with the sum of
// magnitudes below 2^(gmin + 51): a margin of one bit for the rounding of the
// combined magnitude sum), the pbbs-order fold otherwise.
__global__ __launch_bounds__(256) void k_objective_long_final(CsrView g,
                                                              const uint32_t* __restrict__ ids,
                                                              double offset,
                                                              const uint32_t* __restrict__ longq,
                                                              const unsigned int* __restrict__ nlong,
                                                              const double* __restrict__ acc,
                                                              const int* __restrict__ gmin,
                                                              const unsigned long long* __restrict__ nz,
                                                              double* __restrict__ e) {
  const int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nw = (int64_t)gridDim.x * blockDim.x / kWave;
  for (int64_t qq = q; qq < (int64_t)*nlong; qq += nw) {
    const uint32_t v = longq[qq];
    const double sum = acc[2 * qq], sabs = acc[2 * qq + 1];
    const int gm = gmin[qq];
    double tot = sum;
    if (nz[qq] > 2 && !(gm >= (1 << 19) || sabs < ldexp(1.0, gm + 51))) {
      const uint32_t vc = ids[v];
      const int64_t vb = g.off[v], vd = g.off[v + 1] - vb;
      int64_t bs = (int64_t)ceil(sqrt((double)vd)) * 4;
      if (bs < 1024) bs = 1024;
      bool have = false;
      for (int64_t s0 = 0; s0 < vd; s0 += bs) {
        const double blk = ObjFold(g, ids, vc, offset, vb, s0, min(s0 + bs, vd), 0.0, false);
        tot = have ? tot + blk : blk;
        have = true;
      }
    }
    if (LaneId() == 0) e[v] = tot;
  }
}

// Per-vertex intra-cluster edge count and per-cluster degree sums
// (PMC.cc:43-61).
__device__ __forceinline__ double WaveIntraCount(const CsrView& g, const uint32_t* __restrict__ ids,
                                                 uint32_t vc, int64_t vb, int64_t vd) {
  const int lane = LaneId();
  unsigned long long cnt = 0;
  for (int64_t p0 = lane; p0 < vd; p0 += 4 * kWave) {
    uint32_t u[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) u[k] = p0 + k * kWave < vd ? g.nbr[vb + p0 + k * kWave] : 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) cnt += (p0 + k * kWave < vd && ids[u[k]] == vc) ? 1ull : 0ull;
  }
  for (int s2 = 32; s2 > 0; s2 >>= 1) cnt += __shfl_xor(cnt, s2);
  return (double)cnt;
}

__global__ __launch_bounds__(256) void k_modularity_terms(CsrView g, const uint32_t* __restrict__ ids,
                                                          double* __restrict__ intra,
                                                          double* __restrict__ dsum,
                                                          uint32_t* __restrict__ midq,
                                                          unsigned int* __restrict__ nmid) {
  // Integer counts: any order is exact.  Degrees above 64 are counted by the
  // whole wave, rows above kLongRow by k_modularity_long.
  const int lane = LaneId();
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < g.n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = base + threadIdx.x;
    const bool in = v < g.n;
    uint32_t cv = 0;
    int64_t b = 0, deg = 0;
    if (in) {
      cv = ids[v];
      b = g.off[v];
      deg = g.off[v + 1] - b;
      if (deg <= 64) {
        int64_t cnt = 0;
        for (int64_t p0 = b; p0 < b + deg; p0 += 8) {
          uint32_t u[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) u[k] = p0 + k < b + deg ? g.nbr[p0 + k] : 0u;
#pragma unroll
          for (int k = 0; k < 8; ++k) cnt += (p0 + k < b + deg && ids[u[k]] == cv) ? 1 : 0;
        }
        intra[v] = (double)cnt;
      }
    }
    WaveKeyedAdd(in, cv, (double)deg, dsum);
    uint64_t big = __ballot(in && deg > 64 && deg <= kLongRow);
    if (midq) {  // queued for k_modularity_mid
      if (big) {
        const int leader = __ffsll((unsigned long long)big) - 1;
        unsigned int pos = 0;
        if (lane == leader) pos = atomicAdd(nmid, (unsigned int)__popcll(big));
        pos = __shfl(pos, leader);
        const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        if ((big >> lane) & 1ull) midq[pos + __popcll(big & lt)] = (uint32_t)v;
      }
      continue;
    }
    const int64_t wave_v0 = base + (threadIdx.x & ~(kWave - 1));
    while (big) {
      const int l = __ffsll((unsigned long long)big) - 1;
      big &= big - 1;
      const double cnt = WaveIntraCount(g, ids, __shfl(cv, l), __shfl(b, l), __shfl(deg, l));
      if (lane == 0) intra[wave_v0 + l] = cnt;
    }
  }
}

// The rows k_modularity_terms queued, one wave each over the whole grid.
__global__ __launch_bounds__(256) void k_modularity_mid(CsrView g, const uint32_t* __restrict__ ids,
                                                        const uint32_t* __restrict__ midq,
                                                        const unsigned int* __restrict__ nmid,
                                                        double* __restrict__ intra) {
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / kWave);
  const int64_t count = (int64_t)*nmid;
  for (int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; q < count; q += nwaves) {
    const uint32_t v = midq[q];
    const int64_t vb = g.off[v];
    const double cnt = WaveIntraCount(g, ids, ids[v], vb, g.off[v + 1] - vb);
    if (LaneId() == 0) intra[v] = cnt;
  }
}

// Rows longer than kLongRow (intra[v] zeroed beforehand): chunk counts added
// atomically, exact for integer counts below 2^53.
__global__ __launch_bounds__(256) void k_modularity_long(CsrView g, const uint32_t* __restrict__ ids,
                                                         const uint32_t* __restrict__ longq,
                                                         const unsigned int* __restrict__ nlong,
                                                         double* __restrict__ intra) {
  __shared__ unsigned long long su[256 / kWave];
  ForLongChunks(longq, *nlong, g.off, [&](uint32_t, uint32_t v, int64_t e0, int64_t e1) {
    const uint32_t cv = ids[v];
    unsigned long long c = 0;
    for (int64_t p = e0 + threadIdx.x; p < e1; p += blockDim.x) c += ids[g.nbr[p]] == cv ? 1ull : 0ull;
    c = BlockReduce(c, su, [](unsigned long long a, unsigned long long b) { return a + b; });
    if (threadIdx.x == 0 && c) atomicAdd(&intra[v], (double)c);
  });
}

// Rows longer than kLongRow, listed for the long kernels.
__global__ __launch_bounds__(256) void k_list_long_rows(int64_t n, const int64_t* __restrict__ off,
                                                        uint32_t* __restrict__ longq,
                                                        unsigned int* __restrict__ nlong) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    if (off[v + 1] - off[v] > kLongRow) longq[atomicAdd(nlong, 1u)] = (uint32_t)v;
}

__device__ __forceinline__ int GraphOf(const int64_t* gbase, int G, int64_t v) {
  int lo = 0, hi = G - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (gbase[mid] <= v) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// (gamma * d_c * d_c) / total of the graph owning cluster id c.
__global__ __launch_bounds__(256) void k_modularity_penalty(CsrView g, const double* __restrict__ dsum,
                                                            int G, const int64_t* __restrict__ gbase,
                                                            const double* __restrict__ gamma,
                                                            double* __restrict__ t) {
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < g.n;
       c += (int64_t)gridDim.x * blockDim.x) {
    const int gi = GraphOf(gbase, G, c);
    const double total = (double)(g.off[gbase[gi + 1]] - g.off[gbase[gi]]);
    const double d = dsum[c];
    t[c] = (gamma[gi] * d * d) / total;
  }
}

// Per-chunk sums of a[] and b[]: chunk j covers [cs[j], cs[j + 1]) (chunks
// start at their graph's first vertex, so a graph's partials do not depend on
// what else is in the batch).
__global__ __launch_bounds__(256) void k_chunk_sums(const int64_t* __restrict__ cs,
                                                    const double* __restrict__ a,
                                                    const double* __restrict__ b,
                                                    double* __restrict__ pa,
                                                    double* __restrict__ pb) {
  __shared__ double sm[4];
  const int64_t j = blockIdx.x;
  double x = 0.0, y = 0.0;
  for (int64_t v = cs[j] + threadIdx.x; v < cs[j + 1]; v += blockDim.x) {
    x += a[v];
    y += b[v];
  }
  x = BlockReduce(x, sm, [](double p, double q) { return p + q; });
  y = BlockReduce(y, sm, [](double p, double q) { return p + q; });
  if (threadIdx.x == 0) {
    pa[j] = x;
    pb[j] = y;
  }
}

// One workgroup per graph sums a[] and b[] over the graph's vertex range in a
// fixed order.
__global__ __launch_bounds__(256) void k_graph_sums(const double* __restrict__ a,
                                                    const double* __restrict__ b,
                                                    const int64_t* __restrict__ gbase,
                                                    double* __restrict__ sa, double* __restrict__ sb) {
  __shared__ double sm[4];
  const int gi = blockIdx.x;
  double x = 0.0, y = 0.0;
  for (int64_t v = gbase[gi] + threadIdx.x; v < gbase[gi + 1]; v += blockDim.x) {
    x += a[v];
    y += b[v];
  }
  x = BlockReduce(x, sm, [](double p, double q) { return p + q; });
  y = BlockReduce(y, sm, [](double p, double q) { return p + q; });
  if (threadIdx.x == 0) {
    sa[gi] = x;
    sb[gi] = y;
  }
}

__global__ void k_objective_final(int G, const double* __restrict__ esum,
                                  const double* __restrict__ rsum, const double* __restrict__ res,
                                  double* __restrict__ out, const unsigned int* __restrict__ gate) {
  for (int gi = threadIdx.x; gi < G; gi += blockDim.x) {
    double objective = esum[gi];
    objective -= res[gi] * rsum[gi] / 2;
    out[gi] = (gate && *gate) ? __longlong_as_double(0x7FF8000000000000ll) : objective;
  }
}

__global__ void k_modularity_final(CsrView g, int G, const int64_t* __restrict__ gbase,
                                   const double* __restrict__ intra, const double* __restrict__ tsum,
                                   double* __restrict__ out) {
  for (int gi = threadIdx.x; gi < G; gi += blockDim.x) {
    const double total = (double)(g.off[gbase[gi + 1]] - g.off[gbase[gi]]);
    double q = intra[gi];
    q -= tsum[gi];
    out[gi] = q / total;
  }
}

template <typename V>
void SegmentSums(DeviceAllocator* al, const V* vals, const int64_t* start, int64_t nseg,
                 int64_t nvals, V* out, hipStream_t s) {
  if (nseg <= 0) return;
  DevBuf<uint32_t> longq(al, nseg);
  DevBuf<unsigned long long> nlong(al, 1);
  PCC_HIP_CHECK(hipMemsetAsync(nlong.get(), 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL((k_seg_short<V>), dim3(Grid(nseg, 256, 8192)), dim3(256), 0, s, vals, start,
                     nseg, out, longq.get(), nlong.get());
  PCC_HIP_CHECK(hipGetLastError());
  if (nvals <= 1024) return;
  unsigned long long hl = 0;
  ReadBack(&hl, nlong.get(), sizeof(hl), s);
  const int64_t nl = (int64_t)hl;
  if (nl == 0) return;
  DevBuf<uint32_t> nblk(al, nl);
  DevBuf<int64_t> boff(al, nl + 1);
  hipLaunchKernelGGL(k_seg_nblocks, dim3(Grid(nl, 256, 4096)), dim3(256), 0, s, start,
                     longq.get(), nl, nblk.get());
  PCC_HIP_CHECK(hipGetLastError());
  ExclusiveScanU32toI64(al, nblk.get(), boff.get(), nl, boff.get() + nl, s);
  int64_t total = 0;
  ReadBack(&total, boff.get() + nl, sizeof(int64_t), s);
  // Each run has len > 1024 terms and at most ceil(len / 1024) blocks.
  if (total <= 0 || total > nvals / 1024 + nl)
    throw HipError{hipErrorUnknown, "SegmentSums: inconsistent block counts", __FILE__, __LINE__};
  DevBuf<V> part(al, total);
  constexpr int kWavesPerBlock = 256 / kWave;
  hipLaunchKernelGGL((k_seg_blocks<V>), dim3((unsigned)((total + kWavesPerBlock - 1) / kWavesPerBlock)),
                     dim3(256), 0, s, vals, start, longq.get(), boff.get(), nl, part.get());
  hipLaunchKernelGGL((k_seg_top<V>), dim3((unsigned)((nl + kWavesPerBlock - 1) / kWavesPerBlock)),
                     dim3(256), 0, s, longq.get(), boff.get(), nl, part.get(), out);
  PCC_HIP_CHECK(hipGetLastError());
}

}  // namespace

// Segment boundaries of sorted keys: returns the number of segments and
// fills start[0..nseg] (start[nseg] = n).
int64_t Segments(DeviceAllocator* al, const uint64_t* keys, int64_t n, DevBuf<int64_t>* start,
                 hipStream_t s) {
  if (n == 0) {
    start->Reset(al, 1);
    hipLaunchKernelGGL(k_set_i64, dim3(1), dim3(1), 0, s, start->get(), (int64_t)0);
    return 0;
  }
  DevBuf<uint8_t> head(al, n);
  DevBuf<uint32_t> rank(al, n), total(al, 1);
  const int grid = Grid(n, 256, 8192);
  hipLaunchKernelGGL(k_heads, dim3(grid), dim3(256), 0, s, keys, n, head.get());
  ExclusiveScanU8toU32(al, head.get(), rank.get(), n, total.get(), s);
  uint32_t nseg = 0;
  ReadBack(&nseg, total.get(), sizeof(uint32_t), s);
  start->Reset(al, (int64_t)nseg + 1);
  hipLaunchKernelGGL(k_seg_starts, dim3(grid), dim3(256), 0, s, head.get(), rank.get(), n,
                     start->get());
  hipLaunchKernelGGL(k_set_i64, dim3(1), dim3(1), 0, s, start->get() + nseg, n);
  PCC_HIP_CHECK(hipGetLastError());
  return (int64_t)nseg;
}

bool GraphIsSymmetric(DeviceAllocator* al, const CsrView& g, hipStream_t s) {
  if (g.m == 0) return true;
  const int64_t nchunks = (g.m + kEdgeChunk - 1) / kEdgeChunk;
  DevBuf<unsigned long long> acc(al, 1);
  PCC_HIP_CHECK(hipMemsetAsync(acc.get(), 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL(k_sym_fingerprint, dim3(Grid(nchunks, 1, 8192)), dim3(kEdgeThreads), 0, s, g,
                     nchunks, acc.get());
  PCC_HIP_CHECK(hipGetLastError());
  unsigned long long h = 1;
  ReadBack(&h, acc.get(), sizeof(h), s);
  return h == 0;
}

void LaunchGraphMaxIds(int64_t n, const uint32_t* ids, const uint32_t* vgraph, uint32_t* gmax,
                       hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_graph_max, dim3(Grid(n, 256, vgraph ? 8192 : 1024)), dim3(256), 0, s, n, ids, vgraph, gmax);
  PCC_HIP_CHECK(hipGetLastError());
}

void SegmentSumsF32(DeviceAllocator* al, const float* vals, const int64_t* start, int64_t nseg,
                    int64_t nvals, float* out, hipStream_t s) {
  SegmentSums<float>(al, vals, start, nseg, nvals, out, s);
}
void SegmentSumsF64(DeviceAllocator* al, const double* vals, const int64_t* start, int64_t nseg,
                    int64_t nvals, double* out, hipStream_t s) {
  SegmentSums<double>(al, vals, start, nseg, nvals, out, s);
}

void CompressGraph(DeviceAllocator* al, const CsrView& g, const uint32_t* ids, const double* nw,
                   const uint32_t* vgraph, const int64_t* gbase, const int64_t* newbase,
                   int64_t n2, DevCsr* out, DevBuf<double>* out_nw, hipStream_t s, int method,
                   int64_t chunk_terms) {
  if (method == 0 &&
      CompressGraphGrouped(al, g, ids, nw, vgraph, gbase, newbase, n2, out, out_nw, s, chunk_terms))
    return;
  CompressGraphSorted(al, g, ids, nw, vgraph, gbase, newbase, n2, out, out_nw, s);
}

void CompressGraphSorted(DeviceAllocator* al, const CsrView& g, const uint32_t* ids,
                         const double* nw, const uint32_t* vgraph, const int64_t* gbase,
                         const int64_t* newbase, int64_t n2, DevCsr* out,
                         DevBuf<double>* out_nw, hipStream_t s) {
  const int64_t n = g.n;
  const GraphMap gm{vgraph, gbase, newbase};
  const int vbits = BitsFor((uint64_t)(n2 > 1 ? n2 - 1 : 1));

  // Kept edges in CSR order.
  const int64_t nchunks = (g.m + kEdgeChunk - 1) / kEdgeChunk;
  int64_t mk = 0;
  DevBuf<uint64_t> keys;
  DevBuf<float> vals;
  if (nchunks > 0) {
    DevBuf<uint32_t> ccnt(al, nchunks);
    DevBuf<int64_t> coff(al, nchunks + 1);
    const int grid = (int)std::min<int64_t>(nchunks, 1 << 16);
    hipLaunchKernelGGL(k_count_kept, dim3(grid), dim3(kEdgeThreads), 0, s, g, ids, gm, nchunks,
                       ccnt.get());
    PCC_HIP_CHECK(hipGetLastError());
    ExclusiveScanU32toI64(al, ccnt.get(), coff.get(), nchunks, coff.get() + nchunks, s);
    ReadBack(&mk, coff.get() + nchunks, sizeof(int64_t), s);
    keys.Reset(al, mk);
    vals.Reset(al, mk);
    if (mk > 0) {
      hipLaunchKernelGGL(k_emit_kept, dim3(grid), dim3(kEdgeThreads), 0, s, g, ids, gm, nchunks,
                         coff.get(), vbits, keys.get(), vals.get());
      PCC_HIP_CHECK(hipGetLastError());
      RadixSortPairsF32(al, keys.get(), vals.get(), mk, 0, 2 * vbits, s);
    }
  }

  // Runs of equal (c(u), c(v)) -> compressed edges.
  DevBuf<int64_t> start;
  const int64_t m2 = Segments(al, keys.get(), mk, &start, s);
  out->n = n2;
  out->m = m2;
  out->off.Reset(al, n2 + 1);
  out->nbr.Reset(al, m2 > 0 ? m2 : 1);
  out->w.Reset(al, m2 > 0 ? m2 : 1);
  DevBuf<uint32_t> cnt(al, n2);
  PCC_HIP_CHECK(hipMemsetAsync(cnt.get(), 0, sizeof(uint32_t) * (size_t)n2, s));
  if (m2 > 0) {
    SegmentSums<float>(al, vals.get(), start.get(), m2, mk, out->w.get(), s);
    hipLaunchKernelGGL(k_emit_compressed, dim3(Grid(m2, 256, 8192)), dim3(256), 0, s, keys.get(),
                       start.get(), m2, vbits, out->nbr.get(), cnt.get());
    PCC_HIP_CHECK(hipGetLastError());
  }
  ExclusiveScanU32toI64(al, cnt.get(), out->off.get(), n2, out->off.get() + n2, s);
  int64_t mend = -1;
  ReadBack(&mend, out->off.get() + n2, sizeof(int64_t), s);
  if (mend != m2)
    throw HipError{hipErrorUnknown, "CompressGraph: contracted offsets do not cover the edges",
                   __FILE__, __LINE__};
  keys.Release();
  vals.Release();
  start.Release();

  // Node weights: stable sort of (id, weight) by id, ordered sums.
  DevBuf<uint64_t> nkeys(al, n);
  DevBuf<double> nvals(al, n);
  hipLaunchKernelGGL(k_node_keys, dim3(Grid(n, 256, 8192)), dim3(256), 0, s, n, ids, gm, n2, nw,
                     nkeys.get(), nvals.get());
  PCC_HIP_CHECK(hipGetLastError());
  RadixSortPairsF64(al, nkeys.get(), nvals.get(), n, 0, BitsFor((uint64_t)n2), s);
  DevBuf<int64_t> nstart;
  const int64_t ng = Segments(al, nkeys.get(), n, &nstart, s);
  DevBuf<double> gsum(al, ng > 0 ? ng : 1);
  SegmentSums<double>(al, nvals.get(), nstart.get(), ng, n, gsum.get(), s);
  out_nw->Reset(al, n2);
  PCC_HIP_CHECK(hipMemsetAsync(out_nw->get(), 0, sizeof(double) * (size_t)n2, s));
  if (ng > 0) {
    hipLaunchKernelGGL(k_scatter_node_weights, dim3(Grid(ng, 256, 8192)), dim3(256), 0, s,
                       nkeys.get(), nstart.get(), gsum.get(), ng, n2, out_nw->get());
    PCC_HIP_CHECK(hipGetLastError());
  }
}

void LaunchFlattenShift(int64_t n, uint32_t* ids, const uint32_t* vgraph0, const uint8_t* sel,
                        const int64_t* shift, const uint32_t* next_ids, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_flatten_shift, dim3(Grid(n, 256, 8192)), dim3(256), 0, s, n, ids, vgraph0,
                     sel, shift, next_ids);
  PCC_HIP_CHECK(hipGetLastError());
}

void ResetClusteringBatch(DeviceAllocator* al, int64_t n, int64_t n_next, const uint32_t* next_ids,
                          const uint32_t* vgraph, const int64_t* gbase, const int64_t* nbase,
                          const uint8_t* sel, const StateView& st, hipStream_t s) {
  if (n <= 0) return;
  DevBuf<uint8_t> present(al, n_next > 0 ? n_next : 1);
  DevBuf<uint32_t> prank(al, n_next > 0 ? n_next : 1), flat(al, n);
  PCC_HIP_CHECK(hipMemsetAsync(present.get(), 0, (size_t)(n_next > 0 ? n_next : 1), s));
  const int grid = Grid(n, 256, 8192);
  hipLaunchKernelGGL(k_batch_flat, dim3(grid), dim3(256), 0, s, n, st.ids, next_ids, vgraph, gbase,
                     nbase, sel, flat.get(), present.get());
  ExclusiveScanU8toU32(al, present.get(), prank.get(), n_next, nullptr, s);
  hipLaunchKernelGGL(k_batch_clear, dim3(grid), dim3(256), 0, s, n, vgraph, sel, st);
  hipLaunchKernelGGL(k_batch_assign, dim3(grid), dim3(256), 0, s, n, flat.get(), prank.get(),
                     vgraph, gbase, nbase, sel, st);
  PCC_HIP_CHECK(hipGetLastError());
}

void LaunchFlatten(int64_t n, const uint32_t* ids, const uint32_t* compressed, uint32_t* out,
                   hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_flatten, dim3(Grid(n, 256, 8192)), dim3(256), 0, s, n, ids, compressed, out);
  PCC_HIP_CHECK(hipGetLastError());
}

void ResetClusteringFromIds(DeviceAllocator* al, int64_t n, const uint32_t* flat,
                            const StateView& st, hipStream_t s) {
  if (n == 0) return;
  DevBuf<uint8_t> present(al, n);
  DevBuf<uint32_t> rank(al, n);
  PCC_HIP_CHECK(hipMemsetAsync(present.get(), 0, (size_t)n, s));
  const int grid = Grid(n, 256, 8192);
  hipLaunchKernelGGL(k_mark_present, dim3(grid), dim3(256), 0, s, n, flat, present.get());
  ExclusiveScanU8toU32(al, present.get(), rank.get(), n, nullptr, s);
  hipLaunchKernelGGL(k_clear_state, dim3(grid), dim3(256), 0, s, n, st);
  hipLaunchKernelGGL(k_assign_ranked, dim3(grid), dim3(256), 0, s, n, flat, rank.get(), st);
  PCC_HIP_CHECK(hipGetLastError());
}

void LaunchSingletons(int64_t n, const StateView& st, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_singletons, dim3(Grid(n, 256, 8192)), dim3(256), 0, s, n, st);
  PCC_HIP_CHECK(hipGetLastError());
}

void LaunchSetClustering(int64_t n, const uint32_t* init, const StateView& st, hipStream_t s) {
  if (n == 0) return;
  const int grid = Grid(n, 256, 8192);
  hipLaunchKernelGGL(k_clear_state, dim3(grid), dim3(256), 0, s, n, st);
  hipLaunchKernelGGL(k_assign_init, dim3(grid), dim3(256), 0, s, n, init, st);
  PCC_HIP_CHECK(hipGetLastError());
}

void LaunchDegrees(const CsrView& g, double* nw, hipStream_t s) {
  if (g.n == 0) return;
  hipLaunchKernelGGL(k_degrees, dim3(Grid(g.n, 256, 8192)), dim3(256), 0, s, g, nw);
  PCC_HIP_CHECK(hipGetLastError());
}

// Per-graph sums of a[] and b[] in a fixed order that depends only on the
// graph: chunks of 4096 vertices from the graph's first vertex, summed per
// workgroup, then the chunk sums per graph.  A graph in a batch therefore gets
// the same bits as the same graph alone.
void GraphSums(DeviceAllocator* al, const double* a, const double* b, int G,
               const int64_t* gbase, double* sa, double* sb, hipStream_t s) {
  constexpr int64_t kChunk = 4096;
  std::vector<int64_t> hb((size_t)G + 1);
  ReadBack(hb.data(), gbase, sizeof(int64_t) * hb.size(), s);
  // Chunk starts; graphs are contiguous, so each chunk ends where the next
  // one (or the last graph) ends.
  std::vector<int64_t> cs, coff((size_t)G + 1);
  for (int gi = 0; gi < G; ++gi) {
    coff[(size_t)gi] = (int64_t)cs.size();
    for (int64_t v = hb[(size_t)gi]; v < hb[(size_t)gi + 1]; v += kChunk) cs.push_back(v);
  }
  coff[(size_t)G] = (int64_t)cs.size();
  const int64_t nch = (int64_t)cs.size();
  cs.push_back(hb[(size_t)G]);
  DevBuf<int64_t> dcs(al, nch + 1), dcoff(al, (int64_t)G + 1);
  DevBuf<double> pa(al, nch > 0 ? nch : 1), pb(al, nch > 0 ? nch : 1);
  PCC_HIP_CHECK(hipMemcpyAsync(dcs.get(), cs.data(), sizeof(int64_t) * cs.size(),
                               hipMemcpyHostToDevice, s));
  PCC_HIP_CHECK(hipMemcpyAsync(dcoff.get(), coff.data(), sizeof(int64_t) * coff.size(),
                               hipMemcpyHostToDevice, s));
  if (nch > 0)
    hipLaunchKernelGGL(k_chunk_sums, dim3((unsigned)nch), dim3(256), 0, s, dcs.get(), a, b,
                       pa.get(), pb.get());
  hipLaunchKernelGGL(k_graph_sums, dim3(G), dim3(256), 0, s, pa.get(), pb.get(), dcoff.get(), sa,
                     sb);
  PCC_HIP_CHECK(hipGetLastError());
  PCC_HIP_CHECK(hipStreamSynchronize(s));  // cs / coff stay alive until the copies ran
}

void ComputeObjective(DeviceAllocator* al, const CsrView& g, const uint32_t* ids, const double* nw,
                      int G, const int64_t* gbase, const double* res, double offset,
                      double* out_dev, hipStream_t s) {
  const int64_t n = g.n;
  const int64_t nn = n > 0 ? n : 1;
  DevBuf<double> K(al, nn), e(al, nn), r(al, nn), se(al, G), sr(al, G);
  PCC_HIP_CHECK(hipMemsetAsync(K.get(), 0, sizeof(double) * (size_t)nn, s));
  PCC_HIP_CHECK(hipMemsetAsync(e.get(), 0, sizeof(double) * (size_t)nn, s));
  PCC_HIP_CHECK(hipMemsetAsync(r.get(), 0, sizeof(double) * (size_t)nn, s));
  if (n > 0) {
    const int64_t maxlong = g.m / kLongRow + 1;
    DevBuf<uint32_t> longq(al, maxlong);
    DevBuf<unsigned int> nlong(al, 1);
    DevBuf<double> acc(al, 2 * maxlong);
    DevBuf<int> gmin(al, maxlong);
    DevBuf<unsigned long long> nz(al, maxlong);
    PCC_HIP_CHECK(hipMemsetAsync(nlong.get(), 0, sizeof(unsigned int), s));
    PCC_HIP_CHECK(hipMemsetAsync(acc.get(), 0, sizeof(double) * 2 * (size_t)maxlong, s));
    PCC_HIP_CHECK(hipMemsetD32Async((hipDeviceptr_t)gmin.get(), 1 << 20, (size_t)maxlong, s));
    PCC_HIP_CHECK(hipMemsetAsync(nz.get(), 0, sizeof(unsigned long long) * (size_t)maxlong, s));
    const int grid = Grid(n, 256, 8192);
    hipLaunchKernelGGL(k_cluster_weights, dim3(grid), dim3(256), 0, s, n, ids, nw, K.get(),
                       (const unsigned int*)nullptr);
    // Large graphs queue their rows of kObjThreadDeg..kLongRow edges for
    // k_objective_mid (the same per-row computation, spread over the grid).
    const bool mid = n > kObjQueueN;
    DevBuf<uint32_t> midq;
    DevBuf<unsigned int> nmid;
    if (mid) {
      midq.Reset(al, n);
      nmid.Reset(al, 1);
      PCC_HIP_CHECK(hipMemsetAsync(nmid.get(), 0, sizeof(unsigned int), s));
    }
    hipLaunchKernelGGL(k_objective_terms, dim3(grid), dim3(256), 0, s, g, ids, nw, K.get(), offset,
                       e.get(), r.get(), longq.get(), nlong.get(), (const unsigned int*)nullptr,
                       midq.get(), nmid.get());
    if (mid)
      hipLaunchKernelGGL(k_objective_mid, dim3(2048), dim3(256), 0, s, g, ids, offset, midq.get(),
                         nmid.get(), e.get());
    if (g.m > kLongRow) {
      hipLaunchKernelGGL(k_objective_long, dim3(1024), dim3(256), 0, s, g, ids, offset, longq.get(),
                         nlong.get(), acc.get(), gmin.get(), nz.get());
      hipLaunchKernelGGL(k_objective_long_final, dim3(64), dim3(256), 0, s, g, ids, offset,
                         longq.get(), nlong.get(), acc.get(), gmin.get(), nz.get(), e.get());
    }
    PCC_HIP_CHECK(hipGetLastError());
  }
  GraphSums(al, e.get(), r.get(), G, gbase, se.get(), sr.get(), s);
  hipLaunchKernelGGL(k_objective_final, dim3(1), dim3(256), 0, s, G, se.get(), sr.get(), res,
                     out_dev, (const unsigned int*)nullptr);
  PCC_HIP_CHECK(hipGetLastError());
}

void ComputeModularity(DeviceAllocator* al, const CsrView& g, const uint32_t* ids, int G,
                       const int64_t* gbase, const double* gamma, double* out_dev, hipStream_t s) {
  const int64_t n = g.n;
  const int64_t nn = n > 0 ? n : 1;
  DevBuf<double> dsum(al, nn), t(al, nn), intra(al, nn), si(al, G), st(al, G);
  PCC_HIP_CHECK(hipMemsetAsync(dsum.get(), 0, sizeof(double) * (size_t)nn, s));
  PCC_HIP_CHECK(hipMemsetAsync(t.get(), 0, sizeof(double) * (size_t)nn, s));
  PCC_HIP_CHECK(hipMemsetAsync(intra.get(), 0, sizeof(double) * (size_t)nn, s));
  if (n > 0) {
    const int grid = Grid(n, 256, 8192);
    const bool mid = n > kObjQueueN;
    DevBuf<uint32_t> midq;
    DevBuf<unsigned int> nmid;
    if (mid) {
      midq.Reset(al, n);
      nmid.Reset(al, 1);
      PCC_HIP_CHECK(hipMemsetAsync(nmid.get(), 0, sizeof(unsigned int), s));
    }
    hipLaunchKernelGGL(k_modularity_terms, dim3(grid), dim3(256), 0, s, g, ids, intra.get(),
                       dsum.get(), midq.get(), nmid.get());
    if (mid)
      hipLaunchKernelGGL(k_modularity_mid, dim3(2048), dim3(256), 0, s, g, ids, midq.get(),
                         nmid.get(), intra.get());
    if (g.m > kLongRow) {
      const int64_t maxlong = g.m / kLongRow + 1;
      DevBuf<uint32_t> longq(al, maxlong);
      DevBuf<unsigned int> nlong(al, 1);
      PCC_HIP_CHECK(hipMemsetAsync(nlong.get(), 0, sizeof(unsigned int), s));
      hipLaunchKernelGGL(k_list_long_rows, dim3(grid), dim3(256), 0, s, n, g.off, longq.get(),
                         nlong.get());
      hipLaunchKernelGGL(k_modularity_long, dim3(1024), dim3(256), 0, s, g, ids, longq.get(),
                         nlong.get(), intra.get());
    }
    hipLaunchKernelGGL(k_modularity_penalty, dim3(grid), dim3(256), 0, s, g, dsum.get(), G, gbase,
                       gamma, t.get());
    PCC_HIP_CHECK(hipGetLastError());
  }
  GraphSums(al, intra.get(), t.get(), G, gbase, si.get(), st.get(), s);
  hipLaunchKernelGGL(k_modularity_final, dim3(1), dim3(256), 0, s, g, G, gbase, si.get(), st.get(),
                     out_dev);
  PCC_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------- allocation-free objective --
// ComputeObjective for one graph with every work array carved out of a fixed
// per-device array instead of allocations: callers that cannot allocate
// (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc) get the same kernels, in the same order, as the allocating
// path, so the same value bit for bit.  Work memory per call: the cluster
// weights and the two per-vertex term arrays (24 B per vertex), the chunk
// partial sums of the fixed-order reduction, and the long-row accumulators
// (32 B per row of more than kLongRow edges).

namespace {

constexpr int64_t kSumChunk = 4096;  // GraphSums' chunk of vertices
constexpr int64_t kNaMaxChunks = (kNoAllocMaxVertices + kSumChunk - 1) / kSumChunk;
constexpr int64_t kNaMaxLong = kNoAllocMaxEdges / kLongRow + 1;

constexpr size_t Align256(size_t x) { return (x + 255) / 256 * 256; }

// Byte offsets of the work arrays inside the scratch array.
constexpr size_t kNaK = 0;
constexpr size_t kNaE = kNaK + Align256(8 * (size_t)kNoAllocMaxVertices);
constexpr size_t kNaR = kNaE + Align256(8 * (size_t)kNoAllocMaxVertices);
constexpr size_t kNaPa = kNaR + Align256(8 * (size_t)kNoAllocMaxVertices);
constexpr size_t kNaPb = kNaPa + Align256(8 * (size_t)kNaMaxChunks);
constexpr size_t kNaCs = kNaPb + Align256(8 * (size_t)kNaMaxChunks);
constexpr size_t kNaSmall = kNaCs + Align256(8 * (size_t)(kNaMaxChunks + 1));
// small: coff[2] (int64), se, sr, res (double), nlong, bad (uint32)
constexpr size_t kNaLongq = kNaSmall + 256;
constexpr size_t kNaAcc = kNaLongq + Align256(4 * (size_t)kNaMaxLong);
constexpr size_t kNaGmin = kNaAcc + Align256(16 * (size_t)kNaMaxLong);
constexpr size_t kNaNz = kNaGmin + Align256(4 * (size_t)kNaMaxLong);
constexpr size_t kNaBytes = kNaNz + Align256(8 * (size_t)kNaMaxLong);

}  // namespace

}  // namespace pcc

// The per-device work memory of pcc_compute_objective_noalloc (about 50 MiB;
// every device the library runs on keeps its copy until the process ends).
// Host code must not pass this name to kernels: it stands for the symbol, not
// for a device address.  NoAllocScratch() returns the address.
static __device__ __align__(256) unsigned char scratch[pcc::kNaBytes];

namespace pcc {

namespace {

std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};

struct NaWork {
  double* K;
  double* e;
  double* r;
  double* pa;
  double* pb;
  int64_t* cs;
  int64_t* coff;
  double* se;
  double* sr;
  double* res;
  unsigned int* nlong;
  unsigned int* bad;
  uint32_t* longq;
  double* acc;
  int* gmin;
  unsigned long long* nz;
};

NaWork Carve(void* base) {
  unsigned char* b = static_cast<unsigned char*>(base);
  NaWork w;
  w.K = reinterpret_cast<double*>(b + kNaK);
  w.e = reinterpret_cast<double*>(b + kNaE);
  w.r = reinterpret_cast<double*>(b + kNaR);
  w.pa = reinterpret_cast<double*>(b + kNaPa);
  w.pb = reinterpret_cast<double*>(b + kNaPb);
  w.cs = reinterpret_cast<int64_t*>(b + kNaCs);
  w.coff = reinterpret_cast<int64_t*>(b + kNaSmall);
  w.se = reinterpret_cast<double*>(b + kNaSmall + 16);
  w.sr = reinterpret_cast<double*>(b + kNaSmall + 24);
  w.res = reinterpret_cast<double*>(b + kNaSmall + 32);
  w.nlong = reinterpret_cast<unsigned int*>(b + kNaSmall + 40);
  w.bad = reinterpret_cast<unsigned int*>(b + kNaSmall + 44);
  w.longq = reinterpret_cast<uint32_t*>(b + kNaLongq);
  w.acc = reinterpret_cast<double*>(b + kNaAcc);
  w.gmin = reinterpret_cast<int*>(b + kNaGmin);
  w.nz = reinterpret_cast<unsigned long long*>(b + kNaNz);
  return w;
}

// Clears what the kernels accumulate into (as the allocating path's memsets
// do, but as a kernel: memset nodes are not safe in captured graphs) and
// writes the chunk starts of the single graph's fixed-order sum.
__global__ __launch_bounds__(256) void k_na_prepare(int64_t n, int64_t nlongmax, int64_t nch,
                                                    double res, NaWork w) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = tid; v < n; v += stride) {
    w.K[v] = 0.0;
    w.e[v] = 0.0;
    w.r[v] = 0.0;
  }
  for (int64_t q = tid; q < nlongmax; q += stride) {
    w.acc[2 * q] = 0.0;
    w.acc[2 * q + 1] = 0.0;
    w.gmin[q] = 1 << 20;
    w.nz[q] = 0;
  }
  for (int64_t j = tid; j <= nch; j += stride) w.cs[j] = j < nch ? j * kSumChunk : n;
  if (tid == 0) {
    w.coff[0] = 0;
    w.coff[1] = nch;
    w.se[0] = 0.0;
    w.sr[0] = 0.0;
    w.res[0] = res;
    *w.nlong = 0;
    *w.bad = 0;
  }
}

// bad = 1 if a cluster id is outside [0, n) (the kernels that index by
// cluster id then do nothing and the objective is NaN).
__global__ __launch_bounds__(256) void k_na_check_ids(int64_t n, const uint32_t* __restrict__ ids,
                                                      unsigned int* __restrict__ bad) {
  bool out = false;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    out = out || (int64_t)ids[v] >= n;
  if (__ballot(out) && LaneId() == 0) atomicOr(bad, 1u);
}

}  // namespace

void* NoAllocScratch(int device) {
  if (device < 0 || device >= kMaxDevices)
    throw HipError{hipErrorInvalidDevice, "device index beyond the scratch table", __FILE__,
                   __LINE__};
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    // With `device` current: each device holds its own copy of the array.
    void* p = nullptr;
    PCC_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

void ComputeObjectiveNoAlloc(void* scratch_dev, const CsrView& g, const uint32_t* ids,
                             const double* nw, double res, double offset, double* out_dev,
                             hipStream_t s) {
  const int64_t n = g.n;
  const NaWork w = Carve(scratch_dev);
  const int64_t nch = (n + kSumChunk - 1) / kSumChunk;
  const int64_t nlongmax = g.m / kLongRow + 1;
  const int64_t prep = std::max<int64_t>(std::max<int64_t>(n, nlongmax), nch + 1);
  hipLaunchKernelGGL(k_na_prepare, dim3(Grid(prep, 256, 8192)), dim3(256), 0, s, n, nlongmax, nch,
                     res, w);
  if (n > 0) {
    const int grid = Grid(n, 256, 8192);
    hipLaunchKernelGGL(k_na_check_ids, dim3(grid), dim3(256), 0, s, n, ids, w.bad);
    hipLaunchKernelGGL(k_cluster_weights, dim3(grid), dim3(256), 0, s, n, ids, nw, w.K,
                       (const unsigned int*)w.bad);
    hipLaunchKernelGGL(k_objective_terms, dim3(grid), dim3(256), 0, s, g, ids, nw, w.K, offset,
                       w.e, w.r, w.longq, w.nlong, (const unsigned int*)w.bad, (uint32_t*)nullptr,
                       (unsigned int*)nullptr);  // each wave does its own rows
    if (g.m > kLongRow) {
      hipLaunchKernelGGL(k_objective_long, dim3(1024), dim3(256), 0, s, g, ids, offset, w.longq,
                         w.nlong, w.acc, w.gmin, w.nz);
      hipLaunchKernelGGL(k_objective_long_final, dim3(64), dim3(256), 0, s, g, ids, offset,
                         w.longq, w.nlong, w.acc, w.gmin, w.nz, w.e);
    }
    hipLaunchKernelGGL(k_chunk_sums, dim3((unsigned)nch), dim3(256), 0, s, w.cs, w.e, w.r, w.pa,
                       w.pb);
  }
  hipLaunchKernelGGL(k_graph_sums, dim3(1), dim3(256), 0, s, w.pa, w.pb, w.coff, w.se, w.sr);
  hipLaunchKernelGGL(k_objective_final, dim3(1), dim3(256), 0, s, 1, w.se, w.sr, w.res, out_dev,
                     (const unsigned int*)w.bad);
  PCC_HIP_CHECK(hipGetLastError());
}

}  // namespace pcc
