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

#include <cmath>

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

// 16 lanes per vertex count the kept edges (PGU.cc:93-103).
__global__ __launch_bounds__(256) void k_count_kept(CsrView g, const uint32_t* __restrict__ ids,
                                                    GraphMap gm, uint32_t* __restrict__ kept) {
  const int sub = threadIdx.x & 15;
  const int64_t stride = (int64_t)gridDim.x * (blockDim.x / 16);
  for (int64_t u = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 16; u < g.n; u += stride) {
    const uint32_t cu = ids[u];
    uint32_t c = 0;
    const int64_t lim = gm.Shift(u) == INT64_MIN ? g.off[u] : g.off[u + 1];
    for (int64_t p = g.off[u] + sub; p < lim; p += 16) {
      const uint32_t v = g.nbr[p];
      if ((int64_t)v <= u || ids[v] != cu) ++c;
    }
    for (int s = 8; s > 0; s >>= 1) c += __shfl_xor(c, s, 16);
    if (sub == 0) kept[u] = c;
  }
}

// Writes the kept edges of each vertex, in CSR order, as (key, weight).
__global__ __launch_bounds__(256) void k_emit_kept(CsrView g, const uint32_t* __restrict__ ids,
                                                   GraphMap gm,
                                                   const int64_t* __restrict__ kept_off, int vbits,
                                                   uint64_t* __restrict__ keys,
                                                   float* __restrict__ vals) {
  const int sub = threadIdx.x & 15;
  const int wl = LaneId();
  const uint64_t gmask = 0xFFFFull << (wl & ~15);
  const uint64_t lt = wl == 0 ? 0ull : (~0ull >> (64 - wl));
  const int64_t stride = (int64_t)gridDim.x * (blockDim.x / 16);
  for (int64_t ub = ((int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63)) / 16; ub < g.n;
       ub += stride) {
    const int64_t u = ub + (wl >> 4);
    const int64_t shift = u < g.n ? gm.Shift(u) : INT64_MIN;
    const bool active = u < g.n && shift != INT64_MIN;
    const int64_t b = active ? g.off[u] : 0, e = active ? g.off[u + 1] : 0;
    const uint32_t cu = active ? ids[u] : 0u;
    int64_t pos = active ? kept_off[u] : 0;
    const int64_t len = e - b;
    int64_t maxlen = len;
    for (int s = 32; s > 0; s >>= 1) maxlen = max(maxlen, (int64_t)__shfl_xor(maxlen, s));
    for (int64_t c0 = 0; c0 < maxlen; c0 += 16) {
      const int64_t p = b + c0 + sub;
      bool keep = false;
      uint32_t v = 0, cv = 0;
      float w = 0.0f;
      if (active && p < e) {
        v = g.nbr[p];
        cv = ids[v];
        w = g.w[p];
        keep = (int64_t)v <= u || cv != cu;
      }
      const uint64_t bal = __ballot(keep) & gmask;
      if (keep) {
        const int64_t q = pos + __popcll(bal & lt);
        keys[q] = ((uint64_t)(cu + shift) << vbits) | (uint64_t)(cv + shift);
        vals[q] = w;
      }
      pos += __popcll(bal);
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

// Segment sums, pbbslib reduce order.  Runs of at most 1024 terms are folded
// left by one thread; longer runs are queued for k_seg_long.
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

template <typename V>
__global__ __launch_bounds__(64) void k_seg_long(const V* __restrict__ vals,
                                                 const int64_t* __restrict__ start,
                                                 const uint32_t* __restrict__ longq,
                                                 const unsigned long long* __restrict__ nlong,
                                                 V* __restrict__ out, V* __restrict__ scratch,
                                                 int64_t cap) {
  const int lane = threadIdx.x;
  V* bufa = scratch + (size_t)blockIdx.x * cap;
  V* bufb = bufa + cap / 2;
  const int64_t count = (int64_t)*nlong;
  for (int64_t q = blockIdx.x; q < count; q += gridDim.x) {
    const uint32_t sgi = longq[q];
    const V* src = vals + start[sgi];
    int64_t len = start[sgi + 1] - start[sgi];
    V* dst = bufa;
    while (true) {
      int64_t bs = (int64_t)ceil(sqrt((double)len)) * 4;
      if (bs < 1024) bs = 1024;
      const int64_t l = (len + bs - 1) / bs;
      if (l <= 1) {
        if (lane == 0) {
          V r = src[0];
          for (int64_t i = 1; i < len; ++i) r += src[i];
          out[sgi] = r;
        }
        break;
      }
      for (int64_t blk = lane; blk < l; blk += 64) {
        const int64_t s = blk * bs;
        const int64_t e = (s + bs < len) ? s + bs : len;
        V r = src[s];
        for (int64_t i = s + 1; i < e; ++i) r += src[i];
        dst[blk] = r;
      }
      __syncthreads();
      src = dst;
      len = l;
      dst = (dst == bufa) ? bufb : bufa;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_emit_compressed(const uint64_t* __restrict__ keys,
                                                         const int64_t* __restrict__ start,
                                                         int64_t nseg, int vbits,
                                                         uint32_t* __restrict__ nbr,
                                                         uint32_t* __restrict__ cnt) {
  const uint64_t vmask = (1ull << vbits) - 1ull;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nseg;
       s += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[start[s]];
    nbr[s] = (uint32_t)(k & vmask);
    atomicAdd(&cnt[k >> vbits], 1u);
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
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    atomicMax(&gmax[vgraph ? vgraph[v] : 0], ids[v]);
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

__global__ __launch_bounds__(256) void k_cluster_weights(int64_t n, const uint32_t* __restrict__ ids,
                                                         const double* __restrict__ nw,
                                                         double* __restrict__ K) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&K[ids[v]], nw[v]);
}

// Per-vertex terms of ComputeObjective (PCCI.cc:124-145): the intra-cluster
// half weights folded in pbbs order, and k_v (K_c(v) - k_v).
__global__ __launch_bounds__(256) void k_objective_terms(CsrView g, const uint32_t* __restrict__ ids,
                                                         const double* __restrict__ nw,
                                                         const double* __restrict__ K,
                                                         double offset, double* __restrict__ e,
                                                         double* __restrict__ r) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t cv = ids[v];
    const int64_t b = g.off[v], deg = g.off[v + 1] - b;
    double tot = 0.0;
    if (deg > 0) {
      int64_t bs = (int64_t)ceil(sqrt((double)deg)) * 4;
      if (bs < 1024) bs = 1024;
      bool have = false;
      for (int64_t s = 0; s < deg; s += bs) {
        const int64_t en = (s + bs < deg) ? s + bs : deg;
        double blk = 0.0;
        for (int64_t p = s; p < en; ++p) {
          const float w = g.w[b + p];
          const double x = ids[g.nbr[b + p]] == cv ? (w - offset) / 2 : 0.0;
          blk = (p == s) ? x : blk + x;
        }
        tot = have ? tot + blk : blk;
        have = true;
      }
    }
    e[v] = tot;
    r[v] = nw[v] * (K[cv] - nw[v]);
  }
}

// Per-vertex intra-cluster edge count and per-cluster degree sums
// (PMC.cc:43-61).
__global__ __launch_bounds__(256) void k_modularity_terms(CsrView g, const uint32_t* __restrict__ ids,
                                                          double* __restrict__ intra,
                                                          double* __restrict__ dsum) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t cv = ids[v];
    int64_t cnt = 0;
    for (int64_t p = g.off[v]; p < g.off[v + 1]; ++p) cnt += ids[g.nbr[p]] == cv ? 1 : 0;
    intra[v] = (double)cnt;
    atomicAdd(&dsum[cv], (double)(g.off[v + 1] - g.off[v]));
  }
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
                                  double* __restrict__ out) {
  for (int gi = threadIdx.x; gi < G; gi += blockDim.x) {
    double objective = esum[gi];
    objective -= res[gi] * rsum[gi] / 2;
    out[gi] = objective;
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
  if (nvals > 1024) {
    constexpr int kRegions = 16;
    const int64_t cap = 2 * (nvals / 1024 + 2);
    DevBuf<V> scratch(al, cap * kRegions);
    hipLaunchKernelGGL((k_seg_long<V>), dim3(kRegions), dim3(64), 0, s, vals, start, longq.get(),
                       nlong.get(), out, scratch.get(), cap);
    PCC_HIP_CHECK(hipGetLastError());
  }
}

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
  PCC_HIP_CHECK(hipMemcpyAsync(&nseg, total.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  PCC_HIP_CHECK(hipStreamSynchronize(s));
  start->Reset(al, (int64_t)nseg + 1);
  hipLaunchKernelGGL(k_seg_starts, dim3(grid), dim3(256), 0, s, head.get(), rank.get(), n,
                     start->get());
  hipLaunchKernelGGL(k_set_i64, dim3(1), dim3(1), 0, s, start->get() + nseg, n);
  PCC_HIP_CHECK(hipGetLastError());
  return (int64_t)nseg;
}

}  // namespace

void LaunchGraphMaxIds(int64_t n, const uint32_t* ids, const uint32_t* vgraph, uint32_t* gmax,
                       hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_graph_max, dim3(Grid(n, 256, 8192)), dim3(256), 0, s, n, ids, vgraph, gmax);
  PCC_HIP_CHECK(hipGetLastError());
}

void CompressGraph(DeviceAllocator* al, const CsrView& g, const uint32_t* ids, const double* nw,
                   const uint32_t* vgraph, const int64_t* gbase, const int64_t* newbase,
                   int64_t n2, DevCsr* out, DevBuf<double>* out_nw, hipStream_t s) {
  const int64_t n = g.n;
  const GraphMap gm{vgraph, gbase, newbase};
  const int vbits = BitsFor((uint64_t)(n2 > 1 ? n2 - 1 : 1));

  // Kept edges in CSR order.
  DevBuf<uint32_t> kept(al, n);
  DevBuf<int64_t> kept_off(al, n + 1);
  hipLaunchKernelGGL(k_count_kept, dim3(Grid(n, 16, 8192)), dim3(256), 0, s, g, ids, gm, kept.get());
  ExclusiveScanU32toI64(al, kept.get(), kept_off.get(), n, kept_off.get() + n, s);
  int64_t mk = 0;
  PCC_HIP_CHECK(hipMemcpyAsync(&mk, kept_off.get() + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  PCC_HIP_CHECK(hipStreamSynchronize(s));
  DevBuf<uint64_t> keys(al, mk);
  DevBuf<float> vals(al, mk);
  if (mk > 0) {
    hipLaunchKernelGGL(k_emit_kept, dim3(Grid(n, 16, 8192)), dim3(256), 0, s, g, ids, gm,
                       kept_off.get(), vbits, keys.get(), vals.get());
    PCC_HIP_CHECK(hipGetLastError());
    RadixSortPairsF32(al, keys.get(), vals.get(), mk, 0, 2 * vbits, s);
  }
  kept.Release();
  kept_off.Release();

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
    const int grid = Grid(n, 256, 8192);
    hipLaunchKernelGGL(k_cluster_weights, dim3(grid), dim3(256), 0, s, n, ids, nw, K.get());
    hipLaunchKernelGGL(k_objective_terms, dim3(grid), dim3(256), 0, s, g, ids, nw, K.get(), offset,
                       e.get(), r.get());
    PCC_HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_graph_sums, dim3(G), dim3(256), 0, s, e.get(), r.get(), gbase, se.get(),
                     sr.get());
  hipLaunchKernelGGL(k_objective_final, dim3(1), dim3(256), 0, s, G, se.get(), sr.get(), res,
                     out_dev);
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
    hipLaunchKernelGGL(k_modularity_terms, dim3(grid), dim3(256), 0, s, g, ids, intra.get(),
                       dsum.get());
    hipLaunchKernelGGL(k_modularity_penalty, dim3(grid), dim3(256), 0, s, g, dsum.get(), G, gbase,
                       gamma, t.get());
    PCC_HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_graph_sums, dim3(G), dim3(256), 0, s, intra.get(), t.get(), gbase, si.get(),
                     st.get());
  hipLaunchKernelGGL(k_modularity_final, dim3(1), dim3(256), 0, s, g, G, gbase, si.get(), st.get(),
                     out_dev);
  PCC_HIP_CHECK(hipGetLastError());
}

}  // namespace pcc
