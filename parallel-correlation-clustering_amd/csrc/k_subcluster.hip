// k_subcluster.hip -- the subcluster numbering of CompressSubclusters
// (clustering/clusterers/parallel-correlation-clusterer.cc:436-474) with
// ComputeSubcluster (parallel-correlation-clusterer-internal.cc:1101-1120):
// every cluster of a level is split into the components of a subgraph of its
// own edges, and the components become the next level's vertices.
//
//   CONNECTIVITY (PCCI.cc:846-895): all edges between two members;
//   TRIANGLE (PCCI.cc:1015-1083, scores from CountTriangles, PCCI.cc:786-799):
//     edges (u, v) between two members with |N(u) & N(v)| / (deg u + deg v)
//     > 0.25, i.e. 4 |N(u) & N(v)| > deg u + deg v (the quotient of two
//     integers below 2^32 cannot round onto 1/4 from above).
//
// The reference numbers a cluster's members 0..size-1 in ascending vertex
// order and labels each member with the union-find root of its component;
// GBBS's SimpleUnionAsyncStruct hooks the larger root under the smaller, so
// the label is the position of the component's smallest vertex.  This file
// runs the same union-find on vertex ids, lock-free: a root is hooked only by
// a CAS from itself to a smaller root and finds halve paths by CAS, so parent
// pointers only decrease and every component ends rooted at its smallest
// vertex, whatever the order.  A cluster spans size ids (CONNECTIVITY) or
// 1 + its largest label (TRIANGLE); spans are laid out in ascending cluster
// id order, sub[v] = span start + label, and id j of the r-th non-empty
// cluster's span starts the next level in cluster r.

#include "device_common.h"
#include "kernels.h"

namespace pcc {

namespace {

int ScGrid(int64_t items, int per_block, int max_blocks) {
  int64_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (int)b;
}

__device__ __forceinline__ uint32_t UfLoad(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ uint32_t UfFind(uint32_t* uf, uint32_t x) {
  uint32_t px = UfLoad(uf + x);
  while (px != x) {
    const uint32_t ppx = UfLoad(uf + px);
    if (ppx != px) atomicCAS(uf + x, px, ppx);
    x = px;
    px = ppx;
  }
  return x;
}

__device__ void UfUnite(uint32_t* uf, uint32_t a, uint32_t b) {
  while (true) {
    a = UfFind(uf, a);
    b = UfFind(uf, b);
    if (a == b) return;
    if (a < b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(uf + a, a, b) == a) return;
  }
}

// Common entries of two strictly ascending adjacency lists: each entry of the
// shorter list is searched in the rest of the longer one.
__device__ uint32_t CommonNeighbours(const uint32_t* __restrict__ nbr, int64_t a0, int64_t a1,
                                     int64_t b0, int64_t b1) {
  if (a1 - a0 > b1 - b0) {
    int64_t t = a0;
    a0 = b0;
    b0 = t;
    t = a1;
    a1 = b1;
    b1 = t;
  }
  uint32_t common = 0;
  int64_t lo = b0;
  for (int64_t i = a0; i < a1 && lo < b1; ++i) {
    const uint32_t x = nbr[i];
    int64_t hi = b1;
    while (lo < hi) {  // first position >= x
      const int64_t mid = (lo + hi) >> 1;
      if (nbr[mid] < x)
        lo = mid + 1;
      else
        hi = mid;
    }
    if (lo < b1 && nbr[lo] == x) {
      ++common;
      ++lo;
    }
  }
  return common;
}

__global__ __launch_bounds__(256) void k_sc_init(int64_t n, const uint32_t* __restrict__ ids,
                                                 uint32_t* __restrict__ cnt,
                                                 uint32_t* __restrict__ uf,
                                                 uint64_t* __restrict__ keys) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t c = ids[v];
    uf[v] = (uint32_t)v;
    atomicAdd(&cnt[c], 1u);
    keys[v] = ((uint64_t)c << 32) | (uint64_t)v;
  }
}

// Rows must be strictly ascending for the triangle scores (GBBS's merge
// intersect assumes it; contracted levels are built that way).
__global__ __launch_bounds__(256) void k_sc_sorted(CsrView g, uint32_t* __restrict__ bad) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p + 1 < g.m;
       p += (int64_t)gridDim.x * blockDim.x) {
    if (g.nbr[p] >= g.nbr[p + 1]) {
      // p and p + 1 in one row unless a row ends at p + 1.
      const int64_t u = VertexOfEdge(g.off, 0, g.n, p);
      if (g.off[u + 1] > p + 1) atomicOr(bad, 1u);
    }
  }
}

__global__ __launch_bounds__(256) void k_sc_rank(int64_t n, const uint64_t* __restrict__ keys,
                                                 const uint32_t* __restrict__ start,
                                                 uint32_t* __restrict__ pos) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[i];
    pos[(uint32_t)k] = (uint32_t)(i - (int64_t)start[k >> 32]);
  }
}

// Unions over the cluster-internal edges, kScItems consecutive CSR edges per
// thread.  half: the graph is symmetric, so each edge is taken once (u < v).
// TRIANGLE scores an edge in the thread when its shorter row has at most
// kScShortRow entries; longer ones are queued (count only when longq is
// null) for k_sc_tri_long, one wave per edge.
constexpr int kScItems = 16;
constexpr int64_t kScShortRow = 16;
__global__ __launch_bounds__(256) void k_sc_union(CsrView g, const uint32_t* __restrict__ ids,
                                                  int method, int half, uint32_t* __restrict__ uf,
                                                  int64_t* __restrict__ longq,
                                                  unsigned long long* __restrict__ nlong) {
  const int64_t nchunks = (g.m + kScItems - 1) / kScItems;
  for (int64_t ch = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; ch < nchunks;
       ch += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p0 = ch * kScItems;
    const int64_t p1 = min(p0 + (int64_t)kScItems, g.m);
    int64_t u = VertexOfEdge(g.off, 0, g.n, p0);
    int64_t uend = g.off[u + 1];
    for (int64_t p = p0; p < p1; ++p) {
      while (p >= uend) {
        ++u;
        uend = g.off[u + 1];
      }
      const uint32_t v = g.nbr[p];
      if (v == (uint32_t)u || (half && v < (uint32_t)u) || ids[v] != ids[u]) continue;
      if (method == 1) {
        const int64_t du = uend - g.off[u];
        const int64_t dv = g.off[(int64_t)v + 1] - g.off[v];
        if (min(du, dv) > kScShortRow) {
          const unsigned long long q = atomicAdd(nlong, 1ull);
          if (longq) longq[q] = p;
          continue;
        }
        if (!longq) continue;  // counting pass
        const uint32_t common = CommonNeighbours(g.nbr, g.off[u], uend, g.off[v], g.off[(int64_t)v + 1]);
        if (!(4 * (int64_t)common > du + dv)) continue;
      }
      UfUnite(uf, (uint32_t)u, v);
    }
  }
}

// One wave per queued TRIANGLE edge: the lanes search the entries of the
// shorter row in the longer one independently.
__global__ __launch_bounds__(256) void k_sc_tri_long(CsrView g, const int64_t* __restrict__ longq,
                                                     int64_t nq, uint32_t* __restrict__ uf) {
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / kWave);
  const int lane = LaneId();
  for (int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; q < nq; q += waves) {
    const int64_t p = longq[q];
    const int64_t u = VertexOfEdge(g.off, 0, g.n, p);
    const uint32_t v = g.nbr[p];
    int64_t a0 = g.off[u], a1 = g.off[u + 1], b0 = g.off[v], b1 = g.off[(int64_t)v + 1];
    const int64_t du = a1 - a0, dv = b1 - b0;
    if (du > dv) {
      int64_t t = a0;
      a0 = b0;
      b0 = t;
      t = a1;
      a1 = b1;
      b1 = t;
    }
    uint32_t common = 0;
    for (int64_t i = a0 + lane; i < a1; i += kWave) {
      const uint32_t x = g.nbr[i];
      int64_t lo = b0, hi = b1;
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (g.nbr[mid] < x)
          lo = mid + 1;
        else
          hi = mid;
      }
      common += (lo < b1 && g.nbr[lo] == x) ? 1u : 0u;
    }
    for (int s = 32; s > 0; s >>= 1) common += __shfl_xor(common, s);
    if (lane == 0 && 4 * (int64_t)common > du + dv) UfUnite(uf, (uint32_t)u, v);
  }
}

// Labels: position of the component's root; TRIANGLE spans = 1 + the largest
// label of the cluster.
__global__ __launch_bounds__(256) void k_sc_label(int64_t n, const uint32_t* __restrict__ ids,
                                                  uint32_t* __restrict__ uf,
                                                  const uint32_t* __restrict__ pos,
                                                  uint32_t* __restrict__ lab,
                                                  uint32_t* __restrict__ span) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t r = UfFind(uf, (uint32_t)v);
    const uint32_t l = pos[r];
    lab[v] = l;
    if (span) atomicMax(&span[ids[v]], l + 1);
  }
}

__global__ __launch_bounds__(256) void k_sc_sub(int64_t n, const uint32_t* __restrict__ ids,
                                                const uint32_t* __restrict__ lab,
                                                const uint32_t* __restrict__ offset,
                                                uint32_t* __restrict__ sub) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    sub[v] = offset[ids[v]] + lab[v];
}

// mark[span start of every non-empty cluster] = 1 (starts are < n2: a
// non-empty cluster holds an id at or after its start).
__global__ __launch_bounds__(256) void k_sc_mark(int64_t n, const uint32_t* __restrict__ cnt,
                                                 const uint32_t* __restrict__ offset,
                                                 uint8_t* __restrict__ mark) {
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x)
    if (cnt[c] > 0) mark[offset[c]] = 1;
}

__global__ __launch_bounds__(256) void k_sc_parent(int64_t n2, const uint8_t* __restrict__ mark,
                                                   const uint32_t* __restrict__ ex,
                                                   uint32_t* __restrict__ parent) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n2;
       j += (int64_t)gridDim.x * blockDim.x)
    parent[j] = ex[j] + mark[j] - 1u;
}

// Edges of every vertex that leave its cluster (outdeg zeroed by the
// caller), kScItems consecutive CSR edges per thread, one atomic per row run.
__global__ __launch_bounds__(256) void k_sc_outdeg(CsrView g, const uint32_t* __restrict__ ids,
                                                   uint32_t* __restrict__ outdeg) {
  const int64_t nchunks = (g.m + kScItems - 1) / kScItems;
  for (int64_t ch = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; ch < nchunks;
       ch += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p0 = ch * kScItems;
    const int64_t p1 = min(p0 + (int64_t)kScItems, g.m);
    int64_t u = VertexOfEdge(g.off, 0, g.n, p0);
    int64_t uend = g.off[u + 1];
    uint32_t cu = ids[u], k = 0;
    for (int64_t p = p0; p < p1; ++p) {
      while (p >= uend) {
        if (k) atomicAdd(&outdeg[u], k);
        k = 0;
        ++u;
        uend = g.off[u + 1];
        cu = ids[u];
      }
      k += ids[g.nbr[p]] != cu ? 1u : 0u;
    }
    if (k) atomicAdd(&outdeg[u], k);
  }
}

// SLOW_CONNECTIVITY (PCCI.cc:897-1013).  Scratch arrays are indexed by the
// position q in the (cluster, vertex)-sorted member list, so clusters never
// share entries; k_sc_slow_init sets them up for every member at once:
// group value t of a member = its local index, size 1, its edges leaving the
// cluster, and cand = the member has at least size - 1 of them (the
// reference's test, PCCI.cc:929-935, which depends on nothing that moves).
__global__ __launch_bounds__(256) void k_sc_slow_init(int64_t n, const uint64_t* __restrict__ members,
                                                      const uint32_t* __restrict__ start,
                                                      const uint32_t* __restrict__ cnt,
                                                      const uint32_t* __restrict__ outdeg,
                                                      uint32_t* __restrict__ tmp,
                                                      uint32_t* __restrict__ gsize,
                                                      uint64_t* __restrict__ gout,
                                                      uint32_t* __restrict__ gfirst,
                                                      uint32_t* __restrict__ single,
                                                      uint8_t* __restrict__ cand,
                                                      uint32_t* __restrict__ wcnt) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = members[q];
    const uint32_t c = (uint32_t)(k >> 32), v = (uint32_t)k;
    const uint32_t j = (uint32_t)(q - (int64_t)start[c]);
    const uint32_t size = cnt[c];
    tmp[q] = j;
    gsize[q] = 1;
    gout[q] = outdeg[v];
    gfirst[q] = j;
    single[q] = 0;
    cand[q] = size > 1 && outdeg[v] >= size - 1 ? 1 : 0;
    wcnt[q] = 0;
  }
}

// One wave per cluster, in the reference's member order.  The reference
// rebuilds the groups for every member; here each group value t keeps its
// size, its edges leaving the cluster (the sum over its members) and its
// smallest member, updated by lane 0 when a member joins.  Member i (a
// candidate still alone) counts its edges into each group (lanes over its
// row, atomics into wcnt), then takes the last group t != its own, in
// ascending t, with leaving(t) >= |t| (size - |t|) and
// weight - fma(offset, |t|, res |t|) >= 0 and maximal: each lane keeps the
// largest surplus of its t's (the later t on ties) and the wave takes the
// largest surplus, then the largest t.  Only candidates' groups can be
// valid: a singleton group {t} is valid iff t has at least size - 1 leaving
// edges (the candidate test), and a larger group grew from a valid singleton
// by joins, so both loops run over the cluster's candidates, compacted into
// clist in ascending order.  The state lanes share is read with agent-scope
// loads (no stale L1 lines) after a fence.
template <typename T>
__device__ __forceinline__ T ScLoad(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(256) void k_sc_slow(int64_t n, CsrView g, const uint32_t* __restrict__ ids,
                                                 const uint32_t* __restrict__ cnt,
                                                 const uint32_t* __restrict__ start,
                                                 const uint64_t* __restrict__ members,
                                                 const uint32_t* __restrict__ pos,
                                                 const uint32_t* __restrict__ outdeg, double offset,
                                                 double res, uint32_t* tmp, uint32_t* gsize,
                                                 uint64_t* gout, uint32_t* gfirst,
                                                 uint32_t* single,
                                                 const uint8_t* __restrict__ cand,
                                                 uint32_t* wcnt, uint32_t* clist) {
  const int lane = LaneId();
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / kWave);
  for (int64_t ci = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; ci < n; ci += waves) {
    const uint32_t size = cnt[ci];
    if (size <= 1) continue;
    const int64_t base = start[ci];
    uint32_t nc = 0;
    for (uint32_t i0 = 0; i0 < size; i0 += kWave) {
      const bool f = i0 + lane < size && cand[base + i0 + lane] != 0;
      const uint64_t m = __ballot(f);
      if (f) clist[base + nc + (uint32_t)__popcll(m & below)] = i0 + lane;
      nc += (uint32_t)__popcll(m);
    }
    if (nc < 2) continue;  // a lone candidate has only its own group
    __threadfence();
    for (uint32_t ck = 0; ck < nc; ++ck) {
      {
        const uint32_t i = ScLoad(&clist[base + ck]);
        if (ScLoad(&single[base + i])) continue;
        const uint32_t v = (uint32_t)members[base + i];
        const int64_t r0 = g.off[v], r1 = g.off[(int64_t)v + 1];
        for (int64_t p = r0 + lane; p < r1; p += kWave) {
          const uint32_t u = g.nbr[p];
          if (ids[u] == (uint32_t)ci) atomicAdd(&wcnt[base + ScLoad(&tmp[base + pos[u]])], 1u);
        }
        __threadfence();
        double best = -1.0;
        uint32_t bt = kNoneId;
        for (uint32_t kk = lane; kk < nc; kk += kWave) {
          const uint32_t t = ScLoad(&clist[base + kk]);
          const uint32_t gs = ScLoad(&gsize[base + t]);
          if (gs == 0 || t == i) continue;
          if (ScLoad(&gout[base + t]) < (uint64_t)gs * (uint64_t)(size - gs)) continue;
          const double w = (double)ScLoad(&wcnt[base + t]);
          const double cs = (double)gs;
          const double off = __fma_rn(offset, cs, __dmul_rn(res, cs));
          if (w < off) continue;
          const double surplus = __dsub_rn(w, off);
          if (surplus >= best) {
            best = surplus;
            bt = t;
          }
        }
        for (int sh = 32; sh > 0; sh >>= 1) {
          const double ob = __shfl_xor(best, sh);
          const uint32_t ot = __shfl_xor(bt, sh);
          if (ot != kNoneId && (bt == kNoneId || ob > best || (ob == best && ot > bt))) {
            best = ob;
            bt = ot;
          }
        }
        __threadfence();
        for (int64_t p = r0 + lane; p < r1; p += kWave) {
          const uint32_t u = g.nbr[p];
          if (ids[u] == (uint32_t)ci) atomicExch(&wcnt[base + ScLoad(&tmp[base + pos[u]])], 0u);
        }
        if (bt != kNoneId && lane == 0) {
          const uint32_t f = ScLoad(&gfirst[base + bt]);
          atomicExch(&single[base + i], 1u);
          atomicExch(&single[base + f], 1u);
          atomicExch(&tmp[base + i], bt);
          atomicExch(&gsize[base + i], 0u);
          atomicAdd(&gsize[base + bt], 1u);
          atomicAdd((unsigned long long*)&gout[base + bt], (unsigned long long)outdeg[v]);
          atomicExch(&gfirst[base + bt], f < i ? f : i);
        }
        __threadfence();
      }
    }
  }
}

// Labels = group values; span = 1 + the largest (PCCI.cc:1006-1012).
__global__ __launch_bounds__(256) void k_sc_slow_labels(int64_t n, const uint64_t* __restrict__ members,
                                                        const uint32_t* __restrict__ tmp,
                                                        uint32_t* __restrict__ lab,
                                                        uint32_t* __restrict__ span) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = members[q];
    const uint32_t t = tmp[q];
    lab[(uint32_t)k] = t;
    atomicMax(&span[k >> 32], t + 1);
  }
}

int ScBits(uint64_t maxval) {
  int b = 0;
  while (b < 63 && (maxval >> b) != 0) ++b;
  return b < 1 ? 1 : b;
}

}  // namespace

bool RowsAscending(DeviceAllocator* al, const CsrView& g, hipStream_t s) {
  if (g.m <= 1) return true;
  DevBuf<uint32_t> bad(al, 1);
  PCC_HIP_CHECK(hipMemsetAsync(bad.get(), 0, 4, s));
  hipLaunchKernelGGL(k_sc_sorted, dim3(ScGrid(g.m, 256, 8192)), dim3(256), 0, s, g, bad.get());
  PCC_HIP_CHECK(hipGetLastError());
  uint32_t hbad = 0;
  PCC_HIP_CHECK(hipMemcpyAsync(&hbad, bad.get(), 4, hipMemcpyDeviceToHost, s));
  PCC_HIP_CHECK(hipStreamSynchronize(s));
  return hbad == 0;
}

int64_t SubclusterIds(DeviceAllocator* al, const CsrView& g, const uint32_t* ids, int method,
                      DevBuf<uint32_t>* sub, DevBuf<uint32_t>* parent, double offset,
                      double res, hipStream_t s, bool check_rows, bool half) {
  const int64_t n = g.n;
  if (n <= 0) return 0;
  const int grid = ScGrid(n, 256, 8192);
  if (method == 1 && check_rows && !RowsAscending(al, g, s)) return -1;
  DevBuf<uint32_t> cnt(al, n), start(al, n), uf(al, n), pos(al, n), lab(al, n);
  DevBuf<uint64_t> keys(al, n);
  DevBuf<float> dummy(al, n);
  PCC_HIP_CHECK(hipMemsetAsync(cnt.get(), 0, 4 * (size_t)n, s));
  hipLaunchKernelGGL(k_sc_init, dim3(grid), dim3(256), 0, s, n, ids, cnt.get(), uf.get(),
                     keys.get());
  PCC_HIP_CHECK(hipGetLastError());
  ExclusiveScanU32toU32(al, cnt.get(), start.get(), n, nullptr, s);
  // Cluster ids are vertex-range ids (< n): sort on the id bits only; the low
  // word (the vertex) is already ascending within every id.
  RadixSortPairsF32(al, keys.get(), dummy.get(), n, 32, 32 + ScBits((uint64_t)(n - 1)), s);
  dummy.Release();
  hipLaunchKernelGGL(k_sc_rank, dim3(grid), dim3(256), 0, s, n, keys.get(), start.get(),
                     pos.get());
  PCC_HIP_CHECK(hipGetLastError());
  DevBuf<uint32_t> span;
  DevBuf<uint32_t> toff;
  const uint32_t* offset_of = start.get();  // CONNECTIVITY: span = size
  if (method == 3) {
    span.Reset(al, n);
    toff.Reset(al, n);
    PCC_HIP_CHECK(hipMemsetAsync(span.get(), 0, 4 * (size_t)n, s));
    DevBuf<uint32_t> outdeg(al, n), tmp(al, n), gsize(al, n), gfirst(al, n), wcnt(al, n);
    DevBuf<uint64_t> gout(al, n);
    DevBuf<uint32_t> single(al, n), clist(al, n);
    DevBuf<uint8_t> cand(al, n);
    PCC_HIP_CHECK(hipMemsetAsync(outdeg.get(), 0, 4 * (size_t)n, s));
    if (g.m > 0)
      hipLaunchKernelGGL(k_sc_outdeg, dim3(ScGrid((g.m + kScItems - 1) / kScItems, 256, 8192)),
                         dim3(256), 0, s, g, ids, outdeg.get());
    hipLaunchKernelGGL(k_sc_slow_init, dim3(grid), dim3(256), 0, s, n, keys.get(), start.get(),
                       cnt.get(), outdeg.get(), tmp.get(), gsize.get(), gout.get(), gfirst.get(),
                       single.get(), cand.get(), wcnt.get());
    hipLaunchKernelGGL(k_sc_slow, dim3(ScGrid(n, 256 / kWave, 1 << 16)), dim3(256), 0, s, n, g, ids,
                       cnt.get(), start.get(), keys.get(), pos.get(), outdeg.get(), offset, res,
                       tmp.get(), gsize.get(), gout.get(), gfirst.get(), single.get(), cand.get(),
                       wcnt.get(), clist.get());
    hipLaunchKernelGGL(k_sc_slow_labels, dim3(grid), dim3(256), 0, s, n, keys.get(), tmp.get(),
                       lab.get(), span.get());
    PCC_HIP_CHECK(hipGetLastError());
    ExclusiveScanU32toU32(al, span.get(), toff.get(), n, nullptr, s);
    offset_of = toff.get();
  }
  keys.Release();
  if (method != 3 && g.m > 0) {
    const int64_t nchunks = (g.m + kScItems - 1) / kScItems;
    const int ugrid = ScGrid(nchunks, 256, 8192);
    const int h = half ? 1 : 0;
    if (method == 1) {
      // Count the long-row edges, then queue them while the short ones unite.
      DevBuf<unsigned long long> nlong(al, 1);
      PCC_HIP_CHECK(hipMemsetAsync(nlong.get(), 0, sizeof(unsigned long long), s));
      hipLaunchKernelGGL(k_sc_union, dim3(ugrid), dim3(256), 0, s, g, ids, method, h, uf.get(),
                         nullptr, nlong.get());
      PCC_HIP_CHECK(hipGetLastError());
      unsigned long long nq = 0;
      PCC_HIP_CHECK(hipMemcpyAsync(&nq, nlong.get(), sizeof(nq), hipMemcpyDeviceToHost, s));
      PCC_HIP_CHECK(hipStreamSynchronize(s));
      DevBuf<int64_t> longq(al, nq > 0 ? (int64_t)nq : 1);
      PCC_HIP_CHECK(hipMemsetAsync(nlong.get(), 0, sizeof(unsigned long long), s));
      hipLaunchKernelGGL(k_sc_union, dim3(ugrid), dim3(256), 0, s, g, ids, method, h, uf.get(),
                         longq.get(), nlong.get());
      PCC_HIP_CHECK(hipGetLastError());
      if (nq > 0) {
        hipLaunchKernelGGL(k_sc_tri_long, dim3(ScGrid((int64_t)nq, 256 / kWave, 1 << 16)), dim3(256),
                           0, s, g, longq.get(), (int64_t)nq, uf.get());
        PCC_HIP_CHECK(hipGetLastError());
      }
    } else {
      hipLaunchKernelGGL(k_sc_union, dim3(ugrid), dim3(256), 0, s, g, ids, method, h, uf.get(),
                         nullptr, nullptr);
      PCC_HIP_CHECK(hipGetLastError());
    }
  }
  if (method == 1) {
    span.Reset(al, n);
    toff.Reset(al, n);
    PCC_HIP_CHECK(hipMemsetAsync(span.get(), 0, 4 * (size_t)n, s));
  }
  if (method != 3) {
    hipLaunchKernelGGL(k_sc_label, dim3(grid), dim3(256), 0, s, n, ids, uf.get(), pos.get(),
                       lab.get(), span.get());
    PCC_HIP_CHECK(hipGetLastError());
  }
  if (method == 1) {
    ExclusiveScanU32toU32(al, span.get(), toff.get(), n, nullptr, s);
    offset_of = toff.get();
  }
  sub->Reset(al, n);
  hipLaunchKernelGGL(k_sc_sub, dim3(grid), dim3(256), 0, s, n, ids, lab.get(), offset_of,
                     sub->get());
  PCC_HIP_CHECK(hipGetLastError());
  DevBuf<uint32_t> dmax(al, 1);
  ReduceMaxU32(al, sub->get(), n, dmax.get(), s);
  uint32_t hmax = 0;
  PCC_HIP_CHECK(hipMemcpyAsync(&hmax, dmax.get(), 4, hipMemcpyDeviceToHost, s));
  PCC_HIP_CHECK(hipStreamSynchronize(s));
  const int64_t n2 = (int64_t)hmax + 1;
  DevBuf<uint8_t> mark(al, n2);
  DevBuf<uint32_t> ex(al, n2);
  PCC_HIP_CHECK(hipMemsetAsync(mark.get(), 0, (size_t)n2, s));
  hipLaunchKernelGGL(k_sc_mark, dim3(grid), dim3(256), 0, s, n, cnt.get(), offset_of, mark.get());
  PCC_HIP_CHECK(hipGetLastError());
  ExclusiveScanU8toU32(al, mark.get(), ex.get(), n2, nullptr, s);
  parent->Reset(al, n2);
  hipLaunchKernelGGL(k_sc_parent, dim3(ScGrid(n2, 256, 8192)), dim3(256), 0, s, n2, mark.get(),
                     ex.get(), parent->get());
  PCC_HIP_CHECK(hipGetLastError());
  return n2;
}

}  // namespace pcc
