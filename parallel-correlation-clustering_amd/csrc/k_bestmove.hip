// k_bestmove.hip -- degree-binned best-move kernels for gfx950.
//
// Restates ClusteringHelper::EfficientBestMove
// (clustering/clusterers/parallel-correlation-clusterer-internal.cc:448-583),
// the sync decision of BestMovesForVertexSubset
// (clustering/clusterers/parallel-correlation-clusterer.cc:140-156) and
// ClusteringHelper::AsyncMove (parallel-correlation-clusterer-internal.cc:371-403).
//
// For vertex v with cluster c(v) and node weight k:
//   S_c    = sum over neighbours u != v with c(u) == c of (w_uv - offset)
//   change = 0 - NetWeight(S_{c(v)}, k * (K_{c(v)} - k))
//   delta_c = change + NetWeight(S_c, k * K_c)
// The best target is the candidate with the largest delta (smallest id on a
// tie) if that delta exceeds change, else "new cluster" (id n) with change.
//
// Bins and their exactness:
//   bins 0-4 (deg <= 64): G lanes per vertex; every lane sums the weights of
//     its cluster by walking the group's lanes in adjacency order with
//     shuffles, so each S_c is accumulated in exactly the reference's order.
//   bins 5-7: LDS (or global) open-addressing hash with FP64 atomics.  The
//     kernel also checks that every partial sum is exactly representable; if
//     not, the vertex goes to the ordered fallback kernel, which accumulates
//     in adjacency order (and the hub-vertex block order of
//     curr_together_seq).  Results are therefore bit-identical to a
//     sequential evaluation in every case.

#include "device_common.h"
#include "kernels.h"

namespace pcc {

namespace {

__device__ __forceinline__ int BinOf(int64_t deg, int64_t wave_max, int64_t block_max) {
  if (deg <= 4) return 0;
  if (deg <= 8) return 1;
  if (deg <= 16) return 2;
  if (deg <= 32) return 3;
  if (deg <= 64) return 4;
  if (deg <= wave_max) return 5;
  if (deg <= block_max) return 6;
  return 7;
}

__device__ __forceinline__ uint64_t LaneMaskLt() {
  const int l = LaneId();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__global__ __launch_bounds__(256) void k_bin_frontier(int64_t n, const int64_t* __restrict__ off,
                                                      const uint8_t* __restrict__ frontier,
                                                      BinThresholds th, BinLayout layout,
                                                      uint32_t* __restrict__ list,
                                                      RoundCounters* __restrict__ rc) {
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = base + threadIdx.x;
    const bool in = v < n && (frontier == nullptr || frontier[v] != 0);
    int b = -1;
    int64_t deg = 0;
    if (in) {
      deg = off[v + 1] - off[v];
      b = BinOf(deg, th.wave_max, th.block_max);
    }
    for (int bb = 0; bb < kBins; ++bb) {
      const uint64_t mask = __ballot(b == bb);
      if (mask == 0) continue;
      const int leader = __ffsll((unsigned long long)mask) - 1;
      unsigned long long te = (b == bb) ? (unsigned long long)deg : 0ull;
      for (int s = 32; s > 0; s >>= 1) te += __shfl_xor(te, s);
      unsigned long long pos = 0;
      if (LaneId() == leader) {
        pos = atomicAdd(&rc->cnt[bb], (unsigned long long)__popcll(mask));
        atomicAdd(&rc->te[bb], te);
      }
      pos = __shfl(pos, leader);
      if (b == bb) {
        const uint32_t r = (uint32_t)__popcll(mask & LaneMaskLt());
        list[layout.base[bb] + (int64_t)pos + r] = (uint32_t)v;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_bin_capacity(int64_t n, const int64_t* __restrict__ off,
                                                      BinThresholds th,
                                                      unsigned long long* __restrict__ caps) {
  __shared__ unsigned long long h[kBins];
  if (threadIdx.x < kBins) h[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    atomicAdd(&h[BinOf(off[v + 1] - off[v], th.wave_max, th.block_max)], 1ull);
  }
  __syncthreads();
  if (threadIdx.x < kBins && h[threadIdx.x]) atomicAdd(&caps[threadIdx.x], h[threadIdx.x]);
}

// Decision of EfficientBestMove from the best candidate (low-degree path,
// PCCI.cc:490-510; hub path PCCI.cc:565-582, whose only difference is the
// empty-candidate case).
__device__ __forceinline__ void Decide(Cand best, double change, bool hub, uint32_t new_id,
                                       uint32_t* t, double* d) {
  if (best.c == kNoneId) {
    if (hub && !(change >= 0.0)) {
      *t = 0;
      *d = 0.0;
    } else {
      *t = new_id;
      *d = change;
    }
    return;
  }
  if (best.d > change) {
    *t = best.c;
    *d = best.d;
  } else {
    *t = new_id;
    *d = change;
  }
}

// Resolution of the vertex's graph.
__device__ __forceinline__ double ResOf(const MoveParams& mp, uint32_t v) {
  return mp.vgraph ? mp.gres[mp.vgraph[v]] : mp.res;
}

// Sync epilogue: singleton rule then delta > 0 (PCC.cc:145-156).  Targets are
// cluster ids or new_id, so "target < n" of the reference is target != new_id.
__device__ __forceinline__ void SyncEpilogue(uint32_t v, uint32_t t, double d, uint32_t new_id,
                                             const StateView& st, const MoveOut& out) {
  if (out.dbg_target) out.dbg_target[v] = t;
  if (out.dbg_delta) out.dbg_delta[v] = d;
  const uint32_t cur = st.ids[v];
  if (t != new_id && st.sizes[t] == 1 && st.sizes[cur] == 1 && cur >= t) {
    t = cur;
    d = 0.0;
  }
  if (d > 0) {
    out.moves[v] = t;
    out.moved[v] = 1;
  } else {
    out.moves[v] = kNoneId;
    out.moved[v] = 0;
  }
}

// Async apply (AsyncMove, PCCI.cc:371-403).  Moves to a new cluster claim an
// empty id first (probing v, v+1, ... like the reference), so a failed probe
// leaves the state untouched.
__device__ __forceinline__ void AsyncEpilogue(uint32_t v, uint32_t t, double d,
                                              const MoveParams& mp, const StateView& st,
                                              const MoveOut& out) {
  const uint32_t cur = st.ids[v];
  if (mp.async_singleton_rule && t != mp.new_id && st.sizes[t] == 1 && st.sizes[cur] == 1 &&
      cur >= t)
    d = 0.0;
  if (!(d > 0)) {
    out.moved[v] = 0;
    return;
  }
  const double k = st.nw[v];
  if (t != mp.new_id) {
    atomicSub(&st.sizes[cur], 1u);
    atomicAdd(&st.K[cur], -1 * k);
    atomicAdd(&st.sizes[t], 1u);
    atomicAdd(&st.K[t], k);
    st.ids[v] = t;
    out.moved_clusters[t] = 1;
    out.moved[v] = 1;
    return;
  }
  // Empty ids are searched within the vertex's own graph.
  uint32_t lo = 0, hi = mp.n;
  if (mp.vgraph) {
    const uint32_t g = mp.vgraph[v];
    lo = (uint32_t)mp.gbase[g];
    hi = (uint32_t)mp.gbase[g + 1];
  }
  uint32_t i = v;
  for (uint32_t probe = lo; probe < hi; ++probe) {
    if (__hip_atomic_load(&st.sizes[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0 &&
        atomicCAS(&st.sizes[i], 0u, 1u) == 0u) {
      atomicSub(&st.sizes[cur], 1u);
      atomicAdd(&st.K[cur], -1 * k);
      atomicAdd(&st.K[i], k);
      st.ids[v] = i;
      out.moved_clusters[i] = 1;
      out.moved[v] = 1;
      return;
    }
    i = (i + 1 == hi) ? lo : i + 1;
  }
  out.moved[v] = 0;
}

// Sub-round of a vertex in async mode (a fixed hash of the id).
__device__ __forceinline__ bool InSubRound(const MoveParams& mp, uint32_t v) {
  if (mp.sub_count <= 1) return true;
  uint32_t h = v * 0x9E3779B1u;
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  return h % mp.sub_count == mp.sub_index;
}

// Item -> list position; with a salt, a bijection of [0, count) that changes
// the processing order (vertexMapPermute, PCC.cc:69-109).
__device__ __forceinline__ int64_t PermuteIndex(int64_t i, int64_t count, uint64_t salt) {
  if (salt == 0 || count <= 1) return i;
  // i -> (i * p + salt) mod count with p = (2^61 - 1) mod count; 2^61 - 1 is
  // prime and larger than count, so gcd(p, count) = 1.  count < 2^32 keeps
  // the product below 2^64.
  const uint64_t cnt = (uint64_t)count;
  const uint64_t p = 2305843009213693951ull % cnt;
  return (int64_t)(((uint64_t)i * p + salt % cnt) % cnt);
}

// ------------------------------------------------------------ small bins --

template <int G, bool kAsync>
__global__ __launch_bounds__(256) void k_bm_small(CsrView g, StateView st, MoveParams mp,
                                                  const uint32_t* __restrict__ list,
                                                  int64_t count, MoveOut out, uint64_t salt) {
  constexpr int kGroups = 256 / G;
  const int lane = threadIdx.x & (G - 1);
  const int wlane = LaneId();
  const int gbase_lane = wlane & ~(G - 1);
  const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1ull) << gbase_lane);
  for (int64_t gb = (int64_t)blockIdx.x * kGroups; gb < count;
       gb += (int64_t)gridDim.x * kGroups) {
    const int64_t gi = gb + threadIdx.x / G;
    const uint32_t v = gi < count ? list[PermuteIndex(gi, count, salt)] : 0u;
    const bool active = gi < count && (!kAsync || InSubRound(mp, v));
    int64_t start = 0;
    int deg = 0;
    if (active) {
      start = g.off[v];
      deg = (int)(g.off[v + 1] - start);
    }
    uint32_t c = kNoneId;
    float wf = 0.0f;
    if (lane < deg) {
      const uint32_t u = g.nbr[start + lane];
      wf = g.w[start + lane];
      if (u != v) c = st.ids[u];
    }
    const uint32_t cv = active ? st.ids[v] : kNoneId;
    const int jmax = WaveMaxInt(deg);
    double S = 0.0;
    bool first = c != kNoneId;
    for (int j = 0; j < jmax; ++j) {
      const uint32_t cj = __shfl(c, j, G);
      const float wj = __shfl(wf, j, G);
      if (c != kNoneId && cj == c) {
        double x = (double)wj;
        x -= mp.offset;
        S += x;
        if (j < lane) first = false;
      }
    }
    // T = S_{c(v)}: the sum held by the first lane of the vertex's own cluster.
    const uint64_t tm = __ballot(first && c == cv) & gmask;
    const int tl = tm ? (__ffsll((unsigned long long)tm) - 1) : wlane;
    const double Ts = __shfl(S, tl);
    const double T = tm ? Ts : 0.0;
    const double k = active ? st.nw[v] : 0.0;
    const double Kcv = active ? st.K[cv] : 0.0;
    const double res = active ? ResOf(mp, v) : 0.0;
    double change = 0;
    change -= NetWeight(T, k * (Kcv - k), res);
    double mw = 0;
    mw += k;
    Cand cand{0.0, kNoneId};
    if (first) {
      cand.c = c;
      cand.d = change + NetWeight(S, mw * st.K[c], res);
    }
    cand = GroupBest(cand, G);
    if (active && lane == 0) {
      uint32_t t;
      double d;
      Decide(cand, change, false, mp.new_id, &t, &d);
      if (kAsync)
        AsyncEpilogue(v, t, d, mp, st, out);
      else
        SyncEpilogue(v, t, d, mp.new_id, st, out);
    }
  }
}

// ------------------------------------------------------------ hash bins --

template <typename KeyPtr, typename ValPtr>
__device__ __forceinline__ void HashAdd(KeyPtr keys, ValPtr vals, uint32_t mask, uint32_t c,
                                        double x) {
  uint32_t h = HashId(c) & mask;
  while (true) {
    const uint32_t k = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == c) {
      atomicAdd(&vals[h], x);
      return;
    }
    if (k == kNoneId) {
      const uint32_t old = atomicCAS(&keys[h], kNoneId, c);
      if (old == kNoneId || old == c) {
        atomicAdd(&vals[h], x);
        return;
      }
    }
    h = (h + 1) & mask;
  }
}

template <typename KeyPtr>
__device__ __forceinline__ int HashFindOrInsert(KeyPtr keys, uint32_t mask, uint32_t c) {
  uint32_t h = HashId(c) & mask;
  while (true) {
    const uint32_t k = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == c) return (int)h;
    if (k == kNoneId) {
      const uint32_t old = atomicCAS(&keys[h], kNoneId, c);
      if (old == kNoneId || old == c) return (int)h;
    }
    h = (h + 1) & mask;
  }
}

template <typename KeyPtr, typename ValPtr>
__device__ __forceinline__ double HashGet(KeyPtr keys, ValPtr vals, uint32_t mask, uint32_t c) {
  uint32_t h = HashId(c) & mask;
  while (true) {
    const uint32_t k = keys[h];
    if (k == c) return vals[h];
    if (k == kNoneId) return 0.0;
    h = (h + 1) & mask;
  }
}

__device__ __forceinline__ uint32_t NextPow2(uint64_t x) {
  uint32_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

template <int THREADS>
__device__ __forceinline__ Cand BlockBest(Cand x, double* sd, uint32_t* sc) {
  x = GroupBest(x, kWave);
  constexpr int NW = THREADS / kWave;
  if (NW == 1) return x;
  __syncthreads();
  if (LaneId() == 0) {
    sd[threadIdx.x / kWave] = x.d;
    sc[threadIdx.x / kWave] = x.c;
  }
  __syncthreads();
  Cand r{sd[0], sc[0]};
  for (int i = 1; i < NW; ++i) r = PickBetter(r, Cand{sd[i], sc[i]});
  __syncthreads();
  return r;
}

// THREADS cooperating threads per vertex; TABLE > 0: LDS table of TABLE
// entries, TABLE == 0: per-workgroup region of the global scratch table.
template <int THREADS, int TABLE, bool kAsync>
__global__ __launch_bounds__(THREADS) void k_bm_hash(CsrView g, StateView st, MoveParams mp,
                                                     const uint32_t* __restrict__ list,
                                                     int64_t count, MoveOut out,
                                                     uint32_t* __restrict__ fb_list,
                                                     RoundCounters* __restrict__ rc,
                                                     HashScratch hs, uint64_t salt) {
  constexpr int NW = THREADS / kWave;
  __shared__ uint32_t skeys[TABLE > 0 ? TABLE : 1];
  __shared__ double svals[TABLE > 0 ? TABLE : 1];
  __shared__ double sd[NW];
  __shared__ uint32_t sc[NW];
  __shared__ int si[NW];
  __shared__ double sdd[NW];
  uint32_t* keys;
  double* vals;
  if (TABLE > 0) {
    keys = skeys;
    vals = svals;
  } else {
    keys = hs.keys + (size_t)blockIdx.x * hs.cap;
    vals = hs.vals + (size_t)blockIdx.x * hs.cap;
  }
  for (int64_t i = blockIdx.x; i < count; i += gridDim.x) {
    const uint32_t v = list[PermuteIndex(i, count, salt)];
    if (kAsync && !InSubRound(mp, v)) continue;  // uniform across the workgroup
    const int64_t start = g.off[v];
    const int64_t deg = g.off[v + 1] - start;
    // Table sized per vertex (load factor <= 1/2) so that clearing and
    // scanning it costs O(deg), not O(TABLE).
    const uint32_t cap = min(TABLE > 0 ? (uint32_t)TABLE : hs.cap,
                             max(64u, NextPow2((uint64_t)(2 * deg + 2))));
    const uint32_t mask = cap - 1;
    for (uint32_t s = threadIdx.x; s < cap; s += THREADS) {
      keys[s] = kNoneId;
      vals[s] = 0.0;
    }
    __syncthreads();
    const uint32_t cv = st.ids[v];
    const bool hub = deg > kHubDegree;
    int gmin = 1 << 20;
    double sabs = 0.0;
    for (int64_t p = threadIdx.x; p < deg; p += THREADS) {
      const uint32_t u = g.nbr[start + p];
      if (u == v) continue;
      const float wf = g.w[start + p];
      double x;
      if (hub) {
        float wh = wf;
        wh -= mp.offset;  // float arithmetic on the hub path (PCCI.cc:529)
        x = (double)wh;
      } else {
        x = (double)wf;
        x -= mp.offset;
      }
      const uint32_t c = st.ids[u];
      if (!kAsync) {
        gmin = min(gmin, GranExp(x));
        sabs += fabs(x);
      }
      if (TABLE > 0)
        HashAdd(skeys, svals, mask, c, x);
      else
        HashAdd(keys, vals, mask, c, x);
    }
    __syncthreads();
    if (!kAsync) {
      const int gm = BlockReduce(gmin, si, [](int a, int b) { return min(a, b); });
      const double sa = BlockReduce(sabs, sdd, [](double a, double b) { return a + b; });
      const bool exact = !mp.force_ordered && SumIsOrderFree(gm, sa);
      if (!exact) {
        if (threadIdx.x == 0) {
          const unsigned long long pos = atomicAdd(&rc->fb_count, 1ull);
          fb_list[pos] = v;
        }
        __syncthreads();
        continue;
      }
    }
    const double T = HashGet(keys, vals, mask, cv);
    const double k = st.nw[v];
    const double res = ResOf(mp, v);
    double change = 0;
    change -= NetWeight(T, k * (st.K[cv] - k), res);
    double mw = 0;
    mw += k;
    Cand best{0.0, kNoneId};
    for (uint32_t s = threadIdx.x; s < cap; s += THREADS) {
      const uint32_t c = keys[s];
      if (c != kNoneId) {
        const Cand x{change + NetWeight(vals[s], mw * st.K[c], res), c};
        best = PickBetter(best, x);
      }
    }
    best = BlockBest<THREADS>(best, sd, sc);
    if (threadIdx.x == 0) {
      uint32_t t;
      double d;
      Decide(best, change, hub, mp.new_id, &t, &d);
      if (kAsync)
        AsyncEpilogue(v, t, d, mp, st, out);
      else
        SyncEpilogue(v, t, d, mp.new_id, st, out);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------- ordered fallback --

// One wave per vertex.  Chunks of 64 neighbours are folded into the running
// per-cluster sums strictly in adjacency order: within a chunk, the first
// lane holding a cluster adds the chunk's matching weights in lane order to
// the sum carried over from earlier chunks.  For hub vertices the
// own-cluster total follows curr_together_seq's pbbs block order.
__global__ __launch_bounds__(64) void k_bm_ordered(CsrView g, StateView st, MoveParams mp,
                                                   const uint32_t* __restrict__ fb_list,
                                                   const RoundCounters* __restrict__ rc,
                                                   MoveOut out, HashScratch hs) {
  uint32_t* keys = hs.keys + (size_t)blockIdx.x * hs.cap;
  double* vals = hs.vals + (size_t)blockIdx.x * hs.cap;
  const int lane = threadIdx.x;
  const int64_t count = (int64_t)rc->fb_count;
  for (int64_t i = blockIdx.x; i < count; i += gridDim.x) {
    const uint32_t v = fb_list[i];
    const int64_t start = g.off[v];
    const int64_t deg = g.off[v + 1] - start;
    const uint32_t cap = min(hs.cap, NextPow2((uint64_t)(2 * deg + 2)));
    const uint32_t mask = cap - 1;
    for (uint32_t s = lane; s < cap; s += 64) {
      keys[s] = kNoneId;
      vals[s] = 0.0;
    }
    __syncthreads();
    const uint32_t cv = st.ids[v];
    const bool hub = deg > kHubDegree;
    int64_t bs = 0;
    if (hub) {
      bs = (int64_t)ceil(sqrt((double)deg)) * 4;
      if (bs < 1024) bs = 1024;
    }
    double t_blk = 0.0, t_tot = 0.0;
    bool have_tot = false;
    for (int64_t s0 = 0; s0 < deg; s0 += 64) {
      const int64_t p = s0 + lane;
      uint32_t c = kNoneId;
      double x = 0.0;
      if (p < deg) {
        const uint32_t u = g.nbr[start + p];
        const float wf = g.w[start + p];
        if (hub) {
          float wh = wf;
          wh -= mp.offset;
          x = (double)wh;
        } else {
          x = (double)wf;
          x -= mp.offset;
        }
        if (u != v) c = st.ids[u];
      }
      int slot = -1;
      if (c != kNoneId) slot = HashFindOrInsert(keys, mask, c);
      __syncthreads();
      double acc = (slot >= 0) ? vals[slot] : 0.0;
      bool first = slot >= 0;
      const int nj = (int)min<int64_t>(64, deg - s0);
      for (int j = 0; j < nj; ++j) {
        const uint32_t cj = __shfl(c, j);
        const double xj = __shfl(x, j);
        if (slot >= 0 && cj == c) {
          acc += xj;
          if (j < lane) first = false;
        }
        if (hub) {
          const double aj = (cj != kNoneId && cj == cv) ? xj : 0.0;
          const int64_t pj = s0 + j;
          if (pj % bs == 0) {
            if (pj > 0) {
              t_tot = have_tot ? t_tot + t_blk : t_blk;
              have_tot = true;
            }
            t_blk = aj;
          } else {
            t_blk += aj;
          }
        }
      }
      if (first) vals[slot] = acc;
      __syncthreads();
    }
    double T;
    if (hub) {
      T = have_tot ? t_tot + t_blk : t_blk;
      if (deg == 0) T = 0.0;
    } else {
      T = HashGet(keys, vals, mask, cv);
    }
    const double k = st.nw[v];
    const double res = ResOf(mp, v);
    double change = 0;
    change -= NetWeight(T, k * (st.K[cv] - k), res);
    double mw = 0;
    mw += k;
    Cand best{0.0, kNoneId};
    for (uint32_t s = lane; s < cap; s += 64) {
      const uint32_t c = keys[s];
      if (c != kNoneId) {
        const Cand x{change + NetWeight(vals[s], mw * st.K[c], res), c};
        best = PickBetter(best, x);
      }
    }
    best = GroupBest(best, 64);
    if (lane == 0) {
      uint32_t t;
      double d;
      Decide(best, change, hub, mp.new_id, &t, &d);
      SyncEpilogue(v, t, d, mp.new_id, st, out);
    }
    __syncthreads();
  }
}

int GridFor(int64_t items, int per_block, int max_blocks) {
  int64_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (int)b;
}

}  // namespace

void LaunchBinFrontier(const CsrView& g, const uint8_t* frontier, const BinThresholds& th,
                       const BinLayout& layout, uint32_t* list, RoundCounters* rc, hipStream_t s) {
  if (g.n == 0) return;
  const int grid = GridFor(g.n, 256, 8192);
  hipLaunchKernelGGL(k_bin_frontier, dim3(grid), dim3(256), 0, s, g.n, g.off, frontier, th, layout,
                     list, rc);
  PCC_HIP_CHECK(hipGetLastError());
}

void LaunchBinCapacity(const CsrView& g, const BinThresholds& th, unsigned long long* caps,
                       hipStream_t s) {
  if (g.n == 0) return;
  const int grid = GridFor(g.n, 256, 4096);
  hipLaunchKernelGGL(k_bin_capacity, dim3(grid), dim3(256), 0, s, g.n, g.off, th, caps);
  PCC_HIP_CHECK(hipGetLastError());
}

template <int G, bool A>
static void LaunchSmall(int64_t count, const CsrView& g, const StateView& st, const MoveParams& mp,
                        const uint32_t* list, const MoveOut& out, uint64_t salt, hipStream_t s) {
  const int grid = GridFor(count, 256 / G, 1 << 20);
  hipLaunchKernelGGL((k_bm_small<G, A>), dim3(grid), dim3(256), 0, s, g, st, mp, list, count, out,
                     salt);
}

template <int THREADS, int TABLE, bool A>
static void LaunchHash(int64_t count, int max_grid, const CsrView& g, const StateView& st,
                       const MoveParams& mp, const uint32_t* list, const MoveOut& out,
                       uint32_t* fb_list, RoundCounters* rc, const HashScratch& hs, uint64_t salt,
                       hipStream_t s) {
  const int grid = GridFor(count, 1, max_grid);
  hipLaunchKernelGGL((k_bm_hash<THREADS, TABLE, A>), dim3(grid), dim3(THREADS), 0, s, g, st, mp,
                     list, count, out, fb_list, rc, hs, salt);
}

void LaunchBestMoveBin(int bin, int64_t count, bool async, const CsrView& g, const StateView& st,
                       const MoveParams& mp, const uint32_t* list, const MoveOut& out,
                       uint32_t* fb_list, RoundCounters* rc, const HashScratch& hs,
                       uint64_t salt, hipStream_t s) {
  if (count <= 0) return;
  switch (bin) {
    case 0:
      async ? LaunchSmall<4, true>(count, g, st, mp, list, out, salt, s)
            : LaunchSmall<4, false>(count, g, st, mp, list, out, salt, s);
      break;
    case 1:
      async ? LaunchSmall<8, true>(count, g, st, mp, list, out, salt, s)
            : LaunchSmall<8, false>(count, g, st, mp, list, out, salt, s);
      break;
    case 2:
      async ? LaunchSmall<16, true>(count, g, st, mp, list, out, salt, s)
            : LaunchSmall<16, false>(count, g, st, mp, list, out, salt, s);
      break;
    case 3:
      async ? LaunchSmall<32, true>(count, g, st, mp, list, out, salt, s)
            : LaunchSmall<32, false>(count, g, st, mp, list, out, salt, s);
      break;
    case 4:
      async ? LaunchSmall<64, true>(count, g, st, mp, list, out, salt, s)
            : LaunchSmall<64, false>(count, g, st, mp, list, out, salt, s);
      break;
    case 5:
      async ? LaunchHash<64, 1024, true>(count, 1 << 20, g, st, mp, list, out, fb_list, rc, hs, salt, s)
            : LaunchHash<64, 1024, false>(count, 1 << 20, g, st, mp, list, out, fb_list, rc, hs, salt, s);
      break;
    case 6:
      async ? LaunchHash<512, 8192, true>(count, 1 << 20, g, st, mp, list, out, fb_list, rc, hs, salt, s)
            : LaunchHash<512, 8192, false>(count, 1 << 20, g, st, mp, list, out, fb_list, rc, hs, salt, s);
      break;
    default:
      async ? LaunchHash<512, 0, true>(count, hs.regions, g, st, mp, list, out, fb_list, rc, hs, salt, s)
            : LaunchHash<512, 0, false>(count, hs.regions, g, st, mp, list, out, fb_list, rc, hs, salt, s);
      break;
  }
  PCC_HIP_CHECK(hipGetLastError());
}

void LaunchBestMoveOrdered(const CsrView& g, const StateView& st, const MoveParams& mp,
                           const uint32_t* fb_list, const RoundCounters* rc, const MoveOut& out,
                           const HashScratch& hs, hipStream_t s) {
  hipLaunchKernelGGL(k_bm_ordered, dim3(hs.regions), dim3(64), 0, s, g, st, mp, fb_list, rc, out,
                     hs);
  PCC_HIP_CHECK(hipGetLastError());
}

}  // namespace pcc
