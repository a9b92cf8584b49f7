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
//   bins 5-6: one wave / one workgroup per vertex, LDS open-addressing hash
//     with FP64 atomics.
//   bin 7: each vertex is split into hash partitions of its neighbour
//     clusters, one workgroup and one LDS table per partition; the
//     partitions' (cluster, sum) entries are then reduced by a selection
//     kernel, one workgroup per vertex.
//   In bins 5-7 every hash slot also counts its terms and flags terms finer
//   than the vertex's granularity threshold (ThresholdFor).  The sum of a
//   slot with three or more terms including a fine one may depend on the
//   summation order; those slots alone are recomputed in adjacency order (and
//   the hub-vertex block order of curr_together_seq for the own cluster) by
//   RepairInexactSlots.  Every other slot's sum is exact in any order, so
//   results are bit-identical to a sequential evaluation in every case.  The
//   ordered kernel k_bm_ordered evaluates whole vertices in adjacency order; it
//   serves the forced-ordered test mode and the (load-factor-excluded) case of
//   a full partition table.

#include <algorithm>

#include "device_common.h"
#include "kernels.h"

namespace pcc {

namespace {

// Upper degrees of the hash bins' ranges (a range can be empty when the
// thresholds are set below its nominal bound).
struct HashBounds {
  int64_t b[6];  // bins 5..10: deg <= b[bin - 5]
};
__host__ __device__ __forceinline__ HashBounds HashBoundsOf(int64_t wave_max, int64_t block_max) {
  HashBounds h;
  const int64_t w1 = wave_max < 127 ? wave_max : 127;
  int64_t w2 = wave_max < 255 ? wave_max : 255;
  w2 = w2 > w1 ? w2 : w1;
  int64_t k1 = block_max < 1023 ? block_max : 1023;
  k1 = k1 > wave_max ? k1 : wave_max;
  int64_t k2 = block_max < 2047 ? block_max : 2047;
  k2 = k2 > k1 ? k2 : k1;
  h.b[0] = w1;
  h.b[1] = w2;
  h.b[2] = wave_max;
  h.b[3] = k1;
  h.b[4] = k2;
  h.b[5] = block_max;
  return h;
}

__device__ __forceinline__ int BinOf(int64_t deg, int64_t wave_max, int64_t block_max) {
  if (deg <= 4) return 0;
  if (deg <= 8) return 1;
  if (deg <= 16) return 2;
  if (deg <= 32) return 3;
  if (deg <= 64) return 4;
  if (deg > block_max) return kHubBin;
  const HashBounds h = HashBoundsOf(wave_max, block_max);
  // The first bound >= deg names the bin (unrolled: a variable index into h
  // would put it in scratch memory).
  int b = kWaveBin0;
  bool above = true;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    above = above && deg > h.b[j];
    b += above ? 1 : 0;
  }
  return b;
}

__device__ __forceinline__ uint64_t LaneMaskLt() {
  const int l = LaneId();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ unsigned long long WaveSumU64(unsigned long long x) {
  for (int s = 32; s > 0; s >>= 1) x += __shfl_xor(x, s);
  return x;
}

// ------------------------------------------------------------- binning --

// Each workgroup bins a chunk of kBinChunk vertices: every thread counts its
// kBinItems vertices per bin in registers (counts packed four to a 64-bit
// word, 16 bits each: a chunk holds at most 4096 vertices), one block scan
// of the packed words ranks the threads within each bin, one global
// reservation per bin and chunk places the chunk, and the threads write
// their vertices (in thread order within a bin; nothing depends on the order
// of a bin's list).
constexpr int kBinItems = 16;
constexpr int kBinThreads = 256;
constexpr int64_t kBinChunk = (int64_t)kBinItems * kBinThreads;
static_assert(kBins <= 12 && kBinChunk < 65536, "three 64-bit words of 16-bit counts");
constexpr int64_t kBinSmallN = 65536;  // one vertex per thread and chunk up to here

__device__ __forceinline__ uint64_t PackedField(uint64_t w0, uint64_t w1, uint64_t w2, int b) {
  const uint64_t w = b < 4 ? w0 : (b < 8 ? w1 : w2);
  return (w >> (16 * (b & 3))) & 0xFFFFull;
}
__device__ __forceinline__ void PackedAdd(uint64_t& w0, uint64_t& w1, uint64_t& w2, int b) {
  const uint64_t inc = 1ull << (16 * (b & 3));
  w0 += b < 4 ? inc : 0ull;
  w1 += (b >= 4 && b < 8) ? inc : 0ull;
  w2 += b >= 8 ? inc : 0ull;
}

// ITEMS vertices per thread and chunk (kBinItems; 1 on small graphs, whose
// few chunks would otherwise leave all but a few CUs idle).
template <int ITEMS>
__global__ __launch_bounds__(kBinThreads) void k_bin_frontier(
    int64_t n, const int64_t* __restrict__ off, const uint8_t* __restrict__ frontier,
    BinThresholds th, BinLayout layout, uint32_t* __restrict__ list,
    RoundCounters* __restrict__ rc) {
  constexpr int64_t kChunk = (int64_t)ITEMS * kBinThreads;
  __shared__ unsigned long long s_scan[kBinThreads / kWave];
  __shared__ unsigned long long s_te[kBins], s_base[kBins];
  for (int64_t c0 = (int64_t)blockIdx.x * kChunk; c0 < n;
       c0 += (int64_t)gridDim.x * kChunk) {
    if (threadIdx.x < kBins) s_te[threadIdx.x] = 0;
    int8_t bin[ITEMS];
    unsigned long long dg[ITEMS];
    uint64_t p0 = 0, p1 = 0, p2 = 0;
    // All flags first, then all offsets of the listed ones, so that the
    // loads of a thread's vertices are in flight together rather than one
    // flag-then-offset chain after another.
    bool in[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const int64_t v = c0 + (int64_t)k * kBinThreads + threadIdx.x;
      in[k] = v < n && (frontier == nullptr || frontier[v] != 0);
    }
    int64_t o0[ITEMS], o1[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const int64_t v = c0 + (int64_t)k * kBinThreads + threadIdx.x;
      o0[k] = in[k] ? off[v] : 0;
      o1[k] = in[k] ? off[v + 1] : 0;
    }
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const unsigned long long deg = (unsigned long long)(o1[k] - o0[k]);
      const int b = in[k] ? BinOf((int64_t)deg, th.wave_max, th.block_max) : -1;
      bin[k] = (int8_t)b;
      dg[k] = deg;
      if (b >= 0) PackedAdd(p0, p1, p2, b);
    }
    // Thread ranks within each bin (the scans synchronize first, so s_te is
    // clear before the sums below).
    unsigned long long t0 = 0, t1 = 0, t2 = 0;
    uint64_t r0 = BlockExclusiveScan<unsigned long long>(p0, s_scan, &t0);
    uint64_t r1 = BlockExclusiveScan<unsigned long long>(p1, s_scan, &t1);
    uint64_t r2 = BlockExclusiveScan<unsigned long long>(p2, s_scan, &t2);
    // Degree sums per bin by LDS atomics (cheaper than twelve wave sums of
    // per-thread registers).
#pragma unroll
    for (int k = 0; k < ITEMS; ++k)
      if (bin[k] >= 0) atomicAdd(&s_te[bin[k]], dg[k]);
    __syncthreads();
    if (threadIdx.x < kBins) {
      const int j = threadIdx.x;
      const uint64_t cnt = PackedField(t0, t1, t2, j);
      s_base[j] = cnt ? atomicAdd(&rc->cnt[j], (unsigned long long)cnt) : 0ull;
      if (s_te[j]) atomicAdd(&rc->te[j], s_te[j]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const int b = bin[k];
      if (b < 0) continue;
      const int64_t v = c0 + (int64_t)k * kBinThreads + threadIdx.x;
      const uint64_t pos = PackedField(r0, r1, r2, b);
      list[layout.base[b] + (int64_t)s_base[b] + (int64_t)pos] = (uint32_t)v;
      PackedAdd(r0, r1, r2, b);
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_bin_capacity(int64_t n, const int64_t* __restrict__ off,
                                                      BinThresholds th, const double* __restrict__ nw,
                                                      const double* __restrict__ gres, int64_t G,
                                                      unsigned long long* __restrict__ caps) {
  __shared__ unsigned long long h[kLevelStats];
  if (threadIdx.x < kLevelStats) h[threadIdx.x] = 0;
  __syncthreads();
  unsigned long long dmax = 0;
  bool neg = false;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = base + threadIdx.x;
    int b = -1;
    unsigned long long deg = 0;
    if (v < n) {
      deg = (unsigned long long)(off[v + 1] - off[v]);
      b = BinOf((int64_t)deg, th.wave_max, th.block_max);
      dmax = deg > dmax ? deg : dmax;
      if (nw) neg = neg || !(nw[v] >= 0.0);
    }
    uint32_t present = WaveOr(b >= 0 ? (1u << b) : 0u);
    while (present) {
      const int bb = __ffs(present) - 1;
      present &= present - 1;
      const uint64_t mask = __ballot(b == bb);
      const unsigned long long te = WaveSumU64(b == bb ? deg : 0ull);
      if (LaneId() == 0) {
        atomicAdd(&h[bb], (unsigned long long)__popcll(mask));
        atomicAdd(&h[kBins + bb], te);
      }
    }
  }
  if (gres && blockIdx.x == 0)
    for (int64_t i = threadIdx.x; i < G; i += blockDim.x) neg = neg || !(gres[i] >= 0.0);
  for (int sh = 32; sh > 0; sh >>= 1) {
    const unsigned long long o = __shfl_xor(dmax, sh);
    dmax = o > dmax ? o : dmax;
  }
  const bool any_neg = __ballot(neg) != 0;
  if (LaneId() == 0) {
    if (dmax) atomicMax(&h[2 * kBins], dmax);
    if (any_neg) h[2 * kBins + 1] = 1;
  }
  __syncthreads();
  if (threadIdx.x < 2 * kBins && h[threadIdx.x]) atomicAdd(&caps[threadIdx.x], h[threadIdx.x]);
  if (threadIdx.x == 2 * kBins && h[threadIdx.x]) atomicMax(&caps[threadIdx.x], h[threadIdx.x]);
  if (threadIdx.x == 2 * kBins + 1 && h[threadIdx.x]) caps[threadIdx.x] = 1;
}

// Every bin-7 vertex appended to list (order unspecified).
__global__ __launch_bounds__(256) void k_collect_hubs(int64_t n, const int64_t* __restrict__ off,
                                                      int64_t block_max,
                                                      uint32_t* __restrict__ list,
                                                      unsigned long long* __restrict__ count) {
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = base + threadIdx.x;
    const bool hub = v < n && off[v + 1] - off[v] > block_max;
    const uint64_t mask = __ballot(hub);
    if (mask == 0) continue;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    unsigned long long pos = 0;
    if (LaneId() == leader) pos = atomicAdd(count, (unsigned long long)__popcll(mask));
    pos = __shfl(pos, leader);
    if (hub) list[pos + __popcll(mask & LaneMaskLt())] = (uint32_t)v;
  }
}

// Splitting the binned list by async sub-round: list item t of the bins'
// concatenation lies in bin b; its (b, SubRoundOf(v, S)) segment is counted
// (k_sub_count) and, after the host turned counts into segment starts in
// cur, filled (k_sub_scatter).  Each workgroup handles a slab of items and
// reserves its share of every segment with one atomic.
constexpr int kSubThreads = 256;
constexpr int64_t kSubSlab = 16 * kSubThreads;
constexpr int kMaxSub = 1024;

struct BinCounts {
  int64_t base[kBins];
  int64_t count[kBins];
};

__device__ __forceinline__ int BinOfItem(const BinCounts& bc, int64_t t, int64_t* r) {
  int b = 0;
  while (t >= bc.count[b]) t -= bc.count[b++];
  *r = t;
  return b;
}

// Also sums the degrees of each sub-round's bin-7 vertices into te7, which
// sizes that sub-round's hub launches.
__global__ __launch_bounds__(kSubThreads) void k_sub_count(const uint32_t* __restrict__ list,
                                                           BinCounts bc, int64_t total, uint32_t S,
                                                           const int64_t* __restrict__ off,
                                                           uint32_t* __restrict__ cnt,
                                                           unsigned long long* __restrict__ te7) {
  __shared__ uint32_t h[kBins * kMaxSub];
  __shared__ unsigned long long h7[kMaxSub];
  for (uint32_t k = threadIdx.x; k < kBins * S; k += kSubThreads) h[k] = 0;
  for (uint32_t k = threadIdx.x; k < S; k += kSubThreads) h7[k] = 0;
  __syncthreads();
  for (int64_t t = (int64_t)blockIdx.x * kSubThreads + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * kSubThreads) {
    int64_t r;
    const int b = BinOfItem(bc, t, &r);
    const uint32_t v = list[bc.base[b] + r];
    const uint32_t j = SubRoundOf(v, S);
    atomicAdd(&h[b * S + j], 1u);
    if (b == kBins - 1) atomicAdd(&h7[j], (unsigned long long)(off[v + 1] - off[v]));
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < kBins * S; k += kSubThreads)
    if (h[k]) atomicAdd(&cnt[k], h[k]);
  for (uint32_t k = threadIdx.x; k < S; k += kSubThreads)
    if (h7[k]) atomicAdd(&te7[k], h7[k]);
}

__global__ __launch_bounds__(kSubThreads) void k_sub_scatter(const uint32_t* __restrict__ list,
                                                             BinCounts bc, int64_t total,
                                                             uint32_t S,
                                                             uint32_t* __restrict__ cur,
                                                             uint32_t* __restrict__ out) {
  __shared__ uint32_t h[kBins * kMaxSub];
  __shared__ uint32_t hb[kBins * kMaxSub];
  for (int64_t s0 = (int64_t)blockIdx.x * kSubSlab; s0 < total;
       s0 += (int64_t)gridDim.x * kSubSlab) {
    const int64_t s1 = min(total, s0 + kSubSlab);
    for (uint32_t k = threadIdx.x; k < kBins * S; k += kSubThreads) h[k] = 0;
    __syncthreads();
    for (int64_t t = s0 + threadIdx.x; t < s1; t += kSubThreads) {
      int64_t r;
      const int b = BinOfItem(bc, t, &r);
      atomicAdd(&h[b * S + SubRoundOf(list[bc.base[b] + r], S)], 1u);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < kBins * S; k += kSubThreads) {
      hb[k] = h[k] ? atomicAdd(&cur[k], h[k]) : 0u;
      h[k] = 0;
    }
    __syncthreads();
    for (int64_t t = s0 + threadIdx.x; t < s1; t += kSubThreads) {
      int64_t r;
      const int b = BinOfItem(bc, t, &r);
      const uint32_t v = list[bc.base[b] + r];
      const uint32_t k = b * S + SubRoundOf(v, S);
      out[bc.base[b] + hb[k] + atomicAdd(&h[k], 1u)] = v;
    }
    __syncthreads();
  }
}

// The value EfficientBestMove adds for an edge of weight wf: (double)w -
// offset on the low-degree path, the float w -= offset on the hub path
// (PCCI.cc:466, 529).
__device__ __forceinline__ double EdgeValue(float wf, bool hub, double offset) {
  if (hub) {
    float wh = wf;
    wh -= offset;
    return (double)wh;
  }
  double x = (double)wf;
  x -= offset;
  return x;
}

// gscale value of a vertex with more than 64 and at most kLongRow edges
// (rows [vs, vs + vd)), computed by one wave: PackScale(ThresholdFor(B),
// MagExpFor(M)) with B and M the sum and the maximum over its edges of
// |w| + |offset|.
__device__ __forceinline__ int32_t WaveGranScale(const float* __restrict__ w, double offset,
                                                 double ao, int64_t vs, int64_t vd) {
  // Besides B and M, the smallest granularity of the values the kernels
  // add (EdgeValue): with none below the threshold the vertex gets
  // kNoFineThr.
  const bool hub = vd > kHubDegree;
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0, mx = 0;
  int gmin = 1 << 20;
  int64_t p = LaneId();
  for (; p + 3 * kWave < vd; p += 4 * kWave) {
    const float w0 = w[vs + p], w1 = w[vs + p + kWave];
    const float w2 = w[vs + p + 2 * kWave], w3 = w[vs + p + 3 * kWave];
    const double a0 = fabs((double)w0) + ao;
    const double a1 = fabs((double)w1) + ao;
    const double a2 = fabs((double)w2) + ao;
    const double a3 = fabs((double)w3) + ao;
    s0 += a0;
    s1 += a1;
    s2 += a2;
    s3 += a3;
    mx = fmax(mx, fmax(fmax(a0, a1), fmax(a2, a3)));
    gmin = min(gmin, min(min(GranExp(EdgeValue(w0, hub, offset)), GranExp(EdgeValue(w1, hub, offset))),
                         min(GranExp(EdgeValue(w2, hub, offset)), GranExp(EdgeValue(w3, hub, offset)))));
  }
  for (; p < vd; p += kWave) {
    const float wq = w[vs + p];
    const double a = fabs((double)wq) + ao;
    s0 += a;
    mx = fmax(mx, a);
    gmin = min(gmin, GranExp(EdgeValue(wq, hub, offset)));
  }
  const double s = WaveSum((s0 + s1) + (s2 + s3));
  for (int sh = 32; sh > 0; sh >>= 1) mx = fmax(mx, __shfl_xor(mx, sh));
  gmin = WaveMinInt(gmin);
  int thr = ThresholdFor(s);
  if (!(gmin < thr)) thr = kNoFineThr;
  return PackScale(thr, MagExpFor(mx));
}

// gscale[v] for deg > 64.  Rows longer than kLongRow are queued for
// k_gran_long; the others are done by one wave each, here (midq null: the
// wave of their 64 vertices in turn) or, queued on midq, by k_gran_mid, which
// spreads them over all waves (vertices of similar id, and so of similar
// degree on RMAT-like graphs, would otherwise load a few waves with all the
// large rows).
__global__ __launch_bounds__(256) void k_gran_threshold(int64_t n, const int64_t* __restrict__ off,
                                                        const float* __restrict__ w, double offset,
                                                        int32_t* __restrict__ gscale,
                                                        uint32_t* __restrict__ longq,
                                                        unsigned int* __restrict__ nlong,
                                                        uint32_t* __restrict__ midq,
                                                        unsigned int* __restrict__ nmid) {
  const double ao = fabs(offset);
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = base + threadIdx.x;
    int64_t st = 0, deg = 0;
    if (v < n) {
      st = off[v];
      deg = off[v + 1] - st;
      if (deg > kLongRow) longq[atomicAdd(nlong, 1u)] = (uint32_t)v;  // k_gran_long
    }
    uint64_t big = __ballot(v < n && deg > 64 && deg <= kLongRow);
    if (midq) {
      if (big) {
        const int leader = __ffsll((unsigned long long)big) - 1;
        unsigned int pos = 0;
        if (LaneId() == leader) pos = atomicAdd(nmid, (unsigned int)__popcll(big));
        pos = __shfl(pos, leader);
        if ((big >> LaneId()) & 1ull)
          midq[pos + __popcll(big & LaneMaskLt())] = (uint32_t)v;
      }
      continue;
    }
    const int64_t wave_v0 = base + (threadIdx.x & ~(kWave - 1));
    while (big) {
      const int l = __ffsll((unsigned long long)big) - 1;
      big &= big - 1;
      const int32_t sc = WaveGranScale(w, offset, ao, __shfl(st, l), __shfl(deg, l));
      if (LaneId() == 0) gscale[wave_v0 + l] = sc;
    }
  }
}

// The queued rows of k_gran_threshold, one wave each.
__global__ __launch_bounds__(256) void k_gran_mid(const int64_t* __restrict__ off,
                                                  const float* __restrict__ w, double offset,
                                                  const uint32_t* __restrict__ midq,
                                                  const unsigned int* __restrict__ nmid,
                                                  int32_t* __restrict__ gscale) {
  const double ao = fabs(offset);
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / kWave);
  const int64_t count = (int64_t)*nmid;
  for (int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; q < count; q += nwaves) {
    const uint32_t v = midq[q];
    const int64_t vs = off[v];
    const int32_t sc = WaveGranScale(w, offset, ao, vs, off[v + 1] - vs);
    if (LaneId() == 0) gscale[v] = sc;
  }
}

// The long-row accumulators of k_gran_long: sums and maxima zero, granularity
// minima above every exponent (> 2^20), no row queued.
__global__ __launch_bounds__(256) void k_gran_clear(int64_t maxlong, unsigned int* __restrict__ nlong,
                                                    double* __restrict__ acc_s,
                                                    unsigned long long* __restrict__ acc_m,
                                                    int* __restrict__ acc_g) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < maxlong;
       i += (int64_t)gridDim.x * blockDim.x) {
    acc_s[i] = 0.0;
    acc_m[i] = 0ull;
    acc_g[i] = 0x7F7F7F7F;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *nlong = 0u;
}

// Rows longer than kLongRow: chunk sums and maxima combined with atomics (the
// bound B only has to hold, not to be reproducible; ThresholdFor leaves a
// margin for the rounding of any summation order), then packed.
__global__ __launch_bounds__(256) void k_gran_long(const int64_t* __restrict__ off,
                                                   const float* __restrict__ w, double offset,
                                                   const uint32_t* __restrict__ longq,
                                                   const unsigned int* __restrict__ nlong,
                                                   double* __restrict__ acc_s,
                                                   unsigned long long* __restrict__ acc_m,
                                                   int* __restrict__ acc_g) {
  __shared__ double sm[256 / kWave];
  const double ao = fabs(offset);
  ForLongChunks(longq, *nlong, off, [&](uint32_t q, uint32_t, int64_t e0, int64_t e1) {
    double s = 0, mx = 0;
    int gmin = 1 << 20;
    for (int64_t p = e0 + threadIdx.x; p < e1; p += blockDim.x) {
      const float wp = w[p];
      const double a = fabs((double)wp) + ao;
      s += a;
      mx = fmax(mx, a);
      gmin = min(gmin, GranExp(EdgeValue(wp, true, offset)));  // rows > kLongRow are hubs
    }
    s = BlockReduce(s, sm, [](double x, double y) { return x + y; });
    mx = BlockReduce(mx, sm, [](double x, double y) { return fmax(x, y); });
    const double gm = BlockReduce((double)gmin, sm, [](double x, double y) { return fmin(x, y); });
    if (threadIdx.x == 0) {
      atomicAdd(&acc_s[q], s);
      // Non-negative doubles order like their bit patterns.
      atomicMax(&acc_m[q], (unsigned long long)__double_as_longlong(mx));
      atomicMin(&acc_g[q], (int)gm);
    }
  });
}

__global__ __launch_bounds__(256) void k_gran_long_final(const uint32_t* __restrict__ longq,
                                                         const unsigned int* __restrict__ nlong,
                                                         const double* __restrict__ acc_s,
                                                         const unsigned long long* __restrict__ acc_m,
                                                         const int* __restrict__ acc_g,
                                                         int32_t* __restrict__ gscale) {
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < *nlong; q += gridDim.x * blockDim.x) {
    int thr = ThresholdFor(acc_s[q]);
    if (!(acc_g[q] < thr)) thr = kNoFineThr;  // no fine value in the row
    gscale[longq[q]] = PackScale(thr, MagExpFor(__longlong_as_double((long long)acc_m[q])));
  }
}

// Bin-7 prefix sums of a list, one workgroup.
constexpr int kScanThreads = 1024;
constexpr int kHubScanItems = 8;
// THREADS = kScanThreads, or one wave for short lists (no barriers to wait on).
template <int THREADS>
__global__ __launch_bounds__(THREADS) void k_hub_scan(const int64_t* __restrict__ off,
                                                           const uint32_t* __restrict__ list,
                                                           int64_t count, int64_t part_edges,
                                                           int64_t* __restrict__ part_off,
                                                           int64_t* __restrict__ ent_off,
                                                           uint32_t* __restrict__ item_of,
                                                           const unsigned long long* __restrict__ dcount) {
  if (dcount) count = (int64_t)*dcount;
  __shared__ int64_t sp[THREADS / kWave], se[THREADS / kWave];
  // kHubScanItems consecutive items per thread and pass, their list entries
  // and then their row offsets loaded together.
  constexpr int IT = kHubScanItems;
  int64_t carry_p = 0, carry_e = 0;
  for (int64_t b0 = 0; b0 < count; b0 += (int64_t)THREADS * IT) {
    const int64_t i0 = b0 + (int64_t)threadIdx.x * IT;
    uint32_t v[IT];
#pragma unroll
    for (int k = 0; k < IT; ++k) v[k] = i0 + k < count ? list[i0 + k] : 0u;
    int64_t deg[IT], parts[IT];
    int64_t lp = 0, le = 0;
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      deg[k] = i0 + k < count ? off[v[k] + 1] - off[v[k]] : 0;
      parts[k] = i0 + k < count ? (deg[k] > 0 ? (deg[k] + part_edges - 1) / part_edges : 1) : 0;
      lp += parts[k];
      le += deg[k];
    }
    int64_t tp = 0, te = 0;
    int64_t rp = carry_p + BlockExclusiveScan(lp, sp, &tp);
    int64_t re = carry_e + BlockExclusiveScan(le, se, &te);
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      if (i0 + k >= count) break;
      part_off[i0 + k] = rp;
      ent_off[i0 + k] = re;
      for (int64_t q = 0; q < parts[k]; ++q) item_of[rp + q] = (uint32_t)(i0 + k);
      rp += parts[k];
      re += deg[k];
    }
    carry_p += tp;
    carry_e += te;
  }
  if (threadIdx.x == 0) {
    part_off[count] = carry_p;
    ent_off[count] = carry_e;
  }
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
  // Empty ids are searched within the vertex's own graph: v itself first, as
  // the reference's probe does, then the round's list of empty ids, then the
  // probe v + 1, v + 2, ... (PCCI.cc:387-399).
  uint32_t lo = 0, hi = mp.n, gi = 0;
  if (mp.vgraph) {
    gi = mp.vgraph[v];
    lo = (uint32_t)mp.gbase[gi];
    hi = (uint32_t)mp.gbase[gi + 1];
  }
  auto claim = [&](uint32_t i) {
    if (__hip_atomic_load(&st.sizes[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
        atomicCAS(&st.sizes[i], 0u, 1u) != 0u)
      return false;
    atomicSub(&st.sizes[cur], 1u);
    atomicAdd(&st.K[cur], -1 * k);
    atomicAdd(&st.K[i], k);
    st.ids[v] = i;
    out.moved_clusters[i] = 1;
    out.moved[v] = 1;
    return true;
  };
  if (claim(v)) return;
  if (mp.free_ids) {
    while (true) {
      const unsigned long long x = atomicAdd(&mp.free_cur[gi], 1ull);
      if (x >= mp.free_end[gi]) break;
      if (claim(mp.free_ids[x])) return;
    }
  }
  uint32_t i = v;
  for (uint32_t probe = lo; probe < hi; ++probe) {
    i = (i + 1 == hi) ? lo : i + 1;
    if (claim(i)) return;
  }
  out.moved[v] = 0;
}

// Sub-round of a vertex in async mode (a fixed hash of the id).
__device__ __forceinline__ bool InSubRound(const MoveParams& mp, uint32_t v) {
  if (mp.sub_count <= 1) return true;
  return SubRoundOf(v, mp.sub_count) == mp.sub_index;
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

// G lanes per listed vertex; workgroup `block` of `nblocks` working on the
// list (a launch of its own, or one bin's range of the fused launch).
template <int G, bool kAsync>
__device__ __forceinline__ void SmallBinBody(const CsrView& g, const StateView& st,
                                             const MoveParams& mp,
                                             const uint32_t* __restrict__ list, int64_t count,
                                             const MoveOut& out, uint64_t salt, int64_t block,
                                             int64_t nblocks) {
  constexpr int kGroups = 256 / G;
  const int lane = threadIdx.x & (G - 1);
  const int wlane = LaneId();
  const int gbase_lane = wlane & ~(G - 1);
  const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1ull) << gbase_lane);
  for (int64_t gb = XcdItem(block, nblocks) * kGroups; gb < count;
       gb += nblocks * kGroups) {
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
    // (These vertices are not hubs, so a pruned candidate is the same as none:
    // both decide "new cluster" with delta = change.)
    if (first && !(mp.prune && S <= 0.0)) {
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

template <int G, bool kAsync>
__global__ __launch_bounds__(256) void k_bm_small(CsrView g, StateView st, MoveParams mp,
                                                  const uint32_t* __restrict__ list,
                                                  int64_t count, MoveOut out, uint64_t salt,
                                                  const unsigned long long* __restrict__ dcount) {
  if (dcount) count = (int64_t)*dcount;
  SmallBinBody<G, kAsync>(g, st, mp, list, count, out, salt, blockIdx.x, gridDim.x);
}

// Bins 0-4 in one launch: workgroups [end[b-1], end[b]) take bin b (ranges
// of whole multiples of 8 workgroups, so a workgroup's index within its bin
// keeps the XCD it has in the grid).  In sync rounds the bins read the same
// state and write disjoint vertices, so running them together changes no
// result; it saves four dependent launches per round on small graphs.
// Workgroups from end[4] on do the round's clears (RoundResetJob).
struct SmallBins {
  const uint32_t* list[kWaveBin0];
  int64_t count[kWaveBin0];
  const unsigned long long* dcount[kWaveBin0];
  uint32_t end[kWaveBin0];
  RoundResetJob reset;
  uint32_t reset_blocks;
};

template <bool kAsync>
__global__ __launch_bounds__(256) void k_bm_small_all(CsrView g, StateView st, MoveParams mp,
                                                      SmallBins sb, MoveOut out, uint64_t salt) {
  if (blockIdx.x >= sb.end[kWaveBin0 - 1]) {
    const RoundResetJob& r = sb.reset;
    RoundResetSpan((int64_t)(blockIdx.x - sb.end[kWaveBin0 - 1]) * blockDim.x + threadIdx.x,
                   (int64_t)sb.reset_blocks * blockDim.x, r.n, r.a, r.b, (uint8_t*)nullptr, r.rc,
                   r.counters, (unsigned long long*)nullptr, kBins);
    return;
  }
  int b = 0;
  while (b < kWaveBin0 - 1 && blockIdx.x >= sb.end[b]) ++b;
  const uint32_t b0 = b ? sb.end[b - 1] : 0u;
  const int64_t blk = (int64_t)blockIdx.x - b0, nblk = (int64_t)sb.end[b] - b0;
  const int64_t count = sb.dcount[b] ? (int64_t)*sb.dcount[b] : sb.count[b];
  switch (b) {
    case 0: SmallBinBody<4, kAsync>(g, st, mp, sb.list[0], count, out, salt, blk, nblk); break;
    case 1: SmallBinBody<8, kAsync>(g, st, mp, sb.list[1], count, out, salt, blk, nblk); break;
    case 2: SmallBinBody<16, kAsync>(g, st, mp, sb.list[2], count, out, salt, blk, nblk); break;
    case 3: SmallBinBody<32, kAsync>(g, st, mp, sb.list[3], count, out, salt, blk, nblk); break;
    default: SmallBinBody<64, kAsync>(g, st, mp, sb.list[4], count, out, salt, blk, nblk); break;
  }
}

// ------------------------------------------------------------ hash bins --

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

// Bounded insert: the slot of c, or -1 when all cap slots hold other keys.
// *fresh tells whether this call stored the key.
template <typename KeyPtr>
__device__ __forceinline__ int HashSlot(KeyPtr keys, uint32_t mask, uint32_t c, bool* fresh) {
  uint32_t h = HashId(c) & mask;
  *fresh = false;
  // Compare-and-swap first: one LDS round trip per probe whether the slot is
  // free, holds c, or holds another key.
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    const uint32_t old = atomicCAS(&keys[h], kNoneId, c);
    if (old == kNoneId) {
      *fresh = true;
      return (int)h;
    }
    if (old == c) return (int)h;
    h = (h + 1) & mask;
  }
  return -1;
}

// Occupied slots of an LDS table in insertion order, occ[0 .. *n), at most
// cap_list of them; a table whose list fills up reports an overflow.
struct SlotList {
  uint16_t* occ;
  uint32_t* n;
  uint32_t cap_list;
  // Appends the fresh slots of a wave with one counter atomic per wave (a
  // shared counter hit by every inserting lane serializes); called by all
  // lanes of the wave.  False for a lane whose slot did not fit.
  __device__ __forceinline__ bool AddWave(bool fresh, int slot) const {
    const uint64_t fm = __ballot(fresh);
    if (fm == 0) return true;
    const int leader = __ffsll((unsigned long long)fm) - 1;
    uint32_t base = 0;
    if (LaneId() == leader) base = atomicAdd(n, (uint32_t)__popcll(fm));
    base = __shfl(base, leader);
    if (!fresh) return true;
    const uint32_t k = base + (uint32_t)__popcll(fm & LaneMaskLt());
    if (k >= cap_list) return false;
    occ[k] = (uint16_t)slot;
    return true;
  }
};

template <typename KeyPtr>
__device__ __forceinline__ int HashFind(KeyPtr keys, uint32_t mask, uint32_t c) {
  uint32_t h = HashId(c) & mask;
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    const uint32_t k = keys[h];
    if (k == c) return (int)h;
    if (k == kNoneId) return -1;
    h = (h + 1) & mask;
  }
  return -1;
}


// Block size of the pbbs reduce of curr_together_seq (hub path, PCCI.cc:545):
// max(1024, 4 * ceil(sqrt(deg))).
__device__ __forceinline__ int64_t PbbsBlock(int64_t deg) {
  const int64_t bs = (int64_t)ceil(sqrt((double)deg)) * 4;
  return bs < 1024 ? 1024 : bs;
}

// Chunks of THREADS edges whose gathers the order repair issues together
// (one for the 1024-thread kernel, whose 128 VGPRs per thread are taken).
template <int THREADS>
constexpr int kRepairChunks = THREADS >= 1024 ? 1 : 8;

// LDS staging of the order repair: one chunk of THREADS edges.
template <int THREADS>
// Per staged edge: its table slot (< 2^16), its pbbs block of the hub path's
// own-cluster order (position / PbbsBlock(deg) < 2^16 for any degree below
// 2^32) and its weight as stored (the fold recomputes EdgeValue).
struct RepairStage {
  uint16_t slot[THREADS];
  uint16_t blk[THREADS];
  float w[THREADS];
  uint32_t wsum[THREADS / kWave];
  double T;
};

// Recomputes, in adjacency order, the sums of the slots whose summation order
// matters (SlotInexact), after Accumulate filled the table.  Edges of those
// slots are compacted chunk by chunk in adjacency order and folded by one
// thread, so each such slot ends up with exactly the sequential sum.  own is
// the slot of c(v) (or -1); the return value is the own-cluster total: on the
// hub path the pbbs block order of curr_together_seq (block sums folded left,
// PCCI.cc:516-545), otherwise the slot's sum.
template <int THREADS, class Src>
__device__ __forceinline__ double RepairInexactSlots(const CsrView& g, const MoveParams& mp,
                                                  Src src, int64_t start, int64_t deg, bool hub,
                                                  const uint32_t* keys, double* vals,
                                                  const uint32_t* cnt, uint32_t cap,
                                                  uint32_t parts, uint32_t part, int own, int thr,
                                                  int mexp, const uint16_t* occ, uint32_t nocc,
                                                  RepairStage<THREADS>* stg) {
  const uint32_t mask = cap - 1;
  for (uint32_t k = threadIdx.x; k < nocc; k += THREADS) {
    const uint32_t s = occ[k];
    if (SlotInexact(cnt[s], thr, mexp)) vals[s] = 0.0;
  }
  __syncthreads();
  const bool own_ordered = hub && own >= 0 && SlotInexact(cnt[own], thr, mexp);
  const int64_t bs = PbbsBlock(deg);
  const int lane = LaneId(), w = threadIdx.x / kWave;
  // Fold state of thread 0.
  double t_blk = 0.0, t_tot = 0.0;
  int64_t cur_blk = -1;
  bool have_tot = false;
  // U chunks of THREADS edges per trip: their neighbour and cluster gathers
  // are issued together (a hub row is long and its repair would otherwise
  // pay two dependent global loads per chunk), then the chunks are folded
  // one after another in adjacency order.
  constexpr int U = kRepairChunks<THREADS>;
  for (int64_t q0 = 0; q0 < deg; q0 += (int64_t)THREADS * U) {
    uint32_t c[U];
    int slot[U];
    float x[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int64_t q = q0 + (int64_t)k * THREADS + threadIdx.x;
      c[k] = q < deg ? src.Neighbour(q) : kNoneId;
    }
#pragma unroll
    for (int k = 0; k < U; ++k) c[k] = c[k] != kNoneId ? src.Cluster(c[k]) : kNoneId;
    bool any = false;
#pragma unroll
    for (int k = 0; k < U; ++k) {
      slot[k] = -1;
      if (c[k] != kNoneId && (parts == 1 || PartOf(c[k], parts) == part)) {
        const int s = HashFind(keys, mask, c[k]);
        if (s >= 0 && SlotInexact(cnt[s], thr, mexp)) slot[k] = s;
      }
      any = any || slot[k] >= 0;
    }
    if (!__syncthreads_or(any)) continue;
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int64_t q = q0 + (int64_t)k * THREADS + threadIdx.x;
      x[k] = slot[k] >= 0 ? g.w[start + q] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
      if (!__syncthreads_or(slot[k] >= 0)) continue;
      const uint64_t m = __ballot(slot[k] >= 0);
      if (lane == 0) stg->wsum[w] = (uint32_t)__popcll(m);
      __syncthreads();
      uint32_t before = 0, total = 0;
      for (int i = 0; i < THREADS / kWave; ++i) {
        const uint32_t t = stg->wsum[i];
        if (i < w) before += t;
        total += t;
      }
      if (slot[k] >= 0) {
        const uint32_t j = before + (uint32_t)__popcll(m & LaneMaskLt());
        stg->slot[j] = (uint16_t)slot[k];
        // Only the own cluster's edges on the hub path need their block.
        stg->blk[j] = (own_ordered && slot[k] == own)
                          ? (uint16_t)((uint32_t)(q0 + (int64_t)k * THREADS + threadIdx.x) /
                                       (uint32_t)bs)
                          : (uint16_t)0;
        stg->w[j] = x[k];
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        for (uint32_t j = 0; j < total; ++j) {
          const uint32_t sl = stg->slot[j];
          const double xx = EdgeValue(stg->w[j], hub, mp.offset);
          vals[sl] += xx;
          if (own_ordered && (int)sl == own) {
            const int64_t blk = stg->blk[j];
            if (blk != cur_blk) {
              if (cur_blk >= 0) {
                t_tot = have_tot ? t_tot + t_blk : t_blk;
                have_tot = true;
              }
              cur_blk = blk;
              t_blk = xx;
            } else {
              t_blk += xx;
            }
          }
        }
      }
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    double T = 0.0;
    if (own_ordered) {
      T = t_tot;
      if (cur_blk >= 0) T = have_tot ? t_tot + t_blk : t_blk;
    } else if (own >= 0) {
      T = vals[own];
    }
    stg->T = T;
  }
  __syncthreads();
  return stg->T;
}

// Counts a term of a slot (capped) and records how fine it is.
__device__ __forceinline__ void RecordTerm(uint32_t* cnt, int slot, double x, int thr) {
  if (thr == kNoFineThr) return;  // every total of the vertex is exact
  if ((cnt[slot] & kCountMask) < kCountCap) atomicAdd(&cnt[slot], 1u);
  const uint32_t fb = FineBits(GranExp(x), thr);
  if (fb) atomicOr(&cnt[slot], fb);
}

// Where the hashed kernels take the cluster of v's q-th neighbour from: the
// adjacency and the cluster ids, one gather per edge (kNoneId for
// self-loops).
struct AdjacencyClusters {
  const uint32_t* nbr;  // v's adjacency
  const uint32_t* ids;
  uint32_t v;
  __device__ __forceinline__ uint32_t Neighbour(int64_t q) const { return nbr[q]; }
  __device__ __forceinline__ uint32_t Cluster(uint32_t u) const {
    return u != v ? ids[u] : kNoneId;
  }
};

// Accumulate over one partition's bucket of (cluster, weight bits) pairs:
// every pair belongs to the partition, so no filtering is needed.  The first
// P * THREADS pairs arrive in pre (loaded before the table was cleared); the
// rest are loaded Q * THREADS at a time, all of a trip's loads issued before
// its inserts.
template <int THREADS, bool kAsync, int P>
__device__ __forceinline__ bool AccumulateBucket(const MoveParams& mp, const uint2 (&pre)[P],
                                                 const uint2* __restrict__ bk, uint32_t nb,
                                                 bool hub, int thr, uint32_t* keys, double* vals,
                                                 uint32_t* cnt, uint32_t mask, const SlotList& sl) {
  bool overflow = false;
  auto insert = [&](bool act, uint2 e) {
    const double x = act ? EdgeValue(__uint_as_float(e.y), hub, mp.offset) : 0.0;
    bool fresh = false;
    const int slot = act ? HashSlot(keys, mask, e.x, &fresh) : -1;
    const bool listed = sl.AddWave(fresh, slot);  // every lane of the wave
    if (!act) return;
    if (slot < 0 || !listed) {
      overflow = true;
      return;
    }
    atomicAdd(&vals[slot], x);
    if (!kAsync) RecordTerm(cnt, slot, x, thr);
  };
#pragma unroll
  for (int r = 0; r < P; ++r) insert(r * THREADS + threadIdx.x < nb, pre[r]);
  constexpr int Q = 4;
  for (uint32_t k0 = P * THREADS; k0 < nb; k0 += Q * THREADS) {
    uint2 e[Q];
#pragma unroll
    for (int r = 0; r < Q; ++r) {
      const uint32_t k = k0 + r * THREADS + threadIdx.x;
      e[r] = k < nb ? bk[k] : make_uint2(kNoneId, 0u);
    }
#pragma unroll
    for (int r = 0; r < Q; ++r) insert(k0 + r * THREADS + threadIdx.x < nb, e[r]);
  }
  return overflow;
}


// Per-vertex values the hashed kernels read besides the adjacency.
struct HashVertex {
  uint32_t v = 0;
  int64_t start = 0, deg = 0;
  uint32_t cv = 0;   // current cluster
  double k = 0;      // node weight
  int32_t scale = 0; // order-free test scale (sync)
  double Kcv = 0;    // weight of the current cluster
};

// Bins 5 and 6: THREADS cooperating threads and a TABLE-entry LDS table per
// vertex.  Each thread holds E = TABLE / (2 THREADS) edges of a vertex in
// registers (the launch ranges keep deg <= TABLE / 2), and the workgroup runs
// a software pipeline over its vertices: the adjacency and per-vertex loads of
// the next vertex are issued before the current vertex's inserts, its cluster
// gathers before the current vertex's decision, and (sync) thread 0 finishes
// a vertex's epilogue while the next one is being inserted, so the memory
// latencies overlap work instead of adding up vertex after vertex.  In sync
// mode the state is fixed for the round, so reading ahead changes nothing; in
// async mode it is one more interleaving of the concurrent moves.
// At least 4 waves per SIMD (128 VGPRs), so the LDS budget, not registers,
// sets how many workgroups share a CU.
// The items of a hash launch: one bin's list, or (sync rounds with few
// hashed vertices) the lists of several bins one after another, so that one
// launch of the largest table covers them; a vertex's result does not depend
// on which table size handles it.
struct ItemLists {
  const uint32_t* list[kHubBin - kWaveBin0];
  int64_t end[kHubBin - kWaveBin0];  // cumulative item counts
  int n;                             // lists in use
  __device__ __forceinline__ uint32_t At(int64_t i) const {
    int b = 0;
    while (b < n - 1 && i >= end[b]) ++b;
    return list[b][i - (b ? end[b - 1] : 0)];
  }
};
// One bin's list.
struct OneList {
  const uint32_t* list;
  __device__ __forceinline__ uint32_t At(int64_t i) const { return list[i]; }
};

template <int THREADS, int TABLE, bool kAsync, typename Items>
__device__ __forceinline__ void HashBinBody(CsrView g, StateView st, MoveParams mp, Items il,
                                            int64_t count, int64_t dlo, int64_t dhi, MoveOut out,
                                            uint32_t* __restrict__ fb_list,
                                            RoundCounters* __restrict__ rc, uint64_t salt,
                                            const unsigned long long* __restrict__ dcount) {
  if (dcount) count = (int64_t)*dcount;
  constexpr int NW = THREADS / kWave;
  constexpr int E = TABLE / (2 * THREADS);
  static_assert(E >= 1 && E * THREADS * 2 == TABLE, "table holds two slots per edge slot");
  constexpr uint32_t kList = TABLE * 3 / 4;
  __shared__ uint32_t keys[TABLE];
  __shared__ double vals[TABLE];
  __shared__ uint32_t cnt[TABLE];
  __shared__ uint16_t occ[kList];
  __shared__ uint32_t nocc;
  __shared__ double sd[NW];
  __shared__ uint32_t sc[NW];
  __shared__ RepairStage<THREADS> stg;
  // The table is cleared once; after each vertex only its occupied slots are
  // reset, so a vertex costs O(its distinct clusters), not O(TABLE).
  for (uint32_t s = threadIdx.x; s < TABLE; s += THREADS) {
    keys[s] = kNoneId;
    vals[s] = 0.0;
    cnt[s] = 0;
  }
  if (threadIdx.x == 0) nocc = 0;
  __syncthreads();
  const SlotList sl{occ, &nocc, kList};

  // The workgroup's next vertex in (dlo, dhi] at or after item i (uniform).
  // Vertices routed to the ordered kernel are queued on the way.
  auto next_item = [&](int64_t& i, HashVertex& h) -> bool {
    for (; i < count; i += gridDim.x) {
      const uint32_t v = il.At(PermuteIndex(i, count, salt));
      if (kAsync && !InSubRound(mp, v)) continue;
      const int64_t start = g.off[v];
      const int64_t deg = g.off[v + 1] - start;
      if (deg <= dlo || deg > dhi) continue;  // another launch's degree range
      if (deg > (int64_t)E * THREADS || (!kAsync && mp.force_ordered)) {
        if (threadIdx.x == 0) {
          if (kAsync)
            out.moved[v] = 0;
          else
            fb_list[atomicAdd(&rc->fb_count, 1ull)] = v;
        }
        continue;
      }
      h.v = v;
      h.start = start;
      h.deg = deg;
      return true;
    }
    return false;
  };
  auto load_vertex = [&](bool ok, HashVertex& h, uint32_t* nb, float* wf) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int64_t q = threadIdx.x + (int64_t)e * THREADS;
      const bool in = ok && q < h.deg;
      nb[e] = in ? g.nbr[h.start + q] : kNoneId;
      wf[e] = in ? g.w[h.start + q] : 0.0f;
    }
    if (ok) {
      h.cv = st.ids[h.v];
      h.k = st.nw[h.v];
      h.scale = kAsync ? 0 : mp.gscale[h.v];
    }
  };
  auto gather = [&](bool ok, HashVertex& h, const uint32_t* nb, uint32_t* cc) {
#pragma unroll
    for (int e = 0; e < E; ++e)
      cc[e] = (nb[e] != kNoneId && nb[e] != h.v) ? st.ids[nb[e]] : kNoneId;
    if (ok) h.Kcv = st.K[h.cv];
  };
  // Sync epilogue of a decided vertex, split so that its loads are issued one
  // stage before their use (SyncEpilogue's rule).
  bool pend = false;
  uint32_t pv = 0, pt = 0, pcv = 0, psz_t = 0, psz_c = 0;
  double pd = 0.0;

  // Consecutive list items (nearby vertex ids, which share neighbourhoods in
  // RMAT-like graphs) go to workgroups of one XCD.
  int64_t i = XcdItem(blockIdx.x, gridDim.x);
  HashVertex cur, nxt;
  bool have = next_item(i, cur);
  uint32_t cc[E];
  float wf[E];
  {
    uint32_t nb[E];
    load_vertex(have, cur, nb, wf);
    gather(have, cur, nb, cc);
  }
  int64_t i2 = i + gridDim.x;
  bool have2 = have && next_item(i2, nxt);
  while (have) {
    // Stage 1: the next vertex's adjacency and values; the pending
    // epilogue's cluster sizes.
    uint32_t nb2[E];
    float wf2[E];
    load_vertex(have2, nxt, nb2, wf2);
    if (!kAsync && pend && threadIdx.x == 0) {
      psz_t = pt != mp.new_id ? st.sizes[pt] : 0u;
      psz_c = st.sizes[pcv];
    }
    // Stage 2: this vertex's inserts (probe range sized for load <= 1/2).
    const int64_t deg = cur.deg;
    const uint32_t cap = min((uint32_t)TABLE, max(64u, NextPow2((uint64_t)(2 * deg + 2))));
    const uint32_t mask = cap - 1;
    const bool hub = deg > kHubDegree;
    const int thr = ScaleThr(cur.scale), mexp = ScaleMexp(cur.scale);
    bool overflow = false;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const bool act = cc[e] != kNoneId;
      const double x = act ? EdgeValue(wf[e], hub, mp.offset) : 0.0;
      bool fresh = false;
      const int slot = act ? HashSlot(keys, mask, cc[e], &fresh) : -1;
      const bool listed = sl.AddWave(fresh, slot);
      if (!act) continue;
      if (slot < 0 || !listed) {
        overflow = true;
        continue;
      }
      atomicAdd(&vals[slot], x);
      if (!kAsync) RecordTerm(cnt, slot, x, thr);
    }
    const bool ovf = __syncthreads_or(overflow);
    // Stage 3: the next vertex's cluster gathers, the item after it, and the
    // pending epilogue.
    uint32_t cc2[E];
    gather(have2, nxt, nb2, cc2);
    int64_t i3 = i2 + gridDim.x;
    HashVertex nx2;
    const bool have3 = have2 && next_item(i3, nx2);
    if (!kAsync && pend && threadIdx.x == 0) {
      uint32_t t = pt;
      double d = pd;
      if (t != mp.new_id && psz_t == 1 && psz_c == 1 && pcv >= t) {
        t = pcv;
        d = 0.0;
      }
      const bool mv = d > 0;
      out.moves[pv] = mv ? t : kNoneId;
      out.moved[pv] = mv ? 1 : 0;
      pend = false;
    }
    // Stage 4: this vertex's decision.
    const uint32_t v = cur.v;
    const uint32_t n_occ = min(nocc, kList);
    if (ovf) {  // cannot happen at load <= 1/2; kept as a guard
      if (threadIdx.x == 0) {
        if (kAsync)
          out.moved[v] = 0;
        else
          fb_list[atomicAdd(&rc->fb_count, 1ull)] = v;
      }
    } else {
      const uint32_t cv = cur.cv;
      const int own = HashFind(keys, mask, cv);
      double T = own >= 0 ? vals[own] : 0.0;
      if (!kAsync && thr != kNoFineThr) {
        bool inexact = false;
        for (uint32_t k = threadIdx.x; k < n_occ; k += THREADS)
          inexact = inexact || SlotInexact(cnt[occ[k]], thr, mexp);
        if (__syncthreads_or(inexact)) {
          const AdjacencyClusters src{g.nbr + cur.start, st.ids, v};
          T = RepairInexactSlots<THREADS>(g, mp, src, cur.start, deg, hub, keys, vals, cnt, cap, 1,
                                          0, own, thr, mexp, occ, n_occ, &stg);
          if (threadIdx.x == 0) atomicAdd(&rc->repair_count, 1ull);
        }
      }
      const double k = cur.k;
      const double res = ResOf(mp, v);
      double change = 0;
      change -= NetWeight(T, k * (cur.Kcv - k), res);
      double mw = 0;
      mw += k;
      // At most E slots per thread (n_occ <= deg <= E * THREADS): their
      // cluster weights are loaded together.
      uint32_t c[E];
      double Kc[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const uint32_t q = threadIdx.x + (uint32_t)e * THREADS;
        c[e] = q < n_occ ? keys[occ[q]] : kNoneId;
        if (c[e] != kNoneId && mp.prune && vals[occ[q]] <= 0.0) c[e] = kNoneId;  // cannot win
        Kc[e] = c[e] != kNoneId ? st.K[c[e]] : 0.0;
      }
      Cand best{0.0, kNoneId};
#pragma unroll
      for (int e = 0; e < E; ++e) {
        if (c[e] == kNoneId) continue;
        const uint32_t q = threadIdx.x + (uint32_t)e * THREADS;
        const Cand x{change + NetWeight(vals[occ[q]], mw * Kc[e], res), c[e]};
        best = PickBetter(best, x);
      }
      best = BlockBest<THREADS>(best, sd, sc);
      if (threadIdx.x == 0) {
        // Only pruned candidates: a candidate exists, none beats staying.
        if (best.c == kNoneId && n_occ > 0) best = Cand{-INFINITY, 0u};
        uint32_t t;
        double d;
        Decide(best, change, hub, mp.new_id, &t, &d);
        if (kAsync) {
          AsyncEpilogue(v, t, d, mp, st, out);
        } else {
          if (out.dbg_target) out.dbg_target[v] = t;
          if (out.dbg_delta) out.dbg_delta[v] = d;
          pend = true;
          pv = v;
          pt = t;
          pd = d;
          pcv = cv;
        }
      }
    }
    __syncthreads();
    // Back to an empty table.  An overflowing vertex may have stored keys
    // past the list; then the whole probe range is reset.
    if (ovf) {
      for (uint32_t s = threadIdx.x; s < cap; s += THREADS) {
        keys[s] = kNoneId;
        vals[s] = 0.0;
        cnt[s] = 0;
      }
    } else {
      for (uint32_t q = threadIdx.x; q < n_occ; q += THREADS) {
        const uint32_t s = occ[q];
        keys[s] = kNoneId;
        vals[s] = 0.0;
        cnt[s] = 0;
      }
    }
    if (threadIdx.x == 0) nocc = 0;
    __syncthreads();
    // Rotate the pipeline.
    have = have2;
    cur = nxt;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      cc[e] = cc2[e];
      wf[e] = wf2[e];
    }
    i2 = i3;
    have2 = have3;
    nxt = nx2;
  }
  if (!kAsync && pend && threadIdx.x == 0)
    SyncEpilogue(pv, pt, pd, mp.new_id, st, out);
}

template <int THREADS, int TABLE, bool kAsync>
__global__ __launch_bounds__(THREADS, 4) void k_bm_hash(CsrView g, StateView st, MoveParams mp,
                                                     const uint32_t* __restrict__ list,
                                                     int64_t count, int64_t dlo, int64_t dhi,
                                                     MoveOut out,
                                                     uint32_t* __restrict__ fb_list,
                                                     RoundCounters* __restrict__ rc,
                                                     uint64_t salt,
                                                     const unsigned long long* __restrict__ dcount) {
  HashBinBody<THREADS, TABLE, kAsync>(g, st, mp, OneList{list}, count, dlo, dhi, out, fb_list, rc,
                                      salt, dcount);
}

// Sync rounds: the lists of several bins in one launch.
template <int THREADS, int TABLE>
__global__ __launch_bounds__(THREADS, 4) void k_bm_hash_lists(CsrView g, StateView st, MoveParams mp,
                                                           ItemLists il, int64_t count,
                                                           int64_t dlo, int64_t dhi, MoveOut out,
                                                           uint32_t* __restrict__ fb_list,
                                                           RoundCounters* __restrict__ rc) {
  HashBinBody<THREADS, TABLE, false>(g, st, mp, il, count, dlo, dhi, out, fb_list, rc, 0, nullptr);
}

// Accumulate over v's whole adjacency (partition `part` of `parts`), the
// cluster of every neighbour gathered again: the fallback of a partition
// whose bucket overflowed (a heavy cluster), rare enough not to need the
// compaction queue of the buffered form.
template <int THREADS, bool kAsync>
__device__ __forceinline__ bool AccumulateGathered(const CsrView& g, const StateView& st,
                                                   const MoveParams& mp, uint32_t v, int64_t start,
                                                   int64_t deg, bool hub, int thr, uint32_t* keys,
                                                   double* vals, uint32_t* cnt, uint32_t mask,
                                                   uint32_t parts, uint32_t part,
                                                   const SlotList& sl,
                                                   const uint32_t* __restrict__ cb) {
  bool overflow = false;
  // U chunks of THREADS edges per trip: all their neighbour loads, then all
  // their cluster gathers, then the inserts, so a long row costs one pair of
  // memory round trips per U chunks rather than per chunk.
  constexpr int U = 8;
  for (int64_t q0 = 0; q0 < deg; q0 += (int64_t)U * THREADS) {
    uint32_t c[U];
    float wf[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int64_t q = q0 + (int64_t)k * THREADS + threadIdx.x;
      c[k] = q < deg ? (cb ? cb[q] : g.nbr[start + q]) : kNoneId;
      wf[k] = q < deg ? g.w[start + q] : 0.0f;
    }
    if (!cb) {
#pragma unroll
      for (int k = 0; k < U; ++k) c[k] = (c[k] != kNoneId && c[k] != v) ? st.ids[c[k]] : kNoneId;
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const bool act = c[k] != kNoneId && (parts == 1 || PartOf(c[k], parts) == part);
      const double x = act ? EdgeValue(wf[k], hub, mp.offset) : 0.0;
      bool fresh = false;
      const int slot = act ? HashSlot(keys, mask, c[k], &fresh) : -1;
      const bool listed = sl.AddWave(fresh, slot);
      if (!act) continue;
      if (slot < 0 || !listed) {
        overflow = true;
        continue;
      }
      atomicAdd(&vals[slot], x);
      if (!kAsync) RecordTerm(cnt, slot, x, thr);
    }
  }
  return overflow;
}

// Bin 7, gather pass: the cluster of every neighbour of the listed vertices,
// one workgroup per partition of edges, appended with its weight to the
// bucket of the partition the cluster hashes to (HubScratch::bkt): counted
// per bucket in LDS, then one global reservation per bucket and workgroup.
// The partition pass then reads only its bucket; the rare passes that need
// a vertex's whole adjacency in order (an overflowing bucket, the in-order
// repair of inexact sums) gather the clusters again.
constexpr int kGatherThreads = 256;
constexpr int kGatherU = 8;
constexpr int kGatherMaxBuckets = 4096;

template <bool kAsync>
__global__ __launch_bounds__(kGatherThreads) void k_hub_gather(CsrView g, StateView st,
                                                               MoveParams mp,
                                                               const uint32_t* __restrict__ list,
                                                               int64_t count, HubScratch hb,
                                                               const unsigned long long* __restrict__ dcount) {
  if (dcount) count = (int64_t)*dcount;
  __shared__ uint32_t lh[kGatherMaxBuckets];
  if (!kAsync && mp.force_ordered) return;
  const int64_t t = XcdChunkItem(blockIdx.x);
  if (t >= hb.part_off[count]) return;
  const int64_t i = hb.item_of[t];
  const uint32_t v = list[i];
  const int64_t p0 = hb.part_off[i];
  const int64_t parts = hb.part_off[i + 1] - p0;
  const int64_t part = t - p0;
  const int64_t start = g.off[v], deg = g.off[v + 1] - start;
  const int32_t scale = kAsync ? 0 : mp.gscale[v];
  // Every item's record, also of vertices outside the sub-round: the
  // partition pass tests the sub-round on the record's vertex.
  if (threadIdx.x == 0) {
    HubPart h;
    h.start = start;
    h.deg = deg;
    h.v = v;
    h.item = (uint32_t)i;
    h.part = (uint32_t)part;
    h.parts = (uint32_t)parts;
    h.scale = scale;
    h.cv = st.ids[v];
    h.kv = st.nw[v];
    h.res = ResOf(mp, v);
    hb.hpart[t] = h;
  }
  if (kAsync && !InSubRound(mp, v)) return;
  const int64_t per = (deg + parts - 1) / parts;
  const int64_t e0 = part * per, e1 = min(deg, (part + 1) * per);
  const bool local = parts <= kGatherMaxBuckets;
  // Sync rounds flag the partitions that receive a fine value (bin-7
  // vertices take the hub path's float values, EdgeValue(w, true, offset)).
  const int thr = kAsync ? kNoFineThr : ScaleThr(scale);
  constexpr int U = kGatherU;
  for (int64_t q0 = e0; q0 < e1; q0 += (int64_t)kGatherThreads * U) {
    uint32_t c[U];
    float w[U];
    uint32_t b[U], pos[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int64_t q = q0 + threadIdx.x + (int64_t)k * kGatherThreads;
      c[k] = q < e1 ? g.nbr[start + q] : v;
      w[k] = q < e1 ? g.w[start + q] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < U; ++k) c[k] = c[k] != v ? st.ids[c[k]] : kNoneId;
    if (kAsync && hb.cbuf) {
      uint32_t* cb = hb.cbuf + hb.ent_off[i];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const int64_t q = q0 + threadIdx.x + (int64_t)k * kGatherThreads;
        if (q < e1) cb[q] = c[k];
      }
    }
#pragma unroll
    for (int k = 0; k < U; ++k) b[k] = c[k] != kNoneId ? PartOf(c[k], (uint32_t)parts) : 0u;
    if (local) {
      __syncthreads();
      for (int64_t j = threadIdx.x; j < parts; j += kGatherThreads) lh[j] = 0;
      __syncthreads();
#pragma unroll
      for (int k = 0; k < U; ++k)
        if (c[k] != kNoneId) pos[k] = atomicAdd(&lh[b[k]], 1u);
      __syncthreads();
      for (int64_t j = threadIdx.x; j < parts; j += kGatherThreads) {
        const uint32_t n = lh[j];
        lh[j] = n ? atomicAdd(&hb.bcur[p0 + j], n) : 0u;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < U; ++k)
        if (c[k] != kNoneId) pos[k] += lh[b[k]];
    } else {
#pragma unroll
      for (int k = 0; k < U; ++k)
        if (c[k] != kNoneId) pos[k] = atomicAdd(&hb.bcur[p0 + b[k]], 1u);
    }
#pragma unroll
    for (int k = 0; k < U; ++k)
      if (c[k] != kNoneId && pos[k] < (uint32_t)hb.bcap)
        hb.bkt[(p0 + b[k]) * hb.bcap + pos[k]] = make_uint2(c[k], __float_as_uint(w[k]));
    if (thr != kNoFineThr) {
#pragma unroll
      for (int k = 0; k < U; ++k)
        if (c[k] != kNoneId && GranExp(EdgeValue(w[k], true, mp.offset)) < thr)
          atomicOr(&hb.pfine[p0 + b[k]], 1u);
    }
  }
}

// Top two candidate values (see HubScratch::px1): x1 with the smallest
// cluster holding it, x2 the largest value strictly below x1.
struct Top2 {
  double x1 = -INFINITY;
  uint32_t c1 = kNoneId;
  double x2 = -INFINITY;
};
// An empty Top2 (c1 = kNoneId) has x1 = x2 = -inf, so the general rule covers
// it: written with selects, no divergent branches.
__device__ __forceinline__ Top2 MergeTop2(Top2 a, Top2 b) {
  const bool ga = a.x1 > b.x1, gb = b.x1 > a.x1;
  Top2 r;
  r.x1 = ga ? a.x1 : b.x1;
  r.c1 = ga ? a.c1 : (gb ? b.c1 : min(a.c1, b.c1));
  r.x2 = ga ? fmax(a.x2, b.x1) : (gb ? fmax(b.x2, a.x1) : fmax(a.x2, b.x2));
  return r;
}
// MergeTop2(t, Top2{x, c, -inf}).
__device__ __forceinline__ Top2 AddTop2(Top2 t, double x, uint32_t c) {
  const bool gt = x > t.x1, eq = x == t.x1;
  Top2 r;
  r.x1 = gt ? x : t.x1;
  r.c1 = gt ? c : (eq ? min(t.c1, c) : t.c1);
  r.x2 = gt ? t.x1 : (eq ? t.x2 : fmax(t.x2, x));
  return r;
}
__device__ __forceinline__ Top2 WaveTop2(Top2 x) {
  for (int s = kWave >> 1; s > 0; s >>= 1) {
    Top2 y;
    y.x1 = __shfl_xor(x.x1, s);
    y.c1 = __shfl_xor(x.c1, s);
    y.x2 = __shfl_xor(x.x2, s);
    x = MergeTop2(x, y);
  }
  return x;
}
template <int THREADS>
__device__ __forceinline__ Top2 BlockTop2(Top2 x, Top2* sm) {
  x = WaveTop2(x);
  __syncthreads();
  if (LaneId() == 0) sm[threadIdx.x / kWave] = x;
  __syncthreads();
  Top2 r = sm[0];
  for (int i = 1; i < THREADS / kWave; ++i) r = MergeTop2(r, sm[i]);
  return r;
}

// Bin 7, partition pass: work item t (partition p of list item i) builds the
// LDS table of v's neighbour clusters in partition p and reduces them to the
// partition's top two candidate values; the partition holding c(v) records
// S_{c(v)}.
constexpr int kHubThreads = 512;

template <int TABLE, bool kAsync>
__global__ __launch_bounds__(kHubThreads) void k_bm_hub_part(CsrView g, StateView st,
                                                             MoveParams mp,
                                                             const uint32_t* __restrict__ list,
                                                             int64_t count, HubScratch hb,
                                                             RoundCounters* __restrict__ rc,
                                                             const unsigned long long* __restrict__ dcount) {
  if (dcount) count = (int64_t)*dcount;
  __shared__ uint32_t keys[TABLE];
  __shared__ double vals[TABLE];
  __shared__ uint32_t cnt[TABLE];
  __shared__ RepairStage<kHubThreads> stg;
  constexpr uint32_t kList = TABLE * 3 / 4;
  __shared__ uint16_t occ[kList];
  __shared__ uint32_t nocc;
  if (!kAsync && mp.force_ordered) return;
  const int64_t t = XcdChunkItem(blockIdx.x);
  if (t >= hb.part_off[count]) return;
  // One round of loads: the partition's record, its bucket's fill and fine
  // flag, and its first P * kHubThreads pairs, speculatively (slots past the
  // fill, within the bucket's capacity, hold stale pairs and are masked).
  const HubPart hp = hb.hpart[t];
  const uint32_t nb = hb.bcur[t];
  const uint32_t fine = hb.pfine[t];
  const uint2* bk = hb.bkt + t * hb.bcap;
  constexpr int P = 2048 / kHubThreads;
  uint2 pre[P];
#pragma unroll
  for (int r = 0; r < P; ++r) {
    const int64_t k = (int64_t)r * kHubThreads + threadIdx.x;
    pre[r] = k < hb.bcap ? bk[k] : make_uint2(kNoneId, 0u);
  }
  // The whole table is cleared (its probe range needs the record).
  for (uint32_t s = threadIdx.x; s < (uint32_t)TABLE; s += kHubThreads) {
    keys[s] = kNoneId;
    vals[s] = 0.0;
    cnt[s] = 0;
  }
  if (threadIdx.x == 0) nocc = 0;
  const uint32_t v = hp.v;
  if (kAsync && !InSubRound(mp, v)) return;
  const uint32_t cv = hp.cv;
  const double kv = hp.kv;
  __syncthreads();
  const int64_t i = hp.item;
  const uint32_t parts = hp.parts;
  const uint32_t part = hp.part;
  const int64_t start = hp.start;
  const int64_t deg = hp.deg;
  const int64_t per = (deg + parts - 1) / parts;
  const uint32_t cap = min((uint32_t)TABLE, max(64u, NextPow2((uint64_t)(2 * per + 2))));
  const uint32_t mask = cap - 1;
  const SlotList sl{occ, &nocc, kList};
  const bool hub = deg > kHubDegree;
  const int32_t scale = kAsync ? 0 : hp.scale;
  // A partition that received no fine value keeps no per-slot records.
  const int thr = (kAsync || fine != 0) ? ScaleThr(scale) : kNoFineThr;
  const int mexp = ScaleMexp(scale);
  const AdjacencyClusters src{g.nbr + start, st.ids, v};
  // The partition's bucket when it did not overflow; otherwise the whole
  // adjacency, gathered and filtered.
  const bool overflow =
      nb <= (uint32_t)hb.bcap
          ? AccumulateBucket<kHubThreads, kAsync, P>(mp, pre, bk, nb, hub, thr, keys, vals, cnt,
                                                     mask, sl)
          : AccumulateGathered<kHubThreads, kAsync>(g, st, mp, v, start, deg, hub, thr, keys, vals,
                                                    cnt, mask, parts, part, sl,
                                                    kAsync && hb.cbuf ? hb.cbuf + hb.ent_off[i]
                                                                      : nullptr);
  const bool ovf = __syncthreads_or(overflow);
  // Every thread has read bcur[t] and pfine[t] before the barrier.
  if (threadIdx.x == 0) {
    hb.bcur[t] = 0;
    hb.pfine[t] = 0;
  }
  if (ovf) {
    if (threadIdx.x == 0) atomicOr(&hb.hflag[i], 2u);
    return;
  }
  const int own = PartOf(cv, parts) == part ? HashFind(keys, mask, cv) : -1;
  double T = own >= 0 ? vals[own] : 0.0;
  const uint32_t n_occ = min(nocc, kList);
  if (!kAsync && thr != kNoFineThr) {
    bool inexact = false;
    for (uint32_t k = threadIdx.x; k < n_occ; k += kHubThreads)
      inexact = inexact || SlotInexact(cnt[occ[k]], thr, mexp);
    if (__syncthreads_or(inexact)) {
      T = RepairInexactSlots<kHubThreads>(g, mp, src, start, deg, hub, keys, vals, cnt, cap,
                                          parts, part, own, thr, mexp, occ, n_occ, &stg);
      if (threadIdx.x == 0) atomicAdd(&rc->repair_count, 1ull);
    }
  }
  if (threadIdx.x == 0 && own >= 0) {
    hb.tval[i] = T;
    atomicOr(&hb.hflag[i], 1u);
  }
  // The partition's candidates, reduced to the top two values of
  // x_c = NetWeight(S_c, k_v K_c, res); the decision adds the own-cluster
  // term, which only the partition holding c(v) knows, afterwards.
  const double res = hp.res;
  double mw = 0;
  mw += kv;
  // A thread's (at most KS) candidate clusters have their weights loaded
  // together, not one dependent load per loop trip.
  constexpr int KS = (int)((kList + kHubThreads - 1) / kHubThreads);
  uint32_t cc[KS];
  double sv[KS], kc[KS];
#pragma unroll
  for (int e = 0; e < KS; ++e) {
    const uint32_t k = threadIdx.x + (uint32_t)e * kHubThreads;
    cc[e] = kNoneId;
    sv[e] = 0.0;
    kc[e] = 0.0;
    if ((uint32_t)e * kHubThreads >= n_occ) continue;  // the same for the whole workgroup
    if (k < n_occ) {
      const uint32_t s = occ[k];
      sv[e] = vals[s];
      if (!(mp.prune && sv[e] <= 0.0)) cc[e] = keys[s];  // cannot win (see MoveParams::prune)
    }
    kc[e] = cc[e] != kNoneId ? st.K[cc[e]] : 0.0;
  }
  Top2 top;
#pragma unroll
  for (int e = 0; e < KS; ++e) {
    if ((uint32_t)e * kHubThreads >= n_occ) continue;
    if (cc[e] != kNoneId) top = AddTop2(top, NetWeight(sv[e], mw * kc[e], res), cc[e]);
  }
  top = BlockTop2<kHubThreads>(top, reinterpret_cast<Top2*>(&stg));
  if (threadIdx.x == 0) {
    hb.px1[t] = top.x1;
    hb.pc1[t] = top.c1;
    hb.px2[t] = top.x2;
    if (n_occ > 0) atomicOr(&hb.hflag[i], 4u);  // a candidate exists
  }
}

// Own-cluster term of EfficientBestMove for a bin-7 list item.
__device__ __forceinline__ double HubChange(const StateView& st, const MoveParams& mp,
                                            const HubScratch& hb, int64_t i, uint32_t v,
                                            uint32_t flag) {
  const uint32_t cv = st.ids[v];
  const double T = (flag & 1u) ? hb.tval[i] : 0.0;
  const double k = st.nw[v];
  double change = 0;
  change -= NetWeight(T, k * (st.K[cv] - k), ResOf(mp, v));
  return change;
}

// Bin 7, decision pass: one wave per list item merges the partitions' top two
// values, adds the own-cluster term and applies EfficientBestMove's decision:
// the best delta is fl(change + x1) for the smallest cluster with x = x1,
// unless a smaller x2 rounds to the same delta, in which case a larger id
// with a smaller x could tie; such vertices (sync) go to the ordered kernel.
// Then the sync or async epilogue runs and the item's flags are reset.
template <bool kAsync>
__global__ __launch_bounds__(256) void k_bm_hub_final(CsrView g, StateView st, MoveParams mp,
                                                      const uint32_t* __restrict__ list,
                                                      int64_t count, HubScratch hb, MoveOut out,
                                                      uint32_t* __restrict__ fb_list,
                                                      RoundCounters* __restrict__ rc,
                                                      const unsigned long long* __restrict__ dcount) {
  if (dcount) count = (int64_t)*dcount;
  const int lane = LaneId();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * blockDim.x / kWave;
  for (int64_t i = wave; i < count; i += nwaves) {
    const uint32_t v = list[i];
    const uint32_t flag = hb.hflag[i];
    if (lane == 0) hb.hflag[i] = 0;
    if (kAsync && !InSubRound(mp, v)) continue;
    if (!kAsync && (mp.force_ordered || (flag & 2u))) {
      if (lane == 0) fb_list[atomicAdd(&rc->fb_count, 1ull)] = v;
      continue;
    }
    if (kAsync && (flag & 2u)) {  // table overflow: no move this time
      if (lane == 0) out.moved[v] = 0;
      continue;
    }
    Top2 top;
    for (int64_t t = hb.part_off[i] + lane; t < hb.part_off[i + 1]; t += kWave)
      top = MergeTop2(top, Top2{hb.px1[t], hb.pc1[t], hb.px2[t]});
    top = WaveTop2(top);
    if (lane == 0) {
      const int64_t deg = g.off[v + 1] - g.off[v];
      const double change = HubChange(st, mp, hb, i, v, flag);
      Cand best{0.0, kNoneId};
      bool tied = false;
      if (top.c1 != kNoneId) {
        best = Cand{change + top.x1, top.c1};
        tied = top.x2 > -INFINITY && change + top.x2 == best.d;
      } else if (flag & 4u) {
        best = Cand{-INFINITY, 0u};  // only pruned candidates: none beats staying
      }
      if (!kAsync && tied) {
        fb_list[atomicAdd(&rc->fb_count, 1ull)] = v;
        continue;
      }
      uint32_t tt;
      double d;
      Decide(best, change, deg > kHubDegree, mp.new_id, &tt, &d);
      if (kAsync)
        AsyncEpilogue(v, tt, d, mp, st, out);
      else
        SyncEpilogue(v, tt, d, mp.new_id, st, out);
    }
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
  if (g.n <= kBinSmallN) {
    hipLaunchKernelGGL(k_bin_frontier<1>, dim3(GridFor(g.n, kBinThreads, 2048)), dim3(kBinThreads),
                       0, s, g.n, g.off, frontier, th, layout, list, rc);
  } else {
    hipLaunchKernelGGL(k_bin_frontier<kBinItems>, dim3(GridFor(g.n, (int)kBinChunk, 2048)),
                       dim3(kBinThreads), 0, s, g.n, g.off, frontier, th, layout, list, rc);
  }
  PCC_HIP_CHECK(hipGetLastError());
}

void LaunchBinCapacity(const CsrView& g, const BinThresholds& th, const double* nw,
                       const double* gres, int64_t G, unsigned long long* caps, hipStream_t s) {
  if (g.n == 0 && !gres) return;
  const int grid = GridFor(g.n, g.n <= kBinSmallN ? 256 : 256 * 16, 1024);
  hipLaunchKernelGGL(k_bin_capacity, dim3(grid), dim3(256), 0, s, g.n, g.off, th, nw, gres, G,
                     caps);
  PCC_HIP_CHECK(hipGetLastError());
}

void LaunchSubCount(const CsrView& g, const uint32_t* list, const BinLayout& layout,
                    const int64_t counts[kBins], uint32_t S, uint32_t* cnt,
                    unsigned long long* te7, hipStream_t s) {
  BinCounts bc;
  int64_t total = 0;
  for (int b = 0; b < kBins; ++b) {
    bc.base[b] = layout.base[b];
    bc.count[b] = counts[b];
    total += counts[b];
  }
  if (total == 0) return;
  hipLaunchKernelGGL(k_sub_count, dim3(GridFor(total, (int)kSubSlab, 1024)), dim3(kSubThreads), 0, s,
                     list, bc, total, S, g.off, cnt, te7);
  PCC_HIP_CHECK(hipGetLastError());
}

void LaunchSubScatter(const uint32_t* list, const BinLayout& layout, const int64_t counts[kBins],
                      uint32_t S, uint32_t* cur, uint32_t* out, hipStream_t s) {
  BinCounts bc;
  int64_t total = 0;
  for (int b = 0; b < kBins; ++b) {
    bc.base[b] = layout.base[b];
    bc.count[b] = counts[b];
    total += counts[b];
  }
  if (total == 0) return;
  hipLaunchKernelGGL(k_sub_scatter, dim3(GridFor(total, (int)kSubSlab, 1024)), dim3(kSubThreads), 0,
                     s, list, bc, total, S, cur, out);
  PCC_HIP_CHECK(hipGetLastError());
}

void LaunchGranThreshold(DeviceAllocator* al, const CsrView& g, double offset, int32_t* gscale,
                         hipStream_t s, int64_t max_deg) {
  if (g.n == 0) return;
  // Long rows exist only above kLongRow edges (k_gran_threshold queues them).
  const bool longs = g.m > kLongRow && (max_deg < 0 || max_deg > kLongRow);
  const int64_t maxlong = longs ? g.m / kLongRow + 1 : 1;
  DevBuf<uint32_t> longq(al, maxlong);
  DevBuf<unsigned int> nlong(al, 1);
  DevBuf<double> acc_s(al, maxlong);
  DevBuf<unsigned long long> acc_m(al, maxlong);
  DevBuf<int> acc_g(al, maxlong);
  if (longs)
    hipLaunchKernelGGL(k_gran_clear, dim3(GridFor(maxlong, 256, 1024)), dim3(256), 0, s, maxlong,
                       nlong.get(), acc_s.get(), acc_m.get(), acc_g.get());
  const int grid = GridFor(g.n, 256, 8192);
  // Large graphs queue their rows of 65..kLongRow edges for k_gran_mid.
  const bool mid = g.n > kBinSmallN;
  DevBuf<uint32_t> midq;
  DevBuf<unsigned int> nmid;
  if (mid) {
    midq.Reset(al, g.n);
    nmid.Reset(al, 1);
    ZeroSpans z{};
    z.p[0] = reinterpret_cast<uint8_t*>(nmid.get());
    z.bytes[0] = sizeof(unsigned int);
    z.n = 1;
    LaunchZeroSpans(z, s);
  }
  hipLaunchKernelGGL(k_gran_threshold, dim3(grid), dim3(256), 0, s, g.n, g.off, g.w, offset,
                     gscale, longq.get(), nlong.get(), mid ? midq.get() : nullptr,
                     mid ? nmid.get() : nullptr);
  if (mid)
    hipLaunchKernelGGL(k_gran_mid, dim3(2048), dim3(256), 0, s, g.off, g.w, offset, midq.get(),
                       nmid.get(), gscale);
  if (longs) {
    hipLaunchKernelGGL(k_gran_long, dim3(1024), dim3(256), 0, s, g.off, g.w, offset, longq.get(),
                       nlong.get(), acc_s.get(), acc_m.get(), acc_g.get());
    hipLaunchKernelGGL(k_gran_long_final, dim3(16), dim3(256), 0, s, longq.get(), nlong.get(),
                       acc_s.get(), acc_m.get(), acc_g.get(), gscale);
  }
  PCC_HIP_CHECK(hipGetLastError());
}

void LaunchHubScan(const CsrView& g, const uint32_t* list, int64_t count, int64_t part_edges,
                   int64_t* part_off, int64_t* ent_off, uint32_t* item_of, hipStream_t s,
                   const unsigned long long* dcount) {
  if (count <= (int64_t)kWave * kHubScanItems)
    hipLaunchKernelGGL(k_hub_scan<kWave>, dim3(1), dim3(kWave), 0, s, g.off, list, count,
                       part_edges, part_off, ent_off, item_of, dcount);
  else
    hipLaunchKernelGGL(k_hub_scan<kScanThreads>, dim3(1), dim3(kScanThreads), 0, s, g.off, list,
                       count, part_edges, part_off, ent_off, item_of, dcount);
  PCC_HIP_CHECK(hipGetLastError());
}

void LaunchCollectHubs(const CsrView& g, const BinThresholds& th, uint32_t* list,
                       unsigned long long* count, hipStream_t s) {
  if (g.n == 0) return;
  const int grid = GridFor(g.n, 256, 8192);
  hipLaunchKernelGGL(k_collect_hubs, dim3(grid), dim3(256), 0, s, g.n, g.off, th.block_max, list,
                     count);
  PCC_HIP_CHECK(hipGetLastError());
}

// A device-driven round sizes the grid by the bin's capacity, so it is
// capped (the kernel strides): a late round's short list then does not pay
// for a grid of the level's whole bin.
constexpr int kDeviceSmallGrid = 4096;

template <int G, bool A>
static void LaunchSmall(int64_t count, const CsrView& g, const StateView& st, const MoveParams& mp,
                        const uint32_t* list, const MoveOut& out, uint64_t salt, hipStream_t s,
                        const unsigned long long* dcount) {
  const int grid = GridFor(count, 256 / G, dcount ? kDeviceSmallGrid : 1 << 20);
  hipLaunchKernelGGL((k_bm_small<G, A>), dim3(grid), dim3(256), 0, s, g, st, mp, list, count, out,
                     salt, dcount);
}

// One hash bin: the kernel whose table fits the bin's degree range (dlo, dhi]
// at load factor <= 1/2 (smaller tables leave LDS for more workgroups per CU);
// one workgroup per listed vertex up to 8192, striding beyond.
template <int THREADS, int TABLE, bool A>
static void LaunchHash(int64_t count, int64_t dlo, int64_t dhi, int64_t hash_waves,
                       const CsrView& g, const StateView& st, const MoveParams& mp,
                       const ItemLists& il, const MoveOut& out, uint32_t* fb_list,
                       RoundCounters* rc, uint64_t salt, hipStream_t s,
                       const unsigned long long* dcount) {
  if (dhi <= dlo || count <= 0) return;
  int max_grid = 8192;
  if (hash_waves > 0) {
    // Workgroups resident at once: 256 CUs, each limited by its 160 KB of
    // LDS (the table: 16 bytes a slot) and its 32 wave slots.
    constexpr int64_t kLdsPerGroup = (int64_t)TABLE * 16 + 4096;
    int64_t per_cu = std::min<int64_t>((160 * 1024) / kLdsPerGroup, 32 / (THREADS / 64));
    if (per_cu < 1) per_cu = 1;
    max_grid = (int)std::min<int64_t>(8192, 256 * per_cu * hash_waves);
  }
  const int grid = GridFor(count, 1, max_grid);
  if constexpr (!A) {
    if (il.n > 1) {  // merged lists (sync rounds only)
      hipLaunchKernelGGL((k_bm_hash_lists<THREADS, TABLE>), dim3(grid), dim3(THREADS), 0, s, g, st,
                         mp, il, count, dlo, dhi, out, fb_list, rc);
      return;
    }
  }
  hipLaunchKernelGGL((k_bm_hash<THREADS, TABLE, A>), dim3(grid), dim3(THREADS), 0, s, g, st, mp,
                     il.list[0], count, dlo, dhi, out, fb_list, rc, salt, dcount);
}

// The kernel of hash list k (0 = bin 5) over the items il, degrees (lo, hi].
template <bool A>
static void LaunchHashList(int k, int64_t count, int64_t lo, int64_t hi, int64_t hw,
                           const CsrView& g, const StateView& st, const MoveParams& mp,
                           const ItemLists& il, const MoveOut& out, uint32_t* fb_list,
                           RoundCounters* rc, uint64_t salt, hipStream_t s,
                           const unsigned long long* dcount) {
  switch (k) {
    case 0: LaunchHash<64, 256, A>(count, lo, hi, hw, g, st, mp, il, out, fb_list, rc, salt, s, dcount); break;
    case 1: LaunchHash<64, 512, A>(count, lo, hi, hw, g, st, mp, il, out, fb_list, rc, salt, s, dcount); break;
    case 2: LaunchHash<128, 1024, A>(count, lo, hi, hw, g, st, mp, il, out, fb_list, rc, salt, s, dcount); break;
    case 3: LaunchHash<256, 2048, A>(count, lo, hi, hw, g, st, mp, il, out, fb_list, rc, salt, s, dcount); break;
    case 4: LaunchHash<512, 4096, A>(count, lo, hi, hw, g, st, mp, il, out, fb_list, rc, salt, s, dcount); break;
    default:
      LaunchHash<1024, 8192, A>(count, lo, hi, hw, g, st, mp, il, out, fb_list, rc, salt, s, dcount);
  }
}

template <bool A>
static void LaunchHashBin(int bin, int64_t count, const BinThresholds& th, const CsrView& g,
                          const StateView& st, const MoveParams& mp, const uint32_t* list,
                          const MoveOut& out, uint32_t* fb_list, RoundCounters* rc,
                          uint64_t salt, hipStream_t s, const unsigned long long* dcount) {
  const HashBounds h = HashBoundsOf(th.wave_max, th.block_max);
  const int k = bin - kWaveBin0;
  const int64_t lo = k == 0 ? 64 : h.b[k - 1], hi = h.b[k];
  ItemLists il{};
  il.list[0] = list;
  il.end[0] = count;
  il.n = 1;
  LaunchHashList<A>(k, count, lo, hi, th.hash_waves, g, st, mp, il, out, fb_list, rc, salt, s,
                    dcount);
}

// Sync rounds with at most this many hashed vertices take one launch of the
// largest table they need instead of one launch per table size.
constexpr int64_t kMergeHashItems = 1024;

void LaunchBestMoveBin(int bin, int64_t count, bool async, const BinThresholds& th,
                       const CsrView& g, const StateView& st, const MoveParams& mp,
                       const uint32_t* list, const MoveOut& out, uint32_t* fb_list,
                       RoundCounters* rc, uint64_t salt, hipStream_t s,
                       const DeviceCounts* dc) {
  if (count <= 0) return;
  const unsigned long long* dcount = dc ? &dc->cnt[bin] : nullptr;
  switch (bin) {
    case 0:
      async ? LaunchSmall<4, true>(count, g, st, mp, list, out, salt, s, dcount)
            : LaunchSmall<4, false>(count, g, st, mp, list, out, salt, s, dcount);
      break;
    case 1:
      async ? LaunchSmall<8, true>(count, g, st, mp, list, out, salt, s, dcount)
            : LaunchSmall<8, false>(count, g, st, mp, list, out, salt, s, dcount);
      break;
    case 2:
      async ? LaunchSmall<16, true>(count, g, st, mp, list, out, salt, s, dcount)
            : LaunchSmall<16, false>(count, g, st, mp, list, out, salt, s, dcount);
      break;
    case 3:
      async ? LaunchSmall<32, true>(count, g, st, mp, list, out, salt, s, dcount)
            : LaunchSmall<32, false>(count, g, st, mp, list, out, salt, s, dcount);
      break;
    case 4:
      async ? LaunchSmall<64, true>(count, g, st, mp, list, out, salt, s, dcount)
            : LaunchSmall<64, false>(count, g, st, mp, list, out, salt, s, dcount);
      break;
    case 5:
    case 6:
    case 7:
    case 8:
    case 9:
    case 10:
      async ? LaunchHashBin<true>(bin, count, th, g, st, mp, list, out, fb_list, rc, salt, s, dcount)
            : LaunchHashBin<false>(bin, count, th, g, st, mp, list, out, fb_list, rc, salt, s, dcount);
      break;
    default:
      throw HipError{hipErrorInvalidValue, "LaunchBestMoveBin: the hub bin goes through LaunchBestMoveHub",
                     __FILE__, __LINE__};
  }
  PCC_HIP_CHECK(hipGetLastError());
}

void LaunchBestMoveBins(const int64_t counts[kHubBin], const uint32_t* const lists[kHubBin],
                        bool async, const BinThresholds& th, const CsrView& g,
                        const StateView& st, const MoveParams& mp, const MoveOut& out,
                        uint32_t* fb_list, RoundCounters* rc, uint64_t salt, hipStream_t s,
                        const DeviceCounts* dc, const RoundResetJob* reset) {
  if (reset && async) {
    LaunchRoundReset(reset->n, reset->a, reset->b, nullptr, reset->rc, reset->counters, nullptr, s);
    reset = nullptr;
  }
  if (async) {
    // Async rounds keep one launch per bin, each seeing the moves of the
    // ones before: running the bins of a sub-round together lowered the
    // objective of the async modularity runs by about half a percent.
    for (int b = 0; b < kHubBin; ++b)
      if (counts[b] > 0)
        LaunchBestMoveBin(b, counts[b], true, th, g, st, mp, lists[b], out, fb_list, rc, salt, s,
                          dc);
    return;
  }
  SmallBins sb{};
  uint32_t blocks = 0;
  for (int b = 0; b < kWaveBin0; ++b) {
    sb.list[b] = lists[b];
    sb.count[b] = counts[b];
    sb.dcount[b] = dc ? &dc->cnt[b] : nullptr;
    if (counts[b] > 0) {
      const int per = 256 >> (2 + b);  // G = 4 << b lanes per vertex
      const int64_t grid = GridFor(counts[b], per, dc ? kDeviceSmallGrid : 1 << 20);
      blocks += (uint32_t)((grid + 7) / 8 * 8);
    }
    sb.end[b] = blocks;
  }
  if (reset) {
    sb.reset = *reset;
    const int64_t work = (reset->a || reset->b) ? (reset->n + 15) / 16 : 1;
    sb.reset_blocks = (uint32_t)GridFor(work, 256, 2048);
  }
  if (blocks + sb.reset_blocks > 0) {
    hipLaunchKernelGGL(k_bm_small_all<false>, dim3(blocks + sb.reset_blocks), dim3(256), 0, s, g,
                       st, mp, sb, out, salt);
    PCC_HIP_CHECK(hipGetLastError());
  }
  int64_t hashed = 0;
  int top = -1;
  for (int b = kWaveBin0; b < kHubBin; ++b)
    if (counts[b] > 0) {
      hashed += counts[b];
      top = b;
    }
  if (top < 0) return;
  if (dc || hashed > kMergeHashItems) {
    for (int b = kWaveBin0; b < kHubBin; ++b)
      if (counts[b] > 0)
        LaunchBestMoveBin(b, counts[b], false, th, g, st, mp, lists[b], out, fb_list, rc, salt, s,
                          dc);
    return;
  }
  const HashBounds h = HashBoundsOf(th.wave_max, th.block_max);
  ItemLists il{};
  for (int b = kWaveBin0; b < kHubBin; ++b)
    if (counts[b] > 0) {
      il.list[il.n] = lists[b];
      il.end[il.n] = (il.n ? il.end[il.n - 1] : 0) + counts[b];
      ++il.n;
    }
  LaunchHashList<false>(top - kWaveBin0, hashed, 64, h.b[top - kWaveBin0], th.hash_waves, g, st,
                        mp, il, out, fb_list, rc, salt, s, nullptr);
  PCC_HIP_CHECK(hipGetLastError());
}

// The partition table holds 2 * part_edges slots (load factor <= 1/2), from
// 2048 to 8192: smaller partitions leave LDS for more workgroups per CU.
template <bool A>
static void LaunchHubPart(int64_t items, const CsrView& g, const StateView& st,
                          const MoveParams& mp, const uint32_t* list, int64_t count,
                          const HubScratch& hb, RoundCounters* rc, hipStream_t s,
                          const unsigned long long* dcount) {
  const dim3 grid((unsigned)items), block(kHubThreads);
  if (hb.part_edges <= 1024)
    hipLaunchKernelGGL((k_bm_hub_part<2048, A>), grid, block, 0, s, g, st, mp, list, count, hb, rc, dcount);
  else if (hb.part_edges <= 2048)
    hipLaunchKernelGGL((k_bm_hub_part<4096, A>), grid, block, 0, s, g, st, mp, list, count, hb, rc, dcount);
  else
    hipLaunchKernelGGL((k_bm_hub_part<8192, A>), grid, block, 0, s, g, st, mp, list, count, hb, rc, dcount);
}

void LaunchBestMoveHub(int64_t count, int64_t te, bool async, const CsrView& g,
                       const StateView& st, const MoveParams& mp, const uint32_t* list,
                       const HubScratch& hb, const MoveOut& out, uint32_t* fb_list,
                       RoundCounters* rc, hipStream_t s, const DeviceCounts* dc) {
  if (count <= 0) return;
  const unsigned long long* dcount = dc ? &dc->cnt[kHubBin] : nullptr;
  // sum_i ceil(deg_i / E) < te / E + count bounds the partition items; the
  // grid is a whole number of XCD chunks (XcdChunkItem).
  int64_t items = count + te / hb.part_edges + 1;
  items = (items + kXcdGrid - 1) / kXcdGrid * kXcdGrid;
  if (items > (int64_t)INT32_MAX)
    throw HipError{hipErrorInvalidValue, "LaunchBestMoveHub: too many partitions", __FILE__,
                   __LINE__};
  const int fin = GridFor(count, 256 / kWave, 1 << 16);
  if (async) {
    hipLaunchKernelGGL(k_hub_gather<true>, dim3((unsigned)items), dim3(256), 0, s, g, st, mp, list,
                       count, hb, dcount);
    LaunchHubPart<true>(items, g, st, mp, list, count, hb, rc, s, dcount);
    hipLaunchKernelGGL(k_bm_hub_final<true>, dim3(fin), dim3(256), 0, s, g, st, mp, list, count, hb,
                       out, fb_list, rc, dcount);
  } else {
    hipLaunchKernelGGL(k_hub_gather<false>, dim3((unsigned)items), dim3(256), 0, s, g, st, mp, list,
                       count, hb, dcount);
    LaunchHubPart<false>(items, g, st, mp, list, count, hb, rc, s, dcount);
    hipLaunchKernelGGL(k_bm_hub_final<false>, dim3(fin), dim3(256), 0, s, g, st, mp, list, count,
                       hb, out, fb_list, rc, dcount);
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
