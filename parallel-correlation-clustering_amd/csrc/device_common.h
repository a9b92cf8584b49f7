// device_common.h -- gfx950 device helpers shared by the kernel files.
//
// Wave64 throughout: every cross-lane helper assumes warpSize == 64.
#ifndef PCC_AMD_DEVICE_COMMON_H_
#define PCC_AMD_DEVICE_COMMON_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcc {

constexpr uint32_t kNoneId = 0xFFFFFFFFu;
// EfficientBestMove switches to its hub path above this degree
// (parallel-correlation-clusterer-internal.cc:456).
constexpr int64_t kHubDegree = 1000;
constexpr int kWave = 64;

__device__ __forceinline__ int LaneId() { return (int)__lane_id(); }

// EdgeSum::NetWeight (correlation-clusterer-util.h:36-39) as one fused
// multiply-add, matching the reference binary's contraction of w - res * x.
__device__ __forceinline__ double NetWeight(double weight, double x, double res) {
  return __fma_rn(-res, x, weight);
}

// Exponent of the lowest set bit of a finite double (its granularity): x is an
// integer multiple of 2^GranExp(x).  Returns a large value for zero.
__device__ __forceinline__ int GranExp(double x) {
  const uint64_t b = (uint64_t)__double_as_longlong(x);
  const uint64_t mant = b & 0xFFFFFFFFFFFFFull;
  const int e = (int)((b >> 52) & 0x7FF);
  if (e == 0) {
    if (mant == 0) return 1 << 20;
    return -1074 + __builtin_ctzll(mant);
  }
  return e - 1075 + __builtin_ctzll(mant | (1ull << 52));
}

// Order-free summation test of the hashed bins.  Per vertex: B bounds the
// sum and M the largest of the magnitudes of the values the vertex adds into
// its per-cluster totals; thr = ThresholdFor(B) and every |value| < 2^mexp.
// A value x is "fine" if GranExp(x) < thr.  A per-cluster total whose values
// are all multiples of 2^g, and whose partial sums stay below 2^(g + 50) in
// magnitude, is exact in every summation order; so is any total of at most two
// values (floating-point addition commutes).
__device__ __forceinline__ int ThresholdFor(double b) {
  if (!(b < INFINITY)) return 32767;  // NaN / inf: no value counts as coarse
  if (b == 0.0) return -32768;
  int e = 0;
  (void)frexp(b * (1.0 + 0x1p-20), &e);  // b' = f * 2^e, f in [0.5, 1)
  return (e - 1) - 50;                    // floor(log2 b') - 50
}
__device__ __forceinline__ int MagExpFor(double m) {
  if (!(m < INFINITY)) return 32767;
  if (m == 0.0) return -32768;
  int e = 0;
  (void)frexp(m * (1.0 + 0x1p-20), &e);  // m' < 2^e
  return e;
}
// Threshold of a vertex none of whose values is fine (GranExp(x) >= the
// threshold its bound B gives, for every value x): all its per-cluster totals
// are exact in every summation order, so the hashed kernels keep no per-slot
// term records for it.  ThresholdFor(0) gives the same value.
constexpr int kNoFineThr = -32768;

__device__ __forceinline__ int32_t PackScale(int thr, int mexp) {
  return (int32_t)(((uint32_t)(thr & 0xFFFF) << 16) | (uint32_t)(mexp & 0xFFFF));
}
__device__ __forceinline__ int ScaleThr(int32_t s) { return (int16_t)((uint32_t)s >> 16); }
__device__ __forceinline__ int ScaleMexp(int32_t s) { return (int16_t)((uint32_t)s & 0xFFFF); }

// Per-slot term record of the hashed bins: bits 0..23 count the terms
// (capped well below 2^24), bit 24 + i is set when a term has
// GranExp < thr - 8 i (i < 8).
constexpr uint32_t kCountMask = 0xFFFFFFu;
constexpr uint32_t kCountCap = 0xFFF000u;
__device__ __forceinline__ uint32_t FineBits(int ge, int thr) {
  if (ge >= thr) return 0u;
  int j = (thr - 1 - ge) >> 3;
  if (j > 7) j = 7;
  return ((2u << j) - 1u) << 24;
}
// True when the slot's sum may depend on the summation order: at least three
// terms, a fine one among them, and partial sums (below count * 2^mexp) that
// may exceed 2^(g + 50) for the granularity 2^g the fine bits guarantee.
__device__ __forceinline__ bool SlotInexact(uint32_t cnt, int thr, int mexp) {
  const uint32_t bits = cnt >> 24, count = cnt & kCountMask;
  if (bits == 0 || count <= 2) return false;
  const int j = 31 - __clz((int)bits);
  if (j >= 7) return true;
  const int e = thr - 8 * (j + 1) + 50 - mexp;
  if (e >= 31) return false;
  if (e <= 0) return true;
  return count >= (1u << e);
}

__device__ __forceinline__ uint32_t HashId(uint32_t c) {
  uint32_t h = c * 0x9E3779B1u;
  return h ^ (h >> 15);
}

// Async sub-round of vertex v among S (a fixed hash of the id).
__device__ __forceinline__ uint32_t SubRoundOf(uint32_t v, uint32_t S) {
  uint32_t h = v * 0x9E3779B1u;
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  return h % S;
}

// Partition of cluster c among `parts` partitions (a hash independent of
// HashId's low bits, range-reduced by multiply-shift).
__device__ __forceinline__ uint32_t PartOf(uint32_t c, uint32_t parts) {
  uint32_t h = c ^ (c >> 16);
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return (uint32_t)(((uint64_t)h * parts) >> 32);
}

// Workgroups are dispatched round-robin over the 8 XCDs of the chip.  This
// bijection of [0, grid) gives each XCD a contiguous range of work items, so
// consecutive items share an L2 (for items of similar cost).
__device__ __forceinline__ int64_t XcdItem(int64_t block, int64_t grid) {
  constexpr int64_t kXcds = 8;
  const int64_t x = block % kXcds, j = block / kXcds;
  const int64_t q = grid / kXcds, r = grid % kXcds;
  return x * q + (x < r ? x : r) + j;
}

// Items of very different cost: chunks of kXcdChunk consecutive items go to
// the XCDs in turn, so a run of expensive items (the partitions of one large
// vertex) is spread over every XCD while each chunk still shares an L2.  A
// bijection of [0, grid) when grid is a multiple of kXcdGrid.
constexpr int64_t kXcdChunk = 16;
constexpr int64_t kXcdGrid = 8 * kXcdChunk;
__device__ __forceinline__ int64_t XcdChunkItem(int64_t block) {
  constexpr int64_t kXcds = 8;
  const int64_t x = block % kXcds, j = block / kXcds;
  return (x + kXcds * (j / kXcdChunk)) * kXcdChunk + (j % kXcdChunk);
}

// Selection order of EfficientBestMove among candidate clusters
// (parallel-correlation-clusterer-internal.cc:493-503, 566-575): larger delta,
// then smaller cluster id.
__device__ __forceinline__ bool Better(double d1, uint32_t c1, double d2, uint32_t c2) {
  return d1 > d2 || (d1 == d2 && c1 < c2);
}

struct Cand {
  double d;
  uint32_t c;  // kNoneId = no candidate
};

__device__ __forceinline__ Cand PickBetter(Cand a, Cand b) {
  if (a.c == kNoneId) return b;
  if (b.c == kNoneId) return a;
  return Better(a.d, a.c, b.d, b.c) ? a : b;
}

// Reduction of candidates over aligned groups of `width` lanes (power of 2).
__device__ __forceinline__ Cand GroupBest(Cand x, int width) {
  for (int s = width >> 1; s > 0; s >>= 1) {
    Cand y;
    y.d = __shfl_xor(x.d, s, width);
    y.c = __shfl_xor(x.c, s, width);
    x = PickBetter(x, y);
  }
  return x;
}

__device__ __forceinline__ int WaveMinInt(int x) {
  for (int s = 32; s > 0; s >>= 1) x = min(x, __shfl_xor(x, s));
  return x;
}
__device__ __forceinline__ double WaveSum(double x) {
  for (int s = 32; s > 0; s >>= 1) x += __shfl_xor(x, s);
  return x;
}
__device__ __forceinline__ uint32_t WaveOr(uint32_t x) {
  for (int s = 32; s > 0; s >>= 1) x |= __shfl_xor(x, s);
  return x;
}
__device__ __forceinline__ int WaveMaxInt(int x) {
  for (int s = 32; s > 0; s >>= 1) x = max(x, __shfl_xor(x, s));
  return x;
}

// Source vertex of edge p (p < m): the largest u with off[u] <= p, searched in
// [lo, n).
__device__ __forceinline__ int64_t VertexOfEdge(const int64_t* __restrict__ off, int64_t lo,
                                                int64_t n, int64_t p) {
  // The next few vertices first (the common case), then bisection.
  for (int k = 0; k < 4 && lo + 1 < n; ++k) {
    if (off[lo + 1] > p) return lo;
    ++lo;
  }
  int64_t hi = n;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (off[mid] <= p)
      lo = mid;
    else
      hi = mid;
  }
  return lo < n ? lo : n - 1;  // off[n] == m > p keeps lo < n; guard anyway
}

// Block-wide helpers; `smem` needs blockDim.x / 64 slots.
template <typename T, typename Op>
__device__ __forceinline__ T BlockReduce(T x, T* smem, Op op) {
  for (int s = 32; s > 0; s >>= 1) x = op(x, __shfl_xor(x, s));
  const int w = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  __syncthreads();
  if (LaneId() == 0) smem[w] = x;
  __syncthreads();
  T r = smem[0];
  for (int i = 1; i < nw; ++i) r = op(r, smem[i]);
  __syncthreads();
  return r;
}

// Rows longer than kLongRow edges are split over workgroups: chunk j of the
// q-th queued row goes to workgroup (j + 7 q) mod gridDim.x, so one long row
// does not leave a single wave working while the rest of the grid idles.
constexpr int64_t kLongRow = 16384;
constexpr int64_t kLongChunk = 4096;
template <class F>
__device__ __forceinline__ void ForLongChunks(const uint32_t* __restrict__ rows, uint32_t nrows,
                                              const int64_t* __restrict__ off, F f) {
  const int64_t G = gridDim.x;
  for (uint32_t q = 0; q < nrows; ++q) {
    const uint32_t v = rows[q];
    const int64_t st = off[v], deg = off[v + 1] - st;
    const int64_t nch = (deg + kLongChunk - 1) / kLongChunk;
    int64_t j = ((int64_t)blockIdx.x - (7 * (int64_t)q) % G + G) % G;
    for (; j < nch; j += G) {
      const int64_t b = j * kLongChunk;
      f(q, v, st + b, st + (b + kLongChunk < deg ? b + kLongChunk : deg));
    }
  }
}

// A round's clears, over threads tid = 0, stride, ...: byte arrays a, b, c
// (n bytes each, any may be null), the bin counts (counters & 1), the
// fallback / new-cluster / repair counters folded into fb_total
// (counters & 2), and one 64-bit accumulator.
template <typename Counters>
__device__ __forceinline__ void RoundResetSpan(int64_t tid, int64_t stride, int64_t n,
                                               uint8_t* __restrict__ a, uint8_t* __restrict__ b,
                                               uint8_t* __restrict__ c, Counters* __restrict__ rc,
                                               int counters, unsigned long long* __restrict__ acc,
                                               int bins) {
  const int64_t n16 = n / 16;
  const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
  for (int64_t i = tid; i < n16; i += stride) {
    if (a) reinterpret_cast<uint4*>(a)[i] = zero;
    if (b) reinterpret_cast<uint4*>(b)[i] = zero;
    if (c) reinterpret_cast<uint4*>(c)[i] = zero;
  }
  for (int64_t i = n16 * 16 + tid; i < n; i += stride) {
    if (a) a[i] = 0;
    if (b) b[i] = 0;
    if (c) c[i] = 0;
  }
  if (tid < bins && (counters & 1)) {
    rc->cnt[tid] = 0;
    rc->te[tid] = 0;
  }
  if (tid == 0) {
    if (counters & 2) {
      rc->fb_total += rc->fb_count + rc->repair_count;
      rc->fb_count = 0;
      rc->new_count = 0;
      rc->repair_count = 0;
    }
    if (acc) *acc = 0;
  }
}

// Exclusive block scan; `smem` needs blockDim.x / 64 slots.  *total receives
// the block sum.
template <typename T>
__device__ __forceinline__ T BlockExclusiveScan(T x, T* smem, T* total) {
  const int lane = LaneId();
  const int w = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  T inc = x;
  for (int s = 1; s < kWave; s <<= 1) {
    const T y = __shfl_up(inc, s);
    if (lane >= s) inc += y;
  }
  __syncthreads();
  if (lane == kWave - 1) smem[w] = inc;
  __syncthreads();
  T before = 0, tot = 0;
  for (int i = 0; i < nw; ++i) {
    const T t = smem[i];
    if (i < w) before += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return before + inc - x;
}

}  // namespace pcc

#endif  // PCC_AMD_DEVICE_COMMON_H_
