// k_prims.hip -- device primitives on the path: exclusive scan, max reduce,
// fixed-order FP64 sum and a stable LSD radix sort of (key, value) pairs.
//
// These replace the pbbslib scan / filter / sample_sort calls the reference
// makes (parallel/parallel-sequence-ops.h:44-247).  Stability of the radix
// sort reproduces the stable sample_sort order (PSO.h:227-232).

#include <algorithm>
#include <cstring>

#include "device_common.h"
#include "kernels.h"

namespace pcc {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

template <typename T>
__device__ __forceinline__ T WaveInclusiveScan(T x) {
  const int lane = LaneId();
  for (int s = 1; s < 64; s <<= 1) {
    const T y = __shfl_up(x, s);
    if (lane >= s) x += y;
  }
  return x;
}

// Block exclusive scan of per-thread values; returns the block total too.
template <typename T>
__device__ __forceinline__ T BlockExclusiveScan(T x, T* total) {
  __shared__ T wsum[kScanThreads / 64];
  const T inc = WaveInclusiveScan(x);
  const int w = threadIdx.x / 64;
  if (LaneId() == 63) wsum[w] = inc;
  __syncthreads();
  T wbase = 0, tot = 0;
  for (int i = 0; i < kScanThreads / 64; ++i) {
    if (i < w) wbase += wsum[i];
    tot += wsum[i];
  }
  __syncthreads();
  *total = tot;
  return wbase + inc - x;
}

template <typename Tin, typename Tout>
__global__ __launch_bounds__(kScanThreads) void k_scan_tile_sums(const Tin* __restrict__ in,
                                                                 int64_t n, Tout* __restrict__ sums,
                                                                 const unsigned long long* gate) {
  if (gate && *gate == 0) return;
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  Tout s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j)
    if (base + j < n) s += (Tout)in[base + j];
  Tout tot;
  BlockExclusiveScan(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

template <typename Tin, typename Tout>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const Tin* __restrict__ in,
                                                             Tout* __restrict__ out, int64_t n,
                                                             const Tout* __restrict__ tile_off,
                                                             Tout* __restrict__ total,
                                                             const unsigned long long* gate) {
  if (gate && *gate == 0) return;
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  Tout v[kScanItems];
  Tout s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    v[j] = (base + j < n) ? (Tout)in[base + j] : (Tout)0;
    s += v[j];
  }
  Tout tot;
  Tout run = BlockExclusiveScan(s, &tot) + (tile_off ? tile_off[blockIdx.x] : (Tout)0);
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    if (base + j < n) out[base + j] = run;
    run += v[j];
  }
  if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) *total = run;
}

// One workgroup walking the tiles in order with a running offset: a small
// scan in one launch instead of three.
template <typename Tin, typename Tout>
__global__ __launch_bounds__(kScanThreads) void k_scan_serial(const Tin* __restrict__ in,
                                                              Tout* __restrict__ out, int64_t n,
                                                              Tout* __restrict__ total,
                                                              const unsigned long long* gate) {
  if (gate && *gate == 0) return;
  Tout carry = 0;
  for (int64_t t0 = 0; t0 < n; t0 += kScanTile) {
    const int64_t base = t0 + (int64_t)threadIdx.x * kScanItems;
    Tout v[kScanItems];
    Tout s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
      v[j] = (base + j < n) ? (Tout)in[base + j] : (Tout)0;
      s += v[j];
    }
    Tout tot;
    Tout run = BlockExclusiveScan(s, &tot) + carry;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
      if (base + j < n) out[base + j] = run;
      run += v[j];
    }
    carry += tot;
  }
  if (total && threadIdx.x == 0) *total = carry;
}

// Scans of up to this many elements run as one k_scan_serial launch.
constexpr int64_t kScanSerialMax = 8 * kScanTile;

// work (may be null): ScanWorkElems(n) elements for the tile sums of every
// level of the recursion, so the scan allocates nothing (and can run inside a
// captured graph); null draws them from the allocator.
template <typename Tin, typename Tout>
void ScanImpl(DeviceAllocator* al, const Tin* in, Tout* out, int64_t n, Tout* total, hipStream_t s,
              const unsigned long long* gate, Tout* work = nullptr) {
  if (n <= 0) {
    if (total) PCC_HIP_CHECK(hipMemsetAsync(total, 0, sizeof(Tout), s));
    return;
  }
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles == 1) {
    hipLaunchKernelGGL((k_scan_apply<Tin, Tout>), dim3(1), dim3(kScanThreads), 0, s, in, out, n,
                       (const Tout*)nullptr, total, gate);
    PCC_HIP_CHECK(hipGetLastError());
    return;
  }
  if (n <= kScanSerialMax) {
    hipLaunchKernelGGL((k_scan_serial<Tin, Tout>), dim3(1), dim3(kScanThreads), 0, s, in, out, n,
                       total, gate);
    PCC_HIP_CHECK(hipGetLastError());
    return;
  }
  DevBuf<Tout> sums_b, offs_b;
  Tout* sums = work;
  Tout* offs = work ? work + tiles : nullptr;
  if (!work) {
    sums_b.Reset(al, tiles);
    offs_b.Reset(al, tiles);
    sums = sums_b.get();
    offs = offs_b.get();
  }
  hipLaunchKernelGGL((k_scan_tile_sums<Tin, Tout>), dim3((unsigned)tiles), dim3(kScanThreads), 0,
                     s, in, n, sums, gate);
  PCC_HIP_CHECK(hipGetLastError());
  ScanImpl<Tout, Tout>(al, sums, offs, tiles, (Tout*)nullptr, s, gate,
                       work ? work + 2 * tiles : nullptr);
  hipLaunchKernelGGL((k_scan_apply<Tin, Tout>), dim3((unsigned)tiles), dim3(kScanThreads), 0, s,
                     in, out, n, (const Tout*)offs, total, gate);
  PCC_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------ reductions --

__global__ __launch_bounds__(256) void k_max_u32(const uint32_t* __restrict__ in, int64_t n,
                                                 uint32_t* __restrict__ out) {
  uint32_t m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    m = max(m, in[i]);
  for (int s = 32; s > 0; s >>= 1) m = max(m, (uint32_t)__shfl_xor(m, s));
  if (LaneId() == 0) atomicMax(out, m);
}

// Fixed-order sum: per-block sums in a fixed tree, then one block folds them.
__global__ __launch_bounds__(256) void k_sum_partial(const double* __restrict__ in, int64_t n,
                                                     double* __restrict__ part) {
  __shared__ double sm[4];
  double x = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x += in[i];
  x = BlockReduce(x, sm, [](double a, double b) { return a + b; });
  if (threadIdx.x == 0) part[blockIdx.x] = x;
}

__global__ __launch_bounds__(256) void k_sum_final(const double* __restrict__ part, int np,
                                                   double* __restrict__ out) {
  __shared__ double sm[4];
  double x = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) x += part[i];
  x = BlockReduce(x, sm, [](double a, double b) { return a + b; });
  if (threadIdx.x == 0) *out = x;
}

// ------------------------------------------------------------ radix sort --

constexpr int kRadixThreads = 512;
constexpr int kRadixItems = 8;
constexpr int kRadixTile = kRadixThreads * kRadixItems;
constexpr int kRadixWaves = kRadixThreads / 64;

__global__ __launch_bounds__(kRadixThreads) void k_radix_hist(const uint64_t* __restrict__ keys,
                                                              int64_t n, int shift, uint32_t dmask,
                                                              int64_t tiles,
                                                              uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[256];
  for (int i = threadIdx.x; i < 256; i += kRadixThreads) h[i] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kRadixTile;
  for (int j = 0; j < kRadixItems; ++j) {
    const int64_t idx = base + (int64_t)j * kRadixThreads + threadIdx.x;
    if (idx < n) atomicAdd(&h[(uint32_t)(keys[idx] >> shift) & dmask], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < 256; d += kRadixThreads)
    hist[(int64_t)d * tiles + blockIdx.x] = h[d];
}

// Stable scatter of one tile.  Wave w owns items [w*512, (w+1)*512) of the
// tile, visited in 8 rounds of 64 consecutive items, so per-wave running
// counters give ranks in input order; a per-digit prefix across waves and
// digits places each item, and the tile is staged in LDS so that the global
// writes of each digit run are contiguous.
template <typename V>
__global__ __launch_bounds__(kRadixThreads) void k_radix_scatter(
    const uint64_t* __restrict__ kin, const V* __restrict__ vin, uint64_t* __restrict__ kout,
    V* __restrict__ vout, int64_t n, int shift, int bits, int64_t tiles,
    const uint32_t* __restrict__ offsets) {
  __shared__ uint64_t sk[kRadixTile];
  __shared__ V sv[kRadixTile];
  __shared__ uint32_t wcnt[kRadixWaves][256];
  __shared__ uint32_t dstart[256];
  __shared__ uint32_t gbase[256];
  const uint32_t dmask = (1u << bits) - 1u;
  const int w = threadIdx.x / 64;
  const int lane = LaneId();
  const int64_t tile0 = (int64_t)blockIdx.x * kRadixTile;
  for (int i = threadIdx.x; i < kRadixWaves * 256; i += kRadixThreads) wcnt[i / 256][i % 256] = 0;
  __syncthreads();
  uint64_t k[kRadixItems];
  V val[kRadixItems];
  uint32_t rank[kRadixItems];
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int r = 0; r < kRadixItems; ++r) {
    const int64_t idx = tile0 + (int64_t)w * 64 * kRadixItems + (int64_t)r * 64 + lane;
    const bool ok = idx < n;
    k[r] = ok ? kin[idx] : ~0ull;
    val[r] = ok ? vin[idx] : V(0);
    const uint32_t dig = (uint32_t)(k[r] >> shift) & dmask;
    uint64_t peers = __ballot(ok);
    for (int b = 0; b < bits; ++b) {
      const uint64_t bal = __ballot((dig >> b) & 1u);
      peers &= ((dig >> b) & 1u) ? bal : ~bal;
    }
    uint32_t before = 0;
    if (ok) before = wcnt[w][dig];
    rank[r] = before + (uint32_t)__popcll(peers & lt);
    if (ok && (peers & lt) == 0) wcnt[w][dig] = before + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  if (threadIdx.x < 256) {
    const int d = threadIdx.x;
    uint32_t run = 0;
    for (int ww = 0; ww < kRadixWaves; ++ww) {
      const uint32_t c = wcnt[ww][d];
      wcnt[ww][d] = run;
      run += c;
    }
    dstart[d] = run;  // digit total for now
    gbase[d] = (d <= (int)dmask) ? offsets[(int64_t)d * tiles + blockIdx.x] : 0u;
  }
  __syncthreads();
  if (threadIdx.x < 64) {  // exclusive scan of the 256 digit totals
    uint32_t t[4];
    uint32_t s = 0;
    for (int j = 0; j < 4; ++j) {
      t[j] = dstart[threadIdx.x * 4 + j];
      s += t[j];
    }
    uint32_t inc = s;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(inc, o);
      if (lane >= o) inc += y;
    }
    uint32_t run = inc - s;
    for (int j = 0; j < 4; ++j) {
      dstart[threadIdx.x * 4 + j] = run;
      run += t[j];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRadixItems; ++r) {
    const int64_t idx = tile0 + (int64_t)w * 64 * kRadixItems + (int64_t)r * 64 + lane;
    if (idx < n) {
      const uint32_t dig = (uint32_t)(k[r] >> shift) & dmask;
      const uint32_t pos = dstart[dig] + wcnt[w][dig] + rank[r];
      sk[pos] = k[r];
      sv[pos] = val[r];
    }
  }
  __syncthreads();
  const int64_t cnt = (n - tile0) < kRadixTile ? (n - tile0) : kRadixTile;
  for (int i = threadIdx.x; i < cnt; i += kRadixThreads) {
    const uint64_t kk = sk[i];
    const uint32_t d = (uint32_t)(kk >> shift) & dmask;
    const uint32_t gpos = gbase[d] + ((uint32_t)i - dstart[d]);
    kout[gpos] = kk;
    vout[gpos] = sv[i];
  }
}

__global__ void k_copy_u64(const uint64_t* __restrict__ a, uint64_t* __restrict__ b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}
template <typename V>
__global__ void k_copy_v(const V* __restrict__ a, V* __restrict__ b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

template <typename V>
void RadixSortImpl(DeviceAllocator* al, uint64_t* keys, V* vals, int64_t n, int begin_bit,
                   int end_bit, hipStream_t s) {
  if (n <= 1 || end_bit <= begin_bit) return;
  if (n >= (int64_t)0xFFFFFFFFll) throw AllocError{(size_t)n};
  const int64_t tiles = (n + kRadixTile - 1) / kRadixTile;
  DevBuf<uint64_t> k2(al, n);
  DevBuf<V> v2(al, n);
  DevBuf<uint32_t> hist(al, 256 * tiles), offs(al, 256 * tiles);
  uint64_t* ka = keys;
  V* va = vals;
  uint64_t* kb = k2.get();
  V* vb = v2.get();
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    const int bits = (end_bit - shift) < 8 ? (end_bit - shift) : 8;
    const uint32_t dmask = (1u << bits) - 1u;
    // k_radix_hist writes all 256 counts of every tile.
    hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)tiles), dim3(kRadixThreads), 0, s, ka, n,
                       shift, dmask, tiles, hist.get());
    PCC_HIP_CHECK(hipGetLastError());
    ScanImpl<uint32_t, uint32_t>(al, hist.get(), offs.get(), 256 * tiles, nullptr, s, nullptr);
    hipLaunchKernelGGL((k_radix_scatter<V>), dim3((unsigned)tiles), dim3(kRadixThreads), 0, s, ka,
                       va, kb, vb, n, shift, bits, tiles, offs.get());
    PCC_HIP_CHECK(hipGetLastError());
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (ka != keys) {
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_copy_u64, dim3(grid), dim3(256), 0, s, ka, keys, n);
    hipLaunchKernelGGL((k_copy_v<V>), dim3(grid), dim3(256), 0, s, va, vals, n);
    PCC_HIP_CHECK(hipGetLastError());
  }
}

}  // namespace

int64_t ScanWorkElems(int64_t n) {
  int64_t e = 0;
  for (int64_t t = (n + kScanTile - 1) / kScanTile; t > 1; t = (t + kScanTile - 1) / kScanTile)
    e += 2 * t;
  return e;
}

void ExclusiveScanU8toU32(DeviceAllocator* al, const uint8_t* in, uint32_t* out, int64_t n,
                          uint32_t* total, hipStream_t s, const unsigned long long* gate,
                          uint32_t* work) {
  ScanImpl<uint8_t, uint32_t>(al, in, out, n, total, s, gate, work);
}
void ExclusiveScanU32toI64(DeviceAllocator* al, const uint32_t* in, int64_t* out, int64_t n,
                           int64_t* total, hipStream_t s) {
  ScanImpl<uint32_t, int64_t>(al, in, out, n, total, s, nullptr);
}
void ExclusiveScanU32toU32(DeviceAllocator* al, const uint32_t* in, uint32_t* out, int64_t n,
                           uint32_t* total, hipStream_t s) {
  ScanImpl<uint32_t, uint32_t>(al, in, out, n, total, s, nullptr);
}

void ReduceMaxU32(DeviceAllocator* al, const uint32_t* in, int64_t n, uint32_t* out,
                  hipStream_t s) {
  (void)al;
  PCC_HIP_CHECK(hipMemsetAsync(out, 0, sizeof(uint32_t), s));
  if (n <= 0) return;
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 2048);
  hipLaunchKernelGGL(k_max_u32, dim3(grid), dim3(256), 0, s, in, n, out);
  PCC_HIP_CHECK(hipGetLastError());
}

void SumF64(DeviceAllocator* al, const double* in, int64_t n, double* out, hipStream_t s) {
  constexpr int kParts = 1024;
  DevBuf<double> part(al, kParts);
  hipLaunchKernelGGL(k_sum_partial, dim3(kParts), dim3(256), 0, s, in, n, part.get());
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s, part.get(), kParts, out);
  PCC_HIP_CHECK(hipGetLastError());
}

void RadixSortPairsF32(DeviceAllocator* al, uint64_t* keys, float* vals, int64_t n, int begin_bit,
                       int end_bit, hipStream_t s) {
  RadixSortImpl<float>(al, keys, vals, n, begin_bit, end_bit, s);
}
void RadixSortPairsF64(DeviceAllocator* al, uint64_t* keys, double* vals, int64_t n,
                       int begin_bit, int end_bit, hipStream_t s) {
  RadixSortImpl<double>(al, keys, vals, n, begin_bit, end_bit, s);
}
void RadixSortPairsU32(DeviceAllocator* al, uint64_t* keys, uint32_t* vals, int64_t n,
                       int begin_bit, int end_bit, hipStream_t s) {
  RadixSortImpl<uint32_t>(al, keys, vals, n, begin_bit, end_bit, s);
}

void ReadBack(void* host, const void* dev, size_t bytes, hipStream_t s) {
  // A copy into pageable memory is staged by the runtime; one into pinned
  // memory is a plain DMA (about 15 instead of 23 us for a few bytes after a
  // kernel).  The staging buffer lives as long as its thread's process.
  thread_local void* stage = nullptr;
  thread_local size_t cap = 0;
  if (bytes == 0) return;
  if (bytes > cap) {
    const size_t want = std::max<size_t>(bytes, 4096);
    void* p = nullptr;
    PCC_HIP_CHECK(hipHostMalloc(&p, want, hipHostMallocDefault));
    if (stage) {
      PCC_HIP_CHECK(hipStreamSynchronize(s));
      (void)hipHostFree(stage);
    }
    stage = p;
    cap = want;
  }
  PCC_HIP_CHECK(hipMemcpyAsync(stage, dev, bytes, hipMemcpyDeviceToHost, s));
  PCC_HIP_CHECK(hipStreamSynchronize(s));
  std::memcpy(host, stage, bytes);
}

}  // namespace pcc
