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

// A set of doubles whose magnitudes sum to sabs and whose granularities are
// all >= 2^gmin has every partial sum exactly representable when
// sabs < 2^(gmin + 53); the check keeps one bit of margin for the rounding of
// sabs itself.  Then every summation order gives the same result.
__device__ __forceinline__ bool SumIsOrderFree(int gmin, double sabs) {
  if (sabs == 0.0) return true;
  if (gmin >= (1 << 19)) return true;
  return sabs < ldexp(1.0, gmin + 52);
}

__device__ __forceinline__ uint32_t HashId(uint32_t c) {
  uint32_t h = c * 0x9E3779B1u;
  return h ^ (h >> 15);
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
__device__ __forceinline__ int WaveMaxInt(int x) {
  for (int s = 32; s > 0; s >>= 1) x = max(x, __shfl_xor(x, s));
  return x;
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

}  // namespace pcc

#endif  // PCC_AMD_DEVICE_COMMON_H_
