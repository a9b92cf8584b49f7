// engine.cc -- host orchestration of the multilevel Louvain path.
//
// Restates ParallelCorrelationClusterer::RefineClusters_subroutine
// (clustering/clusterers/parallel-correlation-clusterer.cc:476-629),
// IterateBestMoves (:340-365) and BestMovesForVertexSubset (:115-338) over
// device-resident CSR and cluster state, for one graph or a batch of graphs.
// One HIP stream per handle; the host synchronizes once per best-move round to
// read the next frontier's size (the loop condition of IterateBestMoves) and
// the per-bin counts that size the next launches (ALL_MOVE rounds, whose
// frontier never changes, read nothing back).

#include "engine.h"

#include "device_common.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <new>
#include <set>

namespace pcc {

// ------------------------------------------------------------- allocator --

namespace {
// Block sizes: powers of two up to 1 MiB, then four classes per octave, so a
// large buffer wastes at most a quarter of its size.
size_t RoundSize(size_t b) {
  size_t r = 256;
  while (r < b && r < ((size_t)1 << 20)) r <<= 1;
  if (r >= b) return r;
  while ((r << 1) < b) r <<= 1;  // r < b <= 2r
  const size_t q = r / 4;
  return r + ((b - r + q - 1) / q) * q;
}
}  // namespace

namespace {
// Every allocator of the process, for the out-of-memory release.
std::mutex g_alloc_mu;
std::set<DeviceAllocator*>& Allocators() {
  static std::set<DeviceAllocator*> s;
  return s;
}
}  // namespace

DeviceAllocator::DeviceAllocator() {
  std::lock_guard<std::mutex> l(g_alloc_mu);
  Allocators().insert(this);
}

DeviceAllocator::~DeviceAllocator() {
  {
    std::lock_guard<std::mutex> l(g_alloc_mu);
    Allocators().erase(this);
  }
  ReleaseAll();
}

size_t DeviceAllocator::BytesInUse() const {
  std::lock_guard<std::mutex> l(mu_);
  return in_use_;
}

size_t DeviceAllocator::ReleaseFree() {
  std::lock_guard<std::mutex> l(mu_);
  size_t bytes = 0;
  for (auto& kv : free_)
    for (void* q : kv.second) {
      (void)hipFree(q);
      bytes += kv.first;
    }
  free_.clear();
  return bytes;
}

void* DeviceAllocator::Alloc(size_t bytes) {
  const size_t sz = RoundSize(bytes);
  void* p = nullptr;
  {
    std::lock_guard<std::mutex> l(mu_);
    // The smallest cached block of at least sz bytes, if it is less than
    // twice as large.
    for (auto it = free_.lower_bound(sz); it != free_.end() && it->first < 2 * sz; ++it) {
      if (it->second.empty()) continue;
      p = it->second.back();
      it->second.pop_back();
      live_[p] = it->first;
      in_use_ += it->first;
      return p;
    }
  }
  // No lock is held here: the release below locks other allocators (always
  // registry first, then one allocator), never the other way round.
  if (hipMalloc(&p, sz) != hipSuccess) {
    (void)hipGetLastError();
    (void)ReleaseFree();
    if (hipMalloc(&p, sz) != hipSuccess) {
      (void)hipGetLastError();
      {
        std::lock_guard<std::mutex> l(g_alloc_mu);
        for (DeviceAllocator* a : Allocators())
          if (a != this) (void)a->ReleaseFree();
      }
      if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        throw AllocError{bytes};
      }
    }
  }
  std::lock_guard<std::mutex> l(mu_);
  live_[p] = sz;
  in_use_ += sz;
  return p;
}

void DeviceAllocator::Free(void* p) {
  std::lock_guard<std::mutex> l(mu_);
  auto it = live_.find(p);
  if (it == live_.end()) return;
  in_use_ -= it->second;
  free_[it->second].push_back(p);
  live_.erase(it);
}

void DeviceAllocator::ReleaseAll() {
  std::lock_guard<std::mutex> l(mu_);
  for (auto& kv : free_)
    for (void* q : kv.second) (void)hipFree(q);
  free_.clear();
  for (auto& kv : live_) (void)hipFree(kv.first);
  live_.clear();
  in_use_ = 0;
}

// --------------------------------------------------------------- handles --

int Guard(const char* where, const std::function<int()>& body) {
  try {
    return body();
  } catch (const HipError& e) {
    return SetError(PCC_INTERNAL, std::string(where) + ": HIP error " + hipGetErrorString(e.err) +
                                      " at " + e.file + ":" + std::to_string(e.line) + " (" +
                                      e.what + ")");
  } catch (const AllocError& e) {
    return SetError(PCC_RESOURCE_EXHAUSTED,
                    std::string(where) + ": device allocation of " + std::to_string(e.bytes) +
                        " bytes failed (free HBM exhausted)");
  } catch (const std::bad_alloc&) {
    return SetError(PCC_RESOURCE_EXHAUSTED, std::string(where) + ": host allocation failed");
  }
}

int NewHandle(pcc_graph** out) {
  auto h = std::make_unique<pcc_graph>();
  PCC_HIP_CHECK(hipGetDevice(&h->device));
  PCC_HIP_CHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  *out = h.release();
  return PCC_OK;
}

namespace {

int Grid(int64_t items, int per_block, int max_blocks) {
  int64_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (int)b;
}

__global__ void k_validate(int64_t n, int64_t m, const int64_t* off, const uint32_t* nbr,
                           unsigned long long* bad) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x)
    if ((int64_t)nbr[i] >= n) atomicAdd(bad, 1ull);
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    if (off[v + 1] < off[v]) atomicAdd(bad, 1ull);
}

__global__ void k_fill_ones(float* w, int64_t m) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x)
    w[i] = 1.0f;
}

__global__ void k_iota(uint32_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (uint32_t)i;
}

// vgraph[v] = g with base[g] <= v < base[g + 1].
__global__ void k_fill_vgraph(int64_t n, const int64_t* __restrict__ base, int G,
                              uint32_t* __restrict__ vgraph) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = G - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (base[mid] <= v) lo = mid; else hi = mid - 1;
    }
    vgraph[v] = (uint32_t)lo;
  }
}

__global__ void k_sel_frontier(int64_t n, const uint32_t* __restrict__ vgraph,
                               const uint8_t* __restrict__ sel, uint8_t* __restrict__ fr) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    fr[v] = sel[vgraph ? vgraph[v] : 0u];
}

// Final per-vertex ids, local to each graph (out) and as vertex ids of the
// whole batch (global: local + the graph's first vertex).
__global__ void k_output_ids(int64_t n, const uint32_t* __restrict__ vgraph0,
                             const int64_t* __restrict__ base0,
                             const uint8_t* __restrict__ use_ids0,
                             const uint32_t* __restrict__ ids0,
                             const uint32_t* __restrict__ chain,
                             const int64_t* __restrict__ sub, uint32_t* __restrict__ out,
                             uint32_t* __restrict__ global) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t o = vgraph0 ? vgraph0[v] : 0u;
    const int64_t x = use_ids0[o] ? (int64_t)ids0[v] : (int64_t)chain[v];
    const uint32_t local = (uint32_t)(x - sub[o]);
    out[v] = local;
    global[v] = local + (uint32_t)(vgraph0 ? base0[o] : 0);
  }
}

template <typename T>
DevBuf<T> Upload(DeviceAllocator* al, const std::vector<T>& h, hipStream_t s) {
  DevBuf<T> d(al, (int64_t)std::max<size_t>(h.size(), 1));
  if (!h.empty())
    PCC_HIP_CHECK(hipMemcpyAsync(d.get(), h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
  return d;
}

}  // namespace

int UploadCsr(pcc_graph* h, int64_t n, int64_t m, const int64_t* off, const uint32_t* nbrs,
              const float* w) {
  if (n < 0 || m < 0 || n >= (int64_t)0xFFFFFFF0ll)
    return SetError(PCC_INVALID_ARGUMENT, "vertex count must lie in [0, 2^32 - 16)");
  if (!off || (m > 0 && !nbrs)) return SetError(PCC_INVALID_ARGUMENT, "null CSR array");
  if (off[0] != 0 || off[n] != m)
    return SetError(PCC_INVALID_ARGUMENT, "offsets must start at 0 and end at m");
  hipStream_t s = h->stream;
  h->csr.n = n;
  h->csr.m = m;
  h->csr.off.Reset(&h->alloc, n + 1);
  h->csr.nbr.Reset(&h->alloc, m > 0 ? m : 1);
  h->csr.w.Reset(&h->alloc, m > 0 ? m : 1);
  PCC_HIP_CHECK(hipMemcpyAsync(h->csr.off.get(), off, sizeof(int64_t) * (size_t)(n + 1),
                               hipMemcpyHostToDevice, s));
  if (m > 0) {
    PCC_HIP_CHECK(hipMemcpyAsync(h->csr.nbr.get(), nbrs, sizeof(uint32_t) * (size_t)m,
                                 hipMemcpyHostToDevice, s));
    if (w) {
      PCC_HIP_CHECK(hipMemcpyAsync(h->csr.w.get(), w, sizeof(float) * (size_t)m,
                                   hipMemcpyHostToDevice, s));
    } else {
      hipLaunchKernelGGL(k_fill_ones, dim3(1024), dim3(256), 0, s, h->csr.w.get(), m);
    }
  }
  DevBuf<unsigned long long> bad(&h->alloc, 1);
  PCC_HIP_CHECK(hipMemsetAsync(bad.get(), 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL(k_validate, dim3(2048), dim3(256), 0, s, n, m, h->csr.off.get(),
                     h->csr.nbr.get(), bad.get());
  unsigned long long nbad = 0;
  ReadBack(&nbad, bad.get(), sizeof(nbad), s);
  if (nbad) {
    h->csr = DevCsr();
    return SetError(PCC_INVALID_ARGUMENT,
                    "CSR has out-of-range neighbour ids or decreasing offsets");
  }
  if (h->gbase.empty()) h->gbase = {0, n};
  h->goff.resize(h->gbase.size());
  for (size_t i = 0; i < h->gbase.size(); ++i) h->goff[i] = off[h->gbase[i]];
  h->symmetric = -1;
  h->finished = true;
  return PCC_OK;
}

int CheckMovesConfig(pcc_graph* h, const pcc_config& cfg) {
  if (cfg.clustering_moves_method != PCC_LOUVAIN &&
      cfg.clustering_moves_method != PCC_DEFAULT_CLUSTER_MOVES)
    return SetError(PCC_UNIMPLEMENTED,
                    "Correlation clustering moves must be DEFAULT_CLUSTER_MOVES or LOUVAIN.");
  if (cfg.move_method < PCC_NBHR_CLUSTER_MOVE || cfg.move_method > PCC_NBHR_MOVE)
    return SetError(PCC_INVALID_ARGUMENT, "unknown move_method");
  const int subcl = cfg.subclustering_method;
  if (subcl == 0) return PCC_OK;
  if (subcl < 1 || subcl > 3) return SetError(PCC_INVALID_ARGUMENT, "unknown subclustering_method");
  // The async in-round form charges a set spanning several clusters to its
  // first member's cluster (PCCI.cc:413-418) and marks moved clusters by
  // subcluster index (PCC.cc:217-219), past the end of the array.
  if (cfg.clustering_moves_method != PCC_LOUVAIN && cfg.async)
    return SetError(PCC_UNIMPLEMENTED,
                    "async subcluster moves inside DEFAULT_CLUSTER_MOVES rounds are not on this path");
  if (h->num_graphs() > 1)
    return SetError(PCC_UNIMPLEMENTED, "subclustering of batched graphs is not on this path");
  // Triangle scores merge adjacency lists (GBBS's intersect assumes them
  // ascending); contracted levels are built ascending.
  if (subcl == 1 && !RowsAscending(&h->alloc, h->csr.View(), h->stream))
    return SetError(PCC_INVALID_ARGUMENT,
                    "TRIANGLE_SUBCLUSTERING needs strictly ascending adjacency lists");
  return PCC_OK;
}

namespace {
// flag = 1 if a weight is not a finite integer; mag += sum |nw| (an
// order-dependent double sum, compared against 2^53 with a margin).
__global__ void k_weights_exact(const double* __restrict__ nw, int64_t n,
                                unsigned long long* __restrict__ flag, double* __restrict__ mag) {
  double local = 0.0;
  bool bad = false;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const double x = nw[v];
    bad = bad || !(x == floor(x)) || isinf(x);
    local += fabs(x);
  }
  if (bad) atomicOr(flag, 1ull);
  for (int s = 32; s > 0; s >>= 1) local += __shfl_xor(local, s);
  if ((threadIdx.x & (kWave - 1)) == 0 && local != 0.0) atomicAdd(mag, local);
}
}  // namespace

namespace {
__global__ void k_mark_present(int64_t n, const uint32_t* __restrict__ ids,
                               uint8_t* __restrict__ present) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    present[ids[v]] = 1;
}

__global__ void k_count_bytes(int64_t n, const uint8_t* __restrict__ flag,
                              unsigned long long* __restrict__ total) {
  unsigned long long c = 0;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    c += flag[v];
  for (int s = 32; s > 0; s >>= 1) c += __shfl_xor(c, s);
  if ((threadIdx.x & (kWave - 1)) == 0 && c) atomicAdd(total, c);
}

__global__ void k_fill_f64(double* __restrict__ p, int64_t n, double x) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    p[v] = x;
}
}  // namespace

int64_t CountDistinctIds(DeviceAllocator* al, int64_t n, const uint32_t* ids, hipStream_t s) {
  if (n <= 0) return 0;
  DevBuf<uint8_t> present(al, n);
  DevBuf<unsigned long long> total(al, 1);
  PCC_HIP_CHECK(hipMemsetAsync(present.get(), 0, (size_t)n, s));
  PCC_HIP_CHECK(hipMemsetAsync(total.get(), 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL(k_mark_present, dim3(Grid(n, 256, 8192)), dim3(256), 0, s, n, ids,
                     present.get());
  hipLaunchKernelGGL(k_count_bytes, dim3(Grid(n, 256, 2048)), dim3(256), 0, s, n, present.get(),
                     total.get());
  PCC_HIP_CHECK(hipGetLastError());
  unsigned long long h = 0;
  ReadBack(&h, total.get(), sizeof(h), s);
  return (int64_t)h;
}

void LaunchFillF64(double* p, int64_t n, double x, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_f64, dim3(Grid(n, 256, 8192)), dim3(256), 0, s, p, n, x);
  PCC_HIP_CHECK(hipGetLastError());
}

bool NodeWeightsExact(DeviceAllocator* al, const double* nw, int64_t n, hipStream_t s) {
  if (n <= 0) return true;
  struct Acc {
    unsigned long long flag;
    double mag;
  };
  DevBuf<Acc> acc(al, 1);
  PCC_HIP_CHECK(hipMemsetAsync(acc.get(), 0, sizeof(Acc), s));
  Acc* a = acc.get();
  hipLaunchKernelGGL(k_weights_exact, dim3(Grid(n, 256, 2048)), dim3(256), 0, s, nw, n, &a->flag,
                     &a->mag);
  PCC_HIP_CHECK(hipGetLastError());
  Acc h;
  ReadBack(&h, a, sizeof(h), s);
  // Relative error of the double sum is far below 2^-20 for any n < 2^32.
  return h.flag == 0 && h.mag * (1.0 + 0x1p-20) < 0x1p53;
}

void LaunchIota(uint32_t* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_iota, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 8192)), dim3(256),
                     0, s, p, n);
  PCC_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- engine --

Engine::Engine(pcc_graph* h, const pcc_config& cfg, pcc_stats* stats)
    : h_(h), cfg_(cfg), s_(h->stream), al_(&h->alloc), stats_(stats) {}

Engine::~Engine() {
  for (auto& t : pending_) {
    (void)hipEventDestroy(t.first.a);
    (void)hipEventDestroy(t.first.b);
  }
}

Engine::Workspace::Workspace(DeviceAllocator* al, int64_t n, hipStream_t s, uint8_t* also_zero,
                             int64_t also_bytes)
    : na(al, n) {
  const int64_t nn = n > 0 ? n : 1;
  list.Reset(al, nn);
  list2.Reset(al, nn);
  sub_cnt.Reset(al, kBins * 1024);
  sub_cur.Reset(al, kBins * 1024);
  sub_te.Reset(al, 1024);
  fb_list.Reset(al, nn);
  moves.Reset(al, nn);
  moved.Reset(al, nn);
  flags.Reset(al, nn);
  next.Reset(al, nn);
  new_flag.Reset(al, nn);
  rc.Reset(al, 1);
  dc.Reset(al, 1);
  dstats.Reset(al, 1);
  // new_flag is kept all-zero between rounds by the new-cluster assignment.
  ZeroSpans z{};
  z.p[0] = new_flag.get();
  z.bytes[0] = nn;
  z.p[1] = reinterpret_cast<uint8_t*>(rc.get());
  z.bytes[1] = sizeof(RoundCounters);
  z.n = 2;
  if (also_zero && also_bytes > 0) {
    z.p[2] = also_zero;
    z.bytes[2] = also_bytes;
    z.n = 3;
  }
  LaunchZeroSpans(z, s);
}

void Engine::ResetCounters(RoundCounters* rc, bool all) {
  if (all) {
    PCC_HIP_CHECK(hipMemsetAsync(rc, 0, sizeof(RoundCounters), s_));
  } else {
    PCC_HIP_CHECK(hipMemsetAsync(rc, 0, offsetof(RoundCounters, fb_count), s_));
  }
}

void Engine::ReadCounters(RoundCounters* rc, int64_t counts[kBins], int64_t te[kBins],
                          RoundCounters* full) {
  RoundCounters hrc;
  ReadBack(&hrc, rc, sizeof(hrc), s_);
  for (int b = 0; b < kBins; ++b) {
    counts[b] = (int64_t)hrc.cnt[b];
    te[b] = (int64_t)hrc.te[b];
  }
  if (full) *full = hrc;
  CollectTimers();
}

Timer Engine::StartTimer() {
  Timer t;
  if (!h_->timing) return t;
  PCC_HIP_CHECK(hipEventCreate(&t.a));
  PCC_HIP_CHECK(hipEventCreate(&t.b));
  PCC_HIP_CHECK(hipEventRecord(t.a, s_));
  return t;
}

void Engine::StopTimer(Timer t, int kind) {
  if (!t.a) return;
  PCC_HIP_CHECK(hipEventRecord(t.b, s_));
  pending_.emplace_back(t, kind);
}

// kind: 0..7 best-move bin, 8 ordered fallback, 9 apply + frontier,
// 10 contraction, 11 all best-move kernels of a round.
void Engine::CollectTimers() {
  for (auto& p : pending_) {
    float ms = 0;
    PCC_HIP_CHECK(hipEventSynchronize(p.first.b));
    PCC_HIP_CHECK(hipEventElapsedTime(&ms, p.first.a, p.first.b));
    kind_seconds_[p.second] += ms * 1e-3;
    kind_launches_[p.second] += 1;
    (void)hipEventDestroy(p.first.a);
    (void)hipEventDestroy(p.first.b);
  }
  pending_.clear();
}

void Engine::PrepareLevel(const CsrView& g, const StateView& st, const Layout& lay,
                          LevelCtx* c) {
  c->g = g;
  c->st = st;
  // Bin capacities, the largest degree and the sign check of the node
  // weights and resolutions, read back together.
  DevBuf<unsigned long long> caps(al_, kLevelStats);
  PCC_HIP_CHECK(hipMemsetAsync(caps.get(), 0, sizeof(unsigned long long) * kLevelStats, s_));
  LaunchBinCapacity(g, h_->th, st.nw, lay.batch ? lay.gres.get() : nullptr, lay.G, caps.get(), s_);
  unsigned long long hcaps[kLevelStats];
  ReadBack(hcaps, caps.get(), sizeof(hcaps), s_);
  int64_t run = 0;
  for (int b = 0; b < kBins; ++b) {
    c->layout.base[b] = run;
    c->caps[b] = (int64_t)hcaps[b];
    c->caps_te[b] = (int64_t)hcaps[kBins + b];
    run += (int64_t)hcaps[b];
  }
  bool hashed = false;
  for (int b = kWaveBin0; b < kBins; ++b) hashed = hashed || c->caps[b] > 0;
  // Ordered fallback scratch, regions sized for the largest degree.
  const int64_t maxdeg = hashed ? (int64_t)hcaps[2 * kBins] : 0;
  if (maxdeg >= (int64_t)1 << 30)
    throw HipError{hipErrorInvalidValue, "vertex degree >= 2^30 is not supported", __FILE__,
                   __LINE__};
  uint32_t cap = 64;
  while (cap < (uint64_t)(2 * maxdeg + 2)) cap <<= 1;
  c->hs = HashScratch();
  c->hs.cap = cap;
  c->hs.regions = !hashed ? 1 : (cap >= (1u << 22) ? 16 : (cap >= (1u << 20) ? 32 : 64));
  c->hkeys.Reset(al_, (int64_t)c->hs.cap * c->hs.regions);
  c->hvals.Reset(al_, (int64_t)c->hs.cap * c->hs.regions);
  c->hs.keys = c->hkeys.get();
  c->hs.vals = c->hvals.get();
  if (hashed) {
    c->gscale.Reset(al_, g.n);
    LaunchGranThreshold(al_, g, cfg_.edge_weight_offset, c->gscale.get(), s_, maxdeg);
  }
  c->hb = HubScratch();
  const int64_t c7 = c->caps[kBins - 1], t7 = c->caps_te[kBins - 1];
  if (c7 > 0) {
    c->hb_part.Reset(al_, c7 + 1);
    c->hb_ent.Reset(al_, c7 + 1);
    c->hb_flag.Reset(al_, c7);
    c->hb_tval.Reset(al_, c7);
    // Partition items of any sub-list: sum ceil(deg / E) < t7 / E + c7.
    const int64_t items = c7 + t7 / h_->hub_part_edges + 1;
    c->hb_px1.Reset(al_, items);
    c->hb_pc1.Reset(al_, items);
    c->hb_px2.Reset(al_, items);
    c->hb_item_of.Reset(al_, items);
    c->hb_hpart.Reset(al_, items);
    int64_t bcap = HubBucketCap(h_->hub_part_edges);
    if (cfg_.async) {
      // Async rows concentrate on a few heavy clusters (with unit weights the
      // clusters of hubs grow fast), and a partition whose bucket overflows
      // streams its vertex's whole row.  Buckets of 8 partitions' edges keep
      // most of those pairs in the bucket (RMAT scale 24 async: bin 7 from
      // 0.60 to 0.53-0.57 s per Cluster call); they take 64 bytes per bin-7 edge,
      // so only when that fits in a quarter of the free device memory.
      const int64_t big = HubBucketCapAsync(h_->hub_part_edges);
      size_t free_b = 0, total_b = 0;
      if (hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
          (double)items * (double)big * (double)sizeof(uint2) < 0.25 * (double)free_b)
        bcap = big;
    }
    c->hb_bkt.Reset(al_, items * bcap);
    c->hb_bcur.Reset(al_, items);
    c->hb_pfine.Reset(al_, items);
    PCC_HIP_CHECK(hipMemsetAsync(c->hb_flag.get(), 0, sizeof(uint32_t) * (size_t)c7, s_));
    PCC_HIP_CHECK(hipMemsetAsync(c->hb_bcur.get(), 0, sizeof(uint32_t) * (size_t)items, s_));
    PCC_HIP_CHECK(hipMemsetAsync(c->hb_pfine.get(), 0, sizeof(uint32_t) * (size_t)items, s_));
    c->hb.part_edges = h_->hub_part_edges;
    c->hb.part_off = c->hb_part.get();
    c->hb.ent_off = c->hb_ent.get();
    c->hb.hflag = c->hb_flag.get();
    c->hb.tval = c->hb_tval.get();
    c->hb.px1 = c->hb_px1.get();
    c->hb.pc1 = c->hb_pc1.get();
    c->hb.px2 = c->hb_px2.get();
    c->hb.bkt = c->hb_bkt.get();
    c->hb.bcur = c->hb_bcur.get();
    c->hb.pfine = c->hb_pfine.get();
    c->hb.bcap = bcap;
    c->hb.item_of = c->hb_item_of.get();
    c->hb.hpart = c->hb_hpart.get();
    if (cfg_.async) {
      c->hb_cbuf.Reset(al_, t7 > 0 ? t7 : 1);
      c->hb.cbuf = c->hb_cbuf.get();
    }
    if (cfg_.move_method != PCC_ALL_MOVE) {
      c->hubs.Reset(al_, c7);
      c->hubs_part.Reset(al_, c7 + 1);
      c->hubs_ent.Reset(al_, c7 + 1);
      c->hubs_item_of.Reset(al_, items);
      DevBuf<unsigned long long> cnt(al_, 1);
      PCC_HIP_CHECK(hipMemsetAsync(cnt.get(), 0, sizeof(unsigned long long), s_));
      LaunchCollectHubs(g, h_->th, c->hubs.get(), cnt.get(), s_);
      LaunchHubScan(g, c->hubs.get(), c7, h_->hub_part_edges, c->hubs_part.get(),
                    c->hubs_ent.get(), c->hubs_item_of.get(), s_);
      c->hub_count = c7;
      c->hub_te = t7;
    }
  }
  c->pull = false;
  if (cfg_.move_method != PCC_ALL_MOVE && h_->pull_fraction < 1024) {
    // Contracted levels of a symmetric graph are symmetric (an inter-cluster
    // edge keeps both directions, an intra-cluster one becomes a self-loop).
    if (h_->symmetric < 0) h_->symmetric = GraphIsSymmetric(al_, h_->csr.View(), s_) ? 1 : 0;
    c->pull = h_->symmetric == 1;
    if (c->pull) c->vol.Reset(al_, 1);
  }
  MoveParams& mp = c->mp;
  mp = MoveParams();
  mp.res = cfg_.resolution;
  mp.offset = cfg_.edge_weight_offset;
  mp.n = (uint32_t)g.n;
  mp.force_ordered = h_->force_ordered ? 1 : 0;
  mp.gscale = c->gscale.get();
  if (lay.batch) {
    // Batches use a sentinel outside every graph's id range.
    mp.new_id = 0xFFFFFFFEu;
    mp.vgraph = lay.vgraph.get();
    mp.gres = lay.gres.get();
    mp.gbase = lay.dbase.get();
  } else {
    mp.new_id = (uint32_t)g.n;
  }
  // Candidate pruning needs every resolution and node weight >= 0.
  const bool nonneg = hcaps[2 * kBins + 1] == 0 && (lay.batch || mp.res >= 0.0);
  mp.prune = nonneg ? 1 : 0;
}

void Engine::BinFrontier(LevelCtx& c, Workspace& ws, const uint8_t* frontier, int64_t counts[kBins],
                         int64_t te[kBins]) {
  ResetCounters(ws.rc.get(), true);
  LaunchBinFrontier(c.g, frontier, h_->th, c.layout, ws.list.get(), ws.rc.get(), s_);
  ReadCounters(ws.rc.get(), counts, te, nullptr);
}

void Engine::LaunchMoves(LevelCtx& c, Workspace& ws, const MoveParams& mp, bool async,
                         const uint32_t* const lists[kBins], const int64_t counts[kBins],
                         const int64_t te[kBins], const MoveOut& out, uint64_t salt,
                         bool scan_hubs, const RoundResetJob* reset) {
  if (!h_->timing) {
    LaunchBestMoveBins(counts, lists, async, h_->th, c.g, c.st, mp, out, ws.fb_list.get(),
                       ws.rc.get(), salt, s_, nullptr, reset);
  } else if (reset) {
    LaunchRoundReset(reset->n, reset->a, reset->b, nullptr, reset->rc, reset->counters, nullptr, s_);
  }
  // With timing, one launch per bin and one timer per reported group (the
  // three ranges of a hash group together).
  for (int b0 = 0; b0 < kHubBin && h_->timing;) {
    const int grp = StatGroupOf(b0);
    int b1 = b0;
    bool any = false;
    for (; b1 < kHubBin && StatGroupOf(b1) == grp; ++b1) any = any || counts[b1] > 0;
    if (any) {
      Timer tb = StartTimer();
      for (int b = b0; b < b1; ++b)
        if (counts[b] > 0)
          LaunchBestMoveBin(b, counts[b], async, h_->th, c.g, c.st, mp, lists[b], out,
                            ws.fb_list.get(), ws.rc.get(), salt, s_);
      StopTimer(tb, grp);
    }
    b0 = b1;
  }
  const int64_t c7 = counts[kBins - 1];
  if (c7 > 0) {
    const uint32_t* list7 = lists[kBins - 1];
    Timer tb = StartTimer();
    if (scan_hubs)
      LaunchHubScan(c.g, list7, c7, c.hb.part_edges, c.hb.part_off, c.hb.ent_off, c.hb.item_of, s_);
    LaunchBestMoveHub(c7, te[kBins - 1], async, c.g, c.st, mp, list7, c.hb, out,
                      ws.fb_list.get(), ws.rc.get(), s_);
    StopTimer(tb, StatGroupOf(kHubBin));
  }
}

int64_t Engine::RoundStep(LevelCtx& c, Workspace& ws, const uint32_t* vgraph, int G,
                          uint8_t* gmoved, int64_t counts[kBins], int64_t te[kBins]) {
  const CsrView& g = c.g;
  const StateView& st = c.st;
  const int64_t n = g.n;
  const bool async = cfg_.async != 0;
  rounds_ += 1;
  int64_t F = 0, TE = 0;
  for (int b = 0; b < kBins; ++b) {
    F += counts[b];
    TE += te[b];
    bin_edges_[StatGroupOf(b)] += te[b];
    bin_vertices_[StatGroupOf(b)] += counts[b];
  }
  edges_scanned_ += TE;
  frontier_vertices_ += F;
  bestmove_bytes_ += 12.0 * (double)TE + 40.0 * (double)F;

  MoveOut out;
  out.moves = ws.moves.get();
  out.moved = ws.moved.get();
  out.moved_clusters = ws.flags.get();
  // Pull can only win when the seeds' volume may exceed the threshold.  With
  // NBHR_MOVE the seeds are movers, a subset of the frontier, so a frontier
  // of volume TE + F below the threshold rules pull out for this round.
  const double pull_thr = ((double)g.m + (double)n) * h_->pull_fraction / 1024.0;
  const bool pull = c.pull && (cfg_.move_method != PCC_NBHR_MOVE || (double)(TE + F) > pull_thr);
  // One launch clears what the round accumulates into: the modified /
  // destination clusters (only async best moves write them before the
  // apply), the next frontier, the movers when the pull expansion of
  // NBHR_MOVE reads every vertex's flag, the round counters (the host has
  // read this frontier's bin counts already) and the seed volume.
  // Without pull the clears touch nothing the best-move kernels write, so
  // sync rounds do them in the bins-0-4 launch.
  const RoundResetJob reset{n, ws.flags.get(),
                            cfg_.move_method != PCC_ALL_MOVE ? ws.next.get() : nullptr,
                            ws.rc.get(), 3};
  const bool fuse_reset = !async && !pull && !h_->timing;
  if (!fuse_reset)
    LaunchRoundReset(n, ws.flags.get(), cfg_.move_method != PCC_ALL_MOVE ? ws.next.get() : nullptr,
                     (pull && cfg_.move_method == PCC_NBHR_MOVE) ? ws.moved.get() : nullptr,
                     ws.rc.get(), 3, pull ? c.vol.get() : nullptr, s_);
  const uint64_t salt =
      (async && cfg_.permute) ? (cfg_.permute_seed * 0x9E3779B97F4A7C15ull + (uint64_t)rounds_ + 1) : 0;
  Timer all = StartTimer();
  // Async rounds run as sub-rounds over disjoint hashed subsets of the
  // frontier, each launch seeing the moves of the previous ones.  The
  // quality of an async round falls as the fraction of the frontier deciding
  // at once grows; 1/64 of the frontier per sub-round (one vertex each when
  // the frontier is smaller than 64) keeps the objective at the level of a
  // sequential run.  Small frontiers matter most: on contracted levels a few
  // heavy vertices moving at once can merge into one oversized cluster.
  int subs = 1;
  if (async) {
    subs = h_->async_subrounds > 0 ? h_->async_subrounds
                                   : (int)std::min<int64_t>(64, std::max<int64_t>(1, F));
  }
  const uint32_t* lists[kBins];
  int64_t scount[kBins];
  for (int b = 0; b < kBins; ++b) {
    lists[b] = ws.list.get() + c.layout.base[b];
    scount[b] = counts[b];
  }
  std::vector<uint32_t> hcnt, hoff;
  std::vector<unsigned long long> hte7;
  int64_t ste[kBins];
  for (int b = 0; b < kBins; ++b) ste[b] = te[b];
  if (subs > 1) {
    // Group every bin's list by sub-round so that each launch sees only its
    // own vertices.
    const size_t K = (size_t)kBins * (size_t)subs;
    PCC_HIP_CHECK(hipMemsetAsync(ws.sub_cnt.get(), 0, sizeof(uint32_t) * K, s_));
    PCC_HIP_CHECK(hipMemsetAsync(ws.sub_te.get(), 0, sizeof(unsigned long long) * (size_t)subs, s_));
    LaunchSubCount(g, ws.list.get(), c.layout, counts, (uint32_t)subs, ws.sub_cnt.get(),
                   ws.sub_te.get(), s_);
    hcnt.resize(K);
    hoff.resize(K);
    hte7.resize((size_t)subs);
    PCC_HIP_CHECK(hipMemcpyAsync(hcnt.data(), ws.sub_cnt.get(), sizeof(uint32_t) * K,
                                 hipMemcpyDeviceToHost, s_));
    ReadBack(hte7.data(), ws.sub_te.get(), sizeof(unsigned long long) * (size_t)subs, s_);
    for (int b = 0; b < kBins; ++b) {
      uint32_t run = 0;
      for (int j = 0; j < subs; ++j) {
        hoff[(size_t)b * subs + j] = run;
        run += hcnt[(size_t)b * subs + j];
      }
    }
    PCC_HIP_CHECK(hipMemcpyAsync(ws.sub_cur.get(), hoff.data(), sizeof(uint32_t) * K,
                                 hipMemcpyHostToDevice, s_));
    LaunchSubScatter(ws.list.get(), c.layout, counts, (uint32_t)subs, ws.sub_cur.get(),
                     ws.list2.get(), s_);
  }
  FreeIds fids;
  if (async) PrepareFreeIds(al_, st, (uint32_t)n, G, c.mp.gbase, &fids, s_);
  for (int sub = 0; sub < subs; ++sub) {
    MoveParams mp = c.mp;
    mp.async_singleton_rule = async ? h_->async_singleton_rule : 0;
    if (async) {
      mp.free_ids = fids.ids.get();
      mp.free_cur = fids.cur.get();
      mp.free_end = fids.end.get();
    }
    if (subs > 1) {
      for (int b = 0; b < kBins; ++b) {
        const size_t k = (size_t)b * subs + sub;
        lists[b] = ws.list2.get() + c.layout.base[b] + hoff[k];
        scount[b] = hcnt[k];
      }
      ste[kBins - 1] = (int64_t)hte7[(size_t)sub];
    }
    LaunchMoves(c, ws, mp, async, lists, scount, ste, out, salt, subs > 1 || sub == 0,
                fuse_reset ? &reset : nullptr);
  }
  if (subs > 1) {
    // The frontier expansion walks bin 7 of the whole list.
    PCC_HIP_CHECK(hipStreamSynchronize(s_));  // hoff / hcnt stay valid until here
    if (counts[kBins - 1] > 0)
      LaunchHubScan(g, ws.list.get() + c.layout.base[kBins - 1], counts[kBins - 1],
                    c.hb.part_edges, c.hb.part_off, c.hb.ent_off, c.hb.item_of, s_);
  }
  int64_t hashed_count = 0;
  for (int b = kWaveBin0; b < kBins; ++b) hashed_count += counts[b];
  // Sync rounds list fallback vertices only from the hub bin or when every
  // hashed vertex is forced through the ordered kernel (each hash bin's
  // degrees fit its table, so its overflow guard never fires).
  if (!async && (c.mp.force_ordered ? hashed_count > 0 : counts[kBins - 1] > 0)) {
    Timer tf = StartTimer();
    LaunchBestMoveOrdered(g, st, c.mp, ws.fb_list.get(), ws.rc.get(), out, c.hs, s_);
    StopTimer(tf, 8);
  }
  StopTimer(all, 11);

  Timer ta = StartTimer();
  if (!async) {
    // MoveNodesToCluster (PCCI.cc:256-369).
    if (exact_nw_)
      LaunchApply(counts, c.layout, ws.list.get(), st, c.mp.new_id, ws.moves.get(), ws.flags.get(),
                  ws.new_flag.get(), ws.rc.get(), s_);
    else
      LaunchApplyOrdered(al_, counts, c.layout, ws.list.get(), st, (uint32_t)n, c.mp.new_id,
                         ws.moves.get(), ws.flags.get(), ws.new_flag.get(), ws.rc.get(), s_);
    RunNewClusterAssign(al_, st, (uint32_t)n, c.mp.vgraph, c.mp.gbase, ws.new_flag.get(),
                        ws.flags.get(), ws.rc.get(), s_);
  }
  StopTimer(ta, 9);
  if (cfg_.clustering_moves_method == PCC_DEFAULT_CLUSTER_MOVES && cfg_.subclustering_method != 0 &&
      !async) {
    // Subcluster moves, then moves of the clusters as they were before them
    // (PCC.cc:175-297, sync): both are set moves of vertex groups.
    Timer tm = StartTimer();
    DevBuf<uint32_t> pre(al_, n), sub, parent;
    PCC_HIP_CHECK(hipMemcpyAsync(pre.get(), st.ids, 4 * (size_t)n, hipMemcpyDeviceToDevice, s_));
    (void)SubclusterIds(al_, g, st.ids, cfg_.subclustering_method, &sub, &parent,
                        cfg_.edge_weight_offset, cfg_.resolution, s_, false,
                        h_->symmetric == 1);
    parent.Release();
    const int64_t sub_moves = RunGroupMoves(al_, g, st, c.mp, sub.get(), ws.flags.get(),
                                            ws.new_flag.get(), ws.rc.get(), s_);
    cluster_moves_ += sub_moves;
    sub.Release();
    if (sub_moves > 0) {
      cluster_moves_ += RunGroupMoves(al_, g, st, c.mp, pre.get(), ws.flags.get(),
                                      ws.new_flag.get(), ws.rc.get(), s_);
    } else {
      // Nothing moved: the earlier clusters are the current ones, whose set
      // moves the whole-cluster kernels compute (class-1-together and class-2
      // terms are empty, so both give the same values).
      cluster_moves_ += RunClusterMoves(al_, g, st, c.mp, G, counts, c.layout, ws.list.get(),
                                        false, 1, 0, ws.flags.get(), ws.new_flag.get(),
                                        ws.rc.get(), s_);
    }
    StopTimer(tm, 12);
  } else if (cfg_.clustering_moves_method == PCC_DEFAULT_CLUSTER_MOVES) {
    // Whole-cluster moves after the vertex moves (PCC.cc:167-297); the
    // clusters they modify join the vertex phase's in ws.flags.
    Timer tm = StartTimer();
    cluster_moves_ += RunClusterMoves(al_, g, st, c.mp, G, counts, c.layout, ws.list.get(), async,
                                      (uint32_t)h_->cluster_subrounds,
                                      (uint32_t)rounds_ * 0x9E3779B9u,
                                      ws.flags.get(), ws.new_flag.get(), ws.rc.get(), s_);
    StopTimer(tm, 12);
  }
  ta = StartTimer();
  if (cfg_.move_method == PCC_ALL_MOVE) {  // PCC.cc:299-307: frontier unchanged
    // The next round's bins are this round's, so nothing is read back: the
    // host keeps launching while the device works.
    StopTimer(ta, 9);
    return F;
  }
  // Next frontier = neighbours of the seeds (PCC.cc:309-331).
  const bool by_cluster = cfg_.move_method == PCC_NBHR_CLUSTER_MOVE;
  const uint32_t* seed_ids = by_cluster ? st.ids : nullptr;
  ExpandDir dir;
  if (pull) {
    // Direction by seed volume: the pull kernels or the push kernels run,
    // the other side returns at once.
    LaunchSeedVolume(g, seed_ids, ws.flags.get(), ws.moved.get(), c.vol.get(), s_);
    dir.vol = c.vol.get();
    dir.thr = (unsigned long long)pull_thr;
  }
  if (by_cluster) {
    LaunchExpandClusters(g, st.ids, ws.flags.get(), h_->th.block_max, c.Hubs(), h_->hub_part_edges,
                         dir, ws.next.get(), s_);
  } else {
    LaunchExpandMoved(counts, te, c.layout, ws.list.get(), c.hb.part_off, c.hb.item_of,
                      c.hb.part_edges, g, ws.moved.get(), dir, ws.next.get(), s_);
  }
  if (pull)
    LaunchPullFrontier(g, seed_ids, ws.flags.get(), ws.moved.get(), h_->th.block_max, c.Hubs(),
                       h_->hub_part_edges, dir, ws.next.get(), s_);
  LaunchGraphFlags(n, ws.next.get(), vgraph, gmoved, s_);
  LaunchBinFrontier(g, ws.next.get(), h_->th, c.layout, ws.list.get(), ws.rc.get(), s_);
  StopTimer(ta, 9);
  ReadCounters(ws.rc.get(), counts, te, nullptr);
  int64_t f = 0;
  for (int b = 0; b < kBins; ++b) f += counts[b];
  return f;
}

bool Engine::DeviceRounds() const {
  return h_->device_rounds != 0 && !cfg_.async && cfg_.clustering_moves_method == PCC_LOUVAIN &&
         cfg_.subclustering_method == 0 && exact_nw_;
}

void Engine::RoundStepDevice(LevelCtx& c, Workspace& ws, const uint32_t* vgraph, uint8_t* gmoved) {
  const CsrView& g = c.g;
  const StateView& st = c.st;
  const int64_t n = g.n;
  const bool all_move = cfg_.move_method == PCC_ALL_MOVE;
  const bool nbhr = cfg_.move_method == PCC_NBHR_MOVE;
  DeviceCounts* dc = ws.dc.get();
  const double pull_thr = ((double)g.m + (double)n) * h_->pull_fraction / 1024.0;
  // The counts binned by the previous round become this round's; the ALL_MOVE
  // frontier never changes, so its counts stay.
  LaunchRoundBegin(ws.rc.get(), dc, ws.dstats.get(), all_move, nbhr, pull_thr, s_);
  LaunchRoundReset(n, ws.flags.get(), all_move ? nullptr : ws.next.get(),
                   (c.pull && nbhr) ? ws.moved.get() : nullptr, ws.rc.get(), 2,
                   c.pull ? c.vol.get() : nullptr, s_, &dc->pull_possible);
  MoveOut out;
  out.moves = ws.moves.get();
  out.moved = ws.moved.get();
  out.moved_clusters = ws.flags.get();
  Timer all = StartTimer();
  if (!h_->timing) {
    const uint32_t* lists[kHubBin];
    for (int b = 0; b < kHubBin; ++b) lists[b] = ws.list.get() + c.layout.base[b];
    LaunchBestMoveBins(c.caps, lists, false, h_->th, g, st, c.mp, out, ws.fb_list.get(),
                       ws.rc.get(), 0, s_, dc);
  }
  for (int b0 = 0; b0 < kHubBin && h_->timing;) {
    const int grp = StatGroupOf(b0);
    int b1 = b0;
    bool any = false;
    for (; b1 < kHubBin && StatGroupOf(b1) == grp; ++b1) any = any || c.caps[b1] > 0;
    if (any) {
      Timer tb = StartTimer();
      for (int b = b0; b < b1; ++b)
        if (c.caps[b] > 0)
          LaunchBestMoveBin(b, c.caps[b], false, h_->th, g, st, c.mp,
                            ws.list.get() + c.layout.base[b], out, ws.fb_list.get(), ws.rc.get(), 0,
                            s_, dc);
      StopTimer(tb, grp);
    }
    b0 = b1;
  }
  const int64_t c7 = c.caps[kBins - 1];
  if (c7 > 0) {
    const uint32_t* list7 = ws.list.get() + c.layout.base[kBins - 1];
    Timer tb = StartTimer();
    LaunchHubScan(g, list7, c7, c.hb.part_edges, c.hb.part_off, c.hb.ent_off, c.hb.item_of, s_,
                  &dc->cnt[kBins - 1]);
    LaunchBestMoveHub(c7, c.caps_te[kBins - 1], false, g, st, c.mp, list7, c.hb, out,
                      ws.fb_list.get(), ws.rc.get(), s_, dc);
    StopTimer(tb, StatGroupOf(kHubBin));
  }
  int64_t hashed = 0;
  for (int b = kWaveBin0; b < kBins; ++b) hashed += c.caps[b];
  if (c.mp.force_ordered ? hashed > 0 : c.caps[kBins - 1] > 0) {  // as in RoundStep
    Timer tf = StartTimer();
    LaunchBestMoveOrdered(g, st, c.mp, ws.fb_list.get(), ws.rc.get(), out, c.hs, s_);
    StopTimer(tf, 8);
  }
  StopTimer(all, 11);
  Timer ta = StartTimer();
  LaunchApply(c.caps, c.layout, ws.list.get(), st, c.mp.new_id, ws.moves.get(), ws.flags.get(),
              ws.new_flag.get(), ws.rc.get(), s_, dc);
  RunNewClusterAssign(al_, st, (uint32_t)n, c.mp.vgraph, c.mp.gbase, ws.new_flag.get(),
                      ws.flags.get(), ws.rc.get(), s_, &ws.na);
  if (all_move) {  // PCC.cc:299-307: frontier unchanged
    StopTimer(ta, 9);
    return;
  }
  // Next frontier = neighbours of the seeds (PCC.cc:309-331).
  const bool by_cluster = cfg_.move_method == PCC_NBHR_CLUSTER_MOVE;
  const uint32_t* seed_ids = by_cluster ? st.ids : nullptr;
  ExpandDir dir;
  if (c.pull) {
    LaunchSeedVolume(g, seed_ids, ws.flags.get(), ws.moved.get(), c.vol.get(), s_,
                     &dc->pull_possible);
    dir.vol = c.vol.get();
    dir.thr = (unsigned long long)pull_thr;
  }
  if (by_cluster) {
    LaunchExpandClusters(g, st.ids, ws.flags.get(), h_->th.block_max, c.Hubs(), h_->hub_part_edges,
                         dir, ws.next.get(), s_);
  } else {
    LaunchExpandMoved(c.caps, c.caps_te, c.layout, ws.list.get(), c.hb.part_off, c.hb.item_of,
                      c.hb.part_edges, g, ws.moved.get(), dir, ws.next.get(), s_, dc);
  }
  if (c.pull)
    LaunchPullFrontier(g, seed_ids, ws.flags.get(), ws.moved.get(), h_->th.block_max, c.Hubs(),
                       h_->hub_part_edges, dir, ws.next.get(), s_);
  LaunchGraphFlags(n, ws.next.get(), vgraph, gmoved, s_);
  LaunchBinFrontier(g, ws.next.get(), h_->th, c.layout, ws.list.get(), ws.rc.get(), s_);
  StopTimer(ta, 9);
}

std::vector<uint8_t> Engine::IterateBestMoves(const Level& L, const std::vector<uint8_t>* sel,
                                              int64_t inner) {
  const int64_t n = L.g.n;
  const int G = L.lay.G;
  LevelCtx* cp = nullptr;
  for (auto& e : ctx_cache_)
    if (e.first == &L) cp = e.second.get();
  std::unique_ptr<LevelCtx> fresh;
  if (!cp) {
    fresh = std::make_unique<LevelCtx>();
    PrepareLevel(L.g, L.State(), L.lay, fresh.get());
    cp = fresh.get();
    if (cfg_.refine) ctx_cache_.emplace_back(&L, std::move(fresh));
  }
  LevelCtx& c = *cp;
  DevBuf<uint8_t> gmoved(al_, G);
  Workspace ws(al_, n, s_, gmoved.get(), G);
  // Round 0 frontier: every vertex of the selected graphs (PCC.cc:346-352).
  bool all = true;
  if (sel)
    for (uint8_t x : *sel) all = all && x;
  DevBuf<uint8_t> fr0, dsel;
  if (!all) {
    dsel = Upload(al_, *sel, s_);
    fr0.Reset(al_, n > 0 ? n : 1);
    if (n > 0)
      hipLaunchKernelGGL(k_sel_frontier, dim3(Grid(n, 256, 8192)), dim3(256), 0, s_, n,
                         L.lay.vgraph.get(), dsel.get(), fr0.get());
  }
  int64_t counts[kBins], te[kBins];
  int64_t rounds = 0;
  if (DeviceRounds() && inner > 0) {
    // Every round enqueued back to back; the frontier, its bins and the loop
    // statistics stay on device.  Rounds after the frontier empties run as
    // no-ops (every list is empty), as the reference's loop would have
    // stopped there (PCC.cc:356-363).
    ResetCounters(ws.rc.get(), true);
    LaunchBinFrontier(c.g, all ? nullptr : fr0.get(), h_->th, c.layout, ws.list.get(), ws.rc.get(),
                      s_);
    LaunchRoundStatsInit(ws.dstats.get(), s_);
    if (h_->timing || inner == 1 || h_->device_rounds == 2) {
      for (int64_t it = 0; it < inner; ++it) RoundStepDevice(c, ws, L.lay.vgraph.get(), gmoved.get());
    } else {
      // One round captured as a graph (no memset nodes: every clear is a
      // kernel) and replayed for each round.
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      PCC_HIP_CHECK(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
      try {
        RoundStepDevice(c, ws, L.lay.vgraph.get(), gmoved.get());
      } catch (...) {
        (void)hipStreamEndCapture(s_, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      PCC_HIP_CHECK(hipStreamEndCapture(s_, &graph));
      const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (ie != hipSuccess) {
        (void)hipGraphDestroy(graph);
        PCC_HIP_CHECK(ie);
      }
      hipError_t le = hipSuccess;
      for (int64_t it = 0; it < inner && le == hipSuccess; ++it) le = hipGraphLaunch(exec, s_);
      const hipError_t se = hipStreamSynchronize(s_);
      (void)hipGraphExecDestroy(exec);
      (void)hipGraphDestroy(graph);
      PCC_HIP_CHECK(le);
      PCC_HIP_CHECK(se);
    }
    DeviceRoundStats hs;
    RoundCounters hrc;
    PCC_HIP_CHECK(hipMemcpyAsync(&hs, ws.dstats.get(), sizeof(hs), hipMemcpyDeviceToHost, s_));
    ReadBack(&hrc, ws.rc.get(), sizeof(hrc), s_);
    CollectTimers();
    rounds = (int64_t)hs.rounds;
    rounds_ += rounds;
    edges_scanned_ += (int64_t)hs.edges;
    frontier_vertices_ += (int64_t)hs.verts;
    bestmove_bytes_ += 12.0 * (double)hs.edges + 40.0 * (double)hs.verts;
    for (int k = 0; k < kStatGroups; ++k) {
      bin_edges_[k] += (int64_t)hs.bin_edges[k];
      bin_vertices_[k] += (int64_t)hs.bin_verts[k];
    }
    fallback_ += (int64_t)(hrc.fb_total + hrc.fb_count + hrc.repair_count);
  } else {
    if (all) {
      // Every vertex: the bins hold the level's capacities, known already
      // (the round counters are zero from the workspace's set-up).
      LaunchBinFrontier(c.g, nullptr, h_->th, c.layout, ws.list.get(), ws.rc.get(), s_);
      for (int b = 0; b < kBins; ++b) {
        counts[b] = c.caps[b];
        te[b] = c.caps_te[b];
      }
    } else {
      BinFrontier(c, ws, fr0.get(), counts, te);
    }
    bool local_moved = true;
    for (int64_t it = 0; it < inner && local_moved; ++it) {
      const int64_t fsize = RoundStep(c, ws, L.lay.vgraph.get(), G, gmoved.get(), counts, te);
      local_moved = fsize != 0;
      ++rounds;
    }
    // Ordered-path and repaired vertices of every round of this call, added
    // up on device (read once per Run, no synchronization here).
    if (!fb_acc_.get()) {
      fb_acc_.Reset(al_, 1);
      ZeroSpans z{};
      z.p[0] = reinterpret_cast<uint8_t*>(fb_acc_.get());
      z.bytes[0] = sizeof(unsigned long long);
      z.n = 1;
      LaunchZeroSpans(z, s_);
    }
    LaunchAddFallbacks(ws.rc.get(), fb_acc_.get(), s_);
  }
  std::vector<uint8_t> moved((size_t)G, 0);
  if (cfg_.move_method == PCC_ALL_MOVE) {
    // The frontier stays every vertex, so each selected non-empty graph reports
    // "moved" (PCC.cc:299-307, 361-362).
    for (int p = 0; p < G; ++p) {
      const bool selected = !sel || (*sel)[(size_t)p];
      moved[(size_t)p] = (rounds > 0 && selected && L.lay.base[(size_t)p + 1] > L.lay.base[(size_t)p]) ? 1 : 0;
    }
  } else {
    ReadBack(moved.data(), gmoved.get(), (size_t)G, s_);
  }
  return moved;
}

int Engine::Run(DevBuf<double> nw0, const double* gres, const uint32_t* init, uint32_t* host_out,
                std::vector<int64_t>* out_iter, DevBuf<uint32_t>* dev_global) {
  const CsrView g0 = h_->csr.View();
  const int64_t n = g0.n;
  const int G0 = h_->num_graphs();
  int64_t num_iterations = 0, num_inner = 0;
  if (cfg_.clustering_moves_method == PCC_LOUVAIN) {  // PCC.cc:493-500
    num_iterations = cfg_.louvain_num_iterations > 0 ? cfg_.louvain_num_iterations : 10;
    num_inner = cfg_.louvain_num_inner_iterations > 0 ? cfg_.louvain_num_inner_iterations : 10;
  } else if (cfg_.clustering_moves_method == PCC_DEFAULT_CLUSTER_MOVES) {  // PCC.cc:488-492
    num_iterations = 1;
    num_inner = cfg_.num_iterations > 0 ? cfg_.num_iterations : 20;
  } else {
    return SetError(PCC_UNIMPLEMENTED,
                    "Correlation clustering moves must be DEFAULT_CLUSTER_MOVES or LOUVAIN.");
  }
  // CompressSubclusters (PCC.cc:564-567) runs between LOUVAIN levels and the
  // subcluster moves inside sync DEFAULT_CLUSTER_MOVES rounds (PCC.cc:184-249).
  PCC_RETURN_IF_ERROR(CheckMovesConfig(h_, cfg_));
  const int subcl = cfg_.subclustering_method;
  exact_nw_ = NodeWeightsExact(al_, nw0.get(), n, s_);
  // The set moves accumulate group and cluster weights with atomics.
  if (!exact_nw_ && cfg_.clustering_moves_method == PCC_DEFAULT_CLUSTER_MOVES)
    return SetError(PCC_INVALID_ARGUMENT,
                    "DEFAULT_CLUSTER_MOVES needs integer node weights whose magnitudes sum "
                    "below 2^53");

  hipEvent_t e0, e1;
  PCC_HIP_CHECK(hipEventCreate(&e0));
  PCC_HIP_CHECK(hipEventCreate(&e1));
  PCC_HIP_CHECK(hipEventRecord(e0, s_));

  const bool batch = G0 > 1;
  const std::vector<int64_t>& base0 = h_->gbase;
  std::vector<double> res0((size_t)G0, cfg_.resolution);
  if (gres) res0.assign(gres, gres + G0);

  std::vector<std::unique_ptr<Level>> levels;
  levels.emplace_back(new Level());
  {
    Level& L = *levels.back();
    L.g = g0;
    L.nw = std::move(nw0);
    L.ids.Reset(al_, n > 0 ? n : 1);
    L.sizes.Reset(al_, n > 0 ? n : 1);
    L.K.Reset(al_, n > 0 ? n : 1);
    L.lay.G = G0;
    L.lay.batch = batch;
    L.lay.base = base0;
    for (int o = 0; o < G0; ++o) L.lay.orig.push_back(o);
    if (batch) {
      L.lay.dbase = Upload(al_, base0, s_);
      L.lay.vgraph.Reset(al_, n);
      PCC_HIP_CHECK(hipMemcpyAsync(L.lay.vgraph.get(), h_->vgraph.get(), 4 * (size_t)n,
                                   hipMemcpyDeviceToDevice, s_));
      L.lay.gres = Upload(al_, res0, s_);
    }
    if (init) {
      LaunchSetClustering(n, init, L.State(), s_);
      if (!exact_nw_) OrderedClusterWeights(al_, n, L.State(), s_);
    } else {
      LaunchSingletons(n, L.State(), s_);
    }
  }
  const uint32_t* vgraph0 = batch ? h_->vgraph.get() : nullptr;
  DevBuf<uint32_t> cluster_ids(al_, n > 0 ? n : 1);
  LaunchIota(cluster_ids.get(), n, s_);
  std::vector<int64_t> iter_o((size_t)G0, 0);      // final level of each graph
  std::vector<int64_t> cur_base(base0.begin(), base0.end() - 1);  // space of cluster_ids
  std::vector<uint8_t> live((size_t)G0, 1);

  int64_t iter = 0;
  for (iter = 0; iter < num_iterations; ++iter) {
    Level& L = *levels[(size_t)iter];
    const int Gp = L.lay.G;
    const std::vector<uint8_t> moved = IterateBestMoves(L, nullptr, num_inner);
    std::vector<uint8_t> sel_o((size_t)G0, 0);
    bool any = false;
    for (int p = 0; p < Gp; ++p) {
      const int o = L.lay.orig[(size_t)p];
      if (!moved[(size_t)p]) {  // PCC.cc:543-547
        iter_o[(size_t)o] = iter - 1;
        live[(size_t)o] = 0;
      } else {
        sel_o[(size_t)o] = 1;
        iter_o[(size_t)o] = iter;
        any = true;
      }
    }
    const bool last = iter == num_iterations - 1;
    if (!cfg_.refine && any && (!subcl || last)) {  // PCC.cc:550
      const bool every = std::all_of(sel_o.begin(), sel_o.end(), [](uint8_t x) { return x != 0; });
      DevBuf<uint8_t> dsel;
      if (!every) dsel = Upload(al_, sel_o, s_);
      LaunchFlattenShift(n, cluster_ids.get(), vgraph0, every ? nullptr : dsel.get(), nullptr,
                         L.ids.get(), s_);
    }
    // With subclustering a level starts from its parent clustering, which is
    // the result when nothing moves there.
    if (subcl && !cfg_.refine && !any && iter > 0)
      LaunchFlatten(n, cluster_ids.get(), L.ids.get(), cluster_ids.get(), s_);
    if (last || !any) break;  // PCC.cc:556
    if (subcl) {
      // CompressSubclusters (PCC.cc:436-474, 564-567): the subclusters become
      // the next level's vertices and start in their parent cluster.  The
      // next level's vertex of a vertex is its subcluster id, and that is what
      // the flattens compose (the reference composes the cluster ids there,
      // PCC.cc:550 and 594-595, indexing the next level by the wrong vertex).
      Timer tc = StartTimer();
      DevBuf<uint32_t> parent;
      const int64_t n2 = SubclusterIds(al_, L.g, L.ids.get(), subcl, &L.map, &parent,
                                       cfg_.edge_weight_offset, cfg_.resolution, s_, false,
                                       h_->symmetric == 1);
      auto next = std::make_unique<Level>();
      next->lay.G = 1;
      next->lay.batch = false;
      next->lay.base = {0, n2};
      next->lay.orig.push_back(0);
      CompressGraph(al_, L.g, L.map.get(), L.nw.get(), nullptr, nullptr, nullptr, n2, &next->own,
                    &next->nw, s_, h_->contract_method,
                    h_->contract_chunk);
      next->g = next->own.View();
      next->ids.Reset(al_, n2);
      next->sizes.Reset(al_, n2);
      next->K.Reset(al_, n2);
      LaunchSetClustering(n2, parent.get(), next->State(), s_);
      if (!exact_nw_) OrderedClusterWeights(al_, n2, next->State(), s_);
      StopTimer(tc, 10);
      if (!cfg_.refine) {
        LaunchFlatten(n, cluster_ids.get(), L.map.get(), cluster_ids.get(), s_);
        if (iter > 0) {
          L.own = DevCsr();
          L.ids.Release();
          L.sizes.Release();
          L.K.Release();
          L.nw.Release();
        }
        L.map.Release();
      }
      levels.push_back(std::move(next));
      continue;
    }
    // CompressGraph (PCC.cc:569-571) of the graphs that moved, and singleton
    // helpers with aggregated node weights (PCC.cc:583-585).
    Timer tc = StartTimer();
    DevBuf<uint32_t> gmax(al_, Gp);
    PCC_HIP_CHECK(hipMemsetAsync(gmax.get(), 0, 4 * (size_t)Gp, s_));
    LaunchGraphMaxIds(L.g.n, L.ids.get(), L.lay.vgraph.get(), gmax.get(), s_);
    std::vector<uint32_t> hmax((size_t)Gp);
    ReadBack(hmax.data(), gmax.get(), 4 * (size_t)Gp, s_);
    std::vector<int64_t> newbase((size_t)Gp, -1);
    auto next = std::make_unique<Level>();
    next->lay.batch = batch;
    int64_t n2 = 0;
    next->lay.base.push_back(0);
    std::vector<double> res_next;
    std::vector<int64_t> shift_o((size_t)G0, 0);
    for (int p = 0; p < Gp; ++p) {
      if (!moved[(size_t)p]) continue;
      const int o = L.lay.orig[(size_t)p];
      newbase[(size_t)p] = n2;
      shift_o[(size_t)o] = n2 - L.lay.base[(size_t)p];
      n2 += (int64_t)hmax[(size_t)p] - L.lay.base[(size_t)p] + 1;  // 1 + max id (PCCI.cc:713-717)
      next->lay.orig.push_back(o);
      next->lay.base.push_back(n2);
      res_next.push_back(res0[(size_t)o]);
    }
    next->lay.G = (int)next->lay.orig.size();
    if (batch) {
      DevBuf<int64_t> dnb = Upload(al_, newbase, s_);
      CompressGraph(al_, L.g, L.ids.get(), L.nw.get(), L.lay.vgraph.get(), L.lay.dbase.get(),
                    dnb.get(), n2, &next->own, &next->nw, s_, h_->contract_method,
                    h_->contract_chunk);
      next->lay.dbase = Upload(al_, next->lay.base, s_);
      next->lay.vgraph.Reset(al_, n2 > 0 ? n2 : 1);
      if (n2 > 0)
        hipLaunchKernelGGL(k_fill_vgraph, dim3(Grid(n2, 256, 8192)), dim3(256), 0, s_, n2,
                           next->lay.dbase.get(), next->lay.G, next->lay.vgraph.get());
      next->lay.gres = Upload(al_, res_next, s_);
    } else {
      CompressGraph(al_, L.g, L.ids.get(), L.nw.get(), nullptr, nullptr, nullptr, n2, &next->own,
                    &next->nw, s_, h_->contract_method,
                    h_->contract_chunk);
    }
    next->g = next->own.View();
    next->ids.Reset(al_, n2);
    next->sizes.Reset(al_, n2);
    next->K.Reset(al_, n2);
    LaunchSingletons(n2, next->State(), s_);
    StopTimer(tc, 10);
    if (!cfg_.refine) {
      // cluster_ids of the continuing graphs move to the next level's numbering.
      if (batch) {
        DevBuf<uint8_t> dsel = Upload(al_, sel_o, s_);
        DevBuf<int64_t> dshift = Upload(al_, shift_o, s_);
        LaunchFlattenShift(n, cluster_ids.get(), vgraph0, dsel.get(), dshift.get(), nullptr, s_);
      }
      for (int p = 0; p < Gp; ++p)
        if (moved[(size_t)p]) cur_base[(size_t)L.lay.orig[(size_t)p]] = newbase[(size_t)p];
      if (iter > 0) {  // only the current level is needed without refinement
        L.own = DevCsr();
        L.ids.Release();
        L.sizes.Release();
        L.K.Release();
        L.nw.Release();
      }
    }
    levels.push_back(std::move(next));
  }

  std::vector<uint8_t> use_ids0((size_t)G0, 0);
  std::vector<int64_t> sub(base0.begin(), base0.end() - 1);
  if (cfg_.refine) {  // PCC.cc:589-607
    int64_t top = -1;
    for (int o = 0; o < G0; ++o) top = std::max(top, iter_o[(size_t)o] - 1);
    for (int64_t i = top; i >= 0; --i) {
      Level& Li = *levels[(size_t)i];
      Level& Ln = *levels[(size_t)i + 1];
      std::vector<int> pos_next((size_t)G0, -1);
      for (int p = 0; p < Ln.lay.G; ++p) pos_next[(size_t)Ln.lay.orig[(size_t)p]] = p;
      std::vector<uint8_t> sel((size_t)Li.lay.G, 0);
      std::vector<int64_t> nbase((size_t)Li.lay.G, 0);
      bool anysel = false;
      for (int p = 0; p < Li.lay.G; ++p) {
        const int o = Li.lay.orig[(size_t)p];
        if (iter_o[(size_t)o] - 1 >= i && pos_next[(size_t)o] >= 0) {
          sel[(size_t)p] = 1;
          nbase[(size_t)p] = Ln.lay.base[(size_t)pos_next[(size_t)o]];
          anysel = true;
        }
      }
      if (!anysel) continue;
      if (subcl)  // the next level's vertex of each vertex is its subcluster
        PCC_HIP_CHECK(hipMemcpyAsync(Li.ids.get(), Li.map.get(), 4 * (size_t)Li.g.n,
                                     hipMemcpyDeviceToDevice, s_));
      if (batch) {
        DevBuf<uint8_t> dsel = Upload(al_, sel, s_);
        DevBuf<int64_t> dnb = Upload(al_, nbase, s_);
        ResetClusteringBatch(al_, Li.g.n, Ln.g.n, Ln.ids.get(), Li.lay.vgraph.get(),
                             Li.lay.dbase.get(), dnb.get(), dsel.get(), Li.State(), s_);
      } else {
        ResetClusteringBatch(al_, Li.g.n, Ln.g.n, Ln.ids.get(), nullptr, nullptr, nullptr,
                             nullptr, Li.State(), s_);
      }
      if (!exact_nw_) OrderedClusterWeights(al_, Li.g.n, Li.State(), s_);
      IterateBestMoves(Li, &sel, num_inner);
    }
    for (int o = 0; o < G0; ++o) use_ids0[(size_t)o] = iter_o[(size_t)o] > 0 ? 1 : 0;
    // refine with iter == 0 keeps the identity cluster_ids (PCC.cc:526-530, 589).
  } else {
    sub = cur_base;
  }
  PCC_HIP_CHECK(hipEventRecord(e1, s_));
  PCC_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  PCC_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  CollectTimers();
  if (n > 0) {
    DevBuf<uint8_t> duse = Upload(al_, use_ids0, s_);
    DevBuf<int64_t> dsub = Upload(al_, sub, s_);
    DevBuf<uint32_t> out(al_, n), global;
    DevBuf<uint32_t>& gl = dev_global ? *dev_global : global;
    gl.Reset(al_, n);
    hipLaunchKernelGGL(k_output_ids, dim3(Grid(n, 256, 8192)), dim3(256), 0, s_, n, vgraph0,
                       batch ? levels[0]->lay.dbase.get() : (const int64_t*)nullptr, duse.get(),
                       levels[0]->ids.get(), cluster_ids.get(), dsub.get(), out.get(), gl.get());
    PCC_HIP_CHECK(hipGetLastError());
    PCC_HIP_CHECK(hipMemcpyAsync(host_out, out.get(), sizeof(uint32_t) * (size_t)n,
                                 hipMemcpyDeviceToHost, s_));
  }
  PCC_HIP_CHECK(hipStreamSynchronize(s_));
  if (fb_acc_.get()) {
    unsigned long long fb = 0;
    ReadBack(&fb, fb_acc_.get(), sizeof(fb), s_);
    fallback_ += (int64_t)fb;
  }
  if (out_iter) *out_iter = iter_o;
  if (stats_) {
    int64_t lv = 0;
    for (int o = 0; o < G0; ++o) lv = std::max(lv, iter_o[(size_t)o] + 1);
    stats_->levels = lv;
    stats_->rounds = rounds_;
    stats_->edges_scanned = edges_scanned_;
    stats_->frontier_vertices = frontier_vertices_;
    stats_->fallback_vertices = fallback_;
    stats_->cluster_seconds = (double)ms * 1e-3;
    stats_->bestmove_bytes = bestmove_bytes_;
    stats_->bestmove_seconds = kind_seconds_[11];
    stats_->bestmove_launches = 0;
    for (int b = 0; b <= 8; ++b) stats_->bestmove_launches += kind_launches_[b];
    stats_->apply_seconds = kind_seconds_[9];
    stats_->contract_seconds = kind_seconds_[10];
    for (int b = 0; b < kStatGroups; ++b) {
      stats_->bin_seconds[b] = kind_seconds_[b];
      stats_->bin_edges[b] = bin_edges_[b];
      stats_->bin_vertices[b] = bin_vertices_[b];
      stats_->bin_launches[b] = kind_launches_[b];
    }
    stats_->ordered_seconds = kind_seconds_[8];
    stats_->cluster_move_seconds = kind_seconds_[12];
    stats_->cluster_moves = cluster_moves_;
  }
  return PCC_OK;
}

}  // namespace pcc
