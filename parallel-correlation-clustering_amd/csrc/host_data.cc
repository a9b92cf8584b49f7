// host_data.cc -- host-side data formats on either side of the clustering path:
// the text-format CorrelationClustererConfig parser, the GBBS adjacency graph
// reader/writer, and the seeded SBM / RMAT generators used for benchmark and
// test inputs.  Plain C++17 with std::thread; no device code.

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "pcc_amd.h"
#include "status.h"

namespace pcc {

namespace {
thread_local std::string t_last_error;
}  // namespace

int SetError(int code, const std::string& msg) {
  t_last_error = msg;
  return code;
}
void ClearError() { t_last_error.clear(); }
const char* LastError() { return t_last_error.c_str(); }

namespace {

// ---------------------------------------------------------------- threads --

int HostThreads() {
  const char* env = std::getenv("PCC_HOST_THREADS");
  int t = env ? std::atoi(env) : 0;
  if (t <= 0) t = (int)std::thread::hardware_concurrency();
  if (t <= 0) t = 1;
  return std::min(t, 64);
}

template <class F>
void ParallelFor(int64_t n, F f) {
  const int t = (int)std::min<int64_t>(HostThreads(), std::max<int64_t>(1, n / 4096));
  if (t <= 1) {
    f(0, 0, n);
    return;
  }
  std::vector<std::thread> th;
  for (int i = 0; i < t; ++i) {
    const int64_t s = n * i / t, e = n * (i + 1) / t;
    th.emplace_back([=] { f(i, s, e); });
  }
  for (auto& x : th) x.join();
}

// ------------------------------------------------------------------ PRNG --
// SplitMix64 (Steele, Lea, Flood 2014) as a counter-based hash.

inline uint64_t Mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t Hash3(uint64_t seed, uint64_t a, uint64_t b) {
  return Mix(Mix(Mix(seed) ^ a) ^ b);
}
// Uniform in (0, 1].
inline double Unit(uint64_t h) { return (double)((h >> 11) + 1) * 0x1.0p-53; }

constexpr uint64_t kWeightSalt = 0x57454947485453ull;  // "WEIGHTS"

inline float PairWeight(uint64_t seed, uint32_t u, uint32_t v, double lo, double hi) {
  if (lo == hi) return (float)lo;
  const uint32_t a = std::min(u, v), b = std::max(u, v);
  const double x = Unit(Hash3(seed ^ kWeightSalt, a, b));  // (0, 1]
  return (float)(lo + (hi - lo) * (1.0 - x));              // [lo, hi)
}

// ----------------------------------------------------- parallel radix sort --

void RadixSortU64(std::vector<uint64_t>& keys, int bits) {
  const int64_t n = (int64_t)keys.size();
  std::vector<uint64_t> tmp((size_t)n);
  const int T = (int)std::min<int64_t>(HostThreads(), std::max<int64_t>(1, n / 65536));
  for (int shift = 0; shift < bits; shift += 8) {
    std::vector<std::vector<int64_t>> hist(T, std::vector<int64_t>(256, 0));
    auto chunk = [&](int t, int64_t* s, int64_t* e) {
      *s = n * t / T;
      *e = n * (t + 1) / T;
    };
    {
      std::vector<std::thread> th;
      for (int t = 0; t < T; ++t)
        th.emplace_back([&, t] {
          int64_t s, e;
          chunk(t, &s, &e);
          auto& h = hist[t];
          for (int64_t i = s; i < e; ++i) h[(keys[(size_t)i] >> shift) & 255]++;
        });
      for (auto& x : th) x.join();
    }
    int64_t run = 0;
    for (int d = 0; d < 256; ++d)
      for (int t = 0; t < T; ++t) {
        const int64_t c = hist[t][d];
        hist[t][d] = run;
        run += c;
      }
    {
      std::vector<std::thread> th;
      for (int t = 0; t < T; ++t)
        th.emplace_back([&, t] {
          int64_t s, e;
          chunk(t, &s, &e);
          auto& h = hist[t];
          for (int64_t i = s; i < e; ++i) {
            const uint64_t k = keys[(size_t)i];
            tmp[(size_t)h[(k >> shift) & 255]++] = k;
          }
        });
      for (auto& x : th) x.join();
    }
    keys.swap(tmp);
  }
}

int BitsFor(uint64_t maxval) {
  int b = 0;
  while (b < 64 && (maxval >> b) != 0) ++b;
  return b;
}

// Builds a symmetric CSR from sorted, deduplicated (u << 32 | v) keys.
int KeysToCsr(const std::vector<uint64_t>& keys, int64_t n, uint64_t wseed, double lo,
              double hi, pcc_host_csr* out) {
  const int64_t m = (int64_t)keys.size();
  out->n = n;
  out->m = m;
  out->offsets = (int64_t*)std::malloc(sizeof(int64_t) * (size_t)(n + 1));
  out->nbrs = (uint32_t*)std::malloc(sizeof(uint32_t) * (size_t)std::max<int64_t>(m, 1));
  out->weights = (float*)std::malloc(sizeof(float) * (size_t)std::max<int64_t>(m, 1));
  if (!out->offsets || !out->nbrs || !out->weights) {
    pcc_host_csr_free(out);
    return SetError(PCC_RESOURCE_EXHAUSTED, "host allocation failed for generated graph");
  }
  std::memset(out->offsets, 0, sizeof(int64_t) * (size_t)(n + 1));
  for (int64_t i = 0; i < m; ++i) out->offsets[(keys[(size_t)i] >> 32) + 1]++;
  for (int64_t v = 0; v < n; ++v) out->offsets[v + 1] += out->offsets[v];
  ParallelFor(m, [&](int, int64_t s, int64_t e) {
    for (int64_t i = s; i < e; ++i) {
      const uint32_t u = (uint32_t)(keys[(size_t)i] >> 32);
      const uint32_t v = (uint32_t)keys[(size_t)i];
      out->nbrs[i] = v;
      out->weights[i] = PairWeight(wseed, u, v, lo, hi);
    }
  });
  return PCC_OK;
}

void SortUnique(std::vector<uint64_t>& keys, int64_t n) {
  const int vb = BitsFor((uint64_t)std::max<int64_t>(n - 1, 1));
  // Keys are (u << 32 | v); sort the low vb bits of v, then the vb bits of u.
  std::vector<uint64_t> packed(keys.size());
  ParallelFor((int64_t)keys.size(), [&](int, int64_t s, int64_t e) {
    for (int64_t i = s; i < e; ++i)
      packed[(size_t)i] = ((keys[(size_t)i] >> 32) << vb) | (keys[(size_t)i] & 0xFFFFFFFFull);
  });
  RadixSortU64(packed, 2 * vb);
  const uint64_t mask = (vb == 64) ? ~0ull : ((1ull << vb) - 1);
  size_t w = 0;
  for (size_t i = 0; i < packed.size(); ++i) {
    if (i > 0 && packed[i] == packed[i - 1]) continue;
    keys[w++] = ((packed[i] >> vb) << 32) | (packed[i] & mask);
  }
  keys.resize(w);
}

// Triangular pair index -> (i, j) with i < j, idx = j*(j-1)/2 + i.
inline void TriPair(uint64_t idx, uint64_t* i, uint64_t* j) {
  uint64_t jj = (uint64_t)((1.0 + std::sqrt(1.0 + 8.0 * (double)idx)) / 2.0);
  while (jj * (jj - 1) / 2 > idx) --jj;
  while ((jj + 1) * jj / 2 <= idx) ++jj;
  *j = jj;
  *i = idx - jj * (jj - 1) / 2;
}

int GenSbm(const std::vector<int64_t>& sizes, double p_in, double p_out, double lo, double hi,
           uint64_t seed, pcc_host_csr* out) {
  const int64_t B = (int64_t)sizes.size();
  std::vector<int64_t> start((size_t)B + 1, 0);
  for (int64_t b = 0; b < B; ++b) {
    if (sizes[(size_t)b] < 0) return SetError(PCC_INVALID_ARGUMENT, "negative block size");
    start[(size_t)b + 1] = start[(size_t)b] + sizes[(size_t)b];
  }
  const int64_t n = start[(size_t)B];
  if (n <= 0 || n >= (int64_t)0xFFFFFFFFll)
    return SetError(PCC_INVALID_ARGUMENT, "SBM vertex count out of range");
  if (!(p_in >= 0 && p_in <= 1 && p_out >= 0 && p_out <= 1))
    return SetError(PCC_INVALID_ARGUMENT, "SBM probabilities must lie in [0, 1]");
  std::vector<uint32_t> block_of((size_t)n);
  for (int64_t b = 0; b < B; ++b)
    std::fill(block_of.begin() + start[(size_t)b], block_of.begin() + start[(size_t)b + 1], (uint32_t)b);
  // Pairs are drawn by geometric skips over a pair index space (Batagelj &
  // Brandes 2005): each diagonal block with p_in on its own stream, and the
  // whole upper triangle with p_out split into a fixed number of chunks (so
  // the output does not depend on the thread count), rejecting pairs that
  // fall inside a block.
  constexpr int64_t kChunks = 256;
  const uint64_t tri = (uint64_t)n * (uint64_t)(n - 1) / 2;
  const int64_t jobs = B + kChunks;
  auto skip_stream = [](uint64_t st, double p, uint64_t count, auto&& emit) {
    if (p <= 0 || count == 0) return;
    const double lq = (p < 1) ? std::log1p(-p) : 0.0;
    uint64_t idx = 0;
    bool first = true;
    while (true) {
      uint64_t skip = 0;
      if (p < 1) {
        st = Mix(st);
        const double gsk = std::floor(std::log(Unit(st)) / lq);
        if (gsk >= (double)count) break;
        skip = (uint64_t)gsk;
      }
      idx = first ? skip : idx + 1 + skip;
      first = false;
      if (idx >= count) break;
      emit(idx);
    }
  };
  const int T = HostThreads();
  std::vector<std::vector<uint64_t>> parts((size_t)T);
  std::atomic<int64_t> next{0};
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t)
    th.emplace_back([&, t] {
      auto& out_keys = parts[(size_t)t];
      while (true) {
        const int64_t q = next.fetch_add(1);
        if (q >= jobs) break;
        if (q < B) {
          const uint64_t sa = (uint64_t)sizes[(size_t)q];
          const uint64_t base = (uint64_t)start[(size_t)q];
          skip_stream(Hash3(seed, (uint64_t)q, (uint64_t)q), p_in, sa * (sa - (sa ? 1 : 0)) / 2,
                      [&](uint64_t idx) {
                        uint64_t i, j;
                        TriPair(idx, &i, &j);
                        i += base;
                        j += base;
                        out_keys.push_back((i << 32) | j);
                        out_keys.push_back((j << 32) | i);
                      });
        } else {
          const uint64_t c = (uint64_t)(q - B);
          const uint64_t lo_i = tri * c / kChunks, hi_i = tri * (c + 1) / kChunks;
          skip_stream(Hash3(seed ^ 0x4F555445ull, c, (uint64_t)kChunks), p_out, hi_i - lo_i,
                      [&](uint64_t idx) {
                        uint64_t i, j;
                        TriPair(lo_i + idx, &i, &j);
                        if (block_of[(size_t)i] == block_of[(size_t)j]) return;
                        out_keys.push_back((i << 32) | j);
                        out_keys.push_back((j << 32) | i);
                      });
        }
      }
    });
  for (auto& x : th) x.join();
  size_t total = 0;
  for (auto& p : parts) total += p.size();
  std::vector<uint64_t> keys;
  keys.reserve(total);
  for (auto& p : parts) {
    keys.insert(keys.end(), p.begin(), p.end());
    std::vector<uint64_t>().swap(p);
  }
  SortUnique(keys, n);
  return KeysToCsr(keys, n, seed, lo, hi, out);
}

// ------------------------------------------------------- config parsing --

struct Lexer {
  const char* p;
  const char* end;
  void Skip() {
    while (p < end) {
      if (std::isspace((unsigned char)*p) || *p == ',' || *p == ';') {
        ++p;
      } else if (*p == '#') {
        while (p < end && *p != '\n') ++p;
      } else {
        break;
      }
    }
  }
  bool Eof() {
    Skip();
    return p >= end;
  }
  bool Peek(char c) {
    Skip();
    return p < end && *p == c;
  }
  bool Eat(char c) {
    if (Peek(c)) {
      ++p;
      return true;
    }
    return false;
  }
  bool Ident(std::string* out) {
    Skip();
    const char* s = p;
    while (p < end && (std::isalnum((unsigned char)*p) || *p == '_' || *p == '.')) ++p;
    if (p == s) return false;
    out->assign(s, p);
    return true;
  }
  bool Scalar(std::string* out) {
    Skip();
    if (p < end && (*p == '"' || *p == '\'')) {
      const char q = *p++;
      const char* s = p;
      while (p < end && *p != q) ++p;
      if (p >= end) return false;
      out->assign(s, p);
      ++p;
      return true;
    }
    const char* s = p;
    while (p < end && (std::isalnum((unsigned char)*p) || *p == '_' || *p == '.' || *p == '-' ||
                       *p == '+'))
      ++p;
    if (p == s) return false;
    out->assign(s, p);
    return true;
  }
};

bool ParseDouble(const std::string& s, double* v) {
  if (s.empty()) return false;
  std::string t = s;
  if (!t.empty() && (t.back() == 'f' || t.back() == 'F') && t.find_first_of("xX") == std::string::npos &&
      t != "inf" && t != "-inf")
    t.pop_back();
  char* e = nullptr;
  *v = std::strtod(t.c_str(), &e);
  return e && *e == '\0';
}
bool ParseInt(const std::string& s, int64_t* v) {
  if (s.empty()) return false;
  char* e = nullptr;
  *v = std::strtoll(s.c_str(), &e, 0);
  return e && *e == '\0';
}
bool ParseBool(const std::string& s, uint8_t* v) {
  if (s == "true" || s == "t" || s == "True" || s == "1") {
    *v = 1;
    return true;
  }
  if (s == "false" || s == "f" || s == "False" || s == "0") {
    *v = 0;
    return true;
  }
  return false;
}
bool ParseEnum(const std::string& s, const char* const* names, int count, int32_t* v) {
  for (int i = 0; i < count; ++i)
    if (s == names[i]) {
      *v = i;
      return true;
    }
  int64_t x;
  if (ParseInt(s, &x) && x >= 0 && x < count) {
    *v = (int32_t)x;
    return true;
  }
  return false;
}

// Skips a balanced message body after '{' or '<' was consumed.
bool SkipMessage(Lexer& lx, char close) {
  int depth = 1;
  while (lx.p < lx.end && depth > 0) {
    const char c = *lx.p;
    if (c == '"' || c == '\'') {
      std::string dummy;
      if (!lx.Scalar(&dummy)) return false;
      continue;
    }
    if (c == '{' || c == '<') ++depth;
    if (c == '}' || c == '>') --depth;
    ++lx.p;
  }
  (void)close;
  return depth == 0;
}

bool OpenMessage(Lexer& lx, char* close) {
  lx.Eat(':');
  if (lx.Eat('{')) {
    *close = '}';
    return true;
  }
  if (lx.Eat('<')) {
    *close = '>';
    return true;
  }
  return false;
}

const char* const kMoveMethods[] = {"NBHR_CLUSTER_MOVE", "ALL_MOVE", "NBHR_MOVE"};
const char* const kMovesMethods[] = {"DEFAULT_CLUSTER_MOVES", "LOUVAIN"};
const char* const kSubclustering[] = {"NONE_SUBCLUSTERING", "TRIANGLE_SUBCLUSTERING",
                                      "CONNECTIVITY_SUBCLUSTERING",
                                      "SLOW_CONNECTIVITY_SUBCLUSTERING"};
const char* const kPreclustering[] = {"NONE_PRECLUSTERING", "KCORE_PRECLUSTERING"};

int ParseLouvain(Lexer& lx, char close, pcc_config* cfg) {
  while (!lx.Eat(close)) {
    std::string name, val;
    if (!lx.Ident(&name)) return SetError(PCC_INVALID_ARGUMENT, "expected LouvainConfig field");
    if (!lx.Eat(':')) return SetError(PCC_INVALID_ARGUMENT, "expected ':' after " + name);
    if (!lx.Scalar(&val)) return SetError(PCC_INVALID_ARGUMENT, "missing value for " + name);
    int64_t x;
    if (!ParseInt(val, &x)) return SetError(PCC_INVALID_ARGUMENT, "bad integer for " + name);
    if (name == "num_iterations")
      cfg->louvain_num_iterations = x;
    else if (name == "num_inner_iterations")
      cfg->louvain_num_inner_iterations = x;
    else
      return SetError(PCC_INVALID_ARGUMENT, "unknown LouvainConfig field " + name);
  }
  return PCC_OK;
}

// CorrelationClustererConfig body (clustering/config.proto:94-165).
int ParseCorrelation(Lexer& lx, char close, pcc_config* cfg) {
  while (close ? !lx.Eat(close) : !lx.Eof()) {
    if (lx.Eof()) return SetError(PCC_INVALID_ARGUMENT, "unterminated message");
    std::string name;
    if (!lx.Ident(&name)) return SetError(PCC_INVALID_ARGUMENT, "expected a field name");
    if (name == "louvain_config" || name == "kcore_config") {
      char c2;
      if (!OpenMessage(lx, &c2)) return SetError(PCC_INVALID_ARGUMENT, "expected '{' after " + name);
      if (name == "louvain_config") {
        PCC_RETURN_IF_ERROR(ParseLouvain(lx, c2, cfg));
      } else if (!SkipMessage(lx, c2)) {
        return SetError(PCC_INVALID_ARGUMENT, "unterminated kcore_config");
      }
      continue;
    }
    std::string val;
    if (!lx.Eat(':')) return SetError(PCC_INVALID_ARGUMENT, "expected ':' after " + name);
    if (!lx.Scalar(&val)) return SetError(PCC_INVALID_ARGUMENT, "missing value for " + name);
    bool ok = true;
    int64_t x = 0;
    if (name == "resolution") {
      ok = ParseDouble(val, &cfg->resolution);
    } else if (name == "edge_weight_offset") {
      ok = ParseDouble(val, &cfg->edge_weight_offset);
    } else if (name == "random_seed") {
      ok = ParseInt(val, &x) && x >= 0 && x <= 0xFFFFFFFFll;
      cfg->random_seed = (uint32_t)x;
    } else if (name == "num_iterations") {
      ok = ParseInt(val, &cfg->num_iterations);
    } else if (name == "clustering_moves_method") {
      ok = ParseEnum(val, kMovesMethods, 2, &cfg->clustering_moves_method);
    } else if (name == "subclustering_method") {
      ok = ParseEnum(val, kSubclustering, 4, &cfg->subclustering_method);
    } else if (name == "preclustering_method") {
      ok = ParseEnum(val, kPreclustering, 2, &cfg->preclustering_method);
    } else if (name == "move_method") {
      ok = ParseEnum(val, kMoveMethods, 3, &cfg->move_method);
    } else if (name == "refine") {
      ok = ParseBool(val, &cfg->refine);
    } else if (name == "async") {
      ok = ParseBool(val, &cfg->async);
    } else if (name == "connect_stats") {
      ok = ParseBool(val, &cfg->connect_stats);
    } else if (name == "all_iter") {
      ok = ParseBool(val, &cfg->all_iter);
    } else if (name == "permute") {
      ok = ParseBool(val, &cfg->permute);
    } else {
      return SetError(PCC_INVALID_ARGUMENT, "unknown CorrelationClustererConfig field " + name);
    }
    if (!ok) return SetError(PCC_INVALID_ARGUMENT, "bad value '" + val + "' for " + name);
  }
  return PCC_OK;
}

}  // namespace
}  // namespace pcc

using namespace pcc;

extern "C" {

const char* pcc_last_error(void) { return pcc::LastError(); }

void pcc_config_default(pcc_config* cfg) {
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->clustering_moves_method = PCC_LOUVAIN;   // config.proto:127-128
  cfg->move_method = PCC_NBHR_CLUSTER_MOVE;     // config.proto:159-160
  cfg->permute_seed = 0x5EEDull;
}

int pcc_config_parse_text(const char* text, pcc_config* out) {
  ClearError();
  if (!text || !out) return SetError(PCC_INVALID_ARGUMENT, "null argument");
  pcc_config cfg;
  pcc_config_default(&cfg);
  Lexer lx{text, text + std::strlen(text)};
  // Either a ClustererConfig (clustering/config.proto:188-196) or the bare
  // CorrelationClustererConfig body.
  const char* save = lx.p;
  std::string name;
  bool wrapped = false;
  if (lx.Ident(&name) &&
      (name == "correlation_clusterer_config" || name == "affinity_clusterer_config" ||
       name == "parameter_string")) {
    wrapped = true;
  }
  lx.p = save;
  if (wrapped) {
    while (!lx.Eof()) {
      if (!lx.Ident(&name)) return SetError(PCC_INVALID_ARGUMENT, "expected a ClustererConfig field");
      if (name == "correlation_clusterer_config") {
        char close;
        if (!OpenMessage(lx, &close))
          return SetError(PCC_INVALID_ARGUMENT, "expected '{' after correlation_clusterer_config");
        PCC_RETURN_IF_ERROR(ParseCorrelation(lx, close, &cfg));
      } else if (name == "affinity_clusterer_config") {
        char close;
        if (!OpenMessage(lx, &close) || !SkipMessage(lx, close))
          return SetError(PCC_INVALID_ARGUMENT, "malformed affinity_clusterer_config");
      } else if (name == "parameter_string") {
        std::string v;
        if (!lx.Eat(':') || !lx.Scalar(&v))
          return SetError(PCC_INVALID_ARGUMENT, "malformed parameter_string");
      } else {
        return SetError(PCC_INVALID_ARGUMENT, "unknown ClustererConfig field " + name);
      }
    }
  } else {
    PCC_RETURN_IF_ERROR(ParseCorrelation(lx, 0, &cfg));
  }
  *out = cfg;
  return PCC_OK;
}

void pcc_host_csr_free(pcc_host_csr* csr) {
  if (!csr) return;
  std::free(csr->offsets);
  std::free(csr->nbrs);
  std::free(csr->weights);
  csr->offsets = nullptr;
  csr->nbrs = nullptr;
  csr->weights = nullptr;
  csr->n = csr->m = 0;
}

int pcc_gen_sbm(int64_t n, int64_t num_blocks, double p_in, double p_out, double w_lo,
                double w_hi, uint64_t seed, pcc_host_csr* out) {
  ClearError();
  if (!out || n <= 0 || num_blocks <= 0 || num_blocks > n)
    return SetError(PCC_INVALID_ARGUMENT, "bad SBM shape");
  std::memset(out, 0, sizeof(*out));
  std::vector<int64_t> sizes((size_t)num_blocks, n / num_blocks);
  sizes.back() += n - (n / num_blocks) * num_blocks;
  return GenSbm(sizes, p_in, p_out, w_lo, w_hi, seed, out);
}

int pcc_gen_sbm_blocks(int64_t num_blocks, const int64_t* block_sizes, double p_in,
                       double p_out, double w_lo, double w_hi, uint64_t seed,
                       pcc_host_csr* out) {
  ClearError();
  if (!out || num_blocks <= 0 || !block_sizes) return SetError(PCC_INVALID_ARGUMENT, "bad SBM shape");
  std::memset(out, 0, sizeof(*out));
  std::vector<int64_t> sizes(block_sizes, block_sizes + num_blocks);
  return GenSbm(sizes, p_in, p_out, w_lo, w_hi, seed, out);
}

int pcc_gen_rmat(int scale, double edge_factor, double a, double b, double c, double w_lo,
                 double w_hi, uint64_t seed, pcc_host_csr* out) {
  ClearError();
  if (!out || scale < 1 || scale > 31 || edge_factor <= 0 || a < 0 || b < 0 || c < 0 ||
      a + b + c > 1)
    return SetError(PCC_INVALID_ARGUMENT, "bad RMAT parameters");
  std::memset(out, 0, sizeof(*out));
  const int64_t n = (int64_t)1 << scale;
  const int64_t draws = (int64_t)std::llround(edge_factor * (double)n);
  std::vector<uint64_t> keys;
  try {
    keys.assign((size_t)(2 * draws), ~0ull);
  } catch (...) {
    return SetError(PCC_RESOURCE_EXHAUSTED, "host allocation failed for RMAT draws");
  }
  const double ab = a + b, abc = a + b + c;
  ParallelFor(draws, [&](int, int64_t s, int64_t e) {
    for (int64_t i = s; i < e; ++i) {
      uint64_t st = Hash3(seed, 0x524D4154ull, (uint64_t)i);
      uint64_t u = 0, v = 0;
      for (int l = 0; l < scale; ++l) {
        st = Mix(st);
        const double r = 1.0 - Unit(st);  // [0, 1)
        const uint64_t bu = r >= ab ? 1 : 0;
        const uint64_t bv = (r >= a && r < ab) || r >= abc ? 1 : 0;
        u = (u << 1) | bu;
        v = (v << 1) | bv;
      }
      if (u != v) {
        keys[(size_t)(2 * i)] = (u << 32) | v;
        keys[(size_t)(2 * i + 1)] = (v << 32) | u;
      }
    }
  });
  // Drop self-loop slots.
  size_t w = 0;
  for (size_t i = 0; i < keys.size(); ++i)
    if (keys[i] != ~0ull) keys[w++] = keys[i];
  keys.resize(w);
  SortUnique(keys, n);
  return KeysToCsr(keys, n, seed, w_lo, w_hi, out);
}

int pcc_read_adjacency_graph(const char* path, int float_weighted, pcc_host_csr* out) {
  ClearError();
  if (!path || !out) return SetError(PCC_INVALID_ARGUMENT, "null argument");
  std::memset(out, 0, sizeof(*out));
  FILE* f = std::fopen(path, "rb");
  if (!f) return SetError(PCC_NOT_FOUND, std::string("Unable to open file: ") + path);
  std::string data;
  {
    char buf[1 << 16];
    size_t r;
    while ((r = std::fread(buf, 1, sizeof(buf), f)) > 0) data.append(buf, r);
    std::fclose(f);
  }
  const char* p = data.c_str();
  const char* end = p + data.size();
  auto next_tok = [&](std::string* t) {
    while (p < end && std::isspace((unsigned char)*p)) ++p;
    const char* s = p;
    while (p < end && !std::isspace((unsigned char)*p)) ++p;
    t->assign(s, p);
    return p > s;
  };
  std::string tok;
  if (!next_tok(&tok)) return SetError(PCC_INVALID_ARGUMENT, "empty graph file");
  const bool weighted_header = tok == "WeightedAdjacencyGraph";
  if (tok != "AdjacencyGraph" && !weighted_header)
    return SetError(PCC_INVALID_ARGUMENT, "not a GBBS adjacency graph: header " + tok);
  if (float_weighted && !weighted_header)
    return SetError(PCC_INVALID_ARGUMENT, "--float_weighted needs a WeightedAdjacencyGraph");
  int64_t n = 0, m = 0;
  if (!next_tok(&tok) || !ParseInt(tok, &n) || !next_tok(&tok) || !ParseInt(tok, &m) || n < 0 ||
      m < 0 || n >= (int64_t)0xFFFFFFFFll)
    return SetError(PCC_INVALID_ARGUMENT, "bad n/m in graph header");
  out->n = n;
  out->m = m;
  out->offsets = (int64_t*)std::malloc(sizeof(int64_t) * (size_t)(n + 1));
  out->nbrs = (uint32_t*)std::malloc(sizeof(uint32_t) * (size_t)std::max<int64_t>(m, 1));
  out->weights = (float*)std::malloc(sizeof(float) * (size_t)std::max<int64_t>(m, 1));
  if (!out->offsets || !out->nbrs || !out->weights) {
    pcc_host_csr_free(out);
    return SetError(PCC_RESOURCE_EXHAUSTED, "host allocation failed reading graph");
  }
  auto fail = [&](const std::string& msg) {
    pcc_host_csr_free(out);
    return SetError(PCC_INVALID_ARGUMENT, msg);
  };
  for (int64_t i = 0; i < n; ++i) {
    int64_t x;
    if (!next_tok(&tok) || !ParseInt(tok, &x) || x < 0 || x > m) return fail("bad offset");
    out->offsets[i] = x;
    if (i > 0 && x < out->offsets[i - 1]) return fail("offsets not monotone");
  }
  out->offsets[n] = m;
  for (int64_t i = 0; i < m; ++i) {
    int64_t x;
    if (!next_tok(&tok) || !ParseInt(tok, &x) || x < 0 || x >= n) return fail("bad edge target");
    out->nbrs[i] = (uint32_t)x;
  }
  for (int64_t i = 0; i < m; ++i) {
    if (weighted_header) {
      double x;
      if (!next_tok(&tok) || !ParseDouble(tok, &x)) return fail("bad edge weight");
      // An unweighted read of a weighted file ignores the weights, as
      // read_unweighted_symmetric_graph does; weights become 1.0f.
      out->weights[i] = float_weighted ? (float)x : 1.0f;
    } else {
      out->weights[i] = 1.0f;
    }
  }
  return PCC_OK;
}

int pcc_write_adjacency_graph(const char* path, const pcc_host_csr* csr, int float_weighted) {
  ClearError();
  if (!path || !csr) return SetError(PCC_INVALID_ARGUMENT, "null argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return SetError(PCC_NOT_FOUND, std::string("Unable to open file: ") + path);
  std::fprintf(f, "%s\n%lld\n%lld\n", float_weighted ? "WeightedAdjacencyGraph" : "AdjacencyGraph",
               (long long)csr->n, (long long)csr->m);
  for (int64_t i = 0; i < csr->n; ++i) std::fprintf(f, "%lld\n", (long long)csr->offsets[i]);
  for (int64_t i = 0; i < csr->m; ++i) std::fprintf(f, "%u\n", csr->nbrs[i]);
  if (float_weighted)
    for (int64_t i = 0; i < csr->m; ++i) std::fprintf(f, "%.9g\n", (double)csr->weights[i]);
  std::fclose(f);
  return PCC_OK;
}

const char* pcc_version(void) { return "pcc_amd 0.1 (gfx950)"; }

}  // extern "C"
