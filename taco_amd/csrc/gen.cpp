// Seeded synthetic CSR inputs for the hot-path workloads (host C++/OpenMP).
//
// Every random draw is a pure function of (seed, stream, counter) through
// splitmix64 finalisers, so results do not depend on thread count or order
// and the same generator can be restated on the device.  Compiled with g++
// (libgomp, the OpenMP runtime PyTorch also loads).

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "taco_hip.h"

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

struct Rng {
  uint64_t key;
  uint64_t ctr = 0;
  Rng(uint64_t seed, uint64_t stream) : key(mix64(seed ^ mix64(stream))) {}
  uint64_t next() { return mix64(key + ctr++); }
  // Unbiased enough for n << 2^64: multiply-high of a 64-bit draw.
  uint32_t below(uint32_t n) {
    return (uint32_t)(((unsigned __int128)next() * n) >> 64);
  }
  double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

constexpr uint64_t kCountStream = 0x636f756e74ull;   // "count"
constexpr uint64_t kColumnStream = 0x636f6c73ull;    // "cols"

// k distinct values in [0, w), sorted, appended to out.  Draws k candidates,
// sorts, removes duplicates and tops up with fresh draws until k remain.
// For k > w/2 the complement (w - k values) is drawn instead.
void sample_distinct(Rng& rng, uint32_t w, uint32_t k, std::vector<int32_t>& out,
                     std::vector<int32_t>& tmp) {
  out.clear();
  if (k == 0) return;
  if (k >= w) {
    for (uint32_t c = 0; c < w; ++c) out.push_back((int32_t)c);
    return;
  }
  bool complement = (uint64_t)k * 2 > w;
  uint32_t want = complement ? w - k : k;
  tmp.clear();
  while (tmp.size() < want) {
    size_t need = want - tmp.size();
    for (size_t q = 0; q < need; ++q) tmp.push_back((int32_t)rng.below(w));
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
  }
  if (!complement) {
    out.swap(tmp);
    return;
  }
  size_t t = 0;
  for (uint32_t c = 0; c < w; ++c) {
    if (t < tmp.size() && tmp[t] == (int32_t)c) {
      ++t;
      continue;
    }
    out.push_back((int32_t)c);
  }
}

}  // namespace

extern "C" {

void taco_hip_gen_uniform_f64(int64_t n, uint64_t seed, uint64_t stream,
                              double lo, double hi, double* out) {
  const uint64_t key = mix64(seed ^ mix64(stream));
  const double span = hi - lo;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double u = (double)(mix64(key + (uint64_t)i) >> 11) *
               (1.0 / 9007199254740992.0);
    out[i] = lo + span * u;
  }
}

int64_t taco_hip_gen_row_counts_zipf(int32_t m, double alpha, int32_t kmin,
                                     int32_t kmax, double mean, uint64_t seed,
                                     int32_t* counts) {
  if (m <= 0 || kmin < 0 || kmax < kmin) return -1;
  // Inverse CDF over k in [kmin, kmax] with weight k^-alpha (k = 0 gets
  // weight 0 when kmin == 0).
  const int32_t span = kmax - kmin + 1;
  std::vector<double> cdf((size_t)span);
  double acc = 0;
  for (int32_t q = 0; q < span; ++q) {
    int32_t k = kmin + q;
    acc += k > 0 ? std::pow((double)k, -alpha) : 0.0;
    cdf[(size_t)q] = acc;
  }
  const double total = acc;
  int64_t sum = 0;
#pragma omp parallel for schedule(static) reduction(+ : sum)
  for (int32_t i = 0; i < m; ++i) {
    Rng rng(seed, kCountStream + (uint64_t)i);
    double u = rng.unit() * total;
    size_t q = (size_t)(std::upper_bound(cdf.begin(), cdf.end(), u) -
                        cdf.begin());
    if (q >= (size_t)span) q = (size_t)span - 1;
    counts[i] = kmin + (int32_t)q;
    sum += counts[i];
  }
  if (mean <= 0 || sum == 0) return sum;
  // Rescale to the requested mean: k' = clip(round(k * scale), kmin, kmax),
  // scale computed from the exact integer sum.
  const double scale = mean * (double)m / (double)sum;
  int64_t out = 0;
#pragma omp parallel for schedule(static) reduction(+ : out)
  for (int32_t i = 0; i < m; ++i) {
    double v = std::nearbyint((double)counts[i] * scale);
    if (v < kmin) v = kmin;
    if (v > kmax) v = kmax;
    counts[i] = (int32_t)v;
    out += counts[i];
  }
  return out;
}

int64_t taco_hip_gen_row_counts_uniform(int32_t m, int32_t kmin, int32_t kmax,
                                        uint64_t seed, int32_t* counts) {
  if (m < 0 || kmax < kmin || kmin < 0) return -1;
  int64_t sum = 0;
  const uint32_t span = (uint32_t)(kmax - kmin + 1);
#pragma omp parallel for schedule(static) reduction(+ : sum)
  for (int32_t i = 0; i < m; ++i) {
    Rng rng(seed, kCountStream + (uint64_t)i);
    counts[i] = kmin + (int32_t)rng.below(span);
    sum += counts[i];
  }
  return sum;
}

int taco_hip_counts_to_pos(int32_t m, const int32_t* counts, int32_t* pos) {
  int64_t acc = 0;
  pos[0] = 0;
  for (int32_t i = 0; i < m; ++i) {
    acc += counts[i];
    if (acc > INT32_MAX) return 1;
    pos[i + 1] = (int32_t)acc;
  }
  return 0;
}

int taco_hip_gen_csr_columns(int32_t m, int32_t n, const int32_t* pos,
                             int32_t structure, int32_t band, uint64_t seed,
                             int32_t* crd) {
  int bad = 0;
#pragma omp parallel reduction(| : bad)
  {
    std::vector<int32_t> cols, tmp;
#pragma omp for schedule(dynamic, 256)
    for (int32_t i = 0; i < m; ++i) {
      const int32_t cnt = pos[i + 1] - pos[i];
      int64_t lo = 0, hi = (int64_t)n - 1;  // inclusive window
      if (structure == 1) {
        lo = std::max<int64_t>(0, (int64_t)i - band);
        hi = std::min<int64_t>((int64_t)n - 1, (int64_t)i + band);
      } else if (structure == 2) {
        int64_t b = band > 0 ? band : 8;
        lo = ((int64_t)i / b) * b;
        hi = std::min<int64_t>((int64_t)n - 1, lo + b - 1);
      }
      const int64_t w = hi - lo + 1;
      if (cnt < 0 || w < 0 || cnt > w) {
        bad = 1;
        continue;
      }
      Rng rng(seed, kColumnStream + (uint64_t)i);
      sample_distinct(rng, (uint32_t)w, (uint32_t)cnt, cols, tmp);
      int32_t* dst = crd + pos[i];
      for (int32_t q = 0; q < cnt; ++q) dst[q] = (int32_t)(lo + cols[(size_t)q]);
    }
  }
  return bad;
}

}  // extern "C"
