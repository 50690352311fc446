// Seeded synthetic data for tests and benchmarks (host code).
//
// Counter-based: element e of stream s under seed S is a pure function of
// (S, s, e) -- splitmix64 of a per-stream key advanced by e golden-ratio
// steps -- so the output does not depend on the thread count or on the order
// of generation. Gaussians use Box-Muller in double (cos and sin of one
// uniform pair fill two consecutive elements of a row).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "common.hpp"
#include "rplsh.h"

namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ULL;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
  return mix64(mix64(seed + kGamma) ^ (stream * 0xD1B54A32D192ED03ULL));
}

inline uint64_t draw(uint64_t key, uint64_t ctr) { return mix64(key + (ctr + 1) * kGamma); }

// [0, 1) with 24 random bits: exactly representable in float.
inline float unit24(uint64_t u) { return (float)(u >> 40) * (1.0f / 16777216.0f); }

// (0, 1] and [0, 1) with 53 bits for Box-Muller.
inline double open01(uint64_t u) { return ((double)(u >> 11) + 1.0) * (1.0 / 9007199254740992.0); }
inline double half01(uint64_t u) { return (double)(u >> 11) * (1.0 / 9007199254740992.0); }

// Standard normals for row i: pairs (2p, 2p+1) from counter i*pairs + p.
void gaussian_row(uint64_t key, int64_t i, int d, double* z) {
  const int64_t pairs = (d + 1) / 2;
  for (int64_t p = 0; p < pairs; ++p) {
    uint64_t c = (uint64_t)(i * pairs + p);
    double u1 = open01(draw(key, 2 * c));
    double u2 = half01(draw(key, 2 * c + 1));
    double r = std::sqrt(-2.0 * std::log(u1));
    double a = 6.283185307179586 * u2;
    z[2 * p] = r * std::cos(a);
    if (2 * p + 1 < d) z[2 * p + 1] = r * std::sin(a);
  }
}

template <typename F>
void parallel_rows(int64_t n, F&& f) {
  unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  int64_t nt = std::min<int64_t>(hw, std::max<int64_t>(1, n / 4096));
  if (nt <= 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> ts;
  for (int64_t t = 0; t < nt; ++t) {
    int64_t a = n * t / nt, b = n * (t + 1) / nt;
    ts.emplace_back([&f, a, b] { f(a, b); });
  }
  for (auto& t : ts) t.join();
}

}  // namespace

using namespace rplsh;

extern "C" {

rplsh_status rplsh_gen_uniform(float* out, int64_t n, int d, uint64_t seed, float lo, float hi) {
  return guard([&] {
    RPLSH_REQUIRE(out || n == 0, "out is NULL");
    RPLSH_REQUIRE(n >= 0 && d > 0, "bad shape");
    const uint64_t key = stream_key(seed, 0);
    const float span = hi - lo;
    parallel_rows(n, [&](int64_t a, int64_t b) {
      for (int64_t i = a; i < b; ++i)
        for (int j = 0; j < d; ++j) {
          uint64_t e = (uint64_t)(i * d + j);
          out[i * d + j] = lo + span * unit24(draw(key, e));
        }
    });
  });
}

rplsh_status rplsh_gen_gaussian(float* out, int64_t n, int d, uint64_t seed, int unit_normalize) {
  return guard([&] {
    RPLSH_REQUIRE(out || n == 0, "out is NULL");
    RPLSH_REQUIRE(n >= 0 && d > 0, "bad shape");
    const uint64_t key = stream_key(seed, 3);
    parallel_rows(n, [&](int64_t a, int64_t b) {
      std::vector<double> z(d);
      for (int64_t i = a; i < b; ++i) {
        gaussian_row(key, i, d, z.data());
        double s = 1.0;
        if (unit_normalize) {
          double ss = 0.0;
          for (int j = 0; j < d; ++j) ss += z[j] * z[j];
          s = ss > 0.0 ? 1.0 / std::sqrt(ss) : 1.0;
        }
        for (int j = 0; j < d; ++j) out[i * d + j] = (float)(z[j] * s);
      }
    });
  });
}

rplsh_status rplsh_gen_mixture(float* out, int64_t n, int d, int centers, float sigma,
                               uint64_t center_seed, uint64_t point_seed, double dup_fraction) {
  return guard([&] {
    RPLSH_REQUIRE(out || n == 0, "out is NULL");
    RPLSH_REQUIRE(n >= 0 && d > 0 && centers > 0, "bad shape");
    RPLSH_REQUIRE(dup_fraction >= 0.0 && dup_fraction < 1.0, "dup_fraction must be in [0,1)");
    std::vector<float> cen((size_t)centers * d);
    const uint64_t ckey = stream_key(center_seed, 1);
    for (int64_t e = 0; e < (int64_t)centers * d; ++e) cen[e] = -1.0f + 2.0f * unit24(draw(ckey, e));
    const uint64_t akey = stream_key(point_seed, 2);
    const uint64_t nkey = stream_key(point_seed, 3);
    const uint64_t fkey = stream_key(point_seed, 4);
    const uint64_t skey = stream_key(point_seed, 5);
    auto is_dup = [&](int64_t i) {
      return dup_fraction > 0.0 && half01(draw(fkey, (uint64_t)i)) < dup_fraction;
    };
    parallel_rows(n, [&](int64_t a, int64_t b) {
      std::vector<double> z(d);
      for (int64_t i = a; i < b; ++i) {
        int c = (int)(draw(akey, (uint64_t)i) % (uint64_t)centers);
        gaussian_row(nkey, i, d, z.data());
        const float* cv = cen.data() + (size_t)c * d;
        for (int j = 0; j < d; ++j) out[i * d + j] = (float)((double)cv[j] + (double)sigma * z[j]);
      }
    });
    if (dup_fraction > 0.0 && n > 1) {
      // Second phase: duplicate rows copy a non-duplicate source row, which
      // phase one already wrote and this phase never overwrites.
      parallel_rows(n, [&](int64_t a, int64_t b) {
        for (int64_t i = a; i < b; ++i) {
          if (!is_dup(i)) continue;
          int64_t src = (int64_t)(draw(skey, (uint64_t)i) % (uint64_t)n);
          int64_t tries = 0;
          while (is_dup(src) && tries < n) {
            src = (src + 1) % n;
            ++tries;
          }
          if (!is_dup(src)) std::memcpy(out + i * d, out + src * d, sizeof(float) * d);
        }
      });
    }
  });
}

}  // extern "C"
