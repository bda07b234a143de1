// Seeded synthetic inputs: Voronoi mosaics with oriented sinusoid textures and
// Gaussian noise, and the 13x32 texton codebook trained by k-means++ on the
// image's own Schmid responses.  The reference reads real photographs and a
// codebook file (SPFeatures.cpp:317-323); these generators produce inputs of
// the same formats from a seed so that every test and benchmark is repeatable.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "gpb_common.hpp"

namespace gpb {
namespace {

// SplitMix64 finaliser used as a counter-based generator: value(stream, i).
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

enum Stream : uint64_t { kSiteX = 1, kSiteY, kColour, kTexture, kNoise, kClass, kKMeans };

struct Rng {
  uint64_t key;
  Rng(uint64_t seed, uint64_t stream) : key(mix64(mix64(seed) ^ (stream * 0xD1B54A32D192ED03ull))) {}
  uint64_t bits(uint64_t i) const { return mix64(key + i * 0x9E3779B97F4A7C15ull); }
  double uniform(uint64_t i) const { return double(bits(i) >> 11) * (1.0 / 9007199254740992.0); }
  // Box-Muller on the pair (2i, 2i+1).
  double normal(uint64_t i) const {
    const double u1 = uniform(2 * i), u2 = uniform(2 * i + 1);
    return std::sqrt(-2.0 * std::log1p(-u1)) * std::cos(2.0 * M_PI * u2);
  }
};

struct ClassStats {
  double mean_lo, mean_hi, p_lo, p_hi, amp;
};

}  // namespace

void synth_voronoi(const SynthParams& prm, uint8_t* rgb, int* region_out) {
  const int W = prm.width, H = prm.height, R = prm.regions;
  const Rng rsx(prm.seed, kSiteX), rsy(prm.seed, kSiteY), rcol(prm.seed, kColour), rtex(prm.seed, kTexture),
      rnoise(prm.seed, kNoise), rcls(prm.seed, kClass);
  std::vector<double> sx(R), sy(R), mean(3 * size_t(R)), cth(R), sth(R), per(R), phase(R), amp(R);
  // Three region classes (used with classes == 3): dark fine texture, mid
  // colours with medium texture, bright coarse texture.
  const ClassStats cls3[3] = {{0, 120, 4, 8, 30}, {80, 200, 8, 12, 20}, {135, 255, 12, 16, 10}};
  const ClassStats cls1 = {0, 255, 4, 16, 20};
  for (int r = 0; r < R; r++) {
    sx[r] = W * rsx.uniform(r);
    sy[r] = H * rsy.uniform(r);
    const ClassStats& c = (prm.classes == 3) ? cls3[std::min(2, int(3 * rcls.uniform(r)))] : cls1;
    for (int ch = 0; ch < 3; ch++) mean[3 * size_t(r) + ch] = c.mean_lo + (c.mean_hi - c.mean_lo) * rcol.uniform(3 * uint64_t(r) + ch);
    per[r] = c.p_lo + (c.p_hi - c.p_lo) * rtex.uniform(3 * uint64_t(r));
    const double th = M_PI * rtex.uniform(3 * uint64_t(r) + 1);
    cth[r] = std::cos(th);
    sth[r] = std::sin(th);
    phase[r] = 2 * M_PI * rtex.uniform(3 * uint64_t(r) + 2);
    amp[r] = c.amp;
  }
  // Bucket grid for exact nearest-site search (ties go to the lower index).
  const double cell = std::max(1.0, std::sqrt(double(W) * H / std::max(1, R)));
  const int gw = std::max(1, int(std::ceil(W / cell))), gh = std::max(1, int(std::ceil(H / cell)));
  std::vector<std::vector<int>> grid(size_t(gw) * gh);
  for (int r = 0; r < R; r++) {
    const int cx = std::min(gw - 1, int(sx[r] / cell)), cy = std::min(gh - 1, int(sy[r] / cell));
    grid[size_t(cy) * gw + cx].push_back(r);
  }
#pragma omp parallel for schedule(dynamic, 8)
  for (int y = 0; y < H; y++) {
    for (int x = 0; x < W; x++) {
      const int cx = std::min(gw - 1, int(x / cell)), cy = std::min(gh - 1, int(y / cell));
      double best = std::numeric_limits<double>::infinity();
      int bi = -1;
      for (int ring = 0;; ring++) {
        // Every site outside rings <= ring is at least (ring) * cell away.
        if (bi >= 0) {
          const double lim = double(ring - 1) * cell;
          if (lim > 0 && lim * lim > best) break;
        }
        if (ring > gw + gh) break;
        for (int gy = cy - ring; gy <= cy + ring; gy++) {
          if (gy < 0 || gy >= gh) continue;
          for (int gx = cx - ring; gx <= cx + ring; gx++) {
            if (gx < 0 || gx >= gw) continue;
            if (std::max(std::abs(gx - cx), std::abs(gy - cy)) != ring) continue;
            for (int r : grid[size_t(gy) * gw + gx]) {
              const double ddx = x - sx[r], ddy = y - sy[r];
              const double d = ddx * ddx + ddy * ddy;
              if (d < best || (d == best && r < bi)) {
                best = d;
                bi = r;
              }
            }
          }
        }
      }
      const size_t p = size_t(y) * W + x;
      if (region_out) region_out[p] = bi;
      const double tex = amp[bi] * std::sin(2 * M_PI * (x * cth[bi] + y * sth[bi]) / per[bi] + phase[bi]);
      for (int ch = 0; ch < 3; ch++) {
        double v = mean[3 * size_t(bi) + ch] + tex + prm.noise_sigma * rnoise.normal(3 * uint64_t(p) + ch);
        v = std::min(255.0, std::max(0.0, v));
        rgb[3 * p + ch] = uint8_t(std::nearbyint(v));
      }
    }
  }
}

// k-means++ seeding followed by Lloyd iterations (FP64) on the 13-d Schmid
// responses of a strided pixel sample.  Empty clusters keep their centre.
void make_codebook(const uint32_t* packed, int width, int height, uint64_t seed, int iters, int max_samples,
                   float* codebook) {
  const size_t N = size_t(width) * height;
  const size_t S = std::min(N, size_t(std::max(1, max_samples)));
  std::vector<double> pts(S * kSchmidNF);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < long(S); i++) {
    const size_t p = size_t(i) * N / S;
    float r[kSchmidNF];
    schmid_response_at(packed, width, height, int(p % width), int(p / width), r);
    for (int f = 0; f < kSchmidNF; f++) pts[size_t(i) * kSchmidNF + f] = r[f];
  }
  const Rng rk(seed, kKMeans);
  const int K = kTextonK;
  std::vector<double> C(size_t(K) * kSchmidNF);
  std::vector<double> d2(S, std::numeric_limits<double>::infinity());
  auto dist2 = [&](size_t i, const double* c) {
    double s = 0;
    for (int f = 0; f < kSchmidNF; f++) {
      const double d = pts[i * kSchmidNF + f] - c[f];
      s += d * d;
    }
    return s;
  };
  size_t first = std::min(S - 1, size_t(rk.uniform(0) * double(S)));
  std::copy(&pts[first * kSchmidNF], &pts[first * kSchmidNF] + kSchmidNF, &C[0]);
  for (int k = 1; k < K; k++) {
    double tot = 0;
    for (size_t i = 0; i < S; i++) {
      d2[i] = std::min(d2[i], dist2(i, &C[size_t(k - 1) * kSchmidNF]));
      tot += d2[i];
    }
    size_t pick = S - 1;
    if (tot > 0) {
      double target = rk.uniform(uint64_t(k)) * tot, run = 0;
      for (size_t i = 0; i < S; i++) {
        run += d2[i];
        if (run > target) {
          pick = i;
          break;
        }
      }
    } else {
      pick = size_t(rk.uniform(uint64_t(k)) * double(S)) % S;
    }
    std::copy(&pts[pick * kSchmidNF], &pts[pick * kSchmidNF] + kSchmidNF, &C[size_t(k) * kSchmidNF]);
  }
  std::vector<int> assign(S, 0);
  for (int it = 0; it < iters; it++) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(S); i++) {
      double best = std::numeric_limits<double>::infinity();
      int bk = 0;
      for (int k = 0; k < K; k++) {
        const double d = dist2(size_t(i), &C[size_t(k) * kSchmidNF]);
        if (d < best) {
          best = d;
          bk = k;
        }
      }
      assign[size_t(i)] = bk;
    }
    std::vector<double> sum(size_t(K) * kSchmidNF, 0.0);
    std::vector<size_t> cnt(K, 0);
    for (size_t i = 0; i < S; i++) {
      cnt[size_t(assign[i])]++;
      for (int f = 0; f < kSchmidNF; f++) sum[size_t(assign[i]) * kSchmidNF + f] += pts[i * kSchmidNF + f];
    }
    for (int k = 0; k < K; k++)
      if (cnt[size_t(k)])
        for (int f = 0; f < kSchmidNF; f++) C[size_t(k) * kSchmidNF + f] = sum[size_t(k) * kSchmidNF + f] / double(cnt[size_t(k)]);
  }
  // Codebook layout of the reference YAML: NF rows x K columns.
  for (int f = 0; f < kSchmidNF; f++)
    for (int k = 0; k < K; k++) codebook[f * K + k] = float(C[size_t(k) * kSchmidNF + f]);
}

}  // namespace gpb
