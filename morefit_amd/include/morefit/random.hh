/**
 * @file random.hh
 * Pseudo-random generators with the reference's interface (random.hh:18-120):
 * RandomGenerator::random/randomFloat/randomInt64/setSeed/burnin/exponential/poisson/gaus,
 * Splitmix64 and Xoshiro128pp (setSeed via Splitmix64 >> 32, getSeed, jump, float-rounded
 * random()).  The generators are the published algorithms of Blackman & Vigna
 * (prng.di.unimi.it); Poisson follows the reference's choice of Knuth's product method below
 * mean 25 and a Lorentzian-envelope rejection above (random.hh:42-83).
 */
#ifndef MOREFIT_RANDOM_HH
#define MOREFIT_RANDOM_HH

#include <cmath>
#include <cstdint>
#include <vector>
#include <limits>
#include <stdexcept>

namespace morefit {

class RandomGenerator {
 public:
  virtual ~RandomGenerator() = default;
  virtual double random() = 0;
  virtual float randomFloat() = 0;
  virtual uint64_t randomInt64() = 0;
  virtual void setSeed(uint64_t seed) = 0;
  void burnin(unsigned int n = 10000) {
    for (unsigned int i = 0; i < n; i++) random();
  }
  double exponential(double tau = 1.0) { return -tau * std::log(random()); }
  unsigned int poisson(double mean) {
    if (mean <= 0.0) return 0;
    if (mean < 25.0) {
      const double L = std::exp(-mean);
      double p = 1.0;
      int k = 0;
      do {
        p *= random();
        k++;
      } while (p > L);
      return (unsigned int)(k - 1);
    }
    if (mean < 1e9) {
      const double sq = std::sqrt(2.0 * mean), lm = std::log(mean), g = mean * lm - std::lgamma(mean + 1.0);
      double m, y, r;
      do {
        do {
          y = std::tan(M_PI * random());
          m = sq * y + mean;
        } while (m < 0.0);
        m = std::floor(m);
        r = 0.9 * (1.0 + y * y) * std::exp(m * lm - std::lgamma(m + 1.0) - g);
      } while (random() > r);
      return (unsigned int)m;
    }
    double r = mean + gaus(0.0, 1.0) * std::sqrt(mean) + 0.5;
    if (r > (double)std::numeric_limits<unsigned int>::max()) throw std::overflow_error("Poisson number exceeds integer limit");
    return (unsigned int)r;
  }
  double gaus(double mean = 0.0, double width = 1.0) {  // Marsaglia polar method
    double u, v, s;
    do {
      u = 2.0 * random() - 1.0;
      v = 2.0 * random() - 1.0;
      s = u * u + v * v;
    } while (s >= 1.0 || s == 0.0);
    return mean + width * u * std::sqrt(-2.0 * std::log(s) / s);
  }
};

class Splitmix64 : public RandomGenerator {
  uint64_t state_;

 public:
  explicit Splitmix64(uint64_t seed = 0) : state_(seed) {}
  void setSeed(uint64_t seed) override { state_ = seed; }
  uint64_t randomInt64() override {
    uint64_t z = (state_ += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  double random() override { return (double)(randomInt64() >> 11) * 0x1.0p-53; }
  float randomFloat() override { return (float)random(); }
};

class Xoshiro128pp : public RandomGenerator {
  uint32_t s_[4];
  static uint32_t rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

 public:
  explicit Xoshiro128pp(uint32_t seed[4]) { setSeed(seed); }
  explicit Xoshiro128pp(uint64_t seed = 3457) { setSeed(seed); }
  void setSeed(uint32_t seed[4]) {
    for (int i = 0; i < 4; i++) s_[i] = seed[i];
  }
  void setSeed(uint64_t seed) override {
    Splitmix64 sm(seed);
    for (int i = 0; i < 4; i++) s_[i] = (uint32_t)(sm.randomInt64() >> 32);
  }
  uint32_t getSeed(int idx) const { return s_[idx]; }
  uint32_t randomInt32() {
    const uint32_t r = rotl(s_[0] + s_[3], 7) + s_[0];
    const uint32_t t = s_[1] << 9;
    s_[2] ^= s_[0];
    s_[3] ^= s_[1];
    s_[1] ^= s_[2];
    s_[0] ^= s_[3];
    s_[2] ^= t;
    s_[3] = rotl(s_[3], 11);
    return r;
  }
  uint64_t randomInt64() override { return ((uint64_t)randomInt32() << 32) + randomInt32(); }
  double random() override { return (float)(randomInt32() * 2.3283064365386963e-10); }
  float randomFloat() override { return (float)(randomInt32() * 2.3283064365386963e-10); }
  // advance by 2^64 steps (polynomial of the published jump function)
  void jump() {
    static const uint32_t J[4] = {0x8764000b, 0xf542d2d3, 0x6fa035c3, 0x77f2db5b};
    uint32_t a = 0, b = 0, c = 0, d = 0;
    for (int i = 0; i < 4; i++)
      for (int bit = 0; bit < 32; bit++) {
        if (J[i] & (1u << bit)) {
          a ^= s_[0];
          b ^= s_[1];
          c ^= s_[2];
          d ^= s_[3];
        }
        randomInt32();
      }
    s_[0] = a;
    s_[1] = b;
    s_[2] = c;
    s_[3] = d;
  }
};

// Jump-ahead of xoshiro128++ streams in O(log k).  The state transition, and hence jump(), is
// linear over GF(2) on the 128 state bits, so the state of stream k (k jumps from a seeded
// state s) is J^k s.  xoshiro_jump_powers(L) returns J^(2^i) for i < L, each as 128 columns
// of 4 words: column b is the image of state bit b (word b / 32, bit b % 32).
inline void xoshiro_apply_matrix(const uint32_t* cols, uint32_t s[4]) {
  uint32_t a[4] = {0, 0, 0, 0};
  for (int b = 0; b < 128; b++)
    if ((s[b >> 5] >> (b & 31)) & 1u)
      for (int w = 0; w < 4; w++) a[w] ^= cols[4 * b + w];
  for (int w = 0; w < 4; w++) s[w] = a[w];
}
inline std::vector<uint32_t> xoshiro_jump_powers(int levels) {
  std::vector<uint32_t> m((size_t)levels * 512);
  for (int b = 0; b < 128; b++) {
    uint32_t e[4] = {0, 0, 0, 0};
    e[b >> 5] = 1u << (b & 31);
    Xoshiro128pp r(e);
    r.jump();
    for (int w = 0; w < 4; w++) m[4 * b + w] = r.getSeed(w);
  }
  for (int i = 1; i < levels; i++)
    for (int b = 0; b < 128; b++) {
      uint32_t c[4];
      for (int w = 0; w < 4; w++) c[w] = m[(size_t)(i - 1) * 512 + 4 * b + w];
      xoshiro_apply_matrix(&m[(size_t)(i - 1) * 512], c);
      for (int w = 0; w < 4; w++) m[(size_t)i * 512 + 4 * b + w] = c[w];
    }
  return m;
}
// state of stream k of `base` (k < 2^levels)
inline void xoshiro_stream_state(const std::vector<uint32_t>& powers, const uint32_t base[4], unsigned long long k, uint32_t out[4]) {
  for (int w = 0; w < 4; w++) out[w] = base[w];
  for (int i = 0; k >> i; i++)
    if ((k >> i) & 1ull) xoshiro_apply_matrix(&powers[(size_t)i * 512], out);
}

}  // namespace morefit
#endif
