// Host check of fsg::div_rn (csrc/fdiv.hpp) against IEEE float32 division.
// Built with g++ -ffp-contract=off; every case compares the bits of
// div_rn(a, b, 1.0f / b) with a / b and counts mismatches.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>

#include "../../fujishadergpu_amd/csrc/fdiv.hpp"

static uint32_t bits(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  return u;
}
static float from_bits(uint32_t u) {
  float x;
  std::memcpy(&x, &u, 4);
  return x;
}

static uint64_t s_state = 0x9E3779B97F4A7C15ull;
static uint64_t next_u64() {  // splitmix64
  uint64_t z = (s_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static long long n_cases = 0, n_bad = 0;

static void check(float a, float b) {
  volatile float va = a, vb = b;
  const float want = va / vb;
  const float y = 1.0f / vb;
  const float got = fsg::div_rn(a, b, y);
  n_cases++;
  const bool same = bits(got) == bits(want) || (std::isnan(got) && std::isnan(want));
  // the unguarded form inside its range: the same value (zeros of either sign)
  const float m = std::fabs(a);
  bool same_u = true;
  if (m == 0.0f || (m >= 0x1p-75f && m <= 0x1p75f)) {
    const float gu = fsg::div_rn_unguarded(a, b, y);
    same_u = m == 0.0f ? gu == 0.0f : bits(gu) == bits(want);
  }
  if (!same || !same_u) {
    if (n_bad < 10)
      std::printf("mismatch a=%a b=%a got=%a want=%a unguarded %s\n", (double)a, (double)b,
                  (double)got, (double)want, same_u ? "ok" : "differs");
    n_bad++;
  }
}

// a float with the given exponent, a random significand and a random sign
static float rand_float(int e) {
  const uint32_t mant = (uint32_t)(next_u64() & 0x7FFFFF);
  const uint32_t sign = (uint32_t)(next_u64() & 1) << 31;
  return from_bits(sign | ((uint32_t)(e + 127) << 23) | mant);
}

int main(int argc, char** argv) {
  const long long n_random = argc > 1 ? std::atoll(argv[1]) : 20000000;
  const int n_exhaustive_b = argc > 2 ? std::atoi(argv[2]) : 3;
  // 1. random a, b over the whole fast-path range
  for (long long k = 0; k < n_random; k++) {
    const int ea = (int)(next_u64() % 151) - 75;  // [-75, 75]
    const int eb = (int)(next_u64() % 80) - 40;   // [-40, 39]
    check(rand_float(ea), rand_float(eb));
  }
  // 2. differences of elevations over 2h for pixel sizes in use
  const float hs[] = {0.5f, 1.0f, 2.0f, 5.0f, 10.0f, 12.5f, 30.0f, 90.0f, 0.1f, 1.0f / 3.0f,
                      0.000277777f, 0.00083333f, 27.7f, 7.3f};
  for (float h : hs) {
    for (const float sgn : {1.0f, -1.0f}) {
      const float b = sgn * 2.0f * h;
      for (int k = 0; k < 2000000; k++) {
        const float z0 = (float)((double)(next_u64() >> 11) * 0x1p-53 * 9500.0 - 500.0);
        const float z1 = z0 + (float)((double)(next_u64() >> 11) * 0x1p-53 * 200.0 - 100.0);
        check(z1 - z0, b);
      }
    }
  }
  // 3. every significand of a at a few exponents, for random b
  for (int kb = 0; kb < n_exhaustive_b; kb++) {
    const float b = rand_float((int)(next_u64() % 20) - 10);
    for (const int ea : {-20, 0, 13}) {
      for (uint32_t m = 0; m < (1u << 23); m++)
        check(from_bits(((uint32_t)(ea + 127) << 23) | m), b);
    }
  }
  // b with an all-ones significand (the reciprocal's hard case)
  for (const int eb : {-3, 0, 4}) {
    const float b = from_bits(((uint32_t)(eb + 127) << 23) | 0x7FFFFF);
    for (uint32_t m = 0; m < (1u << 23); m += 7) check(from_bits((127u << 23) | m), b);
  }
  // 4. zeros, subnormals, limits, inf and NaN take the guarded path
  const float specials[] = {0.0f, -0.0f, 1e-40f, -1e-40f, 0x1p-76f, 0x1p76f, 3e38f, -3e38f,
                            INFINITY, -INFINITY, NAN, 0x1p-75f, 0x1p75f, 1.0f};
  for (float a : specials)
    for (float b : {20.0f, -20.0f, 0x1p-40f, 0x1p40f, 0.000555f}) check(a, b);
  std::printf("cases %lld mismatches %lld\n", n_cases, n_bad);
  return n_bad == 0 ? 0 : 1;
}
