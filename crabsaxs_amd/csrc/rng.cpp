// The resampling stream of EnsembleFitter::fit's bootstrap (ensemble.rs:
// 184-216): ChaCha8Rng::seed_from_u64(seed), then rng.gen_range(0..n) for
// every point of every replicate.  Restated from the published algorithms of
// its pinned dependencies (rand_core 0.6 seed_from_u64, rand_chacha 0.3
// ChaCha8 with a 64-bit block counter and 64-bit stream, rand 0.8 uniform
// integer sampling by widening multiply with rejection); none of them is
// vendored by the reference.
#include <cstdint>
#include <cstring>
#include <vector>

#include "capi_internal.hpp"

namespace csx {
namespace {

inline uint32_t rotl32(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }

inline void quarter(uint32_t* x, int a, int b, int c, int d) {
  x[a] += x[b]; x[d] = rotl32(x[d] ^ x[a], 16);
  x[c] += x[d]; x[b] = rotl32(x[b] ^ x[c], 12);
  x[a] += x[b]; x[d] = rotl32(x[d] ^ x[a], 8);
  x[c] += x[d]; x[b] = rotl32(x[b] ^ x[c], 7);
}

// One ChaCha block (Bernstein 2008; RFC 7539 section 2.3): 16 output words
// for key words k[8], 64-bit block counter and 64-bit stream (state words
// 12-15), `rounds` rounds (a multiple of 2).
void chacha_block(const uint32_t k[8], uint64_t counter, uint64_t stream, int rounds,
                  uint32_t out[16]) {
  uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                    k[0], k[1], k[2], k[3], k[4], k[5], k[6], k[7],
                    static_cast<uint32_t>(counter), static_cast<uint32_t>(counter >> 32),
                    static_cast<uint32_t>(stream), static_cast<uint32_t>(stream >> 32)};
  uint32_t x[16];
  std::memcpy(x, s, sizeof x);
  for (int r = 0; r < rounds; r += 2) {
    quarter(x, 0, 4, 8, 12);
    quarter(x, 1, 5, 9, 13);
    quarter(x, 2, 6, 10, 14);
    quarter(x, 3, 7, 11, 15);
    quarter(x, 0, 5, 10, 15);
    quarter(x, 1, 6, 11, 12);
    quarter(x, 2, 7, 8, 13);
    quarter(x, 3, 4, 9, 14);
  }
  for (int i = 0; i < 16; ++i) out[i] = x[i] + s[i];
}

// PCG32 (O'Neill 2014): the 64-bit LCG step state = state * 6364136223846793005
// + increment and the XSH-RR output, here of the state before the step (the
// reference generator's convention) or after it (seed_from_u64's).
inline uint32_t pcg32_next(uint64_t& state, uint64_t increment, bool output_new) {
  const uint64_t old = state;
  state = old * 6364136223846793005ull + increment;
  const uint64_t x = output_new ? state : old;
  const uint32_t xorshifted = static_cast<uint32_t>(((x >> 18) ^ x) >> 27);
  const uint32_t rot = static_cast<uint32_t>(x >> 59);
  return (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
}

// ChaCha8Rng seeded like SeedableRng::seed_from_u64: the 32-byte key is eight
// successive PCG32 (XSH-RR) outputs of the advanced state, little-endian,
// with the increment 11634580027462260723.
class ChaCha8 {
 public:
  explicit ChaCha8(uint64_t state) {
    for (int i = 0; i < 8; ++i) key_[i] = pcg32_next(state, 11634580027462260723ull, true);
  }
  // next_u64: two consecutive keystream words, the first as the low half.
  uint64_t next_u64() {
    if (index_ + 2 > 16) refill();
    const uint64_t lo = buf_[index_], hi = buf_[index_ + 1];
    index_ += 2;
    return lo | (hi << 32);
  }
  // gen_range(0..n), n >= 1: widening multiply, reject the low word above the
  // zone (range << leading_zeros(range)) - 1.
  uint64_t below(uint64_t n) {
    int lz = 0;
    while (lz < 64 && !((n << lz) & 0x8000000000000000ull)) ++lz;
    const uint64_t zone = (n << lz) - 1;
    for (;;) {
      const unsigned __int128 m = static_cast<unsigned __int128>(next_u64()) * n;
      const uint64_t hi = static_cast<uint64_t>(m >> 64), lo = static_cast<uint64_t>(m);
      if (lo <= zone) return hi;
    }
  }

 private:
  void refill() {
    chacha_block(key_, counter_++, 0, 8, buf_);
    index_ = 0;
  }
  uint32_t key_[8];
  uint32_t buf_[16];
  uint64_t counter_ = 0;
  int index_ = 16;
};

}  // namespace
}  // namespace csx

using namespace csx;

extern "C" {

csx_status csx_chacha_block(const uint32_t* key, uint64_t counter, uint64_t stream, int rounds,
                            uint32_t* out) {
  if (!key || !out || rounds <= 0 || rounds % 2 != 0)
    return fail(CSX_INVALID_INPUT, "key, output and an even positive round count are required");
  chacha_block(key, counter, stream, rounds, out);
  return CSX_OK;
}

csx_status csx_pcg32(uint64_t state, uint64_t increment, size_t n, int output_new_state,
                     uint32_t* out) {
  if (n > 0 && !out) return fail(CSX_INVALID_INPUT, "an output buffer is required");
  for (size_t k = 0; k < n; ++k) out[k] = pcg32_next(state, increment, output_new_state != 0);
  return CSX_OK;
}

csx_status csx_bootstrap_indices(uint64_t seed, size_t n, size_t replicates, uint64_t* out) {
  CSX_API_RANGE();
  if (n == 0 || (replicates > 0 && !out))
    return fail(CSX_INVALID_INPUT, "a non-empty curve and an output buffer are required");
  ChaCha8 rng(seed);
  for (size_t r = 0; r < replicates; ++r)
    for (size_t i = 0; i < n; ++i) out[r * n + i] = rng.below(n);
  return CSX_OK;
}

}  // extern "C"
