// Small host helpers shared by the host-side translation units.
#pragma once
#include <cstdint>

namespace dawn {

// SplitMix64 finaliser (Steele, Lea, Flood 2014).
inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Counter-based random: a pure function of (seed, stream, index), so any
// thread may draw any index and results do not depend on scheduling.
inline uint64_t rng_u64(uint64_t seed, uint64_t stream, uint64_t index) {
  return splitmix64(splitmix64(seed * 0xD1B54A32D192ED03ull + stream) ^
                    (index * 0x9E3779B97F4A7C15ull));
}

}  // namespace dawn
