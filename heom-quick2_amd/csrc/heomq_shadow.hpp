// The TFQMR shadow vector, shared by the vector kernels and the operator
// launches that carry the iteration's reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace heomq {

// Shadow vector of TFQMR, generated where it is used instead of stored: a
// +-1 sign per real element from a splitmix64 hash of its index (the same
// bits for any grid and any call).
__device__ inline double shadow(int64_t k, uint64_t seed) {
  uint64_t z = 0x5EED1234ull + 0xD1B54A32D192ED03ull * seed + 0x9E3779B97F4A7C15ull * (uint64_t)(k + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (z >> 63) ? -1.0 : 1.0;
}

}  // namespace heomq
