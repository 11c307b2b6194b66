// rng.h -- the library's random stream.
//
// splitmix64 (Steele, Lea and Flood, OOPSLA 2014) for the stream, Lemire's
// multiply-high map for bounded draws, and the inside-out shuffle that Go's
// rand.Perm uses (pqivf/index.go:805 draws its k-means seeds from one).
// Each (seed, train epoch, role) pair gets its own stream: role 0 is the
// coarse quantizer, role 1+i subquantizer i, kRoleSample the training
// samples. The reference's own stream (Go math/rand seeded by HANN_SEED)
// cannot be reproduced without Go's seeding table.
#pragma once

#include <cstdint>
#include <vector>

namespace hann {

constexpr uint64_t kRoleCoarseSample = 1000003;
constexpr uint64_t kRolePQSample = 1000033;

inline uint64_t hash64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

inline uint64_t rng_next(uint64_t &state) {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t rng_bounded(uint64_t &state, uint64_t n) {
    return (uint64_t)(((unsigned __int128)rng_next(state) * n) >> 64);
}

inline void rng_perm(uint64_t &state, int64_t n, std::vector<int64_t> &out) {
    out.assign((size_t)n, 0);
    for (int64_t i = 0; i < n; i++) {
        int64_t j = (int64_t)rng_bounded(state, (uint64_t)(i + 1));
        out[i] = out[j];
        out[j] = i;
    }
}

inline uint64_t stream_seed(uint64_t seed, uint64_t epoch, uint64_t role) {
    return hash64(hash64(seed ^ hash64(epoch)) ^ hash64(role + 0x632BE59BD9B4E019ull));
}

}  // namespace hann
