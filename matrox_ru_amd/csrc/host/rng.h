// Deterministic random streams shared by the host builder and the synthetic
// input generators: splitmix64 for bits, Box-Muller for normals.  The same
// seed gives the same numbers on every platform and thread count, unlike
// std::normal_distribution, whose algorithm is implementation-defined.
#pragma once
#include <cmath>
#include <cstdint>

namespace mxh {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    // uniform in [0, 1) with 53 random bits
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    // uniform integer in [0, n)
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
};

// Seed of an independent stream derived from (seed, a, b).
inline uint64_t stream_seed(uint64_t seed, uint64_t a, uint64_t b = 0) {
    SplitMix64 g(seed ^ (a * 0xD1B54A32D192ED03ULL) ^ (b * 0x8CB92BA72F3D8DD7ULL));
    g.next();
    return g.next();
}

// Box-Muller: pairs of uniforms -> pairs of N(0,1); element i of a stream is a
// pure function of (seed, i), so blocks can be generated in parallel.
inline void normal_pair(uint64_t seed, uint64_t pair_index, double *z0, double *z1) {
    SplitMix64 g(stream_seed(seed, pair_index, 0x5A5A));
    double u1 = g.uniform();
    double u2 = g.uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    double r = std::sqrt(-2.0 * std::log(u1));
    double t = 6.283185307179586476925286766559 * u2;
    *z0 = r * std::cos(t);
    *z1 = r * std::sin(t);
}

}  // namespace mxh
