// spk_sampler.h -- counter-based sample streams keyed by (pixel, sample, dimension).
//
// spica draws from one MT19937 per thread (sources/samplers/independent.cc,
// sources/core/random.cc) reseeded from time(0) (sources/core/integrator.cc:51,71),
// so a path cannot be replayed.  Here every draw is a pure function of
// (seed, pixel, sample, dim): any path can be recomputed alone, which the
// wavefront stages rely on (a stage never carries RNG state).
//
//   independent: PCG4D hash (Jarzynski & Olano, JCGT 2020) of the 4-tuple.
//   sobol:       2-D Sobol (dimensions 0 and 1) per dimension pair, with the
//                sample index shuffled and each coordinate Owen-scrambled by the
//                hash-based nested uniform scramble of Burley (JCGT 2020).
//
// Dimension layout per path (the order spica consumes its stream,
// integrator.cc:84-85, mis.cc:27-30, path.cc:87,119): 0-1 film, 2-3 lens, then per
// bounce b at base 4 + 8b: +0 light pick, +1..2 light sample, +3..4 NEE BSDF
// sample, +5..6 continuation, +7 Russian roulette.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spk {

enum SamplerKind : int { SAMPLER_INDEPENDENT = 0, SAMPLER_SOBOL = 1 };

__host__ __device__ inline void pcg4d(uint32_t& x, uint32_t& y, uint32_t& z, uint32_t& w) {
    x = x * 1664525u + 1013904223u;
    y = y * 1664525u + 1013904223u;
    z = z * 1664525u + 1013904223u;
    w = w * 1664525u + 1013904223u;
    x += y * w;
    y += z * x;
    z += x * y;
    w += y * z;
    x ^= x >> 16;
    y ^= y >> 16;
    z ^= z >> 16;
    w ^= w >> 16;
    x += y * w;
    y += z * x;
    z += x * y;
    w += y * z;
}

__host__ __device__ inline uint32_t bitrev32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_bitreverse32(x);
#else
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0f0f0f0fu) | ((x & 0x0f0f0f0fu) << 4);
    x = ((x >> 8) & 0x00ff00ffu) | ((x & 0x00ff00ffu) << 8);
    return (x >> 16) | (x << 16);
#endif
}

// Laine-Karras style permutation with Burley's constants, applied to reversed
// bits: flips each bit as a function of the more significant bits only.
__host__ __device__ inline uint32_t owen_scramble(uint32_t x, uint32_t seed) {
    x = bitrev32(x);
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return bitrev32(x);
}

// Sobol dimension 1 (primitive polynomial x + 1): direction v_0 = 2^31,
// v_k = v_{k-1} ^ (v_{k-1} >> 1), so v_k has bit 31 - j set iff C(k, j) is odd, i.e.
// iff j is a bit-subset of k (Lucas).  Bit 31 - j of the point is then the parity of the
// index bits k that are supersets of j: a superset-parity (zeta) transform over the five
// bits of k, one shift-xor per bit, followed by a bit reversal.  Equal to the loop
// x ^= v_k over the set bits k of idx, without its up to 32 data-dependent steps.
__host__ __device__ inline uint32_t sobol_dim1(uint32_t idx) {
    uint32_t a = idx;
    a ^= (a >> 1) & 0x55555555u;
    a ^= (a >> 2) & 0x33333333u;
    a ^= (a >> 4) & 0x0f0f0f0fu;
    a ^= (a >> 8) & 0x00ff00ffu;
    a ^= (a >> 16) & 0x0000ffffu;
    return bitrev32(a);
}

__host__ __device__ inline void dim_pair(uint32_t dim, uint32_t& pair, uint32_t& comp) {
    if (dim < 4) {
        pair = dim >> 1;
        comp = dim & 1u;
        return;
    }
    uint32_t k = (dim - 4) & 7u, b = (dim - 4) >> 3, base = 2 + 4 * b;
    if (k == 0) { pair = base; comp = 0; }
    else if (k == 7) { pair = base; comp = 1; }
    else if (k <= 2) { pair = base + 1; comp = k - 1; }
    else if (k <= 4) { pair = base + 2; comp = k - 3; }
    else { pair = base + 3; comp = k - 5; }
}

__host__ __device__ inline uint32_t sample_u32(int kind, uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim) {
    if (kind == SAMPLER_SOBOL) {
        uint32_t pair, comp;
        dim_pair(dim, pair, comp);
        uint32_t a = pixel, b = pair, c = seed, d = 0x9e3779b9u;
        pcg4d(a, b, c, d);
        uint32_t idx = owen_scramble(sample, a);
        uint32_t x = comp == 0 ? bitrev32(idx) : sobol_dim1(idx);
        return owen_scramble(x, comp == 0 ? b : c);
    }
    uint32_t a = pixel, b = sample, c = dim, d = seed;
    pcg4d(a, b, c, d);
    return a;
}

// 24 significant bits: exactly representable in float and double.
__host__ __device__ inline float u32_to_unit(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// Both coordinates of one Sobol pair share the index shuffle, so a 2-D draw
// costs one hash; for the independent sampler it is two hashes.
__device__ inline float2 sample_2d(int kind, uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim0) {
    if (kind == SAMPLER_SOBOL) {
        uint32_t pair, comp;
        dim_pair(dim0, pair, comp);
        uint32_t a = pixel, b = pair, c = seed, d = 0x9e3779b9u;
        pcg4d(a, b, c, d);
        uint32_t idx = owen_scramble(sample, a);
        uint32_t x0 = owen_scramble(bitrev32(idx), b);
        uint32_t x1 = owen_scramble(sobol_dim1(idx), c);
        return make_float2(u32_to_unit(x0), u32_to_unit(x1));
    }
    return make_float2(u32_to_unit(sample_u32(kind, seed, pixel, sample, dim0)),
                       u32_to_unit(sample_u32(kind, seed, pixel, sample, dim0 + 1)));
}

__device__ inline float sample_1d(int kind, uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim) {
    return u32_to_unit(sample_u32(kind, seed, pixel, sample, dim));
}

}  // namespace spk
