/* fa_rng.h — the counter-based random numbers of the device GA.
 *
 * The reference draws from a thread-local std::mt19937 seeded by
 * std::random_device (LIB/gaboom.cpp:2942-2946), which no two runs share.
 * Here every draw is a pure function of (seed, generation, stream, index):
 * SplitMix64 (Steele, Lea & Flood, OOPSLA 2014) applied to the counter, so a
 * generation is reproducible bit for bit and parallel draws need no state.
 * The same definition serves the kernels and the CPU oracle.
 */
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define FA_RNG_FN __host__ __device__ inline
#else
#define FA_RNG_FN static inline
#endif

#define FA_MAX_RANDOM_VALUE 2147483647 /* LIB/gaboom.h:36 */

FA_RNG_FN uint64_t fa_splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* 64 random bits of draw `k` of `stream` in generation `gen` */
FA_RNG_FN uint64_t fa_rng_bits(uint64_t seed, uint64_t gen, uint64_t stream, uint64_t k)
{
    const uint64_t a = fa_splitmix64(seed ^ fa_splitmix64(gen * 0xD1B54A32D192ED03ull + 1ull));
    return fa_splitmix64(a ^ fa_splitmix64(stream * 0x8CB92BA72F3D8DD7ull + k));
}

/* uniform double in [0, 1) with 53 random bits (the role of RandomDouble()) */
FA_RNG_FN double fa_rng_u01(uint64_t seed, uint64_t gen, uint64_t stream, uint64_t k)
{
    return (double)(fa_rng_bits(seed, gen, stream, k) >> 11) * (1.0 / 9007199254740992.0);
}

/* uniform integer gene in [1, FA_MAX_RANDOM_VALUE] (the role of dice()) */
FA_RNG_FN int32_t fa_rng_gene(uint64_t seed, uint64_t gen, uint64_t stream, uint64_t k)
{
    return (int32_t)(1 + fa_rng_bits(seed, gen, stream, k) % (uint64_t)FA_MAX_RANDOM_VALUE);
}

/* draw layout of one offspring pair (stream = pair index) */
#define FA_DRAW_P1 0
#define FA_DRAW_P2 1
#define FA_DRAW_CROSS 2
#define FA_DRAW_GEN_A 3
#define FA_DRAW_GEN_B 4
#define FA_DRAW_PNT_A 5
#define FA_DRAW_PNT_B 6
#define FA_DRAW_MUT 8 /* child c, gene j, bit i: FA_DRAW_MUT + (c * genes + j) * 32 + i */
