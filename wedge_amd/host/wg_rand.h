/*
 * wg_rand.h -- counter-based splitmix64 used by every synthetic generator
 * (host C here; the device generator restates the same function in HIP).
 */
#ifndef WG_RAND_H
#define WG_RAND_H

#include <stdint.h>

/* Output number `counter` of the splitmix64 stream seeded with `key`
 * (Steele, Lea, Flood, "Fast splittable pseudorandom number generators", 2014). */
static inline uint64_t wg_rand64_inline(uint64_t key, uint64_t counter)
{
    uint64_t z = key + (counter + 1ull) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* One R-MAT draw: per level l (bit scale-1-l), a 32-bit uniform u picks the
 * quadrant: u < t_a -> (0,0); u < t_ab -> (src 0, dst 1); u < t_abc -> (1, 0);
 * else (1, 1).  Levels 2j and 2j+1 use the low and high halves of
 * wg_rand64(key, (i << 5) | j). */
static inline void wg_rmat_draw(uint64_t key, int scale, uint32_t t_a, uint32_t t_ab, uint32_t t_abc,
                                uint64_t i, uint32_t* src, uint32_t* dst)
{
    uint32_t s = 0, d = 0;
    for (int j = 0; 2 * j < scale; ++j) {
        const uint64_t h = wg_rand64_inline(key, (i << 5) | (uint64_t)j);
        for (int half = 0; half < 2; ++half) {
            const int l = 2 * j + half;
            if (l >= scale)
                break;
            const uint32_t u = half ? (uint32_t)(h >> 32) : (uint32_t)h;
            const uint32_t sb = (uint32_t)(u >= t_ab);
            const uint32_t db = (uint32_t)((u >= t_a) & (u < t_ab)) | (uint32_t)(u >= t_abc);
            const int bit = scale - 1 - l;
            s |= sb << bit;
            d |= db << bit;
        }
    }
    *src = s;
    *dst = d;
}

#endif /* WG_RAND_H */
