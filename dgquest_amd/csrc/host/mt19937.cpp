// mt19937.cpp -- the Mersenne Twister MT19937 as published by Matsumoto and
// Nishimura (mt19937ar, 2002): state size 624, shift 397, matrix 0x9908b0df,
// tempering masks 0x9d2c5680 / 0xefc60000, seeding by init_by_array.  The
// constants are the algorithm's definition, so they match the reference's
// src/mt19937ar.c:60-180 and measurement outcomes stay bit-exact per seed.
#include <cstdint>

#include "runtime.hpp"

namespace dgq {
namespace {
constexpr int kN = 624, kM = 397;
uint32_t g_mt[kN];
int g_mti = kN + 1;

void init_genrand(uint32_t s) {
    g_mt[0] = s;
    for (g_mti = 1; g_mti < kN; g_mti++)
        g_mt[g_mti] = 1812433253u * (g_mt[g_mti - 1] ^ (g_mt[g_mti - 1] >> 30)) + (uint32_t)g_mti;
}

uint32_t next_u32() {
    if (g_mti >= kN) {
        if (g_mti == kN + 1) init_genrand(5489u);
        auto twist = [](uint32_t a, uint32_t b, uint32_t c) {
            const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
            return c ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        };
        int k = 0;
        for (; k < kN - kM; k++) g_mt[k] = twist(g_mt[k], g_mt[k + 1], g_mt[k + kM]);
        for (; k < kN - 1; k++) g_mt[k] = twist(g_mt[k], g_mt[k + 1], g_mt[k + kM - kN]);
        g_mt[kN - 1] = twist(g_mt[kN - 1], g_mt[0], g_mt[kM - 1]);
        g_mti = 0;
    }
    uint32_t y = g_mt[g_mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}
}  // namespace

void mt_init_by_array(const unsigned long *key, int len) {
    init_genrand(19650218u);
    int i = 1, j = 0;
    for (int k = (kN > len ? kN : len); k; k--) {
        g_mt[i] = (g_mt[i] ^ ((g_mt[i - 1] ^ (g_mt[i - 1] >> 30)) * 1664525u)) + (uint32_t)key[j] + (uint32_t)j;
        if (++i >= kN) { g_mt[0] = g_mt[kN - 1]; i = 1; }
        if (++j >= len) j = 0;
    }
    for (int k = kN - 1; k; k--) {
        g_mt[i] = (g_mt[i] ^ ((g_mt[i - 1] ^ (g_mt[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
        if (++i >= kN) { g_mt[0] = g_mt[kN - 1]; i = 1; }
    }
    g_mt[0] = 0x80000000u;
}

double mt_genrand_real1() { return next_u32() * (1.0 / 4294967295.0); }

}  // namespace dgq
