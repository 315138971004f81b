// mt19937.cpp -- the Mersenne Twister MT19937 as published by Matsumoto and
// Nishimura (mt19937ar, 2002): state size 624, shift 397, matrix 0x9908b0df,
// tempering masks 0x9d2c5680 / 0xefc60000, seeding by init_by_array.  The
// constants are the algorithm's definition, so they match the reference's
// src/mt19937ar.c:60-180 and measurement outcomes stay bit-exact per seed.
#include "mt19937.hpp"

#include <vector>

#include "runtime.hpp"

namespace dgq {
namespace {
constexpr int kN = Mt19937::kN, kM = 397;
}

void Mt19937::init_genrand(uint32_t s) {
    mt[0] = s;
    for (mti = 1; mti < kN; mti++) mt[mti] = 1812433253u * (mt[mti - 1] ^ (mt[mti - 1] >> 30)) + (uint32_t)mti;
}

void Mt19937::init_by_array(const uint32_t *key, int len) {
    init_genrand(19650218u);
    int i = 1, j = 0;
    for (int k = (kN > len ? kN : len); k; k--) {
        mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1664525u)) + key[j] + (uint32_t)j;
        if (++i >= kN) { mt[0] = mt[kN - 1]; i = 1; }
        if (++j >= len) j = 0;
    }
    for (int k = kN - 1; k; k--) {
        mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
        if (++i >= kN) { mt[0] = mt[kN - 1]; i = 1; }
    }
    mt[0] = 0x80000000u;
}

uint32_t Mt19937::next_u32() {
    if (mti >= kN) {
        if (mti == kN + 1) init_genrand(5489u);
        auto twist = [](uint32_t a, uint32_t b, uint32_t c) {
            const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
            return c ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        };
        int k = 0;
        for (; k < kN - kM; k++) mt[k] = twist(mt[k], mt[k + 1], mt[k + kM]);
        for (; k < kN - 1; k++) mt[k] = twist(mt[k], mt[k + 1], mt[k + kM - kN]);
        mt[kN - 1] = twist(mt[kN - 1], mt[0], mt[kM - 1]);
        mti = 0;
    }
    uint32_t y = mt[mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

double Mt19937::real1() { return next_u32() * (1.0 / 4294967295.0); }

double Mt19937::res53() {
    const uint32_t a = next_u32() >> 5, b = next_u32() >> 6;
    return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
}

// the process-global generator of the QuEST API (src/mt19937ar.c:57-58)
static Mt19937 g_api_rng;

void mt_init_by_array(const unsigned long *key, int len) {
    // an empty key (the reference would read key[0] regardless) seeds as {0}
    std::vector<uint32_t> k((size_t)(len > 0 ? len : 1), 0u);
    for (int i = 0; i < len; i++) k[i] = (uint32_t)key[i];
    g_api_rng.init_by_array(k.data(), (int)k.size());
}

double mt_genrand_real1() { return g_api_rng.real1(); }

}  // namespace dgq
