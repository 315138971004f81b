// mt19937.hpp -- one MT19937 generator (mt19937ar algorithm) per object, so
// every served request draws its samples from its own seeded stream.
#pragma once
#include <cstdint>

namespace dgq {

struct Mt19937 {
    static constexpr int kN = 624;
    uint32_t mt[kN];
    int mti = kN + 1;

    void init_genrand(uint32_t s);
    void init_by_array(const uint32_t *key, int len);
    uint32_t next_u32();   // genrand_int32
    double real1();        // genrand_real1, [0, 1]
    double res53();        // genrand_res53, [0, 1) with 53-bit resolution
};

}  // namespace dgq
