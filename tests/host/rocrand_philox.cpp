// rocrand_philox.cpp -- test harness: evaluates Philox4x32-10 with rocRAND's
// open-source implementation (rocrand_philox4x32_10.h) for (counter, key)
// pairs read from stdin, to cross-check the oracle's restatement.
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>

struct Probe : rocrand_device::philox4x32_10_engine {
    uint4 block(uint4 ctr, uint2 key) { return ten_rounds(ctr, key); }
};

int main() {
    Probe p;
    unsigned c0, c1, c2, c3, k0, k1;
    while (std::scanf("%u %u %u %u %u %u", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        uint4 ctr;
        ctr.x = c0; ctr.y = c1; ctr.z = c2; ctr.w = c3;
        uint2 key;
        key.x = k0; key.y = k1;
        const uint4 o = p.block(ctr, key);
        std::printf("%u %u %u %u\n", o.x, o.y, o.z, o.w);
    }
    return 0;
}
