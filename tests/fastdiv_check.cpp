// Exhaustive-in-part check of FastDiv (mr-lbm_amd/csrc/fastdiv.h) against
// the hardware division: every divisor 1..4096 and a set of large ones,
// with dividends at the edges (0, d-1, d, k d +- 1, 2^32 - 1) and from a
// fixed-seed generator. Prints "ok" on success.
#include <cstdio>
#include <cstdint>
#include <vector>

#include "../mr-lbm_amd/csrc/fastdiv.h"

int main() {
    std::vector<uint32_t> ds;
    for (uint32_t d = 1; d <= 4096; d++) ds.push_back(d);
    for (uint32_t d : {65535u, 65536u, 65537u, 1000003u, 16777215u, 16777216u, 2147483647u, 2147483648u,
                       2147483649u, 4294967295u})
        ds.push_back(d);
    uint64_t x = 0x9E3779B97F4A7C15ull;
    auto rnd = [&]() {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return (uint32_t)(x >> 16);
    };
    long checked = 0;
    for (uint32_t d : ds) {
        const mrlbm::FastDiv f = mrlbm::make_fastdiv(d);
        std::vector<uint32_t> ns = {0u, d - 1, d, d + 1, 2 * d - 1, 2 * d, 4294967295u, 4294967294u,
                                    4294967295u / d * d, 4294967295u / d * d - 1};
        for (int i = 0; i < 2000; i++) ns.push_back(rnd());
        for (int i = 0; i < 200; i++) ns.push_back(rnd() % (1u << 22));
        for (uint32_t n : ns) {
            if (f.div(n) != n / d) {
                std::printf("FAIL d=%u n=%u got %u want %u\n", d, n, f.div(n), n / d);
                return 1;
            }
            checked++;
        }
    }
    std::printf("%ld divisions ok\n", checked);
    return 0;
}
