// Host build of the block-count quotient (matter-tests_amd/csrc/mpm_math.hpp: block_div_magic, block_div) that
// the block kernels use for block coordinates. For every divisor d in [1, 65536] and a set of larger ones, checks
// the multiplier's bounds and block_div(n) == n / d at the quotient boundaries q d - 1, q d, q d + 1 for a spread
// of q, at 0 and at the largest numerator 2^31 - 1, and prints the number of failures.
#include <cstdio>
#include "../../matter-tests_amd/csrc/mpm_types.hpp"
#include "../../matter-tests_amd/csrc/mpm_math.hpp"
using namespace mpm;

static unsigned long long fails = 0, checks = 0;
static void one(uint32_t n, uint32_t d, uint32_t m, uint32_t l) {
    if (n >= 0x80000000u) return;
    ++checks;
    if (block_div(n, m, l) != n / d) {
        if (fails < 10) std::printf("d=%u n=%u got %u want %u\n", d, n, block_div(n, m, l), n / d);
        ++fails;
    }
}
static void divisor(uint32_t d) {
    uint32_t m, l;
    block_div_magic(d, m, l);
    const unsigned long long p = 1ull << (31 + l);
    const unsigned long long md = (unsigned long long)m * d;
    if (!(md >= p && md - p < d && d <= (1ull << l))) {
        if (fails < 10) std::printf("d=%u bad multiplier m=%u l=%u\n", d, m, l);
        ++fails;
    }
    one(0, d, m, l);
    one(0x7FFFFFFFu, d, m, l);
    const uint32_t qmax = 0x7FFFFFFFu / d;
    for (uint32_t k = 0; k < 64; ++k) {
        const uint32_t q = (uint32_t)((unsigned long long)qmax * k / 63);
        const unsigned long long n = (unsigned long long)q * d;
        if (n > 0) one((uint32_t)(n - 1), d, m, l);
        if (n < 0x80000000ull) one((uint32_t)n, d, m, l);
        if (n + 1 < 0x80000000ull) one((uint32_t)(n + 1), d, m, l);
    }
}
int main() {
    for (uint32_t d = 1; d <= 65536; ++d) divisor(d);
    for (uint32_t d = 65537; d < 0x80000000u && d != 0; d = d * 3 + 7) divisor(d);
    const uint32_t extra[] = {0x7FFFFFFFu, 0x40000000u, 0x40000001u, 0x3FFFFFFFu, 1000003u, 2147483629u};
    for (uint32_t d : extra) divisor(d);
    std::printf("checks %llu fails %llu\n", checks, fails);
    return fails ? 1 : 0;
}
