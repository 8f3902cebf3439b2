// Writes Philox4x32-10 known-answer vectors from rocRAND's host-callable
// engine (the readable header /opt/rocm/include/rocrand/rocrand_philox4x32_10.h).
// Build and run:  hipcc -O1 -o /tmp/philox_kat tests/golden/make_philox_vectors.cpp && /tmp/philox_kat > tests/golden/philox_kat.json
// engine(seed, subsequence, offset) starts at counter {offset/4 (64-bit), subsequence (64-bit)}
// with key = seed, so words 0..3 are Philox(counter, key).
#include <rocrand/rocrand_philox4x32_10.h>
#include <stdio.h>

int main()
{
    struct Case { unsigned long long seed, sub, ctr01; } cases[] = {
        {0x5EEDull, 0, 0}, {0x5EEDull, 0, 1}, {0x5EEDull, 0, 32767}, {0x5EEDull, 0, (63ull << 32) | 12345},
        {0xdeadbeefcafef00dull, 0, 7}, {0xdeadbeefcafef00dull, 3, 99}, {0ull, 0, 0}, {~0ull, 0, ~0ull >> 2}};
    printf("[\n");
    int n = sizeof(cases) / sizeof(cases[0]);
    for (int i = 0; i < n; ++i) {
        rocrand_device::philox4x32_10_engine e(cases[i].seed, cases[i].sub, cases[i].ctr01 * 4ull);
        unsigned int w[4];
        for (int k = 0; k < 4; ++k) w[k] = e();
        unsigned long long c = cases[i].ctr01, s = cases[i].sub;
        printf("  {\"key\": [%u, %u], \"ctr\": [%u, %u, %u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
               (unsigned)cases[i].seed, (unsigned)(cases[i].seed >> 32), (unsigned)c, (unsigned)(c >> 32),
               (unsigned)s, (unsigned)(s >> 32), w[0], w[1], w[2], w[3], i + 1 < n ? "," : "");
    }
    printf("]\n");
    return 0;
}
