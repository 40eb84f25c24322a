// Prints word 0 of rocRAND's Philox-4x32-10 engine for each "seed tick id"
// line on stdin: engine(seed, subsequence = id, offset = 4 * tick).next().
// Test infrastructure: a second independent source oracle for the SIR draw.
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
int main() {
    unsigned long long seed, tick, id;
    while (std::scanf("%llu %llu %llu", &seed, &tick, &id) == 3) {
        rocrand_device::philox4x32_10_engine eng(seed, id, 4ull * tick);
        unsigned int w0 = eng.next();
        unsigned int w1 = eng.next();
        std::printf("%u %u\n", w0, w1);
    }
    return 0;
}
