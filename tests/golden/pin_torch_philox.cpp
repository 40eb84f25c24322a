// Prints word 0 of torch's host Philox-4x32-10 engine for each "seed tick id"
// line on stdin: at::Philox4_32(seed, subsequence = id, offset = tick)().
// Test infrastructure: an independent source oracle for the SIR draw.
#include <ATen/core/PhiloxRNGEngine.h>
#include <cstdio>
#include <cinttypes>
int main() {
    unsigned long long seed, tick, id;
    while (std::scanf("%llu %llu %llu", &seed, &tick, &id) == 3) {
        at::Philox4_32 eng(seed, id, tick);
        uint32_t w0 = eng();
        uint32_t w1 = eng();
        std::printf("%u %u\n", w0, w1);
    }
    return 0;
}
