// Exhaustive CPU check of the exact-arithmetic shortcuts in birp_cpp_amd/csrc/exact_math.h.
//   scale <lo_bits> <hi_bits>     every float bit pattern in [lo, hi): birp_scale_sample == chain
//   div <b>                       every non-negative finite float a: birp_div_rn(a,b,RN(1/b)) == a/b
//   divrange <b>                  the same for every float a in [2^-64, 2^64) (the kernels' range)
//   div64range <b>                birp_div_rn64(a, RN64(1/b)) == a/b for every float a in [2^-64, 2^64)
// Prints "mismatches N checked M" and exits 0 when N == 0. Threads split the range.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>
#include <atomic>
#include "../../birp_cpp_amd/csrc/exact_math.h"

static float f_of(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t u_of(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 4;
    std::atomic<unsigned long long> bad{0}, checked{0};
    std::vector<std::thread> th;
    if (!strcmp(argv[1], "scale")) {
        uint64_t lo = strtoull(argv[2], 0, 0), hi = strtoull(argv[3], 0, 0);
        for (unsigned t = 0; t < nt; t++)
            th.emplace_back([&, t] {
                unsigned long long b = 0, c = 0;
                for (uint64_t u = lo + t; u < hi; u += nt) {
                    float v = f_of((uint32_t)u);
                    if (v != v) continue;
                    float x = birp_scale_sample(v), y = birp_scale_chain(v);
                    if (u_of(x) != u_of(y)) b++;
                    c++;
                }
                bad += b; checked += c;
            });
    } else if (!strcmp(argv[1], "div") || !strcmp(argv[1], "divrange") || !strcmp(argv[1], "div64range")) {
        float bb = strtof(argv[2], 0);
        volatile float one = 1.0f;
        volatile double done = 1.0;
        float rb = one / bb;
        double rb64 = done / (double)bb;
        const bool ranged = strcmp(argv[1], "div") != 0;
        const bool d64 = !strcmp(argv[1], "div64range");
        const uint64_t lo = ranged ? (uint64_t)(127 - 64) << 23 : 0, hi = ranged ? (uint64_t)(127 + 64) << 23 : 0x7F800000ull;
        for (unsigned t = 0; t < nt; t++)
            th.emplace_back([&, t] {
                unsigned long long b = 0, c = 0;
                for (uint64_t u = lo + t; u < hi; u += nt) {
                    float a = f_of((uint32_t)u);
                    volatile float q = a / bb;
                    float r = d64 ? birp_div_rn64(a, rb64) : birp_div_rn(a, bb, rb);
                    if (u_of(r) != u_of(q)) b++;
                    c++;
                }
                bad += b; checked += c;
            });
    } else return 2;
    for (auto& x : th) x.join();
    printf("mismatches %llu checked %llu\n", (unsigned long long)bad, (unsigned long long)checked);
    return bad == 0 ? 0 : 1;
}
