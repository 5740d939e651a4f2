// CPU check of div_by_recip and div_rn (supereight2_amd/csrc/se8_internal.hpp), the voxel-position
// division of the raycaster and the carver, against IEEE division (x86 divss, rounded to nearest).
// The header is compiled unchanged as host C++ against tests/cpp/simt's HIP stand-ins; the host
// and the device both evaluate the fmas with one rounding, so the host's results are the device's.
//
// Without underflow or overflow a quotient's rounding depends only on the significands of x and
// b, so for a divisor every x in [1, 2) covers every x of the exact range.  Checked:
//   * every x significand against 16 map resolutions, 8 divisors with extreme significands and
//     1024 divisors with random significands (8.7e9 quotients),
//   * x over the exponent range [-60, 60] and both signs for a stride of significands,
//   * div_rn's guarded cases: zero, signed zero, x below 2^-60 and subnormal, above 2^60,
//     infinities and NaN.
// usage: div_check [random divisors]   prints per failing case "FAIL ..." and "OK" when all agree
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "se8_internal.hpp"

namespace se8 {
void set_error(const std::string&) {}
} // namespace se8

static float f_of(uint32_t u)
{
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

static uint32_t u_of(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

// every significand of x in [1, 2) against b: the number of quotients that differ from x / b
static long sweep_significands(float b)
{
    const float inv_b = 1.0f / b;
    long bad = 0;
    for (uint32_t m = 0; m < 0x800000u; m++) {
        const float x = f_of(0x3f800000u | m);
        bad += u_of(se8::div_by_recip(x, b, inv_b)) != u_of(x / b);
    }
    return bad;
}

int main(int argc, char** argv)
{
    const int n_random = argc > 1 ? std::atoi(argv[1]) : 1024;
    std::vector<float> divisors = {0.01f, 0.02f,   0.005f, 0.04f, 0.1f,   0.05f,  0.025f, 0.0125f,
                                   0.08f, 0.2f,    1.0f,   0.03f, 0.015f, 0.0075f, 0.001f, 0.5f};
    // extreme significands: all ones, one ulp above 1, one ulp above 1.5, 2 - ulp, 1.5 - ulp
    for (uint32_t u : {0x3f7fffffu, 0x3f800001u, 0x3fc00001u, 0x3fffffffu, 0x3fbfffffu, 0x3c7fffffu,
                       0x3c000001u, 0x3c23d70au})
        divisors.push_back(f_of(u));
    uint64_t s = 0x9e3779b97f4a7c15ull;
    for (int i = 0; i < n_random; i++) {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        const uint32_t sig = (uint32_t) (s >> 41) & 0x7fffffu;
        const int e = (int) ((s >> 20) % 16) - 12; // exponents 2^-12 .. 2^3
        divisors.push_back(std::ldexp(f_of(0x3f800000u | sig), e));
    }
    // significand sweeps, split over threads
    const unsigned n_threads = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    std::vector<long> bad(divisors.size(), 0);
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < n_threads; t++)
        pool.emplace_back([&, t] {
            for (size_t i = t; i < divisors.size(); i += n_threads)
                bad[i] = sweep_significands(divisors[i]);
        });
    for (auto& th : pool)
        th.join();
    long total_bad = 0;
    for (size_t i = 0; i < divisors.size(); i++)
        if (bad[i]) {
            std::printf("FAIL significands b=%a: %ld quotients differ\n", divisors[i], bad[i]);
            total_bad += bad[i];
        }
    // the exponent range of the exact region, both signs, for a stride of significands
    long bad_exp = 0, n_exp = 0;
    for (size_t i = 0; i < 24; i++) {
        const float b = divisors[i], inv_b = 1.0f / b;
        for (int e = -60; e <= 59; e++)
            for (uint32_t m = 0; m < 0x800000u; m += 4099) {
                for (float sign : {1.0f, -1.0f}) {
                    const float x = sign * std::ldexp(f_of(0x3f800000u | m), e);
                    n_exp++;
                    if (u_of(se8::div_by_recip(x, b, inv_b)) != u_of(x / b) || u_of(se8::div_rn(x, b, inv_b)) != u_of(x / b)) {
                        if (bad_exp++ < 10)
                            std::printf("FAIL exponent range b=%a x=%a\n", b, x);
                    }
                }
            }
    }
    // div_rn outside the exact range falls back to the division itself
    long bad_guard = 0;
    const float specials[] = {0.0f, -0.0f, 0x1p-61f, -0x1p-61f, 0x1p-100f, 0x1p-130f, -0x1p-149f, f_of(0x00000001u),
                              0x1p61f, -0x1p100f, 0x1.fffffep127f, INFINITY, -INFINITY, NAN, 0x1.fffffep-61f};
    for (size_t i = 0; i < 24; i++) {
        const float b = divisors[i], inv_b = 1.0f / b;
        for (float x : specials) {
            const float want = x / b, got = se8::div_rn(x, b, inv_b);
            if (!(u_of(got) == u_of(want) || (std::isnan(got) && std::isnan(want)))) {
                bad_guard++;
                std::printf("FAIL guard b=%a x=%a: %a vs %a\n", b, x, got, want);
            }
        }
    }
    std::printf("divisors %zu: %ld significand sweeps differ; exponent range %ld/%ld differ; guards %ld differ\n",
                divisors.size(), total_bad, bad_exp, n_exp, bad_guard);
    if (total_bad || bad_exp || bad_guard)
        return 1;
    std::printf("OK\n");
    return 0;
}
