// CPU check of opf_div (div_device.inc, the division of the runtime-compiled
// sweeps) against the IEEE division of the host: for every power-of-two divisor
// of both signs and every exponent (normal, subnormal, zero, infinity), random
// dividends of every class (normal, subnormal, zero, infinity, NaN) must give
// the same bits; so must a sample of divisors that are not powers of two.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

static inline long long __double_as_longlong(double d) { long long b; std::memcpy(&b, &d, 8); return b; }
static inline double __longlong_as_double(long long b) { double d; std::memcpy(&d, &b, 8); return d; }
static inline int __float_as_int(float f) { int b; std::memcpy(&b, &f, 4); return b; }
static inline float __int_as_float(int b) { float f; std::memcpy(&f, &b, 4); return f; }
#define __device__
#define __forceinline__ inline
#define OPF_SHARED_SRC(...) __VA_ARGS__
#include "div_device.inc"

int main() {
    std::mt19937_64 rng(11);
    long long checks = 0, bad = 0;
    auto same64 = [](double a, double b) { return __double_as_longlong(a) == __double_as_longlong(b) || (std::isnan(a) && std::isnan(b)); };
    auto same32 = [](float a, float b) { return __float_as_int(a) == __float_as_int(b) || (std::isnan(a) && std::isnan(b)); };
    auto dividend64 = [&](int t) {
        uint64_t b = rng();
        if (t < 8) b = (b & 0x800fffffffffffffULL) | ((uint64_t)(t & 1 ? 0 : 2047) << 52);  // subnormal/zero, inf/NaN
        if (t == 8) b &= 0x8000000000000000ULL;                                               // signed zero
        return __longlong_as_double((long long)b);
    };
    auto dividend32 = [&](int t) {
        uint32_t b = (uint32_t)rng();
        if (t < 8) b = (b & 0x807fffffu) | ((uint32_t)(t & 1 ? 0 : 255) << 23);
        if (t == 8) b &= 0x80000000u;
        return __int_as_float((int)b);
    };
    for (uint64_t e = 0; e < 2048; ++e)
        for (uint64_t sg = 0; sg < 2; ++sg) {
            const double y = __longlong_as_double((long long)((sg << 63) | (e << 52)));
            for (int t = 0; t < 1500; ++t) {
                const double x = dividend64(t);
                ++checks;
                if (!same64(opf_div(x, y), x / y) && bad++ < 10) printf("FAIL f64 x=%a y=%a\n", x, y);
            }
        }
    for (uint32_t e = 0; e < 256; ++e)
        for (uint32_t sg = 0; sg < 2; ++sg) {
            const float y = __int_as_float((int)((sg << 31) | (e << 23)));
            for (int t = 0; t < 12000; ++t) {
                const float x = dividend32(t);
                ++checks;
                if (!same32(opf_div(x, y), x / y) && bad++ < 10) printf("FAIL f32 x=%a y=%a\n", x, y);
            }
        }
    for (int t = 0; t < 2000000; ++t) {  // divisors that are not powers of two
        const double x = dividend64(t % 64), y = dividend64(t % 61);
        const float xf = dividend32(t % 64), yf = dividend32(t % 61);
        checks += 2;
        if (!same64(opf_div(x, y), x / y) && bad++ < 10) printf("FAIL f64 x=%a y=%a\n", x, y);
        if (!same32(opf_div(xf, yf), xf / yf) && bad++ < 10) printf("FAIL f32 x=%a y=%a\n", xf, yf);
    }
    printf("%lld checks, %lld mismatches: %s\n", checks, bad, bad ? "FAILED" : "passed");
    return bad ? 1 : 0;
}
