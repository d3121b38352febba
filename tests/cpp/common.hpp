// Shared helpers of the C++ tests (the shape of the reference's
// tests/common.hpp: a TestFailed exception and a main() that returns 1 on
// the first failure).
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <exception>
#include <iostream>
#include <string>
#include <vector>

class TestFailed : public std::exception {
    std::string message;

   public:
    explicit TestFailed(std::string msg) : message(std::move(msg)) {}
    const char* what() const noexcept override { return message.c_str(); }
};

// Deterministic generator for test inputs (splitmix64).
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        const double u1 = uniform() + 1e-300, u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

// max |a - b| / max |b| over complex values.
inline double peak_normalised_error(const std::complex<float>* a, const std::complex<float>* b,
                                    size_t n) {
    double num = 0, den = 0;
    for (size_t i = 0; i < n; i++) {
        num = std::max(num, (double)std::abs(std::complex<double>(a[i]) - std::complex<double>(b[i])));
        den = std::max(den, (double)std::abs(std::complex<double>(b[i])));
    }
    return num / den;
}

#define RUN_TEST(fn)                                                     \
    do {                                                                 \
        fn();                                                            \
        std::cout << "'" #fn "' passed." << std::endl;                   \
    } while (0)

// Directory of the committed fixtures (tests/golden), found from the path of
// the running test program (tests/cpp/bin/<program>).
#include <climits>
#include <unistd.h>
inline std::string golden_dir() {
    char buf[PATH_MAX];
    const ssize_t n = readlink("/proc/self/exe", buf, sizeof buf - 1);
    if (n <= 0) return "tests/golden";
    std::string p(buf, (size_t)n);
    for (int k = 0; k < 3; k++) p = p.substr(0, p.find_last_of('/'));  // -> tests
    return p + "/golden";
}
