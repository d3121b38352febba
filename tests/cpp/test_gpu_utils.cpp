// The reference's tests/gpu_utils_test.cpp known-answer tests, run on the
// MI355X C ABI: vector sum, fft shift + normalisation, image averaging.
#include <cstdlib>
#include <cstring>

#include "common.hpp"
#include "pacer_imager_mi355x.h"

static blink_ctx* g_ctx = nullptr;

static void ok(int rc) {
    if (rc != BLINK_OK) throw TestFailed(blink_last_error(g_ctx));
}

// tests/gpu_utils_test.cpp:20-57: 5 x 3000 values rand() % 50, exact sums.
void test_vector_sum() {
    const int n_images = 5, image_size = 3000;
    MemoryBuffer<float> data(image_size * n_images);
    std::vector<double> ref(n_images, 0.0);
    srand(1);
    for (int i = 0; i < n_images; i++)
        for (int p = 0; p < image_size; p++) {
            const float val = rand() % 50;
            data[i * image_size + p] = val;
            ref[i] += val;
        }
    data.to_gpu();
    MemoryBuffer<double> sums(n_images, MemoryType::DEVICE);
    ok(blink_vector_sum(g_ctx, data.data(), image_size, n_images, sums.data(), nullptr));
    sums.to_cpu();
    for (int i = 0; i < n_images; i++)
        if (sums[i] != ref[i]) throw TestFailed("test_vector_sum: sums differ");
}

// tests/gpu_utils_test.cpp:60-87: 3 images of 2x2, fnorm = 1.
void test_fft_shift_and_norm() {
    std::complex<float> in[]{{0, 0}, {1, 0}, {2, 0}, {3, 0}, {0, 0}, {1, 0},
                             {2, 0}, {3, 0}, {0, 0}, {1, 0}, {2, 0}, {3, 0}};
    std::complex<float> ref[]{{3, 0}, {2, 0}, {1, 0}, {0, 0}, {3, 0}, {2, 0},
                              {1, 0}, {0, 0}, {3, 0}, {2, 0}, {1, 0}, {0, 0}};
    MemoryBuffer<std::complex<float>> buf(12);
    for (int i = 0; i < 12; i++) buf[i] = in[i];
    MemoryBuffer<double> fnorm(3);
    fnorm[0] = fnorm[1] = fnorm[2] = 1;
    fnorm.to_gpu();
    buf.to_gpu();
    ok(blink_fft_shift_and_norm(g_ctx, (float*)buf.data(), 2, 2, 3, fnorm.data(), 3, nullptr));
    buf.to_cpu();
    for (int x = 0; x < 12; x++)
        if (buf[x] != ref[x]) throw TestFailed("test_fft_shift_and_norm: wrong output");
}

// tests/gpu_utils_test.cpp:91-124: 4 images of 512^2 with values i % 100.
void test_averaging_kernel() {
    const int n_images = 4, side = 512, n = side * side;
    MemoryBuffer<std::complex<float>> imgs((size_t)n_images * n);
    std::vector<std::complex<float>> ref(n);
    for (int k = 0; k < n_images; k++)
        for (int i = 0; i < n; i++) {
            const float val = i % 100;
            imgs[(size_t)k * n + i] = std::complex<float>(val, 0);
            ref[i] += std::complex<float>(val, 0);
        }
    for (int i = 0; i < n; i++) ref[i] /= n_images;
    imgs.to_gpu();
    Images input{std::move(imgs), ObservationInfo{}, 2, 2, (unsigned)side, 0, 0, 0, 0};
    MemoryBuffer<std::complex<float>> avg(n, MemoryType::DEVICE);
    ok(blink_image_averaging(g_ctx, (const float*)input.data(), n, (int)input.size(), 0,
                             (float*)avg.data(), nullptr));
    avg.to_cpu();
    for (int i = 0; i < n; i++)
        if (avg[i] != ref[i]) throw TestFailed("test_averaging_kernel: output differs");
}

int main() {
    if (blink_create(0, &g_ctx) != BLINK_OK) {
        std::cerr << "no GPU" << std::endl;
        return 1;
    }
    int rc = 0;
    try {
        RUN_TEST(test_vector_sum);
        RUN_TEST(test_fft_shift_and_norm);
        RUN_TEST(test_averaging_kernel);
        std::cout << "All tests passed." << std::endl;
    } catch (std::exception& ex) {
        std::cerr << ex.what() << std::endl;
        rc = 1;
    }
    blink_destroy(g_ctx);
    return rc;
}
