// FITS input and output of the host library (no GPU): the correlation-matrix
// reader the imager's command line uses (apps/pacer_imager_main.cpp:720-800)
// and the dirty-image writer (Images::to_fits_files).
//
// With a path argument, prints "nx ny sum first last" of that FITS image so
// that tests/test_cpp_host.py can compare the reader with an independent one.
#include <cstdio>
#include <cstring>
#include <string>

#include "common.hpp"
#include "pacer_imager_mi355x.h"

static const char* kRe = "/eda2_chan204/chan_204_20230205T034618_simul_vis_XX_CorrMatrix_RE.fits";
static const char* kIm = "/eda2_chan204/chan_204_20230205T034618_simul_vis_XX_CorrMatrix_IM.fits";

// The simulated EDA2 correlation matrix is Hermitian off the diagonal: RE
// symmetric, IM antisymmetric (its diagonal carries a constant 0.529 from the
// simulation); a wrong byte order or axis order breaks this.
void test_read_correlation_matrix() {
    const std::string g = golden_dir();
    FitsImage re = read_fits_image(g + kRe), im = read_fits_image(g + kIm);
    if (re.nx != 256 || re.ny != 256 || im.nx != 256 || im.ny != 256)
        throw TestFailed("correlation matrices must be 256 x 256");
    double norm = 0;
    for (long y = 0; y < 256; y++)
        for (long x = 0; x < 256; x++) {
            if (re.getXY(x, y) != re.getXY(y, x)) throw TestFailed("RE is not symmetric");
            if (x != y && im.getXY(x, y) != -im.getXY(y, x))
                throw TestFailed("IM is not antisymmetric");
            norm = std::max(norm, std::abs(re.getXY(x, y)));
        }
    if (!(norm > 0) || !std::isfinite(norm)) throw TestFailed("correlation matrix is empty");
}

void test_visibilities_from_matrices() {
    const std::string g = golden_dir();
    FitsImage re = read_fits_image(g + kRe), im = read_fits_image(g + kIm);
    ObservationInfo info;
    Visibilities xcorr = visibilities_from_correlation_matrices(re, im, nullptr, nullptr, info);
    if (info.nAntennas != 256 || xcorr.size() != 32896u * 4) throw TestFailed("wrong shape");
    for (size_t i = 0; i < 256; i++)
        for (size_t j = 0; j <= i; j++) {
            const std::complex<float>* rec = xcorr.at(0, 0, i, j);
            if (rec[0] != std::complex<float>((float)re.getXY((long)j, (long)i),
                                              (float)im.getXY((long)j, (long)i)))
                throw TestFailed("XX differs from the lower triangle of the matrices");
            if (rec[1] != 0.0f || rec[2] != 0.0f || rec[3] != 0.0f)
                throw TestFailed("XY, YX and YY must be 0 without YY matrices");
        }
}

// Images::to_fits_files writes the real part as BITPIX -32; read it back.
void test_write_read_round_trip() {
    const unsigned n = 12;
    MemoryBuffer<std::complex<float>> buf((size_t)n * n * 2);
    for (size_t i = 0; i < buf.size(); i++) buf[i] = std::complex<float>(0.25f * (float)i - 3.0f, 1.0f);
    ObservationInfo info;
    Images images{std::move(buf), info, 1, 2, n, 10.0, -26.7, 0.5, 0.5};
    char dir[] = "/tmp/blink_fits_XXXXXX";
    if (!mkdtemp(dir)) throw TestFailed("mkdtemp failed");
    images.to_fits_files(dir);
    for (unsigned ch = 0; ch < 2; ch++) {
        char name[512];
        std::snprintf(name, sizeof name, "%s/dirty_image_t%04u_ch%04u_real.fits", dir, 0u, ch);
        FitsImage back = read_fits_image(name);
        if (back.nx != (long)n || back.ny != (long)n) throw TestFailed("round trip: wrong size");
        for (size_t i = 0; i < (size_t)n * n; i++)
            if (back.data[i] != (double)images.image(0, ch)[i].real())
                throw TestFailed("round trip: values differ");
        std::remove(name);
    }
    rmdir(dir);
}

int main(int argc, char** argv) {
    try {
        if (argc > 1) {
            FitsImage img = read_fits_image(argv[1]);
            double sum = 0;
            for (double v : img.data) sum += v;
            std::printf("%ld %ld %.17g %.17g %.17g\n", img.nx, img.ny, sum, img.data.front(),
                        img.data.back());
            return 0;
        }
        RUN_TEST(test_read_correlation_matrix);
        RUN_TEST(test_visibilities_from_matrices);
        RUN_TEST(test_write_read_round_trip);
    } catch (const std::exception& e) {
        std::cerr << "FAILED: " << e.what() << std::endl;
        return 1;
    }
    std::cout << "All tests passed." << std::endl;
    return 0;
}
