// Times CPacerImagerMI355X::run_batched(), the reference-facing batched call,
// on a 128-tile-like array: run by bench.py as a child process, prints one JSON
// line.  Every call goes through the class as a caller would use it: flags,
// UVW and frequencies on the host, their upload, a new device image buffer
// (returned in Images and freed when the next call's result replaces it),
// and the replayed hipGraph updated for the new output address.
//
// usage: bench_run_batched n_ant n_ch n_t n_pixels calls warmup average
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "common.hpp"
#include "pacer_imager_mi355x.h"

int main(int argc, char** argv) {
    if (argc != 8) {
        std::fprintf(stderr, "usage: %s n_ant n_ch n_t n_pixels calls warmup average\n", argv[0]);
        return 2;
    }
    const int n_ant = std::atoi(argv[1]), n_ch = std::atoi(argv[2]), n_t = std::atoi(argv[3]);
    const int N = std::atoi(argv[4]), calls = std::atoi(argv[5]), warmup = std::atoi(argv[6]);
    const bool average = std::atoi(argv[7]) != 0;
    CPacerImager::SetDebugLevel(IMAGER_ERROR_LEVEL);
    // Gaussian layout, sigma 300 m, clipped at 1.5 km
    SplitMix64 g(1002);
    std::vector<double> x, y, z, cable;
    while ((int)x.size() < n_ant) {
        const double px = 300.0 * g.normal(), py = 300.0 * g.normal();
        if (px * px + py * py > 1500.0 * 1500.0) continue;
        x.push_back(px);
        y.push_back(py);
        z.push_back(0.5 * g.normal());
        cable.push_back(50.0 + 250.0 * g.uniform());
    }
    const int n_bl = n_ant * (n_ant + 1) / 2;
    ObservationInfo info{VCS_OBSERVATION_INFO};
    info.nAntennas = n_ant;
    info.nFrequencies = n_ch;
    info.nTimesteps = n_t;
    info.coarseChannel = 133;
    info.frequencyResolution = 0.04;
    info.coarseChannelBandwidth = 1.28;
    MemoryBuffer<std::complex<float>> buf((size_t)n_t * n_ch * n_bl * 4);
    for (int t = 0; t < n_t; t++)
        for (int ch = 0; ch < n_ch; ch++) {
            const double lam = 299792458.0 / ((133 * 1.28 - 0.64 + 0.04 * ch) * 1e6);
            for (int bl = 0; bl < n_bl; bl++) {
                const int a1 = (int)(-0.5 + std::sqrt(0.25 + 2.0 * bl)), a2 = bl - (a1 + 1) * a1 / 2;
                const double u = -(x[a1] - x[a2]), v = -(y[a1] - y[a2]);
                const std::complex<double> s =
                    3.0 * std::exp(std::complex<double>(0, -6.283185307179586 * (0.1 * u - 0.2 * v) / lam));
                std::complex<float>* rec = buf.data() + (((size_t)t * n_ch + ch) * n_bl + bl) * 4;
                for (int p = 0; p < 4; p++)
                    rec[p] = std::complex<float>((float)(0.5 * g.normal()), (float)(0.5 * g.normal()));
                rec[0] += std::complex<float>(s);
                rec[3] += std::complex<float>(0.9 * s);
            }
        }
    Visibilities xcorr(std::move(buf), info, 1, 1);
    xcorr.to_gpu();  // the visibilities are device-resident, as in a GPU pipeline
    CPacerImagerMI355X imager("", N, {}, average, Polarization::XX, 2.0f);
    imager.SetAntennaPositions(x, y, z, cable);
    Images keep;
    for (int i = 0; i < warmup; i++) keep = imager.run_batched(xcorr);
    hip_check(hipDeviceSynchronize(), "sync");
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < calls; i++) keep = imager.run_batched(xcorr);
    hip_check(hipDeviceSynchronize(), "sync");
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const int n_img = n_ch * n_t;
    std::printf("{\"calls\": %d, \"images_per_call\": %d, \"ms_per_call\": %.4f, "
                "\"images_per_s\": %.1f, \"average_images\": %s}\n",
                calls, n_img, 1e3 * s / calls, calls * n_img / s, average ? "true" : "false");
    return 0;
}
