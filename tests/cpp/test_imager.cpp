// The shape of the reference's tests/imager_test.cpp for CPacerImagerMI355X:
// the fft_shift known-answer tests, then imaging runs checked against the CPU
// oracle (oracle/blink_oracle.c) on identical synthetic inputs.
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "blink_oracle.h"
#include "common.hpp"
#include "pacer_imager_mi355x.h"

// tests/imager_test.cpp:21-40
void test_fft_shift_simple() {
    std::complex<float> one[]{{0, 0}, {1, 0}, {2, 0}, {3, 0}};
    std::complex<float> ref_one[]{{3, 0}, {2, 0}, {1, 0}, {0, 0}};
    fft_shift(one, 2, 2);
    for (int x = 0; x < 4; x++)
        if (one[x] != ref_one[x]) throw TestFailed("test_fft_shift_simple: test 1, wrong output.");
    std::complex<float> two[]{{0, 0}, {1, 0}, {2, 0}, {3, 0}, {4, 0}, {5, 0}};
    std::complex<float> ref_two[]{{5, 0}, {4, 0}, {3, 0}, {2, 0}, {1, 0}, {0, 0}};
    fft_shift(two, 3, 2);
    for (int x = 0; x < 6; x++)
        if (two[x] != ref_two[x]) throw TestFailed("test_fft_shift_simple: test 2, wrong output.");
}

struct Array {
    int n_ant;
    std::vector<double> x, y, z, cable;
};

static Array make_array(int n_ant, double radius, uint64_t seed) {
    SplitMix64 g(seed);
    Array a{n_ant, {}, {}, {}, {}};
    for (int i = 0; i < n_ant; i++) {
        const double r = radius * std::sqrt(g.uniform()), th = 6.283185307179586 * g.uniform();
        a.x.push_back(r * std::cos(th));
        a.y.push_back(r * std::sin(th));
        a.z.push_back(0.5 * g.normal());
        a.cable.push_back(50.0 + 250.0 * g.uniform());
    }
    return a;
}

// Point sources + noise, all 4 pol products, [t][ch][bl][4].
static Visibilities make_vis(const Array& a, int n_ch, int n_t, const std::vector<double>& freq,
                             uint64_t seed, ObservationInfo& info) {
    const int n_bl = a.n_ant * (a.n_ant + 1) / 2;
    SplitMix64 g(seed);
    double l[4], m[4], s[4];
    for (int k = 0; k < 4; k++) {
        l[k] = 0.6 * (g.uniform() - 0.5);
        m[k] = 0.6 * (g.uniform() - 0.5);
        s[k] = 1.0 + 9.0 * g.uniform();
    }
    info.nAntennas = a.n_ant;
    info.nFrequencies = n_ch;
    info.nTimesteps = n_t;
    info.frequencyResolution = 0.04;
    info.coarseChannelBandwidth = 1.28;
    info.coarseChannel = 133;
    MemoryBuffer<std::complex<float>> buf((size_t)n_t * n_ch * n_bl * 4);
    for (int t = 0; t < n_t; t++)
        for (int ch = 0; ch < n_ch; ch++) {
            const double lam = 299792458.0 / freq[ch];
            for (int bl = 0; bl < n_bl; bl++) {
                const int a1 = (int)(-0.5 + std::sqrt(0.25 + 2.0 * bl)), a2 = bl - (a1 + 1) * a1 / 2;
                const double u = -(a.x[a1] - a.x[a2]), v = -(a.y[a1] - a.y[a2]);
                std::complex<double> xx = 0;
                for (int k = 0; k < 4; k++)
                    xx += s[k] * std::exp(std::complex<double>(0, -6.283185307179586 * (u * l[k] + v * m[k]) / lam));
                std::complex<float>* rec = buf.data() + (((size_t)t * n_ch + ch) * n_bl + bl) * 4;
                for (int p = 0; p < 4; p++)
                    rec[p] = std::complex<float>((float)(0.1 * g.normal()), (float)(0.1 * g.normal()));
                rec[0] += std::complex<float>(xx);
                rec[3] += std::complex<float>(0.9 * xx);
            }
        }
    return Visibilities(std::move(buf), info, 1, 1);
}

static void check_against_oracle(CPacerImagerMI355X& imager, Images& images, const Array& a,
                                 const std::vector<float>& vis0, const std::vector<double>& freq,
                                 int n_ch, int n_t, bool geom, bool cable, int pol, double tol,
                                 const char* what) {
    const int n = imager.n_pixels;
    const size_t gs = (size_t)n * n, n_img = (size_t)n_ch * n_t;
    const int n_bl = a.n_ant * (a.n_ant + 1) / 2;
    std::vector<float> u(n_bl), v(n_bl), w(n_bl);
    oracle_zenith_uvw(a.n_ant, a.x.data(), a.y.data(), a.z.data(), u.data(), v.data(), w.data());
    std::vector<int> flags(n_bl);
    oracle_baseline_flags(a.n_ant, imager.m_FlaggedAntennas.data(),
                          (int)imager.m_FlaggedAntennas.size(), flags.data());
    double um, vm, du, dv;
    oracle_uv_max(u.data(), v.data(), flags.data(), n_bl, &um, &vm);
    oracle_delta_uv(imager.oversampling_factor, um, vm, n, &du, &dv);
    if (du != imager.delta_u || dv != imager.delta_v) throw TestFailed(std::string(what) + ": delta_u/v differ");
    std::vector<float> vis(vis0), grids(2 * gs * n_img), counts(gs * n_img), ref(2 * gs * n_img);
    oracle_run_f32(vis.data(), a.n_ant, n_ch, n_t, u.data(), v.data(), w.data(), flags.data(),
                   freq.data(), a.cable.data(), geom, cable, du, dv, imager.min_uv, n, pol,
                   grids.data(), counts.data(), ref.data(), 0);
    images.to_cpu();
    const double err = peak_normalised_error(images.data(), (const std::complex<float>*)ref.data(), gs * n_img);
    std::cout << what << ": peak-normalised error vs oracle " << err << std::endl;
    if (!(err < tol)) throw TestFailed(std::string(what) + ": image differs from the oracle");
}

// tests/imager_test.cpp:42-58 images real MWA data with the CPU and GPU imagers
// and compares nothing; here the imager is run on synthetic data and the
// images are compared with the oracle's CPU restatement of the reference.
void test_imager_run() {
    const int n_ch = 3, n_t = 2, N = 128;
    Array a = make_array(48, 17.5, 11);
    std::vector<double> freq;
    for (int ch = 0; ch < n_ch; ch++) freq.push_back((133 * 1.28 - 0.64 + 0.04 * ch) * 1e6);
    ObservationInfo info;
    Visibilities xcorr = make_vis(a, n_ch, n_t, freq, 5, info);
    std::vector<float> vis0((float*)xcorr.data(), (float*)xcorr.data() + 2 * xcorr.size());
    std::vector<int> flagged{3, 17};
    CPacerImagerMI355X imager("", N, flagged, false, Polarization::XX, 2.0f);
    imager.SetAntennaPositions(a.x, a.y, a.z, a.cable);
    Images images = imager.run(xcorr);
    if (images.size() != (size_t)n_ch * n_t || images.side_size != (unsigned)N)
        throw TestFailed("test_imager_run: wrong image shape");
    check_against_oracle(imager, images, a, vis0, freq, n_ch, n_t, true, true, ORACLE_POL_XX, 1e-5,
                         "run (corrections in place + gridding + image)");
    // visibilities were corrected in place, as the reference does
    xcorr.to_cpu();
    std::vector<float> corrected(vis0);
    std::vector<float> w(a.n_ant * (a.n_ant + 1) / 2), u(w.size()), v(w.size());
    oracle_zenith_uvw(a.n_ant, a.x.data(), a.y.data(), a.z.data(), u.data(), v.data(), w.data());
    oracle_geometric_corrections(corrected.data(), a.n_ant, n_ch, n_t, w.data(), freq.data());
    oracle_cable_corrections(corrected.data(), a.n_ant, n_ch, n_t, a.cable.data(), freq.data());
    double worst = 0;
    for (size_t i = 0; i + 1 < corrected.size(); i += 2) {
        const std::complex<double> g(((float*)xcorr.data())[i], ((float*)xcorr.data())[i + 1]);
        const std::complex<double> r(corrected[i], corrected[i + 1]);
        worst = std::max(worst, std::abs(g - r) / std::max(std::abs(r), 1e-30));
    }
    if (!(worst < 2e-6)) throw TestFailed("test_imager_run: corrected visibilities differ");
}

void test_imager_batched_and_pols() {
    const int n_ch = 2, n_t = 2, N = 256;
    Array a = make_array(64, 17.5, 21);
    std::vector<double> freq{159.375e6, 160.0e6};
    ObservationInfo info;
    Visibilities xcorr = make_vis(a, n_ch, n_t, freq, 6, info);
    std::vector<float> vis0((float*)xcorr.data(), (float*)xcorr.data() + 2 * xcorr.size());
    const Polarization pols[] = {Polarization::XX, Polarization::YY, Polarization::I, Polarization::V};
    for (int p = 0; p < 4; p++) {
        CPacerImagerMI355X imager("", N, {}, false, pols[p], 2.0f);
        imager.SetAntennaPositions(a.x, a.y, a.z, a.cable);
        imager.m_fFrequencyMHz = -1;
        // get_frequency_hz: coarse channel 133, so override with explicit values
        Visibilities copy(MemoryBuffer<std::complex<float>>(xcorr.size()), info, 1, 1);
        std::memcpy((void*)copy.data(), vis0.data(), sizeof(float) * vis0.size());
        copy.obsInfo.frequencyResolution = 0.625;
        copy.obsInfo.coarseChannelBandwidth = 1.25;
        copy.obsInfo.coarseChannel = 128;  // 159.375 MHz + 0.625 ch
        Images images = imager.run_batched(copy);
        check_against_oracle(imager, images, a, vis0, freq, n_ch, n_t, true, true, p, 1e-5,
                             "run_batched");
    }
}

// run_batched() with average_images (only the mean image is formed) against
// the oracle's images and their complex mean (image_averaging_gpu,
// src/gpu/gpu_utils.cpp:143-161), and with weighting "U" on an array that
// leaves cells empty: the reference's 0/0 cells make every pixel NaN
// (src/pacer_imager.cpp:421-428).
void test_imager_batched_average_and_uniform() {
    const int n_ch = 3, n_t = 2, N = 256;
    Array a = make_array(48, 17.5, 41);
    std::vector<double> freq{159.375e6, 160.0e6, 160.625e6};
    ObservationInfo info;
    Visibilities xcorr = make_vis(a, n_ch, n_t, freq, 8, info);
    std::vector<float> vis0((float*)xcorr.data(), (float*)xcorr.data() + 2 * xcorr.size());
    const size_t gs = (size_t)N * N, n_img = (size_t)n_ch * n_t;
    const int n_bl = a.n_ant * (a.n_ant + 1) / 2;
    auto fresh = [&]() {
        Visibilities copy(MemoryBuffer<std::complex<float>>(xcorr.size()), info, 1, 1);
        std::memcpy((void*)copy.data(), vis0.data(), sizeof(float) * vis0.size());
        copy.obsInfo.frequencyResolution = 0.625;
        copy.obsInfo.coarseChannelBandwidth = 1.25;
        copy.obsInfo.coarseChannel = 128;  // 159.375 MHz + 0.625 ch
        return copy;
    };
    for (int transform : {BLINK_TRANSFORM_FFT, BLINK_TRANSFORM_DFT_FP16X3}) {
        CPacerImagerMI355X imager("", N, {2}, true, Polarization::XX, 2.0f);
        imager.SetAntennaPositions(a.x, a.y, a.z, a.cable);
        imager.m_nBatchTransform = transform;
        Visibilities copy = fresh();
        Images avg = imager.run_batched(copy);
        if (avg.size() != 1 || avg.side_size != (unsigned)N)
            throw TestFailed("run_batched with average_images must return one image");
        std::vector<float> u(n_bl), v(n_bl), w(n_bl);
        oracle_zenith_uvw(a.n_ant, a.x.data(), a.y.data(), a.z.data(), u.data(), v.data(), w.data());
        std::vector<int> flags(n_bl);
        oracle_baseline_flags(a.n_ant, imager.m_FlaggedAntennas.data(),
                              (int)imager.m_FlaggedAntennas.size(), flags.data());
        std::vector<float> vis(vis0), grids(2 * gs * n_img), counts(gs * n_img),
            images(2 * gs * n_img), ref(2 * gs);
        oracle_run_f32(vis.data(), a.n_ant, n_ch, n_t, u.data(), v.data(), w.data(), flags.data(),
                       freq.data(), a.cable.data(), 1, 1, imager.delta_u, imager.delta_v,
                       imager.min_uv, N, ORACLE_POL_XX, grids.data(), counts.data(),
                       images.data(), 0);
        oracle_image_averaging_complex(images.data(), gs, n_img, ref.data());
        avg.to_cpu();
        const double err = peak_normalised_error(avg.data(), (const std::complex<float>*)ref.data(), gs);
        std::cout << "run_batched average (transform " << transform
                  << "): peak-normalised error vs oracle " << err << std::endl;
        if (!(err < 1e-5)) throw TestFailed("run_batched average differs from the oracle");
    }
    CPacerImagerMI355X uni("", N, {}, false, Polarization::XX, 2.0f, -std::numeric_limits<double>::infinity(),
                           std::numeric_limits<double>::infinity(), "U");
    uni.SetAntennaPositions(a.x, a.y, a.z, a.cable);
    Visibilities copy = fresh();
    Images images = uni.run_batched(copy);
    images.to_cpu();
    for (size_t i = 0; i < images.size() * gs; i++)
        if (!std::isnan(images.data()[i].real()))
            throw TestFailed("\"U\" with empty cells must give NaN images");
}

// src/pacer_imager.cpp:135: image() before grid() throws std::runtime_error.
void test_error_paths() {
    CPacerImagerMI355X imager("", 64, {}, false);
    ObservationInfo info;
    bool thrown = false;
    try {
        imager.image(info);
    } catch (const std::runtime_error&) {
        thrown = true;
    }
    if (!thrown) throw TestFailed("image() before grid() must throw std::runtime_error");
    Array a = make_array(8, 10, 3);
    imager.SetAntennaPositions(a.x, a.y, a.z, a.cable);
    std::vector<double> freq{150e6};
    Visibilities xcorr = make_vis(a, 1, 1, freq, 1, info);
    imager.m_nConvolvingKernelSize = 7;
    imager.weighting = "U";
    thrown = false;
    try {
        imager.run_batched(xcorr);
    } catch (const std::invalid_argument&) {
        thrown = true;
    }
    if (!thrown) throw TestFailed("run_batched must refuse convolutional gridding with \"U\"");
    imager.weighting = "";
    imager._gridMode = GridMode::GaussianKernel;
    thrown = false;
    try {
        imager.run(xcorr);
    } catch (const std::runtime_error&) {
        thrown = true;
    }
    if (!thrown) throw TestFailed("kernels other than Kaiser-Bessel must be refused");
}

// tests/conv_test.cpp:22-36 runs the CPU imager with m_nConvolvingKernelSize =
// 10 on MWA data and only writes FITS files; here run() with the Kaiser-Bessel
// kernel (sizes 7 and 10, with and without the sinc, image averaging on) is
// compared with the oracle's restatement of the convolutional gridding.
void test_imager_kaiser_bessel() {
    const int n_ch = 2, n_t = 2, N = 128;
    Array a = make_array(40, 17.5, 31);
    std::vector<double> freq{150.625e6, 151.25e6};  // exact in binary, as get_frequency_hz gives them
    ObservationInfo info;
    Visibilities xcorr0 = make_vis(a, n_ch, n_t, freq, 3, info);
    std::vector<float> vis0((float*)xcorr0.data(), (float*)xcorr0.data() + 2 * xcorr0.size());
    const int n_bl = a.n_ant * (a.n_ant + 1) / 2;
    const size_t gs = (size_t)N * N, n_img = (size_t)n_ch * n_t;
    const int sizes[] = {7, 10};
    const GridMode modes[] = {GridMode::KaiserBesselKernel, GridMode::KaiserBesselWithoutSinc};
    for (int K : sizes)
        for (GridMode mode : modes) {
            CPacerImagerMI355X imager("", N, {5}, true, Polarization::XX, 2.0f);
            imager.SetAntennaPositions(a.x, a.y, a.z, a.cable);
            imager.m_fFrequencyMHz = -1;
            imager.m_nConvolvingKernelSize = K;
            imager._gridMode = mode;
            Visibilities xcorr(MemoryBuffer<std::complex<float>>(xcorr0.size()), info, 1, 1);
            std::memcpy((void*)xcorr.data(), vis0.data(), sizeof(float) * vis0.size());
            xcorr.obsInfo.frequencyResolution = 0.625;
            xcorr.obsInfo.coarseChannelBandwidth = 1.25;
            xcorr.obsInfo.coarseChannel = 121;  // 150.625 MHz + 0.625 ch
            // the same through run_batched (fused, the per-image images never formed)
            CPacerImagerMI355X batched("", N, {5}, true, Polarization::XX, 2.0f);
            batched.SetAntennaPositions(a.x, a.y, a.z, a.cable);
            batched.m_fFrequencyMHz = -1;
            batched.m_nConvolvingKernelSize = K;
            batched._gridMode = mode;
            Visibilities xb(MemoryBuffer<std::complex<float>>(xcorr0.size()), info, 1, 1);
            std::memcpy((void*)xb.data(), vis0.data(), sizeof(float) * vis0.size());
            xb.obsInfo.frequencyResolution = 0.625;
            xb.obsInfo.coarseChannelBandwidth = 1.25;
            xb.obsInfo.coarseChannel = 121;
            Images avg_b = batched.run_batched(xb);
            Images avg = imager.run(xcorr);
            if (avg.size() != 1 || avg_b.size() != 1)
                throw TestFailed("average_images must return one image");
            // oracle: corrections, convolutional gridding, image, complex average
            std::vector<float> u(n_bl), v(n_bl), w(n_bl);
            oracle_zenith_uvw(a.n_ant, a.x.data(), a.y.data(), a.z.data(), u.data(), v.data(),
                              w.data());
            std::vector<int> flags(n_bl);
            oracle_baseline_flags(a.n_ant, imager.m_FlaggedAntennas.data(),
                                  (int)imager.m_FlaggedAntennas.size(), flags.data());
            std::vector<float> vis(vis0), grids(2 * gs * n_img), counts(gs * n_img),
                images(2 * gs * n_img), ref(2 * gs);
            oracle_geometric_corrections(vis.data(), a.n_ant, n_ch, n_t, w.data(), freq.data());
            oracle_cable_corrections(vis.data(), a.n_ant, n_ch, n_t, a.cable.data(), freq.data());
            oracle_gridding_kb_f32(vis.data(), a.n_ant, n_ch, n_t, u.data(), v.data(),
                                   flags.data(), freq.data(), imager.delta_u, imager.delta_v,
                                   imager.min_uv, N, ORACLE_POL_XX, 0, K,
                                   mode == GridMode::KaiserBesselKernel, grids.data(),
                                   counts.data());
            oracle_image_f32(grids.data(), counts.data(), N, (int)n_img, images.data(), 0);
            oracle_image_averaging_complex(images.data(), gs, n_img, ref.data());
            avg.to_cpu();
            avg_b.to_cpu();
            const double err = peak_normalised_error(avg.data(),
                                                     (const std::complex<float>*)ref.data(), gs);
            const double err_b = peak_normalised_error(avg_b.data(),
                                                       (const std::complex<float>*)ref.data(), gs);
            std::cout << "Kaiser-Bessel K=" << K << (mode == GridMode::KaiserBesselKernel ? " sinc" : "")
                      << ": peak-normalised error of the average vs oracle " << err
                      << " (run), " << err_b << " (run_batched)" << std::endl;
            if (!(err < 1e-5)) throw TestFailed("Kaiser-Bessel run differs from the oracle");
            if (!(err_b < 1e-5)) throw TestFailed("Kaiser-Bessel run_batched differs from the oracle");
        }
}

// The simulated EDA2 correlation matrices the reference's command line images
// (scripts/simul/eda2/20230205/templates, apps/pacer_imager_main.cpp:720-800),
// read from FITS and imaged at 180^2 with the antenna file's positions, by
// run() and by run_batched(), against the oracle's pipeline on the same input.
void test_eda2_correlation_matrices() {
    const std::string g = golden_dir() + "/eda2_chan204/";
    FitsImage re = read_fits_image(g + "chan_204_20230205T034618_simul_vis_XX_CorrMatrix_RE.fits");
    FitsImage im = read_fits_image(g + "chan_204_20230205T034618_simul_vis_XX_CorrMatrix_IM.fits");
    ObservationInfo info;
    Visibilities xcorr = visibilities_from_correlation_matrices(re, im, nullptr, nullptr, info);
    std::vector<float> vis0((float*)xcorr.data(), (float*)xcorr.data() + 2 * xcorr.size());
    CAntennaPositions ants;
    if (ants.ReadAntennaPositions((g + "antenna_locations.txt").c_str()) != 256)
        throw TestFailed("antenna_locations.txt must list 256 antennas");
    Array a;
    a.n_ant = 256;
    for (const InputMapping& m : ants) {
        a.x.push_back(m.x);
        a.y.push_back(m.y);
        a.z.push_back(m.z);
        a.cable.push_back(m.cableLenDelta);
    }
    const std::vector<double> freq{159.375e6};  // channel 204 x 0.78125 MHz
    // run(), then run_batched() with the FFT and both split-precision DFTs
    const int transforms[] = {-1, BLINK_TRANSFORM_FFT, BLINK_TRANSFORM_DFT_BF16X3,
                              BLINK_TRANSFORM_DFT_FP16X3};
    const char* names[] = {"EDA2 matrices, run", "EDA2 matrices, run_batched FFT",
                           "EDA2 matrices, run_batched DFT bf16x3",
                           "EDA2 matrices, run_batched DFT fp16x3"};
    for (int k = 0; k < 4; k++) {
        CPacerImagerMI355X imager("", 180, {}, false, Polarization::XX, 2.0f);
        imager.m_MetaData.m_AntennaPositions.ReadAntennaPositions((g + "antenna_locations.txt").c_str());
        imager.UpdateFlags();
        imager.m_fFrequencyMHz = 159.375;
        if (transforms[k] >= 0) imager.m_nBatchTransform = transforms[k];
        Visibilities copy(MemoryBuffer<std::complex<float>>(xcorr.size()), info, 1, 1);
        std::memcpy((void*)copy.data(), vis0.data(), sizeof(float) * vis0.size());
        Images images = transforms[k] >= 0 ? imager.run_batched(copy) : imager.run(copy);
        check_against_oracle(imager, images, a, vis0, freq, 1, 1, true, true, ORACLE_POL_XX,
                             1e-5, names[k]);
    }
}

// The command line's set-up, written against the base class as
// apps/pacer_imager_main.cpp:653-718 writes Initialise(CPacerImager&): read
// the antenna list, update the flags, zenith UVW when no metadata file is
// given.
static void initialise(CPacerImager& imager, const std::string& antenna_file) {
    const int n = imager.m_MetaData.m_AntennaPositions.ReadAntennaPositions(antenna_file.c_str(),
                                                                             false);
    if (n != 256) throw TestFailed("initialise: the antenna list must hold 256 antennas");
    imager.UpdateFlags();
    imager.CalculateUVW();
}

// The command line's flow on the EDA2 correlation matrices
// (apps/pacer_imager_main.cpp:509-584): the GPU class built with the command
// line's constructor arguments, set up through a CPacerImager&, corrections
// switched off, an ObservationInfo{VCS_OBSERVATION_INFO}, the matrices turned
// into Visibilities, then two rounds of m_nConvolvingKernelSize / m_fFrequencyMHz
// / run() / to_fits_files() -- every image against the oracle, every FITS file
// read back.  The metadata readers the command line calls when a metafits
// file is given throw (this library takes the array from the antenna list).
void test_command_line_flow() {
    const std::string g = golden_dir() + "/eda2_chan204/";
    std::vector<int> flagged{12, 200};
    CPacerImagerMI355X imager("", 180, flagged, false, Polarization::XX);
    CPacerImager& base = imager;
    initialise(base, g + "antenna_locations.txt");
    if (imager.m_Baselines != 256 * 257 / 2) throw TestFailed("CalculateUVW: wrong baseline count");
    imager.apply_geom_correction = false;
    imager.apply_cable_correction = false;
    FitsImage re = read_fits_image(g + "chan_204_20230205T034618_simul_vis_XX_CorrMatrix_RE.fits");
    FitsImage im = read_fits_image(g + "chan_204_20230205T034618_simul_vis_XX_CorrMatrix_IM.fits");
    ObservationInfo obsInfo{VCS_OBSERVATION_INFO};
    if (obsInfo.coarseChannel != 109 || obsInfo.frequencyResolution != 0.040)
        throw TestFailed("VCS_OBSERVATION_INFO does not carry the command line's values");
    Visibilities xcorr = visibilities_from_correlation_matrices(re, im, nullptr, nullptr, obsInfo);
    std::vector<float> vis0((float*)xcorr.data(), (float*)xcorr.data() + 2 * xcorr.size());
    Array a;
    a.n_ant = 256;
    for (const InputMapping& m : imager.m_MetaData.m_AntennaPositions) {
        a.x.push_back(m.x);
        a.y.push_back(m.y);
        a.z.push_back(m.z);
        a.cable.push_back(m.cableLenDelta);
    }
    const std::string out = "/tmp/blink_cli_flow";
    if (std::system(("mkdir -p " + out).c_str()) != 0) throw TestFailed("cannot create " + out);
    for (int i = 0; i < 2; i++) {
        imager.m_nConvolvingKernelSize = 0;
        imager.m_fFrequencyMHz = 159.375;
        Images image = base.run(xcorr);
        check_against_oracle(imager, image, a, vis0, {159.375e6}, 1, 1, false, false,
                             ORACLE_POL_XX, 1e-5, i == 0 ? "command line flow, image 0"
                                                         : "command line flow, image 1");
        image.to_fits_files(out);
        FitsImage back = read_fits_image(out + "/dirty_image_t0000_ch0000_real.fits");
        if (back.nx != 180 || back.ny != 180) throw TestFailed("FITS image of the wrong size");
        for (size_t k = 0; k < back.data.size(); k += 997)
            if ((float)back.data[k] != image.data()[k].real())
                throw TestFailed("FITS image differs from the image");
    }
    bool threw = false;
    try {
        imager.m_MetaData.ReadMetaData((g + "antenna_locations.txt").c_str());
    } catch (const std::runtime_error&) {
        threw = true;
    }
    if (!threw) throw TestFailed("ReadMetaData must throw");
    threw = false;
    try {
        CPacerImagerMI355X with_meta(g + "antenna_locations.txt", 180, {});
    } catch (const std::runtime_error&) {
        threw = true;
    }
    if (!threw) throw TestFailed("a metadata file given to the constructor must throw");
    if (CObsMetadata::ux2gps(1534699078.0) != 1534699078.0 - 315964783.0)
        throw TestFailed("ux2gps");
}

int main() {
    CPacerImagerMI355X::SetDebugLevel(IMAGER_WARNING_LEVEL);
    try {
        RUN_TEST(test_fft_shift_simple);
        RUN_TEST(test_imager_run);
        RUN_TEST(test_imager_batched_and_pols);
        RUN_TEST(test_imager_batched_average_and_uniform);
        RUN_TEST(test_error_paths);
        RUN_TEST(test_imager_kaiser_bessel);
        RUN_TEST(test_eda2_correlation_matrices);
        RUN_TEST(test_command_line_flow);
    } catch (std::exception& ex) {
        std::cerr << ex.what() << std::endl;
        return 1;
    }
    std::cout << "All tests passed." << std::endl;
    return 0;
}
