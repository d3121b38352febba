// CPacerImagerMI355X: the CPacerImager call surface over the MI355X C ABI.
#include "pacer_imager_mi355x.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace {

constexpr double kPi = 3.14159265358979323846;

// 80-character FITS header card.
std::string card(const std::string& key, const std::string& value, bool quoted = false) {
    char buf[81];
    if (quoted) {
        std::string v = "'" + value;
        while (v.size() < 9) v += ' ';
        v += "'";
        std::snprintf(buf, sizeof buf, "%-8s= %-70s", key.c_str(), v.c_str());
    } else {
        std::snprintf(buf, sizeof buf, "%-8s= %20s%50s", key.c_str(), value.c_str(), "");
    }
    return std::string(buf, 80);
}

std::string num(double v) {
    char b[32];
    std::snprintf(b, sizeof b, "%.10g", v);
    return b;
}

}  // namespace

#define PRINTF_INFO \
    if (CPacerImager::m_ImagerDebugLevel >= IMAGER_INFO_LEVEL) std::printf

// ----------------------------------------------------------------- class

CPacerImagerMI355X::CPacerImagerMI355X(const std::string metadata_file, int n_pixels,
                                       const std::vector<int>& flagged_antennas,
                                       bool average_images, Polarization pol_to_image,
                                       float oversampling_factor, double min_uv, double max_uv,
                                       const char* weighting, int device)
    : CPacerImager(metadata_file, n_pixels, flagged_antennas, average_images, pol_to_image,
                   oversampling_factor, min_uv, max_uv, weighting) {
    const int rc = blink_create(device, &ctx_);
    if (rc != BLINK_OK) throw std::runtime_error("CPacerImagerMI355X: cannot create a GPU context");
}

CPacerImagerMI355X::~CPacerImagerMI355X() {
    // device buffers are released by their destructors; the context last
    grids = MemoryBuffer<std::complex<float>>();
    grids_counters = MemoryBuffer<float>();
    if (ctx_) blink_destroy(ctx_);
}

void CPacerImagerMI355X::check(int rc) const {
    if (rc == BLINK_OK) return;
    const std::string msg = blink_last_error(ctx_);
    if (rc == BLINK_ERR_INVALID_ARGUMENT || rc == BLINK_ERR_UNSUPPORTED)
        throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// Copies n host values to the device buffer unless they equal the values the
// last copy uploaded (kept in `shadow`): an unchanged input is not copied
// again, so a call does not wait (hipMemcpy) for the work still reading it.
template <typename T>
static void upload_changed(MemoryBuffer<T>& dev, std::vector<T>& shadow, const T* host, size_t n,
                           const char* what) {
    if (dev.size() == n && shadow.size() == n &&
        (n == 0 || std::memcmp(shadow.data(), host, sizeof(T) * n) == 0))
        return;
    if (dev.size() != n) dev.allocate(n, MemoryType::DEVICE);
    if (n) hip_check(hipMemcpy(dev.data(), host, sizeof(T) * n, hipMemcpyHostToDevice), what);
    shadow.assign(host, host + n);
}

void CPacerImagerMI355X::UpdateAntennaFlags(int n_ant) {
    const size_t n_bl = (size_t)n_ant * (n_ant + 1) / 2;
    if (m_FlaggedBaselines.size() != n_bl) UpdateFlags();
    if (m_FlaggedBaselines.size() != n_bl)
        throw std::invalid_argument("the array has " + std::to_string(m_FlaggedBaselines.size()) +
                                    " baselines but the visibilities have " + std::to_string(n_bl));
    upload_changed(baseline_flags_gpu, flags_up_, m_FlaggedBaselines.data(), n_bl, "flags H2D");
}

void CPacerImagerMI355X::upload_inputs(const Visibilities& xcorr) {
    const int n_ant = (int)xcorr.obsInfo.nAntennas;
    const size_t n_bl = (size_t)n_ant * (n_ant + 1) / 2;
    if (u_cpu.size() != n_bl)
        throw std::invalid_argument("UVW are given for " + std::to_string(u_cpu.size()) +
                                    " baselines but the visibilities have " + std::to_string(n_bl));
    UpdateAntennaFlags(n_ant);
    upload_changed(u_gpu, u_up_, u_cpu.data(), u_cpu.size(), "u H2D");
    upload_changed(v_gpu, v_up_, v_cpu.data(), v_cpu.size(), "v H2D");
    upload_changed(w_gpu, w_up_, w_cpu.data(), w_cpu.size(), "w H2D");
    upload_changed(frequencies_gpu, freq_up_, frequencies.data(), frequencies.size(),
                   "frequencies H2D");
}

// blink_grid_kernel of m_nConvolvingKernelSize / _gridMode, 0 for nearest-cell
// gridding (src/pacer_imager.cpp:378-381, :617-623: NearestNeighbourGridding
// also selects the Kaiser-Bessel table there).
int CPacerImagerMI355X::gridding_kernel_mode() const {
    if (m_nConvolvingKernelSize <= 0) return 0;
    if (_gridMode == GridMode::KaiserBesselWithoutSinc) return BLINK_KERNEL_KAISER_BESSEL_NO_SINC;
    if (_gridMode == GridMode::KaiserBesselKernel || _gridMode == GridMode::NearestNeighbourGridding)
        return BLINK_KERNEL_KAISER_BESSEL;
    throw std::runtime_error("only the Kaiser-Bessel gridding kernels are implemented");
}

void CPacerImagerMI355X::gridding(Visibilities& xcorr) {
    const int kernel_mode = gridding_kernel_mode();
    upload_inputs(xcorr);
    xcorr.to_gpu();
    const size_t n_img = n_gridded_channels * n_gridded_intervals;
    const size_t buffer_size = (size_t)n_pixels * n_pixels * n_img;
    if (grids.size() != buffer_size || grids_counters.size() != buffer_size) {
        grids.allocate(buffer_size, MemoryType::DEVICE);
        grids_counters.allocate(buffer_size, MemoryType::DEVICE);
        hip_check(hipMemset(grids.data(), 0, sizeof(std::complex<float>) * buffer_size), "memset");
        hip_check(hipMemset(grids_counters.data(), 0, sizeof(float) * buffer_size), "memset");
    }
    const int uniform = std::strcmp(weighting, "U") == 0 ? BLINK_WEIGHTING_UNIFORM
                                                         : BLINK_WEIGHTING_NATURAL;
    if (kernel_mode)
        check(blink_gridding_kernel(ctx_, (const float*)xcorr.data(),
                                    (int)xcorr.obsInfo.nAntennas, (int)n_gridded_channels,
                                    (int)n_gridded_intervals, u_gpu.data(), v_gpu.data(),
                                    baseline_flags_gpu.data(), frequencies_gpu.data(), delta_u,
                                    delta_v, min_uv, n_pixels, (int)pol_to_image, uniform,
                                    m_nConvolvingKernelSize, kernel_mode, (float*)grids.data(),
                                    grids_counters.data(), nullptr));
    else
        check(blink_gridding(ctx_, (const float*)xcorr.data(), (int)xcorr.obsInfo.nAntennas,
                             (int)n_gridded_channels, (int)n_gridded_intervals, u_gpu.data(),
                             v_gpu.data(), baseline_flags_gpu.data(), frequencies_gpu.data(),
                             delta_u, delta_v, min_uv, n_pixels, (int)pol_to_image, uniform,
                             BLINK_GRID_TILED, (float*)grids.data(), grids_counters.data(),
                             nullptr));
}

Images CPacerImagerMI355X::image(ObservationInfo& obs_info) {
    if (!grids_counters || !grids)
        throw std::runtime_error{"Grids are not initialised and cannot be imaged."};
    const size_t n_img = n_gridded_channels * n_gridded_intervals;
    MemoryBuffer<std::complex<float>> images_buffer((size_t)n_pixels * n_pixels * n_img,
                                                    MemoryType::DEVICE);
    check(blink_image(ctx_, (float*)grids.data(), grids_counters.data(), n_pixels, (int)n_img,
                      (float*)images_buffer.data(), 1, nullptr));
    const double ra_deg = m_MetaData.raHrs * 15.00;
    const double dec_deg = m_MetaData.decDegs;
    const double pixscale_ra = (1.00 / (oversampling_factor * u_max)) * (180.00 / kPi);
    const double pixscale_dec = (1.00 / (oversampling_factor * v_max)) * (180.00 / kPi);
    Images images{std::move(images_buffer), obs_info, (unsigned)n_gridded_intervals,
                  (unsigned)n_gridded_channels, (unsigned)n_pixels, ra_deg, dec_deg,
                  pixscale_ra, pixscale_dec};
    if (!average_images) return images;
    MemoryBuffer<std::complex<float>> avg((size_t)n_pixels * n_pixels, MemoryType::DEVICE);
    check(blink_image_averaging(ctx_, (const float*)images.data(), (int64_t)images.image_size(),
                                (int)images.size(), 0, (float*)avg.data(), nullptr));
    return Images{std::move(avg), images.obsInfo, 1, 1, images.side_size, images.ra_deg,
                  images.dec_deg, images.pixscale_ra, images.pixscale_dec};
}

void CPacerImagerMI355X::ApplyGeometricCorrections(Visibilities& xcorr, MemoryBuffer<float>& w,
                                                   MemoryBuffer<double>& freqs) {
    xcorr.to_gpu();
    const size_t n_bl = xcorr.n_baselines();
    if (w.size() != n_bl) throw std::invalid_argument("w has the wrong number of baselines");
    upload_changed(w_gpu, w_up_, w.data(), n_bl, "w H2D");
    upload_changed(frequencies_gpu, freq_up_, freqs.data(), freqs.size(), "frequencies H2D");
    check(blink_apply_geometric_corrections(ctx_, (float*)xcorr.data(), (int)xcorr.obsInfo.nAntennas,
                                            (int)xcorr.nFrequencies, (int)xcorr.integration_intervals(),
                                            w_gpu.data(), frequencies_gpu.data(), nullptr));
}

void CPacerImagerMI355X::ApplyCableCorrections(Visibilities& xcorr, MemoryBuffer<double>& cables,
                                               MemoryBuffer<double>& freqs) {
    xcorr.to_gpu();
    if (cables.size() != xcorr.obsInfo.nAntennas)
        throw std::invalid_argument("cable lengths have the wrong number of antennas");
    upload_changed(cable_lengths_gpu, cable_up_, cables.data(), cables.size(), "cables H2D");
    upload_changed(frequencies_gpu, freq_up_, freqs.data(), freqs.size(), "frequencies H2D");
    check(blink_apply_cable_corrections(ctx_, (float*)xcorr.data(), (int)xcorr.obsInfo.nAntennas,
                                        (int)xcorr.nFrequencies, (int)xcorr.integration_intervals(),
                                        cable_lengths_gpu.data(), frequencies_gpu.data(), nullptr));
}

Images CPacerImagerMI355X::run_batched(Visibilities& xcorr) {
    const int kernel_mode = gridding_kernel_mode();
    UpdateFlags();
    CalculateUVW();
    fill_frequencies(xcorr);
    const size_t n_ant = xcorr.obsInfo.nAntennas;
    fill_cable_lengths(n_ant);
    set_grid_scalars(xcorr);
    upload_inputs(xcorr);
    upload_changed(cable_lengths_gpu, cable_up_, cable_lengths.data(), n_ant, "cables H2D");
    // the visibilities are read in place on the device; host visibilities are
    // copied (run() moves them to the device, as the reference's GPU path does)
    const float* vis = (const float*)xcorr.data();
    if (!xcorr.on_gpu()) {
        if (vis_gpu_.size() != xcorr.size()) vis_gpu_.allocate(xcorr.size(), MemoryType::DEVICE);
        hip_check(hipMemcpy(vis_gpu_.data(), xcorr.data(), sizeof(std::complex<float>) * xcorr.size(),
                            hipMemcpyHostToDevice), "visibilities H2D");
        vis = (const float*)vis_gpu_.data();
    }
    const size_t n_img = n_gridded_channels * n_gridded_intervals;
    const size_t image_size = (size_t)n_pixels * n_pixels;
    // with average_images only the mean is formed, unless a DFT transform or
    // an odd grid side needs the images
    const bool keep_images =
        !average_images || m_nBatchTransform != BLINK_TRANSFORM_FFT || n_pixels % 2 != 0;
    MemoryBuffer<std::complex<float>> images_buffer(keep_images ? image_size * n_img : 0,
                                                    MemoryType::DEVICE);
    MemoryBuffer<std::complex<float>> avg(average_images ? image_size : 0, MemoryType::DEVICE);
    blink_image_set st{};
    st.vis = vis;
    st.n_ant = (int)n_ant;
    st.n_ch = (int)n_gridded_channels;
    st.n_t = (int)n_gridded_intervals;
    st.u_m = u_gpu.data();
    st.v_m = v_gpu.data();
    st.w_m = w_gpu.data();
    st.bl_flags = baseline_flags_gpu.data();
    st.freq_hz = frequencies_gpu.data();
    st.cable_m = cable_lengths_gpu.data();
    st.corrections = (apply_geom_correction ? BLINK_CORR_GEOMETRIC : 0u) |
                     (apply_cable_correction ? BLINK_CORR_CABLE : 0u);
    st.delta_u = delta_u;
    st.delta_v = delta_v;
    st.min_uv = min_uv;
    st.n_pixels = n_pixels;
    st.pol = (int)pol_to_image;
    st.images = keep_images ? (float*)images_buffer.data() : nullptr;
    st.count_sums = nullptr;
    st.transform = m_nBatchTransform;
    st.weighting = std::strcmp(weighting, "U") == 0 ? BLINK_WEIGHTING_UNIFORM
                                                    : BLINK_WEIGHTING_NATURAL;
    st.average = average_images ? (float*)avg.data() : nullptr;
    st.kernel_size = kernel_mode ? m_nConvolvingKernelSize : 0;
    st.kernel_mode = kernel_mode;
    check(blink_image_batch(ctx_, &st, 1, nullptr));
    const double pixscale_ra = (1.00 / (oversampling_factor * u_max)) * (180.00 / kPi);
    const double pixscale_dec = (1.00 / (oversampling_factor * v_max)) * (180.00 / kPi);
    if (average_images)  // the mean image, as image() returns it (src/pacer_imager.cpp:203-205)
        return Images{std::move(avg), xcorr.obsInfo, 1, 1, (unsigned)n_pixels,
                      m_MetaData.raHrs * 15.0, m_MetaData.decDegs, pixscale_ra, pixscale_dec};
    return Images{std::move(images_buffer), xcorr.obsInfo, (unsigned)n_gridded_intervals,
                  (unsigned)n_gridded_channels, (unsigned)n_pixels, m_MetaData.raHrs * 15.0,
                  m_MetaData.decDegs, pixscale_ra, pixscale_dec};
}

// ------------------------------------------------------------ FITS input

FitsImage read_fits_image(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open FITS file " + path);
    long bitpix = 0, naxis = -1, nx = 0, ny = 0;
    double bscale = 1.0, bzero = 0.0;
    bool end = false;
    size_t header_bytes = 0;
    char block[2880];
    while (!end) {
        if (!f.read(block, sizeof block)) throw std::runtime_error("truncated FITS header: " + path);
        header_bytes += sizeof block;
        for (int c = 0; c < 36 && !end; c++) {
            const std::string cardtxt(block + 80 * c, 80);
            const std::string key = cardtxt.substr(0, 8);
            if (key.compare(0, 3, "END") == 0 && key.find_first_not_of(' ', 3) == std::string::npos) {
                end = true;
                break;
            }
            if (cardtxt.compare(8, 2, "= ") != 0) continue;
            std::string val = cardtxt.substr(10);
            const size_t slash = val.find('/');
            if (slash != std::string::npos) val = val.substr(0, slash);
            auto num = [&]() { return std::strtod(val.c_str(), nullptr); };
            if (key == "BITPIX  ") bitpix = (long)num();
            else if (key == "NAXIS   ") naxis = (long)num();
            else if (key == "NAXIS1  ") nx = (long)num();
            else if (key == "NAXIS2  ") ny = (long)num();
            else if (key == "BSCALE  ") bscale = num();
            else if (key == "BZERO   ") bzero = num();
        }
    }
    if (naxis != 2 || nx <= 0 || ny <= 0)
        throw std::runtime_error("FITS file " + path + " is not a 2-D image");
    const int bytes = bitpix == 8 ? 1 : bitpix == 16 ? 2 : (bitpix == 32 || bitpix == -32) ? 4
                      : (bitpix == 64 || bitpix == -64) ? 8 : 0;
    if (!bytes) throw std::runtime_error("unsupported BITPIX in " + path);
    const size_t n = (size_t)nx * (size_t)ny;
    std::vector<unsigned char> raw(n * bytes);
    if (!f.read(reinterpret_cast<char*>(raw.data()), (std::streamsize)raw.size()))
        throw std::runtime_error("truncated FITS data: " + path);
    FitsImage img;
    img.nx = nx;
    img.ny = ny;
    img.data.resize(n);
    for (size_t i = 0; i < n; i++) {
        uint64_t b = 0;  // big-endian sample
        for (int k = 0; k < bytes; k++) b = (b << 8) | raw[i * bytes + k];
        double v;
        if (bitpix == -64) {
            std::memcpy(&v, &b, 8);
        } else if (bitpix == -32) {
            const uint32_t b32 = (uint32_t)b;
            float x;
            std::memcpy(&x, &b32, 4);
            v = x;
        } else if (bitpix == 8) {
            v = (double)(uint8_t)b;
        } else if (bitpix == 16) {
            v = (double)(int16_t)(uint16_t)b;
        } else if (bitpix == 32) {
            v = (double)(int32_t)(uint32_t)b;
        } else {
            v = (double)(int64_t)b;
        }
        img.data[i] = bzero + bscale * v;
    }
    (void)header_bytes;
    return img;
}

Visibilities visibilities_from_correlation_matrices(const FitsImage& re, const FitsImage& im,
                                                    const FitsImage* re_yy,
                                                    const FitsImage* im_yy,
                                                    ObservationInfo& info) {
    const long n_ant = re.GetXSize();
    auto square = [&](const FitsImage& m) { return m.nx == n_ant && m.ny == n_ant; };
    if (n_ant <= 0 || !square(re) || !square(im) || (re_yy && !square(*re_yy)) ||
        (im_yy && !square(*im_yy)))
        throw std::invalid_argument("correlation matrices must be n_ant x n_ant of one size");
    // the command line's defaults (apps/pacer_imager_main.cpp:736-746)
    info.nAntennas = (unsigned)n_ant;
    info.nFrequencies = 1;
    info.nPolarizations = 2;
    info.nTimesteps = 1;
    info.timeResolution = 1.0;
    info.frequencyResolution = 0.040;
    info.coarseChannelBandwidth = 1.28;
    info.coarseChannel = 109;
    info.startTime = 1534699078;
    info.coarse_channel_index = 0;
    const size_t n_bl = (size_t)n_ant * (n_ant + 1) / 2;
    MemoryBuffer<std::complex<float>> buf(n_bl * 4);
    std::memset((void*)buf.data(), 0, sizeof(std::complex<float>) * n_bl * 4);
    Visibilities xcorr(std::move(buf), info, 1, 1);
    xcorr.nFrequencies = 1;
    const bool yy = re_yy && im_yy;
    for (long i = 0; i < n_ant; i++)
        for (long j = 0; j <= i; j++) {
            std::complex<float>* rec = xcorr.at(0, 0, (size_t)i, (size_t)j);
            rec[0] = std::complex<float>((float)re.getXY(j, i), (float)im.getXY(j, i));
            if (yy) rec[3] = std::complex<float>((float)re_yy->getXY(j, i), (float)im_yy->getXY(j, i));
        }
    return xcorr;
}

// ------------------------------------------------------------ FITS output

void Images::to_fits_files(const std::string& dir) {
    to_cpu();
    const size_t n = image_size();
    for (unsigned t = 0; t < n_intervals; t++) {
        for (unsigned ch = 0; ch < n_channels; ch++) {
            std::string hdr;
            hdr += card("SIMPLE", "T");
            hdr += card("BITPIX", "-32");
            hdr += card("NAXIS", "2");
            hdr += card("NAXIS1", std::to_string(side_size));
            hdr += card("NAXIS2", std::to_string(side_size));
            hdr += card("CRPIX1", std::to_string(side_size / 2 + 1));
            hdr += card("CDELT1", num(pixscale_ra));
            hdr += card("CRVAL1", num(ra_deg));
            hdr += card("CTYPE1", "RA---SIN", true);
            hdr += card("CRPIX2", std::to_string(side_size / 2 + 1));
            hdr += card("CDELT2", num(pixscale_dec));
            hdr += card("CRVAL2", num(dec_deg));
            hdr += card("CTYPE2", "DEC--SIN", true);
            hdr += std::string("END") + std::string(77, ' ');
            hdr.resize(((hdr.size() + 2879) / 2880) * 2880, ' ');
            std::vector<unsigned char> body(n * 4);
            const std::complex<float>* img = data() + ((size_t)t * n_channels + ch) * n;
            for (size_t i = 0; i < n; i++) {
                float v = img[i].real();
                uint32_t b;
                std::memcpy(&b, &v, 4);
                body[4 * i] = (unsigned char)(b >> 24);
                body[4 * i + 1] = (unsigned char)(b >> 16);
                body[4 * i + 2] = (unsigned char)(b >> 8);
                body[4 * i + 3] = (unsigned char)b;
            }
            body.resize(((body.size() + 2879) / 2880) * 2880, 0);
            char name[256];
            std::snprintf(name, sizeof name, "%s/dirty_image_t%04u_ch%04u_real.fits", dir.c_str(), t, ch);
            std::ofstream f(name, std::ios::binary);
            f.write(hdr.data(), (std::streamsize)hdr.size());
            f.write(reinterpret_cast<const char*>(body.data()), (std::streamsize)body.size());
            if (!f) throw std::runtime_error(std::string("cannot write ") + name);
        }
    }
}
