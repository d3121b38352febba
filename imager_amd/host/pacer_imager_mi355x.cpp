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

bool file_exists(const std::string& f) {
    std::ifstream s(f);
    return s.good();
}

bool is_fits(const std::string& f) {
    std::ifstream s(f, std::ios::binary);
    char head[9] = {0};
    s.read(head, 8);
    return std::strncmp(head, "SIMPLE  ", 8) == 0;
}

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

int CPacerImagerMI355X::m_ImagerDebugLevel = IMAGER_INFO_LEVEL;
int CPacerImagerMI355X::m_SaveFilesLevel = 0;
bool CPacerImagerMI355X::m_bPrintImageStatistics = false;

#define PRINTF_INFO \
    if (CPacerImagerMI355X::m_ImagerDebugLevel >= IMAGER_INFO_LEVEL) std::printf
#define PRINTF_WARNING \
    if (CPacerImagerMI355X::m_ImagerDebugLevel >= IMAGER_WARNING_LEVEL) std::printf

// ------------------------------------------------------------- utilities

void fft_shift(std::complex<float>* image, size_t image_x_side, size_t image_y_side) {
    blink_ctx* ctx = nullptr;
    if (blink_create(0, &ctx) != BLINK_OK) throw std::runtime_error("fft_shift: no GPU context");
    const size_t n = image_x_side * image_y_side;
    void* d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(std::complex<float>) * n);
    if (e == hipSuccess) e = hipMemcpy(d, image, sizeof(std::complex<float>) * n, hipMemcpyHostToDevice);
    int rc = e == hipSuccess ? blink_fft_shift_and_norm(ctx, (float*)d, (int)image_x_side,
                                                         (int)image_y_side, 1, nullptr, 1, nullptr)
                             : BLINK_ERR_HIP;
    if (rc == BLINK_OK) e = hipMemcpy(image, d, sizeof(std::complex<float>) * n, hipMemcpyDeviceToHost);
    std::string msg = rc == BLINK_OK ? "" : blink_last_error(ctx);
    if (d) (void)hipFree(d);
    blink_destroy(ctx);
    if (rc != BLINK_OK) throw std::runtime_error("fft_shift: " + msg);
    hip_check(e, "fft_shift");
}

int CAntennaList::ReadAntennaPositions(const char* filename, bool /*bConvertToXYZ*/) {
    std::ifstream f(filename ? filename : "");
    if (!f) {
        std::printf("ERROR: filename %s does not exist\n", filename ? filename : "(null)");
        return -1;
    }
    clear();
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream is(line);
        InputMapping in;
        if (!(is >> in.szAntName >> in.x >> in.y >> in.z)) continue;
        push_back(in);
    }
    return (int)size();
}

int CAntennaList::CalculateUVW(MemoryBuffer<float>& u, MemoryBuffer<float>& v,
                               MemoryBuffer<float>& w, bool /*bIncludeAutos*/) const {
    const size_t n_ant = size();
    if (n_ant == 0) return 0;
    const size_t n_bl = n_ant * (n_ant + 1) / 2;
    if (u.size() != n_bl) u.allocate(n_bl);
    if (v.size() != n_bl) v.allocate(n_bl);
    if (w.size() != n_bl) w.allocate(n_bl);
    for (size_t bl = 0; bl < n_bl; bl++) {
        const unsigned a1 = (unsigned)(-0.5 + std::sqrt(0.25 + 2.0 * (double)bl));
        const unsigned a2 = (unsigned)(bl - ((size_t)(a1 + 1) * a1) / 2);
        const InputMapping& p = (*this)[a1];
        const InputMapping& q = (*this)[a2];
        u[bl] = (float)(-(p.x - q.x));
        v[bl] = (float)(-(p.y - q.y));
        w[bl] = (float)(-(p.z - q.z));
    }
    return (int)n_bl;
}

// ----------------------------------------------------------------- class

CPacerImagerMI355X::CPacerImagerMI355X(const std::string metadata_file, int n_pixels,
                                       const std::vector<int>& flagged_antennas,
                                       bool average_images, Polarization pol_to_image,
                                       float oversampling_factor, double min_uv, double max_uv,
                                       const char* weighting, int device)
    : average_images(average_images),
      oversampling_factor(oversampling_factor),
      n_pixels(n_pixels),
      min_uv(min_uv),
      max_uv(max_uv),
      weighting(weighting ? weighting : ""),
      pol_to_image(pol_to_image) {
    const int rc = blink_create(device, &ctx_);
    if (rc != BLINK_OK) throw std::runtime_error("CPacerImagerMI355X: cannot create a GPU context");
    if (!metadata_file.empty() && file_exists(metadata_file)) {
        if (is_fits(metadata_file)) {
            PRINTF_WARNING("WARNING : %s is a FITS metadata file; set the array with "
                           "SetAntennaPositions() or SetUVW()\n", metadata_file.c_str());
        } else {
            PRINTF_INFO("INFO : reading antenna positions from %s\n", metadata_file.c_str());
            m_MetaData.m_AntennaPositions.ReadAntennaPositions(metadata_file.c_str());
        }
    }
    m_FlaggedAntennas = flagged_antennas;
    UpdateFlags();
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

void CPacerImagerMI355X::SetDebugLevel(int debug_level) { m_ImagerDebugLevel = debug_level; }
void CPacerImagerMI355X::SetFileLevel(int filesave_level) { m_SaveFilesLevel = filesave_level; }

void CPacerImagerMI355X::SetAntennaPositions(const std::vector<double>& x,
                                             const std::vector<double>& y,
                                             const std::vector<double>& z,
                                             const std::vector<double>& cable_len_delta) {
    if (x.size() != y.size() || x.size() != z.size() ||
        (!cable_len_delta.empty() && cable_len_delta.size() != x.size()))
        throw std::invalid_argument("SetAntennaPositions: arrays of different lengths");
    m_MetaData.m_AntennaPositions.clear();
    for (size_t i = 0; i < x.size(); i++) {
        InputMapping in;
        in.x = x[i];
        in.y = y[i];
        in.z = z[i];
        in.cableLenDelta = cable_len_delta.empty() ? 0.0 : cable_len_delta[i];
        m_MetaData.m_AntennaPositions.push_back(in);
    }
    uvw_given_ = false;
    cable_lengths = MemoryBuffer<double>();
    UpdateFlags();
}

void CPacerImagerMI355X::SetUVW(const float* u, const float* v, const float* w, size_t n_bl) {
    u_cpu.allocate(n_bl);
    v_cpu.allocate(n_bl);
    w_cpu.allocate(n_bl);
    std::memcpy(u_cpu.data(), u, sizeof(float) * n_bl);
    std::memcpy(v_cpu.data(), v, sizeof(float) * n_bl);
    std::memcpy(w_cpu.data(), w, sizeof(float) * n_bl);
    uvw_given_ = true;
    m_Baselines = (int)n_bl;
}

// src/pacer_imager.cpp:60-86: flagged = either antenna in the list, or an auto.
int CPacerImagerMI355X::UpdateFlags() {
    int n_ant = (int)m_MetaData.m_AntennaPositions.size();
    if (n_ant == 0 && uvw_given_) n_ant = (int)std::lround((-1.0 + std::sqrt(1.0 + 8.0 * m_Baselines)) / 2.0);
    const int n_baselines = (n_ant * (n_ant + 1)) / 2;
    m_FlaggedBaselines.assign(n_baselines, 0);
    for (int bl = 0; bl < n_baselines; ++bl) {
        const unsigned a1 = (unsigned)(-0.5 + std::sqrt(0.25 + 2.0 * bl));
        const unsigned a2 = bl - ((a1 + 1) * a1) / 2u;
        bool listed = false;
        for (int a : m_FlaggedAntennas)
            if (a == (int)a1 || a == (int)a2) {
                listed = true;
                break;
            }
        m_FlaggedBaselines[bl] = (listed || a1 == a2) ? 1 : 0;
    }
    return 0;
}

// src/pacer_imager.cpp:210-229
bool CPacerImagerMI355X::CalculateUVW() {
    if (!uvw_given_)
        m_Baselines = m_MetaData.m_AntennaPositions.CalculateUVW(u_cpu, v_cpu, w_cpu, m_bIncludeAutos);
    u_max = -std::numeric_limits<double>::infinity();
    v_max = -std::numeric_limits<double>::infinity();
    for (size_t i = 0; i < u_cpu.size() && i < m_FlaggedBaselines.size(); i++) {
        if (m_FlaggedBaselines[i] == 0) {
            if (u_max < std::abs(u_cpu[i])) u_max = std::abs(u_cpu[i]);
            if (v_max < std::abs(v_cpu[i])) v_max = std::abs(v_cpu[i]);
        }
    }
    u_min = -u_max;
    v_min = -v_max;
    return m_Baselines > 0;
}

// src/pacer_imager.cpp:528-555
double CPacerImagerMI355X::get_frequency_hz(const Visibilities& vis, int fine_channel,
                                            bool cotter_compatible) {
    const double fine_ch_bw = vis.obsInfo.frequencyResolution * vis.nAveragedChannels;
    const double coarse_ch_bw = vis.obsInfo.coarseChannelBandwidth;
    const double coarse_central = vis.obsInfo.coarseChannel * coarse_ch_bw;
    double mhz = coarse_central - coarse_ch_bw / 2.00 + fine_ch_bw * fine_channel;
    if (!cotter_compatible) mhz += fine_ch_bw / 2.00;
    return mhz * 1e6;
}

void CPacerImagerMI355X::UpdateAntennaFlags(int n_ant) {
    const size_t n_bl = (size_t)n_ant * (n_ant + 1) / 2;
    if (m_FlaggedBaselines.size() != n_bl) UpdateFlags();
    if (m_FlaggedBaselines.size() != n_bl)
        throw std::invalid_argument("the array has " + std::to_string(m_FlaggedBaselines.size()) +
                                    " baselines but the visibilities have " + std::to_string(n_bl));
    if (baseline_flags_gpu.size() != n_bl) baseline_flags_gpu.allocate(n_bl, MemoryType::DEVICE);
    hip_check(hipMemcpy(baseline_flags_gpu.data(), m_FlaggedBaselines.data(), sizeof(int) * n_bl,
                        hipMemcpyHostToDevice), "flags H2D");
}

void CPacerImagerMI355X::upload_inputs(const Visibilities& xcorr) {
    const int n_ant = (int)xcorr.obsInfo.nAntennas;
    const size_t n_bl = (size_t)n_ant * (n_ant + 1) / 2;
    if (u_cpu.size() != n_bl)
        throw std::invalid_argument("UVW are given for " + std::to_string(u_cpu.size()) +
                                    " baselines but the visibilities have " + std::to_string(n_bl));
    UpdateAntennaFlags(n_ant);
    auto up = [&](MemoryBuffer<float>& d, MemoryBuffer<float>& h) {
        if (d.size() != h.size()) d.allocate(h.size(), MemoryType::DEVICE);
        hip_check(hipMemcpy(d.data(), h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice), "uvw H2D");
    };
    up(u_gpu, u_cpu);
    up(v_gpu, v_cpu);
    up(w_gpu, w_cpu);
    if (frequencies_gpu.size() != frequencies.size())
        frequencies_gpu.allocate(frequencies.size(), MemoryType::DEVICE);
    hip_check(hipMemcpy(frequencies_gpu.data(), frequencies.data(), sizeof(double) * frequencies.size(),
                        hipMemcpyHostToDevice), "frequencies H2D");
}

// src/pacer_imager.cpp:445-526 (fix_metafits and the metafits UVW are outside
// this library; see the header).
void CPacerImagerMI355X::grid(Visibilities& xcorr) {
    UpdateFlags();
    CalculateUVW();
    if (frequencies.size() != xcorr.nFrequencies) frequencies.allocate(xcorr.nFrequencies);
    for (size_t ch = 0; ch < xcorr.nFrequencies; ch++)
        frequencies[ch] = m_fFrequencyMHz > 0 ? m_fFrequencyMHz * 1e6
                                              : get_frequency_hz(xcorr, (int)ch, true);
    if (apply_geom_correction) ApplyGeometricCorrections(xcorr, w_cpu, frequencies);
    if (apply_cable_correction) {
        const size_t n_ant = xcorr.obsInfo.nAntennas;
        if (cable_lengths.size() != n_ant) {
            cable_lengths.allocate(n_ant);
            for (size_t a = 0; a < n_ant; a++)
                cable_lengths[a] = a < m_MetaData.m_AntennaPositions.size()
                                       ? m_MetaData.m_AntennaPositions[a].cableLenDelta
                                       : 0.0;
        }
        ApplyCableCorrections(xcorr, cable_lengths, frequencies);
    }
    n_gridded_intervals = xcorr.integration_intervals();
    n_gridded_channels = xcorr.nFrequencies;
    delta_u = oversampling_factor * (u_max) / n_pixels;
    delta_v = oversampling_factor * (v_max) / n_pixels;
    pixsize_in_radians = 1.00 / (oversampling_factor * u_max);
    PRINTF_INFO("delta_u = %.8f (u_max = %.8f), delta_v = %.8f (v_max = %.8f)\n", delta_u, u_max,
                delta_v, v_max);
    gridding(xcorr);
}

Images CPacerImagerMI355X::run(Visibilities& xcorr) {
    grid(xcorr);
    return image(xcorr.obsInfo);
}

void CPacerImagerMI355X::gridding(Visibilities& xcorr) {
    int kernel_mode = 0;
    if (m_nConvolvingKernelSize > 0) {
        if (_gridMode == GridMode::KaiserBesselWithoutSinc)
            kernel_mode = BLINK_KERNEL_KAISER_BESSEL_NO_SINC;
        else if (_gridMode == GridMode::KaiserBesselKernel ||
                 _gridMode == GridMode::NearestNeighbourGridding)
            kernel_mode = BLINK_KERNEL_KAISER_BESSEL;
        else
            throw std::runtime_error("only the Kaiser-Bessel gridding kernels are implemented");
    }
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
    if (w_gpu.size() != n_bl) w_gpu.allocate(n_bl, MemoryType::DEVICE);
    hip_check(hipMemcpy(w_gpu.data(), w.data(), sizeof(float) * n_bl, hipMemcpyHostToDevice), "w H2D");
    if (frequencies_gpu.size() != freqs.size()) frequencies_gpu.allocate(freqs.size(), MemoryType::DEVICE);
    hip_check(hipMemcpy(frequencies_gpu.data(), freqs.data(), sizeof(double) * freqs.size(),
                        hipMemcpyHostToDevice), "frequencies H2D");
    check(blink_apply_geometric_corrections(ctx_, (float*)xcorr.data(), (int)xcorr.obsInfo.nAntennas,
                                            (int)xcorr.nFrequencies, (int)xcorr.integration_intervals(),
                                            w_gpu.data(), frequencies_gpu.data(), nullptr));
}

void CPacerImagerMI355X::ApplyCableCorrections(Visibilities& xcorr, MemoryBuffer<double>& cables,
                                               MemoryBuffer<double>& freqs) {
    xcorr.to_gpu();
    if (cables.size() != xcorr.obsInfo.nAntennas)
        throw std::invalid_argument("cable lengths have the wrong number of antennas");
    if (cable_lengths_gpu.size() != cables.size())
        cable_lengths_gpu.allocate(cables.size(), MemoryType::DEVICE);
    hip_check(hipMemcpy(cable_lengths_gpu.data(), cables.data(), sizeof(double) * cables.size(),
                        hipMemcpyHostToDevice), "cables H2D");
    if (frequencies_gpu.size() != freqs.size()) frequencies_gpu.allocate(freqs.size(), MemoryType::DEVICE);
    hip_check(hipMemcpy(frequencies_gpu.data(), freqs.data(), sizeof(double) * freqs.size(),
                        hipMemcpyHostToDevice), "frequencies H2D");
    check(blink_apply_cable_corrections(ctx_, (float*)xcorr.data(), (int)xcorr.obsInfo.nAntennas,
                                        (int)xcorr.nFrequencies, (int)xcorr.integration_intervals(),
                                        cable_lengths_gpu.data(), frequencies_gpu.data(), nullptr));
}

Images CPacerImagerMI355X::run_batched(Visibilities& xcorr) {
    if (std::strcmp(weighting, "U") == 0)
        throw std::invalid_argument("run_batched images with natural weighting; use run() for \"U\"");
    if (m_nConvolvingKernelSize > 0)
        throw std::invalid_argument("run_batched grids to the nearest cell; use run() for "
                                    "convolutional gridding");
    UpdateFlags();
    CalculateUVW();
    if (frequencies.size() != xcorr.nFrequencies) frequencies.allocate(xcorr.nFrequencies);
    for (size_t ch = 0; ch < xcorr.nFrequencies; ch++)
        frequencies[ch] = m_fFrequencyMHz > 0 ? m_fFrequencyMHz * 1e6
                                              : get_frequency_hz(xcorr, (int)ch, true);
    const size_t n_ant = xcorr.obsInfo.nAntennas;
    if (cable_lengths.size() != n_ant) {
        cable_lengths.allocate(n_ant);
        for (size_t a = 0; a < n_ant; a++)
            cable_lengths[a] = a < m_MetaData.m_AntennaPositions.size()
                                   ? m_MetaData.m_AntennaPositions[a].cableLenDelta
                                   : 0.0;
    }
    n_gridded_intervals = xcorr.integration_intervals();
    n_gridded_channels = xcorr.nFrequencies;
    delta_u = oversampling_factor * (u_max) / n_pixels;
    delta_v = oversampling_factor * (v_max) / n_pixels;
    pixsize_in_radians = 1.00 / (oversampling_factor * u_max);
    upload_inputs(xcorr);
    if (cable_lengths_gpu.size() != n_ant) cable_lengths_gpu.allocate(n_ant, MemoryType::DEVICE);
    hip_check(hipMemcpy(cable_lengths_gpu.data(), cable_lengths.data(), sizeof(double) * n_ant,
                        hipMemcpyHostToDevice), "cables H2D");
    xcorr.to_gpu();
    const size_t n_img = n_gridded_channels * n_gridded_intervals;
    MemoryBuffer<std::complex<float>> images_buffer((size_t)n_pixels * n_pixels * n_img,
                                                    MemoryType::DEVICE);
    blink_image_set st{};
    st.vis = (const float*)xcorr.data();
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
    st.images = (float*)images_buffer.data();
    st.count_sums = nullptr;
    st.transform = m_nBatchTransform;
    check(blink_image_batch(ctx_, &st, 1, nullptr));
    const double pixscale_ra = (1.00 / (oversampling_factor * u_max)) * (180.00 / kPi);
    const double pixscale_dec = (1.00 / (oversampling_factor * v_max)) * (180.00 / kPi);
    if (average_images) {
        // the average of every image, as image() returns it (complex, the
        // GPU path's averaging, src/gpu/gpu_utils.cpp:143-161)
        MemoryBuffer<std::complex<float>> avg((size_t)n_pixels * n_pixels, MemoryType::DEVICE);
        check(blink_image_averaging(ctx_, (const float*)images_buffer.data(),
                                    (int64_t)n_pixels * n_pixels, (int)n_img, 0,
                                    (float*)avg.data(), nullptr));
        return Images{std::move(avg), xcorr.obsInfo, 1, 1, (unsigned)n_pixels,
                      m_MetaData.raHrs * 15.0, m_MetaData.decDegs, pixscale_ra, pixscale_dec};
    }
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
