// CPacerImager base: state, orchestration and host-side helpers of the
// imaging call surface (src/pacer_imager.cpp); the stages are implemented by
// the GPU subclass (pacer_imager_mi355x.cpp).
#include "pacer_imager.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

int CPacerImager::m_ImagerDebugLevel = IMAGER_INFO_LEVEL;
int CPacerImager::m_SaveFilesLevel = 0;
bool CPacerImager::m_bPrintImageStatistics = false;

#define PRINTF_INFO \
    if (CPacerImager::m_ImagerDebugLevel >= IMAGER_INFO_LEVEL) std::printf

namespace {

bool file_exists(const std::string& f) {
    std::ifstream s(f);
    return s.good();
}

// Destination of index p under the shift: p <-> p + side/2 + side%2 for
// p < side/2 (src/pacer_imager.cpp:30-35).
size_t shifted(size_t p, size_t side) { return (p + side / 2 + side % 2) % side; }

}  // namespace

void fft_shift(std::complex<float>* image, size_t nx, size_t ny) {
    for (size_t y = 0; y < ny; y++) {
        std::complex<float>* row = image + y * nx;
        for (size_t x = 0; x < nx / 2; x++) std::swap(row[x], row[shifted(x, nx)]);
    }
    for (size_t y = 0; y < ny / 2; y++) {
        std::complex<float>* a = image + y * nx;
        std::complex<float>* b = image + shifted(y, ny) * nx;
        for (size_t x = 0; x < nx; x++) std::swap(a[x], b[x]);
    }
}

// ------------------------------------------------------------ metadata

int CAntennaPositions::ReadAntennaPositions(const char* filename, bool /*bConvertToXYZ*/) {
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

int CAntennaPositions::CalculateUVW(MemoryBuffer<float>& u, MemoryBuffer<float>& v,
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

bool CObsMetadata::ReadMetaData(const char* filename) {
    throw std::runtime_error(std::string("ReadMetaData(") + (filename ? filename : "") +
                             "): metafits and text metadata files are not read by this library; "
                             "give the array with m_AntennaPositions.ReadAntennaPositions(), "
                             "SetAntennaPositions() or SetUVW()");
}

bool CObsMetadata::fix_metafits(double obsid, double /*inttime_sec*/) {
    throw std::runtime_error("fix_metafits(" + std::to_string(obsid) +
                             "): metafits pointing corrections are not part of this library");
}

double CObsMetadata::ux2gps(double uxtime) { return uxtime - 315964783.00; }
double CObsMetadata::gps2ux(double gpstime) { return gpstime + 315964783.00; }

// ----------------------------------------------------------------- class

CPacerImager::CPacerImager(const std::string metadata_file, int n_pixels,
                           const std::vector<int>& flagged_antennas, bool average_images,
                           Polarization pol_to_image, float oversampling_factor, double min_uv,
                           double max_uv, const char* weighting)
    : average_images(average_images),
      oversampling_factor(oversampling_factor),
      n_pixels(n_pixels),
      min_uv(min_uv),
      max_uv(max_uv),
      weighting(weighting ? weighting : ""),
      pol_to_image(pol_to_image) {
    if (!metadata_file.empty() && file_exists(metadata_file)) {
        PRINTF_INFO("INFO : reading meta data from file %s\n", metadata_file.c_str());
        m_MetaData.ReadMetaData(metadata_file.c_str());
    }
    m_FlaggedAntennas = flagged_antennas;
    UpdateFlags();
}

CPacerImager::~CPacerImager() = default;

void CPacerImager::SetDebugLevel(int debug_level) { m_ImagerDebugLevel = debug_level; }
void CPacerImager::SetFileLevel(int filesave_level) { m_SaveFilesLevel = filesave_level; }

void CPacerImager::SetAntennaPositions(const std::vector<double>& x, const std::vector<double>& y,
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

void CPacerImager::SetUVW(const float* u, const float* v, const float* w, size_t n_bl) {
    u_cpu.allocate(n_bl);
    v_cpu.allocate(n_bl);
    w_cpu.allocate(n_bl);
    std::memcpy(u_cpu.data(), u, sizeof(float) * n_bl);
    std::memcpy(v_cpu.data(), v, sizeof(float) * n_bl);
    std::memcpy(w_cpu.data(), w, sizeof(float) * n_bl);
    uvw_given_ = true;
    m_Baselines = (int)n_bl;
}

int CPacerImager::UpdateFlags() {
    int n_ant = (int)m_MetaData.m_AntennaPositions.size();
    if (n_ant == 0 && uvw_given_)
        n_ant = (int)std::lround((-1.0 + std::sqrt(1.0 + 8.0 * m_Baselines)) / 2.0);
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

bool CPacerImager::CalculateUVW() {
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

double CPacerImager::get_frequency_hz(const Visibilities& vis, int fine_channel,
                                      bool cotter_compatible) {
    const double fine_ch_bw = vis.obsInfo.frequencyResolution * vis.nAveragedChannels;
    const double coarse_ch_bw = vis.obsInfo.coarseChannelBandwidth;
    const double coarse_central = vis.obsInfo.coarseChannel * coarse_ch_bw;
    double mhz = coarse_central - coarse_ch_bw / 2.00 + fine_ch_bw * fine_channel;
    if (!cotter_compatible) mhz += fine_ch_bw / 2.00;
    return mhz * 1e6;
}

void CPacerImager::fill_frequencies(const Visibilities& xcorr) {
    if (frequencies.size() != xcorr.nFrequencies) frequencies.allocate(xcorr.nFrequencies);
    for (size_t ch = 0; ch < xcorr.nFrequencies; ch++)
        frequencies[ch] = m_fFrequencyMHz > 0 ? m_fFrequencyMHz * 1e6
                                              : get_frequency_hz(xcorr, (int)ch, COTTER_COMPATIBLE);
}

void CPacerImager::fill_cable_lengths(size_t n_ant) {
    if (cable_lengths.size() == n_ant) return;
    cable_lengths.allocate(n_ant);
    for (size_t a = 0; a < n_ant; a++)
        cable_lengths[a] = a < m_MetaData.m_AntennaPositions.size()
                               ? m_MetaData.m_AntennaPositions[a].cableLenDelta
                               : 0.0;
}

void CPacerImager::set_grid_scalars(const Visibilities& xcorr) {
    n_gridded_intervals = xcorr.integration_intervals();
    n_gridded_channels = xcorr.nFrequencies;
    delta_u = oversampling_factor * (u_max) / n_pixels;
    delta_v = oversampling_factor * (v_max) / n_pixels;
    pixsize_in_radians = 1.00 / (oversampling_factor * u_max);
    PRINTF_INFO("delta_u = %.8f (u_max = %.8f), delta_v = %.8f (v_max = %.8f)\n", delta_u, u_max,
                delta_v, v_max);
}

// The pointing fix of src/pacer_imager.cpp:447 (fix_metafits) is not done:
// UVW are zenith UVW (see the header).
void CPacerImager::grid(Visibilities& xcorr) {
    UpdateFlags();
    CalculateUVW();
    fill_frequencies(xcorr);
    if (apply_geom_correction) ApplyGeometricCorrections(xcorr, w_cpu, frequencies);
    if (apply_cable_correction) {
        fill_cable_lengths(xcorr.obsInfo.nAntennas);
        ApplyCableCorrections(xcorr, cable_lengths, frequencies);
    }
    set_grid_scalars(xcorr);
    gridding(xcorr);
}

Images CPacerImager::run(Visibilities& xcorr) {
    grid(xcorr);
    return image(xcorr.obsInfo);
}
