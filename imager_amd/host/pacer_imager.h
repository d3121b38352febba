// CPacerImager: the imaging call surface of the reference imager
// (src/pacer_imager.h:30-172) as an abstract base class.
//
// It holds the same state and public members and runs the same orchestration
// (run -> grid -> corrections -> gridding, then image; src/pacer_imager.cpp:
// 440-526), and leaves the four stages that the reference's GPU class
// CPacerImagerHip overrides (src/gpu/pacer_imager_hip.h:46-54) pure virtual:
//   gridding(), image(), ApplyGeometricCorrections(), ApplyCableCorrections().
// This library has no CPU implementation of them: CPacerImagerMI355X
// (pacer_imager_mi355x.h) implements them on the GPU.  Code written against
// CPacerImager& -- the command line's Initialise(CPacerImager&),
// apps/pacer_imager_main.cpp:653 -- takes a CPacerImagerMI355X.
//
// Metadata: the array comes from the text antenna list
// (m_MetaData.m_AntennaPositions.ReadAntennaPositions, the format of
// config/eda2/antenna_locations.txt), SetAntennaPositions() or SetUVW().
// metafits / PAL / libnova geometry (src/observation_metadata.cpp,
// src/pacer_geometry.h) is not part of this library: ReadMetaData() and
// fix_metafits() throw std::runtime_error.
#ifndef PACER_IMAGER_H_BLINK
#define PACER_IMAGER_H_BLINK

#include <complex>
#include <limits>
#include <string>
#include <vector>

#include "astroio.hpp"
#include "gridding.hpp"
#include "memory_buffer.hpp"

#define COTTER_COMPATIBLE true

// Verbosity levels of src/pacer_imager_defs.h:24-30.
#define IMAGER_DEBUG_LEVEL 5
#define IMAGER_INFO_LEVEL 3
#define IMAGER_BENCHMARK_LEVEL 2
#define IMAGER_WARNING_LEVEL 1
#define IMAGER_ERROR_LEVEL 0

// FFT shift of one host image in place (src/pacer_imager.h:27): columns
// x <-> (x + nx/2 + nx%2) % nx for x < nx/2, then rows likewise
// (src/pacer_imager.cpp:110-128; odd sides keep the reference's swap order).
void fft_shift(std::complex<float>* image, size_t image_x_side, size_t image_y_side);

// One antenna of the array (the fields of the reference's InputMapping that
// the path uses).
struct InputMapping {
    std::string szAntName;
    double x = 0, y = 0, z = 0;   // metres
    double cableLenDelta = 0;     // metres
    int flag = 0;
};

// The array (src/antenna_positions.h): antennas in input order.
class CAntennaPositions : public std::vector<InputMapping> {
   public:
    // "name x y z" per line, '#' comments (config/eda2/antenna_locations.txt).
    // Positions are kept as given: bConvertToXYZ (the reference's local ->
    // Earth-centred rotation, src/antenna_positions.cpp:142-157) is not
    // applied, so UVW are the zenith differences of the listed positions.
    int ReadAntennaPositions(const char* filename, bool bConvertToXYZ = false);
    // Zenith UVW in metres, u = -(x_a1 - x_a2) etc. formed in double, stored as
    // float (src/antenna_positions.cpp:81-89).
    int CalculateUVW(MemoryBuffer<float>& u, MemoryBuffer<float>& v, MemoryBuffer<float>& w,
                     bool bIncludeAutos = false) const;
};

// Observation metadata (src/observation_metadata.h:26-140): the antenna list
// and the pointing.
class CObsMetadata {
   public:
    CAntennaPositions m_AntennaPositions;
    double raHrs = 0, decDegs = 0;

    // metafits / text metadata readers: not part of this library, throw
    // std::runtime_error naming the file.
    bool ReadMetaData(const char* filename);
    bool fix_metafits(double obsid, double inttime_sec = 1.00);
    // GPS seconds of a unix time (src/observation_metadata.cpp:34-37).
    static double ux2gps(double uxtime);
    static double gps2ux(double gpstime);
};

class CPacerImager {
   protected:
    MemoryBuffer<double> cable_lengths;
    MemoryBuffer<double> frequencies;
    MemoryBuffer<float> grids_counters;
    MemoryBuffer<std::complex<float>> grids;
    size_t n_gridded_channels{0};
    size_t n_gridded_intervals{0};
    MemoryBuffer<float> u_cpu;
    MemoryBuffer<float> v_cpu;
    MemoryBuffer<float> w_cpu;
    bool uvw_given_{false};

    // frequencies of every channel of xcorr (get_frequency_hz, or
    // m_fFrequencyMHz for all), src/pacer_imager.cpp:452-461
    void fill_frequencies(const Visibilities& xcorr);
    // cable lengths of the antennas from the antenna list (0 when absent)
    void fill_cable_lengths(size_t n_ant);
    // n_gridded_*, delta_u / delta_v and the pixel size, src/pacer_imager.cpp:495-504
    void set_grid_scalars(const Visibilities& xcorr);

   public:
    static int m_ImagerDebugLevel;
    static int m_SaveFilesLevel;
    static bool m_bPrintImageStatistics;

    bool m_bIncludeAutos{false};
    bool autofix_metadata{true};
    bool constant_uvw{false};
    bool apply_geom_correction{true};
    bool apply_cable_correction{true};
    bool average_images{false};
    float oversampling_factor{2.0f};

    CObsMetadata m_MetaData;
    std::vector<int> m_FlaggedAntennas;
    std::vector<int> m_FlaggedBaselines;

    double u_min{0}, u_max{0};
    double v_min{0}, v_max{0};
    double w_min{0}, w_max{0};
    double delta_u{0}, delta_v{0};

    int m_Baselines{0};
    int n_pixels;
    double min_uv{-std::numeric_limits<double>::infinity()};
    double max_uv{std::numeric_limits<double>::infinity()};
    const char* weighting = "";

    double m_PixscaleAtZenith{0};
    double pixsize_in_radians{0};

    Polarization pol_to_image{Polarization::XX};
    double m_fFrequencyMHz{-1};
    // convolutional gridding: kernel size in cells (> 0 enables it) and kernel
    // (src/pacer_imager.h:168-169)
    GridMode _gridMode{GridMode::KaiserBesselKernel};
    int m_nConvolvingKernelSize{0};

    // src/pacer_imager.cpp:90-108: reads the antenna list when metadata_file is
    // a text file (a FITS metafits file is reported, not read), then the flags.
    CPacerImager(const std::string metadata_file, int n_pixels,
                 const std::vector<int>& flagged_antennas, bool average_images = false,
                 Polarization pol_to_image = Polarization::XX, float oversampling_factor = 2.0f,
                 double min_uv = -std::numeric_limits<double>::infinity(),
                 double max_uv = std::numeric_limits<double>::infinity(),
                 const char* weighting = "");
    virtual ~CPacerImager();
    CPacerImager(const CPacerImager&) = delete;
    CPacerImager& operator=(const CPacerImager&) = delete;

    static void SetDebugLevel(int debug_level);
    static void SetFileLevel(int filesave_level);

    // Array description (in place of the metafits reader).
    void SetAntennaPositions(const std::vector<double>& x, const std::vector<double>& y,
                             const std::vector<double>& z,
                             const std::vector<double>& cable_len_delta = {});
    // Constant UVW per baseline, metres (in place of CalculateUVW's geometry).
    void SetUVW(const float* u, const float* v, const float* w, size_t n_baselines);

    // src/pacer_imager.cpp:60-86: a baseline is flagged when either antenna is
    // in m_FlaggedAntennas, or when it is an auto-correlation.
    int UpdateFlags();
    // src/pacer_imager.cpp:210-229: zenith UVW of the antenna list (unless set
    // with SetUVW) and u_max / v_max over the unflagged baselines.
    bool CalculateUVW();

    virtual void gridding(Visibilities& xcorr) = 0;
    // src/pacer_imager.cpp:440-443
    Images run(Visibilities& xcorr);
    // src/pacer_imager.cpp:445-526
    void grid(Visibilities& xcorr);
    virtual Images image(ObservationInfo& obsInfo) = 0;

    virtual void ApplyGeometricCorrections(Visibilities& xcorr, MemoryBuffer<float>& w,
                                           MemoryBuffer<double>& frequencies) = 0;
    virtual void ApplyCableCorrections(Visibilities& xcorr, MemoryBuffer<double>& cable_lengths,
                                       MemoryBuffer<double>& frequencies) = 0;

    // src/pacer_imager.cpp:528-555
    double get_frequency_hz(const Visibilities& xcorr, int fine_channel, bool cotter_compatible);
};

#endif
