// CPacerImagerMI355X: the imaging call surface of the reference's CPacerImager
// (src/pacer_imager.h:30-172) on the MI355X library (include/blink_mi355x.h).
//
// Same constructor, public members and methods as CPacerImager; the four
// virtuals that the reference's GPU class CPacerImagerHip overrides
// (src/gpu/pacer_imager_hip.h:46-54) run on the GPU through the C ABI:
//   gridding()                  -> blink_gridding (LDS-tiled, per-image counts), or
//                                  blink_gridding_kernel when m_nConvolvingKernelSize > 0
//   image()                     -> blink_image (+ blink_image_averaging)
//   ApplyGeometricCorrections() -> blink_apply_geometric_corrections (in place)
//   ApplyCableCorrections()     -> blink_apply_cable_corrections (in place)
// plus run_batched(), which images every (timestep, channel) of a
// Visibilities object in one fused, graph-replayed call (blink_image_batch)
// without modifying the visibilities.
//
// Differences from the reference that a caller can observe:
//   * metadata: antenna positions and cable lengths come from
//     SetAntennaPositions(), ReadAntennaPositions() (the text format of
//     config/eda2/antenna_locations.txt, local east/north/up metres) or SetUVW();
//     metafits / PAL / libnova based UVW (src/observation_metadata.cpp,
//     src/pacer_geometry.h) is outside this library;
//   * UVW are zenith UVW of the positions as given (the reference converts
//     text positions to Earth-centred XYZ first, src/antenna_positions.cpp:142-157);
//   * counts are kept per image, as the reference's CPU path does
//     (src/pacer_imager.cpp:387,408; its GPU path counts timestep 0 only);
//   * the geometric correction uses w[bl] (src/gpu/corrections_gpu.cpp:61);
//   * run_batched() images with natural weighting and nearest-cell gridding
//     only; it throws std::invalid_argument for weighting "U" or
//     m_nConvolvingKernelSize > 0 (run() handles both).
#ifndef PACER_IMAGER_MI355X_H
#define PACER_IMAGER_MI355X_H

#include <complex>
#include <limits>
#include <string>
#include <vector>

#include "../../include/blink_mi355x.h"
#include "astroio_types.hpp"
#include "gridding.hpp"

// Verbosity levels of src/pacer_imager_defs.h:24-30.
#define IMAGER_DEBUG_LEVEL 5
#define IMAGER_INFO_LEVEL 3
#define IMAGER_BENCHMARK_LEVEL 2
#define IMAGER_WARNING_LEVEL 1
#define IMAGER_ERROR_LEVEL 0

// FFT shift of one host image (src/pacer_imager.h:27), done on the GPU.
void fft_shift(std::complex<float>* image, size_t image_x_side, size_t image_y_side);

// Reads the 2-D primary image of a FITS file; throws std::runtime_error on a
// missing file or an unsupported layout.
FitsImage read_fits_image(const std::string& path);

// Visibilities of one channel and one timestep from correlation matrices, as
// the imager's command line builds them (ConvertFits2XCorr,
// apps/pacer_imager_main.cpp:720-800): n_ant = the matrices' side; for every
// antenna pair j <= i, XX = re.getXY(j, i) + i im.getXY(j, i) and YY likewise
// from re_yy / im_yy when given (else 0); XY = YX = 0.  `info` gets the
// command line's defaults for the fields the matrices do not carry.
Visibilities visibilities_from_correlation_matrices(const FitsImage& re, const FitsImage& im,
                                                    const FitsImage* re_yy,
                                                    const FitsImage* im_yy,
                                                    ObservationInfo& info);

// One antenna of the array (the fields of the reference's InputMapping that
// the path uses).
struct InputMapping {
    std::string szAntName;
    double x = 0, y = 0, z = 0;   // metres
    double cableLenDelta = 0;     // metres
    int flag = 0;
};

class CAntennaList : public std::vector<InputMapping> {
   public:
    // "name x y z" per line, '#' comments (config/eda2/antenna_locations.txt).
    int ReadAntennaPositions(const char* filename, bool bConvertToXYZ = false);
    // Zenith UVW in metres, u = -(x_a1 - x_a2) etc. (src/antenna_positions.cpp:81-89).
    int CalculateUVW(MemoryBuffer<float>& u, MemoryBuffer<float>& v, MemoryBuffer<float>& w,
                     bool bIncludeAutos = false) const;
};

struct CObsMetadataLite {
    CAntennaList m_AntennaPositions;
    double raHrs = 0, decDegs = 0;
};

class CPacerImagerMI355X {
   protected:
    MemoryBuffer<double> cable_lengths;
    MemoryBuffer<double> frequencies;
    MemoryBuffer<float> grids_counters;              // device
    MemoryBuffer<std::complex<float>> grids;         // device
    size_t n_gridded_channels{0};
    size_t n_gridded_intervals{0};
    MemoryBuffer<float> u_cpu;
    MemoryBuffer<float> v_cpu;
    MemoryBuffer<float> w_cpu;

    // device copies of the per-call inputs
    MemoryBuffer<float> u_gpu, v_gpu, w_gpu;
    MemoryBuffer<int> baseline_flags_gpu;
    MemoryBuffer<double> frequencies_gpu, cable_lengths_gpu;
    bool uvw_given_{false};
    blink_ctx* ctx_{nullptr};

    void check(int rc) const;
    void UpdateAntennaFlags(int n_ant);
    void upload_inputs(const Visibilities& xcorr);

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

    CObsMetadataLite m_MetaData;
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
    // 2-D transform of run_batched(): BLINK_TRANSFORM_FFT (in-house FP32 FFT),
    // or the split-precision DFT-as-GEMM on MFMA for n_pixels <= 512,
    // BLINK_TRANSFORM_DFT_BF16X3 / BLINK_TRANSFORM_DFT_FP16X3 (no reference
    // counterpart; the reference always uses FFTW / hipFFT)
    int m_nBatchTransform{BLINK_TRANSFORM_FFT};

    CPacerImagerMI355X(const std::string metadata_file, int n_pixels,
                       const std::vector<int>& flagged_antennas, bool average_images = false,
                       Polarization pol_to_image = Polarization::XX,
                       float oversampling_factor = 2.0f,
                       double min_uv = -std::numeric_limits<double>::infinity(),
                       double max_uv = std::numeric_limits<double>::infinity(),
                       const char* weighting = "", int device = 0);
    virtual ~CPacerImagerMI355X();
    CPacerImagerMI355X(const CPacerImagerMI355X&) = delete;
    CPacerImagerMI355X& operator=(const CPacerImagerMI355X&) = delete;

    static void SetDebugLevel(int debug_level);
    static void SetFileLevel(int filesave_level);

    // Array description (replaces the metafits reader).
    void SetAntennaPositions(const std::vector<double>& x, const std::vector<double>& y,
                             const std::vector<double>& z,
                             const std::vector<double>& cable_len_delta = {});
    // Constant UVW per baseline, metres (replaces CalculateUVW's geometry).
    void SetUVW(const float* u, const float* v, const float* w, size_t n_baselines);

    int UpdateFlags();
    bool CalculateUVW();

    virtual void gridding(Visibilities& xcorr);
    Images run(Visibilities& xcorr);
    void grid(Visibilities& xcorr);
    virtual Images image(ObservationInfo& obsInfo);

    virtual void ApplyGeometricCorrections(Visibilities& xcorr, MemoryBuffer<float>& w,
                                           MemoryBuffer<double>& frequencies);
    virtual void ApplyCableCorrections(Visibilities& xcorr, MemoryBuffer<double>& cable_lengths,
                                       MemoryBuffer<double>& frequencies);

    double get_frequency_hz(const Visibilities& xcorr, int fine_channel, bool cotter_compatible);

    // Every (timestep, channel) of xcorr imaged in one fused call: corrections
    // applied on the fly as enabled by apply_*_correction, natural weighting.
    // The visibilities are not modified.  Images stay in device memory.
    Images run_batched(Visibilities& xcorr);

    blink_ctx* context() const { return ctx_; }
};

#endif
