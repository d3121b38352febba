// CPacerImagerMI355X: the reference's CPacerImager (src/pacer_imager.h:30-172)
// on the MI355X library (include/blink_mi355x.h).
//
// It derives from CPacerImager (pacer_imager.h: the same constructor, public
// members and orchestration) and implements on the GPU, through the C ABI,
// the four virtuals that the reference's GPU class CPacerImagerHip overrides
// (src/gpu/pacer_imager_hip.h:46-54):
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
//   * metadata: see pacer_imager.h (antenna list, SetAntennaPositions, SetUVW;
//     no metafits);
//   * counts are kept per image, as the reference's CPU path does
//     (src/pacer_imager.cpp:387,408; its GPU path counts timestep 0 only);
//   * the geometric correction uses w[bl] (src/gpu/corrections_gpu.cpp:61).
#ifndef PACER_IMAGER_MI355X_H
#define PACER_IMAGER_MI355X_H

#include <complex>
#include <limits>
#include <string>
#include <vector>

#include "../../include/blink_mi355x.h"
#include "astroio.hpp"
#include "gridding.hpp"
#include "memory_buffer.hpp"
#include "pacer_imager.h"

// The 2-D primary image of a FITS file (the part of msfitslib's CBgFits the
// imager's inputs use): BITPIX 8, 16, 32, 64, -32 or -64, NAXIS = 2, BSCALE /
// BZERO applied.  data[y * nx + x]; getXY(x, y) as CBgFits::getXY.
struct FitsImage {
    long nx = 0, ny = 0;
    std::vector<double> data;
    double getXY(long x, long y) const { return data[(size_t)y * (size_t)nx + (size_t)x]; }
    long GetXSize() const { return nx; }
    long GetYSize() const { return ny; }
};

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

class CPacerImagerMI355X : public CPacerImager {
   protected:
    // device copies of the per-call inputs; grids and grids_counters (base)
    // live in device memory
    MemoryBuffer<float> u_gpu, v_gpu, w_gpu;
    MemoryBuffer<int> baseline_flags_gpu;
    MemoryBuffer<double> frequencies_gpu, cable_lengths_gpu;
    // the host values of the last upload of each (upload_changed)
    std::vector<float> u_up_, v_up_, w_up_;
    std::vector<int> flags_up_;
    std::vector<double> freq_up_, cable_up_;
    blink_ctx* ctx_{nullptr};
    // run_batched(): a device copy of the visibilities when the caller's are
    // on the host (reused while the shape is unchanged)
    MemoryBuffer<std::complex<float>> vis_gpu_;

    void check(int rc) const;
    int gridding_kernel_mode() const;
    void UpdateAntennaFlags(int n_ant);
    void upload_inputs(const Visibilities& xcorr);

   public:
    // 2-D transform of run_batched(): BLINK_TRANSFORM_FFT (in-house FP32 FFT),
    // or the split-precision DFT-as-GEMM on MFMA for n_pixels <= 512,
    // BLINK_TRANSFORM_DFT_BF16X3 / BLINK_TRANSFORM_DFT_FP16X3 (no reference
    // counterpart; the reference always uses FFTW / hipFFT)
    int m_nBatchTransform{BLINK_TRANSFORM_FFT};

    // The reference's constructor (src/pacer_imager.h:98-100) plus the HIP
    // device to run on.
    CPacerImagerMI355X(const std::string metadata_file, int n_pixels,
                       const std::vector<int>& flagged_antennas, bool average_images = false,
                       Polarization pol_to_image = Polarization::XX,
                       float oversampling_factor = 2.0f,
                       double min_uv = -std::numeric_limits<double>::infinity(),
                       double max_uv = std::numeric_limits<double>::infinity(),
                       const char* weighting = "", int device = 0);
    ~CPacerImagerMI355X() override;

    void gridding(Visibilities& xcorr) override;
    Images image(ObservationInfo& obsInfo) override;
    void ApplyGeometricCorrections(Visibilities& xcorr, MemoryBuffer<float>& w,
                                   MemoryBuffer<double>& frequencies) override;
    void ApplyCableCorrections(Visibilities& xcorr, MemoryBuffer<double>& cable_lengths,
                               MemoryBuffer<double>& frequencies) override;

    // Every (timestep, channel) of xcorr imaged in one fused call
    // (blink_image_batch): corrections applied on the fly as enabled by
    // apply_*_correction, the weighting, and, with average_images, only the
    // mean image (the per-image images are not formed).  The visibilities
    // are not modified (host visibilities are copied to the device, not
    // moved).  Images stay in device memory.  Convolutional gridding
    // (m_nConvolvingKernelSize > 0) runs with natural weighting and the FFT
    // transform; other combinations throw std::invalid_argument.
    Images run_batched(Visibilities& xcorr);

    blink_ctx* context() const { return ctx_; }
};

#endif
