// The astroio containers of the imaging path's call surface, under astroio's
// header name (astroio.hpp): ObservationInfo (the fields set at
// apps/pacer_imager_main.cpp:738-748), VCS_OBSERVATION_INFO, and Visibilities
// (at(t, ch, bl), at(t, ch, a1, a2), nFrequencies, integration_intervals(),
// obsInfo, nIntegrationSteps, nAveragedChannels and the constructor of
// apps/pacer_imager_main.cpp:762).  astroio is an un-vendored dependency of
// the reference imager; a build uses either these headers or astroio's,
// never both.  Including this header also declares Images (images.hpp).
#ifndef BLINK_ASTROIO_HPP
#define BLINK_ASTROIO_HPP

#include <complex>
#include <cstdint>
#include <string>

#include "memory_buffer.hpp"

struct ObservationInfo {
    unsigned int nAntennas = 0;
    unsigned int nFrequencies = 0;
    unsigned int nPolarizations = 2;
    unsigned int nTimesteps = 0;
    double timeResolution = 0.0;          // s
    double frequencyResolution = 0.0;     // MHz (fine channel)
    double coarseChannelBandwidth = 0.0;  // MHz
    unsigned int coarseChannel = 0;
    double startTime = 0.0;               // unix seconds
    unsigned int coarse_channel_index = 0;
    std::string metadata_file;
};

// Brace initialiser of an ObservationInfo with the values the imager's
// command line gives a single correlation matrix
// (apps/pacer_imager_main.cpp:736-748): ObservationInfo info{VCS_OBSERVATION_INFO};
#define VCS_OBSERVATION_INFO 0u, 1u, 2u, 1u, 1.0, 0.040, 1.28, 109u, 1534699078.0, 0u, ""

// Cross-correlations [t][ch][bl][4 pol products], bl = a1(a1+1)/2 + a2, a2 <= a1.
class Visibilities : public MemoryBuffer<std::complex<float>> {
   public:
    ObservationInfo obsInfo;
    unsigned int nIntegrationSteps = 1;
    unsigned int nAveragedChannels = 1;
    unsigned int nFrequencies = 0;

    Visibilities() = default;
    Visibilities(MemoryBuffer<std::complex<float>>&& data, const ObservationInfo& info,
                 unsigned int n_integration_steps, unsigned int n_averaged_channels)
        : MemoryBuffer<std::complex<float>>(std::move(data)),
          obsInfo(info),
          nIntegrationSteps(n_integration_steps),
          nAveragedChannels(n_averaged_channels),
          nFrequencies(info.nFrequencies / (n_averaged_channels ? n_averaged_channels : 1)) {}

    size_t n_baselines() const { return (size_t)obsInfo.nAntennas * (obsInfo.nAntennas + 1) / 2; }
    size_t integration_intervals() const {
        const unsigned s = nIntegrationSteps ? nIntegrationSteps : 1;
        return (obsInfo.nTimesteps + s - 1) / s;
    }
    std::complex<float>* at(size_t t, size_t ch, size_t bl) {
        return data() + ((t * nFrequencies + ch) * n_baselines() + bl) * 4;
    }
    // a2 <= a1 (lower triangle); nullptr otherwise.
    std::complex<float>* at(size_t t, size_t ch, size_t a1, size_t a2) {
        if (a2 > a1) return nullptr;
        return at(t, ch, a1 * (a1 + 1) / 2 + a2);
    }
};

#include "images.hpp"

#endif
