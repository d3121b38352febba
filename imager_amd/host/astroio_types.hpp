// Minimal, self-contained stand-ins for the astroio containers that the
// imaging path's call surface uses (astroio is an external, un-vendored
// dependency of the reference: astroio.hpp, memory_buffer.hpp, images.hpp).
// Only the members the reference imager and its CLI touch are provided:
//   MemoryBuffer<T>: allocate / data / size / operator[] / operator bool /
//                    to_gpu / to_cpu / on_gpu / from_dump / dump
//                    (used at src/pacer_imager.cpp:253-260, src/gpu/*.cpp,
//                    tests/gridding_test.cpp:22-30)
//   ObservationInfo: the fields set at apps/pacer_imager_main.cpp:738-748
//   Visibilities:    at(t, ch, bl), at(t, ch, a1, a2), nFrequencies,
//                    integration_intervals(), obsInfo, nIntegrationSteps,
//                    nAveragedChannels, the constructor of
//                    apps/pacer_imager_main.cpp:762
//   Images:          the constructor of src/pacer_imager.cpp:199-201,
//                    image_size(), size(), side_size, ra/dec, pixscales,
//                    to_fits_files() (apps/pacer_imager_main.cpp:578)
#pragma once

#include <hip/hip_runtime_api.h>

#include <complex>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

enum class MemoryType { HOST, DEVICE };

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
class MemoryBuffer {
   public:
    MemoryBuffer() = default;
    explicit MemoryBuffer(size_t n, MemoryType type = MemoryType::HOST) { allocate(n, type); }
    MemoryBuffer(const MemoryBuffer&) = delete;
    MemoryBuffer& operator=(const MemoryBuffer&) = delete;
    MemoryBuffer(MemoryBuffer&& o) noexcept { *this = std::move(o); }
    MemoryBuffer& operator=(MemoryBuffer&& o) noexcept {
        if (this != &o) {
            release();
            ptr_ = o.ptr_;
            n_ = o.n_;
            on_gpu_ = o.on_gpu_;
            o.ptr_ = nullptr;
            o.n_ = 0;
            o.on_gpu_ = false;
        }
        return *this;
    }
    ~MemoryBuffer() { release(); }

    void allocate(size_t n, MemoryType type = MemoryType::HOST) {
        release();
        n_ = n;
        on_gpu_ = type == MemoryType::DEVICE;
        if (n == 0) return;
        if (on_gpu_) {
            hip_check(hipMalloc(reinterpret_cast<void**>(&ptr_), sizeof(T) * n), "hipMalloc");
        } else {
            ptr_ = new T[n]();
        }
    }

    T* data() { return ptr_; }
    const T* data() const { return ptr_; }
    size_t size() const { return n_; }
    explicit operator bool() const { return ptr_ != nullptr; }
    bool on_gpu() const { return on_gpu_; }
    T& operator[](size_t i) { return ptr_[i]; }
    const T& operator[](size_t i) const { return ptr_[i]; }

    // Moves the contents to device memory (no-op when already there).
    void to_gpu() {
        if (on_gpu_ || !ptr_) return;
        T* d = nullptr;
        hip_check(hipMalloc(reinterpret_cast<void**>(&d), sizeof(T) * n_), "hipMalloc");
        hip_check(hipMemcpy(d, ptr_, sizeof(T) * n_, hipMemcpyHostToDevice), "hipMemcpy H2D");
        delete[] ptr_;
        ptr_ = d;
        on_gpu_ = true;
    }

    // Moves the contents to host memory (no-op when already there).
    void to_cpu() {
        if (!on_gpu_ || !ptr_) return;
        T* h = new T[n_];
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hip_check(hipMemcpy(h, ptr_, sizeof(T) * n_, hipMemcpyDeviceToHost), "hipMemcpy D2H");
        hip_check(hipFree(ptr_), "hipFree");
        ptr_ = h;
        on_gpu_ = false;
    }

    static MemoryBuffer from_dump(const std::string& filename) {
        std::ifstream f(filename, std::ios::binary | std::ios::ate);
        if (!f) throw std::runtime_error("cannot open " + filename);
        const std::streamsize bytes = f.tellg();
        f.seekg(0);
        MemoryBuffer b(static_cast<size_t>(bytes) / sizeof(T));
        f.read(reinterpret_cast<char*>(b.data()), bytes);
        return b;
    }

    void dump(const std::string& filename) const {
        if (on_gpu_) throw std::runtime_error("dump: buffer is on the GPU, call to_cpu() first");
        std::ofstream f(filename, std::ios::binary);
        f.write(reinterpret_cast<const char*>(ptr_), sizeof(T) * n_);
        if (!f) throw std::runtime_error("cannot write " + filename);
    }

   private:
    void release() {
        if (!ptr_) return;
        if (on_gpu_) (void)hipFree(ptr_);
        else delete[] ptr_;
        ptr_ = nullptr;
        n_ = 0;
    }
    T* ptr_ = nullptr;
    size_t n_ = 0;
    bool on_gpu_ = false;
};

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

// Dirty images [t][ch][y][x] complex64 plus the sky metadata.
class Images : public MemoryBuffer<std::complex<float>> {
   public:
    ObservationInfo obsInfo;
    unsigned int n_intervals = 0, n_channels = 0, side_size = 0;
    double ra_deg = 0, dec_deg = 0, pixscale_ra = 0, pixscale_dec = 0;

    Images() = default;
    Images(MemoryBuffer<std::complex<float>>&& data, const ObservationInfo& info,
           unsigned int n_int, unsigned int n_ch, unsigned int side, double ra, double dec,
           double psx, double psy)
        : MemoryBuffer<std::complex<float>>(std::move(data)),
          obsInfo(info),
          n_intervals(n_int),
          n_channels(n_ch),
          side_size(side),
          ra_deg(ra),
          dec_deg(dec),
          pixscale_ra(psx),
          pixscale_dec(psy) {}

    size_t image_size() const { return (size_t)side_size * side_size; }
    size_t size() const { return (size_t)n_intervals * n_channels; }
    std::complex<float>* image(size_t t, size_t ch) { return data() + (t * n_channels + ch) * image_size(); }

    // One FITS file per image with the real part as a 2-D float32 array and
    // a sine-projection WCS (CRPIX = N/2 + 1, CDELT = pixel scale in degrees).
    // Writes <dir>/dirty_image_t<t>_ch<ch>_real.fits.  Copies to the host.
    void to_fits_files(const std::string& dir);
};
