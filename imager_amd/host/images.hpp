// Images: the astroio container of dirty images, under astroio's header name
// (images.hpp): the constructor of src/pacer_imager.cpp:199-201, image_size(),
// size(), side_size, ra/dec, pixel scales and to_fits_files()
// (apps/pacer_imager_main.cpp:578).  A build uses either these headers or
// astroio's, never both.
#ifndef BLINK_IMAGES_HPP
#define BLINK_IMAGES_HPP

#include <complex>
#include <string>

#include "astroio.hpp"

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

#endif
