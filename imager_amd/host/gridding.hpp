// Polarisation selector of the imaging call surface (the reference's
// src/gridding.hpp:4): XX, YY, Stokes I = (XX+YY)/2, Stokes V = (i/2)(XY-YX).
#ifndef __GRIDDING_H__
#define __GRIDDING_H__

enum class Polarization { XX, YY, I, V };

// Interpolation of visibilities onto the uv grid (the reference's GridMode,
// src/pacer_imager_defs.h:99-137).  Only the Kaiser-Bessel kernels are
// implemented, as in the reference (src/pacer_imager.cpp:617-623), where
// NearestNeighbourGridding also selects the Kaiser-Bessel table; the
// kernel is used when m_nConvolvingKernelSize > 0.
enum class GridMode {
    NearestNeighbourGridding,
    KaiserBesselKernel,
    KaiserBesselWithoutSinc,
    RectangularKernel,
    GaussianKernel,
    GaussianKernelWithoutSinc,
    BlackmanNuttallKernel,
    BlackmanNuttallKernelWithoutSinc,
    BlackmanHarrisKernel
};

#endif
