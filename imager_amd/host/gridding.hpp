// Polarisation selector of the imaging call surface (the reference's
// src/gridding.hpp:4): XX, YY, Stokes I = (XX+YY)/2, Stokes V = (i/2)(XY-YX).
#ifndef __GRIDDING_H__
#define __GRIDDING_H__

enum class Polarization { XX, YY, I, V };

#endif
