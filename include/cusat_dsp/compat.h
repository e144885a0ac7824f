/*
 * Complex baseband sample type of the cusat_dsp C++ API (the type the
 * reference's headers use, lib/cusat_dsp/include/cusat_dsp/compat.h:18).
 */
#ifndef CUSAT_AMD_DSP_COMPAT_H
#define CUSAT_AMD_DSP_COMPAT_H

#include <complex>

typedef std::complex<float> complexf;

#endif
