// iolet_math.h -- iolet formulas evaluated by host code built with the
// system C++ compiler (iolet_math.cpp), so that complex division, the
// complex exponential and the power function round as in the reference's
// own build of the same expressions.
#pragma once

#include <cstdint>

namespace hlb {

// K(x) of InOutLetWomersleyVelocity::GetVelocity (src/lb/iolets/
// InOutLetWomersleyVelocity.cc:24-41): the velocity at x and step t is
// normal * -Re(K(x) exp(i omega t)), omega = 2 pi / period.
void womersley_coefficient(const double position[3], const double normal[3], double radius, double period,
                           double womersley_number, double pressure_gradient_amplitude, const double x[3],
                           double* re, double* im);
// exp(i omega t) of the same expression.
void womersley_time_factor(double period, int64_t t, double* re, double* im);

}  // namespace hlb
