// iolet_math.cpp -- host-side iolet formulas built with the system C++
// compiler (not the HIP compiler): the Womersley profile's complex arithmetic
// and the file iolets' tables (hlb_iolet_file_table).
#include "iolet_math.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "../../include/hemepure_amd.h"
#include "errors.h"

namespace hlb {
namespace {

constexpr double kPiM = 3.14159265358979323846264338327950288;  // src/constants.h PI
constexpr double kCs2M = 1.0 / 3.0;
constexpr double kMmHgToPa = 133.3223874;   // src/constants.h:24
constexpr double kBloodDensity = 1000.0;    // src/constants.h:25
constexpr double kReferencePressure = 0.0;  // src/constants.h:23

typedef std::complex<double> Complex;

// util::BesselJ0ComplexArgument (src/util/Bessel.cc:18-45): the power series
// summed until a term's squared modulus is at most tolSq = 1e-12.
Complex BesselJ0(const Complex& z) {
  const double tolSq = 1e-12;
  Complex sum(1.0, 0.0);
  double fact = 1;
  Complex zSqOver4_pow(1.0, 0.0);
  Complex term(1.0, 0.0);
  for (unsigned i = 1; term.real() * term.real() + term.imag() * term.imag() > tolSq; ++i) {
    fact *= i;
    zSqOver4_pow *= -0.25 * z * z;
    term = zSqOver4_pow / (fact * fact);
    sum += term;
  }
  return sum;
}

// util::NumericalFunctions::LinearInterpolate (src/util/utilityFunctions.h:96-134).
double LinearInterpolate(const std::vector<double>& x, const std::vector<double>& y, double targetX) {
  size_t lower = 0;
  while (!(targetX >= x[lower] && targetX <= x[lower + 1])) ++lower;
  return y[lower] + (targetX - x[lower]) / (x[lower + 1] - x[lower]) * (y[lower + 1] - y[lower]);
}

}  // namespace

void womersley_coefficient(const double position[3], const double normal[3], double radius, double period,
                           double womersley_number, double pressure_gradient_amplitude, const double x[3],
                           double* re, double* im) {
  const Complex i(0, 1);
  const Complex iPowThreeHalves = std::pow(i, 1.5);
  const double dx = x[0] - position[0], dy = x[1] - position[1], dz = x[2] - position[2];
  const double z = dx * normal[0] + dy * normal[1] + dz * normal[2];
  const double r = std::sqrt((dx * dx + dy * dy + dz * dz) - z * z);
  const double omega = 2.0 * kPiM / period;
  const double density = 1.0;
  const Complex besselNumer = BesselJ0(iPowThreeHalves * womersley_number * r / radius);
  const Complex besselDenom = BesselJ0(iPowThreeHalves * womersley_number);
  const Complex k = pressure_gradient_amplitude / (density * omega) * (1.0 - besselNumer / besselDenom);
  *re = k.real();
  *im = k.imag();
}

void womersley_time_factor(double period, int64_t t, double* re, double* im) {
  const Complex i(0, 1);
  const double omega = 2.0 * kPiM / period;
  const Complex e = std::exp(i * omega * double(t));
  *re = e.real();
  *im = e.imag();
}

}  // namespace hlb

using namespace hlb;

// InOutLetFile::CalculateTable (src/lb/iolets/InOutLetFile.cc:37-101) and
// InOutLetFileVelocity::CalculateTable (InOutLetFileVelocity.cc:31-90).
extern "C" int hlb_iolet_file_table(const char* path, int velocity, int64_t total_steps, double time_step,
                                    double voxel_size, double* out, int64_t out_len, double* density_min) {
  const std::string who("hlb_iolet_file_table");
  if (!path || !out) return fail(HLB_EINVAL, who + ": null argument");
  if (total_steps < 1 || out_len < total_steps + 1)
    return fail(HLB_EINVAL, who + ": the table needs total_steps + 1 entries");
  if (!(time_step > 0.0) || !(voxel_size > 0.0)) return fail(HLB_EINVAL, who + ": time step and voxel size must be positive");
  std::map<double, double> timeValuePairs;
  {
    std::ifstream datafile(path);
    if (!datafile) return fail(HLB_EINVAL, who + ": cannot open " + std::string(path));
    double timeTemp = 0.0, valueTemp = 0.0;
    while (datafile.good()) {
      datafile >> timeTemp >> valueTemp;
      timeValuePairs[timeTemp] = valueTemp;
    }
  }
  if (timeValuePairs.size() < 2) return fail(HLB_EINVAL, who + ": the file needs at least two time points");
  const double end = total_steps * time_step;
  std::vector<double> times, values;
  double vMin = timeValuePairs.begin()->second, vMax = vMin;
  for (const auto& entry : timeValuePairs) {
    if (entry.first > end) {
      // an interpolated end point at the end of the run, then stop
      const double time_diff = end - times.back();
      const double time_diff_ratio = time_diff / (entry.first - times.back());
      const double diff = entry.second - values.back();
      const double final_value = values.back() + time_diff_ratio * diff;
      times.push_back(end);
      vMin = std::min(vMin, final_value);
      vMax = std::max(vMax, final_value);
      values.push_back(final_value);
      break;
    }
    vMin = std::min(vMin, entry.second);
    vMax = std::max(vMax, entry.second);
    times.push_back(entry.first);
    values.push_back(entry.second);
  }
  const int stepsInProfile = (int)(times.back() / time_step);
  if (times.size() < 2 || stepsInProfile <= 0)
    return fail(HLB_EINVAL, who + ": the profile must span at least one time step before the run ends");
  // UnitConverter (src/util/UnitConverter.cc:15-45)
  const double latticeMass = kBloodDensity * voxel_size * voxel_size * voxel_size;
  const double latticePressure = latticeMass / (voxel_size * time_step * time_step);
  const double latticeSpeed = voxel_size / time_step;
  auto density = [&](double p) {
    return (kCs2M + (p - kReferencePressure) * kMmHgToPa / latticePressure) / kCs2M;
  };
  for (int64_t timeStep = 0; timeStep <= total_steps; ++timeStep) {
    const double point = times.front() + (static_cast<double>((unsigned)timeStep % (unsigned)stepsInProfile) /
                                          static_cast<double>(total_steps)) *
                                             (times.back() - times.front());
    const double v = LinearInterpolate(times, values, point);
    out[timeStep] = velocity ? v / latticeSpeed : density(v);
  }
  if (density_min) *density_min = velocity ? 1.0 : density(vMin);
  return HLB_OK;
}
