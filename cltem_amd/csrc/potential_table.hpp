// Tabulated parameterised potentials: per element, V(r) as a piecewise cubic in r^2 over
// buckets read straight off the float bit pattern of r^2.  Bucket i covers 1/32 of an
// octave of r^2 in [2^kOctaveLo, 2^kOctaveHi) A^2, where r^2 is linear, so
//   i = (bits(r2) >> (23 - kSubBits)) - ((127 + kOctaveLo) << kSubBits),
//   t = (low 23 - kSubBits mantissa bits) / 2^(23 - kSubBits) in [0, 1),
//   V = ((c3 t + c2) t + c1) t + c0.
// The coefficients come from interpolating, in double, the reference formulas at the four
// Chebyshev nodes of each bucket: no transcendental function on the device, one 16-byte
// load per evaluation.
#pragma once

namespace mse_tab {

constexpr int kOctaveLo = -20;  // r^2 >= 2^-20 A^2: r_min = px / 4 >= 0.001 A, px >= 0.004 A
constexpr int kOctaveHi = 8;    // r^2 < 256 A^2: 8 A laterally and up to 12 A in depth
constexpr int kSubBits = 5;
constexpr int kBuckets = (kOctaveHi - kOctaveLo) << kSubBits;
constexpr int kMantBits = 23 - kSubBits;

// out[4 * kBuckets] = (c0, c1, c2, c3) per bucket of element Z (1-based) for the projected
// (full3d = false) or full-3D potential of parameterisation `form` (0 Kirkland, 1 Peng,
// 2 Lobato), parameters laid out as in the reference's params files.
void build(int form, bool full3d, const double* params, int i_count, int Z, float* out);

// The reference formulas in double (kernels/transmission_potentials_projected_d.cl:71-181,
// kernels/transmission_potentials_full_3d_d.cl:57-134) at distance r.
double vproj_ref(int form, const double* params, int i_count, int Z, double r);
double v3d_ref(int form, const double* params, int i_count, int Z, double r);

}  // namespace mse_tab
