// CGS constants of the simulation, usable in host and device code.
//
// The values and the expressions that form them follow the reference's unit
// table (include/hamunits.h:8-68) so that every derived constant rounds to
// the same double: the LOS maps are compared with the CPU restatement at
// 1e-10 relative and the limit checks must flip at the same samples.
#pragma once

namespace hamx {
namespace cgs {

constexpr double erg = 1.;
constexpr double cm = 1.;
constexpr double sec = 1.;
constexpr double kelvin = 1.;

constexpr double pi = 3.141592653589793238462643383279502884197;
constexpr double halfpi = 1.570796326794896619231321691639751442099;
constexpr double twothirds = 0.666666666666666666666666666666666666667;
constexpr double onethird = 0.333333333333333333333333333333333333333;
constexpr double twopi = 6.283185307179586476925286766559005768394;
constexpr double sqrtwo = 1.4142135623730950488016887242096980785697;
constexpr double sqrthree = 1.7320508075688772935274463415058723669428;
constexpr double rad = (pi / 180.);

constexpr double esu = 1.; // sqrt(erg*cm)
constexpr double m = (1.e2 * cm);
constexpr double pc = (3.0856775806e+18 * cm);
constexpr double kpc = (3.0856775806e+21 * cm);
constexpr double eV = (1.60217733e-12 * erg);
constexpr double GeV = (1.e9 * eV);
constexpr double GHz = (1.e9 / sec);
constexpr double gauss = 1.; // sqrt(erg/cm^3)
constexpr double mugauss = (1.e-6 * gauss);
constexpr double ccm = (cm * cm * cm);

constexpr double c_light = (2.99792458e+10 * cm / sec);
constexpr double h_planck = (6.626075540e-27 * erg * sec);
constexpr double qe = (4.80320425e-10 * esu);
constexpr double mec2 = (0.51099907e-3 * GeV);
constexpr double mec = (mec2 / c_light);
constexpr double kB = (1.380622e-16 * erg / kelvin);

// HEALPix "undefined pixel" marker and the threshold the interpolation uses
// to skip such pixels (hamdis.h:95,689).
constexpr double undef = -1.6375e30;
constexpr double undef_threshold = -1.63749e30;

} // namespace cgs
} // namespace hamx
