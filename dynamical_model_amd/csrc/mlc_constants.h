// mlc_constants.h -- physical constants of the MOLOCH core (host and device).
//
// Values and the way they are composed follow mod_constants.F90 of the
// reference: egrav :81, amd :103, amw :105, boltzk :88, navgdr :90,
// rgasmol :126, rgas :130, rwat :131, cpd :141, cvd :145, wlhv :167,
// regrav :178, rcpd :179, rovcp :180, rdrcv :181, cpovr :182, tzero :193,
// p00 :225, lrate :231, mathpi :247, degrad :254, raddeg :255,
// earthrad :275, eomeg :279, eomeg2 :280, ep1 :296, ep2 :299.
// Each value is a compile-time double computed with one rounding per
// operation, as a Fortran compiler folds a PARAMETER.
#pragma once

namespace mlc {

constexpr double kEgrav = 9.80665;
constexpr double kAmd = 28.96454;
constexpr double kAmw = 18.01528;
constexpr double kBoltzk = 1.3806490e-23;
constexpr double kNavgdr = 6.02214076e23;
constexpr double kRgasmol = kNavgdr * kBoltzk;
constexpr double kRgas = (kRgasmol / kAmd) * 1000.0;
constexpr double kRwat = (kRgasmol / kAmw) * 1000.0;
constexpr double kCpd = 3.5 * kRgas;
constexpr double kCvd = 2.5 * kRgas;
constexpr double kRcpd = 1.0 / kCpd;
constexpr double kRovcp = kRgas * kRcpd;
constexpr double kRdrcv = kRgas / kCvd;
constexpr double kCpovr = kCpd / kRgas;
constexpr double kWlhv = 2500610.0;
constexpr double kTzero = 273.15;
constexpr double kP00 = 1.0e5;
constexpr double kLrate = 0.00649;
constexpr double kMathpi = 3.1415926535897932384626433832795029;
constexpr double kDegrad = kMathpi / 180.0;
constexpr double kRaddeg = 180.0 / kMathpi;
constexpr double kEarthrad = 6.371229e6;
constexpr double kEomeg = 7.2921159e-5;
constexpr double kEomeg2 = 2.0 * kEomeg;
constexpr double kEp1 = kAmd / kAmw - 1.0;
constexpr double kEp2 = kAmw / kAmd;
constexpr double kRegrav = 1.0 / kEgrav;

// module parameters of nompi/mod_moloch.F90:97-107
constexpr double kMinden = 1.0e-30;
constexpr double kXdamp = 0.0625;
constexpr double kNupaitq = 0.05;

}  // namespace mlc
