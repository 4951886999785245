// mlc_internal.h -- helpers shared by the translation units of the library.
#pragma once

// Record an error message for mlc_last_error() and return the code.
int mlc_fail(int code, const char *fmt, ...);

// Host copies of reference functions used during set-up.
double gzita_host(double zita, double ztop, double a0);  // mod_zita.F90:110-120
double pfwsat_host(double t, double p);                   // nompi/mod_moloch.F90:119-181
