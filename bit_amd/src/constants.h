// constants.h -- compile-time constants of the reference library interface
// (constants.h of ongbw/bit; same names and values, so that code written
// against the reference compiles unchanged).
#ifndef constants_h
#define constants_h

const int MAXPART = 1000000;  // unused by the library (reference: maximum particles)
const int NUMCHILD = 8;       // children per tree node
const int MAXTERM = 12;       // reference cap on Taylor terms (nterms <= MAXTERM-1);
                              // this library accepts nterms <= 13 (BIT_MAX_NTERMS)
const int MAXCHILD = 100000;  // unused (reference: maximum trees)
const double pi = 3.14159265359;
const double inv2pi = 0.5 / pi;

// physical constants (SI), as the reference uses them to scale fields
const double perm = 8.85418781762e-12;
const double elec = 1.60217733e-19;
const double invmass = 1.0 / 9.1093897e-31;
const double mu = 4e-7 * pi;
const double c = 3.0e8;
const double charge_constant = inv2pi / perm;

const double eps = 1e-14;  // squared distance below which a pair is skipped
const double Bfield = 10.0;
const double endcap = -50.0;

#endif
