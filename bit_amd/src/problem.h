// problem.h -- default treecode parameters of the reference interface
// (problem.h of ongbw/bit): MAC parameter ACC, regularization DEL = delta^2,
// kernel family KERNEL_ORDER (only order 0 is implemented, as the reference's
// default).
#ifndef problem_h
#define problem_h

const double ACC = 0.5;
const double DEL = 0.01;
const int KERNEL_ORDER = 0;

#endif
