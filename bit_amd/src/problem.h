// problem.h -- default treecode parameters of the reference interface
// (problem.h of ongbw/bit): MAC parameter ACC, regularization DEL = delta^2,
// order KERNEL_ORDER of the regularized kernel family the field paths
// (treeforce, freespace_particles_tree) use; 0 is the reference's default.
#ifndef problem_h
#define problem_h

const double ACC = 0.5;
const double DEL = 0.01;
const int KERNEL_ORDER = 0;

#endif
