/* Test problem: the td-gauss-slda testsuite case (testsuite/td-gauss-slda)
 * written as a project for the 1D td code: 12-point x axis with a 12x12
 * transverse box, spacing 1.15, SLDA, a time-switched Gaussian V_ext. */
#define NX 12
#define NY 12
#define NZ 12
#define DX 1.15
#define DY 1.15
#define DZ 1.15

#define FUNCTIONAL SLDA
#define ENABLE_V_EXT

#define MAX_USER_PARAMS 32
#define DENSEPSILON 1.0e-8
#define SLDA_STABILIZATION_RETAIN_ABOVE_DENSITY 1.0e-5
#define SLDA_STABILIZATION_EXCLUDE_BELOW_DENISTY 1.0e-7
