/* Test problem for the project build: every callback of the problem
 * interface switched on, on the td-gauss-bdg lattice (12^3, spacing 1.15,
 * BdG functional) for the 3D td code. */
#define NX 12
#define NY 12
#define NZ 12
#define DX 1.15
#define DY 1.15
#define DZ 1.15

#define FUNCTIONAL BDG
#define ENABLE_V_EXT
#define ENABLE_DELTA_EXT
#define ENABLE_MODIFY_POTENTIALS
#define ENABLE_MODIFY_ENERGIES
#define ENABLE_MODIFY_DENSITIES

#define MAX_USER_PARAMS 32
#define DENSEPSILON 1.0e-8
#define SLDA_STABILIZATION_RETAIN_ABOVE_DENSITY 1.0e-5
#define SLDA_STABILIZATION_EXCLUDE_BELOW_DENISTY 1.0e-7
