/* Names a problem's device code uses between this header and
 * tdwslda_functionals_framework_disable.h (the pair its problem-definition.h
 * includes around modify_densities, modify_potentials, modify_energies and
 * a custom functional; reference interface:
 * hpc-engine/tdwslda_functionals_framework_enable.h).
 *
 *   params, extra_data, extra_data_size  the problem's user parameters and
 *                                        extra data as the engine holds them
 *   NUMBER_ELEMENT                       points of the evolved lattice
 *   VOLUME_ELEMENT                       volume element of the energy sums
 *   decode_ixyz2ixiyiz(ixyz, ix, iy, iz, i)
 *                                        lattice coordinates of point ixyz
 *                                        (iy = 0 in 1D, iz = 0 in 1D / 2D)
 */
#define params dc_params
#define extra_data dc_extra_data
#define extra_data_size dc_extra_data_size
#define NUMBER_ELEMENT TDWP_NPTS
#define VOLUME_ELEMENT TDWP_VE
#define decode_ixyz2ixiyiz(ixyz, _ix, _iy, _iz, i) TDWP_DECODE(ixyz, _ix, _iy, _iz, i)
