// tdw_hooks.hpp -- the table through which a project build (a problem's
// predefines.h and problem-definition.h compiled into the library, as the
// reference compiles them into its kernel translation unit,
// hpc-engine/pca_kernels.cu:28-42) hands its callbacks to the engine.
//
// The project translation unit (csrc/project/tdw_project.hip) defines
// tdw_project_table(); the engine declares it weak, so the plain library
// (no project) sees a null table and uses its built-in problem.  Every
// callback that runs on the device is wrapped in a grid-sized kernel of the
// project unit that writes its values into engine-owned fields; the engine's
// own kernels then read those fields.  All launches go on the stream given.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "wslda_td.h"

// the constants the reference keeps in __constant__ memory
// (wslda_cuda_utils.h:21-37), mirrored into the project unit's dc_* symbols
struct tdw_project_consts {
    double mu_a, mu_b, ec, t0, dt, kF, eF, nF, sclgth;
    void *extra_data;
    size_t extra_size;
    double params[TDW_MAX_USER_PARAMS];
};

struct tdw_project {
    int codedim;             // 1, 2, 3: the td code the problem is built for
    int nx, ny, nz;          // NX, NY, NZ of predefines.h
    double dx, dy, dz;       // DX, DY, DZ
    int functional;          // FUNCTIONAL (TDW_SLDA ...)
    unsigned flags;          // TDW_PROJ_* switches (wslda_td.h)
    // the compiled settings as engine parameters (lattice, functional,
    // integration and regularisation scheme, stabilisation densities,
    // spin symmetry); fields predefines.h does not fix are left alone
    void (*fill_params)(tdw_params *p);
    int (*set_consts)(hipStream_t s, const tdw_project_consts *k);
    // v_ext(ix, iy, iz, it, SPINA / SPINB) into va, vb (N each)
    int (*vext)(hipStream_t s, int it, double *va, double *vb);
    // delta_ext(ix, iy, iz, it, delta) into dext (N complex)
    int (*delta_ext)(hipStream_t s, int it, const double2 *delta, double2 *dext);
    // velocity_ext(ix, iy, iz, it, spin, X/Y/ZAXIS) into v (3 N: x, y, z)
    int (*velocity_ext)(hipStream_t s, int it, int spin, double *v);
    // scattering_length(ix, iy, iz, it) into a (N)
    int (*sclgth)(hipStream_t s, int it, double *a);
    // modify_potentials<<<>>>(it, densities, potentials) on the 12 N blocks
    int (*modify_potentials)(hipStream_t s, int it, double *dens, double *pots);
    // modify_energies<<<>>>(it, densities, potentials, energy) with energy
    // the 7 N per-point contributions [EKIN .. EVELEXT]
    int (*modify_energies)(hipStream_t s, int it, double *dens, double *pots, double *energy);
    // host modify_densities(it, h_densities, params, extra_data_size,
    // h_extra_data, d_densities, d_extra_data) on 12 N blocks
    void (*modify_densities)(int it, double *h_dens, double *params, size_t extra_size, void *h_extra,
                             double *d_dens, void *d_extra);
};

extern "C" const tdw_project *tdw_project_table(void) __attribute__((weak));
