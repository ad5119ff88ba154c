// tdw_project.hip -- compiles a problem's predefines.h and problem-definition.h
// into the library, as the reference compiles them into its kernel unit
// (hpc-engine/pca_kernels.cu:28-42, td-project-template/problem-definition.h).
//
// Built only for a project (wslda_amd/build.py --project DIR --codedim D):
// DIR holds the problem's predefines.h and problem-definition.h, CODEDIM is
// the td code (1, 2, 3) the problem is written for.  This unit provides the
// names the problem's code uses (Complex with the thrust::complex<double>
// API, the dc_* / hc_* constants, params and extra data, lattice macros,
// wslda_density / wslda_potential views, the input structure), then wraps
// each device callback into a grid-sized kernel that the engine launches
// through tdw_project_table() (tdw_hooks.hpp).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <thrust/complex.h>

#include "tdw_hooks.hpp"
#include "wslda_td_project.h"

typedef thrust::complex<double> Complex;
#define double_complex Complex

#ifndef CODEDIM
#error "CODEDIM (1, 2 or 3) must be given when a project is built"
#endif
#define TDWSLDA 1

// ids shared with the reference's settings (pca_settings.h:69-73, 211-215, 222-223)
#define SLDA 111
#define ASLDA 112
#define BDG 113
#define CUSTOMEDF 114
#define SLDAE 115
#define AB3AM4 34
#define AB4AM5 45
#define AB5AM5 55
#define SPHERICAL_CUTOFF 88
#define CUBIC_CUTOFF 89

#include "predefines.h"

#ifndef MAX_USER_PARAMS
#define MAX_USER_PARAMS 32
#endif
static_assert(MAX_USER_PARAMS <= TDW_MAX_USER_PARAMS, "MAX_USER_PARAMS above the engine's parameter block");
#ifndef FUNCTIONAL
#define FUNCTIONAL SLDA
#endif
#if FUNCTIONAL == CUSTOMEDF
#error "FUNCTIONAL CUSTOMEDF is not supported by this engine"
#endif
#ifndef INTEGRATION_SCHEME
#define INTEGRATION_SCHEME AB4AM5
#endif
#ifdef SLDAE_FORCE_A1
#define SLDA_FORCE_A1
#endif

// lattice and bookkeeping names of the reference's settings
#define DXYZ (DX * DY * DZ)
#define DXY (DX * DY)
#define NXYZ (NX * NY * NZ)
#define NXY (NX * NY)
#define LX (DX * NX)
#define LY (DY * NY)
#define LZ (DZ * NZ)
#define LXYZ (LX * LY * LZ)
#define LXY (LX * LY)
#define SPINA 0
#define SPINB 1
#define XAXIS 0
#define YAXIS 1
#define ZAXIS 2
#define ENERGYITEMS 7
#define EKIN 0
#define EPOT 1
#define EPAIR 2
#define ECURRENT 3
#define EPOTEXT 4
#define EPAIREXT 5
#define EVELEXT 6
#define NPARTA 7
#define NPARTB 8
#define LZA 9
#define LZB 10

#if CODEDIM == 3
#define TDWP_NPTS NXYZ
#define TDWP_VE (DXYZ)
#elif CODEDIM == 2
#define TDWP_NPTS NXY
#define TDWP_VE (DXYZ * NZ)
#elif CODEDIM == 1
#define TDWP_NPTS NX
#define TDWP_VE (DXYZ * NY * NZ)
#else
#error "CODEDIM must be 1, 2 or 3"
#endif

// point index -> lattice coordinates (x slowest); i is scratch
#define ixyz2ixiyiz(ixyz, _ix, _iy, _iz, i) \
    i = (ixyz);                             \
    _ix = i / (NY * NZ);                    \
    i = i - _ix * NY * NZ;                  \
    _iy = i / NZ;                           \
    _iz = i - _iy * NZ;
#define ixy2ixiy2d(ixy, _ix, _iy) \
    _ix = (ixy) / NY;             \
    _iy = (ixy) - _ix * NY;
#if CODEDIM == 3
#define TDWP_DECODE(ixyz, _ix, _iy, _iz, i) ixyz2ixiyiz(ixyz, _ix, _iy, _iz, i)
#elif CODEDIM == 2
#define TDWP_DECODE(ixyz, _ix, _iy, _iz, i) \
    ixy2ixiy2d(ixyz, _ix, _iy) _iz = 0;
#else
#define TDWP_DECODE(ixyz, _ix, _iy, _iz, i) \
    _ix = (ixyz);                           \
    _iy = 0;                                \
    _iz = 0;
#endif

// device constants (wslda_cuda_utils.h:21-33) and their host copies
__constant__ double dc_mu_a;
__constant__ double dc_mu_b;
__constant__ double dc_ec;
__constant__ double dc_t0;
__constant__ double dc_dt;
__constant__ double dc_kF;
__constant__ double dc_eF;
__constant__ double dc_nF;
__constant__ double dc_sclgth;
__constant__ void *dc_extra_data;
__constant__ size_t dc_extra_data_size;
__constant__ double dc_params[MAX_USER_PARAMS];
static double hc_mu_a = 0.0, hc_mu_b = 0.0, hc_ec = 0.0, hc_t0 = 0.0, hc_dt = 0.0;
static double hc_kF = 0.0, hc_eF = 0.0, hc_nF = 0.0, hc_sclgth = 0.0;

// pointer views of the 12 N density / potential blocks (wslda_potdens.h:53-98)
typedef struct {
    int nx, ny, nz, datadim, blocklength;
    Complex *nu;
    double *rho_a, *rho_b, *tau_a, *tau_b;
    double *j_a_x, *j_a_y, *j_a_z, *j_b_x, *j_b_y, *j_b_z;
} wslda_density;
typedef struct {
    int nx, ny, nz, datadim, blocklength;
    Complex *delta;
    double *alpha_a, *alpha_b, *V_a, *V_b;
    double *A_a_x, *A_a_y, *A_a_z, *A_b_x, *A_b_y, *A_b_z;
    double *mu;
} wslda_potential;

// the run inputs a problem's host code reads (input->sclgth, input->akF, ...);
// the driver sets them through tdw_project_input()
typedef tdw_project_inputs metadata_t;
metadata_t md;
metadata_t *input = &md;

extern "C" tdw_project_inputs *tdw_project_input(void) { return &md; }

#include "problem-definition.h"

namespace {

constexpr int PT = 256;
constexpr long NPT = TDWP_NPTS;
inline dim3 grid_of() { return dim3((unsigned)((NPT + PT - 1) / PT)); }

__device__ __forceinline__ void coords(long i, int &ix, int &iy, int &iz)
{
    long t;
    (void)t;
    TDWP_DECODE(i, ix, iy, iz, t);
}

wslda_density dens_view(double *d)
{
    wslda_density v;
    v.nx = NX; v.ny = CODEDIM >= 2 ? NY : 1; v.nz = CODEDIM >= 3 ? NZ : 1;
    v.datadim = CODEDIM; v.blocklength = (int)NPT;
    v.nu = (Complex *)d;
    v.rho_a = d + 2 * NPT; v.tau_a = d + 3 * NPT;
    v.j_a_x = d + 4 * NPT; v.j_a_y = d + 5 * NPT; v.j_a_z = d + 6 * NPT;
    v.rho_b = d + 7 * NPT; v.tau_b = d + 8 * NPT;
    v.j_b_x = d + 9 * NPT; v.j_b_y = d + 10 * NPT; v.j_b_z = d + 11 * NPT;
    return v;
}

wslda_potential pots_view(double *p)
{
    wslda_potential v;
    v.nx = NX; v.ny = CODEDIM >= 2 ? NY : 1; v.nz = CODEDIM >= 3 ? NZ : 1;
    v.datadim = CODEDIM; v.blocklength = (int)NPT;
    v.V_a = p; v.V_b = p + NPT;
    v.delta = (Complex *)(p + 2 * NPT);
    v.alpha_a = p + 4 * NPT; v.alpha_b = p + 5 * NPT;
    v.A_a_x = p + 6 * NPT; v.A_a_y = p + 7 * NPT; v.A_a_z = p + 8 * NPT;
    v.A_b_x = p + 9 * NPT; v.A_b_y = p + 10 * NPT; v.A_b_z = p + 11 * NPT;
    v.mu = nullptr;
    return v;
}

#ifdef ENABLE_V_EXT
__global__ void kp_vext(int it, double *va, double *vb)
{
    const long i = (long)blockIdx.x * PT + threadIdx.x;
    if (i >= NPT) return;
    int ix, iy, iz;
    coords(i, ix, iy, iz);
    va[i] = v_ext(ix, iy, iz, it, SPINA, dc_params, dc_extra_data_size, dc_extra_data);
    vb[i] = v_ext(ix, iy, iz, it, SPINB, dc_params, dc_extra_data_size, dc_extra_data);
}
int h_vext(hipStream_t s, int it, double *va, double *vb)
{
    hipLaunchKernelGGL(kp_vext, grid_of(), dim3(PT), 0, s, it, va, vb);
    return (int)hipGetLastError();
}
#endif

#ifdef ENABLE_DELTA_EXT
__global__ void kp_delta_ext(int it, const double2 *delta, double2 *out)
{
    const long i = (long)blockIdx.x * PT + threadIdx.x;
    if (i >= NPT) return;
    int ix, iy, iz;
    coords(i, ix, iy, iz);
    const Complex d(delta[i].x, delta[i].y);
    const Complex e = delta_ext(ix, iy, iz, it, d, dc_params, dc_extra_data_size, dc_extra_data);
    out[i] = make_double2(e.real(), e.imag());
}
int h_delta_ext(hipStream_t s, int it, const double2 *delta, double2 *out)
{
    hipLaunchKernelGGL(kp_delta_ext, grid_of(), dim3(PT), 0, s, it, delta, out);
    return (int)hipGetLastError();
}
#endif

// velocity_ext is evaluated for the output even without ENABLE_VELOCITY_EXT
// (pca_kernels.cu:1149-1180), which enables it in H
__global__ void kp_velocity_ext(int it, int spin, double *v)
{
    const long i = (long)blockIdx.x * PT + threadIdx.x;
    if (i >= NPT) return;
    int ix, iy, iz;
    coords(i, ix, iy, iz);
    v[i] = velocity_ext(ix, iy, iz, it, spin, XAXIS, dc_params, dc_extra_data_size, dc_extra_data);
    v[NPT + i] = CODEDIM >= 2 ? velocity_ext(ix, iy, iz, it, spin, YAXIS, dc_params, dc_extra_data_size, dc_extra_data)
                              : 0.0;
    v[2 * NPT + i] =
        CODEDIM >= 3 ? velocity_ext(ix, iy, iz, it, spin, ZAXIS, dc_params, dc_extra_data_size, dc_extra_data) : 0.0;
}
int h_velocity_ext(hipStream_t s, int it, int spin, double *v)
{
    hipLaunchKernelGGL(kp_velocity_ext, grid_of(), dim3(PT), 0, s, it, spin, v);
    return (int)hipGetLastError();
}

#if FUNCTIONAL == BDG || FUNCTIONAL == SLDAE
__global__ void kp_sclgth(int it, double *a)
{
    const long i = (long)blockIdx.x * PT + threadIdx.x;
    if (i >= NPT) return;
    int ix, iy, iz;
    coords(i, ix, iy, iz);
    a[i] = scattering_length(ix, iy, iz, it, dc_params, dc_extra_data_size, dc_extra_data);
}
int h_sclgth(hipStream_t s, int it, double *a)
{
    hipLaunchKernelGGL(kp_sclgth, grid_of(), dim3(PT), 0, s, it, a);
    return (int)hipGetLastError();
}
#endif

#ifdef ENABLE_MODIFY_POTENTIALS
int h_modify_potentials(hipStream_t s, int it, double *dens, double *pots)
{
    hipLaunchKernelGGL(modify_potentials, grid_of(), dim3(PT), 0, s, it, dens_view(dens), pots_view(pots));
    return (int)hipGetLastError();
}
#endif

#ifdef ENABLE_MODIFY_ENERGIES
int h_modify_energies(hipStream_t s, int it, double *dens, double *pots, double *energy)
{
    hipLaunchKernelGGL(modify_energies, grid_of(), dim3(PT), 0, s, it, dens_view(dens), pots_view(pots), energy);
    return (int)hipGetLastError();
}
#endif

#ifdef ENABLE_MODIFY_DENSITIES
void h_modify_densities(int it, double *h_dens, double *prm, size_t esize, void *h_extra, double *d_dens,
                        void *d_extra)
{
    modify_densities(it, dens_view(h_dens), prm, esize, h_extra, dens_view(d_dens), d_extra);
}
#endif

int h_set_consts(hipStream_t s, const tdw_project_consts *k)
{
#define TDWP_SYM(sym, val)                                                                                  \
    do {                                                                                                    \
        hipError_t e = hipMemcpyToSymbolAsync(HIP_SYMBOL(sym), &(val), sizeof(val), 0,                      \
                                              hipMemcpyHostToDevice, s);                                    \
        if (e != hipSuccess) return (int)e;                                                                 \
    } while (0)
    static tdw_project_consts held;   // the async copies read from here
    hipError_t e0 = hipStreamSynchronize(s);
    if (e0 != hipSuccess) return (int)e0;
    held = *k;
    TDWP_SYM(dc_mu_a, held.mu_a);
    TDWP_SYM(dc_mu_b, held.mu_b);
    TDWP_SYM(dc_ec, held.ec);
    TDWP_SYM(dc_t0, held.t0);
    TDWP_SYM(dc_dt, held.dt);
    TDWP_SYM(dc_kF, held.kF);
    TDWP_SYM(dc_eF, held.eF);
    TDWP_SYM(dc_nF, held.nF);
    TDWP_SYM(dc_sclgth, held.sclgth);
    TDWP_SYM(dc_extra_data, held.extra_data);
    TDWP_SYM(dc_extra_data_size, held.extra_size);
    hipError_t e = hipMemcpyToSymbolAsync(HIP_SYMBOL(dc_params), held.params, sizeof(double) * MAX_USER_PARAMS, 0,
                                          hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return (int)e;
#undef TDWP_SYM
    hc_mu_a = k->mu_a; hc_mu_b = k->mu_b; hc_ec = k->ec; hc_t0 = k->t0; hc_dt = k->dt;
    hc_kF = k->kF; hc_eF = k->eF; hc_nF = k->nF; hc_sclgth = k->sclgth;
    return 0;
}

void h_fill_params(tdw_params *p)
{
    p->dim = CODEDIM;
    p->nx = NX; p->ny = NY; p->nz = NZ;
    p->dx = DX; p->dy = DY; p->dz = DZ;
    p->functional = FUNCTIONAL;
    p->scheme = INTEGRATION_SCHEME;
#if defined(REGULARIZATION_SCHEME) && REGULARIZATION_SCHEME == CUBIC_CUTOFF
    p->regularization = TDW_CUBIC_CUTOFF;
#else
    p->regularization = TDW_SPHERICAL_CUTOFF;
#endif
#ifdef INCLUDE_MF_CORRECTION_FROM_REG_SCHEME
    p->mf_correction = 1;
#else
    p->mf_correction = 0;
#endif
#ifdef SLDA_STABILIZATION_EXCLUDE_BELOW_DENISTY
    p->p_nmin = SLDA_STABILIZATION_EXCLUDE_BELOW_DENISTY;
#else
    p->p_nmin = 1.0e-7;
#endif
#ifdef SLDA_STABILIZATION_RETAIN_ABOVE_DENSITY
    p->p_nmax = SLDA_STABILIZATION_RETAIN_ABOVE_DENSITY;
#else
    p->p_nmax = 1.0e-5;
#endif
#ifdef DENSEPSILON
    p->denseps = DENSEPSILON;
#endif
#ifdef SPINSYMMETRY_MODE
    p->spinsymmetry = 1;
#else
    p->spinsymmetry = 0;
#endif
#if FUNCTIONAL == SLDAE && defined(SLDA_FORCE_A1)
    p->force_a1 = 1;
#else
    p->force_a1 = 0;
#endif
}

const tdw_project g_table = {
    CODEDIM, NX, NY, NZ, DX, DY, DZ, FUNCTIONAL,
    0u
#ifdef ENABLE_V_EXT
        | TDW_PROJ_V_EXT
#endif
#ifdef ENABLE_DELTA_EXT
        | TDW_PROJ_DELTA_EXT
#endif
#ifdef ENABLE_VELOCITY_EXT
        | TDW_PROJ_VELOCITY_EXT
#endif
#ifdef ENABLE_MODIFY_POTENTIALS
        | TDW_PROJ_MODIFY_POTENTIALS
#endif
#ifdef ENABLE_MODIFY_ENERGIES
        | TDW_PROJ_MODIFY_ENERGIES
#endif
#ifdef ENABLE_MODIFY_DENSITIES
        | TDW_PROJ_MODIFY_DENSITIES
#endif
    ,
    h_fill_params,
    h_set_consts,
#ifdef ENABLE_V_EXT
    h_vext,
#else
    nullptr,
#endif
#ifdef ENABLE_DELTA_EXT
    h_delta_ext,
#else
    nullptr,
#endif
    h_velocity_ext,
#if FUNCTIONAL == BDG || FUNCTIONAL == SLDAE
    h_sclgth,
#else
    nullptr,
#endif
#ifdef ENABLE_MODIFY_POTENTIALS
    h_modify_potentials,
#else
    nullptr,
#endif
#ifdef ENABLE_MODIFY_ENERGIES
    h_modify_energies,
#else
    nullptr,
#endif
#ifdef ENABLE_MODIFY_DENSITIES
    h_modify_densities,
#else
    nullptr,
#endif
};

}  // namespace

extern "C" const tdw_project *tdw_project_table(void) { return &g_table; }
