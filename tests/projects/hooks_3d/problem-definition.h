/* Test problem exercising every callback (parameters are used as given;
 * process_params leaves them alone):
 *   params[0], [1]  V_ext = p0 exp(p1 x^2) for spin a, its negative for b
 *   params[2]       delta_ext = p2 delta + (p3 + i p4) exp(2 pi i y / LY)
 *   params[5]       scattering length a (1 + p5) on the half ix >= NX/2
 *   params[6]       modify_potentials: V_b += p6 cos(2 pi z / LZ)
 *   params[7]       modify_energies: E_pot += p7 rho_a dV
 *   params[8]       modify_densities: rho_b *= (1 + p8), through the host
 * hooks_calls / hooks_last_it record the host calls of modify_densities. */
#include <stdlib.h>

extern "C" {
int hooks_calls = 0;
int hooks_last_it = -1;
}

extern "C" void process_params(double *params, double *kF, double *mu, size_t extra_data_size, void *extra_data) {}

__device__ double v_ext(int ix, int iy, int iz, int it, int spin, double *params, size_t extra_data_size,
                        void *extra_data)
{
    const double x = DX * (ix - NX / 2);
    const double v = params[0] * exp(params[1] * x * x);
    return spin == SPINA ? v : -v;
}

__device__ Complex delta_ext(int ix, int iy, int iz, int it, Complex delta, double *params, size_t extra_data_size,
                             void *extra_data)
{
    const double y = DY * (iy - NY / 2);
    const double ph = 2.0 * M_PI * y / LY;
    return params[2] * delta + Complex(params[3], params[4]) * Complex(cos(ph), sin(ph));
}

__device__ double velocity_ext(int ix, int iy, int iz, int it, int spin, int coordinate, double *params,
                               size_t extra_data_size, void *extra_data)
{
    return 0.0;
}

__device__ double scattering_length(int ix, int iy, int iz, int it, double *params, size_t extra_data_size,
                                    void *extra_data)
{
    return ix >= NX / 2 ? dc_sclgth * (1.0 + params[5]) : dc_sclgth;
}

#include "tdwslda_functionals_framework_enable.h"
extern "C" void modify_densities(int it, wslda_density h_densities, double *params, size_t extra_data_size,
                                 void *h_extra_data, wslda_density d_densities, void *d_extra_data)
{
    hooks_calls++;
    hooks_last_it = it;
    memcopy_gpu2host(d_densities.rho_b, h_densities.rho_b, NUMBER_ELEMENT * sizeof(double));
    for (int i = 0; i < NUMBER_ELEMENT; i++) h_densities.rho_b[i] *= 1.0 + params[8];
    memcopy_host2gpu(h_densities.rho_b, d_densities.rho_b, NUMBER_ELEMENT * sizeof(double));
}
#include "tdwslda_functionals_framework_disable.h"

#include "tdwslda_functionals_framework_enable.h"
__global__ void modify_potentials(int it, wslda_density h_densities, wslda_potential h_potentials)
{
    size_t ixyz = threadIdx.x + blockIdx.x * blockDim.x;
    int ix, iy, iz, i;
    if (ixyz < NUMBER_ELEMENT) {
        decode_ixyz2ixiyiz(ixyz, ix, iy, iz, i);
        const double z = DZ * (iz - NZ / 2);
        h_potentials.V_b[ixyz] += params[6] * cos(2.0 * M_PI * z / LZ);
    }
}
#include "tdwslda_functionals_framework_disable.h"

#include "tdwslda_functionals_framework_enable.h"
__global__ void modify_energies(int it, wslda_density h_densities, wslda_potential h_potentials, double *energy)
{
    size_t ixyz = threadIdx.x + blockIdx.x * blockDim.x;
    if (ixyz < NUMBER_ELEMENT) energy[EPOT * NUMBER_ELEMENT + ixyz] += params[7] * h_densities.rho_a[ixyz] * VOLUME_ELEMENT;
}
#include "tdwslda_functionals_framework_disable.h"

extern "C" size_t get_extra_data_size(double *params) { return 0; }

extern "C" int load_extra_data(size_t extra_data_size, void *extra_data, double *params) { return 0; }
