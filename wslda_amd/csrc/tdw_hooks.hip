// tdw_hooks.hip -- the engine's side of a project build: the problem's
// callbacks (compiled in by csrc/project/tdw_project.hip, table in
// tdw_hooks.hpp) turned into the fields the engine's kernels read, at the
// points where the reference evaluates them:
//   v_ext             inside the functional and the V_ext energy
//                     (tdwslda_functionals.h:72-73, 393-394, 561-562;
//                     pca_kernels.cu:137)
//   scattering_length inside the BdG and SLDAE functionals and the SLDAE
//                     energy (tdwslda_functionals.h:397, 585, 867)
//   delta_ext         added to Delta inside the H kernels only, never stored
//                     with the potentials (pca_kernels.cu:274-275, 428-429),
//                     and the Delta_ext energy (:159)
//   velocity_ext      added to the current terms of H (pca_kernels.cu:315-327,
//                     624-637, 743-757) and the velocity energy (:180-185)
//   modify_potentials after every potentials evaluation (pca_kernels.cu:90-92)
//   modify_energies   on the per-point energies before their sums
//                     (tdwslda_energy_generic.h:99-101)
//   modify_densities  before every potentials evaluation of a whole step
//                     (pca.c:1443-1447, 1515-1531)
#include <string.h>

#include "tdw_hooks.hpp"
#include "tdw_internal.hpp"
#include "tdw_physics.hpp"

namespace tdw {

static constexpr int NT = 256;

const tdw_project *project_table()
{
    return tdw_project_table ? tdw_project_table() : nullptr;
}

int project_consts(tdw_ctx *c)
{
    if (!c->proj) return 0;
    tdw_project_consts k;
    memset(&k, 0, sizeof(k));
    k.mu_a = c->p.mu_a;
    k.mu_b = c->p.mu_b;
    k.ec = c->p.ec;
    k.t0 = c->p.t0;
    k.dt = c->p.dt;
    k.kF = c->p.kF;
    k.eF = 0.5 * c->p.kF * c->p.kF;
    k.nF = c->p.kF * c->p.kF * c->p.kF / (3.0 * M_PI * M_PI);
    k.sclgth = c->p.sclgth;
    k.extra_data = c->extra_data;
    k.extra_size = c->extra_size;
    for (int j = 0; j < TDW_MAX_USER_PARAMS; j++) k.params[j] = c->user_params[j];
    return c->proj->set_consts(c->stream, &k);
}

int project_attach(tdw_ctx *c)
{
    const tdw_project *P = project_table();
    c->proj = P;
    if (!P) return 0;
    const tdw_params *p = &c->p;
    // the lattice and functional are compiled into the problem's code
    if (p->dim != P->codedim || p->nx != P->nx || p->ny != P->ny || p->nz != P->nz ||
        p->functional != P->functional)
        return TDW_EINVAL;
    // an external velocity enters H through the current terms, which only the
    // varying-mass step applies; the reference switches its constant-mass
    // (FAST) step off for it (pca_settings.h:218-221)
    if ((P->flags & TDW_PROJ_VELOCITY_EXT) && !varying_mass(c)) return TDW_EINVAL;
    const size_t N = c->N;
    TDW_HIP(hipMalloc(&c->d_uext, sizeof(double) * 2 * N));
    TDW_HIP(hipMalloc(&c->d_agrid, sizeof(double) * N));
    TDW_HIP(hipMalloc(&c->d_potsH, sizeof(double) * 12 * N));
    TDW_HIP(hipMalloc(&c->d_dext, sizeof(double2) * N));
    TDW_HIP(hipMalloc(&c->d_vel, sizeof(double) * 6 * N));
    TDW_HIP(hipMalloc(&c->d_epp, sizeof(double) * 7 * N));
    TDW_HIP(hipHostMalloc((void **)&c->h_densx, sizeof(double) * 12 * N, hipHostMallocDefault));
    return project_consts(c);
}

void project_release(tdw_ctx *c)
{
    void *bufs[] = {c->d_uext, c->d_agrid, c->d_potsH, c->d_dext, c->d_vel, c->d_epp};
    for (void *b : bufs)
        if (b) hipFree(b);
    if (c->h_densx) hipHostFree(c->h_densx);
}

const double *vext_grid(const tdw_ctx *c, int it)
{
    if (!c->proj || !c->proj->vext) return nullptr;
    if (c->proj->vext(c->stream, it, c->d_uext, c->d_uext + c->N) != 0) return nullptr;
    return c->d_uext;
}

const double *sclgth_grid(const tdw_ctx *c, int it)
{
    if (!c->proj || !c->proj->sclgth) return nullptr;
    if (c->proj->sclgth(c->stream, it, c->d_agrid) != 0) return nullptr;
    return c->d_agrid;
}

const double2 *dext_field(tdw_ctx *c, int it, const double *pots)
{
    if (!c->proj || !(c->proj->flags & TDW_PROJ_DELTA_EXT)) return nullptr;
    if (c->proj->delta_ext(c->stream, it, (const double2 *)(pots + 2 * c->N), c->d_dext) != 0) return nullptr;
    return c->d_dext;
}

const double *vel_field(tdw_ctx *c, int it)
{
    if (!c->proj || !(c->proj->flags & TDW_PROJ_VELOCITY_EXT)) return nullptr;
    if (c->proj->velocity_ext(c->stream, it, 0, c->d_vel) != 0) return nullptr;
    if (c->proj->velocity_ext(c->stream, it, 1, c->d_vel + 3 * c->N) != 0) return nullptr;
    return c->d_vel;
}

// Delta += Delta_ext; A_s -= v_s on the evolved axes.  In the reference's H
// the velocity enters as +i v_s beside -i A_s in the gradient term and as
// +div(v_s)/2 beside -div(A_s)/2 in the current term (pca_kernels.cu:315-327
// against :287-296), i.e. as A_s - v_s, whose divergence the H path forms.
__global__ void k_hview(long N, int dim, double *__restrict__ ph, const double2 *__restrict__ dext,
                        const double *__restrict__ vel)
{
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= N) return;
    if (dext) {
        ph[2 * N + 2 * i] += dext[i].x;
        ph[2 * N + 2 * i + 1] += dext[i].y;
    }
    if (vel) {
        for (int k = 0; k < dim; k++) {
            ph[(6 + k) * N + i] -= vel[k * N + i];
            ph[(9 + k) * N + i] -= vel[(3 + k) * N + i];
        }
    }
}

const double *h_view(tdw_ctx *c, int it, const double *pots)
{
    if (!c->proj || !(c->proj->flags & (TDW_PROJ_DELTA_EXT | TDW_PROJ_VELOCITY_EXT))) return pots;
    const long N = (long)c->N;
    const double2 *de = dext_field(c, it, pots);
    const double *ve = vel_field(c, it);
    if (hipMemcpyAsync(c->d_potsH, pots, sizeof(double) * 12 * N, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
        return nullptr;
    hipLaunchKernelGGL(k_hview, dim3((N + NT - 1) / NT), dim3(NT), 0, c->stream, N, c->dim, c->d_potsH, de, ve);
    if (hipGetLastError() != hipSuccess) return nullptr;
    return c->d_potsH;
}

int project_modify_potentials(tdw_ctx *c, int it, const double *dens, double *pots)
{
    if (!c->proj || !c->proj->modify_potentials) return 0;
    return c->proj->modify_potentials(c->stream, it, (double *)dens, pots);
}

int step_potentials(tdw_ctx *c, int it, double *dens, double *pots)
{
    if (c->proj && c->proj->modify_densities) {
        // the driver's host call between the density reduction and the
        // potentials: the device densities are final here
        TDW_HIP(hipStreamSynchronize(c->stream));
        c->proj->modify_densities(it, c->h_densx, c->user_params, c->extra_size, nullptr, dens, c->extra_data);
        TDW_HIP(hipDeviceSynchronize());
    }
    return potentials(c, it, dens, pots);
}

}  // namespace tdw

extern "C" int tdw_project_info(tdw_params *p, unsigned *flags)
{
    const tdw_project *P = tdw::project_table();
    if (!P) return 0;
    if (p) P->fill_params(p);
    if (flags) *flags = P->flags;
    return 1;
}
