// tdw_physics.hpp -- launchers of the modular-path kernels.
#pragma once
#include "tdw_internal.hpp"

namespace tdw {

// Volume element of the norm / <H> sums: DXYZ in 3D (pca_kernels.cu:913),
// DX in 1D (ccpca_kernels.cu:871), DX*DY in 2D (cpca_kernels.cu:579).
inline double dvol_norm(const tdw_ctx *c)
{
    if (c->dim == 1) return c->p.dx;
    if (c->dim == 2) return c->p.dx * c->p.dy;
    return c->p.dx * c->p.dy * c->p.dz;
}
// VOLUME_ELEMENT of the energy sums (tdwslda_functionals_framework_enable.h:7-16).
inline double dvol_energy(const tdw_ctx *c)
{
    double dxyz = c->p.dx * c->p.dy * c->p.dz;
    if (c->dim == 1) return dxyz * c->p.ny * c->p.nz;
    if (c->dim == 2) return dxyz * c->p.nz;
    return dxyz;
}
// Transverse normalisation of the densities: LY*LZ in 1D
// (ccpca_dens.cu:131-141), LZ in 2D (cpca_dens.cu:132-143).
inline double dens_transverse(const tdw_ctx *c)
{
    if (c->dim == 1) return (c->p.dy * c->p.ny) * (c->p.dz * c->p.nz);
    if (c->dim == 2) return c->p.dz * c->p.nz;
    return 1.0;
}

int abm1(tdw_ctx *c, int n, double2 *y, double2 *f1, double2 *f2, double2 *f3, double2 *f4, double2 *f5);
int abm4(tdw_ctx *c, int n, const double2 *hy, double2 *y, const double2 *fl);
int norms(tdw_ctx *c, int n, const double2 *wf, double *d_nrm);
int normalize(tdw_ctx *c, int n, double2 *wf, double *d_nrm);
int laplace(tdw_ctx *c, int m, const double2 *in, double2 *out);
int derivatives(tdw_ctx *c, int m, const double2 *in, double2 *ddx, double2 *ddy, double2 *ddz, double2 *lap);
int laplace_real(tdw_ctx *c, const double *f, double *lap);
int densities(tdw_ctx *c, int n, const double2 *wf, const double2 *ddx, const double2 *ddy,
              const double2 *ddz, const double2 *lap, const double *fbeta, const double *wgt,
              const double *kt, const int *cnt, double *dens, int grad);
int density_tau(tdw_ctx *c, double *dens);
int vext_eval(tdw_ctx *c, int it, double *va, double *vb);
int potentials(tdw_ctx *c, int it, const double *dens, double *pots);
int apply_h(tdw_ctx *c, int it, int n, const double2 *in, double2 *out, const double2 *lap,
            const double *pots, const double *kt, const double *useqpe, double *d_qpe);
int energy(tdw_ctx *c, int it, const double *dens, const double *pots, double *h_e);
int taylor(tdw_ctx *c, int order, double dt, int n, const double2 *hpsi, double2 *upd, double2 *contr);

}  // namespace tdw
