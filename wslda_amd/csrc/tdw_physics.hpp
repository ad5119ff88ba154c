// tdw_physics.hpp -- launchers of the modular-path kernels.
#pragma once
#include "tdw_functional.hpp"
#include "tdw_internal.hpp"

namespace tdw {

// ---------------------------------------------------------------------------
// external potential of the built-in problem: Gaussians (times a smooth
// on/off switch) plus a harmonic trap, in coordinates centred on the lattice
// ---------------------------------------------------------------------------
struct VextDev {
    tdw_vext v;
    double s;          // switch value at this time
    int dim, nx, ny, nz;
    double dx, dy, dz;
    const double *grid;   // project build: V_ext of spin a, b (2 N); overrides the above
    long npts;
};

__device__ __forceinline__ void decode(int dim, int ny, int nz, long i, int &ix, int &iy, int &iz)
{
    if (dim == 1) { ix = (int)i; iy = 0; iz = 0; }
    else if (dim == 2) { ix = (int)(i / ny); iy = (int)(i - (long)ix * ny); iz = 0; }
    else {
        ix = (int)(i / ((long)ny * nz));
        long r = i - (long)ix * ny * nz;
        iy = (int)(r / nz);
        iz = (int)(r - (long)iy * nz);
    }
}

__device__ __forceinline__ void vext_point(const VextDev &V, long i, double &va, double &vb)
{
    if (V.grid) {
        va = V.grid[i];
        vb = V.grid[V.npts + i];
        return;
    }
    int ix, iy, iz;
    decode(V.dim, V.ny, V.nz, i, ix, iy, iz);
    const double x = V.dx * (ix - V.nx / 2);
    const double y = V.dim >= 2 ? V.dy * (iy - V.ny / 2) : 0.0;
    const double z = V.dim >= 3 ? V.dz * (iz - V.nz / 2) : 0.0;
    double g = 0.0;
    for (int k = 0; k < V.v.ngauss; k++) {
        const double X = x - V.v.cx[k], Y = y - V.v.cy[k], Z = z - V.v.cz[k];
        g += V.v.amp[k] * V.s * exp(V.v.wx[k] * (X * X) + V.v.wy[k] * (Y * Y) + V.v.wz[k] * (Z * Z));
    }
    const double trap = 0.5 * V.v.omega * V.v.omega * (x * x + y * y + z * z);
    va = g + trap;
    vb = (V.v.flip_b ? -g : g) + trap;
}

// V_ext of the context at step it (switch value evaluated on the host; in a
// project build the problem's v_ext field of that step, zero without
// ENABLE_V_EXT)
VextDev make_vext(const tdw_ctx *c, int it);
// project build (tdw_hooks.hip): the engine's side of the problem's callbacks
int project_attach(tdw_ctx *c);
void project_release(tdw_ctx *c);
int project_consts(tdw_ctx *c);
const double *vext_grid(const tdw_ctx *c, int it);
const double *sclgth_grid(const tdw_ctx *c, int it);
const double2 *dext_field(tdw_ctx *c, int it, const double *pots);
const double *vel_field(tdw_ctx *c, int it);
// the potentials H reads: pots itself, or a copy with Delta + Delta_ext and
// A_s - v_s when the problem has those callbacks (null on a launch error)
const double *h_view(tdw_ctx *c, int it, const double *pots);
int project_modify_potentials(tdw_ctx *c, int it, const double *dens, double *pots);
// modify_densities (host) before the potentials of a whole step
int step_potentials(tdw_ctx *c, int it, double *dens, double *pots);
// functional parameters of the context
FParams fparams(const tdw_ctx *c);

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
// abm1 reading y from yin and writing the prediction to y (may equal yin);
// with cout the partial corrector goes there and the history is only read
int abm1_to(tdw_ctx *c, int n, const double2 *yin, double2 *y, double2 *cout, double2 *f1, double2 *f2, double2 *f3,
            double2 *f4, double2 *f5);
int abm4(tdw_ctx *c, int n, const double2 *hy, double2 *y, const double2 *fl);
int norms(tdw_ctx *c, int n, const double2 *wf, double *d_nrm);
int normalize(tdw_ctx *c, int n, double2 *wf, double *d_nrm);
// the same two with the caller's partial-sum buffer (n * norm_chunks(c)
// doubles) instead of the context's
int norm_chunks(const tdw_ctx *c);
int norms_with(tdw_ctx *c, int n, const double2 *wf, double *d_nrm, double *part);
int normalize_with(tdw_ctx *c, int n, double2 *wf, double *d_nrm, double *part);
// the no-allocation scratch of this device (tdw_noalloc.hip), set up once per
// device by tdw_create with the device current
int noalloc_setup(int device);
int laplace(tdw_ctx *c, int m, const double2 *in, double2 *out);
int derivatives(tdw_ctx *c, int m, const double2 *in, double2 *ddx, double2 *ddy, double2 *ddz, double2 *lap);
int laplace_real(tdw_ctx *c, const double *f, double *lap);
int densities(tdw_ctx *c, int n, const double2 *wf, const double2 *ddx, const double2 *ddy,
              const double2 *ddz, const double2 *lap, const double *fbeta, const double *wgt,
              const double *kt, const int *cnt, double *dens, int grad);
// rho, nu (no gradients) from FP32 wavefunctions, sums in FP64
int densities32(tdw_ctx *c, int n, const float2 *wf, const double *fbeta, const double *wgt, const double *kt,
                const int *cnt, double *dens);
int density_tau(tdw_ctx *c, double *dens);
// spin-symmetric mode: fold the u contributions kept in spin a into spin b,
// double nu, copy b to a (no-op when the mode is off)
int spin_fold(tdw_ctx *c, double *dens);
int copy_f64(tdw_ctx *c, const double *src, double *dst, long n);
// grid-field helpers (tdw_grid.hip): spectral gradient, per-component
// derivative and divergence of real fields (evolved axes only), the
// Fermi-Dirac filter FD(k^2/2) of n real fields or one complex field, and the
// per-state overlaps <w1_a | w2_a> dV
int grid_gradient(tdw_ctx *c, const double *f, double *const df[3]);
int grid_dvec(tdw_ctx *c, const double *const f[3], double *const df[3]);
int grid_div(tdw_ctx *c, const double *const f[3], double *div);
int grid_filter_d(tdw_ctx *c, int n, const double *in, double *out, double mu, double T);
int grid_filter_c(tdw_ctx *c, const double2 *in, double2 *out, double mu, double T);
// hkf_mode filtering of V, Delta, A (and alpha with with_alpha) in place
int hkf_fields(tdw_ctx *c, double *pots, bool with_alpha);
int grid_overlap(tdw_ctx *c, int n, const double2 *w1, const double2 *w2, double *re, double *im);
int grid_sum(tdw_ctx *c, const double *a, long n, double *out);
// quantum friction: the densities U_a, U_b, D (4N doubles) and the friction
// terms added to / removed from V_a, V_b, Delta around an H application
int qf_densities(tdw_ctx *c, int n, const double2 *wf, const double2 *lap, const double *fbeta, const double *wgt,
                 const double *kt, const int *cnt, double *out);
int qf_add(tdw_ctx *c, double *qf, double *pots, double qfswitch, double pccoeff);
int qf_remove(tdw_ctx *c, const double *qf, double *pots);
int vext_eval(tdw_ctx *c, int it, double *va, double *vb);
// cccoeff: the current-corrections coefficient of the ASLDA vector
// potentials and current terms (1 after the reference's default ramp)
int potentials(tdw_ctx *c, int it, const double *dens, double *pots, double cccoeff = 1.0);
int apply_h(tdw_ctx *c, int it, int n, const double2 *in, double2 *out, const double2 *lap,
            const double *pots, const double *kt, const double *useqpe, double *d_qpe);
int apply_h_only(tdw_ctx *c, int n, const double2 *in, double2 *out, const double2 *lap, const double *pots,
                 const double *kt);
// the same two on 2D lattices with a one-pass plane kernel (h2d_supported):
// the Laplacian of in is formed in LDS; out must differ from in
int apply_h_plane(tdw_ctx *c, int it, int n, const double2 *in, double2 *out, const double *pots, const double *kt,
                  const double *useqpe, double *d_qpe);
// With dotw, also d_dot[a] = dscale Re<in_a, dotw_a> (both components).
int apply_h_only_plane(tdw_ctx *c, int n, const double2 *in, double2 *out, const double *pots, const double *kt,
                       const double2 *dotw = nullptr, double dscale = 0.0, double *d_dot = nullptr);
// ASLDA: out = alpha_s psi_s (in place allowed)
int alpha_wf(tdw_ctx *c, int n, const double2 *in, double2 *out, const double *pots);
// ASLDA H application with the <H> subtraction: ddx..ddz the gradients of in
// along the evolved axes, lap its Laplacian, alap lap(alpha psi) (NULL: formed
// here).  out may equal lap or alap (pointwise), not in.
// subtract = false (two-pass form only, alap NULL on a supported lattice):
// out = H in without the subtraction, <H> in d_qpe
int apply_h_aslda(tdw_ctx *c, int n, const double2 *in, double2 *out, const double2 *ddx, const double2 *ddy,
                  const double2 *ddz, const double2 *lap, const double2 *alap, const double *pots, const double *kt,
                  const double *useqpe, double *d_qpe, bool subtract = true);
// the corrector with the predictor's <H> subtraction folded in:
// y = -i dt a0 (hy - qpe y) + fl
int abm4q(tdw_ctx *c, int n, const double2 *hy, double2 *y, const double2 *fl, const double *qpe);
// ASLDA H on 3D lattices in an x-pencil and a yz-plane pass (tdw_aslda.hip):
// out = H in without the subtraction, <H> into d_qpe; T state-sized scratch,
// hx the grid fields of aslda_fields, part n*NX doubles; with sc (n
// per-state factors) H acts on sc_a in_a
bool aslda_h_fused_supported(const tdw_ctx *c);
// hx = [lap alpha_a, lap alpha_b, div A_a, div A_b] of the potentials into c->d_hx
int aslda_fields(tdw_ctx *c, const double *pots);
int aslda_h_fused(tdw_ctx *c, int n, const double2 *in, double2 *out, double2 *T, const double *pots,
                  const double *hx, double *part, double *d_qpe, const double *sc = nullptr);
// The ASLDA step on 3D lattices of the two-pass kernels (FP64) with the ABM
// updates, normalisations and <H> subtractions carried as per-state scalars:
// it leaves y unnormalised and f[0] without its subtraction (c->pending = 2)
// until the next such step or fused_finalize applies them.
bool aslda_step_supported(const tdw_ctx *c);
int aslda_fused_step(tdw_ctx *c, tdw_state *s, int it, int grad_last);
// Densities with gradients (rho, nu, tau before its lap(rho) correction, j)
// on 3D lattices in an x and a yz pass that keep the gradients in LDS
// (tdw_aslda.hip); the same dens block as densities(..., grad = 1)
bool dens_grad_fused_supported(const tdw_ctx *c);
int dens_grad_fused(tdw_ctx *c, int n, const double2 *wf, const double *fbeta, const double *wgt, double *dens);
int energy(tdw_ctx *c, int it, const double *dens, const double *pots, double *h_e);
int taylor(tdw_ctx *c, int order, double dt, int n, const double2 *hpsi, double2 *upd, double2 *contr);

}  // namespace tdw
