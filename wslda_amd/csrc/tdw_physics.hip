// tdw_physics.hip -- pointwise and reduction kernels of the modular path:
// ABM predictor/corrector (tdwslda_ode_integrator.h:26-220), normalisation
// (pca_kernels.cu:913-976), densities (pca_dens.cu:13-198 and the 1D/2D
// twins), tau (pca_dens.cu:275-338), potentials (tdwslda_functionals.h),
// H application with the <H> subtraction (pca_kernels.cu:233-495, 552-594,
// 806-832) and the energy sums (tdwslda_energy_generic.h:21-110).
#include "tdw_functional.hpp"
#include "tdw_internal.hpp"
#include "tdw_physics.hpp"
#include "tdw_hooks.hpp"

namespace tdw {

static constexpr int NT = 256;

// ---------------------------------------------------------------------------
// ABM
// ---------------------------------------------------------------------------
template <int SCH>
__global__ void __launch_bounds__(NT)
k_abm1(long tot, double dt, const double2 *yin, double2 *y, double2 *cout, double2 *f1, double2 *f2, double2 *f3,
       double2 *f4, double2 *f5)
{
    long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= tot) return;
    const double2 yk = yin[i];
    const double2 a = f1[i], b = f2[i], d = f3[i];
    double2 sp, sc;
    if (SCH == 34) {
        sp = a * (23. / 12.) - b * (16. / 12.) + d * (5. / 12.);
        sc = a * (19. / 24.) - b * (5. / 24.) + d * (1. / 24.);
    } else if (SCH == 45) {
        const double2 e = f4[i];
        sp = a * (55. / 24.) - b * (59. / 24.) + d * (37. / 24.) - e * (9. / 24.);
        sc = a * (646. / 720.) - b * (264. / 720.) + d * (106. / 720.) - e * (19. / 720.);
    } else {
        const double2 e = f4[i], g = f5[i];
        sp = a * (1901. / 720.) - b * (2774. / 720.) + d * (2616. / 720.) - e * (1274. / 720.) + g * (251. / 720.);
        sc = a * (646. / 720.) - b * (264. / 720.) + d * (106. / 720.) - e * (19. / 720.);
    }
    // y + (-i dt) * s
    y[i] = yk + imul(sp, -dt);
    const double2 c = yk + imul(sc, -dt);
    if (cout) cout[i] = c;
    else if (SCH == 34) f3[i] = c;
    else if (SCH == 45) f4[i] = c;
    else f5[i] = c;
}

__global__ void __launch_bounds__(NT)
k_abm4(long tot, double dta0, const double2 *__restrict__ hy, double2 *__restrict__ y,
       const double2 *__restrict__ fl)
{
    long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= tot) return;
    // (-i dt) * hy * a0 + flast
    y[i] = imul(hy[i], -dta0) + fl[i];
}

int abm1_to(tdw_ctx *c, int n, const double2 *yin, double2 *y, double2 *cout, double2 *f1, double2 *f2, double2 *f3,
            double2 *f4, double2 *f5)
{
    TDW_TIMED(c, "abm_step1");
    long tot = 2L * n * (long)c->N;
    dim3 g((tot + NT - 1) / NT);
    double dt = c->p.dt;
#define TDW_A1(S) hipLaunchKernelGGL(k_abm1<S>, g, dim3(NT), 0, c->stream, tot, dt, yin, y, cout, f1, f2, f3, f4, f5)
    switch (c->p.scheme) {
    case TDW_AB3AM4: TDW_A1(34); break;
    case TDW_AB4AM5: TDW_A1(45); break;
    case TDW_AB5AM5: TDW_A1(55); break;
    default: return TDW_EINVAL;
    }
#undef TDW_A1
    return (int)hipGetLastError();
}

int abm1(tdw_ctx *c, int n, double2 *y, double2 *f1, double2 *f2, double2 *f3, double2 *f4, double2 *f5)
{
    return abm1_to(c, n, y, y, nullptr, f1, f2, f3, f4, f5);
}

// corrector with the <H> subtraction of the predictor's H folded in:
// y = -i dt a0 (hy - qpe y) + fl, y holding the normalised prediction
__global__ void __launch_bounds__(NT)
k_abm4q(long tot, double dta0, const double2 *__restrict__ hy, double2 *y, const double2 *__restrict__ fl,
        const double *__restrict__ qpe, int n, long N)
{
    long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= tot) return;
    const long a = (i / N) % n;
    y[i] = imul(hy[i] - y[i] * qpe[a], -dta0) + fl[i];
}

int abm4q(tdw_ctx *c, int n, const double2 *hy, double2 *y, const double2 *fl, const double *qpe)
{
    TDW_TIMED(c, "abm_step4q");
    long tot = 2L * n * (long)c->N;
    double a0 = (c->p.scheme == TDW_AB3AM4) ? 9. / 24. : 251. / 720.;
    hipLaunchKernelGGL(k_abm4q, dim3((tot + NT - 1) / NT), dim3(NT), 0, c->stream, tot, c->p.dt * a0, hy, y, fl, qpe, n,
                       (long)c->N);
    return (int)hipGetLastError();
}

int abm4(tdw_ctx *c, int n, const double2 *hy, double2 *y, const double2 *fl)
{
    TDW_TIMED(c, "abm_step4");
    long tot = 2L * n * (long)c->N;
    double a0 = (c->p.scheme == TDW_AB3AM4) ? 9. / 24. : 251. / 720.;
    hipLaunchKernelGGL(k_abm4, dim3((tot + NT - 1) / NT), dim3(NT), 0, c->stream, tot, c->p.dt * a0, hy, y, fl);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// per-state reductions: norms and <psi|H psi>
// grid (nchunk, n); block b of state a sums points [b*chunk, (b+1)*chunk).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NT)
k_norm_part(const double2 *__restrict__ wf, int n, long N, long chunk, double *__restrict__ part)
{
    __shared__ double red[NT / 64];
    const int a = blockIdx.y;
    const long i0 = (long)blockIdx.x * chunk, i1 = min(N, i0 + chunk);
    const double2 *u = wf + (long)a * N, *v = wf + ((long)n + a) * N;
    double s = 0.0;
    for (long i = i0 + threadIdx.x; i < i1; i += NT) s += cnorm(u[i]) + cnorm(v[i]);
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[(long)a * gridDim.x + blockIdx.x] = s;
}

__global__ void k_rowsum(const double *__restrict__ part, int n, int nchunk, double scale,
                         double *__restrict__ out)
{
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    double s = 0.0;
    for (int k = 0; k < nchunk; k++) s += part[(long)a * nchunk + k];
    out[a] = s * scale;
}

__global__ void __launch_bounds__(NT)
k_scale_states(double2 *__restrict__ wf, int n, long N, const double *__restrict__ nrm)
{
    long i = (long)blockIdx.x * NT + threadIdx.x;
    long tot = 2L * n * N;
    if (i >= tot) return;
    long a = (i / N) % n;
    double s = 1.0 / sqrt(nrm[a]);
    wf[i] = wf[i] * s;
}

static long pick_chunk(const tdw_ctx *c, int *nchunk)
{
    long N = (long)c->N;
    long chunk = 4096;
    if (chunk > N) chunk = N;
    *nchunk = (int)((N + chunk - 1) / chunk);
    return chunk;
}

int norm_chunks(const tdw_ctx *c)
{
    int nchunk;
    pick_chunk(c, &nchunk);
    return nchunk;
}

// every (chunk, state) slot of part is written before the row sums read it
int norms_with(tdw_ctx *c, int n, const double2 *wf, double *d_nrm, double *part)
{
    TDW_TIMED(c, "norms");
    int nchunk;
    long chunk = pick_chunk(c, &nchunk);
    hipLaunchKernelGGL(k_norm_part, dim3(nchunk, n), dim3(NT), 0, c->stream, wf, n, (long)c->N, chunk, part);
    hipLaunchKernelGGL(k_rowsum, dim3((n + 255) / 256), dim3(256), 0, c->stream, part, n, nchunk, dvol_norm(c),
                       d_nrm);
    return (int)hipGetLastError();
}

int normalize_with(tdw_ctx *c, int n, double2 *wf, double *d_nrm, double *part)
{
    TDW_TRY(norms_with(c, n, wf, d_nrm, part));
    long tot = 2L * n * (long)c->N;
    TDW_TIMED(c, "scale");
    hipLaunchKernelGGL(k_scale_states, dim3((tot + NT - 1) / NT), dim3(NT), 0, c->stream, wf, n, (long)c->N, d_nrm);
    return (int)hipGetLastError();
}

int norms(tdw_ctx *c, int n, const double2 *wf, double *d_nrm)
{
    TDW_TRY(ensure_part(c, (size_t)n * norm_chunks(c)));
    return norms_with(c, n, wf, d_nrm, c->d_part);
}

int normalize(tdw_ctx *c, int n, double2 *wf, double *d_nrm)
{
    TDW_TRY(ensure_part(c, (size_t)n * norm_chunks(c)));
    return normalize_with(c, n, wf, d_nrm, c->d_part);
}

// ---------------------------------------------------------------------------
// densities: partial sums over state chunks, then a fixed-order reduction.
// NV = 4 (rho_a, rho_b, nu) or 12 (+ tau_a, tau_b, j_a, j_b).
// ---------------------------------------------------------------------------
template <int DIM, bool GRAD, typename W = double2>
__global__ void __launch_bounds__(NT)
k_dens_part(const W *__restrict__ wf, const double2 *__restrict__ ddx, const double2 *__restrict__ ddy,
            const double2 *__restrict__ ddz, const double2 *__restrict__ lap, const double *__restrict__ fbeta,
            const double *__restrict__ wgt, const double *__restrict__ kt, const int *__restrict__ cnt,
            int n, long N, int chunk, int cubic, double piz, double *__restrict__ part)
{
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= N) return;
    const int a0 = blockIdx.y * chunk, a1 = min(n, a0 + chunk);
    double na = 0, nb = 0, ta = 0, tb = 0, jax = 0, jay = 0, jaz = 0, jbx = 0, jby = 0, jbz = 0;
    double2 nuu = make_double2(0.0, 0.0);
    for (int a = a0; a < a1; a++) {
        double fb = fbeta[a] * 10000.;
        double fbm = 10000. - fb;
        if (wgt) { fb *= wgt[a]; fbm *= wgt[a]; }
        double w = 1.0, k2t = 0.0;
        if (DIM == 1) {
            double ky = kt[a], kz = kt[n + a];
            k2t = kz * kz + ky * ky;
            w = (double)cnt[a];
        } else if (DIM == 2) {
            double kz = kt[a];
            k2t = kz * kz;
            w = (fabs(kz) < 1.0e-12) ? 1.0 : 2.0;
            if (cubic && fabs(kz + piz) < 1.0e-12) w = 1.0;
        }
        const double2 u = to64(wf[(long)a * N + i]), v = to64(wf[((long)n + a) * N + i]);
        const double fw = fb * w, fmw = fbm * w;
        na += cnorm(u) * fw;
        nb += cnorm(v) * fmw;
        double2 uv = cmul(u, make_double2(v.x, -v.y));
        nuu = nuu + uv * ((fbm - fb) / 2.0 * w);
        if (GRAD) {
            const double2 ux = ddx[(long)a * N + i], vx = ddx[((long)n + a) * N + i];
            double2 lu = lap[(long)a * N + i], lv = lap[((long)n + a) * N + i];
            if (DIM < 3) { lu = lu - u * k2t; lv = lv - v * k2t; }
            ta += cmulc(u, lu).x * fw;
            tb += cmulc(v, lv).x * fmw;
            jax += cmulc(u, ux).y * fw;
            jbx -= cmulc(v, vx).y * fmw;
            if (DIM >= 2) {
                const double2 uy = ddy[(long)a * N + i], vy = ddy[((long)n + a) * N + i];
                jay += cmulc(u, uy).y * fw;
                jby -= cmulc(v, vy).y * fmw;
            }
            if (DIM == 3) {
                const double2 uz = ddz[(long)a * N + i], vz = ddz[((long)n + a) * N + i];
                jaz += cmulc(u, uz).y * fw;
                jbz -= cmulc(v, vz).y * fmw;
            }
        }
    }
    const int NV = GRAD ? 12 : 4;
    double *p = part + (long)blockIdx.y * NV * N + i;
    p[0 * N] = na;
    p[1 * N] = nb;
    p[2 * N] = nuu.x;
    p[3 * N] = nuu.y;
    if (GRAD) {
        p[4 * N] = ta;  p[5 * N] = tb;
        p[6 * N] = jax; p[7 * N] = jay; p[8 * N] = jaz;
        p[9 * N] = jbx; p[10 * N] = jby; p[11 * N] = jbz;
    }
}

template <bool GRAD>
__global__ void __launch_bounds__(NT)
k_dens_final(const double *__restrict__ part, int nchunk, long N, double tr, double *__restrict__ dens)
{
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= N) return;
    const int NV = GRAD ? 12 : 4;
    double s[12];
#pragma unroll
    for (int v = 0; v < 12; v++) s[v] = 0.0;
    for (int k = 0; k < nchunk; k++) {
        const double *p = part + (long)k * NV * N + i;
#pragma unroll
        for (int v = 0; v < NV; v++) s[v] += p[v * N];
    }
    const double sc = 10000.;
    dens[2 * i] = s[2] / sc / tr;          // nu (complex, first 2N doubles)
    dens[2 * i + 1] = s[3] / sc / tr;
    dens[2 * N + i] = s[0] / sc / tr;      // rho_a
    dens[7 * N + i] = s[1] / sc / tr;      // rho_b
    dens[3 * N + i] = GRAD ? s[4] / sc / tr : 0.0;
    dens[8 * N + i] = GRAD ? s[5] / sc / tr : 0.0;
    dens[4 * N + i] = GRAD ? s[6] / sc / tr : 0.0;
    dens[5 * N + i] = GRAD ? s[7] / sc / tr : 0.0;
    dens[6 * N + i] = GRAD ? s[8] / sc / tr : 0.0;
    dens[9 * N + i] = GRAD ? s[9] / sc / tr : 0.0;
    dens[10 * N + i] = GRAD ? s[10] / sc / tr : 0.0;
    dens[11 * N + i] = GRAD ? s[11] / sc / tr : 0.0;
}

// spin-symmetric mode (SPINSYMMETRY_MODE, pca_dens.cu:55-58, 71-90, 112-118):
// the reference accumulates both components of every state into spin b, nu
// without the factor 1/2, and copies b to a.  The partial sums above keep the
// u and v contributions apart, so the fold is b = a + b for rho, tau_acc and
// j, nu *= 2 (exact), then a = b.
__global__ void k_spin_fold(double *__restrict__ dens, long N)
{
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= N) return;
    dens[2 * i] *= 2.0;
    dens[2 * i + 1] *= 2.0;
#pragma unroll
    for (int f = 0; f < 5; f++) {
        const double v = dens[(2 + f) * N + i] + dens[(7 + f) * N + i];
        dens[(2 + f) * N + i] = v;
        dens[(7 + f) * N + i] = v;
    }
}

int spin_fold(tdw_ctx *c, double *dens)
{
    if (!c->p.spinsymmetry) return 0;
    const long N = (long)c->N;
    hipLaunchKernelGGL(k_spin_fold, dim3((N + NT - 1) / NT), dim3(NT), 0, c->stream, dens, N);
    return (int)hipGetLastError();
}

// a <- b copy of n doubles (symmetrize_densities_device, tau in spin-symmetric mode)
__global__ void k_copy_f64(const double *__restrict__ src, double *__restrict__ dst, long n)
{
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

int copy_f64(tdw_ctx *c, const double *src, double *dst, long n)
{
    hipLaunchKernelGGL(k_copy_f64, dim3((n + NT - 1) / NT), dim3(NT), 0, c->stream, src, dst, n);
    return (int)hipGetLastError();
}

template <typename W>
static int densities_t(tdw_ctx *c, int n, const W *wf, const double2 *ddx, const double2 *ddy,
                       const double2 *ddz, const double2 *lap, const double *fbeta, const double *wgt,
                       const double *kt, const int *cnt, double *dens, int grad)
{
    TDW_TIMED(c, "densities");
    const long N = (long)c->N;
    int chunk = 128;
    long nblk = (N + NT - 1) / NT;
    // keep >= ~2048 blocks in flight
    while (chunk > 8 && nblk * ((n + chunk - 1) / chunk) < 2048) chunk >>= 1;
    int nchunk = (n + chunk - 1) / chunk;
    const int NV = grad ? 12 : 4;
    TDW_TRY(ensure_part(c, (size_t)nchunk * NV * N));
    double piz = M_PI / c->p.dz;
    int cubic = c->p.regularization == TDW_CUBIC_CUTOFF;
    dim3 g(nblk, nchunk);
#define TDW_DP(D, G)                                                                                      \
    hipLaunchKernelGGL((k_dens_part<D, G, W>), g, dim3(NT), 0, c->stream, wf, ddx, ddy, ddz, lap, fbeta, wgt, \
                       kt, cnt, n, N, chunk, cubic, piz, c->d_part)
    if (c->dim == 1) { if (grad) TDW_DP(1, true); else TDW_DP(1, false); }
    else if (c->dim == 2) { if (grad) TDW_DP(2, true); else TDW_DP(2, false); }
    else { if (grad) TDW_DP(3, true); else TDW_DP(3, false); }
#undef TDW_DP
    double tr = dens_transverse(c);
    if (grad) hipLaunchKernelGGL(k_dens_final<true>, dim3(nblk), dim3(NT), 0, c->stream, c->d_part, nchunk, N, tr, dens);
    else hipLaunchKernelGGL(k_dens_final<false>, dim3(nblk), dim3(NT), 0, c->stream, c->d_part, nchunk, N, tr, dens);
    TDW_TRY((int)hipGetLastError());
    return spin_fold(c, dens);
}

int densities(tdw_ctx *c, int n, const double2 *wf, const double2 *ddx, const double2 *ddy,
              const double2 *ddz, const double2 *lap, const double *fbeta, const double *wgt,
              const double *kt, const int *cnt, double *dens, int grad)
{
    return densities_t<double2>(c, n, wf, ddx, ddy, ddz, lap, fbeta, wgt, kt, cnt, dens, grad);
}

int densities32(tdw_ctx *c, int n, const float2 *wf, const double *fbeta, const double *wgt, const double *kt,
                const int *cnt, double *dens)
{
    return densities_t<float2>(c, n, wf, nullptr, nullptr, nullptr, nullptr, fbeta, wgt, kt, cnt, dens, 0);
}

// ---------------------------------------------------------------------------
// quantum-friction densities (pca_dens.cu:360-450, cpca_dens.cu:385-490,
// ccpca_dens.cu:384-500), with the weights f and 1 - f (no 1e4 scale) and
// L psi = lap psi - k_t^2 psi in 1D / 2D:
//   U_a = sum f w Im(u* L u),  U_b = -sum (1 - f) w Im(v* L v),
//   D = -1/2 sum (v* L u + (L v)* u)(1 - 2 f) w,
// each divided by the transverse length.  w: 1 (3D), 1 or 2 by k_z (2D, no
// cubic-cutoff exception in this kernel of the reference), cnt (1D).
// Spin-symmetric mode: U_b = U_a + U_b, D doubled, U_a = U_b.
// out: [U_a (N), U_b (N), D (N complex)].
// ---------------------------------------------------------------------------
template <int DIM>
__global__ void __launch_bounds__(NT)
k_qf_part(const double2 *__restrict__ wf, const double2 *__restrict__ lap, const double *__restrict__ fbeta,
          const double *__restrict__ wgt, const double *__restrict__ kt, const int *__restrict__ cnt, int n, long N,
          int chunk, double *__restrict__ part)
{
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= N) return;
    const int a0 = blockIdx.y * chunk, a1 = min(n, a0 + chunk);
    double ua = 0.0, ub = 0.0;
    double2 d = make_double2(0.0, 0.0);
    for (int a = a0; a < a1; a++) {
        double fb = fbeta[a], fbm = 1.0 - fb;
        if (wgt) { fb *= wgt[a]; fbm *= wgt[a]; }
        double w = 1.0, k2t = 0.0;
        if (DIM == 1) {
            const double ky = kt[a], kz = kt[n + a];
            k2t = kz * kz + ky * ky;
            w = (double)cnt[a];
        } else if (DIM == 2) {
            const double kz = kt[a];
            k2t = kz * kz;
            w = (fabs(kz) < 1.0e-12) ? 1.0 : 2.0;
        }
        const double2 u = wf[(long)a * N + i], v = wf[((long)n + a) * N + i];
        double2 lu = lap[(long)a * N + i], lv = lap[((long)n + a) * N + i];
        if (DIM < 3) { lu = lu - u * k2t; lv = lv - v * k2t; }
        ua += cmulc(u, lu).y * fb * w;
        ub -= cmulc(v, lv).y * fbm * w;
        d = d + (cmulc(v, lu) + cmulc(lv, u)) * ((fbm - fb) * w);
    }
    double *p = part + (long)blockIdx.y * 4 * N + i;
    p[0] = ua;
    p[N] = ub;
    p[2 * N] = d.x;
    p[3 * N] = d.y;
}

__global__ void __launch_bounds__(NT)
k_qf_final(const double *__restrict__ part, int nchunk, long N, double tr, int ss, double *__restrict__ out)
{
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= N) return;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < nchunk; k++) {
        const double *p = part + (long)k * 4 * N + i;
#pragma unroll
        for (int v = 0; v < 4; v++) s[v] += p[v * N];
    }
    if (ss) {
        s[1] += s[0];
        s[0] = s[1];
        s[2] *= 2.0;
        s[3] *= 2.0;
    }
    out[i] = s[0] / tr;
    out[N + i] = s[1] / tr;
    out[2 * N + 2 * i] = -0.5 * s[2] / tr;
    out[2 * N + 2 * i + 1] = -0.5 * s[3] / tr;
}

int qf_densities(tdw_ctx *c, int n, const double2 *wf, const double2 *lap, const double *fbeta, const double *wgt,
                 const double *kt, const int *cnt, double *out)
{
    TDW_TIMED(c, "qf_densities");
    const long N = (long)c->N;
    int chunk = 128;
    const long nblk = (N + NT - 1) / NT;
    while (chunk > 8 && nblk * ((n + chunk - 1) / chunk) < 2048) chunk >>= 1;
    const int nchunk = (n + chunk - 1) / chunk;
    TDW_TRY(ensure_part(c, (size_t)nchunk * 4 * N));
    dim3 g(nblk, nchunk);
    if (c->dim == 1)
        hipLaunchKernelGGL(k_qf_part<1>, g, dim3(NT), 0, c->stream, wf, lap, fbeta, wgt, kt, cnt, n, N, chunk, c->d_part);
    else if (c->dim == 2)
        hipLaunchKernelGGL(k_qf_part<2>, g, dim3(NT), 0, c->stream, wf, lap, fbeta, wgt, kt, cnt, n, N, chunk, c->d_part);
    else
        hipLaunchKernelGGL(k_qf_part<3>, g, dim3(NT), 0, c->stream, wf, lap, fbeta, wgt, kt, cnt, n, N, chunk, c->d_part);
    hipLaunchKernelGGL(k_qf_final, dim3(nblk), dim3(NT), 0, c->stream, c->d_part, nchunk, N, dens_transverse(c),
                       c->p.spinsymmetry, out);
    return (int)hipGetLastError();
}

// Friction terms into the mean fields before H and their removal after it
// (pca_kernels.cu:497-539, Eq. (3) of arXiv:1305.6891):
//   V_s -= U_s qfalpha / n_F,
//   Delta += qfbeta |Delta| e^{i phi_Delta} sin(phi_Delta - phi_D) + i qfgamma Delta,
// with the original V_a, V_b, Delta saved over the friction densities (as
// the reference does) and copied back by k_qf_remove.
__global__ void k_qf_add(double *__restrict__ qf, double *__restrict__ pots, long N, double qa, double qb,
                         double qg, double nF)
{
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= N) return;
    double2 *dl = reinterpret_cast<double2 *>(pots + 2 * N);
    double2 *dq = reinterpret_cast<double2 *>(qf + 2 * N);
    const double2 del = dl[i], dd = dq[i];
    const double phd = atan2(del.y, del.x), phn = atan2(dd.y, dd.x);
    const double va = pots[i], vb = pots[N + i];
    pots[i] = va - qf[i] * qa / nF;
    pots[N + i] = vb - qf[N + i] * qa / nF;
    double2 nd = del + make_double2(cos(phd), sin(phd)) * (qb * hypot(del.x, del.y) * sin(phd - phn));
    nd = nd + cmul(make_double2(0.0, qg), del);
    dl[i] = nd;
    qf[i] = va;
    qf[N + i] = vb;
    dq[i] = del;
}
__global__ void k_qf_remove(const double *__restrict__ qf, double *__restrict__ pots, long N)
{
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= N) return;
    pots[i] = qf[i];
    pots[N + i] = qf[N + i];
    pots[2 * N + 2 * i] = qf[2 * N + 2 * i];
    pots[2 * N + 2 * i + 1] = qf[2 * N + 2 * i + 1];
}

int qf_add(tdw_ctx *c, double *qf, double *pots, double qfswitch, double pccoeff)
{
    const long N = (long)c->N;
    const double kF = c->p.kF;
    const double nF = kF * kF * kF / (3. * M_PI * M_PI);
    hipLaunchKernelGGL(k_qf_add, dim3((N + NT - 1) / NT), dim3(NT), 0, c->stream, qf, pots, N,
                       c->p.qfalpha * qfswitch, c->p.qfbeta * qfswitch, c->p.qfgamma * qfswitch * pccoeff, nF);
    return (int)hipGetLastError();
}
int qf_remove(tdw_ctx *c, const double *qf, double *pots)
{
    const long N = (long)c->N;
    hipLaunchKernelGGL(k_qf_remove, dim3((N + NT - 1) / NT), dim3(NT), 0, c->stream, qf, pots, N);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// tau = max(0, lap(rho)/2 - tau_acc)
// ---------------------------------------------------------------------------
__global__ void k_real2cplx(const double *__restrict__ f, double2 *__restrict__ z, long N)
{
    long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i < N) z[i] = make_double2(f[i], 0.0);
}
__global__ void k_cplx_re(const double2 *__restrict__ z, double *__restrict__ f, long N)
{
    long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i < N) f[i] = z[i].x;
}
__global__ void k_tau(const double2 *__restrict__ lr, double *__restrict__ tau, long N)
{
    long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= N) return;
    double t = 0.5 * lr[i].x - tau[i];
    if (t < 0.0) t = 0.0;
    tau[i] = t;
}

int laplace_real_c(tdw_ctx *c, const double *f, double2 *out_c)
{
    long N = (long)c->N;
    hipLaunchKernelGGL(k_real2cplx, dim3((N + NT - 1) / NT), dim3(NT), 0, c->stream, f, c->d_rtmp, N);
    return laplace(c, 1, c->d_rtmp, out_c);
}

int laplace_real(tdw_ctx *c, const double *f, double *lap)
{
    long N = (long)c->N;
    TDW_TRY(ensure_T(c, (size_t)N));
    TDW_TRY(laplace_real_c(c, f, c->d_T));
    hipLaunchKernelGGL(k_cplx_re, dim3((N + NT - 1) / NT), dim3(NT), 0, c->stream, c->d_T, lap, N);
    return (int)hipGetLastError();
}

int density_tau(tdw_ctx *c, double *dens)
{
    TDW_TIMED(c, "tau");
    TimingOff quiet(c);
    long N = (long)c->N;
    TDW_TRY(ensure_T(c, (size_t)N));
    // b component first, then a (pca_dens.cu:317-333); in spin-symmetric
    // mode tau_a is a copy of tau_b (pca_dens.cu:329-331)
    TDW_TRY(laplace_real_c(c, dens + 7 * N, c->d_T));
    hipLaunchKernelGGL(k_tau, dim3((N + NT - 1) / NT), dim3(NT), 0, c->stream, c->d_T, dens + 8 * N, N);
    if (c->p.spinsymmetry) return copy_f64(c, dens + 8 * N, dens + 3 * N, N);
    TDW_TRY(laplace_real_c(c, dens + 2 * N, c->d_T));
    hipLaunchKernelGGL(k_tau, dim3((N + NT - 1) / NT), dim3(NT), 0, c->stream, c->d_T, dens + 3 * N, N);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// external potential of the built-in problem
// ---------------------------------------------------------------------------
static double switch_function(double t, double T, double alpha)
{
    return 0.5 * (1.0 + tanh(alpha * tan(M_PI_2 * (2.0 * t / T - 1.0))));
}
static double smooth_step(double t, double t0, double t1, double T, double alpha)
{
    if (t <= t0 || t >= t1) return 0.0;
    if (t >= t0 + T && t <= t1 - T) return 1.0;
    if (t > t0 && t < t0 + T) return switch_function(t - t0, T, alpha);
    return 1.0 - switch_function(t - t1 + T, T, alpha);
}

VextDev make_vext(const tdw_ctx *c, int it)
{
    VextDev V;
    V.v = c->p.vext;
    double t = c->p.t0 + c->p.dt * it;
    V.s = c->p.vext.switch_on ? smooth_step(t, c->p.vext.t_start, c->p.vext.t_stop, c->p.vext.t_switch, 1.0) : 1.0;
    V.dim = c->dim; V.nx = c->nx; V.ny = c->dim >= 2 ? c->ny : 1; V.nz = c->dim >= 3 ? c->nz : 1;
    V.dx = c->p.dx; V.dy = c->p.dy; V.dz = c->p.dz;
    V.grid = nullptr;
    V.npts = (long)c->N;
    if (c->proj) {
        // the problem's v_ext replaces the built-in family; without
        // ENABLE_V_EXT the external potential is zero (pca_kernels.cu:32-36)
        V.v.ngauss = 0;
        V.v.omega = 0.0;
        V.grid = vext_grid(c, it);
    }
    return V;
}

__global__ void k_vext(VextDev V, long N, double *__restrict__ va, double *__restrict__ vb)
{
    long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= N) return;
    double a, b;
    vext_point(V, i, a, b);
    va[i] = a;
    vb[i] = b;
}

int vext_eval(tdw_ctx *c, int it, double *va, double *vb)
{
    long N = (long)c->N;
    hipLaunchKernelGGL(k_vext, dim3((N + NT - 1) / NT), dim3(NT), 0, c->stream, make_vext(c, it), N, va, vb);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// potentials
// ---------------------------------------------------------------------------
FParams fparams(const tdw_ctx *c)
{
    FParams f;
    f.p_nmin = c->p.p_nmin; f.p_nmax = c->p.p_nmax; f.denseps = c->p.denseps;
    f.mu_a = c->p.mu_a; f.mu_b = c->p.mu_b; f.ec = c->p.ec; f.dx = c->p.dx;
    f.cubic = c->p.regularization == TDW_CUBIC_CUTOFF;
    if (f.cubic) { f.R0 = 2.442749607806335 / (4.0 * M_PI * M_PI); f.R1 = 0.31786000; }
    else { f.R0 = 1. / (2.0 * M_PI * M_PI); f.R1 = 0.5; }
    f.mf = c->p.mf_correction;
    f.bdg = c->p.functional == TDW_BDG;
    f.aslda = c->p.functional == TDW_ASLDA;
    f.sldae = c->p.functional == TDW_SLDAE;
    f.a1 = c->p.force_a1 != 0;
    f.sclgth = c->p.sclgth;
    f.agrid = nullptr;
    return f;
}

__global__ void __launch_bounds__(NT)
k_potentials(FParams f0, VextDev V, long N, int dim, double cc, const double *__restrict__ dens,
             double *__restrict__ pots)
{
    long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= N) return;
    const double na = dens[2 * N + i], nb = dens[7 * N + i];
    const double2 nu = make_double2(dens[2 * i], dens[2 * i + 1]);
    double ua, ub;
    vext_point(V, i, ua, ub);
    FParams f = f0;
    if (f0.agrid) f.sclgth = f0.agrid[i];
    double Va = pots[i], Vb = pots[N + i];
    double2 d;
    if (f.aslda || f.sldae) {
        double ja[3] = {0.0, 0.0, 0.0}, jb[3] = {0.0, 0.0, 0.0}, Aa[3], Ab[3], ala, alb;
#pragma unroll
        for (int k = 0; k < 3; k++)
            if (k < dim) {
                ja[k] = dens[(4 + k) * N + i];
                jb[k] = dens[(9 + k) * N + i];
            }
        if (f.sldae)
            sldae_point(f, na, nb, nu, dens[3 * N + i], dens[8 * N + i], ja, jb, dim, cc, ua, ub, Va, Vb, d, ala,
                        alb, Aa, Ab);
        else
            aslda_point(f, na, nb, nu, dens[3 * N + i], dens[8 * N + i], ja, jb, dim, cc, ua, ub, Va, Vb, d, ala,
                        alb, Aa, Ab);
        // the effective masses and vector potentials are saved only with
        // CURRENT_CORRECTIONS (not in SLDA_FORCE_A1 builds)
        if (!(f.sldae && f.a1)) {
            pots[4 * N + i] = ala;
            pots[5 * N + i] = alb;
#pragma unroll
            for (int k = 0; k < 3; k++)
                if (k < dim) {
                    pots[(6 + k) * N + i] = Aa[k];
                    pots[(9 + k) * N + i] = Ab[k];
                }
        }
    } else if (f.bdg) bdg_point(f, na, nb, nu, ua, ub, pots[4 * N + i], pots[5 * N + i], Va, Vb, d);
    else slda_point(f, na, nb, nu, ua, ub, Va, Vb, d);
    pots[i] = Va;
    pots[N + i] = Vb;
    pots[2 * N + 2 * i] = d.x;
    pots[2 * N + 2 * i + 1] = d.y;
}

int potentials(tdw_ctx *c, int it, const double *dens, double *pots, double cccoeff)
{
    TDW_TIMED(c, "potentials");
    long N = (long)c->N;
    FParams f = fparams(c);
    f.agrid = sclgth_grid(c, it);
    hipLaunchKernelGGL(k_potentials, dim3((N + NT - 1) / NT), dim3(NT), 0, c->stream, f, make_vext(c, it), N,
                       c->dim, cccoeff, dens, pots);
    TDW_HIP(hipGetLastError());
    return project_modify_potentials(c, it, dens, pots);
}

// ---------------------------------------------------------------------------
// H application: out = H in ; part = per-block Re<in|H in> ; then <H> and
// the subtraction out -= <H> in.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NT)
k_apply_h(const double2 *__restrict__ in, double2 *out, const double2 *lap, const double *__restrict__ pots,
          const double *__restrict__ kt, int n, long N, long chunk, int dim, int bdg, double mu_a,
          double mu_b, double *__restrict__ part)
{
    __shared__ double red[NT / 64];
    const int a = blockIdx.y;
    const long i0 = (long)blockIdx.x * chunk, i1 = min(N, i0 + chunk);
    double k2t = 0.0;
    if (dim == 1) k2t = kt[n + a] * kt[n + a] + kt[a] * kt[a];
    else if (dim == 2) k2t = kt[a] * kt[a];
    const long ou = (long)a * N, ov = ((long)n + a) * N;
    double s = 0.0;
    for (long i = i0 + threadIdx.x; i < i1; i += NT) {
        const double aa = bdg ? -0.5 : pots[4 * N + i] * -0.5;
        const double ab = bdg ? -0.5 : pots[5 * N + i] * -0.5;
        const double ja = -1.0 * aa * k2t, jb = -1.0 * ab * k2t;
        const double Va = pots[i] - mu_a + ja, Vb = pots[N + i] - mu_b + jb;
        const double2 D = make_double2(pots[2 * N + 2 * i], pots[2 * N + 2 * i + 1]);
        const double2 tu = in[ou + i], tv = in[ov + i];
        const double2 lu = lap[ou + i], lv = lap[ov + i];
        double2 u = tu * Va + cmul(tv, D) + lu * aa;
        double2 v = tv * (-Vb) + cmulc(D, tu) - lv * ab;
        out[ou + i] = u;
        out[ov + i] = v;
        s += cmulc(tu, u).x + cmulc(tv, v).x;
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[(long)a * gridDim.x + blockIdx.x] = s;
}

__global__ void __launch_bounds__(NT)
k_sub_qpe(const double2 *__restrict__ in, double2 *__restrict__ out, int n, long N, const double *__restrict__ qpe)
{
    long i = (long)blockIdx.x * NT + threadIdx.x;
    long tot = 2L * n * N;
    if (i >= tot) return;
    long a = (i / N) % n;
    out[i] = out[i] - in[i] * qpe[a];
}

int apply_h(tdw_ctx *c, int it, int n, const double2 *in, double2 *out, const double2 *lap,
            const double *pots, const double *kt, const double *useqpe, double *d_qpe)
{
    TDW_TIMED(c, "apply_h");
    (void)it;
    const long N = (long)c->N;
    int nchunk;
    long chunk = pick_chunk(c, &nchunk);
    TDW_TRY(ensure_part(c, (size_t)n * nchunk));
    hipLaunchKernelGGL(k_apply_h, dim3(nchunk, n), dim3(NT), 0, c->stream, in, out, lap, pots, kt, n, N,
                       chunk, c->dim, c->p.functional == TDW_BDG, c->p.mu_a, c->p.mu_b, c->d_part);
    const double *q = useqpe;
    if (!q) {
        hipLaunchKernelGGL(k_rowsum, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d_part, n, nchunk,
                           dvol_norm(c), d_qpe);
        q = d_qpe;
    } else if (d_qpe && d_qpe != useqpe) {
        TDW_HIP(hipMemcpyAsync(d_qpe, useqpe, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    }
    long tot = 2L * n * N;
    hipLaunchKernelGGL(k_sub_qpe, dim3((tot + NT - 1) / NT), dim3(NT), 0, c->stream, in, out, n, N, q);
    return (int)hipGetLastError();
}

// qpe[a] = dv (part[a] + part[n + a]): <H> from the per-plane sums of h2d
__global__ void k_plane_pairs(const double *__restrict__ part, int n, double dv, double *__restrict__ qpe)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < n) qpe[a] = dv * (part[a] + part[n + a]);
}

// apply_h for 2D lattices with a one-pass plane kernel: the Laplacian of in
// is formed in LDS and not stored (lap of apply_h is not needed)
int apply_h_plane(tdw_ctx *c, int it, int n, const double2 *in, double2 *out, const double *pots, const double *kt,
                  const double *useqpe, double *d_qpe)
{
    (void)it;
    TDW_TRY(ensure_part(c, (size_t)2 * n));
    if (useqpe) {
        // a known <H>: subtracted inside the plane pass
        TDW_TRY(h2d(c, n, in, out, pots, kt, c->d_part, nullptr, nullptr, useqpe));
        if (d_qpe && d_qpe != useqpe)
            TDW_HIP(hipMemcpyAsync(d_qpe, useqpe, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
        return 0;
    }
    TDW_TRY(h2d(c, n, in, out, pots, kt, c->d_part));
    TDW_TIMED(c, "apply_h");
    hipLaunchKernelGGL(k_plane_pairs, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d_part, n, dvol_norm(c),
                       d_qpe);
    const long N = (long)c->N, tot = 2L * n * N;
    hipLaunchKernelGGL(k_sub_qpe, dim3((tot + NT - 1) / NT), dim3(NT), 0, c->stream, in, out, n, N, d_qpe);
    return (int)hipGetLastError();
}

int apply_h_only_plane(tdw_ctx *c, int n, const double2 *in, double2 *out, const double *pots, const double *kt,
                       const double2 *dotw, double dscale, double *d_dot)
{
    TDW_TRY(ensure_part(c, (size_t)4 * n));
    TDW_TRY(h2d(c, n, in, out, pots, kt, c->d_part, dotw, dotw ? c->d_part + 2 * n : nullptr));
    if (dotw)
        hipLaunchKernelGGL(k_plane_pairs, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d_part + 2 * n, n, dscale,
                           d_dot);
    return (int)hipGetLastError();
}

// out = H in, without the <H> sums and subtraction (the adjoint applies H to
// its own vectors and subtracts in its next pass)
int apply_h_only(tdw_ctx *c, int n, const double2 *in, double2 *out, const double2 *lap, const double *pots,
                 const double *kt)
{
    TDW_TIMED(c, "apply_h");
    const long N = (long)c->N;
    int nchunk;
    long chunk = pick_chunk(c, &nchunk);
    TDW_TRY(ensure_part(c, (size_t)n * nchunk));
    hipLaunchKernelGGL(k_apply_h, dim3(nchunk, n), dim3(NT), 0, c->stream, in, out, lap, pots, kt, n, N, chunk, c->dim,
                       c->p.functional == TDW_BDG, c->p.mu_a, c->p.mu_b, c->d_part);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// ASLDA: Hamiltonian with the density-dependent effective mass and the
// current terms (pca_kernels.cu:233-391 and 674-822 without
// FAST_CONST_EFFECTIVE_MASS_MODE, with CURRENT_CORRECTIONS; 1D / 2D twins
// ccpca_kernels.cu:209-350, cpca_kernels.cu:231-385):
//   (H psi)_u = (V_a - mu_a + alpha_a k_t^2/2 - i/2 div A_a) u + Delta v
//               + lap(alpha_a) u / 4 - i A_a . grad u
//               - alpha_a lap(u) / 4 - lap(alpha_a u) / 4
// and the conjugate structure for v, then the <H> subtraction.
// ---------------------------------------------------------------------------
// out = alpha_s psi_s (multiply_wf_by_alpha, pca_kernels.cu:981-1026); in
// place allowed
__global__ void __launch_bounds__(NT)
k_alpha_wf(const double2 *in, double2 *out, const double *__restrict__ pots, int n, long N)
{
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    const long tot = 2L * n * N;
    if (i >= tot) return;
    const long j = i % N;
    out[i] = in[i] * pots[(i < (long)n * N ? 4 : 5) * N + j];
}

int alpha_wf(tdw_ctx *c, int n, const double2 *in, double2 *out, const double *pots)
{
    TDW_TIMED(c, "alpha_wf");
    const long tot = 2L * n * (long)c->N;
    hipLaunchKernelGGL(k_alpha_wf, dim3((tot + NT - 1) / NT), dim3(NT), 0, c->stream, in, out, pots, n, (long)c->N);
    return (int)hipGetLastError();
}

// acc = Re(dd) (first) or acc += Re(dd)
__global__ void k_re_acc(const double2 *__restrict__ dd, double *__restrict__ acc, long N, int first)
{
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= N) return;
    acc[i] = first ? dd[i].x : acc[i] + dd[i].x;
}

static int ensure_aslda(tdw_ctx *c)
{
    const size_t N = c->N;
    if (!c->d_hx) TDW_HIP(hipMalloc(&c->d_hx, sizeof(double) * 4 * N));
    if (!c->d_ascr) TDW_HIP(hipMalloc(&c->d_ascr, sizeof(double2) * 4 * N));
    return 0;
}

// hx = [lap alpha_a, lap alpha_b, div A_a, div A_b] of the potentials
// (compute_laplace_real_f and compute_derivative_real_vector_f with the
// Nyquist derivative set to zero, pca_kernels.cu:700-760)
int aslda_fields(tdw_ctx *c, const double *pots)
{
    TDW_TRY(ensure_aslda(c));
    TDW_TIMED(c, "aslda_fields");
    TimingOff quiet(c);
    const long N = (long)c->N;
    TDW_TRY(laplace_real(c, pots + 4 * N, c->d_hx));
    TDW_TRY(laplace_real(c, pots + 5 * N, c->d_hx + N));
    double2 *g = c->d_ascr;
    for (int sp = 0; sp < 2; sp++) {
        const double *A = pots + (sp ? 9 : 6) * N;
        double *acc = c->d_hx + (2 + sp) * N;
        for (int d = 0; d < c->dim; d++) {
            hipLaunchKernelGGL(k_real2cplx, dim3((N + NT - 1) / NT), dim3(NT), 0, c->stream, A + d * N, c->d_rtmp, N);
            TDW_TRY(derivatives(c, 1, c->d_rtmp, g, g + N, g + 2 * N, g + 3 * N));
            hipLaunchKernelGGL(k_re_acc, dim3((N + NT - 1) / NT), dim3(NT), 0, c->stream, g + d * N, acc, N, d == 0);
        }
    }
    return (int)hipGetLastError();
}

__global__ void __launch_bounds__(NT)
k_apply_h_aslda(const double2 *__restrict__ in, double2 *out, const double2 *__restrict__ gx,
                const double2 *__restrict__ gy, const double2 *__restrict__ gz, const double2 *lap,
                const double2 *alap, const double *__restrict__ pots, const double *__restrict__ hx,
                const double *__restrict__ kt, int n, long N, long chunk, int dim, double mu_a, double mu_b,
                double *__restrict__ part)
{
    __shared__ double red[NT / 64];
    const int a = blockIdx.y;
    const long i0 = (long)blockIdx.x * chunk, i1 = min(N, i0 + chunk);
    double k2t = 0.0;
    if (dim == 1) k2t = kt[n + a] * kt[n + a] + kt[a] * kt[a];
    else if (dim == 2) k2t = kt[a] * kt[a];
    const long ou = (long)a * N, ov = ((long)n + a) * N;
    const double2 *G[3] = {gx, gy, gz};
    double s = 0.0;
    for (long i = i0 + threadIdx.x; i < i1; i += NT) {
        const double aa = pots[4 * N + i] * -0.25, ab = pots[5 * N + i] * -0.25;
        const double la = 0.25 * hx[i], lb = 0.25 * hx[N + i];
        const double ca = -0.5 * hx[2 * N + i], cb = -0.5 * hx[3 * N + i];
        const double ja = -2.0 * aa * k2t, jb = -2.0 * ab * k2t;
        const double2 D = make_double2(pots[2 * N + 2 * i], pots[2 * N + 2 * i + 1]);
        const double2 tu = in[ou + i], tv = in[ov + i];
        const double2 lu = lap[ou + i], lv = lap[ov + i];
        const double2 alu = alap[ou + i], alv = alap[ov + i];
        double2 u = cmul(tu, make_double2(pots[i] - mu_a + ja, ca));
        double2 v = make_double2(0.0, 0.0) - cmul(tv, make_double2(pots[N + i] - mu_b + jb, -1.0 * cb));
        u = u + cmul(tv, D);
        v = v + cmulc(D, tu);
        u = u + tu * la;
        v = v - tv * lb;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (d >= dim) break;
            const double Aa = pots[(6 + d) * N + i], Ab = pots[(9 + d) * N + i];
            u = u + cmul(G[d][ou + i], make_double2(0.0, -1.0 * Aa));
            v = v - cmul(G[d][ov + i], make_double2(0.0, Ab));
        }
        u = u + lu * aa;
        v = v - lv * ab;
        u = u + alu * (-0.25);
        v = v - alv * (-0.25);
        out[ou + i] = u;
        out[ov + i] = v;
        s += cmulc(tu, u).x + cmulc(tv, v).x;
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[(long)a * gridDim.x + blockIdx.x] = s;
}

int apply_h_aslda(tdw_ctx *c, int n, const double2 *in, double2 *out, const double2 *ddx, const double2 *ddy,
                  const double2 *ddz, const double2 *lap, const double2 *alap, const double *pots, const double *kt,
                  const double *useqpe, double *d_qpe, bool subtract)
{
    const long N = (long)c->N;
    if (!alap && aslda_h_fused_supported(c)) {
        // two passes from psi, alpha and A (the context's state-sized array
        // holds the x terms)
        const size_t need = (size_t)2 * n * N;
        if (need > c->aw_cap) {
            TDW_HIP(hipStreamSynchronize(c->stream));
            if (c->d_aw) TDW_HIP(hipFree(c->d_aw));
            c->d_aw = nullptr;
            c->aw_cap = 0;
            TDW_HIP(hipMalloc(&c->d_aw, need * sizeof(double2)));
            c->aw_cap = need;
        }
        TDW_TRY(aslda_fields(c, pots));
        TDW_TRY(ensure_part(c, (size_t)n * c->nx + n));
        double *qcalc = useqpe ? c->d_part + (size_t)n * c->nx : d_qpe;
        TDW_TRY(aslda_h_fused(c, n, in, out, c->d_aw, pots, c->d_hx, c->d_part, qcalc));
        const double *q = useqpe ? useqpe : d_qpe;
        if (useqpe && d_qpe && d_qpe != useqpe)
            TDW_HIP(hipMemcpyAsync(d_qpe, useqpe, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
        if (!subtract) return (int)hipGetLastError();
        TDW_TIMED(c, "sub_qpe");
        const long tot = 2L * n * N;
        hipLaunchKernelGGL(k_sub_qpe, dim3((tot + NT - 1) / NT), dim3(NT), 0, c->stream, in, out, n, N, q);
        return (int)hipGetLastError();
    }
    if (!alap) {
        // lap(alpha psi) into the context's array
        const size_t need = (size_t)2 * n * N;
        if (need > c->aw_cap) {
            TDW_HIP(hipStreamSynchronize(c->stream));
            if (c->d_aw) TDW_HIP(hipFree(c->d_aw));
            c->d_aw = nullptr;
            c->aw_cap = 0;
            TDW_HIP(hipMalloc(&c->d_aw, need * sizeof(double2)));
            c->aw_cap = need;
        }
        TDW_TRY(alpha_wf(c, n, in, c->d_aw, pots));
        TDW_TRY(laplace(c, 2 * n, c->d_aw, c->d_aw));
        alap = c->d_aw;
    }
    TDW_TRY(aslda_fields(c, pots));
    TDW_TIMED(c, "apply_h");
    int nchunk;
    long chunk = pick_chunk(c, &nchunk);
    TDW_TRY(ensure_part(c, (size_t)n * nchunk));
    hipLaunchKernelGGL(k_apply_h_aslda, dim3(nchunk, n), dim3(NT), 0, c->stream, in, out, ddx,
                       c->dim >= 2 ? ddy : ddx, c->dim >= 3 ? ddz : ddx, lap, alap, pots, c->d_hx, kt, n, N, chunk,
                       c->dim, c->p.mu_a, c->p.mu_b, c->d_part);
    const double *q = useqpe;
    if (!q) {
        hipLaunchKernelGGL(k_rowsum, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d_part, n, nchunk,
                           dvol_norm(c), d_qpe);
        q = d_qpe;
    } else if (d_qpe && d_qpe != useqpe) {
        TDW_HIP(hipMemcpyAsync(d_qpe, useqpe, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    }
    const long tot = 2L * n * N;
    hipLaunchKernelGGL(k_sub_qpe, dim3((tot + NT - 1) / NT), dim3(NT), 0, c->stream, in, out, n, N, q);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// energy
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NT)
k_energy(FParams f0, VextDev V, long N, int dim, double VE, const double *__restrict__ dens,
         const double *__restrict__ pots, const double2 *__restrict__ dext, const double *__restrict__ vel,
         double *__restrict__ pp, double *__restrict__ part)
{
    __shared__ double red[NT / 64];
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    double e[TDW_NITEMS];
#pragma unroll
    for (int k = 0; k < TDW_NITEMS; k++) e[k] = 0.0;
    if (i < N) {
        FParams f = f0;
        if (f0.agrid) f.sclgth = f0.agrid[i];
        const double na = dens[2 * N + i], nb = dens[7 * N + i];
        const double tx1 = dens[4 * N + i], tx2 = dens[9 * N + i];
        const double ty1 = dim >= 2 ? dens[5 * N + i] : 0.0, ty2 = dim >= 2 ? dens[10 * N + i] : 0.0;
        const double tz1 = dim >= 3 ? dens[6 * N + i] : 0.0, tz2 = dim >= 3 ? dens[11 * N + i] : 0.0;
        const double ja2 = tx1 * tx1 + ty1 * ty1 + tz1 * tz1, jb2 = tx2 * tx2 + ty2 * ty2 + tz2 * tz2;
        const double ta = dens[3 * N + i] - p_reg(f, na) * ja2 / na;
        const double tb = dens[8 * N + i] - p_reg(f, nb) * jb2 / nb;
        if (f.sldae) {
            // tdwslda_functionals.h:815-889: A (tau~_a + tau~_b)/2, 3/5 B e_F n
            const double nt = na + nb;
            const double kF = pow(3. * M_PI * M_PI * nt, 1. / 3.);
            const double x = fabs(f.sclgth * kF);
            e[TDW_EKIN] = sldae::alpha0(x, f.a1) * (ta + tb) / 2. * VE;
            e[TDW_EPOT] = (3. / 5.) * sldae::bfun0(x, f.a1) * (kF * kF / 2.) * nt * VE;
        } else if (f.aslda) {
            const double p = polar(f, na, nb);
            e[TDW_EKIN] = 0.5 * (aslda::alpha(p) * ta + aslda::alpha(-1.0 * p) * tb) * VE;
            e[TDW_EPOT] = aslda::funD(na + nb, p) * VE;
        } else if (!f.bdg) {
            const double p = polar(f, na, nb);
            e[TDW_EKIN] = 0.5 * (alpha_f(p) * ta + alpha_f(-1.0 * p) * tb) * VE;
            e[TDW_EPOT] = funD(f, na, nb) * VE;
        } else {
            e[TDW_EKIN] = 0.5 * (pots[4 * N + i] * ta + pots[5 * N + i] * tb) * VE;
        }
        const double2 d = make_double2(pots[2 * N + 2 * i], pots[2 * N + 2 * i + 1]);
        const double2 nu = make_double2(dens[2 * i], dens[2 * i + 1]);
        e[TDW_EPAIR] = (d.x * nu.x + d.y * nu.y) * (-1.0) * VE;
        e[TDW_ECURRENT] = p_reg(f, na) * ja2 / (2. * na) * VE + p_reg(f, nb) * jb2 / (2. * nb) * VE;
        double ua, ub;
        vext_point(V, i, ua, ub);
        e[TDW_EPOTEXT] = (na * ua + nb * ub) * VE;
        // problem callbacks: Re(conj(nu) Delta_ext) (-2) and -(j_a.v_a + j_b.v_b)
        // (pca_kernels.cu:159, 180-185)
        if (dext) e[TDW_EPAIREXT] = (nu.x * dext[i].x + nu.y * dext[i].y) * (-2.0) * VE;
        if (vel)
            e[TDW_EVELEXT] = (tx1 * vel[i] + ty1 * vel[N + i] + tz1 * vel[2 * N + i] + tx2 * vel[3 * N + i] +
                              ty2 * vel[4 * N + i] + tz2 * vel[5 * N + i]) *
                             (-1.0) * VE;
        if (pp) {
            // modify_energies acts on the per-point contributions; their sums
            // are formed after it (k_pp_sum)
#pragma unroll
            for (int k = 0; k < TDW_EVELEXT + 1; k++) {
                pp[(long)k * N + i] = e[k];
                e[k] = 0.0;
            }
        }
        e[TDW_NPARTA] = na;
        e[TDW_NPARTB] = nb;
        if (dim >= 2) {
            int ix, iy, iz;
            decode(dim, V.ny, V.nz, i, ix, iy, iz);
            const double x = (double)(ix - V.nx / 2) * V.dx, y = (double)(iy - V.ny / 2) * V.dy;
            e[TDW_LZA] = (x * dens[5 * N + i] - y * dens[4 * N + i]) * VE;
            e[TDW_LZB] = (x * dens[10 * N + i] - y * dens[9 * N + i]) * VE;
        }
    }
#pragma unroll
    for (int k = 0; k < TDW_NITEMS; k++) {
        double s = block_sum(e[k], red);
        if (threadIdx.x == 0) part[(long)k * gridDim.x + blockIdx.x] = s;
    }
}

// block sums of the per-point energy rows [EKIN .. EVELEXT] into the rows of
// part that k_energy left zero
__global__ void __launch_bounds__(NT) k_pp_sum(long N, const double *__restrict__ pp, double *__restrict__ part)
{
    __shared__ double red[NT / 64];
    const long i = (long)blockIdx.x * NT + threadIdx.x;
#pragma unroll
    for (int k = 0; k < TDW_EVELEXT + 1; k++) {
        double s = block_sum(i < N ? pp[(long)k * N + i] : 0.0, red);
        if (threadIdx.x == 0) part[(long)k * gridDim.x + blockIdx.x] = s;
    }
}

int energy(tdw_ctx *c, int it, const double *dens, const double *pots, double *h_e)
{
    const long N = (long)c->N;
    const int nb = (int)((N + NT - 1) / NT);
    TDW_TRY(ensure_part(c, (size_t)TDW_NITEMS * nb));
    double VE = dvol_energy(c);
    FParams f = fparams(c);
    f.agrid = sclgth_grid(c, it);
    const double2 *dext = dext_field(c, it, pots);
    const double *vel = vel_field(c, it);
    const bool modify = c->proj && c->proj->modify_energies;
    double *pp = modify ? c->d_epp : nullptr;
    hipLaunchKernelGGL(k_energy, dim3(nb), dim3(NT), 0, c->stream, f, make_vext(c, it), N, c->dim, VE, dens, pots,
                       dext, vel, pp, c->d_part);
    if (modify) {
        TDW_HIP(hipGetLastError());
        TDW_TRY(c->proj->modify_energies(c->stream, it, (double *)dens, (double *)pots, pp));
        hipLaunchKernelGGL(k_pp_sum, dim3(nb), dim3(NT), 0, c->stream, N, pp, c->d_part);
    }
    hipLaunchKernelGGL(k_rowsum, dim3(1), dim3(64), 0, c->stream, c->d_part, TDW_NITEMS, nb, 1.0, c->d_esum);
    TDW_HIP(hipGetLastError());
    double e[TDW_NITEMS];
    TDW_HIP(hipMemcpyAsync(e, c->d_esum, sizeof(e), hipMemcpyDeviceToHost, c->stream));
    TDW_HIP(hipStreamSynchronize(c->stream));
    e[TDW_NPARTA] *= VE;
    e[TDW_NPARTB] *= VE;
    for (int k = 0; k < TDW_NITEMS; k++) h_e[k] = e[k];
    return 0;
}

// ---------------------------------------------------------------------------
// Taylor self-start contribution (pca_kernels.cu:1032-1085)
// ---------------------------------------------------------------------------
__global__ void k_taylor(long tot, double tc, const double2 *__restrict__ hpsi, double2 *__restrict__ upd,
                         double2 *__restrict__ contr)
{
    long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= tot) return;
    double2 u = imul(hpsi[i], tc);
    contr[i] = u;
    upd[i] = upd[i] + u;
}

int taylor(tdw_ctx *c, int order, double dt, int n, const double2 *hpsi, double2 *upd, double2 *contr)
{
    long tot = 2L * n * (long)c->N;
    hipLaunchKernelGGL(k_taylor, dim3((tot + NT - 1) / NT), dim3(NT), 0, c->stream, tot, -1.0 * dt / (double)order,
                       hpsi, upd, contr);
    return (int)hipGetLastError();
}

}  // namespace tdw
