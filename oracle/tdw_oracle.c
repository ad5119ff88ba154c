/*
 * tdw_oracle.c -- CPU FP64 restatement of the W-SLDA td step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline).  The product
 * never links it.  Each routine cites the reference file:line it follows;
 * the pinning of each routine is listed in tdw_oracle.h.
 *
 * Complex numbers are stored as interleaved (re, im) doubles, exactly the
 * memory layout of cufftDoubleComplex / C99 double complex in the reference.
 * The state array of n quasiparticle states is [u_0..u_{n-1}, v_0..v_{n-1}],
 * each of tdwo_npts() points (pca.c:307, pca_kernels.cu:337-338).
 */
#include "tdw_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef double complex zc;

/* Functional constants (pca_settings.h:80-110): SLDA and BdG force the
 * effective mass to A0 = 1 (SLDA_FORCE_A1, FAST_CONST_EFFECTIVE_MASS_MODE);
 * ASLDA uses A0 = 1.094, A1 = 0.156, A2 = -0.532 with CURRENT_CORRECTIONS
 * and the density-dependent effective mass. */
static double A0of(const tdwo_cfg *c) { return c->functional == TDWO_ASLDA ? 1.094 : 1.000; }
static double A1of(const tdwo_cfg *c) { return c->functional == TDWO_ASLDA ? 0.156 : 0.0; }
static double A2of(const tdwo_cfg *c) { return c->functional == TDWO_ASLDA ? -0.532 : 0.0; }
#define A0c A0of(c)
#define A1c A1of(c)
#define A2c A2of(c)
#define G0c 0.357
#define G1c 0.642
#define P_ALPHAc 1.0
#define GAMMA0c (-11.11)
#define UD_SCITERS 15
#define UD_MIX_COEFF 0.75
#define DENS_FACTOR_M 10000.

static int g_threads = 0;
void tdwo_set_threads(int nthreads) { g_threads = nthreads; }
int tdwo_get_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}
#ifdef _OPENMP
#define OMP_NT num_threads(tdwo_get_threads())
#endif

size_t tdwo_npts(const tdwo_cfg *c)
{
    if (c->dim == 1) return (size_t)c->nx;
    if (c->dim == 2) return (size_t)c->nx * c->ny;
    return (size_t)c->nx * c->ny * c->nz;
}

/* Volume element of the evolved lattice used by normalisation and <H>
 * (pca_kernels.cu:913 uses DXYZ; ccpca_kernels.cu:871 DX; cpca DX*DY). */
static double dvol_norm(const tdwo_cfg *c)
{
    if (c->dim == 1) return c->dx;
    if (c->dim == 2) return c->dx * c->dy;
    return c->dx * c->dy * c->dz;
}
/* VOLUME_ELEMENT of tdwslda_functionals_framework_enable.h:7-16. */
static double dvol_energy(const tdwo_cfg *c)
{
    double dxyz = c->dx * c->dy * c->dz;
    if (c->dim == 1) return dxyz * c->ny * c->nz;
    if (c->dim == 2) return dxyz * c->nz;
    return dxyz;
}
/* Density normalisation of the transverse plane waves:
 * 1D divides by LY*LZ (ccpca_dens.cu:131-141), 2D by LZ (cpca_dens.cu:132). */
static double dens_transverse(const tdwo_cfg *c)
{
    if (c->dim == 1) return (c->dy * c->ny) * (c->dz * c->nz);
    if (c->dim == 2) return c->dz * c->nz;
    return 1.0;
}

/* =========================================================================
 * Mixed-radix FFT (replaces cuFFT Z2Z; pca_derivative.cu:41-101).
 * Stockham autosort, decimation in time: stage s with radix R and span Ns
 * (product of the earlier radices) maps x[j + t*n/R] -> y[(j/Ns)*Ns*R +
 * j%Ns + t*Ns] after multiplying by exp(sign*2*pi*i*t*(j%Ns)/(Ns*R)).
 * Twiddles are tabulated per stage from cos/sin in double precision.
 * ========================================================================= */
#define MAXSTAGES 16
typedef struct {
    int n, nst;
    int R[MAXSTAGES], Ns[MAXSTAGES];
    zc *tw[2][MAXSTAGES];   /* [sign][stage][k*R + t] */
} fftplan;
#define MAXPLANS 16
static fftplan g_plans[MAXPLANS];
static int g_nplans = 0;

static int pick_radix(int n)
{
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    if (n % 3 == 0) return 3;
    if (n % 5 == 0) return 5;
    for (int p = 7; p * p <= n; p += 2) if (n % p == 0) return p;
    return n;
}

static const fftplan *get_plan(int n)
{
    int i;
    for (i = 0; i < g_nplans; i++) if (g_plans[i].n == n) return &g_plans[i];
    fftplan *p = NULL;
#ifdef _OPENMP
#pragma omp critical(tdwo_plan)
#endif
    {
        for (i = 0; i < g_nplans; i++) if (g_plans[i].n == n) { p = &g_plans[i]; break; }
        if (!p && g_nplans < MAXPLANS) {
            fftplan *q = &g_plans[g_nplans];
            q->n = n;
            q->nst = 0;
            int rem = n, Ns = 1;
            while (rem > 1) {
                int R = pick_radix(rem);
                int s = q->nst++;
                q->R[s] = R;
                q->Ns[s] = Ns;
                for (int sg = 0; sg < 2; sg++) {
                    double sgn = sg == 0 ? -1.0 : 1.0;
                    q->tw[sg][s] = (zc *)malloc(sizeof(zc) * (size_t)Ns * R);
                    for (int k = 0; k < Ns; k++)
                        for (int t = 0; t < R; t++) {
                            double a = 2.0 * M_PI * (double)(t * k) / (double)(Ns * R);
                            q->tw[sg][s][k * R + t] = cos(a) + I * (sgn * sin(a));
                        }
                }
                Ns *= R;
                rem /= R;
            }
            g_nplans++;
            p = q;
        }
    }
    return p;
}

/* small DFT of length R in place; sign = -1 forward, +1 inverse */
static inline void dft_small(zc *v, int R, double sgn)
{
    if (R == 2) {
        zc a = v[0], b = v[1];
        v[0] = a + b; v[1] = a - b;
    } else if (R == 4) {
        zc a = v[0] + v[2], b = v[0] - v[2], c = v[1] + v[3], d = (v[1] - v[3]) * (sgn * I);
        v[0] = a + c; v[1] = b + d; v[2] = a - c; v[3] = b - d;
    } else if (R == 3) {
        const double c1 = -0.5, s1 = 0.86602540378443864676 * sgn;
        zc a = v[0], b = v[1], c = v[2];
        zc t1 = b + c, t2 = (b - c) * (I * s1);
        v[0] = a + t1;
        v[1] = a + c1 * t1 + t2;
        v[2] = a + c1 * t1 - t2;
    } else {
        zc t[64];
        for (int q = 0; q < R; q++) {
            zc acc = 0;
            for (int j = 0; j < R; j++) {
                double a = 2.0 * M_PI * (double)((j * q) % R) / (double)R;
                acc += v[j] * (cos(a) + I * (sgn * sin(a)));
            }
            t[q] = acc;
        }
        for (int q = 0; q < R; q++) v[q] = t[q];
    }
}

/* 1D transform of a pencil of length n with the given stride, in place. */
static void fft_pencil(zc *z, int n, size_t stride, int sign, zc *bx, zc *by)
{
    const fftplan *pl = get_plan(n);
    int sg = sign < 0 ? 0 : 1;
    double sgn = sign < 0 ? -1.0 : 1.0;
    for (int i = 0; i < n; i++) bx[i] = z[(size_t)i * stride];
    zc *x = bx, *y = by;
    for (int s = 0; s < pl->nst; s++) {
        int R = pl->R[s], Ns = pl->Ns[s], nR = n / R;
        const zc *tw = pl->tw[sg][s];
        zc v[64];
        for (int j = 0; j < nR; j++) {
            int k = j % Ns;
            for (int t = 0; t < R; t++) v[t] = x[j + t * nR] * tw[k * R + t];
            dft_small(v, R, sgn);
            int base = (j / Ns) * Ns * R + k;
            for (int t = 0; t < R; t++) y[base + t * Ns] = v[t];
        }
        zc *tmp = x; x = y; y = tmp;
    }
    for (int i = 0; i < n; i++) z[(size_t)i * stride] = x[i];
}

static void fft_one(const tdwo_cfg *c, zc *z, int sign)
{
    zc bi[512], bo[512];
    if (c->dim == 1) { fft_pencil(z, c->nx, 1, sign, bi, bo); return; }
    if (c->dim == 2) {
        int nx = c->nx, ny = c->ny;
        for (int ix = 0; ix < nx; ix++) fft_pencil(z + (size_t)ix * ny, ny, 1, sign, bi, bo);
        for (int iy = 0; iy < ny; iy++) fft_pencil(z + iy, nx, (size_t)ny, sign, bi, bo);
        return;
    }
    int nx = c->nx, ny = c->ny, nz = c->nz;
    for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++)
            fft_pencil(z + ((size_t)ix * ny + iy) * nz, nz, 1, sign, bi, bo);
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++)
            fft_pencil(z + (size_t)ix * ny * nz + iz, ny, (size_t)nz, sign, bi, bo);
    for (int iy = 0; iy < ny; iy++)
        for (int iz = 0; iz < nz; iz++)
            fft_pencil(z + (size_t)iy * nz + iz, nx, (size_t)ny * nz, sign, bi, bo);
}

static void prepare_plans(const tdwo_cfg *c)
{
    get_plan(c->nx);
    if (c->dim >= 2) get_plan(c->ny);
    if (c->dim >= 3) get_plan(c->nz);
}

void tdwo_fft(const tdwo_cfg *c, double *z, int sign)
{
    prepare_plans(c);
    fft_one(c, (zc *)z, sign);
}

/* Lattice momentum of index i on an axis of n points and length L in the
 * standard FFT order (pca_derivative.cu:150-160). */
static double kval(int i, int n, double L)
{
    return (i < n / 2) ? 2. * M_PI / L * (double)i : 2. * M_PI / L * (double)(i - n);
}

static void decode(const tdwo_cfg *c, size_t ixyz, int *ix, int *iy, int *iz)
{
    if (c->dim == 1) { *ix = (int)ixyz; *iy = 0; *iz = 0; return; }
    if (c->dim == 2) { *ix = (int)(ixyz / c->ny); *iy = (int)(ixyz - (size_t)(*ix) * c->ny); *iz = 0; return; }
    size_t i = ixyz;
    *ix = (int)(i / ((size_t)c->ny * c->nz));
    i -= (size_t)(*ix) * c->ny * c->nz;
    *iy = (int)(i / c->nz);
    *iz = (int)(i - (size_t)(*iy) * c->nz);
}

/* k-space factors: -k^2/N and i*k/N (Nyquist zeroed) per point
 * (pca_derivative.cu:150-176, 1D/2D twins use only the evolved axes). */
static void kfactors(const tdwo_cfg *c, size_t ixyz, double *k2n, double *kxn, double *kyn, double *kzn)
{
    int ix, iy, iz;
    decode(c, ixyz, &ix, &iy, &iz);
    double N = (double)tdwo_npts(c);
    double kx = kval(ix, c->nx, c->dx * c->nx);
    double ky = c->dim >= 2 ? kval(iy, c->ny, c->dy * c->ny) : 0.0;
    double kz = c->dim >= 3 ? kval(iz, c->nz, c->dz * c->nz) : 0.0;
    *k2n = -1.0 * (kx * kx + ky * ky + kz * kz) / N;
    double gx = kx / N, gy = ky / N, gz = kz / N;
    if (ix == c->nx / 2) gx = 0.0;
    if (c->dim >= 2 && iy == c->ny / 2) gy = 0.0;
    if (c->dim >= 3 && iz == c->nz / 2) gz = 0.0;
    *kxn = gx; *kyn = gy; *kzn = gz;
}

void tdwo_laplace(const tdwo_cfg *c, int m, const double *in, double *out)
{
    size_t N = tdwo_npts(c);
    prepare_plans(c);
    double *k2 = (double *)malloc(sizeof(double) * N);
    for (size_t i = 0; i < N; i++) { double a, b, d; kfactors(c, i, &k2[i], &a, &b, &d); }
#ifdef _OPENMP
#pragma omp parallel for schedule(static) OMP_NT
#endif
    for (int a = 0; a < m; a++) {
        zc *z = (zc *)out + (size_t)a * N;
        if (out != in) memcpy(z, (const zc *)in + (size_t)a * N, sizeof(zc) * N);
        fft_one(c, z, -1);
        for (size_t i = 0; i < N; i++) z[i] *= k2[i];
        fft_one(c, z, +1);
    }
    free(k2);
}

void tdwo_derivatives(const tdwo_cfg *c, int m, const double *in,
                      double *ddx, double *ddy, double *ddz, double *lap)
{
    size_t N = tdwo_npts(c);
    prepare_plans(c);
    double *kf = (double *)malloc(sizeof(double) * 4 * N);
    for (size_t i = 0; i < N; i++) kfactors(c, i, &kf[4 * i], &kf[4 * i + 1], &kf[4 * i + 2], &kf[4 * i + 3]);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) OMP_NT
#endif
    for (int a = 0; a < m; a++) {
        zc *zl = (zc *)lap + (size_t)a * N;
        zc *zx = (zc *)ddx + (size_t)a * N;
        zc *zy = ddy ? (zc *)ddy + (size_t)a * N : NULL;
        zc *zz = ddz ? (zc *)ddz + (size_t)a * N : NULL;
        if ((const double *)zl != in + 2 * (size_t)a * N)
            memcpy(zl, (const zc *)in + (size_t)a * N, sizeof(zc) * N);
        fft_one(c, zl, -1);
        for (size_t i = 0; i < N; i++) {
            zc v = zl[i];
            zx[i] = I * kf[4 * i + 1] * v;
            if (zy) zy[i] = I * kf[4 * i + 2] * v;
            if (zz) zz[i] = I * kf[4 * i + 3] * v;
            zl[i] = kf[4 * i] * v;
        }
        fft_one(c, zl, +1);
        fft_one(c, zx, +1);
        if (zy && c->dim >= 2) fft_one(c, zy, +1);
        if (zz && c->dim >= 3) fft_one(c, zz, +1);
    }
    free(kf);
}

void tdwo_laplace_real(const tdwo_cfg *c, const double *f, double *lap)
{
    /* D2Z / Z2D pair of pca_derivative.cu:640-668 equals the real part of the
     * complex transform of a real field. */
    size_t N = tdwo_npts(c);
    zc *z = (zc *)malloc(sizeof(zc) * N);
    for (size_t i = 0; i < N; i++) z[i] = f[i];
    tdwo_laplace(c, 1, (const double *)z, (double *)z);
    for (size_t i = 0; i < N; i++) lap[i] = creal(z[i]);
    free(z);
}

/* Real-field derivatives along the evolved axes, Nyquist derivative zero
 * (pca_derivative.cu:385-488, 490-598, 673-782): the D2Z / Z2D pairs equal
 * the real part of the complex transform of the real field. */
static void real_derivs(const tdwo_cfg *c, const double *f, zc *g /* 4N */)
{
    size_t N = tdwo_npts(c);
    zc *z = (zc *)malloc(sizeof(zc) * N);
    for (size_t i = 0; i < N; i++) z[i] = f[i];
    tdwo_derivatives(c, 1, (const double *)z, (double *)g, (double *)(g + N), (double *)(g + 2 * N),
                     (double *)(g + 3 * N));
    free(z);
}

void tdwo_gradient_real(const tdwo_cfg *c, const double *f, double *dx, double *dy, double *dz)
{
    size_t N = tdwo_npts(c);
    zc *g = (zc *)malloc(sizeof(zc) * 4 * N);
    real_derivs(c, f, g);
    double *out[3] = {dx, dy, dz};
    for (int d = 0; d < c->dim; d++)
        for (size_t i = 0; i < N; i++) out[d][i] = creal(g[d * N + i]);
    free(g);
}

void tdwo_dvec_real(const tdwo_cfg *c, const double *fx, const double *fy, const double *fz,
                    double *dfx, double *dfy, double *dfz)
{
    size_t N = tdwo_npts(c);
    zc *g = (zc *)malloc(sizeof(zc) * 4 * N);
    const double *f[3] = {fx, fy, fz};
    double *out[3] = {dfx, dfy, dfz};
    for (int d = 0; d < c->dim; d++) {
        real_derivs(c, f[d], g);
        for (size_t i = 0; i < N; i++) out[d][i] = creal(g[d * N + i]);
    }
    free(g);
}

void tdwo_div_real(const tdwo_cfg *c, const double *fx, const double *fy, const double *fz, double *div)
{
    size_t N = tdwo_npts(c);
    zc *g = (zc *)malloc(sizeof(zc) * 4 * N);
    double *acc = (double *)calloc(N, sizeof(double));
    const double *f[3] = {fx, fy, fz};
    for (int d = 0; d < c->dim; d++) {
        real_derivs(c, f[d], g);
        for (size_t i = 0; i < N; i++) acc[i] += creal(g[d * N + i]);
    }
    memcpy(div, acc, sizeof(double) * N);
    free(acc);
    free(g);
}

/* Fermi-Dirac high-frequency filter of one complex field: F^-1[FD(k^2/2)/N F]
 * with FD(e) = 1/(exp((e - mu)/T) + 1) (pca_derivative.cu:784-963). */
void tdwo_hf_filter_c(const tdwo_cfg *c, const double *in, double *out, double mu, double T)
{
    size_t N = tdwo_npts(c);
    zc *z = (zc *)malloc(sizeof(zc) * N);
    memcpy(z, in, sizeof(zc) * N);
    prepare_plans(c);
    fft_one(c, z, -1);
    for (size_t i = 0; i < N; i++) {
        double k2n, a, b, d;
        kfactors(c, i, &k2n, &a, &b, &d);
        double ek = -0.5 * k2n * (double)N;   /* k^2 / 2 */
        z[i] *= 1.0 / (exp((ek - mu) / T) + 1.0) / (double)N;
    }
    fft_one(c, z, +1);
    memcpy(out, z, sizeof(zc) * N);
    free(z);
}

void tdwo_hf_filter_d(const tdwo_cfg *c, int n, const double *in, double *out, double mu, double T)
{
    size_t N = tdwo_npts(c);
    zc *z = (zc *)malloc(sizeof(zc) * N);
    for (int a = 0; a < n; a++) {
        for (size_t i = 0; i < N; i++) z[i] = in[(size_t)a * N + i];
        tdwo_hf_filter_c(c, (const double *)z, (double *)z, mu, T);
        for (size_t i = 0; i < N; i++) out[(size_t)a * N + i] = creal(z[i]);
    }
    free(z);
}

/* <w1_a|w2_a> dV = dV sum (conj(u1) u2 + conj(v1) v2) (pca_kernels.cu:853-911) */
void tdwo_overlap(const tdwo_cfg *c, int n, const double *w1, const double *w2, double *re, double *im)
{
    size_t N = tdwo_npts(c);
    const zc *A = (const zc *)w1, *B = (const zc *)w2;
    double dv = c->dx * (c->dim >= 2 ? c->dy : 1.0) * (c->dim >= 3 ? c->dz : 1.0);
    for (int a = 0; a < n; a++) {
        zc s = 0;
        for (size_t i = 0; i < N; i++)
            s += conj(A[(size_t)a * N + i]) * B[(size_t)a * N + i] +
                 conj(A[((size_t)n + a) * N + i]) * B[((size_t)n + a) * N + i];
        if (re) re[a] = creal(s) * dv;
        if (im) im[a] = cimag(s) * dv;
    }
}

/* =========================================================================
 * ABM integrator (tdwslda_ode_integrator.h:26-220)
 * ========================================================================= */
void tdwo_abm_step1(const tdwo_cfg *c, int n, double *y, double *f1, double *f2,
                    double *f3, double *f4, double *f5)
{
    size_t tot = 2 * (size_t)n * tdwo_npts(c);
    zc dti = -I * c->dt;
    zc *Y = (zc *)y, *F1 = (zc *)f1, *F2 = (zc *)f2, *F3 = (zc *)f3, *F4 = (zc *)f4, *F5 = (zc *)f5;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) OMP_NT
#endif
    for (size_t i = 0; i < tot; i++) {
        zc yk = Y[i];
        if (c->scheme == TDWO_AB3AM4) {
            zc a = F1[i], b = F2[i], d = F3[i];
            Y[i] = yk + dti * (a * (23. / 12.) - b * (16. / 12.) + d * (5. / 12.));
            F3[i] = yk + dti * (a * (19. / 24.) - b * (5. / 24.) + d * (1. / 24.));
        } else if (c->scheme == TDWO_AB4AM5) {
            zc a = F1[i], b = F2[i], d = F3[i], e = F4[i];
            Y[i] = yk + dti * (a * (55. / 24.) - b * (59. / 24.) + d * (37. / 24.) - e * (9. / 24.));
            F4[i] = yk + dti * (a * (646. / 720.) - b * (264. / 720.) + d * (106. / 720.) - e * (19. / 720.));
        } else {
            zc a = F1[i], b = F2[i], d = F3[i], e = F4[i], g = F5[i];
            Y[i] = yk + dti * (a * (1901. / 720.) - b * (2774. / 720.) + d * (2616. / 720.) - e * (1274. / 720.) + g * (251. / 720.));
            F5[i] = yk + dti * (a * (646. / 720.) - b * (264. / 720.) + d * (106. / 720.) - e * (19. / 720.));
        }
    }
}

void tdwo_abm_step4(const tdwo_cfg *c, int n, const double *hy, double *y, const double *flast)
{
    size_t tot = 2 * (size_t)n * tdwo_npts(c);
    zc dti = -I * c->dt;
    double a0 = (c->scheme == TDWO_AB3AM4) ? 9. / 24. : 251. / 720.;
    const zc *H = (const zc *)hy, *FL = (const zc *)flast;
    zc *Y = (zc *)y;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) OMP_NT
#endif
    for (size_t i = 0; i < tot; i++) Y[i] = dti * H[i] * a0 + FL[i];
}

/* =========================================================================
 * Normalisation (pca_kernels.cu:913-976)
 * ========================================================================= */
void tdwo_norms(const tdwo_cfg *c, int n, const double *wf, double *norm)
{
    size_t N = tdwo_npts(c);
    double dv = dvol_norm(c);
    const zc *W = (const zc *)wf;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) OMP_NT
#endif
    for (int a = 0; a < n; a++) {
        const zc *u = W + (size_t)a * N, *v = W + ((size_t)n + a) * N;
        double s = 0.0;
        for (size_t i = 0; i < N; i++)
            s += (creal(u[i]) * creal(u[i]) + cimag(u[i]) * cimag(u[i]) +
                  creal(v[i]) * creal(v[i]) + cimag(v[i]) * cimag(v[i])) * dv;
        norm[a] = s;
    }
}

void tdwo_normalize(const tdwo_cfg *c, int n, double *wf)
{
    size_t N = tdwo_npts(c);
    double *nrm = (double *)malloc(sizeof(double) * (size_t)n);
    tdwo_norms(c, n, wf, nrm);
    zc *W = (zc *)wf;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) OMP_NT
#endif
    for (int a = 0; a < n; a++) {
        double s = 1.0 / sqrt(nrm[a]);
        zc *u = W + (size_t)a * N, *v = W + ((size_t)n + a) * N;
        for (size_t i = 0; i < N; i++) { u[i] *= s; v[i] *= s; }
    }
    free(nrm);
}

/* =========================================================================
 * Densities (pca_dens.cu:13-198, ccpca_dens.cu:14-146, cpca_dens.cu:14-145)
 * ========================================================================= */
static inline double nrm2(zc z) { return creal(z) * creal(z) + cimag(z) * cimag(z); }

void tdwo_densities(const tdwo_cfg *c, int n, const double *wf,
                    const double *ddx, const double *ddy, const double *ddz,
                    const double *lap, const double *fbeta, const double *kt,
                    const int *cnt, double *dens, int gradients)
{
    size_t N = tdwo_npts(c);
    const zc *W = (const zc *)wf;
    const zc *DX = (const zc *)ddx, *DY = (const zc *)ddy, *DZ = (const zc *)ddz, *L = (const zc *)lap;
    zc *nu = (zc *)dens;
    double *rho_a = dens + 2 * N, *tau_a = dens + 3 * N, *jax_ = dens + 4 * N, *jay_ = dens + 5 * N, *jaz_ = dens + 6 * N;
    double *rho_b = dens + 7 * N, *tau_b = dens + 8 * N, *jbx_ = dens + 9 * N, *jby_ = dens + 10 * N, *jbz_ = dens + 11 * N;
    double tr = dens_transverse(c);
    double piy = M_PI / c->dy, piz = M_PI / c->dz;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) OMP_NT
#endif
    for (size_t i = 0; i < N; i++) {
        double na = 0, nb = 0, ta = 0, tb = 0, jax = 0, jay = 0, jaz = 0, jbx = 0, jby = 0, jbz = 0;
        zc nuu = 0;
        for (int a = 0; a < n; a++) {
            double fb = fbeta[a] * DENS_FACTOR_M;
            double fbm = DENS_FACTOR_M - fb;
            double w = 1.0, k2t = 0.0;
            if (c->dim == 1) {
                double ky = kt[a], kz = kt[n + a];
                k2t = kz * kz + ky * ky;
                w = (double)cnt[a];
            } else if (c->dim == 2) {
                double kz = kt[a];
                k2t = kz * kz;
                w = 2.0;
                if (fabs(kz) < 1.0e-12) w = 1.0;
                if (c->regularization == TDWO_CUBIC && fabs(kz + piz) < 1.0e-12) w = 1.0;
            }
            (void)piy;
            zc u = W[(size_t)a * N + i], v = W[((size_t)n + a) * N + i];
            if (c->spinsymmetry) {
                /* SPINSYMMETRY_MODE (pca_dens.cu:55-58, 71-90): both
                 * components go to spin b, nu without the 1/2 */
                nb += (nrm2(v) * fbm + nrm2(u) * fb) * w;
                nuu += u * conj(v) * (fbm - fb) * w;
                if (gradients) {
                    zc ux = DX[(size_t)a * N + i], vx = DX[((size_t)n + a) * N + i];
                    zc lu = L[(size_t)a * N + i], lv = L[((size_t)n + a) * N + i];
                    tb += creal(conj(u) * (lu - u * k2t)) * fb * w + creal(conj(v) * (lv - v * k2t)) * fbm * w;
                    jbx += cimag(conj(u) * ux) * fb * w;
                    jbx -= cimag(conj(v) * vx) * fbm * w;
                    if (c->dim >= 2) {
                        zc uy = DY[(size_t)a * N + i], vy = DY[((size_t)n + a) * N + i];
                        jby += cimag(conj(u) * uy) * fb * w;
                        jby -= cimag(conj(v) * vy) * fbm * w;
                    }
                    if (c->dim == 3) {
                        zc uz = DZ[(size_t)a * N + i], vz = DZ[((size_t)n + a) * N + i];
                        jbz += cimag(conj(u) * uz) * fb * w;
                        jbz -= cimag(conj(v) * vz) * fbm * w;
                    }
                }
                continue;
            }
            if (c->dim == 3) {
                na += nrm2(u) * fb;
                nb += nrm2(v) * fbm;
                nuu += u * conj(v) * (fbm - fb) / 2.0;
            } else {
                na += nrm2(u) * fb * w;
                nb += nrm2(v) * fbm * w;
                nuu += u * conj(v) * (fbm - fb) / 2.0 * w;
            }
            if (gradients) {
                zc ux = DX[(size_t)a * N + i], vx = DX[((size_t)n + a) * N + i];
                zc lu = L[(size_t)a * N + i], lv = L[((size_t)n + a) * N + i];
                if (c->dim == 3) {
                    zc uy = DY[(size_t)a * N + i], uz = DZ[(size_t)a * N + i];
                    zc vy = DY[((size_t)n + a) * N + i], vz = DZ[((size_t)n + a) * N + i];
                    ta += creal(conj(u) * lu) * fb;
                    jax += cimag(conj(u) * ux) * fb;
                    jay += cimag(conj(u) * uy) * fb;
                    jaz += cimag(conj(u) * uz) * fb;
                    tb += creal(conj(v) * lv) * fbm;
                    jbx -= cimag(conj(v) * vx) * fbm;
                    jby -= cimag(conj(v) * vy) * fbm;
                    jbz -= cimag(conj(v) * vz) * fbm;
                } else if (c->dim == 2) {
                    zc uy = DY[(size_t)a * N + i], vy = DY[((size_t)n + a) * N + i];
                    ta += creal(conj(u) * (lu - u * k2t)) * fb * w;
                    jax += cimag(conj(u) * ux) * fb * w;
                    jay += cimag(conj(u) * uy) * fb * w;
                    tb += creal(conj(v) * (lv - v * k2t)) * fbm * w;
                    jbx -= cimag(conj(v) * vx) * fbm * w;
                    jby -= cimag(conj(v) * vy) * fbm * w;
                } else {
                    ta += creal(conj(u) * (lu - u * k2t)) * fb * w;
                    jax += cimag(conj(u) * ux) * fb * w;
                    tb += creal(conj(v) * (lv - v * k2t)) * fbm * w;
                    jbx -= cimag(conj(v) * vx) * fbm * w;
                }
            }
        }
        if (c->spinsymmetry) {
            na = nb;
            ta = tb;
            jax = jbx;
            jay = jby;
            jaz = jbz;
        }
        double sc = DENS_FACTOR_M;
        rho_a[i] = na / sc / tr;
        rho_b[i] = nb / sc / tr;
        tau_a[i] = ta / sc / tr;
        tau_b[i] = tb / sc / tr;
        nu[i] = nuu / sc / tr;
        jax_[i] = jax / sc / tr; jay_[i] = jay / sc / tr; jaz_[i] = jaz / sc / tr;
        jbx_[i] = jbx / sc / tr; jby_[i] = jby / sc / tr; jbz_[i] = jbz / sc / tr;
    }
}

void tdwo_density_tau(const tdwo_cfg *c, double *dens)
{
    /* pca_dens.cu:275-338: tau = max(0, lap(rho)/2 - tau_acc) */
    size_t N = tdwo_npts(c);
    double *lr = (double *)malloc(sizeof(double) * N);
    double *rho[2] = {dens + 7 * N, dens + 2 * N};
    double *tau[2] = {dens + 8 * N, dens + 3 * N};
    for (int s = 0; s < 2; s++) {
        if (s == 1 && c->spinsymmetry) {
            /* SPINSYMMETRY_MODE: tau_a is a copy of tau_b (pca_dens.cu:329-331) */
            memcpy(tau[1], tau[0], sizeof(double) * N);
            break;
        }
        tdwo_laplace_real(c, rho[s], lr);
        for (size_t i = 0; i < N; i++) {
            double t = 0.5 * lr[i] - tau[s][i];
            if (t < 0.0) t = 0.0;
            tau[s][i] = t;
        }
    }
    free(lr);
}

/* =========================================================================
 * Quantum friction (pca_dens.cu:360-480, cpca_dens.cu:385-490,
 * ccpca_dens.cu:384-500; pca_kernels.cu:497-539)
 * ========================================================================= */
void tdwo_qf_densities(const tdwo_cfg *c, int n, const double *wf, const double *lap, const double *fbeta,
                       const double *kt, const int *cnt, double *out)
{
    size_t N = tdwo_npts(c);
    const zc *W = (const zc *)wf, *L = (const zc *)lap;
    double tr = dens_transverse(c);
    zc *D = (zc *)(out + 2 * N);
    for (size_t i = 0; i < N; i++) {
        double Ua = 0.0, Ub = 0.0;
        zc Dl = 0.0;
        for (int a = 0; a < n; a++) {
            double fb = fbeta[a], fbm = 1.0 - fb, w = 1.0, k2t = 0.0;
            if (c->dim == 1) {
                k2t = kt[n + a] * kt[n + a] + kt[a] * kt[a];
                w = (double)cnt[a];
            } else if (c->dim == 2) {
                k2t = kt[a] * kt[a];
                w = fabs(kt[a]) < 1.0e-12 ? 1.0 : 2.0;
            }
            zc u = W[(size_t)a * N + i], v = W[((size_t)n + a) * N + i];
            zc lu = L[(size_t)a * N + i] - u * k2t, lv = L[((size_t)n + a) * N + i] - v * k2t;
            if (c->spinsymmetry) {
                Ub += (cimag(conj(u) * lu) * fb - cimag(conj(v) * lv) * fbm) * w;
                Dl += (conj(v) * lu + conj(lv) * u) * (fbm - fb) * w * 2.0;
            } else {
                Ua += cimag(conj(u) * lu) * fb * w;
                Ub -= cimag(conj(v) * lv) * fbm * w;
                Dl += (conj(v) * lu + conj(lv) * u) * (fbm - fb) * w;
            }
        }
        if (c->spinsymmetry) Ua = Ub;
        out[i] = Ua / tr;
        out[N + i] = Ub / tr;
        D[i] = -0.5 * Dl / tr;
    }
}

void tdwo_qf_add(const tdwo_cfg *c, double *qf, double *pots, double qfswitch, double pccoeff)
{
    size_t N = tdwo_npts(c);
    double qa = c->qfalpha * qfswitch, qb = c->qfbeta * qfswitch, qg = c->qfgamma * qfswitch * pccoeff;
    double nF = c->kF * c->kF * c->kF / (3. * M_PI * M_PI);
    zc *dl = (zc *)(pots + 2 * N), *dq = (zc *)(qf + 2 * N);
    for (size_t i = 0; i < N; i++) {
        double phd = carg(dl[i]), phn = carg(dq[i]);
        double va = pots[i], vb = pots[N + i];
        zc del = dl[i];
        pots[i] -= qf[i] * qa / nF;
        pots[N + i] -= qf[N + i] * qa / nF;
        dl[i] += qb * cabs(del) * (cos(phd) + I * sin(phd)) * sin(phd - phn);
        dl[i] += I * qg * del;
        qf[i] = va;
        qf[N + i] = vb;
        dq[i] = del;
    }
}

void tdwo_qf_remove(const tdwo_cfg *c, const double *qf, double *pots)
{
    size_t N = tdwo_npts(c);
    memcpy(pots, qf, sizeof(double) * 4 * N);
}

/* =========================================================================
 * Energy density functional helpers (pca_edf.h:17-352), SLDA with A0=1.
 * ========================================================================= */
static double p_reg(const tdwo_cfg *c, double n)
{
    if (n <= c->p_nmin) return 0.0;
    if (n >= c->p_nmax) return 1.0;
    return 0.5 + 0.5 * tanh(P_ALPHAc * tan((M_PI * (n - c->p_nmin)) / (c->p_nmax - c->p_nmin) - M_PI / 2.0));
}
static double der_p_reg(const tdwo_cfg *c, double n)
{
    if (n <= c->p_nmin) return 0.0;
    if (n >= c->p_nmax) return 0.0;
    double a = (M_PI * (n - c->p_nmin)) / (c->p_nmax - c->p_nmin) - M_PI / 2.0;
    return 0.5 * P_ALPHAc * M_PI / pow(cosh(P_ALPHAc * tan(a)) * cos(a), 2) / (c->p_nmax - c->p_nmin);
}
static double polarization(const tdwo_cfg *c, double na, double nb)
{
    double pr = p_reg(c, na + nb);
    if (pr == 0.0) return 0.0;
    return pr * (na - nb) / (na + nb);
}
static double der_pol_na(const tdwo_cfg *c, double na, double nb)
{
    double np = na + nb, pr = p_reg(c, np), dpr = der_p_reg(c, np), r = 0.0;
    if (pr != 0.0) r += pr * (2.0 * nb) / (np * np);
    if (dpr != 0.0) r += dpr * (na - nb) / np;
    return r;
}
static double der_pol_nb(const tdwo_cfg *c, double na, double nb)
{
    double np = na + nb, pr = p_reg(c, np), dpr = der_p_reg(c, np), r = 0.0;
    if (pr != 0.0) r += pr * (-2.0 * na) / (np * np);
    if (dpr != 0.0) r += dpr * (na - nb) / np;
    return r;
}
static double alpha_f(const tdwo_cfg *c, double p)
{
    double p2 = p * p, p4 = p2 * p2;
    return A0c + A1c * p * (1.0 - (2.0 / 3.0) * p2 + (1.0 / 5.0) * p4) + A2c * p2 * (1.0 - p2 + (1.0 / 3.0) * p4);
}
static double alpha_plus(const tdwo_cfg *c, double p) { return 0.5 * (alpha_f(c, p) + alpha_f(c, -p)); }
static double alpha_minus(const tdwo_cfg *c, double p) { return 0.5 * (alpha_f(c, p) - alpha_f(c, -p)); }
static double der_alpha_plus_na(const tdwo_cfg *c, double na, double nb)
{
    double p = polarization(c, na, nb), p2 = p * p, p4 = p2 * p2;
    return (2.0 * A2c * p * (1.0 - p2 + (1.0 / 3.0) * p4) - 2.0 * A2c * p * p2 * (1.0 - (2.0 / 3.0) * p2)) * der_pol_na(c, na, nb);
}
static double der_alpha_plus_nb(const tdwo_cfg *c, double na, double nb)
{
    double p = polarization(c, na, nb), p2 = p * p, p4 = p2 * p2;
    return (2.0 * A2c * p * (1.0 - p2 + (1.0 / 3.0) * p4) - 2.0 * A2c * p * p2 * (1.0 - (2.0 / 3.0) * p2)) * der_pol_nb(c, na, nb);
}
static double der_alpha_minus_na(const tdwo_cfg *c, double na, double nb)
{
    double p = polarization(c, na, nb), p2 = p * p, p4 = p2 * p2;
    return (A1c * (1.0 - (2.0 / 3.0) * p2 + (1.0 / 5.0) * p4) - 4.0 * A1c * p2 * ((1.0 / 3.0) - (1.0 / 5.0) * p2)) *
           der_pol_na(c, na, nb);
}
static double der_alpha_minus_nb(const tdwo_cfg *c, double na, double nb)
{
    double p = polarization(c, na, nb), p2 = p * p, p4 = p2 * p2;
    return (A1c * (1.0 - (2.0 / 3.0) * p2 + (1.0 / 5.0) * p4) - 4.0 * A1c * p2 * ((1.0 / 3.0) - (1.0 / 5.0) * p2)) *
           der_pol_nb(c, na, nb);
}
static double funG(double p) { return G0c + G1c * p * p; }
static double funD(const tdwo_cfg *c, double na, double nb)
{
    double p = polarization(c, na, nb);
    return (1.0 / (20.0 * M_PI * M_PI)) * pow(6.0 * M_PI * M_PI * (na + nb), 5.0 / 3.0) *
           (funG(p) - alpha_f(c, p) * pow((1.0 + p) / 2.0, 5.0 / 3.0) - alpha_f(c, -1.0 * p) * pow((1.0 - p) / 2.0, 5.0 / 3.0));
}
static double der_funD_common(const tdwo_cfg *c, double na, double nb, double dpol)
{
    double p = polarization(c, na, nb), p2 = p * p, p4 = p2 * p2;
    double base = funG(p) - alpha_f(c, p) * pow((1.0 + p) / 2.0, 5.0 / 3.0) - alpha_f(c, -1.0 * p) * pow((1.0 - p) / 2.0, 5.0 / 3.0);
    double da_p = A1c * (1.0 - (2.0 / 3.0) * p2 + (1.0 / 5.0) * p4) - 4.0 * A1c * p2 * ((1.0 / 3.0) - (1.0 / 5.0) * p2) +
                  2.0 * A2c * p * (1.0 - p2 + (1.0 / 3.0) * p4) - 2.0 * A2c * p * p2 * (1.0 - (2.0 / 3.0) * p2);
    double da_m = -1.0 * A1c * (1.0 - (2.0 / 3.0) * p2 + (1.0 / 5.0) * p4) + 4.0 * A1c * p2 * ((1.0 / 3.0) - (1.0 / 5.0) * p2) +
                  2.0 * A2c * p * (1.0 - p2 + (1.0 / 3.0) * p4) - 2.0 * A2c * p * p2 * (1.0 - (2.0 / 3.0) * p2);
    return (1.0 / 2.0) * pow(6.0 * M_PI * M_PI * (na + nb), 2.0 / 3.0) * base +
           (1.0 / (20.0 * M_PI * M_PI)) * pow(6.0 * M_PI * M_PI * (na + nb), 5.0 / 3.0) *
               (2.0 * G1c * p - da_p * pow((1.0 + p) / 2.0, 5.0 / 3.0) - (5.0 / 6.0) * alpha_f(c, p) * pow((1.0 + p) / 2.0, 2.0 / 3.0) -
                da_m * pow((1.0 - p) / 2.0, 5.0 / 3.0) + (5.0 / 6.0) * alpha_f(c, -1.0 * p) * pow((1.0 - p) / 2.0, 2.0 / 3.0)) *
               dpol;
}
static double der_funD_na(const tdwo_cfg *c, double na, double nb) { return der_funD_common(c, na, nb, der_pol_na(c, na, nb)); }
static double der_funD_nb(const tdwo_cfg *c, double na, double nb) { return der_funD_common(c, na, nb, der_pol_nb(c, na, nb)); }
static double tildeC(const tdwo_cfg *c, double na, double nb)
{
    double p = polarization(c, na, nb);
    return alpha_plus(c, p) * pow(na + nb + c->denseps, 1.0 / 3.0) / (GAMMA0c * 1.0);
}
static double der_tildeC_na(const tdwo_cfg *c, double na, double nb)
{
    double p = polarization(c, na, nb);
    return (der_alpha_plus_na(c, na, nb) * pow(na + nb + c->denseps, 1.0 / 3.0) +
            (1.0 / 3.0) * alpha_plus(c, p) * pow(na + nb + c->denseps, -2.0 / 3.0)) / (GAMMA0c * 1.0);
}
static double der_tildeC_nb(const tdwo_cfg *c, double na, double nb)
{
    double p = polarization(c, na, nb);
    return (der_alpha_plus_nb(c, na, nb) * pow(na + nb + c->denseps, 1.0 / 3.0) +
            (1.0 / 3.0) * alpha_plus(c, p) * pow(na + nb + c->denseps, -2.0 / 3.0)) / (GAMMA0c * 1.0);
}

/* Regularisation constants (pca_settings.h:226-239). */
static void reg_coeffs(const tdwo_cfg *c, double *R0, double *R1)
{
    if (c->regularization == TDWO_CUBIC) { *R0 = 2.442749607806335 / (4.0 * M_PI * M_PI); *R1 = 0.31786000; }
    else { *R0 = 1. / (2.0 * M_PI * M_PI); *R1 = 0.5; }
}

/* =========================================================================
 * External potential family (td-gauss problem-definition.h v_ext and the
 * switch functions of that file).
 * ========================================================================= */
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

void tdwo_vext_eval(const tdwo_cfg *c, int it, double *va, double *vb)
{
    size_t N = tdwo_npts(c);
    const tdwo_vext *e = &c->vext;
    double t = c->t0 + c->dt * it;
    double s = e->switch_on ? smooth_step(t, e->t_start, e->t_stop, e->t_switch, 1.0) : 1.0;
    for (size_t i = 0; i < N; i++) {
        int ix, iy, iz;
        decode(c, i, &ix, &iy, &iz);
        double x = c->dx * (ix - c->nx / 2);
        double y = c->dim >= 2 ? c->dy * (iy - c->ny / 2) : 0.0;
        double z = c->dim >= 3 ? c->dz * (iz - c->nz / 2) : 0.0;
        double g = 0.0;
        for (int k = 0; k < e->ngauss; k++) {
            double X = x - e->cx[k], Y = y - e->cy[k], Z = z - e->cz[k];
            g += e->amp[k] * s * exp(e->wx[k] * pow(X, 2) + e->wy[k] * pow(Y, 2) + e->wz[k] * pow(Z, 2));
        }
        double trap = 0.5 * e->omega * e->omega * (x * x + y * y + z * z);
        va[i] = g + trap;
        vb[i] = (e->flip_b ? -g : g) + trap;
    }
}

/* =========================================================================
 * Potentials (tdwslda_functionals.h:47-306 SLDA, :373-452 BdG)
 * ========================================================================= */
/* SLDAE functional parameters (sldae_functional.c:14-330, default set of
 * sldae_parameters.h:1-75): Pade[4/4] fits in x = |a kF|. */
static double se_pade(double x, const double *q /* a1..a4, b1..b4 */)
{
    double x2 = x * x, x3 = x2 * x, x4 = x3 * x;
    return (q[0] * x + q[1] * x2 + q[2] * x3 + q[3] * x4) / (1. + q[4] * x + q[5] * x2 + q[6] * x3 + q[7] * x4);
}
static double se_pade_d(double x, const double *q)
{
    double x2 = x * x, x3 = x2 * x, x4 = x3 * x;
    double N_ = q[0] * x + q[1] * x2 + q[2] * x3 + q[3] * x4;
    double D_ = 1. + q[4] * x + q[5] * x2 + q[6] * x3 + q[7] * x4;
    double dN = q[0] + 2. * q[1] * x + 3. * q[2] * x2 + 4. * q[3] * x3;
    double dD = q[4] + 2. * q[5] * x + 3. * q[6] * x2 + 4. * q[7] * x3;
    return -(dD * N_) / (D_ * D_) + dN / D_;
}
static const double SE_ALPHA[8] = {0.0, -0.2081560, -0.1358830, -0.0162071, 2.09838, 1.738, 0.668549, 0.101508};
static const double SE_BETA[8] = {1.58442, 17.2973, 0.820671, 3.48966, 13.9448, 10.1914, 11.9429, 3.48966};
static const double SE_IGAM[8] = {6.15853, 0.156722, 0.76668, 0.0823379, 1.48431, 2.07468, 0.737836, 0.0823379};
static const double SE_B[8] = {1.29871, 6.02212, 0.376541, 0.810972, 6.47079, 3.73389, 3.47301, 0.810972};
static const double SE_C[8] = {0.267952, 0.754923, 0.330699, 0.0900897, 2.61695, 2.71855, 0.550587, 0.0900897};
#define SE_BF (-0.281863)
#define SE_CF (-14.95950)
/* SLDA_FORCE_A1 set (sldae_parameters_A1.h:1-36): A = 1 */
static const double SE1_BETA[8] = {0.797400, 0.943961, 1.037043, 0.523445, 1.227062, 2.126210, 1.734795, 0.523445};
static const double SE1_IGAM[8] = {2.713374, 1.441196, 3.700251, 1.877961, 0.634303, 3.522427, 3.701347, 1.877961};
static const double SE1_B[8] = {0.698112, 0.034678, 0.455138, -0.031327, 0.353905, 0.806444, 0.403679, -0.031327};
static const double SE1_C[8] = {0.370145, 1.041299, 1.737347, 1.085400, 2.458364, 5.604947, 3.828124, 1.085400};
#define SE1_BF (-0.481506)
#define SE1_CF (-10.532397)
static int g_se_a1 = 0;   /* parameter set of the current call (set by the entry points) */
static double se_alpha(double x) { return g_se_a1 ? 1. : 1. + se_pade(x, SE_ALPHA); }
static double se_alpha_d(double x) { return g_se_a1 ? 0. : se_pade_d(x, SE_ALPHA); }
static double se_beta(double x) { return g_se_a1 ? SE1_BF * se_pade(x, SE1_BETA) : SE_BF * se_pade(x, SE_BETA); }
static double se_igamma(double x) { return g_se_a1 ? (1. / SE1_CF) * se_pade(x, SE1_IGAM) : (1. / SE_CF) * se_pade(x, SE_IGAM); }
static double se_bfun(double x) { return g_se_a1 ? SE1_BF * se_pade(x, SE1_B) : SE_BF * se_pade(x, SE_B); }
static double se_cfun(double x) { return g_se_a1 ? SE1_CF * se_pade(x, SE1_C) : SE_CF * se_pade(x, SE_C); }
static double se_mu_ufg(double x)
{
    double U = 5. / 24., V = U / tan(3. * M_PI / 16. * (1. - 0.3582341));
    double s0 = atan(x * U / (1. + x * V));
    double s1 = U / (1. + 2. * V * x + (U * U + V * V) * x * x);
    return 1. - 16. / (3. * M_PI) * s0 - 16. / (15. * M_PI) * x * s1;
}
/* af, af', beta, eF, C~, C~', B at total density nt (tdwslda_functionals.h:580-618) */
typedef struct { double af, af_p, beta, eF, ct, ct_p, bf; } se_local;
static se_local se_coeffs(double nt, double ntreg, double as)
{
    se_local L;
    double kF = pow(3. * M_PI * M_PI * nt, 1. / 3.);
    L.eF = kF * kF / 2.;
    double x = fabs(as * kF);
    double dx = ntreg > 0.0 ? ntreg * x / (3. * nt) : 0.0;
    L.beta = se_beta(x);
    double ig = se_igamma(x);
    L.af = se_alpha(x);
    L.bf = se_bfun(x);
    double cf = se_cfun(x), cf_p = 0.0;
    L.af_p = 0.0;
    if (ntreg > 0.0) {
        L.af_p = dx * se_alpha_d(x);
        cf_p = ntreg * cf / (3. * nt) * (1. - cf * ig);
    }
    L.ct = L.af * pow(nt, 1. / 3.) / cf;
    double igeff = (1 / cf) * (1. - 3. * nt / cf * cf_p);
    L.ct_p = ntreg > 0. ? pow(ntreg, 2. / 3.) * igeff * L.af / (3. * pow(nt, 2. / 3.)) : 0.;
    L.ct_p += L.af_p * pow(nt, 1. / 3.) / cf;
    return L;
}

void tdwo_potentials(const tdwo_cfg *c, int it, const double *dens, double *pots)
{
    g_se_a1 = c->force_a1;
    size_t N = tdwo_npts(c);
    const zc *nu = (const zc *)dens;
    const double *rho_a = dens + 2 * N, *rho_b = dens + 7 * N;
    double *Va_ = pots, *Vb_ = pots + N, *al_a = pots + 4 * N, *al_b = pots + 5 * N;
    zc *delta = (zc *)(pots + 2 * N);
    double *ua = (double *)malloc(sizeof(double) * N), *ub = (double *)malloc(sizeof(double) * N);
    tdwo_vext_eval(c, it, ua, ub);
    double R0, R1;
    reg_coeffs(c, &R0, &R1);
    double DXs = c->dx;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) OMP_NT
#endif
    for (size_t i = 0; i < N; i++) {
        double na = rho_a[i], nb = rho_b[i];
        zc lnu = nu[i];
        if (c->functional == TDWO_SLDAE) {
            /* tdwslda_functionals.h:530-812, CURRENT_CORRECTIONS, cccoeff = 1 */
            double nt = na + nb;
            se_local L = se_coeffs(nt, p_reg(c, nt), c->sclgth);
            const double *ja = dens + 4 * N, *jb = dens + 9 * N;
            double *Aa = pots + 6 * N, *Ab = pots + 9 * N;
            /* SLDA_FORCE_A1: FAST_CONST_EFFECTIVE_MASS_MODE, no CURRENT_CORRECTIONS */
            const int cc_on = !c->force_a1;
            double fa = cc_on ? p_reg(c, na) : 0.0, fb = cc_on ? p_reg(c, nb) : 0.0, j2a = 0.0, j2b = 0.0;
            if (cc_on) {
                al_a[i] = L.af;
                al_b[i] = L.af;
            }
            for (int d = 0; d < c->dim && cc_on; d++) {
                Aa[d * N + i] = fa == 0.0 ? 0.0 : fa * (1. - L.af) * ja[d * N + i] / na;
                Ab[d * N + i] = fb == 0.0 ? 0.0 : fb * (1. - L.af) * jb[d * N + i] / nb;
                j2a += ja[d * N + i] * ja[d * N + i];
                j2b += jb[d * N + i] * jb[d * N + i];
            }
            double taus = dens[3 * N + i] + dens[8 * N + i];
            double Va_const = ua[i] + L.af_p * taus / 2. + L.beta * L.eF;
            double Vb_const = ub[i] + L.af_p * taus / 2. + L.beta * L.eF;
            if (fa != 0.0) {
                double t7 = fa * j2a / (2.0 * na);
                Va_const += ((L.af - 1.0) / na - L.af_p) * t7;
                Va_const -= der_p_reg(c, na) * (L.af - 1.0) * j2a / (2.0 * na);
                Vb_const -= L.af_p * t7;
            }
            if (fb != 0.0) {
                double t7 = fb * j2b / (2.0 * nb);
                Va_const -= L.af_p * t7;
                Vb_const += ((L.af - 1.0) / nb - L.af_p) * t7;
                Vb_const -= der_p_reg(c, nb) * (L.af - 1.0) * j2b / (2.0 * nb);
            }
            double Va = Va_[i], Vb = Vb_[i];
            zc ldelta = 0;
            double lkF = pow(3.0 * M_PI * M_PI * nt, 1.0 / 3.0);
            for (int k = 0; k < UD_SCITERS; k++) {
                double lmu = (c->mu_a - Va + c->mu_b - Vb) / 2.;
                zc p0 = csqrt(CMPLX(2.0 * lmu / L.af, 0.0));
                if (cimag(p0) < 0.) p0 *= -1.;
                zc kc;
                if (c->regularization == TDWO_CUBIC) kc = M_PI / DXs;
                else {
                    kc = csqrt(CMPLX(2.0 * (c->ec + lmu) / L.af, 0.0));
                    if (cimag(kc) < 0.) kc *= -1.;
                }
                zc wz = clog((kc + p0) / (kc - p0));
                if (cimag(wz) < 0.) wz += CMPLX(0.0, 2. * M_PI);
                wz = kc * R0 * (1. - (p0 / kc) * R1 * wz);
                double g_eff = L.af / (L.ct - creal(wz));
                ldelta = -lnu * g_eff;
                double dn = creal(conj(ldelta) * lnu), d2 = nrm2(ldelta);
                double Vanew = Va_const - (L.af_p / L.af) * dn - L.ct_p / L.af * d2;
                double Vbnew = Vb_const - (L.af_p / L.af) * dn - L.ct_p / L.af * d2;
                if (c->mf_correction) {
                    double bcoeff = creal(p0 / (lkF + 1.0e-12));
                    double lx = bcoeff * lkF * DXs / M_PI;
                    if (g_eff < -1.0e-10 && lx > 1.0e-10) {
                        double Lam0 = bcoeff * R0 / lx * (1.0 - R1 * lx * log((1.0 + lx) / (1.0 - lx)));
                        double dLam0 = (-2. * bcoeff * R0 * R1) / (1. - lx * lx) - bcoeff * R0 / lx / lx;
                        double dLam_dn = (lkF / (3. * nt * L.af)) * (Lam0 + dLam0 * bcoeff * DXs * lkF / M_PI);
                        Vanew += dLam_dn * d2;
                        Vbnew += dLam_dn * d2;
                    }
                }
                Va = UD_MIX_COEFF * Vanew + (1. - UD_MIX_COEFF) * Va;
                Vb = UD_MIX_COEFF * Vbnew + (1. - UD_MIX_COEFF) * Vb;
            }
            Va_[i] = Va;
            Vb_[i] = Vb;
            delta[i] = ldelta;
            continue;
        }
        if (c->functional == TDWO_SLDA || c->functional == TDWO_ASLDA) {
            double t1p = polarization(c, na, nb);
            double alph_plus = alpha_plus(c, t1p);
            double Va_const = ua[i], Vb_const = ub[i];
            /* SLDA: FAST_CONST_EFFECTIVE_MASS_MODE, d(alpha)/dn = 0 */
            double dalphp_dna = 0.0, dalphp_dnb = 0.0;
            if (c->functional == TDWO_ASLDA) {
                /* CURRENT_CORRECTIONS and the density-dependent effective
                 * mass (tdwslda_functionals.h:57-227, cccoeff = 1) */
                double alph_minus = alpha_minus(c, t1p);
                double aA = alpha_f(c, t1p), aB = alpha_f(c, -t1p);
                al_a[i] = aA;
                al_b[i] = aB;
                const double *ja = dens + 4 * N, *jb = dens + 9 * N;
                double *Aa = pots + 6 * N, *Ab = pots + 9 * N;
                double fa = p_reg(c, na), fb = p_reg(c, nb);
                for (int d = 0; d < c->dim; d++) {
                    Aa[d * N + i] = fa == 0.0 ? 0.0 : fa * (1. - aA) * ja[d * N + i] / na;
                    Ab[d * N + i] = fb == 0.0 ? 0.0 : fb * (1. - aB) * jb[d * N + i] / nb;
                }
                double taua = dens[3 * N + i], taub = dens[8 * N + i];
                double dalphm_dna = der_alpha_minus_na(c, na, nb), dalphm_dnb = der_alpha_minus_nb(c, na, nb);
                Va_const += dalphm_dna * (taua - taub) / 2.0;
                Vb_const += dalphm_dnb * (taua - taub) / 2.0;
                dalphp_dna = der_alpha_plus_na(c, na, nb);
                dalphp_dnb = der_alpha_plus_nb(c, na, nb);
                Va_const += dalphp_dna * (taua + taub) / 2.0;
                Vb_const += dalphp_dnb * (taua + taub) / 2.0;
                Va_const += der_funD_na(c, na, nb);
                Vb_const += der_funD_nb(c, na, nb);
                double j2a = 0.0, j2b = 0.0;
                for (int d = 0; d < c->dim; d++) {
                    j2a += ja[d * N + i] * ja[d * N + i];
                    j2b += jb[d * N + i] * jb[d * N + i];
                }
                if (fa != 0.0) {
                    double t7 = fa * j2a / (2.0 * na);
                    Va_const += ((alph_plus + alph_minus - 1.0) / na - (dalphp_dna + dalphm_dna)) * t7;
                    Va_const -= der_p_reg(c, na) * (alph_plus + alph_minus - 1.0) * j2a / (2.0 * na);
                    Vb_const -= (dalphp_dnb + dalphm_dnb) * t7;
                }
                if (fb != 0.0) {
                    double t7 = fb * j2b / (2.0 * nb);
                    Va_const -= (dalphp_dna - dalphm_dna) * t7;
                    Vb_const += ((alph_plus - alph_minus - 1.0) / nb - (dalphp_dnb - dalphm_dnb)) * t7;
                    Vb_const -= der_p_reg(c, nb) * (alph_plus - alph_minus - 1.0) * j2b / (2.0 * nb);
                }
            } else {
                Va_const += der_funD_na(c, na, nb);
                Vb_const += der_funD_nb(c, na, nb);
            }
            double t1 = dalphp_dna / alph_plus, t2 = dalphp_dnb / alph_plus;
            double t3 = der_tildeC_na(c, na, nb) / alph_plus;
            double t4 = der_tildeC_nb(c, na, nb) / alph_plus;
            double t5 = tildeC(c, na, nb);
            double Va = Va_[i], Vb = Vb_[i];
            zc ldelta = 0;
            double lkF = 0.0;
            if (c->mf_correction) lkF = pow(3.0 * M_PI * M_PI * (na + nb), 1.0 / 3.0);
            for (int k = 0; k < UD_SCITERS; k++) {
                double t7 = (c->mu_a - Va + c->mu_b - Vb) / 2.0;
                zc p0 = csqrt(CMPLX(2.0 * t7 / alph_plus, 0.0));
                if (cimag(p0) < 0.) p0 *= -1.;
                zc kc;
                if (c->regularization == TDWO_CUBIC) kc = M_PI / DXs;
                else {
                    kc = csqrt(CMPLX(2.0 * (c->ec + t7) / alph_plus, 0.0));
                    if (cimag(kc) < 0.) kc *= -1.;
                }
                zc wz = clog((kc + p0) / (kc - p0));
                if (cimag(wz) < 0.) wz += CMPLX(0.0, 2. * M_PI);
                wz = kc * R0 * (1. - (p0 / kc) * R1 * wz);
                wz = alph_plus / (t5 - wz);
                ldelta = lnu * (-1.0 * creal(wz));
                double t6 = creal(conj(ldelta) * lnu);
                double t7b = nrm2(ldelta);
                double Vanew = Va_const - t1 * t6 - t3 * t7b;
                double Vbnew = Vb_const - t2 * t6 - t4 * t7b;
                if (c->mf_correction) {
                    double bcoeff = creal(p0 / (lkF + 1.0e-12));
                    double lx = bcoeff * lkF * DXs / M_PI;
                    if (creal(wz) < -1.0e-10 && lx > 1.0e-10) {
                        double Lam0 = bcoeff * R0 / lx * (1.0 - R1 * lx * log((1.0 + lx) / (1.0 - lx)));
                        double dLam0 = (-2. * bcoeff * R0 * R1) / (1. - lx * lx) - bcoeff * R0 / lx / lx;
                        double dLam_dn = (lkF / (3. * (na + nb) * alph_plus)) * (Lam0 + dLam0 * bcoeff * DXs * lkF / M_PI);
                        Vanew += dLam_dn * t7b;
                        Vbnew += dLam_dn * t7b;
                    }
                }
                Va = UD_MIX_COEFF * Vanew + (1.0 - UD_MIX_COEFF) * Va;
                Vb = UD_MIX_COEFF * Vbnew + (1.0 - UD_MIX_COEFF) * Vb;
            }
            Va_[i] = Va;
            Vb_[i] = Vb;
            delta[i] = ldelta;
        } else {
            double Va = ua[i], Vb = ub[i];
            double alph_plus = 0.5 * (al_a[i] + al_b[i]);
            double t5 = 1.0 / (4.0 * M_PI * c->sclgth * alph_plus);
            double lkF = 0.0, nab = na + nb;
            if (c->mf_correction) lkF = pow(3.0 * M_PI * M_PI * nab, 1.0 / 3.0);
            double t7 = (c->mu_a - Va_[i] + c->mu_b - Vb_[i]) / 2.0;
            zc p0 = csqrt(CMPLX(2.0 * t7, 0.0));
            if (cimag(p0) < 0.) p0 *= -1.;
            zc kc;
            if (c->regularization == TDWO_CUBIC) kc = M_PI / DXs;
            else {
                kc = csqrt(CMPLX(2.0 * (c->ec + t7), 0.0));
                if (cimag(kc) < 0.) kc *= -1.;
            }
            zc wz = clog((kc + p0) / (kc - p0));
            if (cimag(wz) < 0.) wz += CMPLX(0.0, 2. * M_PI);
            wz = kc * R0 / alph_plus * (1. - (p0 / kc) * R1 * wz);
            wz = 1.0 / (t5 - wz);
            zc ldelta = lnu * (-1.0 * creal(wz));
            if (c->mf_correction) {
                double bcoeff = creal(p0 / (lkF + 1.0e-12));
                double lx = bcoeff * lkF * DXs / M_PI;
                if (creal(wz) < -1.0e-10 && lx > 1.0e-10) {
                    double Lam0 = bcoeff * R0 / lx * (1.0 - R1 * lx * log((1.0 + lx) / (1.0 - lx)));
                    double dLam0 = (-2. * bcoeff * R0 * R1) / (1. - lx * lx) - bcoeff * R0 / lx / lx;
                    double dLam_dn = (lkF / (3. * nab * alph_plus)) * (Lam0 + dLam0 * bcoeff * DXs * lkF / M_PI);
                    double t7b = nrm2(ldelta);
                    Va += dLam_dn * t7b;
                    Vb += dLam_dn * t7b;
                }
            }
            Va_[i] = Va;
            Vb_[i] = Vb;
            delta[i] = ldelta;
        }
    }
    free(ua);
    free(ub);
}

/* =========================================================================
 * Hamiltonian (pca_kernels.cu:233-391 FAST mode, :393-495 BdG;
 * 1D/2D transverse kinetic energy ccpca_kernels.cu:303-312, cpca:330-342;
 * qpe and subtraction pca_kernels.cu:552-594, 806-832)
 * ========================================================================= */
void tdwo_apply_h(const tdwo_cfg *c, int n, const double *in, double *out,
                  const double *lap, const double *pots, const double *kt,
                  const double *useqpe, double *qpe)
{
    size_t N = tdwo_npts(c);
    const zc *W = (const zc *)in, *L = (const zc *)lap;
    zc *O = (zc *)out;
    const double *Va = pots, *Vb = pots + N, *al_a = pots + 4 * N, *al_b = pots + 5 * N;
    const zc *D = (const zc *)(pots + 2 * N);
    double dv = dvol_norm(c);
    double *e = (double *)malloc(sizeof(double) * (size_t)n);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) OMP_NT
#endif
    for (int a = 0; a < n; a++) {
        double k2t = 0.0;
        if (c->dim == 1) k2t = kt[n + a] * kt[n + a] + kt[a] * kt[a];
        else if (c->dim == 2) k2t = kt[a] * kt[a];
        double ssum = 0.0;
        for (size_t i = 0; i < N; i++) {
            double aa, ab;
            if (c->functional == TDWO_SLDA) { aa = al_a[i] * -0.5; ab = al_b[i] * -0.5; }
            else { aa = -0.5; ab = -0.5; }
            double ja = -1.0 * aa * k2t, jb = -1.0 * ab * k2t;
            zc tu = W[(size_t)a * N + i], tv = W[((size_t)n + a) * N + i];
            zc u = 0, v = 0;
            u += tu * CMPLX(Va[i] - c->mu_a + ja, 0.0);
            v -= tv * CMPLX(Vb[i] - c->mu_b + jb, -1.0 * 0.0);
            u += tv * D[i];
            v += tu * conj(D[i]);
            zc lu = L[(size_t)a * N + i], lv = L[((size_t)n + a) * N + i];
            u += lu * aa;
            v -= lv * ab;
            O[(size_t)a * N + i] = u;
            O[((size_t)n + a) * N + i] = v;
            if (!useqpe) ssum += creal(conj(tu) * u + conj(tv) * v) * dv;
        }
        e[a] = useqpe ? useqpe[a] : ssum;
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static) OMP_NT
#endif
    for (int a = 0; a < n; a++) {
        double ea = e[a];
        for (size_t i = 0; i < N; i++) {
            O[(size_t)a * N + i] -= W[(size_t)a * N + i] * ea;
            O[((size_t)n + a) * N + i] -= W[((size_t)n + a) * N + i] * ea;
        }
    }
    if (qpe) memcpy(qpe, e, sizeof(double) * (size_t)n);
    free(e);
}

/* =========================================================================
 * Hamiltonian with the density-dependent effective mass and the current
 * corrections (ASLDA; pca_kernels.cu:233-391 and 674-790 without
 * FAST_CONST_EFFECTIVE_MASS_MODE, with CURRENT_CORRECTIONS; 1D / 2D twins
 * ccpca_kernels.cu:209-340, cpca_kernels.cu:231-370):
 *   (H psi)_u = (V_a - mu_a + alpha_a k_t^2/2 - i/2 div A_a) u + Delta v
 *               + lap(alpha_a) u / 4 - i A_a . grad u
 *               - alpha_a lap(u) / 4 - lap(alpha_a u) / 4
 * and the conjugate structure for v.  alap holds lap(alpha psi) (the
 * reference's alphawf_laplace), ddx..ddz the gradients of psi along the
 * evolved axes.  Other functionals: tdwo_apply_h.
 * ========================================================================= */
void tdwo_apply_h_full(const tdwo_cfg *c, int n, const double *in, double *out, const double *ddx,
                       const double *ddy, const double *ddz, const double *lap, const double *alap,
                       const double *pots, const double *kt, const double *useqpe, double *qpe)
{
    if (c->functional != TDWO_ASLDA && !(c->functional == TDWO_SLDAE && !c->force_a1)) {
        tdwo_apply_h(c, n, in, out, lap, pots, kt, useqpe, qpe);
        return;
    }
    size_t N = tdwo_npts(c);
    const zc *W = (const zc *)in, *L = (const zc *)lap, *AL = (const zc *)alap;
    const zc *G[3] = {(const zc *)ddx, (const zc *)ddy, (const zc *)ddz};
    zc *O = (zc *)out;
    const double *Va = pots, *Vb = pots + N, *al_a = pots + 4 * N, *al_b = pots + 5 * N;
    const double *Aa = pots + 6 * N, *Ab = pots + 9 * N;
    const zc *D = (const zc *)(pots + 2 * N);
    double dv = dvol_norm(c);
    /* laplace of the effective masses and div A (compute_laplace_real_f,
     * compute_derivative_real_vector_f: Nyquist derivatives zero) */
    double *lal = (double *)malloc(sizeof(double) * 2 * N), *cj = (double *)malloc(sizeof(double) * 2 * N);
    tdwo_laplace_real(c, al_a, lal);
    tdwo_laplace_real(c, al_b, lal + N);
    {
        double *z = (double *)calloc(2 * N, sizeof(double));
        double *d3[3];
        for (int d = 0; d < 3; d++) d3[d] = (double *)malloc(sizeof(double) * 2 * N);
        double *lz = (double *)malloc(sizeof(double) * 2 * N);
        for (int sp = 0; sp < 2; sp++) {
            const double *A = sp ? Ab : Aa;
            for (size_t i = 0; i < N; i++) cj[sp * N + i] = 0.0;
            for (int d = 0; d < c->dim; d++) {
                for (size_t i = 0; i < N; i++) { z[2 * i] = A[d * N + i]; z[2 * i + 1] = 0.0; }
                tdwo_derivatives(c, 1, z, d3[0], d3[1], d3[2], lz);
                for (size_t i = 0; i < N; i++) cj[sp * N + i] += d3[d][2 * i];
            }
            for (size_t i = 0; i < N; i++) cj[sp * N + i] *= -0.5;
        }
        free(z); free(lz);
        for (int d = 0; d < 3; d++) free(d3[d]);
    }
    double *e = (double *)malloc(sizeof(double) * (size_t)n);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) OMP_NT
#endif
    for (int a = 0; a < n; a++) {
        double k2t = 0.0;
        if (c->dim == 1) k2t = kt[n + a] * kt[n + a] + kt[a] * kt[a];
        else if (c->dim == 2) k2t = kt[a] * kt[a];
        double ssum = 0.0;
        for (size_t i = 0; i < N; i++) {
            double aa = al_a[i] * -0.25, ab = al_b[i] * -0.25;
            double la = 0.25 * lal[i], lb = 0.25 * lal[N + i];
            double ja = -2.0 * aa * k2t, jb = -2.0 * ab * k2t;
            size_t iu = (size_t)a * N + i, iv = ((size_t)n + a) * N + i;
            zc tu = W[iu], tv = W[iv];
            zc u = 0, v = 0;
            u += tu * CMPLX(Va[i] - c->mu_a + ja, cj[i]);
            v -= tv * CMPLX(Vb[i] - c->mu_b + jb, -1.0 * cj[N + i]);
            u += tv * D[i];
            v += tu * conj(D[i]);
            u += tu * la;
            v -= tv * lb;
            for (int d = 0; d < c->dim; d++) {
                u += G[d][iu] * CMPLX(0.0, -1. * Aa[d * N + i]);
                v -= G[d][iv] * CMPLX(0.0, 1. * Ab[d * N + i]);
            }
            u += L[iu] * aa;
            v -= L[iv] * ab;
            u += AL[iu] * (-0.25);
            v -= AL[iv] * (-0.25);
            O[iu] = u;
            O[iv] = v;
            if (!useqpe) ssum += creal(conj(tu) * u + conj(tv) * v) * dv;
        }
        e[a] = useqpe ? useqpe[a] : ssum;
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static) OMP_NT
#endif
    for (int a = 0; a < n; a++) {
        double ea = e[a];
        for (size_t i = 0; i < N; i++) {
            O[(size_t)a * N + i] -= W[(size_t)a * N + i] * ea;
            O[((size_t)n + a) * N + i] -= W[((size_t)n + a) * N + i] * ea;
        }
    }
    if (qpe) memcpy(qpe, e, sizeof(double) * (size_t)n);
    free(e); free(lal); free(cj);
}

/* alpha_a u, alpha_b v (multiply_wf_by_alpha, pca_kernels.cu:981-1026) */
void tdwo_alpha_wf(const tdwo_cfg *c, int n, const double *in, const double *pots, double *out)
{
    size_t N = tdwo_npts(c);
    const zc *W = (const zc *)in;
    zc *O = (zc *)out;
    const double *al_a = pots + 4 * N, *al_b = pots + 5 * N;
    for (int a = 0; a < n; a++)
        for (size_t i = 0; i < N; i++) {
            O[(size_t)a * N + i] = W[(size_t)a * N + i] * al_a[i];
            O[((size_t)n + a) * N + i] = W[((size_t)n + a) * N + i] * al_b[i];
        }
}

/* =========================================================================
 * Energy (tdwslda_energy_generic.h:21-110, tdwslda_functionals.h:308-366,
 * :455-510, pca_kernels.cu:121-209)
 * ========================================================================= */
void tdwo_energy(const tdwo_cfg *c, int it, const double *dens, const double *pots, double *E)
{
    g_se_a1 = c->force_a1;
    size_t N = tdwo_npts(c);
    const zc *nu = (const zc *)dens;
    const double *rho_a = dens + 2 * N, *tau_a = dens + 3 * N, *jax = dens + 4 * N, *jay = dens + 5 * N, *jaz = dens + 6 * N;
    const double *rho_b = dens + 7 * N, *tau_b = dens + 8 * N, *jbx = dens + 9 * N, *jby = dens + 10 * N, *jbz = dens + 11 * N;
    const double *al_a = pots + 4 * N, *al_b = pots + 5 * N;
    const zc *delta = (const zc *)(pots + 2 * N);
    double VE = dvol_energy(c);
    double *ua = (double *)malloc(sizeof(double) * N), *ub = (double *)malloc(sizeof(double) * N);
    tdwo_vext_eval(c, it, ua, ub);
    double s[TDWO_NITEMS];
    for (int k = 0; k < TDWO_NITEMS; k++) s[k] = 0.0;
    for (size_t i = 0; i < N; i++) {
        double na = rho_a[i], nb = rho_b[i];
        double tx1 = jax[i], tx2 = jbx[i];
        double ty1 = c->dim >= 2 ? jay[i] : 0.0, ty2 = c->dim >= 2 ? jby[i] : 0.0;
        double tz1 = c->dim >= 3 ? jaz[i] : 0.0, tz2 = c->dim >= 3 ? jbz[i] : 0.0;
        double ta = tau_a[i] - p_reg(c, na) * (tx1 * tx1 + ty1 * ty1 + tz1 * tz1) / na;
        double tb = tau_b[i] - p_reg(c, nb) * (tx2 * tx2 + ty2 * ty2 + tz2 * tz2) / nb;
        if (c->functional == TDWO_SLDAE) {
            /* tdwslda_functionals.h:815-889 */
            double nt = na + nb, kF = pow(3. * M_PI * M_PI * nt, 1. / 3.), x = fabs(c->sclgth * kF);
            s[TDWO_EKIN] += se_alpha(x) * (ta + tb) / 2. * VE;
            s[TDWO_EPOT] += (3. / 5.) * se_bfun(x) * (kF * kF / 2.) * nt * VE;
        } else if (c->functional == TDWO_SLDA || c->functional == TDWO_ASLDA) {
            double p = polarization(c, na, nb);
            s[TDWO_EKIN] += 0.5 * (alpha_f(c, p) * ta + alpha_f(c, -1.0 * p) * tb) * VE;
            s[TDWO_EPOT] += funD(c, na, nb) * VE;
        } else {
            s[TDWO_EKIN] += 0.5 * (al_a[i] * ta + al_b[i] * tb) * VE;
            s[TDWO_EPOT] += 0.0;
        }
        s[TDWO_EPAIR] += creal(delta[i] * conj(nu[i])) * (-1.0) * VE;
        s[TDWO_ECURRENT] += p_reg(c, na) * (tx1 * tx1 + ty1 * ty1 + tz1 * tz1) / (2. * na) * VE +
                            p_reg(c, nb) * (tx2 * tx2 + ty2 * ty2 + tz2 * tz2) / (2. * nb) * VE;
        s[TDWO_EPOTEXT] += (na * ua[i] + nb * ub[i]) * VE;
        s[TDWO_NPARTA] += na;
        s[TDWO_NPARTB] += nb;
        if (c->dim >= 2) {
            int ix, iy, iz;
            decode(c, i, &ix, &iy, &iz);
            double x = (double)(ix - c->nx / 2) * c->dx, y = (double)(iy - c->ny / 2) * c->dy;
            s[TDWO_LZA] += (x * jay[i] - y * jax[i]) * VE;
            s[TDWO_LZB] += (x * jby[i] - y * jbx[i]) * VE;
        }
    }
    s[TDWO_NPARTA] *= VE;
    s[TDWO_NPARTB] *= VE;
    for (int k = 0; k < TDWO_NITEMS; k++) E[k] = s[k];
    free(ua);
    free(ub);
}

/* =========================================================================
 * PECE step (pca.c:1401-1570, FAST_CONST_EFFECTIVE_MASS_MODE, no quantum
 * friction, no spin symmetry).
 * ========================================================================= */
static void rotate_history(const tdwo_cfg *c, tdwo_state *s)
{
    /* pca.c:1543-1562 */
    int k = c->scheme == TDWO_AB3AM4 ? 3 : c->scheme == TDWO_AB4AM5 ? 4 : 5;
    double *tmp = s->f[k - 1];
    for (int j = k - 1; j > 0; j--) s->f[j] = s->f[j - 1];
    s->f[0] = tmp;
}

static void half_densities(const tdwo_cfg *c, tdwo_state *s, int grad)
{
    int n = s->n;
    if (grad) tdwo_derivatives(c, 2 * n, s->wf, s->ddx, s->ddy, s->ddz, s->lap);
    else tdwo_laplace(c, 2 * n, s->wf, s->lap);
    tdwo_densities(c, n, s->wf, s->ddx, s->ddy, s->ddz, s->lap, s->fbeta, s->kt, s->cnt, s->dens, grad);
    if (grad) tdwo_density_tau(c, s->dens);
}

/* One H application of a half-step.  ASLDA (not FAST_CONST_EFFECTIVE_MASS
 * _MODE, pca.c:1443-1465): gradients every half-step, then
 * lap(alpha psi) and the full Hamiltonian. */
static void half_h(const tdwo_cfg *c, tdwo_state *s, double *out)
{
    int n = s->n;
    if (c->functional == TDWO_ASLDA || (c->functional == TDWO_SLDAE && !c->force_a1)) {
        size_t N = tdwo_npts(c);
        double *al = (double *)malloc(sizeof(double) * 4 * (size_t)n * N);
        tdwo_alpha_wf(c, n, s->wf, s->pots, al);
        tdwo_laplace(c, 2 * n, al, al);
        tdwo_apply_h_full(c, n, s->wf, out, s->ddx, s->ddy, s->ddz, s->lap, al, s->pots, s->kt, NULL, s->qpe);
        free(al);
    } else {
        tdwo_apply_h(c, n, s->wf, out, s->lap, s->pots, s->kt, NULL, s->qpe);
    }
}

/* hkf_mode filtering of the mean fields before an H application: alpha_a,
 * alpha_b in the varying-mass modes (pca.c:1448-1452), V_a, V_b
 * (pca_kernels.cu:733-737), Delta for hkf_mode >= 2 (:758-762) and the
 * vector potentials A_a, A_b with current corrections (:786-790), each in
 * place with the Fermi-Dirac filter. */
static void hkf_fields(const tdwo_cfg *c, double *pots)
{
    if (c->hkf_mode < 1) return;
    const size_t N = tdwo_npts(c);
    const int vm = c->functional == TDWO_ASLDA || (c->functional == TDWO_SLDAE && !c->force_a1);
    if (vm) tdwo_hf_filter_d(c, 2, pots + 4 * N, pots + 4 * N, c->hkf_mu, c->hkf_T);
    tdwo_hf_filter_d(c, 2, pots, pots, c->hkf_mu, c->hkf_T);
    if (c->hkf_mode >= 2) tdwo_hf_filter_c(c, pots + 2 * N, pots + 2 * N, c->hkf_mu, c->hkf_T);
    if (vm) tdwo_hf_filter_d(c, 6, pots + 6 * N, pots + 6 * N, c->hkf_mu, c->hkf_T);
}

int tdwo_step(const tdwo_cfg *c, tdwo_state *s, int it, int grad_last)
{
    int n = s->n;
    int last = c->scheme == TDWO_AB3AM4 ? 2 : c->scheme == TDWO_AB4AM5 ? 3 : 4;
    const int full = c->functional == TDWO_ASLDA || (c->functional == TDWO_SLDAE && !c->force_a1);
    /* predictor */
    tdwo_abm_step1(c, n, s->wf, s->f[0], s->f[1], s->f[2], s->f[3], s->f[4]);
    tdwo_normalize(c, n, s->wf);
    half_densities(c, s, full);
    tdwo_potentials(c, it + 1, s->dens, s->pots);
    hkf_fields(c, s->pots);
    {
        /* the H output goes to lap (the reference's buffer), which the
         * full Hamiltonian still reads: stage it */
        size_t N = tdwo_npts(c);
        double *h = (double *)malloc(sizeof(double) * 4 * (size_t)n * N);
        half_h(c, s, h);
        memcpy(s->lap, h, sizeof(double) * 4 * (size_t)n * N);
        free(h);
    }
    /* corrector */
    tdwo_abm_step4(c, n, s->lap, s->wf, s->f[last]);
    tdwo_normalize(c, n, s->wf);
    half_densities(c, s, grad_last || full);
    tdwo_potentials(c, it + 1, s->dens, s->pots);
    hkf_fields(c, s->pots);
    rotate_history(c, s);
    half_h(c, s, s->f[0]);
    return 0;
}

/* Taylor term of the self-start (pca_kernels.cu:1032-1085): contr =
 * (-i dt / order) hpsi, upd += contr, over 2 n N complex values. */
void tdwo_taylor(const tdwo_cfg *c, int order, double dt, int n, const double *hpsi, double *upd, double *contr)
{
    const size_t tot = 2 * (size_t)n * tdwo_npts(c);
    const zc *H = (const zc *)hpsi;
    zc *U = (zc *)upd, *T = (zc *)contr;
    const double tc = -1.0 * dt / (double)order;
    for (size_t i = 0; i < tot; i++) {
        const zc v = H[i] * (I * tc);
        T[i] = v;
        U[i] += v;
    }
}

/* Derivatives the self-start takes of an array before H is applied to it
 * (pca.c:1119-1126, 1179-1186): all of them with gradients, the Laplacian
 * alone otherwise; lap(alpha psi) into al with a varying mass
 * (pca.c:1152-1154, 1188-1191). */
static void ss_derivs(const tdwo_cfg *c, tdwo_state *s, const double *in, int grad, int vm, double *al)
{
    const int n = s->n;
    if (grad) tdwo_derivatives(c, 2 * n, in, s->ddx, s->ddy, s->ddz, s->lap);
    else tdwo_laplace(c, 2 * n, in, s->lap);
    if (vm) {
        tdwo_alpha_wf(c, n, in, s->pots, al);
        tdwo_laplace(c, 2 * n, al, al);
    }
}

/* apply_hamiltonian as the self-start calls it (qfswitch 0): the mean-field
 * filter acts on V, Delta and A in place inside every call
 * (pca_kernels.cu:733-764, 786-790), then H - <H> (useqpe NULL: <H> into
 * qpe; otherwise the given values). */
static void ss_apply_h(const tdwo_cfg *c, tdwo_state *s, const double *in, double *out, int vm, const double *al,
                       const double *useqpe, double *qpe)
{
    const int n = s->n;
    if (c->hkf_mode >= 1) {
        const size_t N = tdwo_npts(c);
        tdwo_hf_filter_d(c, 2, s->pots, s->pots, c->hkf_mu, c->hkf_T);
        if (c->hkf_mode >= 2) tdwo_hf_filter_c(c, s->pots + 2 * N, s->pots + 2 * N, c->hkf_mu, c->hkf_T);
        if (vm) tdwo_hf_filter_d(c, 6, s->pots + 6 * N, s->pots + 6 * N, c->hkf_mu, c->hkf_T);
    }
    if (vm) tdwo_apply_h_full(c, n, in, out, s->ddx, s->ddy, s->ddz, s->lap, al, s->pots, s->kt, useqpe, qpe);
    else tdwo_apply_h(c, n, in, out, s->lap, s->pots, s->kt, useqpe, qpe);
}

/* Densities and potentials of an array, at iteration 0 (pca.c:1127-1159,
 * 1223-1251): densities (with gradients, tau), the functional, and with a
 * varying mass the filter of alpha. */
static void ss_fields(const tdwo_cfg *c, tdwo_state *s, const double *wf, int grad, int vm)
{
    const int n = s->n;
    tdwo_densities(c, n, wf, s->ddx, s->ddy, s->ddz, s->lap, s->fbeta, s->kt, s->cnt, s->dens, grad);
    if (grad) tdwo_density_tau(c, s->dens);
    tdwo_potentials(c, 0, s->dens, s->pots);
    if (vm && c->hkf_mode >= 1) {
        const size_t N = tdwo_npts(c);
        tdwo_hf_filter_d(c, 2, s->pots + 4 * N, s->pots + 4 * N, c->hkf_mu, c->hkf_T);
    }
}

/* Self-start (pca.c:1086-1392, selfstart 1): K = 3, 4, 5 steps (AB3AM4,
 * AB4AM5, AB5AM5), each the midpoint rule with exp(-i (H - <H>) dt) taken as
 * its Taylor polynomial of order E = 4, 5, 5:
 *   predictor: densities / potentials of psi (it = 0), F = (H - <H>) psi
 *     kept as this step's history entry, psi_m = sum_k (-i dt/2 (H - <H>))^k
 *     psi / k!, normalised;
 *   corrector: densities / potentials of psi_m, then psi <- sum_k (-i dt (H
 *     - <H>))^k psi / k!, normalised, with H of the midpoint potentials and
 *     <H> of its first application.
 * The gradients are taken only with a varying effective mass (pca.c:1117,
 * 1213: no quantum friction here).  The history entries go through a host
 * buffer whose shift copies slot j to slot j + 1 for increasing j
 * (pca.c:1166), so every slot but the first ends up holding the same entry:
 * fkm1 = F_{K-1} and fkm2 .. fkmK = F_{K-2}; the restatement keeps that.
 * The iteration counter is not advanced (pca.c:1144, 1240).  Afterwards:
 * derivatives, densities with gradients, tau and the potentials of the new
 * state (pca.c:1336-1356).  f[0..2] are used as scratch (fkm1..fkm3). */
int tdwo_selfstart(const tdwo_cfg *c, tdwo_state *s)
{
    const int n = s->n;
    const size_t N = tdwo_npts(c);
    const size_t tot = 2 * (size_t)n * N;
    const int K = c->scheme == TDWO_AB3AM4 ? 3 : c->scheme == TDWO_AB4AM5 ? 4 : 5;
    const int E = c->scheme == TDWO_AB3AM4 ? 4 : 5;
    const int vm = c->functional == TDWO_ASLDA || (c->functional == TDWO_SLDAE && !c->force_a1);
    const int grad = vm;
    const double dt = c->dt;
    double *m1 = s->f[0], *m2 = s->f[1], *m3 = s->f[2];
    double *al = vm ? (double *)malloc(sizeof(zc) * tot) : NULL;
    double *qpe0 = (double *)malloc(sizeof(double) * (size_t)n);
    double *slots = (double *)malloc(sizeof(zc) * tot * (size_t)K);
    for (int step = 0; step < K; step++) {
        /* predictor: half a step from psi */
        ss_derivs(c, s, s->wf, grad, 0, NULL);
        ss_fields(c, s, s->wf, grad, vm);
        if (vm) ss_derivs(c, s, s->wf, grad, vm, al);
        memcpy(m3, s->wf, sizeof(zc) * tot);
        ss_apply_h(c, s, m3, m1, vm, al, NULL, qpe0);
        for (int j = 0; j < K - 1; j++)
            memcpy(slots + 2 * tot * (size_t)(j + 1), slots + 2 * tot * (size_t)j, sizeof(zc) * tot);
        memcpy(slots, m1, sizeof(zc) * tot);
        tdwo_taylor(c, 1, 0.5 * dt, n, m1, m3, m2);
        for (int k = 1; k < E; k++) {
            ss_derivs(c, s, m2, grad, vm, al);
            ss_apply_h(c, s, m2, m1, vm, al, qpe0, NULL);
            tdwo_taylor(c, k + 1, 0.5 * dt, n, m1, m3, m2);
        }
        tdwo_normalize(c, n, m3);
        /* corrector: the whole step from psi with the midpoint's potentials */
        ss_derivs(c, s, m3, grad, 0, NULL);
        ss_fields(c, s, m3, grad, vm);
        ss_derivs(c, s, s->wf, grad, vm, al);
        ss_apply_h(c, s, s->wf, m1, vm, al, NULL, qpe0);
        tdwo_taylor(c, 1, dt, n, m1, s->wf, m2);
        for (int k = 1; k < E; k++) {
            ss_derivs(c, s, m2, grad, vm, al);
            ss_apply_h(c, s, m2, m1, vm, al, qpe0, NULL);
            tdwo_taylor(c, k + 1, dt, n, m1, s->wf, m2);
        }
        tdwo_normalize(c, n, s->wf);
    }
    for (int j = 0; j < K; j++) memcpy(s->f[j], slots + 2 * tot * (size_t)j, sizeof(zc) * tot);
    memcpy(s->qpe, qpe0, sizeof(double) * (size_t)n);
    /* measurement of the new state (pca.c:1336-1356) */
    tdwo_derivatives(c, 2 * n, s->wf, s->ddx, s->ddy, s->ddz, s->lap);
    tdwo_densities(c, n, s->wf, s->ddx, s->ddy, s->ddz, s->lap, s->fbeta, s->kt, s->cnt, s->dens, 1);
    tdwo_density_tau(c, s->dens);
    tdwo_potentials(c, 0, s->dens, s->pots);
    free(slots);
    free(qpe0);
    free(al);
    return 0;
}

void tdwo_initial_measure(const tdwo_cfg *c, tdwo_state *s)
{
    tdwo_normalize(c, s->n, s->wf);
    half_densities(c, s, 1);
}

/* =========================================================================
 * Uniform solution (pca_uniform.h:70-502, 636-850, 2044-2560)
 * ========================================================================= */
double tdwo_fbeta(double E, double beta)
{
    double bE = beta * E;
    if (bE > 50.) return 0.0;
    if (bE < -50.) return 1.0;
    return 1.0 / (exp(bE) + 1.0);
}

static double *klist(int n, double L)
{
    double *k = (double *)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; i++) k[i] = kval(i, n, L);
    return k;
}

static void bdg_branch(double eta_a, double eta_b, double delta, int branch, double *ek, double *uk, double *vk)
{
    double sq = sqrt((eta_b - eta_a) * (eta_b - eta_a) + 4.0 * (eta_a * eta_b + delta * delta));
    double e, v;
    if (branch == 0) {
        e = -0.5 * (eta_b - eta_a) + 0.5 * sq;
        v = delta * delta / (pow(0.5 * (eta_a + eta_b) + 0.5 * sq, 2) + delta * delta);
    } else {
        e = -0.5 * (eta_b - eta_a) - 0.5 * sq;
        v = delta * delta / (pow(0.5 * (eta_a + eta_b) - 0.5 * sq, 2) + delta * delta);
    }
    double u = 1.0 - v;
    v = sqrt(v);
    u = sqrt(u);
    if (eta_b + e < 0.0) v = -1.0 * v;
    *ek = e; *uk = u; *vk = v;
}

/* solve_uniform_problem_bdg (pca_uniform.h:1105-1470): BdG functional with
 * bare coupling 4*pi*a (a = sclgth), alpha = 1, no density functional in the
 * mean fields (only the mean-field correction of the regularisation), and
 * the chemical-potential update mu += muchange*(n0 - n) (:1431-1432). */
static int uniform_solve_bdg(const tdwo_cfg *c, double n0_a, double n0_b, double kc_in, double Tstart,
                             double Tstop, double dT, double eps, double scmix, double muchange, int maxiter,
                             tdwo_uniform *U)
{
    int NX = c->nx, NY = c->ny, NZ = c->nz;
    double LX = c->dx * NX, LY = c->dy * NY, LZ = c->dz * NZ, LXYZ = LX * LY * LZ;
    double *kkx = klist(NX, LX), *kky = klist(NY, LY), *kkz = klist(NZ, LZ);
    size_t NXYZ = (size_t)NX * NY * NZ;
    double *kk2 = (double *)malloc(sizeof(double) * NXYZ);
    size_t i = 0;
    for (int ix = 0; ix < NX; ix++) for (int iy = 0; iy < NY; iy++) for (int iz = 0; iz < NZ; iz++, i++)
        kk2[i] = kkx[ix] * kkx[ix] + kky[iy] * kky[iy] + kkz[iz] * kkz[iz];
    double eF_a = pow(6.0 * M_PI * M_PI * n0_a, 2.0 / 3.0) / 2.0;
    double eF_b = pow(6.0 * M_PI * M_PI * n0_b, 2.0 / 3.0) / 2.0;
    double eF_avg = pow(3.0 * M_PI * M_PI * (n0_a + n0_b), 2.0 / 3.0) / 2.0;
    double gbare = 4.0 * M_PI * c->sclgth;
    double kc = kc_in;
    const double alph_a = 1.0, alph_b = 1.0;
    double mu_a = 0.96087287 * eF_a, mu_b = 0.96087287 * eF_b;
    double tau_a = pow(6.0 * M_PI * M_PI * n0_a, 5.0 / 3.0) / (10.0 * M_PI * M_PI);
    double tau_b = pow(6.0 * M_PI * M_PI * n0_b, 5.0 / 3.0) / (10.0 * M_PI * M_PI);
    double delta = (n0_a < n0_b) ? 0.21 * pow(6.0 * M_PI * M_PI * n0_a, 2.0 / 3.0) / 2.0
                                 : 0.21 * pow(6.0 * M_PI * M_PI * n0_b, 2.0 / 3.0) / 2.0;
    double nu = -1.0 * delta / gbare;
    double V_a = 0.0, V_b = 0.0, V_a_old, V_b_old, tau_a_old, tau_b_old, delta_old, nu_old;
    double n_a = 0, n_b = 0, S = 0, beta = 1.0, T;
    double ec = c->ec;
    double R0, R1;
    reg_coeffs(c, &R0, &R1);
    int iter = 0, nwf = 0;
    double dsT = 1.0, TT;
    double Ts = fabs(Tstart), Tp = fabs(Tstop), DT = fabs(dT);
    if (Ts < Tp) dsT = -1.0;
    for (TT = Ts * dsT; TT >= Tp * dsT; TT -= DT) {
        T = TT * dsT;
        if (T < 0.0) break;
        beta = (T <= 1.0e-16) ? 1.0e16 : 1.0 / (T * eF_avg);
        for (iter = 0; iter < maxiter; iter++) {
            tau_a_old = tau_a; tau_b_old = tau_b; delta_old = delta; nu_old = nu;
            V_a_old = V_a; V_b_old = V_b;
            V_a = 0.0;
            V_b = 0.0;
            double mu_p = (mu_a - V_a_old + mu_b - V_b_old) / 2.0;
            zc p0 = csqrt(CMPLX(2.0 * mu_p, 0.0));
            if (cimag(p0) < 0.) p0 *= -1.;
            if (c->regularization == TDWO_CUBIC) kc = M_PI / c->dx;
            else ec = kc * kc / 2.0 - mu_p;
            zc wz = clog((kc + p0) / (kc - p0));
            if (cimag(wz) < 0.) wz += I * 2. * M_PI;
            wz = kc * R0 * (1. - (p0 / kc) * R1 * wz);
            double g_eff = creal(1.0 / (1.0 / gbare - wz));
            delta = -1.0 * g_eff * nu;
            if (c->mf_correction) {
                double kF = pow(3.0 * M_PI * M_PI * (n0_a + n0_b), 1.0 / 3.0);
                double bcoeff = creal(p0) / (kF + 1.0e-12);
                double x = bcoeff * kF * c->dx / M_PI;
                if (g_eff < -1.0e-10 && x > 1.0e-10) {
                    double Lam_0 = bcoeff * R0 / x * (1.0 - R1 * x * log((1.0 + x) / (1.0 - x)));
                    double dLam_0_dx = (-2. * bcoeff * R0 * R1) / (1. - x * x) - bcoeff * R0 / x / x;
                    double dLam_dn = (kF / (3. * (n0_a + n0_b))) * (Lam_0 + dLam_0_dx * bcoeff * c->dx * kF / M_PI);
                    V_a += dLam_dn * pow(delta, 2);
                    V_b += dLam_dn * pow(delta, 2);
                }
            }
            S = 0.0; n_a = 0.0; n_b = 0.0; tau_a = 0.0; tau_b = 0.0; nu = 0.0; nwf = 0;
            size_t ixyz = 0;
            for (int ix = 0; ix < NX; ix++) for (int iy = 0; iy < NY; iy++) for (int iz = 0; iz < NZ; iz++, ixyz++) {
                if (c->regularization != TDWO_CUBIC) {
                    if ((c->dim == 2 || c->dim == 1) && iz == NZ / 2) continue;
                    if (c->dim == 1 && iy == NY / 2) continue;
                }
                double eta_a = alph_a * kk2[ixyz] / 2.0 + V_a - mu_a;
                double eta_b = alph_b * kk2[ixyz] / 2.0 + V_b - mu_b;
                for (int br = 0; br < 2; br++) {
                    double ek, uk, vk;
                    bdg_branch(eta_a, eta_b, delta, br, &ek, &uk, &vk);
                    if (c->spinsymmetry) {
                        /* spin-symmetric mode keeps the E > 0 half of the
                         * spectrum and folds spin a into b (pca_uniform.h:325-337) */
                        if (ek > 0.0 && ek < ec) {
                            double fp = tdwo_fbeta(ek, beta), fm = tdwo_fbeta(-1.0 * ek, beta);
                            n_b += vk * vk * fm + uk * uk * fp;
                            tau_b += kk2[ixyz] * (vk * vk * fm + uk * uk * fp);
                            nu += uk * vk * (fm - fp) * 2.0;
                            if (fp > 1.0e-16) S -= 2. * fp * log(fp);
                            if (fm > 1.0e-16) S -= 2. * fm * log(fm);
                            nwf++;
                        }
                    } else if (ek > -ec && ek < ec) {
                        double fp = tdwo_fbeta(ek, beta), fm = tdwo_fbeta(-1.0 * ek, beta);
                        n_a += uk * uk * fp;
                        tau_a += kk2[ixyz] * uk * uk * fp;
                        n_b += vk * vk * fm;
                        tau_b += kk2[ixyz] * vk * vk * fm;
                        nu += uk * vk * (fm - fp);
                        if (fp > 1.0e-16) S -= fp * log(fp);
                        if (fm > 1.0e-16) S -= fm * log(fm);
                        nwf++;
                    }
                }
            }
            if (c->spinsymmetry) {  /* pca_uniform.h:397-402 */
                n_b /= LXYZ; n_a = n_b; tau_b /= LXYZ; tau_a = tau_b; nu /= 2.0 * LXYZ;
            } else {
                n_a /= LXYZ; n_b /= LXYZ; tau_a /= LXYZ; tau_b /= LXYZ; nu /= 2.0 * LXYZ;
            }
            int conv = 1;
            if (fabs(n0_a - n_a) > eps) conv = 0;
            if (fabs(n0_b - n_b) > eps) conv = 0;
            if (fabs(delta - delta_old) > eps) conv = 0;
            if (conv) break;
            tau_a = (1. - scmix) * tau_a_old + scmix * tau_a;
            tau_b = (1. - scmix) * tau_b_old + scmix * tau_b;
            delta = (1. - scmix) * delta_old + scmix * delta;
            nu = (1. - scmix) * nu_old + scmix * nu;
            mu_a += muchange * (n0_a - n_a);
            mu_b += muchange * (n0_b - n_b);
            if (c->spinsymmetry) mu_b = mu_a;  /* pca_uniform.h:1432 */
        }
        eF_a = pow(6.0 * M_PI * M_PI * n_a, 2.0 / 3.0) / 2.0;
        eF_b = pow(6.0 * M_PI * M_PI * n_b, 2.0 / 3.0) / 2.0;
        eF_avg = pow(3.0 * M_PI * M_PI * (n_a + n_b), 2.0 / 3.0) / 2.0;
    }
    U->n0_a = n_a; U->n0_b = n_b; U->mu_a = mu_a; U->mu_b = mu_b; U->V_a = V_a; U->V_b = V_b;
    U->alph_a = alph_a; U->alph_b = alph_b; U->kc = kc; U->ec = ec; U->nu = nu; U->delta = delta;
    U->tau_a = tau_a; U->tau_b = tau_b; U->beta = beta; U->S = S; U->nwf = nwf; U->iters = iter;
    U->ekin = (0.5 * alph_a * tau_a + 0.5 * alph_b * tau_b) * LXYZ;
    U->epot = 0.0;
    U->epair = -1.0 * delta * nu * LXYZ;
    free(kkx); free(kky); free(kkz); free(kk2);
    return iter == maxiter ? 1 : 0;
}

int tdwo_uniform_solve(const tdwo_cfg *c, double n0_a, double n0_b, double kc_in,
                       double Tstart, double Tstop, double dT, double eps,
                       double scmix, double muchange, int maxiter, int mu_rule, tdwo_uniform *U)
{
    if (c->functional == TDWO_BDG)
        return uniform_solve_bdg(c, n0_a, n0_b, kc_in, Tstart, Tstop, dT, eps, scmix, muchange, maxiter, U);
    g_se_a1 = c->force_a1;
    int NX = c->nx, NY = c->ny, NZ = c->nz;
    double LX = c->dx * NX, LY = c->dy * NY, LZ = c->dz * NZ, LXYZ = LX * LY * LZ;
    double *kkx = klist(NX, LX), *kky = klist(NY, LY), *kkz = klist(NZ, LZ);
    size_t NXYZ = (size_t)NX * NY * NZ;
    double *kk2 = (double *)malloc(sizeof(double) * NXYZ);
    size_t i = 0;
    for (int ix = 0; ix < NX; ix++) for (int iy = 0; iy < NY; iy++) for (int iz = 0; iz < NZ; iz++, i++)
        kk2[i] = kkx[ix] * kkx[ix] + kky[iy] * kky[iy] + kkz[iz] * kkz[iz];
    double p = polarization(c, n0_a, n0_b);
    double alph_a = alpha_f(c, p), alph_b = alpha_f(c, -p);
    double alph_plus = alpha_plus(c, p);
    /* pca_uniform.h:157-163 (all zero for SLDA's A1 = A2 = 0) */
    double dalphm_dna = der_alpha_minus_na(c, n0_a, n0_b), dalphm_dnb = der_alpha_minus_nb(c, n0_a, n0_b);
    double dalphp_dna = der_alpha_plus_na(c, n0_a, n0_b), dalphp_dnb = der_alpha_plus_nb(c, n0_a, n0_b);
    double dtC_dna = der_tildeC_na(c, n0_a, n0_b), dtC_dnb = der_tildeC_nb(c, n0_a, n0_b);
    double dD_dna = der_funD_na(c, n0_a, n0_b), dD_dnb = der_funD_nb(c, n0_a, n0_b);
    double tC = tildeC(c, n0_a, n0_b), Dn = funD(c, n0_a, n0_b);
    double eF_a = pow(6.0 * M_PI * M_PI * n0_a, 2.0 / 3.0) / 2.0;
    double eF_b = pow(6.0 * M_PI * M_PI * n0_b, 2.0 / 3.0) / 2.0;
    double eF_avg = pow(3.0 * M_PI * M_PI * (n0_a + n0_b), 2.0 / 3.0) / 2.0;
    double kc = kc_in;
    double mu_a = 0.37 * eF_a, mu_b = 0.37 * eF_b;
    /* SLDAE: solve_uniform_problem_sldae (pca_uniform.h:1532-2043) */
    const int se = c->functional == TDWO_SLDAE;
    se_local SL = se_coeffs(n0_a + n0_b, 1.0, c->sclgth);
    if (se) {
        double x = fabs(c->sclgth * pow(3.0 * M_PI * M_PI * (n0_a + n0_b), 1.0 / 3.0));
        alph_a = alph_b = alph_plus = SL.af;
        tC = SL.ct;
        Dn = (3. / 5.) * SL.bf * (n0_a + n0_b) * SL.eF;
        mu_a = se_mu_ufg(x) * eF_a;
        mu_b = se_mu_ufg(x) * eF_b;
    }
    double tau_a = pow(6.0 * M_PI * M_PI * n0_a, 5.0 / 3.0) / (10.0 * M_PI * M_PI);
    double tau_b = pow(6.0 * M_PI * M_PI * n0_b, 5.0 / 3.0) / (10.0 * M_PI * M_PI);
    double delta = (n0_a < n0_b) ? 0.5 * pow(6.0 * M_PI * M_PI * n0_a, 2.0 / 3.0) / 2.0
                                 : 0.5 * pow(6.0 * M_PI * M_PI * n0_b, 2.0 / 3.0) / 2.0;
    double nu = -1.0 * delta * tC / alph_plus;
    double V_a = 0.0, V_b = 0.0, V_a_old, V_b_old, tau_a_old, tau_b_old, delta_old, nu_old;
    double n_a = 0, n_b = 0, S = 0, beta = 1.0, T;
    double ec = c->ec;
    double R0, R1;
    reg_coeffs(c, &R0, &R1);
    int iter = 0, nwf = 0;
    double dsT = 1.0, TT;
    double Ts = fabs(Tstart), Tp = fabs(Tstop), DT = fabs(dT);
    if (Ts < Tp) dsT = -1.0;
    for (TT = Ts * dsT; TT >= Tp * dsT; TT -= DT) {
        T = TT * dsT;
        if (T < 0.0) break;
        if (T <= 1.0e-16) beta = 1.0e16;
        else beta = 1.0 / (T * eF_avg);
        for (iter = 0; iter < maxiter; iter++) {
            tau_a_old = tau_a; tau_b_old = tau_b; delta_old = delta; nu_old = nu;
            V_a_old = V_a; V_b_old = V_b;
            double tau_p = tau_a + tau_b, tau_m = tau_a - tau_b;
            V_a = dalphm_dna * tau_m / 2.0 + dalphp_dna * (tau_p / 2.0 - delta * nu / alph_plus) - dtC_dna * delta * delta / alph_plus + dD_dna;
            V_b = dalphm_dnb * tau_m / 2.0 + dalphp_dnb * (tau_p / 2.0 - delta * nu / alph_plus) - dtC_dnb * delta * delta / alph_plus + dD_dnb;
            if (se) {
                V_a = SL.af_p * tau_p / 2. + SL.beta * SL.eF - SL.ct_p * delta * delta / SL.af - SL.af_p * delta * nu / SL.af;
                V_b = V_a;
            }
            double mu_p = (mu_a - V_a_old + mu_b - V_b_old) / 2.0;
            zc p0 = csqrt(CMPLX(2.0 * mu_p / alph_plus, 0.0));
            if (cimag(p0) < 0.) p0 *= -1.;
            if (c->regularization == TDWO_CUBIC) kc = M_PI / c->dx;
            else ec = alph_plus * kc * kc / 2.0 - mu_p;
            zc wz = clog((kc + p0) / (kc - p0));
            if (cimag(wz) < 0.) wz += I * 2. * M_PI;
            wz = kc * R0 * (1. - (p0 / kc) * R1 * wz);
            double g_eff = se ? alph_plus / (tC - creal(wz)) : creal(alph_plus / (tC - wz));
            delta = -1.0 * g_eff * nu;
            if (c->mf_correction) {
                double kF = pow(3.0 * M_PI * M_PI * (n0_a + n0_b), 1.0 / 3.0);
                double bcoeff = creal(p0) / (kF + 1.0e-12);
                double x = bcoeff * kF * c->dx / M_PI;
                if (g_eff < -1.0e-10 && x > 1.0e-10) {
                    double Lam_0 = bcoeff * R0 / x * (1.0 - R1 * x * log((1.0 + x) / (1.0 - x)));
                    double dLam_0_dx = (-2. * bcoeff * R0 * R1) / (1. - x * x) - bcoeff * R0 / x / x;
                    double dLam_dn = (kF / (3. * (n0_a + n0_b) * alph_plus)) * (Lam_0 + dLam_0_dx * bcoeff * c->dx * kF / M_PI);
                    V_a += dLam_dn * pow(delta, 2);
                    V_b += dLam_dn * pow(delta, 2);
                }
            }
            S = 0.0; n_a = 0.0; n_b = 0.0; tau_a = 0.0; tau_b = 0.0; nu = 0.0; nwf = 0;
            size_t ixyz = 0;
            for (int ix = 0; ix < NX; ix++) for (int iy = 0; iy < NY; iy++) for (int iz = 0; iz < NZ; iz++, ixyz++) {
                if (c->regularization != TDWO_CUBIC) {
                    if ((c->dim == 2 || c->dim == 1) && iz == NZ / 2) continue;
                    if (c->dim == 1 && iy == NY / 2) continue;
                }
                double eta_a = alph_a * kk2[ixyz] / 2.0 + V_a - mu_a;
                double eta_b = alph_b * kk2[ixyz] / 2.0 + V_b - mu_b;
                for (int br = 0; br < 2; br++) {
                    double ek, uk, vk;
                    bdg_branch(eta_a, eta_b, delta, br, &ek, &uk, &vk);
                    if (c->spinsymmetry) {
                        /* spin-symmetric mode keeps the E > 0 half of the
                         * spectrum and folds spin a into b (pca_uniform.h:325-337) */
                        if (ek > 0.0 && ek < ec) {
                            double fp = tdwo_fbeta(ek, beta), fm = tdwo_fbeta(-1.0 * ek, beta);
                            n_b += vk * vk * fm + uk * uk * fp;
                            tau_b += kk2[ixyz] * (vk * vk * fm + uk * uk * fp);
                            nu += uk * vk * (fm - fp) * 2.0;
                            if (fp > 1.0e-16) S -= 2. * fp * log(fp);
                            if (fm > 1.0e-16) S -= 2. * fm * log(fm);
                            nwf++;
                        }
                    } else if (ek > -ec && ek < ec) {
                        double fp = tdwo_fbeta(ek, beta), fm = tdwo_fbeta(-1.0 * ek, beta);
                        n_a += uk * uk * fp;
                        tau_a += kk2[ixyz] * uk * uk * fp;
                        n_b += vk * vk * fm;
                        tau_b += kk2[ixyz] * vk * vk * fm;
                        nu += uk * vk * (fm - fp);
                        if (fp > 1.0e-16) S -= fp * log(fp);
                        if (fm > 1.0e-16) S -= fm * log(fm);
                        nwf++;
                    }
                }
            }
            if (c->spinsymmetry) {  /* pca_uniform.h:397-402 */
                n_b /= LXYZ; n_a = n_b; tau_b /= LXYZ; tau_a = tau_b; nu /= 2.0 * LXYZ;
            } else {
                n_a /= LXYZ; n_b /= LXYZ; tau_a /= LXYZ; tau_b /= LXYZ; nu /= 2.0 * LXYZ;
            }
            int conv = 1;
            if (fabs(n0_a - n_a) > eps) conv = 0;
            if (fabs(n0_b - n_b) > eps) conv = 0;
            if (fabs(delta - delta_old) > eps) conv = 0;
            if (conv) break;
            tau_a = (1. - scmix) * tau_a_old + scmix * tau_a;
            tau_b = (1. - scmix) * tau_b_old + scmix * tau_b;
            delta = (1. - scmix) * delta_old + scmix * delta;
            nu = (1. - scmix) * nu_old + scmix * nu;
            if (mu_rule == 1) {
                mu_a += muchange * (Ts / T) * (n0_a - n_a);
                mu_b += muchange * (Ts / T) * (n0_b - n_b);
            } else {
                mu_a += muchange * (n0_a - n_a) / n0_a * eF_a;
                mu_b += muchange * (n0_b - n_b) / n0_b * eF_b;
            }
            if (c->spinsymmetry) mu_b = mu_a;  /* pca_uniform.h:429 */
        }
        eF_a = pow(6.0 * M_PI * M_PI * n_a, 2.0 / 3.0) / 2.0;
        eF_b = pow(6.0 * M_PI * M_PI * n_b, 2.0 / 3.0) / 2.0;
        eF_avg = pow(3.0 * M_PI * M_PI * (n_a + n_b), 2.0 / 3.0) / 2.0;
    }
    U->n0_a = n_a; U->n0_b = n_b; U->mu_a = mu_a; U->mu_b = mu_b; U->V_a = V_a; U->V_b = V_b;
    U->alph_a = alph_a; U->alph_b = alph_b; U->kc = kc; U->ec = ec; U->nu = nu; U->delta = delta;
    U->tau_a = tau_a; U->tau_b = tau_b; U->beta = beta; U->S = S; U->nwf = nwf; U->iters = iter;
    U->ekin = (0.5 * alph_a * tau_a + 0.5 * alph_b * tau_b) * LXYZ;
    U->epot = Dn * LXYZ;
    U->epair = -1.0 * delta * nu * LXYZ;
    free(kkx); free(kky); free(kkz); free(kk2);
    return iter == maxiter ? 1 : 0;
}

/* 1D transverse modes: unique ky^2+kz^2 with degeneracy
 * (wslda_wavevectors.h:21-80, spherical cutoff skips the Nyquist rows). */
typedef struct { int cnt; double ky[4096], kz[4096]; int w[4096]; } kmodes1d;
static void fill_kmodes1d(const tdwo_cfg *c, kmodes1d *km)
{
    double *kky = klist(c->ny, c->dy * c->ny), *kkz = klist(c->nz, c->dz * c->nz);
    double lk2[4096];
    km->cnt = 0;
    for (int iy = 0; iy < c->ny; iy++) for (int iz = 0; iz < c->nz; iz++) {
        double k2 = kky[iy] * kky[iy] + kkz[iz] * kkz[iz];
        if (c->regularization != TDWO_CUBIC) {
            if (c->ny > 1 && iy == c->ny / 2) continue;
            if (c->nz > 1 && iz == c->nz / 2) continue;
        }
        int i, has = 0;
        for (i = 0; i < km->cnt; i++) if (fabs(lk2[i] - k2) < 1.0e-12) { has = 1; break; }
        if (has) km->w[i]++;
        else { km->w[km->cnt] = 1; km->ky[km->cnt] = kky[iy]; km->kz[km->cnt] = kkz[iz]; lk2[km->cnt] = k2; km->cnt++; }
    }
    free(kky); free(kkz);
}
static int kmode1d_weight(const kmodes1d *km, double ky, double kz)
{
    for (int i = 0; i < km->cnt; i++)
        if (fabs(km->ky[i] - ky) < 1.0e-12 && fabs(km->kz[i] - kz) < 1.0e-12) return km->w[i];
    return 0;
}

static int uniform_wf_generic(const tdwo_cfg *c, const tdwo_uniform *U, int idxfrom, int idxto,
                              double *wf, double *En, double *kt, int *cnt, int countonly)
{
    int NX = c->nx, NY = c->ny, NZ = c->nz;
    double *kkx = klist(NX, c->dx * NX), *kky = klist(NY, c->dy * NY), *kkz = klist(NZ, c->dz * NZ);
    size_t Nl = tdwo_npts(c);
    double vol = c->dim == 1 ? c->dx * NX : c->dim == 2 ? (c->dx * NX) * (c->dy * NY) : (c->dx * NX) * (c->dy * NY) * (c->dz * NZ);
    double sqv = sqrt(vol);
    kmodes1d *km = NULL;
    if (c->dim == 1) { km = (kmodes1d *)malloc(sizeof(kmodes1d)); fill_kmodes1d(c, km); }
    int nloc = idxto - idxfrom;
    zc *W = (zc *)wf;
    int nwf = -1;
    for (int ix = 0; ix < NX; ix++) for (int iy = 0; iy < NY; iy++) for (int iz = 0; iz < NZ; iz++) {
        int deg = 1;
        if (c->dim == 1) deg = kmode1d_weight(km, kky[iy], kkz[iz]);
        if (c->dim == 2) {
            int sel = (iz == 0) ? 1 : (iz < NZ / 2) ? 2 : 0;
            if (iz == NZ / 2) sel = c->regularization == TDWO_CUBIC ? 1 : 0;
            deg = sel;
        }
        if (deg <= 0) continue;
        double k2 = kkx[ix] * kkx[ix] + kky[iy] * kky[iy] + kkz[iz] * kkz[iz];
        double eta_a = U->alph_a * k2 / 2.0 + U->V_a - U->mu_a;
        double eta_b = U->alph_b * k2 / 2.0 + U->V_b - U->mu_b;
        for (int br = 0; br < 2; br++) {
            double ek, uk, vk;
            bdg_branch(eta_a, eta_b, U->delta, br, &ek, &uk, &vk);
            /* spin-symmetric mode: E > 0 states only (pca_uniform.h:727-733) */
            double elo = c->spinsymmetry ? 0.0 : -U->ec;
            if (!(ek > elo && ek < U->ec)) continue;
            nwf++;
            if (countonly || nwf < idxfrom || nwf >= idxto) continue;
            int l = nwf - idxfrom;
            zc *u = W + (size_t)l * Nl, *v = W + ((size_t)nloc + l) * Nl;
            for (size_t j = 0; j < Nl; j++) {
                int jx, jy, jz;
                decode(c, j, &jx, &jy, &jz);
                double ph = (double)jx * c->dx * kkx[ix];
                if (c->dim >= 2) ph += (double)jy * c->dy * kky[iy];
                if (c->dim >= 3) ph += (double)jz * c->dz * kkz[iz];
                zc e = cexp(I * ph);
                u[j] = e * uk / sqv;
                v[j] = e * vk / sqv;
            }
            if (En) En[l] = ek;
            if (c->dim == 1) { kt[l] = kky[iy]; kt[nloc + l] = kkz[iz]; cnt[l] = deg; }
            if (c->dim == 2) kt[l] = kkz[iz];
        }
    }
    free(kkx); free(kky); free(kkz);
    if (km) free(km);
    return nwf + 1;
}

int tdwo_uniform_count(const tdwo_cfg *c, const tdwo_uniform *u)
{
    return uniform_wf_generic(c, u, 0, 0, NULL, NULL, NULL, NULL, 1);
}

int tdwo_uniform_wf(const tdwo_cfg *c, const tdwo_uniform *u, int idxfrom, int idxto,
                    double *wf, double *En, double *kt, int *cnt)
{
    uniform_wf_generic(c, u, idxfrom, idxto, wf, En, kt, cnt, 0);
    return 0;
}
