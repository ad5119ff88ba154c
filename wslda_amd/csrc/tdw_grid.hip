// tdw_grid.hip -- grid-sized helpers of the reference's derivative and
// kernel layers that act on one field rather than on the state array:
//
//   compute_gradient_real_f            pca_derivative.cu:385-488
//   compute_derivative_real_vector_f   pca_derivative.cu:490-598
//   compute_divergence_real_vector_f   pca_derivative.cu:673-782
//   high_frequency_filter_{d,c,massive_d}  pca_derivative.cu:784-963
//   compute_ovelap                     pca_kernels.cu:853-911
//   (1D / 2D twins: ccpca_derivative.cu:394-800, cpca_derivative.cu:420-880,
//    ccpca_kernels.cu:805-860, cpca_kernels.cu:855-910)
//
// The derivatives use the state path's spectral transform on one field
// (real input carried as complex, the Nyquist derivative zeroed as in the
// reference's D2Z kernels); the real part of the result is the Z2D output.
// The Fermi-Dirac filter FD(k^2/2) is not separable over the axes, so it
// runs as a forward DFT along each axis, the pointwise weight FD/N, and the
// inverse DFT: one thread per output point, L complex products per axis,
// twiddles from the context's exp(-2 pi i m / L) tables.  All of this is
// grid-sized (N up to ~10^5 points).
#include "tdw_device.hpp"
#include "tdw_internal.hpp"
#include "tdw_physics.hpp"

namespace tdw {

static constexpr int GT = 256;

static int ensure_grid(tdw_ctx *c)
{
    // 6 N complex: the field as complex, 4 derivative outputs, one spare
    if (!c->d_gscr) TDW_HIP(hipMalloc(&c->d_gscr, sizeof(double2) * 6 * c->N));
    return 0;
}

static inline dim3 gblocks(long n) { return dim3((unsigned)((n + GT - 1) / GT)); }

__global__ void k_g_real2cplx(const double *__restrict__ f, double2 *__restrict__ z, long N)
{
    const long i = (long)blockIdx.x * GT + threadIdx.x;
    if (i < N) z[i] = make_double2(f[i], 0.0);
}
// out = Re(a)  (acc = 0)  or  out += Re(a)  (acc = 1)
__global__ void k_g_re(const double2 *__restrict__ a, double *out, long N, int acc)
{
    const long i = (long)blockIdx.x * GT + threadIdx.x;
    if (i >= N) return;
    out[i] = acc ? out[i] + a[i].x : a[i].x;
}
__global__ void k_g_copy(const double *__restrict__ a, double *__restrict__ b, long N)
{
    const long i = (long)blockIdx.x * GT + threadIdx.x;
    if (i < N) b[i] = a[i];
}

// derivatives of one real field along the evolved axes: g[d] = d f / dx_d
static int real_field_derivs(tdw_ctx *c, const double *f, double2 *g)
{
    const long N = (long)c->N;
    double2 *z = c->d_gscr + 4 * N;
    hipLaunchKernelGGL(k_g_real2cplx, gblocks(N), dim3(GT), 0, c->stream, f, z, N);
    return derivatives(c, 1, z, g, c->dim >= 2 ? g + N : nullptr, c->dim >= 3 ? g + 2 * N : nullptr, g + 3 * N);
}

int grid_gradient(tdw_ctx *c, const double *f, double *const df[3])
{
    TDW_TRY(ensure_grid(c));
    const long N = (long)c->N;
    for (int d = 0; d < c->dim; d++)
        if (!df[d]) return TDW_EINVAL;
    TDW_TRY(real_field_derivs(c, f, c->d_gscr));
    for (int d = 0; d < c->dim; d++)
        hipLaunchKernelGGL(k_g_re, gblocks(N), dim3(GT), 0, c->stream, c->d_gscr + d * N, df[d], N, 0);
    return (int)hipGetLastError();
}

// df[d] = d f[d] / dx_d (outputs may alias the inputs)
int grid_dvec(tdw_ctx *c, const double *const f[3], double *const df[3])
{
    TDW_TRY(ensure_grid(c));
    const long N = (long)c->N;
    for (int d = 0; d < c->dim; d++)
        if (!f[d] || !df[d]) return TDW_EINVAL;
    for (int d = 0; d < c->dim; d++) {
        TDW_TRY(real_field_derivs(c, f[d], c->d_gscr));
        hipLaunchKernelGGL(k_g_re, gblocks(N), dim3(GT), 0, c->stream, c->d_gscr + d * N, df[d], N, 0);
    }
    return (int)hipGetLastError();
}

// div = sum_d d f[d] / dx_d, summed in a scratch field and written last
// (div may alias an input)
int grid_div(tdw_ctx *c, const double *const f[3], double *div)
{
    TDW_TRY(ensure_grid(c));
    const long N = (long)c->N;
    if (!div) return TDW_EINVAL;
    for (int d = 0; d < c->dim; d++)
        if (!f[d]) return TDW_EINVAL;
    double *acc = reinterpret_cast<double *>(c->d_gscr + 5 * N);
    for (int d = 0; d < c->dim; d++) {
        TDW_TRY(real_field_derivs(c, f[d], c->d_gscr));
        hipLaunchKernelGGL(k_g_re, gblocks(N), dim3(GT), 0, c->stream, c->d_gscr + d * N, acc, N, d > 0);
    }
    hipLaunchKernelGGL(k_g_copy, gblocks(N), dim3(GT), 0, c->stream, acc, div, N);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fermi-Dirac high-frequency filter
// ---------------------------------------------------------------------------
// One output per thread: out[.., k, ..] = sum_j in[.., j, ..] w^(j k) along an
// axis of L points and stride 'stride'; w = tw[1] (forward) or its conjugate.
__global__ void k_dft_axis(const double2 *__restrict__ in, double2 *__restrict__ out, long N, int L, long stride,
                           const double2 *__restrict__ tw, int inverse)
{
    const long o = (long)blockIdx.x * GT + threadIdx.x;
    if (o >= N) return;
    const int k = (int)((o / stride) % L);
    const long base = o - (long)k * stride;
    const double sg = inverse ? -1.0 : 1.0;
    double2 s = make_double2(0.0, 0.0);
    int m = 0;   // (j * k) mod L
    for (int j = 0; j < L; j++) {
        const double2 w = tw[m];
        s = s + cmul(in[base + (long)j * stride], make_double2(w.x, sg * w.y));
        m += k;
        if (m >= L) m -= L;
    }
    out[o] = s;
}

// z *= FD(k^2/2) / N, FD(e) = 1 / (exp((e - mu) / T) + 1)
// (pca_derivative.cu:784-827, :861-904; 1D / 2D: only the evolved axes)
__global__ void k_fd_weight(double2 *z, long N, int dim, int nx, int ny, int nz, double Lx, double Ly, double Lz,
                            double mu, double T)
{
    const long i = (long)blockIdx.x * GT + threadIdx.x;
    if (i >= N) return;
    int ix, iy = 0, iz = 0;
    if (dim == 1) ix = (int)i;
    else if (dim == 2) { ix = (int)(i / ny); iy = (int)(i % ny); }
    else { ix = (int)(i / ((long)ny * nz)); iy = (int)((i / nz) % ny); iz = (int)(i % nz); }
    const double kx = kval(ix, nx, Lx);
    const double ky = dim >= 2 ? kval(iy, ny, Ly) : 0.0;
    const double kz = dim >= 3 ? kval(iz, nz, Lz) : 0.0;
    const double ek = 0.5 * (kx * kx + ky * ky + kz * kz);
    const double w = 1.0 / (exp((ek - mu) / T) + 1.0) / (double)N;
    z[i] = z[i] * w;
}

// out = F^-1[FD F[in]] of one complex field; in and out may be the same
static int fd_filter(tdw_ctx *c, const double2 *in, double2 *out, double mu, double T)
{
    const long N = (long)c->N;
    const int L[3] = {c->nx, c->ny, c->nz};
    long stride[3];
    if (c->dim == 1) { stride[0] = 1; }
    else if (c->dim == 2) { stride[0] = c->ny; stride[1] = 1; }
    else { stride[0] = (long)c->ny * c->nz; stride[1] = c->nz; stride[2] = 1; }
    double2 *A = c->d_gscr, *B = c->d_gscr + N;
    const double2 *src = in;
    double2 *dst = A;
    for (int d = 0; d < c->dim; d++) {
        hipLaunchKernelGGL(k_dft_axis, gblocks(N), dim3(GT), 0, c->stream, src, dst, N, L[d], stride[d], c->d_tw[d], 0);
        src = dst;
        dst = (dst == A) ? B : A;
    }
    double2 *spec = const_cast<double2 *>(src);
    hipLaunchKernelGGL(k_fd_weight, gblocks(N), dim3(GT), 0, c->stream, spec, N, c->dim, c->nx, c->ny, c->nz,
                       c->Lphys[0], c->Lphys[1], c->Lphys[2], mu, T);
    src = spec;
    dst = (spec == A) ? B : A;
    for (int d = c->dim - 1; d >= 0; d--) {
        double2 *o = (d == 0) ? out : dst;
        hipLaunchKernelGGL(k_dft_axis, gblocks(N), dim3(GT), 0, c->stream, src, o, N, L[d], stride[d], c->d_tw[d], 1);
        src = o;
        dst = (dst == A) ? B : A;
    }
    return (int)hipGetLastError();
}

int grid_filter_c(tdw_ctx *c, const double2 *in, double2 *out, double mu, double T)
{
    TDW_TRY(ensure_grid(c));
    return fd_filter(c, in, out, mu, T);
}

int grid_filter_d(tdw_ctx *c, int n, const double *in, double *out, double mu, double T)
{
    TDW_TRY(ensure_grid(c));
    const long N = (long)c->N;
    double2 *z = c->d_gscr + 2 * N;
    for (int a = 0; a < n; a++) {
        hipLaunchKernelGGL(k_g_real2cplx, gblocks(N), dim3(GT), 0, c->stream, in + (long)a * N, z, N);
        TDW_TRY(fd_filter(c, z, z, mu, T));
        hipLaunchKernelGGL(k_g_re, gblocks(N), dim3(GT), 0, c->stream, z, out + (long)a * N, N, 0);
    }
    return (int)hipGetLastError();
}

// hkf_mode filtering of the mean fields, in place: V_a, V_b (hkf_mode >= 1,
// pca_kernels.cu:733-737), Delta (>= 2, :758-762), and with a varying
// effective mass the vector potentials A_a, A_b (:786-790) and, when the
// caller stands in for the driver, alpha_a, alpha_b (pca.c:1448-1452)
int hkf_fields(tdw_ctx *c, double *pots, bool with_alpha)
{
    if (c->p.hkf_mode < 1) return 0;
    const long N = (long)c->N;
    const double mu = c->p.hkf_mu, T = c->p.hkf_T;
    if (!(T > 0.0)) return TDW_EINVAL;
    const bool vm = varying_mass(c);
    if (vm && with_alpha) TDW_TRY(grid_filter_d(c, 2, pots + 4 * N, pots + 4 * N, mu, T));
    TDW_TRY(grid_filter_d(c, 2, pots, pots, mu, T));
    if (c->p.hkf_mode >= 2) TDW_TRY(grid_filter_c(c, (const double2 *)(pots + 2 * N), (double2 *)(pots + 2 * N), mu, T));
    if (vm) TDW_TRY(grid_filter_d(c, 6, pots + 6 * N, pots + 6 * N, mu, T));
    return 0;
}

// ---------------------------------------------------------------------------
// overlaps <wf1_a | wf2_a> = dV sum (conj(u1) u2 + conj(v1) v2), one block per
// state (pca_kernels.cu:853-911)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(GT)
k_overlap(const double2 *__restrict__ w1, const double2 *__restrict__ w2, int n, long N, double dv,
          double *re, double *im)
{
    __shared__ double red[GT / 64];
    const int a = blockIdx.x;
    const long ou = (long)a * N, ov = ((long)n + a) * N;
    double sr = 0.0, si = 0.0;
    for (long i = threadIdx.x; i < N; i += GT) {
        const double2 p = cmulc(w1[ou + i], w2[ou + i]) + cmulc(w1[ov + i], w2[ov + i]);
        sr += p.x;
        si += p.y;
    }
    sr = block_sum(sr, red);
    si = block_sum(si, red);
    if (threadIdx.x == 0) {
        if (re) re[a] = sr * dv;
        if (im) im[a] = si * dv;
    }
}

int grid_overlap(tdw_ctx *c, int n, const double2 *w1, const double2 *w2, double *re, double *im)
{
    if (n <= 0) return 0;
    double dv = c->p.dx;
    if (c->dim >= 2) dv *= c->p.dy;
    if (c->dim >= 3) dv *= c->p.dz;
    hipLaunchKernelGGL(k_overlap, dim3(n), dim3(GT), 0, c->stream, w1, w2, n, (long)c->N, dv, re, im);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// sum of n doubles into out[0] (local_reductionR, wslda_cuda_utils.h:250-278):
// block partials over a grid-stride loop, then one block over the partials in
// a fixed order; out may be the input array
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(GT) k_gsum_part(const double *__restrict__ a, long n, double *__restrict__ part)
{
    __shared__ double red[GT / 64];
    double s = 0.0;
    for (long i = (long)blockIdx.x * GT + threadIdx.x; i < n; i += (long)gridDim.x * GT) s += a[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ void __launch_bounds__(GT) k_gsum_final(const double *__restrict__ part, int nb, double *out)
{
    __shared__ double red[GT / 64];
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += GT) s += part[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[0] = s;
}

int grid_sum(tdw_ctx *c, const double *a, long n, double *out)
{
    if (n < 0 || !a || !out) return TDW_EINVAL;
    long nb = (n + GT - 1) / GT;
    if (nb > 1024) nb = 1024;
    if (nb < 1) nb = 1;
    TDW_TRY(ensure_part(c, (size_t)nb));
    hipLaunchKernelGGL(k_gsum_part, dim3((unsigned)nb), dim3(GT), 0, c->stream, a, n, c->d_part);
    hipLaunchKernelGGL(k_gsum_final, dim3(1), dim3(GT), 0, c->stream, c->d_part, (int)nb, out);
    return (int)hipGetLastError();
}

}  // namespace tdw
