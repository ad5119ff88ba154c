// tdw_fft.hip -- spectral Laplacian / gradient passes of the modular path.
//
// Outer pass: pencils along x (stride S = N/nx).  A workgroup takes TP
// consecutive pencils (consecutive inner offsets j, so the loads of one x-row
// are contiguous runs of TP complex values), transforms them in LDS and
// writes T = D_x(in) (and ddx = G_x(in)).
// Inner pass: contiguous blocks of S = ny*nz (3D) or ny (2D) points; per block
// the z- and y-transforms run in LDS and out = T + D_y(in) + D_z(in).
// 1D problems have no inner axis: the outer pass writes the Laplacian.
#include "tdw_fft.hpp"
#include "tdw_regfft.hpp"

namespace tdw {

static constexpr int NTHR = 256;

// --------------------------------------------------------------------------
template <int MODE>
__global__ void __launch_bounds__(NTHR)
k_outer(const double2 *in, double2 *T, double2 *__restrict__ ddx,
        int M, long N, int S, int TP, tdw_axis ax, const double2 *__restrict__ twg, double Lp)
{
    extern __shared__ double2 sm[];
    const int L = ax.L;
    double2 *tw = sm;
    double2 *A = sm + L;
    double2 *B = A + TP * L;
    double2 *D = B + TP * L;
    const int tid = threadIdx.x;
    for (int k = tid; k < L; k += NTHR) tw[k] = twg[k];
    const long P0 = (long)blockIdx.x * TP;
    const long Ptot = (long)M * S;
    const bool contig = (S == 1);
    for (int e = tid; e < TP * L; e += NTHR) {
        int p, i;
        if (contig) { i = e % L; p = e / L; }
        else { p = e % TP; i = e / TP; }
        long P = P0 + p;
        double2 v = make_double2(0.0, 0.0);
        if (P < Ptot) {
            long a = P / S, j = P - a * S;
            v = in[a * N + j + (long)i * S];
        }
        A[p + i * TP] = v;
    }
    __syncthreads();
    double2 *X = lds_fft(A, B, TP, 1, TP, ax, tw, -1, tid, NTHR);
    double2 *Y = (X == A) ? B : A;
    for (int e = tid; e < TP * L; e += NTHR) {
        int i = e / TP;
        double2 v = X[e];
        if (MODE == 1) D[e] = imul(v, kfac1(i, L, Lp));
        X[e] = v * kfac2(i, L, Lp);
    }
    __syncthreads();
    double2 *R = lds_fft(X, Y, TP, 1, TP, ax, tw, +1, tid, NTHR);
    for (int e = tid; e < TP * L; e += NTHR) {
        int p, i;
        if (contig) { i = e % L; p = e / L; }
        else { p = e % TP; i = e / TP; }
        long P = P0 + p;
        if (P < Ptot) {
            long a = P / S, j = P - a * S;
            T[a * N + j + (long)i * S] = R[p + i * TP];
        }
    }
    if (MODE == 1) {
        __syncthreads();
        // A and B are free again: ping-pong D <-> A
        double2 *R2 = lds_fft(D, A, TP, 1, TP, ax, tw, +1, tid, NTHR);
        for (int e = tid; e < TP * L; e += NTHR) {
            int p, i;
            if (contig) { i = e % L; p = e / L; }
            else { p = e % TP; i = e / TP; }
            long P = P0 + p;
            if (P < Ptot) {
                long a = P / S, j = P - a * S;
                ddx[a * N + j + (long)i * S] = R2[p + i * TP];
            }
        }
    }
}

// --------------------------------------------------------------------------
// Inner pass.  NAX = 1 (2D: y) or 2 (3D: z then y).  KM = max elements per
// thread.  For 3D one block = one yz-plane (TB = 1); for 2D TB y-lines.
// in, T and out may alias: each element is read and written by one thread.
template <int MODE, int NAX, int KM>
__global__ void __launch_bounds__(NTHR)
k_inner(const double2 *in, const double2 *T, double2 *out,
        double2 *__restrict__ ddy, double2 *__restrict__ ddz, long nblk, int S, int TB, int ny,
        int nz, tdw_axis axy, tdw_axis axz, const double2 *__restrict__ twy,
        const double2 *__restrict__ twz, double Ly, double Lz)
{
    extern __shared__ double2 sm[];
    const int E = TB * S;
    double2 *ty = sm;
    double2 *tz = ty + ny;
    double2 *A = tz + (NAX == 2 ? nz : 0);
    double2 *B = A + E;
    double2 *D = B + E;
    const int tid = threadIdx.x;
    for (int k = tid; k < ny; k += NTHR) ty[k] = twy[k];
    if (NAX == 2)
        for (int k = tid; k < nz; k += NTHR) tz[k] = twz[k];
    const long off = (long)blockIdx.x * E;
    const long lim = nblk * (long)S;   // total elements
    double2 X[KM], O[KM];
#pragma unroll
    for (int k = 0; k < KM; k++) {
        int e = tid + k * NTHR;
        X[k] = make_double2(0.0, 0.0);
        O[k] = make_double2(0.0, 0.0);
        if (e < E && off + e < lim) {
            X[k] = in[off + e];
            if (T) O[k] = T[off + e];
        }
    }
    // one axis: forward FFT, spectral factor(s), inverse FFT(s).  The axis
    // plan is passed as the kernel argument itself (not a runtime selection
    // between the two), so that its stage tables stay in kernel-argument
    // memory instead of a private copy.
    auto run_axis = [&](const tdw_axis &ax, const double2 *tw, double Lp, int P, int ps, int is, bool isz) {
        const int L = ax.L;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KM; k++) {
            int e = tid + k * NTHR;
            if (e < E) A[e] = X[k];
        }
        __syncthreads();
        double2 *F = lds_fft(A, B, P, ps, is, ax, tw, -1, tid, NTHR);
        double2 *G = (F == A) ? B : A;
        for (int e = tid; e < E; e += NTHR) {
            int i = isz ? (e % nz) : (NAX == 2 ? (e / nz) % ny : e % ny);
            double2 v = F[e];
            if (MODE == 1) D[e] = imul(v, kfac1(i, L, Lp));
            F[e] = v * kfac2(i, L, Lp);
        }
        __syncthreads();
        double2 *R = lds_fft(F, G, P, ps, is, ax, tw, +1, tid, NTHR);
#pragma unroll
        for (int k = 0; k < KM; k++) {
            int e = tid + k * NTHR;
            if (e < E) O[k] = O[k] + R[e];
        }
        if (MODE == 1) {
            __syncthreads();
            double2 *R2 = lds_fft(D, A, P, ps, is, ax, tw, +1, tid, NTHR);
            double2 *dd = isz ? ddz : ddy;
            for (int e = tid; e < E; e += NTHR)
                if (off + e < lim) dd[off + e] = R2[e];
        }
    };
    // axis geometry: 3D: z (contiguous) then y (stride nz); 2D: y (contiguous)
    if (NAX == 2) {
        run_axis(axz, tz, Lz, ny, nz, 1, true);
        run_axis(axy, ty, Ly, nz, 1, nz, false);
    } else {
        run_axis(axy, ty, Ly, TB, ny, 1, false);
    }
#pragma unroll
    for (int k = 0; k < KM; k++) {
        int e = tid + k * NTHR;
        if (e < E && off + e < lim) out[off + e] = O[k];
    }
}

// --------------------------------------------------------------------------
static int pick_tp(const tdw_ctx *c, int mode)
{
    const size_t per = sizeof(double2) * (size_t)c->nx * (mode ? 3 : 2);
    int tp = c->S == 1 ? 8 : 32;
    while (tp > 1 && ((c->S > 1 && (c->S % tp) != 0) ||
                      sizeof(double2) * (size_t)c->nx + per * tp > (size_t)c->max_lds))
        tp >>= 1;
    return tp;
}

int launch_outer(tdw_ctx *c, int m, const double2 *in, double2 *T, double2 *ddx, int mode)
{
    const int L = c->nx, S = c->S, TP = pick_tp(c, mode);
    long Ptot = (long)m * S;
    long nblk = (Ptot + TP - 1) / TP;
    size_t lds = sizeof(double2) * ((size_t)L + (size_t)(mode ? 3 : 2) * TP * L);
    if (lds > (size_t)c->max_lds) return TDW_ESIZE;
    TDW_TIMED(c, mode ? "fft_outer_deriv" : "fft_outer");
    if (mode == 0)
        hipLaunchKernelGGL(k_outer<0>, dim3(nblk), dim3(NTHR), lds, c->stream, in, T, ddx, m,
                           (long)c->N, S, TP, c->ax[0], c->d_tw[0], c->Lphys[0]);
    else
        hipLaunchKernelGGL(k_outer<1>, dim3(nblk), dim3(NTHR), lds, c->stream, in, T, ddx, m,
                           (long)c->N, S, TP, c->ax[0], c->d_tw[0], c->Lphys[0]);
    return (int)hipGetLastError();
}

template <int MODE, int NAX>
static int launch_inner_k(tdw_ctx *c, long nblk_in, int TB, const double2 *in, const double2 *T,
                          double2 *out, double2 *ddy, double2 *ddz)
{
    const int S = c->S, E = TB * S;
    const int km = (E + NTHR - 1) / NTHR;
    size_t lds = sizeof(double2) * ((size_t)c->ny + (NAX == 2 ? c->nz : 0) + (size_t)(MODE ? 3 : 2) * E);
    if (lds > (size_t)c->max_lds) return TDW_ESIZE;
    long grid = (nblk_in + TB - 1) / TB;
    TDW_TIMED(c, MODE ? "fft_inner_deriv" : "fft_inner");
    int nz = NAX == 2 ? c->nz : 1;
    const double2 *twz = NAX == 2 ? c->d_tw[2] : c->d_tw[1];
#define TDW_LI(KMV)                                                                                    \
    hipLaunchKernelGGL((k_inner<MODE, NAX, KMV>), dim3(grid), dim3(NTHR), lds, c->stream, in, T, out,  \
                       ddy, ddz, nblk_in, S, TB, c->ny, nz, c->ax[1], c->ax[NAX == 2 ? 2 : 1],       \
                       c->d_tw[1], twz, c->Lphys[1], c->Lphys[NAX == 2 ? 2 : 1])
    if (km <= 1) TDW_LI(1);
    else if (km <= 2) TDW_LI(2);
    else if (km <= 4) TDW_LI(4);
    else if (km <= 8) TDW_LI(8);
    else if (km <= 16) TDW_LI(16);
    else if (km <= 32) TDW_LI(32);
    else return TDW_ESIZE;
#undef TDW_LI
    return (int)hipGetLastError();
}

int launch_inner(tdw_ctx *c, int m, const double2 *in, const double2 *T, double2 *out,
                 double2 *ddy, double2 *ddz, int mode)
{
    if (c->dim == 1) return TDW_EINVAL;
    long nblk = (long)m * c->nx;   // inner blocks of S points
    if (c->dim == 3) {
        if (mode == 0) return launch_inner_k<0, 2>(c, nblk, 1, in, T, out, ddy, ddz);
        return launch_inner_k<1, 2>(c, nblk, 1, in, T, out, ddy, ddz);
    }
    int TB = 1;
    while (TB * c->S < 256 && TB < 64) TB <<= 1;
    if (mode == 0) return launch_inner_k<0, 1>(c, nblk, TB, in, T, out, ddy, ddz);
    return launch_inner_k<1, 1>(c, nblk, TB, in, T, out, ddy, ddz);
}


// --------------------------------------------------------------------------
// 2D Laplacian in one pass: one workgroup per NX x NY plane held in LDS
// (row stride NY+1), lap = F^-1[-(kx^2 + ky^2)/N F in] with in-place
// two-factor transforms (L = R1*R2, n = n1*R2 + n2, frequency k1 + R1*k2 at
// position k1*R2 + k2 after the forward stages).
// --------------------------------------------------------------------------
template <int L, int P, int PS, int IS>
__device__ __forceinline__ void fwd2(double2 *A, const double2 *tw, int tid)
{
    constexpr int R1 = rf::fac2<L>::R1, R2 = rf::fac2<L>::R2;
    for (int b = tid; b < P * R2; b += NTHR) {
        const int n2 = b / P, q = b - n2 * P;
        double2 *pa = A + q * PS;
        double2 v[R1];
#pragma unroll
        for (int n1 = 0; n1 < R1; n1++) v[n1] = pa[(n1 * R2 + n2) * IS];
        rf::small_dft<R1, -1, double2>::run(v);
#pragma unroll
        for (int k1 = 1; k1 < R1; k1++) v[k1] = cmul(v[k1], tw[(n2 * k1) % L]);
#pragma unroll
        for (int k1 = 0; k1 < R1; k1++) pa[(k1 * R2 + n2) * IS] = v[k1];
    }
    __syncthreads();
    if constexpr (R2 == 1) return;   // the R1-point DFT was the whole transform
    for (int b = tid; b < P * R1; b += NTHR) {
        const int k1 = b / P, q = b - k1 * P;
        double2 *pa = A + q * PS;
        double2 v[R2];
#pragma unroll
        for (int n2 = 0; n2 < R2; n2++) v[n2] = pa[(k1 * R2 + n2) * IS];
        rf::small_dft<R2, -1, double2>::run(v);
#pragma unroll
        for (int k2 = 0; k2 < R2; k2++) pa[(k1 * R2 + k2) * IS] = v[k2];
    }
    __syncthreads();
}

template <int L, int P, int PS, int IS>
__device__ __forceinline__ void inv2(double2 *A, const double2 *tw, int tid)
{
    constexpr int R1 = rf::fac2<L>::R1, R2 = rf::fac2<L>::R2;
    if constexpr (R2 > 1) {
        for (int b = tid; b < P * R1; b += NTHR) {
            const int k1 = b / P, q = b - k1 * P;
            double2 *pa = A + q * PS;
            double2 v[R2];
#pragma unroll
            for (int k2 = 0; k2 < R2; k2++) v[k2] = pa[(k1 * R2 + k2) * IS];
            rf::small_dft<R2, +1, double2>::run(v);
#pragma unroll
            for (int n2 = 0; n2 < R2; n2++) {
                const double2 w = tw[(n2 * k1) % L];
                pa[(k1 * R2 + n2) * IS] = n2 == 0 ? v[n2] : cmul(v[n2], make_double2(w.x, -w.y));
            }
        }
        __syncthreads();
    }
    for (int b = tid; b < P * R2; b += NTHR) {
        const int n2 = b / P, q = b - n2 * P;
        double2 *pa = A + q * PS;
        double2 v[R1];
#pragma unroll
        for (int k1 = 0; k1 < R1; k1++) v[k1] = pa[(k1 * R2 + n2) * IS];
        rf::small_dft<R1, +1, double2>::run(v);
#pragma unroll
        for (int n1 = 0; n1 < R1; n1++) pa[(n1 * R2 + n2) * IS] = v[n1];
    }
    __syncthreads();
}

// frequency index held at position p of an axis after fwd2
template <int L> __device__ __forceinline__ int perm_freq(int p)
{
    constexpr int R1 = rf::fac2<L>::R1, R2 = rf::fac2<L>::R2;
    const int k1 = p / R2, k2 = p - k1 * R2;
    return k1 + R1 * k2;
}

// The 2D Laplacian of the plane held in LDS (row stride NY + 1), in place:
// forward transforms along y then x, -(ksx mx^2 + ksy my^2) at the permuted
// frequencies, inverse along x then y.
template <int NX, int NY>
__device__ __forceinline__ void lap_plane(double2 *A, const double2 *twx, const double2 *twy, int tid, double ksx,
                                          double ksy)
{
    constexpr int NYP = NY + 1;
    fwd2<NY, NX, NYP, 1>(A, twy, tid);
    fwd2<NX, NY, 1, NYP>(A, twx, tid);
    for (int e = tid; e < NX * NY; e += NTHR) {
        const int px = e / NY, py = e - px * NY;
        const int fx = perm_freq<NX>(px), fy = perm_freq<NY>(py);
        const int mx = (fx < NX / 2) ? fx : fx - NX, my = (fy < NY / 2) ? fy : fy - NY;
        A[px * NYP + py] = A[px * NYP + py] * (-(ksx * (double)(mx * mx) + ksy * (double)(my * my)));
    }
    __syncthreads();
    inv2<NX, NY, 1, NYP>(A, twx, tid);
    inv2<NY, NX, NYP, 1>(A, twy, tid);
}

template <int NX, int NY>
__device__ __forceinline__ void load_plane(double2 *A, double2 *twx, double2 *twy, const double2 *src,
                                           const double2 *__restrict__ twx_g, const double2 *__restrict__ twy_g,
                                           int tid)
{
    constexpr int NYP = NY + 1;
    for (int k = tid; k < NX; k += NTHR) twx[k] = twx_g[k];
    for (int k = tid; k < NY; k += NTHR) twy[k] = twy_g[k];
    for (int e = tid; e < NX * NY; e += NTHR) A[(e / NY) * NYP + (e % NY)] = src[e];
    __syncthreads();
}

template <int NX, int NY>
__global__ void __launch_bounds__(NTHR)
k_lap2d(const double2 *in, double2 *out, long N, const double2 *__restrict__ twx_g,
        const double2 *__restrict__ twy_g, double ksx, double ksy)
{
    constexpr int NYP = NY + 1;
    extern __shared__ double2 lds2[];
    __shared__ double2 twx[NX], twy[NY];
    double2 *A = lds2;
    const int tid = threadIdx.x;
    const long r0 = (long)blockIdx.x * N;
    load_plane<NX, NY>(A, twx, twy, in + r0, twx_g, twy_g, tid);
    lap_plane<NX, NY>(A, twx, twy, tid, ksx, ksy);
    for (int e = tid; e < NX * NY; e += NTHR) out[r0 + e] = A[(e / NY) * NYP + (e % NY)];
}

// H applied to one plane (one component of one state) per workgroup, with
// the Laplacian kept in LDS: plane p < n is u_p, p >= n is v_{p-n}; the
// other component is read pointwise.  The pointwise part is that of
// apply_hamiltonian_2d (as in k_apply_h), with kz the state's transverse
// momentum and alpha = 1 for BdG:
//   (H psi)_u =  (V_a - mu_a + alpha_a kz^2/2) u + Delta v  - alpha_a/2 lap u
//   (H psi)_v = -(V_b - mu_b + alpha_b kz^2/2) v + Delta* u + alpha_b/2 lap v
// part[p] = Re sum_i conj(psi_p) (H psi)_p; with dotw, also
// part2[p] = Re sum_i conj(psi_p) dotw_p.  With esub, out = (H - esub_a) psi
// (the <H> subtraction with a known <H>, in the same pass).  out must not
// alias in (the other component is read from in).
template <int NX, int NY>
__global__ void __launch_bounds__(NTHR)
k_h2d(const double2 *__restrict__ in, double2 *__restrict__ out, long N, int n, const double2 *__restrict__ twx_g,
      const double2 *__restrict__ twy_g, double ksx, double ksy, const double *__restrict__ pots,
      const double *__restrict__ kt, int bdg, double mu_a, double mu_b, double *__restrict__ part,
      const double2 *__restrict__ dotw, double *__restrict__ part2, const double *__restrict__ esub)
{
    constexpr int NYP = NY + 1, E = NX * NY, KM = (E + NTHR - 1) / NTHR;
    extern __shared__ double2 lds2[];
    __shared__ double2 twx[NX], twy[NY];
    __shared__ double red[NTHR / 64];
    double2 *A = lds2;
    const int tid = threadIdx.x;
    const int p = blockIdx.x;
    const int comp = p >= n ? 1 : 0;
    const int a = p - comp * n;
    const long own = (long)p * N, oth = (long)(comp ? a : n + a) * N;
    for (int k = tid; k < NX; k += NTHR) twx[k] = twx_g[k];
    for (int k = tid; k < NY; k += NTHR) twy[k] = twy_g[k];
    // the plane stays in registers for the pointwise part; the other
    // component is fetched while the transforms run
    double2 X[KM], Y[KM];
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int e = tid + k * NTHR;
        if (e < E) {
            X[k] = in[own + e];
            A[(e / NY) * NYP + (e % NY)] = X[k];
        }
    }
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int e = tid + k * NTHR;
        if (e < E) Y[k] = in[oth + e];
    }
    __syncthreads();
    lap_plane<NX, NY>(A, twx, twy, tid, ksx, ksy);
    const double k2t = kt ? kt[a] * kt[a] : 0.0;
    const double es = esub ? esub[a] : 0.0;
    double s = 0.0, s2 = 0.0;
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int e = tid + k * NTHR;
        if (e >= E) continue;
        const double2 lp = A[(e / NY) * NYP + (e % NY)];
        const double2 x = X[k], y = Y[k];
        const double2 D = make_double2(pots[2 * N + 2 * e], pots[2 * N + 2 * e + 1]);
        double2 r;
        if (!comp) {
            const double aa = bdg ? -0.5 : pots[4 * N + e] * -0.5;
            const double Va = pots[e] - mu_a + (-1.0 * aa * k2t);
            r = x * Va + cmul(y, D) + lp * aa;
        } else {
            const double ab = bdg ? -0.5 : pots[5 * N + e] * -0.5;
            const double Vb = pots[N + e] - mu_b + (-1.0 * ab * k2t);
            r = x * (-Vb) + cmulc(D, y) - lp * ab;
        }
        s += cmulc(x, r).x;
        if (esub) r = r - x * es;
        out[own + e] = r;
        if (dotw) s2 += cmulc(x, dotw[own + e]).x;
    }
    s = block_sum(s, red);
    if (tid == 0) part[p] = s;
    if (dotw) {
        s2 = block_sum(s2, red);
        if (tid == 0) part2[p] = s2;
    }
}

typedef void (*lap2d_fn)(const double2 *, double2 *, long, const double2 *, const double2 *, double, double);

static bool lap2d_for(const tdw_ctx *c, lap2d_fn *f, size_t *lds)
{
    if (c->dim != 2 || c->nx != c->ny) return false;
    switch (c->nx) {
    case 8: *f = k_lap2d<8, 8>; break;
    case 12: *f = k_lap2d<12, 12>; break;
    case 16: *f = k_lap2d<16, 16>; break;
    case 24: *f = k_lap2d<24, 24>; break;
    case 32: *f = k_lap2d<32, 32>; break;
    case 64: *f = k_lap2d<64, 64>; break;
    default: return false;
    }
    *lds = sizeof(double2) * (size_t)c->nx * (c->ny + 1);
    return *lds <= (size_t)c->max_lds;
}

// 2D Laplacian of m arrays in one pass; *done = 0 if the lattice has no
// one-pass kernel (the caller then runs the outer / inner passes)
int lap2d(tdw_ctx *c, int m, const double2 *in, double2 *out, int *done)
{
    lap2d_fn f;
    size_t lds;
    *done = 0;
    if (!lap2d_for(c, &f, &lds)) return 0;
    *done = 1;
    TDW_TIMED(c, "lap2d");
    const double N = (double)c->N;
    const double ksx = pow(2.0 * M_PI / c->Lphys[0], 2) / N, ksy = pow(2.0 * M_PI / c->Lphys[1], 2) / N;
    hipLaunchKernelGGL(f, dim3(m), dim3(NTHR), lds, c->stream, in, out, (long)c->N, c->d_tw[0], c->d_tw[1], ksx,
                       ksy);
    TDW_HIP(hipGetLastError());
    return 0;
}

typedef void (*h2d_fn)(const double2 *, double2 *, long, int, const double2 *, const double2 *, double, double,
                       const double *, const double *, int, double, double, double *, const double2 *, double *,
                       const double *);

static bool h2d_for(const tdw_ctx *c, h2d_fn *f, size_t *lds)
{
    if (c->dim != 2 || c->nx != c->ny) return false;
    switch (c->nx) {
    case 8: *f = k_h2d<8, 8>; break;
    case 12: *f = k_h2d<12, 12>; break;
    case 16: *f = k_h2d<16, 16>; break;
    case 24: *f = k_h2d<24, 24>; break;
    case 32: *f = k_h2d<32, 32>; break;
    case 64: *f = k_h2d<64, 64>; break;
    default: return false;
    }
    *lds = sizeof(double2) * (size_t)c->nx * (c->ny + 1);
    return *lds <= (size_t)c->max_lds;
}

bool h2d_supported(const tdw_ctx *c)
{
    // the plane kernel applies the constant-effective-mass Hamiltonian
    if (varying_mass(c)) return false;
    h2d_fn f;
    size_t lds;
    return h2d_for(c, &f, &lds);
}

int h2d(tdw_ctx *c, int n, const double2 *in, double2 *out, const double *pots, const double *kt, double *part,
        const double2 *dotw, double *part2, const double *esub)
{
    h2d_fn f;
    size_t lds;
    if (!h2d_for(c, &f, &lds) || in == out) return TDW_EINVAL;
    TDW_TIMED(c, "h2d");
    const double N = (double)c->N;
    const double ksx = pow(2.0 * M_PI / c->Lphys[0], 2) / N, ksy = pow(2.0 * M_PI / c->Lphys[1], 2) / N;
    hipLaunchKernelGGL(f, dim3(2 * n), dim3(NTHR), lds, c->stream, in, out, (long)c->N, n, c->d_tw[0], c->d_tw[1],
                       ksx, ksy, pots, kt, c->p.functional == TDW_BDG, c->p.mu_a, c->p.mu_b, part, dotw, part2,
                       esub);
    TDW_HIP(hipGetLastError());
    return 0;
}

}  // namespace tdw
