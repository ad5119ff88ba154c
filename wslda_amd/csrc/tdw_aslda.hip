// tdw_aslda.hip -- the ASLDA Hamiltonian on 3D lattices in two passes.
//
// H of the current-correction path (pca_kernels.cu:233-391 without
// FAST_CONST_EFFECTIVE_MASS_MODE, with CURRENT_CORRECTIONS):
//   (H psi)_u = (V_a - mu_a - i/2 div A_a + lap(alpha_a)/4) u + Delta v
//               + sum_d K_d,   K_d = -(alpha_a d_d^2 u + d_d^2(alpha_a u))/4 - i A_a,d d_d u
//   (H psi)_v = -(V_b - mu_b + i/2 div A_b + lap(alpha_b)/4) v + conj(Delta) u
//               + sum_d K'_d,  K'_d = +(alpha_b d_d^2 v + d_d^2(alpha_b v))/4 - i A_b,d d_d v
// Every K_d needs only spectral derivatives along axis d of psi and of
// alpha psi, so the x terms are formed in an x-pencil pass (blocks of
// pencils, in-place two-factor FFT in LDS) and stored in T, and the y, z
// terms with the pointwise terms in a yz-plane pass (the same FFT) that
// reads psi and T and writes H psi with the <psi|H psi> partials.  This
// replaces alpha*psi, the batched lap(alpha psi) through a scratch array,
// and the H kernel that reads psi, its three gradients, its Laplacian and
// lap(alpha psi): 5 S per state instead of 14 S.  The gradient of psi along
// an axis has its Nyquist mode set to zero, as the reference's derivatives
// do (pca_derivative.cu:169-171).
#include "tdw_fft.hpp"
#include "tdw_fused.hpp"
#include "tdw_physics.hpp"
#include "tdw_regfft.hpp"

#include <type_traits>

namespace tdw {

static constexpr int AT = 256;

// storage of the passes' inputs and outputs: FP64, or FP32 in the
// split-precision step (the transforms in FP32, the sums in FP64)
template <bool SP> using ac_t = typename std::conditional<SP, float2, double2>::type;

// ---------------------------------------------------------------------------
// In-place derivatives along one axis of P pencils held in LDS (L = R1*R2,
// n = n1*R2 + n2, frequency f = k1 + R1*k2 at position k1*R2 + n2 between
// the stages; consecutive lanes take consecutive pencils, so layouts whose
// pencils start one element apart, or a padded row apart, are read without
// LDS bank conflicts).  A receives the second derivative (factor -k2s m^2, the
// inverse transform's 1/L folded into k2s); with GRAD, B (same layout)
// receives the first derivative (factor i k1s m, Nyquist mode zero).  Pencil q starts at
// (q / QD) * QH + (q % QD) * QL, element stride IS.
// ---------------------------------------------------------------------------
template <int L, int P, int QD, int QH, int QL, int IS, int NT, bool GRAD, typename C = double2>
__device__ __forceinline__ void axis_d12(C *A, C *B, const C *tw, int tid, double k1s, double k2s)
{
    typedef decltype(C().x) R;
    constexpr int R1 = rf::fac2<L>::R1, R2 = rf::fac2<L>::R2;
    static_assert(R2 > 1, "axis_d12 needs L = R1 * R2 with R2 > 1");
    // stage 1: R1-point DFTs over n1, twiddle W_L^{n2 k1}
    for (int b = tid; b < P * R2; b += NT) {
        const int n2 = b / P, q = b - n2 * P;
        C *pa = A + (q / QD) * QH + (q % QD) * QL;
        C v[R1];
#pragma unroll
        for (int n1 = 0; n1 < R1; n1++) v[n1] = pa[(n1 * R2 + n2) * IS];
        rf::small_dft<R1, -1, C>::run(v);
#pragma unroll
        for (int k1 = 1; k1 < R1; k1++) v[k1] = cmul(v[k1], tw[(n2 * k1) % L]);
#pragma unroll
        for (int k1 = 0; k1 < R1; k1++) pa[(k1 * R2 + n2) * IS] = v[k1];
    }
    __syncthreads();
    // stage 2: R2-point DFTs over n2, the derivative factors, inverse R2 DFTs
    for (int b = tid; b < P * R1; b += NT) {
        const int k1 = b / P, q = b - k1 * P;
        const int off = (q / QD) * QH + (q % QD) * QL;
        C *pa = A + off;
        C v[R2], g[R2];
#pragma unroll
        for (int n2 = 0; n2 < R2; n2++) v[n2] = pa[(k1 * R2 + n2) * IS];
        rf::small_dft<R2, -1, C>::run(v);
#pragma unroll
        for (int k2 = 0; k2 < R2; k2++) {
            const int f = k1 + R1 * k2;
            const int m = (f < L / 2) ? f : f - L;
            if constexpr (GRAD) g[k2] = (f == L / 2) ? cvt<C>(make_double2(0.0, 0.0)) : imul(v[k2], (R)(k1s * (double)m));
            v[k2] = v[k2] * (R)(-k2s * (double)(m * m));
        }
        rf::small_dft<R2, +1, C>::run(v);
        if constexpr (GRAD) rf::small_dft<R2, +1, C>::run(g);
#pragma unroll
        for (int n2 = 0; n2 < R2; n2++) {
            C wc = tw[(n2 * k1) % L];
            wc.y = -wc.y;
            pa[(k1 * R2 + n2) * IS] = n2 == 0 ? v[n2] : cmul(v[n2], wc);
            if constexpr (GRAD) B[off + (k1 * R2 + n2) * IS] = n2 == 0 ? g[n2] : cmul(g[n2], wc);
        }
    }
    __syncthreads();
    // stage 3: inverse R1-point DFTs over k1
    for (int b = tid; b < P * R2; b += NT) {
        const int n2 = b / P, q = b - n2 * P;
        const int off = (q / QD) * QH + (q % QD) * QL;
        C v[R1];
        C *pa = A + off;
#pragma unroll
        for (int k1 = 0; k1 < R1; k1++) v[k1] = pa[(k1 * R2 + n2) * IS];
        rf::small_dft<R1, +1, C>::run(v);
#pragma unroll
        for (int n1 = 0; n1 < R1; n1++) pa[(n1 * R2 + n2) * IS] = v[n1];
        if constexpr (GRAD) {
            C *pb = B + off;
#pragma unroll
            for (int k1 = 0; k1 < R1; k1++) v[k1] = pb[(k1 * R2 + n2) * IS];
            rf::small_dft<R1, +1, C>::run(v);
#pragma unroll
            for (int n1 = 0; n1 < R1; n1++) pb[(n1 * R2 + n2) * IS] = v[n1];
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// x pass: T = K_x (u) / K'_x (v).  One workgroup per (state component, block
// of PX consecutive x pencils); the block sits in LDS as A[q + i*PX] (pencil
// q, point i), so the loads and stores of consecutive lanes are consecutive
// in memory.  k1s = 2 pi / (Lx * LX), k2s = (2 pi / Lx)^2 / LX.
// ---------------------------------------------------------------------------
template <int LX, int PX, bool SP>
__global__ void __launch_bounds__(AT)
ah_x(const ac_t<SP> *__restrict__ y, ac_t<SP> *__restrict__ T, const double *__restrict__ pots, int n, long N, int S,
     const ac_t<SP> *__restrict__ twx_g, double k1s, double k2s, const double *__restrict__ sc)
{
    typedef ac_t<SP> C;
    typedef decltype(C().x) R;
    constexpr int E = LX * PX;
    constexpr int KM = (E + AT - 1) / AT;
    __shared__ C twx[LX];
    __shared__ C A[E], B[E];
    const int tid = threadIdx.x;
    for (int k = tid; k < LX; k += AT) twx[k] = twx_g[k];
    const int nb = S / PX;
    const long a2 = blockIdx.x / nb;
    const int jb = (int)(blockIdx.x - a2 * nb);
    const bool vcomp = a2 >= n;
    const long base = a2 * N + (long)jb * PX;
    const long fb = (long)jb * PX;
    const double *alpha = pots + (vcomp ? 5 : 4) * N;
    const double *Ax = pots + (vcomp ? 9 : 6) * N;
    const double sg = vcomp ? 0.25 : -0.25;
    const double ys = sc ? sc[vcomp ? a2 - n : a2] : 1.0;
    C X[KM], acc[KM];
    double al[KM];
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int e = tid + k * AT;
        if (e < E) {
            const int i = e / PX, q = e - i * PX;
            const long off = (long)i * S + q;
            X[k] = y[base + off] * (R)ys;
            al[k] = alpha[fb + off];
            A[e] = X[k];
        }
    }
    __syncthreads();
    axis_d12<LX, PX, PX, 0, 1, PX, AT, true, C>(A, B, twx, tid, k1s, k2s);
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int e = tid + k * AT;
        if (e < E) {
            const int i = e / PX, q = e - i * PX;
            acc[k] = A[e] * (R)(sg * al[k]) + imul(B[e], (R)(-Ax[fb + (long)i * S + q]));
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int e = tid + k * AT;
        if (e < E) A[e] = X[k] * (R)al[k];
    }
    __syncthreads();
    axis_d12<LX, PX, PX, 0, 1, PX, AT, false, C>(A, nullptr, twx, tid, k1s, k2s);
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int e = tid + k * AT;
        if (e < E) {
            const int i = e / PX, q = e - i * PX;
            T[base + (long)i * S + q] = acc[k] + A[e] * (R)sg;
        }
    }
}

constexpr int nt_ayz(int E)
{
    return E % 256 == 0 ? 256 : E % 192 == 0 ? 192 : E <= 64 ? 64 : E <= 128 ? 128 : E <= 192 ? 192 : 256;
}

// ---------------------------------------------------------------------------
// yz pass: one workgroup per (state, x plane); both components of the plane
// in LDS (A and B: 2 planes of row stride LZ + 1).  out = H psi,
// part[block] = Re <psi|H psi> of the
// plane.  hx = [lap alpha_a, lap alpha_b, div A_a, div A_b].
// ---------------------------------------------------------------------------
template <int LY, int LZ, bool SP>
__global__ void __launch_bounds__(nt_ayz(LY *LZ), (LY * LZ <= 576) ? 3 : 2)
ah_yz(const ac_t<SP> *__restrict__ y, const ac_t<SP> *__restrict__ T, ac_t<SP> *__restrict__ out,
      const double *__restrict__ pots, const double *__restrict__ hx, int n, long N, int NX,
      const ac_t<SP> *__restrict__ twy_g, const ac_t<SP> *__restrict__ twz_g, double k1y, double k2y, double k1z,
      double k2z, double mu_a, double mu_b, double *__restrict__ part, const double *__restrict__ sc)
{
    typedef ac_t<SP> C;
    typedef decltype(C().x) R;
    constexpr int E = LY * LZ;
    constexpr int EP = LY * (LZ + 1);   // padded plane in LDS: row stride LZ + 1
    constexpr int NT = nt_ayz(E);
    constexpr int KM = (E + NT - 1) / NT;
    __shared__ C twy[LY], twz[LZ];
    __shared__ double red[NT / 64];
    extern __shared__ double2 lds_raw[];
    C *A = (C *)lds_raw, *B = (C *)lds_raw + 2 * EP;
    const int tid = threadIdx.x;
    for (int k = tid; k < LY; k += NT) twy[k] = twy_g[k];
    for (int k = tid; k < LZ; k += NT) twz[k] = twz_g[k];
    const int a = blockIdx.x / NX;
    const int x = blockIdx.x - a * NX;
    const long gu = (long)a * N + (long)x * E, gv = ((long)n + a) * N + (long)x * E;
    const long g0 = (long)x * E;
    const double ys = sc ? sc[a] : 1.0;
    C Xu[KM], Xv[KM], Ou[KM], Ov[KM];
    double al[KM], bl[KM];
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int e = tid + k * NT;
        if (e < E) {
            Xu[k] = y[gu + e] * (R)ys;
            Xv[k] = y[gv + e] * (R)ys;
            Ou[k] = T[gu + e];
            Ov[k] = T[gv + e];
            al[k] = pots[4 * N + g0 + e];
            bl[k] = pots[5 * N + g0 + e];
        }
    }
    // axis 0: z (pencils (c, iy): offset q*LZ, stride 1); axis 1: y (pencils
    // (c, iz): offset (q/LZ)*E + q%LZ, stride LZ)
#pragma unroll
    for (int ax = 0; ax < 2; ax++) {
        const int d = ax == 0 ? 2 : 1;   // lattice axis of the vector potential component
        // psi: d^2 into A, d into B
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KM; k++) {
            const int e = tid + k * NT;
            if (e < E) {
                const int pe = (e / LZ) * (LZ + 1) + e % LZ;
                A[pe] = Xu[k];
                A[EP + pe] = Xv[k];
            }
        }
        __syncthreads();
        if (ax == 0) axis_d12<LZ, 2 * LY, LY, EP, LZ + 1, 1, NT, true, C>(A, B, twz, tid, k1z, k2z);
        else axis_d12<LY, 2 * LZ, LZ, EP, 1, LZ + 1, NT, true, C>(A, B, twy, tid, k1y, k2y);
#pragma unroll
        for (int k = 0; k < KM; k++) {
            const int e = tid + k * NT;
            if (e < E) {
                const int pe = (e / LZ) * (LZ + 1) + e % LZ;
                const double Aa = pots[(6 + d) * N + g0 + e], Ab = pots[(9 + d) * N + g0 + e];
                Ou[k] = Ou[k] + A[pe] * (R)(-0.25 * al[k]) + imul(B[pe], (R)(-Aa));
                Ov[k] = Ov[k] + A[EP + pe] * (R)(0.25 * bl[k]) + imul(B[EP + pe], (R)(-Ab));
            }
        }
        // alpha psi: d^2 into A
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KM; k++) {
            const int e = tid + k * NT;
            if (e < E) {
                const int pe = (e / LZ) * (LZ + 1) + e % LZ;
                A[pe] = Xu[k] * (R)al[k];
                A[EP + pe] = Xv[k] * (R)bl[k];
            }
        }
        __syncthreads();
        if (ax == 0) axis_d12<LZ, 2 * LY, LY, EP, LZ + 1, 1, NT, false, C>(A, nullptr, twz, tid, k1z, k2z);
        else axis_d12<LY, 2 * LZ, LZ, EP, 1, LZ + 1, NT, false, C>(A, nullptr, twy, tid, k1y, k2y);
#pragma unroll
        for (int k = 0; k < KM; k++) {
            const int e = tid + k * NT;
            if (e < E) {
                const int pe = (e / LZ) * (LZ + 1) + e % LZ;
                Ou[k] = Ou[k] + A[pe] * (R)-0.25;
                Ov[k] = Ov[k] + A[EP + pe] * (R)0.25;
            }
        }
    }
    // the pointwise terms, the partner component and <psi|H psi> in FP64
    double ssum = 0.0;
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int e = tid + k * NT;
        if (e < E) {
            const long gi = g0 + e;
            const double la = 0.25 * hx[gi], lb = 0.25 * hx[N + gi];
            const double ca = -0.5 * hx[2 * N + gi], cb = -0.5 * hx[3 * N + gi];
            const double2 D = make_double2(pots[2 * N + 2 * gi], pots[2 * N + 2 * gi + 1]);
            const double2 tu = to64(Xu[k]), tv = to64(Xv[k]);
            double2 u = cmul(tu, make_double2(pots[gi] - mu_a, ca));
            double2 v = make_double2(0.0, 0.0) - cmul(tv, make_double2(pots[N + gi] - mu_b, -1.0 * cb));
            u = u + cmul(tv, D);
            v = v + cmulc(D, tu);
            u = u + tu * la;
            v = v - tv * lb;
            u = u + to64(Ou[k]);
            v = v + to64(Ov[k]);
            out[gu + e] = cvt<C>(u);
            out[gv + e] = cvt<C>(v);
            ssum += cmulc(tu, u).x + cmulc(tv, v).x;
        }
    }
    ssum = block_sum(ssum, red);
    if (tid == 0) part[blockIdx.x] = ssum;
}

// ---------------------------------------------------------------------------
// Densities with gradients without storing the gradients: an x pass and a
// yz pass over the state, each workgroup looping over a chunk of G states
// and summing, per lattice point, the terms of calculate_densities
// (pca_dens.cu:220-267: rho, nu, Re psi* lap psi for tau, Im psi* grad psi
// for j; weights f*1e4 and (1-f)*1e4) that its axis contributes.  Partial
// sums per chunk, then a fixed-order reduction over the chunks.
// ---------------------------------------------------------------------------
// x pass: both components of PX consecutive x pencils in LDS (u at A[0..E),
// v at A[E..2E), pencil q, point i at q + i*PX).  Fields per chunk:
// rho_a, rho_b, Re nu, Im nu, tau_a (x), tau_b (x), j_a,x, j_b,x.
template <int LX, int PX, bool SP>
__global__ void __launch_bounds__(AT, (LX * PX <= 512) ? 3 : 2)
dg_x(const ac_t<SP> *__restrict__ y, const double *__restrict__ fbeta, const double *__restrict__ wgt, int n, long N,
     int S, int G, const ac_t<SP> *__restrict__ twx_g, double k1s, double k2s, double *__restrict__ part)
{
    typedef ac_t<SP> C;
    constexpr int E = LX * PX;
    constexpr int KM = (E + AT - 1) / AT;
    __shared__ C twx[LX];
    __shared__ C A[2 * E], B[2 * E];
    const int tid = threadIdx.x;
    for (int k = tid; k < LX; k += AT) twx[k] = twx_g[k];
    const int nb = S / PX;
    const int ch = blockIdx.x / nb;
    const int jb = blockIdx.x - ch * nb;
    const int a0 = ch * G, a1 = min(n, a0 + G);
    const long fb0 = (long)jb * PX;
    double na[KM], nbv[KM], nux[KM], nuy[KM], ta[KM], tb[KM], ja[KM], jbx[KM];
#pragma unroll
    for (int k = 0; k < KM; k++) na[k] = nbv[k] = nux[k] = nuy[k] = ta[k] = tb[k] = ja[k] = jbx[k] = 0.0;
    C Xu[KM], Xv[KM];
    auto load = [&](int a, C *U, C *V) {
        const long bu = (long)a * N + fb0, bv = ((long)n + a) * N + fb0;
#pragma unroll
        for (int k = 0; k < KM; k++) {
            const int e = tid + k * AT;
            if (e < E) {
                const int i = e / PX, q = e - i * PX;
                U[k] = y[bu + (long)i * S + q];
                V[k] = y[bv + (long)i * S + q];
            }
        }
    };
    for (int a = a0; a < a1; a++) {
        double fb = fbeta[a] * 10000.;
        double fbm = 10000. - fb;
        if (wgt) { fb *= wgt[a]; fbm *= wgt[a]; }
        load(a, Xu, Xv);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KM; k++) {
            const int e = tid + k * AT;
            if (e < E) {
                A[e] = Xu[k];
                A[E + e] = Xv[k];
            }
        }
        __syncthreads();
        axis_d12<LX, 2 * PX, PX, E, 1, PX, AT, true, C>(A, B, twx, tid, k1s, k2s);
#pragma unroll
        for (int k = 0; k < KM; k++) {
            const int e = tid + k * AT;
            if (e < E) {
                const double2 u = to64(Xu[k]), v = to64(Xv[k]);
                na[k] += cnorm(u) * fb;
                nbv[k] += cnorm(v) * fbm;
                const double2 uv = cmul(u, make_double2(v.x, -v.y)) * ((fbm - fb) / 2.0);
                nux[k] += uv.x;
                nuy[k] += uv.y;
                ta[k] += cmulc(u, to64(A[e])).x * fb;
                tb[k] += cmulc(v, to64(A[E + e])).x * fbm;
                ja[k] += cmulc(u, to64(B[e])).y * fb;
                jbx[k] -= cmulc(v, to64(B[E + e])).y * fbm;
            }
        }
    }
    double *p = part + (long)ch * 8 * N;
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int e = tid + k * AT;
        if (e < E) {
            const int i = e / PX, q = e - i * PX;
            const long g = fb0 + (long)i * S + q;
            p[0 * N + g] = na[k];
            p[1 * N + g] = nbv[k];
            p[2 * N + g] = nux[k];
            p[3 * N + g] = nuy[k];
            p[4 * N + g] = ta[k];
            p[5 * N + g] = tb[k];
            p[6 * N + g] = ja[k];
            p[7 * N + g] = jbx[k];
        }
    }
}

// yz pass: one workgroup per (chunk, x plane), both components of the plane
// in LDS.  Fields per chunk: tau_a (y+z), tau_b (y+z), j_a,y, j_a,z, j_b,y,
// j_b,z.
template <int LY, int LZ, bool SP>
__global__ void __launch_bounds__(nt_ayz(LY *LZ), (LY * LZ <= 576) ? 3 : 2)
dg_yz(const ac_t<SP> *__restrict__ y, const double *__restrict__ fbeta, const double *__restrict__ wgt, int n, long N,
      int NX, int G, const ac_t<SP> *__restrict__ twy_g, const ac_t<SP> *__restrict__ twz_g, double k1y, double k2y,
      double k1z, double k2z, double *__restrict__ part)
{
    typedef ac_t<SP> C;
    constexpr int E = LY * LZ;
    constexpr int EP = LY * (LZ + 1);   // padded plane in LDS: row stride LZ + 1
    constexpr int NT = nt_ayz(E);
    constexpr int KM = (E + NT - 1) / NT;
    __shared__ C twy[LY], twz[LZ];
    extern __shared__ double2 lds_raw[];
    C *A = (C *)lds_raw, *B = (C *)lds_raw + 2 * EP;
    const int tid = threadIdx.x;
    for (int k = tid; k < LY; k += NT) twy[k] = twy_g[k];
    for (int k = tid; k < LZ; k += NT) twz[k] = twz_g[k];
    const int ch = blockIdx.x / NX;
    const int x = blockIdx.x - ch * NX;
    const int a0 = ch * G, a1 = min(n, a0 + G);
    const long g0 = (long)x * E;
    double ta[KM], tb[KM], jay[KM], jaz[KM], jby[KM], jbz[KM];
#pragma unroll
    for (int k = 0; k < KM; k++) ta[k] = tb[k] = jay[k] = jaz[k] = jby[k] = jbz[k] = 0.0;
    C Xu[KM], Xv[KM];
    auto load = [&](int a, C *U, C *V) {
        const long gu = (long)a * N + g0, gv = ((long)n + a) * N + g0;
#pragma unroll
        for (int k = 0; k < KM; k++) {
            const int e = tid + k * NT;
            if (e < E) {
                U[k] = y[gu + e];
                V[k] = y[gv + e];
            }
        }
    };
    for (int a = a0; a < a1; a++) {
        double fb = fbeta[a] * 10000.;
        double fbm = 10000. - fb;
        if (wgt) { fb *= wgt[a]; fbm *= wgt[a]; }
        load(a, Xu, Xv);
#pragma unroll
        for (int ax = 0; ax < 2; ax++) {
            __syncthreads();
#pragma unroll
            for (int k = 0; k < KM; k++) {
                const int e = tid + k * NT;
                if (e < E) {
                    const int pe = (e / LZ) * (LZ + 1) + e % LZ;
                    A[pe] = Xu[k];
                    A[EP + pe] = Xv[k];
                }
            }
            __syncthreads();
            if (ax == 0) axis_d12<LZ, 2 * LY, LY, EP, LZ + 1, 1, NT, true, C>(A, B, twz, tid, k1z, k2z);
            else axis_d12<LY, 2 * LZ, LZ, EP, 1, LZ + 1, NT, true, C>(A, B, twy, tid, k1y, k2y);
#pragma unroll
            for (int k = 0; k < KM; k++) {
                const int e = tid + k * NT;
                if (e < E) {
                    const int pe = (e / LZ) * (LZ + 1) + e % LZ;
                    const double2 xu = to64(Xu[k]), xv = to64(Xv[k]);
                    ta[k] += cmulc(xu, to64(A[pe])).x * fb;
                    tb[k] += cmulc(xv, to64(A[EP + pe])).x * fbm;
                    const double gu_ = cmulc(xu, to64(B[pe])).y * fb, gv_ = cmulc(xv, to64(B[EP + pe])).y * fbm;
                    if (ax == 0) { jaz[k] += gu_; jbz[k] -= gv_; }
                    else { jay[k] += gu_; jby[k] -= gv_; }
                }
            }
        }
    }
    double *p = part + (long)ch * 6 * N + g0;
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int e = tid + k * NT;
        if (e < E) {
            p[0 * N + e] = ta[k];
            p[1 * N + e] = tb[k];
            p[2 * N + e] = jay[k];
            p[3 * N + e] = jaz[k];
            p[4 * N + e] = jby[k];
            p[5 * N + e] = jbz[k];
        }
    }
}

// dens (the 12*N block) from the chunk partials of dg_x and dg_yz: one
// thread per (output field, point), each summing its field over the chunks
// in chunk order
__global__ void __launch_bounds__(AT)
dg_final(const double *__restrict__ px, const double *__restrict__ pyz, int nchunk, long N, double tr,
         double *__restrict__ dens)
{
    const long t = (long)blockIdx.x * AT + threadIdx.x;
    if (t >= 12 * N) return;
    const int f = (int)(t / N);
    const long i = t - (long)f * N;
    // source field in the x partials (8 per chunk) and the yz partials (6 per chunk), -1: none
    // output f: 0 Re nu, 1 Im nu, 2 rho_a, 3 tau_a, 4 j_a,x, 5 j_a,y, 6 j_a,z, 7 rho_b, 8 tau_b, 9 j_b,x,
    //           10 j_b,y, 11 j_b,z
    constexpr int sx[12] = {2, 3, 0, 4, 6, -1, -1, 1, 5, 7, -1, -1};
    constexpr int sy[12] = {-1, -1, -1, 0, -1, 2, 3, -1, 1, -1, 4, 5};
    const int fx = sx[f], fy = sy[f];
    double a = 0.0, b = 0.0;
    if (fx >= 0)
        for (int k = 0; k < nchunk; k++) a += px[((long)k * 8 + fx) * N + i];
    if (fy >= 0)
        for (int k = 0; k < nchunk; k++) b += pyz[((long)k * 6 + fy) * N + i];
    const double v = (a + b) / 10000. / tr;
    if (f < 2) dens[2 * i + f] = v;
    else dens[(long)f * N + i] = v;
}

// qpe[a] = dv * sum of the state's NX plane partials
__global__ void ah_qpe(const double *__restrict__ part, int n, int NX, double dv, double *__restrict__ qpe)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    double s = 0.0;
    for (int k = 0; k < NX; k++) s += part[(long)a * NX + k];
    qpe[a] = dv * s;
}

template <bool SP> struct AFns {
    typedef ac_t<SP> C;
    typedef void (*ahx_fn)(const C *, C *, const double *, int, long, int, const C *, double, double, const double *);
    typedef void (*ahyz_fn)(const C *, const C *, C *, const double *, const double *, int, long, int, const C *,
                            const C *, double, double, double, double, double, double, double *, const double *);
    typedef void (*dgx_fn)(const C *, const double *, const double *, int, long, int, int, const C *, double, double,
                           double *);
    typedef void (*dgyz_fn)(const C *, const double *, const double *, int, long, int, int, const C *, const C *,
                            double, double, double, double, double *);

    // x pass for nx points and PX pencils per workgroup (S a multiple of PX)
    static bool ahx(int nx, int S, ahx_fn *f, int *px)
    {
        if (S % 32 == 0) {
            *px = 32;
            switch (nx) {
            case 12: *f = ah_x<12, 32, SP>; return true;
            case 16: *f = ah_x<16, 32, SP>; return true;
            case 24: *f = ah_x<24, 32, SP>; return true;
            case 32: *f = ah_x<32, 32, SP>; return true;
            default: return false;
            }
        }
        if (S % 16 == 0) {
            *px = 16;
            switch (nx) {
            case 12: *f = ah_x<12, 16, SP>; return true;
            case 16: *f = ah_x<16, 16, SP>; return true;
            case 24: *f = ah_x<24, 16, SP>; return true;
            case 32: *f = ah_x<32, 16, SP>; return true;
            default: return false;
            }
        }
        return false;
    }

    static bool ahyz(int ny, int nz, ahyz_fn *f, int *nt)
    {
        if (ny == 12 && nz == 12) { *f = ah_yz<12, 12, SP>; *nt = nt_ayz(144); return true; }
        if (ny == 16 && nz == 16) { *f = ah_yz<16, 16, SP>; *nt = nt_ayz(256); return true; }
        if (ny == 24 && nz == 24) { *f = ah_yz<24, 24, SP>; *nt = nt_ayz(576); return true; }
        if (ny == 32 && nz == 32) { *f = ah_yz<32, 32, SP>; *nt = nt_ayz(1024); return true; }
        return false;
    }

    static bool dgx(int nx, int S, dgx_fn *f, int *px)
    {
        if (S % 32 == 0) {
            *px = 32;
            switch (nx) {
            case 12: *f = dg_x<12, 32, SP>; return true;
            case 16: *f = dg_x<16, 32, SP>; return true;
            case 24: *f = dg_x<24, 16, SP>; *px = 16; return true;   // fewer registers, three waves per SIMD
            case 32: *f = dg_x<32, 16, SP>; *px = 16; return true;   // 2 x 2 x 32 x 32 points would exceed 64 KB
            default: return false;
            }
        }
        if (S % 16 == 0) {
            *px = 16;
            switch (nx) {
            case 12: *f = dg_x<12, 16, SP>; return true;
            case 16: *f = dg_x<16, 16, SP>; return true;
            case 24: *f = dg_x<24, 16, SP>; return true;
            case 32: *f = dg_x<32, 16, SP>; return true;
            default: return false;
            }
        }
        return false;
    }

    static bool dgyz(int ny, int nz, dgyz_fn *f, int *nt)
    {
        if (ny == 12 && nz == 12) { *f = dg_yz<12, 12, SP>; *nt = nt_ayz(144); return true; }
        if (ny == 16 && nz == 16) { *f = dg_yz<16, 16, SP>; *nt = nt_ayz(256); return true; }
        if (ny == 24 && nz == 24) { *f = dg_yz<24, 24, SP>; *nt = nt_ayz(576); return true; }
        if (ny == 32 && nz == 32) { *f = dg_yz<32, 32, SP>; *nt = nt_ayz(1024); return true; }
        return false;
    }

    static const C *tw(const tdw_ctx *c, int ax)
    {
        if constexpr (SP) return c->d_tw32[ax];
        else return c->d_tw[ax];
    }
};

bool dens_grad_fused_supported(const tdw_ctx *c)
{
    if (c->dim != 3) return false;
    AFns<false>::dgx_fn fx;
    AFns<false>::dgyz_fn fy;
    int nt, px;
    if (!AFns<false>::dgx(c->nx, c->S, &fx, &px) || !AFns<false>::dgyz(c->ny, c->nz, &fy, &nt)) return false;
    return sizeof(double2) * 4 * (size_t)c->ny * (c->nz + 1) <= (size_t)c->max_lds;
}

template <bool SP>
static int dens_grad_t(tdw_ctx *c, int n, const ac_t<SP> *wf, const double *fbeta, const double *wgt, double *dens)
{
    typedef AFns<SP> F;
    typename F::dgx_fn fx;
    typename F::dgyz_fn fy;
    int nt, px;
    if (!F::dgx(c->nx, c->S, &fx, &px) || !F::dgyz(c->ny, c->nz, &fy, &nt)) return TDW_ESIZE;
    const long N = (long)c->N;
    const int S = c->S;
    // states per chunk: keep >= ~4096 workgroups in the x pass
    int G = 64;
    while (G > 4 && (long)(S / px) * ((n + G - 1) / G) < 4096) G >>= 1;
    const int nchunk = (n + G - 1) / G;
    TDW_TRY(ensure_part(c, (size_t)nchunk * 14 * N));
    double *px_part = c->d_part, *pyz_part = c->d_part + (size_t)nchunk * 8 * N;
    const double k1 = 2.0 * M_PI / c->Lphys[0] / c->nx, k2 = pow(2.0 * M_PI / c->Lphys[0], 2) / c->nx;
    {
        TDW_TIMED(c, "dens_gx");
        hipLaunchKernelGGL(fx, dim3((unsigned)((long)nchunk * (S / px))), dim3(AT), 0, c->stream, wf, fbeta, wgt, n, N,
                           S, G, F::tw(c, 0), k1, k2, px_part);
    }
    TDW_TIMED(c, "dens_gyz");
    const size_t lds = sizeof(typename F::C) * 4 * (size_t)c->ny * (c->nz + 1);
    const double k1y = 2.0 * M_PI / c->Lphys[1] / c->ny, k2y = pow(2.0 * M_PI / c->Lphys[1], 2) / c->ny;
    const double k1z = 2.0 * M_PI / c->Lphys[2] / c->nz, k2z = pow(2.0 * M_PI / c->Lphys[2], 2) / c->nz;
    hipLaunchKernelGGL(fy, dim3((unsigned)((long)nchunk * c->nx)), dim3(nt), lds, c->stream, wf, fbeta, wgt, n, N,
                       c->nx, G, F::tw(c, 1), F::tw(c, 2), k1y, k2y, k1z, k2z, pyz_part);
    hipLaunchKernelGGL(dg_final, dim3((unsigned)((12 * N + AT - 1) / AT)), dim3(AT), 0, c->stream, px_part, pyz_part,
                       nchunk, N, dens_transverse(c), dens);
    TDW_TRY((int)hipGetLastError());
    return spin_fold(c, dens);
}

int dens_grad_fused(tdw_ctx *c, int n, const double2 *wf, const double *fbeta, const double *wgt, double *dens)
{
    return dens_grad_t<false>(c, n, wf, fbeta, wgt, dens);
}

bool aslda_h_fused_supported(const tdw_ctx *c)
{
    if (c->dim != 3) return false;
    AFns<false>::ahx_fn fx;
    AFns<false>::ahyz_fn fy;
    int nt, px;
    if (!AFns<false>::ahx(c->nx, c->S, &fx, &px) || !AFns<false>::ahyz(c->ny, c->nz, &fy, &nt)) return false;
    return sizeof(double2) * 4 * (size_t)c->ny * (c->nz + 1) <= (size_t)c->max_lds;
}

// out = H in with <H> into qpe (no subtraction); T: 2*n*N scratch of the
// passes' precision; hx: the fields of aslda_fields; part: n*NX doubles; sc
// (n per-state factors, or null): H acts on sc_a in_a
template <bool SP>
static int aslda_h_t(tdw_ctx *c, int n, const ac_t<SP> *in, ac_t<SP> *out, ac_t<SP> *T, const double *pots,
                     const double *hx, double *part, double *d_qpe, const double *sc)
{
    typedef AFns<SP> F;
    typename F::ahx_fn fx;
    typename F::ahyz_fn fy;
    int nt, px;
    if (!F::ahx(c->nx, c->S, &fx, &px) || !F::ahyz(c->ny, c->nz, &fy, &nt)) return TDW_ESIZE;
    const long N = (long)c->N;
    const int S = c->S;
    {
        TDW_TIMED(c, "ah_x");
        const long nblk = 2L * n * (S / px);
        const double k1 = 2.0 * M_PI / c->Lphys[0] / c->nx, k2 = pow(2.0 * M_PI / c->Lphys[0], 2) / c->nx;
        hipLaunchKernelGGL(fx, dim3((unsigned)nblk), dim3(AT), 0, c->stream, in, T, pots, n, N, S, F::tw(c, 0), k1, k2,
                           sc);
    }
    {
        TDW_TIMED(c, "ah_yz");
        const size_t lds = sizeof(typename F::C) * 4 * (size_t)c->ny * (c->nz + 1);
        const double k1y = 2.0 * M_PI / c->Lphys[1] / c->ny, k2y = pow(2.0 * M_PI / c->Lphys[1], 2) / c->ny;
        const double k1z = 2.0 * M_PI / c->Lphys[2] / c->nz, k2z = pow(2.0 * M_PI / c->Lphys[2], 2) / c->nz;
        hipLaunchKernelGGL(fy, dim3((unsigned)((long)n * c->nx)), dim3(nt), lds, c->stream, in, T, out, pots, hx, n, N,
                           c->nx, F::tw(c, 1), F::tw(c, 2), k1y, k2y, k1z, k2z, c->p.mu_a, c->p.mu_b, part, sc);
        hipLaunchKernelGGL(ah_qpe, dim3((n + 255) / 256), dim3(256), 0, c->stream, part, n, c->nx, dvol_norm(c), d_qpe);
    }
    return (int)hipGetLastError();
}

int aslda_h_fused(tdw_ctx *c, int n, const double2 *in, double2 *out, double2 *T, const double *pots,
                  const double *hx, double *part, double *d_qpe, const double *sc)
{
    return aslda_h_t<false>(c, n, in, out, T, pots, hx, part, d_qpe, sc);
}

// ---------------------------------------------------------------------------
// The ASLDA step with its streaming work folded into two passes
// (pca.c:1443-1465 and :1510-1535 in the current-correction form).  The
// modular form spends 21 S per state-step on the predictor (8), two norm +
// scale pairs (3 + 3), the corrector (4) and the history entry's <H>
// subtraction (3).  Here the normalisation and the subtraction are carried as
// per-state scalars into the next pass over the state, as the fused SLDA step
// carries them:
//   a_pred  completes the newest history entry (F1 = f1 - e y_k, y_k = s y,
//           written back), ABM predictor y_p = y_k - i dt sum b_j F_j and the
//           partial corrector c = y_k - i dt sum a_j F_j (into the oldest
//           slot, which it has read), norm partials of y_p;
//   dens    densities of s_p y_p (weights f s_p^2, gradients in the passes);
//   H       the two passes on s_p y_p, <H>_p;
//   a_corr  y_c = -i dt a0 (H - <H>_p s_p y_p) + c, norm partials of y_c;
//   dens, H again on s_c y_c into the newest slot, without the subtraction.
// 9 + 2 + 5 + 4 + 2 + 5 = 27 S per state-step (AB5AM5).  The stored y stays
// unnormalised and f[0] unsubtracted (pending 2) until the next step's a_pred
// or fused_finalize (y *= s, f0 -= e y) applies them.
// ---------------------------------------------------------------------------
static constexpr int PT = 256;

// grid (nchunk, n): block b of state a runs points [b chunk, (b+1) chunk) of
// both components; part[a * nchunk + b] = sum |y_p|^2 over them.  sk, ek
// (pending): y_k = sk_a y and f1 -= ek_a y_k.  The other output goes to the
// oldest history slot (f3 / f4 / f5), read first by the same thread:
//   plain: y <- y_p, slot <- c = y_k - i dt sum a_j F_j (FP64)
//   split: y <- c (FP64), slot <- the FP32 copy of y_p (the history is FP32)
template <int SCH, bool SP>
__global__ void __launch_bounds__(PT)
k_apred(double2 *y, ac_t<SP> *f1, const ac_t<SP> *f2, ac_t<SP> *f3, ac_t<SP> *f4, ac_t<SP> *f5,
        const double *__restrict__ sk, const double *__restrict__ ek, int n, long N, long chunk, double dt,
        double *__restrict__ part)
{
    typedef ac_t<SP> HC;
    __shared__ double red[PT / 64];
    const int a = blockIdx.y;
    const long i0 = (long)blockIdx.x * chunk, i1 = min(N, i0 + chunk);
    const double s = sk ? sk[a] : 1.0, e = sk ? ek[a] : 0.0;
    HC *fold = SCH == 34 ? f3 : SCH == 45 ? f4 : f5;
    double nrm = 0.0;
    for (long i = i0 + threadIdx.x; i < i1; i += PT) {
        double q = 0.0;
#pragma unroll
        for (int comp = 0; comp < 2; comp++) {
            const long idx = (comp ? (long)n + a : (long)a) * N + i;
            const double2 yk = y[idx] * s;
            double2 F1 = to64(f1[idx]);
            if (sk) {
                F1 = F1 - yk * e;
                f1[idx] = cvt<HC>(F1);
            }
            const double2 F2 = to64(f2[idx]), F3 = to64(f3[idx]);
            const double2 F4 = SCH >= 45 ? to64(f4[idx]) : make_double2(0.0, 0.0);
            const double2 F5 = SCH >= 55 ? to64(f5[idx]) : make_double2(0.0, 0.0);
            double2 sp, sc;
            abm_sums<SCH>(F1, F2, F3, F4, F5, sp, sc);
            const double2 yp = yk + imul(sp, -dt);
            if constexpr (SP) {
                y[idx] = yk + imul(sc, -dt);
                fold[idx] = to32(yp);
            } else {
                y[idx] = yp;
                fold[idx] = yk + imul(sc, -dt);
            }
            q += cnorm(yp);
        }
        nrm += q;
    }
    nrm = block_sum(nrm, red);
    if (threadIdx.x == 0) part[(long)a * gridDim.x + blockIdx.x] = nrm;
}

// y_c = -i dta0 (hy - ep_a yh) + c with yh = sp_a y_p (the normalised
// prediction), norm partials of y_c as k_apred's.
//   plain: y holds y_p, cin c;  split: y holds c, cin the FP32 y_p, and the
//   FP32 copy of y_c goes to y32 (hy's buffer: each element of hy is read by
//   the thread that then writes it)
template <bool SP>
__global__ void __launch_bounds__(PT)
k_acorr(double2 *__restrict__ y, const ac_t<SP> *hy, const ac_t<SP> *__restrict__ cin,
        const double *__restrict__ sp, const double *__restrict__ ep, int n, long N, long chunk, double dta0,
        double *__restrict__ part, float2 *y32)
{
    __shared__ double red[PT / 64];
    const int a = blockIdx.y;
    const long i0 = (long)blockIdx.x * chunk, i1 = min(N, i0 + chunk);
    const double s = sp[a], e = ep[a];
    double nrm = 0.0;
    for (long i = i0 + threadIdx.x; i < i1; i += PT) {
        double q = 0.0;
#pragma unroll
        for (int comp = 0; comp < 2; comp++) {
            const long idx = (comp ? (long)n + a : (long)a) * N + i;
            const double2 yh = (SP ? to64(cin[idx]) : y[idx]) * s;
            const double2 cc = SP ? y[idx] : to64(cin[idx]);
            const double2 yc = imul(to64(hy[idx]) - yh * e, -dta0) + cc;
            y[idx] = yc;
            if constexpr (SP) y32[idx] = to32(yc);
            q += cnorm(yc);
        }
        nrm += q;
    }
    nrm = block_sum(nrm, red);
    if (threadIdx.x == 0) part[(long)a * gridDim.x + blockIdx.x] = nrm;
}

// s = 1/sqrt(dv sum_k part[a nchunk + k]), s2 = s^2 (the normalisation of
// pca_kernels.cu:955-976 as a per-state factor)
__global__ void k_ascale(const double *__restrict__ part, int n, int nchunk, double dv, double *__restrict__ s,
                         double *__restrict__ s2)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    double t = 0.0;
    for (int k = 0; k < nchunk; k++) t += part[(long)a * nchunk + k];
    const double sc = 1.0 / sqrt(t * dv);
    s[a] = sc;
    s2[a] = sc * sc;
}

bool aslda_step_supported(const tdw_ctx *c)
{
    return varying_mass(c) && (c->p.precision == TDW_FP64 || c->p.precision == TDW_SPLIT) &&
           aslda_h_fused_supported(c) && dens_grad_fused_supported(c);
}

static int ascale(tdw_ctx *c, int n, int nchunk, double *s, double *s2)
{
    TDW_TIMED(c, "a_norm");
    hipLaunchKernelGGL(k_ascale, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d_part, n, nchunk, dvol_norm(c), s,
                       s2);
    return (int)hipGetLastError();
}

// densities, tau, potentials and H of the state s_a y (H into out, <H> into
// e; T the x pass's scratch)
template <bool SP>
static int aslda_half(tdw_ctx *c, tdw_state *st, int it, const ac_t<SP> *y, const double *s, const double *s2,
                      ac_t<SP> *out, double *e, ac_t<SP> *T)
{
    const int n = st->n;
    TDW_TRY(dens_grad_t<SP>(c, n, y, c->d_fbeta, s2, st->dens));
    TDW_TRY(density_tau(c, st->dens));
    TDW_TRY(step_potentials(c, it, st->dens, st->pots));
    TDW_TRY(hkf_fields(c, st->pots, true));
    const double *hp = h_view(c, it, st->pots);
    if (!hp) return TDW_EINVAL;
    TDW_TRY(aslda_fields(c, hp));
    TDW_TRY(ensure_part(c, (size_t)n * c->nx + n));
    return aslda_h_t<SP>(c, n, y, out, T, hp, c->d_hx, c->d_part, e, s);
}

// FP64 -> FP32 copy (the measurement step's FP64 H psi into the split history)
__global__ void __launch_bounds__(PT) k_ato32(const double2 *__restrict__ src, float2 *__restrict__ dst, long tot)
{
    const long i = (long)blockIdx.x * PT + threadIdx.x;
    if (i < tot) dst[i] = to32(src[i]);
}

template <int SCH, bool SP>
static void launch_apred(tdw_ctx *c, dim3 g, double2 *y, ac_t<SP> **f, const double *sk, const double *ek, int n,
                         long N, long chunk)
{
    hipLaunchKernelGGL((k_apred<SCH, SP>), g, dim3(PT), 0, c->stream, y, f[0], f[1], f[2], f[3], f[4], sk, ek, n, N,
                       chunk, c->p.dt, c->d_part);
}

template <bool SP> static int aslda_step_t(tdw_ctx *c, tdw_state *s, int it, int grad_last)
{
    typedef ac_t<SP> HC;
    if (!s->ddx || c->pending == 1) return TDW_EINVAL;
    // the split form's measurement step uses ddy and ddz as FP64 scratch:
    // refused before anything is written
    if (SP && grad_last && (!s->ddy || !s->ddz)) return TDW_EINVAL;
    const int n = s->n;
    const long N = (long)c->N;
    const int h = c->p.scheme == TDW_AB3AM4 ? 3 : c->p.scheme == TDW_AB4AM5 ? 4 : 5;
    TDW_TRY(fused_state_vecs(c));
    const long chunk = N < 4096 ? N : 4096;
    const int nchunk = (int)((N + chunk - 1) / chunk);
    TDW_TRY(ensure_part(c, (size_t)n * nchunk));
    double2 *y = (double2 *)s->wf;
    HC **f = (HC **)s->f;
    // H of the prediction in lap (FP32 in the split form, then the FP32
    // copy of the corrected state); ddx is the x pass's scratch
    HC *H = (HC *)s->lap, *T = (HC *)s->ddx;
    double *qpe = s->qpe ? s->qpe : c->d_qpe;
    const bool pend = c->pending == 2;
    const dim3 g(nchunk, n);
    // ---- predictor ----
    {
        TDW_TIMED(c, "a_pred");
        const double *sk = pend ? c->d_scale : nullptr, *ek = pend ? c->d_qpe : nullptr;
        if (h == 3) launch_apred<34, SP>(c, g, y, f, sk, ek, n, N, chunk);
        else if (h == 4) launch_apred<45, SP>(c, g, y, f, sk, ek, n, N, chunk);
        else launch_apred<55, SP>(c, g, y, f, sk, ek, n, N, chunk);
        TDW_HIP(hipGetLastError());
    }
    c->pending = 0;
    TDW_TRY(ascale(c, n, nchunk, c->d_sp, c->d_s2));
    // the prediction: y itself (plain) or its FP32 copy in the oldest slot (split)
    const HC *yp = SP ? (const HC *)f[h - 1] : (const HC *)y;
    TDW_TRY(aslda_half<SP>(c, s, it + 1, yp, c->d_sp, c->d_s2, H, c->d_ep, T));
    // ---- corrector ----
    {
        TDW_TIMED(c, "a_corr");
        const double a0 = (c->p.scheme == TDW_AB3AM4) ? 9. / 24. : 251. / 720.;
        TDW_TRY(ensure_part(c, (size_t)n * nchunk));
        hipLaunchKernelGGL(k_acorr<SP>, g, dim3(PT), 0, c->stream, y, H, f[h - 1], c->d_sp, c->d_ep, n, N, chunk,
                           c->p.dt * a0, c->d_part, SP ? (float2 *)H : nullptr);
        TDW_HIP(hipGetLastError());
    }
    TDW_TRY(ascale(c, n, nchunk, c->d_scale, c->d_s2));
    // the corrected state: y (plain) or its FP32 copy in H's buffer (split)
    const HC *yc = SP ? (const HC *)H : (const HC *)y;
    tdw_complex *tmp = s->f[h - 1];
    for (int j = h - 1; j > 0; j--) s->f[j] = s->f[j - 1];
    s->f[0] = tmp;
    if (SP && grad_last) {
        // measurement step of the split form: the densities (tau and j with
        // them) and H psi of the corrected state in FP64, from y itself (ddy
        // holds H psi, ddz the x pass's scratch), H psi then rounded into the
        // FP32 history
        double2 *H64 = (double2 *)s->ddy, *T64 = (double2 *)s->ddz;
        TDW_TRY(aslda_half<false>(c, s, it + 1, y, c->d_scale, c->d_s2, H64, c->d_qpe, T64));
        const long tot = 2L * n * N;
        TDW_TIMED(c, "a_to32");
        hipLaunchKernelGGL(k_ato32, dim3((tot + PT - 1) / PT), dim3(PT), 0, c->stream, H64, (float2 *)s->f[0], tot);
        TDW_HIP(hipGetLastError());
    } else {
        TDW_TRY(aslda_half<SP>(c, s, it + 1, yc, c->d_scale, c->d_s2, (HC *)s->f[0], c->d_qpe, T));
    }
    if (qpe != c->d_qpe) TDW_HIP(hipMemcpyAsync(qpe, c->d_qpe, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    c->pending = 2;
    // a measurement step leaves the arrays as the reference's step does
    return grad_last ? fused_finalize(c, s) : 0;
}

int aslda_fused_step(tdw_ctx *c, tdw_state *s, int it, int grad_last)
{
    if (c->p.precision == TDW_SPLIT) return aslda_step_t<true>(c, s, it, grad_last);
    return aslda_step_t<false>(c, s, it, grad_last);
}

}  // namespace tdw
