// tdw_fused.hip -- fused PECE step for 3D lattices (kernel families a + b).
//
// The reference step (pca.c:1401-1570) touches the state arrays ~57 times per
// step.  Here one step is six passes, none of which stores a derivative:
//   f_pred  (x-pencils, one lane per pencil, FFT in registers): completes the
//           newest history entry (its x kinetic part, from the lane's pencil
//           of y in Fourier space, and the pending <H> subtraction), applies
//           the pending normalisation of y, ABM predictor
//           (tdwslda_ode_integrator.h:26-135), norm partials and the x part
//           of <psi_p|H psi_p> (Parseval).
//   dens    densities of the prediction, normalisation folded into the
//           occupation weights (f*s^2).
//   f_hyz   (yz-planes, FFTs in LDS): D_y + D_z, then H psi without its x
//           kinetic part (pca_kernels.cu:233-391) and the <psi|H psi>
//           partials.
//   f_corr  (x-pencils): the x kinetic part and the <H> subtraction applied
//           to the prediction's pencil in Fourier space, AM corrector
//           y = c - i dt a0 (H psi - <H> psi) (:138-220), norm partials and
//           the x part of <psi|H psi> of the new y.
//   dens, f_hyz again: H of the corrected state, without its x kinetic part,
//   into the newest history slot.
// Normalisation (pca_kernels.cu:955-976) and the <H> subtraction
// (pca_kernels.cu:806-832) are carried as per-state scalars (s = 1/sqrt(norm),
// e = <H>) into the next pass instead of extra passes over the state: the
// stored y stays unnormalised and the newest history entry lacks its x
// kinetic part and the subtraction until the next predictor (or
// tdw_finalize) adds them.
//
// Split precision: the history and H are FP32; y and every ABM update of it
// stay FP64.  The predictor writes the FP64 partial corrector c into y and
// an FP32 copy of the prediction into the history slot it consumed; the
// densities and the yz pass read that copy.  The corrector reads c and the
// copy, writes the FP64 y and an FP32 copy of it (into H's buffer, which it
// has consumed) for the second density and yz passes.
#include "tdw_fft.hpp"
#include "tdw_fused.hpp"
#include "tdw_physics.hpp"
#include "tdw_regfft.hpp"

#include <algorithm>
#include <type_traits>

namespace tdw {

static constexpr int FT = 256;

// Storage type of the history entries, the partial corrector and the T / H
// scratch: FP64 in the plain form; FP32 in the split-precision form, where
// the state y stays FP64 and every update of it is done in FP64.
template <bool SP> using hc_t = typename std::conditional<SP, float2, double2>::type;

template <int LX, typename C>
__device__ __forceinline__ void dxx_inplace(C (&r)[LX], double kscale)
{
    // D_x = F^-1 (-k^2/LX) F ; kscale = (2 pi / Lx)^2 / LX
    typedef decltype(r[0].x) R;
    rf::fft<LX, -1, C>(r);
#pragma unroll
    for (int i = 0; i < LX; i++) {
        const int m = (i < LX / 2) ? i : i - LX;
        r[i] = r[i] * (R)(-kscale * (double)(m * m));
    }
    rf::fft<LX, +1, C>(r);
}

// wave: S % 64 == 0, so every wave lies inside one component and Ptot % 64 == 0
__device__ __forceinline__ void put_norm_part(double nrm, bool wave, long P, long Ptot, double *part)
{
    if (wave) {
        nrm = wave_sum(nrm);
        if ((threadIdx.x & 63) == 0 && P < Ptot) part[P >> 6] = nrm;
    } else if (P < Ptot) {
        part[P] = nrm;
    }
}

// the x part of the kinetic term, sigma (-1/2 D_x) (sigma = +1 on u rows,
// -1 on v rows), applied in place to a pencil held in registers:
// F^-1 (sigma kscale m^2 / 2 + add) F, with add = -<H> / LX when the <H>
// subtraction rides along
template <int LX, typename C>
__device__ __forceinline__ void xkin_inplace(C (&r)[LX], double sig_kscale_half, double add)
{
    typedef decltype(r[0].x) R;
    rf::fft<LX, -1, C>(r);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int m = 0; m < LX; m++) {
        const int mm = (m < LX / 2) ? m : m - LX;
        r[m] = r[m] * (R)fma(sig_kscale_half, (double)(mm * mm), add);
    }
    __builtin_amdgcn_sched_barrier(0);
    rf::fft<LX, +1, C>(r);
}

// sum_x conj(y) sigma (-1/2 D_x) y of a pencil by Parseval, sigma kscale / 2
// sum_m m^2 |Y_m|^2 (r is transformed in place)
template <int LX, typename C>
__device__ __forceinline__ double xkin_energy(C (&r)[LX], double sig_kscale_half)
{
    rf::fft<LX, -1, C>(r);
    double x = 0.0;
#pragma unroll
    for (int m = 0; m < LX; m++) {
        const int mm = (m < LX / 2) ? m : m - LX;
        x += (double)(mm * mm) * cnorm(to64(r[m]));
    }
    return x * sig_kscale_half;
}

// ---------------------------------------------------------------------------
// predictor x-pass
//   plain: y <- y_p (FP64), fold <- c = y_k - i dt sum_j a_j f_j (FP64)
//   split: y <- c (FP64),   fold <- y_p (FP32 copy)
// With pending (the previous fused step's state): y_k = s y, and the newest
// history entry f1 holds H psi_k without the x part of its kinetic term and
// without the <H> subtraction; both are added here,
//   F1 = f1 + s sigma (-1/2 D_x) y - e y_k,
// the D_x from the lane's pencil of y (rounded to FP32 in the split form, as
// the operator is), and F1 is stored back.
// No D_x of the prediction is stored: the corrector applies the x part of
// the kinetic term itself, and the x part of <psi_p|H psi_p> comes from here
// by Parseval, unscaled by the state's normalisation (xpart, laid out as the
// norm partials).
// ---------------------------------------------------------------------------
template <int LX, int SCH, bool SP>
__global__ void __launch_bounds__(FT)
f_pred(double2 *__restrict__ y, hc_t<SP> *f1, const hc_t<SP> *f2, const hc_t<SP> *f3, const hc_t<SP> *f4,
       const hc_t<SP> *f5, hc_t<SP> *fold, double *__restrict__ xpart,
       const double *__restrict__ sk, const double *__restrict__ ek, int pending, int n, long N, int S, double dt,
       double kscale, int wave, double *__restrict__ part)
{
    typedef hc_t<SP> HC;
    const long P = (long)blockIdx.x * FT + threadIdx.x;
    const long Ptot = 2L * n * S;
    const bool act = P < Ptot;
    double nrm = 0.0, xk = 0.0;
    if (act) {
        const long a2 = P / S, j = P - a2 * S;
        const int a = (int)(a2 < n ? a2 : a2 - n);
        const double s = pending ? sk[a] : 1.0;
        const double e = pending ? ek[a] : 0.0;
        const double skh = (a2 < n ? 0.5 : -0.5) * kscale;
        const long base = a2 * N + j;
        HC r[LX];
        if (pending) {
            // the pencil of y (read again in the loop below, from cache)
#pragma unroll
            for (int i = 0; i < LX; i++) r[i] = cvt<HC>(y[base + (long)i * S]);
            xkin_inplace<LX>(r, skh, 0.0);
            __builtin_amdgcn_sched_barrier(0);
        }
        // the loop's addresses are formed here, not held across the transforms
        long lb = base;
        int ls = S;
        asm volatile("" : "+v"(lb), "+s"(ls));
#pragma unroll
        for (int i = 0; i < LX; i++) {
            const long idx = lb + (long)i * ls;
            const double2 yk = y[idx] * s;
            double2 F1 = to64(f1[idx]);
            if (pending) {
                F1 = F1 + to64(r[i]) * s - yk * e;
                f1[idx] = cvt<HC>(F1);
            }
            const double2 F2 = to64(f2[idx]), F3 = to64(f3[idx]);
            const double2 F4 = (SCH >= 45) ? to64(f4[idx]) : make_double2(0.0, 0.0);
            const double2 F5 = (SCH >= 55) ? to64(f5[idx]) : make_double2(0.0, 0.0);
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
            r[i] = cvt<HC>(yp);
            nrm += cnorm(yp);
            if (i % 4 == 3) __builtin_amdgcn_sched_barrier(0);
        }
        xk = xkin_energy<LX>(r, skh);
    }
    put_norm_part(nrm, wave != 0, P, Ptot, part);
    put_norm_part(xk, wave != 0, P, Ptot, xpart);
}

// ---------------------------------------------------------------------------
// corrector x-pass: y = -i dt a0 (H - <H>) psi_p + c
//   plain: y holds y_p, cin c;  split: y holds c, cin the FP32 y_p, and the
//   FP32 copy of the new y goes to y32 (H's buffer: each element of hin is
//   read by the thread that then writes it)
// hin holds H psi_p without the x part of the kinetic term (the yz pass of
// the prediction reads no D_x); that part and the <H> subtraction act on the
// lane's pencil of y_p at once: (-/+ 1/2 D_x - <H>) y_p (u / v rows) is one
// multiplier, sigma kscale m^2 / 2 - <H> / LX, between a forward and an
// inverse transform.  Then (H - <H>) psi_p = hin + s (that pencil).  The x
// part of <psi|H psi> of the new y goes to xpart by Parseval (the history
// refresh's yz pass reads no D_x either; the next predictor completes the
// entry).
// ---------------------------------------------------------------------------
template <int LX, bool SP>
__global__ void __launch_bounds__(FT)
f_corr(double2 *__restrict__ y, const hc_t<SP> *hin, const hc_t<SP> *__restrict__ cin,
       double *__restrict__ xpart, const double *__restrict__ sp_, const double *__restrict__ ep_, int n, long N,
       int S, double dta0, double kscale, int wave, double *__restrict__ part, float2 *y32)
{
    typedef hc_t<SP> HC;
    const long P = (long)blockIdx.x * FT + threadIdx.x;
    const long Ptot = 2L * n * S;
    const bool act = P < Ptot;
    double nrm = 0.0, xk = 0.0;
    if (act) {
        const long a2 = P / S, j = P - a2 * S;
        const int a = (int)(a2 < n ? a2 : a2 - n);
        const double s = sp_[a], e = ep_[a];
        const double skh = (a2 < n ? 0.5 : -0.5) * kscale;
        const long base = a2 * N + j;
        HC r[LX];
#pragma unroll
        for (int i = 0; i < LX; i++) {
            if constexpr (SP) r[i] = cin[base + (long)i * S];   // the FP32 copy of y_p
            else r[i] = y[base + (long)i * S];                  // y_p
        }
        xkin_inplace<LX>(r, skh, -e / LX);
        // keep the loop's loads out of the transforms, and form the loop's
        // addresses anew rather than holding the preload's across them
        // (registers)
        __builtin_amdgcn_sched_barrier(0);
        long lb = base;
        int ls = S;
        asm volatile("" : "+v"(lb), "+s"(ls));
#pragma unroll
        for (int i = 0; i < LX; i++) {
            const long idx = lb + (long)i * ls;
            const double2 cc = SP ? y[idx] : to64(cin[idx]);
            const double2 fp = to64(hin[idx]) + to64(r[i]) * s;   // (H - <H>) psi_p
            const double2 yn = imul(fp, -dta0) + cc;
            y[idx] = yn;
            r[i] = cvt<HC>(yn);
            if constexpr (SP) y32[idx] = to32(yn);
            nrm += cnorm(yn);
            // bound how far the scheduler lifts the next points' loads
            // (registers: r[] is live across the loop)
            if (i % 4 == 3) __builtin_amdgcn_sched_barrier(0);
        }
        xk = xkin_energy<LX>(r, skh);
    }
    put_norm_part(nrm, wave != 0, P, Ptot, part);
    put_norm_part(xk, wave != 0, P, Ptot, xpart);
}

// s = 1/sqrt(norm * dV), s2 = s^2 per state (and norm * dV itself into nrm
// when given); part holds `per` entries per component
__global__ void f_norm_final(const double *__restrict__ part, int n, int per, double dv, double *__restrict__ s,
                             double *__restrict__ s2, double *__restrict__ nrm_out)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    double su = 0.0, sv = 0.0;
    const double *pu = part + (long)a * per, *pv = part + ((long)n + a) * per;
    for (int k = 0; k < per; k++) su += pu[k];
    for (int k = 0; k < per; k++) sv += pv[k];
    const double nrm = (su + sv) * dv;
    const double sc = 1.0 / sqrt(nrm);
    s[a] = sc;
    s2[a] = sc * sc;
    if (nrm_out) nrm_out[a] = nrm;
}

// ---------------------------------------------------------------------------
// yz-plane pass: lap = s (T + D_y y + D_z y); H psi; <psi|H psi> partials
// ---------------------------------------------------------------------------
// threads of the yz-plane pass for a plane of E points: a multiple of 64
// that leaves few idle lanes when each thread owns E/NT points
constexpr int nt_hyz(int E)
{
    return E % 256 == 0 ? 256 : E % 192 == 0 ? 192 : E <= 64 ? 64 : E <= 128 ? 128 : E <= 192 ? 192 : 256;
}

// In-place transforms along one axis of P pencils held in LDS, for
// L = R1*R2: with n = n1*R2 + n2 and f = k1 + R1*k2,
//   X[f] = sum_n2 W_R2^{n2 k2} W_L^{n2 k1} sum_n1 x[n1 R2 + n2] W_R1^{n1 k1},
// stage 1 (per n2): R1-point DFT over n1, twiddle W_L^{n2 k1}, stored at k1*R2 + n2;
// stage 2 (per k1): R2-point DFT over n2 -> X[k1 + R1 k2], times mul(m(f), q)
//   for pencil q, inverse R2-point DFT and the conjugate twiddle, stored at
//   k1*R2 + n2;
// stage 3 (per n2): inverse R1-point DFT over k1, stored at n1*R2 + n2.
// MODE 0 runs all three (with mul = -ks m^2 the second derivative); MODE 1
// stops after the forward R2-point DFT and leaves X[k1 + R1 k2] at position
// k1*R2 + k2; MODE 2 starts from that layout and runs the inverse half.
// Every stage reads and writes the same positions, so one buffer suffices.
// Consecutive lanes take consecutive pencils: with pencils one element or one
// padded row apart, a wave's LDS accesses fall in distinct banks.

// frequency index m(f) in (-L/2, L/2] of the DFT output at position p (MODE 1 layout)
template <int L> __host__ __device__ constexpr int axis_freq(int p)
{
    constexpr int R1 = rf::fac2<L>::R1, R2 = rf::fac2<L>::R2;
    const int f = R2 > 1 ? p / R2 + R1 * (p % R2) : p;
    return f < L / 2 ? f : f - L;
}

struct KsMul {
    double ks;
    __device__ double operator()(int m, int) const { return -ks * (double)(m * m); }
};

template <int L, int P, int QD, int QH, int QL, int IS, int NT, int MODE, typename C, typename MUL>
__device__ __forceinline__ void dft_axis(C *A, const C *tw, int tid, MUL mul)
{
    typedef decltype(C().x) Rl;
    constexpr int R1 = rf::fac2<L>::R1, R2 = rf::fac2<L>::R2;
    // stage 1
    if constexpr (R2 > 1 && MODE != 2) {
        for (int b = tid; b < P * R2; b += NT) {
            const int n2 = b / P, q = b - n2 * P;
            C *pa = A + (q / QD) * QH + (q % QD) * QL;
            C v[R1];
#pragma unroll
            for (int n1 = 0; n1 < R1; n1++) v[n1] = pa[(n1 * R2 + n2) * IS];
            rf::dft_inplace<R1, -1, C>(v);
#pragma unroll
            for (int k1 = 1; k1 < R1; k1++) v[k1] = cmul(v[k1], tw[(n2 * k1) % L]);
#pragma unroll
            for (int k1 = 0; k1 < R1; k1++) pa[(k1 * R2 + n2) * IS] = v[k1];
        }
        __syncthreads();
    }
    // stage 2 (for R2 == 1 the whole transform)
    for (int b = tid; b < P * R1; b += NT) {
        const int k1 = b / P, q = b - k1 * P;
        C *pa = A + (q / QD) * QH + (q % QD) * QL;
        if constexpr (R2 > 1) {
            C v[R2];
#pragma unroll
            for (int n2 = 0; n2 < R2; n2++) v[n2] = pa[(k1 * R2 + n2) * IS];
            if constexpr (MODE != 2) {
                rf::dft_inplace<R2, -1, C>(v);
                if constexpr (MODE == 1) {
#pragma unroll
                    for (int k2 = 0; k2 < R2; k2++) pa[(k1 * R2 + k2) * IS] = v[k2];
                    continue;
                }
#pragma unroll
                for (int k2 = 0; k2 < R2; k2++) {
                    const int f = k1 + R1 * k2;
                    const int m = (f < L / 2) ? f : f - L;
                    v[k2] = v[k2] * (Rl)mul(m, q);
                }
            }
            rf::dft_inplace<R2, +1, C>(v);
#pragma unroll
            for (int n2 = 0; n2 < R2; n2++) {
                C w = tw[(n2 * k1) % L];
                w.y = -w.y;
                pa[(k1 * R2 + n2) * IS] = n2 == 0 ? v[n2] : cmul(v[n2], w);
            }
        } else if (k1 == 0) {
            // R2 == 1: one group per pencil holds the whole pencil
            C v[R1];
#pragma unroll
            for (int n = 0; n < R1; n++) v[n] = pa[n * IS];
            if constexpr (MODE != 2) {
                rf::dft_inplace<R1, -1, C>(v);
                if constexpr (MODE == 0) {
#pragma unroll
                    for (int f = 0; f < R1; f++) {
                        const int m = (f < L / 2) ? f : f - L;
                        v[f] = v[f] * (Rl)mul(m, q);
                    }
                }
            }
            if constexpr (MODE != 1) rf::dft_inplace<R1, +1, C>(v);
#pragma unroll
            for (int n = 0; n < R1; n++) pa[n * IS] = v[n];
        }
    }
    __syncthreads();
    // stage 3
    if constexpr (R2 > 1 && MODE != 1) {
        for (int b = tid; b < P * R2; b += NT) {
            const int n2 = b / P, q = b - n2 * P;
            C *pa = A + (q / QD) * QH + (q % QD) * QL;
            C v[R1];
#pragma unroll
            for (int k1 = 0; k1 < R1; k1++) v[k1] = pa[(k1 * R2 + n2) * IS];
            rf::dft_inplace<R1, +1, C>(v);
#pragma unroll
            for (int n1 = 0; n1 < R1; n1++) pa[(n1 * R2 + n2) * IS] = v[n1];
        }
        __syncthreads();
    }
}

// second derivative along one axis, in place (ks = (2 pi / L_phys)^2 / L)
template <int L, int P, int QD, int QH, int QL, int IS, int NT, typename C>
__device__ __forceinline__ void lap_axis(C *A, const C *tw, int tid, double ks)
{
    dft_axis<L, P, QD, QH, QL, IS, NT, 0>(A, tw, tid, KsMul{ks});
}

struct HParams {
    double mu_a, mu_b;
    int bdg;
};

// TX: add T (D_x of the plane, from an x-pass) to the transverse terms (H
// of arbitrary vectors, fused_apply_h); without it the output lacks the x
// part of the kinetic term (the fused step's yz passes: its x-passes add
// that part themselves)
template <int LY, int LZ, bool SP, bool TX, int NT = nt_hyz(LY * LZ)>
__global__ void __launch_bounds__(NT)
f_hyz(const hc_t<SP> *__restrict__ y, const hc_t<SP> *__restrict__ T, hc_t<SP> *__restrict__ hout,
      const double *__restrict__ pots, const double *__restrict__ sc, int n, long N, int NX,
      const hc_t<SP> *__restrict__ twy_g, const hc_t<SP> *__restrict__ twz_g, double kys, double kzs, HParams hp,
      double *__restrict__ part)
{
    typedef hc_t<SP> C;
    typedef decltype(C().x) R;
    constexpr int E = LY * LZ;              // points of one plane
    constexpr int EP = LY * (LZ + 1);       // the plane in LDS, row stride LZ + 1
    constexpr int KM = (E + NT - 1) / NT;   // owned points per thread
    __shared__ C twy[LY], twz[LZ];
    __shared__ double red[NT / 64];
    extern __shared__ double2 lds_raw[];
    C *A = (C *)lds_raw;
    const int tid = threadIdx.x;
    for (int k = tid; k < LY; k += NT) twy[k] = twy_g[k];
    for (int k = tid; k < LZ; k += NT) twz[k] = twz_g[k];
    const int a = blockIdx.x / NX;
    const int x = blockIdx.x - a * NX;
    const long gu = (long)a * N + (long)x * E, gv = ((long)n + a) * N + (long)x * E;
    C Xu[KM], Xv[KM], Ou[KM], Ov[KM];
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int e = tid + k * NT;
        if (e < E) {
            Xu[k] = y[gu + e];
            Xv[k] = y[gv + e];
            if constexpr (TX) {
                Ou[k] = T[gu + e];
                Ov[k] = T[gv + e];
            } else {
                Ou[k] = cvt<C>(make_double2(0.0, 0.0));
                Ov[k] = cvt<C>(make_double2(0.0, 0.0));
            }
        }
    }
    // z axis: pencils (c, iy), offset c*EP + iy*(LZ+1), stride 1; y axis:
    // pencils (c, iz), offset c*EP + iz, stride LZ+1
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
        // second derivative along the axis, in place
        if (ax == 0) lap_axis<LZ, 2 * LY, LY, EP, LZ + 1, 1, NT, C>(A, twz, tid, kzs);
        else lap_axis<LY, 2 * LZ, LZ, EP, 1, LZ + 1, NT, C>(A, twy, tid, kys);
        const C *Rr = A;
#pragma unroll
        for (int k = 0; k < KM; k++) {
            const int e = tid + k * NT;
            if (e < E) {
                const int pe = (e / LZ) * (LZ + 1) + e % LZ;
                Ou[k] = Ou[k] + Rr[pe];
                Ov[k] = Ov[k] + Rr[EP + pe];
            }
        }
    }
    // H psi on owned points (pointwise terms and the sums in FP64); sc null:
    // the input is the state itself (scale 1)
    const double s = sc ? sc[a] : 1.0;
    double ssum = 0.0;
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int e = tid + k * NT;
        if (e < E) {
            const long gi = (long)x * E + e;
            const double2 pu = to64(Xu[k]) * s, pv = to64(Xv[k]) * s;
            const double2 lu = to64(Ou[k]) * s, lv = to64(Ov[k]) * s;
            const double aa = hp.bdg ? -0.5 : pots[4 * N + gi] * -0.5;
            const double ab = hp.bdg ? -0.5 : pots[5 * N + gi] * -0.5;
            const double Va = pots[gi] - hp.mu_a, Vb = pots[N + gi] - hp.mu_b;
            const double2 D = make_double2(pots[2 * N + 2 * gi], pots[2 * N + 2 * gi + 1]);
            const double2 hu = pu * Va + cmul(pv, D) + lu * aa;
            const double2 hv = pv * (-Vb) + cmulc(D, pu) - lv * ab;
            hout[gu + e] = cvt<C>(hu);
            hout[gv + e] = cvt<C>(hv);
            ssum += cmulc(pu, hu).x + cmulc(pv, hv).x;
        }
    }
    ssum = block_sum(ssum, red);
    if (tid == 0) part[blockIdx.x] = ssum;
}

// multiplier of the y pass of a 2D transform: -(k_y^2 + k_z^2) / (LY LZ),
// with the z part per position of the z-transformed row in zt
struct YZMul {
    double ky;
    const double *zt;
    int lz;
    __device__ double operator()(int m, int q) const { return -(ky * (double)(m * m) + zt[q % lz]); }
};

// yz pass with D_y + D_z as one 2D transform in place (forward along z;
// forward, scale and inverse along y; inverse along z), NC components in LDS
// at a time.  Only the plane values stay in registers (T is read when H psi
// is formed), so more workgroups share a CU than with the per-axis form
// above: planes above 1024 points (the 32 x 96 tube section, where that form
// fits one workgroup per CU) and the FP64 24^2 and 32^2 planes.
template <int LY, int LZ, bool SP, bool TX, int NC, int NT>
__global__ void __launch_bounds__(NT)
f_hyz_big(const hc_t<SP> *__restrict__ y, const hc_t<SP> *__restrict__ T, hc_t<SP> *__restrict__ hout,
          const double *__restrict__ pots, const double *__restrict__ sc, int n, long N, int NX,
          const hc_t<SP> *__restrict__ twy_g, const hc_t<SP> *__restrict__ twz_g, double kys, double kzs,
          HParams hp, double *__restrict__ part)
{
    typedef hc_t<SP> C;
    constexpr int E = LY * LZ;              // points of one plane
    constexpr int EP = LY * (LZ + 1);       // one component's plane in LDS, row stride LZ + 1
    constexpr int KM = E / NT;              // owned points per thread
    static_assert(E % NT == 0 && (NC == 1 || NC == 2), "plane split");
    __shared__ C twy[LY], twz[LZ];
    __shared__ double zt[LZ];
    __shared__ double red[NT / 64];
    extern __shared__ double2 lds_raw[];
    C *A = (C *)lds_raw;
    const int tid = threadIdx.x;
    for (int k = tid; k < LY; k += NT) twy[k] = twy_g[k];
    for (int k = tid; k < LZ; k += NT) twz[k] = twz_g[k];
    // k_z^2 / (LY LZ) at each position of a z-transformed row
    for (int k = tid; k < LZ; k += NT) {
        const int m = axis_freq<LZ>(k);
        zt[k] = kzs / LY * (double)(m * m);
    }
    const int a = blockIdx.x / NX;
    const int x = blockIdx.x - a * NX;
    // workgroup-uniform bases (scalar registers) and 32-bit offsets e
    const long gu = (long)a * N + (long)x * E, gv = ((long)n + a) * N + (long)x * E;
    const C *yu = y + gu, *yv = y + gv, *Tu = T + gu, *Tv = T + gv;
    C *hu_o = hout + gu, *hv_o = hout + gv;
    const double *Vab = pots + (long)x * E, *Vbb = pots + N + (long)x * E;
    const double *Db = pots + 2 * N + 2L * x * E, *aab = pots + 4 * N + (long)x * E, *abb = pots + 5 * N + (long)x * E;
    C Xu[KM], Xv[KM];
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int e = tid + k * NT;
        Xu[k] = yu[e];
        Xv[k] = yv[e];
    }
    const double s = sc ? sc[a] : 1.0;
    double ssum = 0.0;
#pragma unroll
    for (int c0 = 0; c0 < 2; c0 += NC) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KM; k++) {
            const int e = tid + k * NT;
            const int pe = (e / LZ) * (LZ + 1) + e % LZ;
            if constexpr (NC == 2) {
                A[pe] = Xu[k];
                A[EP + pe] = Xv[k];
            } else {
                A[pe] = c0 ? Xv[k] : Xu[k];
            }
        }
        __syncthreads();
        // z pencils (c, iy): offset c*EP + iy*(LZ+1), stride 1; y pencils
        // (c, iz): offset c*EP + iz, stride LZ+1
        dft_axis<LZ, NC * LY, LY, EP, LZ + 1, 1, NT, 1>(A, twz, tid, KsMul{0.0});
        dft_axis<LY, NC * LZ, LZ, EP, 1, LZ + 1, NT, 0>(A, twy, tid, YZMul{kys / LZ, zt, LZ});
        dft_axis<LZ, NC * LY, LY, EP, LZ + 1, 1, NT, 2>(A, twz, tid, KsMul{0.0});
        // H psi on owned points for the components in LDS (pointwise terms
        // and the sums in FP64)
#pragma unroll
        for (int k = 0; k < KM; k++) {
            const int e = tid + k * NT;
            const int pe = (e / LZ) * (LZ + 1) + e % LZ;
            const double2 pu = to64(Xu[k]) * s, pv = to64(Xv[k]) * s;
            const double2 D = make_double2(Db[2 * e], Db[2 * e + 1]);
#pragma unroll
            for (int cc = 0; cc < NC; cc++) {
                const int cmp = c0 + cc;
                if (cmp == 0) {
                    const double2 lu = to64(TX ? Tu[e] + A[cc * EP + pe] : A[cc * EP + pe]) * s;
                    const double aa = hp.bdg ? -0.5 : aab[e] * -0.5;
                    const double Va = Vab[e] - hp.mu_a;
                    const double2 hu = pu * Va + cmul(pv, D) + lu * aa;
                    hu_o[e] = cvt<C>(hu);
                    ssum += cmulc(pu, hu).x;
                } else {
                    const double2 lv = to64(TX ? Tv[e] + A[cc * EP + pe] : A[cc * EP + pe]) * s;
                    const double ab = hp.bdg ? -0.5 : abb[e] * -0.5;
                    const double Vb = Vbb[e] - hp.mu_b;
                    const double2 hv = pv * (-Vb) + cmulc(D, pu) - lv * ab;
                    hv_o[e] = cvt<C>(hv);
                    ssum += cmulc(pv, hv).x;
                }
            }
        }
    }
    ssum = block_sum(ssum, red);
    if (tid == 0) part[blockIdx.x] = ssum;
}

// T = D_x in (x-pencils, one lane per pencil): the x-pass of an H
// application to arbitrary vectors (the adjoint's)
template <int LX>
__global__ void __launch_bounds__(FT)
f_dx(const double2 *__restrict__ in, double2 *__restrict__ T, int n, long N, int S, double kscale)
{
    const long P = (long)blockIdx.x * FT + threadIdx.x;
    if (P >= 2L * n * S) return;
    const long a2 = P / S, j = P - a2 * S;
    const long base = a2 * N + j;
    double2 r[LX];
#pragma unroll
    for (int i = 0; i < LX; i++) r[i] = in[base + (long)i * S];
    dxx_inplace<LX>(r, kscale);
#pragma unroll
    for (int i = 0; i < LX; i++) T[base + (long)i * S] = r[i];
}

// out -= e_a in (the <H> subtraction, pca_kernels.cu:806-832)
__global__ void __launch_bounds__(FT)
f_sub(const double2 *__restrict__ in, double2 *__restrict__ out, const double *__restrict__ e, int n, long N)
{
    const long i = (long)blockIdx.x * FT + threadIdx.x;
    if (i >= 2L * n * N) return;
    const long a = (i / N) % n;
    out[i] = out[i] - in[i] * e[a];
}

// <H> per state from the yz pass's partials (per per state); with xpart,
// plus the x part of the kinetic term from the predictor (xper entries per
// component, unscaled: times s2 = s^2 of the state)
__global__ void f_e_final(const double *__restrict__ part, int n, int per, double dv, double *__restrict__ e,
                          double *__restrict__ e2, const double *__restrict__ xpart, int xper,
                          const double *__restrict__ s2)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    double s = 0.0;
    for (int k = 0; k < per; k++) s += part[(long)a * per + k];
    if (xpart) {
        double x = 0.0;
        const double *pu = xpart + (long)a * xper, *pv = xpart + ((long)n + a) * xper;
        for (int k = 0; k < xper; k++) x += pu[k];
        for (int k = 0; k < xper; k++) x += pv[k];
        s += x * s2[a];
    }
    e[a] = s * dv;
    if (e2) e2[a] = s * dv;
}

// y *= s ; f0 -= e * y   (applies the pending normalisation and <H> subtraction)
template <bool SP>
__global__ void __launch_bounds__(FT)
f_fixup(double2 *__restrict__ y, hc_t<SP> *__restrict__ f0, const double *__restrict__ s,
        const double *__restrict__ e, int n, long N, int do_f)
{
    const long i = (long)blockIdx.x * FT + threadIdx.x;
    if (i >= 2L * n * N) return;
    const long a = (i / N) % n;
    const double2 v = y[i] * s[a];
    y[i] = v;
    if (do_f) f0[i] = cvt<hc_t<SP>>(to64(f0[i]) - v * e[a]);
}

// The pending state of a fused step made explicit (tdw_finalize): y *= s,
// and the newest history entry completed, f0 += s sigma (-1/2 D_x) y - e y_k
// (x-pencils, as the predictor does it)
template <int LX, bool SP>
__global__ void __launch_bounds__(FT)
f_fixup_x(double2 *__restrict__ y, hc_t<SP> *__restrict__ f0, const double *__restrict__ sk,
          const double *__restrict__ ek, int n, long N, int S, double kscale)
{
    typedef hc_t<SP> HC;
    const long P = (long)blockIdx.x * FT + threadIdx.x;
    if (P >= 2L * n * S) return;
    const long a2 = P / S, j = P - a2 * S;
    const int a = (int)(a2 < n ? a2 : a2 - n);
    const double s = sk[a], e = ek[a];
    const long base = a2 * N + j;
    HC r[LX];
#pragma unroll
    for (int i = 0; i < LX; i++) r[i] = cvt<HC>(y[base + (long)i * S]);
    xkin_inplace<LX>(r, (a2 < n ? 0.5 : -0.5) * kscale, 0.0);
#pragma unroll
    for (int i = 0; i < LX; i++) {
        const long idx = base + (long)i * S;
        const double2 v = y[idx] * s;
        y[idx] = v;
        f0[idx] = cvt<HC>(to64(f0[idx]) + to64(r[i]) * s - v * e);
    }
}

// FP64 -> FP32 copy (the measurement step's exact H psi into the split history)
__global__ void __launch_bounds__(FT) f_to32(const double2 *__restrict__ src, float2 *__restrict__ dst, long tot)
{
    const long i = (long)blockIdx.x * FT + threadIdx.x;
    if (i < tot) dst[i] = to32(src[i]);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int ensure_f(tdw_ctx *c, size_t nd)
{
    if (nd <= c->fpart_cap) return 0;
    TDW_HIP(hipStreamSynchronize(c->stream));
    if (c->d_fpart) TDW_HIP(hipFree(c->d_fpart));
    c->d_fpart = nullptr;
    c->fpart_cap = 0;
    TDW_HIP(hipMalloc(&c->d_fpart, nd * sizeof(double)));
    c->fpart_cap = nd;
    return 0;
}

int fused_state_vecs(tdw_ctx *c)
{
    if (c->d_sp) return 0;
    TDW_HIP(hipMalloc(&c->d_sp, sizeof(double) * c->nwf));
    TDW_HIP(hipMalloc(&c->d_s2, sizeof(double) * c->nwf));
    TDW_HIP(hipMalloc(&c->d_ep, sizeof(double) * c->nwf));
    return 0;
}

void fused_release(tdw_ctx *c)
{
    if (c->d_fpart) hipFree(c->d_fpart);
    if (c->d_sp) hipFree(c->d_sp);
    if (c->d_s2) hipFree(c->d_s2);
    if (c->d_ep) hipFree(c->d_ep);
    c->d_fpart = nullptr;
    c->d_sp = c->d_s2 = c->d_ep = nullptr;
}

// the big-plane yz pass: both components in LDS (32 x 97 planes: 49.7 KB in
// FP32, 99 KB in FP64); 512 threads in FP32 (114 VGPRs: two workgroups per
// CU), 1024 in FP64 (the registers allow one workgroup per CU either way)
template <bool SP> struct HyzBig {
    static constexpr int NC = 2;
    static constexpr int NT = SP ? 512 : 1024;
};

template <bool SP> struct FusedFns {
    typedef hc_t<SP> HC;
    typedef void (*pred_fn)(double2 *, HC *, const HC *, const HC *, const HC *, const HC *, HC *, double *,
                            const double *, const double *, int, int, long, int, double, double, int, double *);
    typedef void (*corr_fn)(double2 *, const HC *, const HC *, double *, const double *, const double *, int, long,
                            int, double, double, int, double *, float2 *);
    typedef void (*fix_fn)(double2 *, HC *, const double *, const double *, int, long, int, double);
    typedef void (*hyz_fn)(const HC *, const HC *, HC *, const double *, const double *, int, long, int,
                           const HC *, const HC *, double, double, HParams, double *);

    template <int LX> static pred_fn pick_pred(int sch)
    {
        if (sch == TDW_AB3AM4) return f_pred<LX, 34, SP>;
        if (sch == TDW_AB4AM5) return f_pred<LX, 45, SP>;
        return f_pred<LX, 55, SP>;
    }

    static bool xpass(int nx, int sch, pred_fn *p, corr_fn *cf)
    {
        switch (nx) {
        case 8: *p = pick_pred<8>(sch); *cf = f_corr<8, SP>; return true;
        case 12: *p = pick_pred<12>(sch); *cf = f_corr<12, SP>; return true;
        case 16: *p = pick_pred<16>(sch); *cf = f_corr<16, SP>; return true;
        case 24: *p = pick_pred<24>(sch); *cf = f_corr<24, SP>; return true;
        case 32: *p = pick_pred<32>(sch); *cf = f_corr<32, SP>; return true;
        default: return false;
        }
    }

    static bool fix(int nx, fix_fn *f)
    {
        switch (nx) {
        case 8: *f = f_fixup_x<8, SP>; return true;
        case 12: *f = f_fixup_x<12, SP>; return true;
        case 16: *f = f_fixup_x<16, SP>; return true;
        case 24: *f = f_fixup_x<24, SP>; return true;
        case 32: *f = f_fixup_x<32, SP>; return true;
        default: return false;
        }
    }

    template <bool TX> static bool hyz_t(int ny, int nz, hyz_fn *h)
    {
        if (ny == 12 && nz == 12) { *h = f_hyz<12, 12, SP, TX>; return true; }
        // FP64 24^2 and 32^2 planes: the in-place 2D transform (90 / 106
        // VGPRs against 110 / 141 for the per-axis form: one more wave per SIMD)
        if (ny == 24 && nz == 24) {
            if constexpr (SP) *h = f_hyz<24, 24, SP, TX>;
            else *h = f_hyz_big<24, 24, SP, TX, 2, nt_hyz(24 * 24)>;
            return true;
        }
        if (ny == 32 && nz == 32) {
            if constexpr (SP) *h = f_hyz<32, 32, SP, TX>;
            else *h = f_hyz_big<32, 32, SP, TX, 2, nt_hyz(32 * 32)>;
            return true;
        }
        if (ny == 16 && nz == 16) { *h = f_hyz<16, 16, SP, TX>; return true; }
        if (ny == 8 && nz == 8) { *h = f_hyz<8, 8, SP, TX>; return true; }
        if (ny == 8 && nz == 12) { *h = f_hyz<8, 12, SP, TX>; return true; }
        if (ny == 12 && nz == 6) { *h = f_hyz<12, 6, SP, TX>; return true; }
        if (ny == 32 && nz == 96) { *h = f_hyz_big<32, 96, SP, TX, HyzBig<SP>::NC, HyzBig<SP>::NT>; return true; }
        return false;
    }

    // tx: the pass adds T (fused_apply_h); without it, the fused step's passes
    static bool hyz(int ny, int nz, bool tx, hyz_fn *h) { return tx ? hyz_t<true>(ny, nz, h) : hyz_t<false>(ny, nz, h); }
};

static bool split_mode(const tdw_ctx *c) { return c->p.precision == TDW_SPLIT; }

// planes above 1024 points (the tube section) run f_hyz_big with its own
// thread counts; FP64 24^2 and 32^2 planes run it with nt_hyz's
static bool hyz_big(const tdw_ctx *c) { return c->ny * c->nz > 1024; }

// threads of the yz pass in the precision SP of its arrays
template <bool SP> static int hyz_nt(const tdw_ctx *c)
{
    if (!hyz_big(c)) return nt_hyz(c->ny * c->nz);
    return HyzBig<SP>::NT;
}

// dynamic LDS of the yz pass: the components it holds at once
template <bool SP> static size_t hyz_lds(const tdw_ctx *c)
{
    const size_t plane = (SP ? sizeof(float2) : sizeof(double2)) * (size_t)c->ny * (c->nz + 1);
    if (!hyz_big(c)) return 2 * plane;
    return HyzBig<SP>::NC * plane;
}

bool fused_supported(const tdw_ctx *c)
{
    if (c->dim != 3) return false;
    // the fused passes apply the constant-effective-mass Hamiltonian
    if (varying_mass(c)) return false;
    if (c->p.precision != TDW_FP64 && c->p.precision != TDW_SPLIT) return false;
    FusedFns<false>::pred_fn p;
    FusedFns<false>::corr_fn cf;
    FusedFns<false>::hyz_fn h;
    if (!FusedFns<false>::xpass(c->nx, c->p.scheme, &p, &cf)) return false;
    if (!FusedFns<false>::hyz(c->ny, c->nz, true, &h)) return false;
    // the FP64 yz pass runs in both forms (H of FP64 vectors, gradient steps)
    return hyz_lds<false>(c) <= (size_t)c->max_lds && hyz_lds<true>(c) <= (size_t)c->max_lds;
}

static int nhist(const tdw_ctx *c) { return c->p.scheme == TDW_AB3AM4 ? 3 : c->p.scheme == TDW_AB4AM5 ? 4 : 5; }

template <bool SP> static int fixup(tdw_ctx *c, double2 *y, void *f0, int n, int do_f)
{
    const long tot = 2L * n * (long)c->N;
    TDW_TIMED(c, "f_fixup");
    hipLaunchKernelGGL(f_fixup<SP>, dim3((tot + FT - 1) / FT), dim3(FT), 0, c->stream, y, (hc_t<SP> *)f0,
                       c->d_scale, c->d_qpe, n, (long)c->N, do_f);
    return (int)hipGetLastError();
}

template <bool SP> static int fixup_x(tdw_ctx *c, tdw_state *s)
{
    typename FusedFns<SP>::fix_fn f;
    if (!FusedFns<SP>::fix(c->nx, &f)) return TDW_EINVAL;
    const int n = s->n, S = c->S;
    const double kxs = pow(2.0 * M_PI / c->Lphys[0], 2) / c->nx;
    TDW_TIMED(c, "f_fixup");
    hipLaunchKernelGGL(f, dim3((2L * n * S + FT - 1) / FT), dim3(FT), 0, c->stream, (double2 *)s->wf,
                       (hc_t<SP> *)s->f[0], c->d_scale, c->d_qpe, n, (long)c->N, S, kxs);
    return (int)hipGetLastError();
}

int fused_finalize(tdw_ctx *c, tdw_state *s)
{
    const int p = c->pending;
    if (!p) return 0;
    c->pending = 0;
    // the ASLDA step's pending state lacks only the scale and the subtraction
    if (p == 2) {
        if (split_mode(c)) return fixup<true>(c, (double2 *)s->wf, s->f[0], s->n, 1);
        return fixup<false>(c, (double2 *)s->wf, s->f[0], s->n, 1);
    }
    if (split_mode(c)) return fixup_x<true>(c, s);
    return fixup_x<false>(c, s);
}

static int xpass_norm(tdw_ctx *c, int n, double *s_out, double *s2_out, int wave, double *nrm_out = nullptr)
{
    const int per = wave ? c->S / 64 : c->S;
    TDW_TIMED(c, "f_norm");
    hipLaunchKernelGGL(f_norm_final, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d_fpart, n, per,
                       dvol_norm(c), s_out, s2_out, nrm_out);
    return (int)hipGetLastError();
}

// The yz pass and <H>.  With xpart (the prediction's pass) T is not read:
// the x part of the kinetic term enters <H> from the predictor's partials
// (xper per component, times s2) and H psi from the corrector.
template <bool SP>
static int hyz(tdw_ctx *c, int n, const hc_t<SP> *y, const hc_t<SP> *T, hc_t<SP> *out, const double *pots,
               const double *scale, double *e_out, double *e_out2, const double *xpart = nullptr, int xper = 0,
               const double *s2 = nullptr)
{
    HParams hp{c->p.mu_a, c->p.mu_b, c->p.functional == TDW_BDG};
    const double kys = pow(2.0 * M_PI / c->Lphys[1], 2) / c->ny;
    const double kzs = pow(2.0 * M_PI / c->Lphys[2], 2) / c->nz;
    typename FusedFns<SP>::hyz_fn h;
    FusedFns<SP>::hyz(c->ny, c->nz, xpart == nullptr, &h);
    const size_t lds = hyz_lds<SP>(c);
    const hc_t<SP> *twy, *twz;
    if constexpr (SP) { twy = c->d_tw32[1]; twz = c->d_tw32[2]; }
    else { twy = c->d_tw[1]; twz = c->d_tw[2]; }
    // the norm partials are consumed already: reuse the partial buffer
    {
        TDW_TIMED(c, "f_hyz");
        hipLaunchKernelGGL(h, dim3((long)n * c->nx), dim3(hyz_nt<SP>(c)), lds, c->stream, y, T, out, pots,
                           scale, n,
                           (long)c->N, c->nx, twy, twz, kys, kzs, hp, c->d_fpart);
        TDW_HIP(hipGetLastError());
    }
    TDW_TIMED(c, "f_e");
    hipLaunchKernelGGL(f_e_final, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d_fpart, n, c->nx,
                       dvol_norm(c), e_out, e_out2, xpart, xper, s2);
    return (int)hipGetLastError();
}

// tape (FP64 only, no gradient step): also record what the adjoint's
// backward step reads (norms, <H>, densities and potentials of both halves,
// in the tape's slots); cbuf: the partial corrector's buffer instead of the
// oldest history entry (which then stays intact); f_next: the new history
// entry's buffer instead of the dropped oldest one.
static int tape_copy(tdw_ctx *c, const double *src, double *dst, long nd)
{
    return (int)hipMemcpyAsync(dst, src, sizeof(double) * nd, hipMemcpyDeviceToDevice, c->stream);
}

template <bool SP>
static int fused_step_t(tdw_ctx *c, tdw_state *s, int it, int grad_last, tdw_tape_step *tape = nullptr,
                        double2 *cbuf = nullptr, tdw_complex *f_next = nullptr)
{
    typedef hc_t<SP> HC;
    if ((tape || cbuf || f_next) && (SP || grad_last)) return TDW_EINVAL;
    // a gradient step's FP64 H psi and scratch live in ddx (and ddy):
    // refused before anything is written
    if (grad_last && (!s->ddx || !s->ddy || !s->ddz || !s->lap)) return TDW_EINVAL;
    const int n = s->n, h = nhist(c);
    const long N = (long)c->N;
    const int S = c->S;
    const int wave = (S % 64) == 0;
    TDW_TRY(fused_state_vecs(c));
    // partial sums: the norm partials (then the yz pass's <H> partials) at
    // the start of d_fpart, the predictor's x-kinetic partials behind both
    const int per = wave ? S / 64 : S;
    const size_t xoff = std::max((size_t)2 * n * per, (size_t)n * c->nx);
    TDW_TRY(ensure_f(c, xoff + (size_t)2 * n * per));
    double *xpart = c->d_fpart + xoff;
    typename FusedFns<SP>::pred_fn fp;
    typename FusedFns<SP>::corr_fn fc;
    FusedFns<SP>::xpass(c->nx, c->p.scheme, &fp, &fc);
    const double kxs = pow(2.0 * M_PI / c->Lphys[0], 2) / c->nx;
    const long nblk = (2L * n * S + FT - 1) / FT;
    // H of the prediction lives in lap (in the split form FP32 values, and
    // then the FP32 copy of the corrected state); ddx is no longer written
    // by the fused passes (T is only the yz kernels' argument)
    double2 *y = (double2 *)s->wf;
    HC *T = (HC *)s->ddx, *H = (HC *)s->lap;
    HC **f = (HC **)s->f;
    HC *fold = cbuf ? (HC *)cbuf : f[h - 1];
    double *qpe = s->qpe ? s->qpe : c->d_qpe;

    // ---- predictor ----
    {
        TDW_TIMED(c, "f_pred");
        hipLaunchKernelGGL(fp, dim3(nblk), dim3(FT), 0, c->stream, y, f[0], f[1], f[2], f[3], f[4], fold,
                           xpart, c->d_scale, c->d_qpe, c->pending, n, N, S, c->p.dt, kxs, wave, c->d_fpart);
        TDW_HIP(hipGetLastError());
    }
    c->pending = 0;
    TDW_TRY(xpass_norm(c, n, c->d_sp, c->d_s2, wave, tape ? tape->scal : nullptr));
    // the prediction: y itself (plain) or its FP32 copy in the consumed
    // history slot (split)
    const HC *yp = SP ? (const HC *)f[h - 1] : (const HC *)y;
    if constexpr (SP)
        TDW_TRY(densities32(c, n, (const float2 *)yp, c->d_fbeta, c->d_s2, c->d_kt, c->d_cnt, s->dens));
    else
        TDW_TRY(densities(c, n, y, nullptr, nullptr, nullptr, nullptr, c->d_fbeta, c->d_s2, c->d_kt, c->d_cnt,
                          s->dens, 0));
    if (tape) TDW_TRY(tape_copy(c, s->dens, tape->dens_p, 12 * N));
    TDW_TRY(step_potentials(c, it + 1, s->dens, s->pots));
    TDW_TRY(hkf_fields(c, s->pots, true));
    if (tape) TDW_TRY(tape_copy(c, s->pots, tape->pots_p, 12 * N));
    const double *hp = h_view(c, it + 1, s->pots);
    if (!hp) return TDW_EINVAL;
    TDW_TRY(hyz<SP>(c, n, yp, T, H, hp, c->d_sp, c->d_ep, tape ? tape->scal + n : nullptr, xpart, per, c->d_s2));

    // ---- corrector ----
    const double a0 = (c->p.scheme == TDW_AB3AM4) ? 9. / 24. : 251. / 720.;
    {
        TDW_TIMED(c, "f_corr");
        hipLaunchKernelGGL(fc, dim3(nblk), dim3(FT), 0, c->stream, y, H, fold, xpart, c->d_sp, c->d_ep, n, N,
                           S, c->p.dt * a0, kxs, wave, c->d_fpart, SP ? (float2 *)H : nullptr);
        TDW_HIP(hipGetLastError());
    }
    TDW_TRY(xpass_norm(c, n, c->d_scale, c->d_s2, wave, tape ? tape->scal + 2 * n : nullptr));
    if (!grad_last) {
        // the corrected state: y (plain) or its FP32 copy in H's buffer (split)
        const HC *yc = SP ? (const HC *)H : (const HC *)y;
        if constexpr (SP)
            TDW_TRY(densities32(c, n, (const float2 *)yc, c->d_fbeta, c->d_s2, c->d_kt, c->d_cnt, s->dens));
        else
            TDW_TRY(densities(c, n, y, nullptr, nullptr, nullptr, nullptr, c->d_fbeta, c->d_s2, c->d_kt, c->d_cnt,
                              s->dens, 0));
        if (tape) TDW_TRY(tape_copy(c, s->dens, tape->dens_c, 12 * N));
        TDW_TRY(step_potentials(c, it + 1, s->dens, s->pots));
        TDW_TRY(hkf_fields(c, s->pots, true));
        if (tape) TDW_TRY(tape_copy(c, s->pots, tape->pots_c, 12 * N));
        hp = h_view(c, it + 1, s->pots);
        if (!hp) return TDW_EINVAL;
        tdw_complex *tmp = f_next ? f_next : s->f[h - 1];
        for (int j = h - 1; j > 0; j--) s->f[j] = s->f[j - 1];
        s->f[0] = tmp;
        double *e2 = tape ? tape->scal + 3 * n : (qpe != c->d_qpe ? qpe : nullptr);
        TDW_TRY(hyz<SP>(c, n, yc, T, (HC *)s->f[0], hp, c->d_scale, c->d_qpe, e2, xpart, per, c->d_s2));
        if (tape && qpe != c->d_qpe) TDW_TRY(tape_copy(c, c->d_qpe, qpe, n));
        // pending: y unscaled, and f[0] without the x kinetic part and the
        // <H> subtraction (the next predictor or tdw_finalize adds both)
        c->pending = 1;
        return 0;
    }
    // measurement step: normalise in place, then the reference's derivative /
    // density / tau / potential / H sequence, in FP64 in both forms.  Where
    // the lattice has the in-pass gradient density kernels, the densities
    // with tau and j come from them and H from the fused x and yz passes
    // (no derivative array is stored); elsewhere the four derivative arrays
    // and the modular density and H kernels.
    TDW_TRY(fixup<SP>(c, y, nullptr, n, 0));
    const bool gpass = dens_grad_fused_supported(c);
    if (gpass) {
        TDW_TRY(dens_grad_fused(c, n, y, c->d_fbeta, nullptr, s->dens));
    } else {
        TDW_TRY(derivatives(c, 2 * n, y, (double2 *)s->ddx, (double2 *)s->ddy, (double2 *)s->ddz,
                            (double2 *)s->lap));
        TDW_TRY(densities(c, n, y, (double2 *)s->ddx, (double2 *)s->ddy, (double2 *)s->ddz, (double2 *)s->lap,
                          c->d_fbeta, nullptr, c->d_kt, c->d_cnt, s->dens, 1));
    }
    TDW_TRY(density_tau(c, s->dens));
    TDW_TRY(step_potentials(c, it + 1, s->dens, s->pots));
    TDW_TRY(hkf_fields(c, s->pots, true));
    hp = h_view(c, it + 1, s->pots);
    if (!hp) return TDW_EINVAL;
    tdw_complex *tmp = s->f[h - 1];
    for (int j = h - 1; j > 0; j--) s->f[j] = s->f[j - 1];
    s->f[0] = tmp;
    if constexpr (SP) {
        // exact H psi - <H> psi into ddx (free after the densities), then FP32 history
        if (gpass)
            TDW_TRY(fused_apply_h(c, n, y, (double2 *)s->ddx, (double2 *)s->ddy, hp, qpe));
        else
            TDW_TRY(apply_h(c, it + 1, n, y, (double2 *)s->ddx, (double2 *)s->lap, hp, c->d_kt, nullptr, qpe));
        const long tot = 2L * n * N;
        TDW_TIMED(c, "f_to32");
        hipLaunchKernelGGL(f_to32, dim3((tot + FT - 1) / FT), dim3(FT), 0, c->stream, (const double2 *)s->ddx,
                           (float2 *)s->f[0], tot);
        TDW_HIP(hipGetLastError());
    } else if (gpass) {
        TDW_TRY(fused_apply_h(c, n, y, (double2 *)s->f[0], (double2 *)s->ddx, hp, qpe));
    } else {
        TDW_TRY(apply_h(c, it + 1, n, y, (double2 *)s->f[0], (double2 *)s->lap, hp, c->d_kt, nullptr, qpe));
    }
    return 0;
}

// H applied to arbitrary FP64 vectors on the fused passes: T = D_x in
// (f_dx), then the yz pass adding T (scale 1).  With qpe: <H> per state into
// qpe and out -= <H> in (apply_h); without: H in only (apply_h_only).  With
// use_e: out -= use_e in, a known <H> (qpe is then not written).
int fused_apply_h(tdw_ctx *c, int n, const double2 *in, double2 *out, double2 *T, const double *pots, double *qpe,
                  const double *use_e)
{
    // FP64 vectors and FP64 passes, in either form of the context
    if (!fused_supported(c)) return TDW_EINVAL;
    TDW_TRY(fused_state_vecs(c));
    TDW_TRY(ensure_f(c, (size_t)n * c->nx));
    const int S = c->S;
    const double kxs = pow(2.0 * M_PI / c->Lphys[0], 2) / c->nx;
    void (*dx)(const double2 *, double2 *, int, long, int, double) = nullptr;
    switch (c->nx) {
    case 8: dx = f_dx<8>; break;
    case 12: dx = f_dx<12>; break;
    case 16: dx = f_dx<16>; break;
    case 24: dx = f_dx<24>; break;
    case 32: dx = f_dx<32>; break;
    default: return TDW_EINVAL;
    }
    {
        TDW_TIMED(c, "f_dx");
        hipLaunchKernelGGL(dx, dim3((2L * n * S + FT - 1) / FT), dim3(FT), 0, c->stream, in, T, n, (long)c->N, S,
                           kxs);
        TDW_HIP(hipGetLastError());
    }
    TDW_TRY(hyz<false>(c, n, in, T, out, pots, nullptr, (qpe && !use_e) ? qpe : c->d_ep, nullptr));
    const double *e = use_e ? use_e : qpe;
    if (e) {
        const long tot = 2L * n * (long)c->N;
        TDW_TIMED(c, "f_sub");
        hipLaunchKernelGGL(f_sub, dim3((tot + FT - 1) / FT), dim3(FT), 0, c->stream, in, out, e, n, (long)c->N);
        TDW_HIP(hipGetLastError());
    }
    return 0;
}

// The adjoint's taped step on the fused FP64 passes: y goes to y_out (NULL:
// in place) and the step is finalised at its end, so y_out holds the
// normalised state and the new history entry is complete, as the modular
// taped step leaves them.
int fused_adj_step(tdw_ctx *c, tdw_state *s, int it, tdw_complex *y_out, double2 *cbuf, tdw_complex *f_next,
                   tdw_tape_step *t)
{
    if (!fused_supported(c) || split_mode(c)) return TDW_EINVAL;
    TDW_TRY(fused_finalize(c, s));
    if (y_out && y_out != s->wf) {
        TDW_HIP(hipMemcpyAsync(y_out, s->wf, sizeof(double2) * 2 * (size_t)s->n * c->N, hipMemcpyDeviceToDevice,
                               c->stream));
        s->wf = y_out;
    }
    TDW_TRY(fused_step_t<false>(c, s, it, 0, t, cbuf, f_next));
    return fused_finalize(c, s);
}

int fused_step(tdw_ctx *c, tdw_state *s, int it, int grad_last)
{
    if (aslda_step_supported(c)) return aslda_fused_step(c, s, it, grad_last);
    if (!fused_supported(c)) {
        // the split form exists only as the fused step: its history arrays
        // hold FP32 values, which the FP64 step cannot continue from
        if (split_mode(c)) return TDW_EINVAL;
        TDW_TRY(fused_finalize(c, s));
        return tdw_step(c, s, it, grad_last);
    }
    if (split_mode(c)) return fused_step_t<true>(c, s, it, grad_last);
    return fused_step_t<false>(c, s, it, grad_last);
}

}  // namespace tdw
