// tdw_fft.hpp -- batched Stockham FFT staged in LDS (kernel family a).
//
// A tile of P pencils of length L lives in LDS; element i of pencil p is at
// buf[p*ps + i*is].  Each stage of radix R and span Ns (product of the
// earlier radices) maps, for butterfly j < L/R,
//     x[j + t*L/R] * W_L^{t*(j%Ns)*L/(Ns*R)}  --DFT_R-->  y[(j/Ns)*Ns*R + j%Ns + t*Ns]
// (Stockham autosort, decimation in time), ping-ponging between two LDS
// buffers with one barrier per stage.  Twiddles come from a table of
// exp(-2*pi*i*k/L), k < L, staged in LDS; the inverse uses the conjugate.
// Spectral derivatives along one axis follow pca_derivative.cu:139-206 with
// the axis-separable factorisation lap = sum_d F_d^-1 (-k_d^2) F_d and
// d/dx_d = F_d^-1 (i k_d, Nyquist zeroed) F_d, equal to the reference's full
// transform in exact arithmetic.
#pragma once
#include "tdw_device.hpp"
#include "tdw_internal.hpp"

namespace tdw {

template <typename C> __device__ __forceinline__ C conj_if(C w, int sign)
{
    return sign > 0 ? mk(w.x, -w.y) : w;
}

// DFT of length R in registers; sign -1 forward, +1 inverse.
template <int R, typename C> __device__ __forceinline__ void dft(C *v, int sign);

template <> __device__ __forceinline__ void dft<2, double2>(double2 *v, int)
{
    double2 a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
}
template <> __device__ __forceinline__ void dft<2, float2>(float2 *v, int)
{
    float2 a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
}

template <typename C> __device__ __forceinline__ void dft4(C *v, int sign)
{
    C a = v[0] + v[2], b = v[0] - v[2], c = v[1] + v[3], e = v[1] - v[3];
    // d = (sign*i) * e
    C d = mk(-sign * e.y, sign * e.x);
    v[0] = a + c;
    v[1] = b + d;
    v[2] = a - c;
    v[3] = b - d;
}
template <> __device__ __forceinline__ void dft<4, double2>(double2 *v, int s) { dft4(v, s); }
template <> __device__ __forceinline__ void dft<4, float2>(float2 *v, int s) { dft4(v, s); }

template <typename C> __device__ __forceinline__ void dft3(C *v, int sign)
{
    typedef decltype(v[0].x) R;
    const R c1 = (R)-0.5, s1 = (R)0.86602540378443864676 * (R)sign;
    C a = v[0], b = v[1], c = v[2];
    C t1 = b + c, t3 = b - c;
    C t2 = mk(-s1 * t3.y, s1 * t3.x);
    C m = a + t1 * c1;
    v[0] = a + t1;
    v[1] = m + t2;
    v[2] = m - t2;
}
template <> __device__ __forceinline__ void dft<3, double2>(double2 *v, int s) { dft3(v, s); }
template <> __device__ __forceinline__ void dft<3, float2>(float2 *v, int s) { dft3(v, s); }

template <typename C> __device__ __forceinline__ void dft8(C *v, int sign)
{
    typedef decltype(v[0].x) R;
    const R h = (R)0.70710678118654752440;
    C e[4] = {v[0], v[2], v[4], v[6]};
    C o[4] = {v[1], v[3], v[5], v[7]};
    dft4(e, sign);
    dft4(o, sign);
    // W8^k = exp(sign*i*pi*k/4)
    C w1 = mk(h * (o[1].x - sign * o[1].y), h * (o[1].y + sign * o[1].x));
    C w2 = mk(-sign * o[2].y, sign * o[2].x);
    C w3 = mk(h * (-o[3].x - sign * o[3].y), h * (-o[3].y + sign * o[3].x));
    v[0] = e[0] + o[0];
    v[4] = e[0] - o[0];
    v[1] = e[1] + w1;
    v[5] = e[1] - w1;
    v[2] = e[2] + w2;
    v[6] = e[2] - w2;
    v[3] = e[3] + w3;
    v[7] = e[3] - w3;
}
template <> __device__ __forceinline__ void dft<8, double2>(double2 *v, int s) { dft8(v, s); }
template <> __device__ __forceinline__ void dft<8, float2>(float2 *v, int s) { dft8(v, s); }

template <int R, typename C>
__device__ __forceinline__ void stage_r(const C *src, C *dst, int P, int ps, int is, int L, int Ns,
                                        const C *tw, int sign, int tid, int nthr)
{
    const int nb = L / R;
    const int tstep = L / (Ns * R);
    for (int b = tid; b < P * nb; b += nthr) {
        const int p = b / nb, j = b - p * nb;
        const int k = j % Ns;
        C v[R];
#pragma unroll
        for (int t = 0; t < R; t++) v[t] = src[p * ps + (j + t * nb) * is];
#pragma unroll
        for (int t = 1; t < R; t++) v[t] = cmul(v[t], conj_if(tw[t * k * tstep], sign));
        dft<R, C>(v, sign);
        const int base = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int t = 0; t < R; t++) dst[p * ps + (base + t * Ns) * is] = v[t];
    }
}

// Full FFT of P pencils; returns the buffer holding the result (a or b).
template <typename C>
__device__ C *lds_fft(C *a, C *b, int P, int ps, int is, const tdw_axis &ax, const C *tw, int sign,
                      int tid, int nthr)
{
    C *src = a, *dst = b;
    for (int s = 0; s < ax.nst; s++) {
        switch (ax.R[s]) {
        case 8: stage_r<8>(src, dst, P, ps, is, ax.L, ax.Ns[s], tw, sign, tid, nthr); break;
        case 4: stage_r<4>(src, dst, P, ps, is, ax.L, ax.Ns[s], tw, sign, tid, nthr); break;
        case 3: stage_r<3>(src, dst, P, ps, is, ax.L, ax.Ns[s], tw, sign, tid, nthr); break;
        default: stage_r<2>(src, dst, P, ps, is, ax.L, ax.Ns[s], tw, sign, tid, nthr); break;
        }
        __syncthreads();
        C *t = src;
        src = dst;
        dst = t;
    }
    return src;
}

// -k^2/L and i*k/L (Nyquist zeroed) for index i along an axis.
__device__ __forceinline__ double kfac2(int i, int L, double Lp)
{
    double k = kval(i, L, Lp);
    return -(k * k) / (double)L;
}
__device__ __forceinline__ double kfac1(int i, int L, double Lp)
{
    if (i == L / 2) return 0.0;
    return kval(i, L, Lp) / (double)L;
}

}  // namespace tdw
