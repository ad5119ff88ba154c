// tdw_regfft.hpp -- FFTs of compile-time length held in registers.
//
// One lane owns one pencil of L complex values.  The transform is a
// mixed-radix decimation-in-time Cooley-Tukey recursion, fully unrolled at
// compile time; twiddle factors are evaluated by constexpr Taylor series, so
// they are immediates in the generated code.
#pragma once
#include "tdw_device.hpp"

namespace tdw {
namespace rf {

constexpr double CPI = 3.141592653589793238462643383279502884;

// sin/cos of x in [-pi/4, pi/4] by Taylor series (error < 1e-17)
constexpr double sin_s(double x)
{
    double x2 = x * x, t = x, s = x;
    for (int n = 1; n < 12; n++) {
        t *= -x2 / ((2.0 * n) * (2.0 * n + 1.0));
        s += t;
    }
    return s;
}
constexpr double cos_s(double x)
{
    double x2 = x * x, t = 1.0, s = 1.0;
    for (int n = 1; n < 12; n++) {
        t *= -x2 / ((2.0 * n - 1.0) * (2.0 * n));
        s += t;
    }
    return s;
}
// cos and sin of 2*pi*k/L, reduced to one octant with exact symmetries.
constexpr double cos2pi(int k, int L)
{
    k = ((k % L) + L) % L;
    // angle = 2 pi k / L in [0, 2pi); use octant index o = floor(8k/L)
    int o = (8 * k) / L;
    double r = 2.0 * CPI * ((double)k / (double)L - (double)o / 8.0);   // in [0, pi/4)
    switch (o) {
    case 0: return cos_s(r);
    case 1: return sin_s(CPI / 4 - r);
    case 2: return -sin_s(r);
    case 3: return -cos_s(CPI / 4 - r);
    case 4: return -cos_s(r);
    case 5: return -sin_s(CPI / 4 - r);
    case 6: return sin_s(r);
    default: return cos_s(CPI / 4 - r);
    }
}
constexpr double sin2pi(int k, int L)
{
    k = ((k % L) + L) % L;
    int o = (8 * k) / L;
    double r = 2.0 * CPI * ((double)k / (double)L - (double)o / 8.0);
    switch (o) {
    case 0: return sin_s(r);
    case 1: return cos_s(CPI / 4 - r);
    case 2: return cos_s(r);
    case 3: return sin_s(CPI / 4 - r);
    case 4: return -sin_s(r);
    case 5: return -cos_s(CPI / 4 - r);
    case 6: return -cos_s(r);
    default: return -sin_s(CPI / 4 - r);
    }
}

template <int L> struct first_radix {
    static constexpr int value = (L % 8 == 0 && L != 8) ? 8 : (L % 4 == 0 && L != 4) ? 4
                               : (L % 3 == 0 && L != 3) ? 3 : (L % 2 == 0 && L != 2) ? 2 : L;
};

// multiply v by W_L^e, W = exp(SIGN*2*pi*i/L), e compile-time
template <int L, int SIGN, int E, typename C> __device__ __forceinline__ C twiddle(C v)
{
    if constexpr (E % L == 0) {
        return v;
    } else {
        typedef decltype(v.x) R;
        constexpr double c = cos2pi(E, L);
        constexpr double s = SIGN * sin2pi(E, L);
        return mk((R)c * v.x - (R)s * v.y, (R)c * v.y + (R)s * v.x);
    }
}

// small DFTs, in place
template <int R, int SIGN, typename C> struct small_dft;
template <int SIGN, typename C> struct small_dft<1, SIGN, C> {
    static __device__ __forceinline__ void run(C *) {}
};
template <int SIGN, typename C> struct small_dft<2, SIGN, C> {
    static __device__ __forceinline__ void run(C *v)
    {
        C a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    }
};
template <int SIGN, typename C> struct small_dft<3, SIGN, C> {
    static __device__ __forceinline__ void run(C *v)
    {
        typedef decltype(v[0].x) R;
        const R c1 = (R)-0.5, s1 = (R)(0.86602540378443864676 * SIGN);
        C a = v[0], b = v[1], c = v[2];
        C t1 = b + c, t3 = b - c;
        C t2 = mk(-s1 * t3.y, s1 * t3.x);
        C m = a + t1 * c1;
        v[0] = a + t1;
        v[1] = m + t2;
        v[2] = m - t2;
    }
};
template <int SIGN, typename C> struct small_dft<4, SIGN, C> {
    static __device__ __forceinline__ void run(C *v)
    {
        C a = v[0] + v[2], b = v[0] - v[2], c = v[1] + v[3], e = v[1] - v[3];
        C d = mk(-SIGN * e.y, SIGN * e.x);
        v[0] = a + c;
        v[1] = b + d;
        v[2] = a - c;
        v[3] = b - d;
    }
};

template <int SIGN, typename C> struct small_dft<8, SIGN, C> {
    static __device__ __forceinline__ void run(C *v)
    {
        typedef decltype(v[0].x) R;
        const R h = (R)0.70710678118654752440;
        C e[4] = {v[0], v[2], v[4], v[6]};
        C o[4] = {v[1], v[3], v[5], v[7]};
        small_dft<4, SIGN, C>::run(e);
        small_dft<4, SIGN, C>::run(o);
        C w1 = mk(h * (o[1].x - SIGN * o[1].y), h * (o[1].y + SIGN * o[1].x));
        C w2 = mk(-SIGN * o[2].y, SIGN * o[2].x);
        C w3 = mk(h * (-o[3].x - SIGN * o[3].y), h * (-o[3].y + SIGN * o[3].x));
        v[0] = e[0] + o[0];
        v[4] = e[0] - o[0];
        v[1] = e[1] + w1;
        v[5] = e[1] - w1;
        v[2] = e[2] + w2;
        v[6] = e[2] - w2;
        v[3] = e[3] + w3;
        v[7] = e[3] - w3;
    }
};

// Prime-factor (Good-Thomas) transform of N = N1 * N2 points, gcd(N1, N2) =
// 1, with no twiddles: input n = (N2 n1 + N1 n2) mod N, output
// k = (N2 i2 k1 + N1 i1 k2) mod N with i2 = N2^-1 mod N1, i1 = N1^-1 mod N2,
// so W_N^{nk} = W_N1^{n1 k1} W_N2^{n2 k2}; N1 DFTs of N2 points, then N2 of
// N1 points, every index fixed at compile time, so the values stay in the
// registers of v (12 = 4 * 3, 24 = 8 * 3).
constexpr int mod_inverse(int a, int m)
{
    for (int x = 1; x < m; x++)
        if ((a * x) % m == 1) return x;
    return 0;
}

template <int N1, int N2, int SIGN, typename C> struct pfa_dft {
    static constexpr int N = N1 * N2;
    static constexpr int I2 = mod_inverse(N2 % N1, N1), I1 = mod_inverse(N1 % N2, N2);
    static __device__ __forceinline__ void run(C *v)
    {
        C a[N1][N2];
#pragma unroll
        for (int n1 = 0; n1 < N1; n1++)
#pragma unroll
            for (int n2 = 0; n2 < N2; n2++) a[n1][n2] = v[(N2 * n1 + N1 * n2) % N];
#pragma unroll
        for (int n1 = 0; n1 < N1; n1++) small_dft<N2, SIGN, C>::run(a[n1]);
#pragma unroll
        for (int k2 = 0; k2 < N2; k2++) {
            C b[N1];
#pragma unroll
            for (int n1 = 0; n1 < N1; n1++) b[n1] = a[n1][k2];
            small_dft<N1, SIGN, C>::run(b);
#pragma unroll
            for (int k1 = 0; k1 < N1; k1++) v[(N2 * I2 * k1 + N1 * I1 * k2) % N] = b[k1];
        }
    }
};

template <int SIGN, typename C> struct small_dft<12, SIGN, C> : pfa_dft<4, 3, SIGN, C> {};
template <int SIGN, typename C> struct small_dft<24, SIGN, C> : pfa_dft<8, 3, SIGN, C> {};

// DFT of length L of x[OFF + STRIDE*m], m < L, written in natural order to
// y[0..L).  Recursion: L = R*M, sub-transforms of stride STRIDE*R.
template <int L, int SIGN, typename C> struct reg_fft {
    static constexpr int R = first_radix<L>::value;
    static constexpr int M = L / R;

    template <int OFF, int STRIDE, int N>
    static __device__ __forceinline__ void run(const C (&x)[N], C (&y)[L])
    {
        if constexpr (M == 1) {
#pragma unroll
            for (int m = 0; m < L; m++) y[m] = x[OFF + STRIDE * m];
            small_dft<L, SIGN, C>::run(y);
        } else {
            C sub[R][M];
            sub_run<0, OFF, STRIDE, N>(x, sub);
            // combine: X[k + M q] = sum_r (W_L^{r k} Y_r[k]) W_R^{r q}
            combine<0>(sub, y);
        }
    }

    template <int r, int OFF, int STRIDE, int N>
    static __device__ __forceinline__ void sub_run(const C (&x)[N], C (&sub)[R][M])
    {
        if constexpr (r < R) {
            reg_fft<M, SIGN, C>::template run<OFF + STRIDE * r, STRIDE * R, N>(x, sub[r]);
            sub_run<r + 1, OFF, STRIDE, N>(x, sub);
        }
    }

    template <int k>
    static __device__ __forceinline__ void combine(C (&sub)[R][M], C (&y)[L])
    {
        if constexpr (k < M) {
            C t[R];
            tw_load<0, k>(sub, t);
            small_dft<R, SIGN, C>::run(t);
#pragma unroll
            for (int q = 0; q < R; q++) y[k + M * q] = t[q];
            combine<k + 1>(sub, y);
        }
    }

    template <int r, int k>
    static __device__ __forceinline__ void tw_load(C (&sub)[R][M], C (&t)[R])
    {
        if constexpr (r < R) {
            t[r] = twiddle<L, SIGN, r * k>(sub[r][k]);
            tw_load<r + 1, k>(sub, t);
        }
    }
};

// in-place transform of a register array
template <int L, int SIGN, typename C> __device__ __forceinline__ void fft(C (&x)[L])
{
    if constexpr (L == 12 || L == 24) {
        // in place, no twiddles (fewer values live at once)
        small_dft<L, SIGN, C>::run(x);
    } else {
        C y[L];
        reg_fft<L, SIGN, C>::template run<0, 1, L>(x, y);
#pragma unroll
        for (int i = 0; i < L; i++) x[i] = y[i];
    }
}

// in-place DFT of R values: the small DFTs for 1, 2, 3, 4, 8 and 12 points,
// the recursive register FFT for other lengths (e.g. 6 = 2 * 3)
template <int R, int SIGN, typename C> __device__ __forceinline__ void dft_inplace(C *v)
{
    if constexpr (R == 1 || R == 2 || R == 3 || R == 4 || R == 8 || R == 12) {
        small_dft<R, SIGN, C>::run(v);
    } else {
        C y[R];
        reg_fft<R, SIGN, C>::template run<0, 1, R>(*reinterpret_cast<C(*)[R]>(v), y);
#pragma unroll
        for (int i = 0; i < R; i++) v[i] = y[i];
    }
}

// L = R1 * R2: the two factors of the in-place transforms in LDS, R1 a
// small DFT (2, 3, 4, 8), R2 a small DFT or a register FFT (96 = 8 * 12)
template <int L> struct fac2 {
    static constexpr int R1 = (L == 8 || L == 4 || L == 3 || L == 2) ? L
                            : (L % 8 == 0) ? 8 : (L % 4 == 0) ? 4 : (L % 3 == 0) ? 3 : 2;
    static constexpr int R2 = L / R1;
};

// -k_i^2 / L for index i of an axis of L points and physical length Lp
__device__ __forceinline__ double k2fac(int i, int L, double Lp)
{
    double k = (i < L / 2) ? 2. * M_PI / Lp * (double)i : 2. * M_PI / Lp * (double)(i - L);
    return -(k * k) / (double)L;
}

}  // namespace rf
}  // namespace tdw
