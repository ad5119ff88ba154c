// tdw_device.hpp -- device-side helpers shared by the td-step kernels.
//
// Complex values are (re, im) pairs in the memory layout of the reference's
// cufftDoubleComplex (hpc-engine/pca_kernels.h:6-8): double2 for FP64 and
// float2 for the FP32 operator of the split-precision mode.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wslda_td.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

namespace tdw {

template <typename T> struct cplx;
template <> struct cplx<double> { typedef double2 type; };
template <> struct cplx<float> { typedef float2 type; };

__device__ __forceinline__ double2 mk(double a, double b) { return make_double2(a, b); }
__device__ __forceinline__ float2 mk(float a, float b) { return make_float2(a, b); }
__device__ __forceinline__ double2 operator+(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 operator-(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 operator*(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 operator+(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 operator-(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 operator*(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

template <typename C> __device__ __forceinline__ C cmul(C a, C b)
{
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
template <typename C> __device__ __forceinline__ C cmulc(C a, C b)
{
    return mk(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
// i * s * a
template <typename C, typename R> __device__ __forceinline__ C imul(C a, R s) { return mk(-s * a.y, s * a.x); }
template <typename C> __device__ __forceinline__ auto cnorm(C a) { return a.x * a.x + a.y * a.y; }

__device__ __forceinline__ double2 to64(float2 a) { return make_double2((double)a.x, (double)a.y); }
__device__ __forceinline__ double2 to64(double2 a) { return a; }
__device__ __forceinline__ float2 to32(double2 a) { return make_float2((float)a.x, (float)a.y); }
__device__ __forceinline__ float2 to32(float2 a) { return a; }
template <typename C> __device__ __forceinline__ C cvt(double2 a);
template <> __device__ __forceinline__ double2 cvt<double2>(double2 a) { return a; }
template <> __device__ __forceinline__ float2 cvt<float2>(double2 a) { return to32(a); }

// ABM coefficient sums of the predictor (sp) and of the corrector's history
// part (sc) over the newest entries a, b, d, e, g (tdwslda_ode_integrator.h:26-97)
template <int SCH>
__device__ __forceinline__ void abm_sums(double2 a, double2 b, double2 d, double2 e, double2 g, double2 &sp,
                                         double2 &sc)
{
    if constexpr (SCH == 34) {
        sp = a * (23. / 12.) - b * (16. / 12.) + d * (5. / 12.);
        sc = a * (19. / 24.) - b * (5. / 24.) + d * (1. / 24.);
    } else if constexpr (SCH == 45) {
        sp = a * (55. / 24.) - b * (59. / 24.) + d * (37. / 24.) - e * (9. / 24.);
        sc = a * (646. / 720.) - b * (264. / 720.) + d * (106. / 720.) - e * (19. / 720.);
    } else {
        sp = a * (1901. / 720.) - b * (2774. / 720.) + d * (2616. / 720.) - e * (1274. / 720.) + g * (251. / 720.);
        sc = a * (646. / 720.) - b * (264. / 720.) + d * (106. / 720.) - e * (19. / 720.);
    }
}

// Lattice momentum of index i on an axis of n points and length L, standard
// FFT order (hpc-engine/pca_derivative.cu:150-160).
__device__ __forceinline__ double kval(int i, int n, double L)
{
    return (i < n / 2) ? 2. * M_PI / L * (double)i : 2. * M_PI / L * (double)(i - n);
}

// ---------------------------------------------------------------------------
// Wave64 reductions (kernel family c): DPP/shuffle tree inside a wavefront,
// then one LDS slot per wave.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide sum; every thread gets the result.  red must hold
// blockDim.x/64 doubles.  Deterministic for a fixed launch shape.
__device__ __forceinline__ double block_sum(double v, double *red)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = blockDim.x >> 6;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < nw; k++) s += red[k];
    return s;
}

} // namespace tdw
