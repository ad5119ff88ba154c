// tdw_dual.hpp -- forward-mode dual numbers for pointwise Jacobians.
//
// Dual<K> carries a value and K tangents; every elementary function below
// propagates the tangents by its exact derivative.  Used by the adjoint to
// differentiate the pointwise SLDA functional (potentials as functions of
// the densities, the previous potentials and the external potential).
#pragma once
#include "tdw_device.hpp"

namespace tdw {

template <int K> struct Dual {
    double v;
    double d[K];
};

template <int K> __host__ __device__ __forceinline__ Dual<K> dconst(double v)
{
    Dual<K> r;
    r.v = v;
#pragma unroll
    for (int k = 0; k < K; k++) r.d[k] = 0.0;
    return r;
}
template <int K> __host__ __device__ __forceinline__ Dual<K> dvar(double v, int k0)
{
    Dual<K> r = dconst<K>(v);
    r.d[k0] = 1.0;
    return r;
}
// r = a*x + b*y on the tangents, value given
template <int K> __host__ __device__ __forceinline__ Dual<K> dlin(double v, double a, const Dual<K> &x)
{
    Dual<K> r;
    r.v = v;
#pragma unroll
    for (int k = 0; k < K; k++) r.d[k] = a * x.d[k];
    return r;
}
template <int K>
__host__ __device__ __forceinline__ Dual<K> dlin2(double v, double a, const Dual<K> &x, double b, const Dual<K> &y)
{
    Dual<K> r;
    r.v = v;
#pragma unroll
    for (int k = 0; k < K; k++) r.d[k] = a * x.d[k] + b * y.d[k];
    return r;
}

template <int K> __host__ __device__ __forceinline__ Dual<K> operator+(const Dual<K> &a, const Dual<K> &b)
{
    return dlin2(a.v + b.v, 1.0, a, 1.0, b);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> operator-(const Dual<K> &a, const Dual<K> &b)
{
    return dlin2(a.v - b.v, 1.0, a, -1.0, b);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> operator-(const Dual<K> &a)
{
    return dlin(-a.v, -1.0, a);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> operator*(const Dual<K> &a, const Dual<K> &b)
{
    return dlin2(a.v * b.v, b.v, a, a.v, b);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> operator/(const Dual<K> &a, const Dual<K> &b)
{
    const double q = a.v / b.v;
    return dlin2(q, 1.0 / b.v, a, -q / b.v, b);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> operator+(const Dual<K> &a, double b)
{
    return dlin(a.v + b, 1.0, a);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> operator+(double b, const Dual<K> &a)
{
    return dlin(a.v + b, 1.0, a);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> operator-(const Dual<K> &a, double b)
{
    return dlin(a.v - b, 1.0, a);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> operator-(double b, const Dual<K> &a)
{
    return dlin(b - a.v, -1.0, a);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> operator*(const Dual<K> &a, double b)
{
    return dlin(a.v * b, b, a);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> operator*(double b, const Dual<K> &a)
{
    return dlin(a.v * b, b, a);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> operator/(const Dual<K> &a, double b)
{
    return dlin(a.v / b, 1.0 / b, a);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> operator/(double b, const Dual<K> &a)
{
    const double q = b / a.v;
    return dlin(q, -q / a.v, a);
}

// value of a scalar (double or dual), for branches
__host__ __device__ __forceinline__ double val(double x) { return x; }
template <int K> __host__ __device__ __forceinline__ double val(const Dual<K> &x) { return x.v; }

// elementary functions: double overloads forward to libm, dual overloads
// apply the chain rule
__host__ __device__ __forceinline__ double ad_pow(double x, double e) { return pow(x, e); }
__host__ __device__ __forceinline__ double ad_sqrt(double x) { return sqrt(x); }
__host__ __device__ __forceinline__ double ad_log(double x) { return log(x); }
__host__ __device__ __forceinline__ double ad_tanh(double x) { return tanh(x); }
__host__ __device__ __forceinline__ double ad_tan(double x) { return tan(x); }
__host__ __device__ __forceinline__ double ad_cosh(double x) { return cosh(x); }
__host__ __device__ __forceinline__ double ad_cos(double x) { return cos(x); }
__host__ __device__ __forceinline__ double ad_atan2(double y, double x) { return atan2(y, x); }
__host__ __device__ __forceinline__ double ad_hypot(double x, double y) { return hypot(x, y); }

template <int K> __host__ __device__ __forceinline__ Dual<K> ad_pow(const Dual<K> &x, double e)
{
    const double p = pow(x.v, e);
    return dlin(p, e * pow(x.v, e - 1.0), x);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> ad_sqrt(const Dual<K> &x)
{
    const double s = sqrt(x.v);
    return dlin(s, 0.5 / s, x);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> ad_log(const Dual<K> &x)
{
    return dlin(log(x.v), 1.0 / x.v, x);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> ad_tanh(const Dual<K> &x)
{
    const double t = tanh(x.v);
    return dlin(t, 1.0 - t * t, x);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> ad_tan(const Dual<K> &x)
{
    const double t = tan(x.v);
    return dlin(t, 1.0 + t * t, x);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> ad_cosh(const Dual<K> &x)
{
    return dlin(cosh(x.v), sinh(x.v), x);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> ad_cos(const Dual<K> &x)
{
    return dlin(cos(x.v), -sin(x.v), x);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> ad_atan2(const Dual<K> &y, const Dual<K> &x)
{
    const double r2 = x.v * x.v + y.v * y.v;
    return dlin2(atan2(y.v, x.v), x.v / r2, y, -y.v / r2, x);
}
template <int K> __host__ __device__ __forceinline__ Dual<K> ad_hypot(const Dual<K> &x, const Dual<K> &y)
{
    const double h = hypot(x.v, y.v);
    return dlin2(h, x.v / h, x, y.v / h, y);
}

// complex numbers over a scalar type R (double or Dual<K>)
template <typename R> struct Cx {
    R x, y;
};
template <typename R> __host__ __device__ __forceinline__ Cx<R> cx(R a, R b)
{
    Cx<R> r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename R> __host__ __device__ __forceinline__ Cx<R> operator+(const Cx<R> &a, const Cx<R> &b)
{
    return cx<R>(a.x + b.x, a.y + b.y);
}
template <typename R> __host__ __device__ __forceinline__ Cx<R> operator-(const Cx<R> &a, const Cx<R> &b)
{
    return cx<R>(a.x - b.x, a.y - b.y);
}
template <typename R> __host__ __device__ __forceinline__ Cx<R> cxmul(const Cx<R> &a, const Cx<R> &b)
{
    return cx<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename R> __host__ __device__ __forceinline__ Cx<R> cxdiv(const Cx<R> &a, const Cx<R> &b)
{
    const R d = b.x * b.x + b.y * b.y;
    return cx<R>((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
template <typename R> __host__ __device__ __forceinline__ Cx<R> cxscale(const Cx<R> &a, double s)
{
    return cx<R>(a.x * s, a.y * s);
}
template <typename R> __host__ __device__ __forceinline__ Cx<R> cxlog(const Cx<R> &z)
{
    return cx<R>(ad_log(ad_hypot(z.x, z.y)), ad_atan2(z.y, z.x));
}

}  // namespace tdw
