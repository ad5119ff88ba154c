// tdw_sldae.hpp -- the SLDAE functional's density-dependent parameters
// (host and device): the Pade[4/4] fits in x = |a k_F| of the effective mass
// alpha(x), the mean-field parameter beta(x), the pairing parameter
// 1/gamma(x) and the functional coefficients A(x), B(x), C(x), and the APS
// chemical potential that seeds the uniform solver.
//
// Restates hpc-engine/sldae_functional.c:14-330 with the default parameter
// set of sldae_parameters.h:1-75 (the td-gauss-sldae build) or, with a1,
// the SLDA_FORCE_A1 set of sldae_parameters_A1.h:1-36 (td-gauss-sldae-a1),
// and the constants of sldae_functional.h:36-56.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace tdw {
namespace sldae {

constexpr double PI_ = 3.14159265358979323846;
constexpr double GSE_UFG = 0.3582341;
constexpr double U_APS = 5.0 / 24.0;
// V = U / tan(3 pi / 16 (1 - GSE_UFG))
__host__ __device__ inline double v_aps() { return U_APS / tan(3.0 * PI_ / 16.0 * (1.0 - GSE_UFG)); }

// fixed values at x = 1e9 (sldae_parameters.h:72-75)
constexpr double AF_UFG = 0.8403360, BF_UFG = -0.281863, CF_UFG = -14.95950;

// Pade[4/4]: (a1 x + a2 x^2 + a3 x^3 + a4 x^4) / (1 + b1 x + ... + b4 x^4)
struct Pade {
    double a1, a2, a3, a4, b1, b2, b3, b4;
};
__host__ __device__ inline double pade0(const Pade &c, double x)
{
    const double x2 = x * x, x3 = x2 * x, x4 = x3 * x;
    return (c.a1 * x + c.a2 * x2 + c.a3 * x3 + c.a4 * x4) / (1. + c.b1 * x + c.b2 * x2 + c.b3 * x3 + c.b4 * x4);
}
__host__ __device__ inline double pade1(const Pade &c, double x)
{
    const double x2 = x * x, x3 = x2 * x, x4 = x3 * x;
    const double num = c.a1 * x + c.a2 * x2 + c.a3 * x3 + c.a4 * x4;
    const double den = 1. + c.b1 * x + c.b2 * x2 + c.b3 * x3 + c.b4 * x4;
    const double dnum = c.a1 + 2. * c.a2 * x + 3. * c.a3 * x2 + 4. * c.a4 * x3;
    const double dden = c.b1 + 2. * c.b2 * x + 3. * c.b3 * x2 + 4. * c.b4 * x3;
    return dnum / den - dden * num / (den * den);
}

// coefficient sets (sldae_functional.c:128-320, sldae_parameters.h:1-70);
// the beta, 1/gamma, B and C fits use a4 = b4 = c
__host__ __device__ inline Pade alpha_fit() { return {0.0, -0.2081560, -0.1358830, -0.0162071, 2.09838, 1.738, 0.668549, 0.101508}; }
__host__ __device__ inline Pade beta_fit() { return {1.58442, 17.2973, 0.820671, 3.48966, 13.9448, 10.1914, 11.9429, 3.48966}; }
__host__ __device__ inline Pade inv_gamma_fit() { return {6.15853, 0.156722, 0.76668, 0.0823379, 1.48431, 2.07468, 0.737836, 0.0823379}; }
__host__ __device__ inline Pade b_fit() { return {1.29871, 6.02212, 0.376541, 0.810972, 6.47079, 3.73389, 3.47301, 0.810972}; }
__host__ __device__ inline Pade c_fit() { return {0.267952, 0.754923, 0.330699, 0.0900897, 2.61695, 2.71855, 0.550587, 0.0900897}; }

// SLDA_FORCE_A1 builds (sldae_parameters_A1.h:1-36): A = 1 and these fits
constexpr double BF1_UFG = -0.481506, CF1_UFG = -10.532397;
__host__ __device__ inline Pade beta_fit_a1() { return {0.797400, 0.943961, 1.037043, 0.523445, 1.227062, 2.126210, 1.734795, 0.523445}; }
__host__ __device__ inline Pade inv_gamma_fit_a1() { return {2.713374, 1.441196, 3.700251, 1.877961, 0.634303, 3.522427, 3.701347, 1.877961}; }
__host__ __device__ inline Pade b_fit_a1() { return {0.698112, 0.034678, 0.455138, -0.031327, 0.353905, 0.806444, 0.403679, -0.031327}; }
__host__ __device__ inline Pade c_fit_a1() { return {0.370145, 1.041299, 1.737347, 1.085400, 2.458364, 5.604947, 3.828124, 1.085400}; }

__host__ __device__ inline double alpha0(double x, bool a1) { return a1 ? 1. : 1. + pade0(alpha_fit(), x); }   // = A(x)
__host__ __device__ inline double alpha1(double x, bool a1) { return a1 ? 0. : pade1(alpha_fit(), x); }
__host__ __device__ inline double beta0(double x, bool a1)
{
    return a1 ? BF1_UFG * pade0(beta_fit_a1(), x) : BF_UFG * pade0(beta_fit(), x);
}
__host__ __device__ inline double inv_gamma0(double x, bool a1)
{
    return a1 ? (1. / CF1_UFG) * pade0(inv_gamma_fit_a1(), x) : (1. / CF_UFG) * pade0(inv_gamma_fit(), x);
}
__host__ __device__ inline double bfun0(double x, bool a1)
{
    return a1 ? BF1_UFG * pade0(b_fit_a1(), x) : BF_UFG * pade0(b_fit(), x);
}
__host__ __device__ inline double cfun0(double x, bool a1)
{
    return a1 ? CF1_UFG * pade0(c_fit_a1(), x) : CF_UFG * pade0(c_fit(), x);
}

// APS interpolation s(x) and the chemical potential mu/e_F
// (sldae_functional.c:17-62)
__host__ __device__ inline double s_aps0(double x)
{
    const double V = v_aps();
    return atan(x * U_APS / (1. + x * V));
}
__host__ __device__ inline double s_aps1(double x)
{
    const double V = v_aps();
    const double uv = U_APS * U_APS + V * V;
    return U_APS / (1. + 2. * V * x + uv * x * x);
}
__host__ __device__ inline double chem_pot0(double x)
{
    return 1. - 16. / (3. * PI_) * s_aps0(x) - 16. / (15. * PI_) * x * s_aps1(x);
}

// Local coefficients of the potentials at total density n with the
// regularisation weight nreg = p_reg(n) (tdwslda_functionals.h:580-618):
// A = alpha, its n-derivative, beta, e_F, C~ and its n-derivative.
struct Local {
    double af, af_p, beta, eF, ctilde, ctilde_p, bf;
};
__host__ __device__ inline Local local(double n, double nreg, double a, bool a1)
{
    Local L;
    const double n13 = pow(n, 1. / 3.), n23 = pow(n, 2. / 3.);
    const double kF = pow(3. * PI_ * PI_ * n, 1. / 3.);
    L.eF = kF * kF / 2.;
    const double x = fabs(a * kF);
    const double dx_dn = nreg > 0.0 ? nreg * x / (3. * n) : 0.0;
    const double alpha = alpha0(x, a1);
    L.beta = beta0(x, a1);
    const double invg = inv_gamma0(x, a1);
    const double alpha_p = dx_dn * alpha1(x, a1);
    L.af = alpha;
    L.bf = bfun0(x, a1);
    const double cf = cfun0(x, a1);
    double cf_p = 0.0;
    L.af_p = 0.0;
    if (nreg > 0.0) {
        L.af_p = alpha_p;
        cf_p = nreg * cf / (3. * n) * (1. - cf * invg);
    }
    L.ctilde = L.af * n13 / cf;
    const double invg_eff = (1. / cf) * (1. - 3. * n / cf * cf_p);
    L.ctilde_p = nreg > 0.0 ? pow(nreg, 2. / 3.) * invg_eff * L.af / (3. * n23) : 0.0;
    L.ctilde_p += L.af_p * n13 / cf;
    return L;
}

}  // namespace sldae
}  // namespace tdw
