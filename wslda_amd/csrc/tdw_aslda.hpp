// tdw_aslda.hpp -- the polarization-dependent pieces of the ASLDA functional,
// shared by the device potentials / energy kernels and the host uniform
// solver.
//
// Restates hpc-engine/pca_edf.h:96-352 with the ASLDA constants of
// pca_settings.h:92-99 (A0 = 1.094, A1 = 0.156, A2 = -0.532, G0 = 0.357,
// G1 = 0.642, gamma0 = -11.11).  Every function takes the polarization p
// (and its density derivative dp) already evaluated, so the same code serves
// the device (FParams) and the host (Edf) regularisations.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace tdw {
namespace aslda {

constexpr double A0 = 1.094, A1 = 0.156, A2 = -0.532;
constexpr double G0 = 0.357, G1 = 0.642, GAMMA0 = -11.11;
constexpr double PI_ = 3.14159265358979323846;

// alpha(p), pca_edf.h:96-108
__host__ __device__ inline double alpha(double p)
{
    const double p2 = p * p, p4 = p2 * p2;
    return A0 + A1 * p * (1.0 - (2.0 / 3.0) * p2 + (1.0 / 5.0) * p4) + A2 * p2 * (1.0 - p2 + (1.0 / 3.0) * p4);
}
// alpha_plus / alpha_minus, pca_edf.h:131-150 (the even / odd parts)
__host__ __device__ inline double alpha_plus(double p) { return 0.5 * (alpha(p) + alpha(-p)); }
__host__ __device__ inline double alpha_minus(double p) { return 0.5 * (alpha(p) - alpha(-p)); }
// d alpha_plus / dp and d alpha_minus / dp (the factors of
// der_alpha_{plus,minus}__der_n{a,b}, pca_edf.h:152-214)
__host__ __device__ inline double dalpha_plus_dp(double p)
{
    const double p2 = p * p, p4 = p2 * p2;
    return 2.0 * A2 * p * (1.0 - p2 + (1.0 / 3.0) * p4) - 2.0 * A2 * p * p2 * (1.0 - (2.0 / 3.0) * p2);
}
__host__ __device__ inline double dalpha_minus_dp(double p)
{
    const double p2 = p * p, p4 = p2 * p2;
    return A1 * (1.0 - (2.0 / 3.0) * p2 + (1.0 / 5.0) * p4) - 4.0 * A1 * p2 * ((1.0 / 3.0) - (1.0 / 5.0) * p2);
}
__host__ __device__ inline double funG(double p) { return G0 + G1 * p * p; }

// D(n_a, n_b) with n = n_a + n_b, pca_edf.h:228-243
__host__ __device__ inline double funD(double n, double p)
{
    return (1.0 / (20.0 * PI_ * PI_)) * pow(6.0 * PI_ * PI_ * n, 5.0 / 3.0) *
           (funG(p) - alpha(p) * pow((1.0 + p) / 2.0, 5.0 / 3.0) - alpha(-1.0 * p) * pow((1.0 - p) / 2.0, 5.0 / 3.0));
}
// dD/dn_s given dp = dp/dn_s, pca_edf.h:245-314
__host__ __device__ inline double der_funD(double n, double p, double dp)
{
    const double base = funG(p) - alpha(p) * pow((1.0 + p) / 2.0, 5.0 / 3.0) -
                        alpha(-1.0 * p) * pow((1.0 - p) / 2.0, 5.0 / 3.0);
    // d alpha(p)/dp and d alpha(-p)/dp
    const double da_p = dalpha_minus_dp(p) + dalpha_plus_dp(p);
    const double da_m = -1.0 * dalpha_minus_dp(p) + dalpha_plus_dp(p);
    return (1.0 / 2.0) * pow(6.0 * PI_ * PI_ * n, 2.0 / 3.0) * base +
           (1.0 / (20.0 * PI_ * PI_)) * pow(6.0 * PI_ * PI_ * n, 5.0 / 3.0) *
               (2.0 * G1 * p - da_p * pow((1.0 + p) / 2.0, 5.0 / 3.0) -
                (5.0 / 6.0) * alpha(p) * pow((1.0 + p) / 2.0, 2.0 / 3.0) - da_m * pow((1.0 - p) / 2.0, 5.0 / 3.0) +
                (5.0 / 6.0) * alpha(-1.0 * p) * pow((1.0 - p) / 2.0, 2.0 / 3.0)) *
               dp;
}
// tilde C and its derivative (ccoeff = 1), pca_edf.h:316-352
__host__ __device__ inline double tildeC(double n, double p, double eps)
{
    return alpha_plus(p) * pow(n + eps, 1.0 / 3.0) / (GAMMA0 * 1.0);
}
__host__ __device__ inline double der_tildeC(double n, double p, double dp, double eps)
{
    return (dalpha_plus_dp(p) * dp * pow(n + eps, 1.0 / 3.0) +
            (1.0 / 3.0) * alpha_plus(p) * pow(n + eps, -2.0 / 3.0)) /
           (GAMMA0 * 1.0);
}

}  // namespace aslda
}  // namespace tdw
