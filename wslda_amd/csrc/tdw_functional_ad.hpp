// tdw_functional_ad.hpp -- the pointwise SLDA functional over a generic
// scalar (double or Dual<K>), for the adjoint's Jacobian of the potentials.
//
// Same formulas, constants and branch decisions as tdw_functional.hpp
// (pca_edf.h:17-352, tdwslda_functionals.h:47-306, FAST constant-mass mode);
// branches are decided on values, so the dual result is the derivative of
// the branch the forward step took.
#pragma once
#include "tdw_dual.hpp"
#include "tdw_functional.hpp"

namespace tdw {
namespace ad {

template <typename R> struct S;
template <> struct S<double> {
    static __host__ __device__ __forceinline__ double c(double v) { return v; }
};
template <int K> struct S<Dual<K>> {
    static __host__ __device__ __forceinline__ Dual<K> c(double v) { return dconst<K>(v); }
};

template <typename R> __host__ __device__ __forceinline__ R p_reg(const FParams &f, const R &n)
{
    if (val(n) <= f.p_nmin) return S<R>::c(0.0);
    if (val(n) >= f.p_nmax) return S<R>::c(1.0);
    return 0.5 + 0.5 * ad_tanh(TDW_P_ALPHA * ad_tan((M_PI * (n - f.p_nmin)) / (f.p_nmax - f.p_nmin) - M_PI / 2.0));
}
template <typename R> __host__ __device__ __forceinline__ R der_p_reg(const FParams &f, const R &n)
{
    if (val(n) <= f.p_nmin || val(n) >= f.p_nmax) return S<R>::c(0.0);
    R a = (M_PI * (n - f.p_nmin)) / (f.p_nmax - f.p_nmin) - M_PI / 2.0;
    R q = ad_cosh(TDW_P_ALPHA * ad_tan(a)) * ad_cos(a);
    return 0.5 * TDW_P_ALPHA * M_PI / (q * q) / (f.p_nmax - f.p_nmin);
}
template <typename R> __host__ __device__ __forceinline__ R polar(const FParams &f, const R &na, const R &nb)
{
    R pr = p_reg(f, na + nb);
    if (val(pr) == 0.0) return S<R>::c(0.0);
    return pr * (na - nb) / (na + nb);
}
template <typename R>
__host__ __device__ __forceinline__ R der_pol(const FParams &f, const R &na, const R &nb, int wrt_b)
{
    R np = na + nb, pr = p_reg(f, np), dpr = der_p_reg(f, np);
    R r = S<R>::c(0.0);
    if (val(pr) != 0.0) r = r + pr * (wrt_b ? (-2.0 * na) : (2.0 * nb)) / (np * np);
    if (val(dpr) != 0.0) r = r + dpr * (na - nb) / np;
    return r;
}
template <typename R> __host__ __device__ __forceinline__ R funG(const R &p) { return TDW_G0 + TDW_G1 * p * p; }

template <typename R>
__host__ __device__ __forceinline__ R der_funD(const FParams &f, const R &na, const R &nb, int wrt_b)
{
    R p = polar(f, na, nb);
    R base = funG(p) - TDW_A0 * ad_pow((1.0 + p) / 2.0, 5.0 / 3.0) - TDW_A0 * ad_pow((1.0 - p) / 2.0, 5.0 / 3.0);
    R dpol = der_pol(f, na, nb, wrt_b);
    return (1.0 / 2.0) * ad_pow(6.0 * M_PI * M_PI * (na + nb), 2.0 / 3.0) * base +
           (1.0 / (20.0 * M_PI * M_PI)) * ad_pow(6.0 * M_PI * M_PI * (na + nb), 5.0 / 3.0) *
               (2.0 * TDW_G1 * p - (5.0 / 6.0) * TDW_A0 * ad_pow((1.0 + p) / 2.0, 2.0 / 3.0) +
                (5.0 / 6.0) * TDW_A0 * ad_pow((1.0 - p) / 2.0, 2.0 / 3.0)) *
               dpol;
}
template <typename R> __host__ __device__ __forceinline__ R tildeC(const FParams &f, const R &na, const R &nb)
{
    return TDW_A0 * ad_pow(na + nb + f.denseps, 1.0 / 3.0) / (TDW_GAMMA0 * 1.0);
}
template <typename R> __host__ __device__ __forceinline__ R der_tildeC(const FParams &f, const R &na, const R &nb)
{
    return ((1.0 / 3.0) * TDW_A0 * ad_pow(na + nb + f.denseps, -2.0 / 3.0)) / (TDW_GAMMA0 * 1.0);
}

template <typename R> __host__ __device__ __forceinline__ Cx<R> csqrt_real(const R &x)
{
    if (val(x) >= 0.0) return cx<R>(ad_sqrt(x), S<R>::c(0.0));
    return cx<R>(S<R>::c(0.0), ad_sqrt(-x));
}

template <typename R>
__host__ __device__ __forceinline__ void reg_wz(const FParams &f, const R &t7, double alph_plus, Cx<R> &wz, Cx<R> &p0)
{
    p0 = csqrt_real(2.0 * t7 / alph_plus);
    if (val(p0.y) < 0.) p0 = cx<R>(-p0.x, -p0.y);
    Cx<R> kc;
    if (f.cubic) kc = cx<R>(S<R>::c(M_PI / f.dx), S<R>::c(0.0));
    else {
        kc = csqrt_real(2.0 * (f.ec + t7) / alph_plus);
        if (val(kc.y) < 0.) kc = cx<R>(-kc.x, -kc.y);
    }
    Cx<R> w = cxlog(cxdiv(kc + p0, kc - p0));
    if (val(w.y) < 0.) w.y = w.y + 2. * M_PI;
    Cx<R> q = cxscale(cxmul(cxdiv(p0, kc), w), f.R1);
    Cx<R> one_m = cx<R>(1.0 - q.x, -q.y);
    wz = cxscale(cxmul(kc, one_m), f.R0);
}

template <typename R>
__host__ __device__ __forceinline__ R mf_corr(const FParams &f, const Cx<R> &p0, const R &wz_re, const R &lkF,
                                              const R &n, double alph_plus)
{
    R bcoeff = p0.x / (lkF + 1.0e-12);
    R lx = bcoeff * lkF * f.dx / M_PI;
    if (val(wz_re) < -1.0e-10 && val(lx) > 1.0e-10) {
        R Lam0 = bcoeff * f.R0 / lx * (1.0 - f.R1 * lx * ad_log((1.0 + lx) / (1.0 - lx)));
        R dLam0 = (-2. * f.R0 * f.R1) * bcoeff / (1. - lx * lx) - bcoeff * f.R0 / lx / lx;
        return (lkF / (3. * n * alph_plus)) * (Lam0 + dLam0 * bcoeff * f.dx * lkF / M_PI);
    }
    return S<R>::c(0.0);
}

// SLDA potentials at one point; Va, Vb enter as the previous potentials (the
// start of the 15-step mixing loop) and leave as the new ones.
template <typename R>
__host__ __device__ __forceinline__ void slda_point(const FParams &f, const R &na, const R &nb, const Cx<R> &nu,
                                                    const R &ua, const R &ub, R &Va, R &Vb, Cx<R> &delta)
{
    const double alph_plus = TDW_A0;
    R Va_const = ua + der_funD(f, na, nb, 0);
    R Vb_const = ub + der_funD(f, na, nb, 1);
    const R t3 = der_tildeC(f, na, nb) / alph_plus;
    const R t5 = tildeC(f, na, nb);
    R lkF = S<R>::c(0.0);
    if (f.mf) lkF = ad_pow(3.0 * M_PI * M_PI * (na + nb), 1.0 / 3.0);
    Cx<R> ld = cx<R>(S<R>::c(0.0), S<R>::c(0.0));
    for (int k = 0; k < TDW_UD_SCITERS; k++) {
        R t7 = (f.mu_a - Va + f.mu_b - Vb) / 2.0;
        Cx<R> wz, p0;
        reg_wz(f, t7, alph_plus, wz, p0);
        Cx<R> g = cxdiv(cx<R>(S<R>::c(alph_plus), S<R>::c(0.0)), cx<R>(t5 - wz.x, -wz.y));
        ld = cx<R>(nu.x * (-1.0 * g.x), nu.y * (-1.0 * g.x));
        R t7b = ld.x * ld.x + ld.y * ld.y;
        R Vanew = Va_const - t3 * t7b;
        R Vbnew = Vb_const - t3 * t7b;
        if (f.mf) {
            R dl = mf_corr(f, p0, g.x, lkF, na + nb, alph_plus);
            Vanew = Vanew + dl * t7b;
            Vbnew = Vbnew + dl * t7b;
        }
        Va = TDW_UD_MIX * Vanew + (1.0 - TDW_UD_MIX) * Va;
        Vb = TDW_UD_MIX * Vbnew + (1.0 - TDW_UD_MIX) * Vb;
    }
    delta = ld;
}

}  // namespace ad
}  // namespace tdw
