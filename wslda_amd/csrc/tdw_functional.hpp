// tdw_functional.hpp -- pointwise SLDA / ASLDA / BdG functional on the device.
//
// Restates hpc-engine/pca_edf.h:17-352 (SLDA with the A0=1 effective mass of
// SLDA_FORCE_A1, pca_settings.h:80-90; ASLDA with the polarization-dependent
// effective mass of tdw_aslda.hpp) and the potential loops of
// tdwslda_functionals.h:47-306 (SLDA, ASLDA) and :373-452 (BdG).
#pragma once
#include "tdw_aslda.hpp"
#include "tdw_sldae.hpp"
#include "tdw_device.hpp"

namespace tdw {

// SLDA constants (pca_settings.h:80-110, 226-239)
#define TDW_A0 1.000
#define TDW_G0 0.357
#define TDW_G1 0.642
#define TDW_P_ALPHA 1.0
#define TDW_GAMMA0 (-11.11)
#define TDW_UD_SCITERS 15
#define TDW_UD_MIX 0.75

struct FParams {
    double p_nmin, p_nmax, denseps;
    double mu_a, mu_b, ec, dx;
    double R0, R1;
    int cubic, mf, bdg, aslda, sldae, a1;
    double sclgth;
    const double *agrid;   // project build: scattering_length() per point (N), or null
};

__device__ __forceinline__ double p_reg(const FParams &f, double n)
{
    if (n <= f.p_nmin) return 0.0;
    if (n >= f.p_nmax) return 1.0;
    return 0.5 + 0.5 * tanh(TDW_P_ALPHA * tan((M_PI * (n - f.p_nmin)) / (f.p_nmax - f.p_nmin) - M_PI / 2.0));
}
__device__ __forceinline__ double der_p_reg(const FParams &f, double n)
{
    if (n <= f.p_nmin || n >= f.p_nmax) return 0.0;
    double a = (M_PI * (n - f.p_nmin)) / (f.p_nmax - f.p_nmin) - M_PI / 2.0;
    double q = cosh(TDW_P_ALPHA * tan(a)) * cos(a);
    return 0.5 * TDW_P_ALPHA * M_PI / (q * q) / (f.p_nmax - f.p_nmin);
}
__device__ __forceinline__ double polar(const FParams &f, double na, double nb)
{
    double pr = p_reg(f, na + nb);
    if (pr == 0.0) return 0.0;
    return pr * (na - nb) / (na + nb);
}
__device__ __forceinline__ double der_pol(const FParams &f, double na, double nb, int wrt_b)
{
    double np = na + nb, pr = p_reg(f, np), dpr = der_p_reg(f, np), r = 0.0;
    if (pr != 0.0) r += pr * (wrt_b ? (-2.0 * na) : (2.0 * nb)) / (np * np);
    if (dpr != 0.0) r += dpr * (na - nb) / np;
    return r;
}
// alpha(p) = A0 for SLDA with A1 = A2 = 0.
__device__ __forceinline__ double alpha_f(double) { return TDW_A0; }
__device__ __forceinline__ double funG(double p) { return TDW_G0 + TDW_G1 * p * p; }

__device__ __forceinline__ double funD(const FParams &f, double na, double nb)
{
    double p = polar(f, na, nb);
    return (1.0 / (20.0 * M_PI * M_PI)) * pow(6.0 * M_PI * M_PI * (na + nb), 5.0 / 3.0) *
           (funG(p) - alpha_f(p) * pow((1.0 + p) / 2.0, 5.0 / 3.0) -
            alpha_f(-1.0 * p) * pow((1.0 - p) / 2.0, 5.0 / 3.0));
}
// dD/dn_a (wrt_b=0) or dD/dn_b (wrt_b=1), pca_edf.h:232-296 with A1=A2=0.
__device__ __forceinline__ double der_funD(const FParams &f, double na, double nb, int wrt_b)
{
    double p = polar(f, na, nb);
    double base = funG(p) - alpha_f(p) * pow((1.0 + p) / 2.0, 5.0 / 3.0) -
                  alpha_f(-1.0 * p) * pow((1.0 - p) / 2.0, 5.0 / 3.0);
    double dpol = der_pol(f, na, nb, wrt_b);
    return (1.0 / 2.0) * pow(6.0 * M_PI * M_PI * (na + nb), 2.0 / 3.0) * base +
           (1.0 / (20.0 * M_PI * M_PI)) * pow(6.0 * M_PI * M_PI * (na + nb), 5.0 / 3.0) *
               (2.0 * TDW_G1 * p - (5.0 / 6.0) * alpha_f(p) * pow((1.0 + p) / 2.0, 2.0 / 3.0) +
                (5.0 / 6.0) * alpha_f(-1.0 * p) * pow((1.0 - p) / 2.0, 2.0 / 3.0)) *
               dpol;
}
__device__ __forceinline__ double tildeC(const FParams &f, double na, double nb)
{
    return TDW_A0 * pow(na + nb + f.denseps, 1.0 / 3.0) / (TDW_GAMMA0 * 1.0);
}
__device__ __forceinline__ double der_tildeC(const FParams &f, double na, double nb)
{
    // d(alpha_plus)/dn = 0 for A2 = 0
    return ((1.0 / 3.0) * TDW_A0 * pow(na + nb + f.denseps, -2.0 / 3.0)) / (TDW_GAMMA0 * 1.0);
}

// complex helpers on double2 for the regularisation (principal branches, as
// thrust::complex<double>)
__device__ __forceinline__ double2 csqrt_real(double x)
{
    // sqrt(x + 0i)
    return x >= 0.0 ? make_double2(sqrt(x), 0.0) : make_double2(0.0, sqrt(-x));
}
__device__ __forceinline__ double2 cdiv(double2 a, double2 b)
{
    double d = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ __forceinline__ double2 clog2(double2 z)
{
    return make_double2(log(hypot(z.x, z.y)), atan2(z.y, z.x));
}

// One evaluation of the regularised coupling (tdwslda_functionals.h:251-265)
// given mu_+ - V_+ = t7 and the alpha_plus scaling; returns wz before the
// final inversion and p0 (for the mean-field correction).
__device__ __forceinline__ void reg_wz(const FParams &f, double t7, double alph_plus, double kscale,
                                       double2 &wz, double2 &p0)
{
    // kscale = alph_plus for SLDA (p0 = sqrt(2 t7/alpha)), 1 for BdG
    p0 = csqrt_real(2.0 * t7 / kscale);
    if (p0.y < 0.) p0 = make_double2(-p0.x, -p0.y);
    double2 kc;
    if (f.cubic) kc = make_double2(M_PI / f.dx, 0.0);
    else {
        kc = csqrt_real(2.0 * (f.ec + t7) / kscale);
        if (kc.y < 0.) kc = make_double2(-kc.x, -kc.y);
    }
    double2 w = clog2(cdiv(kc + p0, kc - p0));
    if (w.y < 0.) w.y += 2. * M_PI;
    // kc*R0*(1 - (p0/kc)*R1*w)
    double2 q = cmul(cdiv(p0, kc), w) * f.R1;
    double2 one_m = make_double2(1.0 - q.x, -q.y);
    wz = cmul(kc, one_m) * f.R0;
}

// Mean-field correction of the regularisation scheme (:268-290); returns
// dLam/dn or 0.
__device__ __forceinline__ double mf_corr(const FParams &f, double2 p0, double wz_re, double lkF,
                                          double n, double alph_plus)
{
    double bcoeff = p0.x / (lkF + 1.0e-12);
    double lx = bcoeff * lkF * f.dx / M_PI;
    if (wz_re < -1.0e-10 && lx > 1.0e-10) {
        double Lam0 = bcoeff * f.R0 / lx * (1.0 - f.R1 * lx * log((1.0 + lx) / (1.0 - lx)));
        double dLam0 = (-2. * bcoeff * f.R0 * f.R1) / (1. - lx * lx) - bcoeff * f.R0 / lx / lx;
        return (lkF / (3. * n * alph_plus)) * (Lam0 + dLam0 * bcoeff * f.dx * lkF / M_PI);
    }
    return 0.0;
}

// SLDA potentials at one point (tdwslda_functionals.h:47-306, FAST mode).
__device__ __forceinline__ void slda_point(const FParams &f, double na, double nb, double2 nu,
                                           double ua, double ub, double &Va, double &Vb,
                                           double2 &delta)
{
    const double alph_plus = TDW_A0;
    double Va_const = ua + der_funD(f, na, nb, 0);
    double Vb_const = ub + der_funD(f, na, nb, 1);
    const double t3 = der_tildeC(f, na, nb) / alph_plus;
    const double t4 = t3;
    const double t5 = tildeC(f, na, nb);
    double lkF = 0.0;
    if (f.mf) lkF = pow(3.0 * M_PI * M_PI * (na + nb), 1.0 / 3.0);
    double2 ld = make_double2(0.0, 0.0);
    for (int k = 0; k < TDW_UD_SCITERS; k++) {
        double t7 = (f.mu_a - Va + f.mu_b - Vb) / 2.0;
        double2 wz, p0;
        reg_wz(f, t7, alph_plus, alph_plus, wz, p0);
        // alph_plus / (t5 - wz)
        double2 g = cdiv(make_double2(alph_plus, 0.0), make_double2(t5 - wz.x, -wz.y));
        ld = nu * (-1.0 * g.x);
        double t6 = ld.x * nu.x + ld.y * nu.y;   // Re(conj(delta) * nu)
        double t7b = ld.x * ld.x + ld.y * ld.y;
        (void)t6;   // t1 = t2 = 0 in FAST mode
        double Vanew = Va_const - t3 * t7b;
        double Vbnew = Vb_const - t4 * t7b;
        if (f.mf) {
            double dl = mf_corr(f, p0, g.x, lkF, na + nb, alph_plus);
            Vanew += dl * t7b;
            Vbnew += dl * t7b;
        }
        Va = TDW_UD_MIX * Vanew + (1.0 - TDW_UD_MIX) * Va;
        Vb = TDW_UD_MIX * Vbnew + (1.0 - TDW_UD_MIX) * Vb;
    }
    delta = ld;
}

// ASLDA potentials at one point (tdwslda_functionals.h:47-306 without
// FAST_CONST_EFFECTIVE_MASS_MODE, with CURRENT_CORRECTIONS): V_a, V_b, Delta,
// the effective masses alpha_a, alpha_b and the vector potentials
// A_s = cc f_r(n_s) (1 - alpha_s) j_s / n_s (dim components; cc is the
// current-corrections coefficient).  tau and j are the densities of the same
// half-step.
__device__ __forceinline__ void aslda_point(const FParams &f, double na, double nb, double2 nu, double taua,
                                            double taub, const double *ja, const double *jb, int dim, double cc,
                                            double ua, double ub, double &Va, double &Vb, double2 &delta,
                                            double &ala, double &alb, double *Aa, double *Ab)
{
    const double n = na + nb;
    const double p = polar(f, na, nb);
    const double dpa = der_pol(f, na, nb, 0), dpb = der_pol(f, na, nb, 1);
    const double alph_plus = aslda::alpha_plus(p), alph_minus = aslda::alpha_minus(p);
    const double aA = aslda::alpha(p), aB = aslda::alpha(-p);
    ala = aA;
    alb = aB;
    const double fa = p_reg(f, na) * cc, fb = p_reg(f, nb) * cc;
    double j2a = 0.0, j2b = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (d >= dim) break;
        Aa[d] = fa == 0.0 ? 0.0 : fa * (1. - aA) * ja[d] / na;
        Ab[d] = fb == 0.0 ? 0.0 : fb * (1. - aB) * jb[d] / nb;
        j2a += ja[d] * ja[d];
        j2b += jb[d] * jb[d];
    }
    const double dalphp_dna = aslda::dalpha_plus_dp(p) * dpa, dalphp_dnb = aslda::dalpha_plus_dp(p) * dpb;
    const double dalphm_dna = aslda::dalpha_minus_dp(p) * dpa, dalphm_dnb = aslda::dalpha_minus_dp(p) * dpb;
    double Va_const = ua, Vb_const = ub;
    Va_const += dalphm_dna * (taua - taub) / 2.0;
    Vb_const += dalphm_dnb * (taua - taub) / 2.0;
    Va_const += dalphp_dna * (taua + taub) / 2.0;
    Vb_const += dalphp_dnb * (taua + taub) / 2.0;
    Va_const += aslda::der_funD(n, p, dpa);
    Vb_const += aslda::der_funD(n, p, dpb);
    if (fa != 0.0) {
        const double t7 = fa * j2a / (2.0 * na);
        Va_const += ((alph_plus + alph_minus - 1.0) / na - (dalphp_dna + dalphm_dna)) * t7;
        Va_const -= cc * der_p_reg(f, na) * (alph_plus + alph_minus - 1.0) * j2a / (2.0 * na);
        Vb_const -= (dalphp_dnb + dalphm_dnb) * t7;
    }
    if (fb != 0.0) {
        const double t7 = fb * j2b / (2.0 * nb);
        Va_const -= (dalphp_dna - dalphm_dna) * t7;
        Vb_const += ((alph_plus - alph_minus - 1.0) / nb - (dalphp_dnb - dalphm_dnb)) * t7;
        Vb_const -= cc * der_p_reg(f, nb) * (alph_plus - alph_minus - 1.0) * j2b / (2.0 * nb);
    }
    const double t1 = dalphp_dna / alph_plus, t2 = dalphp_dnb / alph_plus;
    const double t3 = aslda::der_tildeC(n, p, dpa, f.denseps) / alph_plus;
    const double t4 = aslda::der_tildeC(n, p, dpb, f.denseps) / alph_plus;
    const double t5 = aslda::tildeC(n, p, f.denseps);
    double lkF = 0.0;
    if (f.mf) lkF = pow(3.0 * M_PI * M_PI * n, 1.0 / 3.0);
    double2 ld = make_double2(0.0, 0.0);
    for (int k = 0; k < TDW_UD_SCITERS; k++) {
        double t7 = (f.mu_a - Va + f.mu_b - Vb) / 2.0;
        double2 wz, p0;
        reg_wz(f, t7, alph_plus, alph_plus, wz, p0);
        double2 g = cdiv(make_double2(alph_plus, 0.0), make_double2(t5 - wz.x, -wz.y));
        ld = nu * (-1.0 * g.x);
        const double t6 = ld.x * nu.x + ld.y * nu.y;   // Re(conj(delta) * nu)
        const double t7b = ld.x * ld.x + ld.y * ld.y;
        double Vanew = Va_const - t1 * t6 - t3 * t7b;
        double Vbnew = Vb_const - t2 * t6 - t4 * t7b;
        if (f.mf) {
            double dl = mf_corr(f, p0, g.x, lkF, n, alph_plus);
            Vanew += dl * t7b;
            Vbnew += dl * t7b;
        }
        Va = TDW_UD_MIX * Vanew + (1.0 - TDW_UD_MIX) * Va;
        Vb = TDW_UD_MIX * Vbnew + (1.0 - TDW_UD_MIX) * Vb;
    }
    delta = ld;
}

// SLDAE potentials at one point (tdwslda_functionals.h:530-812, with
// CURRENT_CORRECTIONS): alpha_a = alpha_b = A(x), x = |a k_F(n)|, the vector
// potentials as in ASLDA, V_s = U_ext + A' (tau_a + tau_b)/2 + beta e_F
// + current terms - (A'/A) Re(Delta* nu) - C~'/A |Delta|^2, and
// Delta = -nu A / (C~ - Lambda) with Lambda = Re of the regularised integral.
__device__ __forceinline__ void sldae_point(const FParams &f, double na, double nb, double2 nu, double taua,
                                            double taub, const double *ja, const double *jb, int dim, double cc,
                                            double ua, double ub, double &Va, double &Vb, double2 &delta,
                                            double &ala, double &alb, double *Aa, double *Ab)
{
    const double n = na + nb;
    const sldae::Local L = sldae::local(n, p_reg(f, n), f.sclgth, f.a1);
    const double af = L.af, af_p = L.af_p;
    ala = af;
    alb = af;
    // SLDA_FORCE_A1: FAST_CONST_EFFECTIVE_MASS_MODE, no current corrections
    const double fa = f.a1 ? 0.0 : p_reg(f, na) * cc, fb = f.a1 ? 0.0 : p_reg(f, nb) * cc;
    double j2a = 0.0, j2b = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (d >= dim) break;
        Aa[d] = fa == 0.0 ? 0.0 : fa * (1. - af) * ja[d] / na;
        Ab[d] = fb == 0.0 ? 0.0 : fb * (1. - af) * jb[d] / nb;
        j2a += ja[d] * ja[d];
        j2b += jb[d] * jb[d];
    }
    double Va_const = ua + af_p * (taua + taub) / 2.;
    double Vb_const = ub + af_p * (taua + taub) / 2.;
    Va_const += L.beta * L.eF;
    Vb_const += L.beta * L.eF;
    if (fa != 0.0) {
        const double t7 = fa * j2a / (2.0 * na);
        Va_const += ((af - 1.0) / na - af_p) * t7;
        Va_const -= cc * der_p_reg(f, na) * (af - 1.0) * j2a / (2.0 * na);
        Vb_const -= af_p * t7;
    }
    if (fb != 0.0) {
        const double t7 = fb * j2b / (2.0 * nb);
        Va_const -= af_p * t7;
        Vb_const += ((af - 1.0) / nb - af_p) * t7;
        Vb_const -= cc * der_p_reg(f, nb) * (af - 1.0) * j2b / (2.0 * nb);
    }
    double lkF = 0.0;
    if (f.mf) lkF = pow(3.0 * M_PI * M_PI * n, 1.0 / 3.0);
    double2 ld = make_double2(0.0, 0.0);
    for (int k = 0; k < TDW_UD_SCITERS; k++) {
        const double lmu = (f.mu_a - Va + f.mu_b - Vb) / 2.0;
        double2 wz, p0;
        reg_wz(f, lmu, af, af, wz, p0);
        const double g_eff = af / (L.ctilde - wz.x);
        ld = nu * (-1.0 * g_eff);
        const double dn = ld.x * nu.x + ld.y * nu.y;   // Re(conj(delta) * nu)
        const double d2 = ld.x * ld.x + ld.y * ld.y;
        double Vanew = Va_const - (af_p / af) * dn - L.ctilde_p / af * d2;
        double Vbnew = Vb_const - (af_p / af) * dn - L.ctilde_p / af * d2;
        if (f.mf) {
            const double dl = mf_corr(f, p0, g_eff, lkF, n, af);
            Vanew += dl * d2;
            Vbnew += dl * d2;
        }
        Va = TDW_UD_MIX * Vanew + (1.0 - TDW_UD_MIX) * Va;
        Vb = TDW_UD_MIX * Vbnew + (1.0 - TDW_UD_MIX) * Vb;
    }
    delta = ld;
}

// BdG potentials at one point (tdwslda_functionals.h:373-452).
__device__ __forceinline__ void bdg_point(const FParams &f, double na, double nb, double2 nu,
                                          double ua, double ub, double al_a, double al_b,
                                          double &Va, double &Vb, double2 &delta)
{
    const double alph_plus = 0.5 * (al_a + al_b);
    const double t5 = 1.0 / (4.0 * M_PI * f.sclgth * alph_plus);
    double nab = na + nb;
    double lkF = f.mf ? pow(3.0 * M_PI * M_PI * nab, 1.0 / 3.0) : 0.0;
    double t7 = (f.mu_a - Va + f.mu_b - Vb) / 2.0;
    double2 wz, p0;
    reg_wz(f, t7, alph_plus, 1.0, wz, p0);
    wz = wz * (1.0 / alph_plus);
    double2 g = cdiv(make_double2(1.0, 0.0), make_double2(t5 - wz.x, -wz.y));
    double2 ld = nu * (-1.0 * g.x);
    double nVa = ua, nVb = ub;
    if (f.mf) {
        double dl = mf_corr(f, p0, g.x, lkF, nab, alph_plus);
        double t7b = ld.x * ld.x + ld.y * ld.y;
        nVa += dl * t7b;
        nVb += dl * t7b;
    }
    Va = nVa;
    Vb = nVb;
    delta = ld;
}

}  // namespace tdw
