// tdw_uniform.cpp -- uniform SLDA / ASLDA solution and plane-wave BdG states
// (host).
//
// Restates hpc-engine/pca_uniform.h: fbeta (:70-78), solve_uniform_problem
// (:86-502, SLDA with the A0=1 effective mass or ASLDA with the
// polarization-dependent one of tdw_aslda.hpp, with or without spin symmetry),
// create_uniform_wf (:636-850), create_uniform_wf_2d_generic (:2044-2287),
// create_uniform_wf_1d_generic (:2321-2560) and the 1D transverse-mode table
// of wslda_wavevectors.h:21-80.
#include <complex>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "tdw_aslda.hpp"
#include "tdw_sldae.hpp"
#include "wslda_td.h"

typedef std::complex<double> zc;

namespace {

const double PI = 3.14159265358979323846;

double kval(int i, int n, double L)
{
    return (i < n / 2) ? 2. * PI / L * (double)i : 2. * PI / L * (double)(i - n);
}

struct Edf {
    double pmin, pmax, eps;
    double preg(double n) const
    {
        if (n <= pmin) return 0.0;
        if (n >= pmax) return 1.0;
        return 0.5 + 0.5 * tanh(1.0 * tan((PI * (n - pmin)) / (pmax - pmin) - PI / 2.0));
    }
    double dpreg(double n) const
    {
        if (n <= pmin || n >= pmax) return 0.0;
        double a = (PI * (n - pmin)) / (pmax - pmin) - PI / 2.0;
        return 0.5 * PI / pow(cosh(tan(a)) * cos(a), 2) / (pmax - pmin);
    }
    double pol(double na, double nb) const
    {
        double pr = preg(na + nb);
        return pr == 0.0 ? 0.0 : pr * (na - nb) / (na + nb);
    }
    double dpol(double na, double nb, int b) const
    {
        double np = na + nb, pr = preg(np), dpr = dpreg(np), r = 0.0;
        if (pr != 0.0) r += pr * (b ? -2.0 * na : 2.0 * nb) / (np * np);
        if (dpr != 0.0) r += dpr * (na - nb) / np;
        return r;
    }
    static double G(double p) { return 0.357 + 0.642 * p * p; }
    double D(double na, double nb) const
    {
        double p = pol(na, nb);
        return (1.0 / (20.0 * PI * PI)) * pow(6.0 * PI * PI * (na + nb), 5.0 / 3.0) *
               (G(p) - pow((1.0 + p) / 2.0, 5.0 / 3.0) - pow((1.0 - p) / 2.0, 5.0 / 3.0));
    }
    double dD(double na, double nb, int b) const
    {
        double p = pol(na, nb);
        double base = G(p) - pow((1.0 + p) / 2.0, 5.0 / 3.0) - pow((1.0 - p) / 2.0, 5.0 / 3.0);
        return 0.5 * pow(6.0 * PI * PI * (na + nb), 2.0 / 3.0) * base +
               (1.0 / (20.0 * PI * PI)) * pow(6.0 * PI * PI * (na + nb), 5.0 / 3.0) *
                   (2.0 * 0.642 * p - (5.0 / 6.0) * pow((1.0 + p) / 2.0, 2.0 / 3.0) +
                    (5.0 / 6.0) * pow((1.0 - p) / 2.0, 2.0 / 3.0)) *
                   dpol(na, nb, b);
    }
    double tC(double na, double nb) const { return pow(na + nb + eps, 1.0 / 3.0) / (-11.11); }
    double dtC(double na, double nb) const { return ((1.0 / 3.0) * pow(na + nb + eps, -2.0 / 3.0)) / (-11.11); }
};

void branch(double ea, double eb, double d, int br, double *E, double *u, double *v)
{
    double sq = sqrt((eb - ea) * (eb - ea) + 4.0 * (ea * eb + d * d));
    double e, vv;
    if (br == 0) {
        e = -0.5 * (eb - ea) + 0.5 * sq;
        vv = d * d / (pow(0.5 * (ea + eb) + 0.5 * sq, 2) + d * d);
    } else {
        e = -0.5 * (eb - ea) - 0.5 * sq;
        vv = d * d / (pow(0.5 * (ea + eb) - 0.5 * sq, 2) + d * d);
    }
    double uu = 1.0 - vv;
    vv = sqrt(vv);
    uu = sqrt(uu);
    if (eb + e < 0.0) vv = -vv;
    *E = e;
    *u = uu;
    *v = vv;
}

}  // namespace

extern "C" double tdw_fbeta(double E, double beta)
{
    double bE = beta * E;
    if (bE > 50.) return 0.0;
    if (bE < -50.) return 1.0;
    return 1.0 / (exp(bE) + 1.0);
}

extern "C" int tdw_uniform_solve(const tdw_params *c, double n0_a, double n0_b, double kc_in, double Tstart,
                                 double Tstop, double dT, double eps, double scmix, double muchange, int maxiter,
                                 int mu_rule, tdw_uniform *U)
{
    if (!c || !U || (c->functional != TDW_SLDA && c->functional != TDW_ASLDA && c->functional != TDW_BDG &&
                     c->functional != TDW_SLDAE))
        return TDW_EINVAL;
    // BdG (solve_uniform_problem_bdg, pca_uniform.h:1105-1470): bare coupling
    // 4 pi a, no density functional in the mean fields, mu += muchange (n0 - n)
    const bool bdg = c->functional == TDW_BDG;
    const bool ss = c->spinsymmetry != 0;
    const double gbare = 4.0 * PI * c->sclgth;
    const int NX = c->nx, NY = c->ny, NZ = c->nz;
    const double LX = c->dx * NX, LY = c->dy * NY, LZ = c->dz * NZ, LXYZ = LX * LY * LZ;
    std::vector<double> kk2((size_t)NX * NY * NZ);
    {
        size_t i = 0;
        for (int ix = 0; ix < NX; ix++)
            for (int iy = 0; iy < NY; iy++)
                for (int iz = 0; iz < NZ; iz++, i++) {
                    double a = kval(ix, NX, LX), b = kval(iy, NY, LY), d = kval(iz, NZ, LZ);
                    kk2[i] = a * a + b * b + d * d;
                }
    }
    Edf f{c->p_nmin, c->p_nmax, c->denseps};
    const bool cubic = c->regularization == TDW_CUBIC_CUTOFF;
    const double R0 = cubic ? 2.442749607806335 / (4.0 * PI * PI) : 1. / (2.0 * PI * PI);
    const double R1 = cubic ? 0.31786000 : 0.5;
    // ASLDA: alpha and its density derivatives at the target densities
    // (pca_uniform.h:157-163); SLDA / BdG: alpha = 1
    const bool as = c->functional == TDW_ASLDA;
    const double p0r = as ? f.pol(n0_a, n0_b) : 0.0;
    const double dpa = as ? f.dpol(n0_a, n0_b, 0) : 0.0, dpb = as ? f.dpol(n0_a, n0_b, 1) : 0.0;
    // SLDAE (solve_uniform_problem_sldae, pca_uniform.h:1532-2043): A, beta,
    // C~ of x = |a k_F| at the total density, alpha_a = alpha_b = A
    const bool se = c->functional == TDW_SLDAE;
    const double nt0 = n0_a + n0_b;
    const tdw::sldae::Local SL = tdw::sldae::local(nt0, 1.0, c->sclgth, c->force_a1 != 0);
    const double kFt = pow(3.0 * PI * PI * nt0, 1.0 / 3.0);
    const double xse = fabs(c->sclgth * kFt);
    const double alph_plus = as ? tdw::aslda::alpha_plus(p0r) : se ? SL.af : 1.0;
    const double alph_a = as ? tdw::aslda::alpha(p0r) : se ? SL.af : 1.0;
    const double alph_b = as ? tdw::aslda::alpha(-p0r) : se ? SL.af : 1.0;
    const double dalphm_dna = as ? tdw::aslda::dalpha_minus_dp(p0r) * dpa : 0.0;
    const double dalphm_dnb = as ? tdw::aslda::dalpha_minus_dp(p0r) * dpb : 0.0;
    const double dalphp_dna = as ? tdw::aslda::dalpha_plus_dp(p0r) * dpa : 0.0;
    const double dalphp_dnb = as ? tdw::aslda::dalpha_plus_dp(p0r) * dpb : 0.0;
    const double n0 = n0_a + n0_b;
    const double dtC_dna = as ? tdw::aslda::der_tildeC(n0, p0r, dpa, c->denseps) : f.dtC(n0_a, n0_b);
    const double dtC_dnb = as ? tdw::aslda::der_tildeC(n0, p0r, dpb, c->denseps) : dtC_dna;
    const double dD_dna = as ? tdw::aslda::der_funD(n0, p0r, dpa) : f.dD(n0_a, n0_b, 0);
    const double dD_dnb = as ? tdw::aslda::der_funD(n0, p0r, dpb) : f.dD(n0_a, n0_b, 1);
    const double tC = as ? tdw::aslda::tildeC(n0, p0r, c->denseps) : se ? SL.ctilde : f.tC(n0_a, n0_b);
    const double Dn = as ? tdw::aslda::funD(n0, p0r) : se ? (3. / 5.) * SL.bf * nt0 * SL.eF : f.D(n0_a, n0_b);
    double eF_a = pow(6.0 * PI * PI * n0_a, 2.0 / 3.0) / 2.0;
    double eF_b = pow(6.0 * PI * PI * n0_b, 2.0 / 3.0) / 2.0;
    double eF_avg = pow(3.0 * PI * PI * (n0_a + n0_b), 2.0 / 3.0) / 2.0;
    double kc = kc_in, ec = c->ec;
    const double mu0 = bdg ? 0.96087287 : 0.37, d0 = bdg ? 0.21 : 0.5;
    double mu_a = mu0 * eF_a, mu_b = mu0 * eF_b;
    if (se) {
        mu_a = tdw::sldae::chem_pot0(xse) * eF_a;
        mu_b = tdw::sldae::chem_pot0(xse) * eF_b;
    }
    double tau_a = pow(6.0 * PI * PI * n0_a, 5.0 / 3.0) / (10.0 * PI * PI);
    double tau_b = pow(6.0 * PI * PI * n0_b, 5.0 / 3.0) / (10.0 * PI * PI);
    double delta = (n0_a < n0_b) ? d0 * pow(6.0 * PI * PI * n0_a, 2.0 / 3.0) / 2.0
                                 : d0 * pow(6.0 * PI * PI * n0_b, 2.0 / 3.0) / 2.0;
    double nu = bdg ? -1.0 * delta / gbare : -1.0 * delta * tC / alph_plus;
    double V_a = 0, V_b = 0, n_a = 0, n_b = 0, S = 0, beta = 1.0;
    int iter = 0, nwf = 0;
    const double Ts = fabs(Tstart), Tp = fabs(Tstop), DT = fabs(dT);
    const double dsT = (Ts < Tp) ? -1.0 : 1.0;
    for (double TT = Ts * dsT; TT >= Tp * dsT; TT -= DT) {
        const double T = TT * dsT;
        if (T < 0.0) break;
        beta = (T <= 1.0e-16) ? 1.0e16 : 1.0 / (T * eF_avg);
        for (iter = 0; iter < maxiter; iter++) {
            double tau_a_old = tau_a, tau_b_old = tau_b, delta_old = delta, nu_old = nu;
            double V_a_old = V_a, V_b_old = V_b;
            if (se) {
                V_a = SL.af_p * (tau_a + tau_b) / 2. + SL.beta * SL.eF - SL.ctilde_p * delta * delta / SL.af -
                      SL.af_p * delta * nu / SL.af;
                V_b = V_a;
            } else if (as) {
                const double tau_p = tau_a + tau_b, tau_m = tau_a - tau_b;
                V_a = dalphm_dna * tau_m / 2.0 + dalphp_dna * (tau_p / 2.0 - delta * nu / alph_plus) -
                      dtC_dna * delta * delta / alph_plus + dD_dna;
                V_b = dalphm_dnb * tau_m / 2.0 + dalphp_dnb * (tau_p / 2.0 - delta * nu / alph_plus) -
                      dtC_dnb * delta * delta / alph_plus + dD_dnb;
            } else {
                V_a = bdg ? 0.0 : -dtC_dna * delta * delta / alph_plus + dD_dna;
                V_b = bdg ? 0.0 : -dtC_dnb * delta * delta / alph_plus + dD_dnb;
            }
            double mu_p = (mu_a - V_a_old + mu_b - V_b_old) / 2.0;
            zc p0 = std::sqrt(zc(2.0 * mu_p / alph_plus, 0.0));
            if (p0.imag() < 0.) p0 *= -1.;
            if (cubic) kc = PI / c->dx;
            else ec = alph_plus * kc * kc / 2.0 - mu_p;
            zc wz = std::log((kc + p0) / (kc - p0));
            if (wz.imag() < 0.) wz += zc(0.0, 2. * PI);
            wz = kc * R0 * (1. - (p0 / kc) * R1 * wz);
            double g_eff = bdg ? (1.0 / (1.0 / gbare - wz)).real()
                               : se ? alph_plus / (tC - wz.real()) : (alph_plus / (tC - wz)).real();
            delta = -1.0 * g_eff * nu;
            if (c->mf_correction) {
                double kF = pow(3.0 * PI * PI * (n0_a + n0_b), 1.0 / 3.0);
                double bcoeff = p0.real() / (kF + 1.0e-12);
                double x = bcoeff * kF * c->dx / PI;
                if (g_eff < -1.0e-10 && x > 1.0e-10) {
                    double Lam_0 = bcoeff * R0 / x * (1.0 - R1 * x * log((1.0 + x) / (1.0 - x)));
                    double dLam_0_dx = (-2. * bcoeff * R0 * R1) / (1. - x * x) - bcoeff * R0 / x / x;
                    double dLam_dn = (kF / (3. * (n0_a + n0_b) * alph_plus)) * (Lam_0 + dLam_0_dx * bcoeff * c->dx * kF / PI);
                    V_a += dLam_dn * pow(delta, 2);
                    V_b += dLam_dn * pow(delta, 2);
                }
            }
            S = n_a = n_b = 0.0;
            tau_a = tau_b = nu = 0.0;
            nwf = 0;
            size_t ixyz = 0;
            for (int ix = 0; ix < NX; ix++)
                for (int iy = 0; iy < NY; iy++)
                    for (int iz = 0; iz < NZ; iz++, ixyz++) {
                        if (!cubic) {
                            if ((c->dim == 2 || c->dim == 1) && iz == NZ / 2) continue;
                            if (c->dim == 1 && iy == NY / 2) continue;
                        }
                        double ea = alph_a * kk2[ixyz] / 2.0 + V_a - mu_a;
                        double eb = alph_b * kk2[ixyz] / 2.0 + V_b - mu_b;
                        for (int br = 0; br < 2; br++) {
                            double E, uk, vk;
                            branch(ea, eb, delta, br, &E, &uk, &vk);
                            if (ss) {
                                // spin-symmetric mode: the E > 0 half of the
                                // spectrum, spin a folded into b
                                // (pca_uniform.h:325-337)
                                if (E > 0.0 && E < ec) {
                                    double fp = tdw_fbeta(E, beta), fm = tdw_fbeta(-E, beta);
                                    n_b += vk * vk * fm + uk * uk * fp;
                                    tau_b += kk2[ixyz] * (vk * vk * fm + uk * uk * fp);
                                    nu += uk * vk * (fm - fp) * 2.0;
                                    if (fp > 1.0e-16) S -= 2. * fp * log(fp);
                                    if (fm > 1.0e-16) S -= 2. * fm * log(fm);
                                    nwf++;
                                }
                            } else if (E > -ec && E < ec) {
                                double fp = tdw_fbeta(E, beta), fm = tdw_fbeta(-E, beta);
                                n_a += uk * uk * fp;
                                tau_a += kk2[ixyz] * uk * uk * fp;
                                n_b += vk * vk * fm;
                                tau_b += kk2[ixyz] * vk * vk * fm;
                                nu += uk * vk * (fm - fp);
                                if (fp > 1.0e-16) S -= fp * log(fp);
                                if (fm > 1.0e-16) S -= fm * log(fm);
                                nwf++;
                            }
                        }
                    }
            if (ss) {  // pca_uniform.h:397-402
                n_b /= LXYZ; n_a = n_b; tau_b /= LXYZ; tau_a = tau_b; nu /= 2.0 * LXYZ;
            } else {
                n_a /= LXYZ; n_b /= LXYZ; tau_a /= LXYZ; tau_b /= LXYZ; nu /= 2.0 * LXYZ;
            }
            if (fabs(n0_a - n_a) <= eps && fabs(n0_b - n_b) <= eps && fabs(delta - delta_old) <= eps) break;
            tau_a = (1. - scmix) * tau_a_old + scmix * tau_a;
            tau_b = (1. - scmix) * tau_b_old + scmix * tau_b;
            delta = (1. - scmix) * delta_old + scmix * delta;
            nu = (1. - scmix) * nu_old + scmix * nu;
            if (bdg) {
                mu_a += muchange * (n0_a - n_a);
                mu_b += muchange * (n0_b - n_b);
            } else if (mu_rule == 1) {
                mu_a += muchange * (Ts / T) * (n0_a - n_a);
                mu_b += muchange * (Ts / T) * (n0_b - n_b);
            } else {
                mu_a += muchange * (n0_a - n_a) / n0_a * eF_a;
                mu_b += muchange * (n0_b - n_b) / n0_b * eF_b;
            }
            if (ss) mu_b = mu_a;  // pca_uniform.h:429, 1432
        }
        eF_a = pow(6.0 * PI * PI * n_a, 2.0 / 3.0) / 2.0;
        eF_b = pow(6.0 * PI * PI * n_b, 2.0 / 3.0) / 2.0;
        eF_avg = pow(3.0 * PI * PI * (n_a + n_b), 2.0 / 3.0) / 2.0;
    }
    U->n0_a = n_a; U->n0_b = n_b; U->mu_a = mu_a; U->mu_b = mu_b; U->V_a = V_a; U->V_b = V_b;
    U->alph_a = alph_a; U->alph_b = alph_b; U->kc = kc; U->ec = ec; U->nu = nu; U->delta = delta;
    U->tau_a = tau_a; U->tau_b = tau_b; U->beta = beta; U->S = S; U->nwf = nwf; U->iters = iter;
    U->ekin = (0.5 * alph_a * tau_a + 0.5 * alph_b * tau_b) * LXYZ;
    U->epot = bdg ? 0.0 : Dn * LXYZ;
    U->epair = -1.0 * delta * nu * LXYZ;
    return iter == maxiter ? 1 : 0;
}

namespace {

struct Kmodes1d {
    std::vector<double> ky, kz;
    std::vector<int> w;
};

Kmodes1d kmodes_1d(const tdw_params *c)
{
    Kmodes1d km;
    std::vector<double> lk2;
    const bool cubic = c->regularization == TDW_CUBIC_CUTOFF;
    for (int iy = 0; iy < c->ny; iy++)
        for (int iz = 0; iz < c->nz; iz++) {
            double ky = kval(iy, c->ny, c->dy * c->ny), kz = kval(iz, c->nz, c->dz * c->nz);
            double k2 = ky * ky + kz * kz;
            if (!cubic) {
                if (c->ny > 1 && iy == c->ny / 2) continue;
                if (c->nz > 1 && iz == c->nz / 2) continue;
            }
            size_t i;
            for (i = 0; i < lk2.size(); i++)
                if (fabs(lk2[i] - k2) < 1.0e-12) break;
            if (i < lk2.size()) km.w[i]++;
            else {
                km.ky.push_back(ky);
                km.kz.push_back(kz);
                km.w.push_back(1);
                lk2.push_back(k2);
            }
        }
    return km;
}

int states_generic(const tdw_params *c, const tdw_uniform *U, int idxfrom, int idxto, tdw_complex *wf,
                   double *En, double *kt, int *cnt, int countonly)
{
    const int NX = c->nx, NY = c->ny, NZ = c->nz;
    size_t Nl = c->dim == 1 ? (size_t)NX : c->dim == 2 ? (size_t)NX * NY : (size_t)NX * NY * NZ;
    double vol = c->dim == 1 ? c->dx * NX : c->dim == 2 ? (c->dx * NX) * (c->dy * NY)
                                                       : (c->dx * NX) * (c->dy * NY) * (c->dz * NZ);
    const double sqv = sqrt(vol);
    Kmodes1d km;
    if (c->dim == 1) km = kmodes_1d(c);
    const int nloc = idxto - idxfrom;
    int nwf = -1;
    for (int ix = 0; ix < NX; ix++)
        for (int iy = 0; iy < NY; iy++)
            for (int iz = 0; iz < NZ; iz++) {
                const double kx = kval(ix, NX, c->dx * NX), ky = kval(iy, NY, c->dy * NY), kz = kval(iz, NZ, c->dz * NZ);
                int deg = 1;
                if (c->dim == 1) {
                    deg = 0;
                    for (size_t i = 0; i < km.w.size(); i++)
                        if (fabs(km.ky[i] - ky) < 1.0e-12 && fabs(km.kz[i] - kz) < 1.0e-12) { deg = km.w[i]; break; }
                } else if (c->dim == 2) {
                    deg = (iz == 0) ? 1 : (iz < NZ / 2) ? 2 : 0;
                    if (iz == NZ / 2) deg = c->regularization == TDW_CUBIC_CUTOFF ? 1 : 0;
                }
                if (deg <= 0) continue;
                const double k2 = kx * kx + ky * ky + kz * kz;
                const double ea = U->alph_a * k2 / 2.0 + U->V_a - U->mu_a;
                const double eb = U->alph_b * k2 / 2.0 + U->V_b - U->mu_b;
                for (int br = 0; br < 2; br++) {
                    double E, uk, vk;
                    branch(ea, eb, U->delta, br, &E, &uk, &vk);
                    // spin-symmetric mode: E > 0 states only (pca_uniform.h:727-733)
                    const double elo = c->spinsymmetry ? 0.0 : -U->ec;
                    if (!(E > elo && E < U->ec)) continue;
                    nwf++;
                    if (countonly || nwf < idxfrom || nwf >= idxto) continue;
                    const int l = nwf - idxfrom;
                    tdw_complex *u = wf + (size_t)l * Nl, *v = wf + ((size_t)nloc + l) * Nl;
                    for (size_t j = 0; j < Nl; j++) {
                        int jx, jy = 0, jz = 0;
                        if (c->dim == 1) jx = (int)j;
                        else if (c->dim == 2) { jx = (int)(j / NY); jy = (int)(j % NY); }
                        else { jx = (int)(j / ((size_t)NY * NZ)); jy = (int)((j / NZ) % NY); jz = (int)(j % NZ); }
                        double ph = (double)jx * c->dx * kx;
                        if (c->dim >= 2) ph += (double)jy * c->dy * ky;
                        if (c->dim >= 3) ph += (double)jz * c->dz * kz;
                        zc e = std::exp(zc(0.0, ph));
                        zc uu = e * uk / sqv, vv = e * vk / sqv;
                        u[j].re = uu.real(); u[j].im = uu.imag();
                        v[j].re = vv.real(); v[j].im = vv.imag();
                    }
                    if (En) En[l] = E;
                    if (c->dim == 1 && kt) { kt[l] = ky; kt[nloc + l] = kz; }
                    if (c->dim == 1 && cnt) cnt[l] = deg;
                    if (c->dim == 2 && kt) kt[l] = kz;
                }
            }
    return nwf + 1;
}

}  // namespace

extern "C" int tdw_uniform_count(const tdw_params *c, const tdw_uniform *u)
{
    if (!c || !u) return TDW_EINVAL;
    return states_generic(c, u, 0, 0, nullptr, nullptr, nullptr, nullptr, 1);
}

extern "C" int tdw_uniform_states(const tdw_params *c, const tdw_uniform *u, int idxfrom, int idxto,
                                  tdw_complex *h_wf, double *h_En, double *h_kt, int *h_cnt)
{
    if (!c || !u || idxto <= idxfrom) return TDW_EINVAL;
    states_generic(c, u, idxfrom, idxto, h_wf, h_En, h_kt, h_cnt, 0);
    return 0;
}
