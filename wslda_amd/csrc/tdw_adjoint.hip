// tdw_adjoint.hip -- reverse-mode (adjoint) PECE step and the taped forward
// step it differentiates.
//
// The forward map is the modular step of pca.c:1401-1570 (SLDA, FAST
// constant-mass mode, no quantum friction, non-gradient steps):
//   y_p = y - i dt sum b_j f_j ;  c = y - i dt sum a_j f_j      (abm_step1)
//   yh_p = y_p / n_p ; D_p = dens(yh_p) ; U_p = Phi(D_p, U_prev, V_ext)
//   F_p = H(U_p) yh_p - e_p yh_p ,  e_p = dV Re<yh_p, H yh_p>
//   y_c = -i dt a0 F_p + c                                         (abm_step4)
//   y'  = y_c / n_c ; D' = dens(y') ; U' = Phi(D', U_p, V_ext)
//   f_new = H(U') y' - e' y' ; history rotated
// The backward step applies the transpose of its Jacobian (as a real-linear
// map; gradients are paired by Re sum_r conj(a) b) in reverse order.  H is
// self-adjoint; Phi is differentiated pointwise with dual numbers; the
// derivative of the V_ext parameters is accumulated on the way.
#include "tdw_functional_ad.hpp"
#include "tdw_fused.hpp"
#include "tdw_internal.hpp"
#include "tdw_physics.hpp"

namespace tdw {

static constexpr int AT = 256;
static constexpr int NTHETA = 7 * TDW_MAX_GAUSS;   // amp, cx, cy, cz, wx, wy, wz per Gaussian

struct AbmCoef {
    int h;
    double b[5], a[5], a0;
};

// tdwslda_ode_integrator.h:26-97, 138-220
static AbmCoef abm_coef(int scheme)
{
    AbmCoef k{};
    if (scheme == TDW_AB3AM4) {
        k.h = 3;
        const double b[3] = {23. / 12., -16. / 12., 5. / 12.}, a[3] = {19. / 24., -5. / 24., 1. / 24.};
        for (int j = 0; j < 3; j++) { k.b[j] = b[j]; k.a[j] = a[j]; }
        k.a0 = 9. / 24.;
    } else if (scheme == TDW_AB4AM5) {
        k.h = 4;
        const double b[4] = {55. / 24., -59. / 24., 37. / 24., -9. / 24.};
        const double a[4] = {646. / 720., -264. / 720., 106. / 720., -19. / 720.};
        for (int j = 0; j < 4; j++) { k.b[j] = b[j]; k.a[j] = a[j]; }
        k.a0 = 251. / 720.;
    } else {
        k.h = 5;
        const double b[5] = {1901. / 720., -2774. / 720., 2616. / 720., -1274. / 720., 251. / 720.};
        const double a[4] = {646. / 720., -264. / 720., 106. / 720., -19. / 720.};
        for (int j = 0; j < 5; j++) k.b[j] = b[j];
        for (int j = 0; j < 4; j++) k.a[j] = a[j];
        k.a0 = 251. / 720.;
    }
    return k;
}

// ---------------------------------------------------------------------------
// per-state sums: part[a * gridDim.x + blk] over a chunk of points
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(AT)
k_sdot_part(const double2 *__restrict__ x, const double2 *__restrict__ y, int n, long N, long chunk,
            double *__restrict__ part)
{
    __shared__ double red[AT / 64];
    const int a = blockIdx.y;
    const long i0 = (long)blockIdx.x * chunk, i1 = min(N, i0 + chunk);
    const long ou = (long)a * N, ov = ((long)n + a) * N;
    double s = 0.0;
    for (long i = i0 + threadIdx.x; i < i1; i += AT)
        s += cmulc(x[ou + i], y[ou + i]).x + cmulc(x[ov + i], y[ov + i]).x;
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[(long)a * gridDim.x + blockIdx.x] = s;
}

__global__ void k_rows(const double *__restrict__ part, int n, int nchunk, double scale, double *__restrict__ out)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    double s = 0.0;
    for (int k = 0; k < nchunk; k++) s += part[(long)a * nchunk + k];
    out[a] = s * scale;
}

// ---------------------------------------------------------------------------
// out = acc + (hterm - em_a mu) - 2 s_a (g1 + e_a g2)  (elementwise), and the potential
// gradient of H: with nu = mu - s_a psi,
//   dL/dVa = sum Re(nu_u* psi_u), dL/dVb = -sum Re(nu_v* psi_v),
//   dL/dRe D = sum Re(nu_u* psi_v + nu_v* psi_u),
//   dL/dIm D = sum (-Im(nu_u* psi_v) + Im(nu_v* psi_u))
// summed over a chunk of states per point (point-major: coalesced rows).
// ---------------------------------------------------------------------------
// States are taken CU at a time, all loads of the group issued before its
// stores, so that a wave keeps several states' reads in flight.
template <int CU>
__global__ void __launch_bounds__(AT)
k_cross_upd(double2 *out, const double2 *acc, const double2 *hterm, const double2 *__restrict__ g1,
            const double2 *__restrict__ g2, const double2 *__restrict__ mu, const double2 *__restrict__ psi,
            const double *__restrict__ sv, const double *__restrict__ ev, const double *__restrict__ em, int n, long N,
            int chunk, double *__restrict__ part)
{
    const long i = (long)blockIdx.x * AT + threadIdx.x;
    if (i >= N) return;
    const int a0 = blockIdx.y * chunk, a1 = min(n, a0 + chunk);
    double gVa = 0.0, gVb = 0.0, gDr = 0.0, gDi = 0.0;
    for (int a = a0; a < a1; a += CU) {
        double2 Gu[CU], Gv[CU], Mu[CU], Mv[CU], Hu[CU], Hv[CU], Au[CU], Av[CU], Pu[CU], Pv[CU];
#pragma unroll
        for (int q = 0; q < CU; q++) {
            const int b = a + q;
            if (b >= a1) break;
            const long iu = (long)b * N + i, iv = ((long)n + b) * N + i;
            Pu[q] = psi[iu];
            Pv[q] = psi[iv];
            Gu[q] = g1[iu];
            Gv[q] = g1[iv];
            if (g2) {
                // g2 is psi in the history-refresh call: reuse the loaded values
                const double e = ev ? ev[b] : 0.0;
                Gu[q] = Gu[q] + (g2 == psi ? Pu[q] : g2[iu]) * e;
                Gv[q] = Gv[q] + (g2 == psi ? Pv[q] : g2[iv]) * e;
            }
            Mu[q] = mu[iu];
            Mv[q] = mu[iv];
            Hu[q] = hterm ? hterm[iu] : make_double2(0.0, 0.0);
            Hv[q] = hterm ? hterm[iv] : make_double2(0.0, 0.0);
            Au[q] = acc ? acc[iu] : make_double2(0.0, 0.0);
            Av[q] = acc ? acc[iv] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int q = 0; q < CU; q++) {
            const int b = a + q;
            if (b >= a1) break;
            const long iu = (long)b * N + i, iv = ((long)n + b) * N + i;
            const double s = sv[b];
            double2 ou = Gu[q] * (-2.0 * s), ov = Gv[q] * (-2.0 * s);
            if (hterm) {
                const double m = em ? em[b] : 0.0;
                ou = Hu[q] - Mu[q] * m + ou;
                ov = Hv[q] - Mv[q] * m + ov;
            }
            if (acc) { ou = Au[q] + ou; ov = Av[q] + ov; }
            const double2 nu_u = Mu[q] - Pu[q] * s, nu_v = Mv[q] - Pv[q] * s;
            out[iu] = ou;
            out[iv] = ov;
            gVa += cmulc(nu_u, Pu[q]).x;
            gVb -= cmulc(nu_v, Pv[q]).x;
            const double2 uv = cmulc(nu_u, Pv[q]), vu = cmulc(nu_v, Pu[q]);
            gDr += uv.x + vu.x;
            gDi += -uv.y + vu.y;
        }
    }
    double *p = part + (long)blockIdx.y * 4 * N + i;
    p[0] = gVa;
    p[N] = gVb;
    p[2 * N] = gDr;
    p[3 * N] = gDi;
}

// one thread per (gradient plane k, point i): 4N threads, each summing the
// nchunk state-chunk partials in chunk order
__global__ void k_cross_final(const double *__restrict__ part, int nchunk, long N, const double *__restrict__ add,
                              double *__restrict__ out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 4 * N) return;
    const long k = t / N, i = t - k * N;
    double s = 0.0;
    for (int c = 0; c < nchunk; c++) s += part[((long)c * 4 + k) * N + i];
    if (add) s += add[k * N + i];
    out[k * N + i] = s;
}

// ---------------------------------------------------------------------------
// transpose of the potential update at one point: dual numbers over the
// inputs (rho_a, rho_b, Re nu, Im nu, u_a, u_b, Va_prev, Vb_prev); lU holds
// dL/d(Va, Vb, Re D, Im D) in 4 planes of N.  Also accumulates dL/dtheta of
// the Gaussian parameters through u_a, u_b (block partials).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(AT)
k_pot_vjp(FParams f, VextDev V, long N, const double *__restrict__ dens, const double *__restrict__ pots_prev,
          const double *__restrict__ lU, const double *__restrict__ dens_add, double *__restrict__ lD,
          double *__restrict__ lUprev, double *__restrict__ gpart)
{
    typedef Dual<8> D8;
    __shared__ double red[AT / 64];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double lua = 0.0, lub = 0.0;
    if (i < N) {
        double ua, ub;
        vext_point(V, i, ua, ub);
        const D8 na = dvar<8>(dens[2 * N + i], 0), nb = dvar<8>(dens[7 * N + i], 1);
        const Cx<D8> nu = cx<D8>(dvar<8>(dens[2 * i], 2), dvar<8>(dens[2 * i + 1], 3));
        const D8 dua = dvar<8>(ua, 4), dub = dvar<8>(ub, 5);
        D8 Va = dvar<8>(pots_prev[i], 6), Vb = dvar<8>(pots_prev[N + i], 7);
        Cx<D8> del;
        ad::slda_point(f, na, nb, nu, dua, dub, Va, Vb, del);
        const double l0 = lU[i], l1 = lU[N + i], l2 = lU[2 * N + i], l3 = lU[3 * N + i];
        double g[8];
#pragma unroll
        for (int k = 0; k < 8; k++) g[k] = l0 * Va.d[k] + l1 * Vb.d[k] + l2 * del.x.d[k] + l3 * del.y.d[k];
        double d0 = g[0], d1 = g[1], d2 = g[2], d3 = g[3];
        if (dens_add) {
            d0 += dens_add[i];
            d1 += dens_add[N + i];
            d2 += dens_add[2 * N + i];
            d3 += dens_add[3 * N + i];
        }
        lD[i] = d0;
        lD[N + i] = d1;
        lD[2 * N + i] = d2;
        lD[3 * N + i] = d3;
        lUprev[i] = g[6];
        lUprev[N + i] = g[7];
        lUprev[2 * N + i] = 0.0;
        lUprev[3 * N + i] = 0.0;
        lua = g[4];
        lub = g[5];
    }
    // d(u_a, u_b)/dtheta of each Gaussian (same coordinates as vext_point)
    int ix = 0, iy = 0, iz = 0;
    if (i < N) decode(V.dim, V.ny, V.nz, i, ix, iy, iz);
    const double x = V.dx * (ix - V.nx / 2);
    const double y = V.dim >= 2 ? V.dy * (iy - V.ny / 2) : 0.0;
    const double z = V.dim >= 3 ? V.dz * (iz - V.nz / 2) : 0.0;
    const double lg = lua + (V.v.flip_b ? -lub : lub);   // dL/dg, g = the Gaussian sum
    for (int k = 0; k < TDW_MAX_GAUSS; k++) {
        double t[7] = {0, 0, 0, 0, 0, 0, 0};
        if (k < V.v.ngauss && i < N) {
            const double X = x - V.v.cx[k], Y = y - V.v.cy[k], Z = z - V.v.cz[k];
            const double E = V.s * exp(V.v.wx[k] * (X * X) + V.v.wy[k] * (Y * Y) + V.v.wz[k] * (Z * Z));
            const double A = V.v.amp[k];
            t[0] = lg * E;
            t[1] = lg * A * E * (-2.0 * V.v.wx[k] * X);
            t[2] = lg * A * E * (-2.0 * V.v.wy[k] * Y);
            t[3] = lg * A * E * (-2.0 * V.v.wz[k] * Z);
            t[4] = lg * A * E * (X * X);
            t[5] = lg * A * E * (Y * Y);
            t[6] = lg * A * E * (Z * Z);
        }
        for (int q = 0; q < 7; q++) {
            const double s = block_sum(t[q], red);
            if (threadIdx.x == 0) gpart[(long)blockIdx.x * NTHETA + 7 * k + q] = s;
        }
    }
}

// out[p] += sum over the nblk workgroup partials of parameter p: one
// workgroup per parameter, a fixed strided order per thread and a fixed tree
// (deterministic for a given nblk)
__global__ void __launch_bounds__(AT) k_gsum(const double *__restrict__ gpart, int nblk, double *__restrict__ out)
{
    __shared__ double red[AT / 64];
    const int p = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += AT) s += gpart[(long)b * NTHETA + p];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[p] += s;
}

// ---------------------------------------------------------------------------
// out += dens^T(lD) at psi (per state), and part = Re<psi, out> partials
//   rho_a = sum wa |u|^2, rho_b = sum wb |v|^2, nu = sum cn u v*   (per state
//   weights of k_dens_part: fbeta, transverse degeneracy, 1/transverse length)
// ---------------------------------------------------------------------------
template <int DIM>
__global__ void __launch_bounds__(AT)
k_dens_t(const double2 *__restrict__ psi, double2 *out, const double *__restrict__ lD,
         const double *__restrict__ fbeta, const double *__restrict__ kt, const int *__restrict__ cnt, int n, long N,
         long chunk, double tr, int cubic, double piz, double *__restrict__ part)
{
    __shared__ double red[AT / 64];
    const int a = blockIdx.y;
    const long i0 = (long)blockIdx.x * chunk, i1 = min(N, i0 + chunk);
    double w = 1.0;
    if (DIM == 1) w = (double)cnt[a];
    else if (DIM == 2) {
        const double kz = kt[a];
        w = (fabs(kz) < 1.0e-12) ? 1.0 : 2.0;
        if (cubic && fabs(kz + piz) < 1.0e-12) w = 1.0;
    }
    const double fb = fbeta[a];
    const double wa = fb * w / tr, wb = (1.0 - fb) * w / tr, cn = (1.0 - 2.0 * fb) / 2.0 * w / tr;
    const long ou = (long)a * N, ov = ((long)n + a) * N;
    double s = 0.0;
    for (long i = i0 + threadIdx.x; i < i1; i += AT) {
        const double2 u = psi[ou + i], v = psi[ov + i];
        const double2 ln = make_double2(lD[2 * N + i], lD[3 * N + i]);
        const double2 gu = out[ou + i] + u * (2.0 * wa * lD[i]) + cmul(ln, v) * cn;
        const double2 gv = out[ov + i] + v * (2.0 * wb * lD[N + i]) + cmul(make_double2(ln.x, -ln.y), u) * cn;
        out[ou + i] = gu;
        out[ov + i] = gv;
        s += cmulc(u, gu).x + cmulc(v, gv).x;
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[(long)a * gridDim.x + blockIdx.x] = s;
}

// ---------------------------------------------------------------------------
// corrector normalisation back, and the predictor quantities rebuilt from the
// tape:  ly <- (ly - t1 y1)/n_c (= dL/dy_c = dL/dc) ; w0 <- i dt a0 ly (dL/dF_p) ;
//   w3 <- yh_p = (y0 - i dt sum b f)/n_p ;
//   w4 <- H(U_p) yh_p = i (n_c y1 - c)/(dt a0) + e_p yh_p ;  part: Re<w0, w3>
// ---------------------------------------------------------------------------
template <int H>
__global__ void __launch_bounds__(AT)
k_recon(double2 *ly, const double2 *__restrict__ y1, const double2 *__restrict__ y0, const double2 *__restrict__ f0,
        const double2 *__restrict__ f1, const double2 *__restrict__ f2, const double2 *__restrict__ f3,
        const double2 *__restrict__ f4, double2 *__restrict__ w0, double2 *__restrict__ w3, double2 *__restrict__ w4,
        const double *__restrict__ t1, const double *__restrict__ scal, AbmCoef cf, double dt, int n, long N,
        long chunk, double *__restrict__ part)
{
    __shared__ double red[AT / 64];
    const int a = blockIdx.y;
    const long i0 = (long)blockIdx.x * chunk, i1 = min(N, i0 + chunk);
    const double np = sqrt(scal[a]), ep = scal[n + a], nc = sqrt(scal[2 * n + a]);
    const double tt = t1[a];
    double s = 0.0;
    for (int comp = 0; comp < 2; comp++) {
        const long o = ((long)comp * n + a) * N;
        for (long i = i0 + threadIdx.x; i < i1; i += AT) {
            const long idx = o + i;
            const double2 Y1 = y1[idx];
            const double2 lyc = (ly[idx] - Y1 * tt) * (1.0 / nc);
            ly[idx] = lyc;
            const double2 mu = imul(lyc, dt * cf.a0);
            w0[idx] = mu;
            // H (the scheme's history length) is a template parameter so that
            // F stays in registers
            double2 F[H];
            F[0] = f0[idx];
            F[1] = f1[idx];
            F[2] = f2[idx];
            if constexpr (H >= 4) F[3] = f3[idx];
            if constexpr (H >= 5) F[4] = f4[idx];
            double2 sb = make_double2(0, 0), sa = make_double2(0, 0);
#pragma unroll
            for (int j = 0; j < H; j++) {
                sb = sb + F[j] * cf.b[j];
                sa = sa + F[j] * cf.a[j];
            }
            const double2 Y0 = y0[idx];
            const double2 yh = (Y0 + imul(sb, -dt)) * (1.0 / np);
            const double2 cc = Y0 + imul(sa, -dt);
            const double2 G = imul(Y1 * nc - cc, 1.0 / (dt * cf.a0)) + yh * ep;
            w3[idx] = yh;
            w4[idx] = G;
            s += cmulc(mu, yh).x;
        }
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[(long)a * gridDim.x + blockIdx.x] = s;
}

// predictor normalisation back and the ABM transpose:
//   lyp = (w1 - t2 w3)/n_p ; lf_j <- lf_j (0 for the last) + i dt (b_j lyp + a_j ly) ; ly <- ly + lyp
__global__ void __launch_bounds__(AT)
k_hist_back(double2 *ly, const double2 *__restrict__ w1, const double2 *__restrict__ w3, double2 *lf0, double2 *lf1,
            double2 *lf2, double2 *lf3, double2 *lf4, const double *__restrict__ t2, const double *__restrict__ scal,
            AbmCoef cf, double dt, int n, long N)
{
    const long idx = (long)blockIdx.x * AT + threadIdx.x;
    if (idx >= 2L * n * N) return;
    const long a = (idx / N) % n;
    const double2 lyp = (w1[idx] - w3[idx] * t2[a]) * (1.0 / sqrt(scal[a]));
    const double2 lc = ly[idx];
    double2 *lf[5] = {lf0, lf1, lf2, lf3, lf4};
    for (int j = 0; j < cf.h; j++) {
        const double2 base = (j == cf.h - 1) ? make_double2(0.0, 0.0) : lf[j][idx];
        lf[j][idx] = base + imul(lyp * cf.b[j] + lc * cf.a[j], dt);
    }
    ly[idx] = lc + lyp;
}

// ---------------------------------------------------------------------------
// The density transpose fused into the pass that consumes it, for states
// small enough that one workgroup holds a whole state (N <= RT * KP): phase
// 1 forms g = out + dens^T(lD)(psi) in registers and reduces t = dV Re<psi, g>
// over the state; phase 2 uses g and t.  Same arithmetic as k_dens_t
// followed by k_recon / k_hist_back.
// ---------------------------------------------------------------------------
static constexpr int RT = 1024;

// per-state density weights of k_dens_part (occupation, transverse
// degeneracy, 1/transverse length): rho_a, rho_b and nu factors
__device__ __forceinline__ void dens_wts(int dim, int a, const double *__restrict__ fbeta,
                                         const double *__restrict__ kt, const int *__restrict__ cnt, double tr,
                                         int cubic, double piz, double &wa, double &wb, double &cn)
{
    double w = 1.0;
    if (dim == 1) w = (double)cnt[a];
    else if (dim == 2) {
        const double kz = kt[a];
        w = (fabs(kz) < 1.0e-12) ? 1.0 : 2.0;
        if (cubic && fabs(kz + piz) < 1.0e-12) w = 1.0;
    }
    const double fb = fbeta[a];
    wa = fb * w / tr;
    wb = (1.0 - fb) * w / tr;
    cn = (1.0 - 2.0 * fb) / 2.0 * w / tr;
}

// phase 1 for the points of one thread: gu/gv = out + dens^T(lD) at (u, v)
template <int KP>
__device__ __forceinline__ double dens_t_regs(const double2 *__restrict__ psi, const double2 *out,
                                              const double *__restrict__ lD, long ou, long ov, long N, double wa,
                                              double wb, double cn, double2 (&gu)[KP], double2 (&gv)[KP])
{
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < KP; k++) {
        const long i = threadIdx.x + (long)k * RT;
        if (i < N) {
            const double2 u = psi[ou + i], v = psi[ov + i];
            const double2 ln = make_double2(lD[2 * N + i], lD[3 * N + i]);
            gu[k] = out[ou + i] + u * (2.0 * wa * lD[i]) + cmul(ln, v) * cn;
            gv[k] = out[ov + i] + v * (2.0 * wb * lD[N + i]) + cmul(make_double2(ln.x, -ln.y), u) * cn;
            s += cmulc(u, gu[k]).x + cmulc(v, gv[k]).x;
        }
    }
    return s;
}

// k_dens_t(y1 -> ly) then k_recon, one workgroup per state; svp[a] is
// written directly (no partials)
template <int H, int KP>
__global__ void __launch_bounds__(RT)
k_recon_dt(double2 *ly, const double2 *__restrict__ y1, const double2 *__restrict__ y0,
           const double2 *__restrict__ f0, const double2 *__restrict__ f1, const double2 *__restrict__ f2,
           const double2 *__restrict__ f3, const double2 *__restrict__ f4, double2 *__restrict__ w0,
           double2 *__restrict__ w3, double2 *__restrict__ w4, const double *__restrict__ lD,
           const double *__restrict__ fbeta, const double *__restrict__ kt, const int *__restrict__ cnt, int dim,
           double tr, int cubic, double piz, const double *__restrict__ scal, AbmCoef cf, double dt, int n, long N,
           double dv, double *__restrict__ t1out, double *__restrict__ svp)
{
    __shared__ double red[RT / 64];
    const int a = blockIdx.x;
    const long ou = (long)a * N, ov = ((long)n + a) * N;
    double wa, wb, cn;
    dens_wts(dim, a, fbeta, kt, cnt, tr, cubic, piz, wa, wb, cn);
    double2 gu[KP], gv[KP];
    const double tt = dv * block_sum(dens_t_regs<KP>(y1, ly, lD, ou, ov, N, wa, wb, cn, gu, gv), red);
    if (threadIdx.x == 0) t1out[a] = tt;
    const double np = sqrt(scal[a]), ep = scal[n + a], nc = sqrt(scal[2 * n + a]);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < KP; k++) {
        const long i = threadIdx.x + (long)k * RT;
        if (i >= N) continue;
#pragma unroll
        for (int comp = 0; comp < 2; comp++) {
            const long idx = (comp ? ov : ou) + i;
            const double2 Y1 = y1[idx];
            const double2 lyc = ((comp ? gv[k] : gu[k]) - Y1 * tt) * (1.0 / nc);
            ly[idx] = lyc;
            const double2 mu = imul(lyc, dt * cf.a0);
            w0[idx] = mu;
            double2 F[H];
            F[0] = f0[idx];
            F[1] = f1[idx];
            F[2] = f2[idx];
            if constexpr (H >= 4) F[3] = f3[idx];
            if constexpr (H >= 5) F[4] = f4[idx];
            double2 sb = make_double2(0, 0), sa = make_double2(0, 0);
#pragma unroll
            for (int j = 0; j < H; j++) {
                sb = sb + F[j] * cf.b[j];
                sa = sa + F[j] * cf.a[j];
            }
            const double2 Y0 = y0[idx];
            const double2 yh = (Y0 + imul(sb, -dt)) * (1.0 / np);
            const double2 cc = Y0 + imul(sa, -dt);
            const double2 G = imul(Y1 * nc - cc, 1.0 / (dt * cf.a0)) + yh * ep;
            w3[idx] = yh;
            w4[idx] = G;
            s += cmulc(mu, yh).x;
        }
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) svp[a] = s * dv;
}

// k_dens_t(w3 -> w1) then k_hist_back, one workgroup per state
template <int H, int KP>
__global__ void __launch_bounds__(RT)
k_hist_dt(double2 *ly, const double2 *__restrict__ w1, const double2 *__restrict__ w3, double2 *lf0, double2 *lf1,
          double2 *lf2, double2 *lf3, double2 *lf4, const double *__restrict__ lD, const double *__restrict__ fbeta,
          const double *__restrict__ kt, const int *__restrict__ cnt, int dim, double tr, int cubic, double piz,
          const double *__restrict__ scal, AbmCoef cf, double dt, int n, long N, double dv)
{
    __shared__ double red[RT / 64];
    const int a = blockIdx.x;
    const long ou = (long)a * N, ov = ((long)n + a) * N;
    double wa, wb, cn;
    dens_wts(dim, a, fbeta, kt, cnt, tr, cubic, piz, wa, wb, cn);
    double2 gu[KP], gv[KP];
    const double t2 = dv * block_sum(dens_t_regs<KP>(w3, w1, lD, ou, ov, N, wa, wb, cn, gu, gv), red);
    const double inp = 1.0 / sqrt(scal[a]);
    double2 *lf[5] = {lf0, lf1, lf2, lf3, lf4};
#pragma unroll
    for (int k = 0; k < KP; k++) {
        const long i = threadIdx.x + (long)k * RT;
        if (i >= N) continue;
#pragma unroll
        for (int comp = 0; comp < 2; comp++) {
            const long idx = (comp ? ov : ou) + i;
            const double2 lyp = ((comp ? gv[k] : gu[k]) - w3[idx] * t2) * inp;
            const double2 lc = ly[idx];
#pragma unroll
            for (int j = 0; j < H; j++) {
                const double2 base = (j == H - 1) ? make_double2(0.0, 0.0) : lf[j][idx];
                lf[j][idx] = base + imul(lyp * cf.b[j] + lc * cf.a[j], dt);
            }
            ly[idx] = lc + lyp;
        }
    }
}

__global__ void k_copy_d(const double *__restrict__ s, double *__restrict__ d, long n)
{
    const long i = (long)blockIdx.x * AT + threadIdx.x;
    if (i < n) d[i] = s[i];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static long dot_chunk(const tdw_ctx *c, int *nchunk)
{
    const long N = (long)c->N;
    long chunk = 4096;
    if (chunk > N) chunk = N;
    *nchunk = (int)((N + chunk - 1) / chunk);
    return chunk;
}

static int sdot(tdw_ctx *c, int n, const double2 *x, const double2 *y, double scale, double *out)
{
    int nchunk;
    const long chunk = dot_chunk(c, &nchunk);
    TDW_TRY(ensure_part(c, (size_t)n * nchunk));
    hipLaunchKernelGGL(k_sdot_part, dim3(nchunk, n), dim3(AT), 0, c->stream, x, y, n, (long)c->N, chunk, c->d_part);
    hipLaunchKernelGGL(k_rows, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d_part, n, nchunk, scale, out);
    return (int)hipGetLastError();
}

static int cross_upd(tdw_ctx *c, int n, double2 *out, const double2 *acc, const double2 *hterm, const double2 *g1,
                     const double2 *g2, const double2 *mu, const double2 *psi, const double *sv, const double *ev,
                     const double *em, const double *add, double *lU)
{
    TDW_TIMED(c, "adj_cross");
    const long N = (long)c->N;
    const long nblk = (N + AT - 1) / AT;
    int chunk = 128;
    while (chunk > 8 && nblk * ((n + chunk - 1) / chunk) < 2048) chunk >>= 1;
    const int nchunk = (n + chunk - 1) / chunk;
    TDW_TRY(ensure_part(c, (size_t)nchunk * 4 * N));
    hipLaunchKernelGGL(k_cross_upd<2>, dim3(nblk, nchunk), dim3(AT), 0, c->stream, out, acc, hterm, g1, g2, mu, psi, sv,
                       ev, em, n, N, chunk, c->d_part);
    hipLaunchKernelGGL(k_cross_final, dim3((4 * N + 63) / 64), dim3(64), 0, c->stream, c->d_part, nchunk, N, add, lU);
    return (int)hipGetLastError();
}

static int pot_vjp(tdw_ctx *c, int it, const double *dens, const double *pots_prev, const double *lU,
                   const double *dens_add, double *lD, double *lUprev, double *theta_grad)
{
    TDW_TIMED(c, "adj_pot_vjp");
    const long N = (long)c->N;
    // one wave per workgroup: the kernel is register-bound (one wave per
    // SIMD), so small grids spread over more CUs
    constexpr int PT = 64;
    const int nblk = (int)((N + PT - 1) / PT);
    TDW_TRY(ensure_part(c, (size_t)nblk * NTHETA));
    hipLaunchKernelGGL(k_pot_vjp, dim3(nblk), dim3(PT), 0, c->stream, fparams(c), make_vext(c, it), N, dens,
                       pots_prev, lU, dens_add, lD, lUprev, c->d_part);
    hipLaunchKernelGGL(k_gsum, dim3(NTHETA), dim3(AT), 0, c->stream, c->d_part, nblk, theta_grad);
    return (int)hipGetLastError();
}

static int dens_t(tdw_ctx *c, int n, const double2 *psi, double2 *out, const double *lD, double *tvec)
{
    TDW_TIMED(c, "adj_dens_t");
    int nchunk;
    const long chunk = dot_chunk(c, &nchunk);
    TDW_TRY(ensure_part(c, (size_t)n * nchunk));
    const double tr = dens_transverse(c), piz = M_PI / c->p.dz;
    const int cubic = c->p.regularization == TDW_CUBIC_CUTOFF;
    const long N = (long)c->N;
    dim3 g(nchunk, n);
    if (c->dim == 1)
        hipLaunchKernelGGL(k_dens_t<1>, g, dim3(AT), 0, c->stream, psi, out, lD, c->d_fbeta, c->d_kt, c->d_cnt, n, N,
                           chunk, tr, cubic, piz, c->d_part);
    else if (c->dim == 2)
        hipLaunchKernelGGL(k_dens_t<2>, g, dim3(AT), 0, c->stream, psi, out, lD, c->d_fbeta, c->d_kt, c->d_cnt, n, N,
                           chunk, tr, cubic, piz, c->d_part);
    else
        hipLaunchKernelGGL(k_dens_t<3>, g, dim3(AT), 0, c->stream, psi, out, lD, c->d_fbeta, c->d_kt, c->d_cnt, n, N,
                           chunk, tr, cubic, piz, c->d_part);
    hipLaunchKernelGGL(k_rows, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d_part, n, nchunk, dvol_norm(c),
                       tvec);
    return (int)hipGetLastError();
}

static int copy_d(tdw_ctx *c, const double *s, double *d, long n)
{
    hipLaunchKernelGGL(k_copy_d, dim3((n + AT - 1) / AT), dim3(AT), 0, c->stream, s, d, n);
    return (int)hipGetLastError();
}

// points per thread of the fused one-workgroup-per-state kernels; 0: the
// state does not fit one workgroup (the separate passes run)
static int fused_kp(long N)
{
    for (int kp : {1, 2, 4, 8})
        if (N <= (long)RT * kp) return kp;
    return 0;
}

typedef void (*recon_dt_t)(double2 *, const double2 *, const double2 *, const double2 *, const double2 *,
                           const double2 *, const double2 *, const double2 *, double2 *, double2 *, double2 *,
                           const double *, const double *, const double *, const int *, int, double, int, double,
                           const double *, AbmCoef, double, int, long, double, double *, double *);
typedef void (*hist_dt_t)(double2 *, const double2 *, const double2 *, double2 *, double2 *, double2 *, double2 *,
                          double2 *, const double *, const double *, const double *, const int *, int, double, int,
                          double, const double *, AbmCoef, double, int, long, double);

template <int H> static recon_dt_t recon_dt_kp(int kp)
{
    return kp == 1 ? k_recon_dt<H, 1> : kp == 2 ? k_recon_dt<H, 2> : kp == 4 ? k_recon_dt<H, 4> : k_recon_dt<H, 8>;
}
static recon_dt_t recon_dt_fn(int h, int kp)
{
    return h == 3 ? recon_dt_kp<3>(kp) : h == 4 ? recon_dt_kp<4>(kp) : recon_dt_kp<5>(kp);
}
template <int H> static hist_dt_t hist_dt_kp(int kp)
{
    return kp == 1 ? k_hist_dt<H, 1> : kp == 2 ? k_hist_dt<H, 2> : kp == 4 ? k_hist_dt<H, 4> : k_hist_dt<H, 8>;
}
static hist_dt_t hist_dt_fn(int h, int kp)
{
    return h == 3 ? hist_dt_kp<3>(kp) : h == 4 ? hist_dt_kp<4>(kp) : hist_dt_kp<5>(kp);
}

static int adj_supported(const tdw_ctx *c)
{
    // the backward step differentiates the SLDA step as written: no
    // spin-symmetric folding and no mean-field filter between the functional
    // and H
    // and no problem callbacks of a project build (the gradient is taken
    // with respect to the built-in V_ext parameters)
    return c->p.functional == TDW_SLDA && c->p.precision == TDW_FP64 && !c->p.spinsymmetry && c->p.hkf_mode < 1 &&
           !c->proj;
}

}  // namespace tdw

using namespace tdw;

#define D2(p) ((double2 *)(p))
#define CD2(p) ((const double2 *)(p))

// The taped step.  y_out: where the new y goes (NULL: in place); cbuf: the
// partial corrector's buffer (NULL: the oldest history entry, as abm_step1
// does); f_next: the new history entry's buffer (NULL: the dropped oldest
// entry).  With y_out, cbuf and f_next the step writes no history entry and
// leaves y at its start unchanged.
static int adj_forward(tdw_ctx *c, tdw_state *s, int it, tdw_complex *y_out, double2 *cbuf, tdw_complex *f_next,
                       tdw_tape_step *t)
{
    TDW_TRY(fused_finalize(c, s));
    const int n = s->n;
    const AbmCoef cf = abm_coef(c->p.scheme);
    const int k = cf.h;
    const long N = (long)c->N;
    double *qpe = s->qpe ? s->qpe : c->d_qpe;
    // the call sequence of tdw_step (pca.c:1401-1570), non-gradient form;
    // the densities of a non-gradient step do not read the Laplacian, and 2D
    // lattices with a plane kernel form it inside the H application
    const bool plane = h2d_supported(c);
    const bool fused = !plane && fused_supported(c);
    if (fused && s->ddx && s->ddy) {
        // 3D lattices of the fused step: its passes (finalised at the end of
        // the step), with the tape read from its own per-state scalars
        TDW_TRY(fused_adj_step(c, s, it, y_out, cbuf, f_next, t));
        TDW_HIP(hipMemcpyAsync(qpe, t->scal + 3 * n, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
        return 0;
    }
    if (y_out) {
        TDW_TRY(abm1_to(c, n, CD2(s->wf), D2(y_out), cbuf, D2(s->f[0]), D2(s->f[1]), D2(s->f[2]), D2(s->f[3]),
                        D2(s->f[4])));
        s->wf = y_out;
    } else {
        TDW_TRY(abm1_to(c, n, CD2(s->wf), D2(s->wf), cbuf, D2(s->f[0]), D2(s->f[1]), D2(s->f[2]), D2(s->f[3]),
                        D2(s->f[4])));
    }
    const double2 *flast = cbuf ? cbuf : CD2(s->f[k - 1]);
    // 3D lattices of the fused step: H through its x and yz passes, with
    // ddy as the D_x scratch (the Laplacian array is then not formed)
    double2 *hT = (fused && s->ddy) ? D2(s->ddy) : nullptr;
    TDW_TRY(normalize(c, n, D2(s->wf), t->scal));
    if (!plane && !hT) TDW_TRY(laplace(c, 2 * n, D2(s->wf), D2(s->lap)));
    TDW_TRY(densities(c, n, D2(s->wf), nullptr, nullptr, nullptr, D2(s->lap), c->d_fbeta, nullptr, c->d_kt, c->d_cnt,
                      s->dens, 0));
    TDW_TRY(copy_d(c, s->dens, t->dens_p, 12 * N));
    TDW_TRY(potentials(c, it + 1, s->dens, s->pots));
    TDW_TRY(copy_d(c, s->pots, t->pots_p, 12 * N));
    if (plane) TDW_TRY(apply_h_plane(c, it + 1, n, D2(s->wf), D2(s->lap), s->pots, c->d_kt, nullptr, t->scal + n));
    else if (hT) TDW_TRY(fused_apply_h(c, n, D2(s->wf), D2(s->lap), hT, s->pots, t->scal + n));
    else TDW_TRY(apply_h(c, it + 1, n, D2(s->wf), D2(s->lap), CD2(s->lap), s->pots, c->d_kt, nullptr, t->scal + n));
    TDW_TRY(abm4(c, n, CD2(s->lap), D2(s->wf), flast));
    TDW_TRY(normalize(c, n, D2(s->wf), t->scal + 2 * n));
    if (!plane && !hT) TDW_TRY(laplace(c, 2 * n, D2(s->wf), D2(s->lap)));
    TDW_TRY(densities(c, n, D2(s->wf), nullptr, nullptr, nullptr, D2(s->lap), c->d_fbeta, nullptr, c->d_kt, c->d_cnt,
                      s->dens, 0));
    TDW_TRY(copy_d(c, s->dens, t->dens_c, 12 * N));
    TDW_TRY(potentials(c, it + 1, s->dens, s->pots));
    TDW_TRY(copy_d(c, s->pots, t->pots_c, 12 * N));
    tdw_complex *tmp = f_next ? f_next : s->f[k - 1];
    for (int j = k - 1; j > 0; j--) s->f[j] = s->f[j - 1];
    s->f[0] = tmp;
    if (plane)
        TDW_TRY(apply_h_plane(c, it + 1, n, D2(s->wf), D2(s->f[0]), s->pots, c->d_kt, nullptr, t->scal + 3 * n));
    else if (hT)
        TDW_TRY(fused_apply_h(c, n, D2(s->wf), D2(s->f[0]), hT, s->pots, t->scal + 3 * n));
    else
        TDW_TRY(apply_h(c, it + 1, n, D2(s->wf), D2(s->f[0]), CD2(s->lap), s->pots, c->d_kt, nullptr, t->scal + 3 * n));
    TDW_HIP(hipMemcpyAsync(qpe, t->scal + 3 * n, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

extern "C" int tdw_adj_forward_step(tdw_ctx *c, tdw_state *s, int it, tdw_complex *y_copy, tdw_complex *fnew_copy,
                                    tdw_tape_step *t)
{
    if (!c || !s || !t || s->n > c->nwf || t->n != s->n) return TDW_EINVAL;
    if (!adj_supported(c)) return TDW_EINVAL;
    TDW_TRY(fused_finalize(c, s));
    const size_t sbytes = sizeof(double2) * 2 * (size_t)s->n * c->N;
    if (y_copy) TDW_HIP(hipMemcpyAsync(y_copy, s->wf, sbytes, hipMemcpyDeviceToDevice, c->stream));
    TDW_TRY(adj_forward(c, s, it, nullptr, nullptr, nullptr, t));
    if (fnew_copy) TDW_HIP(hipMemcpyAsync(fnew_copy, s->f[0], sbytes, hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

extern "C" int tdw_adj_forward_step_to(tdw_ctx *c, tdw_state *s, int it, tdw_complex *y_next, tdw_complex *f_next,
                                       tdw_tape_step *t)
{
    if (!c || !s || !t || !y_next || !f_next || !s->ddx || s->n > c->nwf || t->n != s->n) return TDW_EINVAL;
    if (!adj_supported(c)) return TDW_EINVAL;
    const int h = abm_coef(c->p.scheme).h;
    // every buffer the step writes must differ from those it reads (ddy:
    // the D_x scratch of 3D lattices with the fused passes)
    const tdw_complex *bufs[5] = {y_next, f_next, s->ddx, s->lap, s->ddy};
    const int nb = s->ddy ? 5 : 4;
    for (int i = 0; i < nb; i++) {
        if (bufs[i] == s->wf) return TDW_EINVAL;
        for (int j = 0; j < h; j++)
            if (bufs[i] == s->f[j]) return TDW_EINVAL;
        for (int j = i + 1; j < nb; j++)
            if (bufs[i] == bufs[j]) return TDW_EINVAL;
    }
    return adj_forward(c, s, it, y_next, D2(s->ddx), f_next, t);
}

// The history entry taped step `it` left, rebuilt from y after the step and
// the step's tape: H(pots_c) y1 with the H application that step used, minus
// the <H> it recorded (scal[3n..4n)).  2D plane lattices subtract inside the
// plane pass, the others with the step's own subtraction kernel.
extern "C" int tdw_adj_history_entry(tdw_ctx *c, int it, const tdw_tape_step *t, const tdw_complex *y1,
                                     tdw_complex *fnew, tdw_complex *scratch)
{
    if (!c || !t || !y1 || !fnew || !scratch || t->n <= 0 || t->n > c->nwf) return TDW_EINVAL;
    if (fnew == y1 || scratch == y1 || scratch == fnew) return TDW_EINVAL;
    if (!adj_supported(c)) return TDW_EINVAL;
    const int n = t->n;
    const double *e_c = t->scal + 3 * n;
    const bool plane = h2d_supported(c);
    const bool fused = !plane && fused_supported(c);
    if (plane) return apply_h_plane(c, it + 1, n, CD2(y1), D2(fnew), t->pots_c, c->d_kt, e_c, nullptr);
    if (fused) return fused_apply_h(c, n, CD2(y1), D2(fnew), D2(scratch), t->pots_c, nullptr, e_c);
    TDW_TRY(laplace(c, 2 * n, CD2(y1), D2(scratch)));
    return apply_h(c, it + 1, n, CD2(y1), D2(fnew), CD2(scratch), t->pots_c, c->d_kt, e_c, nullptr);
}

extern "C" int tdw_adj_backward_step(tdw_ctx *c, tdw_adj_state *A, int it, const tdw_complex *y0,
                                     const tdw_complex *const *f, const tdw_complex *y1, const tdw_complex *fnew,
                                     const tdw_tape_step *t, const double *pots_prev, const double *dens_grad)
{
    if (!c || !A || !t || !f || A->n > c->nwf || t->n != A->n) return TDW_EINVAL;
    if (!adj_supported(c)) return TDW_EINVAL;
    const int n = A->n;
    const long N = (long)c->N;
    const AbmCoef cf = abm_coef(c->p.scheme);
    const int h = cf.h;
    for (int j = 0; j < h; j++)
        if (!f[j] || !A->lf[j]) return TDW_EINVAL;
    const double dv = dvol_norm(c);
    if (!c->d_adjv) TDW_HIP(hipMalloc(&c->d_adjv, sizeof(double) * 4 * (size_t)c->nwf));
    double *sv1 = c->d_adjv, *t1 = c->d_adjv + c->nwf, *svp = c->d_adjv + 2 * c->nwf, *t2 = c->d_adjv + 3 * c->nwf;
    const double *nrm_p = t->scal, *e_p = t->scal + n, *e_n = t->scal + 3 * n;
    (void)nrm_p;
    double2 *LY = D2(A->ly), *W0 = D2(A->w[0]), *W1 = D2(A->w[1]), *W2 = D2(A->w[2]), *W3 = D2(A->w[3]),
            *W4 = D2(A->w[4]);
    double2 *MU = D2(A->lf[0]);
    double *lUc = A->lgrid, *lD = A->lgrid + 4 * N, *lUo = A->lgrid + 8 * N;
    TimerScope ts(c, "adj_step");

    // history refresh f_new = H(U') y' - e' y'  (mu = dL/df_new)
    const bool plane = h2d_supported(c);
    const bool fused = !plane && fused_supported(c);
    if (plane) TDW_TRY(apply_h_only_plane(c, n, MU, W1, t->pots_c, c->d_kt, CD2(y1), dv, sv1));
    else {
        TDW_TRY(sdot(c, n, MU, CD2(y1), dv, sv1));
        if (fused) TDW_TRY(fused_apply_h(c, n, MU, W1, W2, t->pots_c, nullptr));
        else {
            TDW_TRY(laplace(c, 2 * n, MU, W2));
            TDW_TRY(apply_h_only(c, n, MU, W1, W2, t->pots_c, c->d_kt));
        }
    }
    TDW_TRY(cross_upd(c, n, LY, LY, W1, CD2(fnew), CD2(y1), MU, CD2(y1), sv1, e_n, e_n, A->lpots, lUc));
    // U' = Phi(D', U_p, V_ext)
    TDW_TRY(pot_vjp(c, it + 1, t->dens_c, t->pots_p, lUc, dens_grad, lD, lUo, A->theta_grad));
    // D' = dens(y'), then y' = y_c / n_c; rebuild the predictor side: LY
    // becomes dL/dy_c = dL/dc, W0 = dL/dF_p
    const int kp = fused_kp(N);
    const double tr = dens_transverse(c), piz = M_PI / c->p.dz;
    const int cubic = c->p.regularization == TDW_CUBIC_CUTOFF;
    if (kp) {
        TDW_TIMED(c, "adj_recon");
        const double2 *F[5] = {CD2(f[0]), CD2(f[1]), CD2(f[2]), h >= 4 ? CD2(f[3]) : nullptr,
                               h >= 5 ? CD2(f[4]) : nullptr};
        auto kr = recon_dt_fn(h, kp);
        hipLaunchKernelGGL(kr, dim3(n), dim3(RT), 0, c->stream, LY, CD2(y1), CD2(y0), F[0], F[1], F[2], F[3], F[4],
                           W0, W3, W4, lD, c->d_fbeta, c->d_kt, c->d_cnt, c->dim, tr, cubic, piz, t->scal, cf,
                           c->p.dt, n, N, dv, t1, svp);
        TDW_HIP(hipGetLastError());
    } else {
        TDW_TRY(dens_t(c, n, CD2(y1), LY, lD, t1));
        int nchunk;
        const long chunk = dot_chunk(c, &nchunk);
        TDW_TRY(ensure_part(c, (size_t)n * nchunk));
        const double2 *F[5] = {CD2(f[0]), CD2(f[1]), CD2(f[2]), h >= 4 ? CD2(f[3]) : nullptr,
                               h >= 5 ? CD2(f[4]) : nullptr};
        TDW_TIMED(c, "adj_recon");
        auto kr = h == 3 ? k_recon<3> : h == 4 ? k_recon<4> : k_recon<5>;
        hipLaunchKernelGGL(kr, dim3(nchunk, n), dim3(AT), 0, c->stream, LY, CD2(y1), CD2(y0), F[0], F[1], F[2], F[3],
                           F[4], W0, W3, W4, t1, t->scal, cf, c->p.dt, n, N, chunk, c->d_part);
        hipLaunchKernelGGL(k_rows, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d_part, n, nchunk, dv, svp);
        TDW_HIP(hipGetLastError());
    }
    // F_p = H(U_p) yh_p - e_p yh_p
    if (plane) TDW_TRY(apply_h_only_plane(c, n, W0, W1, t->pots_p, c->d_kt));
    else if (fused) TDW_TRY(fused_apply_h(c, n, W0, W1, W2, t->pots_p, nullptr));
    else {
        TDW_TRY(laplace(c, 2 * n, W0, W2));
        TDW_TRY(apply_h_only(c, n, W0, W1, W2, t->pots_p, c->d_kt));
    }
    TDW_TRY(cross_upd(c, n, W1, nullptr, W1, W4, nullptr, W0, W3, svp, nullptr, e_p, lUo, lUc));
    // U_p = Phi(D_p, U_prev, V_ext)
    TDW_TRY(pot_vjp(c, it + 1, t->dens_p, pots_prev, lUc, nullptr, lD, A->lpots, A->theta_grad));
    // yh_p = y_p / n_p, then the ABM predictor / partial corrector
    if (!kp) TDW_TRY(dens_t(c, n, W3, W1, lD, t2));
    tdw_complex *mu_store = A->lf[0];
    for (int j = 0; j < h - 1; j++) A->lf[j] = A->lf[j + 1];
    A->lf[h - 1] = mu_store;
    if (kp) {
        TDW_TIMED(c, "adj_hist");
        auto kh = hist_dt_fn(h, kp);
        hipLaunchKernelGGL(kh, dim3(n), dim3(RT), 0, c->stream, LY, W1, W3, D2(A->lf[0]), D2(A->lf[1]), D2(A->lf[2]),
                           h >= 4 ? D2(A->lf[3]) : nullptr, h >= 5 ? D2(A->lf[4]) : nullptr, lD, c->d_fbeta,
                           c->d_kt, c->d_cnt, c->dim, tr, cubic, piz, t->scal, cf, c->p.dt, n, N, dv);
        TDW_HIP(hipGetLastError());
    } else {
        TDW_TIMED(c, "adj_hist");
        const long tot = 2L * n * N;
        hipLaunchKernelGGL(k_hist_back, dim3((tot + AT - 1) / AT), dim3(AT), 0, c->stream, LY, W1, W3, D2(A->lf[0]),
                           D2(A->lf[1]), D2(A->lf[2]), h >= 4 ? D2(A->lf[3]) : nullptr,
                           h >= 5 ? D2(A->lf[4]) : nullptr, t2, t->scal, cf, c->p.dt, n, N);
        TDW_HIP(hipGetLastError());
    }
    return 0;
}
