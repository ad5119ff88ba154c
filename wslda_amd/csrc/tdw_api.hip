// tdw_api.hip -- host side: context, FFT plans, the generic tdw_* layer, the
// modular PECE step and the reference-named drop-in entry points.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "tdw_internal.hpp"
#include "tdw_physics.hpp"
#include "tdw_fused.hpp"

using namespace tdw;

namespace tdw {

static hipEvent_t ev_get(tdw_ctx *c)
{
    if (!c->evpool) c->evpool = new std::vector<hipEvent_t>();
    if (!c->evpool->empty()) {
        hipEvent_t e = c->evpool->back();
        c->evpool->pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
}

int timer_begin(tdw_ctx *c, const char *name)
{
    if (!c->timers) c->timers = new std::vector<tdw_timer_rec>();
    tdw_timer_rec r;
    r.name = name;
    r.a = ev_get(c);
    r.b = ev_get(c);
    hipEventRecord(r.a, c->stream);
    c->timers->push_back(r);
    return (int)c->timers->size() - 1;
}

void timer_end(tdw_ctx *c, int idx)
{
    if (!c->timers || idx < 0 || idx >= (int)c->timers->size()) return;
    hipEventRecord((*c->timers)[idx].b, c->stream);
}

int ensure_part(tdw_ctx *c, size_t nd)
{
    if (nd <= c->part_cap) return 0;
    TDW_HIP(hipStreamSynchronize(c->stream));   // in-flight kernels may still read it
    if (c->d_part) TDW_HIP(hipFree(c->d_part));
    c->d_part = nullptr;
    c->part_cap = 0;
    TDW_HIP(hipMalloc(&c->d_part, nd * sizeof(double)));
    c->part_cap = nd;
    return 0;
}

int ensure_T(tdw_ctx *c, size_t ne)
{
    if (ne <= c->T_cap) return 0;
    TDW_HIP(hipStreamSynchronize(c->stream));
    if (c->d_T) TDW_HIP(hipFree(c->d_T));
    c->d_T = nullptr;
    c->T_cap = 0;
    TDW_HIP(hipMalloc(&c->d_T, ne * sizeof(double2)));
    c->T_cap = ne;
    return 0;
}

// Laplacian of m arrays.  1D: one outer pass.  2D planes that fit in LDS:
// one pass.  Otherwise outer pass into T (the output array when it differs
// from the input), inner pass into out.
int laplace(tdw_ctx *c, int m, const double2 *in, double2 *out)
{
    if (m <= 0) return 0;
    if (c->dim == 1) return launch_outer(c, m, in, out, nullptr, 0);
    if (c->dim == 2) {
        // one pass per plane where the plane fits in LDS (in place allowed)
        int done = 0;
        TDW_TRY(lap2d(c, m, in, out, &done));
        if (done) return 0;
    }
    if (out != in) {
        TDW_TRY(launch_outer(c, m, in, out, nullptr, 0));
        return launch_inner(c, m, in, out, out, nullptr, nullptr, 0);
    }
    // in place: batch through the scratch array
    const size_t N = c->N;
    int batch = m;
    const size_t cap = (size_t)1 << 26;   // 64 Mi complex = 1 GiB
    if ((size_t)batch * N > cap) batch = (int)(cap / N) > 0 ? (int)(cap / N) : 1;
    TDW_TRY(ensure_T(c, (size_t)batch * N));
    for (int a0 = 0; a0 < m; a0 += batch) {
        int mm = (m - a0 < batch) ? m - a0 : batch;
        TDW_TRY(launch_outer(c, mm, in + (size_t)a0 * N, c->d_T, nullptr, 0));
        TDW_TRY(launch_inner(c, mm, in + (size_t)a0 * N, c->d_T, out + (size_t)a0 * N, nullptr, nullptr, 0));
    }
    return 0;
}

int derivatives(tdw_ctx *c, int m, const double2 *in, double2 *ddx, double2 *ddy, double2 *ddz, double2 *lap)
{
    if (m <= 0) return 0;
    if (lap == in || ddx == in) return TDW_EINVAL;
    TDW_TRY(launch_outer(c, m, in, lap, ddx, 1));
    if (c->dim == 1) return 0;
    return launch_inner(c, m, in, lap, lap, ddy, c->dim == 3 ? ddz : nullptr, 1);
}

}  // namespace tdw

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
static int plan_axis(int L, tdw_axis *a)
{
    a->L = L;
    a->nst = 0;
    int rem = L, Ns = 1;
    while (rem > 1) {
        int R;
        if (rem % 8 == 0) R = 8;
        else if (rem % 4 == 0) R = 4;
        else if (rem % 3 == 0) R = 3;
        else if (rem % 2 == 0) R = 2;
        else return TDW_ESIZE;
        if (a->nst >= TDW_MAXST) return TDW_ESIZE;
        a->R[a->nst] = R;
        a->Ns[a->nst] = Ns;
        a->nst++;
        Ns *= R;
        rem /= R;
    }
    return 0;
}

static int upload_twiddles(tdw_ctx *c, int k, int L)
{
    double2 *h = (double2 *)malloc(sizeof(double2) * L);
    float2 *hf = (float2 *)malloc(sizeof(float2) * L);
    for (int t = 0; t < L; t++) {
        double a = 2.0 * M_PI * (double)t / (double)L;
        h[t] = make_double2(cos(a), -sin(a));
        hf[t] = make_float2((float)cos(a), (float)-sin(a));
    }
    int e = 0;
    if (hipMalloc(&c->d_tw[k], sizeof(double2) * L) != hipSuccess) e = TDW_ENOMEM;
    if (!e && hipMalloc(&c->d_tw32[k], sizeof(float2) * L) != hipSuccess) e = TDW_ENOMEM;
    if (!e) e = (int)hipMemcpy(c->d_tw[k], h, sizeof(double2) * L, hipMemcpyHostToDevice);
    if (!e) e = (int)hipMemcpy(c->d_tw32[k], hf, sizeof(float2) * L, hipMemcpyHostToDevice);
    free(h);
    free(hf);
    return e;
}

extern "C" int tdw_create(const tdw_params *p, int nwf, int device, tdw_ctx **out)
{
    if (!p || !out || nwf <= 0) return TDW_EINVAL;
    if (p->dim < 1 || p->dim > 3 || p->nx < 2) return TDW_EINVAL;
    if (p->functional != TDW_SLDA && p->functional != TDW_ASLDA && p->functional != TDW_BDG &&
        p->functional != TDW_SLDAE)
        return TDW_EINVAL;
    if (p->scheme != TDW_AB3AM4 && p->scheme != TDW_AB4AM5 && p->scheme != TDW_AB5AM5) return TDW_EINVAL;
    TDW_HIP(hipSetDevice(device));
    tdw_ctx *c = (tdw_ctx *)calloc(1, sizeof(tdw_ctx));
    if (!c) return TDW_ENOMEM;
    c->p = *p;
    c->dim = p->dim;
    c->nx = p->nx;
    c->ny = p->dim >= 2 ? p->ny : 1;
    c->nz = p->dim >= 3 ? p->nz : 1;
    c->N = (size_t)c->nx * c->ny * c->nz;
    c->S = (int)(c->N / c->nx);
    c->nwf = nwf;
    c->device = device;
    c->Lphys[0] = p->nx * p->dx;
    c->Lphys[1] = p->ny * p->dy;
    c->Lphys[2] = p->nz * p->dz;
    int e = plan_axis(c->nx, &c->ax[0]);
    if (!e && c->dim >= 2) e = plan_axis(c->ny, &c->ax[1]);
    if (!e && c->dim >= 3) e = plan_axis(c->nz, &c->ax[2]);
    if (!e) e = (int)hipDeviceGetAttribute(&c->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
    if (!e && c->max_lds > 160 * 1024) c->max_lds = 160 * 1024;
    if (!e) e = (int)hipStreamCreate(&c->stream);
    if (!e) e = upload_twiddles(c, 0, c->nx);
    if (!e && c->dim >= 2) e = upload_twiddles(c, 1, c->ny);
    if (!e && c->dim >= 3) e = upload_twiddles(c, 2, c->nz);
    if (!e) e = (int)hipMalloc(&c->d_fbeta, sizeof(double) * nwf);
    if (!e) e = (int)hipMalloc(&c->d_kt, sizeof(double) * 2 * nwf);
    if (!e) e = (int)hipMalloc(&c->d_cnt, sizeof(int) * nwf);
    if (!e) e = (int)hipMalloc(&c->d_nrm, sizeof(double) * nwf);
    if (!e) e = (int)hipMalloc(&c->d_qpe, sizeof(double) * nwf);
    if (!e) e = (int)hipMalloc(&c->d_scale, sizeof(double) * nwf);
    if (!e) e = (int)hipMalloc(&c->d_esum, sizeof(double) * 64);
    if (!e) e = (int)hipMalloc(&c->d_rtmp, sizeof(double2) * c->N);
    if (!e) e = (int)hipMalloc(&c->d_vext, sizeof(double) * 2 * c->N);
    if (!e) e = (int)hipMemset(c->d_kt, 0, sizeof(double) * 2 * nwf);
    if (!e) e = noalloc_setup(device);
    if (!e) {
        int *ones = (int *)malloc(sizeof(int) * nwf);
        for (int i = 0; i < nwf; i++) ones[i] = 1;
        e = (int)hipMemcpy(c->d_cnt, ones, sizeof(int) * nwf, hipMemcpyHostToDevice);
        free(ones);
    }
    if (!e) e = project_attach(c);
    if (e) {
        tdw_destroy(c);
        return e;
    }
    *out = c;
    return 0;
}

extern "C" int tdw_destroy(tdw_ctx *c)
{
    if (!c) return 0;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (int k = 0; k < 3; k++) {
        if (c->d_tw[k]) hipFree(c->d_tw[k]);
        if (c->d_tw32[k]) hipFree(c->d_tw32[k]);
    }
    void *bufs[] = {c->d_fbeta, c->d_kt, c->d_cnt, c->d_part, c->d_nrm, c->d_qpe, c->d_esum,
                    c->d_T, c->d_rtmp, c->d_vext, c->d_scale, c->d_adjv, c->d_hx, c->d_ascr, c->d_aw,
                    c->d_gscr};
    for (void *b : bufs)
        if (b) hipFree(b);
    project_release(c);
    fused_release(c);
    if (c->timers) {
        for (auto &r : *c->timers) { hipEventDestroy(r.a); hipEventDestroy(r.b); }
        delete c->timers;
    }
    if (c->evpool) {
        for (auto e : *c->evpool) hipEventDestroy(e);
        delete c->evpool;
    }
    if (c->stream) hipStreamDestroy(c->stream);
    if (tdw_bound() == c) tdw_bind(nullptr);
    free(c);
    return 0;
}

extern "C" void *tdw_stream(tdw_ctx *c) { return c ? (void *)c->stream : nullptr; }

extern "C" int tdw_timing(tdw_ctx *c, int enable)
{
    if (!c) return TDW_EINVAL;
    c->timing = enable ? 1 : 0;
    return 0;
}

extern "C" int tdw_timing_report(tdw_ctx *c, int max, char *names, double *total_ms, int *count)
{
    if (!c) return TDW_EINVAL;
    TDW_HIP(hipStreamSynchronize(c->stream));
    int nout = 0;
    if (c->timers) {
        for (auto &r : *c->timers) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) {
                (void)hipGetLastError();   // do not leave the query's error behind
                ms = 0.f;
            }
            int k;
            for (k = 0; k < nout; k++)
                if (strncmp(names + 32 * k, r.name, 31) == 0) break;
            if (k == nout) {
                if (nout >= max) continue;
                strncpy(names + 32 * k, r.name, 31);
                names[32 * k + 31] = 0;
                total_ms[k] = 0.0;
                count[k] = 0;
                nout++;
            }
            total_ms[k] += ms;
            count[k] += 1;
            c->evpool->push_back(r.a);
            c->evpool->push_back(r.b);
        }
        c->timers->clear();
    }
    return nout;
}

extern "C" int tdw_sync(tdw_ctx *c)
{
    if (!c) return TDW_EINVAL;
    TDW_HIP(hipStreamSynchronize(c->stream));
    return (int)hipGetLastError();
}

extern "C" size_t tdw_npts(const tdw_ctx *c) { return c ? c->N : 0; }

extern "C" int tdw_set_params(tdw_ctx *c, const tdw_params *p)
{
    if (!c || !p) return TDW_EINVAL;
    if (p->dim != c->p.dim || p->nx != c->p.nx || (c->dim >= 2 && p->ny != c->p.ny) ||
        (c->dim >= 3 && p->nz != c->p.nz))
        return TDW_EINVAL;
    if (c->proj && p->functional != c->p.functional) return TDW_EINVAL;
    c->p = *p;
    return project_consts(c);
}

extern "C" int tdw_set_states(tdw_ctx *c, const double *h_fbeta, const double *h_kt, const int *h_cnt)
{
    if (!c) return TDW_EINVAL;
    const int n = c->nwf;
    if (h_fbeta) TDW_HIP(hipMemcpy(c->d_fbeta, h_fbeta, sizeof(double) * n, hipMemcpyHostToDevice));
    if (h_kt) TDW_HIP(hipMemcpy(c->d_kt, h_kt, sizeof(double) * (c->dim == 1 ? 2 * n : n), hipMemcpyHostToDevice));
    if (h_cnt) TDW_HIP(hipMemcpy(c->d_cnt, h_cnt, sizeof(int) * n, hipMemcpyHostToDevice));
    return 0;
}

#define D2(x) ((double2 *)(x))
#define CD2(x) ((const double2 *)(x))

extern "C" int tdw_abm_step1(tdw_ctx *c, int n, tdw_complex *y, tdw_complex *f1, tdw_complex *f2,
                             tdw_complex *f3, tdw_complex *f4, tdw_complex *f5)
{
    if (!c) return TDW_EINVAL;
    return abm1(c, n, D2(y), D2(f1), D2(f2), D2(f3), D2(f4), D2(f5));
}
extern "C" int tdw_abm_step4(tdw_ctx *c, int n, const tdw_complex *hy, tdw_complex *y, const tdw_complex *fl)
{
    if (!c) return TDW_EINVAL;
    return abm4(c, n, CD2(hy), D2(y), CD2(fl));
}
extern "C" int tdw_normalize(tdw_ctx *c, int n, tdw_complex *wf, double *nrm)
{
    if (!c || n > c->nwf) return TDW_EINVAL;
    return normalize(c, n, D2(wf), nrm ? nrm : c->d_nrm);
}
extern "C" int tdw_laplace(tdw_ctx *c, int m, const tdw_complex *in, tdw_complex *out)
{
    if (!c) return TDW_EINVAL;
    return laplace(c, m, CD2(in), D2(out));
}
extern "C" int tdw_derivatives(tdw_ctx *c, int m, const tdw_complex *in, tdw_complex *ddx, tdw_complex *ddy,
                               tdw_complex *ddz, tdw_complex *lap)
{
    if (!c) return TDW_EINVAL;
    return derivatives(c, m, CD2(in), D2(ddx), D2(ddy), D2(ddz), D2(lap));
}
extern "C" int tdw_laplace_real(tdw_ctx *c, const double *f, double *lap)
{
    if (!c) return TDW_EINVAL;
    return laplace_real(c, f, lap);
}
extern "C" int tdw_densities(tdw_ctx *c, int n, const tdw_complex *wf, const tdw_complex *ddx,
                             const tdw_complex *ddy, const tdw_complex *ddz, const tdw_complex *lap,
                             const double *weights, double *dens, int gradients)
{
    if (!c || n > c->nwf) return TDW_EINVAL;
    if (gradients && !ddx && !ddy && !ddz && !lap) {
        // gradients formed inside the density passes
        if (!dens_grad_fused_supported(c)) return TDW_EINVAL;
        return dens_grad_fused(c, n, CD2(wf), c->d_fbeta, weights, dens);
    }
    return densities(c, n, CD2(wf), CD2(ddx), CD2(ddy), CD2(ddz), CD2(lap), c->d_fbeta, weights, c->d_kt,
                     c->d_cnt, dens, gradients);
}
extern "C" int tdw_density_tau(tdw_ctx *c, double *dens)
{
    if (!c) return TDW_EINVAL;
    return density_tau(c, dens);
}
extern "C" int tdw_potentials(tdw_ctx *c, int it, const double *dens, double *pots)
{
    if (!c) return TDW_EINVAL;
    return potentials(c, it, dens, pots);
}
extern "C" int tdw_apply_h(tdw_ctx *c, int it, int n, const tdw_complex *in, tdw_complex *out,
                           const tdw_complex *lap, const double *pots, const double *useqpe, double *qpe)
{
    if (!c || n > c->nwf || in == out) return TDW_EINVAL;
    // ASLDA's H needs the gradients and lap(alpha psi): tdw_apply_h_full
    if (varying_mass(c)) return TDW_EINVAL;
    const double *hp = h_view(c, it, pots);
    if (!hp) return TDW_EINVAL;
    return apply_h(c, it, n, CD2(in), D2(out), CD2(lap), hp, c->d_kt, useqpe, qpe ? qpe : c->d_qpe);
}
extern "C" int tdw_apply_h_full(tdw_ctx *c, int it, int n, const tdw_complex *in, tdw_complex *out,
                                const tdw_complex *ddx, const tdw_complex *ddy, const tdw_complex *ddz,
                                const tdw_complex *lap, const tdw_complex *alap, const double *pots,
                                const double *useqpe, double *qpe)
{
    if (!c || n > c->nwf || in == out) return TDW_EINVAL;
    const double *hp = h_view(c, it, pots);
    if (!hp) return TDW_EINVAL;
    if (!varying_mass(c))
        return apply_h(c, it, n, CD2(in), D2(out), CD2(lap), hp, c->d_kt, useqpe, qpe ? qpe : c->d_qpe);
    if (!ddx || (c->dim >= 2 && !ddy) || (c->dim >= 3 && !ddz) || out == ddx || out == ddy || out == ddz)
        return TDW_EINVAL;
    return apply_h_aslda(c, n, CD2(in), D2(out), CD2(ddx), CD2(ddy), CD2(ddz), CD2(lap), CD2(alap), hp, c->d_kt,
                         useqpe, qpe ? qpe : c->d_qpe);
}
extern "C" int tdw_alpha_wf(tdw_ctx *c, int n, const tdw_complex *in, const double *pots, tdw_complex *out)
{
    if (!c || n > c->nwf) return TDW_EINVAL;
    return alpha_wf(c, n, CD2(in), D2(out), pots);
}
extern "C" int tdw_set_user_params(tdw_ctx *c, const double *h_params)
{
    if (!c || !h_params) return TDW_EINVAL;
    for (int k = 0; k < TDW_MAX_USER_PARAMS; k++) c->user_params[k] = h_params[k];
    return project_consts(c);
}
extern "C" int tdw_get_user_params(const tdw_ctx *c, double *h_params)
{
    if (!c || !h_params) return TDW_EINVAL;
    for (int k = 0; k < TDW_MAX_USER_PARAMS; k++) h_params[k] = c->user_params[k];
    return 0;
}
extern "C" int tdw_energy(tdw_ctx *c, int it, const double *dens, const double *pots, double *h_e)
{
    if (!c) return TDW_EINVAL;
    return energy(c, it, dens, pots, h_e);
}
extern "C" int tdw_vext_eval(tdw_ctx *c, int it, double *va, double *vb)
{
    if (!c) return TDW_EINVAL;
    return vext_eval(c, it, va, vb);
}

// ---------------------------------------------------------------------------
// modular PECE step: the call sequence of pca.c:1401-1570
// ---------------------------------------------------------------------------
// skip_lap: a non-gradient half-step whose H application forms the
// Laplacian itself (apply_h_plane); the densities do not read it then
static int half_densities(tdw_ctx *c, tdw_state *s, int grad, bool skip_lap = false)
{
    const int n = s->n;
    if (grad) TDW_TRY(derivatives(c, 2 * n, D2(s->wf), D2(s->ddx), D2(s->ddy), D2(s->ddz), D2(s->lap)));
    else if (!skip_lap) TDW_TRY(laplace(c, 2 * n, D2(s->wf), D2(s->lap)));
    TDW_TRY(densities(c, n, D2(s->wf), D2(s->ddx), D2(s->ddy), D2(s->ddz), D2(s->lap), c->d_fbeta, nullptr,
                      c->d_kt, c->d_cnt, s->dens, grad));
    if (grad) TDW_TRY(density_tau(c, s->dens));
    return 0;
}

static int nhist(const tdw_ctx *c)
{
    return c->p.scheme == TDW_AB3AM4 ? 3 : c->p.scheme == TDW_AB4AM5 ? 4 : 5;
}

// The ASLDA step (not FAST_CONST_EFFECTIVE_MASS_MODE, pca.c:1443-1465 and
// :1510-1535): every half-step computes the derivatives, tau and j, the
// potentials with the effective masses and vector potentials, lap(alpha psi)
// and the full Hamiltonian.
// 3D lattices with the two-pass kernels form the gradients inside the
// density and H passes (tdw_aslda.hip) and store none of them.
static int aslda_densities(tdw_ctx *c, tdw_state *s)
{
    if (aslda_h_fused_supported(c) && dens_grad_fused_supported(c)) {
        TDW_TRY(dens_grad_fused(c, s->n, D2(s->wf), c->d_fbeta, nullptr, s->dens));
        return density_tau(c, s->dens);
    }
    return half_densities(c, s, 1);
}

static int step_aslda(tdw_ctx *c, tdw_state *s, int it, int k, double *qpe)
{
    const int n = s->n;
    // predictor
    TDW_TRY(abm1(c, n, D2(s->wf), D2(s->f[0]), D2(s->f[1]), D2(s->f[2]), D2(s->f[3]), D2(s->f[4])));
    TDW_TRY(normalize(c, n, D2(s->wf), c->d_nrm));
    TDW_TRY(aslda_densities(c, s));
    TDW_TRY(step_potentials(c, it + 1, s->dens, s->pots));
    TDW_TRY(hkf_fields(c, s->pots, true));
    // two-pass lattices fold the predictor's <H> subtraction into the corrector
    const bool fold = aslda_h_fused_supported(c);
    const double *hp = h_view(c, it + 1, s->pots);
    if (!hp) return TDW_EINVAL;
    TDW_TRY(apply_h_aslda(c, n, D2(s->wf), D2(s->lap), D2(s->ddx), D2(s->ddy), D2(s->ddz), D2(s->lap), nullptr,
                          hp, c->d_kt, nullptr, qpe, !fold));
    // corrector
    if (fold) TDW_TRY(abm4q(c, n, CD2(s->lap), D2(s->wf), CD2(s->f[k - 1]), qpe));
    else TDW_TRY(abm4(c, n, CD2(s->lap), D2(s->wf), CD2(s->f[k - 1])));
    TDW_TRY(normalize(c, n, D2(s->wf), c->d_nrm));
    TDW_TRY(aslda_densities(c, s));
    TDW_TRY(step_potentials(c, it + 1, s->dens, s->pots));
    TDW_TRY(hkf_fields(c, s->pots, true));
    tdw_complex *tmp = s->f[k - 1];
    for (int j = k - 1; j > 0; j--) s->f[j] = s->f[j - 1];
    s->f[0] = tmp;
    hp = h_view(c, it + 1, s->pots);
    if (!hp) return TDW_EINVAL;
    return apply_h_aslda(c, n, D2(s->wf), D2(s->f[0]), D2(s->ddx), D2(s->ddy), D2(s->ddz), D2(s->lap), nullptr,
                         hp, c->d_kt, nullptr, qpe);
}

extern "C" int tdw_step(tdw_ctx *c, tdw_state *s, int it, int grad_last)
{
    if (!c || !s || s->n > c->nwf) return TDW_EINVAL;
    // the split-precision form keeps FP32 history entries and exists only as
    // the fused step
    if (c->p.precision == TDW_SPLIT) return TDW_EINVAL;
    TDW_TRY(fused_finalize(c, s));
    const int n = s->n, k = nhist(c);
    double *qpe = s->qpe ? s->qpe : c->d_qpe;
    if (varying_mass(c)) {
        if (!s->ddx || (c->dim >= 2 && !s->ddy) || (c->dim >= 3 && !s->ddz)) return TDW_EINVAL;
        return step_aslda(c, s, it, k, qpe);
    }
    // 2D lattices with a one-pass plane kernel apply H with the Laplacian
    // formed in LDS on non-gradient half-steps (lap then holds H y_p, not the
    // Laplacian, after the step)
    const bool plane = h2d_supported(c);
    // predictor
    TDW_TRY(abm1(c, n, D2(s->wf), D2(s->f[0]), D2(s->f[1]), D2(s->f[2]), D2(s->f[3]), D2(s->f[4])));
    TDW_TRY(normalize(c, n, D2(s->wf), c->d_nrm));
    TDW_TRY(half_densities(c, s, 0, plane));
    TDW_TRY(step_potentials(c, it + 1, s->dens, s->pots));
    TDW_TRY(hkf_fields(c, s->pots, true));
    const double *hp = h_view(c, it + 1, s->pots);
    if (!hp) return TDW_EINVAL;
    if (plane) TDW_TRY(apply_h_plane(c, it + 1, n, D2(s->wf), D2(s->lap), hp, c->d_kt, nullptr, qpe));
    else TDW_TRY(apply_h(c, it + 1, n, D2(s->wf), D2(s->lap), CD2(s->lap), hp, c->d_kt, nullptr, qpe));
    // corrector
    TDW_TRY(abm4(c, n, CD2(s->lap), D2(s->wf), CD2(s->f[k - 1])));
    TDW_TRY(normalize(c, n, D2(s->wf), c->d_nrm));
    const bool plane_c = plane && !grad_last;
    TDW_TRY(half_densities(c, s, grad_last, plane_c));
    TDW_TRY(step_potentials(c, it + 1, s->dens, s->pots));
    TDW_TRY(hkf_fields(c, s->pots, true));
    tdw_complex *tmp = s->f[k - 1];
    for (int j = k - 1; j > 0; j--) s->f[j] = s->f[j - 1];
    s->f[0] = tmp;
    hp = h_view(c, it + 1, s->pots);
    if (!hp) return TDW_EINVAL;
    if (plane_c) TDW_TRY(apply_h_plane(c, it + 1, n, D2(s->wf), D2(s->f[0]), hp, c->d_kt, nullptr, qpe));
    else TDW_TRY(apply_h(c, it + 1, n, D2(s->wf), D2(s->f[0]), CD2(s->lap), hp, c->d_kt, nullptr, qpe));
    return 0;
}

extern "C" int tdw_initial_measure(tdw_ctx *c, tdw_state *s)
{
    if (!c || !s) return TDW_EINVAL;
    c->pending = 0;
    TDW_TRY(normalize(c, s->n, D2(s->wf), c->d_nrm));
    return half_densities(c, s, 1);
}

// ---------------------------------------------------------------------------
// self-start: the Taylor-series midpoint steps of pca.c:1086-1392
// ---------------------------------------------------------------------------
// derivatives of an array before H acts on it (pca.c:1119-1126, 1179-1186)
static int ss_derivs(tdw_ctx *c, tdw_state *s, const double2 *in, bool grad)
{
    const int n = s->n;
    if (grad) return derivatives(c, 2 * n, in, D2(s->ddx), D2(s->ddy), D2(s->ddz), D2(s->lap));
    return laplace(c, 2 * n, in, D2(s->lap));
}
// densities and potentials at iteration 0; with a varying mass the filter of
// alpha (pca.c:1127-1159, 1223-1251)
static int ss_fields(tdw_ctx *c, tdw_state *s, const double2 *wf, bool grad, bool vm)
{
    TDW_TRY(densities(c, s->n, wf, CD2(s->ddx), CD2(s->ddy), CD2(s->ddz), CD2(s->lap), c->d_fbeta, nullptr,
                      c->d_kt, c->d_cnt, s->dens, grad));
    if (grad) TDW_TRY(density_tau(c, s->dens));
    TDW_TRY(step_potentials(c, 0, s->dens, s->pots));
    if (vm && c->p.hkf_mode >= 1) {
        const long N = (long)c->N;
        TDW_TRY(grid_filter_d(c, 2, s->pots + 4 * N, s->pots + 4 * N, c->p.hkf_mu, c->p.hkf_T));
    }
    return 0;
}
// apply_hamiltonian as the self-start calls it (no friction): V, Delta, A
// filtered in place inside every call (pca_kernels.cu:733-764, 786-790), then
// H in - <H> in (useqpe NULL: <H> into qpe)
static int ss_apply_h(tdw_ctx *c, tdw_state *s, const double2 *in, double2 *out, bool vm, const double *useqpe,
                      double *qpe)
{
    TDW_TRY(hkf_fields(c, s->pots, false));
    const double *hp = h_view(c, 0, s->pots);
    if (!hp) return TDW_EINVAL;
    if (vm)
        return apply_h_aslda(c, s->n, in, out, CD2(s->ddx), CD2(s->ddy), CD2(s->ddz), CD2(s->lap), nullptr, hp,
                             c->d_kt, useqpe, qpe);
    return apply_h(c, 0, s->n, in, out, CD2(s->lap), hp, c->d_kt, useqpe, qpe);
}
static int copy_state(tdw_ctx *c, int n, const tdw_complex *src, tdw_complex *dst)
{
    if (src == dst) return 0;
    return (int)hipMemcpyAsync(dst, src, sizeof(double2) * 2 * (size_t)n * c->N, hipMemcpyDeviceToDevice, c->stream);
}

extern "C" int tdw_selfstart(tdw_ctx *c, tdw_state *s, tdw_complex *scratch0, tdw_complex *scratch1)
{
    if (!c || !s || s->n > c->nwf || !s->wf || !s->lap || !s->dens || !s->pots) return TDW_EINVAL;
    if (c->p.precision != TDW_FP64) return TDW_EINVAL;
    const int K = nhist(c), E = c->p.scheme == TDW_AB3AM4 ? 4 : 5;
    for (int j = 0; j < K; j++)
        if (!s->f[j]) return TDW_EINVAL;
    if ((K <= 4 && !scratch0) || (K == 3 && !scratch1)) return TDW_EINVAL;
    const bool vm = varying_mass(c), grad = vm;
    if (grad && (!s->ddx || (c->dim >= 2 && !s->ddy) || (c->dim >= 3 && !s->ddz))) return TDW_EINVAL;
    TDW_TRY(fused_finalize(c, s));
    const int n = s->n;
    const double dt = c->p.dt;
    double2 *wf = D2(s->wf), *m1 = D2(s->f[0]), *m2 = D2(s->f[1]), *m3 = D2(s->f[2]);
    // the history entries the reference keeps (see below): F_{K-2} in B1,
    // F_{K-1} in B0 -- spare history slots where the scheme has them
    tdw_complex *B1 = K == 5 ? s->f[4] : K == 4 ? s->f[3] : scratch1;
    tdw_complex *B0 = K == 5 ? s->f[3] : scratch0;
    double *qpe0 = s->qpe ? s->qpe : c->d_qpe;
    for (int step = 0; step < K; step++) {
        // predictor: half a step from psi
        TDW_TRY(ss_derivs(c, s, wf, grad));
        TDW_TRY(ss_fields(c, s, wf, grad, vm));
        TDW_TRY(copy_state(c, n, s->wf, s->f[2]));
        TDW_TRY(ss_apply_h(c, s, m3, m1, vm, nullptr, qpe0));
        if (step == K - 2) TDW_TRY(copy_state(c, n, s->f[0], B1));
        if (step == K - 1) TDW_TRY(copy_state(c, n, s->f[0], B0));
        TDW_TRY(taylor(c, 1, 0.5 * dt, n, m1, m3, m2));
        for (int k = 1; k < E; k++) {
            TDW_TRY(ss_derivs(c, s, m2, grad));
            TDW_TRY(ss_apply_h(c, s, m2, m1, vm, qpe0, qpe0));
            TDW_TRY(taylor(c, k + 1, 0.5 * dt, n, m1, m3, m2));
        }
        TDW_TRY(normalize(c, n, m3, c->d_nrm));
        // corrector: the whole step from psi with the midpoint's potentials
        TDW_TRY(ss_derivs(c, s, m3, grad));
        TDW_TRY(ss_fields(c, s, m3, grad, vm));
        TDW_TRY(ss_derivs(c, s, wf, grad));
        TDW_TRY(ss_apply_h(c, s, wf, m1, vm, nullptr, qpe0));
        TDW_TRY(taylor(c, 1, dt, n, m1, wf, m2));
        for (int k = 1; k < E; k++) {
            TDW_TRY(ss_derivs(c, s, m2, grad));
            TDW_TRY(ss_apply_h(c, s, m2, m1, vm, qpe0, qpe0));
            TDW_TRY(taylor(c, k + 1, dt, n, m1, wf, m2));
        }
        TDW_TRY(normalize(c, n, wf, c->d_nrm));
    }
    // history as the reference leaves it: its host buffer shifts slot j to
    // slot j + 1 for increasing j (pca.c:1166), so fkm1 = F_{K-1} and every
    // other slot holds F_{K-2}
    TDW_TRY(copy_state(c, n, B0, s->f[0]));
    for (int j = 1; j < K; j++) TDW_TRY(copy_state(c, n, B1, s->f[j]));
    // the new state's measurement (pca.c:1336-1356)
    TDW_TRY(derivatives(c, 2 * n, wf, D2(s->ddx), D2(s->ddy), D2(s->ddz), D2(s->lap)));
    TDW_TRY(densities(c, n, wf, CD2(s->ddx), CD2(s->ddy), CD2(s->ddz), CD2(s->lap), c->d_fbeta, nullptr, c->d_kt,
                      c->d_cnt, s->dens, 1));
    TDW_TRY(density_tau(c, s->dens));
    return step_potentials(c, 0, s->dens, s->pots);
}

extern "C" int tdw_step_fused(tdw_ctx *c, tdw_state *s, int it, int grad_last)
{
    if (!c || !s || s->n > c->nwf) return TDW_EINVAL;
    return fused_step(c, s, it, grad_last);
}

extern "C" int tdw_finalize(tdw_ctx *c, tdw_state *s)
{
    if (!c || !s) return TDW_EINVAL;
    return fused_finalize(c, s);
}

extern "C" int tdw_fused_supported(tdw_ctx *c) { return c ? (int)fused_supported(c) : 0; }

// ---------------------------------------------------------------------------
// drop-in layer
// ---------------------------------------------------------------------------
static tdw_ctx *g_bound = nullptr;

extern "C" int tdw_bind(tdw_ctx *c)
{
    g_bound = c;
    return 0;
}
extern "C" tdw_ctx *tdw_bound(void) { return g_bound; }

#define BOUND()                        \
    tdw_ctx *c = g_bound;              \
    if (!c) return TDW_ENOCTX;

extern "C" int set_gpu(int deviceId) { return (int)hipSetDevice(deviceId); }
extern "C" int gpu_malloc(size_t size, void **ptr) { return (int)hipMalloc(ptr, size); }
extern "C" int gpu_free(void *ptr) { return (int)hipFree(ptr); }
extern "C" int host_malloc_pl(size_t size, void **ptr) { return (int)hipHostMalloc(ptr, size, hipHostMallocDefault); }
extern "C" int host_free_pl(void *ptr) { return (int)hipHostFree(ptr); }
extern "C" int memcopy_host2gpu(void *h, void *d, size_t size)
{
    if (g_bound) TDW_HIP(hipStreamSynchronize(g_bound->stream));
    return (int)hipMemcpy(d, h, size, hipMemcpyHostToDevice);
}
extern "C" int memcopy_gpu2host(void *d, void *h, size_t size)
{
    if (g_bound) TDW_HIP(hipStreamSynchronize(g_bound->stream));
    return (int)hipMemcpy(h, d, size, hipMemcpyDeviceToHost);
}
extern "C" int memcopy_gpu2gpu(void *src, void *dst, size_t size)
{
    if (g_bound) TDW_HIP(hipStreamSynchronize(g_bound->stream));
    return (int)hipMemcpy(dst, src, size, hipMemcpyDeviceToDevice);
}
extern "C" int memcopy_const(double mu_a, double mu_b, double ec, double t0, double dt, double kF)
{
    BOUND();
    c->p.kF = kF;
    c->p.mu_a = mu_a;
    c->p.mu_b = mu_b;
    c->p.ec = ec;
    c->p.t0 = t0;
    c->p.dt = dt;
    return project_consts(c);
}
// the user parameter array (MAX_USER_PARAMS doubles) of the problem's
// callbacks (wslda_cuda_utils.h:80-90): kept on the bound context and, in a
// project build, copied to the callbacks' dc_params
extern "C" int memcopy_const_params(double *params)
{
    BOUND();
    return tdw_set_user_params(c, params);
}
extern "C" int memcopy_const_sclgth(double a)
{
    BOUND();
    c->p.sclgth = a;
    return project_consts(c);
}

extern "C" int create_cufftPlans(int batch_size, int nwfip, size_t *workSize)
{
    BOUND();
    (void)batch_size;
    if (nwfip > c->nwf) return TDW_EINVAL;
    // the caller's work area carries the <H> values and the energies back
    // (pca.c:1575, :1590) and holds the mm_get_* zones: kernels and
    // derivatives (mm_get_size_of_total_workspace) followed by the fftONE
    // zone, which the reference sizes as a one-array plan's work on top of
    // the total (pca_derivative.cu:86-98); here that zone is 2 N complex
    // values (TDW_FFTONE_BYTES_PER_POINT bytes per point)
    const size_t ret = sizeof(double) * (size_t)(nwfip > TDW_NITEMS ? nwfip : TDW_NITEMS);
    const size_t zones = mm_get_size_of_total_workspace(c->N) + (size_t)TDW_FFTONE_BYTES_PER_POINT * c->N;
    if (workSize) *workSize = ret > zones ? ret : zones;
    return 0;
}
extern "C" int set_workspace_for_cufftPlan(void *workArea)
{
    BOUND();
    c->ref_workarea = workArea;
    return 0;
}
extern "C" int compute_derivatives(int n, tdw_complex *wf, tdw_complex *dx, tdw_complex *dy, tdw_complex *dz,
                                   tdw_complex *lap, int nthreads)
{
    BOUND();
    (void)nthreads;
    return derivatives(c, n, CD2(wf), D2(dx), D2(dy), D2(dz), D2(lap));
}
extern "C" int compute_laplace(int n, tdw_complex *wf, tdw_complex *lap, int nthreads)
{
    BOUND();
    (void)nthreads;
    return laplace(c, n, CD2(wf), D2(lap));
}
extern "C" int compute_laplace_real_f(double *f, double *lap, int nthreads)
{
    BOUND();
    (void)nthreads;
    return laplace_real(c, f, lap);
}
extern "C" int calculate_densities(int n, tdw_complex *wf, tdw_complex *dx, tdw_complex *dy, tdw_complex *dz,
                                   tdw_complex *lap, double *fbeta, double *weights, double *dens, int grad,
                                   int nthreads)
{
    BOUND();
    (void)nthreads;
    return densities(c, n, CD2(wf), CD2(dx), CD2(dy), CD2(dz), CD2(lap), fbeta, weights, c->d_kt, c->d_cnt, dens,
                     grad);
}
extern "C" int calculate_densities_1d(int n, tdw_complex *wf, tdw_complex *dx, tdw_complex *lap, double *kkyz,
                                      double *fbeta, double *weights, int *cnt, double *dens, int grad,
                                      int nthreads)
{
    BOUND();
    (void)nthreads;
    return densities(c, n, CD2(wf), CD2(dx), nullptr, nullptr, CD2(lap), fbeta, weights, kkyz, cnt, dens, grad);
}
extern "C" int calculate_densities_2d(int n, tdw_complex *wf, tdw_complex *dx, tdw_complex *dy, tdw_complex *lap,
                                      double *kkz, double *fbeta, double *weights, double *dens, int grad,
                                      int nthreads)
{
    BOUND();
    (void)nthreads;
    return densities(c, n, CD2(wf), CD2(dx), CD2(dy), nullptr, CD2(lap), fbeta, weights, kkz, c->d_cnt, dens, grad);
}
// quantum-friction densities into d_qf_densities = [U_a, U_b, D] (4 N
// doubles, the driver passes d_densities + 12 N): pca_dens.cu:459-480,
// 1D ccpca_dens.cu:470-495 (kkyz, cnt), 2D cpca_dens.cu:471-490 (kkz)
extern "C" int calculate_quantum_friction_densities(int n, tdw_complex *wf, tdw_complex *lap, double *fbeta,
                                                    double *weights, double *qf, int nthreads)
{
    BOUND();
    (void)nthreads;
    if (n < 0 || n > c->nwf || !wf || !lap || !fbeta || !qf) return TDW_EINVAL;
    return qf_densities(c, n, CD2(wf), CD2(lap), fbeta, weights, c->d_kt, c->d_cnt, qf);
}
extern "C" int calculate_quantum_friction_densities_1d(int n, tdw_complex *wf, tdw_complex *lap, double *kkyz,
                                                       double *fbeta, double *weights, int *cnt, double *qf,
                                                       int nthreads)
{
    BOUND();
    (void)nthreads;
    if (c->dim != 1 || n < 0 || n > c->nwf || !wf || !lap || !fbeta || !qf || !kkyz || !cnt) return TDW_EINVAL;
    return qf_densities(c, n, CD2(wf), CD2(lap), fbeta, weights, kkyz, cnt, qf);
}
extern "C" int calculate_quantum_friction_densities_2d(int n, tdw_complex *wf, tdw_complex *lap, double *kkz,
                                                       double *fbeta, double *weights, double *qf, int nthreads)
{
    BOUND();
    (void)nthreads;
    if (c->dim != 2 || n < 0 || n > c->nwf || !wf || !lap || !fbeta || !qf || !kkz) return TDW_EINVAL;
    return qf_densities(c, n, CD2(wf), CD2(lap), fbeta, weights, kkz, nullptr, qf);
}
extern "C" int density_caculate_tau(double *dens, int nthreads)
{
    BOUND();
    (void)nthreads;
    return density_tau(c, dens);
}
// Grid-field helpers of the derivative layer (pca_derivative.cu:385-963 and
// the 1D / 2D twins): only the evolved axes' arguments are used, as in the
// 1D / 2D reference files; outputs may alias inputs.
extern "C" int compute_gradient_real_f(double *f, double *df_dx, double *df_dy, double *df_dz, int nthreads)
{
    BOUND();
    (void)nthreads;
    double *const df[3] = {df_dx, df_dy, df_dz};
    return grid_gradient(c, f, df);
}
extern "C" int compute_derivative_real_vector_f(double *fx, double *fy, double *fz, double *dfx_dx, double *dfy_dy,
                                                double *dfz_dz, int nthreads)
{
    BOUND();
    (void)nthreads;
    const double *const f[3] = {fx, fy, fz};
    double *const df[3] = {dfx_dx, dfy_dy, dfz_dz};
    return grid_dvec(c, f, df);
}
extern "C" int compute_divergence_real_vector_f(double *fx, double *fy, double *fz, double *divf, int nthreads)
{
    BOUND();
    (void)nthreads;
    const double *const f[3] = {fx, fy, fz};
    return grid_div(c, f, divf);
}
extern "C" int high_frequency_filter_d(double *in, double *out, double fd_mu, double fd_T, int nthreads)
{
    BOUND();
    (void)nthreads;
    if (!in || !out) return TDW_EINVAL;
    return grid_filter_d(c, 1, in, out, fd_mu, fd_T);
}
extern "C" int high_frequency_filter_c(tdw_complex *in, tdw_complex *out, double fd_mu, double fd_T, int nthreads)
{
    BOUND();
    (void)nthreads;
    if (!in || !out) return TDW_EINVAL;
    return grid_filter_c(c, CD2(in), D2(out), fd_mu, fd_T);
}
extern "C" int high_frequency_filter_massive_d(int n, double *in, double *out, double fd_mu, double fd_T,
                                               int nthreads)
{
    BOUND();
    (void)nthreads;
    if (n < 0 || ((!in || !out) && n > 0)) return TDW_EINVAL;
    return grid_filter_d(c, n, in, out, fd_mu, fd_T);
}
// overlaps of state pairs (pca_kernels.cu:877-911); workarea is not needed
// (the sums run in the context).  As in the reference, overlap_im receives
// the real part too (it copies the first of the two reduced sums into both
// outputs, :899-907); tdw_overlap gives the imaginary part.
extern "C" int compute_ovelap(int n, tdw_complex *wf1, tdw_complex *wf2, double *overlap_re, double *overlap_im,
                              double *workarea, int nthreads)
{
    BOUND();
    (void)workarea;
    (void)nthreads;
    if (n < 0 || n > c->nwf || ((!wf1 || !wf2) && n > 0)) return TDW_EINVAL;
    double *re = overlap_re ? overlap_re : overlap_im;
    if (!re) return 0;
    TDW_TRY(grid_overlap(c, n, CD2(wf1), CD2(wf2), re, nullptr));
    if (overlap_re && overlap_im && n > 0)
        TDW_HIP(hipMemcpyAsync(overlap_im, overlap_re, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    return 0;
}
extern "C" int tdw_overlap(tdw_ctx *c, int n, const tdw_complex *wf1, const tdw_complex *wf2, double *re, double *im)
{
    if (!c || n < 0 || n > c->nwf || ((!wf1 || !wf2) && n > 0)) return TDW_EINVAL;
    return grid_overlap(c, n, CD2(wf1), CD2(wf2), re, im);
}
// sum of size doubles into partial_sums[0] (wslda_cuda_utils.h:250-278)
extern "C" int local_reductionR(double *array, int size, double *partial_sums, int threads, int mode)
{
    BOUND();
    (void)threads;
    if (mode != 0) return TDW_EINVAL;   // the reference defines mode 0 (sum) only
    return grid_sum(c, array, (long)size, partial_sums);
}
// the caller's block of extra device data for its own code
// (wslda_cuda_utils.h:47-52); kept on the bound context
extern "C" int memcopy_extra_data(size_t extra_data_size, void *extra_data)
{
    BOUND();
    c->extra_data = extra_data;
    c->extra_size = extra_data_size;
    return project_consts(c);
}
// workspace views of the work area given to set_workspace_for_cufftPlan
// (tdwslda_memory_management.cu:31-94): [kernels 24 n doubles]
// [derivatives 8 n doubles][fft ONE]; the caller allocates at least
// mm_get_size_of_total_workspace(N) bytes, as pca.c:99, 928-930 does
extern "C" size_t mm_get_size_of_kernels_workspace(size_t n) { return sizeof(double) * n * 24; }
extern "C" size_t mm_get_size_of_derivatives_workspace(size_t n) { return sizeof(double) * n * 8; }
extern "C" size_t mm_get_size_of_total_workspace(size_t n)
{
    return mm_get_size_of_derivatives_workspace(n) + mm_get_size_of_kernels_workspace(n);
}
static double *workarea_of_bound()
{
    return g_bound ? (double *)g_bound->ref_workarea : nullptr;
}
extern "C" void *mm_get_pointer_to_kernels_workspace(size_t n)
{
    (void)n;
    return workarea_of_bound();
}
extern "C" void *mm_get_pointer_to_derivatives_workspace(size_t n)
{
    double *w = workarea_of_bound();
    return w ? (void *)(w + 24 * n) : nullptr;
}
extern "C" void *mm_get_pointer_to_total_workspace(size_t n)
{
    (void)n;
    return workarea_of_bound();
}
extern "C" void *mm_get_pointer_to_fftONE_workspace(size_t n)
{
    double *w = workarea_of_bound();
    return w ? (void *)(w + 32 * n) : nullptr;
}
// rho_b, tau_b, j_b <- rho_a, tau_a, j_a (pca_dens.cu:343-359: 5 N doubles
// from d_densities + 2N to d_densities + 7N)
extern "C" int symmetrize_densities_device(double *dens)
{
    BOUND();
    if (!dens) return TDW_EINVAL;
    const long N = (long)c->N;
    return copy_f64(c, dens + 2 * N, dens + 7 * N, 5 * N);
}
extern "C" int compute_potentials(int it, double *dens, double *pots, double cccoeff, int nthreads)
{
    BOUND();
    (void)nthreads;
    return potentials(c, it, dens, pots, cccoeff);
}
extern "C" int compute_energy(int it, double *dens, double *pots, double *work, int nthreads)
{
    BOUND();
    (void)nthreads;
    double e[TDW_NITEMS];
    TDW_TRY(energy(c, it, dens, pots, e));
    if (work) TDW_HIP(hipMemcpy(work, e, sizeof(e), hipMemcpyHostToDevice));
    return 0;
}

static int apply_h_ref(tdw_ctx *c, int it, int n, tdw_complex *in, tdw_complex *out, tdw_complex *dx,
                       tdw_complex *dy, tdw_complex *dz, tdw_complex *lap, tdw_complex *alap, const double *kt,
                       double *pots, double qfswitch, double *useqpe, double *dens, double pccoeff)
{
    // quantum friction: the friction densities sit at dens + 12N
    // (pca_kernels.cu:714-733, 835-845); V_a, V_b, Delta are modified for
    // this H application and restored after it
    double *qf = nullptr;
    if (qfswitch > 0.0) {
        if (!dens || c->p.kF <= 0.0) return TDW_EINVAL;
        qf = dens + 12 * c->N;
        TDW_TRY(qf_add(c, qf, pots, qfswitch, pccoeff));
    }
    // mean-field filtering, in place, after the friction terms
    // (pca_kernels.cu:733-764, 786-790); alpha is the driver's to filter
    TDW_TRY(hkf_fields(c, pots, false));
    // Delta_ext and the external velocity act inside this H application only
    // (pca_kernels.cu:274-275, 315-327)
    const double *hp = h_view(c, it, pots);
    if (!hp) return TDW_EINVAL;
    if (varying_mass(c)) {
        if (!dx || (c->dim >= 2 && !dy) || (c->dim >= 3 && !dz) || !alap || in == out) return TDW_EINVAL;
        TDW_TRY(apply_h_aslda(c, n, CD2(in), D2(out), CD2(dx), CD2(dy), CD2(dz), CD2(lap), CD2(alap), hp, kt,
                              useqpe, c->d_qpe));
    } else {
        TDW_TRY(apply_h(c, it, n, CD2(in), D2(out), CD2(lap), hp, kt, useqpe, c->d_qpe));
    }
    if (qf) TDW_TRY(qf_remove(c, qf, pots));
    // <H> is left at the head of the work area (pca_kernels.cu:812-822)
    if (c->ref_workarea)
        TDW_HIP(hipMemcpyAsync(c->ref_workarea, useqpe ? useqpe : c->d_qpe, sizeof(double) * n,
                               hipMemcpyDeviceToDevice, c->stream));
    return 0;
}
extern "C" int apply_hamiltonian(int it, int n, tdw_complex *in, tdw_complex *out, tdw_complex *dx, tdw_complex *dy,
                                 tdw_complex *dz, tdw_complex *lap, tdw_complex *alap, double *dens, double *pots,
                                 double qfswitch, double *useqpe, double pccoeff, int nthreads)
{
    BOUND();
    (void)nthreads;
    return apply_h_ref(c, it, n, in, out, dx, dy, dz, lap, alap, c->d_kt, pots, qfswitch, useqpe, dens, pccoeff);
}
extern "C" int apply_hamiltonian_1d(int it, int n, tdw_complex *in, tdw_complex *out, tdw_complex *dx, double *kkyz,
                                    tdw_complex *lap, tdw_complex *alap, double *dens, double *pots, double qfswitch,
                                    double *useqpe, double pccoeff, int nthreads)
{
    BOUND();
    (void)nthreads;
    return apply_h_ref(c, it, n, in, out, dx, nullptr, nullptr, lap, alap, kkyz, pots, qfswitch, useqpe, dens,
                       pccoeff);
}
extern "C" int apply_hamiltonian_2d(int it, int n, tdw_complex *in, tdw_complex *out, tdw_complex *dx,
                                    tdw_complex *dy, double *kkz, tdw_complex *lap, tdw_complex *alap, double *dens,
                                    double *pots, double qfswitch, double *useqpe, double pccoeff, int nthreads)
{
    BOUND();
    (void)nthreads;
    return apply_h_ref(c, it, n, in, out, dx, dy, nullptr, lap, alap, kkz, pots, qfswitch, useqpe, dens, pccoeff);
}
extern "C" int multiply_wf_by_alpha(int n, tdw_complex *wf_in, tdw_complex *wf_out, double *pots, int nthreads)
{
    BOUND();
    (void)nthreads;
    return alpha_wf(c, n, CD2(wf_in), D2(wf_out), pots);
}
extern "C" int normalize_wf(int n, tdw_complex *wf, int nthreads)
{
    BOUND();
    (void)nthreads;
    return normalize(c, n, D2(wf), c->d_nrm);
}
extern "C" int abm_step1(int n, tdw_complex *y, tdw_complex *f1, tdw_complex *f2, tdw_complex *f3, tdw_complex *f4,
                         tdw_complex *f5, int scheme, int nthreads)
{
    BOUND();
    (void)nthreads;
    if (scheme != TDW_AB3AM4 && scheme != TDW_AB4AM5 && scheme != TDW_AB5AM5) return -999999;
    c->p.scheme = scheme;
    return abm1(c, n, D2(y), D2(f1), D2(f2), D2(f3), D2(f4), D2(f5));
}
extern "C" int abm_step4(int n, tdw_complex *in, tdw_complex *out, tdw_complex *flast, int scheme, int nthreads)
{
    BOUND();
    (void)nthreads;
    if (scheme != TDW_AB3AM4 && scheme != TDW_AB4AM5 && scheme != TDW_AB5AM5) return -999999;
    c->p.scheme = scheme;
    return abm4(c, n, CD2(in), D2(out), CD2(flast));
}
extern "C" int taylor_expansion_contribution(int order, double dt, int n, tdw_complex *hpsi, tdw_complex *upd,
                                             tdw_complex *contr, int nthreads)
{
    BOUND();
    (void)nthreads;
    return taylor(c, order, dt, n, CD2(hpsi), D2(upd), D2(contr));
}
extern "C" int get_v_ext(int datadim, int spin, int it, double *data)
{
    BOUND();
    (void)datadim;
    TDW_TRY(vext_eval(c, it, c->d_vext, c->d_vext + c->N));
    TDW_HIP(hipStreamSynchronize(c->stream));
    return (int)hipMemcpy(data, c->d_vext + (spin ? c->N : 0), sizeof(double) * c->N, hipMemcpyDeviceToHost);
}
