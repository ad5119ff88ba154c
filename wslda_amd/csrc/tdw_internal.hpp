// tdw_internal.hpp -- library-internal context and launcher declarations.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "wslda_td.h"

#include <vector>

#define TDW_MAXST 8

struct tdw_project;   // a project build's callback table (tdw_hooks.hpp)

struct tdw_timer_rec {
    const char *name;
    hipEvent_t a, b;
};

// Stockham plan of one lattice axis: radices R[s] applied with span Ns[s]
// (product of the earlier radices).
struct tdw_axis {
    int L;
    int nst;
    int R[TDW_MAXST];
    int Ns[TDW_MAXST];
};

struct tdw_ctx {
    tdw_params p;
    int dim;
    int nx, ny, nz;   // evolved axes: nx always, ny if dim>=2, nz if dim==3
    size_t N;         // evolved points
    int S;            // inner block = N / nx (ny*nz, ny, or 1)
    int nwf;
    int device;
    int max_lds;      // bytes of LDS one workgroup may allocate
    hipStream_t stream;
    tdw_axis ax[3];            // 0:x 1:y 2:z
    double2 *d_tw[3];          // exp(-2*pi*i*k/L), k < L, per axis
    float2 *d_tw32[3];
    double Lphys[3];           // nx*dx, ny*dy, nz*dz
    // per-state data
    double *d_fbeta;
    double *d_kt;
    int *d_cnt;
    // scratch
    double *d_part;  size_t part_cap;   // partial sums (doubles)
    double *d_nrm;                       // n norms
    double *d_qpe;                       // n <H>
    double *d_esum;                      // energies (TDW_NITEMS)
    double2 *d_T;    size_t T_cap;       // laplacian scratch (complex elements)
    double2 *d_rtmp;                     // N complex (real-field transforms)
    double *d_vext;                      // 2N
    void *ref_workarea;                  // set_workspace_for_cufftPlan()
    // fused-path state
    double *d_scale;                     // n: 1/sqrt(norm) of the stored (unnormalised) y
    double *d_sp, *d_s2, *d_ep;          // n: predictor scale, scale^2, predictor <H>
    double *d_fpart;  size_t fpart_cap;  // partial sums of the fused kernels
    int pending;                         // y and f[0] carry a pending scale / <H> fix-up
    // adjoint
    double *d_adjv;                      // 4n per-state scalars of the backward step
    // ASLDA (allocated on first use)
    double *d_hx;                        // 4N: lap alpha_a, lap alpha_b, div A_a, div A_b
    double2 *d_ascr;                     // 4N complex: derivative scratch of the real fields
    double2 *d_aw;   size_t aw_cap;      // lap(alpha psi) (complex elements)
    // grid-field helpers (tdw_grid.hip, allocated on first use)
    double2 *d_gscr;                     // 6N complex
    // memcopy_extra_data(): the caller's device block, kept for its own code
    void *extra_data;
    size_t extra_size;
    // memcopy_const_params(): the problem's user parameters
    double user_params[TDW_MAX_USER_PARAMS];
    // project build: the problem's callbacks and the fields they fill
    // (tdw_hooks.hip; null without a project)
    const tdw_project *proj;
    double *d_uext;                      // 2N V_ext of spins a, b
    double *d_agrid;                     // N scattering length
    double *d_potsH;                     // 12N potentials as H reads them
    double2 *d_dext;                     // N Delta_ext
    double *d_vel;                       // 6N velocity of spins a, b
    double *d_epp;                       // 7N per-point energies (modify_energies)
    double *h_densx;                     // 12N pinned host densities (modify_densities)
    // per-kernel timing
    int timing;
    std::vector<tdw_timer_rec> *timers;  // pending records
    std::vector<hipEvent_t> *evpool;
};

namespace tdw {
// The functionals whose effective mass varies in space: the step takes
// derivatives on every half-step and H carries the alpha, A terms
// (ASLDA; SLDAE unless built with SLDA_FORCE_A1).
inline bool varying_mass(const tdw_ctx *c)
{
    return c->p.functional == TDW_ASLDA || (c->p.functional == TDW_SLDAE && !c->p.force_a1);
}
// Brackets one kernel launch with events when timing is enabled.
// Records nest (a timed launcher may call other timed launchers): each scope
// closes its own record.
int timer_begin(tdw_ctx *c, const char *name);
void timer_end(tdw_ctx *c, int idx);
struct TimerScope {
    tdw_ctx *c;
    int idx;
    TimerScope(tdw_ctx *cc, const char *name) : c(cc), idx(-1) { if (c->timing) idx = timer_begin(c, name); }
    ~TimerScope() { if (idx >= 0) timer_end(c, idx); }
};
// Suspends the per-kernel records inside a timed helper, so that the
// grid-sized transforms it launches are not averaged with the state-sized
// launches of the same kernels.
struct TimingOff {
    tdw_ctx *c;
    int saved;
    explicit TimingOff(tdw_ctx *cc) : c(cc), saved(cc->timing) { c->timing = 0; }
    ~TimingOff() { c->timing = saved; }
};
#define TDW_CAT2(a, b) a##b
#define TDW_CAT(a, b) TDW_CAT2(a, b)
#define TDW_TIMED(ctx, name) tdw::TimerScope TDW_CAT(_tdw_ts_, __LINE__)(ctx, name)
int ensure_part(tdw_ctx *c, size_t ndoubles);
int ensure_T(tdw_ctx *c, size_t nelem);

// FFT launchers (tdw_fft.hip).  mode 0: laplacian, 1: derivatives.
// Outer (x) pass: T = D_x(in) [and ddx = G_x(in)] for m arrays.
int launch_outer(tdw_ctx *c, int m, const double2 *in, double2 *T, double2 *ddx, int mode);
int lap2d(tdw_ctx *c, int m, const double2 *in, double2 *out, int *done);
// 2D H application with the Laplacian kept in LDS (one pass per plane);
// part: 2n per-plane sums Re<psi_p, (H psi)_p>; with dotw (state-sized),
// part2: 2n per-plane sums Re<psi_p, dotw_p>; with esub (n per-state <H>),
// out = H psi - esub psi.  h2d_supported says whether the lattice has the
// kernel; in and out must differ.
bool h2d_supported(const tdw_ctx *c);
int h2d(tdw_ctx *c, int n, const double2 *in, double2 *out, const double *pots, const double *kt, double *part,
        const double2 *dotw = nullptr, double *part2 = nullptr, const double *esub = nullptr);
// Inner (y[,z]) pass: out = T + D_y(in) + D_z(in) [ddy, ddz].  T may be NULL.
int launch_inner(tdw_ctx *c, int m, const double2 *in, const double2 *T, double2 *out,
                 double2 *ddy, double2 *ddz, int mode);
}  // namespace tdw

#define TDW_TRY(x)                                  \
    do {                                            \
        int _e = (int)(x);                          \
        if (_e != 0) return _e;                     \
    } while (0)
#define TDW_HIP(x)                                  \
    do {                                            \
        hipError_t _e = (x);                        \
        if (_e != hipSuccess) return (int)_e;       \
    } while (0)
