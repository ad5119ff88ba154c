// tdw_fused.hpp -- fused PECE step (fewer HBM passes per step).
#pragma once
#include "tdw_internal.hpp"

namespace tdw {
bool fused_supported(const tdw_ctx *c);
int fused_step(tdw_ctx *c, tdw_state *s, int it, int grad_last);
// Applies the pending normalisation of y and <H> subtraction of the newest
// history entry, leaving the arrays exactly as the reference step does.
int fused_finalize(tdw_ctx *c, tdw_state *s);
void fused_release(tdw_ctx *c);
// the per-state scalars of the fused steps (d_sp, d_s2, d_ep), allocated on first use
int fused_state_vecs(tdw_ctx *c);
// The adjoint's taped step (tape t) on the fused FP64 passes, finalised at
// its end: y to y_out (NULL: in place), the partial corrector to cbuf (NULL:
// the oldest history entry), the new entry to f_next (NULL: the dropped
// oldest entry).
int fused_adj_step(tdw_ctx *c, tdw_state *s, int it, tdw_complex *y_out, double2 *cbuf, tdw_complex *f_next,
                   tdw_tape_step *t);
// H applied to arbitrary FP64 vectors (2n rows) on the fused x and yz
// passes, T a state-sized scratch; with qpe, <H> per state into qpe and
// out -= <H> in (as apply_h), without it H in only (as apply_h_only).  3D
// lattices that fused_supported covers, FP64 contexts.
int fused_apply_h(tdw_ctx *c, int n, const double2 *in, double2 *out, double2 *T, const double *pots, double *qpe,
                  const double *use_e = nullptr);
}  // namespace tdw
