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
}  // namespace tdw
