// tdw_fused.hpp -- fused PECE step (fewer HBM passes per step).
#pragma once
#include "tdw_internal.hpp"

namespace tdw {
int fused_step(tdw_ctx *c, tdw_state *s, int it, int grad_last);
void fused_release(tdw_ctx *c);
}  // namespace tdw
