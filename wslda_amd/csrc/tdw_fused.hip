// tdw_fused.hip -- fused PECE step.
#include "tdw_fused.hpp"
#include "tdw_physics.hpp"

namespace tdw {

int fused_step(tdw_ctx *c, tdw_state *s, int it, int grad_last)
{
    return tdw_step(c, s, it, grad_last);
}

void fused_release(tdw_ctx *c) { (void)c; }

}  // namespace tdw
