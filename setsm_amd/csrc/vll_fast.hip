// vll_fast.hip — precision form of VerticalLineLocus (placeholder until the
// LDS kernel lands).
#include "smx_internal.hpp"

namespace smx {
int launch_vll_fast(smx_job *j, int half, const CellArgs &A)
{
    (void)j; (void)half; (void)A;
    return fail(SMX_E_UNSUPPORTED, "SMX_PREC_FAST not built yet");
}
}  // namespace smx
