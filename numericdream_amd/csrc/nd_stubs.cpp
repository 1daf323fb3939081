// nd_stubs.cpp - entry points whose passes are not built yet report ND_E_UNSUPPORTED.
#include "nd_ctx.h"

extern "C" {

nd_status nd_path_trace(nd_ctx* c, const void*, size_t, uint32_t, nd_rect, const nd_image_out*, nd_trace_stats*) {
    return nd_fail(c, ND_E_UNSUPPORTED, "the wavefront path trace is not built yet");
}

}  // extern "C"

extern "C" nd_status nd_adjoint(nd_ctx* c, const nd_uniforms*, uint32_t, const nd_adjoint_io*, nd_batch_stats*) {
    return nd_fail(c, ND_E_UNSUPPORTED, "the adjoint pass is not built yet");
}
