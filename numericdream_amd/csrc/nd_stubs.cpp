// nd_stubs.cpp - entry points whose passes are not built yet report ND_E_UNSUPPORTED.
#include "nd_ctx.h"

extern "C" {

nd_status nd_path_trace(nd_ctx* c, const void*, size_t, uint32_t, nd_rect, const nd_image_out*, nd_trace_stats*) {
    return nd_fail(c, ND_E_UNSUPPORTED, "the wavefront path trace is not built yet");
}

}  // extern "C"

