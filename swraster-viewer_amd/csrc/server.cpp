// server.cpp -- render server: continuous batching over resident scenes.
// (filled in below the drop-in path; entry points return SR_ERR_UNSUPPORTED
// until then)

#include "runtime.h"
#include "swraster.h"

using namespace sr;

extern "C" {
SR_API int sr_server_create(const sr_server_opts *, sr_server **) { return fail(SR_ERR_UNSUPPORTED, "server not built"); }
SR_API int sr_server_destroy(sr_server *) { return SR_OK; }
SR_API int sr_server_add_scene(sr_server *, const sr_scene *, int32_t *) { return fail(SR_ERR_UNSUPPORTED, "server not built"); }
SR_API int sr_server_submit(sr_server *, const sr_request *, uint64_t *) { return fail(SR_ERR_UNSUPPORTED, "server not built"); }
SR_API int sr_server_poll(sr_server *, sr_completion *, int32_t, int32_t *) { return fail(SR_ERR_UNSUPPORTED, "server not built"); }
SR_API int sr_server_read_pixels(sr_server *, uint64_t, uint32_t *) { return fail(SR_ERR_UNSUPPORTED, "server not built"); }
SR_API int sr_server_read_vbuffer(sr_server *, uint64_t, uint32_t *, float *) { return fail(SR_ERR_UNSUPPORTED, "server not built"); }
SR_API int sr_server_release(sr_server *, uint64_t) { return fail(SR_ERR_UNSUPPORTED, "server not built"); }
SR_API int sr_server_drain(sr_server *) { return fail(SR_ERR_UNSUPPORTED, "server not built"); }
SR_API double sr_server_clock(const sr_server *) { return 0.0; }
}
