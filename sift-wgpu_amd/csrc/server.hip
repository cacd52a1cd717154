// server.hip -- continuous-batching request loop (placeholder until the
// serving path lands; the entry points report SIFT_ERR_INVALID).
#include "sift_context.h"

namespace sh {
struct Server {};
}  // namespace sh

extern "C" {
int sift_hip_server_start(sift_hip_ctx*, const sift_server_config_t*) { return SIFT_ERR_INVALID; }
int sift_hip_submit(sift_hip_ctx*, const sift_image_t*, uint64_t) { return SIFT_ERR_INVALID; }
int sift_hip_poll(sift_hip_ctx*, sift_completion_t*, size_t, size_t* n, int) {
    if (n) *n = 0;
    return SIFT_ERR_INVALID;
}
int sift_hip_server_stop(sift_hip_ctx*) { return SIFT_OK; }
}
