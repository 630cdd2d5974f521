// stubs.hip -- operators not built yet report VO_ERR_UNSUPPORTED.
#include "vo_internal.hpp"

namespace vo {





uint64_t queue_submit(vo_ctx *ctx, vo_command *) { ctx->last_error = "queue not built"; ctx->last_status = VO_ERR_QUEUE; return 0; }
int queue_wait(vo_ctx *, uint64_t, uint32_t) { return VO_ERR_QUEUE; }
int queue_poll(vo_ctx *, uint64_t) { return -VO_ERR_QUEUE; }
int queue_sync(vo_ctx *) { return VO_ERR_QUEUE; }

} // namespace vo

extern "C" {
int vo_queue_start(vo_ctx *) { return VO_ERR_UNSUPPORTED; }
int vo_queue_stop(vo_ctx *, uint32_t) { return VO_OK; }
int vo_queue_running(const vo_ctx *) { return 0; }
}
