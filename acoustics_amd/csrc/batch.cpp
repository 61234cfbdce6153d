// batch.cpp -- placeholder entry points for batches of scenes (replaced by
// the persistent batched path).
#include "device_scene.h"

extern "C" {

void fdtd_batch_options_default(fdtd_batch_options *o)
{
    if (!o) return;
    o->stream = nullptr;
    o->steps_per_graph = 0;
    o->tile_config = -1;
    o->persistent_groups = 0;
}
fdtd_status fdtd_batch_create(const fdtd_scene *, const int *, int, const fdtd_batch_options *, fdtd_batch **out)
{
    if (out) *out = nullptr;
    fdtd::set_last_error("batches are not built yet");
    return FDTD_ERR_UNSUPPORTED;
}
fdtd_status fdtd_batch_run(fdtd_batch *) { return FDTD_ERR_UNSUPPORTED; }
fdtd_status fdtd_batch_step(fdtd_batch *, int) { return FDTD_ERR_UNSUPPORTED; }
fdtd_status fdtd_batch_steps_done(fdtd_batch *, int, int64_t *) { return FDTD_ERR_UNSUPPORTED; }
fdtd_status fdtd_batch_read_listener_series(fdtd_batch *, int, double *) { return FDTD_ERR_UNSUPPORTED; }
fdtd_status fdtd_batch_destroy(fdtd_batch *) { return FDTD_ERR_UNSUPPORTED; }

void fdtd_server_options_default(fdtd_server_options *o)
{
    if (!o) return;
    o->hbm_budget_bytes = 0;
    o->max_active = 0;
    o->steps_per_block = 0;
    o->device = -1;
}
fdtd_status fdtd_server_create(const fdtd_server_options *, fdtd_server **out)
{
    if (out) *out = nullptr;
    return FDTD_ERR_UNSUPPORTED;
}
fdtd_status fdtd_server_submit(fdtd_server *, const fdtd_scene *, int, int64_t *) { return FDTD_ERR_UNSUPPORTED; }
fdtd_status fdtd_server_poll(fdtd_server *, int64_t) { return FDTD_ERR_UNSUPPORTED; }
fdtd_status fdtd_server_wait(fdtd_server *, int64_t, int) { return FDTD_ERR_UNSUPPORTED; }
fdtd_status fdtd_server_take_result(fdtd_server *, int64_t, double *, double *) { return FDTD_ERR_UNSUPPORTED; }
fdtd_status fdtd_server_cancel(fdtd_server *, int64_t) { return FDTD_ERR_UNSUPPORTED; }
fdtd_status fdtd_server_get_stats(fdtd_server *, fdtd_server_stats *) { return FDTD_ERR_UNSUPPORTED; }
int64_t fdtd_scene_device_bytes(const fdtd_scene *) { return -1; }
fdtd_status fdtd_server_destroy(fdtd_server *) { return FDTD_ERR_UNSUPPORTED; }

}  // extern "C"
