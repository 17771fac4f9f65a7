// Continuous-batching planning server (transitional stubs).
#include "mp_plan.h"
extern "C" {
int mp_server_create(const mp_server_params*, mp_server**) { return MP_E_STATE; }
int mp_server_submit(mp_server*, int, const mp_scene*, const float*, const float*, uint64_t, int64_t*) { return MP_E_STATE; }
int mp_server_poll(mp_server*, int64_t, mp_path*) { return MP_E_STATE; }
int mp_server_wait_idle(mp_server*, double) { return MP_E_STATE; }
int mp_server_stats_get(mp_server*, mp_server_stats*) { return MP_E_STATE; }
int mp_server_destroy(mp_server*) { return MP_E_STATE; }
}
