// Request loop (filled in below).
#include "qj_internal.h"

extern "C" {
qj_status qj_server_create(const qj_server_config*, qj_server**) { return QJ_ERR_INVALID; }
double qj_server_now(qj_server*) { return 0; }
qj_status qj_server_submit(qj_server*, const qj_request*, uint64_t*) { return QJ_ERR_INVALID; }
qj_status qj_server_poll(qj_server*, qj_completion*, int32_t, int32_t*, double) { return QJ_ERR_INVALID; }
qj_status qj_server_release(qj_server*, uint64_t) { return QJ_ERR_INVALID; }
qj_status qj_server_destroy(qj_server*) { return QJ_ERR_INVALID; }
}
