// serve.cpp -- placeholder.
#include "../../include/jssp.h"
extern "C" {
int jssp_server_create(jssp_engine*, const jssp_server_config*, jssp_server** out) { if (out) *out = nullptr; return JSSP_ERR_INVALID; }
void jssp_server_destroy(jssp_server*) {}
int jssp_serve_submit(jssp_server*, const jssp_instance*, const jssp_params*, uint64_t*) { return JSSP_ERR_INVALID; }
int jssp_serve_step(jssp_server*, int32_t, int32_t*) { return JSSP_ERR_INVALID; }
int jssp_serve_poll(jssp_server*, uint64_t, jssp_result*) { return JSSP_ERR_INVALID; }
int jssp_serve_cancel(jssp_server*, uint64_t) { return JSSP_ERR_INVALID; }
int jssp_serve_drain(jssp_server*) { return JSSP_ERR_INVALID; }
int jssp_serve_times(jssp_server*, uint64_t, double*, double*, double*) { return JSSP_ERR_INVALID; }
}
