// Request server for the served form (continuous batching on one engine).
#pragma once
#include <cstdint>

#include "../../include/nbody.h"

namespace nbody {

class Server;

Server* server_create(const nbody_server_config& cfg);
uint64_t server_submit(Server* s, const nbody_request& req);
void server_poll(Server* s, int timeout_ms, nbody_event* ev, double* x,
                 double* y, double* z, double* vx, double* vy, double* vz,
                 int64_t capacity);
void server_cancel(Server* s, uint64_t id);
void server_stats(Server* s, double* out);
void server_destroy(Server* s);

}  // namespace nbody
