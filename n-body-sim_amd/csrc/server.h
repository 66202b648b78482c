// Request server for the served form (continuous batching on one engine).
#pragma once
#include <cstddef>
#include <cstdint>
#include <map>
#include <vector>

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

// One admission pass of the server: sizes[k] are the body counts of the
// queued requests in arrival order, bypassed[k] how often each was bypassed
// so far.  Placement is replayed on `free` (a copy of the arena's free
// ranges) with the arena's first-fit rule, and at most free_slots requests
// are taken.  A request that does not fit is bypassed once by each later
// request taken past it; once one has been bypassed max_bypass times, nothing
// behind it is taken.  Returns the positions taken (ascending) and updates
// bypassed.
std::vector<size_t> admission_pass(const std::vector<int64_t>& sizes,
                                   std::vector<int>& bypassed,
                                   std::map<int64_t, int64_t> free,
                                   int free_slots, int max_bypass);

}  // namespace nbody
