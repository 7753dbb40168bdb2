// batch.hip -- batched solves and the serving loop (filled in below the operator layer).
#include <hip/hip_runtime.h>

#include "op.h"

using namespace cosmo;

extern "C" {
int cosmo_batch_create(int, const cosmo_system *, int, cosmo_batch *) {
  return fail(COSMO_EUNSUP, "cosmo_batch_create: not built yet");
}
int cosmo_batch_solve(cosmo_batch, const cosmo_dav_params *, const int *, double *, double *, cosmo_dav_stats *) {
  return fail(COSMO_EUNSUP, "cosmo_batch_solve: not built yet");
}
int cosmo_batch_destroy(cosmo_batch) { return COSMO_OK; }
int cosmo_server_create(const cosmo_server_params *, cosmo_server *) {
  return fail(COSMO_EUNSUP, "cosmo_server_create: not built yet");
}
int cosmo_server_submit(cosmo_server, cosmo_system, int, double, int64_t *) {
  return fail(COSMO_EUNSUP, "not built yet");
}
int cosmo_server_poll(cosmo_server, cosmo_result *, int) { return fail(COSMO_EUNSUP, "not built yet"); }
int cosmo_server_drain(cosmo_server, cosmo_result *, int, int *) { return fail(COSMO_EUNSUP, "not built yet"); }
int cosmo_server_destroy(cosmo_server) { return COSMO_OK; }
}
