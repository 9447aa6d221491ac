// su3gf_error.h -- the C ABI's error state, shared with the host-only
// translation units (gf_last_error() returns the message of the last failure).
#pragma once
#include "su3gf.h"

namespace su3gf {
gf_status set_error(gf_status st, const char *msg);
// gf_batch_run_sa with GF_SA_ONE_LAUNCH in mode (su3gf_api.hip): on stream,
// or on a stream of its own when stream is null
gf_status batch_sa_one_launch(int n, const gf_lattice *lats, void *const *U, int gauge, float tmax, float tmin,
                              int steps, int micro, int mode, const int *seeds, uint32_t *counters, int *temperatures,
                              void *stream);
}
