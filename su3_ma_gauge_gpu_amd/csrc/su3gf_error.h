// su3gf_error.h -- the C ABI's error state, shared with the host-only
// translation units (gf_last_error() returns the message of the last failure).
#pragma once
#include "su3gf.h"

namespace su3gf {
gf_status set_error(gf_status st, const char *msg);
}
