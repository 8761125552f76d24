// Library-internal helpers shared by the translation units of
// libmorphompm_hip.so (not exported: hidden visibility).
#pragma once

#include <cstdint>

#include "../../include/morphompm_hip.h"

#define MPM_HIDDEN __attribute__((visibility("hidden")))

namespace mpm_internal {

// Writes the per-thread message mpm_last_error() returns; returns `code`.
MPM_HIDDEN int set_err(int code, const char* fmt, ...);
// MPM_ELAUNCH with the HIP error of the last launch, else MPM_OK.
MPM_HIDDEN int check_launch(const char* what);
// MPM_OK or the error code (message set) for a batch the kernels cannot run.
MPM_HIDDEN int validate_batch(const mpm_batch* b);

}  // namespace mpm_internal
