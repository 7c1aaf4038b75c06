// device_context.h — the process-wide device context the drop-in classes
// share (one uepm_ctx per process, on device UEPM_DEVICE or 0).
#pragma once

#include <stdexcept>
#include <string>

#include "uepm.h"

namespace uepm::pseudopotential {

/* Creates the context on first use; throws std::runtime_error when no device
 * can be opened (there is no CPU fallback). */
uepm_ctx* device_context();

/* Maps a C-ABI status to the reference's exception types:
 * INVALID_ARG -> std::invalid_argument, OUT_OF_RANGE -> std::out_of_range,
 * everything else -> std::runtime_error. */
void throw_on_error(int status, const char* where);

}  // namespace uepm::pseudopotential
