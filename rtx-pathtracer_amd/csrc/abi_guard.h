// Exception barrier of the C ABI: every extern "C" entry point of
// librtxamd.so is a function-try-block whose handler returns the status
// these helpers give, so no C++ exception (std::bad_alloc from a host
// vector, a std::length_error, ...) crosses the ABI.
#pragma once

#include "rtx_amd.h"

namespace rtx {
// Call from inside a catch handler: records the message on ctx (may be NULL)
// and returns RTX_E_MEMORY for std::bad_alloc, RTX_E_DEVICE otherwise.
int abiFail(rtx_ctx *ctx) noexcept;
// The same status without a context.
int abiStatus() noexcept;
} // namespace rtx
