// Error reporting shared by the host-only sources of libqsim_gfx950.so
// (no HIP headers): fail() records a printf-style message for
// qsim_last_error() and returns the status code.
#pragma once

namespace qsim {
int fail(int code, const char* fmt, ...);
void clear_error();
}  // namespace qsim
