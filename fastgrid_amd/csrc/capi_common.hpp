// Shared error plumbing of the C ABI: every entry point returns an FG_* code
// and records a message for fg_last_error() on the calling thread.
#pragma once

#include <string>

namespace fgc {
int fail(int code, const std::string &msg);
void clearError();
}  // namespace fgc
