// Library-wide hooks defined in engine.hip for the other translation units:
// the error text returned by a2d_last_error, the lock every C entry point
// holds, and the device selected by a2d_init (made current; fails -1 when
// a2d_init has not been called).
#pragma once
#include <mutex>
#include <string>

namespace a2d {
void set_error_text(const std::string& s);
std::recursive_mutex& api_mutex();
int require_device();
}  // namespace a2d
