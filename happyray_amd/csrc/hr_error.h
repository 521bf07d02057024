// hr_error.h -- the per-thread error message behind hr_last_error(), shared by
// the library's translation units.
#pragma once

#include <string>

#include "../../include/happyray.h"

namespace hr {
// records msg as the calling thread's last error and returns s
hr_status set_error(hr_status s, const std::string& msg);
}  // namespace hr
