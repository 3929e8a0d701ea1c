// host_error.h -- the library's error text for calls without a context
// (desalm_last_error(NULL)).  Host code only.
#pragma once

#include <string>

namespace desalm {

// Records msg as the error desalm_last_error(NULL) returns; returns -1.
int host_fail(const std::string& msg);

}  // namespace desalm
