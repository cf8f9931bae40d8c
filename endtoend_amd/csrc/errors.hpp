// errors.hpp -- status codes and the thread-local "last error" text behind cccs_last_error().
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/cccs.h"

namespace cccs {

int set_error(int code, const std::string& msg);
void clear_error();

// Evaluates a HIP call; on failure records the message and returns CCCS_EDEVICE from the caller.
#define CCCS_HIP(call)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return ::cccs::set_error(CCCS_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CCCS_TRY(expr)                   \
    do {                                 \
        int s_ = (expr);                 \
        if (s_ != CCCS_OK) return s_;    \
    } while (0)

}  // namespace cccs
