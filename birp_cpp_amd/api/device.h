// device.h -- the process-wide GPU the C++ API renders on, and C-ABI error translation.
#pragma once
#include <string>

#include "birp.h"

namespace birp_api {
// The device of the C++ API: ordinal $BIRP_DEVICE (default 0), opened on first use and kept
// for the life of the process. Throws std::runtime_error when no GPU can be opened.
birp_device* device();
// Maps a C-ABI status to the reference's exception types: BIRP_ERR_INVALID_ARGUMENT ->
// std::invalid_argument (DataCube.cpp:136,179), BIRP_ERR_SINGULAR -> std::runtime_error
// ("Determinant is 0", Helper.cpp:243), anything else -> std::runtime_error.
void check(int status, const char* where);
}  // namespace birp_api
