// Declarations shared by the library's host sources (not part of the C ABI).
#pragma once
#include <cstdint>

// Sets kore_last_error() of the calling host thread; returns code.
int32_t kore_internal_set_error(int32_t code, const char* msg);
