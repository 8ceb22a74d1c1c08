// Shared error plumbing for the tensormesh_amd C ABI.
#pragma once
#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/tensormesh_amd.h"

namespace tmx {

// Per-thread message of the last failing call (read by tm_last_error()).
void set_error(const char* fmt, ...);

}  // namespace tmx

#define TM_REQUIRE(cond, code, ...)   \
  do {                                \
    if (!(cond)) {                    \
      tmx::set_error(__VA_ARGS__);     \
      return (code);                  \
    }                                 \
  } while (0)
