// host_common.hpp -- helpers for the host-only (g++) translation units.
#pragma once

#include "isai_types.h"

namespace isai {
void set_error(const char *fmt, ...);
void clear_error();
}  // namespace isai

#define ISAI_HOST_REQUIRE(cond, ...)                                       \
    do {                                                                   \
        if (!(cond)) {                                                     \
            ::isai::set_error(__VA_ARGS__);                                \
            return ISAI_ERR_INVALID_ARG;                                   \
        }                                                                  \
    } while (0)
