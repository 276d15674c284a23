// gh_error.hpp -- the library's error type and argument check, without HIP
// headers (shared by the HIP translation units and the host-only I/O code).
#pragma once
#include <stdexcept>
#include <string>

#include "glimpse_hip.h"

namespace gh {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_error(const std::string &m);

#define GH_CHECK(cond, code, msg)                                                               \
    do {                                                                                        \
        if (!(cond)) throw ::gh::Error((code), (msg));                                          \
    } while (0)

}  // namespace gh
