// The library's error type. Raised inside the library; every extern "C" entry
// catches it and returns its negative code, so nothing is thrown across the C
// ABI. Free of HIP, so host-only parts (the engine's scheduler) use it too.
#pragma once
#include <stdexcept>
#include <string>

namespace apg {

struct ApgError : std::runtime_error {
    int code;
    ApgError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

}  // namespace apg
