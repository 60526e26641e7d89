#pragma once
// Error channel of the C++ layer: every C-ABI status other than DY_OK becomes
// a dyphur::Error (a std::runtime_error), where the reference raises SYCL's
// asynchronous exceptions from wait_and_throw (compute/include/compute/stream.hpp:10,26).
#include <dyphur_hip.h>

#include <stdexcept>
#include <string>

namespace dyphur {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string &what)
        : std::runtime_error(what + ": " + dy_status_string(status)), status_(status) {}
    int status() const noexcept { return status_; }

private:
    int status_;
};

inline void check(int status, const char *what) {
    if (status != DY_OK) throw Error(status, what);
}

}  // namespace dyphur
