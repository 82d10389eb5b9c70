// Small shared pieces of the host library: error type and thread-local message.
#pragma once
#include <stdexcept>
#include <string>

namespace tis_host {

// Input files missing or malformed (maps to TIS_EIO at the C boundary).
struct IoError : std::runtime_error {
    explicit IoError(const std::string &m) : std::runtime_error(m) {}
};

void set_error(const std::string &msg);

}  // namespace tis_host
