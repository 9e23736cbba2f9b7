// OpenMM-shaped stand-in (used only when OpenMM is not installed).
#pragma once

#include <stdexcept>
#include <string>

namespace OpenMM {

class OpenMMException : public std::runtime_error {
public:
    explicit OpenMMException(const std::string& message) : std::runtime_error(message) {}
};

}  // namespace OpenMM
