// Error type of the library's host code.  Plain C++ (no HIP), so the host
// routines that use it can also be compiled and checked by a host compiler.
#pragma once

#include <string>
#include <utility>

namespace dmrg {

// Thrown by host code inside the library; every extern "C" entry point
// catches it and turns it into a status code plus dmrg_last_error().
struct Error {
  std::string msg;
  explicit Error(std::string m) : msg(std::move(m)) {}
};

}  // namespace dmrg
