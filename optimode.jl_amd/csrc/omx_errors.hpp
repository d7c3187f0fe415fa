// Error types shared by the host side of the omx library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdexcept>
#include <string>
#include "omx.h"

namespace omx {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

class HipError : public Error {
 public:
  HipError(hipError_t e, const char* expr, const char* file, int line)
      : Error(OMX_ERR_HIP, std::string("HIP error ") + hipGetErrorString(e) + " in " + expr + " (" + file + ":" +
                               std::to_string(line) + ")") {}
};

}  // namespace omx
