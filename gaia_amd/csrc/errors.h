// Error reporting shared by the host-side objects behind the C ABI: a failure carries a
// GVBD_ERR_* status and a message; the C-ABI entry points catch it and return the status.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/gaia_vbd.h"

namespace gvbd {

struct Failure {
    int code;
    std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Failure{code, msg}; }

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(GVBD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

inline void require(bool cond, const char* msg) {
    if (!cond) fail(GVBD_ERR_INVALID, msg);
}
inline void require(bool cond, const std::string& msg) {
    if (!cond) fail(GVBD_ERR_INVALID, msg);
}

}  // namespace gvbd
