// cr_host.h -- host-side helpers shared by the C ABI translation units: the
// thread-local error message behind cr_last_error, status macros and a device buffer
// that grows on demand.
#pragma once

#include "../../include/conductorrigid.h"

#include <hip/hip_runtime.h>

#include <string>

namespace cr {

inline thread_local std::string g_error;

inline int set_error(int code, const std::string& msg)
{
    g_error = msg;
    return code;
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    hipError_t ensure(size_t b)
    {
        if (b <= bytes && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, b > 0 ? b : 16);
        if (e == hipSuccess) bytes = b;
        return e;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

}  // namespace cr

#define CR_HIP(call)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return cr::set_error(CR_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

#define CR_RET(call)                    \
    do {                                \
        int r_ = (call);                \
        if (r_ != CR_OK) return r_;     \
    } while (0)
