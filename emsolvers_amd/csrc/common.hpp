// Shared host-side helpers of libems: per-thread error message, HIP status
// checks and small device-buffer RAII.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdio>
#include <string>

#include "ems/ems.h"

namespace ems {

void set_error(const std::string &msg);
void clear_error();

// Returns EMS_ERR_HIP with the HIP error text recorded.
ems_status hip_fail(hipError_t e, const char *what, const char *file, int line);

#define EMS_HIP(call)                                                       \
    do {                                                                    \
        hipError_t _e = (call);                                             \
        if (_e != hipSuccess) return ::ems::hip_fail(_e, #call, __FILE__, __LINE__); \
    } while (0)

#define EMS_TRY(call)                          \
    do {                                       \
        ems_status _s = (call);                \
        if (_s != EMS_OK) return _s;           \
    } while (0)

inline ems_status invalid(const std::string &msg)
{
    set_error(msg);
    return EMS_ERR_INVALID_ARG;
}

inline ems_status not_supported(const std::string &msg)
{
    set_error(msg);
    return EMS_ERR_NOT_SUPPORTED;
}

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes)
    {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    ems_status reserve(size_t n)
    {
        if (n <= bytes && p) return EMS_OK;
        release();
        if (n == 0) return EMS_OK;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
        }
        bytes = n;
        return EMS_OK;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

}  // namespace ems
