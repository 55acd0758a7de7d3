// Library-wide entry points: error reporting, version, device count.
#include "common.hpp"

namespace ems {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

ems_status hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    char buf[512];
    std::snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", (int)e,
                  hipGetErrorString(e), what, file, line);
    g_last_error = buf;
    if (e == hipErrorOutOfMemory) return EMS_ERR_NO_MEMORY;
    return EMS_ERR_HIP;
}

}  // namespace ems

extern "C" const char *ems_last_error(void) { return ems::g_last_error.c_str(); }

extern "C" const char *ems_version(void) { return "emsolvers_amd 0.1.0 (gfx950)"; }

extern "C" int ems_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}
