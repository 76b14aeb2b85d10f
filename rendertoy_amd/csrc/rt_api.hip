// rt_api.hip — library-level entry points and the error plumbing of the C ABI.
#include <mutex>
#include <string>

#include "rt_common.h"

namespace rt {

namespace {
thread_local std::string g_last_error;
std::mutex g_cu_mutex;
int g_cu_count[64] = {0};
}  // namespace

int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return RT_OK;
    return set_error(RT_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

int check_launch(const char* what) { return check_hip(hipGetLastError(), what); }

// Compute-unit count of the current device (cached per device id).
int device_cu_count() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    std::lock_guard<std::mutex> lock(g_cu_mutex);
    if (g_cu_count[dev] == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        g_cu_count[dev] = cus;
    }
    return g_cu_count[dev];
}

}  // namespace rt

extern "C" int rt_abi_version(void) { return RT_ABI_VERSION; }

extern "C" const char* rt_last_error(void) { return rt::g_last_error.c_str(); }
