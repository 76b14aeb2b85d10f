// rt_api.hip — library-level entry points and the error plumbing of the C ABI.
#include <mutex>
#include <string>

#include "rt_common.h"

namespace rt {

namespace {
thread_local std::string g_last_error;
std::mutex g_cu_mutex;
int g_cu_count[64] = {0};
std::mutex g_tuning_mutex;
rt_tuning g_tuning = {RT_TRAVERSAL_QUAD, RT_TRAVERSAL_LANE, 0u, (uint32_t)RT_LANE_STACK_MAX, 256u,
                      RT_BUILD_LBVH, RT_TRAVERSAL_LANE, 0u, 32u};
}  // namespace

rt_tuning current_tuning() {
    std::lock_guard<std::mutex> lock(g_tuning_mutex);
    return g_tuning;
}

// Range checks shared by rt_set_tuning and the per-call tunings; empty string = valid.
std::string tuning_problem(const rt_tuning& t) {
    auto alg_ok = [](int32_t a) { return a >= RT_TRAVERSAL_PACKET && a <= RT_TRAVERSAL_HYBRID; };
    if (!(alg_ok(t.primary) || t.primary == RT_TRAVERSAL_AUTO || t.primary == RT_TRAVERSAL_QUAD) ||
        !alg_ok(t.shadow))
        return "traversal must be one of RT_TRAVERSAL_* (RT_TRAVERSAL_AUTO / QUAD for camera rays only)";
    if (t.query != RT_TRAVERSAL_BINARY && t.query != RT_TRAVERSAL_LANE)
        return "query must be RT_TRAVERSAL_BINARY or RT_TRAVERSAL_LANE";
    if (t.packet_budget < 1) return "packet_budget must be >= 1";
    if (t.super_tile > 64) return "super_tile must be 0..64";
    if (t.lane_stack < 1 || t.lane_stack > RT_LANE_STACK_MAX) return "lane_stack must be 1..RT_LANE_STACK_MAX";
    if (t.builder != RT_BUILD_LBVH && t.builder != RT_BUILD_PLOC)
        return "builder must be RT_BUILD_LBVH or RT_BUILD_PLOC";
    // the camera hand-off queue these two fields drive is switched off: its runs at 1920x1080 did
    // not finish, and no cause has been shown
    if (t.handoff_steps != 0) return "handoff_steps must be 0 (the camera hand-off queue is switched off)";
    if (t.item_steps < 1 || t.item_steps > (1u << 20)) return "item_steps must be 1..2^20";
    return std::string();
}

int resolve_tuning(const rt_tuning* in, rt_tuning& out, const char* what) {
    if (!in) {
        out = current_tuning();
        return RT_OK;
    }
    const std::string why = tuning_problem(*in);
    if (!why.empty()) return set_error(RT_EINVAL, std::string(what) + ": tuning: " + why);
    out = *in;
    return RT_OK;
}

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

#ifndef RT_BUILD_ID
#define RT_BUILD_ID "unknown"
#endif

extern "C" int rt_abi_version(void) { return RT_ABI_VERSION; }

extern "C" const char* rt_build_id(void) { return RT_BUILD_ID; }

extern "C" const char* rt_last_error(void) { return rt::g_last_error.c_str(); }

extern "C" int rt_get_tuning(rt_tuning* out) {
    if (!out) return rt::set_error(RT_EINVAL, "rt_get_tuning: out is NULL");
    *out = rt::current_tuning();
    return RT_OK;
}

extern "C" int rt_set_tuning(const rt_tuning* t) {
    if (!t) return rt::set_error(RT_EINVAL, "rt_set_tuning: tuning is NULL");
    const std::string why = rt::tuning_problem(*t);
    if (!why.empty()) return rt::set_error(RT_EINVAL, "rt_set_tuning: " + why);
    std::lock_guard<std::mutex> lock(rt::g_tuning_mutex);
    rt::g_tuning = *t;
    return RT_OK;
}
