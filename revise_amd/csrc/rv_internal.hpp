// Internal helpers shared by the library's translation units: status codes,
// the thread-local error message, and HIP call checking.
#pragma once

#include "../../include/revise_hip.h"
#include "share_staging.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace rv {

void set_error(const std::string& msg);
void clear_error();

inline int fail(int code, const std::string& msg)
{
    set_error(msg);
    return code;
}

inline int hip_fail(hipError_t e, const char* what, const char* file, int line)
{
    char buf[512];
    std::snprintf(buf, sizeof buf, "%s failed: %s (%d) at %s:%d", what, hipGetErrorString(e), (int)e,
                  file, line);
    set_error(buf);
    return e == hipErrorOutOfMemory ? RV_ERR_OOM : RV_ERR_HIP;
}

inline hipStream_t S(rv_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// getBoundRect of a device ARGB image (render.hip): bound_rect_launch
// leaves {minX, minY, maxX, maxY} in d_buf (16 bytes of device memory) in
// stream order, crop_launch packs that rectangle's pixels into d_part (cap
// pixels of device or pinned host memory; rect_out, if not null, gets the
// rectangle),
// bound_rect_read waits for the stream and returns {x, y, w, h}.
int bound_rect_launch(const uint32_t* d_pixels, uint32_t w, uint32_t h, int* d_buf, hipStream_t s);
int crop_launch(const uint32_t* d_pixels, uint32_t w, uint32_t h, const int* d_buf, uint32_t* d_part,
                uint64_t cap, int* rect_out, hipStream_t s);
int bound_rect_read(const int* d_buf, int32_t rect[4], int* nonempty, hipStream_t s);

} // namespace rv

#define RV_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t rv_e_ = (call);                                                     \
        if (rv_e_ != hipSuccess)                                                       \
            return ::rv::hip_fail(rv_e_, #call, __FILE__, __LINE__);                   \
    } while (0)

#define RV_LAUNCH_CHECK()                                                              \
    do {                                                                               \
        hipError_t rv_e_ = hipGetLastError();                                          \
        if (rv_e_ != hipSuccess)                                                       \
            return ::rv::hip_fail(rv_e_, "kernel launch", __FILE__, __LINE__);         \
    } while (0)

#define RV_REQUIRE(cond, msg)                                                          \
    do {                                                                               \
        if (!(cond))                                                                   \
            return ::rv::fail(RV_ERR_INVALID, msg);                                    \
    } while (0)

// Guard for C-ABI bodies: no C++ exception may cross the boundary.
#define RV_TRY try {
#define RV_CATCH                                                                       \
    }                                                                                  \
    catch (const std::bad_alloc&)                                                      \
    {                                                                                  \
        return ::rv::fail(RV_ERR_OOM, "host allocation failed");                       \
    }                                                                                  \
    catch (const std::exception& ex)                                                   \
    {                                                                                  \
        return ::rv::fail(RV_ERR_INTERNAL, ex.what());                                 \
    }                                                                                  \
    catch (...)                                                                        \
    {                                                                                  \
        return ::rv::fail(RV_ERR_INTERNAL, "unknown exception");                       \
    }
