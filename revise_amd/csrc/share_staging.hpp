// The worker's staged share launch (frame.hip), declared for host-only
// translation units as well.
#pragma once

#include "../../include/revise_hip.h"

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace rv {

// Staging of a share's block table: pinned host bytes (and their device
// address) and device bytes of `cap` each (grown on demand), and the event
// of the last copy out of the host bytes.
struct ShareStaging {
    void* host = nullptr;
    void* host_dev = nullptr; // the device's address of `host`
    void* dev = nullptr;
    uint64_t cap = 0;
    hipEvent_t copied = nullptr;
    void release();
};

// rv_render_share_layout with the block table and the share kernel's screen
// bins staged through `st` and copied in stream order: no synchronisation.
// d_rect (or null): 4 ints of device memory that receive the crop rectangle
// of the finished viewport, {minX, minY, maxX, maxY} in viewport rows (maxX
// -1 when every pixel is transparent), as bound_rect_launch computes it;
// *rect_done tells whether the launch wrote it (no launch for no blocks).
// init != ~0: the viewport is taken as filled with init (its clear value)
// and is written whole, not read.  record_copied false: st.copied is not
// recorded here, and the caller records it (share_staging_copied) after
// its own launches that follow on the stream -- an event recorded between
// two launches delays the second by several microseconds.
int render_share_staged(const rv_block_params* params, const float* const* bricks, const float* const* bricks2,
                        uint32_t nblocks, uint32_t layout, uint32_t field_mode, float iso_opacity,
                        const float* levels_rel, uint32_t level_count, float smin, float smax, const float* d_tf,
                        uint32_t tf_n, uint32_t* d_viewport, uint32_t vp_w, uint32_t vp_h, ShareStaging& st,
                        rv_stream_t stream, int* d_rect = nullptr, bool* rect_done = nullptr,
                        uint32_t init = ~0u, bool record_copied = true);

// Records st.copied on the stream (after render_share_staged with
// record_copied false and the launches the caller added behind it).
inline hipError_t share_staging_copied(ShareStaging& st, hipStream_t stream)
{
    return st.copied ? hipEventRecord(st.copied, stream) : hipSuccess;
}

} // namespace rv
