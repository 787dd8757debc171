// Shared by the frame kernels (frame.hip, adjoint.hip): the frame plan
// handle, the screen-tile block loop and launch constants.
#pragma once

#include "interp_device.hpp"
#include "render_common.hpp"
#include "rv_internal.hpp"

#include <vector>

namespace rv {

constexpr uint32_t kT = 512;      // threads per workgroup = pixels per screen tile (with bricks)
constexpr uint32_t kRT = 256;     // the same for frames of resident bricks and the adjoint (16 x 16)
constexpr uint32_t kMaxTf = 1024; // colour table entries held in LDS

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Visit, in front-to-back order, the blocks whose view rectangles meet the
// screen tile [x0,x1) x [y0,y1): the rectangles are tested NT at a time (NT =
// workgroup size) and the hits compacted into LDS (list, NT entries) with
// wave ballots.  per_block runs on every thread of the workgroup with the
// block index and returns whether that thread's pixel is done; the loop stops
// once every pixel is.
template <uint32_t NT, class F>
__device__ void for_tile_blocks(const uint4* __restrict__ rects, const uint32_t* __restrict__ f2b,
                                uint32_t nblocks, uint32_t x0, uint32_t y0, uint32_t x1,
                                uint32_t y1, uint32_t* list, uint32_t* wave_cnt, F per_block)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    bool finished = false;
    for (uint32_t c = 0; c < nblocks && !finished; c += NT) {
        const uint32_t i = c + tid;
        uint32_t b = 0;
        bool hit = false;
        if (i < nblocks) {
            b = f2b[i];
            const uint4 r = rects[b];
            hit = r.x < r.z && r.x < x1 && r.z > x0 && r.y < y1 && r.w > y0;
        }
        const uint64_t ball = __ballot(hit);
        if (lane == 0)
            wave_cnt[wave] = (uint32_t)__popcll(ball);
        __syncthreads();
        uint32_t off = 0, total = 0;
        for (uint32_t w = 0; w < NT / 64u; ++w) {
            const uint32_t n = wave_cnt[w];
            off += w < wave ? n : 0u;
            total += n;
        }
        if (hit)
            list[off + (uint32_t)__popcll(ball & ((1ull << lane) - 1ull))] = b;
        __syncthreads();
        for (uint32_t k = 0; k < total; ++k) {
            const bool done = per_block(uniform(list[k]));
            if (__syncthreads_and(done)) {
                finished = true;
                break;
            }
        }
        __syncthreads(); // list and wave_cnt are reused by the next chunk
    }
}

// Replaces *dst with a device copy of src (bytes may be 0: *dst = NULL).
int frame_upload(void** dst, const void* src, size_t bytes);

} // namespace rv

// The frame plan handle (opaque in the C ABI).
struct rv_frame_plan {
    using TileTask = rv::TileTask;
    uint32_t nblocks = 0;
    std::vector<rv_frame_block> blocks;
    rv_frame_desc desc{};
    bool fp16 = false;
    TileTask* d_tasks = nullptr;
    uint32_t n_interp = 0;
    const void** d_bricks = nullptr;
    uint32_t* d_counters = nullptr;
    uint32_t ncams = 0;
    std::vector<float> views;
    rv_block_params* d_params = nullptr;
    uint4* d_rects = nullptr;
    uint32_t* d_f2b = nullptr;
    uint32_t grid = 0;
    uint32_t tw = 32, th = 16;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t capture_stream = nullptr;
    bool can_interp = true;
    float** d_grad_bricks = nullptr; // per block fp32 field gradient (adjoint), or NULL
    uint32_t adj_grid = 0;
    uint32_t render_grid = 0; // resident-brick frames
    // adjoint tape: per camera, per block, the offset of the block's view
    // rectangle in the tape (float4 per pixel); the blocks' 16 x 16 tiles
    // as (block, tile) pairs, grouped by block
    std::vector<uint64_t> rec_total, item_off;
    uint64_t* d_rec_off = nullptr;
    uint32_t* d_items = nullptr;
    float4* d_tape = nullptr;
    uint64_t tape_cap = 0;
    bool tape_on = false;
    int64_t tape_cam = -1;

    ~rv_frame_plan()
    {
        if (graph_exec)
            (void)hipGraphExecDestroy(graph_exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        if (capture_stream)
            (void)hipStreamDestroy(capture_stream);
        for (void* p : {(void*)d_tasks, (void*)d_bricks, (void*)d_counters,
                        (void*)d_params, (void*)d_rects, (void*)d_f2b, (void*)d_grad_bricks,
                        (void*)d_rec_off, (void*)d_items, (void*)d_tape})
            if (p)
                (void)hipFree(p);
    }
};
