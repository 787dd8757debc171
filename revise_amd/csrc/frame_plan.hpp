// Shared by the frame kernels (frame.hip, adjoint.hip): the frame plan
// handle, the screen bins and launch constants.
#pragma once

#include "interp_device.hpp"
#include "render_common.hpp"
#include "rv_internal.hpp"

#include <functional>
#include <memory>
#include <vector>

namespace rv {

constexpr uint32_t kT = 512;      // threads per workgroup of the interpolate-and-render kernel
constexpr uint32_t kRT = 256;     // the same for frames of resident bricks and the adjoint (16 x 16)
constexpr uint32_t kMaxTf = 1024; // colour table entries held in LDS

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Run fn(begin, end) over [0, n) on up to 16 host threads (std::thread),
// in contiguous chunks; returns the first nonzero status.
int parallel_for(uint64_t n, const std::function<int(uint64_t, uint64_t)>& fn);

// Screen bins of one wave footprint (bw x bh = 64 pixels): per camera, the
// blocks whose view rectangles meet each bin, front to back.
struct FrameBins {
    uint32_t bw = 0, bh = 0, nbx = 0, nby = 0;
    uint32_t* d_off = nullptr;            // per camera nbins+1 offsets into its list
    uint32_t* d_list = nullptr;           // block indices
    std::vector<uint64_t> list_base;      // per camera: first entry in d_list
    ~FrameBins()
    {
        if (d_off)
            (void)hipFree(d_off);
        if (d_list)
            (void)hipFree(d_list);
    }
};

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
    bool tiled = false; // fp32 bricks in 4 x 4 x 4 micro-bricks (RV_BRICK_TILED4)
    bool split = false; // interpolating frames as two launches (bricks, then the resident kernel)
    bool interp_small = false; // every brick tile fits the small LDS class (65 x 9 x 9 nodes)
    TileTask* d_tasks = nullptr;
    uint32_t n_interp = 0;
    const void** d_bricks = nullptr;
    uint32_t* d_counters = nullptr;
    uint32_t ncams = 0;
    std::vector<float> views;
    rv_block_params* d_params = nullptr;
    uint32_t grid = 0;
    uint32_t tw = 8, th = 8; // wave footprint of the screen bins
    std::vector<uint4> h_rects;  // per camera, per block view rectangle
    std::vector<uint32_t> h_f2b; // per camera, front-to-back block order
    std::vector<std::unique_ptr<rv::FrameBins>> bins;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t capture_stream = nullptr;
    bool can_interp = true;
    float** d_grad_bricks = nullptr; // per block fp32 field gradient (adjoint), or NULL
    uint32_t adj_grid = 0;
    uint32_t rec_grid = 0;    // colour-table-only adjoint (one lane per record)
    uint32_t render_grid = 0; // resident-brick frames
    bool render_grouped = false; // resident-brick frames: 2 x 2 bins per workgroup (else the queue)
    uint32_t cus = 0, grid_max_wgs = 0, render_max_wgs = 0; // CUs, resident workgroups per CU
    uint32_t frame_rw = rv::kT / 64u; // waves per frame_kernel workgroup that render
    uint32_t field_mode = RV_FIELD_ARGB;
    float iso_opacity = 1.f;
    rv::LevelsArg levels{};
    uint32_t mode_max_wgs = 0; // resident interpolating workgroups per CU in field_mode
    uint32_t schedule = RV_FRAME_SCHEDULE_BINS; // how a wave walks its screen bin (Argb)
    uint32_t lanes_max_wgs = 0; // resident interpolating workgroups per CU, per-lane walk
    uint32_t hybrid_min = 16;   // hybrid schedule: per-lane walk for bins of at least this many blocks
    // secondary field (ValueOnIsosurface(s)): per block brick pointers, and
    // the tiles of both fields for interpolating frames
    const void** d_bricks2 = nullptr;
    TileTask* d_tasks2 = nullptr;
    uint32_t n_interp2 = 0;
    float smin = 0.f, smax = 1.f;
    // adjoint tape: per camera, per block, the offset of the block's view
    // rectangle in the tape (float4 per pixel); the blocks' 8 x 8 tiles
    // as (block, tile) pairs, grouped by block
    std::vector<uint64_t> rec_total, item_off, chunk_off;
    uint64_t* d_rec_off = nullptr;
    uint32_t* d_items = nullptr;
    uint32_t* d_rec_chunk = nullptr; // per camera, per 64 records: the block of the chunk's first
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
        for (void* p : {(void*)d_tasks, (void*)d_bricks, (void*)d_counters, (void*)d_bricks2, (void*)d_tasks2,
                        (void*)d_params, (void*)d_grad_bricks,
                        (void*)d_rec_off, (void*)d_items, (void*)d_rec_chunk, (void*)d_tape})
            if (p)
                (void)hipFree(p);
    }
};
