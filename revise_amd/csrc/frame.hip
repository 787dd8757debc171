// Batched persistent frame: interpolation of every brick and ray-casting of
// the whole frame in one launch per frame, graph-capturable.
//
// Work is handed out by two atomic queue counters:
//   counters[0]: brick tiles (the interpolation tile of interp.hip), taken
//                by whole workgroups;
//   counters[3]: screen bins of bw x bh = 64 pixels, taken by single waves.
// A bin's wave composites, for each of its pixels, the per-block ray
// segments of the blocks listed for the bin (every block whose view
// rectangle meets it), front to back, in registers (blendBehind of the
// reference, cuda_blend.hpp:24-33), and leaves the list once every lane's
// pixel is opaque (a wave ballot).  Each block segment is sampled and
// classified exactly like the per-block renderer (renderArgb.hpp:30-55), so
// the result is the reference's compositing of the same segments without
// the per-block 8-bit viewport round trip.  Waves never wait for each other
// inside the screen phase.
//
// Bricks and screen bins synchronize once per workgroup: a workgroup that
// interpolated brick tiles publishes them (device-scope fence and one
// release add of its tile count) when the brick queue runs dry; it then
// waits (acquire) until all n_interp brick tiles are published.  On this
// chip every release/acquire at agent scope writes back / invalidates the
// XCD's L2, so signalling per brick tile costs more than the interpolation
// itself.  A workgroup waits only after the brick queue is empty, and a
// workgroup holding a brick tile never waits, so every wait ends; a wait is
// additionally bounded (it sets an error flag and gives up after ~2^22
// sleeps).
//
// The bins' block lists are built on the host per camera (counting sort of
// the view rectangles over the bins, in front-to-back order).

#include "frame_plan.hpp"
#include "interp_device.hpp"
#include "order.hpp"
#include "render_common.hpp"
#include "rv_internal.hpp"

#include <algorithm>
#include <chrono>
#include <memory>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace rv {

int fill_u32(uint32_t* d, uint64_t n, uint32_t value, hipStream_t s);

namespace {

struct FrameArgs {
    const TileTask* tasks;         // brick tiles of all blocks
    uint32_t n_interp;             // brick tiles of this frame (0: bricks are resident)
    const rv_block_params* params; // per block
    const uint32_t* bin_off;       // per bin: its first entry in bin_list (nbins + 1)
    const uint32_t* bin_list;      // per bin: block indices, front to back
    const void* const* bricks;     // per block, float or half
    const float4* tf;
    uint32_t tf_n;
    uint32_t W, H;
    uint32_t bw_log2, bw, bh, nbx, nbins; // wave footprint bw x bh = 64
    uint32_t render_waves;         // waves per workgroup that take screen bins (frame_kernel)
    float iso_opacity;             // Isosurface(s) modes: alpha scale of the level's colour
    LevelsArg lev;                 // Isosurfaces mode: the levels
    float acc_cutoff;
    uint32_t* argb;                // W*H, rows bottom-up like the reference viewport
    float4* rgba;                  // optional float output, same layout
    uint32_t* counters;            // [0] brick queue, [1] error, [2] brick tiles published, [3] bin queue
    float4* tape;                  // adjoint record per (block, pixel of its rectangle), or NULL
    const uint64_t* rec_off;       // per block: offset of its rectangle in the tape
    uint32_t lanes_min;            // per-lane walk kernels: bins with at least this many blocks
    const void* const* bricks2;    // secondary field bricks (ValueOnIsosurface(s) modes)
    float smin, smax;              // secondary field range
    uint32_t fast;                 // Argb without the tape: the fast sample sequence (march_fast)
    uint32_t grouped;              // render_frame_kernel: 2 x 2 bins per workgroup (else the queue)
};

constexpr uint32_t kSpinLimit = 1u << 22;

// One wave: the pixels of screen bin `bin`, in field mode kMode.
template <class BrickT, uint32_t kMode>
__device__ void render_bin(const FrameArgs& a, uint32_t bin, const float4* tfs, const float* lvs, const uint2* yz)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t bx = bin % a.nbx, by = bin / a.nbx;
    const uint32_t vx = bx * a.bw + (lane & (a.bw - 1u)), vy = by * a.bh + (lane >> a.bw_log2);
    const bool inside = vx < a.W && vy < a.H;
    const uint32_t k1 = a.bin_off[bin + 1];
    if (kMode == kModeMip) {
        // the largest normalised value over the blocks: order does not matter
        uint32_t bits = 0x7fc00000u; // NaN: none yet
        for (uint32_t k = a.bin_off[bin]; k < k1; ++k) {
            const uint32_t blk = uniform(a.bin_list[k]);
            const rv_block_params p = a.params[blk];
            if (inside && vx >= p.view_x && vx < p.view_x + p.view_w && vy >= p.view_y &&
                vy < p.view_y + p.view_h) {
                const RaySeg r = setup_ray(p, vx, vy);
                if (r.hit)
                    bits = mip_merge(bits, march_mip(p, reinterpret_cast<const BrickT*>(a.bricks[blk]), r), p);
            }
        }
        if (inside) {
            const uint32_t pix = (a.H - 1u - vy) * a.W + vx;
            a.argb[pix] = bits;
            if (a.rgba) {
                const float t = __uint_as_float(bits);
                a.rgba[pix] = make_float4(t, t, t, t);
            }
        }
        return;
    }
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    float trans = 1.f; // product of (1 - alpha) in front of the current block (tape only)
    bool opaque = !inside;
    for (uint32_t k = a.bin_off[bin]; k < k1; ++k) {
        const uint32_t blk = uniform(a.bin_list[k]);
        const rv_block_params p = a.params[blk];
        const bool mine = !opaque && vx >= p.view_x && vx < p.view_x + p.view_w && vy >= p.view_y &&
                          vy < p.view_y + p.view_h;
        // Argb: the march with the whole wave (lanes off the block get no
        // samples), so that it shares overlong segments (march_fast,
        // march_impl)
        bool wave_argb = false;
        if constexpr (kMode == kModeArgb)
            wave_argb = true;
        if (wave_argb) {
            RaySeg r = setup_ray(p, vx, vy);
            if (!mine) {
                r.hit = false;
                r.sample_count = 0;
            }
            const BrickT* brick = reinterpret_cast<const BrickT*>(a.bricks[blk]);
            float4 seg;
            if (a.tape) {
                // the pixel's state in front of this block, for the adjoint
                if (r.hit)
                    a.tape[a.rec_off[blk] + (vy - p.view_y) * p.view_w + (vx - p.view_x)] =
                        make_float4(acc.x * acc.w, acc.y * acc.w, acc.z * acc.w, trans);
                float ts;
                seg = march_impl<BrickT, true, true, false>(p, brick, tfs, (int)a.tf_n, r, &ts);
                if (r.hit)
                    trans = trans * ts;
            } else if (a.fast) {
                seg = march_fast<BrickT, true>(p, brick, tfs, (int)a.tf_n, r, yz);
            } else {
                seg = march_impl<BrickT, false, true, false>(p, brick, tfs, (int)a.tf_n, r, nullptr);
            }
            if (r.hit) {
                blend_behind(acc, seg);
                opaque = acc.w > a.acc_cutoff;
            }
        } else if (mine) {
            const RaySeg r = setup_ray(p, vx, vy);
            if (r.hit) {
                const BrickT* brick = reinterpret_cast<const BrickT*>(a.bricks[blk]);
                float4 seg;
                if (multi_level(kMode) && a.lev.n == 0) {
                    seg = make_float4(0.f, 0.f, 0.f, 0.f); // no levels: nothing to draw
                } else if (kMode == kModeValueOnIsosurface || kMode == kModeValueOnIsosurfaces) {
                    const Secondary<BrickT> sec{reinterpret_cast<const BrickT*>(a.bricks2[blk]), a.smin,
                                                a.smax};
                    seg = segment<kMode>(p, brick, tfs, (int)a.tf_n, a.iso_opacity, r, lvs, a.lev.n, sec);
                } else {
                    seg = segment<kMode>(p, brick, tfs, (int)a.tf_n, a.iso_opacity, r, lvs, a.lev.n);
                }
                blend_behind(acc, seg);
                opaque = acc.w > a.acc_cutoff;
            }
        }
        if (__all(opaque))
            break;
    }
    if (inside) {
        const uint32_t pix = (a.H - 1u - vy) * a.W + vx;
        a.argb[pix] = ((unsigned)(acc.w * 255.999f) << 24) + ((unsigned)(acc.x * 255.999f) << 16) +
                      ((unsigned)(acc.y * 255.999f) << 8) + (unsigned)(acc.z * 255.999f);
        if (a.rgba)
            a.rgba[pix] = acc;
    }
}

// The same bin with each lane walking its own pixel's blocks (Argb): a lane
// scans the bin's front-to-back list for the next block whose view
// rectangle holds its pixel and marches it with that block's parameters, so
// the lanes of a wave march different blocks at once.  Every pixel composes
// the same segments in the same order as render_bin, so the results are
// identical; the gain is for blocks a few pixels wide, which cover a small
// part of a bin (the wave-wide walk then marches few of its lanes).
template <class BrickT>
__device__ void render_bin_lanes(const FrameArgs& a, uint32_t bin, const float4* tfs, const uint2* yz)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t bx = bin % a.nbx, by = bin / a.nbx;
    const uint32_t vx = bx * a.bw + (lane & (a.bw - 1u)), vy = by * a.bh + (lane >> a.bw_log2);
    const bool inside = vx < a.W && vy < a.H;
    const uint32_t k1 = a.bin_off[bin + 1];
    uint32_t k = a.bin_off[bin];
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    float trans = 1.f;
    bool live = inside; // the lane may still meet blocks
    for (;;) {
        uint32_t blk = 0;
        bool found = false;
        while (live && k < k1) {
            const uint32_t b = a.bin_list[k];
            ++k;
            // view_x, view_y, view_w, view_h
            const uint32_t* vr = reinterpret_cast<const uint32_t*>(a.params + b) + 2;
            const uint32_t x0 = vr[0], y0 = vr[1], w = vr[2], h = vr[3];
            if (vx >= x0 && vx < x0 + w && vy >= y0 && vy < y0 + h) {
                blk = b;
                found = true;
                break;
            }
        }
        if (!found)
            live = false;
        if (!__any(found))
            break;
        {
            // the march with the whole wave, each lane on its block (lanes
            // without one: block 0, no samples), so that it shares overlong
            // segments (march_fast, march_impl)
            const rv_block_params p = a.params[blk];
            RaySeg r = setup_ray(p, vx, vy);
            if (!found) {
                r.hit = false;
                r.sample_count = 0;
            }
            const BrickT* brick = reinterpret_cast<const BrickT*>(a.bricks[blk]);
            float4 seg;
            if (a.tape) {
                if (r.hit)
                    a.tape[a.rec_off[blk] + (vy - p.view_y) * p.view_w + (vx - p.view_x)] =
                        make_float4(acc.x * acc.w, acc.y * acc.w, acc.z * acc.w, trans);
                float ts;
                seg = march_impl<BrickT, true, true, true>(p, brick, tfs, (int)a.tf_n, r, &ts);
                if (r.hit)
                    trans = trans * ts;
            } else if (a.fast) {
                seg = march_fast<BrickT, true, true>(p, brick, tfs, (int)a.tf_n, r, yz);
            } else {
                seg = march_impl<BrickT, false, true, true>(p, brick, tfs, (int)a.tf_n, r, nullptr);
            }
            if (r.hit) {
                blend_behind(acc, seg);
                if (acc.w > a.acc_cutoff)
                    live = false;
            }
        }
    }
    if (inside) {
        const uint32_t pix = (a.H - 1u - vy) * a.W + vx;
        a.argb[pix] = ((unsigned)(acc.w * 255.999f) << 24) + ((unsigned)(acc.x * 255.999f) << 16) +
                      ((unsigned)(acc.y * 255.999f) << 8) + (unsigned)(acc.z * 255.999f);
        if (a.rgba)
            a.rgba[pix] = acc;
    }
}

// Screen phase of a wave: take bins from the bin queue until it is empty.
// kLanes: the per-lane walk (Argb only) for bins listing at least
// a.lanes_min blocks, the bin walk for the others.
template <class BrickT, uint32_t kMode, bool kLanes>
__device__ void render_bins(const FrameArgs& a, const float4* tfs, const float* lvs, const uint2* yz)
{
    for (;;) {
        uint32_t bin = 0;
        if ((threadIdx.x & 63u) == 0u)
            bin = atomicAdd(a.counters + 3, 1u);
        bin = uniform(bin);
        if (bin >= a.nbins)
            break;
        if constexpr (kLanes) {
            if (a.bin_off[bin + 1] - a.bin_off[bin] >= a.lanes_min) {
                render_bin_lanes<BrickT>(a, bin, tfs, yz);
                continue;
            }
        }
        render_bin<BrickT, kMode>(a, bin, tfs, lvs, yz);
    }
}

// Make this workgroup's brick stores visible device-wide and add its tile
// count to the published total.
__device__ __forceinline__ void publish_tiles(uint32_t* published, uint32_t tiles)
{
    if (tiles == 0u) // uniform over the workgroup
        return;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_fetch_add(published, tiles, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

template <class BrickT, uint32_t kMode, bool kLanes>
__global__ __launch_bounds__(kT) void frame_kernel(FrameArgs a)
{
    __shared__ __align__(16) float A[kTileFloats]; // interpolation tile, then the colour table
    __shared__ float lat[64];
    __shared__ float lvs[kMaxLevels];
    __shared__ uint2 yz_s[2 * kYZTabEntries]; // TILED4 y/z offsets (fast Argb)
    __shared__ uint32_t item_s;
    float4* tfs = reinterpret_cast<float4*>(A);
    uint32_t my_tiles = 0; // brick tiles done by this workgroup
    for (;;) {
        if (threadIdx.x == 0)
            item_s = atomicAdd(a.counters, 1u);
        __syncthreads();
        const uint32_t item = item_s;
        __syncthreads();
        if (item >= a.n_interp)
            break;
        const TileTask k = a.tasks[item];
        interp_tile<kT>(k.job, k.local, k.e0, k.e_own, k.e1, k.idx_off, k.lds_off, A, lat);
        ++my_tiles; // (the queue barrier above protects A before its next use)
    }
    if (a.n_interp) {
        publish_tiles(a.counters + 2, my_tiles);
        if (threadIdx.x == 0) {
            uint32_t spins = 0;
            while (__hip_atomic_load(a.counters + 2, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) <
                   a.n_interp) {
                __builtin_amdgcn_s_sleep(8);
                if (++spins > kSpinLimit) {
                    atomicOr(a.counters + 1, 1u);
                    break;
                }
            }
        }
        __syncthreads();
    }
    stage_tf(tfs, a.tf, a.tf_n, threadIdx.x, kT);
    if (multi_level(kMode) && threadIdx.x < a.lev.n)
        lvs[threadIdx.x] = a.lev.v[threadIdx.x];
    const uint2* yz = nullptr;
    if constexpr (std::is_same<BrickT, Tiled4>::value && kMode == kModeArgb) {
        stage_yz_tables(yz_s, threadIdx.x, kT);
        yz = yz_s;
    }
    __syncthreads();
    if ((threadIdx.x >> 6) < a.render_waves) // fewer concurrent rays can be faster (vector L1)
        render_bins<BrickT, kMode, kLanes>(a, tfs, lvs, yz);
}

// Frames of resident bricks, screen bins only.  a.grouped: workgroup g
// takes the 2 x 2 screen bins of group g (bins of the wave footprint, group
// rows of gx = ceil(nbx / 2)), one per wave, so the four waves that share a
// CU's vector L1 march neighbouring pixels through the same bricks (the
// dynamic LDS caps the workgroups per CU); else a persistent grid whose
// waves take the next free bin from the queue, which balances better where
// bins differ much in cost.
template <class BrickT, uint32_t kMode, bool kLanes>
__global__ __launch_bounds__(kRT) void render_frame_kernel(FrameArgs a)
{
    extern __shared__ float4 tfs_dyn[]; // tf_n + 1 entries
    __shared__ float lvs[kMaxLevels];
    __shared__ uint2 yz_s[2 * kYZTabEntries]; // TILED4 y/z offsets (fast Argb)
    stage_tf(tfs_dyn, a.tf, a.tf_n, threadIdx.x, kRT);
    if (multi_level(kMode) && threadIdx.x < a.lev.n)
        lvs[threadIdx.x] = a.lev.v[threadIdx.x];
    const uint2* yz = nullptr;
    if constexpr (std::is_same<BrickT, Tiled4>::value && kMode == kModeArgb) {
        stage_yz_tables(yz_s, threadIdx.x, kRT);
        yz = yz_s;
    }
    __syncthreads();
    if (!a.grouped) {
        render_bins<BrickT, kMode, kLanes>(a, tfs_dyn, lvs, yz);
        return;
    }
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t gx = (a.nbx + 1u) / 2u, nby = a.nbins / a.nbx;
    const uint32_t bx = (blockIdx.x % gx) * 2u + (wave & 1u), by = (blockIdx.x / gx) * 2u + (wave >> 1);
    if (bx >= a.nbx || by >= nby) // wave-uniform
        return;
    const uint32_t bin = by * a.nbx + bx;
    if constexpr (kLanes) {
        if (a.bin_off[bin + 1] - a.bin_off[bin] >= a.lanes_min) {
            render_bin_lanes<BrickT>(a, bin, tfs_dyn, yz);
            return;
        }
    }
    render_bin<BrickT, kMode>(a, bin, tfs_dyn, lvs, yz);
}

// Worker share in one launch: the blocks (back to front) of one share are
// ray-cast into the viewport with exactly the per-block operation of the
// per-block renderer -- pixel = blendAbove(pixel, segment) through the 8-bit
// ARGB round trip -- so the result equals one launch per block in that order,
// bit for bit.  A workgroup owns 16 x 8 pixels, two 8 x 8 squares with two
// waves each.  Blocks are independent (a segment depends on its block and
// pixel only; only the 8-bit composite is ordered), so a square's blocks are
// split between its two waves: the back wave composites the farther ones in
// order while the front wave marches the nearest ones (at most kShareFront,
// at most half) into LDS, and after a barrier the back wave composites those
// in order.  A square whose rays cross many blocks so takes about half the
// time: the launch is otherwise bounded by its longest workgroups.  The
// blocks meeting a square come from its screen bin's list (worker launches,
// bins of 16 x 16 pixels, two workgroups each, taken in bin_order) or from
// the view rectangles tested 64 at a time with ballots.  The fast sequence
// marches each block with the whole wave (lanes off the block get no
// samples), so that the wave continues the overlong segment of a ray that
// grazes a block's edge (march_fast).  The dynamic LDS is padded so that at
// most kShareWgsPerCu workgroups share a CU: fewer concurrent rays keep more
// of each brick neighbourhood in the vector L1.
constexpr uint32_t kShareT = 256;
#ifndef RV_SHARE_SPLIT
#define RV_SHARE_SPLIT 2
#endif
constexpr uint32_t kShareSplit = RV_SHARE_SPLIT;         // waves per 8 x 8 square
constexpr uint32_t kShareSquares = 4u / kShareSplit;     // squares per workgroup
constexpr uint32_t kShareParts = 4u / kShareSquares;     // workgroups per 16 x 16 bin
constexpr uint32_t kShareFront = 16;                     // blocks per front wave at most
constexpr uint32_t kShareSlots = (kShareSplit - 1u) * kShareFront; // per square
#ifndef RV_SHARE_WGS_PER_CU
#define RV_SHARE_WGS_PER_CU 4
#endif
constexpr uint32_t kShareWgsPerCu = RV_SHARE_WGS_PER_CU;

template <uint32_t kMode, class BrickT>
__global__ __launch_bounds__(kShareT) void share_kernel(const rv_block_params* __restrict__ params,
                                                        const uint4* __restrict__ rects,
                                                        const void* const* __restrict__ bricks,
                                                        const void* const* __restrict__ bricks2,
                                                        const uint32_t* __restrict__ bin_off,
                                                        const uint32_t* __restrict__ bin_list,
                                                        const uint32_t* __restrict__ bin_order,
                                                        float smin, float smax,
                                                        uint32_t nblocks, const float4* __restrict__ tf,
                                                        uint32_t tf_n, float iso_opacity, LevelsArg lev,
                                                        uint32_t* __restrict__ viewport,
                                                        uint32_t W, uint32_t H, int* __restrict__ rect_acc,
                                                        uint32_t init)
{
    extern __shared__ float4 tfs_share[]; // tf_n + 1 entries (and padding)
    __shared__ float lvs[kMaxLevels];
    __shared__ uint2 yz_s[2 * kYZTabEntries]; // TILED4 y/z offsets (fast Argb)
    __shared__ float4 seg_s[kShareSquares][kShareSlots][64]; // the front waves' segments
    __shared__ uint64_t seg_m[kShareSquares][kShareSlots];   // ... and the lanes that composite them
    stage_tf(tfs_share, tf, tf_n, threadIdx.x, kShareT);
    if (multi_level(kMode) && threadIdx.x < lev.n)
        lvs[threadIdx.x] = lev.v[threadIdx.x];
    if constexpr (std::is_same<BrickT, Tiled4>::value && kMode == kModeArgbFast)
        stage_yz_tables(yz_s, threadIdx.x, kShareT);
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t sq = wave % kShareSquares; // this wave's 8 x 8 square
    const uint32_t rank = wave / kShareSquares; // 0: the square's back wave
    const bool front = rank != 0u;
    // the workgroup's squares: part (blockIdx.y % kShareParts) of the screen
    // bin (blockIdx.x, blockIdx.y / kShareParts), or with bin_order (bins by
    // decreasing block count, so the costly bins start first) part
    // lin % kShareParts of bin bin_order[lin / kShareParts]
    uint32_t bx = blockIdx.x, by = blockIdx.y / kShareParts, part = blockIdx.y % kShareParts;
    if (bin_order) {
        const uint32_t lin = blockIdx.y * gridDim.x + blockIdx.x;
        const uint32_t b = bin_order[lin / kShareParts];
        part = lin % kShareParts;
        by = b / gridDim.x;
        bx = b - by * gridDim.x;
    }
    const uint32_t sib = part * kShareSquares + sq; // the square within its bin, x first
    const uint32_t x0 = bx * 16u + (sib & 1u) * 8u, y0 = by * 16u + (sib >> 1) * 8u;
    const bool active = x0 < W && y0 < H; // wave-uniform
    const uint32_t x1 = min(x0 + 8u, W), y1 = min(y0 + 8u, H);
    const uint32_t vx = x0 + (lane & 7u), vy = y0 + (lane >> 3);
    const bool inside = active && vx < W && vy < H;
    const uint32_t pix = inside ? (H - 1u - vy) * W + vx : 0u;
    // init (unless ~0): the viewport's clear value, which the launch writes
    // instead of reading (every pixel belongs to one back wave)
    uint32_t value = inside && !front ? (init != ~0u ? init : viewport[pix]) : 0u;
    // the segment of one block (wave-uniform) at this lane's pixel; *hit:
    // whether the per-block renderer composites it there
    auto seg_of = [&](uint32_t blk, const rv_block_params& p, const RaySeg& r) -> float4 {
        const BrickT* brick = reinterpret_cast<const BrickT*>(bricks[blk]);
        Secondary<BrickT> sec{nullptr, smin, smax};
        if (kMode == kModeValueOnIsosurface || kMode == kModeValueOnIsosurfaces)
            sec.brick = reinterpret_cast<const BrickT*>(bricks2[blk]);
        if constexpr (kMode == kModeArgbFast)
            return march_fast<BrickT, true>(p, brick, tfs_share, (int)tf_n, r, yz_s);
        else if constexpr (kMode == kModeArgb)
            return march_impl<BrickT, false, true, false>(p, brick, tfs_share, (int)tf_n, r, nullptr);
        else
            return segment<kMode>(p, brick, tfs_share, (int)tf_n, iso_opacity, r, lvs, lev.n, sec);
    };
    auto wave_ray = [&](const rv_block_params& p, bool inrect) {
        RaySeg r = setup_ray(p, vx, vy);
        if (!inrect) {
            r.hit = false;
            r.sample_count = 0;
        }
        return r;
    };
    // one block composited at the wave's pixels (the back wave)
    auto draw = [&](uint32_t blk) {
        const rv_block_params p = params[blk];
        const bool inrect = inside && vx >= p.view_x && vx < p.view_x + p.view_w && vy >= p.view_y &&
                            vy < p.view_y + p.view_h;
        if constexpr (kMode == kModeArgbFast || kMode == kModeArgb) {
            const RaySeg r = wave_ray(p, inrect);
            const float4 c = seg_of(blk, p, r);
            if (r.hit)
                value = blend_above_u32(value, c);
            return;
        }
        if (inrect) {
            const RaySeg r = setup_ray(p, vx, vy);
            if (r.hit) {
                if constexpr (kMode == kModeMip) {
                    const BrickT* brick = reinterpret_cast<const BrickT*>(bricks[blk]);
                    value = mip_merge(value, march_mip(p, brick, r), p);
                } else if (multi_level(kMode) && lev.n == 0) {
                    value = 0u; // renderIsosurfaces.hpp:33-34
                } else {
                    value = blend_above_u32(value, seg_of(blk, p, r));
                }
            }
        }
    };
    // one block's segment into LDS slot j (the front wave)
    auto stash = [&](uint32_t blk, uint32_t j) {
        const rv_block_params p = params[blk];
        const bool inrect = inside && vx >= p.view_x && vx < p.view_x + p.view_w && vy >= p.view_y &&
                            vy < p.view_y + p.view_h;
        bool hit = false;
        float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
        if constexpr (kMode == kModeArgbFast || kMode == kModeArgb) {
            const RaySeg r = wave_ray(p, inrect);
            c = seg_of(blk, p, r);
            hit = r.hit;
        } else if (inrect) {
            const RaySeg r = setup_ray(p, vx, vy);
            if (r.hit) {
                c = seg_of(blk, p, r);
                hit = true;
            }
        }
        seg_s[sq][j][lane] = c;
        const uint64_t m = __ballot(hit);
        if (lane == 0u)
            seg_m[sq][j] = m;
    };
    // blocks whose view rectangle {x0, y0, x1, y1} meets this wave's square
    auto meets = [&](const uint4& r) { return r.x < r.z && r.x < x1 && r.z > x0 && r.y < y1 && r.w > y0; };
    // the split: MIP merges are not blends, and Isosurfaces without levels
    // clear the pixel per block (the back wave does all of those)
    const bool split = kMode != kModeMip && !(multi_level(kMode) && lev.n == 0);
    // a square's n blocks: front wave r takes the f blocks of hits
    // [nb + (r - 1) f, nb + r f) into slots (r - 1) f .., the back wave the
    // first nb = n - (kShareSplit - 1) f, with f = min(kShareFront, n /
    // kShareSplit)
    uint32_t nf = 0; // blocks all front waves take
    if (active) {
        if (bin_off) {
            const uint32_t bin = by * gridDim.x + bx;
            const uint32_t k0 = bin_off[bin], k1 = bin_off[bin + 1u];
            uint32_t n = 0;
            if (split)
                for (uint32_t k = k0; k < k1; ++k)
                    n += meets(rects[bin_list[k]]) ? 1u : 0u;
            const uint32_t f = min(kShareFront, n / kShareSplit);
            nf = (kShareSplit - 1u) * f;
            const uint32_t nb = n - nf;
            const uint32_t lo = nb + (rank - 1u) * f, hi = lo + f; // a front wave's hits
            uint32_t h = 0;
            for (uint32_t k = k0; k < k1 && !(front && (f == 0u || h >= hi)); ++k) {
                const uint32_t blk = bin_list[k];
                if (!meets(rects[blk]))
                    continue;
                if (!front) {
                    if (split && h == nb)
                        break;
                    draw(blk);
                } else if (h >= lo && h < hi) {
                    stash(blk, h - nb); // < nf: n counted these blocks
                }
                ++h;
            }
        } else {
            // every block's view rectangle, 64 at a time, tested with a ballot
            uint32_t n = 0;
            if (split)
                for (uint32_t c = 0; c < nblocks; c += 64u)
                    n += (uint32_t)__builtin_popcountll(__ballot(c + lane < nblocks && meets(rects[c + lane])));
            const uint32_t f = min(kShareFront, n / kShareSplit);
            nf = (kShareSplit - 1u) * f;
            const uint32_t nb = n - nf;
            const uint32_t lo = nb + (rank - 1u) * f, hi = lo + f; // a front wave's hits
            uint32_t h = 0;
            for (uint32_t c = 0; c < nblocks && !(!front && split && h == nb) && !(front && (f == 0u || h >= hi));
                 c += 64u) {
                const uint32_t i = c + lane;
                uint64_t m = __ballot(i < nblocks && meets(rects[i]));
                while (m) {
                    const uint32_t blk = c + (uint32_t)__builtin_ctzll(m);
                    m &= m - 1ull;
                    if (!front) {
                        if (split && h == nb)
                            break;
                        draw(blk);
                    } else if (h >= lo && h < hi) {
                        stash(blk, h - nb); // < nf: n counted these blocks
                    }
                    ++h;
                }
            }
        }
    }
    // one square per workgroup: its waves agree on nf, and with nothing
    // handed over the front waves leave at once instead of at a barrier
    if (kShareSquares > 1u || nf != 0u)
        __syncthreads();
    if (front)
        return;
    for (uint32_t j = 0; j < nf; ++j)
        if ((seg_m[sq][j] >> lane) & 1ull)
            value = blend_above_u32(value, seg_s[sq][j][lane]);
    if (inside)
        viewport[pix] = value;
    if (rect_acc && active) {
        // the crop rectangle of the finished viewport (what bound_rect_kernel
        // computes, in viewport rows): this wave's non-transparent pixels
        // from one ballot (bit 8 r + c = pixel (x0 + c, y0 + r)), then an
        // atomic only where it moves a bound (the bounds only shrink toward
        // their result, so a stale read costs an atomic, never a bound)
        const uint64_t m = __ballot(inside && (value >> 24) != 0u);
        if (m != 0ull && lane == 0u) {
            uint32_t cols = 0u, rows = 0u;
#pragma unroll
            for (uint32_t r = 0; r < 8u; ++r) {
                const uint32_t b = (uint32_t)(m >> (8u * r)) & 0xFFu;
                cols |= b;
                rows |= (b != 0u ? 1u : 0u) << r;
            }
            const int mnx = (int)(x0 + (uint32_t)__builtin_ctz(cols));
            const int mxx = (int)(x0 + 31u - (uint32_t)__builtin_clz(cols));
            const int mny = (int)(H - 1u - (y0 + 31u - (uint32_t)__builtin_clz(rows)));
            const int mxy = (int)(H - 1u - (y0 + (uint32_t)__builtin_ctz(rows)));
            if (mnx < __hip_atomic_load(rect_acc + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMin(rect_acc + 0, mnx);
            if (mny < __hip_atomic_load(rect_acc + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMin(rect_acc + 1, mny);
            if (mxx > __hip_atomic_load(rect_acc + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMax(rect_acc + 2, mxx);
            if (mxy > __hip_atomic_load(rect_acc + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMax(rect_acc + 3, mxy);
        }
    }
}

// The bricks of an interpolating frame as a launch of their own (a split
// frame): one workgroup per brick tile, consecutive tiles on one XCD; the
// resident-brick kernel follows in the stream.  Two LDS classes, as in the
// standalone batch: 512 threads for tiles of up to 65 x 17 x 17 nodes, 256
// for plans whose tiles all fit 65 x 9 x 9 (bricks of depth <= 4 with
// 16-cell tiles), which keeps more workgroups per CU.
template <uint32_t kFloats, uint32_t NT>
__global__ __launch_bounds__(NT) void frame_interp_kernel(const TileTask* __restrict__ tasks, uint32_t n)
{
    __shared__ __align__(16) float A[kFloats];
    __shared__ float lat[64];
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= n)
        return;
    const TileTask k = tasks[t];
    interp_tile<NT>(k.job, k.local, k.e0, k.e_own, k.e1, k.idx_off, k.lds_off, A, lat);
}

__global__ void frame_reset_kernel(uint32_t* counters, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (i != 1u)
            counters[i] = 0u;
}

} // namespace
} // namespace rv

using namespace rv;

int rv::frame_upload(void** dst, const void* src, size_t bytes)
{
    if (*dst)
        (void)hipFree(*dst);
    *dst = nullptr;
    if (!bytes)
        return RV_OK;
    RV_HIP(hipMalloc(dst, bytes));
    RV_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return RV_OK;
}

int rv::parallel_for(uint64_t n, const std::function<int(uint64_t, uint64_t)>& fn)
{
    const uint64_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const uint64_t nt = std::min<uint64_t>(hw, n);
    if (nt <= 1)
        return n ? fn(0, n) : RV_OK;
    std::vector<int> rc(nt, RV_OK);
    std::vector<std::string> err(nt);
    std::vector<std::thread> th;
    th.reserve(nt);
    for (uint64_t t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            rc[t] = fn(n * t / nt, n * (t + 1) / nt);
            if (rc[t])
                err[t] = rv_last_error(); // the message is thread-local
        });
    for (auto& x : th)
        x.join();
    for (uint64_t t = 0; t < nt; ++t)
        if (rc[t]) {
            set_error(err[t]);
            return rc[t];
        }
    return RV_OK;
}

namespace {

// Screen bins of the wave footprint bw x bh for every camera of the plan:
// a counting sort of the blocks' view rectangles over the bins, in
// front-to-back order, so each bin's list is front to back.
int build_bins(rv_frame_plan* P, uint32_t bw, uint32_t bh, FrameBins** out)
{
    auto B = std::make_unique<FrameBins>();
    B->bw = bw;
    B->bh = bh;
    B->nbx = (P->desc.vp_w + bw - 1) / bw;
    B->nby = (P->desc.vp_h + bh - 1) / bh;
    const uint64_t nbins = (uint64_t)B->nbx * B->nby, nb = P->nblocks;
    std::vector<std::vector<uint32_t>> off(P->ncams), list(P->ncams);
    int rc = parallel_for(P->ncams, [&](uint64_t c0, uint64_t c1) -> int {
        for (uint64_t c = c0; c < c1; ++c) {
            const uint4* rects = P->h_rects.data() + c * nb;
            const uint32_t* f2b = P->h_f2b.data() + c * nb;
            std::vector<uint32_t>& o = off[c];
            o.assign(nbins + 1, 0u);
            for (uint64_t i = 0; i < nb; ++i) {
                const uint4 r = rects[i];
                if (r.x >= r.z || r.y >= r.w)
                    continue;
                for (uint32_t y = r.y / bh; y <= (r.w - 1) / bh; ++y)
                    for (uint32_t x = r.x / bw; x <= (r.z - 1) / bw; ++x)
                        ++o[(uint64_t)y * B->nbx + x + 1];
            }
            for (uint64_t k = 0; k < nbins; ++k)
                o[k + 1] += o[k];
            if (o[nbins] >= 0x80000000u)
                return fail(RV_ERR_INVALID, "rv_frame: too many (block, bin) pairs for one camera");
            std::vector<uint32_t> cur(o.begin(), o.end() - 1);
            std::vector<uint32_t>& l = list[c];
            l.resize(o[nbins]);
            for (uint64_t i = 0; i < nb; ++i) {
                const uint32_t b = f2b[i];
                const uint4 r = rects[b];
                if (r.x >= r.z || r.y >= r.w)
                    continue;
                for (uint32_t y = r.y / bh; y <= (r.w - 1) / bh; ++y)
                    for (uint32_t x = r.x / bw; x <= (r.z - 1) / bw; ++x)
                        l[cur[(uint64_t)y * B->nbx + x]++] = b;
            }
        }
        return RV_OK;
    });
    if (rc)
        return rc;
    uint64_t total = 0;
    B->list_base.resize(P->ncams);
    for (uint32_t c = 0; c < P->ncams; ++c) {
        B->list_base[c] = total;
        total += list[c].size();
    }
    RV_HIP(hipMalloc((void**)&B->d_off, 4ull * (nbins + 1) * P->ncams));
    RV_HIP(hipMalloc((void**)&B->d_list, 4ull * std::max<uint64_t>(total, 1)));
    for (uint32_t c = 0; c < P->ncams; ++c) {
        RV_HIP(hipMemcpy(B->d_off + (nbins + 1) * c, off[c].data(), 4ull * (nbins + 1),
                         hipMemcpyHostToDevice));
        if (!list[c].empty())
            RV_HIP(hipMemcpy(B->d_list + B->list_base[c], list[c].data(), 4ull * list[c].size(),
                             hipMemcpyHostToDevice));
    }
    P->bins.push_back(std::move(B));
    *out = P->bins.back().get();
    return RV_OK;
}

// The bins of footprint bw x bh, built on first use.
int get_bins(rv_frame_plan* P, uint32_t bw, uint32_t bh, FrameBins** out)
{
    for (auto& b : P->bins)
        if (b->bw == bw && b->bh == bh) {
            *out = b.get();
            return RV_OK;
        }
    return build_bins(P, bw, bh, out);
}

template <uint32_t kMode, bool kLanes>
void launch_sched(const rv_frame_plan* P, const FrameArgs& a, int interpolate, uint32_t grid,
                  hipStream_t s)
{
    if (!interpolate) {
        // grouped: one workgroup per 2 x 2 bins, the colour table's LDS
        // padded so that at most render_grid / cus workgroups share a CU;
        // else the persistent grid of render_grid workgroups
        uint32_t wgs = P->render_grid;
        size_t lds = sizeof(float4) * (a.tf_n + 1u);
        if (a.grouped) {
            wgs = ((a.nbx + 1u) / 2u) * ((a.nbins / a.nbx + 1u) / 2u);
            const uint32_t per_cu = std::max<uint32_t>(1u, P->render_grid / std::max<uint32_t>(P->cus, 1u));
            constexpr size_t kStatic = sizeof(float) * kMaxLevels + sizeof(uint2) * 2u * kYZTabEntries + 1024u;
            const size_t cap = (160u * 1024u) / per_cu;
            lds = std::max<size_t>(lds, std::min<size_t>(65536u, cap > kStatic ? cap - kStatic : 0u));
        }
        if (P->fp16)
            hipLaunchKernelGGL((render_frame_kernel<__half, kMode, kLanes>), dim3(wgs), dim3(kRT), lds, s, a);
        else if (P->tiled)
            hipLaunchKernelGGL((render_frame_kernel<Tiled4, kMode, kLanes>), dim3(wgs), dim3(kRT), lds, s, a);
        else
            hipLaunchKernelGGL((render_frame_kernel<float, kMode, kLanes>), dim3(wgs), dim3(kRT), lds, s, a);
    } else if (P->fp16) {
        hipLaunchKernelGGL((frame_kernel<__half, kMode, kLanes>), dim3(grid), dim3(kT), 0, s, a);
    } else if (P->tiled) {
        hipLaunchKernelGGL((frame_kernel<Tiled4, kMode, kLanes>), dim3(grid), dim3(kT), 0, s, a);
    } else {
        hipLaunchKernelGGL((frame_kernel<float, kMode, kLanes>), dim3(grid), dim3(kT), 0, s, a);
    }
}

template <uint32_t kMode>
void launch_mode(const rv_frame_plan* P, const FrameArgs& a, int interpolate, uint32_t grid,
                 hipStream_t s)
{
    if constexpr (kMode == kModeArgb) {
        if (P->schedule != RV_FRAME_SCHEDULE_BINS) {
            launch_sched<kMode, true>(P, a, interpolate, grid, s);
            return;
        }
    }
    launch_sched<kMode, false>(P, a, interpolate, grid, s);
}

// Resident workgroups per CU of the interpolating kernel in a field mode.
template <uint32_t kMode, bool kLanes = false>
int frame_max_wgs(bool fp16, bool tiled, int* per_cu)
{
    if (fp16)
        RV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, frame_kernel<__half, kMode, kLanes>,
                                                            kT, 0));
    else if (tiled)
        RV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, frame_kernel<Tiled4, kMode, kLanes>,
                                                            kT, 0));
    else
        RV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, frame_kernel<float, kMode, kLanes>,
                                                            kT, 0));
    return RV_OK;
}

bool valid_footprint(uint32_t w, uint32_t h)
{
    return w * h == 64u && w >= 1u && (w & (w - 1u)) == 0u;
}

int launch_frame(rv_frame_plan* P, uint32_t cam, uint32_t* d_argb, float* d_rgba, int interpolate,
                 hipStream_t s, uint32_t tw, uint32_t th)
{
    FrameBins* B = nullptr;
    int rc = get_bins(P, tw, th, &B);
    if (rc)
        return rc;
    FrameArgs a{};
    const bool value_mode = P->field_mode == RV_FIELD_VALUE_ON_ISOSURFACE ||
                            P->field_mode == RV_FIELD_VALUE_ON_ISOSURFACES;
    if (value_mode && interpolate && !P->d_tasks2)
        return fail(RV_ERR_INVALID, "rv_frame_render: the secondary field has no sparse values to "
                                    "interpolate (rv_frame_set_secondary)");
    a.tasks = value_mode ? P->d_tasks2 : P->d_tasks;
    a.n_interp = interpolate ? (value_mode ? P->n_interp2 : P->n_interp) : 0u;
    a.bricks2 = P->d_bricks2;
    a.smin = P->smin;
    a.smax = P->smax;
    a.params = P->d_params + (size_t)cam * P->nblocks;
    const uint64_t nbins = (uint64_t)B->nbx * B->nby;
    a.bin_off = B->d_off + (nbins + 1) * cam;
    a.bin_list = B->d_list + B->list_base[cam];
    a.bricks = P->d_bricks;
    a.tf = reinterpret_cast<const float4*>(P->desc.d_tf);
    a.tf_n = P->desc.tf_n;
    a.W = P->desc.vp_w;
    a.H = P->desc.vp_h;
    a.bw = tw;
    a.bh = th;
    a.bw_log2 = (uint32_t)__builtin_ctz(tw);
    a.nbx = B->nbx;
    a.nbins = (uint32_t)nbins;
    a.render_waves = P->frame_rw;
    a.grouped = P->render_grouped ? 1u : 0u;
    a.acc_cutoff = P->desc.acc_cutoff;
    a.argb = d_argb;
    a.rgba = reinterpret_cast<float4*>(d_rgba);
    a.counters = P->d_counters;
    if (P->tape_on && d_rgba && P->field_mode == RV_FIELD_ARGB) {
        // records of blocks a pixel never reaches keep T = -1
        const uint64_t nrec = P->rec_total[cam];
        rc = fill_u32(reinterpret_cast<uint32_t*>(P->d_tape), 4 * nrec, 0xBF800000u, s);
        if (rc)
            return rc;
        a.tape = P->d_tape;
        a.rec_off = P->d_rec_off + (size_t)cam * P->nblocks;
        P->tape_cam = cam;
    }
    a.iso_opacity = P->iso_opacity;
    a.lev = P->levels;
    a.fast = P->desc.exact == 0u && a.tape == nullptr ? 1u : 0u;
    hipLaunchKernelGGL(frame_reset_kernel, dim3(1), dim3(64), 0, s, P->d_counters, 4u);
    RV_LAUNCH_CHECK();
    if (interpolate && P->split) {
        if (a.n_interp) {
            // xcd_tile needs a grid of a multiple of 8 workgroups
            if (P->interp_small)
                hipLaunchKernelGGL((frame_interp_kernel<kTileFloats8, 256u>), dim3((a.n_interp + 7u) & ~7u),
                                   dim3(256u), 0, s, a.tasks, a.n_interp);
            else
                hipLaunchKernelGGL((frame_interp_kernel<kTileFloats, kT>), dim3((a.n_interp + 7u) & ~7u),
                                   dim3(kT), 0, s, a.tasks, a.n_interp);
            RV_LAUNCH_CHECK();
        }
        a.n_interp = 0u;
        interpolate = 0;
    }
    // the interpolating kernel waits for its bricks: its grid never exceeds
    // what the mode's kernel keeps resident
    uint32_t cap = P->mode_max_wgs;
    if (P->schedule != RV_FRAME_SCHEDULE_BINS && P->field_mode == RV_FIELD_ARGB)
        cap = std::min(cap, P->lanes_max_wgs);
    a.lanes_min = P->schedule == RV_FRAME_SCHEDULE_HYBRID ? P->hybrid_min : 0u;
    const uint32_t grid = std::min(P->grid, P->cus * cap);
    switch (P->field_mode) {
    case RV_FIELD_ARGB_LIGHT:
        launch_mode<kModeArgbLight>(P, a, interpolate, grid, s);
        break;
    case RV_FIELD_ISOSURFACE:
        launch_mode<kModeIsosurface>(P, a, interpolate, grid, s);
        break;
    case RV_FIELD_MAX_INTENSITY_PROJECTION:
        launch_mode<kModeMip>(P, a, interpolate, grid, s);
        break;
    case RV_FIELD_ISOSURFACES:
        launch_mode<kModeIsosurfaces>(P, a, interpolate, grid, s);
        break;
    case RV_FIELD_DOMAIN_VOXELS:
        launch_mode<kModeDomainVoxels>(P, a, interpolate, grid, s);
        break;
    case RV_FIELD_VALUE_ON_ISOSURFACE:
        launch_mode<kModeValueOnIsosurface>(P, a, interpolate, grid, s);
        break;
    case RV_FIELD_VALUE_ON_ISOSURFACES:
        launch_mode<kModeValueOnIsosurfaces>(P, a, interpolate, grid, s);
        break;
    default:
        launch_mode<kModeArgb>(P, a, interpolate, grid, s);
        break;
    }
    RV_LAUNCH_CHECK();
    return RV_OK;
}

} // namespace

extern "C" int rv_frame_plan_create(const rv_frame_block* blocks, uint32_t nblocks,
                                    const rv_frame_desc* desc, rv_frame_plan** out)
{
    RV_TRY
    RV_REQUIRE(out && desc, "rv_frame_plan_create: null pointer");
    *out = nullptr;
    RV_REQUIRE(nblocks > 0 && blocks, "rv_frame_plan_create: no blocks");
    RV_REQUIRE(desc->vp_w > 0 && desc->vp_h > 0 && desc->vp_w <= 16384 && desc->vp_h <= 16384,
               "rv_frame_plan_create: bad viewport");
    RV_REQUIRE(desc->d_tf && desc->tf_n >= 1 && desc->tf_n <= kMaxTf,
               "rv_frame_plan_create: colour table must have 1..1024 entries");
    RV_REQUIRE(desc->fmax > desc->fmin && desc->quality > 0.f, "rv_frame_plan_create: bad range");
    auto* P = new rv_frame_plan;
    std::unique_ptr<rv_frame_plan> guard(P);
    P->nblocks = nblocks;
    P->blocks.assign(blocks, blocks + nblocks);
    P->desc = *desc;
    const bool fp16 = blocks[0].brick16 != nullptr;
    RV_REQUIRE(desc->layout == RV_BRICK_LINEAR || (desc->layout == RV_BRICK_TILED4 && !fp16),
               "rv_frame_plan_create: unknown brick layout (TILED4 bricks are fp32)");
    P->tiled = desc->layout == RV_BRICK_TILED4;
    std::vector<Job> jobs(nblocks);
    std::vector<const void*> bricks(nblocks);
    bool small = true;
    for (uint32_t i = 0; i < nblocks; ++i) {
        const rv_frame_block& b = blocks[i];
        RV_REQUIRE(b.depth <= kMaxDepth, "rv_frame_plan_create: depth must be <= 8");
        small = small && tile_floats(geom_of(b.depth)) <= kTileFloats8;
        RV_REQUIRE(b.level <= 16, "rv_frame_plan_create: level too large");
        for (int k = 0; k < 3; ++k)
            RV_REQUIRE(b.index[k] < (1u << b.level), "rv_frame_plan_create: index outside level");
        RV_REQUIRE((b.brick16 != nullptr) == fp16 && (b.brick != nullptr) != fp16,
                   "rv_frame_plan_create: all blocks must use the same brick type");
        RV_REQUIRE(!P->tiled || (reinterpret_cast<uintptr_t>(b.brick) & 15u) == 0u,
                   "rv_frame_plan_create: TILED4 bricks must be 16-byte aligned");
        jobs[i] = Job{b.sparse, static_cast<const uint8_t*>(b.plan), b.brick,
                      reinterpret_cast<__half*>(b.brick16), b.depth, P->tiled ? 1u : 0u};
        bricks[i] = fp16 ? (const void*)b.brick16 : (const void*)b.brick;
        if (!b.sparse || !b.plan)
            P->can_interp = false;
    }
    P->fp16 = fp16;
    P->interp_small = small;
    int rc;
    if (P->can_interp) {
        std::vector<TileTask> tasks;
        if ((rc = build_tile_tasks(jobs.data(), nblocks, tasks)))
            return rc;
        RV_REQUIRE(tasks.size() < 0x40000000ull, "rv_frame_plan_create: too many brick tiles");
        P->n_interp = (uint32_t)tasks.size();
        if ((rc = frame_upload((void**)&P->d_tasks, tasks.data(), sizeof(TileTask) * tasks.size())))
            return rc;
    }
    if ((rc = frame_upload((void**)&P->d_bricks, bricks.data(), sizeof(void*) * nblocks)))
        return rc;
    RV_HIP(hipMalloc((void**)&P->d_counters, 32)); // [4]: the adjoint's work queue
    RV_HIP(hipMemset(P->d_counters, 0, 32));
    // persistent grid: every workgroup resident at once
    int dev = 0, cus = 0, per_cu = 0;
    RV_HIP(hipGetDevice(&dev));
    RV_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    int lanes_per_cu = 0;
    int orc = frame_max_wgs<kModeArgb, true>(fp16, P->tiled, &lanes_per_cu);
    if (orc)
        return orc;
    RV_REQUIRE(lanes_per_cu >= 1, "rv_frame_plan_create: frame kernel cannot be resident");
    P->lanes_max_wgs = (uint32_t)lanes_per_cu;
    orc = frame_max_wgs<kModeArgb, false>(fp16, P->tiled, &per_cu);
    if (orc)
        return orc;
    RV_REQUIRE(per_cu >= 1, "rv_frame_plan_create: frame kernel cannot be resident");
    P->cus = (uint32_t)cus;
    P->grid_max_wgs = (uint32_t)per_cu;
    P->mode_max_wgs = (uint32_t)per_cu;
    P->grid = (uint32_t)(cus * per_cu);
    if (fp16)
        RV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, render_frame_kernel<__half, kModeArgb, false>, kRT, sizeof(float4) * (desc->tf_n + 1u)));
    else if (P->tiled)
        RV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, render_frame_kernel<Tiled4, kModeArgb, false>, kRT, sizeof(float4) * (desc->tf_n + 1u)));
    else
        RV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, render_frame_kernel<float, kModeArgb, false>, kRT, sizeof(float4) * (desc->tf_n + 1u)));
    RV_REQUIRE(per_cu >= 1, "rv_frame_plan_create: frame kernel cannot be resident");
    P->render_max_wgs = (uint32_t)per_cu;
    P->render_grid = (uint32_t)(cus * std::min(per_cu, 4)); // 16 waves per CU until tuned
    if (desc->tile_w && desc->tile_h) {
        RV_REQUIRE(valid_footprint(desc->tile_w, desc->tile_h),
                   "rv_frame_plan_create: tile_w*tile_h must be 64 with tile_w a power of two");
        P->tw = desc->tile_w;
        P->th = desc->tile_h;
    }
    *out = guard.release();
    return RV_OK;
    RV_CATCH
}

extern "C" int rv_frame_plan_set_cameras(rv_frame_plan* P, const float* views16, uint32_t ncams)
{
    RV_TRY
    RV_REQUIRE(P && views16 && ncams > 0, "rv_frame_plan_set_cameras: bad arguments");
    if (P->graph_exec) {
        (void)hipGraphExecDestroy(P->graph_exec);
        P->graph_exec = nullptr;
    }
    if (P->graph) {
        (void)hipGraphDestroy(P->graph);
        P->graph = nullptr;
    }
    P->bins.clear();
    P->ncams = 0;
    const uint32_t nb = P->nblocks;
    std::vector<rv_block_params> params((size_t)ncams * nb);
    std::vector<uint4>& rects = P->h_rects;
    std::vector<uint32_t>& f2b = P->h_f2b;
    rects.assign((size_t)ncams * nb, make_uint4(0, 0, 0, 0));
    f2b.assign((size_t)ncams * nb, 0u);
    std::vector<uint32_t> keys(4ull * nb);
    for (uint32_t i = 0; i < nb; ++i) {
        keys[4 * i] = P->blocks[i].level;
        for (int k = 0; k < 3; ++k)
            keys[4 * i + 1 + k] = P->blocks[i].index[k];
    }
    const rv_frame_desc& d = P->desc;
    int rc = parallel_for((uint64_t)ncams * nb, [&](uint64_t j0, uint64_t j1) -> int {
        for (uint64_t j = j0; j < j1; ++j) {
            const uint64_t c = j / nb, i = j % nb;
            rv_block_params& p = params[j];
            const int live = block_params(views16 + 16 * c, d.vp_w, d.vp_h, d.fov_y_deg,
                                          P->blocks[i].level, P->blocks[i].index, P->blocks[i].depth,
                                          d.fmin, d.fmax, d.threshold_rel, d.quality, d.alpha_cutoff, p);
            p.flags = d.exact ? RV_RENDER_EXACT : 0u;
            rects[j] = live ? make_uint4(p.view_x, p.view_y, p.view_x + p.view_w, p.view_y + p.view_h)
                            : make_uint4(0, 0, 0, 0);
        }
        return RV_OK;
    });
    if (rc)
        return rc;
    rc = parallel_for(ncams, [&](uint64_t c0, uint64_t c1) -> int {
        std::vector<uint32_t> order;
        for (uint64_t c = c0; c < c1; ++c) {
            float eye[3];
            eye_from_view(views16 + 16 * c, eye);
            const int r = mixed_back_to_front(keys.data(), nb, eye, order);
            if (r)
                return r;
            for (uint32_t i = 0; i < nb; ++i)
                f2b[c * nb + i] = order[nb - 1 - i];
        }
        return RV_OK;
    });
    if (rc)
        return rc;
    if ((rc = frame_upload((void**)&P->d_params, params.data(), sizeof(rv_block_params) * params.size())))
        return rc;
    // adjoint tables: tape offsets of the view rectangles, (block, 8 x 8 tile) pairs
    std::vector<uint64_t> rec_off((size_t)ncams * nb);
    std::vector<uint32_t> items, chunks;
    P->rec_total.assign(ncams, 0);
    P->item_off.assign(ncams + 1, 0);
    P->chunk_off.assign(ncams + 1, 0);
    for (uint32_t c = 0; c < ncams; ++c) {
        uint64_t off = 0;
        for (uint32_t i = 0; i < nb; ++i) {
            const uint4 r = rects[(size_t)c * nb + i];
            rec_off[(size_t)c * nb + i] = off;
            const uint64_t w = r.z - r.x, h = r.w - r.y;
            // the 64-record chunks whose first record is this block's
            for (uint64_t j = (off + 63) / 64; j * 64 < off + w * h; ++j)
                chunks.push_back(i);
            off += w * h;
            const uint32_t tx = (uint32_t)((w + 7) / 8), ty = (uint32_t)((h + 7) / 8);
            for (uint32_t t = 0; t < tx * ty; ++t) {
                items.push_back(i);
                items.push_back(t);
            }
        }
        P->rec_total[c] = off;
        P->item_off[c + 1] = items.size() / 2;
        P->chunk_off[c + 1] = chunks.size();
    }
    if ((rc = frame_upload((void**)&P->d_rec_off, rec_off.data(), 8ull * rec_off.size())))
        return rc;
    if ((rc = frame_upload((void**)&P->d_rec_chunk, chunks.data(), 4ull * chunks.size())))
        return rc;
    if ((rc = frame_upload((void**)&P->d_items, items.data(), 4ull * items.size())))
        return rc;
    P->tape_cam = -1;
    P->ncams = ncams;
    FrameBins* B = nullptr;
    if ((rc = get_bins(P, P->tw, P->th, &B)))
        return rc;
    if (P->tape_on) {
        const int rc2 = rv_frame_set_tape(P, 1);
        if (rc2)
            return rc2;
    }
    P->views.assign(views16, views16 + 16ull * ncams);
    return RV_OK;
    RV_CATCH
}

extern "C" int rv_frame_set_tape(rv_frame_plan* P, int enable)
{
    RV_TRY
    RV_REQUIRE(P, "rv_frame_set_tape: null plan");
    P->tape_cam = -1;
    if (!enable) {
        P->tape_on = false;
        return RV_OK;
    }
    RV_REQUIRE(P->ncams > 0, "rv_frame_set_tape: set cameras first");
    uint64_t need = 1;
    for (uint64_t v : P->rec_total)
        need = std::max(need, v);
    if (need > P->tape_cap) {
        if (P->d_tape)
            RV_HIP(hipFree(P->d_tape));
        P->d_tape = nullptr;
        P->tape_cap = 0;
        RV_HIP(hipMalloc((void**)&P->d_tape, sizeof(float4) * need));
        P->tape_cap = need;
    }
    P->tape_on = true;
    return RV_OK;
    RV_CATCH
}

extern "C" int rv_frame_render(rv_frame_plan* P, uint32_t cam, uint32_t* d_argb, float* d_rgba,
                               int interpolate, rv_stream_t stream)
{
    RV_REQUIRE(P && d_argb, "rv_frame_render: null pointer");
    RV_REQUIRE(cam < P->ncams, "rv_frame_render: camera index out of range (set cameras first)");
    RV_REQUIRE(!interpolate || P->can_interp, "rv_frame_render: blocks have no sparse values/plans");
    return launch_frame(P, cam, d_argb, d_rgba, interpolate, S(stream), P->tw, P->th);
}

extern "C" int rv_frame_graph_build(rv_frame_plan* P, uint32_t* d_argb, uint64_t frame_stride,
                                    int interpolate, rv_stream_t stream)
{
    RV_REQUIRE(P && d_argb && P->ncams > 0, "rv_frame_graph_build: bad arguments");
    RV_REQUIRE(!interpolate || P->can_interp, "rv_frame_graph_build: blocks have no sparse values/plans");
    if (P->graph_exec) {
        (void)hipGraphExecDestroy(P->graph_exec);
        P->graph_exec = nullptr;
    }
    if (P->graph) {
        (void)hipGraphDestroy(P->graph);
        P->graph = nullptr;
    }
    // Capture on a stream of the plan's own (the legacy default stream cannot
    // be captured); the graph is replayed on the caller's stream.
    (void)stream;
    if (!P->capture_stream)
        RV_HIP(hipStreamCreateWithFlags(&P->capture_stream, hipStreamNonBlocking));
    hipStream_t s = P->capture_stream;
    FrameBins* B = nullptr; // built before the capture (host work and copies)
    int brc = get_bins(P, P->tw, P->th, &B);
    if (brc)
        return brc;
    RV_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = RV_OK;
    for (uint32_t c = 0; c < P->ncams && !rc; ++c)
        rc = launch_frame(P, c, d_argb + frame_stride * c, nullptr, interpolate, s, P->tw, P->th);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g)
            (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess)
        return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    P->graph = g;
    RV_HIP(hipGraphInstantiate(&P->graph_exec, g, nullptr, nullptr, 0));
    return RV_OK;
}

extern "C" int rv_frame_graph_launch(rv_frame_plan* P, rv_stream_t stream)
{
    RV_REQUIRE(P && P->graph_exec, "rv_frame_graph_launch: no graph built");
    RV_HIP(hipGraphLaunch(P->graph_exec, S(stream)));
    return RV_OK;
}

extern "C" int rv_frame_status(rv_frame_plan* P, uint32_t* error_flags)
{
    RV_REQUIRE(P && error_flags, "rv_frame_status: null pointer");
    uint32_t v = 0;
    RV_HIP(hipMemcpy(&v, P->d_counters + 1, 4, hipMemcpyDeviceToHost));
    *error_flags = v;
    return RV_OK;
}

extern "C" int rv_frame_get_tile(const rv_frame_plan* P, uint32_t* tile_w, uint32_t* tile_h)
{
    RV_REQUIRE(P && tile_w && tile_h, "rv_frame_get_tile: null pointer");
    *tile_w = P->tw;
    *tile_h = P->th;
    return RV_OK;
}

extern "C" int rv_frame_set_schedule(rv_frame_plan* P, uint32_t schedule)
{
    RV_REQUIRE(P, "rv_frame_set_schedule: null plan");
    RV_REQUIRE(schedule == RV_FRAME_SCHEDULE_BINS || schedule == RV_FRAME_SCHEDULE_LANES ||
                   schedule == RV_FRAME_SCHEDULE_HYBRID,
               "rv_frame_set_schedule: unknown schedule");
    if (schedule != P->schedule) {
        // a captured graph holds the previous schedule's kernels
        if (P->graph_exec) {
            (void)hipGraphExecDestroy(P->graph_exec);
            P->graph_exec = nullptr;
        }
        if (P->graph) {
            (void)hipGraphDestroy(P->graph);
            P->graph = nullptr;
        }
        P->schedule = schedule;
    }
    return RV_OK;
}

extern "C" int rv_frame_get_schedule(const rv_frame_plan* P, uint32_t* schedule)
{
    RV_REQUIRE(P && schedule, "rv_frame_get_schedule: null pointer");
    *schedule = P->schedule;
    return RV_OK;
}

extern "C" int rv_frame_set_hybrid_min(rv_frame_plan* P, uint32_t blocks)
{
    RV_REQUIRE(P, "rv_frame_set_hybrid_min: null plan");
    if (blocks != P->hybrid_min) {
        if (P->graph_exec) {
            (void)hipGraphExecDestroy(P->graph_exec);
            P->graph_exec = nullptr;
        }
        if (P->graph) {
            (void)hipGraphDestroy(P->graph);
            P->graph = nullptr;
        }
        P->hybrid_min = blocks;
    }
    return RV_OK;
}

extern "C" int rv_frame_set_grouped(rv_frame_plan* P, int grouped)
{
    RV_REQUIRE(P, "rv_frame_set_grouped: null plan");
    if ((grouped != 0) != P->render_grouped) {
        // a captured graph holds the previous launch shape
        if (P->graph_exec) {
            (void)hipGraphExecDestroy(P->graph_exec);
            P->graph_exec = nullptr;
        }
        if (P->graph) {
            (void)hipGraphDestroy(P->graph);
            P->graph = nullptr;
        }
        P->render_grouped = grouped != 0;
    }
    return RV_OK;
}

// A captured graph holds the launch structure, grid and bins it was captured
// with; every change of them drops it (rv_frame_graph_launch then fails until
// rv_frame_graph_build runs again).
static void drop_graph(rv_frame_plan* P)
{
    if (P->graph_exec) {
        (void)hipGraphExecDestroy(P->graph_exec);
        P->graph_exec = nullptr;
    }
    if (P->graph) {
        (void)hipGraphDestroy(P->graph);
        P->graph = nullptr;
    }
}

extern "C" int rv_frame_set_split(rv_frame_plan* P, int split)
{
    RV_REQUIRE(P, "rv_frame_set_split: null plan");
    if ((split != 0) != P->split) {
        drop_graph(P);
        P->split = split != 0;
    }
    return RV_OK;
}

extern "C" int rv_frame_set_exact(rv_frame_plan* P, int exact)
{
    RV_REQUIRE(P, "rv_frame_set_exact: null plan");
    if ((exact != 0) != (P->desc.exact != 0u)) {
        drop_graph(P);
        P->desc.exact = exact != 0 ? 1u : 0u;
    }
    return RV_OK;
}

extern "C" int rv_frame_get_split(const rv_frame_plan* P, int* split)
{
    RV_REQUIRE(P && split, "rv_frame_get_split: null pointer");
    *split = P->split ? 1 : 0;
    return RV_OK;
}

// Autotuner: time every candidate wave footprint of the screen bins on the
// plan's camera 0 (hip events, median of 3 after a warm-up) and keep the
// fastest; the bins of the other footprints are released.  Results
// are cached in a small JSON file keyed by device name and workload
// signature, so later runs reuse the choice without timing.
extern "C" int rv_frame_autotune(rv_frame_plan* P, uint32_t* d_argb, int interpolate,
                                 const char* cache_path, rv_stream_t stream, uint32_t* tile_w,
                                 uint32_t* tile_h)
{
    RV_TRY
    RV_REQUIRE(P && d_argb && P->ncams > 0, "rv_frame_autotune: bad arguments");
    RV_REQUIRE(!interpolate || P->can_interp, "rv_frame_autotune: blocks have no sparse values/plans");
    // tuning changes the grid and footprint and releases the bins of the
    // footprints not chosen, which a graph captured before may hold
    drop_graph(P);
    static const uint32_t cand[][2] = {{8, 8}, {16, 4}, {4, 16}, {32, 2}};
    int dev = 0;
    RV_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    RV_HIP(hipGetDeviceProperties(&prop, dev));
    std::ostringstream sig;
    sig << prop.gcnArchName << "|bins6|" << P->nblocks << "|" << P->desc.vp_w << "x" << P->desc.vp_h
        << "|" << (P->fp16 ? "f16" : "f32") << "|interp" << interpolate << "|d" << P->blocks[0].depth
        << "|mode" << P->field_mode << "|layout" << P->desc.layout << "|small" << (P->interp_small ? 1 : 0);
    const std::string key = sig.str();
    if (cache_path && *cache_path) {
        std::ifstream f(cache_path);
        std::string line;
        while (std::getline(f, line)) {
            // one entry per line: {"key": "...", "tile": [w, h], ...}
            const std::string k = "\"key\": \"" + key + "\"";
            if (line.find(k) == std::string::npos)
                continue;
            unsigned w = 0, h = 0, g = 0;
            const size_t t = line.find("\"tile\": [");
            const size_t gp = line.find("\"wgs_per_cu\": ");
            if (t != std::string::npos && std::sscanf(line.c_str() + t + 9, "%u, %u", &w, &h) == 2 &&
                valid_footprint(w, h) && gp != std::string::npos &&
                std::sscanf(line.c_str() + gp + 14, "%u", &g) == 1 && g >= 1 &&
                g <= (interpolate ? P->grid_max_wgs : P->render_max_wgs)) {
                unsigned rw = 0, sch = 0, hmin = 0;
                const size_t sp = line.find("\"schedule\": ");
                const size_t hp = line.find("\"hybrid_min\": ");
                if (hp != std::string::npos)
                    (void)std::sscanf(line.c_str() + hp + 14, "%u", &hmin);
                if (sp != std::string::npos && std::sscanf(line.c_str() + sp + 12, "%u", &sch) == 1 &&
                    (sch == RV_FRAME_SCHEDULE_BINS || P->field_mode == RV_FIELD_ARGB) &&
                    rv_frame_set_schedule(P, sch) == RV_OK)
                    (void)rv_frame_set_hybrid_min(P, hmin);
                const size_t spl = line.find("\"split\": ");
                unsigned split = 0;
                if (interpolate && spl != std::string::npos &&
                    std::sscanf(line.c_str() + spl + 9, "%u", &split) == 1)
                    (void)rv_frame_set_split(P, (int)split); // drops a graph of the other structure
                const size_t gr = line.find("\"grouped\": ");
                unsigned grouped = 0;
                if (!interpolate && gr != std::string::npos &&
                    std::sscanf(line.c_str() + gr + 12, "%u", &grouped) == 1)
                    P->render_grouped = grouped != 0;
                const size_t rp = line.find("\"render_waves\": ");
                if (interpolate && rp != std::string::npos &&
                    std::sscanf(line.c_str() + rp + 16, "%u", &rw) == 1 && rw >= 1 && rw <= kT / 64u)
                    P->frame_rw = rw;
                P->tw = w;
                P->th = h;
                (interpolate ? P->grid : P->render_grid) = P->cus * g;
                if (tile_w)
                    *tile_w = w;
                if (tile_h)
                    *tile_h = h;
                return RV_OK;
            }
        }
    }
    hipStream_t s = S(stream);
    hipEvent_t e0, e1;
    RV_HIP(hipEventCreate(&e0));
    RV_HIP(hipEventCreate(&e1));
    uint32_t& grid = interpolate ? P->grid : P->render_grid;
    const uint32_t gmax = interpolate ? P->grid_max_wgs : P->render_max_wgs;
    std::vector<uint32_t> wgs;
    for (uint32_t g : {1u, 2u, 3u, 4u, 6u, 8u, gmax})
        if (g <= gmax && std::find(wgs.begin(), wgs.end(), g) == wgs.end())
            wgs.push_back(g);
    std::vector<uint32_t> rws = {kT / 64u};
    if (interpolate)
        rws = {4u, 6u, kT / 64u};
    // schedules: bin walk, per-lane walk, and the hybrid with the per-lane
    // walk for bins listing at least 48 blocks, as indices into kSched
    static const uint32_t kSched[3][2] = {{RV_FRAME_SCHEDULE_BINS, 0u},
                                          {RV_FRAME_SCHEDULE_LANES, 0u},
                                          {RV_FRAME_SCHEDULE_HYBRID, 48u}};
    std::vector<uint32_t> scheds = {0u};
    if (P->field_mode == RV_FIELD_ARGB)
        scheds = {0u, 1u, 2u};
    std::ostringstream times;
    if (interpolate)
        (void)rv_frame_set_split(P, 0); // the candidates below are one-launch frames
    // median of 3 timed frames of camera cam in a configuration
    auto time_cfg = [&](uint32_t sch, uint32_t tw, uint32_t th, uint32_t g, uint32_t rw, uint32_t grp,
                        uint32_t cam, float* ms) -> int {
        grid = P->cus * g;
        if (!interpolate)
            P->render_grouped = grp != 0;
        if (interpolate)
            P->frame_rw = rw;
        (void)rv_frame_set_schedule(P, kSched[sch][0]);
        (void)rv_frame_set_hybrid_min(P, kSched[sch][1]);
        float t[3];
        int rc = launch_frame(P, cam, d_argb, nullptr, interpolate, s, tw, th);
        for (int r = 0; r < 3 && !rc; ++r) {
            (void)hipEventRecord(e0, s);
            rc = launch_frame(P, cam, d_argb, nullptr, interpolate, s, tw, th);
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(&t[r], e0, e1);
        }
        std::sort(t, t + 3);
        *ms = t[1];
        return rc;
    };
    struct Choice {
        float ms = 1e30f;
        uint32_t w = 8, h = 8, g = 1, rw = kT / 64u, grp = 0;
    } pick[6]; // per schedule and bin handout (2 sch + grp)
    // resident frames: the queue's persistent grid and the grouped bins
    std::vector<uint32_t> grps = {0u};
    if (!interpolate)
        grps = {0u, 1u};
    for (uint32_t sch : scheds)
    for (const auto& c : cand)
    for (uint32_t g : wgs)
    for (uint32_t rw : rws)
    for (uint32_t grp : grps) {
        float ms = 0.f;
        const int rc = time_cfg(sch, c[0], c[1], g, rw, grp, 0, &ms);
        if (rc) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return rc;
        }
        static const char* kName[3] = {"B", "L", "H48/"};
        times << (times.tellp() > 0 ? ", " : "") << "\"" << kName[sch] << c[0] << "x" << c[1] << "/" << g;
        if (interpolate)
            times << "/" << rw;
        if (grp)
            times << "/grouped";
        times << "\": " << ms;
        if (ms < pick[2 * sch + grp].ms)
            pick[2 * sch + grp] = Choice{ms, c[0], c[1], g, rw, grp};
    }
    // The best configuration of each schedule and bin handout on camera 0
    // is compared again over up to 4 cameras spread through the plan's set:
    // how much a per-lane walk or the grouped bins gain depends on the view.
    std::vector<uint32_t> slots;
    for (uint32_t sch : scheds)
        for (uint32_t grp : grps)
            slots.push_back(2 * sch + grp);
    uint32_t bslot = slots[0];
    for (uint32_t k : slots)
        if (pick[k].ms < pick[bslot].ms)
            bslot = k;
    std::vector<uint32_t> cams;
    for (uint32_t q = 0; q < 4; ++q) {
        const uint32_t cam = (uint32_t)((uint64_t)P->ncams * q / 4);
        if (std::find(cams.begin(), cams.end(), cam) == cams.end())
            cams.push_back(cam);
    }
    if (slots.size() > 1 && P->ncams > 1) {
        float tot[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (uint32_t k : slots)
            for (uint32_t cam : cams) {
                float ms = 0.f;
                const Choice& c = pick[k];
                const int rc = time_cfg(k / 2, c.w, c.h, c.g, c.rw, c.grp, cam, &ms);
                if (rc) {
                    (void)hipEventDestroy(e0);
                    (void)hipEventDestroy(e1);
                    return rc;
                }
                tot[k] += ms;
            }
        bslot = slots[0];
        for (uint32_t k : slots) {
            times << ", \"schedule " << k / 2 << (k % 2 ? " grouped" : "") << " over " << cams.size()
                  << " cameras\": " << tot[k];
            if (tot[k] < tot[bslot])
                bslot = k;
        }
    }
    const uint32_t bsched = bslot / 2;
    const Choice& bc = pick[bslot];
    const uint32_t bw = bc.w, bh = bc.h;
    if (interpolate) {
        // the chosen one-launch frame against the same frames split in two
        // launches (interpolation, then the resident kernel with its grid)
        float tot[2] = {0.f, 0.f};
        for (int split = 0; split < 2; ++split) {
            (void)rv_frame_set_split(P, split);
            for (uint32_t cam : cams) {
                float ms = 0.f;
                const int rc = time_cfg(bsched, bw, bh, bc.g, bc.rw, bc.grp, cam, &ms);
                if (rc) {
                    (void)rv_frame_set_split(P, 0);
                    (void)hipEventDestroy(e0);
                    (void)hipEventDestroy(e1);
                    return rc;
                }
                tot[split] += ms;
            }
        }
        (void)rv_frame_set_split(P, tot[1] < tot[0] ? 1 : 0);
        times << ", \"one launch over " << cams.size() << " cameras\": " << tot[0] << ", \"split\": " << tot[1];
    }
    grid = P->cus * bc.g;
    if (interpolate)
        P->frame_rw = bc.rw;
    else
        P->render_grouped = bc.grp != 0;
    (void)rv_frame_set_schedule(P, kSched[bsched][0]);
    (void)rv_frame_set_hybrid_min(P, kSched[bsched][1]);
    const uint32_t bg = bc.g;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    P->tw = bw;
    P->th = bh;
    P->bins.erase(std::remove_if(P->bins.begin(), P->bins.end(),
                                 [&](const std::unique_ptr<FrameBins>& b) { return b->bw != bw || b->bh != bh; }),
                  P->bins.end());
    if (tile_w)
        *tile_w = bw;
    if (tile_h)
        *tile_h = bh;
    if (cache_path && *cache_path) {
        std::ofstream f(cache_path, std::ios::app);
        f << "{\"key\": \"" << key << "\", \"tile\": [" << bw << ", " << bh
          << "], \"wgs_per_cu\": " << bg << ", \"render_waves\": " << P->frame_rw
          << ", \"schedule\": " << P->schedule << ", \"hybrid_min\": " << P->hybrid_min
          << ", \"grouped\": " << (P->render_grouped ? 1 : 0)
          << ", \"split\": " << (P->split ? 1 : 0) << ", \"ms\": {"
          << times.str() << "}}\n";
    }
    return RV_OK;
    RV_CATCH
}

extern "C" int rv_frame_set_secondary(rv_frame_plan* P, const float* const* d_sparse2,
                                      void* const* d_bricks2, float smin, float smax)
{
    RV_TRY
    RV_REQUIRE(P, "rv_frame_set_secondary: null plan");
    if (P->graph_exec) { // a captured graph holds the previous tables
        (void)hipGraphExecDestroy(P->graph_exec);
        P->graph_exec = nullptr;
    }
    if (P->graph) {
        (void)hipGraphDestroy(P->graph);
        P->graph = nullptr;
    }
    if (!d_bricks2) {
        RV_REQUIRE(P->field_mode != RV_FIELD_VALUE_ON_ISOSURFACE &&
                       P->field_mode != RV_FIELD_VALUE_ON_ISOSURFACES,
                   "rv_frame_set_secondary: the plan renders a secondary-field mode");
        for (void** q : {(void**)&P->d_bricks2, (void**)&P->d_tasks2})
            if (*q) {
                (void)hipFree(*q);
                *q = nullptr;
            }
        P->n_interp2 = 0;
        return RV_OK;
    }
    RV_REQUIRE(smax > smin, "rv_frame_set_secondary: the secondary field range must be non-empty");
    std::vector<const void*> bricks(P->nblocks);
    for (uint32_t i = 0; i < P->nblocks; ++i) {
        RV_REQUIRE(d_bricks2[i], "rv_frame_set_secondary: null secondary brick");
        bricks[i] = d_bricks2[i];
    }
    int rc;
    if (d_sparse2) {
        RV_REQUIRE(P->can_interp, "rv_frame_set_secondary: the blocks have no gather plans");
        // the primary field's jobs, then the secondary's: same plans, depths
        std::vector<Job> jobs(2 * (size_t)P->nblocks);
        for (uint32_t i = 0; i < P->nblocks; ++i) {
            const rv_frame_block& b = P->blocks[i];
            RV_REQUIRE(d_sparse2[i], "rv_frame_set_secondary: null secondary sparse values");
            RV_REQUIRE(!P->tiled || (reinterpret_cast<uintptr_t>(d_bricks2[i]) & 15u) == 0u,
                       "rv_frame_set_secondary: TILED4 bricks must be 16-byte aligned");
            jobs[i] = Job{b.sparse, static_cast<const uint8_t*>(b.plan), b.brick,
                          reinterpret_cast<__half*>(b.brick16), b.depth, P->tiled ? 1u : 0u};
            jobs[P->nblocks + i] =
                Job{d_sparse2[i], static_cast<const uint8_t*>(b.plan),
                    P->fp16 ? nullptr : static_cast<float*>(d_bricks2[i]),
                    P->fp16 ? static_cast<__half*>(d_bricks2[i]) : nullptr, b.depth, P->tiled ? 1u : 0u};
        }
        std::vector<TileTask> tasks;
        if ((rc = build_tile_tasks(jobs.data(), (uint32_t)jobs.size(), tasks)))
            return rc;
        RV_REQUIRE(tasks.size() < 0x40000000ull, "rv_frame_set_secondary: too many brick tiles");
        P->n_interp2 = (uint32_t)tasks.size();
        if ((rc = frame_upload((void**)&P->d_tasks2, tasks.data(), sizeof(TileTask) * tasks.size())))
            return rc;
    } else if (P->d_tasks2) {
        (void)hipFree(P->d_tasks2);
        P->d_tasks2 = nullptr;
        P->n_interp2 = 0;
    }
    if ((rc = frame_upload((void**)&P->d_bricks2, bricks.data(), sizeof(void*) * P->nblocks)))
        return rc;
    P->smin = smin;
    P->smax = smax;
    return RV_OK;
    RV_CATCH
}

extern "C" int rv_frame_set_field_mode(rv_frame_plan* P, uint32_t field_mode, float iso_opacity,
                                       const float* levels_rel, uint32_t level_count)
{
    RV_TRY
    RV_REQUIRE(P, "rv_frame_set_field_mode: null plan");
    RV_REQUIRE(levels_rel || level_count == 0, "rv_frame_set_field_mode: null levels");
    int per_cu = 0, rc;
    switch (field_mode) {
    case RV_FIELD_ARGB:
        rc = frame_max_wgs<kModeArgb>(P->fp16, P->tiled, &per_cu);
        break;
    case RV_FIELD_ARGB_LIGHT:
        rc = frame_max_wgs<kModeArgbLight>(P->fp16, P->tiled, &per_cu);
        break;
    case RV_FIELD_ISOSURFACE:
        rc = frame_max_wgs<kModeIsosurface>(P->fp16, P->tiled, &per_cu);
        break;
    case RV_FIELD_MAX_INTENSITY_PROJECTION:
        rc = frame_max_wgs<kModeMip>(P->fp16, P->tiled, &per_cu);
        break;
    case RV_FIELD_ISOSURFACES:
        rc = frame_max_wgs<kModeIsosurfaces>(P->fp16, P->tiled, &per_cu);
        break;
    case RV_FIELD_DOMAIN_VOXELS:
        rc = frame_max_wgs<kModeDomainVoxels>(P->fp16, P->tiled, &per_cu);
        break;
    case RV_FIELD_VALUE_ON_ISOSURFACE:
        RV_REQUIRE(P->d_bricks2, "rv_frame_set_field_mode: set the secondary field first (rv_frame_set_secondary)");
        rc = frame_max_wgs<kModeValueOnIsosurface>(P->fp16, P->tiled, &per_cu);
        break;
    case RV_FIELD_VALUE_ON_ISOSURFACES:
        RV_REQUIRE(P->d_bricks2, "rv_frame_set_field_mode: set the secondary field first (rv_frame_set_secondary)");
        rc = frame_max_wgs<kModeValueOnIsosurfaces>(P->fp16, P->tiled, &per_cu);
        break;
    default:
        return fail(RV_ERR_INVALID, "rv_frame_set_field_mode: unknown field mode");
    }
    if (rc)
        return rc;
    RV_REQUIRE(per_cu >= 1, "rv_frame_set_field_mode: frame kernel cannot be resident");
    if (P->graph_exec) { // a captured graph holds the previous mode's kernels
        (void)hipGraphExecDestroy(P->graph_exec);
        P->graph_exec = nullptr;
    }
    if (P->graph) {
        (void)hipGraphDestroy(P->graph);
        P->graph = nullptr;
    }
    P->field_mode = field_mode;
    P->iso_opacity = iso_opacity;
    P->levels = make_levels(P->desc.fmin, P->desc.fmax, levels_rel, level_count);
    P->mode_max_wgs = (uint32_t)per_cu;
    P->tape_cam = -1;
    return RV_OK;
    RV_CATCH
}

extern "C" void rv_frame_plan_destroy(rv_frame_plan* P) { delete P; }

// The share kernel over a staged block table (device params, view
// rectangles {x0, y0, x1, y1}, brick and secondary-brick pointers).
static int launch_share(const rv_block_params* dp, const uint4* dr, const void* const* db,
                        const void* const* db2, const uint32_t* bin_off, const uint32_t* bin_list,
                        const uint32_t* bin_order,
                        uint32_t nblocks, uint32_t layout, uint32_t field_mode,
                        bool exact, float iso_opacity, const LevelsArg& lev, float smin, float smax,
                        const float* d_tf, uint32_t tf_n, uint32_t* d_viewport, uint32_t vp_w, uint32_t vp_h,
                        rv_stream_t stream, int* d_rect = nullptr, uint32_t init = ~0u)
{
    // LDS: the colour table, padded to cap the workgroups per CU (64 KiB of
    // LDS per workgroup at most)
    // (the static LDS, mostly the front waves' segments, counted in)
    constexpr size_t kStatic = (sizeof(float4) * 64u + 8u) * kShareSquares * kShareSlots +
                               sizeof(uint2) * 2u * kYZTabEntries + 4u * kMaxLevels;
    constexpr size_t kPerWg = (160u * 1024u) / kShareWgsPerCu - 1024u;
    static_assert(kPerWg > kStatic, "the static LDS leaves room for the colour table");
    const size_t lds = std::max<size_t>(sizeof(float4) * (tf_n + 1u), std::min<size_t>(65536, kPerWg - kStatic));
    const auto* tf4 = reinterpret_cast<const float4*>(d_tf);
    const dim3 grid((vp_w + 15) / 16, kShareParts * ((vp_h + 15) / 16)); // kShareParts per 16 x 16 bin
#define RV_SHARE_LAUNCH(M)                                                                         \
    do {                                                                                           \
        if (layout == RV_BRICK_TILED4)                                                             \
            hipLaunchKernelGGL((share_kernel<M, Tiled4>), grid, dim3(kShareT), lds, S(stream), dp, dr, \
                               db, db2, bin_off, bin_list, bin_order, smin, smax, nblocks, tf4, tf_n, iso_opacity, lev, d_viewport, \
                               vp_w, vp_h, d_rect, init);                                                \
        else                                                                                       \
            hipLaunchKernelGGL((share_kernel<M, float>), grid, dim3(kShareT), lds, S(stream), dp, dr, \
                               db, db2, bin_off, bin_list, bin_order, smin, smax, nblocks, tf4, tf_n, iso_opacity, lev, d_viewport, \
                               vp_w, vp_h, d_rect, init);                                                \
    } while (0)
    switch (field_mode) {
    case RV_FIELD_ARGB:
        if (exact)
            RV_SHARE_LAUNCH(kModeArgb);
        else
            RV_SHARE_LAUNCH(kModeArgbFast);
        break;
    case RV_FIELD_ARGB_LIGHT: RV_SHARE_LAUNCH(kModeArgbLight); break;
    case RV_FIELD_ISOSURFACE: RV_SHARE_LAUNCH(kModeIsosurface); break;
    case RV_FIELD_ISOSURFACES: RV_SHARE_LAUNCH(kModeIsosurfaces); break;
    case RV_FIELD_DOMAIN_VOXELS: RV_SHARE_LAUNCH(kModeDomainVoxels); break;
    case RV_FIELD_VALUE_ON_ISOSURFACE: RV_SHARE_LAUNCH(kModeValueOnIsosurface); break;
    case RV_FIELD_VALUE_ON_ISOSURFACES: RV_SHARE_LAUNCH(kModeValueOnIsosurfaces); break;
    default: RV_SHARE_LAUNCH(kModeMip); break;
    }
#undef RV_SHARE_LAUNCH
    RV_LAUNCH_CHECK();
    return RV_OK;
}

// Byte offset of the view rectangles in a staged block table: after the
// params, rounded up to 16 bytes (the 256 bytes of slack in
// rv_share_workspace_bytes cover the rounding).
static uint64_t share_rect_at(uint32_t nblocks)
{
    return (sizeof(rv_block_params) * (uint64_t)nblocks + 15u) & ~15ull;
}

extern "C" uint64_t rv_share_workspace_bytes(uint32_t nblocks)
{
    return (uint64_t)nblocks * (sizeof(rv_block_params) + sizeof(uint4) + 2 * sizeof(void*)) + 256;
}

// The share's block table: params, view rectangles {x0, y0, x1, y1}, brick
// and secondary-brick pointers, laid out as rv_share_workspace_bytes says,
// written to `host` after the arguments are checked.
static int share_table(const rv_block_params* params, const float* const* bricks, const float* const* bricks2,
                       uint32_t nblocks, uint32_t layout, uint32_t field_mode, const float* levels_rel,
                       uint32_t level_count, float smin, float smax, const float* d_tf, uint32_t tf_n,
                       const uint32_t* d_viewport, uint32_t vp_w, uint32_t vp_h, uint8_t* host,
                       uint64_t rect_at)
{
    RV_REQUIRE(layout == RV_BRICK_LINEAR || layout == RV_BRICK_TILED4, "rv_render_share: unknown layout");
    RV_REQUIRE(field_mode <= RV_FIELD_VALUE_ON_ISOSURFACES, "rv_render_share: unknown field mode");
    const bool value = field_mode == RV_FIELD_VALUE_ON_ISOSURFACE ||
                       field_mode == RV_FIELD_VALUE_ON_ISOSURFACES;
    RV_REQUIRE(!value || (bricks2 && smax > smin),
               "rv_render_share: the ValueOnIsosurface(s) modes need the secondary field's bricks "
               "and a non-empty range");
    RV_REQUIRE(levels_rel || level_count == 0, "rv_render_share: null levels");
    RV_REQUIRE(params && bricks && d_tf && d_viewport, "rv_render_share: null pointer");
    RV_REQUIRE(tf_n >= 1 && tf_n <= kMaxTf, "rv_render_share: tf_n must be in [1, 1024]");
    auto* hp = reinterpret_cast<rv_block_params*>(host);
    auto* hr = reinterpret_cast<uint4*>(host + rect_at);
    auto* hb = reinterpret_cast<const void**>(host + rect_at + sizeof(uint4) * nblocks);
    for (uint32_t i = 0; i < nblocks; ++i) {
        const rv_block_params& p = params[i];
        RV_REQUIRE(p.vp_w == vp_w && p.vp_h == vp_h, "rv_render_share: viewport mismatch");
        RV_REQUIRE(p.tex_depth <= 9 && p.nodes == (1u << p.tex_depth) + 1u && p.fmax > p.fmin,
                   "rv_render_share: inconsistent block parameters");
        RV_REQUIRE((uint64_t)p.view_x + p.view_w <= vp_w && (uint64_t)p.view_y + p.view_h <= vp_h,
                   "rv_render_share: view rectangle outside the viewport");
        RV_REQUIRE(bricks[i] || p.view_w == 0 || p.view_h == 0, "rv_render_share: null brick");
        RV_REQUIRE(!value || bricks2[i] || p.view_w == 0 || p.view_h == 0,
                   "rv_render_share: null secondary brick");
        RV_REQUIRE(p.fmin == params[0].fmin && p.fmax == params[0].fmax,
                   "rv_render_share: the blocks of a share have one field range");
        RV_REQUIRE(((p.flags ^ params[0].flags) & RV_RENDER_EXACT) == 0u,
                   "rv_render_share: the blocks of a share have one sample sequence (flags)");
        hp[i] = p;
        hr[i] = make_uint4(p.view_x, p.view_y, p.view_x + p.view_w, p.view_y + p.view_h);
        hb[i] = bricks[i];
        hb[nblocks + i] = value ? bricks2[i] : nullptr;
    }
    return RV_OK;
}

// The staged table from pinned host memory into device memory by the
// compute queue itself (a kernel reading host memory): the share launch
// that follows waits for no copy engine.
// The staged table into device memory; rect (or null) is set to the empty
// crop rectangle the share kernel then shrinks.
__global__ __launch_bounds__(256) static void stage_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                                uint64_t n16, int* __restrict__ rect)
{
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * 256u)
        dst[i] = src[i];
    if (rect && blockIdx.x == 0u && threadIdx.x < 4u)
        rect[threadIdx.x] = threadIdx.x < 2u ? 0x7fffffff : -1;
}

// The share launch over a table staged in pinned host memory and copied in
// stream order, with screen bins: the blocks whose view rectangles meet each
// 16 x 16-pixel bin, in table (back-to-front) order, so a workgroup visits
// only its bin's blocks.  Layout of the staged bytes: params, view
// rectangles (16-byte aligned), brick and secondary-brick pointers, bin
// offsets (nbins + 1), bin lists.  No synchronisation but the wait for the
// previous copy out of the staging buffer.
int rv::render_share_staged(const rv_block_params* params, const float* const* bricks,
                            const float* const* bricks2, uint32_t nblocks, uint32_t layout, uint32_t field_mode,
                            float iso_opacity, const float* levels_rel, uint32_t level_count, float smin,
                            float smax, const float* d_tf, uint32_t tf_n, uint32_t* d_viewport, uint32_t vp_w,
                            uint32_t vp_h, ShareStaging& st, rv_stream_t stream, int* d_rect, bool* rect_done,
                            uint32_t init, bool record_copied)
{
    RV_TRY
    if (rect_done)
        *rect_done = false;
    if (nblocks == 0)
        return RV_OK;
    const uint32_t gx = (vp_w + 15u) / 16u, gy = (vp_h + 15u) / 16u;
    const uint64_t nbins = (uint64_t)gx * gy;
    // bin counts first: the staged size depends on them
    std::vector<uint32_t> off(nbins + 1u, 0u);
    for (uint32_t i = 0; i < nblocks; ++i) {
        const rv_block_params& p = params[i];
        if (p.view_w == 0 || p.view_h == 0 || p.vp_w != vp_w || p.vp_h != vp_h ||
            (uint64_t)p.view_x + p.view_w > vp_w || (uint64_t)p.view_y + p.view_h > vp_h)
            continue; // empty, or rejected by share_table below
        for (uint32_t by = p.view_y / 16u; by <= (p.view_y + p.view_h - 1u) / 16u; ++by)
            for (uint32_t bx = p.view_x / 16u; bx <= (p.view_x + p.view_w - 1u) / 16u; ++bx)
                ++off[(uint64_t)by * gx + bx + 1u];
    }
    for (uint64_t b = 0; b < nbins; ++b)
        off[b + 1u] += off[b];
    const uint64_t entries = off[nbins];
    const uint64_t rect_at = share_rect_at(nblocks);
    const uint64_t off_at = rect_at + (sizeof(uint4) + 2u * sizeof(void*)) * (uint64_t)nblocks;
    const uint64_t list_at = off_at + 4u * (nbins + 1u);
    const uint64_t order_at = list_at + 4u * entries;
    const uint64_t bytes = (order_at + 4u * nbins + 15u) & ~15ull; // copied as 16-byte words
    if (!st.copied)
        RV_HIP(hipEventCreateWithFlags(&st.copied, hipEventDisableTiming));
    // the previous copy out of the staging buffer has run (normally long
    // since: every worker part ends with a stream synchronisation)
    RV_HIP(hipEventSynchronize(st.copied));
    if (bytes > st.cap) {
        RV_HIP(hipStreamSynchronize(S(stream))); // no launch still reads the device table
        if (st.dev)
            (void)hipFree(st.dev);
        if (st.host)
            (void)hipHostFree(st.host);
        st.dev = st.host = nullptr;
        st.cap = 0;
        const uint64_t cap = bytes + bytes / 4u + 4096u;
        RV_HIP(hipMalloc(&st.dev, cap));
        RV_HIP(hipHostMalloc(&st.host, cap, hipHostMallocDefault));
        RV_HIP(hipHostGetDevicePointer(&st.host_dev, st.host, 0));
        st.cap = cap;
    }
    uint8_t* host = static_cast<uint8_t*>(st.host);
    int rc = share_table(params, bricks, bricks2, nblocks, layout, field_mode, levels_rel, level_count, smin, smax,
                         d_tf, tf_n, d_viewport, vp_w, vp_h, host, rect_at);
    if (rc)
        return rc;
    auto* hoff = reinterpret_cast<uint32_t*>(host + off_at);
    auto* hlist = reinterpret_cast<uint32_t*>(host + list_at);
    std::memcpy(hoff, off.data(), 4u * (nbins + 1u));
    // the bins by decreasing block count (a counting sort; equal counts in
    // bin order), the order the share kernel's workgroups take them in
    {
        uint32_t maxc = 0;
        for (uint64_t b = 0; b < nbins; ++b)
            maxc = std::max(maxc, off[b + 1] - off[b]);
        std::vector<uint32_t> start(maxc + 2u, 0u);
        for (uint64_t b = 0; b < nbins; ++b)
            ++start[maxc - (off[b + 1] - off[b]) + 1u];
        for (uint32_t c = 0; c <= maxc; ++c)
            start[c + 1u] += start[c];
        auto* horder = reinterpret_cast<uint32_t*>(host + order_at);
        for (uint64_t b = 0; b < nbins; ++b)
            horder[start[maxc - (off[b + 1] - off[b])]++] = (uint32_t)b;
    }
    for (uint32_t i = 0; i < nblocks; ++i) {
        const rv_block_params& p = params[i];
        if (p.view_w == 0 || p.view_h == 0)
            continue;
        for (uint32_t by = p.view_y / 16u; by <= (p.view_y + p.view_h - 1u) / 16u; ++by)
            for (uint32_t bx = p.view_x / 16u; bx <= (p.view_x + p.view_w - 1u) / 16u; ++bx)
                hlist[off[(uint64_t)by * gx + bx]++] = i;
    }
    uint8_t* d = static_cast<uint8_t*>(st.dev);
    const uint64_t n16 = bytes / 16u;
    hipLaunchKernelGGL(stage_copy_kernel, dim3((uint32_t)std::min<uint64_t>(256u, (n16 + 255u) / 256u)), dim3(256),
                       0, S(stream), static_cast<const uint4*>(st.host_dev), reinterpret_cast<uint4*>(d), n16, d_rect);
    RV_LAUNCH_CHECK();
    const auto* dp = reinterpret_cast<const rv_block_params*>(d);
    const auto* dr = reinterpret_cast<const uint4*>(d + rect_at);
    const auto* db = reinterpret_cast<const void* const*>(d + rect_at + sizeof(uint4) * (uint64_t)nblocks);
    const LevelsArg lev = make_levels(params[0].fmin, params[0].fmax, levels_rel, level_count);
    rc = launch_share(dp, dr, db, db + nblocks, reinterpret_cast<const uint32_t*>(d + off_at),
                        reinterpret_cast<const uint32_t*>(d + list_at), reinterpret_cast<const uint32_t*>(d + order_at),
                        nblocks, layout, field_mode,
                        (params[0].flags & RV_RENDER_EXACT) != 0u, iso_opacity, lev, smin, smax, d_tf, tf_n,
                        d_viewport, vp_w, vp_h, stream, d_rect, init);
    // recorded after the share launch, not between the two launches (a
    // marker there delays the share kernel's start by several us); the
    // next call reuses the staging buffer only after it
    if (record_copied)
        RV_HIP(hipEventRecord(st.copied, S(stream)));
    if (rc == RV_OK && rect_done)
        *rect_done = d_rect != nullptr;
    return rc;
    RV_CATCH
}

void rv::ShareStaging::release()
{
    if (dev)
        (void)hipFree(dev);
    if (host)
        (void)hipHostFree(host);
    if (copied)
        (void)hipEventDestroy(copied);
    dev = host = host_dev = nullptr;
    copied = nullptr;
    cap = 0;
}

extern "C" int rv_render_share_layout(const rv_block_params* params, const float* const* bricks,
                                      const float* const* bricks2, uint32_t nblocks, uint32_t layout,
                                      uint32_t field_mode, float iso_opacity, const float* levels_rel,
                                      uint32_t level_count, float smin, float smax, const float* d_tf,
                                      uint32_t tf_n, uint32_t* d_viewport, uint32_t vp_w, uint32_t vp_h,
                                      void* d_workspace, rv_stream_t stream)
{
    RV_TRY
    RV_REQUIRE(layout == RV_BRICK_LINEAR || layout == RV_BRICK_TILED4, "rv_render_share: unknown layout");
    RV_REQUIRE(field_mode <= RV_FIELD_VALUE_ON_ISOSURFACES, "rv_render_share: unknown field mode");
    if (nblocks == 0)
        return RV_OK;
    RV_REQUIRE(d_workspace, "rv_render_share: null pointer");
    std::vector<uint8_t> host(rv_share_workspace_bytes(nblocks));
    int rc = share_table(params, bricks, bricks2, nblocks, layout, field_mode, levels_rel, level_count, smin, smax,
                         d_tf, tf_n, d_viewport, vp_w, vp_h, host.data(),
                         share_rect_at(nblocks));
    if (rc)
        return rc;
    uint8_t* d = static_cast<uint8_t*>(d_workspace);
    RV_HIP(hipMemcpyAsync(d, host.data(), share_rect_at(nblocks) + (sizeof(uint4) + 2 * sizeof(void*)) * nblocks,
                          hipMemcpyHostToDevice, S(stream)));
    // the host staging buffer must outlive the copy
    RV_HIP(hipStreamSynchronize(S(stream)));
    const auto* dp = reinterpret_cast<const rv_block_params*>(d);
    const auto* dr = reinterpret_cast<const uint4*>(d + share_rect_at(nblocks));
    const auto* db = reinterpret_cast<const void* const*>(d + share_rect_at(nblocks) + sizeof(uint4) * nblocks);
    // the levels are relative to the field's range, the same for every block
    const LevelsArg lev = make_levels(params[0].fmin, params[0].fmax, levels_rel, level_count);
    return launch_share(dp, dr, db, db + nblocks, nullptr, nullptr, nullptr, nblocks, layout, field_mode,
                        (params[0].flags & RV_RENDER_EXACT) != 0u, iso_opacity, lev, smin, smax, d_tf, tf_n,
                        d_viewport, vp_w, vp_h, stream);
    RV_CATCH
}

extern "C" int rv_render_share_secondary(const rv_block_params* params, const float* const* bricks,
                                         const float* const* bricks2, uint32_t nblocks,
                                         uint32_t field_mode, float iso_opacity,
                                         const float* levels_rel, uint32_t level_count, float smin,
                                         float smax, const float* d_tf, uint32_t tf_n,
                                         uint32_t* d_viewport, uint32_t vp_w, uint32_t vp_h,
                                         void* d_workspace, rv_stream_t stream)
{
    return rv_render_share_layout(params, bricks, bricks2, nblocks, RV_BRICK_LINEAR, field_mode,
                                  iso_opacity, levels_rel, level_count, smin, smax, d_tf, tf_n,
                                  d_viewport, vp_w, vp_h, d_workspace, stream);
}

extern "C" int rv_render_share(const rv_block_params* params, const float* const* bricks,
                               uint32_t nblocks, uint32_t field_mode, float iso_opacity,
                               const float* levels_rel, uint32_t level_count, const float* d_tf,
                               uint32_t tf_n, uint32_t* d_viewport, uint32_t vp_w, uint32_t vp_h,
                               void* d_workspace, rv_stream_t stream)
{
    return rv_render_share_secondary(params, bricks, nullptr, nblocks, field_mode, iso_opacity,
                                     levels_rel, level_count, 0.f, 1.f, d_tf, tf_n, d_viewport, vp_w,
                                     vp_h, d_workspace, stream);
}

extern "C" int rv_render_share_argb(const rv_block_params* params, const float* const* bricks,
                                    uint32_t nblocks, const float* d_tf, uint32_t tf_n,
                                    uint32_t* d_viewport, uint32_t vp_w, uint32_t vp_h,
                                    void* d_workspace, rv_stream_t stream)
{
    return rv_render_share(params, bricks, nblocks, RV_FIELD_ARGB, 1.f, nullptr, 0, d_tf, tf_n,
                           d_viewport, vp_w, vp_h, d_workspace, stream);
}

// ---------------------------------------------------------------- share without allocation
// rv_render_share_argb for callers that cannot allocate: the block table
// lives in `scratch`, a fixed array of the library on every GPU, laid out as
// rv_share_workspace_bytes(RV_SHARE_STATIC_MAX_BLOCKS) describes (params,
// then view rectangles, then brick pointers).  Host code never passes the
// array's name to a kernel or a copy: it takes the array's device address on
// each GPU once, with hipGetSymbolAddress, and uses that.

static __device__ __align__(256) uint8_t
    scratch[(uint64_t)RV_SHARE_STATIC_MAX_BLOCKS * (sizeof(rv_block_params) + sizeof(uint4) + 2 * sizeof(void*)) +
            256];

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch[kScratchDevices]; // per device: the address of `scratch` there

// The current device's address of `scratch`, looked up on the first call on
// that device (the device is current then).
int scratch_address(uint8_t** out)
{
    int dev = 0;
    RV_HIP(hipGetDevice(&dev));
    RV_REQUIRE(dev >= 0 && dev < kScratchDevices, "rv_render_share_argb_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        RV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        RV_REQUIRE(p, "rv_render_share_argb_noalloc: no address for the scratch array");
        g_scratch[dev] = static_cast<uint8_t*>(p);
    }
    *out = g_scratch[dev];
    return RV_OK;
}

// View rectangles {x0, y0, x1, y1} of the staged params.
__global__ void share_rects_kernel(const rv_block_params* __restrict__ p, uint4* __restrict__ r, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        r[i] = make_uint4(p[i].view_x, p[i].view_y, p[i].view_x + p[i].view_w, p[i].view_y + p[i].view_h);
}

} // namespace

extern "C" int rv_render_share_argb_noalloc(const rv_block_params* params, const float* const* bricks,
                                            uint32_t nblocks, const float* d_tf, uint32_t tf_n,
                                            uint32_t* d_viewport, uint32_t vp_w, uint32_t vp_h,
                                            rv_stream_t stream)
{
    RV_TRY
    if (nblocks == 0)
        return RV_OK;
    RV_REQUIRE(nblocks <= RV_SHARE_STATIC_MAX_BLOCKS,
               "rv_render_share_argb_noalloc: more blocks than RV_SHARE_STATIC_MAX_BLOCKS (4096); "
               "use rv_render_share_argb with a workspace");
    RV_REQUIRE(params && bricks && d_tf && d_viewport, "rv_render_share_argb_noalloc: null pointer");
    RV_REQUIRE(tf_n >= 1 && tf_n <= kMaxTf, "rv_render_share_argb_noalloc: tf_n must be in [1, 1024]");
    for (uint32_t i = 0; i < nblocks; ++i) {
        const rv_block_params& p = params[i];
        RV_REQUIRE(p.vp_w == vp_w && p.vp_h == vp_h, "rv_render_share_argb_noalloc: viewport mismatch");
        RV_REQUIRE(p.tex_depth <= 9 && p.nodes == (1u << p.tex_depth) + 1u && p.fmax > p.fmin,
                   "rv_render_share_argb_noalloc: inconsistent block parameters");
        RV_REQUIRE((uint64_t)p.view_x + p.view_w <= vp_w && (uint64_t)p.view_y + p.view_h <= vp_h,
                   "rv_render_share_argb_noalloc: view rectangle outside the viewport");
        RV_REQUIRE(bricks[i] || p.view_w == 0 || p.view_h == 0, "rv_render_share_argb_noalloc: null brick");
        RV_REQUIRE(p.fmin == params[0].fmin && p.fmax == params[0].fmax,
                   "rv_render_share_argb_noalloc: the blocks of a share have one field range");
        RV_REQUIRE(((p.flags ^ params[0].flags) & RV_RENDER_EXACT) == 0u,
                   "rv_render_share_argb_noalloc: the blocks of a share have one sample sequence (flags)");
    }
    uint8_t* base = nullptr;
    int rc = scratch_address(&base);
    if (rc)
        return rc;
    auto* dp = reinterpret_cast<rv_block_params*>(base);
    auto* dr = reinterpret_cast<uint4*>(base + share_rect_at(nblocks));
    auto* db = reinterpret_cast<const void**>(base + share_rect_at(nblocks) + sizeof(uint4) * nblocks);
    RV_HIP(hipMemcpyAsync(dp, params, sizeof(rv_block_params) * nblocks, hipMemcpyHostToDevice, S(stream)));
    RV_HIP(hipMemcpyAsync(db, bricks, sizeof(void*) * nblocks, hipMemcpyHostToDevice, S(stream)));
    // the caller's host arrays must outlive the copies
    RV_HIP(hipStreamSynchronize(S(stream)));
    hipLaunchKernelGGL(share_rects_kernel, dim3((nblocks + 255) / 256), dim3(256), 0, S(stream), dp, dr, nblocks);
    RV_LAUNCH_CHECK();
    // Argb reads no secondary bricks: their slot gets the primary table
    return launch_share(dp, dr, db, db, nullptr, nullptr, nullptr, nblocks, RV_BRICK_LINEAR, RV_FIELD_ARGB,
                        (params[0].flags & RV_RENDER_EXACT) != 0u, 1.f, LevelsArg{}, 0.f, 1.f, d_tf, tf_n,
                        d_viewport, vp_w, vp_h, stream);
    RV_CATCH
}
