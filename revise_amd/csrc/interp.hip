// Sparse block-tree nodes -> dense brick interpolation for gfx950.
//
// Semantics: DenseFieldInterpolator<3>::interpolate
// (DenseFieldInterpolator.hpp:198-419): scatter the sparse values into a
// NaN-filled (2^d+1)^3 brick, then for each level l = 0..d-1 fill every still
// undefined edge midpoint (mean of 2), face centre (mean of 4 edge midpoints)
// and cell centre (mean of 6 face centres) of the level-l cubes, propagate the
// 1e30 "no field" marker, and finally turn 1e30 into NaN.  Operand order and
// constants follow the CPU code so results are bit-identical to it.
//
// Decomposition.  A midpoint of a level-l cube depends only on values inside
// that cube, and level-l midpoints never touch coarser lattice nodes.  So the
// brick is cut into tiles of X x T x T cells (X = min(2^d, 64), so x spans
// whole rows of the brick up to depth 6; T = 8 or 16, interp_geom.hpp), one
// workgroup each:
//   1. a tile with a node not present in the sparse data first completes the
//      coarse levels (cubes larger than T) on the stride-T lattice of the
//      whole brick -- at most 17^3 nodes, in LDS -- and keeps the lattice
//      values of its own box;
//   2. it scatters its sparse nodes into LDS: the gather plan (plan.cpp) lists
//      them per tile in increasing sparse index with their LDS offsets, so
//      every global read is coalesced and the permutation happens in LDS;
//   3. it completes the remaining levels in LDS, three barrier-separated
//      phases per level (edges, faces, cells); a tile whose every node is a
//      sparse node skips 1 and 3;
//   4. it stores the nodes it owns (half-open tile box, plus the brick's far
//      faces), converting 1e30 to NaN.
// Nodes on faces shared by two tiles are computed identically by both (their
// operands lie on the shared face), so the tiles are independent.
//
// Memory layout: the dense brick is x fastest; up to depth 6 a tile writes
// (T or T+1) full rows of the brick per z-plane, one contiguous run, stored
// lane-contiguously.

#include "interp_device.hpp"
#include "rv_internal.hpp"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

extern "C" uint64_t rv_interp_plan_bytes(const uint16_t* n2i, uint64_t n, uint32_t depth);
extern "C" int rv_interp_plan_encode(const uint16_t* n2i, uint64_t n, uint32_t depth, void* plan,
                                     uint64_t bytes);

namespace rv {

namespace {

// Workgroup size of the two LDS classes: 512 threads for tiles of up to
// 65 x 17 x 17 nodes, RV_INTERP_SMALL_THREADS for tiles of up to 65 x 9 x 9.
#ifndef RV_INTERP_SMALL_THREADS
#define RV_INTERP_SMALL_THREADS 256
#endif
constexpr uint32_t kBigThreads = 512;
constexpr uint32_t kSmallThreads = RV_INTERP_SMALL_THREADS;

template <uint32_t kFloats, uint32_t NT>
__global__ __launch_bounds__(NT) void interp_single_kernel(Job job, uint32_t ntiles)
{
    __shared__ float A[kFloats];
    __shared__ float lat[64];
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= ntiles)
        return;
    const PlanView P = plan_view(job.plan);
    const uint32_t* h = reinterpret_cast<const uint32_t*>(job.plan);
    interp_tile<NT>(job, t, P.tile_start[t], P.own_end[t], P.tile_start[t + 1], h[6], h[7], A, lat);
}

template <uint32_t kFloats, uint32_t NT>
__global__ __launch_bounds__(NT) void interp_batch_kernel(const TileTask* __restrict__ tasks,
                                                          uint32_t ntiles)
{
    __shared__ float A[kFloats];
    __shared__ float lat[64];
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= ntiles)
        return;
    const TileTask k = tasks[t];
    interp_tile<NT>(k.job, k.local, k.e0, k.e_own, k.e1, k.idx_off, k.lds_off, A, lat);
}

bool big_class(uint32_t depth) { return tile_floats(geom_of(depth)) > kTileFloats8; }

__global__ void fill_u32_kernel(uint32_t* __restrict__ dst, uint64_t n, uint32_t value)
{
    const uint64_t n4 = n >> 2;
    uint4* d4 = reinterpret_cast<uint4*>(dst);
    const uint4 v = make_uint4(value, value, value, value);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n4;
         i += (uint64_t)gridDim.x * blockDim.x)
        d4[i] = v;
    const uint64_t tail = n4 << 2;
    const uint64_t i = tail + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = value;
}

__global__ void ordinals_kernel(const uint16_t* __restrict__ n2i, uint64_t n, uint32_t depth,
                                uint32_t* __restrict__ out)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = ordinal3(n2i[3 * i], n2i[3 * i + 1], n2i[3 * i + 2], depth);
}

uint32_t grid_for(uint64_t n, uint32_t block)
{
    uint64_t g = (n + block - 1) / block;
    if (g < 1)
        g = 1;
    if (g > 65536u * 8u)
        g = 65536u * 8u;
    return (uint32_t)g;
}

uint32_t round8(uint32_t v) { return (v + 7u) & ~7u; }

} // namespace

namespace {

// A plan's tile table as the batch builder reads it: per tile, its entry
// range and the end of its owned entries.
struct PlanTable {
    uint32_t depth = ~0u;
    uint32_t idx_off = 0, lds_off = 0; // header words 6 and 7
    std::vector<uint32_t> start, own;
};

// Reads the tile table from a plan's bytes; get(dst, offset, bytes) copies
// them from wherever the plan lives.
template <class Get>
int read_plan_table(Get get, PlanTable& tb)
{
    uint32_t head[kPlanHeader / 4];
    RV_HIP(get(head, 0, kPlanHeader));
    if (head[0] != kPlanMagic)
        return fail(RV_ERR_INVALID, "interpolation job: plan is not an interpolation plan");
    tb.depth = head[1];
    tb.idx_off = head[6];
    tb.lds_off = head[7];
    if (tb.depth <= kMaxDepth && head[2] == geom_of(tb.depth).ntiles) {
        tb.start.resize(head[2] + 1u);
        RV_HIP(get(tb.start.data(), kPlanHeader, 4ull * tb.start.size()));
        tb.own.resize(head[2]);
        RV_HIP(get(tb.own.data(), head[10], 4ull * tb.own.size()));
        for (uint32_t t = 0; t < head[2]; ++t)
            if (tb.own[t] < tb.start[t] || tb.own[t] > tb.start[t + 1])
                return fail(RV_ERR_INVALID, "interpolation job: corrupt plan");
    }
    return RV_OK;
}

// Tile tables of the plans rv_interp_plan_build uploaded, read from the host
// copy at upload time and dropped by rv_interp_plan_free: building a batch
// over them reads nothing back from the device.  Plans uploaded by the
// caller are read back once per batch.
std::mutex g_plan_mu;
std::map<const void*, PlanTable> g_plan_tables;

} // namespace

int build_tile_tasks(const Job* jobs, uint32_t njobs, std::vector<TileTask>& tasks)
{
    std::map<const uint8_t*, PlanTable> seen;
    tasks.clear();
    for (uint32_t i = 0; i < njobs; ++i) {
        const Job& j = jobs[i];
        const Geom g = geom_of(j.depth);
        auto it = seen.find(j.plan);
        if (it == seen.end()) {
            PlanTable tb;
            bool known = false;
            {
                std::lock_guard<std::mutex> lk(g_plan_mu);
                auto r = g_plan_tables.find(j.plan);
                if (r != g_plan_tables.end()) {
                    tb = r->second;
                    known = true;
                }
            }
            if (!known) {
                const uint8_t* plan = j.plan;
                int rc = read_plan_table(
                    [plan](void* dst, uint64_t off, uint64_t n) {
                        return hipMemcpy(dst, plan + off, n, hipMemcpyDeviceToHost);
                    },
                    tb);
                if (rc)
                    return rc;
            }
            it = seen.emplace(j.plan, std::move(tb)).first;
        }
        const PlanTable& tb = it->second;
        if (tb.depth != j.depth || tb.start.size() != g.ntiles + 1u)
            return fail(RV_ERR_INVALID, "interpolation job: plan was built for another depth");
        for (uint32_t t = 0; t < g.ntiles; ++t)
            tasks.push_back(TileTask{j, t, tb.start[t], tb.start[t + 1], tb.own[t], tb.idx_off,
                                     tb.lds_off});
    }
    return RV_OK;
}

int fill_u32(uint32_t* d, uint64_t n, uint32_t value, hipStream_t s)
{
    if (n == 0)
        return RV_OK;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for((n >> 2) + 1, 256)), dim3(256), 0, s, d, n,
                       value);
    RV_LAUNCH_CHECK();
    return RV_OK;
}

} // namespace rv

using namespace rv;

struct rv_interp_batch {
    TileTask* d_tasks = nullptr;
    uint32_t ntiles = 0;
    uint64_t dense_nodes = 0;
    bool big = false; // some brick needs the 65 x 17 x 17-node LDS class
};

extern "C" uint64_t rv_brick_nodes(uint32_t depth)
{
    if (depth > 15)
        return 0;
    const uint64_t n = (1ull << depth) + 1ull;
    return n * n * n;
}

extern "C" int rv_node_ordinals(const uint16_t* d_n2i, uint64_t n, uint32_t depth, uint32_t* d_out,
                                rv_stream_t stream)
{
    RV_REQUIRE(depth <= 10, "rv_node_ordinals: depth must be <= 10");
    if (n == 0)
        return RV_OK;
    RV_REQUIRE(d_n2i && d_out, "rv_node_ordinals: null pointer");
    hipLaunchKernelGGL(ordinals_kernel, dim3(grid_for(n, 256)), dim3(256), 0, S(stream), d_n2i, n,
                       depth, d_out);
    RV_LAUNCH_CHECK();
    return RV_OK;
}

extern "C" int rv_interp_plan_build(const uint16_t* n2i, uint64_t n, uint32_t depth, void** d_plan,
                                    uint64_t* bytes_out)
{
    RV_TRY
    RV_REQUIRE(d_plan, "rv_interp_plan_build: null output");
    *d_plan = nullptr;
    RV_REQUIRE(depth <= kMaxDepth, "rv_interp_plan_build: depth must be <= 8");
    const uint64_t bytes = rv_interp_plan_bytes(n2i, n, depth);
    if (!bytes)
        return RV_ERR_INVALID; // message set by the encoder
    std::vector<uint8_t> host(bytes);
    int rc = rv_interp_plan_encode(n2i, n, depth, host.data(), bytes);
    if (rc)
        return rc;
    void* d = nullptr;
    RV_HIP(hipMalloc(&d, bytes));
    hipError_t e = hipMemcpy(d, host.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
    }
    PlanTable tb;
    const uint8_t* hp = host.data();
    rc = read_plan_table(
        [hp](void* dst, uint64_t off, uint64_t n) {
            std::memcpy(dst, hp + off, n);
            return hipSuccess;
        },
        tb);
    if (rc) {
        (void)hipFree(d);
        return rc;
    }
    {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        g_plan_tables[d] = std::move(tb);
    }
    *d_plan = d;
    if (bytes_out)
        *bytes_out = bytes;
    return RV_OK;
    RV_CATCH
}

extern "C" int rv_interp_plan_free(void* d_plan)
{
    if (d_plan) {
        {
            std::lock_guard<std::mutex> lk(g_plan_mu);
            g_plan_tables.erase(d_plan);
        }
        RV_HIP(hipFree(d_plan));
    }
    return RV_OK;
}

extern "C" int rv_interp_brick_planned(const float* d_sparse, const void* d_plan, uint32_t depth,
                                       float* d_dense, rv_stream_t stream)
{
    RV_REQUIRE(depth <= kMaxDepth, "rv_interp_brick_planned: depth must be <= 8");
    RV_REQUIRE(d_plan && d_dense, "rv_interp_brick_planned: null pointer");
    Job job{d_sparse, static_cast<const uint8_t*>(d_plan), d_dense, nullptr, depth};
    const Geom g = geom_of(depth);
    if (big_class(depth))
        hipLaunchKernelGGL((interp_single_kernel<kTileFloats, kBigThreads>), dim3(round8(g.ntiles)),
                           dim3(kBigThreads), 0, S(stream), job, g.ntiles);
    else
        hipLaunchKernelGGL((interp_single_kernel<kTileFloats8, kSmallThreads>),
                           dim3(round8(g.ntiles)), dim3(kSmallThreads), 0, S(stream), job, g.ntiles);
    RV_LAUNCH_CHECK();
    return RV_OK;
}

extern "C" int rv_interp_brick(const float* d_sparse, const uint16_t* n2i, uint64_t n,
                               uint32_t depth, float* d_dense, rv_stream_t stream)
{
    RV_TRY
    RV_REQUIRE(depth <= kMaxDepth, "rv_interp_brick: depth must be <= 8");
    const uint64_t bytes = rv_interp_plan_bytes(n2i, n, depth);
    if (!bytes)
        return RV_ERR_INVALID;
    std::vector<uint8_t> host(bytes);
    int rc = rv_interp_plan_encode(n2i, n, depth, host.data(), bytes);
    if (rc)
        return rc;
    void* d = nullptr;
    RV_HIP(hipMallocAsync(&d, bytes, S(stream)));
    hipError_t e = hipMemcpyAsync(d, host.data(), bytes, hipMemcpyHostToDevice, S(stream));
    if (e == hipSuccess)
        rc = rv_interp_brick_planned(d_sparse, d, depth, d_dense, stream);
    const hipError_t e2 = hipFreeAsync(d, S(stream));
    // the host staging buffer must outlive the copy
    const hipError_t e3 = hipStreamSynchronize(S(stream));
    if (e != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__);
    if (rc)
        return rc;
    if (e2 != hipSuccess)
        return hip_fail(e2, "hipFreeAsync", __FILE__, __LINE__);
    if (e3 != hipSuccess)
        return hip_fail(e3, "hipStreamSynchronize", __FILE__, __LINE__);
    return RV_OK;
    RV_CATCH
}

extern "C" uint64_t rv_brick_floats(uint32_t depth, uint32_t layout)
{
    if (depth > 10)
        return 0;
    const uint64_t n = (1ull << depth) + 1ull;
    if (layout == RV_BRICK_TILED4) {
        const uint64_t G = ((n + 1) >> 2) + 1; // x-pairs start at positions 0 .. n + 1
        return G * G * G * 80;
    }
    return layout == RV_BRICK_LINEAR ? n * n * n : 0;
}

extern "C" int rv_interp_batch_create(const rv_brick_job* jobs, uint32_t count,
                                      rv_interp_batch** out)
{
    RV_TRY
    RV_REQUIRE(out, "rv_interp_batch_create: null out");
    *out = nullptr;
    RV_REQUIRE(count == 0 || jobs, "rv_interp_batch_create: null jobs");
    auto b = std::make_unique<rv_interp_batch>();
    std::vector<Job> host(count);
    for (uint32_t i = 0; i < count; ++i) {
        const rv_brick_job& j = jobs[i];
        RV_REQUIRE(j.depth <= kMaxDepth, "rv_interp_batch_create: depth must be <= 8");
        RV_REQUIRE(j.plan, "rv_interp_batch_create: null plan");
        RV_REQUIRE((j.dense != nullptr) != (j.dense16 != nullptr),
                   "rv_interp_batch_create: exactly one of dense/dense16 must be set");
        RV_REQUIRE(j.layout == RV_BRICK_LINEAR || (j.layout == RV_BRICK_TILED4 && j.dense),
                   "rv_interp_batch_create: unknown layout (TILED4 bricks are fp32)");
        RV_REQUIRE(j.layout != RV_BRICK_TILED4 || (reinterpret_cast<uintptr_t>(j.dense) & 15u) == 0u,
                   "rv_interp_batch_create: TILED4 bricks must be 16-byte aligned");
        host[i] = Job{j.sparse, static_cast<const uint8_t*>(j.plan), j.dense,
                      reinterpret_cast<__half*>(j.dense16), j.depth,
                      j.layout == RV_BRICK_TILED4 ? 1u : 0u};
        b->big |= big_class(j.depth);
        b->dense_nodes += rv_brick_nodes(j.depth);
    }
    std::vector<TileTask> tasks;
    int rc = build_tile_tasks(host.data(), count, tasks);
    if (rc)
        return rc;
    RV_REQUIRE(tasks.size() < 0x7FFFFFF0ull, "rv_interp_batch_create: too many tiles");
    b->ntiles = (uint32_t)tasks.size();
    if (!tasks.empty()) {
        RV_HIP(hipMalloc((void**)&b->d_tasks, sizeof(TileTask) * tasks.size()));
        const hipError_t e = hipMemcpy(b->d_tasks, tasks.data(), sizeof(TileTask) * tasks.size(),
                                       hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(b->d_tasks);
            return hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
        }
    }
    *out = b.release();
    return RV_OK;
    RV_CATCH
}

extern "C" int rv_interp_batch_run(rv_interp_batch* b, rv_stream_t stream)
{
    RV_REQUIRE(b, "rv_interp_batch_run: null batch");
    if (b->ntiles == 0)
        return RV_OK;
    if (b->big)
        hipLaunchKernelGGL((interp_batch_kernel<kTileFloats, kBigThreads>), dim3(round8(b->ntiles)),
                           dim3(kBigThreads), 0, S(stream), b->d_tasks, b->ntiles);
    else
        hipLaunchKernelGGL((interp_batch_kernel<kTileFloats8, kSmallThreads>),
                           dim3(round8(b->ntiles)), dim3(kSmallThreads), 0, S(stream), b->d_tasks,
                           b->ntiles);
    RV_LAUNCH_CHECK();
    return RV_OK;
}

extern "C" int rv_interp_batch_info(const rv_interp_batch* b, uint64_t* tiles, uint64_t* nodes)
{
    RV_REQUIRE(b, "rv_interp_batch_info: null batch");
    if (tiles)
        *tiles = b->ntiles;
    if (nodes)
        *nodes = b->dense_nodes;
    return RV_OK;
}

extern "C" void rv_interp_batch_destroy(rv_interp_batch* b)
{
    if (!b)
        return;
    if (b->d_tasks)
        (void)hipFree(b->d_tasks);
    delete b;
}
