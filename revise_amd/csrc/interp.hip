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
// brick is cut into tiles of X x T x T cells (T = min(2^d, 16), X = min(2^d, 64),
// x spans whole rows of the brick):
//   1. every tile workgroup first completes the coarse levels (cubes larger
//      than T) on the stride-T lattice of the whole brick -- at most 17^3
//      nodes, in LDS -- and keeps the lattice values at its own corners;
//   2. it gathers its own (X+1)(T+1)^2 nodes through the gather plan
//      (plan[node] = sparse index or none) into LDS, sets its lattice nodes,
//      and completes the remaining levels in LDS, three barrier-separated
//      phases per level (edges, faces, cells);
//   3. it stores the nodes it owns (half-open tile box, plus the brick's far
//      faces), converting 1e30 to NaN.
// Nodes on faces shared by two tiles are computed identically by both (their
// operands lie on the shared face), so the tiles are independent.
//
// Memory layout: the dense brick is x fastest; a tile writes (T or T+1) full
// rows of the brick per z-plane, i.e. contiguous runs of 16 x 260 B for d = 6.

#include "rv_internal.hpp"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <vector>

namespace rv {

namespace {

constexpr uint32_t kNoIndex = 0xFFFFFFFFu;
constexpr float kNoField = 1e30f;
constexpr uint32_t kThreads = 512;
constexpr uint32_t kMaxDepth = 8;
// Largest tile: 65 x 17 x 17 floats.  The coarse lattice (<= 17^3) reuses it.
constexpr uint32_t kTileFloats = 65u * 17u * 17u;

struct Geom {
    uint32_t d, n, t, T, X, tiles_x, tiles_yz, ntiles, m;
};

__host__ __device__ inline Geom geom_of(uint32_t d)
{
    Geom g;
    g.d = d;
    g.n = (1u << d) + 1u;
    g.t = d < 4u ? d : 4u;
    g.T = 1u << g.t;
    g.X = d < 6u ? (1u << d) : 64u;
    g.tiles_x = (1u << d) / g.X;
    g.tiles_yz = (1u << d) / g.T;
    g.ntiles = g.tiles_x * g.tiles_yz * g.tiles_yz;
    g.m = (1u << d) >> g.t; // lattice cells per axis
    return g;
}

__device__ __forceinline__ uint32_t ordinal3(uint32_t x, uint32_t y, uint32_t z, uint32_t d)
{
    uint32_t r = (z << d) + z + y;
    r = (r << d) + r + x;
    return r;
}

__device__ __forceinline__ bool is_nan(float v) { return v != v; }

struct Job {
    const float* sparse;
    const uint32_t* plan;
    float* dense;
    __half* dense16;
    uint32_t depth;
    uint32_t tile_start;
};

// Complete the levels with cube sizes s_start, s_start/2, ..., 2 (in box units)
// on a box of (X+1) x (T+1) x (T+1) values at A (x fastest).  Three phases per
// level: nodes with one, two, three coordinates an odd multiple of h = s/2.
__device__ void refine_box(float* A, uint32_t X, uint32_t T, uint32_t s_start)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t nwaves = kThreads / 64u;
    const uint32_t rowlen = X + 1u;
    const uint32_t plane = rowlen * (T + 1u);
    for (uint32_t s = s_start; s >= 2u; s >>= 1) {
        const uint32_t h = s >> 1;
        const uint32_t ny = T / h + 1u;   // sub-lattice rows per axis (y and z)
        const uint32_t nx = X / h + 1u;   // sub-lattice nodes per row
        const uint32_t nrows = ny * ny;
        const uint32_t sx = h, sy = h * rowlen, sz = h * plane;
        for (uint32_t phase = 1; phase <= 3; ++phase) {
            for (uint32_t q = wave; q < nrows; q += nwaves) {
                const uint32_t j = q % ny, k = q / ny;
                const uint32_t oy = j & 1u, oz = k & 1u;
                // rows whose y/z parities cannot produce this phase
                if (oy + oz > phase || oy + oz + 1u < phase)
                    continue;
                float* row = A + (k * h) * plane + (j * h) * rowlen;
                for (uint32_t i = lane; i < nx; i += 64u) {
                    const uint32_t ox = i & 1u;
                    if (ox + oy + oz != phase)
                        continue;
                    float* p = row + i * h;
                    if (!is_nan(*p))
                        continue;
                    float v;
                    if (phase == 1u) {
                        const uint32_t st = ox ? sx : (oy ? sy : sz);
                        const float a = p[-(int)st], b = p[st];
                        v = (a == kNoField || b == kNoField) ? kNoField : 0.5f * (a + b);
                    } else if (phase == 2u) {
                        // odd axes d1 < d2
                        const uint32_t st1 = ox ? sx : sy;
                        const uint32_t st2 = oz ? sz : sy;
                        const float s1 = p[-(int)st1], e1 = p[st1];
                        const float s2 = p[-(int)st2], e2 = p[st2];
                        v = (s1 == kNoField || e1 == kNoField || s2 == kNoField || e2 == kNoField)
                                ? kNoField
                                : 0.25f * (((s1 + e1) + s2) + e2);
                    } else {
                        const float s1 = p[-(int)sx], e1 = p[sx];
                        const float s2 = p[-(int)sy], e2 = p[sy];
                        const float s3 = p[-(int)sz], e3 = p[sz];
                        v = (s1 == kNoField || e1 == kNoField || s2 == kNoField ||
                             e2 == kNoField || s3 == kNoField || e3 == kNoField)
                                ? kNoField
                                : (float)(1. / 6) * (((((s1 + e1) + s2) + e2) + s3) + e3);
                    }
                    *p = v;
                }
            }
            __syncthreads();
        }
    }
}

__device__ __forceinline__ float gather(const Job& job, uint32_t ord)
{
    const uint32_t s = __builtin_nontemporal_load(job.plan + ord);
    return s == kNoIndex ? __builtin_nanf("") : job.sparse[s];
}

__device__ void interp_tile(const Job& job, uint32_t local_tile, float* A, float* lat)
{
    const Geom g = geom_of(job.depth);
    const uint32_t d = g.d, X = g.X, T = g.T;
    const uint32_t tx = local_tile % g.tiles_x;
    const uint32_t tyz = local_tile / g.tiles_x;
    const uint32_t ty = tyz % g.tiles_yz, tz = tyz / g.tiles_yz;
    const uint32_t x0 = tx * X, y0 = ty * T, z0 = tz * T;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t nwaves = kThreads / 64u;
    const uint32_t nlat_x = X / T + 1u;

    // 1. coarse levels on the stride-T lattice of the whole brick
    if (g.m >= 2u) {
        const uint32_t mp = g.m + 1u;
        const uint32_t total = mp * mp * mp;
        for (uint32_t q = threadIdx.x; q < total; q += kThreads) {
            const uint32_t i = q % mp, j = (q / mp) % mp, k = q / (mp * mp);
            A[q] = gather(job, ordinal3(i * T, j * T, k * T, d));
        }
        __syncthreads();
        refine_box(A, g.m, g.m, g.m);
        const uint32_t nl = nlat_x * 4u;
        if (threadIdx.x < nl) {
            const uint32_t li = threadIdx.x % nlat_x, c = threadIdx.x / nlat_x;
            const uint32_t I = x0 / T + li, J = y0 / T + (c & 1u), K = z0 / T + (c >> 1);
            lat[threadIdx.x] = A[(K * mp + J) * mp + I];
        }
        __syncthreads();
    }

    // 2. gather the tile
    const uint32_t rowlen = X + 1u;
    const uint32_t tp1 = T + 1u;
    const uint32_t nrows = tp1 * tp1;
    for (uint32_t r = wave; r < nrows; r += nwaves) {
        const uint32_t y = r % tp1, z = r / tp1;
        const uint32_t base = ordinal3(x0, y0 + y, z0 + z, d);
        float* dst = A + r * rowlen;
        for (uint32_t x = lane; x < rowlen; x += 64u)
            dst[x] = gather(job, base + x);
    }
    __syncthreads();
    if (g.m >= 2u) {
        const uint32_t nl = nlat_x * 4u;
        if (threadIdx.x < nl) {
            const uint32_t li = threadIdx.x % nlat_x, c = threadIdx.x / nlat_x;
            const uint32_t y = (c & 1u) * T, z = (c >> 1) * T;
            A[(z * tp1 + y) * rowlen + li * T] = lat[threadIdx.x];
        }
        __syncthreads();
    }

    // 3. finer levels inside the tile
    refine_box(A, X, T, T);

    // 4. store the owned half-open box (+ far faces of the brick)
    const uint32_t ox = (tx + 1u == g.tiles_x) ? rowlen : X;
    const uint32_t oy = (ty + 1u == g.tiles_yz) ? tp1 : T;
    const uint32_t oz = (tz + 1u == g.tiles_yz) ? tp1 : T;
    const uint32_t orows = oy * oz;
    for (uint32_t r = wave; r < orows; r += nwaves) {
        const uint32_t y = r % oy, z = r / oy;
        const uint32_t base = ordinal3(x0, y0 + y, z0 + z, d);
        const float* src = A + (z * tp1 + y) * rowlen;
        for (uint32_t x = lane; x < ox; x += 64u) {
            float v = src[x];
            if (v == kNoField)
                v = __builtin_nanf("");
            if (job.dense16)
                job.dense16[base + x] = __float2half(v);
            else
                __builtin_nontemporal_store(v, job.dense + base + x);
        }
    }
}

// Workgroup -> tile mapping that keeps consecutive tiles (the tiles of one
// brick) on one XCD: workgroups are dispatched round-robin over the 8 XCDs.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t bid, uint32_t nblocks)
{
    const uint32_t per = nblocks >> 3; // nblocks is a multiple of 8
    return (bid & 7u) * per + (bid >> 3);
}

__global__ __launch_bounds__(kThreads) void interp_single_kernel(Job job, uint32_t ntiles)
{
    __shared__ float A[kTileFloats];
    __shared__ float lat[32];
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= ntiles)
        return;
    interp_tile(job, t, A, lat);
}

__global__ __launch_bounds__(kThreads) void interp_batch_kernel(const Job* __restrict__ jobs,
                                                                uint32_t njobs, uint32_t ntiles)
{
    __shared__ float A[kTileFloats];
    __shared__ float lat[32];
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= ntiles)
        return;
    // last job with tile_start <= t
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].tile_start <= t)
            lo = mid;
        else
            hi = mid;
    }
    const Job job = jobs[lo];
    interp_tile(job, t - job.tile_start, A, lat);
}

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

__global__ void plan_scatter_kernel(const uint16_t* __restrict__ n2i, uint64_t n, uint32_t depth,
                                    uint32_t* __restrict__ plan, uint32_t* bad)
{
    const uint32_t npe = (1u << depth) + 1u;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = n2i[3 * i], y = n2i[3 * i + 1], z = n2i[3 * i + 2];
        if (x >= npe || y >= npe || z >= npe) {
            if (bad)
                atomicAdd(bad, 1u);
            continue;
        }
        plan[ordinal3(x, y, z, depth)] = (uint32_t)i;
    }
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
    Job* d_jobs = nullptr;
    uint32_t njobs = 0;
    uint32_t ntiles = 0;
    uint64_t dense_nodes = 0;
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

extern "C" int rv_interp_plan_build(const uint16_t* d_n2i, uint64_t n, uint32_t depth,
                                    uint32_t* d_plan, uint32_t* d_bad_count, rv_stream_t stream)
{
    RV_REQUIRE(depth <= kMaxDepth, "rv_interp_plan_build: depth must be <= 8");
    RV_REQUIRE(d_plan, "rv_interp_plan_build: null plan");
    RV_REQUIRE(n == 0 || d_n2i, "rv_interp_plan_build: null n2i");
    RV_REQUIRE(n <= 0xFFFFFFFEull, "rv_interp_plan_build: too many nodes");
    const uint64_t V = rv_brick_nodes(depth);
    int rc = fill_u32(d_plan, V, kNoIndex, S(stream));
    if (rc)
        return rc;
    if (n) {
        hipLaunchKernelGGL(plan_scatter_kernel, dim3(grid_for(n, 256)), dim3(256), 0, S(stream),
                           d_n2i, n, depth, d_plan, d_bad_count);
        RV_LAUNCH_CHECK();
    }
    return RV_OK;
}

extern "C" int rv_interp_brick_planned(const float* d_sparse, const uint32_t* d_plan,
                                       uint32_t depth, float* d_dense, rv_stream_t stream)
{
    RV_REQUIRE(depth <= kMaxDepth, "rv_interp_brick_planned: depth must be <= 8");
    RV_REQUIRE(d_plan && d_dense, "rv_interp_brick_planned: null pointer");
    Job job{d_sparse, d_plan, d_dense, nullptr, depth, 0};
    const Geom g = geom_of(depth);
    hipLaunchKernelGGL(interp_single_kernel, dim3(round8(g.ntiles)), dim3(kThreads), 0, S(stream),
                       job, g.ntiles);
    RV_LAUNCH_CHECK();
    return RV_OK;
}

extern "C" int rv_interp_brick(const float* d_sparse, const uint16_t* d_n2i, uint64_t n,
                               uint32_t depth, float* d_dense, uint32_t* d_workspace,
                               rv_stream_t stream)
{
    RV_REQUIRE(depth <= kMaxDepth, "rv_interp_brick: depth must be <= 8");
    uint32_t* ws = d_workspace;
    if (!ws)
        RV_HIP(hipMallocAsync((void**)&ws, rv_brick_nodes(depth) * 4, S(stream)));
    int rc = rv_interp_plan_build(d_n2i, n, depth, ws, nullptr, stream);
    if (!rc)
        rc = rv_interp_brick_planned(d_sparse, ws, depth, d_dense, stream);
    if (!d_workspace) {
        hipError_t e = hipFreeAsync(ws, S(stream));
        if (!rc && e != hipSuccess)
            return hip_fail(e, "hipFreeAsync", __FILE__, __LINE__);
    }
    return rc;
}

extern "C" int rv_interp_batch_create(const rv_brick_job* jobs, uint32_t count,
                                      rv_interp_batch** out)
{
    RV_TRY
    RV_REQUIRE(out, "rv_interp_batch_create: null out");
    *out = nullptr;
    RV_REQUIRE(count == 0 || jobs, "rv_interp_batch_create: null jobs");
    std::vector<Job> host(count);
    uint64_t tiles = 0, nodes = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const rv_brick_job& j = jobs[i];
        RV_REQUIRE(j.depth <= kMaxDepth, "rv_interp_batch_create: depth must be <= 8");
        RV_REQUIRE(j.plan, "rv_interp_batch_create: null plan");
        RV_REQUIRE((j.dense != nullptr) != (j.dense16 != nullptr),
                   "rv_interp_batch_create: exactly one of dense/dense16 must be set");
        host[i] = Job{j.sparse, j.plan, j.dense, reinterpret_cast<__half*>(j.dense16), j.depth,
                      (uint32_t)tiles};
        tiles += geom_of(j.depth).ntiles;
        nodes += rv_brick_nodes(j.depth);
    }
    RV_REQUIRE(tiles < 0x7FFFFFF0ull, "rv_interp_batch_create: too many tiles");
    auto* b = new rv_interp_batch;
    b->njobs = count;
    b->ntiles = (uint32_t)tiles;
    b->dense_nodes = nodes;
    if (count) {
        hipError_t e = hipMalloc((void**)&b->d_jobs, sizeof(Job) * count);
        if (e != hipSuccess) {
            delete b;
            return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
        }
        e = hipMemcpy(b->d_jobs, host.data(), sizeof(Job) * count, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(b->d_jobs);
            delete b;
            return hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
        }
    }
    *out = b;
    return RV_OK;
    RV_CATCH
}

extern "C" int rv_interp_batch_run(rv_interp_batch* b, rv_stream_t stream)
{
    RV_REQUIRE(b, "rv_interp_batch_run: null batch");
    if (b->ntiles == 0)
        return RV_OK;
    hipLaunchKernelGGL(interp_batch_kernel, dim3(round8(b->ntiles)), dim3(kThreads), 0, S(stream),
                       b->d_jobs, b->njobs, b->ntiles);
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
    if (b->d_jobs)
        (void)hipFree(b->d_jobs);
    delete b;
}
