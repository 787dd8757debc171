// Device code of the brick interpolation tile, shared by the interpolation
// kernels and the persistent frame kernel.  See interp.hip for the method and
// plan.cpp for the gather plan layout.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "interp_geom.hpp"

namespace rv {

constexpr uint32_t kNoIndex = 0xFFFFFFFFu;
constexpr float kNoField = 1e30f;
constexpr uint32_t kMaxDepth = 8;
// LDS floats of the largest tile, 65 x 17 x 17 nodes (depth 8 always uses it).
constexpr uint32_t kTileFloats = 65u * 17u * 17u;
// LDS floats of a 64 x 8 x 8-cell tile; its coarse lattice (at most 17^3
// nodes, depth <= 7) fits as well.
constexpr uint32_t kTileFloats8 = 65u * 9u * 9u;

__device__ __forceinline__ uint32_t ordinal3(uint32_t x, uint32_t y, uint32_t z, uint32_t d)
{
    uint32_t r = (z << d) + z + y;
    r = (r << d) + r + x;
    return r;
}

__device__ __forceinline__ bool is_nan(float v) { return v != v; }

struct Job {
    const float* sparse;
    const uint8_t* plan;  // encoded gather plan (plan.cpp)
    float* dense;
    __half* dense16;
    uint32_t depth;
    uint32_t tiled;       // dense (fp32) in 4 x 4 x 4 micro-bricks (tiled4_index)
};


// One tile of a batched launch, with its plan entry range looked up on the
// host when the batch is built: a workgroup reads this one record instead of
// searching the job list and reading the plan's tile table.
struct TileTask {
    Job job;
    uint32_t local;   // tile index within its brick
    uint32_t e0, e1;  // plan entries [e0, e1) of the tile
    uint32_t e_own;   // [e0, e_own): the nodes the tile stores
    uint32_t idx_off; // byte offsets of the plan's sparse index and LDS
    uint32_t lds_off; // offset arrays (header words 6 and 7)
};

// Builds the tile list of a batch, job after job (host).  Reads each distinct
// plan's header and tile table back from the device once and checks the plan
// against the job's depth (RV_ERR_INVALID on a mismatch).
int build_tile_tasks(const Job* jobs, uint32_t njobs, std::vector<TileTask>& tasks);

// The odd-axis classes of one refinement phase of a level: phase 1 = edge
// midpoints (one odd coordinate), 2 = face centres (two), 3 = cell centres
// (three).  Class k of a phase has the odd-axis mask below (bit 0 = x, 1 = y,
// 2 = z); its nodes form a cx x cy x cz grid, with qx, qy cubes per axis
// (qy also along z).  Items of a phase are numbered class by class.
__device__ __forceinline__ uint32_t phase_mask(uint32_t phase, uint32_t k)
{
    return phase == 1u ? (1u << k) : (phase == 2u ? (7u ^ (4u >> k)) : 7u);
}

__device__ __forceinline__ uint32_t class_size(uint32_t m, uint32_t qx, uint32_t qy)
{
    return ((m & 1u) ? qx : qx + 1u) * ((m & 2u) ? qy : qy + 1u) * ((m & 4u) ? qy : qy + 1u);
}

// Item q of a phase (classes end at end0 < end1 <= total): its node and the
// value the reference assigns it if it is still undefined (operand order of
// DenseFieldInterpolator.hpp:257-370).  Returns false when the node is
// already defined.  All values stay in registers (no per-class tables, which
// the compiler would place in scratch memory).
__device__ __forceinline__ bool refine_item(float* A, uint32_t phase, uint32_t q, uint32_t qx,
                                            uint32_t qy, uint32_t end0, uint32_t end1, uint32_t h,
                                            uint32_t rowlen, uint32_t plane, float*& out,
                                            float& value)
{
    const uint32_t k = q < end0 ? 0u : (q < end1 ? 1u : 2u);
    const uint32_t m = phase_mask(phase, k);
    const uint32_t base = k == 0u ? 0u : (k == 1u ? end0 : end1);
    const uint32_t cx = (m & 1u) ? qx : qx + 1u;
    const uint32_t cxy = cx * ((m & 2u) ? qy : qy + 1u);
    // r / cxy and r2 / cx through fp32 reciprocals: the quotients are below
    // 64 and (r + 0.5) / d stays 0.5 / d >= 1e-4 away from an integer, far
    // above the rounding error (a few 1e-6), so the truncation is exact.
    const uint32_t r = q - base;
    const uint32_t iz = (uint32_t)(((float)r + 0.5f) * __builtin_amdgcn_rcpf((float)cxy));
    const uint32_t r2 = r - iz * cxy;
    const uint32_t iy = (uint32_t)(((float)r2 + 0.5f) * __builtin_amdgcn_rcpf((float)cx));
    const uint32_t ix = r2 - iy * cx;
    const uint32_t x = ((ix << 1) + (m & 1u)) * h;
    const uint32_t y = ((iy << 1) + ((m >> 1) & 1u)) * h;
    const uint32_t z = ((iz << 1) + (m >> 2)) * h;
    float* p = A + z * plane + y * rowlen + x;
    const int sx = (int)h, sy = (int)(h * rowlen), sz = (int)(h * plane);
    const float cur = *p;
    float v;
    if (phase == 1u) {
        const int st = (m & 1u) ? sx : ((m & 2u) ? sy : sz);
        const float a = p[-st], b = p[st];
        v = (a == kNoField || b == kNoField) ? kNoField : 0.5f * (a + b);
    } else if (phase == 2u) {
        // odd axes d1 < d2
        const int st1 = (m & 1u) ? sx : sy;
        const int st2 = (m & 4u) ? sz : sy;
        const float s1 = p[-st1], e1 = p[st1];
        const float s2 = p[-st2], e2 = p[st2];
        v = (s1 == kNoField || e1 == kNoField || s2 == kNoField || e2 == kNoField)
                ? kNoField
                : 0.25f * (((s1 + e1) + s2) + e2);
    } else {
        const float s1 = p[-sx], e1 = p[sx];
        const float s2 = p[-sy], e2 = p[sy];
        const float s3 = p[-sz], e3 = p[sz];
        v = (s1 == kNoField || e1 == kNoField || s2 == kNoField || e2 == kNoField ||
             s3 == kNoField || e3 == kNoField)
                ? kNoField
                : (float)(1. / 6) * (((((s1 + e1) + s2) + e2) + s3) + e3);
    }
    out = p;
    value = v;
    return is_nan(cur);
}

// Complete the levels with cube sizes s_start, s_start/2, ..., 2 (in box units)
// on a box of (X+1) x (T+1) x (T+1) values at A (x fastest).  Three phases per
// level, separated by barriers; within a phase no node reads another node of
// the same phase, so each thread handles two items at a time with all their
// LDS reads issued before either store.
template <uint32_t NT>
__device__ inline void refine_box(float* A, uint32_t X, uint32_t T, uint32_t s_start)
{
    const uint32_t rowlen = X + 1u;
    const uint32_t plane = rowlen * (T + 1u);
    for (uint32_t s = s_start; s >= 2u; s >>= 1) {
        const uint32_t h = s >> 1;
        for (uint32_t phase = 1; phase <= 3; ++phase) {
            const uint32_t qx = X / s, qy = T / s;
            const uint32_t end0 = class_size(phase_mask(phase, 0), qx, qy);
            const uint32_t end1 = phase == 3u ? end0 : end0 + class_size(phase_mask(phase, 1), qx, qy);
            const uint32_t total = phase == 3u ? end0 : end1 + class_size(phase_mask(phase, 2), qx, qy);
            for (uint32_t q = threadIdx.x; q < total; q += 2u * NT) {
                float *p0, *p1 = nullptr;
                float v0, v1 = 0.f;
                const bool w0 = refine_item(A, phase, q, qx, qy, end0, end1, h, rowlen, plane, p0, v0);
                bool w1 = false;
                if (q + NT < total)
                    w1 = refine_item(A, phase, q + NT, qx, qy, end0, end1, h, rowlen, plane, p1, v1);
                if (w0)
                    *p0 = v0;
                if (w1)
                    *p1 = v1;
            }
            __syncthreads();
        }
    }
}

struct PlanView {
    uint32_t depth, ntiles, nlat, entries;
    const uint32_t* tile_start;
    const uint32_t* own_end;
    const uint32_t* lat;
    const uint32_t* idx;
    const uint16_t* lds;
};

__device__ __forceinline__ PlanView plan_view(const uint8_t* P)
{
    const uint32_t* h = reinterpret_cast<const uint32_t*>(P);
    PlanView v;
    v.depth = h[1];
    v.ntiles = h[2];
    v.nlat = h[3];
    v.entries = h[4];
    v.tile_start = h + kPlanHeader / 4u;
    v.own_end = reinterpret_cast<const uint32_t*>(P + h[10]);
    v.lat = reinterpret_cast<const uint32_t*>(P + h[5]);
    v.idx = reinterpret_cast<const uint32_t*>(P + h[6]);
    v.lds = reinterpret_cast<const uint16_t*>(P + h[7]);
    return v;
}

// One tile of a brick with NT threads (a multiple of 64); A holds
// tile_floats(geom_of(depth)) floats, lat 64 floats; [e0, e1) is the tile's
// plan entry range, [e0, e_own) the nodes it stores.
// idx_off, lds_off: the byte offsets of the plan's entry arrays (header words
// 6 and 7), passed in so that a full tile's first plan loads wait for nothing
// but the caller's own.
template <uint32_t NT>
__device__ inline void interp_tile(const Job& job, uint32_t local_tile, uint32_t e0, uint32_t e_own,
                                   uint32_t e1, uint32_t idx_off, uint32_t lds_off, float* A,
                                   float* lat)
{
    static_assert(NT % 64u == 0u, "whole waves per workgroup");
    const PlanView P = plan_view(job.plan);
    const uint32_t* __restrict__ plan_idx = reinterpret_cast<const uint32_t*>(job.plan + idx_off);
    const uint16_t* __restrict__ plan_lds = reinterpret_cast<const uint16_t*>(job.plan + lds_off);
    const Geom g = geom_of(job.depth);
    const uint32_t d = g.d, X = g.X, T = g.T;
    const uint32_t tx = local_tile % g.tiles_x;
    const uint32_t tyz = local_tile / g.tiles_x;
    const uint32_t ty = tyz % g.tiles_yz, tz = tyz / g.tiles_yz;
    const uint32_t x0 = tx * X, y0 = ty * T, z0 = tz * T;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t nwaves = NT / 64u;
    const uint32_t nlat_x = X / T + 1u;
    const uint32_t rowlen = X + 1u;
    const uint32_t tp1 = T + 1u;
    const uint32_t tile_nodes = rowlen * tp1 * tp1;

    // A "full" tile skips the interpolation and scatters only the nodes it
    // stores; if one of them is NaN (undefined, so the reference interpolates
    // it) the tile is redone as a partial one.
    bool full = e1 - e0 == tile_nodes;
    for (int pass = 0; pass < 2; ++pass) {
        if (!full) {
            // 1. coarse levels on the stride-T lattice of the whole brick
            if (g.m >= 2u) {
                const uint32_t mp = g.m + 1u;
                for (uint32_t q = threadIdx.x; q < P.nlat; q += NT) {
                    const uint32_t s = P.lat[q];
                    A[q] = s == kNoIndex ? __builtin_nanf("") : job.sparse[s];
                }
                __syncthreads();
                refine_box<NT>(A, g.m, g.m, g.m);
                const uint32_t nl = nlat_x * 4u;
                if (threadIdx.x < nl) {
                    const uint32_t li = threadIdx.x % nlat_x, c = threadIdx.x / nlat_x;
                    const uint32_t I = x0 / T + li, J = y0 / T + (c & 1u), K = z0 / T + (c >> 1);
                    lat[threadIdx.x] = A[(K * mp + J) * mp + I];
                }
                __syncthreads();
            }
            for (uint32_t q = threadIdx.x; q < tile_nodes; q += NT)
                A[q] = __builtin_nanf("");
            __syncthreads();
        }

        // 2. scatter the tile's sparse nodes into LDS: plan entries and sparse
        // values are read in increasing sparse index (coalesced), kUnroll per
        // thread in flight.  Both go through the caches: the plan is
        // per-subtree metadata read every frame, and a cache line of sparse
        // values usually holds nodes of two neighbouring tiles (of one brick,
        // so on one XCD), which then read it from L2 once.
        // 512-thread workgroups (2 per CU, bound by LDS) issue 24 loads per
        // thread at a time (about 100 VGPRs, still 4 waves per SIMD): a
        // 65 x 16 x 16 full tile takes two rounds of plan then value loads
        // instead of five.  256-thread workgroups (up to 7 per CU) issue 12
        // (56 VGPRs, still 7 waves per SIMD): a 17^3 brick in two rounds
        // (12.4 -> 11.9 ms for 2.0 M of them; 20 per thread, one round at 5
        // waves per SIMD, measured 12.1 ms).
        constexpr uint32_t kUnroll = NT >= 512u ? 24u : 12u;
        bool saw_nan = false;
        const uint32_t eend = full ? e_own : e1;
        for (uint32_t k0 = e0 + threadIdx.x; k0 < eend; k0 += NT * kUnroll) {
            uint32_t si[kUnroll];
            uint32_t off[kUnroll];
#pragma unroll
            for (uint32_t j = 0; j < kUnroll; ++j) {
                const uint32_t k = k0 + j * NT;
                si[j] = kNoIndex;
                off[j] = 0;
                if (k < eend) {
                    si[j] = plan_idx[k];
                    off[j] = plan_lds[k];
                }
            }
            float v[kUnroll];
#pragma unroll
            for (uint32_t j = 0; j < kUnroll; ++j)
                v[j] = si[j] == kNoIndex ? 0.f : job.sparse[si[j]];
#pragma unroll
            for (uint32_t j = 0; j < kUnroll; ++j)
                if (si[j] != kNoIndex) {
                    A[off[j]] = v[j];
                    saw_nan |= is_nan(v[j]);
                }
        }
        if (full && __syncthreads_or(saw_nan)) {
            full = false;
            continue;
        }
        __syncthreads();
        break;
    }

    if (!full) {
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
        refine_box<NT>(A, X, T, T);
    }

    // 4. store the owned half-open box (+ far faces of the brick)
    if (job.tiled) {
        // Micro-bricks of 4 x 4 x 4 cells, 80 consecutive floats each: 16
        // rows (y, then z) of 5 x-positions, 4m .. 4m + 4 in micro-brick m
        // (neighbours along x share the position 4m + 4, so a cell's x-pair
        // never leaves its row), micro-bricks x fastest over G^3 of them.
        // Along y and z node c is at position c; along x at c + 1, and
        // x-positions 0 and n .. 4G hold the edge nodes 0 and n - 1, so the
        // x-pair (c, c + 1) of every clamped sample lies in the brick.  A tile
        // writes every copy of the positions of the nodes it owns (x: x0 + 1
        // .. x0 + ex; the first x-tile also position 0, the last the edge
        // positions; y/z: its owned box, the last tile up to 4G with padding
        // rows, written NaN and never read).  Along y and z tile boundaries
        // are multiples of 4, so each row belongs to one tile; along x
        // (bricks of several x-tiles) the micro-bricks at a seam are shared
        // with the previous tile, each writing its own floats of them.
        const uint32_t G = ((g.n + 1u) >> 2) + 1u;
        const bool lastx = tx + 1u == g.tiles_x, lasty = ty + 1u == g.tiles_yz, lastz = tz + 1u == g.tiles_yz;
        const uint32_t ex = lastx ? rowlen : X;
        const uint32_t ey = lasty ? tp1 : T;
        const uint32_t ez = lastz ? tp1 : T;
        const uint32_t p_lo = tx == 0u ? 0u : x0 + 1u, p_hi = lastx ? 4u * G + 1u : x0 + ex + 1u;
        const uint32_t mx0 = p_lo == 0u ? 0u : (p_lo - 1u) >> 2, my0 = y0 >> 2, mz0 = z0 >> 2;
        const uint32_t nmx = min((p_hi - 1u) >> 2, G - 1u) + 1u - mx0;
        const uint32_t nmy = (lasty ? G : ((y0 + ey + 3u) >> 2)) - my0;
        const uint32_t nmz = (lastz ? G : ((z0 + ez + 3u) >> 2)) - mz0;
        // one micro-brick row (5 positions along x, 20 bytes) per lane: a
        // wave stores 4 whole micro-bricks (1280 contiguous bytes).  The
        // wave's first micro-brick index is wave-uniform (NT is a multiple
        // of 64), so its decomposition runs on the scalar unit; the lanes'
        // offsets 0..3 are carried along x.
        const uint32_t total = nmx * nmy * nmz * 16u;
        const int nlast = (int)g.n - 1;
        for (uint32_t q = threadIdx.x; q < total; q += NT) {
            const uint32_t mbw = __builtin_amdgcn_readfirstlane(q >> 6) << 2, r = q & 15u;
            uint32_t mx = mbw % nmx, my = (mbw / nmx) % nmy, mz = mbw / (nmx * nmy);
            mx += (q >> 4) & 3u;
            while (mx >= nmx) { // at most 3 steps
                mx -= nmx;
                if (++my == nmy) {
                    my = 0u;
                    ++mz;
                }
            }
            mx += mx0;
            my += my0;
            mz += mz0;
            const uint32_t y = (my << 2) | (r & 3u), z = (mz << 2) | (r >> 2), pb = mx << 2;
            const bool row_ok = y < y0 + ey && z < z0 + ez;
            float o[5];
            bool own[5];
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                const uint32_t pos = pb + (uint32_t)c;
                own[c] = pos >= p_lo && pos < p_hi;
                const int node = min(max((int)pos - 1, 0), nlast); // in [x0, x0 + ex) when owned
                float v = __builtin_nanf("");
                if (row_ok && own[c]) {
                    v = A[((z - z0) * tp1 + (y - y0)) * rowlen + ((uint32_t)node - x0)];
                    if (v == kNoField)
                        v = __builtin_nanf("");
                }
                o[c] = v;
            }
            float* dst = job.dense + ((mz * G + my) * G + mx) * 80u + r * 5u;
            if (own[0] && own[1] && own[2] && own[3] && own[4]) {
                typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));
                const f32x4u head = {o[0], o[1], o[2], o[3]};
                __builtin_nontemporal_store(head, reinterpret_cast<f32x4u*>(dst));
                __builtin_nontemporal_store(o[4], dst + 4);
            } else {
#pragma unroll
                for (int c = 0; c < 5; ++c)
                    if (own[c])
                        __builtin_nontemporal_store(o[c], dst + c);
            }
        }
        return;
    }
    if (g.tiles_x == 1u && g.tiles_yz == 1u) {
        // the tile is the whole brick: LDS and brick layouts coincide, one
        // contiguous run (fp16: two values per 4-byte store when aligned)
        if (job.dense16 && (reinterpret_cast<uintptr_t>(job.dense16) & 3u) == 0u) {
            // stored from the start of the aligned 256-B chunk the brick
            // starts in (lanes before it masked), as the row runs below
            const uint32_t pairs = tile_nodes >> 1;
            __half2* out2 = reinterpret_cast<__half2*>(job.dense16);
            const float2* A2 = reinterpret_cast<const float2*>(A);
            const uint32_t skew = (uint32_t)(reinterpret_cast<uintptr_t>(job.dense16) >> 2) & 63u;
            for (uint32_t qs = threadIdx.x; qs < pairs + skew; qs += NT) {
                if (qs < skew)
                    continue;
                const uint32_t q = qs - skew;
                float2 v = A2[q];
                if (v.x == kNoField)
                    v.x = __builtin_nanf("");
                if (v.y == kNoField)
                    v.y = __builtin_nanf("");
                out2[q] = __halves2half2(__float2half(v.x), __float2half(v.y));
            }
            if ((tile_nodes & 1u) && threadIdx.x == 0) {
                float v = A[tile_nodes - 1u];
                if (v == kNoField)
                    v = __builtin_nanf("");
                job.dense16[tile_nodes - 1u] = __float2half(v);
            }
            return;
        }
        for (uint32_t q = threadIdx.x; q < tile_nodes; q += NT) {
            float v = A[q];
            if (v == kNoField)
                v = __builtin_nanf("");
            if (job.dense16)
                job.dense16[q] = __float2half(v);
            else
                __builtin_nontemporal_store(v, job.dense + q);
        }
        return;
    }
    const uint32_t ox = (tx + 1u == g.tiles_x) ? rowlen : X;
    const uint32_t oy = (ty + 1u == g.tiles_yz) ? tp1 : T;
    const uint32_t oz = (tz + 1u == g.tiles_yz) ? tp1 : T;
    if (g.tiles_x == 1u) {
        // the tile spans whole brick rows: its rows of one z-plane are one
        // contiguous run in the brick and in LDS.  The run is stored from the
        // start of the 64-lane chunk (256 B fp32, 128 B fp16) its first node
        // falls in, lanes before the run masked, so every store instruction
        // of a wave covers one aligned chunk and no cache line is written by
        // two instructions.
        const uint32_t run = oy * rowlen;
        const uint32_t esz = job.dense16 ? 2u : 4u;
        const uintptr_t brick = job.dense16 ? reinterpret_cast<uintptr_t>(job.dense16)
                                            : reinterpret_cast<uintptr_t>(job.dense);
        for (uint32_t z = 0; z < oz; ++z) {
            const uint32_t base = ordinal3(0, y0, z0 + z, d);
            const float* src = A + z * tp1 * rowlen;
            const uint32_t skew = (uint32_t)((brick + (uintptr_t)base * esz) / esz) & 63u;
            for (uint32_t q = threadIdx.x; q < run + skew; q += NT) {
                if (q < skew)
                    continue;
                const uint32_t i = q - skew;
                float v = src[i];
                if (v == kNoField)
                    v = __builtin_nanf("");
                if (job.dense16)
                    job.dense16[base + i] = __float2half(v);
                else
                    __builtin_nontemporal_store(v, job.dense + base + i);
            }
        }
        return;
    }
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

} // namespace rv
