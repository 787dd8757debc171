// Device helpers shared by the ray-cast kernels: the per-sample operations of
// the reference Argb renderer (renderArgb.hpp:30-55, RenderFuncInput.hpp:42-49,
// cuda_blend.hpp:24-44, cuda_argb_transform.hpp:26-43) in a fixed fp32
// operation order, so every kernel of this library produces the same bits.
#pragma once

#include <type_traits>

#include "../../include/revise_hip.h"
#include "interp_geom.hpp"

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

namespace rv {

// pow replacement for 0 <= b <= 1, e > 0 (the reference calls CUDA powf,
// renderArgb.hpp:45).  Method: b = m 2^k with m in [sqrt(1/2), sqrt(2));
// log2 b = k + (2/ln2) atanh(s), s = (m-1)/(m+1), by the odd Taylor series of
// atanh to s^9; 2^y = 2^floor(y) P(y - floor(y)) with P a degree-6 polynomial,
// P(0) = 1, the rest a relative-error minimax fit to 2^f on [0,1] (Lawson's
// reweighted least squares; tests/golden/fit_exp2.py makes the coefficients),
// both polynomials by Horner's rule with fused multiply-adds; relative error
// below 1e-6 over the renderer's range.  Results under 2^-126 are 0.
__device__ __forceinline__ float pow01(float b, float e)
{
    if (!(b >= 0.f))
        return __builtin_nanf("");
    if (b == 0.f)
        return 0.f;
    if (b == 1.f)
        return 1.f;
    uint32_t bits = __float_as_uint(b);
    int k = (int)((bits >> 23) & 0xffu) - 127;
    uint32_t mbits = (bits & 0x007fffffu) | 0x3f800000u;
    if (((bits >> 23) & 0xffu) == 0u) {
        const float bs = b * 4294967296.0f;
        bits = __float_as_uint(bs);
        k = (int)((bits >> 23) & 0xffu) - 127 - 32;
        mbits = (bits & 0x007fffffu) | 0x3f800000u;
    }
    float m = __uint_as_float(mbits);
    if (m >= 1.41421356f) {
        m = m * 0.5f;
        k = k + 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float poly = 0.111111111f;
    poly = fmaf(poly, s2, 0.142857143f);
    poly = fmaf(poly, s2, 0.2f);
    poly = fmaf(poly, s2, 0.333333333f);
    poly = fmaf(poly, s2, 1.0f);
    const float l = (s * poly) * 2.88539008f;
    const float y = e * ((float)k + l);
    if (y < -126.0f)
        return 0.f;
    if (y >= 0.0f)
        return 1.0f;
    const float fl = floorf(y);
    const float f = y - fl;
    float p = 0x1.c54176p-13f;
    p = fmaf(p, f, 0x1.46d64cp-10f);
    p = fmaf(p, f, 0x1.3d0b92p-7f);
    p = fmaf(p, f, 0x1.c68912p-5f);
    p = fmaf(p, f, 0x1.ebfd58p-3f);
    p = fmaf(p, f, 0x1.62e42cp-1f);
    p = fmaf(p, f, 1.0f);
    return ldexpf(p, (int)fl);
}

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

__device__ __forceinline__ float lerpf(float a, float b, float t) { return (1.f - t) * a + t * b; }

// Texel index pair and weight along one axis, unnormalized coordinates with
// texel centres at i + 0.5 and clamp addressing.
__device__ __forceinline__ void axis_taps(float t, int n, int& i0, int& i1, float& f)
{
    float v = t - 0.5f;
    if (!(v > -1.f))
        v = -1.f;
    if (v > (float)n)
        v = (float)n;
    const float fl = floorf(v);
    f = v - fl;
    const int i = (int)fl;
    i0 = clampi(i, 0, n - 1);
    i1 = clampi(i + 1, 0, n - 1);
}

// The two x-neighbours brick[i], brick[i+1] in one 8-byte (fp32) or 4-byte
// (fp16) load: the brick is x fastest, so the corners (x0, x0+1) of a cell
// are adjacent, and a wave's corner loads touch half as many cache lines as
// with one load per corner.
// Bricks live in device memory: loads through global-address-space pointers
// are global loads, not flat ones (no aperture check, no LDS counter).  The
// offsets into a brick are 32-bit byte counts, so a wave-uniform brick
// address stays in scalar registers (global_load with saddr); the ray-casts
// accept bricks of depth <= 9 (a depth-9 fp32 brick is 540 MB).
template <class T>
using gptr = const __attribute__((address_space(1))) T*;

template <class T>
__device__ __forceinline__ gptr<T> global_of(const T* p)
{
    return (gptr<T>)p;
}

// A pair of fp32 (fp16) brick values starts at any 4-byte (2-byte) boundary.
typedef uint64_t pair32_t __attribute__((aligned(4)));
typedef uint32_t pair16_t __attribute__((aligned(2)));

__device__ __forceinline__ void load_pair(const float* __restrict__ b, uint32_t i, float& lo, float& hi)
{
    const uint64_t v = *reinterpret_cast<gptr<pair32_t>>(global_of(reinterpret_cast<const char*>(b)) + (i << 2));
    lo = __uint_as_float((uint32_t)v);
    hi = __uint_as_float((uint32_t)(v >> 32));
}

__device__ __forceinline__ void load_pair(const __half* __restrict__ b, uint32_t i, float& lo, float& hi)
{
    const uint32_t v = *reinterpret_cast<gptr<pair16_t>>(global_of(reinterpret_cast<const char*>(b)) + (i << 1));
    lo = __half2float(__ushort_as_half((unsigned short)(v & 0xFFFFu)));
    hi = __half2float(__ushort_as_half((unsigned short)(v >> 16)));
}

// The 8 corner values of the trilinear fetch at texel coordinates (tx, ty,
// tz): taps and weights of axis_taps, corner c = bit0 x, bit1 y, bit2 z, and
// the index arithmetic the adjoint scatters through (a + {0, dx} + {0, dy} +
// {0, dz}).  x-pairs are loaded at bx = min(x0, n-2) (n >= 2) and the two
// corners picked from the pair, so no load leaves the brick.
struct Taps {
    uint32_t a, dx, dy, dz;
    float fx, fy, fz;
    int x0, y0, z0; // node of corner 000
};

template <class T>
__device__ __forceinline__ Taps fetch_corners(const T* __restrict__ brick, int n, float tx, float ty,
                                              float tz, float v[8])
{
    int x0, x1, y0, y1, z0, z1;
    Taps t;
    axis_taps(tx, n, x0, x1, t.fx);
    axis_taps(ty, n, y0, y1, t.fy);
    axis_taps(tz, n, z0, z1, t.fz);
    t.x0 = x0;
    t.y0 = y0;
    t.z0 = z0;
    t.a = ((uint32_t)z0 * n + y0) * n + x0;
    t.dx = (uint32_t)(x1 - x0);
    t.dy = (uint32_t)(y1 - y0) * n;
    t.dz = (uint32_t)(z1 - z0) * n * n;
    const int bx = x0 < n - 2 ? x0 : n - 2;
    const uint32_t row = t.a - (uint32_t)x0 + (uint32_t)bx;
    const bool l0 = x0 == bx, l1 = x1 == bx;
    const uint32_t off[4] = {0u, t.dy, t.dz, t.dz + t.dy};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float p0, p1;
        load_pair(brick, row + off[k], p0, p1);
        v[2 * k] = l0 ? p0 : p1;
        v[2 * k + 1] = l1 ? p0 : p1;
    }
    return t;
}

// An fp32 brick in 4 x 4 x 4 micro-bricks (tiled4_index): the 8 corners of a
// cell fall into at most 8 micro-bricks of 256 bytes, and the corners of
// neighbouring rays into the same ones, so a wave's corner loads touch fewer
// cache lines than in the x-fastest layout.  The values, and so every
// result, are the same.
struct Tiled4 {
    float v;
};

__device__ __forceinline__ Taps fetch_corners(const Tiled4* __restrict__ brick, int n, float tx, float ty,
                                              float tz, float v[8])
{
    int x0, x1, y0, y1, z0, z1;
    Taps t;
    axis_taps(tx, n, x0, x1, t.fx);
    axis_taps(ty, n, y0, y1, t.fy);
    axis_taps(tz, n, z0, z1, t.fz);
    t.x0 = x0;
    t.y0 = y0;
    t.z0 = z0;
    t.a = ((uint32_t)z0 * n + y0) * n + x0;
    t.dx = (uint32_t)(x1 - x0);
    t.dy = (uint32_t)(y1 - y0) * n;
    t.dz = (uint32_t)(z1 - z0) * n * n;
    // tiled4_index split per axis, (c >> 2) * (micro-brick stride) + (c & 3)
    // * (node stride), the three terms summed (x-positions are nodes + 1).
    // One load per corner: the exact sequence is bound by its arithmetic,
    // and the pair loads of the fast fetch cost it more instructions than
    // they save.
    const uint32_t G = (((uint32_t)n + 1u) >> 2) + 1u, sy = G * 80u, sz = G * G * 80u;
    const uint32_t px0 = (uint32_t)x0 + 1u, px1 = (uint32_t)x1 + 1u;
    const uint32_t X0 = (px0 >> 2) * 80u + (px0 & 3u);
    const uint32_t X1 = (px1 >> 2) * 80u + (px1 & 3u);
    const uint32_t Y0 = __umul24((uint32_t)y0 >> 2, sy) + ((uint32_t)y0 & 3u) * 5u;
    const uint32_t Y1 = __umul24((uint32_t)y1 >> 2, sy) + ((uint32_t)y1 & 3u) * 5u;
    const uint32_t Z0 = __umul24((uint32_t)z0 >> 2, sz) + ((uint32_t)z0 & 3u) * 20u;
    const uint32_t Z1 = __umul24((uint32_t)z1 >> 2, sz) + ((uint32_t)z1 & 3u) * 20u;
    const gptr<char> b = global_of(reinterpret_cast<const char*>(brick));
    const uint32_t o[8] = {X0 + Y0 + Z0, X1 + Y0 + Z0, X0 + Y1 + Z0, X1 + Y1 + Z0,
                           X0 + Y0 + Z1, X1 + Y0 + Z1, X0 + Y1 + Z1, X1 + Y1 + Z1};
#pragma unroll
    for (int k = 0; k < 8; ++k)
        v[k] = *reinterpret_cast<gptr<float>>(b + (o[k] << 2));
    return t;
}

// The colour table into LDS (tid-th of nt threads), with one more entry,
// dst[n] = src[n - 1]: the upper neighbour the fast sequence reads for the
// last entry.  The LDS array holds n + 1 entries.
__device__ __forceinline__ void stage_tf(float4* dst, const float4* src, uint32_t n, uint32_t tid, uint32_t nt)
{
    for (uint32_t i = tid; i <= n && n > 0u; i += nt)
        dst[i] = src[i < n ? i : n - 1u];
}

// Exact fp32 trilinear fetch (the tex3D of RenderFuncInput.hpp:46-49).
__device__ __forceinline__ float trilinear(const float v[8], float fx, float fy, float fz)
{
    const float c00 = lerpf(v[0], v[1], fx);
    const float c10 = lerpf(v[2], v[3], fx);
    const float c01 = lerpf(v[4], v[5], fx);
    const float c11 = lerpf(v[6], v[7], fx);
    const float c0 = lerpf(c00, c10, fy);
    const float c1 = lerpf(c01, c11, fy);
    return lerpf(c0, c1, fz);
}

template <class T>
__device__ __forceinline__ float fetch3(const T* __restrict__ brick, int n, float tx, float ty,
                                        float tz)
{
    float v[8];
    const Taps t = fetch_corners(brick, n, tx, ty, tz, v);
    return trilinear(v, t.fx, t.fy, t.fz);
}

// Pixel of thread tid in a tw x th screen tile: each wave covers an 8 x 8
// square when both extents are multiples of 8 (a compact ray bundle, fewer
// brick cache lines per load), else rows of the tile.
__device__ __forceinline__ void tile_pixel(uint32_t tid, uint32_t tw, uint32_t th, uint32_t& px,
                                           uint32_t& py)
{
    if ((tw & 7u) == 0u && (th & 7u) == 0u) {
        const uint32_t w = tid >> 6, l = tid & 63u, sq = tw >> 3;
        px = (w % sq) * 8u + (l & 7u);
        py = (w / sq) * 8u + (l >> 3);
    } else {
        px = tid % tw;
        py = tid / tw;
    }
}

// Normalized, linear, clamped 1-D fetch of the colour table (tex1D,
// DenseFieldRenderer_cu.cu:573-589) from a float4 table in LDS.
__device__ __forceinline__ float4 fetch_tf(const float4* tf, int N, float u)
{
    float x = u * (float)N - 0.5f;
    if (!(x > -1.f))
        x = -1.f;
    if (x > (float)N)
        x = (float)N;
    const float fl = floorf(x);
    const float f = x - fl;
    const int i = (int)fl;
    const float4 a = tf[clampi(i, 0, N - 1)];
    const float4 b = tf[clampi(i + 1, 0, N - 1)];
    return make_float4(lerpf(a.x, b.x, f), lerpf(a.y, b.y, f), lerpf(a.z, b.z, f),
                       lerpf(a.w, b.w, f));
}

struct RaySeg {
    bool hit;
    int sample_count;
    float px, py, pz;   // first sample position
    float sx, sy, sz;   // step
};

// Ray setup and slab test, DenseFieldRenderer_cu.cu:66-181.
__device__ __forceinline__ RaySeg setup_ray(const rv_block_params& p, uint32_t vx, uint32_t vy)
{
    RaySeg r;
    r.hit = false;
    r.sample_count = 0;
    r.px = r.py = r.pz = r.sx = r.sy = r.sz = 0.f;
    const float x = (-0.5f + (vx + 0.5f) / p.vp_w) * p.w;
    const float y = (-0.5f + (vy + 0.5f) / p.vp_h) * p.h;
    const float rx = (p.n[0] + p.e1[0] * x) + p.e2[0] * y;
    const float ry = (p.n[1] + p.e1[1] * x) + p.e2[1] * y;
    const float rz = (p.n[2] + p.e1[2] * x) + p.e2[2] * y;
    const float rv[3] = {rx, ry, rz};
    const float ev[3] = {p.pos[0], p.pos[1], p.pos[2]};
    float tauEnter = 1.f, tauExit = 0.f;
    const float h = p.half;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!(fabsf(rv[a]) > 1e-10f))
            continue;
        const int o1 = a == 0 ? 1 : 0, o2 = a == 2 ? 1 : 2;
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const float face = side == 0 ? p.center[a] - h : p.center[a] + h;
            const float tau = (face - ev[a]) / rv[a];
            const float q1 = ev[o1] + rv[o1] * tau;
            const float q2 = ev[o2] + rv[o2] * tau;
            if (fabsf(q1 - p.center[o1]) <= h && fabsf(q2 - p.center[o2]) <= h) {
                const bool enter = side == 0 ? (rv[a] > 0) : !(rv[a] > 0);
                if (enter)
                    tauEnter = tau;
                else
                    tauExit = tau;
            }
        }
    }
    if (tauEnter < 1)
        tauEnter = 1;
    if (!(tauExit > tauEnter))
        return r;
    r.hit = true;
    const float dt = tauExit - tauEnter;
    const float dotrr = (rx * rx + ry * ry) + rz * rz;
    const float l1 = (fabsf(rx) + fabsf(ry)) + fabsf(rz);
    r.sample_count = (int)(p.quality * dt * dotrr / (p.texel * l1));
    const float k = dt / (float)r.sample_count;
    r.sx = rx * k;
    r.sy = ry * k;
    r.sz = rz * k;
    r.px = ev[0] + rx * tauEnter;
    r.py = ev[1] + ry * tauEnter;
    r.pz = ev[2] + rz * tauEnter;
    return r;
}

// One sample's colour after the transfer table and the opacity correction
// (renderArgb.hpp:42-45); v must not be NaN.
__device__ __forceinline__ float4 classify(const rv_block_params& p, const float4* tf, int tf_n,
                                           float v)
{
    const float v01 = (v - p.fmin) / (p.fmax - p.fmin);
    float4 c = fetch_tf(tf, tf_n, v01);
    c.w = 1.f - pow01(1.f - c.w * p.opacity_factor, p.i_layer_count);
    return c;
}

// blendBehind (cuda_blend.hpp:24-33): acc is in front of c.
__device__ __forceinline__ void blend_behind(float4& acc, const float4& c)
{
    const float da = c.w * (1.f - acc.w);
    const float a = acc.w + da;
    if (a > 0.f) {
        const float ka = acc.w / a, kb = da / a;
        acc.x = acc.x * ka + c.x * kb;
        acc.y = acc.y * ka + c.y * kb;
        acc.z = acc.z * ka + c.z * kb;
    }
    acc.w = a;
}

// uint32ArgbToFloatRgba + blendAbove + floatRgbaToUint32Argb
// (cuda_argb_transform.hpp:26-43, cuda_blend.hpp:35-44).
__device__ __forceinline__ uint32_t blend_above_u32(uint32_t argb, const float4& in)
{
    float ex = ((argb >> 16) & 0xffu) / 255.f;
    float ey = ((argb >> 8) & 0xffu) / 255.f;
    float ez = (argb & 0xffu) / 255.f;
    const float ew = ((argb >> 24) & 0xffu) / 255.f;
    const float da = ew * (1.f - in.w);
    const float a = in.w + da;
    if (a > 0.f) {
        const float ka = da / a, kb = in.w / a;
        ex = ex * ka + in.x * kb;
        ey = ey * ka + in.y * kb;
        ez = ez * ka + in.z * kb;
    }
    return ((unsigned)(a * 255.999f) << 24) + ((unsigned)(ex * 255.999f) << 16) +
           ((unsigned)(ey * 255.999f) << 8) + (unsigned)(ez * 255.999f);
}

// March one block segment from rgba = 0 (renderArgb.hpp:39-51).
// The march is latency bound (each sample waits for its 8 brick loads), so
// the positions and brick fetches of kBatch consecutive samples are issued
// together and then classified and composited in order; a segment that
// terminates early discards the rest of its batch.  Positions advance by the
// same sequential fp32 additions as the reference loop, so every composited
// sample is bit-identical to the one-at-a-time march.
#ifndef RV_MARCH_BATCH
#define RV_MARCH_BATCH 4
#endif
constexpr int kBatch = RV_MARCH_BATCH;

// A block's parameters as lane l holds them, for the whole wave (every
// 32-bit word read out of lane l).
__device__ __forceinline__ rv_block_params lane_params(const rv_block_params& p, int l)
{
    constexpr int kWords = (int)(sizeof(rv_block_params) / 4);
    uint32_t w[kWords];
    __builtin_memcpy(w, &p, sizeof(w));
#pragma unroll
    for (int k = 0; k < kWords; ++k)
        w[k] = (uint32_t)__builtin_amdgcn_readlane((int)w[k], l);
    rv_block_params u;
    __builtin_memcpy(&u, w, sizeof(w));
    return u;
}

template <class T> __device__ __forceinline__ const T* lane_ptr(const T* q, int l)
{
    const uint64_t v = reinterpret_cast<uint64_t>(q);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
    return reinterpret_cast<const T*>(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ float lane_f(float x, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// The exact sequence's march.  kTrans: also return in *trans the product
// of (1 - alpha) over the composited samples (the segment's transmittance,
// as a product rather than 1 - alpha, for the adjoint's record of the pixel
// state at each block).  kCoop and kLaneBlocks as for the fast
// sequence (march_fast_impl): called by every lane of the wave, each lane
// marches at most the cap, and a lane with samples left is continued by the
// whole wave, 64 samples at a time -- positions by the same chain of
// additions, fetch, trilinear value and classification one sample per lane,
// blendBehind (and the transmittance product) in sample order from values
// read out of the lanes -- bit-identical to the one-lane march.
template <class T, bool kTrans, bool kCoop = false, bool kLaneBlocks = false>
__device__ __forceinline__ float4 march_impl(const rv_block_params& p, const T* __restrict__ brick,
                                             const float4* tf, int tf_n, RaySeg r, float* trans)
{
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    float tr = 1.f;
    const int n = (int)p.nodes;
    int cnt = r.sample_count;
    int cap = 0;
    bool coop = false;
    if constexpr (kCoop) {
        cap = (int)(2.f * p.quality * (float)n);
        cap = max(kBatch, (cap + kBatch - 1) / kBatch * kBatch); // whole batches
        coop = __ballot(true) == ~0ull;                          // the whole wave is here
        if (coop)
            cnt = min(cnt, cap + 1);
    }
    bool term = false;
    for (int i = 1; i < cnt; i += kBatch) {
        // the corner loads of the whole batch are issued before any is used;
        // samples past the segment's end read clamped in-brick corners and
        // are dropped
        float c[kBatch][8];
        Taps t[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            const float tx = (r.px - p.tex_origin[0]) * p.tcf;
            const float ty = (r.py - p.tex_origin[1]) * p.tcf;
            const float tz = (r.pz - p.tex_origin[2]) * p.tcf;
            t[k] = fetch_corners(brick, n, tx, ty, tz, c[k]);
            r.px = r.px + r.sx;
            r.py = r.py + r.sy;
            r.pz = r.pz + r.sz;
        }
        float v[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; ++k)
            v[k] = trilinear(c[k], t[k].fx, t[k].fy, t[k].fz) + (i + k < r.sample_count ? 0.f : __builtin_nanf(""));
        bool done = false;
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            if (!done && v[k] == v[k]) {
                const float4 c = classify(p, tf, tf_n, v[k]);
                blend_behind(acc, c);
                if (kTrans)
                    tr = tr * (1.f - c.w);
                done = acc.w > p.alpha_cutoff;
            }
        }
        if (done) {
            term = true;
            break;
        }
    }
    if constexpr (kCoop) {
        uint64_t left = coop ? __ballot(!term && r.sample_count > cap + 1) : 0ull;
        const int me = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
        while (left) {
            const int L = (int)__builtin_ctzll(left);
            left &= left - 1ull;
            rv_block_params pu = p;
            const T* ub = brick;
            int ucap = cap;
            if constexpr (kLaneBlocks) {
                pu = lane_params(p, L);
                ub = lane_ptr(brick, L);
                ucap = __builtin_amdgcn_readlane(cap, L);
            }
            float Px = lane_f(r.px, L), Py = lane_f(r.py, L), Pz = lane_f(r.pz, L);
            const float Sx = lane_f(r.sx, L), Sy = lane_f(r.sy, L), Sz = lane_f(r.sz, L);
            const int N = __builtin_amdgcn_readlane(r.sample_count, L);
            float4 ua = make_float4(lane_f(acc.x, L), lane_f(acc.y, L), lane_f(acc.z, L), lane_f(acc.w, L));
            float utr = lane_f(tr, L);
            bool udone = false;
            for (int base = ucap + 1; base < N && !udone; base += 64) {
                float qx = 0.f, qy = 0.f, qz = 0.f;
#pragma unroll 1
                for (int j = 0; j < 64; ++j) {
                    if (j == me) {
                        qx = Px;
                        qy = Py;
                        qz = Pz;
                    }
                    Px = Px + Sx;
                    Py = Py + Sy;
                    Pz = Pz + Sz;
                }
                float cc[8];
                const float tx = (qx - pu.tex_origin[0]) * pu.tcf;
                const float ty = (qy - pu.tex_origin[1]) * pu.tcf;
                const float tz = (qz - pu.tex_origin[2]) * pu.tcf;
                const Taps tt = fetch_corners(ub, (int)pu.nodes, tx, ty, tz, cc);
                const float v = trilinear(cc, tt.fx, tt.fy, tt.fz) + (base + me < N ? 0.f : __builtin_nanf(""));
                const bool valid = v == v;
                float4 cs = make_float4(0.f, 0.f, 0.f, 0.f);
                if (valid)
                    cs = classify(pu, tf, tf_n, v);
                uint64_t vm = __ballot(valid);
                while (vm) {
                    const int j = (int)__builtin_ctzll(vm);
                    vm &= vm - 1ull;
                    const float4 cj = make_float4(lane_f(cs.x, j), lane_f(cs.y, j), lane_f(cs.z, j), lane_f(cs.w, j));
                    blend_behind(ua, cj);
                    if (kTrans)
                        utr = utr * (1.f - cj.w);
                    if (ua.w > pu.alpha_cutoff) {
                        udone = true;
                        break;
                    }
                }
            }
            if (me == L) {
                acc = ua;
                tr = utr;
            }
        }
    }
    if (kTrans)
        *trans = tr;
    return acc;
}

template <class T>
__device__ __forceinline__ float4 march(const rv_block_params& p, const T* __restrict__ brick,
                                        const float4* tf, int tf_n, RaySeg r)
{
    return march_impl<T, false>(p, brick, tf, tf_n, r, nullptr);
}

// ---------------------------------------------------------------- fast sample sequence
// The default sequence of the Argb ray-casts (rv_block_params.flags without
// RV_RENDER_EXACT): the same samples as march() -- the same positions (the
// same sequential fp32 steps) and the same 8 corners per sample, so the same
// samples are undefined (a NaN corner makes a sample undefined whatever its
// weight, and a position moved by 1e-6 texel can change which cell, hence
// which corners, it reads) -- with fewer operations for what varies
// continuously:
//  * each trilinear step is one fma, a + t (b - a);
//  * the table coordinate is one fma of the field value (the range
//    normalisation folded in), clamped with one med3;
//  * the opacity correction 1 - (1 - a of)^k is 1 - 2^(k log2(1 - a of))
//    with the hardware log2 and exp2;
//  * compositing accumulates premultiplied colour and alpha front to back,
//    w = (1 - A) a, C += w c, A += w, with no division per sample; the
//    segment's colour is divided by its alpha once at the end (alpha is
//    accumulated itself, not as 1 - transmittance, so a nearly transparent
//    segment keeps an accurate colour).
// The results agree with the exact sequence to a few 1e-6 per channel; a
// segment may end one sample earlier or later at the early-termination test.

// One axis of a sample: the clamped taps of axis_taps with med3s.
__device__ __forceinline__ void fast_axis(float t, int n, uint32_t& i0, uint32_t& i1, float& f)
{
    const float v = __builtin_amdgcn_fmed3f(t - 0.5f, -1.f, (float)n);
    const float fl = floorf(v);
    f = v - fl;
    const int i = (int)fl; // -1 <= i <= n
    i0 = (uint32_t)min(max(i, 0), n - 1);
    i1 = (uint32_t)min(i + 1, n - 1);
}

// The 8 corners of a sample (those of fetch_corners), x-fastest bricks (fp32
// or fp16): the x-pairs are loaded as in fetch_corners.
template <class T>
__device__ __forceinline__ void fetch_fast(const T* __restrict__ brick, int n, float tx, float ty, float tz,
                                           float v[8], float& fx, float& fy, float& fz)
{
    uint32_t x0, x1, y0, y1, z0, z1;
    fast_axis(tx, n, x0, x1, fx);
    fast_axis(ty, n, y0, y1, fy);
    fast_axis(tz, n, z0, z1, fz);
    const uint32_t un = (uint32_t)n;
    const uint32_t bx = x0 < un - 2u ? x0 : un - 2u;
    const uint32_t row = __umul24(__umul24(z0, un) + y0, un) + bx;
    const uint32_t dy = __umul24(y1 - y0, un), dz = __umul24(__umul24(z1 - z0, un), un);
    const bool l0 = x0 == bx, l1 = x1 == bx;
    const uint32_t off[4] = {0u, dy, dz, dz + dy};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float p0, p1;
        load_pair(brick, row + off[k], p0, p1);
        v[2 * k] = l0 ? p0 : p1;
        v[2 * k + 1] = l1 ? p0 : p1;
    }
}

// ... and micro-brick (TILED4) bricks: the per-axis parts of tiled4_index,
// in bytes, summed per corner row as 32-bit offsets from the wave-uniform
// brick address (global loads with a scalar base, no 64-bit arithmetic per
// lane).  Along x the brick holds node c at position c + 1 and the edge
// nodes at the positions beyond, so the pair of a sample is positions
// (i + 1, i + 2) for its unclamped cell i in [-1, n], with no clamp (the
// pair of a clamped sample holds the edge node twice); micro-brick rows
// hold 5 x-positions, so the pair is one 8-byte load inside one row.  Four
// load instructions per sample: fewer L1 tag lookups, which bound this
// march.
__device__ __forceinline__ uint32_t tiled4_x_bytes(uint32_t c) { return __umul24(c >> 2, 320u) + ((c & 3u) << 2); }

__device__ __forceinline__ void tiled4_load_rows(const Tiled4* __restrict__ brick, const uint32_t row[4],
                                                 float v[8])
{
    const gptr<char> b = global_of(reinterpret_cast<const char*>(brick));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint64_t pr = *reinterpret_cast<gptr<pair32_t>>(b + row[k]);
        v[2 * k] = __uint_as_float((uint32_t)pr);
        v[2 * k + 1] = __uint_as_float((uint32_t)(pr >> 32));
    }
}

__device__ __forceinline__ void fetch_fast(const Tiled4* __restrict__ brick, int n, float tx, float ty, float tz,
                                           float v[8], float& fx, float& fy, float& fz)
{
    const float vx = __builtin_amdgcn_fmed3f(tx - 0.5f, -1.f, (float)n);
    const float flx = floorf(vx);
    fx = vx - flx;
    const uint32_t X = tiled4_x_bytes((uint32_t)((int)flx + 1)); // pair at x-positions i + 1, i + 2
    uint32_t y0, y1, z0, z1;
    fast_axis(ty, n, y0, y1, fy);
    fast_axis(tz, n, z0, z1, fz);
    const uint32_t G = (((uint32_t)n + 1u) >> 2) + 1u, sy = G * 320u, sz = G * G * 320u;
    const uint32_t Y0 = __umul24(y0 >> 2, sy) + (y0 & 3u) * 20u;
    const uint32_t Y1 = __umul24(y1 >> 2, sy) + (y1 & 3u) * 20u;
    const uint32_t Z0 = __umul24(z0 >> 2, sz) + (z0 & 3u) * 80u;
    const uint32_t Z1 = __umul24(z1 >> 2, sz) + (z1 & 3u) * 80u;
    const uint32_t XY0 = X + Y0, XY1 = X + Y1;
    const uint32_t row[4] = {XY0 + Z0, XY1 + Z0, XY0 + Z1, XY1 + Z1};
    tiled4_load_rows(brick, row, v);
}

// Corner-row byte offsets of TILED4 bricks along y and z, per depth 0..6
// (n = 2^d + 1 <= 65 nodes), for the fast fetch: entry j in [0, n + 1] of
// depth d holds the offsets of the corners (clamp(j - 1), clamp(j)) of a
// sample whose unclamped cell (fast_axis) is j - 1, so one LDS load replaces
// the clamps and the micro-brick decomposition of both corners.  The y
// tables of depths 0..6 come first, then the z tables, kYZTabEntries each.
constexpr uint32_t kYZTabDepths = 7;
constexpr uint32_t kYZTabEntries = 148; // sum over d = 0..6 of 2^d + 3

__host__ __device__ constexpr uint32_t yz_tab_base(uint32_t d) { return ((1u << d) - 1u) + 3u * d; }

// The tables into LDS (2 kYZTabEntries entries), tid-th of nt threads.
__device__ inline void stage_yz_tables(uint2* tab, uint32_t tid, uint32_t nt)
{
    for (uint32_t k = tid; k < 2u * kYZTabEntries; k += nt) {
        const uint32_t axis = k >= kYZTabEntries ? 1u : 0u;
        const uint32_t e = k - axis * kYZTabEntries;
        uint32_t d = 0;
        while (d + 1u < kYZTabDepths && yz_tab_base(d + 1u) <= e)
            ++d;
        const int n = (1 << d) + 1, j = (int)(e - yz_tab_base(d));
        const uint32_t G = (((uint32_t)n + 1u) >> 2) + 1u;
        const uint32_t ms = axis ? G * G * 320u : G * 320u, ns = axis ? 80u : 20u; // micro-brick, node strides
        const uint32_t c0 = (uint32_t)min(max(j - 1, 0), n - 1), c1 = (uint32_t)min(j, n - 1);
        tab[k] = make_uint2((c0 >> 2) * ms + (c0 & 3u) * ns, (c1 >> 2) * ms + (c1 & 3u) * ns);
    }
}

// fetch_fast of a TILED4 brick of depth <= 6 with the y and z offsets from
// the tables of its depth (ytab, ztab): the same corners and weights.
__device__ __forceinline__ void fetch_fast_tab(const Tiled4* __restrict__ brick, int n, float tx, float ty, float tz,
                                               const uint2* ytab, const uint2* ztab, float v[8], float& fx,
                                               float& fy, float& fz)
{
    const float vx = __builtin_amdgcn_fmed3f(tx - 0.5f, -1.f, (float)n);
    const float vy = __builtin_amdgcn_fmed3f(ty - 0.5f, -1.f, (float)n);
    const float vz = __builtin_amdgcn_fmed3f(tz - 0.5f, -1.f, (float)n);
    const float flx = floorf(vx), fly = floorf(vy), flz = floorf(vz);
    fx = vx - flx;
    fy = vy - fly;
    fz = vz - flz;
    const uint32_t X = tiled4_x_bytes((uint32_t)((int)flx + 1));
    const uint2 Y = ytab[(int)fly + 1], Z = ztab[(int)flz + 1];
    const uint32_t XY0 = X + Y.x, XY1 = X + Y.y;
    const uint32_t row[4] = {XY0 + Z.x, XY1 + Z.x, XY0 + Z.y, XY1 + Z.y};
    tiled4_load_rows(brick, row, v);
}

__device__ __forceinline__ float lerp_fast(float a, float b, float t) { return fmaf(t, b - a, a); }

// The constants of the fast sequence over one block.
struct FastBlock {
    float ox, oy, oz, tcf; // texture coordinate (p - tex_origin) tcf
    int n;                 // nodes per edge
    uint32_t depth;        // tex_depth (the y/z tables of TILED4 bricks)
    float ta, tb, thi;     // table coordinate v ta + tb, clamped to [0, thi]
    float k, of, cut;      // opacity correction exponent and factor, cutoff
};

__device__ __forceinline__ FastBlock fast_block(const rv_block_params& p, int tf_n)
{
    FastBlock b;
    b.ox = p.tex_origin[0];
    b.oy = p.tex_origin[1];
    b.oz = p.tex_origin[2];
    b.tcf = p.tcf;
    b.n = (int)p.nodes;
    b.depth = p.tex_depth;
    // table coordinate u N - 0.5 with u = (v - fmin) / (fmax - fmin)
    b.ta = (float)tf_n / (p.fmax - p.fmin);
    b.tb = fmaf(-p.fmin, b.ta, -0.5f);
    b.thi = (float)(tf_n - 1);
    b.k = p.i_layer_count;
    b.of = p.opacity_factor;
    b.cut = p.alpha_cutoff;
    return b;
}


// A block's constants as lane l holds them, for the whole wave.
__device__ __forceinline__ FastBlock lane_block(const FastBlock& b, int l)
{
    FastBlock u;
    u.ox = lane_f(b.ox, l);
    u.oy = lane_f(b.oy, l);
    u.oz = lane_f(b.oz, l);
    u.tcf = lane_f(b.tcf, l);
    u.n = __builtin_amdgcn_readlane(b.n, l);
    u.depth = (uint32_t)__builtin_amdgcn_readlane((int)b.depth, l);
    u.ta = lane_f(b.ta, l);
    u.tb = lane_f(b.tb, l);
    u.thi = lane_f(b.thi, l);
    u.k = lane_f(b.k, l);
    u.of = lane_f(b.of, l);
    u.cut = lane_f(b.cut, l);
    return u;
}


// One sample's corners and weights at a world position.
template <class T, bool kTab>
__device__ __forceinline__ void fast_fetch_at(const FastBlock& b, const T* __restrict__ brick, const uint2* yz,
                                              float px, float py, float pz, float c[8], float& fx, float& fy,
                                              float& fz)
{
    const float tx = (px - b.ox) * b.tcf, ty = (py - b.oy) * b.tcf, tz = (pz - b.oz) * b.tcf;
    if constexpr (kTab) {
        const uint2* ytab = yz + yz_tab_base(b.depth);
        fetch_fast_tab(brick, b.n, tx, ty, tz, ytab, ytab + kYZTabEntries, c, fx, fy, fz);
    } else {
        fetch_fast(brick, b.n, tx, ty, tz, c, fx, fy, fz);
    }
}

__device__ __forceinline__ float fast_value(const float c[8], float fx, float fy, float fz)
{
    const float c00 = lerp_fast(c[0], c[1], fx), c10 = lerp_fast(c[2], c[3], fx);
    const float c01 = lerp_fast(c[4], c[5], fx), c11 = lerp_fast(c[6], c[7], fx);
    return lerp_fast(lerp_fast(c00, c10, fy), lerp_fast(c01, c11, fy), fz);
}

// The opacity-corrected alpha and the colour of a (non-NaN) value.
__device__ __forceinline__ void fast_classify(const FastBlock& b, const float4* tf, float v, float& al, float& qr,
                                              float& qg, float& qb)
{
    const float x = __builtin_amdgcn_fmed3f(fmaf(v, b.ta, b.tb), 0.f, b.thi);
    const float fl = floorf(x);
    const float f = x - fl;
    // tf[tf_n] repeats tf[tf_n - 1] (stage_tf), so i0 + 1 needs no clamp
    const int i0 = (int)fl;
    const float4 ca = tf[i0], cb = tf[i0 + 1];
    al = 1.f - __builtin_amdgcn_exp2f(b.k * __builtin_amdgcn_logf(fmaf(-lerp_fast(ca.w, cb.w, f), b.of, 1.f)));
    qr = lerp_fast(ca.x, cb.x, f);
    qg = lerp_fast(ca.y, cb.y, f);
    qb = lerp_fast(ca.z, cb.z, f);
}

// Front to back, premultiplied: w = (1 - A) al.
__device__ __forceinline__ void fast_over(float al, float qr, float qg, float qb, float& cr, float& cg, float& cb,
                                          float& A)
{
    const float w = fmaf(-A, al, al);
    cr = fmaf(w, qr, cr);
    cg = fmaf(w, qg, cg);
    cb = fmaf(w, qb, cb);
    A = A + w;
}

// A ray that grazes a block's edge can miss its entry face in fp32 and
// then starts at the eye (tauEnter = 1, as in the reference), so its segment
// is many times the block's length: a few rays per frame at some cameras,
// each holding its wave for thousands of samples.  With kCoop (a call by
// every lane of the wave; lanes with nothing to march pass sample_count 0)
// every lane marches at most the cap (twice the samples of the longest ray
// inside its block); each lane with samples left is then continued by the
// whole wave, 64 samples at a time: the positions by the same chain of fp32
// additions (every lane runs it and keeps its own), the loads, values and
// table lookups one sample per lane, and the composite in sample order from
// values read out of each lane.  Every operation is the one the lane would
// have done itself, so the result is bit-identical to the one-lane march.
// kLaneBlocks: the lanes may be on different blocks (and bricks), so the
// continued lane's block constants and brick are read out of it too.
template <class T, bool kTab, bool kCoop, bool kLaneBlocks>
__device__ __forceinline__ float4 march_fast_impl(const rv_block_params& p, const T* __restrict__ brick,
                                                  const float4* tf, int tf_n, RaySeg r, const uint2* yz)
{
    const FastBlock fb = fast_block(p, tf_n);
    float cr = 0.f, cg = 0.f, cb = 0.f, A = 0.f; // premultiplied colour, alpha
    int cnt = r.sample_count;
    int cap = 0;
    bool coop = false;
    if constexpr (kCoop) {
        cap = (int)(2.f * p.quality * (float)fb.n);
        cap = max(kBatch, (cap + kBatch - 1) / kBatch * kBatch); // whole batches
        coop = __ballot(true) == ~0ull;                          // the whole wave is here
        if (coop)
            cnt = min(cnt, cap + 1);
    }
    bool term = false;
    for (int i = 1; i < cnt; i += kBatch) {
        // the corner loads of the whole batch are issued before any is used;
        // samples past the segment's end read clamped in-brick corners and
        // are dropped
        float c[kBatch][8], fx[kBatch], fy[kBatch], fz[kBatch];
#pragma unroll
        for (int s = 0; s < kBatch; ++s) {
            fast_fetch_at<T, kTab>(fb, brick, yz, r.px, r.py, r.pz, c[s], fx[s], fy[s], fz[s]);
            r.px = r.px + r.sx;
            r.py = r.py + r.sy;
            r.pz = r.pz + r.sz;
        }
        float v[kBatch];
#pragma unroll
        for (int s = 0; s < kBatch; ++s) {
            // + NaN past the end (an add, not a select, so no sample's loads
            // move into a branch)
            v[s] = fast_value(c[s], fx[s], fy[s], fz[s]) + (i + s < r.sample_count ? 0.f : __builtin_nanf(""));
        }
        bool done = false;
#pragma unroll
        for (int s = 0; s < kBatch; ++s) {
            if (!done && v[s] == v[s]) {
                float al, qr, qg, qb;
                fast_classify(fb, tf, v[s], al, qr, qg, qb);
                fast_over(al, qr, qg, qb, cr, cg, cb, A);
                done = A > fb.cut;
            }
        }
        if (done) {
            term = true;
            break;
        }
    }
    if constexpr (kCoop) {
        // lanes with samples left: sample i (from cap + 1 on) sits at the
        // position the chain of i - 1 additions reaches
        uint64_t left = coop ? __ballot(!term && r.sample_count > cap + 1) : 0ull;
        const int me = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
        while (left) {
            const int L = (int)__builtin_ctzll(left);
            left &= left - 1ull;
            FastBlock ub = fb;
            const T* ubrick = brick;
            int ucap = cap;
            if constexpr (kLaneBlocks) {
                ub = lane_block(fb, L);
                ubrick = lane_ptr(brick, L);
                ucap = __builtin_amdgcn_readlane(cap, L);
            }
            float Px = lane_f(r.px, L), Py = lane_f(r.py, L), Pz = lane_f(r.pz, L);
            const float Sx = lane_f(r.sx, L), Sy = lane_f(r.sy, L), Sz = lane_f(r.sz, L);
            const int N = __builtin_amdgcn_readlane(r.sample_count, L);
            float ur = lane_f(cr, L), ug = lane_f(cg, L), ubl = lane_f(cb, L), uA = lane_f(A, L);
            bool udone = false;
            for (int base = ucap + 1; base < N && !udone; base += 64) {
                float qx = 0.f, qy = 0.f, qz = 0.f;
#pragma unroll 1
                for (int j = 0; j < 64; ++j) {
                    if (j == me) {
                        qx = Px;
                        qy = Py;
                        qz = Pz;
                    }
                    Px = Px + Sx;
                    Py = Py + Sy;
                    Pz = Pz + Sz;
                }
                float c[8], fx, fy, fz;
                fast_fetch_at<T, kTab>(ub, ubrick, yz, qx, qy, qz, c, fx, fy, fz);
                const float v = fast_value(c, fx, fy, fz) + (base + me < N ? 0.f : __builtin_nanf(""));
                const bool valid = v == v;
                float al = 0.f, qr = 0.f, qg = 0.f, qb = 0.f;
                if (valid)
                    fast_classify(ub, tf, v, al, qr, qg, qb);
                uint64_t vm = __ballot(valid);
                while (vm) {
                    const int j = (int)__builtin_ctzll(vm);
                    vm &= vm - 1ull;
                    fast_over(lane_f(al, j), lane_f(qr, j), lane_f(qg, j), lane_f(qb, j), ur, ug, ubl, uA);
                    if (uA > ub.cut) {
                        udone = true;
                        break;
                    }
                }
            }
            if (me == L) {
                cr = ur;
                cg = ug;
                cb = ubl;
                A = uA;
            }
        }
    }
    if (!(A > 0.f))
        return make_float4(0.f, 0.f, 0.f, 0.f);
    const float ia = 1.f / A;
    return make_float4(cr * ia, cg * ia, cb * ia, A);
}

// The fast sequence over one block; yz: the LDS offset tables
// (stage_yz_tables) or nullptr.  TILED4 bricks of depth <= 6 read their y
// and z offsets from the tables.  kCoop: called by every lane of the wave
// (lanes with nothing to march pass sample_count 0), so overlong segments
// are shared by the wave (march_fast_impl); kLaneBlocks: the lanes may be
// on different blocks.  A wave whose lanes take both table forms continues
// them one lane each.
template <class T, bool kCoop = false, bool kLaneBlocks = false>
__device__ __forceinline__ float4 march_fast(const rv_block_params& p, const T* __restrict__ brick,
                                             const float4* tf, int tf_n, RaySeg r, const uint2* yz = nullptr)
{
    if constexpr (std::is_same<T, Tiled4>::value) {
        if (yz != nullptr && p.tex_depth < kYZTabDepths)
            return march_fast_impl<T, true, kCoop, kLaneBlocks>(p, brick, tf, tf_n, r, yz);
    }
    return march_fast_impl<T, false, kCoop, kLaneBlocks>(p, brick, tf, tf_n, r, yz);
}

// ---------------------------------------------------------------- field modes
// s3vs::FieldMode (FieldMode.hpp:26-51), the same numbering.
constexpr uint32_t kModeIsosurface = 0, kModeIsosurfaces = 1, kModeMip = 2, kModeArgb = 3,
                   kModeArgbLight = 4, kModeDomainVoxels = 5, kModeValueOnIsosurface = 6,
                   kModeValueOnIsosurfaces = 7;
// Argb with the fast sample sequence (march_fast): a kernel instantiation of
// its own, not a field mode of the ABI.
constexpr uint32_t kModeArgbFast = 8;

// The secondary field of ValueOnIsosurface(s): the block's brick of it (the
// primary brick's layout) and the field's range (RenderKernelParam.hpp:103,
// DenseFieldRenderer_cu.cu:288-335).
template <class T> struct Secondary {
    const T* brick;
    float smin, smax;
};

// Levels are staged for these modes.
__host__ __device__ constexpr bool multi_level(uint32_t mode)
{
    return mode == kModeIsosurfaces || mode == kModeValueOnIsosurfaces;
}

// Isosurface levels of the Isosurfaces mode: absolute values, sorted,
// distinct, at most kMaxLevels (RenderKernelPresentationParam::MaxLevels,
// RenderKernelParam.hpp:75-84; makeLevels, DenseFieldRenderer_cu.cu:597-612).
constexpr uint32_t kMaxLevels = 50;
struct LevelsArg {
    uint32_t n;
    float v[kMaxLevels];
};

// makeLevels on the host: relative levels to absolute (relToAbsFieldValue,
// as the threshold), sorted, duplicates dropped, at most kMaxLevels kept.
inline LevelsArg make_levels(float fmin, float fmax, const float* rel, uint32_t count)
{
    LevelsArg L{};
    float buf[256];
    uint32_t m = 0;
    for (uint32_t i = 0; i < count && m < 256; ++i)
        buf[m++] = fmin + (fmax - fmin) * rel[i];
    for (uint32_t i = 1; i < m; ++i) // insertion sort (few levels)
        for (uint32_t j = i; j > 0 && buf[j] < buf[j - 1]; --j) {
            const float t = buf[j];
            buf[j] = buf[j - 1];
            buf[j - 1] = t;
        }
    for (uint32_t i = 0; i < m && L.n < kMaxLevels; ++i)
        if (L.n == 0 || buf[i] != L.v[L.n - 1])
            L.v[L.n++] = buf[i];
    return L;
}

struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// cuda_vec_op.hpp:48-52
__device__ __forceinline__ F3 normalize3(F3 a)
{
    const float il = 1.f / sqrtf(dot3(a, a));
    return f3(a.x * il, a.y * il, a.z * il);
}
// clamp.hpp:24-34 (max/min ignore a NaN operand)
__device__ __forceinline__ float clamp01(float x) { return fmaxf(0.f, fminf(1.f, x)); }

// The one light of DenseFieldRenderer_cu.cu:440-445: at eye + n + e2 h,
// ambient 0, diffuse 1, specular 0.1.
__device__ __forceinline__ F3 light_pos(const rv_block_params& p)
{
    return (f3(p.pos[0], p.pos[1], p.pos[2]) + f3(p.n[0], p.n[1], p.n[2])) +
           f3(p.e2[0], p.e2[1], p.e2[2]) * p.h;
}

// computeFragColor + blinn (cuda_lighting.hpp:27-95) for one light: the
// normal is the central difference of the field at +-0.1 texel, lit double
// sided; Ka = 1, Ks = 0.5, shininess 100.
template <class T>
__device__ inline float4 frag_color(const rv_block_params& p, const T* __restrict__ brick,
                                    const float4& diffuse, F3 surf, F3 tc)
{
    const int n = (int)p.nodes;
    constexpr float d = 0.1f;
    F3 a, b;
    a.x = fetch3(brick, n, tc.x - d, tc.y, tc.z);
    b.x = fetch3(brick, n, tc.x + d, tc.y, tc.z);
    a.y = fetch3(brick, n, tc.x, tc.y - d, tc.z);
    b.y = fetch3(brick, n, tc.x, tc.y + d, tc.z);
    a.z = fetch3(brick, n, tc.x, tc.y, tc.z - d);
    b.z = fetch3(brick, n, tc.x, tc.y, tc.z + d);
    F3 N = normalize3(a - b);
    const F3 V = normalize3(f3(p.pos[0], p.pos[1], p.pos[2]) - surf);
    const F3 L = normalize3(light_pos(p) - surf);
    if (dot3(L, N) < 0.f)
        N = f3(-N.x, -N.y, -N.z);
    const float diff = fmaxf(dot3(L, N), 0.f);
    const F3 H = normalize3(L + V);
    const float spec = diff > 0.f ? pow01(fmaxf(dot3(H, N), 0.f), 100.f) : 0.f;
    // Ka * ambient + Kd * diffuse * diff + Ks * specular * spec, per channel
    const float ks = 0.5f * 0.1f;
    const float cx = (1.f * 0.f + (diffuse.x * 1.f) * diff) + ks * spec;
    const float cy = (1.f * 0.f + (diffuse.y * 1.f) * diff) + ks * spec;
    const float cz = (1.f * 0.f + (diffuse.z * 1.f) * diff) + ks * spec;
    return make_float4(clamp01(0.f + cx), clamp01(0.f + cy), clamp01(0.f + cz), diffuse.w);
}

// renderArgbLight loop (renderArgbLight.hpp:29-63): Argb classification,
// samples above alpha 1e-5 shaded, composited front to back.
template <class T>
__device__ inline float4 march_light(const rv_block_params& p, const T* __restrict__ brick,
                                     const float4* tf, int tf_n, RaySeg r)
{
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    const int n = (int)p.nodes;
    for (int i = 1; i < r.sample_count; ++i) {
        const F3 pos = f3(r.px, r.py, r.pz);
        const F3 tc = f3((r.px - p.tex_origin[0]) * p.tcf, (r.py - p.tex_origin[1]) * p.tcf,
                         (r.pz - p.tex_origin[2]) * p.tcf);
        r.px = r.px + r.sx;
        r.py = r.py + r.sy;
        r.pz = r.pz + r.sz;
        const float v = fetch3(brick, n, tc.x, tc.y, tc.z);
        if (!(v == v))
            continue;
        const float4 c = classify(p, tf, tf_n, v);
        if (c.w >= 0.00001f) {
            blend_behind(acc, frag_color(p, brick, c, pos, tc));
            if (acc.w > p.alpha_cutoff)
                break;
        }
    }
    return acc;
}

// Diffuse colour of a surface point of ValueOnIsosurface(s): the table
// colour of the secondary field's value there, normalised by its range, its
// alpha scaled by iso_opacity (renderValueOnIsosurface.hpp:44-48).
template <class T>
__device__ __forceinline__ float4 value_diffuse(const Secondary<T>& s, int n, const float4* tf, int tf_n,
                                                float iso_opacity, F3 tc)
{
    const float v = fetch3(s.brick, n, tc.x, tc.y, tc.z);
    float4 d = fetch_tf(tf, tf_n, (v - s.smin) / (s.smax - s.smin));
    d.w = d.w * iso_opacity;
    return d;
}

// renderIsosurface loop (renderIsosurface.hpp:31-61): where consecutive
// samples v1, v2 straddle the level (p.threshold), the surface point is
// placed at the second sample's position plus step * (level - v1) / (v2 - v1),
// as the reference does, and shaded with the table colour of the level, its
// alpha scaled by iso_opacity.  kValue: ValueOnIsosurface
// (renderValueOnIsosurface.hpp:32-62), the colour of the secondary field at
// the point instead.
template <bool kValue, class T>
__device__ inline float4 march_iso(const rv_block_params& p, const T* __restrict__ brick,
                                   const float4* tf, int tf_n, float iso_opacity, RaySeg r,
                                   Secondary<T> sec)
{
    const int n = (int)p.nodes;
    auto value_at = [&](float x, float y, float z) {
        return fetch3(brick, n, (x - p.tex_origin[0]) * p.tcf, (y - p.tex_origin[1]) * p.tcf,
                      (z - p.tex_origin[2]) * p.tcf);
    };
    float v1 = value_at(r.px, r.py, r.pz);
    r.px = r.px + r.sx;
    r.py = r.py + r.sy;
    r.pz = r.pz + r.sz;
    const float lv = p.threshold;
    float4 diffuse = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!kValue) {
        diffuse = fetch_tf(tf, tf_n, (lv - p.fmin) / (p.fmax - p.fmin));
        diffuse.w = diffuse.w * iso_opacity;
    }
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int i = 1; i < r.sample_count; ++i) {
        const float v2 = value_at(r.px, r.py, r.pz);
        if (v1 == v1 && v2 == v2 && (v1 - lv) * (v2 - lv) < 0.f) {
            const float t = (lv - v1) / (v2 - v1);
            const F3 x = f3(r.px, r.py, r.pz) + f3(r.sx, r.sy, r.sz) * t;
            const F3 tc = f3((x.x - p.tex_origin[0]) * p.tcf, (x.y - p.tex_origin[1]) * p.tcf,
                             (x.z - p.tex_origin[2]) * p.tcf);
            const float4 dif = kValue ? value_diffuse(sec, n, tf, tf_n, iso_opacity, tc) : diffuse;
            blend_behind(acc, frag_color(p, brick, dif, x, tc));
            if (acc.w > p.alpha_cutoff)
                break;
        }
        v1 = v2;
        r.px = r.px + r.sx;
        r.py = r.py + r.sy;
        r.pz = r.pz + r.sz;
    }
    return acc;
}

// renderMaxIntensityProjection loop (renderMaxIntensityProjection.hpp:31-37):
// the largest sample at or above p.threshold, or fmin - (fmax - fmin) when
// there is none.
template <class T>
__device__ inline float march_mip(const rv_block_params& p, const T* __restrict__ brick, RaySeg r)
{
    const int n = (int)p.nodes;
    float m = p.fmin - (p.fmax - p.fmin);
    for (int i = 1; i < r.sample_count; ++i) {
        const float v = fetch3(brick, n, (r.px - p.tex_origin[0]) * p.tcf,
                               (r.py - p.tex_origin[1]) * p.tcf, (r.pz - p.tex_origin[2]) * p.tcf);
        r.px = r.px + r.sx;
        r.py = r.py + r.sy;
        r.pz = r.pz + r.sz;
        if (v >= p.threshold && m < v)
            m = v;
    }
    return m;
}

// The viewport update of renderMaxIntensityProjection.hpp:39-48: the pixel
// holds float bits of the normalised maximum (NaN: none yet).
__device__ __forceinline__ uint32_t mip_merge(uint32_t prev_bits, float m, const rv_block_params& p)
{
    if (m < p.fmin)
        return prev_bits;
    const float t = (m - p.fmin) / (p.fmax - p.fmin);
    const float prev = __uint_as_float(prev_bits);
    if (prev == prev && prev > t)
        return prev_bits;
    return __float_as_uint(t);
}

// findInterval (cuda_findInterval.hpp:28-44): 0 <= i <= n with v between
// levels[i-1] and levels[i]; v must not be NaN.
__device__ __forceinline__ int find_interval(float v, const float* lv, uint32_t n)
{
    if (v < lv[0])
        return 0;
    if (v >= lv[n - 1])
        return (int)n;
    int i1 = 0, i2 = (int)n;
    while (i2 - i1 > 1) {
        const int i = (i1 + i2) >> 1;
        if (v < lv[i])
            i2 = i;
        else
            i1 = i;
    }
    return i2;
}

// renderIsosurfaces loop (renderIsosurfaces.hpp:31-103): every level crossed
// between consecutive samples is shaded at its point (placed as in
// march_iso) with the table colour of the level, in the order the ray
// crosses them.  lv: n >= 1 sorted distinct levels (the caller handles n = 0).
// kValue: ValueOnIsosurfaces (renderValueOnIsosurfaces.hpp:33-109), each
// point coloured by the secondary field there.
template <bool kValue, class T>
__device__ inline float4 march_isos(const rv_block_params& p, const T* __restrict__ brick,
                                    const float4* tf, int tf_n, float iso_opacity, const float* lv,
                                    uint32_t nlv, RaySeg r, Secondary<T> sec)
{
    const int n = (int)p.nodes;
    auto value_at = [&](float x, float y, float z) {
        return fetch3(brick, n, (x - p.tex_origin[0]) * p.tcf, (y - p.tex_origin[1]) * p.tcf,
                      (z - p.tex_origin[2]) * p.tcf);
    };
    float v1 = value_at(r.px, r.py, r.pz);
    r.px = r.px + r.sx;
    r.py = r.py + r.sy;
    r.pz = r.pz + r.sz;
    int iint = v1 == v1 ? find_interval(v1, lv, nlv) : -1;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int i = 1; i < r.sample_count; ++i) {
        const float v2 = value_at(r.px, r.py, r.pz);
        const float px = r.px, py = r.py, pz = r.pz;
        r.px = r.px + r.sx;
        r.py = r.py + r.sy;
        r.pz = r.pz + r.sz;
        if (!(v2 == v2)) {
            iint = -1;
            v1 = __builtin_nanf("");
            continue;
        }
        if (!(v1 == v1)) {
            v1 = v2;
            iint = find_interval(v1, lv, nlv);
            continue;
        }
        int il, d;
        if (v2 < v1) {
            if (iint == 0) {
                v1 = v2;
                continue;
            }
            il = iint - 1;
            d = -1;
        } else {
            if (iint == (int)nlv) {
                v1 = v2;
                continue;
            }
            il = iint;
            d = 1;
        }
        while (il >= 0 && il < (int)nlv) {
            const float level = lv[il];
            if (!((v1 - level) * (v2 - level) < 0.f))
                break;
            const float t = (level - v1) / (v2 - v1);
            const F3 x = f3(px, py, pz) + f3(r.sx, r.sy, r.sz) * t;
            const F3 tc = f3((x.x - p.tex_origin[0]) * p.tcf, (x.y - p.tex_origin[1]) * p.tcf,
                             (x.z - p.tex_origin[2]) * p.tcf);
            float4 diffuse;
            if (kValue) {
                diffuse = value_diffuse(sec, n, tf, tf_n, iso_opacity, tc);
            } else {
                diffuse = fetch_tf(tf, tf_n, (level - p.fmin) / (p.fmax - p.fmin));
                diffuse.w = diffuse.w * iso_opacity;
            }
            blend_behind(acc, frag_color(p, brick, diffuse, x, tc));
            iint += d;
            il += d;
        }
        if (acc.w > p.alpha_cutoff)
            break;
        v1 = v2;
    }
    return acc;
}

// renderDomainVoxels loop (renderDomainVoxels.hpp:30-54): every sample where
// the field is defined adds the table colour at 0.5, opacity-corrected.
template <class T>
__device__ inline float4 march_domain(const rv_block_params& p, const T* __restrict__ brick,
                                      const float4* tf, int tf_n, RaySeg r)
{
    const int n = (int)p.nodes;
    float4 diffuse = fetch_tf(tf, tf_n, 0.5f);
    diffuse.w = 1.f - pow01(1.f - diffuse.w * p.opacity_factor, p.i_layer_count);
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int i = 1; i < r.sample_count; ++i) {
        const float v = fetch3(brick, n, (r.px - p.tex_origin[0]) * p.tcf,
                               (r.py - p.tex_origin[1]) * p.tcf, (r.pz - p.tex_origin[2]) * p.tcf);
        r.px = r.px + r.sx;
        r.py = r.py + r.sy;
        r.pz = r.pz + r.sz;
        if (v == v) {
            blend_behind(acc, diffuse);
            if (acc.w > p.alpha_cutoff)
                break;
        }
    }
    return acc;
}

// One block segment of a pixel in field mode kMode, as a colour to blend
// (Argb, ArgbLight, Isosurface(s) and ValueOnIsosurface(s) with nlv >= 1
// levels where they take levels, DomainVoxels).  sec: the secondary field
// of the ValueOnIsosurface(s) modes.
template <uint32_t kMode, class T>
__device__ __forceinline__ float4 segment(const rv_block_params& p, const T* __restrict__ brick,
                                          const float4* tf, int tf_n, float iso_opacity, RaySeg r,
                                          const float* lv = nullptr, uint32_t nlv = 0,
                                          Secondary<T> sec = Secondary<T>{nullptr, 0.f, 1.f})
{
    if (kMode == kModeArgbLight)
        return march_light(p, brick, tf, tf_n, r);
    if (kMode == kModeIsosurface || kMode == kModeValueOnIsosurface)
        return march_iso<kMode == kModeValueOnIsosurface>(p, brick, tf, tf_n, iso_opacity, r, sec);
    if (kMode == kModeIsosurfaces || kMode == kModeValueOnIsosurfaces)
        return march_isos<kMode == kModeValueOnIsosurfaces>(p, brick, tf, tf_n, iso_opacity, lv, nlv,
                                                            r, sec);
    if (kMode == kModeDomainVoxels)
        return march_domain(p, brick, tf, tf_n, r);
    if (kMode == kModeArgbFast)
        return march_fast(p, brick, tf, tf_n, r);
    return march(p, brick, tf, tf_n, r);
}

} // namespace rv
