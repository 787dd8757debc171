// Plain per-block ray-cast (the drop-in for renderKernel<fieldMode>,
// DenseFieldRenderer_cu.cu:62-215, in the Argb, ArgbLight, Isosurface,
// Isosurfaces, DomainVoxels and MaxIntensityProjection modes), viewport clear, crop rectangle, 8-bit
// sub-image blending and the max-intensity compositing for gfx950.
//
// The block kernel runs one thread per pixel of the block's view rectangle in
// 16 x 16 pixel workgroups (one wave = 8 x 8 pixels, so a wave's rays are a
// compact bundle through the brick and share L1/L2 lines of the brick).  The
// colour table is staged in LDS as float4.  Every arithmetic step follows the
// reference's operand order (see render_common.hpp), so the 8-bit viewport is
// bit-identical to the CPU restatement.

#include "render_common.hpp"
#include "rv_internal.hpp"

namespace rv {

int fill_u32(uint32_t* d, uint64_t n, uint32_t value, hipStream_t s);

namespace {

constexpr uint32_t kTfMax = 4096;

template <uint32_t kMode>
__global__ __launch_bounds__(256) void render_block_kernel(rv_block_params p,
                                                           const float* __restrict__ brick,
                                                           Secondary<float> sec,
                                                           const float* __restrict__ tf,
                                                           uint32_t tf_n, float iso_opacity,
                                                           LevelsArg lev,
                                                           uint32_t* __restrict__ viewport)
{
    extern __shared__ float4 tfs[];
    __shared__ float lvs[kMaxLevels];
    stage_tf(tfs, reinterpret_cast<const float4*>(tf), tf_n, threadIdx.x, 256u);
    if (multi_level(kMode) && threadIdx.x < lev.n)
        lvs[threadIdx.x] = lev.v[threadIdx.x];
    __syncthreads();
    uint32_t lx, ly;
    tile_pixel(threadIdx.x, 16u, 16u, lx, ly);
    const uint32_t ox = blockIdx.x * 16u + lx, oy = blockIdx.y * 16u + ly;
    if (ox >= p.view_w || oy >= p.view_h)
        return;
    const uint32_t vx = p.view_x + ox, vy = p.view_y + oy;
    const RaySeg r = setup_ray(p, vx, vy);
    if (!r.hit)
        return;
    const uint32_t pix = (p.vp_h - 1u - vy) * p.vp_w + vx;
    if (kMode == kModeMip) {
        viewport[pix] = mip_merge(viewport[pix], march_mip(p, brick, r), p);
        return;
    }
    if (multi_level(kMode) && lev.n == 0) {
        // renderIsosurfaces.hpp:33-34, renderValueOnIsosurfaces.hpp:35-36:
        // no levels, the pixel is 0
        viewport[pix] = 0u;
        return;
    }
    const float4 rgba = segment<kMode>(p, brick, tfs, (int)tf_n, iso_opacity, r, lvs, lev.n, sec);
    viewport[pix] = blend_above_u32(viewport[pix], rgba);
}

// Bounding rectangle of non-transparent pixels: per-workgroup reduction, then
// 4 atomics into out = {minX, minY, maxX, maxY}.  A workgroup is 64 columns x
// 4 rows; grid.x covers the width, and the workgroups of a column band step
// through the rows (32-bit indices throughout).
__global__ __launch_bounds__(256) void bound_rect_kernel(const uint32_t* __restrict__ pix,
                                                         uint32_t w, uint32_t h, int* out)
{
    __shared__ int red[4][4];
    int mnx = 0x7fffffff, mny = 0x7fffffff, mxx = -1, mxy = -1;
    const uint32_t x = blockIdx.x * 64u + (threadIdx.x & 63u);
    if (x < w) {
#pragma unroll 4
        for (uint32_t y = blockIdx.y * 4u + (threadIdx.x >> 6); y < h; y += gridDim.y * 4u) {
            if (pix[y * w + x] >> 24) {
                mnx = min(mnx, (int)x);
                mxx = max(mxx, (int)x);
                mny = min(mny, (int)y);
                mxy = max(mxy, (int)y);
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        mnx = min(mnx, __shfl_down(mnx, off));
        mny = min(mny, __shfl_down(mny, off));
        mxx = max(mxx, __shfl_down(mxx, off));
        mxy = max(mxy, __shfl_down(mxy, off));
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][wave] = mnx;
        red[1][wave] = mny;
        red[2][wave] = mxx;
        red[3][wave] = mxy;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int wv = 1; wv < 4; ++wv) {
            red[0][0] = min(red[0][0], red[0][wv]);
            red[1][0] = min(red[1][0], red[1][wv]);
            red[2][0] = max(red[2][0], red[2][wv]);
            red[3][0] = max(red[3][0], red[3][wv]);
        }
        if (red[2][0] >= 0) {
            atomicMin(out + 0, red[0][0]);
            atomicMin(out + 1, red[1][0]);
            atomicMax(out + 2, red[2][0]);
            atomicMax(out + 3, red[3][0]);
        }
    }
}

// The pixels of the rectangle the bound_rect_kernel left in rect[] (rows as
// in the viewport), packed w x h into part: the crop of VsWorker_cu.cpp:
// 279-306 without a host round trip.  part may be pinned host memory; a
// rectangle of more than cap pixels writes nothing.  rect_out (optional, may
// be pinned host memory) receives the rectangle {minX, minY, maxX, maxY}.
__global__ __launch_bounds__(256) void crop_kernel(const uint32_t* __restrict__ pix, uint32_t w, uint32_t h,
                                                   const int* __restrict__ rect, uint32_t* __restrict__ part,
                                                   uint64_t cap, int* __restrict__ rect_out)
{
    const int x0 = rect[0], y0 = rect[1], x1 = rect[2], y1 = rect[3];
    if (rect_out && blockIdx.x == 0u && blockIdx.y == 0u && threadIdx.x < 4u)
        rect_out[threadIdx.x] = rect[threadIdx.x];
    if (x1 < 0)
        return; // empty
    const int pw = x1 - x0 + 1;
    if ((uint64_t)pw * (uint64_t)(y1 - y0 + 1) > cap)
        return;
    const int x = x0 + (int)(blockIdx.x * 64u + (threadIdx.x & 63u));
    if (x > x1)
        return;
    for (int y = y0 + (int)(blockIdx.y * 4u + (threadIdx.x >> 6)); y <= y1; y += (int)(gridDim.y * 4u))
        part[(uint32_t)(y - y0) * (uint32_t)pw + (uint32_t)(x - x0)] = pix[(uint32_t)y * w + (uint32_t)x];
}

// S = sum over the view rectangle of (sampleCount - 1) for rays that hit
// the block: the sample count of renderArgb.hpp:40 without early termination.
__global__ __launch_bounds__(256) void sample_count_kernel(rv_block_params p,
                                                           unsigned long long* out)
{
    uint32_t lx, ly;
    tile_pixel(threadIdx.x, 16u, 16u, lx, ly);
    const uint32_t ox = blockIdx.x * 16u + lx, oy = blockIdx.y * 16u + ly;
    unsigned long long c = 0;
    if (ox < p.view_w && oy < p.view_h) {
        const RaySeg r = setup_ray(p, p.view_x + ox, p.view_y + oy);
        if (r.hit && r.sample_count > 1)
            c = (unsigned long long)(r.sample_count - 1);
    }
    for (int off = 32; off > 0; off >>= 1)
        c += __shfl_down(c, off);
    if ((threadIdx.x & 63u) == 0 && c)
        atomicAdd(out, c);
}

__global__ void init_rect_kernel(int* out)
{
    if (threadIdx.x == 0) {
        out[0] = 0x7fffffff;
        out[1] = 0x7fffffff;
        out[2] = -1;
        out[3] = -1;
    }
}

// blendImagePart (BlendImagePart.hpp:61-145): one thread per source pixel.
__global__ __launch_bounds__(256) void blend_part_kernel(uint32_t* __restrict__ dst, int dw,
                                                         int dx, int dy,
                                                         const uint32_t* __restrict__ src, int sw,
                                                         int sh, int just_copy)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= sw || y >= sh)
        return;
    const uint32_t s = src[(size_t)y * sw + x];
    uint32_t& d = dst[(size_t)dw * (dy + y) + dx + x];
    if (just_copy) {
        d = s;
        return;
    }
    const uint32_t as = s >> 24;
    if (as == 0u)
        return;
    if (as == 255u) {
        d = s;
        return;
    }
    const uint32_t dv = d;
    const uint32_t ad = dv >> 24;
    if (ad == 0u) {
        d = s;
        return;
    }
    const uint32_t a = as + ad * (255u - as) / 255u;
    uint32_t out = a << 24;
    for (uint32_t k = 0; k < 3; ++k) {
        const uint32_t cs = (s >> (8 * k)) & 0xffu, cd = (dv >> (8 * k)) & 0xffu;
        const uint32_t c = (cs * as + cd * (a - as)) / a;
        out |= (c & 255u) << (8 * k);
    }
    d = out;
}

// blendImagePartMaxIntensity (BlendImagePart.hpp:148-191): float pixels,
// NaN = none; the larger value wins.
__global__ __launch_bounds__(256) void blend_part_mip_kernel(float* __restrict__ dst, int dw, int dx,
                                                             int dy, const float* __restrict__ src,
                                                             int sw, int sh, int just_copy)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= sw || y >= sh)
        return;
    const float s = src[(size_t)y * sw + x];
    float& d = dst[(size_t)dw * (dy + y) + dx + x];
    if (just_copy) {
        d = s;
        return;
    }
    if (!(s == s))
        return;
    const float dv = d;
    if (!(dv == dv) || dv < s)
        d = s;
}

// makeMaxIntensityProjection (VsRenderer.cpp:542-565): each non-NaN value
// becomes calcColor of the colour vector (ColorTransferFunction.hpp:83-94),
// encoded ARGB8; NaN pixels keep what dst holds.
__global__ __launch_bounds__(256) void mip_colorize_kernel(uint32_t* __restrict__ dst, int dw, int dx,
                                                           int dy, const float* __restrict__ src,
                                                           int sw, int sh,
                                                           const float4* __restrict__ ctv,
                                                           uint32_t nctv)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= sw || y >= sh)
        return;
    const float s = src[(size_t)y * sw + x];
    if (!(s == s))
        return;
    float4 c;
    if (s <= 0.f || nctv == 1u) {
        c = ctv[0];
    } else if (s >= 1.f) {
        c = ctv[nctv - 1u];
    } else {
        const float par = s * (float)(nctv - 1u);
        const uint32_t idx = (uint32_t)par;
        const float t = par - (float)idx;
        const float4 a = ctv[idx], b = ctv[idx + 1u];
        c = make_float4(a.x * (1.f - t) + b.x * t, a.y * (1.f - t) + b.y * t,
                        a.z * (1.f - t) + b.z * t, a.w * (1.f - t) + b.w * t);
    }
    dst[(size_t)dw * (dy + y) + dx + x] =
        ((unsigned)(c.w * 255.999f) << 24) + ((unsigned)(c.x * 255.999f) << 16) +
        ((unsigned)(c.y * 255.999f) << 8) + (unsigned)(c.z * 255.999f);
}

} // namespace
} // namespace rv

using namespace rv;

// The entry points of the field modes share their argument checks.
static int check_block(const rv_block_params* params, const float* d_brick, const float* d_tf,
                       uint32_t tf_n, const uint32_t* d_viewport, const char* who)
{
    if (!(params && d_brick && d_tf && d_viewport))
        return fail(RV_ERR_INVALID, std::string(who) + ": null pointer");
    if (!(tf_n >= 1 && tf_n <= kTfMax))
        return fail(RV_ERR_INVALID, std::string(who) + ": tf_n must be in [1, 4096]");
    const rv_block_params& p = *params;
    if (!(p.tex_depth <= 9 && p.nodes == (1u << p.tex_depth) + 1u))
        return fail(RV_ERR_INVALID,
                    std::string(who) + ": inconsistent tex_depth/nodes (use rv_block_params_setup)");
    if (!(p.fmax > p.fmin))
        return fail(RV_ERR_INVALID, std::string(who) + ": field range must be non-empty");
    if (!((uint64_t)p.view_x + p.view_w <= p.vp_w && (uint64_t)p.view_y + p.view_h <= p.vp_h))
        return fail(RV_ERR_INVALID, std::string(who) + ": view rectangle outside the viewport");
    return RV_OK;
}

extern "C" int rv_render_block_secondary(const rv_block_params* params, uint32_t field_mode,
                                         float iso_opacity, const float* levels_rel,
                                         uint32_t level_count, const float* d_brick,
                                         const float* d_brick2, float smin, float smax,
                                         const float* d_tf, uint32_t tf_n, uint32_t* d_viewport,
                                         rv_stream_t stream)
{
    int rc = check_block(params, d_brick, d_tf, tf_n, d_viewport, "rv_render_block");
    if (rc)
        return rc;
    RV_REQUIRE(levels_rel || level_count == 0, "rv_render_block: null levels");
    const bool value = field_mode == RV_FIELD_VALUE_ON_ISOSURFACE ||
                       field_mode == RV_FIELD_VALUE_ON_ISOSURFACES;
    RV_REQUIRE(!value || d_brick2, "rv_render_block: the ValueOnIsosurface(s) modes need the "
                                   "secondary field's brick");
    RV_REQUIRE(!value || smax > smin, "rv_render_block: the secondary field range must be non-empty");
    const rv_block_params& p = *params;
    if (p.view_w == 0 || p.view_h == 0)
        return RV_OK;
    const LevelsArg lev = make_levels(p.fmin, p.fmax, levels_rel, level_count);
    const Secondary<float> sec{value ? d_brick2 : nullptr, smin, smax};
    dim3 grid((p.view_w + 15) / 16, (p.view_h + 15) / 16);
    const size_t lds = (tf_n + 1u) * sizeof(float4);
#define RV_BLOCK_LAUNCH(M)                                                                         \
    hipLaunchKernelGGL(render_block_kernel<M>, grid, dim3(256), lds, S(stream), p, d_brick, sec, d_tf, \
                       tf_n, iso_opacity, lev, d_viewport)
    switch (field_mode) {
    case RV_FIELD_ARGB:
        if (p.flags & RV_RENDER_EXACT)
            RV_BLOCK_LAUNCH(kModeArgb);
        else
            RV_BLOCK_LAUNCH(kModeArgbFast);
        break;
    case RV_FIELD_ARGB_LIGHT: RV_BLOCK_LAUNCH(kModeArgbLight); break;
    case RV_FIELD_ISOSURFACE: RV_BLOCK_LAUNCH(kModeIsosurface); break;
    case RV_FIELD_ISOSURFACES: RV_BLOCK_LAUNCH(kModeIsosurfaces); break;
    case RV_FIELD_DOMAIN_VOXELS: RV_BLOCK_LAUNCH(kModeDomainVoxels); break;
    case RV_FIELD_MAX_INTENSITY_PROJECTION: RV_BLOCK_LAUNCH(kModeMip); break;
    case RV_FIELD_VALUE_ON_ISOSURFACE: RV_BLOCK_LAUNCH(kModeValueOnIsosurface); break;
    case RV_FIELD_VALUE_ON_ISOSURFACES: RV_BLOCK_LAUNCH(kModeValueOnIsosurfaces); break;
    default: return fail(RV_ERR_INVALID, "rv_render_block: unknown field mode");
    }
#undef RV_BLOCK_LAUNCH
    RV_LAUNCH_CHECK();
    return RV_OK;
}

extern "C" int rv_render_block(const rv_block_params* params, uint32_t field_mode, float iso_opacity,
                               const float* levels_rel, uint32_t level_count, const float* d_brick,
                               const float* d_tf, uint32_t tf_n, uint32_t* d_viewport,
                               rv_stream_t stream)
{
    return rv_render_block_secondary(params, field_mode, iso_opacity, levels_rel, level_count, d_brick,
                                     nullptr, 0.f, 1.f, d_tf, tf_n, d_viewport, stream);
}

extern "C" int rv_clear_viewport_mode(uint32_t* d_viewport, uint32_t vp_w, uint32_t vp_h,
                                      uint32_t field_mode, rv_stream_t stream)
{
    RV_REQUIRE(d_viewport || (uint64_t)vp_w * vp_h == 0, "rv_clear_viewport_mode: null viewport");
    // DenseFieldRenderer_cu.cu:268-277: NaN bits for max intensity projection
    const uint32_t v = field_mode == RV_FIELD_MAX_INTENSITY_PROJECTION ? 0x7fc00000u : 0u;
    return fill_u32(d_viewport, (uint64_t)vp_w * vp_h, v, S(stream));
}

extern "C" int rv_blend_image_part_max_intensity(float* d_dst, int32_t dw, int32_t dh,
                                                 int32_t dst_off_x, int32_t dst_off_y,
                                                 const float* d_src, int32_t sw, int32_t sh,
                                                 int32_t src_x, int32_t src_y, int just_copy,
                                                 rv_stream_t stream)
{
    if (sw <= 0 || sh <= 0)
        return RV_OK;
    RV_REQUIRE(d_dst && d_src, "rv_blend_image_part_max_intensity: null pointer");
    const int dx = src_x - dst_off_x, dy = src_y - dst_off_y;
    RV_REQUIRE(dx >= 0 && dy >= 0 && dx + sw <= dw && dy + sh <= dh,
               "rv_blend_image_part_max_intensity: part outside destination");
    dim3 grid((sw + 63) / 64, (sh + 3) / 4);
    hipLaunchKernelGGL(blend_part_mip_kernel, grid, dim3(256), 0, S(stream), d_dst, dw, dx, dy, d_src,
                       sw, sh, just_copy);
    RV_LAUNCH_CHECK();
    return RV_OK;
}

extern "C" int rv_mip_colorize(uint32_t* d_dst, int32_t dw, int32_t dh, const float* d_src,
                               int32_t sw, int32_t sh, int32_t src_x, int32_t src_y,
                               const float* d_ctv, uint32_t nctv, rv_stream_t stream)
{
    if (sw <= 0 || sh <= 0)
        return RV_OK;
    RV_REQUIRE(d_dst && d_src && d_ctv && nctv >= 1, "rv_mip_colorize: null pointer or empty vector");
    RV_REQUIRE(src_x >= 0 && src_y >= 0 && src_x + sw <= dw && src_y + sh <= dh,
               "rv_mip_colorize: part outside destination");
    dim3 grid((sw + 63) / 64, (sh + 3) / 4);
    hipLaunchKernelGGL(mip_colorize_kernel, grid, dim3(256), 0, S(stream), d_dst, dw, src_x, src_y,
                       d_src, sw, sh, reinterpret_cast<const float4*>(d_ctv), nctv);
    RV_LAUNCH_CHECK();
    return RV_OK;
}

extern "C" int rv_clear_viewport(uint32_t* d_viewport, uint32_t vp_w, uint32_t vp_h,
                                 rv_stream_t stream)
{
    RV_REQUIRE(d_viewport || (uint64_t)vp_w * vp_h == 0, "rv_clear_viewport: null viewport");
    return fill_u32(d_viewport, (uint64_t)vp_w * vp_h, 0u, S(stream));
}

extern "C" int rv_fill_u32(uint32_t* d_dst, uint64_t n, uint32_t value, rv_stream_t stream)
{
    RV_REQUIRE(d_dst || n == 0, "rv_fill_u32: null pointer");
    return fill_u32(d_dst, n, value, S(stream));
}

extern "C" int rv_render_block_argb(const rv_block_params* params, const float* d_brick,
                                    const float* d_tf, uint32_t tf_n, uint32_t* d_viewport,
                                    rv_stream_t stream)
{
    RV_REQUIRE(params && d_brick && d_tf && d_viewport, "rv_render_block_argb: null pointer");
    RV_REQUIRE(tf_n >= 1 && tf_n <= kTfMax, "rv_render_block_argb: tf_n must be in [1, 4096]");
    const rv_block_params& p = *params;
    RV_REQUIRE(p.tex_depth <= 9 && p.nodes == (1u << p.tex_depth) + 1u,
               "rv_render_block_argb: inconsistent tex_depth/nodes (use rv_block_params_setup)");
    RV_REQUIRE(p.fmax > p.fmin, "rv_render_block_argb: field range must be non-empty");
    RV_REQUIRE((uint64_t)p.view_x + p.view_w <= p.vp_w && (uint64_t)p.view_y + p.view_h <= p.vp_h,
               "rv_render_block_argb: view rectangle outside the viewport");
    if (p.view_w == 0 || p.view_h == 0)
        return RV_OK;
    dim3 grid((p.view_w + 15) / 16, (p.view_h + 15) / 16);
    if (p.flags & RV_RENDER_EXACT)
        hipLaunchKernelGGL(render_block_kernel<kModeArgb>, grid, dim3(256), (tf_n + 1u) * sizeof(float4),
                           S(stream), p, d_brick, Secondary<float>{nullptr, 0.f, 1.f}, d_tf, tf_n, 1.f,
                           LevelsArg{}, d_viewport);
    else
        hipLaunchKernelGGL(render_block_kernel<kModeArgbFast>, grid, dim3(256), (tf_n + 1u) * sizeof(float4),
                           S(stream), p, d_brick, Secondary<float>{nullptr, 0.f, 1.f}, d_tf, tf_n, 1.f,
                           LevelsArg{}, d_viewport);
    RV_LAUNCH_CHECK();
    return RV_OK;
}

extern "C" int rv_block_sample_count(const rv_block_params* params, uint64_t* d_count,
                                     rv_stream_t stream)
{
    RV_REQUIRE(params && d_count, "rv_block_sample_count: null pointer");
    const rv_block_params& p = *params;
    if (p.view_w == 0 || p.view_h == 0)
        return RV_OK;
    dim3 grid((p.view_w + 15) / 16, (p.view_h + 15) / 16);
    hipLaunchKernelGGL(sample_count_kernel, grid, dim3(256), 0, S(stream), p,
                       reinterpret_cast<unsigned long long*>(d_count));
    RV_LAUNCH_CHECK();
    return RV_OK;
}

// The crop rectangle into a caller's 16-byte device buffer, in stream
// order: {minX, minY, maxX, maxY}, maxX < 0 when every pixel is transparent.
int rv::bound_rect_launch(const uint32_t* d_pixels, uint32_t w, uint32_t h, int* d_buf, hipStream_t s)
{
    hipLaunchKernelGGL(init_rect_kernel, dim3(1), dim3(64), 0, s, d_buf);
    const uint32_t gx = (w + 63u) / 64u;
    // few workgroups, many rows each: the 4 atomics per workgroup meet on
    // one address each, so their count, not the pixel reads, bounds the
    // kernel when workgroups are many
    uint32_t gy = (h + 3u) / 4u;
    if (gy > 16u)
        gy = 16u;
    if (gx > 0u && gy > 0u)
        hipLaunchKernelGGL(bound_rect_kernel, dim3(gx, gy), dim3(256), 0, s, d_pixels, w, h, d_buf);
    RV_LAUNCH_CHECK();
    return RV_OK;
}

// The rectangle's pixels packed into d_part (room for cap pixels; device
// memory or pinned host memory), in stream order after bound_rect_launch;
// rect_out (or null) receives the rectangle.
int rv::crop_launch(const uint32_t* d_pixels, uint32_t w, uint32_t h, const int* d_buf, uint32_t* d_part,
                    uint64_t cap, int* rect_out, hipStream_t s)
{
    const uint32_t gx = (w + 63u) / 64u;
    uint32_t gy = (h + 3u) / 4u;
    if (gy > 256u)
        gy = 256u;
    if (gx > 0u && gy > 0u)
        hipLaunchKernelGGL(crop_kernel, dim3(gx, gy), dim3(256), 0, s, d_pixels, w, h, d_buf, d_part, cap,
                           rect_out);
    RV_LAUNCH_CHECK();
    return RV_OK;
}

// Reads the rectangle of bound_rect_launch (synchronous) as {x, y, w, h}.
int rv::bound_rect_read(const int* d_buf, int32_t rect[4], int* nonempty, hipStream_t s)
{
    *nonempty = 0;
    rect[0] = rect[1] = rect[2] = rect[3] = 0;
    int host[4];
    hipError_t e = hipMemcpyAsync(host, d_buf, 16, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return hip_fail(e, "rv_bound_rect", __FILE__, __LINE__);
    if (host[2] < 0)
        return RV_OK;
    *nonempty = 1;
    rect[0] = host[0];
    rect[1] = host[1];
    rect[2] = host[2] - host[0] + 1;
    rect[3] = host[3] - host[1] + 1;
    return RV_OK;
}

extern "C" int rv_bound_rect(const uint32_t* d_pixels, uint32_t w, uint32_t h, int32_t rect[4],
                             int* nonempty, rv_stream_t stream)
{
    RV_REQUIRE(d_pixels && rect && nonempty, "rv_bound_rect: null pointer");
    *nonempty = 0;
    rect[0] = rect[1] = rect[2] = rect[3] = 0;
    if ((uint64_t)w * h == 0)
        return RV_OK;
    int* d_out = nullptr;
    RV_HIP(hipMallocAsync((void**)&d_out, 16, S(stream)));
    int rc = bound_rect_launch(d_pixels, w, h, d_out, S(stream));
    if (!rc)
        rc = bound_rect_read(d_out, rect, nonempty, S(stream));
    const hipError_t e = hipFreeAsync(d_out, S(stream));
    if (rc)
        return rc;
    if (e != hipSuccess)
        return hip_fail(e, "hipFreeAsync", __FILE__, __LINE__);
    return RV_OK;
}

extern "C" int rv_blend_image_part(uint32_t* d_dst, int32_t dw, int32_t dh, int32_t dst_off_x,
                                   int32_t dst_off_y, const uint32_t* d_src, int32_t sw,
                                   int32_t sh, int32_t src_x, int32_t src_y, int just_copy,
                                   rv_stream_t stream)
{
    if (sw <= 0 || sh <= 0)
        return RV_OK;
    RV_REQUIRE(d_dst && d_src, "rv_blend_image_part: null pointer");
    const int dx = src_x - dst_off_x, dy = src_y - dst_off_y;
    RV_REQUIRE(dx >= 0 && dy >= 0 && dx + sw <= dw && dy + sh <= dh,
               "rv_blend_image_part: part outside destination");
    dim3 grid((sw + 63) / 64, (sh + 3) / 4);
    hipLaunchKernelGGL(blend_part_kernel, grid, dim3(256), 0, S(stream), d_dst, dw, dx, dy, d_src,
                       sw, sh, just_copy);
    RV_LAUNCH_CHECK();
    return RV_OK;
}
