// Adjoint of the batched frame ray-cast: gradients of an image loss with
// respect to the colour table (tf_n x 4) and the dense field values of every
// brick, by per-ray recomputation (no samples are stored).
//
// Forward model (frame.hip, render_tile): per pixel, the blocks covering it
// are visited front to back; each block contributes the segment of its
// samples s (renderArgb.hpp:39-51): c_s = classify(trilinear(brick, x_s)),
// composited front to back with blendBehind (cuda_blend.hpp:24-33), and the
// segments are composited front to back into the pixel.  blendBehind on
// non-premultiplied colours is front-to-back compositing of premultiplied
// ones, so with a_s = c_s.w and T_s = prod_{s' < s} (1 - a_s') over all kept
// samples of the pixel,
//     P = sum_s T_s a_s c_s.rgb,  A = 1 - T_end,  image = (P / A, A).
// Given dL/dimage per pixel, dL/dP = g.rgb / A and dL/dA = g.w - g.rgb.I.rgb / A,
// and for a kept sample
//     dL/dc_s.rgb = T_s a_s dL/dP,
//     dL/da_s     = T_s (dL/dP . c_s.rgb - dL/dP . Q_s + dL/dA . T_end / T_{s+1}),
// where Q_s = (P - P_{<=s}) / T_{s+1} is what lies behind s, seen from just
// behind it.  One front-to-back re-march per pixel evaluates this, with P and
// T_end taken from the forward image.  Early termination (segment alpha above
// alpha_cutoff, pixel alpha above acc_cutoff) is recomputed with the
// forward's exact fp32 operations, so the adjoint differentiates the same set
// of samples (the terminations are piecewise-constant decisions).
//
// Through classify (renderArgb.hpp:42-45): a = 1 - (1 - t.w of)^ilc with t the
// linearly interpolated table entry, so da/dt.w = ilc of (1 - a) / (1 - t.w of);
// t = (1 - f) tf[i0] + f tf[i1] at x = v01 N - 0.5, dt/dv = (tf[i1] - tf[i0]) N
// / (fmax - fmin) inside the table's range and 0 where x is clamped; v is
// trilinear in the 8 brick corners.  Colour-table gradients accumulate in LDS
// per workgroup and are added to the output once per workgroup; field
// gradients are added to the per-block gradient arrays with float atomics,
// merged in registers while consecutive samples share a cell.

#include "frame_plan.hpp"

#include <hip/hip_fp16.h>

namespace rv {
namespace {

struct AdjArgs {
    const rv_block_params* params;
    const uint4* rects;
    const uint32_t* f2b;
    const void* const* bricks;
    float* const* grad_bricks; // NULL: no field gradient
    uint32_t nblocks;
    const float4* tf;
    uint32_t tf_n;
    uint32_t W, H, tw, th, ntx, nty;
    float acc_cutoff;
    const float4* rgba;        // forward image, rows bottom-up
    const float4* grad_rgba;   // dL/dimage, same layout
    float4* grad_tf;           // tf_n entries, accumulated
};

// Per-pixel state of the re-march.
struct AdjPixel {
    float acc_w;           // pixel alpha, with the forward's exact operations
    float T;               // transmittance in front of the current block
    float Px, Py, Pz;      // premultiplied colour of the samples so far
    float Fx, Fy, Fz, Tend;
    float gPx, gPy, gPz, gA;
};

// Field-gradient accumulator of the current cell (8 corners).
struct CellAcc {
    uint32_t key;  // corner-000 index * 8 + clamp pattern, ~0 = empty
    uint32_t a, dx, dy, dz;
    float g[8];
};

__device__ __forceinline__ void cell_flush(CellAcc& c, float* __restrict__ gb)
{
    if (c.key == 0xFFFFFFFFu)
        return;
    const uint32_t off[8] = {0u, c.dx, c.dy, c.dy + c.dx, c.dz, c.dz + c.dx, c.dz + c.dy,
                             c.dz + c.dy + c.dx};
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (c.g[k] != 0.f)
            unsafeAtomicAdd(gb + c.a + off[k], c.g[k]);
    c.key = 0xFFFFFFFFu;
}

template <class T>
__device__ void adjoint_segment(const rv_block_params& p, const T* __restrict__ brick,
                                float* __restrict__ gbrick, const float4* tfs, float4* gtf,
                                int tf_n, RaySeg r, AdjPixel& st)
{
    const int n = (int)p.nodes;
    const float span = p.fmax - p.fmin;
    float seg_w = 0.f, Tseg = 1.f;
    CellAcc cell;
    cell.key = 0xFFFFFFFFu;
    // Samples are fetched two at a time (positions by the forward's
    // sequential additions), then processed in order.
    constexpr int K = 2;
    bool ended = false;
    for (int i0s = 1; i0s < r.sample_count && !ended; i0s += K) {
        uint32_t ka[K], kdx[K], kdy[K], kdz[K];
        float kfx[K], kfy[K], kfz[K], kv[K][8];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            ka[k] = 0u;
            kdx[k] = kdy[k] = kdz[k] = 0u;
            kfx[k] = kfy[k] = kfz[k] = 0.f;
#pragma unroll
            for (int c = 0; c < 8; ++c)
                kv[k][c] = __builtin_nanf("");
            if (i0s + k < r.sample_count) {
                const float tx = (r.px - p.tex_origin[0]) * p.tcf;
                const float ty = (r.py - p.tex_origin[1]) * p.tcf;
                const float tz = (r.pz - p.tex_origin[2]) * p.tcf;
                int x0, x1, y0, y1, z0, z1;
                axis_taps(tx, n, x0, x1, kfx[k]);
                axis_taps(ty, n, y0, y1, kfy[k]);
                axis_taps(tz, n, z0, z1, kfz[k]);
                const uint32_t a = ((uint32_t)z0 * n + y0) * n + x0;
                const uint32_t dx = (uint32_t)(x1 - x0);
                const uint32_t dy = (uint32_t)(y1 - y0) * n;
                const uint32_t dz = (uint32_t)(z1 - z0) * n * n;
                ka[k] = a;
                kdx[k] = dx;
                kdy[k] = dy;
                kdz[k] = dz;
                kv[k][0] = (float)brick[a];
                kv[k][1] = (float)brick[a + dx];
                kv[k][2] = (float)brick[a + dy];
                kv[k][3] = (float)brick[a + dy + dx];
                kv[k][4] = (float)brick[a + dz];
                kv[k][5] = (float)brick[a + dz + dx];
                kv[k][6] = (float)brick[a + dz + dy];
                kv[k][7] = (float)brick[a + dz + dy + dx];
                r.px = r.px + r.sx;
                r.py = r.py + r.sy;
                r.pz = r.pz + r.sz;
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (ended)
                break;
            const float fx = kfx[k], fy = kfy[k], fz = kfz[k];
            // the trilinear operations of fetch3
            const float c00 = lerpf(kv[k][0], kv[k][1], fx);
            const float c10 = lerpf(kv[k][2], kv[k][3], fx);
            const float c01 = lerpf(kv[k][4], kv[k][5], fx);
            const float c11 = lerpf(kv[k][6], kv[k][7], fx);
            const float c0 = lerpf(c00, c10, fy);
            const float c1 = lerpf(c01, c11, fy);
            const float v = lerpf(c0, c1, fz);
            if (!(v == v))
                continue; // NaN sample, or past the segment's end
            const uint32_t a = ka[k], dx = kdx[k], dy = kdy[k], dz = kdz[k];
            // classify, the operations of classify / fetch_tf
            const float v01 = (v - p.fmin) / span;
            const float xraw = v01 * (float)tf_n - 0.5f;
            float x = xraw;
            if (!(x > -1.f))
                x = -1.f;
            if (x > (float)tf_n)
                x = (float)tf_n;
            const float fl = floorf(x);
            const float f = x - fl;
            const int ti = (int)fl;
            const int i0 = clampi(ti, 0, tf_n - 1), i1 = clampi(ti + 1, 0, tf_n - 1);
            const float4 A0 = tfs[i0], A1 = tfs[i1];
            const float cx = lerpf(A0.x, A1.x, f), cy = lerpf(A0.y, A1.y, f),
                        cz = lerpf(A0.z, A1.z, f), tw = lerpf(A0.w, A1.w, f);
            const float base = 1.f - tw * p.opacity_factor;
            const float alpha = 1.f - pow01(base, p.i_layer_count);

            const float Ts = st.T * Tseg;
            const float wgt = Ts * alpha;
            st.Px = st.Px + wgt * cx;
            st.Py = st.Py + wgt * cy;
            st.Pz = st.Pz + wgt * cz;
            const float Tn = Ts * (1.f - alpha);
            const bool behind = Tn > 1e-30f;
            const float inv = behind ? 1.f / Tn : 0.f;
            const float Qx = (st.Fx - st.Px) * inv, Qy = (st.Fy - st.Py) * inv,
                        Qz = (st.Fz - st.Pz) * inv;
            const float ratio = behind ? st.Tend * inv : 1.f;
            const float g_alpha =
                Ts * (((st.gPx * cx + st.gPy * cy) + st.gPz * cz) -
                      ((st.gPx * Qx + st.gPy * Qy) + st.gPz * Qz) + st.gA * ratio);
            const float gcx = st.gPx * wgt, gcy = st.gPy * wgt, gcz = st.gPz * wgt;
            const float gtw = base > 0.f
                                  ? g_alpha * (p.i_layer_count * p.opacity_factor * (1.f - alpha) / base)
                                  : 0.f;
            // colour table
            const float w0 = 1.f - f;
            atomicAdd(&gtf[i0].x, w0 * gcx);
            atomicAdd(&gtf[i0].y, w0 * gcy);
            atomicAdd(&gtf[i0].z, w0 * gcz);
            atomicAdd(&gtf[i0].w, w0 * gtw);
            atomicAdd(&gtf[i1].x, f * gcx);
            atomicAdd(&gtf[i1].y, f * gcy);
            atomicAdd(&gtf[i1].z, f * gcz);
            atomicAdd(&gtf[i1].w, f * gtw);
            // field
            if (gbrick && xraw > -1.f && xraw <= (float)tf_n) {
                const float dxdv = (float)tf_n / span;
                const float gv = dxdv * ((((A1.x - A0.x) * gcx + (A1.y - A0.y) * gcy) +
                                          (A1.z - A0.z) * gcz) + (A1.w - A0.w) * gtw);
                if (gv != 0.f) {
                    const uint32_t key = a * 8u + (dx ? 1u : 0u) + (dy ? 2u : 0u) + (dz ? 4u : 0u);
                    if (key != cell.key) {
                        cell_flush(cell, gbrick);
                        cell.key = key;
                        cell.a = a;
                        cell.dx = dx;
                        cell.dy = dy;
                        cell.dz = dz;
#pragma unroll
                        for (int c = 0; c < 8; ++c)
                            cell.g[c] = 0.f;
                    }
                    const float ux = 1.f - fx, uy = 1.f - fy, uz = 1.f - fz;
                    const float gz0 = gv * uz, gz1 = gv * fz;
                    const float g00 = gz0 * uy, g10 = gz0 * fy, g01 = gz1 * uy, g11 = gz1 * fy;
                    cell.g[0] += g00 * ux;
                    cell.g[1] += g00 * fx;
                    cell.g[2] += g10 * ux;
                    cell.g[3] += g10 * fx;
                    cell.g[4] += g01 * ux;
                    cell.g[5] += g01 * fx;
                    cell.g[6] += g11 * ux;
                    cell.g[7] += g11 * fx;
                }
            }
            Tseg = Tseg * (1.f - alpha);
            // the forward's segment alpha (blend_behind), for its termination
            const float da = alpha * (1.f - seg_w);
            seg_w = seg_w + da;
            if (seg_w > p.alpha_cutoff)
                ended = true;
        }
    }
    if (gbrick)
        cell_flush(cell, gbrick);
    const float da = seg_w * (1.f - st.acc_w);
    st.acc_w = st.acc_w + da;
    st.T = st.T * Tseg;
}

template <class BrickT>
__global__ __launch_bounds__(kT) void adjoint_kernel(AdjArgs a)
{
    __shared__ float4 tfs[kMaxTf];
    __shared__ float4 gtf[kMaxTf];
    __shared__ uint32_t list[kT];
    __shared__ uint32_t wave_cnt[kT / 64u];
    for (uint32_t i = threadIdx.x; i < a.tf_n; i += kT) {
        tfs[i] = a.tf[i];
        gtf[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();
    const uint32_t ntiles = a.ntx * a.nty;
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint32_t tid = threadIdx.x;
        const uint32_t tx = tile % a.ntx, ty = tile / a.ntx;
        const uint32_t x0 = tx * a.tw, y0 = ty * a.th;
        const uint32_t x1 = min(x0 + a.tw, a.W), y1 = min(y0 + a.th, a.H);
        const uint32_t vx = x0 + tid % a.tw, vy = y0 + tid / a.tw;
        const bool inside = vx < a.W && vy < a.H;
        AdjPixel st{};
        bool done = true;
        if (inside) {
            const uint32_t pix = (a.H - 1u - vy) * a.W + vx;
            const float4 I = a.rgba[pix];
            const float4 g = a.grad_rgba[pix];
            st.acc_w = 0.f;
            st.T = 1.f;
            st.Fx = I.x * I.w;
            st.Fy = I.y * I.w;
            st.Fz = I.z * I.w;
            st.Tend = 1.f - I.w;
            if (I.w > 0.f) {
                st.gPx = g.x / I.w;
                st.gPy = g.y / I.w;
                st.gPz = g.z / I.w;
                st.gA = g.w - ((g.x * I.x + g.y * I.y) + g.z * I.z) / I.w;
            } else {
                st.gA = g.w;
            }
            done = st.gPx == 0.f && st.gPy == 0.f && st.gPz == 0.f && st.gA == 0.f;
        }
        for_tile_blocks(a.rects, a.f2b, a.nblocks, x0, y0, x1, y1, list, wave_cnt,
                        [&](uint32_t blk) {
                            const rv_block_params p = a.params[blk];
                            if (!done && vx >= p.view_x && vx < p.view_x + p.view_w &&
                                vy >= p.view_y && vy < p.view_y + p.view_h) {
                                const RaySeg r = setup_ray(p, vx, vy);
                                if (r.hit) {
                                    adjoint_segment(p, reinterpret_cast<const BrickT*>(a.bricks[blk]),
                                                    a.grad_bricks ? a.grad_bricks[blk] : nullptr,
                                                    tfs, gtf, (int)a.tf_n, r, st);
                                    done = st.acc_w > a.acc_cutoff;
                                }
                            }
                            return done;
                        });
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < a.tf_n; i += kT) {
        const float4 g = gtf[i];
        float* out = reinterpret_cast<float*>(a.grad_tf + i);
        if (g.x != 0.f)
            unsafeAtomicAdd(out + 0, g.x);
        if (g.y != 0.f)
            unsafeAtomicAdd(out + 1, g.y);
        if (g.z != 0.f)
            unsafeAtomicAdd(out + 2, g.z);
        if (g.w != 0.f)
            unsafeAtomicAdd(out + 3, g.w);
    }
}

} // namespace
} // namespace rv

using namespace rv;

extern "C" int rv_frame_set_grad_bricks(rv_frame_plan* P, float* const* d_grad_bricks)
{
    RV_TRY
    RV_REQUIRE(P, "rv_frame_set_grad_bricks: null plan");
    if (!d_grad_bricks) {
        if (P->d_grad_bricks)
            RV_HIP(hipFree(P->d_grad_bricks));
        P->d_grad_bricks = nullptr;
        return RV_OK;
    }
    for (uint32_t i = 0; i < P->nblocks; ++i)
        RV_REQUIRE(d_grad_bricks[i], "rv_frame_set_grad_bricks: null gradient array");
    return frame_upload((void**)&P->d_grad_bricks, d_grad_bricks, sizeof(float*) * P->nblocks);
    RV_CATCH
}

extern "C" int rv_frame_adjoint(rv_frame_plan* P, uint32_t cam, const float* d_rgba,
                                const float* d_grad_rgba, float* d_grad_tf, rv_stream_t stream)
{
    RV_TRY
    RV_REQUIRE(P && d_rgba && d_grad_rgba && d_grad_tf, "rv_frame_adjoint: null pointer");
    RV_REQUIRE(cam < P->ncams, "rv_frame_adjoint: camera index out of range (set cameras first)");
    if (!P->adj_grid) {
        int dev = 0, cus = 0, per_cu = 0;
        RV_HIP(hipGetDevice(&dev));
        RV_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        if (P->fp16)
            RV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, adjoint_kernel<__half>, kT, 0));
        else
            RV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, adjoint_kernel<float>, kT, 0));
        RV_REQUIRE(per_cu >= 1, "rv_frame_adjoint: kernel cannot be resident");
        P->adj_grid = (uint32_t)(cus * per_cu);
    }
    AdjArgs a{};
    a.params = P->d_params + (size_t)cam * P->nblocks;
    a.rects = P->d_rects + (size_t)cam * P->nblocks;
    a.f2b = P->d_f2b + (size_t)cam * P->nblocks;
    a.bricks = P->d_bricks;
    a.grad_bricks = P->d_grad_bricks;
    a.nblocks = P->nblocks;
    a.tf = reinterpret_cast<const float4*>(P->desc.d_tf);
    a.tf_n = P->desc.tf_n;
    a.W = P->desc.vp_w;
    a.H = P->desc.vp_h;
    a.tw = P->tw;
    a.th = P->th;
    a.ntx = (a.W + a.tw - 1) / a.tw;
    a.nty = (a.H + a.th - 1) / a.th;
    a.acc_cutoff = P->desc.acc_cutoff;
    a.rgba = reinterpret_cast<const float4*>(d_rgba);
    a.grad_rgba = reinterpret_cast<const float4*>(d_grad_rgba);
    a.grad_tf = reinterpret_cast<float4*>(d_grad_tf);
    const uint32_t grid = std::min(P->adj_grid, a.ntx * a.nty);
    if (P->fp16)
        hipLaunchKernelGGL(adjoint_kernel<__half>, dim3(grid), dim3(kT), 0, S(stream), a);
    else
        hipLaunchKernelGGL(adjoint_kernel<float>, dim3(grid), dim3(kT), 0, S(stream), a);
    RV_LAUNCH_CHECK();
    return RV_OK;
    RV_CATCH
}
