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
// trilinear in the 8 brick corners.
//
// Order of work.  The forward records, for every pixel and every block its
// ray reaches, the pixel's state in front of the block (premultiplied colour
// and transmittance, 16 bytes; rv_frame_set_tape).  The adjoint then works
// block by block: one wave takes an 8 x 8 tile of one block's view rectangle
// and re-marches only that block's segments, starting from the recorded
// state, so all rays of a wave are in one block at a time, and waves never
// wait for each other.  Colour-table terms are summed in registers while a
// ray's samples interpolate the same two table entries, then into an LDS copy
// of the table shared by the workgroup, added to the output once per
// workgroup.  Field terms are summed per cell in registers (keeping the 4
// corners shared with the next cell when the ray steps to a face neighbour),
// then into the wave's LDS window over one slab of the block, which is added
// to memory with coalesced atomics (adjoint_item).

#include "frame_plan.hpp"

#include <hip/hip_fp16.h>

namespace rv {
namespace {

constexpr int kSlab = 8;                // cells per slab along the slab axis (at most)
constexpr uint32_t kWinWave = 4096;     // floats of a wave's LDS gradient window
constexpr uint32_t kAdjWaves = kRT / 64u;

__device__ __forceinline__ int wave_min(int v)
{
    for (int o = 32; o > 0; o >>= 1)
        v = min(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ int wave_max(int v)
{
    for (int o = 32; o > 0; o >>= 1)
        v = max(v, __shfl_xor(v, o));
    return v;
}

// Orders the LDS accesses of one wave's lanes (a wave's LDS operations
// execute in program order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

struct AdjArgs {
    const rv_block_params* params;
    const void* const* bricks;
    float* const* grad_bricks; // NULL: no field gradient
    const uint32_t* items;     // (block, tile) pairs of this camera
    uint32_t nitems;
    const uint64_t* rec_off;   // per block: its rectangle's offset in the tape
    const float4* tape;        // per (block, pixel): state in front of the block, T < 0: none
    const float4* tf;
    uint32_t tf_n;
    uint32_t W, H;
    const float4* rgba;        // forward image, rows bottom-up
    const float4* grad_rgba;   // dL/dimage, same layout
    float4* grad_tf;           // tf_n entries, accumulated
    const uint32_t* rec_chunk; // per 64 records: the block of the first (colour-table-only kernel)
    uint64_t nrec;             // records of this camera
    uint32_t nblocks;
    uint32_t* queue;           // next work item (zeroed before the launch)
};

// Per-pixel state of the re-march.
struct AdjPixel {
    float T;               // transmittance in front of the current sample's block
    float Px, Py, Pz;      // premultiplied colour of the samples so far
    float Fx, Fy, Fz, Tend;
    float gPx, gPy, gPz, gA;
};

// Colour-table gradient accumulator of the current pair of table entries.
struct TfAcc {
    int i0, i1;  // i0 < 0: empty
    float4 g0, g1;
};

__device__ __forceinline__ void tf_flush(TfAcc& t, float4* gtf)
{
    if (t.i0 < 0)
        return;
    atomicAdd(&gtf[t.i0].x, t.g0.x);
    atomicAdd(&gtf[t.i0].y, t.g0.y);
    atomicAdd(&gtf[t.i0].z, t.g0.z);
    atomicAdd(&gtf[t.i0].w, t.g0.w);
    atomicAdd(&gtf[t.i1].x, t.g1.x);
    atomicAdd(&gtf[t.i1].y, t.g1.y);
    atomicAdd(&gtf[t.i1].z, t.g1.z);
    atomicAdd(&gtf[t.i1].w, t.g1.w);
    t.i0 = -1;
}

// The LDS gradient window of the current slab: node ranges along x, y, z
// (origin and extent), uniform over the workgroup; ok = false: no window.
struct Window {
    int o[3], e[3];
    bool ok;
};

// Field-gradient accumulator of the current cell (8 corners).
struct CellAcc {
    uint32_t key;  // corner-000 index * 8 + clamp pattern, ~0 = empty
    uint32_t a, dx, dy, dz;
    int x0, y0, z0; // node of corner 000
    float g[8];
};

// Adds the cell's 8 corner terms to the window where the corner lies inside
// it (LDS float atomics), else to the brick's gradient in memory.
__device__ __forceinline__ void cell_flush(CellAcc& c, const Window& w, float* win,
                                           float* __restrict__ gb)
{
    if (c.key == 0xFFFFFFFFu)
        return;
    const uint32_t off[8] = {0u, c.dx, c.dy, c.dy + c.dx, c.dz, c.dz + c.dx, c.dz + c.dy,
                             c.dz + c.dy + c.dx};
    const int ddx = c.dx ? 1 : 0, ddy = c.dy ? 1 : 0, ddz = c.dz ? 1 : 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (c.g[k] == 0.f)
            continue;
        const int lx = c.x0 + ((k & 1) ? ddx : 0) - w.o[0];
        const int ly = c.y0 + ((k & 2) ? ddy : 0) - w.o[1];
        const int lz = c.z0 + ((k & 4) ? ddz : 0) - w.o[2];
        if (w.ok && (unsigned)lx < (unsigned)w.e[0] && (unsigned)ly < (unsigned)w.e[1] &&
            (unsigned)lz < (unsigned)w.e[2])
            atomicAdd(win + (lz * w.e[1] + ly) * w.e[0] + lx, c.g[k]);
        else
            unsafeAtomicAdd(gb + c.a + off[k], c.g[k]);
    }
    c.key = 0xFFFFFFFFu;
}

// Moves the accumulator to the cell of taps t.  A ray usually steps to a face
// neighbour, which shares 4 corners with the current cell: those terms stay
// in registers and only the 4 corners left behind are flushed.  A step of
// +-1 in the corner index is a step along x only if both cells are unclamped
// in x (dx != 0), likewise +-n along y (dy != 0) and +-n^2 along z (dz !=
// 0); every other move flushes all 8.
__device__ __forceinline__ void cell_move(CellAcc& c, uint32_t key, const Taps& t, uint32_t n,
                                          const Window& w, float* win, float* __restrict__ gb)
{
    float ng[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (c.key != 0xFFFFFFFFu && t.dx == c.dx && t.dy == c.dy && t.dz == c.dz) {
        const int32_t d = (int32_t)(t.a - c.a);
        const int32_t nn = (int32_t)(n * n);
        float* g = c.g;
        if (t.dx && d == 1) {
            ng[0] = g[1]; ng[2] = g[3]; ng[4] = g[5]; ng[6] = g[7];
            g[1] = g[3] = g[5] = g[7] = 0.f;
        } else if (t.dx && d == -1) {
            ng[1] = g[0]; ng[3] = g[2]; ng[5] = g[4]; ng[7] = g[6];
            g[0] = g[2] = g[4] = g[6] = 0.f;
        } else if (t.dy && d == (int32_t)n) {
            ng[0] = g[2]; ng[1] = g[3]; ng[4] = g[6]; ng[5] = g[7];
            g[2] = g[3] = g[6] = g[7] = 0.f;
        } else if (t.dy && d == -(int32_t)n) {
            ng[2] = g[0]; ng[3] = g[1]; ng[6] = g[4]; ng[7] = g[5];
            g[0] = g[1] = g[4] = g[5] = 0.f;
        } else if (t.dz && d == nn) {
            ng[0] = g[4]; ng[1] = g[5]; ng[2] = g[6]; ng[3] = g[7];
            g[4] = g[5] = g[6] = g[7] = 0.f;
        } else if (t.dz && d == -nn) {
            ng[4] = g[0]; ng[5] = g[1]; ng[6] = g[2]; ng[7] = g[3];
            g[0] = g[1] = g[2] = g[3] = 0.f;
        }
    }
    cell_flush(c, w, win, gb);
#pragma unroll
    for (int k = 0; k < 8; ++k)
        c.g[k] = ng[k];
    c.key = key;
    c.a = t.a;
    c.dx = t.dx;
    c.dy = t.dy;
    c.dz = t.dz;
    c.x0 = t.x0;
    c.y0 = t.y0;
    c.z0 = t.z0;
}

// Per-ray state of the re-march of one segment.
struct SegRay {
    RaySeg r;
    int i;          // next sample
    float seg_w;    // the forward's segment alpha (for its termination)
    float Tseg;     // transmittance through the block so far
    bool ended;
    TfAcc tacc;
    CellAcc cell;
};

// One sample's values before the ray's state enters (the operations of
// fetch3 and classify); v NaN: the sample is skipped.
struct AdjSample {
    float v, xraw, f, cx, cy, cz, base, alpha;
    int i0, i1;
    float4 A0, A1; // the two table entries (field gradient)
    Taps t;        // the corners (field gradient)
};

template <class T>
__device__ __forceinline__ void adj_sample(const rv_block_params& p, const T* __restrict__ brick, const float4* tfs,
                                           int tf_n, float tx, float ty, float tz, AdjSample& s)
{
    const int n = (int)p.nodes;
    const float span = p.fmax - p.fmin;
    float v8[8];
    s.t = fetch_corners(brick, n, tx, ty, tz, v8);
    s.v = trilinear(v8, s.t.fx, s.t.fy, s.t.fz); // the operations of fetch3
    if (!(s.v == s.v))
        return;
    // classify, the operations of classify / fetch_tf
    const float v01 = (s.v - p.fmin) / span;
    s.xraw = v01 * (float)tf_n - 0.5f;
    float x = s.xraw;
    if (!(x > -1.f))
        x = -1.f;
    if (x > (float)tf_n)
        x = (float)tf_n;
    const float fl = floorf(x);
    s.f = x - fl;
    const int ti = (int)fl;
    s.i0 = clampi(ti, 0, tf_n - 1);
    s.i1 = clampi(ti + 1, 0, tf_n - 1);
    s.A0 = tfs[s.i0];
    s.A1 = tfs[s.i1];
    s.cx = lerpf(s.A0.x, s.A1.x, s.f);
    s.cy = lerpf(s.A0.y, s.A1.y, s.f);
    s.cz = lerpf(s.A0.z, s.A1.z, s.f);
    const float tw = lerpf(s.A0.w, s.A1.w, s.f);
    s.base = 1.f - tw * p.opacity_factor;
    s.alpha = 1.f - pow01(s.base, p.i_layer_count);
}

// The ray's state through one kept sample: its colour-table terms (and,
// with gbrick, its field terms), then the segment's transmittance and
// termination.
__device__ __forceinline__ void adj_update(const rv_block_params& p, float* __restrict__ gbrick, float4* gtf,
                                           int tf_n, AdjPixel& st, SegRay& sg, const AdjSample& s,
                                           const Window& w, float* win)
{
    const int n = (int)p.nodes;
    const float span = p.fmax - p.fmin;
    const float cx = s.cx, cy = s.cy, cz = s.cz, alpha = s.alpha, base = s.base, f = s.f;
    const float Ts = st.T * sg.Tseg;
    const float wgt = Ts * alpha;
    st.Px = st.Px + wgt * cx;
    st.Py = st.Py + wgt * cy;
    st.Pz = st.Pz + wgt * cz;
    const float Tn = Ts * (1.f - alpha);
    const bool behind = Tn > 1e-30f;
    // gradient-only quantities use the hardware reciprocal (1 ulp); the
    // classification keeps the forward's exact operations
    const float inv = behind ? __builtin_amdgcn_rcpf(Tn) : 0.f;
    const float Qx = (st.Fx - st.Px) * inv, Qy = (st.Fy - st.Py) * inv, Qz = (st.Fz - st.Pz) * inv;
    const float ratio = behind ? st.Tend * inv : 1.f;
    const float g_alpha = Ts * (((st.gPx * cx + st.gPy * cy) + st.gPz * cz) -
                                ((st.gPx * Qx + st.gPy * Qy) + st.gPz * Qz) + st.gA * ratio);
    const float gcx = st.gPx * wgt, gcy = st.gPy * wgt, gcz = st.gPz * wgt;
    const float gtw = base > 0.f
                          ? g_alpha * (p.i_layer_count * p.opacity_factor * (1.f - alpha) *
                                       __builtin_amdgcn_rcpf(base))
                          : 0.f;
    // colour table
    TfAcc& ta = sg.tacc;
    const float w0 = 1.f - f;
    if (s.i0 != ta.i0 || s.i1 != ta.i1) {
        tf_flush(ta, gtf);
        ta.i0 = s.i0;
        ta.i1 = s.i1;
        ta.g0 = make_float4(0.f, 0.f, 0.f, 0.f);
        ta.g1 = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    ta.g0.x += w0 * gcx;
    ta.g0.y += w0 * gcy;
    ta.g0.z += w0 * gcz;
    ta.g0.w += w0 * gtw;
    ta.g1.x += f * gcx;
    ta.g1.y += f * gcy;
    ta.g1.z += f * gcz;
    ta.g1.w += f * gtw;
    // field
    if (gbrick && s.xraw > -1.f && s.xraw <= (float)tf_n) {
        const float4 A0 = s.A0, A1 = s.A1;
        const Taps& t = s.t;
        const float gv = ((float)tf_n / span) *
                         ((((A1.x - A0.x) * gcx + (A1.y - A0.y) * gcy) + (A1.z - A0.z) * gcz) +
                          (A1.w - A0.w) * gtw);
        if (gv != 0.f) {
            CellAcc& cell = sg.cell;
            const uint32_t key = t.a * 8u + (t.dx ? 1u : 0u) + (t.dy ? 2u : 0u) + (t.dz ? 4u : 0u);
            if (key != cell.key)
                cell_move(cell, key, t, (uint32_t)n, w, win, gbrick);
            const float ux = 1.f - t.fx, uy = 1.f - t.fy, uz = 1.f - t.fz;
            const float gz0 = gv * uz, gz1 = gv * t.fz;
            const float g00 = gz0 * uy, g10 = gz0 * t.fy, g01 = gz1 * uy, g11 = gz1 * t.fy;
            cell.g[0] += g00 * ux;
            cell.g[1] += g00 * t.fx;
            cell.g[2] += g10 * ux;
            cell.g[3] += g10 * t.fx;
            cell.g[4] += g01 * ux;
            cell.g[5] += g01 * t.fx;
            cell.g[6] += g11 * ux;
            cell.g[7] += g11 * t.fx;
        }
    }
    sg.Tseg = sg.Tseg * (1.f - alpha);
    const float da = alpha * (1.f - sg.seg_w);
    sg.seg_w = sg.seg_w + da;
    if (sg.seg_w > p.alpha_cutoff)
        sg.ended = true;
}

// The samples of one ray segment from sample sg.i on, up to the slab end:
// along axis A (0 x, 1 y, 2 z) a sample whose cell index passes `lim`
// (> lim when up, < lim otherwise) is left for the next slab; A < 0: no
// slab limit.
template <class T>
__device__ __forceinline__ void march_adjoint(const rv_block_params& p, const T* __restrict__ brick,
                                              float* __restrict__ gbrick, const float4* tfs,
                                              float4* gtf, int tf_n, AdjPixel& st, SegRay& sg,
                                              int A, bool up, int lim, const Window& w, float* win)
{
    const int n = (int)p.nodes;
    while (!sg.ended && sg.i < sg.r.sample_count) {
        const float tx = (sg.r.px - p.tex_origin[0]) * p.tcf;
        const float ty = (sg.r.py - p.tex_origin[1]) * p.tcf;
        const float tz = (sg.r.pz - p.tex_origin[2]) * p.tcf;
        if (A >= 0) {
            int c0, c1;
            float fdum;
            axis_taps(A == 0 ? tx : (A == 1 ? ty : tz), n, c0, c1, fdum);
            if (up ? c0 > lim : c0 < lim)
                break; // the next slab's
        }
        sg.r.px = sg.r.px + sg.r.sx;
        sg.r.py = sg.r.py + sg.r.sy;
        sg.r.pz = sg.r.pz + sg.r.sz;
        ++sg.i;
        AdjSample smp;
        adj_sample(p, brick, tfs, tf_n, tx, ty, tz, smp);
        if (!(smp.v == smp.v))
            continue;
        adj_update(p, gbrick, gtf, tf_n, st, sg, smp, w, win);
    }
}

// Lane j's sample values (j uniform among the active lanes).
__device__ __forceinline__ AdjSample lane_sample(const AdjSample& s, int j)
{
    AdjSample o;
    o.v = lane_f(s.v, j);
    o.xraw = lane_f(s.xraw, j);
    o.f = lane_f(s.f, j);
    o.cx = lane_f(s.cx, j);
    o.cy = lane_f(s.cy, j);
    o.cz = lane_f(s.cz, j);
    o.base = lane_f(s.base, j);
    o.alpha = lane_f(s.alpha, j);
    o.i0 = __builtin_amdgcn_readlane(s.i0, j);
    o.i1 = __builtin_amdgcn_readlane(s.i1, j);
    o.A0 = make_float4(lane_f(s.A0.x, j), lane_f(s.A0.y, j), lane_f(s.A0.z, j), lane_f(s.A0.w, j));
    o.A1 = make_float4(lane_f(s.A1.x, j), lane_f(s.A1.y, j), lane_f(s.A1.z, j), lane_f(s.A1.w, j));
    o.t.a = (uint32_t)__builtin_amdgcn_readlane((int)s.t.a, j);
    o.t.dx = (uint32_t)__builtin_amdgcn_readlane((int)s.t.dx, j);
    o.t.dy = (uint32_t)__builtin_amdgcn_readlane((int)s.t.dy, j);
    o.t.dz = (uint32_t)__builtin_amdgcn_readlane((int)s.t.dz, j);
    o.t.fx = lane_f(s.t.fx, j);
    o.t.fy = lane_f(s.t.fy, j);
    o.t.fz = lane_f(s.t.fz, j);
    o.t.x0 = __builtin_amdgcn_readlane(s.t.x0, j);
    o.t.y0 = __builtin_amdgcn_readlane(s.t.y0, j);
    o.t.z0 = __builtin_amdgcn_readlane(s.t.z0, j);
    return o;
}

// march_adjoint for the segments left after a capped march (sample index
// capn or later): rays that graze a block's edge and start at the eye
// (thousands of samples).  All lanes of the wave are on block p.  Each such
// ray is continued by the whole wave, 64 samples at a time: every lane runs
// the ray's chain of fp32 position additions and keeps its own sample's
// position, computes that sample's slab test and values (adj_sample), and
// the ray's own lane takes its state through the samples in order
// (adj_update), up to the slab's end, the segment's end or termination --
// the operations of the one-lane march, which it leaves where that march
// would (sample index and position).
template <class T>
__device__ __forceinline__ void continue_adjoint(const rv_block_params& p, const T* __restrict__ brick,
                                                 float* __restrict__ gbrick, const float4* tfs,
                                                 float4* gtf, int tf_n, AdjPixel& st, SegRay& sg,
                                                 int A, bool up, int lim, const Window& w, float* win,
                                                 int capn)
{
    const int n = (int)p.nodes;
    const int me = (int)(threadIdx.x & 63u);
    uint64_t left = __ballot(!sg.ended && sg.i >= capn && sg.i < sg.r.sample_count);
    while (left) {
        const int L = (int)__builtin_ctzll(left);
        left &= left - 1ull;
        float Px = lane_f(sg.r.px, L), Py = lane_f(sg.r.py, L), Pz = lane_f(sg.r.pz, L);
        const float Sx = lane_f(sg.r.sx, L), Sy = lane_f(sg.r.sy, L), Sz = lane_f(sg.r.sz, L);
        const int N = __builtin_amdgcn_readlane(sg.r.sample_count, L);
        int base = __builtin_amdgcn_readlane(sg.i, L); // sample index of lane 0's position
        for (;;) {
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
            const float tx = (qx - p.tex_origin[0]) * p.tcf;
            const float ty = (qy - p.tex_origin[1]) * p.tcf;
            const float tz = (qz - p.tex_origin[2]) * p.tcf;
            bool past = false;
            if (A >= 0) {
                int c0, c1;
                float fdum;
                axis_taps(A == 0 ? tx : (A == 1 ? ty : tz), n, c0, c1, fdum);
                past = up ? c0 > lim : c0 < lim;
            }
            const uint64_t pm = __ballot(past);
            const int kstop = pm ? (int)__builtin_ctzll(pm) : 64; // the first sample of the next slab
            AdjSample smp;
            adj_sample(p, brick, tfs, tf_n, tx, ty, tz, smp);
            const uint64_t vm = __ballot(smp.v == smp.v);
            // the loop and its lane reads run on the whole wave, only the
            // update on lane L: a value read out of lane j inside a branch
            // that only lane L takes can be computed inside that branch by
            // the compiler, which lane j then never executes
            for (int j = 0; j < kstop && base + j < N; ++j) {
                if (__builtin_amdgcn_readlane(sg.ended ? 1 : 0, L))
                    break;
                if (me == L)
                    sg.i = base + j + 1;
                if (!((vm >> j) & 1ull)) // NaN: skipped, as in the one-lane march
                    continue;
                const AdjSample sj = lane_sample(smp, j);
                if (me == L)
                    adj_update(p, gbrick, gtf, tf_n, st, sg, sj, w, win);
            }
            const int iL = __builtin_amdgcn_readlane(sg.i, L);
            const bool uend = __builtin_amdgcn_readlane(sg.ended ? 1 : 0, L) != 0;
            const int used = iL - base; // 0 .. 64 samples taken; the next is sample iL
            const int k = used & 63;
            const float nx = lane_f(qx, k), ny = lane_f(qy, k), nz = lane_f(qz, k);
            if (me == L) {
                sg.r.px = used < 64 ? nx : Px;
                sg.r.py = used < 64 ? ny : Py;
                sg.r.pz = used < 64 ? nz : Pz;
            }
            if (used < 64 || uend || iL >= N)
                break;
            base += 64;
        }
    }
}

// One work item (one block, one 8 x 8 tile of its view rectangle) for the
// lanes of one wave; active: this lane's pixel has a segment in the block.
// With field gradients the block is walked in slabs of kSlab cells along the
// axis the rays advance along most: the wave's rays pause at each slab's far
// side, their field terms go to the wave's LDS window covering the slab's
// part of their frustum, and the window is added to the brick's gradient in
// x-fastest order, so the global atomics of a wave fall into a few cache
// lines (scattered float atomics cost one L2 request per lane).  A window
// above kWinWave floats, or rays advancing both ways along the axis, fall
// back to direct atomics.  The window is zero between items: every flush
// clears what it adds.  In each slab a lane marches its own segment up to
// the cap (twice the samples of the longest ray inside the block), and the
// wave continues longer segments together (continue_adjoint).  All
// decisions are wave-uniform (ballots and butterfly reductions).
template <class T>
__device__ void adjoint_item(const rv_block_params& p, const T* __restrict__ brick,
                             float* __restrict__ gbrick, const float4* tfs, float4* gtf, int tf_n,
                             bool active, const RaySeg& r, AdjPixel& st, float* win)
{
    const int n = (int)p.nodes;
    const uint32_t lane = threadIdx.x & 63u;
    SegRay sg;
    sg.r = r;
    sg.i = 1;
    sg.seg_w = 0.f;
    sg.Tseg = 1.f;
    sg.ended = !active;
    sg.tacc.i0 = -1;
    sg.cell.key = 0xFFFFFFFFu;
    const int full = sg.r.sample_count;
    // the one-lane march's cap: twice the samples of the longest ray inside
    // the block (sample indices start at 1)
    const int capn = min(full, (int)(2.f * p.quality * (float)p.nodes) + 1);
    Window w;
    w.ok = false;
    w.o[0] = w.o[1] = w.o[2] = 0;
    w.e[0] = w.e[1] = w.e[2] = 0;
    // slab axis A (uniform): the axis the view follows most; U, V the others
    const float ax = fabsf(p.n[0]), ay = fabsf(p.n[1]), az = fabsf(p.n[2]);
    const int A = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    const int U = A == 0 ? 1 : 0, V = 3 - A - U;
    bool up = true, slabs = false;
    if (gbrick) {
        const float sA = A == 0 ? r.sx : (A == 1 ? r.sy : r.sz);
        const bool any_up = __ballot(active && sA > 0.f) != 0ull;
        const bool any_down = __ballot(active && !(sA > 0.f)) != 0ull;
        up = any_up;
        slabs = any_up != any_down;
    }
    // slabs of at most kSlab cells, walked in the rays' direction along A;
    // a slab whose window would not fit is halved until it does (or is one
    // cell thick)
    int cur = up ? 0 : n - 1; // next cell index along A
    for (;;) {
        if (slabs && __ballot(!sg.ended && sg.i < sg.r.sample_count) == 0ull)
            break; // every ray of the wave has left the block
        int lo = up ? cur : max(cur - kSlab + 1, 0);
        int hi = up ? min(cur + kSlab - 1, n - 1) : cur; // cells lo..hi along A
        if (slabs) {
            for (;;) {
                // window bounds: the part of each ray's remaining samples in
                // this slab, from its position and step in texel units,
                // widened by one node for the fp32 accumulation of positions
                int umin = 0x7fffffff, umax = -1, vmin = 0x7fffffff, vmax = -1;
                if (!sg.ended && sg.i < sg.r.sample_count) {
                    const float tx = (sg.r.px - p.tex_origin[0]) * p.tcf;
                    const float ty = (sg.r.py - p.tex_origin[1]) * p.tcf;
                    const float tz = (sg.r.pz - p.tex_origin[2]) * p.tcf;
                    const float dx = sg.r.sx * p.tcf, dy = sg.r.sy * p.tcf, dz = sg.r.sz * p.tcf;
                    const float tA = A == 0 ? tx : (A == 1 ? ty : tz), dA = A == 0 ? dx : (A == 1 ? dy : dz);
                    const float tU = U == 0 ? tx : ty, dU = U == 0 ? dx : dy;
                    const float tV = V == 1 ? ty : tz, dV = V == 1 ? dy : dz;
                    const float rem = (float)(sg.r.sample_count - 1 - sg.i);
                    float la = ((float)lo + 0.5f - tA) / dA;
                    float lb = ((float)hi + 1.5f - tA) / dA;
                    if (la > lb) {
                        const float q = la;
                        la = lb;
                        lb = q;
                    }
                    la = fmaxf(la, 0.f);
                    lb = fminf(lb, rem);
                    if (la <= lb) {
                        const float ua = tU + la * dU, ub = tU + lb * dU;
                        const float va = tV + la * dV, vb = tV + lb * dV;
                        umin = max((int)floorf(fminf(ua, ub) - 0.5f) - 1, 0);
                        umax = min((int)floorf(fmaxf(ua, ub) - 0.5f) + 2, n - 1);
                        vmin = max((int)floorf(fminf(va, vb) - 0.5f) - 1, 0);
                        vmax = min((int)floorf(fmaxf(va, vb) - 0.5f) + 2, n - 1);
                    }
                }
                umin = wave_min(umin);
                umax = wave_max(umax);
                vmin = wave_min(vmin);
                vmax = wave_max(vmax);
                const int nl = min(hi + 1, n - 1) - lo + 1;
                const int nu = umax - umin + 1, nv = vmax - vmin + 1;
                w.ok = umax >= umin && vmax >= vmin && nu * nv * nl <= (int)kWinWave;
                if (w.ok) {
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        w.o[j] = j == A ? lo : (j == U ? umin : vmin);
                        w.e[j] = j == A ? nl : (j == U ? nu : nv);
                    }
                }
                if (w.ok || hi == lo || umax < umin || vmax < vmin)
                    break;
                // too wide: a thinner slab
                const int k = (hi - lo + 1) >> 1;
                if (up)
                    hi = lo + k - 1;
                else
                    lo = hi - k + 1;
            }
        }
        // each lane marches at most to sample capn - 1 by itself; the wave
        // continues longer segments together
        sg.r.sample_count = capn;
        march_adjoint(p, brick, gbrick, tfs, gtf, tf_n, st, sg, slabs ? A : -1, up, up ? hi : lo, w,
                      win);
        sg.r.sample_count = full;
        continue_adjoint(p, brick, gbrick, tfs, gtf, tf_n, st, sg, slabs ? A : -1, up, up ? hi : lo, w,
                         win, capn);
        if (gbrick)
            cell_flush(sg.cell, w, win, gbrick); // the slab's last cell
        if (!slabs)
            break;
        cur = up ? hi + 1 : lo - 1;
        if (w.ok) {
            // the window, x fastest, to the brick's gradient; cleared behind
            wave_sync();
            const int ex = w.e[0], exy = w.e[0] * w.e[1], size = exy * w.e[2];
            for (int j = (int)lane; j < size; j += 64) {
                const float g = win[j];
                if (g == 0.f)
                    continue;
                win[j] = 0.f;
                const int lz = j / exy, rxy = j - lz * exy;
                const int ly = rxy / ex, lx = rxy - ly * ex;
                unsafeAtomicAdd(gbrick + ((uint32_t)(w.o[2] + lz) * n + (uint32_t)(w.o[1] + ly)) * n +
                                    (uint32_t)(w.o[0] + lx),
                                g);
            }
            wave_sync();
        }
        if (cur < 0 || cur > n - 1)
            break;
    }
    tf_flush(sg.tacc, gtf);
}

// The re-march's start for the record rec of block p, pixel (lx, ly) of its
// view rectangle: the state in front of the block (tape), the forward image
// and dL/dimage at the pixel; false when the ray never reaches the block,
// carries no gradient, or misses it.
__device__ __forceinline__ bool pixel_start(const AdjArgs& a, const rv_block_params& p, uint64_t recidx,
                                            uint32_t lx, uint32_t ly, AdjPixel& st, RaySeg& r)
{
    const float4 rec = a.tape[recidx];
    if (!(rec.w >= 0.f)) // the pixel's ray does not reach this block
        return false;
    const uint32_t vx = p.view_x + lx, vy = p.view_y + ly;
    const uint32_t pix = (a.H - 1u - vy) * a.W + vx;
    const float4 I = a.rgba[pix];
    const float4 g = a.grad_rgba[pix];
    st.T = rec.w;
    st.Px = rec.x;
    st.Py = rec.y;
    st.Pz = rec.z;
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
    if (st.gPx == 0.f && st.gPy == 0.f && st.gPz == 0.f && st.gA == 0.f)
        return false;
    r = setup_ray(p, vx, vy);
    return r.hit;
}

// Colour-table gradient only (no field gradient arrays): one lane per tape
// record (block, pixel of its view rectangle), records in tape order, so a
// wave's lanes are one block's consecutive pixels, or the pixels of a few
// consecutive blocks, and small blocks fill the wave.  Each lane re-marches
// its segment with its own block's parameters, at most the cap (twice the
// samples of the longest ray inside the block); a record with samples left
// (a ray that grazes a block's edge and starts at the eye: thousands of
// samples) is continued by the whole wave, 64 samples at a time, with its
// block's parameters and brick: the positions by the same chain of fp32
// additions (every lane runs it and keeps its own), each sample's values by
// one lane (adj_sample), and the ray's state through the samples in order
// by the record's own lane, which reads them out of the others
// (adj_update) -- the operations of the one-lane march.
template <class BrickT>
__global__ __launch_bounds__(kRT) void adjoint_records_kernel(AdjArgs a)
{
    extern __shared__ float4 dyn[]; // colour table and its gradient
    float4* tfs = dyn;
    float4* gtf = dyn + a.tf_n;
    for (uint32_t i = threadIdx.x; i < a.tf_n; i += kRT) {
        tfs[i] = a.tf[i];
        gtf[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * kRT;
    const uint32_t lane = threadIdx.x & 63u;
    const int me = (int)lane;
    Window w;
    w.ok = false;
    w.o[0] = w.o[1] = w.o[2] = 0;
    w.e[0] = w.e[1] = w.e[2] = 0;
    // wave-uniform loop: every lane takes part in each round (a lane past
    // the last record, or whose record carries nothing, marches nothing)
    for (uint64_t r0 = (uint64_t)blockIdx.x * kRT + (threadIdx.x - lane); r0 < a.nrec; r0 += stride) {
        const uint64_t rid = r0 + lane;
        const uint64_t rq = rid < a.nrec ? rid : a.nrec - 1u; // (lookups of lanes past the end)
        uint32_t blk = a.rec_chunk[rq >> 6];
        while (blk + 1u < a.nblocks && a.rec_off[blk + 1u] <= rq)
            ++blk;
        const rv_block_params p = a.params[blk];
        const BrickT* brick = reinterpret_cast<const BrickT*>(a.bricks[blk]);
        const uint32_t li = (uint32_t)(rq - a.rec_off[blk]);
        const uint32_t ly = li / p.view_w, lx = li - ly * p.view_w;
        AdjPixel st{};
        SegRay sg;
        const bool ok = rid < a.nrec && pixel_start(a, p, rid, lx, ly, st, sg.r);
        if (!ok)
            sg.r.sample_count = 0;
        sg.i = 1;
        sg.seg_w = 0.f;
        sg.Tseg = 1.f;
        sg.ended = false;
        sg.tacc.i0 = -1;
        sg.cell.key = 0xFFFFFFFFu;
        const int full = sg.r.sample_count;
        const int cap = (int)(2.f * p.quality * (float)p.nodes);
        sg.r.sample_count = min(full, cap + 1);
        march_adjoint(p, brick, nullptr, tfs, gtf, (int)a.tf_n, st, sg, -1, true, 0, w, nullptr);
        sg.r.sample_count = full;
        uint64_t left = __ballot(ok && !sg.ended && sg.i < full);
        while (left) {
            const int L = (int)__builtin_ctzll(left);
            left &= left - 1ull;
            const rv_block_params pu = lane_params(p, L);
            const BrickT* ub = lane_ptr(brick, L);
            float Px = lane_f(sg.r.px, L), Py = lane_f(sg.r.py, L), Pz = lane_f(sg.r.pz, L);
            const float Sx = lane_f(sg.r.sx, L), Sy = lane_f(sg.r.sy, L), Sz = lane_f(sg.r.sz, L);
            const int N = __builtin_amdgcn_readlane(full, L);
            int base = __builtin_amdgcn_readlane(sg.i, L);
            bool uend = false;
            for (; base < N && !uend; base += 64) {
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
                AdjSample smp;
                adj_sample(pu, ub, tfs, (int)a.tf_n, (qx - pu.tex_origin[0]) * pu.tcf,
                           (qy - pu.tex_origin[1]) * pu.tcf, (qz - pu.tex_origin[2]) * pu.tcf, smp);
                const uint64_t vm = __ballot(base + me < N && smp.v == smp.v);
                // lane reads on the whole wave, the update on lane L (as in
                // continue_adjoint)
                for (int j = 0; j < 64 && base + j < N; ++j) {
                    if (__builtin_amdgcn_readlane(sg.ended ? 1 : 0, L))
                        break;
                    if (!((vm >> j) & 1ull)) // NaN: skipped, as in the one-lane march
                        continue;
                    const AdjSample sj = lane_sample(smp, j);
                    if (me == L)
                        adj_update(p, nullptr, gtf, (int)a.tf_n, st, sg, sj, w, nullptr);
                }
                uend = __builtin_amdgcn_readlane(sg.ended ? 1 : 0, L) != 0;
            }
        }
        tf_flush(sg.tacc, gtf);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < a.tf_n; i += kRT) {
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

// Persistent grid; work items (block, tile) in order, tiles of a block
// consecutive; each wave takes the next item from the queue counter, so
// waves whose rays are short take more items, and the waves working at one
// time are on neighbouring tiles of a few blocks.
template <class BrickT>
__global__ __launch_bounds__(kRT) void adjoint_kernel(AdjArgs a)
{
    extern __shared__ float4 dyn[]; // colour table, its gradient, the waves' gradient windows
    float4* tfs = dyn;
    float4* gtf = dyn + a.tf_n;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    float* win = reinterpret_cast<float*>(dyn + 2 * a.tf_n) + wave * kWinWave;
    for (uint32_t i = threadIdx.x; i < a.tf_n; i += kRT) {
        tfs[i] = a.tf[i];
        gtf[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (a.grad_bricks)
        for (uint32_t j = lane; j < kWinWave; j += 64u)
            win[j] = 0.f;
    __syncthreads();
    for (;;) {
        uint32_t item = 0u;
        if (lane == 0u)
            item = atomicAdd(a.queue, 1u);
        item = uniform(__shfl(item, 0));
        if (item >= a.nitems)
            break;
        const uint32_t blk = uniform(a.items[2 * item]), tile = uniform(a.items[2 * item + 1]);
        const rv_block_params p = a.params[blk];
        const uint32_t ntx = (p.view_w + 7u) >> 3;
        const uint32_t lx = (tile % ntx) * 8u + (lane & 7u), ly = (tile / ntx) * 8u + (lane >> 3);
        AdjPixel st{};
        RaySeg r{};
        bool active = false;
        if (lx < p.view_w && ly < p.view_h)
            active = pixel_start(a, p, a.rec_off[blk] + ly * p.view_w + lx, lx, ly, st, r);
        adjoint_item(p, reinterpret_cast<const BrickT*>(a.bricks[blk]),
                     a.grad_bricks ? a.grad_bricks[blk] : nullptr, tfs, gtf, (int)a.tf_n, active, r,
                     st, win);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < a.tf_n; i += kRT) {
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

__global__ void adjoint_queue_reset_kernel(uint32_t* queue)
{
    if (threadIdx.x == 0u)
        *queue = 0u;
}

} // namespace
} // namespace rv

using namespace rv;

static size_t adjoint_lds(uint32_t tf_n) { return 32ull * tf_n + 4ull * kWinWave * kAdjWaves; }

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
            RV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, adjoint_kernel<__half>, kRT,
                                                                adjoint_lds(P->desc.tf_n)));
        else if (P->tiled)
            RV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, adjoint_kernel<Tiled4>, kRT,
                                                                adjoint_lds(P->desc.tf_n)));
        else
            RV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, adjoint_kernel<float>, kRT,
                                                                adjoint_lds(P->desc.tf_n)));
        RV_REQUIRE(per_cu >= 1, "rv_frame_adjoint: kernel cannot be resident");
        P->adj_grid = (uint32_t)(cus * per_cu);
    }
    RV_REQUIRE(P->tape_on && P->tape_cam == (int64_t)cam,
               "rv_frame_adjoint: render this camera with the tape on (rv_frame_set_tape) and a float "
               "image first");
    AdjArgs a{};
    a.params = P->d_params + (size_t)cam * P->nblocks;
    a.bricks = P->d_bricks;
    a.grad_bricks = P->d_grad_bricks;
    a.items = P->d_items + 2 * P->item_off[cam];
    a.nitems = (uint32_t)(P->item_off[cam + 1] - P->item_off[cam]);
    a.rec_off = P->d_rec_off + (size_t)cam * P->nblocks;
    a.tape = P->d_tape;
    a.tf = reinterpret_cast<const float4*>(P->desc.d_tf);
    a.tf_n = P->desc.tf_n;
    a.W = P->desc.vp_w;
    a.H = P->desc.vp_h;
    a.rgba = reinterpret_cast<const float4*>(d_rgba);
    a.grad_rgba = reinterpret_cast<const float4*>(d_grad_rgba);
    a.grad_tf = reinterpret_cast<float4*>(d_grad_tf);
    if (a.nitems == 0)
        return RV_OK;
    // Colour-table-only gradients of blocks a few pixels wide: their 8 x 8
    // tiles leave most lanes of a wave idle, so one lane per record instead
    const uint64_t nrec = P->rec_total[cam];
    if (!a.grad_bricks && nrec < 48ull * a.nitems) {
        a.rec_chunk = P->d_rec_chunk + P->chunk_off[cam];
        a.nrec = nrec;
        a.nblocks = P->nblocks;
        if (!P->rec_grid) {
            int dev = 0, cus = 0, per_cu = 0;
            RV_HIP(hipGetDevice(&dev));
            RV_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            if (P->fp16)
                RV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, adjoint_records_kernel<__half>,
                                                                    kRT, 32ull * P->desc.tf_n));
            else if (P->tiled)
                RV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, adjoint_records_kernel<Tiled4>,
                                                                    kRT, 32ull * P->desc.tf_n));
            else
                RV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, adjoint_records_kernel<float>,
                                                                    kRT, 32ull * P->desc.tf_n));
            RV_REQUIRE(per_cu >= 1, "rv_frame_adjoint: kernel cannot be resident");
            P->rec_grid = (uint32_t)(cus * per_cu);
        }
        const uint64_t need = (a.nrec + kRT - 1) / kRT;
        const uint32_t rgrid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(P->rec_grid, need));
        if (P->fp16)
            hipLaunchKernelGGL(adjoint_records_kernel<__half>, dim3(rgrid), dim3(kRT), 32ull * a.tf_n,
                               S(stream), a);
        else if (P->tiled)
            hipLaunchKernelGGL(adjoint_records_kernel<Tiled4>, dim3(rgrid), dim3(kRT), 32ull * a.tf_n,
                               S(stream), a);
        else
            hipLaunchKernelGGL(adjoint_records_kernel<float>, dim3(rgrid), dim3(kRT), 32ull * a.tf_n,
                               S(stream), a);
        RV_LAUNCH_CHECK();
        return RV_OK;
    }
    const uint32_t grid = P->adj_grid;
    a.queue = P->d_counters + 4; // past the frame kernels' counters 0..3
    hipLaunchKernelGGL(adjoint_queue_reset_kernel, dim3(1), dim3(64), 0, S(stream), a.queue);
    RV_LAUNCH_CHECK();
    if (P->fp16)
        hipLaunchKernelGGL(adjoint_kernel<__half>, dim3(grid), dim3(kRT), adjoint_lds(a.tf_n),
                           S(stream), a);
    else if (P->tiled)
        hipLaunchKernelGGL(adjoint_kernel<Tiled4>, dim3(grid), dim3(kRT), adjoint_lds(a.tf_n), S(stream),
                           a);
    else
        hipLaunchKernelGGL(adjoint_kernel<float>, dim3(grid), dim3(kRT), adjoint_lds(a.tf_n),
                           S(stream), a);
    RV_LAUNCH_CHECK();
    return RV_OK;
    RV_CATCH
}
