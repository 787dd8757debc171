/*
 * revise_hip.h -- C ABI of the MI355X (gfx950) sparse->dense interpolation and
 * volume ray-casting path of ReVisE.
 *
 * The reference (deadmorous/revise) exposes this path as C++ plug-in points;
 * each entry point below names the reference interface it replaces
 * (file:line, paths relative to the reference repository root).
 *
 * Conventions
 *  - Every function returns an int status: RV_OK (0) or an RV_ERR_* code.
 *    No C++ exception crosses this boundary; rv_last_error() returns a
 *    thread-local message for the last failure on the calling thread.
 *  - Pointers named d_* are device pointers on the current HIP device, owned
 *    by the caller.  The library keeps no pointer past a call, except inside
 *    the opaque handles (rv_interp_batch, rv_frame_plan, rv_worker), which
 *    document what they hold.
 *  - Work is enqueued on the given stream (0 = the default stream) and is
 *    asynchronous unless stated otherwise.
 *  - Dense bricks are (2^depth+1)^3 nodes, x fastest, then y, then z
 *    (IndexTransform.hpp:250-257).  Node coordinates (n2i) are uint16 x 3
 *    (BlockTreeNodeCoord.hpp:24).  The "no field" sentinel in sparse values is
 *    1e30f and becomes NaN in dense bricks (BlockTreeFieldProvider.hpp:137-139,
 *    DenseFieldInterpolator.hpp:416-418).
 */
#ifndef REVISE_HIP_H
#define REVISE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* rv_stream_t; /* == hipStream_t */

enum {
    RV_OK = 0,
    RV_ERR_INVALID = 1,     /* bad argument (size, depth, null pointer, ...) */
    RV_ERR_HIP = 2,         /* a HIP runtime call failed */
    RV_ERR_OOM = 3,         /* device allocation failed */
    RV_ERR_CANCELLED = 4,   /* the caller's cancel flag was set */
    RV_ERR_UNSUPPORTED = 5, /* valid request this build does not serve */
    RV_ERR_INTERNAL = 6
};

/* Thread-local message of the last failure on this thread ("" if none). */
const char* rv_last_error(void);
/* Library ABI version (major*10000 + minor*100 + patch). */
int rv_version(void);

/* ------------------------------------------------------------------------
 * Block tree and node index map (host)
 * ------------------------------------------------------------------------ */

/* Depth of a subtree, min(actual depth, max_depth).
 * Replaces CompressedBlockTree::depthUpTo (CompressedBlockTree.hpp:290-300),
 * as used for Metadata subtree depth (Metadata.hpp:414-416).
 * children: nnodes x 8 child indices (0 = none), node 0 = subtree root,
 * child local index bit0 = x, bit1 = y, bit2 = z. */
int rv_subtree_depth(const uint32_t* children, uint32_t nnodes, uint32_t max_depth,
                     uint32_t* depth_out);

/* n2i list of a subtree (first-visit order of block corners in a pre-order
 * walk limited to `depth` levels).  Replaces BlockTreeNodes<3,BT>::BlockTreeNodes
 * (BlockTreeNodes.hpp:66-80, 166-181) / Metadata::blockTreeNodes
 * (Metadata.hpp:255-268).  n2i: cap x 3 uint16.  *count receives the node
 * count; RV_ERR_INVALID if cap is too small or the tree is malformed. */
int rv_subtree_n2i(const uint32_t* children, uint32_t nnodes, uint32_t depth, uint16_t* n2i,
                   uint64_t cap, uint64_t* count);

/* The same list built on the device (depth <= 8): the replacement of the
 * host DFS and hash lookups of BlockTreeNodes (BlockTreeNodes.hpp:166-181)
 * by level-parallel expansion, pre-order positions from subtree sizes, and
 * first sightings on a dense lattice.  d_children: nnodes x 8 uint32 on the
 * device (0 = no child, node 0 the root); d_n2i: cap x 3 uint16 on the
 * device.  Synchronous; same result and errors as rv_subtree_n2i. */
int rv_subtree_n2i_device(const uint32_t* d_children, uint32_t nnodes, uint32_t depth,
                          uint16_t* d_n2i, uint64_t cap, uint64_t* count, rv_stream_t stream);

/* Node -> dense ordinal map on the device: d_out[i] = vertexIndexToOrdinal(n2i[i], depth).
 * Replaces I3::vertexIndexToOrdinal (DenseFieldInterpolator_cu_3.cu:24-30). */
int rv_node_ordinals(const uint16_t* d_n2i, uint64_t n, uint32_t depth, uint32_t* d_out,
                     rv_stream_t stream);

/* ------------------------------------------------------------------------
 * Sparse -> dense interpolation (device)
 * ------------------------------------------------------------------------ */

/* Number of nodes of a dense brick, (2^depth+1)^3. */
uint64_t rv_brick_nodes(uint32_t depth);

/* Gather plan of a subtree, from its n2i (host array, n x 3).  The kernel
 * completes a brick in tiles of at most 64 x 16 x 16 cells; the plan lists,
 * per tile, the sparse nodes inside the tile's closed box in increasing
 * sparse index with their offsets in the tile's LDS array, and the sparse
 * index of each node of the brick's stride-16 lattice.  Its size is about
 * 6 bytes per (node, tile) pair.  A plan depends only on the subtree's n2i,
 * so callers cache it per subtree (as the reference caches BlockTreeNodes,
 * BlockTreeFieldService.hpp:133-137).  RV_ERR_INVALID if a coordinate lies
 * outside the (2^depth+1)^3 brick. */
uint64_t rv_interp_plan_bytes(const uint16_t* n2i, uint64_t n, uint32_t depth);
int rv_interp_plan_encode(const uint16_t* n2i, uint64_t n, uint32_t depth, void* plan,
                          uint64_t bytes);
/* Encode and upload (synchronous); free with rv_interp_plan_free. */
int rv_interp_plan_build(const uint16_t* n2i, uint64_t n, uint32_t depth, void** d_plan,
                         uint64_t* bytes);
int rv_interp_plan_free(void* d_plan);

/* Interpolate one brick from its sparse values through a plan.
 * Replaces DenseFieldInterpolator_cu_3::interpolate
 * (DenseFieldInterpolator_cu_3.cu:197-253; DenseFieldInterpolatorArr_cu_3.cu:197-257),
 * with the CPU semantics of DenseFieldInterpolator<3>::interpolate
 * (DenseFieldInterpolator.hpp:198-419), bit for bit.  depth must be the
 * depth the plan was built for. */
int rv_interp_brick_planned(const float* d_sparse, const void* d_plan, uint32_t depth,
                            float* d_dense, rv_stream_t stream);

/* Plan + interpolate in one call, from the host n2i (the plain drop-in for
 * DenseFieldInterpolatorTpl_cu_3::interpolate(denseDev, sparse, nodes),
 * DenseFieldInterpolatorTpl_cu_3.hpp:103-107, which uploads the n2i per
 * call).  The plan is staged stream-ordered and the call synchronizes the
 * stream before returning. */
int rv_interp_brick(const float* d_sparse, const uint16_t* n2i, uint64_t n, uint32_t depth,
                    float* d_dense, rv_stream_t stream);

/* Dense brick layouts.  LINEAR: (2^depth+1)^3 nodes, x fastest, then y, z
 * (the reference's).  TILED4 (fp32): micro-bricks of 4 x 4 x 4 cells, 80
 * floats each -- 16 rows (y fastest, then z) of 5 x-positions, micro-brick
 * m holding x-positions 4m .. 4m+4 (x neighbours share one position) --
 * ordered x fastest over G^3 micro-bricks, G = floor((n+1)/4) + 1 for n =
 * 2^depth+1 nodes per axis.  Node (x, y, z) is at position (x+1, y, z):
 * index ((z/4 * G + y/4) * G + (x+1)/4) * 80 + z%4 * 20 + y%4 * 5 +
 * (x+1)%4 (and, where (x+1)%4 == 0 and x+1 >= 4, also at the last position
 * of the previous micro-brick's row); x-positions 0 and n .. 4G hold copies
 * of the edge nodes 0 and n-1, other padding is written NaN and never read;
 * bricks 16-byte aligned.  The ray-casts read a cell's 8 corners as 4
 * 8-byte row loads from fewer cache lines; the values are the same. */
enum { RV_BRICK_LINEAR = 0, RV_BRICK_TILED4 = 1 };
/* Floats a brick of this depth takes in a layout. */
uint64_t rv_brick_floats(uint32_t depth, uint32_t layout);

/* One brick of a batch.  dense16 != NULL stores fp16 instead of fp32
 * (exactly one of dense / dense16 is set); layout: RV_BRICK_* (TILED4 is
 * fp32 only). */
typedef struct rv_brick_job {
    const float* sparse;
    const void* plan;
    float* dense;
    uint16_t* dense16;
    uint32_t depth;
    uint32_t layout;
} rv_brick_job;

/* Batched interpolation of many bricks of any depths in one launch.
 * There is no reference counterpart: the reference interpolates one brick per
 * call (DenseFieldRenderer_cu.cu:504-545).  The handle owns a device copy of
 * the job table and its tile index; the bricks stay caller-owned. */
typedef struct rv_interp_batch rv_interp_batch;
int rv_interp_batch_create(const rv_brick_job* jobs, uint32_t count, rv_interp_batch** out);
int rv_interp_batch_run(rv_interp_batch* batch, rv_stream_t stream);
/* Total tiles (workgroups) of one run, for roofline accounting. */
int rv_interp_batch_info(const rv_interp_batch* batch, uint64_t* tiles, uint64_t* dense_nodes);
void rv_interp_batch_destroy(rv_interp_batch* batch);

/* ------------------------------------------------------------------------
 * Per-block ray-cast (device) and its host parameter setup
 * ------------------------------------------------------------------------ */

/* Parameters of one block render: RenderKernelParam (RenderKernelParam.hpp:26-116)
 * restricted to the Argb mode, plus the per-block constants the reference
 * recomputes per pixel (DenseFieldRenderer_cu.cu:164-181, renderArgb.hpp:34-38). */
typedef struct rv_block_params {
    uint32_t vp_w, vp_h;                 /* viewport */
    uint32_t view_x, view_y, view_w, view_h; /* view rectangle (max excluded) */
    float center[3];                     /* cube centre, world */
    float half;                          /* cube half size */
    uint32_t level;                      /* cube level */
    uint32_t tex_depth;                  /* brick depth: 2^tex_depth+1 nodes per edge
                                            (the ray-casts take <= 9) */
    float fmin, fmax;                    /* field range */
    float pos[3], e1[3], e2[3], n[3];    /* eye */
    float w, h, pixel_size;              /* screen */
    float threshold;                     /* absolute threshold */
    float quality;                       /* sampling quality */
    float alpha_cutoff;                  /* early ray termination (reference: 0.996) */
    /* derived */
    uint32_t nodes;                      /* 2^tex_depth+1 */
    float texel;                         /* 2*half/nodes */
    float tex_origin[3];                 /* centre - (half + texel/2) */
    float tcf;                           /* nodes / (2*(half + texel/2)) */
    float i_layer_count;                 /* 100 / (2^(tex_depth+level) * quality) */
    float opacity_factor;                /* (threshold-fmin)/(fmax-fmin) */
    uint32_t flags;                      /* RV_RENDER_*; rv_block_params_setup sets 0 */
} rv_block_params;

/* Sample sequence of the Argb ray-casts.  0 (the default): the fast
 * sequence -- the reference's samples (the same positions and corners, so
 * the same undefined samples), with the table lookup, the opacity correction
 * 1-(1-a*of)^k (hardware log2/exp2) and the front-to-back compositing
 * (premultiplied, no division per sample) evaluated with fewer operations;
 * it agrees with the exact sequence to a few 1e-6 per channel.  RV_RENDER_EXACT: the reference's
 * operation order step by step (renderArgb.hpp:30-55, cuda_blend.hpp:24-44,
 * with pow01 for powf), bit-identical to the CPU restatement.  The other
 * field modes always run their exact sequence. */
#define RV_RENDER_EXACT 1u

/* Host parameter setup of DenseFieldRenderer_cu::Impl::renderDenseField
 * (DenseFieldRenderer_cu.cu:317-453): cube, camera basis, view rectangle,
 * threshold.  view16: row-major world->eye rigid transform (the reference's
 * VsRenderSharedState::cameraTransform, VsRenderSharedState.hpp:29-46).
 * *live = 0 when the view rectangle is empty (the reference returns early). */
int rv_block_params_setup(const float* view16, uint32_t vp_w, uint32_t vp_h, float fov_y_deg,
                          uint32_t level, const uint32_t index[3], uint32_t tex_depth, float fmin,
                          float fmax, float threshold_rel, float quality, float alpha_cutoff,
                          rv_block_params* out, int* live);

/* Clear a viewport to 0.  Replaces DenseFieldRenderer_cu::clearViewport
 * (DenseFieldRenderer_cu.cu:268-277) for the Argb mode. */
int rv_clear_viewport(uint32_t* d_viewport, uint32_t vp_w, uint32_t vp_h, rv_stream_t stream);

/* Ray-cast one block into the ARGB viewport (composited over what is there).
 * Replaces renderKernel<FieldMode::Argb> (DenseFieldRenderer_cu.cu:62-215,
 * renderArgb.hpp:30-55).  d_tf: tf_n x 4 floats (RGBA), sampled like the
 * reference's normalized, linear, clamped 1-D texture. */
int rv_render_block_argb(const rv_block_params* params, const float* d_brick, const float* d_tf,
                         uint32_t tf_n, uint32_t* d_viewport, rv_stream_t stream);

/* Field modes: s3vs::FieldMode (FieldMode.hpp:26-51), the same numbering.
 * All but the secondary-field modes (ValueOnIsosurface, ValueOnIsosurfaces)
 * are rendered. */
enum {
    RV_FIELD_ISOSURFACE = 0,
    RV_FIELD_ISOSURFACES = 1,
    RV_FIELD_MAX_INTENSITY_PROJECTION = 2,
    RV_FIELD_ARGB = 3,
    RV_FIELD_ARGB_LIGHT = 4,
    RV_FIELD_DOMAIN_VOXELS = 5,
    RV_FIELD_VALUE_ON_ISOSURFACE = 6,
    RV_FIELD_VALUE_ON_ISOSURFACES = 7
};

/* Ray-cast one block in a field mode.  Replaces renderKernel<fieldMode>
 * (DenseFieldRenderer_cu.cu:62-215) with renderArgb (renderArgb.hpp:30-55),
 * renderArgbLight (renderArgbLight.hpp:29-63, lighting of
 * cuda_lighting.hpp:27-95 with the light of DenseFieldRenderer_cu.cu:440-445),
 * renderIsosurface (renderIsosurface.hpp:31-61; the level is
 * params->threshold, i.e. set up with threshold_rel = the isosurface level)
 * renderIsosurfaces (renderIsosurfaces.hpp:31-103; levels_rel: level_count
 * isosurfaceLevels relative to the field range, made absolute, sorted and
 * deduplicated, at most 50 kept as makeLevels does,
 * DenseFieldRenderer_cu.cu:597-612; none: the block's pixels become 0),
 * renderDomainVoxels (renderDomainVoxels.hpp:30-54) and
 * renderMaxIntensityProjection (renderMaxIntensityProjection.hpp:31-48;
 * the viewport holds float bits, cleared to NaN by rv_clear_viewport_mode).
 * iso_opacity scales the isosurface colour's alpha (isosurfaceOpacity).
 * Every segment stops at params->alpha_cutoff (the reference: 0.996). */
int rv_render_block(const rv_block_params* params, uint32_t field_mode, float iso_opacity,
                    const float* levels_rel, uint32_t level_count, const float* d_brick,
                    const float* d_tf, uint32_t tf_n, uint32_t* d_viewport, rv_stream_t stream);
/* rv_render_block with the secondary field, which the modes
 * ValueOnIsosurface (renderValueOnIsosurface.hpp:32-62) and
 * ValueOnIsosurfaces (renderValueOnIsosurfaces.hpp:33-109) take: the
 * primary field's isosurface(s), as in Isosurface(s), each surface point
 * coloured by the table at (v2 - smin) / (smax - smin), v2 the secondary
 * field there (d_brick2: the block's brick of it, the primary brick's
 * layout; [smin, smax] its range, secondaryField.minValue/maxValue), alpha
 * times iso_opacity.  Other modes ignore d_brick2. */
int rv_render_block_secondary(const rv_block_params* params, uint32_t field_mode, float iso_opacity,
                              const float* levels_rel, uint32_t level_count, const float* d_brick,
                              const float* d_brick2, float smin, float smax, const float* d_tf,
                              uint32_t tf_n, uint32_t* d_viewport, rv_stream_t stream);

/* clearViewport of a field mode (DenseFieldRenderer_cu.cu:268-277): 0, or NaN
 * bits for MaxIntensityProjection. */
int rv_clear_viewport_mode(uint32_t* d_viewport, uint32_t vp_w, uint32_t vp_h, uint32_t field_mode,
                           rv_stream_t stream);

/* blendImagePartMaxIntensity (BlendImagePart.hpp:148-191) on device float
 * images: the larger value wins, NaN is empty; arguments as
 * rv_blend_image_part. */
int rv_blend_image_part_max_intensity(float* d_dst, int32_t dw, int32_t dh, int32_t dst_off_x,
                                      int32_t dst_off_y, const float* d_src, int32_t sw,
                                      int32_t sh, int32_t src_x, int32_t src_y, int just_copy,
                                      rv_stream_t stream);

/* makeMaxIntensityProjection (VsRenderer.cpp:542-565): the non-NaN values of
 * a float part at (src_x, src_y) become ARGB8 colours of the nctv-entry
 * colour vector d_ctv (s3vs::calcColor, ColorTransferFunction.hpp:83-94) in
 * the dw x dh frame d_dst. */
int rv_mip_colorize(uint32_t* d_dst, int32_t dw, int32_t dh, const float* d_src, int32_t sw,
                    int32_t sh, int32_t src_x, int32_t src_y, const float* d_ctv, uint32_t nctv,
                    rv_stream_t stream);

/* Adds to *d_count (a device uint64) the number of samples the block's rays
 * take without early termination: the sum over its view rectangle of
 * (sampleCount - 1) (DenseFieldRenderer_cu.cu:171-175, renderArgb.hpp:40).
 * Measurement helper: no reference counterpart. */
int rv_block_sample_count(const rv_block_params* params, uint64_t* d_count, rv_stream_t stream);

/* A whole worker share in one launch: blocks given back to front (params
 * from rv_block_params_setup, bricks as device pointers), each ray-cast into
 * the viewport with the same per-pixel operation as rv_render_block_argb, in
 * that order -- the result equals one rv_render_block_argb call per block,
 * bit for bit.  It replaces the per-block renderDenseField loop of
 * VsWorker_cu::renderScenePart (VsWorker_cu.cpp:269-277).  d_workspace:
 * caller-owned device scratch of rv_share_workspace_bytes(nblocks) bytes;
 * the call synchronizes the stream once (after staging the block table). */
uint64_t rv_share_workspace_bytes(uint32_t nblocks);
int rv_render_share_argb(const rv_block_params* params, const float* const* bricks,
                         uint32_t nblocks, const float* d_tf, uint32_t tf_n, uint32_t* d_viewport,
                         uint32_t vp_w, uint32_t vp_h, void* d_workspace, rv_stream_t stream);
/* The same in a field mode of rv_render_block: the result equals one
 * rv_render_block call per block, in the given order. */
int rv_render_share(const rv_block_params* params, const float* const* bricks, uint32_t nblocks,
                    uint32_t field_mode, float iso_opacity, const float* levels_rel,
                    uint32_t level_count, const float* d_tf, uint32_t tf_n, uint32_t* d_viewport,
                    uint32_t vp_w, uint32_t vp_h, void* d_workspace, rv_stream_t stream);
/* ... and with the secondary field (bricks2: per block; the modes of
 * rv_render_block_secondary), bricks of either layout (RV_BRICK_*; the
 * secondary bricks in the same one). */
int rv_render_share_layout(const rv_block_params* params, const float* const* bricks,
                           const float* const* bricks2, uint32_t nblocks, uint32_t layout,
                           uint32_t field_mode, float iso_opacity, const float* levels_rel,
                           uint32_t level_count, float smin, float smax, const float* d_tf,
                           uint32_t tf_n, uint32_t* d_viewport, uint32_t vp_w, uint32_t vp_h,
                           void* d_workspace, rv_stream_t stream);
int rv_render_share_secondary(const rv_block_params* params, const float* const* bricks,
                              const float* const* bricks2, uint32_t nblocks, uint32_t field_mode,
                              float iso_opacity, const float* levels_rel, uint32_t level_count,
                              float smin, float smax, const float* d_tf, uint32_t tf_n,
                              uint32_t* d_viewport, uint32_t vp_w, uint32_t vp_h, void* d_workspace,
                              rv_stream_t stream);

/* rv_render_share_argb without a workspace and without any allocation.  No
 * reference counterpart: it is for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  The block table is staged in a fixed device array of the
 * library, one per GPU, holding at most RV_SHARE_STATIC_MAX_BLOCKS blocks
 * (803,072 bytes, kept by each GPU the library runs on until the process
 * ends); the call makes no hipMalloc/hipFree and no host heap allocation.
 * nblocks above the maximum: RV_ERR_INVALID.  Same result as
 * rv_render_share_argb, and, like it, the call synchronizes the stream once
 * after staging the table; the launch then reads the array until it ends.
 * Every call on a GPU uses that GPU's one array: calls that overlap on one
 * GPU (other threads, or other streams before this launch has finished) race
 * and give wrong images; calls in order on one stream are safe. */
#define RV_SHARE_STATIC_MAX_BLOCKS 4096
int rv_render_share_argb_noalloc(const rv_block_params* params, const float* const* bricks,
                                 uint32_t nblocks, const float* d_tf, uint32_t tf_n,
                                 uint32_t* d_viewport, uint32_t vp_w, uint32_t vp_h,
                                 rv_stream_t stream);

/* Measurement helpers (no reference counterpart): kernels that measure the
 * chip's peaks the path's kernels are held against.  Each enqueues one
 * launch of `blocks` workgroups of 256 threads on `stream` and returns its
 * work count; the caller times the launch.  d_sink: blocks*256 floats the
 * kernels write only in a case that does not occur (it keeps the work live).
 *  - valu: 8 independent non-packed fp32 fma chains per thread, iters long;
 *    *fma_count = lane fmas.
 *  - lds: ds_read_b128 of conflict-free addresses, iters per thread;
 *    *bytes = bytes read.
 *  - gather: iters dword loads per thread from d_src (16 KB, L1-resident),
 *    coalesced (scattered = 0: a wave reads 256 consecutive bytes) or
 *    scattered (1: each lane a different 64-byte sector); *loads = wave load
 *    instructions.
 *  - stream: non-temporal copy of `bytes` (a multiple of 16) from d_src to
 *    d_dst (HBM read + write). */
int rv_probe_valu(float* d_sink, uint32_t blocks, uint32_t iters, uint64_t* fma_count, rv_stream_t stream);
int rv_probe_lds(float* d_sink, uint32_t blocks, uint32_t iters, uint64_t* bytes, rv_stream_t stream);
int rv_probe_gather(const float* d_src, float* d_sink, uint32_t blocks, uint32_t iters, int scattered,
                    uint64_t* loads, rv_stream_t stream);
int rv_probe_stream(const void* d_src, void* d_dst, uint64_t bytes, uint32_t blocks, rv_stream_t stream);

/* Piecewise-linear colour table from control points.  Replaces
 * s3vs::makeColorTransferVector (ColorTransferFunction.hpp:27-69).
 * xs: count sorted positions, colors: count x 4, out: color_count x 4. */
int rv_color_transfer_vector(const float* xs, const float* colors, uint32_t count,
                             uint32_t color_count, float* out);

/* ------------------------------------------------------------------------
 * Block ordering (host)
 * ------------------------------------------------------------------------ */

/* Back-to-front order of the blocks of an index box.  Replaces
 * BackToFrontOrder<MacroBlock<3>>::range(eye) (BackToFrontOrder.hpp:61-359,
 * MacroBlock.hpp:29-67).  out: cap x 3 indices. */
int rv_back_to_front(const uint32_t imin[3], const uint32_t imax[3], const float cmin[3],
                     const float cmax[3], const float eye[3], uint32_t* out, uint64_t cap,
                     uint64_t* count);

/* Eye position of a camera transform.  Replaces s3vs::eyeFromTransform
 * (BlockOrderUtil.hpp:54-59). */
int rv_eye_from_view(const float* view16, float eye[3]);

/* World box of an index box at a level.  Replaces s3vs::boundingBoxOfIndexBox
 * (BlockOrderUtil.hpp:29-51). */
int rv_index_box_coords(uint32_t level, const uint32_t bmin[3], const uint32_t bmax[3],
                        float cmin[3], float cmax[3]);

/* Worker boxes of a level, sorted back to front.  Replaces
 * IndexCubeSplitStorage::getBoxesSorted (IndexCubeSplitStorage.hpp:52-99) over
 * IndexCubeSplitter<3> (IndexCubeSplitter.hpp:35-196).  boxes: cap x 6
 * {min x,y,z, max x,y,z}. */
int rv_index_cube_split_sorted(uint32_t level, uint32_t split_count, const float eye[3],
                               uint32_t* boxes, uint64_t cap, uint64_t* count);

/* Back-to-front order of a set of disjoint blocks of mixed levels.  No
 * reference counterpart (the reference renders one level per frame,
 * VsRenderer.cpp:1067-1083); reduces to rv_back_to_front when all blocks share
 * one level.  blocks: n x 4 {level, x, y, z}; order_out: n indices into blocks. */
int rv_mixed_back_to_front(const uint32_t* blocks, uint64_t n, const float eye[3],
                           uint32_t* order_out);

/* ------------------------------------------------------------------------
 * Image parts: crop and blend (device)
 * ------------------------------------------------------------------------ */

/* Bounding rectangle of pixels with non-zero alpha; synchronous (waits for
 * the stream).  Replaces getBoundRect (VsWorker_cu.cpp:51-88).  rect:
 * {x, y, w, h}; *nonempty = 0 if all pixels are transparent. */
int rv_bound_rect(const uint32_t* d_pixels, uint32_t w, uint32_t h, int32_t rect[4],
                  int* nonempty, rv_stream_t stream);

/* Blend an image part into a destination image, 8-bit integer "over".
 * Replaces s3vs::blendImagePart (BlendImagePart.hpp:61-145).  The part
 * (sw x sh at src_x, src_y) must lie inside dst (dw x dh at dst_off). */
int rv_blend_image_part(uint32_t* d_dst, int32_t dw, int32_t dh, int32_t dst_off_x,
                        int32_t dst_off_y, const uint32_t* d_src, int32_t sw, int32_t sh,
                        int32_t src_x, int32_t src_y, int just_copy, rv_stream_t stream);

/* Fill n uint32 with a value (background fill of FinalComposeFunc,
 * VsRenderer.cpp:475-519). */
int rv_fill_u32(uint32_t* d_dst, uint64_t n, uint32_t value, rv_stream_t stream);

/* floatRgbToUint32Argb (VsRenderer.cpp:115-122). */
uint32_t rv_background_argb(float r, float g, float b);


/* ------------------------------------------------------------------------
 * Batched persistent frame (device)
 * ------------------------------------------------------------------------
 * No single reference counterpart: it replaces the per-block loop of
 * VsWorker_cu::renderScenePart (VsWorker_cu.cpp:254-277) -- one
 * DenseFieldRenderer_cu::renderDenseField call, i.e. one interpolation and one
 * renderKernel launch, per block -- by one launch per frame over all blocks of
 * one GPU's share, of uneven levels, composited front to back in registers
 * (blendBehind, cuda_blend.hpp:24-33) instead of through the 8-bit viewport.
 * The screen is cut into bins of one wave footprint (tile_w x tile_h = 64
 * pixels); per camera, each bin lists the blocks whose view rectangles meet
 * it, front to back (rv_mixed_back_to_front order), and one wave composites
 * one bin.  The handle owns device copies of the block table, per-camera
 * parameters, bin lists, sync counters and an optional captured graph;
 * bricks, sparse values, plans, colour table and outputs stay caller-owned. */

typedef struct rv_frame_block {
    uint32_t level;          /* block (subtree root) level */
    uint32_t index[3];       /* block index at that level */
    uint32_t depth;          /* brick depth */
    const float* sparse;     /* sparse values (for per-frame interpolation), may be NULL */
    const void* plan;        /* gather plan (rv_interp_plan_build), may be NULL */
    float* brick;            /* fp32 brick, or */
    uint16_t* brick16;       /* fp16 brick; one kind for all blocks of a plan */
} rv_frame_block;

typedef struct rv_frame_desc {
    uint32_t vp_w, vp_h;
    float fov_y_deg;
    float fmin, fmax, threshold_rel, quality;
    float alpha_cutoff;      /* per-block segment early termination (reference 0.996) */
    float acc_cutoff;        /* stop compositing a pixel once its alpha exceeds this */
    const float* d_tf;       /* tf_n x 4 colour table on the device */
    uint32_t tf_n;           /* 1..1024 */
    uint32_t tile_w, tile_h; /* wave footprint of a screen bin: tile_w*tile_h == 64,
                                tile_w a power of two; 0,0 = 8x8 */
    uint32_t layout;         /* RV_BRICK_* of the blocks' (and secondary) bricks; TILED4 is
                                fp32 only; field gradient arrays stay LINEAR */
    uint32_t exact;          /* nonzero: the exact Argb sample sequence (RV_RENDER_EXACT); frames
                                recording the adjoint's tape always use it */
} rv_frame_desc;

typedef struct rv_frame_plan rv_frame_plan;

int rv_frame_plan_create(const rv_frame_block* blocks, uint32_t nblocks, const rv_frame_desc* desc,
                         rv_frame_plan** out);
/* Cameras (row-major world->eye transforms, ncams x 16 floats): per camera the
 * host computes every block's parameters (rv_block_params_setup), the
 * front-to-back order and the bin lists (on up to 16 threads), and uploads
 * them.  Synchronous. */
int rv_frame_plan_set_cameras(rv_frame_plan* plan, const float* views16, uint32_t ncams);
/* One frame for camera `cam`: a counter-reset kernel and one persistent
 * launch.  interpolate != 0 rebuilds every brick from its sparse values
 * within the same launch.  d_argb: vp_w*vp_h ARGB8 (rows bottom-up, like the
 * reference viewport); d_rgba: optional vp_w*vp_h*4 float output. */
int rv_frame_render(rv_frame_plan* plan, uint32_t cam, uint32_t* d_argb, float* d_rgba,
                    int interpolate, rv_stream_t stream);
/* Capture the frames of all cameras into one HIP graph (camera c writes
 * d_argb + c*frame_stride) and replay it with rv_frame_graph_launch. */
int rv_frame_graph_build(rv_frame_plan* plan, uint32_t* d_argb, uint64_t frame_stride,
                         int interpolate, rv_stream_t stream);
int rv_frame_graph_launch(rv_frame_plan* plan, rv_stream_t stream);
/* Sticky error flags of the last frames (1: a brick wait gave up); synchronous. */
int rv_frame_status(rv_frame_plan* plan, uint32_t* error_flags);
int rv_frame_get_tile(const rv_frame_plan* plan, uint32_t* tile_w, uint32_t* tile_h);
/* How a wave walks the blocks of its screen bin.  BINS: the wave marches
 * each block of the bin's list with the lanes whose pixels the block
 * covers.  LANES (Argb mode): each lane walks its own pixel's blocks, so a
 * wave marches different blocks at once -- faster when blocks are a few
 * pixels wide.  HYBRID (Argb mode): LANES for bins listing at least
 * rv_frame_set_hybrid_min blocks (default 16), BINS for the others.  All
 * give identical images (and adjoint records). */
enum { RV_FRAME_SCHEDULE_BINS = 0, RV_FRAME_SCHEDULE_LANES = 1, RV_FRAME_SCHEDULE_HYBRID = 2 };
int rv_frame_set_schedule(rv_frame_plan* plan, uint32_t schedule);
int rv_frame_get_schedule(const rv_frame_plan* plan, uint32_t* schedule);
int rv_frame_set_hybrid_min(rv_frame_plan* plan, uint32_t blocks);
/* How the resident-brick kernel hands out its screen bins.  0 (the
 * default): a persistent grid whose waves take the next free bin from a
 * queue.  1: one workgroup per 2 x 2 neighbouring bins, one bin per wave,
 * so the waves that share a CU march neighbouring pixels through the same
 * bricks; the grid's tuned workgroups per CU become a cap (LDS padding).
 * Same images; rv_frame_autotune with interpolate == 0 picks the faster.
 * No counterpart in the reference (its renderer launches one grid per
 * block). */
int rv_frame_set_grouped(rv_frame_plan* plan, int grouped);
/* Interpolating frames as one launch (0: the bricks' tiles and the ray-cast
 * in one persistent kernel that waits for them) or two (1: a launch that
 * interpolates the bricks, then the resident-brick kernel with its own grid).
 * Same images; rv_frame_autotune with interpolate != 0 picks the faster. */
int rv_frame_set_split(rv_frame_plan* plan, int split);
/* The Argb sample sequence of later frames: nonzero the exact one
 * (RV_RENDER_EXACT), 0 the fast one (rv_frame_desc.exact at creation).
 * Frames that record the adjoint's tape use the exact one either way. */
int rv_frame_set_exact(rv_frame_plan* plan, int exact);
int rv_frame_get_split(const rv_frame_plan* plan, int* split);
/* Pick the bins' wave footprint (8x8, 16x4, 4x16, 32x2), the persistent
 * grid (resident workgroups per CU) of the mode's kernel (interpolate or
 * resident bricks) and, in Argb mode, the schedule, by timing candidates
 * on camera 0; the choice is cached (one JSON object per line) in
 * cache_path when it is not NULL.  The footprint and schedule apply to
 * both modes, the grid to the tuned mode.  With interpolate != 0 it also
 * times the chosen frame split in two launches (rv_frame_set_split; the
 * resident-brick kernel keeps its grid, so tune that mode first). */
int rv_frame_autotune(rv_frame_plan* plan, uint32_t* d_argb, int interpolate,
                      const char* cache_path, rv_stream_t stream, uint32_t* tile_w,
                      uint32_t* tile_h);
/* Field mode of later frames (default RV_FIELD_ARGB; levels as in
 * rv_render_block): Argb, ArgbLight, Isosurface, Isosurfaces, DomainVoxels
 * and (with rv_frame_set_secondary) ValueOnIsosurface(s) segments are
 * composited front to back (blendBehind); in
 * MaxIntensityProjection the frame's d_argb receives the float bits of the
 * largest normalised value over the blocks (NaN: none), the order of blocks
 * being immaterial.  The tape and the adjoint are Argb only. */
int rv_frame_set_field_mode(rv_frame_plan* plan, uint32_t field_mode, float iso_opacity,
                            const float* levels_rel, uint32_t level_count);
/* The secondary field of the ValueOnIsosurface(s) modes (set before
 * selecting them): per block its brick (the primary bricks' type and
 * layout) and, for interpolating frames, its sparse values (NULL: resident
 * frames only), and the field's range [smin, smax].  d_bricks2 = NULL
 * removes it. */
int rv_frame_set_secondary(rv_frame_plan* plan, const float* const* d_sparse2, void* const* d_bricks2,
                           float smin, float smax);
void rv_frame_plan_destroy(rv_frame_plan* plan);

/* Adjoint ray-cast (no reference counterpart: the reference renders only).
 * rv_frame_set_tape(plan, 1) makes every later rv_frame_render with a float
 * image record, per pixel and per block its ray reaches, the pixel's state in
 * front of the block (premultiplied colour and transmittance, 16 bytes per
 * pixel of each block's view rectangle; the plan owns the buffer, sized for
 * the largest camera).  rv_frame_adjoint then gives gradients of an image
 * loss L with respect to the colour table and the dense field values for the
 * camera last rendered that way: given d_rgba (that forward image, vp_w*vp_h
 * float4) and d_grad_rgba (dL/d image, same layout), it adds dL/d table to
 * d_grad_tf (tf_n x 4 floats) and dL/d field to the per-block gradient arrays
 * set with rv_frame_set_grad_bricks (fp32, RV_BRICK_LINEAR whatever the
 * bricks' layout; not set = colour table only).  Samples are recomputed per ray segment, none are stored.
 * Outputs accumulate: the caller zeroes them.  Calls on one plan must not
 * overlap. */
int rv_frame_set_tape(rv_frame_plan* plan, int enable);
int rv_frame_set_grad_bricks(rv_frame_plan* plan, float* const* d_grad_bricks);
int rv_frame_adjoint(rv_frame_plan* plan, uint32_t cam, const float* d_rgba,
                     const float* d_grad_rgba, float* d_grad_tf, rv_stream_t stream);

/* ------------------------------------------------------------------------
 * Worker (host + device): the VsWorkerInterface drop-in behind a C ABI
 * ------------------------------------------------------------------------
 * rv_worker bundles a field service (subtrees by level/index with their n2i
 * and sparse values, copied to the device; gather plans cached per subtree),
 * the render shared state, and a VsWorkerHip (include/revise/vs_worker_hip.hpp).
 * It owns its device buffers and one HIP stream. */

typedef struct rv_worker rv_worker;

typedef struct rv_render_state {
    const float* view16;     /* row-major world->eye transform (cameraTransform) */
    uint32_t vp_w, vp_h;     /* viewportSize */
    float fov_y_deg;         /* fovY */
    float quality;           /* renderQuality */
    float threshold_rel;     /* threshold, relative to the field range */
    float alpha_cutoff;      /* early ray termination (reference 0.996) */
    float fmin, fmax;        /* field range (BlockTreeFieldService::fieldRange) */
    float background[3];     /* backgroundColor */
    const float* tf;         /* host colour table, tf_n x 4 */
    uint32_t tf_n;
    uint32_t field_mode;     /* RV_FIELD_* (fieldMode); note 0 is Isosurface */
    float iso_level_rel;     /* isosurfaceLevel, relative to the field range */
    float iso_opacity;       /* isosurfaceOpacity */
    const float* iso_levels_rel; /* isosurfaceLevels (Isosurfaces), relative */
    uint32_t iso_level_count;
    float smin, smax;        /* the secondary field's range (ValueOnIsosurface(s)) */
    uint32_t exact;          /* nonzero: the exact Argb sample sequence (RV_RENDER_EXACT) */
} rv_render_state;

/* VsWorkerInterface::initialize: selects GPU thread_id % device count
 * (VsWorker_cu.cpp:96-104, selectGpu.cu:6-9). */
int rv_worker_create(uint32_t thread_id, rv_worker** out);
/* Register a subtree root of the field (BlockTreeFieldProvider data of one
 * subtree, BlockTreeFieldProvider.hpp:156-168): n2i (n x 3) and sparse values
 * are host arrays, copied. */
int rv_worker_add_subtree(rv_worker* w, uint32_t level, const uint32_t index[3], uint32_t depth,
                          const uint16_t* n2i, uint64_t n, const float* sparse);
/* The secondary field's values of a registered subtree (same n2i order and
 * count), which the ValueOnIsosurface(s) modes colour by: the reference's
 * field service values of secondaryFieldIndex (DenseFieldRenderer_cu.cu:
 * 288-307).  Copied. */
int rv_worker_add_subtree_secondary(rv_worker* w, uint32_t level, const uint32_t index[3],
                                    const float* sparse);
/* setRenderSharedState + updateState (VsWorkerInterface.hpp:38-111). */
int rv_worker_set_state(rv_worker* w, const rv_render_state* state);
/* renderScenePart (VsWorker_cu.cpp:225-313): clear, BackToFrontOrder over the
 * index box, interpolate + ray-cast each block, crop to the non-transparent
 * rectangle.  rect = {x, y, w, h} (w = h = 0 for an empty part); bits
 * receives w*h ARGB pixels (cap pixels available; a larger part is
 * RV_ERR_INVALID and nothing is written).  Pinned host memory receives the
 * part from the GPU directly (one synchronisation); pageable memory through
 * a device buffer and a copy.  *cancel is polled between blocks; when set
 * the call returns RV_ERR_CANCELLED with an empty part.  An absent block
 * gives RV_ERR_INVALID ("Block does not exist"). */
int rv_worker_render_scene_part(rv_worker* w, uint32_t level, const uint32_t box_min[3],
                                const uint32_t box_max[3], const volatile int* cancel,
                                int32_t rect[4], uint32_t* bits, uint64_t cap);
/* One level's frame with split_count worker boxes on this GPU: boxes from
 * rv_index_cube_split_sorted, one part each, assembled pairwise on the device
 * (AssembleFunc, VsRenderer.cpp:403-473, tree of RenderGraphGenerator.cpp:37-129)
 * and composed over the background (FinalComposeFunc, VsRenderer.cpp:475-519).
 * frame: vp_w*vp_h ARGB. */
int rv_worker_compose_frame(rv_worker* w, uint32_t level, uint32_t split_count, uint32_t* frame,
                            const volatile int* cancel);
/* The block a frame output shared memory holds (VsControllerFrameOutputRW,
 * VsControllerFrameOutputRW.hpp:41-75): this header
 * (VsControllerFrameOutputHeader, VsControllerFrameOutputHeader.hpp:30-50;
 * 24 bytes), then vp_w*vp_h ARGB pixels. */
typedef struct rv_frame_output_header {
    uint64_t frame_number; /* ~0: no frame */
    uint64_t duration_ms;  /* rendering time of the frame (when measured) */
    uint32_t level;        /* level of detail of the frame */
    uint32_t pad;
} rv_frame_output_header;
/* VsRenderer::continueRendering (VsRenderer.cpp:1061-1166): the progressive
 * level-of-detail loop.  For each level from level_min to level_max while
 * *cancel is 0, the level's frame (rv_worker_compose_frame with split_count
 * boxes) is written to `out` (writeFrame): the header with *frame_number
 * (then incremented), the level's wall time in ms if measure_time, and the
 * level, followed by the pixels.  A level cancelled midway writes nothing.
 * out_size must hold the header and vp_w*vp_h*4 bytes.  The writes are
 * plain copies: a caller sharing `out` between processes provides the lock
 * (the reference's SharedMemory write accessor).  *frames_written: frames
 * written. */
int rv_worker_render_progressive(rv_worker* w, uint32_t level_min, uint32_t level_max,
                                 uint32_t split_count, const volatile int* cancel, void* out,
                                 uint64_t out_size, uint64_t* frame_number, int measure_time,
                                 uint32_t* frames_written);
/* readFrame (VsControllerFrameOutputRW.hpp:41-60) of such a block; pixels
 * may be NULL (header only). */
int rv_frame_output_read(const void* in, uint64_t in_size, uint32_t vp_w, uint32_t vp_h,
                         rv_frame_output_header* header, uint32_t* pixels);
/* fused != 0 (default): one batched interpolation + one share launch per
 * renderScenePart (bit-identical to the per-block loop); 0: one
 * interpolation + one ray-cast launch per block, like the reference. */
int rv_worker_set_fused(rv_worker* w, int fused);
void rv_worker_destroy(rv_worker* w);

/* ------------------------------------------------------------------------
 * Dataset files (host only: these calls make no device call)
 * ------------------------------------------------------------------------
 * A dataset is <base>.s3dmm-meta (the block tree and its subtree roots,
 * Metadata<3>, Metadata.hpp:182-214, 401-499), one <base>.s3dmm-field#<name>
 * per field (per subtree: float min, max, then its values in n2i order,
 * BlockTreeFieldProvider.hpp:124-168, 216-262) and the text file
 * <base>.s3dmm-fields (BlockTreeFieldService.hpp:177-228).  Time-dependent
 * datasets use <dir>/<stem>/<stem>_tNNNNNN<ext> as the base of frame NNNNNN
 * (filename_util.hpp:36-48). */

typedef struct rv_dataset rv_dataset;

typedef struct rv_dataset_info {
    uint32_t format_version;    /* 1, or 0 for files without the header */
    uint32_t max_subtree_depth; /* maxSubtreeDepth */
    uint32_t max_full_level;    /* maxFullLevel (0 in version 0) */
    uint32_t level_count;       /* levels of the subtree maps */
    float bc_min[3], bc_size;   /* bounding cube of the block tree */
    uint64_t node_count;        /* blocks in the tree */
    uint64_t subtree_count;
    uint64_t field_file_size;   /* bytes of each field file */
} rv_dataset_info;

typedef struct rv_subtree_info {
    uint32_t level, location[3]; /* the root block */
    uint32_t node;               /* subtreeBlockTreePos: the root's node index in the tree */
    uint32_t depth;              /* subtreeDepth (maxSubtreeDepth in version 0) */
    uint64_t values_pos;         /* subtreeValuesPos: byte offset in the field files */
} rv_subtree_info;

typedef struct rv_field_info {
    char name[64];
    float min, max;             /* fieldRange */
    int32_t animated;
} rv_field_info;

/* Metadata<3>(std::istream&) (Metadata.hpp:182-214): reads the whole file and
 * checks it (header, tree node count, every subtree root present in the
 * tree).  RV_ERR_INVALID for a malformed or truncated file. */
int rv_dataset_open(const char* meta_path, rv_dataset** out);
void rv_dataset_close(rv_dataset* ds);
int rv_dataset_get_info(const rv_dataset* ds, rv_dataset_info* info);
/* Subtree i, in the order of their values in the field files: the order of
 * Metadata::generate's walkSubtrees (Metadata.hpp:408-433). */
int rv_dataset_subtree(const rv_dataset* ds, uint64_t i, rv_subtree_info* out);
/* subtreeData(level, blockIndex) (Metadata.hpp:242-248): the subtree whose
 * root is this block; RV_ERR_INVALID if none. */
int rv_dataset_find_subtree(const rv_dataset* ds, uint32_t level, const uint32_t loc[3], uint64_t* i);
/* closestSubtree (Metadata.hpp:216-233): the subtree containing the block,
 * found by climbing from (level, block) to the root. */
int rv_dataset_closest_subtree(const rv_dataset* ds, uint32_t level, const uint32_t block[3],
                               uint64_t* i);
/* The subtree's blocks down to its depth, as the children table that
 * rv_subtree_n2i and rv_subtree_n2i_device take (nnodes x 8, node 0 = the
 * root; BlockTreeNodes(blockTree, root, subtreeDepth), Metadata.hpp:254-267).
 * *nnodes is set also when cap_nodes is too small (then RV_ERR_INVALID). */
int rv_dataset_subtree_children(const rv_dataset* ds, uint64_t i, uint32_t* children,
                                uint64_t cap_nodes, uint64_t* nnodes);
/* BlockTreeFieldProvider::fieldValues (BlockTreeFieldProvider.hpp:156-168):
 * range = the subtree's (min, max), values = n sparse values, n = the n2i
 * count of the subtree. */
int rv_dataset_read_field(const rv_dataset* ds, const char* field_path, uint64_t i, float range[2],
                          float* values, uint64_t n);
/* Register every subtree rooted at levels level_min..level_max with the
 * worker (rv_dataset_subtree_children -> rv_subtree_n2i -> the field file ->
 * rv_worker_add_subtree); *loaded = subtrees registered. */
int rv_dataset_load_worker(const rv_dataset* ds, const char* field_path, rv_worker* w,
                           uint32_t level_min, uint32_t level_max, uint64_t* loaded);
/* The same for a secondary field (rv_worker_add_subtree_secondary) of
 * subtrees already registered. */
int rv_dataset_load_worker_secondary(const rv_dataset* ds, const char* field_path, rv_worker* w,
                                     uint32_t level_min, uint32_t level_max, uint64_t* loaded);
/* The field info file: frame count and per field (name, min, max, animated);
 * *count is set also when cap is too small (then RV_ERR_INVALID). */
int rv_fields_info_read(const char* path, uint32_t* time_steps, rv_field_info* fields, uint32_t cap,
                        uint32_t* count);

/* ------------------------------------------------------------------------
 * Frame output in shared memory (host only)
 * ------------------------------------------------------------------------
 * SharedMemory (ipc/SharedMemory.h:30-280): a System V segment holding
 * its user size (size_t, 8 bytes), a reader/writer control word
 * (int32: readers count up while it is >= 0; a writer swaps 0 for -1000),
 * then the user bytes at offset 12.  Readers of the reference (the s3vs_js
 * addon, ws_sendframe) attach to a segment made here and the other way
 * round.  A segment is removed (IPC_RMID) when its creator closes it; a
 * process still attached keeps its mapping until it detaches. */

typedef struct rv_shm rv_shm;

/* SharedMemory::createNew(bytes): a new segment of `bytes` user bytes, the
 * control word 0, the bytes zero.  Made with IPC_PRIVATE, so it never
 * attaches to another process's segment (the reference draws a random key in
 * 5000..10000 with IPC_CREAT, SharedMemory.h:226-243). */
int rv_shm_create(uint64_t bytes, rv_shm** out);
/* SharedMemory::attachNew(shmid) (SharedMemory.h:44-46, 204-211). */
int rv_shm_attach(int shmid, rv_shm** out);
int rv_shm_id(const rv_shm* shm);
/* size() (SharedMemory.h:85-87): the user bytes. */
uint64_t rv_shm_size(const rv_shm* shm);
/* read / write (SharedMemory.h:89-99) of user bytes [start, start + count),
 * under the read or the write lock. */
int rv_shm_read(const rv_shm* shm, void* dst, uint64_t start, uint64_t count);
int rv_shm_write(rv_shm* shm, const void* src, uint64_t start, uint64_t count);
/* Detach; remove the segment if this handle created it (~SharedMemory,
 * SharedMemory.h:56-66). */
void rv_shm_close(rv_shm* shm);
/* VsControllerFrameOutputRW::writeFrame / readFrame
 * (VsControllerFrameOutputRW.hpp:41-75) on a segment: the header, then
 * frame_size bytes of pixels, under the write (read) lock.  pixels NULL:
 * the header only.  The segment must hold 24 + frame_size user bytes. */
int rv_frame_output_write_shm(rv_shm* shm, uint64_t frame_size, const rv_frame_output_header* header,
                              const void* pixels);
int rv_frame_output_read_shm(const rv_shm* shm, uint64_t frame_size, rv_frame_output_header* header,
                             void* pixels);

/* ------------------------------------------------------------------------
 * Controller (host; drives one rv_worker on a render thread of its own)
 * ------------------------------------------------------------------------
 * VsControllerInterface / VsController + VsRenderer (VsControllerInterface.hpp:
 * 36-178, VsController.cpp:42-196, VsRenderer.cpp:576-1192) for one GPU:
 * every input change cancels the level-of-detail loop in progress, waits for
 * it, applies the change and starts the loop again on the render thread
 * (updateRenderSharedState, VsRenderer.cpp:1177-1192).  Each level's
 * composed frame goes to the frame output segment (continueRendering,
 * VsRenderer.cpp:1061-1166), made anew when the viewport size changes and
 * first written with the "no frame" header (maybeInitFrameOutput,
 * VsRenderer.cpp:81-103).  A failure on the render thread stops the loop;
 * rv_controller_status reports it. */

typedef struct rv_controller rv_controller;

/* VsControllerInterface::newInstance + ComputingCaps: the render thread's
 * worker uses GPU thread_id % device count.  split_count (ComputingCaps::
 * workerThreadPerNodeCount, default 1): worker boxes per level, composed on
 * the device.  measure_time (VsControllerOpts::measureRenderingTime): the
 * frame header carries the level's wall time. */
int rv_controller_create(uint32_t thread_id, uint32_t split_count, int measure_time,
                         rv_controller** out);
/* problemPath / timeFrame / primaryField / secondaryField
 * (VsControllerInput.hpp, VsRenderer.cpp:708-758, 880-902): the dataset's
 * base name as BlockTreeFieldService takes it (BlockTreeFieldService.hpp:
 * 244-272, filename_util.hpp:36-94): <base>.s3dmm-meta etc. when <base> or
 * <base>.s3dmm-meta is a file, frame files under <dir>/<stem>/ when that is
 * a directory.  A field name the dataset lacks (or NULL) selects its first
 * field, as resolveField does (VsRenderer.cpp:57-72); secondary NULL or ""
 * = none.  The subtrees of levels 0..maxFullLevel are read and uploaded on
 * the render thread.  The camera is reset (CameraController.cpp:88-101). */
int rv_controller_set_problem(rv_controller* c, const char* problem_path, uint32_t time_frame,
                              const char* primary_field, const char* secondary_field);
/* The dataset's fields (fieldNames, fieldRange; VsController.cpp:123-130)
 * and frame count (timeFrameCount) of the current problem. */
int rv_controller_field_count(const rv_controller* c, uint32_t* count);
int rv_controller_field_info(const rv_controller* c, uint32_t i, rv_field_info* info);
int rv_controller_time_frame_count(const rv_controller* c, uint32_t* count);
/* The render inputs of VsControllerInput (viewport, fovY, quality, threshold,
 * colour table, field mode, isosurface parameters, background) and the
 * camera: view16 may be NULL to keep the camera.  fmin/fmax and smin/smax
 * are ignored: the ranges are the dataset's (fieldRange of the primary and
 * secondary fields). */
int rv_controller_set_input(rv_controller* c, const rv_render_state* state);
/* setCameraTransform / cameraTransform / resetCameraTransform
 * (VsController.cpp:100-114; CameraController.cpp:88-101): row-major
 * world->eye. */
int rv_controller_set_camera(rv_controller* c, const float* view16);
int rv_controller_camera(const rv_controller* c, float* view16);
int rv_controller_reset_camera(rv_controller* c);
/* renderLevel (VsRenderer.cpp:1075-1079): -1 renders levels 0..maxFullLevel
 * in turn; L >= 0 only min(L, maxFullLevel). */
int rv_controller_set_render_level(rv_controller* c, int32_t level);
/* start (VsController.cpp:91-93): rendering begins with the inputs set. */
int rv_controller_start(rv_controller* c);
/* frameOutput (VsController.cpp:147-151): the segment id, frame size
 * (vp_w * vp_h * 4) and frame width and height.  The segment stays valid
 * until the viewport size changes or the controller is destroyed. */
int rv_controller_frame_output(rv_controller* c, int* shmid, uint64_t* frame_size, uint32_t* width,
                               uint32_t* height);
/* No counterpart in the reference (its callers poll the frame number): waits
 * up to timeout_ms until the render thread has nothing left to do; *idle = 1
 * if so.  *frames: frames written so far. */
int rv_controller_wait_idle(rv_controller* c, uint32_t timeout_ms, int* idle, uint64_t* frames);
/* The render thread's last failure: its status (RV_OK if none) and message. */
int rv_controller_status(const rv_controller* c, char* message, uint64_t cap);
/* kill (VsController.cpp:153-157): cancels, stops the render thread, frees
 * the worker and removes the frame output segment. */
void rv_controller_destroy(rv_controller* c);

#ifdef __cplusplus
}
#endif

#endif /* REVISE_HIP_H */
