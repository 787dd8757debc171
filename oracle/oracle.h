/*
 * oracle.h -- C API of the CPU oracle.
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  The oracle is a plain CPU
 * restatement of the reference (deadmorous/revise) sparse->dense
 * interpolation and Argb ray-cast path.  Only tests/, the smoke() check and the
 * CPU-baseline leg of bench.py may load it, and only as the checker or the
 * timed CPU baseline; the product library (revise_amd) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Paths are
 * relative to the reference repository root.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- block tree / n2i ------------------------------------------------- */

/* Subtree depth = min(actual depth, max_depth).
 * CompressedBlockTree.hpp:290-300 (depthUpTo), Metadata.hpp:414-416. */
uint32_t orc_subtree_depth(const uint32_t* children, uint32_t nnodes, uint32_t max_depth);

/* n2i list of a subtree: first-visit order of block corners in a pre-order
 * DFS limited to `depth` levels.  BlockTreeNodes.hpp:66-80,166-181,
 * BlockTree.hpp:222-239.  children: nnodes x 8 child indices (0 = absent),
 * node 0 is the subtree root, child local index bit0 = x, bit1 = y, bit2 = z.
 * Returns the node count, or -1 if cap is too small / input malformed. */
int64_t orc_subtree_n2i(const uint32_t* children, uint32_t nnodes, uint32_t depth,
                        uint16_t* n2i_out, uint64_t cap);

/* vertexIndexToOrdinal: IndexTransform.hpp:250-257. */
void orc_ordinals(const uint16_t* n2i, uint64_t n, uint32_t depth, uint32_t* out);

/* Dense interpolation: DenseFieldInterpolator.hpp:198-419.
 * dense must hold (2^depth+1)^3 floats. */
int orc_interpolate(const float* sparse, const uint16_t* n2i, uint64_t n, uint32_t depth,
                    float* dense);

/* ---- ordering ----------------------------------------------------------- */

/* UniformIndexRangeSplitter::operator(): UniformIndexRangeSplitter.hpp:27-57.
 * out5 = {before0, before1, at, after0, after1} (empty range -> {0,0}). */
void orc_uniform_split(uint32_t i0, uint32_t i1, float c0, float c1, float x, uint32_t* out5);

/* NonUniformIndexRangeSplitter::operator(): NonUniformIndexRangeSplitter.hpp:31-64. */
void orc_nonuniform_split(uint32_t start, const float* split, uint32_t nsplit,
                          float c0, float c1, float x, uint32_t* out5);

/* BackToFrontOrder<MacroBlock<dim>>::range(eye): BackToFrontOrder.hpp:61-359,
 * MacroBlock.hpp:29-67.  imin/imax: index box (max exclusive), cmin/cmax:
 * coordinate box, eye: dim floats.  Writes count x dim indices.
 * Returns count or -1 if cap too small. */
int64_t orc_b2f_macro(uint32_t dim, const uint32_t* imin, const uint32_t* imax,
                      const float* cmin, const float* cmax, const float* eye,
                      uint32_t* out, uint64_t cap);

/* IndexCubeSplitStorage::getBoxesSorted: IndexCubeSplitStorage.hpp:52-99 with
 * IndexCubeSplitter<3>: IndexCubeSplitter.hpp:35-196.  bbox = [-5,5]^3.
 * Writes boxes as {min0,min1,min2,max0,max1,max2}.  Returns count. */
int64_t orc_index_cube_split_sorted(uint32_t level, uint32_t split_count, const float* eye,
                                    uint32_t* boxes_out, uint64_t cap);

/* IndexCubeSplitter<3>::all_blocks(add_trailing_empty): boxes in index order,
 * empty boxes written as all-zero.  Returns count. */
int64_t orc_index_cube_split_all(uint32_t level, uint32_t split_count, int add_trailing_empty,
                                 uint32_t* boxes_out, uint64_t cap);

/* ---- camera / per-block render parameters --------------------------------- */

typedef struct orc_block_params {
    uint32_t vp_w, vp_h;
    uint32_t view_x, view_y, view_w, view_h;
    float center[3];
    float half;
    uint32_t level;
    uint32_t tex_depth;
    float fmin, fmax;
    float pos[3], e1[3], e2[3], n[3];
    float w, h, pixel_size;
    float threshold;   /* absolute */
    float quality;
    float alpha_cutoff;
} orc_block_params;

/* Host parameter setup of DenseFieldRenderer_cu::Impl::renderDenseField,
 * DenseFieldRenderer_cu.cu:317-453.  view16: row-major world->eye rigid
 * transform (the reference cameraTransform).  Returns 0 when the view
 * rectangle is empty (nothing to render), 1 otherwise. */
int orc_block_params_setup(const float* view16, uint32_t vp_w, uint32_t vp_h, float fov_y_deg,
                           uint32_t level, const uint32_t* index, uint32_t tex_depth,
                           float fmin, float fmax, float threshold_rel, float quality,
                           float alpha_cutoff, orc_block_params* out);

/* renderKernel<Argb> over the view rectangle: DenseFieldRenderer_cu.cu:62-215,
 * renderArgb.hpp:30-55, RenderFuncInput.hpp:42-49, cuda_blend.hpp:24-44,
 * cuda_argb_transform.hpp:26-43.  brick: (2^tex_depth+1)^3 floats. */
void orc_render_block(const orc_block_params* p, const float* brick, const float* tf,
                      uint32_t tf_n, uint32_t* viewport, uint32_t row0, uint32_t row1);

/* renderKernel<fieldMode> over the view rectangle in a field mode (s3vs
 * FieldMode numbering: 0 Isosurface, 1 Isosurfaces (levels relative, made
 * absolute/sorted/unique as makeLevels), 2 MaxIntensityProjection, 3 Argb,
 * 4 ArgbLight, 5 DomainVoxels; renderIsosurfaces.hpp:31-103,
 * renderDomainVoxels.hpp:30-54): renderIsosurface.hpp:31-61, renderMaxIntensityProjection.hpp:
 * 31-48, renderArgb.hpp:30-55, renderArgbLight.hpp:29-63 with
 * cuda_lighting.hpp:27-95 and the light of DenseFieldRenderer_cu.cu:440-445.
 * MaxIntensityProjection keeps float bits in the viewport (NaN: none). */
void orc_render_block_mode(const orc_block_params* p, uint32_t mode, float iso_opacity,
                           const float* levels_rel, uint32_t level_count, const float* brick,
                           const float* tf, uint32_t tf_n, uint32_t* viewport, uint32_t row0,
                           uint32_t row1);
/* The same with the secondary field of modes 6 ValueOnIsosurface and 7
 * ValueOnIsosurfaces (brick2: the block's brick of it, [smin, smax] its range). */
void orc_render_block_mode2(const orc_block_params* p, uint32_t mode, float iso_opacity,
                            const float* levels_rel, uint32_t level_count, const float* brick,
                            const float* brick2, float smin, float smax, const float* tf,
                            uint32_t tf_n, uint32_t* viewport, uint32_t row0, uint32_t row1);

/* blendImagePartMaxIntensity (BlendImagePart.hpp:148-191) of a float part at
 * (dx, dy) of a dw-wide float image. */
void orc_blend_image_part_max_intensity(float* dst, int32_t dw, const float* src, int32_t sw,
                                        int32_t sh, int32_t dx, int32_t dy, int just_copy);

/* makeMaxIntensityProjection (VsRenderer.cpp:542-565): non-NaN values of a
 * float part become ARGB8 colours of the nctv x 4 colour vector. */
void orc_mip_colorize(uint32_t* dst, int32_t dw, const float* src, int32_t sw, int32_t sh,
                      int32_t dx, int32_t dy, const float* ctv, uint32_t nctv);

/* Per-ray float RGBA segment of one block (no viewport blend): the loop of
 * renderArgb.hpp:40-51.  Returns 1 if the pixel is covered, writes rgba[4]. */
int orc_ray_segment(const orc_block_params* p, const float* brick, const float* tf,
                    uint32_t tf_n, uint32_t vx, uint32_t vy, float* rgba);

/* Derived scalars of a block: {texel, tcf, i_layer_count, opacity_factor}. */
void orc_block_derived(const orc_block_params* p, float* out4);

/* Kept samples of one ray segment with their trilinear corners and weights
 * (test support for the adjoint's gradient checks). */
int orc_ray_samples(const orc_block_params* p, const float* brick, const float* tf,
                    uint32_t tf_n, uint32_t vx, uint32_t vy, uint32_t max_samples,
                    uint32_t* corners, float* weights, float* seg_alpha);

/* Reference pow replacement shared by the oracle and the HIP kernels (see
 * the definition for the method).  Exposed for its error-bound test. */
float orc_powf(float b, float e);

/* ---- parts, crop, blend -------------------------------------------------- */

/* getBoundRect over alpha bytes, VsWorker_cu.cpp:51-88.  Returns 0 if empty,
 * else writes {x, y, w, h}. */
int orc_bound_rect(const uint32_t* pix, uint32_t w, uint32_t h, int32_t* rect4);

/* blendImagePart, BlendImagePart.hpp:61-145.  dst: dw x dh, src part: sw x sh
 * at (sx, sy) relative to dst origin (dst_off). */
void orc_blend_image_part(uint32_t* dst, int32_t dw, int32_t dh, const uint32_t* src, int32_t sw,
                          int32_t sh, int32_t src_x, int32_t src_y, int32_t dst_off_x,
                          int32_t dst_off_y, int just_copy);

/* floatRgbToUint32Argb, VsRenderer.cpp:115-122. */
uint32_t orc_background_argb(float r, float g, float b);

/* makeColorTransferVector, ColorTransferFunction.hpp:27-69.  xs/colors:
 * count control points (sorted by x, like std::map), colors count x 4.
 * out: color_count x 4 floats. */
void orc_color_transfer_vector(const float* xs, const float* colors, uint32_t count,
                               uint32_t color_count, float* out);

/* Full back-to-front frame of one worker share (VsWorker_cu.cpp:225-313):
 * clear, BackToFrontOrder over the index box, per block render (bricks given
 * in row-major index order of the box, x fastest). viewport: vp_w*vp_h. */
typedef struct orc_frame_desc {
    const float* view16;
    uint32_t vp_w, vp_h;
    float fov_y_deg;
    uint32_t level;
    uint32_t box_min[3], box_max[3];
    uint32_t tex_depth;
    float fmin, fmax, threshold_rel, quality, alpha_cutoff;
    const float* tf;
    uint32_t tf_n;
} orc_frame_desc;

void orc_render_share(const orc_frame_desc* d, const float* const* bricks, uint32_t* viewport,
                      uint32_t nthreads);

/* Float front-to-back composite of per-block segments for a list of blocks
 * given in back-to-front order (the compositing operator of the reference,
 * evaluated without the per-block 8-bit viewport round trip).  Blocks behind a
 * pixel whose accumulated alpha exceeds acc_cutoff are skipped (acc_cutoff >=
 * 1 composites every block).
 * out_rgba: vp_w*vp_h*4 floats (non-premultiplied, like the reference). */
typedef struct orc_block_ref {
    uint32_t level;
    uint32_t index[3];
    uint32_t tex_depth;
    const float* brick;
    const float* brick2; /* secondary field brick (ValueOnIsosurface(s)), or NULL */
} orc_block_ref;

void orc_render_float_composite(const float* view16, uint32_t vp_w, uint32_t vp_h, float fov_y_deg,
                                const orc_block_ref* blocks_b2f, uint32_t nblocks, float fmin,
                                float fmax, float threshold_rel, float quality, float alpha_cutoff,
                                float acc_cutoff, const float* tf, uint32_t tf_n, float* out_rgba,
                                uint32_t nthreads);
/* The same with the secondary field range of the ValueOnIsosurface(s)
 * modes (the blocks' brick2). */
void orc_render_float_composite_mode2(const float* view16, uint32_t vp_w, uint32_t vp_h,
                                      float fov_y_deg, const orc_block_ref* blocks_b2f,
                                      uint32_t nblocks, float fmin, float fmax, float threshold_rel,
                                      float quality, float alpha_cutoff, float acc_cutoff,
                                      const float* tf, uint32_t tf_n, uint32_t mode,
                                      float iso_opacity, const float* levels_rel,
                                      uint32_t level_count, float smin, float smax, float* out_rgba,
                                      uint32_t nthreads);
/* The same in a field mode: Argb, ArgbLight and Isosurface segments
 * composited front to back; MaxIntensityProjection writes the float of the
 * largest normalised value (NaN: none) to all 4 channels. */
void orc_render_float_composite_mode(const float* view16, uint32_t vp_w, uint32_t vp_h,
                                     float fov_y_deg, const orc_block_ref* blocks_b2f,
                                     uint32_t nblocks, float fmin, float fmax, float threshold_rel,
                                     float quality, float alpha_cutoff, float acc_cutoff,
                                     const float* tf, uint32_t tf_n, uint32_t mode,
                                     float iso_opacity, const float* levels_rel,
                                     uint32_t level_count, float* out_rgba, uint32_t nthreads);

#ifdef __cplusplus
}
#endif
