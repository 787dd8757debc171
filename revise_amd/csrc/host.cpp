// Host-side parts of the path: subtree node lists (n2i), block ordering,
// per-block render parameters, colour table.  Integer results here are
// required to equal the reference's exactly; floating-point parameter setup
// uses the reference's operand order (cited per function).
#include "../../include/revise_hip.h"
#include "order.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace rv {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

inline int fail(int code, const std::string& msg)
{
    set_error(msg);
    return code;
}

} // namespace rv

using namespace rv;

#define H_REQUIRE(cond, msg)                                                           \
    do {                                                                               \
        if (!(cond))                                                                   \
            return ::rv::fail(RV_ERR_INVALID, msg);                                    \
    } while (0)

#define H_TRY try {
#define H_CATCH                                                                        \
    }                                                                                  \
    catch (const std::bad_alloc&)                                                      \
    {                                                                                  \
        return ::rv::fail(RV_ERR_OOM, "host allocation failed");                       \
    }                                                                                  \
    catch (const std::exception& ex)                                                   \
    {                                                                                  \
        return ::rv::fail(RV_ERR_INTERNAL, ex.what());                                 \
    }

extern "C" const char* rv_last_error(void) { return g_last_error.c_str(); }

extern "C" int rv_version(void) { return 10000; }

/* ------------------------------------------------------------------------ */
/* Subtree walk and n2i                                                      */
/* ------------------------------------------------------------------------ */

namespace {

// Pre-order walk of a subtree limited to max_depth levels below the root,
// children in local index order 0..7 (x = bit0): BlockTree::walk(root,
// maxDepth, cb), BlockTree.hpp:222-239.  Iterative, explicit stack.
template <class Visit>
bool walk_subtree(const uint32_t* children, uint32_t nnodes, uint32_t max_depth, Visit&& visit)
{
    struct Frame {
        uint32_t node, level, x, y, z;
    };
    std::vector<Frame> stack;
    stack.push_back({0, 0, 0, 0, 0});
    while (!stack.empty()) {
        const Frame f = stack.back();
        stack.pop_back();
        visit(f.level, f.x, f.y, f.z);
        if (f.level >= max_depth)
            continue;
        // push children in reverse so that local index 0 is visited first
        for (int k = 7; k >= 0; --k) {
            const uint32_t c = children[8 * (size_t)f.node + k];
            if (!c)
                continue;
            if (c >= nnodes)
                return false;
            stack.push_back({c, f.level + 1, (f.x << 1) | (k & 1u), (f.y << 1) | ((k >> 1) & 1u),
                             (f.z << 1) | ((k >> 2) & 1u)});
        }
    }
    return true;
}

} // namespace

extern "C" int rv_subtree_depth(const uint32_t* children, uint32_t nnodes, uint32_t max_depth,
                                uint32_t* depth_out)
{
    H_TRY
    H_REQUIRE(children && nnodes && depth_out, "rv_subtree_depth: bad arguments");
    uint32_t best = 0;
    bool ok = walk_subtree(children, nnodes, max_depth,
                           [&](uint32_t lvl, uint32_t, uint32_t, uint32_t) { best = std::max(best, lvl); });
    H_REQUIRE(ok, "rv_subtree_depth: child index out of range");
    *depth_out = best;
    return RV_OK;
    H_CATCH
}

extern "C" int rv_subtree_n2i(const uint32_t* children, uint32_t nnodes, uint32_t depth,
                              uint16_t* n2i, uint64_t cap, uint64_t* count)
{
    H_TRY
    H_REQUIRE(children && nnodes && count, "rv_subtree_n2i: bad arguments");
    H_REQUIRE(depth <= 15, "rv_subtree_n2i: depth must be <= 15");
    const uint32_t npe = (1u << depth) + 1u;
    const uint32_t mask = (1u << depth) - 1u;
    // BlockTreeNodes: each visited block contributes its 8 corners (local
    // order, x = bit0) at ((location << shift) & mask) + (bit << shift);
    // a corner is appended the first time it is seen (BlockTreeNodes.hpp:66-80,
    // 166-181).  Seen-set: dense bitmap for bricks up to 257^3, hash above.
    const bool dense = depth <= 8;
    std::vector<uint64_t> seen_bits;
    std::unordered_set<uint64_t> seen_hash;
    if (dense)
        seen_bits.assign(((uint64_t)npe * npe * npe + 63) / 64, 0ull);
    uint64_t n = 0;
    bool overflow = false;
    bool ok = walk_subtree(children, nnodes, depth, [&](uint32_t lvl, uint32_t x, uint32_t y, uint32_t z) {
        const uint32_t shift = depth - lvl;
        const uint32_t bx = (x << shift) & mask, by = (y << shift) & mask, bz = (z << shift) & mask;
        for (uint32_t k = 0; k < 8; ++k) {
            const uint32_t cx = bx + ((k & 1u) << shift);
            const uint32_t cy = by + (((k >> 1) & 1u) << shift);
            const uint32_t cz = bz + (((k >> 2) & 1u) << shift);
            const uint64_t key = ((uint64_t)cz * npe + cy) * npe + cx;
            bool fresh;
            if (dense) {
                uint64_t& w = seen_bits[key >> 6];
                const uint64_t bit = 1ull << (key & 63);
                fresh = !(w & bit);
                w |= bit;
            } else {
                fresh = seen_hash.insert(key).second;
            }
            if (!fresh)
                continue;
            if (n < cap) {
                n2i[3 * n] = (uint16_t)cx;
                n2i[3 * n + 1] = (uint16_t)cy;
                n2i[3 * n + 2] = (uint16_t)cz;
            } else {
                overflow = true;
            }
            ++n;
        }
    });
    H_REQUIRE(ok, "rv_subtree_n2i: child index out of range");
    *count = n;
    H_REQUIRE(!overflow, "rv_subtree_n2i: output capacity too small");
    return RV_OK;
    H_CATCH
}

/* ------------------------------------------------------------------------ */
/* Ordering                                                                  */
/* ------------------------------------------------------------------------ */

extern "C" int rv_back_to_front(const uint32_t imin[3], const uint32_t imax[3], const float cmin[3],
                                const float cmax[3], const float eye[3], uint32_t* out,
                                uint64_t cap, uint64_t* count)
{
    H_TRY
    H_REQUIRE(imin && imax && cmin && cmax && eye && count, "rv_back_to_front: null pointer");
    for (int a = 0; a < 3; ++a) {
        H_REQUIRE(imin[a] < imax[a], "rv_back_to_front: index range must be non-empty");
        H_REQUIRE(cmin[a] < cmax[a], "rv_back_to_front: coordinate range must be non-empty");
    }
    AxisSplit s[3];
    for (int a = 0; a < 3; ++a)
        s[a] = UniformSplitter(imin[a], imax[a], cmin[a], cmax[a]).split(eye[a]);
    uint64_t n = 0;
    BackToFront3 it(s);
    uint32_t p[3];
    while (it.next(p)) {
        if (n < cap && out) {
            out[3 * n] = p[0];
            out[3 * n + 1] = p[1];
            out[3 * n + 2] = p[2];
        }
        ++n;
    }
    *count = n;
    H_REQUIRE(n <= cap, "rv_back_to_front: output capacity too small");
    return RV_OK;
    H_CATCH
}

extern "C" int rv_eye_from_view(const float* M, float eye[3])
{
    H_REQUIRE(M && eye, "rv_eye_from_view: null pointer");
    eye_from_view(M, eye);
    return RV_OK;
}

extern "C" int rv_index_box_coords(uint32_t level, const uint32_t bmin[3], const uint32_t bmax[3],
                                   float cmin[3], float cmax[3])
{
    H_REQUIRE(bmin && bmax && cmin && cmax, "rv_index_box_coords: null pointer");
    H_REQUIRE(level <= 20, "rv_index_box_coords: level too large");
    for (int a = 0; a < 3; ++a)
        H_REQUIRE(bmin[a] < bmax[a], "rv_index_box_coords: empty box");
    index_box_coords(level, bmin, bmax, cmin, cmax);
    return RV_OK;
}

extern "C" int rv_index_cube_split_sorted(uint32_t level, uint32_t split_count, const float eye[3],
                                          uint32_t* boxes, uint64_t cap, uint64_t* count)
{
    H_TRY
    H_REQUIRE(eye && count && split_count > 0, "rv_index_cube_split_sorted: bad arguments");
    H_REQUIRE(level <= 10, "rv_index_cube_split_sorted: level too large");
    CubeSplit cs(level, split_count);
    AxisSplit s[3];
    for (int a = 0; a < 3; ++a)
        s[a] = cs.axis[a].split(eye[a]);
    BackToFront3 it(s);
    uint32_t p[3];
    uint64_t n = 0;
    while (it.next(p)) {
        if (n < cap && boxes)
            cs.box(p, boxes + 6 * n);
        ++n;
    }
    *count = n;
    H_REQUIRE(n <= cap, "rv_index_cube_split_sorted: output capacity too small");
    return RV_OK;
    H_CATCH
}

extern "C" int rv_mixed_back_to_front(const uint32_t* blocks, uint64_t n, const float eye[3],
                                      uint32_t* order_out)
{
    H_TRY
    H_REQUIRE((blocks && order_out) || n == 0, "rv_mixed_back_to_front: null pointer");
    H_REQUIRE(eye, "rv_mixed_back_to_front: null eye");
    std::vector<uint32_t> order;
    int rc = mixed_back_to_front(blocks, n, eye, order);
    if (rc)
        return rc;
    std::copy(order.begin(), order.end(), order_out);
    return RV_OK;
    H_CATCH
}

/* ------------------------------------------------------------------------ */
/* Per-block parameters: DenseFieldRenderer_cu.cu:317-453                    */
/* ------------------------------------------------------------------------ */

extern "C" int rv_block_params_setup(const float* M, uint32_t vp_w, uint32_t vp_h, float fov_y_deg,
                                     uint32_t level, const uint32_t index[3], uint32_t tex_depth,
                                     float fmin, float fmax, float threshold_rel, float quality,
                                     float alpha_cutoff, rv_block_params* out, int* live)
{
    H_REQUIRE(M && index && out && live, "rv_block_params_setup: null pointer");
    H_REQUIRE(vp_w > 0 && vp_h > 0 && vp_w <= 32768 && vp_h <= 32768,
              "rv_block_params_setup: bad viewport size");
    H_REQUIRE(level <= 20 && tex_depth <= 10, "rv_block_params_setup: level/depth too large");
    H_REQUIRE(fmax > fmin, "rv_block_params_setup: empty field range");
    H_REQUIRE(quality > 0.f, "rv_block_params_setup: quality must be positive");
    for (int a = 0; a < 3; ++a)
        H_REQUIRE(index[a] < (1u << level), "rv_block_params_setup: index outside level");
    *live = block_params(M, vp_w, vp_h, fov_y_deg, level, index, tex_depth, fmin, fmax,
                         threshold_rel, quality, alpha_cutoff, *out);
    return RV_OK;
}

extern "C" int rv_color_transfer_vector(const float* xs, const float* colors, uint32_t count,
                                        uint32_t color_count, float* out)
{
    H_REQUIRE(out && color_count > 0, "rv_color_transfer_vector: bad arguments");
    H_REQUIRE(count == 0 || (xs && colors), "rv_color_transfer_vector: null control points");
    for (uint32_t i = 1; i < count; ++i)
        H_REQUIRE(xs[i] > xs[i - 1], "rv_color_transfer_vector: xs must increase");
    // makeColorTransferVector, ColorTransferFunction.hpp:27-69
    if (count == 0) {
        std::fill(out, out + 4 * (size_t)color_count, 1.f);
        return RV_OK;
    }
    if (count == 1) {
        for (uint32_t i = 0; i < color_count; ++i)
            std::memcpy(out + 4 * (size_t)i, colors, 16);
        return RV_OK;
    }
    std::fill(out, out + 4 * (size_t)color_count, 0.f);
    const float xmin = xs[0], xmax = xs[count - 1];
    float prev[4];
    std::memcpy(prev, colors, 16);
    std::memcpy(out, colors, 16);
    uint32_t i1 = 0;
    for (uint32_t it = 1; it < count; ++it) {
        const float* next = colors + 4 * (size_t)it;
        const uint32_t i2 = it + 1 == count
                                ? color_count - 1u
                                : (uint32_t)((xs[it] - xmin) / (xmax - xmin) * color_count);
        for (uint32_t i = i1 + 1; i2 > i1 && i <= i2 && i < color_count; ++i) {
            const float t = (float)(i - i1) / (i2 - i1);
            for (int c = 0; c < 4; ++c)
                out[4 * (size_t)i + c] = prev[c] * (1 - t) + next[c] * t;
        }
        std::memcpy(prev, next, 16);
        i1 = i2;
    }
    return RV_OK;
}

extern "C" uint32_t rv_background_argb(float r, float g, float b)
{
    // floatRgbToUint32Argb, VsRenderer.cpp:115-122
    return 0xff000000u + ((uint32_t)std::lround(r * 255) << 16) +
           ((uint32_t)std::lround(g * 255) << 8) + (uint32_t)std::lround(b * 255);
}
