// Camera basis, per-block render parameters, worker-box split and the
// mixed-level painter's order.
#include "order.hpp"

#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>

namespace rv {

void set_error(const std::string& msg);

namespace {

constexpr float kTopHalf = 5.f;                // TopLevelCubeHalfSize, DenseFieldRenderer_cu.cu:339
constexpr float kLcam = 0.01f * 2 * kTopHalf;  // DenseFieldRenderer_cu.cu:353
constexpr double kPi = 3.14159265358979323846;

inline float cube_half(uint32_t level) { return kTopHalf / (int)(1u << level); }
inline float cube_center(uint32_t level, uint32_t idx)
{
    return -kTopHalf + cube_half(level) * (1 + (idx << 1));
}

inline int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// Truncating conversion of a screen coordinate, first limited to +-2^30 so a
// cube corner next to the eye plane cannot overflow the int conversion.
inline int screen_int(float v)
{
    v = v > -1073741824.f ? v : -1073741824.f;
    v = v < 1073741824.f ? v : 1073741824.f;
    return (int)v;
}

} // namespace

CubeSplit::CubeSplit(uint32_t level, uint32_t split_count)
{
    const uint32_t size = 1u << level;
    uint32_t remaining = std::min(1u << (3u * level), split_count);
    uint32_t per_axis[3];
    for (int a = 0; a < 3; ++a) {
        if (remaining <= size) {
            per_axis[a] = remaining;
            remaining = 1;
        } else {
            per_axis[a] = size;
            remaining = std::max(1u, remaining / size);
        }
    }
    for (int a = 0; a < 3; ++a) {
        const uint32_t c = per_axis[a];
        cuts[a].assign(c + 1, 0);
        cuts[a].back() = size;
        std::vector<float> bounds(c + 1);
        bounds.front() = -kTopHalf;
        bounds.back() = kTopHalf;
        const float origin = -kTopHalf, length = 2 * kTopHalf;
        for (uint32_t i = 1; i < c; ++i) {
            const uint32_t idx = i * size / c;
            cuts[a][i] = idx;
            bounds[i] = origin + length * idx / size;
        }
        axis[a] = BoundarySplitter(0u, std::move(bounds));
    }
}

// eyeFromTransform (BlockOrderUtil.hpp:54-59) for a rigid world->eye
// transform M = [R t]: the eye is -R^T t.
void eye_from_view(const float* M, float* eye)
{
    const double t0 = M[3], t1 = M[7], t2 = M[11];
    for (int i = 0; i < 3; ++i)
        eye[i] = (float)(-((double)M[i] * t0 + (double)M[4 + i] * t1 + (double)M[8 + i] * t2));
}

// boundingBoxOfIndexBox (BlockOrderUtil.hpp:29-51).
void index_box_coords(uint32_t level, const uint32_t* bmin, const uint32_t* bmax, float* cmin,
                      float* cmax)
{
    const double h = cube_half(level);
    for (int a = 0; a < 3; ++a) {
        cmin[a] = (float)((double)cube_center(level, bmin[a]) - h);
        cmax[a] = (float)((double)cube_center(level, bmax[a] - 1) + h);
    }
}

// Host setup of DenseFieldRenderer_cu::Impl::renderDenseField
// (DenseFieldRenderer_cu.cu:317-453) plus the per-block constants of the
// kernel (:164-181) and renderArgb (renderArgb.hpp:34-38).  The camera basis
// the reference reads off VisualizationLibrary matrices (:354-365) is, for a
// rigid transform, eye = -R^T t, e1/e2 = normalized rows 0/1 of R and
// n = -Lcam * normalized row 2 of R; it is evaluated in double and rounded
// once to float.
int block_params(const float* M, uint32_t vp_w, uint32_t vp_h, float fov_y_deg, uint32_t level,
                 const uint32_t* index, uint32_t tex_depth, float fmin, float fmax,
                 float threshold_rel, float quality, float alpha_cutoff, rv_block_params& p)
{
    std::memset(&p, 0, sizeof p);
    p.vp_w = vp_w;
    p.vp_h = vp_h;
    p.level = level;
    p.tex_depth = tex_depth;
    p.fmin = fmin;
    p.fmax = fmax;
    p.quality = quality;
    p.alpha_cutoff = alpha_cutoff;
    p.half = cube_half(level);
    for (int a = 0; a < 3; ++a)
        p.center[a] = cube_center(level, index[a]);

    double R[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            R[r][c] = M[4 * r + c];
    eye_from_view(M, p.pos);
    float* axes[3] = {p.e1, p.e2, p.n};
    const double scale[3] = {1.0, 1.0, -(double)kLcam};
    for (int r = 0; r < 3; ++r) {
        const double len = std::sqrt(R[r][0] * R[r][0] + R[r][1] * R[r][1] + R[r][2] * R[r][2]);
        for (int i = 0; i < 3; ++i)
            axes[r][i] = (float)(R[r][i] / len * scale[r]);
    }
    p.h = 2 * kLcam * std::tan(0.5 * fov_y_deg * kPi / 180);
    p.pixel_size = p.h / (int)vp_h;
    p.w = p.pixel_size * (int)vp_w;

    // View rectangle: integer bounding box of the projected cube corners, the
    // whole viewport if a corner is not in front of the eye (:368-409).
    int lo[2] = {0, 0}, hi[2] = {0, 0};
    bool any = false;
    auto add = [&](int x, int y) {
        if (!any) {
            lo[0] = hi[0] = x;
            lo[1] = hi[1] = y;
            any = true;
            return;
        }
        lo[0] = std::min(lo[0], x);
        hi[0] = std::max(hi[0], x);
        lo[1] = std::min(lo[1], y);
        hi[1] = std::max(hi[1], y);
    };
    for (int c = 0; c < 8; ++c) {
        float wv[3];
        for (int a = 0; a < 3; ++a)
            wv[a] = p.center[a] + (float)(((c >> a) & 1) * 2 - 1) * p.half;
        const float ex = M[0] * wv[0] + M[1] * wv[1] + M[2] * wv[2] + M[3] * 1.f;
        const float ey = M[4] * wv[0] + M[5] * wv[1] + M[6] * wv[2] + M[7] * 1.f;
        const float ez = M[8] * wv[0] + M[9] * wv[1] + M[10] * wv[2] + M[11] * 1.f;
        if (ez >= 0) {
            add(0, 0);
            add((int)vp_w, (int)vp_h);
            break;
        }
        add(clampi(screen_int(-ex / ez * kLcam / p.pixel_size) + ((int)vp_w >> 1), 0, (int)vp_w - 1),
            clampi(screen_int(-ey / ez * kLcam / p.pixel_size) + ((int)vp_h >> 1), 0, (int)vp_h - 1));
    }
    p.view_x = (uint32_t)lo[0];
    p.view_y = (uint32_t)lo[1];
    p.view_w = (uint32_t)(hi[0] - lo[0]);
    p.view_h = (uint32_t)(hi[1] - lo[1]);
    p.threshold = fmin + (fmax - fmin) * threshold_rel; // relToAbsFieldValue, :597-601

    p.nodes = (1u << tex_depth) + 1u;
    p.texel = 2 * p.half / p.nodes;
    const double tex_half = p.half + 0.5 * p.texel;
    for (int a = 0; a < 3; ++a)
        p.tex_origin[a] = p.center[a] - (float)tex_half;
    p.tcf = (float)(p.nodes / (2 * tex_half));
    float ilc = 1.f / ((1 << (tex_depth + level)) * quality);
    p.i_layer_count = ilc * 100.f;
    p.opacity_factor = (p.threshold - fmin) / (fmax - fmin);
    return (p.view_w == 0 || p.view_h == 0) ? 0 : 1;
}

// Painter's order of disjoint blocks of mixed levels.  A node of the octree
// is handled as follows: if it is itself a block, emit it; if all blocks
// below it share one level L, emit them in BackToFront3 order over the node's
// index box at level L (skipping absent cells); otherwise order its 8
// children with BackToFront3 over their 2x2x2 box and recurse.  At a single
// level this is exactly the BackToFrontOrder of the whole cube.
int mixed_back_to_front(const uint32_t* blocks, uint64_t n, const float* eye,
                        std::vector<uint32_t>& order)
{
    order.clear();
    if (n == 0)
        return RV_OK;
    // Blocks all of one level L <= 8 (the root node is then "uniform"): the
    // same BackToFront3 walk of the cube's index box at level L, with a dense
    // index table in place of the hash maps.
    bool one_level = blocks[0] <= 8;
    for (uint64_t i = 1; i < n && one_level; ++i)
        one_level = blocks[4 * i] == blocks[0];
    if (one_level) {
        const uint32_t L = blocks[0], m = 1u << L;
        std::vector<uint32_t> at((size_t)m * m * m, ~0u);
        for (uint64_t i = 0; i < n; ++i) {
            const uint32_t x = blocks[4 * i + 1], y = blocks[4 * i + 2], z = blocks[4 * i + 3];
            if (x >= m || y >= m || z >= m) {
                set_error("rv_mixed_back_to_front: block index outside its level");
                return RV_ERR_INVALID;
            }
            uint32_t& slot = at[((size_t)z * m + y) * m + x];
            if (slot != ~0u) {
                set_error("rv_mixed_back_to_front: duplicate block");
                return RV_ERR_INVALID;
            }
            slot = (uint32_t)i;
        }
        const uint32_t bmin[3] = {0, 0, 0}, bmax[3] = {m, m, m};
        float cmin[3], cmax[3];
        index_box_coords(L, bmin, bmax, cmin, cmax);
        AxisSplit sp[3];
        for (int a = 0; a < 3; ++a)
            sp[a] = UniformSplitter(bmin[a], bmax[a], cmin[a], cmax[a]).split(eye[a]);
        BackToFront3 it3(sp);
        uint32_t p[3];
        order.reserve(n);
        while (it3.next(p)) {
            const uint32_t b = at[((size_t)p[2] * m + p[1]) * m + p[0]];
            if (b != ~0u)
                order.push_back(b);
        }
        return RV_OK;
    }
    auto key = [](uint32_t level, uint32_t x, uint32_t y, uint32_t z) {
        return ((uint64_t)level << 60) | ((uint64_t)x << 40) | ((uint64_t)y << 20) | (uint64_t)z;
    };
    struct Info {
        uint32_t min_level = ~0u, max_level = 0;
    };
    std::unordered_map<uint64_t, uint32_t> block_at;
    std::unordered_map<uint64_t, Info> node_info;
    block_at.reserve(n * 2);
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t L = blocks[4 * i], x = blocks[4 * i + 1], y = blocks[4 * i + 2],
                       z = blocks[4 * i + 3];
        if (L > 19 || x >= (1u << L) || y >= (1u << L) || z >= (1u << L)) {
            set_error("rv_mixed_back_to_front: block index outside its level");
            return RV_ERR_INVALID;
        }
        if (!block_at.emplace(key(L, x, y, z), (uint32_t)i).second) {
            set_error("rv_mixed_back_to_front: duplicate block");
            return RV_ERR_INVALID;
        }
        for (uint32_t l = 0; l <= L; ++l) {
            const uint32_t sh = L - l;
            Info& inf = node_info[key(l, x >> sh, y >> sh, z >> sh)];
            inf.min_level = std::min(inf.min_level, L);
            inf.max_level = std::max(inf.max_level, L);
        }
    }
    // Disjointness: a block must not contain another block.
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t L = blocks[4 * i];
        const Info& inf = node_info[key(L, blocks[4 * i + 1], blocks[4 * i + 2], blocks[4 * i + 3])];
        if (inf.max_level != L) {
            set_error("rv_mixed_back_to_front: blocks overlap");
            return RV_ERR_INVALID;
        }
    }
    struct Rec {
        static void visit(uint32_t l, uint32_t x, uint32_t y, uint32_t z, const float* eye,
                          std::unordered_map<uint64_t, uint32_t>& block_at,
                          std::unordered_map<uint64_t, Info>& node_info,
                          std::vector<uint32_t>& order, decltype(key)& key)
        {
            auto it = node_info.find(key(l, x, y, z));
            if (it == node_info.end())
                return;
            auto b = block_at.find(key(l, x, y, z));
            if (b != block_at.end()) {
                order.push_back(b->second);
                return;
            }
            const Info inf = it->second;
            const bool uniform = inf.min_level == inf.max_level;
            const uint32_t L = uniform ? inf.min_level : l + 1;
            const uint32_t sh = L - l;
            const uint32_t bmin[3] = {x << sh, y << sh, z << sh};
            const uint32_t bmax[3] = {(x + 1) << sh, (y + 1) << sh, (z + 1) << sh};
            float cmin[3], cmax[3];
            index_box_coords(L, bmin, bmax, cmin, cmax);
            AxisSplit s[3];
            for (int a = 0; a < 3; ++a)
                s[a] = UniformSplitter(bmin[a], bmax[a], cmin[a], cmax[a]).split(eye[a]);
            BackToFront3 it3(s);
            uint32_t p[3];
            while (it3.next(p)) {
                if (uniform) {
                    auto bb = block_at.find(key(L, p[0], p[1], p[2]));
                    if (bb != block_at.end())
                        order.push_back(bb->second);
                } else {
                    visit(L, p[0], p[1], p[2], eye, block_at, node_info, order, key);
                }
            }
        }
    };
    Rec::visit(0, 0, 0, 0, eye, block_at, node_info, order, key);
    if (order.size() != n) {
        set_error("rv_mixed_back_to_front: internal ordering error");
        return RV_ERR_INTERNAL;
    }
    return RV_OK;
}

} // namespace rv
