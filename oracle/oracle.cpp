/*
 * oracle.cpp -- CPU restatement of the reference sparse->dense interpolation
 * and Argb ray-cast path (deadmorous/revise).
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  It is the checker for the
 * HIP path and the timed CPU baseline; the product library never links it.
 *
 * It is written to follow the reference's algorithms step by step (cited per
 * function), in plain single-threaded C++ per block, with the reference's
 * floating-point operand order.  Where the reference relies on NVIDIA texture
 * hardware (tex3D / tex1D linear filtering) or on CUDA's powf, this file uses
 * exact fp32 trilinear / linear filtering and the documented pow replacement
 * (orc_powf) -- the same definitions the HIP kernels use.
 *
 * Build: oracle/Makefile (g++ -O2 -ffp-contract=off).
 */
#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

/* ------------------------------------------------------------------------ */
/* Block tree walk and n2i.                                                  */
/* ------------------------------------------------------------------------ */

struct Blk {
    uint32_t node;
    uint32_t level; /* relative to the subtree root */
    uint32_t loc[3];
};

/* BlockTree::walk(root, maxDepth, cb): BlockTree.hpp:222-239.  Children are
 * visited in local index order 0..7, x = bit0 (childBlockLocation,
 * BlockTree.hpp:366-372). */
template <class Cb>
void walk_limited(const uint32_t* children, uint32_t nnodes, const Blk& b, uint32_t maxDepth,
                  const Cb& cb, bool& bad)
{
    cb(b);
    for (uint32_t k = 0; k < 8; ++k) {
        uint32_t c = children[8 * (size_t)b.node + k];
        if (!c)
            continue;
        if (c >= nnodes) {
            bad = true;
            return;
        }
        if (maxDepth > 0) {
            Blk ch;
            ch.node = c;
            ch.level = b.level + 1;
            for (int d = 0; d < 3; ++d)
                ch.loc[d] = (b.loc[d] << 1) + ((k >> d) & 1u);
            walk_limited(children, nnodes, ch, maxDepth - 1, cb, bad);
            if (bad)
                return;
        }
    }
}

uint64_t key3(uint32_t x, uint32_t y, uint32_t z)
{
    return (uint64_t)x | ((uint64_t)y << 21) | ((uint64_t)z << 42);
}

} // namespace

extern "C" uint32_t orc_subtree_depth(const uint32_t* children, uint32_t nnodes, uint32_t max_depth)
{
    /* depthUpTo: the deepest relative level reached, capped at max_depth. */
    uint32_t best = 0;
    bool bad = false;
    Blk root{0, 0, {0, 0, 0}};
    if (nnodes == 0)
        return 0;
    walk_limited(children, nnodes, root, max_depth,
                 [&](const Blk& b) {
                     if (b.level > best)
                         best = b.level;
                 },
                 bad);
    return best;
}

extern "C" int64_t orc_subtree_n2i(const uint32_t* children, uint32_t nnodes, uint32_t depth,
                                   uint16_t* n2i_out, uint64_t cap)
{
    if (nnodes == 0 || depth > 15)
        return -1;
    /* BlockTreeNodes ctor: BlockTreeNodes.hpp:66-80.  For each visited block,
     * the 8 corners in local order (x = bit0) are
     * ((location << levelShift) & mask) + (bit << levelShift),
     * BlockTreeNodes.hpp:166-181; new corners are appended. */
    std::unordered_map<uint64_t, uint64_t> i2n;
    uint64_t count = 0;
    bool overflow = false, bad = false;
    const uint32_t mask = (1u << depth) - 1u;
    Blk root{0, 0, {0, 0, 0}};
    walk_limited(children, nnodes, root, depth,
                 [&](const Blk& b) {
                     uint32_t shift = depth - b.level;
                     uint32_t base[3];
                     for (int d = 0; d < 3; ++d)
                         base[d] = (b.loc[d] << shift) & mask;
                     for (uint32_t k = 0; k < 8; ++k) {
                         uint32_t c[3];
                         for (int d = 0; d < 3; ++d)
                             c[d] = base[d] + (((k >> d) & 1u) << shift);
                         uint64_t key = key3(c[0], c[1], c[2]);
                         if (i2n.find(key) == i2n.end()) {
                             i2n[key] = count;
                             if (count < cap) {
                                 n2i_out[3 * count + 0] = (uint16_t)c[0];
                                 n2i_out[3 * count + 1] = (uint16_t)c[1];
                                 n2i_out[3 * count + 2] = (uint16_t)c[2];
                             } else {
                                 overflow = true;
                             }
                             ++count;
                         }
                     }
                 },
                 bad);
    if (bad || overflow)
        return -1;
    return (int64_t)count;
}

/* vertexIndexToOrdinal<3>: IndexTransform.hpp:250-257 (nested
 * (r << level) + r + i form, evaluated with unsigned wrap like the source). */
static inline uint32_t ordinal3(uint32_t x, uint32_t y, uint32_t z, uint32_t level)
{
    uint32_t r = (z << level) + z + y;
    r = (r << level) + r + x;
    return r;
}

extern "C" void orc_ordinals(const uint16_t* n2i, uint64_t n, uint32_t depth, uint32_t* out)
{
    for (uint64_t i = 0; i < n; ++i)
        out[i] = ordinal3(n2i[3 * i], n2i[3 * i + 1], n2i[3 * i + 2], depth);
}

/* ------------------------------------------------------------------------ */
/* Dense interpolation: DenseFieldInterpolator.hpp:198-419.                  */
/* ------------------------------------------------------------------------ */

namespace {

/* HypercubeWalker<3,1>, <3,2>, <3,3> tables (HypercubeWalker.hpp:72-103):
 * origin (x,y,z) and dims. */
const uint32_t kEdges[12][4] = {
    {0, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, 0, 2}, {0, 0, 1, 0}, {0, 0, 1, 1}, {0, 1, 0, 0},
    {0, 1, 0, 2}, {0, 1, 1, 0}, {1, 0, 0, 1}, {1, 0, 0, 2}, {1, 0, 1, 1}, {1, 1, 0, 2},
};
const uint32_t kFaces[6][5] = {
    {0, 0, 0, 0, 1}, {0, 0, 0, 0, 2}, {0, 0, 0, 1, 2},
    {0, 0, 1, 0, 1}, {0, 1, 0, 0, 2}, {1, 0, 0, 1, 2},
};

} // namespace

extern "C" int orc_interpolate(const float* sparse, const uint16_t* n2i, uint64_t n,
                               uint32_t depth, float* dense)
{
    if (depth > 10)
        return -1;
    const uint32_t npe = 1u + (1u << depth);
    const uint64_t total = (uint64_t)npe * npe * npe;
    /* NaN fill + scatter: :207-215 */
    for (uint64_t i = 0; i < total; ++i)
        dense[i] = NAN;
    for (uint64_t s = 0; s < n; ++s) {
        uint32_t x = n2i[3 * s], y = n2i[3 * s + 1], z = n2i[3 * s + 2];
        if (x >= npe || y >= npe || z >= npe)
            return -2;
        dense[ordinal3(x, y, z, depth)] = sparse[s];
    }
    const uint32_t stride[3] = {1u, npe, npe * npe};
    const float noField = 1e30f; /* BlockTreeFieldProvider.hpp:137-139 */

    uint32_t levelShift = depth;
    for (uint32_t level = 0; level < depth; ++level, --levelShift) {
        const uint32_t cubesPerEdge = 1u << level;
        const uint32_t hs = levelShift - 1;
        const uint32_t half[3] = {stride[0] << hs, stride[1] << hs, stride[2] << hs};
        const uint32_t voff = 1u << hs;
        /* cubes in IncMultiIndex order (x fastest): IncMultiIndex.hpp:122-139 */
        for (uint32_t cz = 0; cz < cubesPerEdge; ++cz)
            for (uint32_t cy = 0; cy < cubesPerEdge; ++cy)
                for (uint32_t cx = 0; cx < cubesPerEdge; ++cx) {
                    const uint32_t base[3] = {cx << levelShift, cy << levelShift,
                                              cz << levelShift};
                    auto at = [&](const uint32_t* o, const uint32_t* dims, int nd) {
                        uint32_t idx[3];
                        for (int d = 0; d < 3; ++d)
                            idx[d] = base[d] + (o[d] << levelShift);
                        for (int k = 0; k < nd; ++k)
                            idx[dims[k]] += voff;
                        return ordinal3(idx[0], idx[1], idx[2], depth);
                    };
                    /* 12 edges: :257-275 */
                    for (int e = 0; e < 12; ++e) {
                        uint32_t dims[1] = {kEdges[e][3]};
                        uint32_t m = at(kEdges[e], dims, 1);
                        float& mid = dense[m];
                        if (std::isnan(mid)) {
                            float a = dense[m - half[dims[0]]];
                            float b = dense[m + half[dims[0]]];
                            if (a == noField || b == noField)
                                mid = noField;
                            else
                                mid = 0.5f * (a + b);
                        }
                    }
                    /* 6 faces: :289-311 */
                    for (int f = 0; f < 6; ++f) {
                        uint32_t dims[2] = {kFaces[f][3], kFaces[f][4]};
                        uint32_t m = at(kFaces[f], dims, 2);
                        float& mid = dense[m];
                        if (std::isnan(mid)) {
                            float s1 = dense[m - half[dims[0]]], e1 = dense[m + half[dims[0]]];
                            float s2 = dense[m - half[dims[1]]], e2 = dense[m + half[dims[1]]];
                            if (s1 == noField || e1 == noField || s2 == noField || e2 == noField)
                                mid = noField;
                            else
                                mid = 0.25f * (s1 + e1 + s2 + e2);
                        }
                    }
                    /* cell: :343-370; make_real(1./6) is float(1/6) */
                    {
                        const uint32_t o[3] = {0, 0, 0};
                        uint32_t dims[3] = {0, 1, 2};
                        uint32_t m = at(o, dims, 3);
                        float& mid = dense[m];
                        if (std::isnan(mid)) {
                            float s1 = dense[m - half[0]], e1 = dense[m + half[0]];
                            float s2 = dense[m - half[1]], e2 = dense[m + half[1]];
                            float s3 = dense[m - half[2]], e3 = dense[m + half[2]];
                            if (s1 == noField || e1 == noField || s2 == noField ||
                                e2 == noField || s3 == noField || e3 == noField)
                                mid = noField;
                            else
                                mid = (float)(1. / 6) * (s1 + e1 + s2 + e2 + s3 + e3);
                        }
                    }
                }
    }
    /* 1e30 -> NaN: :416-418 */
    for (uint64_t i = 0; i < total; ++i)
        if (dense[i] == noField)
            dense[i] = NAN;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Splitters and BackToFrontOrder.                                           */
/* ------------------------------------------------------------------------ */

namespace {

struct IR {
    uint32_t a = 0, b = 0;
    bool empty() const { return a == b; }
};
struct Split {
    IR before;
    uint32_t at = ~0u;
    IR after;
};

/* UniformIndexRangeSplitter: UniformIndexRangeSplitter.hpp:27-57 */
struct UniformSplitter {
    IR ir;
    float c0, c1, factor;
    UniformSplitter(IR r, float a, float b) : ir(r), c0(a), c1(b)
    {
        factor = (float)(ir.b - ir.a) / (c1 - c0);
    }
    Split operator()(float x) const
    {
        Split s;
        if (x <= c0) {
            s.after = ir;
            return s;
        }
        if (x >= c1) {
            s.before = ir;
            return s;
        }
        uint32_t at = (uint32_t)((x - c0) * factor) + ir.a;
        s.before = {ir.a, at};
        s.at = at;
        s.after = {at + 1, ir.b};
        return s;
    }
};

/* NonUniformIndexRangeSplitter: NonUniformIndexRangeSplitter.hpp:31-77 */
struct NonUniformSplitter {
    uint32_t start = 0;
    std::vector<float> coords;
    NonUniformSplitter() = default;
    NonUniformSplitter(uint32_t s, const float* split, uint32_t ns, float c0, float c1) : start(s)
    {
        coords.resize(2 + ns);
        coords.front() = c0;
        std::copy(split, split + ns, coords.begin() + 1);
        coords.back() = c1;
    }
    Split operator()(float x) const
    {
        Split s;
        auto it = std::lower_bound(coords.begin(), coords.end(), x);
        uint32_t n = (uint32_t)(coords.size() - 1);
        if (it == coords.begin()) {
            s.after = {start, start + n};
            return s;
        }
        if (it == coords.end()) {
            s.before = {start, start + n};
            return s;
        }
        uint32_t at = start + (uint32_t)(it - coords.begin()) - 1;
        s.before = {start, at};
        s.at = at;
        s.after = {at + 1, start + n};
        return s;
    }
    IR index_range() const { return {start, start + (uint32_t)coords.size() - 1u}; }
};

void put_split(const Split& s, uint32_t* out5)
{
    out5[0] = s.before.empty() ? 0 : s.before.a;
    out5[1] = s.before.empty() ? 0 : s.before.b;
    out5[2] = s.at;
    out5[3] = s.after.empty() ? 0 : s.after.a;
    out5[4] = s.after.empty() ? 0 : s.after.b;
}

/* BackToFrontOrder iterator: BackToFrontOrder.hpp:95-325.  Generic over the
 * dimension; splits[axis] are the per-axis splits at the eye. */
void b2f_visit(uint32_t dim, const Split* splits, const std::function<void(const uint32_t*)>& emit)
{
    static const uint8_t order1[] = {0, 2, 1};
    static const uint8_t order2[] = {0, 2, 6, 8, 1, 3, 5, 7, 4};
    static const uint8_t order3[] = {0,  2,  6,  8,  18, 20, 24, 26, 1,  3,  5,  7,  9, 11,
                                     15, 17, 19, 21, 23, 25, 4,  10, 12, 14, 16, 22, 13};
    const uint8_t* order = dim == 1 ? order1 : dim == 2 ? order2 : order3;
    const size_t norder = dim == 1 ? 3 : dim == 2 ? 9 : 27;

    for (size_t o = 0; o < norder; ++o) {
        uint32_t ord = order[o];
        uint32_t mp[3] = {0, 0, 0};
        if (dim == 1)
            mp[0] = ord;
        else if (dim == 2) {
            mp[0] = ord % 3u;
            mp[1] = ord / 3u;
        } else {
            mp[0] = ord % 3u;
            ord /= 3u;
            mp[1] = ord % 3u;
            mp[2] = ord / 3u;
        }
        bool valid = true;
        for (uint32_t a = 0; a < dim; ++a) {
            const Split& s = splits[a];
            bool v = mp[a] == 0 ? !s.before.empty() : mp[a] == 1 ? s.at != ~0u : !s.after.empty();
            if (!v)
                valid = false;
        }
        if (!valid)
            continue;
        uint32_t mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0}, p[3] = {0, 0, 0};
        for (uint32_t a = 0; a < dim; ++a) {
            const Split& s = splits[a];
            IR r = mp[a] == 0 ? s.before : mp[a] == 1 ? IR{s.at, s.at + 1} : s.after;
            mn[a] = p[a] = r.a;
            mx[a] = r.b;
        }
        for (;;) {
            uint32_t pos[3];
            for (uint32_t a = 0; a < dim; ++a)
                pos[a] = mp[a] == 2 ? mx[a] - 1 + mn[a] - p[a] : p[a];
            emit(pos);
            /* inc_pos: BackToFrontOrder.hpp:238-311 (unsigned arithmetic) */
            bool more = true;
            if (dim == 1) {
                ++p[0];
                more = p[0] < mx[0];
            } else if (dim == 2) {
                uint32_t s = p[0] + p[1];
                if (s + 2 == mx[0] + mx[1])
                    more = false;
                else if (p[0] > mn[0] && p[1] + 1 < mx[1]) {
                    --p[0];
                    ++p[1];
                } else if (s - mn[1] + 1 < mx[0]) {
                    p[0] = s - mn[1] + 1;
                    p[1] = mn[1];
                } else {
                    p[0] = mx[0] - 1;
                    p[1] = s - mx[0] + 2;
                }
            } else {
                uint32_t s = p[0] + p[1] + p[2];
                if (s + 3 == mx[0] + mx[1] + mx[2])
                    more = false;
                else if (p[0] > mn[0] && p[1] + 1 < mx[1]) {
                    --p[0];
                    ++p[1];
                } else if (p[0] + p[1] > mn[0] + mn[1] && p[2] + 1 < mx[2]) {
                    if (s - mn[1] - p[2] < mx[0] + 1) {
                        uint32_t n0 = s - mn[1] - p[2] - 1, n2 = p[2] + 1;
                        p[0] = n0;
                        p[1] = mn[1];
                        p[2] = n2;
                    } else {
                        uint32_t n1 = s - mx[0] - p[2], n2 = p[2] + 1;
                        p[0] = mx[0] - 1;
                        p[1] = n1;
                        p[2] = n2;
                    }
                } else {
                    ++s;
                    uint32_t x0 = s - mn[1] - mn[2];
                    if (x0 < mx[0]) {
                        p[0] = x0;
                        p[1] = mn[1];
                        p[2] = mn[2];
                    } else {
                        uint32_t x1 = s - mx[0] + 1 - mn[2];
                        if (x1 < mx[1]) {
                            p[0] = mx[0] - 1;
                            p[1] = x1;
                            p[2] = mn[2];
                        } else {
                            p[0] = mx[0] - 1;
                            p[1] = mx[1] - 1;
                            p[2] = s + 2 - mx[0] - mx[1];
                        }
                    }
                }
            }
            if (!more)
                break;
        }
    }
}

/* IndexCubeSplitter<3>: IndexCubeSplitter.hpp:35-196 */
struct CubeSplitter {
    uint32_t level, split_count;
    float bmin[3], bmax[3];
    std::vector<uint32_t> split_indices[3];
    NonUniformSplitter splitter[3];

    CubeSplitter(uint32_t lvl, uint32_t cnt, const float* lo, const float* hi)
        : level(lvl), split_count(cnt)
    {
        for (int a = 0; a < 3; ++a) {
            bmin[a] = lo[a];
            bmax[a] = hi[a];
        }
        uint32_t size = 1u << level;
        uint32_t max_split = 1u << (3u * level);
        uint32_t axis_counts[3];
        uint32_t sc = std::min(max_split, split_count);
        for (uint32_t axis = 0; axis < 3; ++axis) {
            if (sc <= size) {
                axis_counts[axis] = sc;
                sc = 1;
            } else {
                axis_counts[axis] = size;
                sc = std::max(1u, sc / size);
            }
        }
        for (uint32_t axis = 0; axis < 3; ++axis) {
            uint32_t c = axis_counts[axis];
            std::vector<uint32_t> si(c + 1);
            si.front() = 0;
            si.back() = size;
            std::vector<float> sco(c - 1);
            float origin = bmin[axis], length = bmax[axis] - bmin[axis];
            for (uint32_t i = 1; i < c; ++i) {
                uint32_t idx = i * size / c;
                si[i] = idx;
                sco[i - 1] = origin + length * idx / size;
            }
            split_indices[axis] = si;
            splitter[axis] =
                NonUniformSplitter(0u, sco.data(), (uint32_t)sco.size(), bmin[axis], bmax[axis]);
        }
    }
    void at(const uint32_t* pos, uint32_t* box6) const
    {
        for (int a = 0; a < 3; ++a) {
            box6[a] = split_indices[a][pos[a]];
            box6[3 + a] = split_indices[a][pos[a] + 1];
        }
    }
    uint32_t end(int a) const { return splitter[a].index_range().b; }
};

} // namespace

extern "C" void orc_uniform_split(uint32_t i0, uint32_t i1, float c0, float c1, float x,
                                  uint32_t* out5)
{
    put_split(UniformSplitter({i0, i1}, c0, c1)(x), out5);
}

extern "C" void orc_nonuniform_split(uint32_t start, const float* split, uint32_t nsplit,
                                     float c0, float c1, float x, uint32_t* out5)
{
    put_split(NonUniformSplitter(start, split, nsplit, c0, c1)(x), out5);
}

extern "C" int64_t orc_b2f_macro(uint32_t dim, const uint32_t* imin, const uint32_t* imax,
                                 const float* cmin, const float* cmax, const float* eye,
                                 uint32_t* out, uint64_t cap)
{
    if (dim < 1 || dim > 3)
        return -1;
    Split splits[3];
    for (uint32_t a = 0; a < dim; ++a)
        splits[a] = UniformSplitter({imin[a], imax[a]}, cmin[a], cmax[a])(eye[a]);
    uint64_t count = 0;
    bool overflow = false;
    b2f_visit(dim, splits, [&](const uint32_t* pos) {
        if (count < cap)
            for (uint32_t a = 0; a < dim; ++a)
                out[dim * count + a] = pos[a];
        else
            overflow = true;
        ++count;
    });
    return overflow ? -1 : (int64_t)count;
}

extern "C" int64_t orc_index_cube_split_all(uint32_t level, uint32_t split_count,
                                            int add_trailing_empty, uint32_t* boxes_out,
                                            uint64_t cap)
{
    const float lo[3] = {-5, -5, -5}, hi[3] = {5, 5, 5};
    CubeSplitter cs(level, split_count, lo, hi);
    uint64_t count = 0;
    uint32_t end[3] = {cs.end(0), cs.end(1), cs.end(2)};
    for (uint32_t z = 0; z < end[2]; ++z)
        for (uint32_t y = 0; y < end[1]; ++y)
            for (uint32_t x = 0; x < end[0]; ++x) {
                uint32_t pos[3] = {x, y, z};
                if (count < cap)
                    cs.at(pos, boxes_out + 6 * count);
                ++count;
            }
    if (add_trailing_empty)
        while (count < split_count) {
            if (count < cap)
                std::fill(boxes_out + 6 * count, boxes_out + 6 * count + 6, 0u);
            ++count;
        }
    return count <= cap ? (int64_t)count : -1;
}

extern "C" int64_t orc_index_cube_split_sorted(uint32_t level, uint32_t split_count,
                                               const float* eye, uint32_t* boxes_out,
                                               uint64_t cap)
{
    const float lo[3] = {-5, -5, -5}, hi[3] = {5, 5, 5};
    CubeSplitter cs(level, split_count, lo, hi);
    Split splits[3];
    for (int a = 0; a < 3; ++a)
        splits[a] = cs.splitter[a](eye[a]);
    uint64_t count = 0;
    bool overflow = false;
    b2f_visit(3, splits, [&](const uint32_t* pos) {
        if (count < cap)
            cs.at(pos, boxes_out + 6 * count);
        else
            overflow = true;
        ++count;
    });
    return overflow ? -1 : (int64_t)count;
}

/* ------------------------------------------------------------------------ */
/* pow replacement.                                                          */
/* ------------------------------------------------------------------------ */

/* The reference evaluates 1 - pow(1 - a*opacity, iLayerCount) with CUDA's
 * device powf (renderArgb.hpp:45), which is not available here.  Both this
 * oracle and the HIP kernels use the same explicit fp32 definition instead,
 * so CPU and GPU agree bit for bit:
 *   pow(b, e) = 2^(e*log2(b)) for 0 < b < 1, 0 at b == 0, 1 at b == 1,
 *   NaN for b < 0 or NaN b;
 *   log2(b) = k + (2/ln2)*atanh(s), b = m*2^k with m in [sqrt(1/2), sqrt(2)),
 *             s = (m-1)/(m+1), atanh by its odd Taylor series to s^9;
 *   2^y     = 2^floor(y) * P(f), f = y - floor(y), P a degree-6 polynomial
 *             fitted (least squares, Lawson-reweighted) to 2^f on [0,1].
 * Both polynomials are evaluated by Horner's rule with fused multiply-adds
 * (fmaf, one rounding each); every other step is a correctly rounded fp32
 * operation with no contraction.  Results below 2^-126 are returned as 0. */
extern "C" float orc_powf(float b, float e)
{
    if (!(b >= 0.f))
        return NAN;
    if (b == 0.f)
        return 0.f;
    if (b == 1.f)
        return 1.f;
    uint32_t bits;
    std::memcpy(&bits, &b, 4);
    int k = (int)((bits >> 23) & 0xffu) - 127;
    uint32_t mbits = (bits & 0x007fffffu) | 0x3f800000u; /* m in [1,2) */
    if (((bits >> 23) & 0xffu) == 0u) {
        /* subnormal input: scale up exactly by 2^32 first */
        float bs = b * 4294967296.0f;
        std::memcpy(&bits, &bs, 4);
        k = (int)((bits >> 23) & 0xffu) - 127 - 32;
        mbits = (bits & 0x007fffffu) | 0x3f800000u;
    }
    float m;
    std::memcpy(&m, &mbits, 4);
    if (m >= 1.41421356f) {
        m = m * 0.5f;
        k = k + 1;
    }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float poly = 0.111111111f;            /* 1/9 */
    poly = std::fmaf(poly, s2, 0.142857143f);      /* 1/7 */
    poly = std::fmaf(poly, s2, 0.2f);              /* 1/5 */
    poly = std::fmaf(poly, s2, 0.333333333f);      /* 1/3 */
    poly = std::fmaf(poly, s2, 1.0f);
    float l = (s * poly) * 2.88539008f;   /* 2/ln2 */
    float y = e * ((float)k + l);
    if (y < -126.0f)
        return 0.f;
    if (y >= 0.0f)
        return 1.0f; /* b < 1 and e > 0 give y < 0; guard rounding */
    float fl = std::floor(y);
    float f = y - fl;
    float p = 0x1.c54176p-13f;
    p = std::fmaf(p, f, 0x1.46d64cp-10f);
    p = std::fmaf(p, f, 0x1.3d0b92p-7f);
    p = std::fmaf(p, f, 0x1.c68912p-5f);
    p = std::fmaf(p, f, 0x1.ebfd58p-3f);
    p = std::fmaf(p, f, 0x1.62e42cp-1f);
    p = std::fmaf(p, f, 1.0f);
    return std::ldexp(p, (int)fl);
}

/* ------------------------------------------------------------------------ */
/* Camera and per-block parameters: DenseFieldRenderer_cu.cu:317-453.        */
/* ------------------------------------------------------------------------ */

namespace {

const float kTopHalf = 5.f;                /* TopLevelCubeHalfSize, :339 */
const float kLcam = 0.01f * 2 * kTopHalf;  /* :353 */
const double kPi = 3.14159265358979323846; /* M_PI */

/* The reference reads the camera basis off VisualizationLibrary matrices
 * (tcam.getInverse(), get3x3, getX/Y/Z, getT), :354-365.  For a rigid
 * world->eye transform M = [R t] those are: eye position -R^T t, screen x/y
 * axes = rows 0/1 of R, viewing axis = row 2 of R.  Evaluated here in double
 * and rounded to float once. */
void camera_basis(const float* M, float* pos, float* e1, float* e2, float* nvec)
{
    double R[3][3], t[3];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c)
            R[r][c] = M[4 * r + c];
        t[r] = M[4 * r + 3];
    }
    for (int i = 0; i < 3; ++i)
        pos[i] = (float)(-(R[0][i] * t[0] + R[1][i] * t[1] + R[2][i] * t[2]));
    auto nrm = [&](int row, double scale, float* out) {
        double l = std::sqrt(R[row][0] * R[row][0] + R[row][1] * R[row][1] + R[row][2] * R[row][2]);
        for (int i = 0; i < 3; ++i)
            out[i] = (float)(R[row][i] / l * scale);
    };
    nrm(0, 1.0, e1);
    nrm(1, 1.0, e2);
    nrm(2, -(double)kLcam, nvec);
}

inline int clampi(int x, int lo, int hi) { return x < lo ? lo : x > hi ? hi : x; }

/* (int) conversion of a screen coordinate, clamped into int range first so
 * that a corner right at the eye plane cannot overflow the conversion. */
inline int to_int_screen(float v)
{
    if (!(v > -1073741824.f))
        v = -1073741824.f;
    if (v > 1073741824.f)
        v = 1073741824.f;
    return (int)v;
}

} // namespace

extern "C" int orc_block_params_setup(const float* M, uint32_t vp_w, uint32_t vp_h,
                                      float fov_y_deg, uint32_t level, const uint32_t* index,
                                      uint32_t tex_depth, float fmin, float fmax,
                                      float threshold_rel, float quality, float alpha_cutoff,
                                      orc_block_params* p)
{
    std::memset(p, 0, sizeof(*p));
    p->vp_w = vp_w;
    p->vp_h = vp_h;
    p->tex_depth = tex_depth;
    p->fmin = fmin;
    p->fmax = fmax;
    p->level = level;
    p->quality = quality;
    p->alpha_cutoff = alpha_cutoff;
    /* cube: :336-349 */
    int cubesPerLevel = 1 << level;
    float half = kTopHalf / cubesPerLevel;
    p->half = half;
    for (int a = 0; a < 3; ++a)
        p->center[a] = -kTopHalf + half * (1 + (index[a] << 1));
    /* camera: :352-365 */
    camera_basis(M, p->pos, p->e1, p->e2, p->n);
    p->h = 2 * kLcam * std::tan(0.5 * fov_y_deg * kPi / 180);
    p->pixel_size = p->h / (int)vp_h;
    p->w = p->pixel_size * (int)vp_w;
    /* view rectangle: :368-409 */
    int bmin[2] = {0, 0}, bmax[2] = {0, 0};
    bool bempty = true;
    auto addpt = [&](int x, int y) {
        if (bempty) {
            bmin[0] = bmax[0] = x;
            bmin[1] = bmax[1] = y;
            bempty = false;
            return;
        }
        int v[2] = {x, y};
        for (int i = 0; i < 2; ++i) {
            if (bmin[i] > v[i])
                bmin[i] = v[i];
            else if (bmax[i] < v[i])
                bmax[i] = v[i];
        }
    };
    for (int c = 0; c < 8; ++c) {
        float d[3];
        for (int a = 0; a < 3; ++a)
            d[a] = (float)(((c >> a) & 1) * 2 - 1) * half;
        float wv[3] = {p->center[0] + d[0], p->center[1] + d[1], p->center[2] + d[2]};
        float ev[3];
        for (int r = 0; r < 3; ++r)
            ev[r] = M[4 * r + 0] * wv[0] + M[4 * r + 1] * wv[1] + M[4 * r + 2] * wv[2] +
                    M[4 * r + 3] * 1.f;
        if (ev[2] >= 0) {
            addpt(0, 0);
            addpt((int)vp_w, (int)vp_h);
            break;
        }
        int sx = clampi(to_int_screen(-ev[0] / ev[2] * kLcam / p->pixel_size) + ((int)vp_w >> 1),
                        0, (int)vp_w - 1);
        int sy = clampi(to_int_screen(-ev[1] / ev[2] * kLcam / p->pixel_size) + ((int)vp_h >> 1),
                        0, (int)vp_h - 1);
        addpt(sx, sy);
    }
    p->view_x = (uint32_t)bmin[0];
    p->view_y = (uint32_t)bmin[1];
    p->view_w = (uint32_t)(bmax[0] - bmin[0]);
    p->view_h = (uint32_t)(bmax[1] - bmin[1]);
    /* threshold (relToAbsFieldValue, :597-601) */
    p->threshold = fmin + (fmax - fmin) * threshold_rel;
    return (p->view_w == 0 || p->view_h == 0) ? 0 : 1;
}

/* ------------------------------------------------------------------------ */
/* Ray-cast: DenseFieldRenderer_cu.cu:62-215, renderArgb.hpp:30-55.          */
/* ------------------------------------------------------------------------ */

namespace {

struct V3 {
    float x, y, z;
};
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

struct Derived {
    uint32_t n;
    float texel, tcf, i_layer_count, opacity_factor;
    V3 origin;
};

Derived derive(const orc_block_params* p)
{
    Derived d;
    d.n = (1u << p->tex_depth) + 1u;
    d.texel = 2 * p->half / d.n; /* :165 */
    double texHalf = p->half + 0.5 * d.texel; /* :172 */
    d.origin = {p->center[0] - (float)texHalf, p->center[1] - (float)texHalf,
                p->center[2] - (float)texHalf};
    d.tcf = (float)(d.n / (2 * texHalf)); /* :176 */
    /* renderArgb.hpp:34-38 */
    float ilc = 1.f / ((1 << (p->tex_depth + p->level)) * p->quality);
    ilc *= 100.f;
    d.i_layer_count = ilc;
    d.opacity_factor = (p->threshold - p->fmin) / (p->fmax - p->fmin);
    return d;
}

inline float lerpf(float a, float b, float t) { return (1.f - t) * a + t * b; }

/* Exact fp32 trilinear filtering of the brick at unnormalized texel
 * coordinates (texel centres at i+0.5, clamp addressing): the texture fetch of
 * RenderFuncInput.hpp:42-49. */
float fetch3(const float* brick, uint32_t n, V3 t)
{
    float c[3] = {t.x - 0.5f, t.y - 0.5f, t.z - 0.5f};
    int i0[3], i1[3];
    float f[3];
    const float hi = (float)n;
    for (int a = 0; a < 3; ++a) {
        float v = c[a];
        if (!(v > -1.f))
            v = -1.f;
        if (v > hi)
            v = hi;
        float fl = std::floor(v);
        f[a] = v - fl;
        int i = (int)fl;
        i0[a] = clampi(i, 0, (int)n - 1);
        i1[a] = clampi(i + 1, 0, (int)n - 1);
    }
    auto at = [&](int x, int y, int z) { return brick[((size_t)z * n + y) * n + x]; };
    float c00 = lerpf(at(i0[0], i0[1], i0[2]), at(i1[0], i0[1], i0[2]), f[0]);
    float c10 = lerpf(at(i0[0], i1[1], i0[2]), at(i1[0], i1[1], i0[2]), f[0]);
    float c01 = lerpf(at(i0[0], i0[1], i1[2]), at(i1[0], i0[1], i1[2]), f[0]);
    float c11 = lerpf(at(i0[0], i1[1], i1[2]), at(i1[0], i1[1], i1[2]), f[0]);
    float c0 = lerpf(c00, c10, f[1]);
    float c1 = lerpf(c01, c11, f[1]);
    return lerpf(c0, c1, f[2]);
}

/* tex1D over the transfer table with normalized coordinates, linear filter,
 * clamp addressing (DenseFieldRenderer_cu.cu:573-589). */
void fetch_tf(const float* tf, uint32_t N, float u, float* out4)
{
    float x = u * (float)N - 0.5f;
    if (!(x > -1.f))
        x = -1.f;
    if (x > (float)N)
        x = (float)N;
    float fl = std::floor(x);
    float f = x - fl;
    int i = (int)fl;
    int a = clampi(i, 0, (int)N - 1), b = clampi(i + 1, 0, (int)N - 1);
    for (int c = 0; c < 4; ++c)
        out4[c] = lerpf(tf[4 * a + c], tf[4 * b + c], f);
}

struct Ray {
    bool hit;
    uint32_t pixel;
    int sampleCount;
    V3 step, pos;
};

/* Ray setup: DenseFieldRenderer_cu.cu:66-181. */
Ray setup_ray(const orc_block_params* p, const Derived& d, uint32_t vx, uint32_t vy)
{
    Ray r{};
    float x = (-0.5f + (vx + 0.5f) / p->vp_w) * p->w;
    float y = (-0.5f + (vy + 0.5f) / p->vp_h) * p->h;
    V3 e1{p->e1[0], p->e1[1], p->e1[2]}, e2{p->e2[0], p->e2[1], p->e2[2]},
        nn{p->n[0], p->n[1], p->n[2]}, eye{p->pos[0], p->pos[1], p->pos[2]};
    V3 ray = add(add(nn, mul(e1, x)), mul(e2, y));
    float tauEnter = 1.f, tauExit = 0.f;
    const float tol = 1e-10f;
    const float h = p->half;
    const float* c = p->center;
    const float rv[3] = {ray.x, ray.y, ray.z};
    const float ev[3] = {eye.x, eye.y, eye.z};
    for (int a = 0; a < 3; ++a) {
        if (!(std::fabs(rv[a]) > tol))
            continue;
        int o1 = a == 0 ? 1 : 0, o2 = a == 2 ? 1 : 2;
        for (int side = 0; side < 2; ++side) {
            float face = side == 0 ? c[a] - h : c[a] + h;
            float tau = (face - ev[a]) / rv[a];
            V3 pt = add(eye, mul(ray, tau));
            float pv[3] = {pt.x, pt.y, pt.z};
            if (std::fabs(pv[o1] - c[o1]) <= h && std::fabs(pv[o2] - c[o2]) <= h) {
                bool enter = side == 0 ? (rv[a] > 0) : !(rv[a] > 0);
                (enter ? tauEnter : tauExit) = tau;
            }
        }
    }
    if (tauEnter < 1)
        tauEnter = 1;
    if (!(tauExit > tauEnter))
        return r;
    r.hit = true;
    r.pixel = (p->vp_h - 1 - vy) * p->vp_w + vx;
    r.sampleCount = (int)(p->quality * (tauExit - tauEnter) * dot(ray, ray) /
                          (d.texel * (std::fabs(ray.x) + std::fabs(ray.y) + std::fabs(ray.z))));
    r.step = mul(ray, (tauExit - tauEnter) / r.sampleCount);
    r.pos = add(eye, mul(ray, tauEnter));
    return r;
}

/* renderArgb loop, renderArgb.hpp:40-51, with blendBehind, cuda_blend.hpp:24-33 */
void march(const orc_block_params* p, const Derived& d, const float* brick, const float* tf,
           uint32_t tf_n, Ray& r, float* rgba)
{
    rgba[0] = rgba[1] = rgba[2] = rgba[3] = 0.f;
    const float range = p->fmax - p->fmin;
    for (int i = 1; i < r.sampleCount; ++i, r.pos = add(r.pos, r.step)) {
        V3 tc = mul(sub(r.pos, d.origin), d.tcf);
        float v = fetch3(brick, d.n, tc);
        if (std::isnan(v))
            continue;
        float v01 = (v - p->fmin) / range;
        float c[4];
        fetch_tf(tf, tf_n, v01, c);
        c[3] = 1.f - orc_powf(1.f - c[3] * d.opacity_factor, d.i_layer_count);
        float da = c[3] * (1.f - rgba[3]);
        float a = rgba[3] + da;
        if (a > 0.f) {
            float ka = rgba[3] / a, kb = da / a;
            for (int k = 0; k < 3; ++k)
                rgba[k] = rgba[k] * ka + c[k] * kb;
        }
        rgba[3] = a;
        if (rgba[3] > p->alpha_cutoff)
            break;
    }
}

/* uint32ArgbToFloatRgba / blendAbove / floatRgbaToUint32Argb:
 * cuda_argb_transform.hpp:26-43, cuda_blend.hpp:35-44 */
uint32_t blend_above_u32(uint32_t argb, const float* rgba)
{
    float ex[4] = {((argb >> 16) & 0xffu) / 255.f, ((argb >> 8) & 0xffu) / 255.f,
                   (argb & 0xffu) / 255.f, ((argb >> 24) & 0xffu) / 255.f};
    float da = ex[3] * (1.f - rgba[3]);
    float a = rgba[3] + da;
    if (a > 0.f) {
        float ka = da / a, kb = rgba[3] / a;
        for (int k = 0; k < 3; ++k)
            ex[k] = ex[k] * ka + rgba[k] * kb;
    }
    ex[3] = a;
    return ((unsigned)(ex[3] * 255.999f) << 24) + ((unsigned)(ex[0] * 255.999f) << 16) +
           ((unsigned)(ex[1] * 255.999f) << 8) + (unsigned)(ex[2] * 255.999f);
}

/* ---- field modes (s3vs::FieldMode numbering, FieldMode.hpp:26-51) ---- */
enum { kIso = 0, kIsos = 1, kMip = 2, kArgb = 3, kArgbLight = 4, kDomain = 5, kValIso = 6, kValIsos = 7 };

/* The secondary field of ValueOnIsosurface(s): the block's brick of it and
 * the field's range; brick == nullptr for the other modes. */
struct Second {
    const float* brick;
    float smin, smax;
};

bool multi_level(uint32_t mode) { return mode == kIsos || mode == kValIsos; }

/* makeLevels, DenseFieldRenderer_cu.cu:597-612: relToAbsFieldValue of each
 * level, sorted, duplicates removed, at most 50 (MaxLevels) kept. */
std::vector<float> make_levels(float fmin, float fmax, const float* rel, uint32_t n)
{
    std::vector<float> v(n);
    for (uint32_t i = 0; i < n; ++i)
        v[i] = fmin + (fmax - fmin) * rel[i];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    if (v.size() > 50)
        v.resize(50);
    return v;
}

/* findInterval, cuda_findInterval.hpp:28-44 */
int find_interval(float v, const std::vector<float>& lv)
{
    const int n = (int)lv.size();
    if (v < lv[0])
        return 0;
    if (v >= lv[n - 1])
        return n;
    int i1 = 0, i2 = n;
    while (i2 - i1 > 1) {
        int i = (i1 + i2) >> 1;
        if (v < lv[i])
            i2 = i;
        else
            i1 = i;
    }
    return i2;
}

/* cuda_vec_op.hpp:48-52 */
V3 normalize3(V3 a)
{
    float il = 1.f / std::sqrt(dot(a, a));
    return {a.x * il, a.y * il, a.z * il};
}

/* clamp.hpp:24-34: max(0, min(1, x)), a NaN operand ignored */
float clamp01(float x) { return std::fmax(0.f, std::fmin(1.f, x)); }

/* computeFragColor + blinn, cuda_lighting.hpp:27-95, with the one light of
 * DenseFieldRenderer_cu.cu:440-445 (at pos + n + e2 h; ambient 0, diffuse 1,
 * specular 0.1); Ka = 1, Ks = 0.5, shininess 100. */
void frag_color(const orc_block_params* p, const Derived& d, const float* brick,
                const float* diffuse, V3 surf, V3 tc, float* out)
{
    const float del = 0.1f;
    V3 a{fetch3(brick, d.n, {tc.x - del, tc.y, tc.z}), fetch3(brick, d.n, {tc.x, tc.y - del, tc.z}),
         fetch3(brick, d.n, {tc.x, tc.y, tc.z - del})};
    V3 b{fetch3(brick, d.n, {tc.x + del, tc.y, tc.z}), fetch3(brick, d.n, {tc.x, tc.y + del, tc.z}),
         fetch3(brick, d.n, {tc.x, tc.y, tc.z + del})};
    V3 N = normalize3(sub(a, b));
    V3 eye{p->pos[0], p->pos[1], p->pos[2]};
    V3 light = add(add(eye, V3{p->n[0], p->n[1], p->n[2]}), mul(V3{p->e2[0], p->e2[1], p->e2[2]}, p->h));
    V3 V = normalize3(sub(eye, surf));
    V3 L = normalize3(sub(light, surf));
    if (dot(L, N) < 0.f)
        N = {-N.x, -N.y, -N.z};
    float diff = std::fmax(dot(L, N), 0.f);
    V3 H = normalize3(add(L, V));
    float spec = diff > 0.f ? orc_powf(std::fmax(dot(H, N), 0.f), 100.f) : 0.f;
    float ks = 0.5f * 0.1f;
    for (int k = 0; k < 3; ++k) {
        float c = (1.f * 0.f + (diffuse[k] * 1.f) * diff) + ks * spec;
        out[k] = clamp01(0.f + c);
    }
    out[3] = diffuse[3];
}

void blend_behind4(float* acc, const float* c)
{
    float da = c[3] * (1.f - acc[3]);
    float a = acc[3] + da;
    if (a > 0.f) {
        float ka = acc[3] / a, kb = da / a;
        for (int k = 0; k < 3; ++k)
            acc[k] = acc[k] * ka + c[k] * kb;
    }
    acc[3] = a;
}

/* renderArgbLight loop, renderArgbLight.hpp:29-63 */
void march_light(const orc_block_params* p, const Derived& d, const float* brick, const float* tf,
                 uint32_t tf_n, Ray& r, float* rgba)
{
    rgba[0] = rgba[1] = rgba[2] = rgba[3] = 0.f;
    const float range = p->fmax - p->fmin;
    for (int i = 1; i < r.sampleCount; ++i, r.pos = add(r.pos, r.step)) {
        V3 tc = mul(sub(r.pos, d.origin), d.tcf);
        float v = fetch3(brick, d.n, tc);
        if (std::isnan(v))
            continue;
        float c[4];
        fetch_tf(tf, tf_n, (v - p->fmin) / range, c);
        c[3] = 1.f - orc_powf(1.f - c[3] * d.opacity_factor, d.i_layer_count);
        if (c[3] >= 0.00001f) {
            float col[4];
            frag_color(p, d, brick, c, r.pos, tc, col);
            blend_behind4(rgba, col);
            if (rgba[3] > p->alpha_cutoff)
                break;
        }
    }
}

/* renderValueOnIsosurface(s): the diffuse colour at a surface point is the
 * table colour of the secondary field there, normalised by its range, alpha
 * times the isosurface opacity (renderValueOnIsosurface.hpp:44-48,
 * renderValueOnIsosurfaces.hpp:77-81). */
void value_diffuse(const Second& s, const Derived& d, const float* tf, uint32_t tf_n,
                   float iso_opacity, V3 tc, float* diffuse)
{
    float v = fetch3(s.brick, d.n, tc);
    fetch_tf(tf, tf_n, (v - s.smin) / (s.smax - s.smin), diffuse);
    diffuse[3] = diffuse[3] * iso_opacity;
}

/* renderIsosurface loop, renderIsosurface.hpp:31-61: the surface point is the
 * second sample's position plus step * (level - v1) / (v2 - v1).  With a
 * secondary field: renderValueOnIsosurface, renderValueOnIsosurface.hpp:32-62. */
void march_iso(const orc_block_params* p, const Derived& d, const float* brick, const float* tf,
               uint32_t tf_n, float iso_opacity, Ray& r, float* rgba, const Second& sec)
{
    rgba[0] = rgba[1] = rgba[2] = rgba[3] = 0.f;
    float v1 = fetch3(brick, d.n, mul(sub(r.pos, d.origin), d.tcf));
    r.pos = add(r.pos, r.step);
    const float lv = p->threshold;
    float diffuse[4];
    fetch_tf(tf, tf_n, (lv - p->fmin) / (p->fmax - p->fmin), diffuse);
    diffuse[3] = diffuse[3] * iso_opacity;
    for (int i = 1; i < r.sampleCount; ++i, r.pos = add(r.pos, r.step)) {
        float v2 = fetch3(brick, d.n, mul(sub(r.pos, d.origin), d.tcf));
        if (!std::isnan(v1) && !std::isnan(v2) && (v1 - lv) * (v2 - lv) < 0.f) {
            float t = (lv - v1) / (v2 - v1);
            V3 x = add(r.pos, mul(r.step, t));
            if (sec.brick)
                value_diffuse(sec, d, tf, tf_n, iso_opacity, mul(sub(x, d.origin), d.tcf), diffuse);
            float col[4];
            frag_color(p, d, brick, diffuse, x, mul(sub(x, d.origin), d.tcf), col);
            blend_behind4(rgba, col);
            if (rgba[3] > p->alpha_cutoff)
                break;
        }
        v1 = v2;
    }
}

/* renderMaxIntensityProjection.hpp:31-37 */
float march_mip(const orc_block_params* p, const Derived& d, const float* brick, Ray& r)
{
    float m = p->fmin - (p->fmax - p->fmin);
    for (int i = 1; i < r.sampleCount; ++i, r.pos = add(r.pos, r.step)) {
        float v = fetch3(brick, d.n, mul(sub(r.pos, d.origin), d.tcf));
        if (v >= p->threshold && m < v)
            m = v;
    }
    return m;
}

/* renderMaxIntensityProjection.hpp:39-48: float bits in the viewport */
uint32_t mip_merge(uint32_t prev_bits, float m, const orc_block_params* p)
{
    if (m < p->fmin)
        return prev_bits;
    float t = (m - p->fmin) / (p->fmax - p->fmin);
    float prev;
    std::memcpy(&prev, &prev_bits, 4);
    if (!std::isnan(prev) && prev > t)
        return prev_bits;
    uint32_t bits;
    std::memcpy(&bits, &t, 4);
    return bits;
}

/* renderIsosurfaces loop, renderIsosurfaces.hpp:31-103 (levels non-empty);
 * with a secondary field renderValueOnIsosurfaces,
 * renderValueOnIsosurfaces.hpp:33-109. */
void march_isos(const orc_block_params* p, const Derived& d, const float* brick, const float* tf,
                uint32_t tf_n, float iso_opacity, const std::vector<float>& lv, Ray& r, float* rgba,
                const Second& sec)
{
    rgba[0] = rgba[1] = rgba[2] = rgba[3] = 0.f;
    const int n = (int)lv.size();
    float v1 = fetch3(brick, d.n, mul(sub(r.pos, d.origin), d.tcf));
    r.pos = add(r.pos, r.step);
    int iint = std::isnan(v1) ? -1 : find_interval(v1, lv);
    for (int i = 1; i < r.sampleCount; ++i, r.pos = add(r.pos, r.step)) {
        float v2 = fetch3(brick, d.n, mul(sub(r.pos, d.origin), d.tcf));
        if (std::isnan(v2)) {
            iint = -1;
            v1 = NAN;
            continue;
        }
        if (std::isnan(v1)) {
            v1 = v2;
            iint = find_interval(v1, lv);
            continue;
        }
        int il, dl;
        if (v2 < v1) {
            if (iint == 0) {
                v1 = v2;
                continue;
            }
            il = iint - 1;
            dl = -1;
        } else {
            if (iint == n) {
                v1 = v2;
                continue;
            }
            il = iint;
            dl = 1;
        }
        while (il >= 0 && il < n) {
            float level = lv[il];
            if (!((v1 - level) * (v2 - level) < 0.f))
                break;
            float t = (level - v1) / (v2 - v1);
            V3 x = add(r.pos, mul(r.step, t));
            float diffuse[4];
            if (sec.brick) {
                value_diffuse(sec, d, tf, tf_n, iso_opacity, mul(sub(x, d.origin), d.tcf), diffuse);
            } else {
                fetch_tf(tf, tf_n, (level - p->fmin) / (p->fmax - p->fmin), diffuse);
                diffuse[3] = diffuse[3] * iso_opacity;
            }
            float col[4];
            frag_color(p, d, brick, diffuse, x, mul(sub(x, d.origin), d.tcf), col);
            blend_behind4(rgba, col);
            iint += dl;
            il += dl;
        }
        if (rgba[3] > p->alpha_cutoff)
            break;
        v1 = v2;
    }
}

/* renderDomainVoxels loop, renderDomainVoxels.hpp:30-54 */
void march_domain(const orc_block_params* p, const Derived& d, const float* brick, const float* tf,
                  uint32_t tf_n, Ray& r, float* rgba)
{
    rgba[0] = rgba[1] = rgba[2] = rgba[3] = 0.f;
    float diffuse[4];
    fetch_tf(tf, tf_n, 0.5f, diffuse);
    diffuse[3] = 1.f - orc_powf(1.f - diffuse[3] * d.opacity_factor, d.i_layer_count);
    for (int i = 1; i < r.sampleCount; ++i, r.pos = add(r.pos, r.step)) {
        float v = fetch3(brick, d.n, mul(sub(r.pos, d.origin), d.tcf));
        if (!std::isnan(v)) {
            blend_behind4(rgba, diffuse);
            if (rgba[3] > p->alpha_cutoff)
                break;
        }
    }
}

void segment_mode(const orc_block_params* p, const Derived& d, const float* brick, const float* tf,
                  uint32_t tf_n, uint32_t mode, float iso_opacity, const std::vector<float>& lv,
                  Ray& r, float* rgba, const Second& sec = Second{nullptr, 0.f, 1.f})
{
    const Second none{nullptr, 0.f, 1.f};
    if (mode == kArgbLight)
        march_light(p, d, brick, tf, tf_n, r, rgba);
    else if (mode == kIso || mode == kValIso)
        march_iso(p, d, brick, tf, tf_n, iso_opacity, r, rgba, mode == kValIso ? sec : none);
    else if (mode == kIsos || mode == kValIsos)
        march_isos(p, d, brick, tf, tf_n, iso_opacity, lv, r, rgba, mode == kValIsos ? sec : none);
    else if (mode == kDomain)
        march_domain(p, d, brick, tf, tf_n, r, rgba);
    else
        march(p, d, brick, tf, tf_n, r, rgba);
}

} // namespace

extern "C" void orc_render_block(const orc_block_params* p, const float* brick, const float* tf,
                                 uint32_t tf_n, uint32_t* viewport, uint32_t row0, uint32_t row1)
{
    Derived d = derive(p);
    for (uint32_t vy = p->view_y; vy < p->view_y + p->view_h; ++vy) {
        if (vy < row0 || vy >= row1)
            continue;
        for (uint32_t vx = p->view_x; vx < p->view_x + p->view_w; ++vx) {
            Ray r = setup_ray(p, d, vx, vy);
            if (!r.hit)
                continue;
            float rgba[4];
            march(p, d, brick, tf, tf_n, r, rgba);
            viewport[r.pixel] = blend_above_u32(viewport[r.pixel], rgba);
        }
    }
}

/* renderKernel<fieldMode> for one block in a field mode (Argb, ArgbLight,
 * Isosurface, MaxIntensityProjection), DenseFieldRenderer_cu.cu:62-215. */
extern "C" void orc_render_block_mode2(const orc_block_params* p, uint32_t mode, float iso_opacity,
                                       const float* levels_rel, uint32_t level_count,
                                       const float* brick, const float* brick2, float smin,
                                       float smax, const float* tf, uint32_t tf_n,
                                       uint32_t* viewport, uint32_t row0, uint32_t row1)
{
    const Second sec{brick2, smin, smax};
    Derived d = derive(p);
    const std::vector<float> lv = make_levels(p->fmin, p->fmax, levels_rel, level_count);
    for (uint32_t vy = p->view_y; vy < p->view_y + p->view_h; ++vy) {
        if (vy < row0 || vy >= row1)
            continue;
        for (uint32_t vx = p->view_x; vx < p->view_x + p->view_w; ++vx) {
            Ray r = setup_ray(p, d, vx, vy);
            if (!r.hit)
                continue;
            if (mode == kMip) {
                viewport[r.pixel] = mip_merge(viewport[r.pixel], march_mip(p, d, brick, r), p);
                continue;
            }
            if (multi_level(mode) && lv.empty()) {
                viewport[r.pixel] = 0u; /* renderIsosurfaces.hpp:33-34 */
                continue;
            }
            float rgba[4];
            segment_mode(p, d, brick, tf, tf_n, mode, iso_opacity, lv, r, rgba, sec);
            viewport[r.pixel] = blend_above_u32(viewport[r.pixel], rgba);
        }
    }
}

extern "C" void orc_render_block_mode(const orc_block_params* p, uint32_t mode, float iso_opacity,
                                      const float* levels_rel, uint32_t level_count,
                                      const float* brick, const float* tf, uint32_t tf_n,
                                      uint32_t* viewport, uint32_t row0, uint32_t row1)
{
    orc_render_block_mode2(p, mode, iso_opacity, levels_rel, level_count, brick, nullptr, 0.f, 1.f, tf,
                           tf_n, viewport, row0, row1);
}

/* blendImagePartMaxIntensity, BlendImagePart.hpp:148-191 */
extern "C" void orc_blend_image_part_max_intensity(float* dst, int32_t dw, const float* src,
                                                   int32_t sw, int32_t sh, int32_t dx, int32_t dy,
                                                   int just_copy)
{
    for (int32_t y = 0; y < sh; ++y)
        for (int32_t x = 0; x < sw; ++x) {
            float s = src[(size_t)y * sw + x];
            float& dv = dst[(size_t)dw * (dy + y) + dx + x];
            if (just_copy) {
                dv = s;
                continue;
            }
            if (std::isnan(s))
                continue;
            if (std::isnan(dv) || dv < s)
                dv = s;
        }
}

/* makeMaxIntensityProjection, VsRenderer.cpp:542-565, with s3vs::calcColor,
 * ColorTransferFunction.hpp:83-94, and floatRgbaToUint32Argb */
extern "C" void orc_mip_colorize(uint32_t* dst, int32_t dw, const float* src, int32_t sw, int32_t sh,
                                 int32_t dx, int32_t dy, const float* ctv, uint32_t nctv)
{
    for (int32_t y = 0; y < sh; ++y)
        for (int32_t x = 0; x < sw; ++x) {
            float s = src[(size_t)y * sw + x];
            if (std::isnan(s))
                continue;
            float c[4];
            if (s <= 0.f || nctv == 1u) {
                for (int k = 0; k < 4; ++k)
                    c[k] = ctv[k];
            } else if (s >= 1.f) {
                for (int k = 0; k < 4; ++k)
                    c[k] = ctv[4 * (nctv - 1) + k];
            } else {
                float par = s * (float)(nctv - 1);
                size_t idx = (size_t)par;
                float t = par - (float)idx;
                for (int k = 0; k < 4; ++k)
                    c[k] = ctv[4 * idx + k] * (1 - t) + ctv[4 * (idx + 1) + k] * t;
            }
            dst[(size_t)dw * (dy + y) + dx + x] =
                ((unsigned)(c[3] * 255.999f) << 24) + ((unsigned)(c[0] * 255.999f) << 16) +
                ((unsigned)(c[1] * 255.999f) << 8) + (unsigned)(c[2] * 255.999f);
        }
}

extern "C" int orc_ray_segment(const orc_block_params* p, const float* brick, const float* tf,
                               uint32_t tf_n, uint32_t vx, uint32_t vy, float* rgba)
{
    Derived d = derive(p);
    rgba[0] = rgba[1] = rgba[2] = rgba[3] = 0.f;
    if (vx < p->view_x || vx >= p->view_x + p->view_w || vy < p->view_y ||
        vy >= p->view_y + p->view_h)
        return 0;
    Ray r = setup_ray(p, d, vx, vy);
    if (!r.hit)
        return 0;
    march(p, d, brick, tf, tf_n, r, rgba);
    return 1;
}

/* Derived per-block scalars of derive(): out = {texel, tcf, i_layer_count,
 * opacity_factor}. */
extern "C" void orc_block_derived(const orc_block_params* p, float* out4)
{
    Derived d = derive(p);
    out4[0] = d.texel;
    out4[1] = d.tcf;
    out4[2] = d.i_layer_count;
    out4[3] = d.opacity_factor;
}

/* The samples march() composites for one pixel and block -- value not NaN,
 * up to and including the one that ends the segment -- with the 8 corner
 * indices of each sample's trilinear fetch (order 000,100,010,110,001,101,
 * 011,111; x fastest) and its weights fx, fy, fz.  Test support for the
 * adjoint: these fp32 decisions fix the set of samples a gradient is taken
 * over.  Returns the number of samples (-1 when above max_samples, 0 when
 * the ray misses) and the segment's alpha in *seg_alpha. */
extern "C" int orc_ray_samples(const orc_block_params* p, const float* brick, const float* tf,
                               uint32_t tf_n, uint32_t vx, uint32_t vy, uint32_t max_samples,
                               uint32_t* corners, float* weights, float* seg_alpha)
{
    Derived d = derive(p);
    *seg_alpha = 0.f;
    if (vx < p->view_x || vx >= p->view_x + p->view_w || vy < p->view_y ||
        vy >= p->view_y + p->view_h)
        return 0;
    Ray r = setup_ray(p, d, vx, vy);
    if (!r.hit)
        return 0;
    const float range = p->fmax - p->fmin;
    const uint32_t n = d.n;
    float acc = 0.f;
    int count = 0;
    for (int i = 1; i < r.sampleCount; ++i, r.pos = add(r.pos, r.step)) {
        V3 tc = mul(sub(r.pos, d.origin), d.tcf);
        float v = fetch3(brick, n, tc);
        if (std::isnan(v))
            continue;
        float v01 = (v - p->fmin) / range;
        float c[4];
        fetch_tf(tf, tf_n, v01, c);
        c[3] = 1.f - orc_powf(1.f - c[3] * d.opacity_factor, d.i_layer_count);
        if ((uint32_t)count >= max_samples)
            return -1;
        /* the taps of fetch3 */
        float t[3] = {tc.x - 0.5f, tc.y - 0.5f, tc.z - 0.5f};
        int i0[3], i1[3];
        for (int a = 0; a < 3; ++a) {
            float q = t[a];
            if (!(q > -1.f))
                q = -1.f;
            if (q > (float)n)
                q = (float)n;
            float fl = std::floor(q);
            weights[3 * count + a] = q - fl;
            int k = (int)fl;
            i0[a] = clampi(k, 0, (int)n - 1);
            i1[a] = clampi(k + 1, 0, (int)n - 1);
        }
        for (int k = 0; k < 8; ++k) {
            const int x = (k & 1) ? i1[0] : i0[0];
            const int y = (k & 2) ? i1[1] : i0[1];
            const int z = (k & 4) ? i1[2] : i0[2];
            corners[8 * count + k] = (uint32_t)(((size_t)z * n + y) * n + x);
        }
        ++count;
        float da = c[3] * (1.f - acc);
        acc = acc + da;
        if (acc > p->alpha_cutoff)
            break;
    }
    *seg_alpha = acc;
    return count;
}

/* ------------------------------------------------------------------------ */
/* Crop, blend, compose.                                                     */
/* ------------------------------------------------------------------------ */

extern "C" int orc_bound_rect(const uint32_t* pix, uint32_t w, uint32_t h, int32_t* rect4)
{
    /* getBoundRect: VsWorker_cu.cpp:51-88 (alpha byte = bits 24..31) */
    int maxX = 0, minX = (int)w, maxY = 0, minY = (int)h;
    for (int y = 0; y < (int)h; ++y)
        for (int x = 0; x < (int)w; ++x)
            if ((pix[(size_t)y * w + x] >> 24) != 0) {
                if (x < minX)
                    minX = x;
                if (x > maxX)
                    maxX = x;
                if (y < minY)
                    minY = y;
                if (y > maxY)
                    maxY = y;
            }
    if (minX > maxX || minY > maxY)
        return 0;
    rect4[0] = minX;
    rect4[1] = minY;
    rect4[2] = maxX - minX + 1;
    rect4[3] = maxY - minY + 1;
    return 1;
}

extern "C" void orc_blend_image_part(uint32_t* dst, int32_t dw, int32_t /*dh*/, const uint32_t* src,
                                     int32_t sw, int32_t sh, int32_t src_x, int32_t src_y,
                                     int32_t dst_off_x, int32_t dst_off_y, int just_copy)
{
    /* blendImagePart: BlendImagePart.hpp:61-145 */
    int dx = src_x - dst_off_x, dy = src_y - dst_off_y;
    auto getByte = [](uint32_t v, uint32_t n) { return (v >> (n << 3)) & 0xFFu; };
    auto setByte = [](uint32_t& d, uint32_t n, uint32_t s) {
        d &= ~(255u << (n << 3));
        d |= ((s & 255u) << (n << 3));
    };
    for (int y = 0; y < sh; ++y) {
        const uint32_t* srow = src + (size_t)y * sw;
        uint32_t* drow = dst + (size_t)dw * (dy + y) + dx;
        if (just_copy) {
            std::copy(srow, srow + sw, drow);
            continue;
        }
        for (int i = 0; i < sw; ++i) {
            uint32_t s = srow[i];
            uint32_t& d = drow[i];
            uint32_t as = getByte(s, 3);
            if (as == 0)
                continue;
            if (as == 255u) {
                d = s;
                continue;
            }
            uint32_t ad = getByte(d, 3);
            if (ad == 0) {
                d = s;
                continue;
            }
            uint32_t a = as + ad * (255u - as) / 255u;
            setByte(d, 3, a);
            for (uint32_t k = 0; k < 3; ++k) {
                uint32_t c = (getByte(s, k) * as + getByte(d, k) * (a - as)) / a;
                setByte(d, k, c);
            }
        }
    }
}

extern "C" uint32_t orc_background_argb(float r, float g, float b)
{
    return 0xff000000u + ((uint32_t)std::lround(r * 255) << 16) +
           ((uint32_t)std::lround(g * 255) << 8) + (uint32_t)std::lround(b * 255);
}

extern "C" void orc_color_transfer_vector(const float* xs, const float* colors, uint32_t count,
                                          uint32_t colorCount, float* ctf)
{
    /* makeColorTransferVector: ColorTransferFunction.hpp:27-69 */
    if (count == 0) {
        std::fill(ctf, ctf + 4 * colorCount, 1.f);
        return;
    }
    if (count == 1) {
        for (uint32_t i = 0; i < colorCount; ++i)
            for (int c = 0; c < 4; ++c)
                ctf[4 * i + c] = colors[c];
        return;
    }
    std::fill(ctf, ctf + 4 * colorCount, 0.f);
    float xmin = xs[0], xmax = xs[count - 1];
    float c1[4];
    for (int c = 0; c < 4; ++c)
        c1[c] = ctf[c] = colors[c];
    uint32_t i1 = 0, icolor = 1;
    for (uint32_t it = 1; it < count; ++it, ++icolor) {
        float x2 = xs[it];
        float c2[4];
        for (int c = 0; c < 4; ++c)
            c2[c] = colors[4 * it + c];
        uint32_t i2 = icolor + 1 == count ? colorCount - 1u
                                          : (uint32_t)((x2 - xmin) / (xmax - xmin) * colorCount);
        if (i2 > i1)
            for (uint32_t i = i1 + 1; i <= i2; ++i) {
                float t = (float)(i - i1) / (i2 - i1);
                for (int c = 0; c < 4; ++c)
                    ctf[4 * i + c] = c1[c] * (1 - t) + c2[c] * t;
            }
        for (int c = 0; c < 4; ++c)
            c1[c] = c2[c];
        i1 = i2;
    }
}

/* ------------------------------------------------------------------------ */
/* Whole worker share and float composite.                                   */
/* ------------------------------------------------------------------------ */

namespace {

/* boundingBoxOfIndexBox, BlockOrderUtil.hpp:29-51 */
void bbox_of_index_box(uint32_t level, const uint32_t* bmin, const uint32_t* bmax, float* cmin,
                       float* cmax)
{
    int cubesPerLevel = 1 << level;
    float half = kTopHalf / cubesPerLevel;
    auto cc = [&](uint32_t idx) { return -kTopHalf + half * (1 + (idx << 1)); };
    for (int a = 0; a < 3; ++a) {
        double lo = (double)cc(bmin[a]) - (double)half;
        double hi = (double)cc(bmax[a] - 1) + (double)half;
        cmin[a] = (float)lo;
        cmax[a] = (float)hi;
    }
}

/* eyeFromTransform, BlockOrderUtil.hpp:54-59: eye = -R^T t */
void eye_of(const float* M, float* eye)
{
    double t[3] = {M[3], M[7], M[11]};
    for (int i = 0; i < 3; ++i)
        eye[i] = (float)(-((double)M[i] * t[0] + (double)M[4 + i] * t[1] + (double)M[8 + i] * t[2]));
}

void parallel_rows(uint32_t H, uint32_t nthreads, const std::function<void(uint32_t, uint32_t)>& fn)
{
    if (nthreads <= 1) {
        fn(0, H);
        return;
    }
    std::vector<std::thread> th;
    uint32_t chunk = (H + nthreads - 1) / nthreads;
    for (uint32_t t = 0; t < nthreads; ++t) {
        uint32_t r0 = t * chunk, r1 = std::min(H, r0 + chunk);
        if (r0 >= r1)
            break;
        th.emplace_back(fn, r0, r1);
    }
    for (auto& x : th)
        x.join();
}

} // namespace

extern "C" void orc_render_share(const orc_frame_desc* d, const float* const* bricks,
                                 uint32_t* viewport, uint32_t nthreads)
{
    /* clearViewport: DenseFieldRenderer_cu.cu:268-277 */
    std::fill(viewport, viewport + (size_t)d->vp_w * d->vp_h, 0u);
    float cmin[3], cmax[3], eye[3];
    bbox_of_index_box(d->level, d->box_min, d->box_max, cmin, cmax);
    eye_of(d->view16, eye);
    Split splits[3];
    for (int a = 0; a < 3; ++a)
        splits[a] = UniformSplitter({d->box_min[a], d->box_max[a]}, cmin[a], cmax[a])(eye[a]);
    std::vector<std::array<uint32_t, 3>> order;
    b2f_visit(3, splits, [&](const uint32_t* pos) { order.push_back({pos[0], pos[1], pos[2]}); });
    std::vector<orc_block_params> params(order.size());
    std::vector<int> live(order.size());
    for (size_t i = 0; i < order.size(); ++i)
        live[i] = orc_block_params_setup(d->view16, d->vp_w, d->vp_h, d->fov_y_deg, d->level,
                                         order[i].data(), d->tex_depth, d->fmin, d->fmax,
                                         d->threshold_rel, d->quality, d->alpha_cutoff, &params[i]);
    uint32_t bx = d->box_max[0] - d->box_min[0], by = d->box_max[1] - d->box_min[1];
    /* Pixels are independent, so each thread renders every block (in order)
     * over its own band of viewport rows; the per-pixel operation sequence
     * is that of the sequential block loop, VsWorker_cu.cpp:269-277. */
    parallel_rows(d->vp_h, nthreads, [&](uint32_t r0, uint32_t r1) {
        for (size_t i = 0; i < order.size(); ++i) {
            if (!live[i])
                continue;
            const auto& o = order[i];
            size_t bi = ((size_t)(o[2] - d->box_min[2]) * by + (o[1] - d->box_min[1])) * bx +
                        (o[0] - d->box_min[0]);
            /* rows of the viewport buffer are bottom-up: pixel row = H-1-vy */
            uint32_t vy0 = d->vp_h - r1, vy1 = d->vp_h - r0;
            orc_render_block(&params[i], bricks[bi], d->tf, d->tf_n, viewport, vy0, vy1);
        }
    });
}

extern "C" void orc_render_float_composite_mode2(const float* view16, uint32_t vp_w, uint32_t vp_h,
                                                 float fov_y_deg, const orc_block_ref* blocks,
                                                 uint32_t nblocks, float fmin, float fmax,
                                                 float threshold_rel, float quality,
                                                 float alpha_cutoff, float acc_cutoff,
                                                 const float* tf, uint32_t tf_n, uint32_t mode,
                                                 float iso_opacity, const float* levels_rel,
                                                 uint32_t level_count, float smin, float smax,
                                                 float* out, uint32_t nthreads)
{
    const std::vector<float> lv = make_levels(fmin, fmax, levels_rel, level_count);
    std::vector<orc_block_params> params(nblocks);
    std::vector<Derived> der(nblocks);
    std::vector<int> live(nblocks);
    for (uint32_t i = 0; i < nblocks; ++i) {
        live[i] = orc_block_params_setup(view16, vp_w, vp_h, fov_y_deg, blocks[i].level,
                                         blocks[i].index, blocks[i].tex_depth, fmin, fmax,
                                         threshold_rel, quality, alpha_cutoff, &params[i]);
        der[i] = derive(&params[i]);
    }
    parallel_rows(vp_h, nthreads, [&](uint32_t r0, uint32_t r1) {
        for (uint32_t row = r0; row < r1; ++row) {
            uint32_t vy = vp_h - 1 - row;
            for (uint32_t vx = 0; vx < vp_w; ++vx) {
                float acc[4] = {0, 0, 0, 0};
                if (mode == kMip) {
                    /* the largest normalised value over the blocks */
                    uint32_t bits = 0x7fc00000u;
                    for (uint32_t bi = 0; bi < nblocks; ++bi) {
                        const orc_block_params* p = &params[bi];
                        if (!live[bi] || vx < p->view_x || vx >= p->view_x + p->view_w ||
                            vy < p->view_y || vy >= p->view_y + p->view_h)
                            continue;
                        Ray r = setup_ray(p, der[bi], vx, vy);
                        if (r.hit)
                            bits = mip_merge(bits, march_mip(p, der[bi], blocks[bi].brick, r), p);
                    }
                    float t;
                    std::memcpy(&t, &bits, 4);
                    float* o = out + 4 * ((size_t)row * vp_w + vx);
                    for (int k = 0; k < 4; ++k)
                        o[k] = t;
                    continue;
                }
                /* front to back: the reverse of the back-to-front list */
                for (uint32_t bi = nblocks; bi-- > 0;) {
                    if (acc[3] > acc_cutoff)
                        break; /* the pixel is opaque: nothing behind shows */
                    if (!live[bi])
                        continue;
                    const orc_block_params* p = &params[bi];
                    if (vx < p->view_x || vx >= p->view_x + p->view_w || vy < p->view_y ||
                        vy >= p->view_y + p->view_h)
                        continue;
                    Ray r = setup_ray(p, der[bi], vx, vy);
                    if (!r.hit)
                        continue;
                    float seg[4];
                    if (multi_level(mode) && lv.empty())
                        continue; /* no levels: nothing drawn */
                    segment_mode(p, der[bi], blocks[bi].brick, tf, tf_n, mode, iso_opacity, lv, r, seg,
                                 Second{blocks[bi].brick2, smin, smax});
                    /* acc (in front) composited over seg: blendBehind(acc, seg) */
                    float da = seg[3] * (1.f - acc[3]);
                    float a = acc[3] + da;
                    if (a > 0.f) {
                        float ka = acc[3] / a, kb = da / a;
                        for (int k = 0; k < 3; ++k)
                            acc[k] = acc[k] * ka + seg[k] * kb;
                    }
                    acc[3] = a;
                }
                float* o = out + 4 * ((size_t)row * vp_w + vx);
                for (int k = 0; k < 4; ++k)
                    o[k] = acc[k];
            }
        }
    });
}

extern "C" void orc_render_float_composite_mode(const float* view16, uint32_t vp_w, uint32_t vp_h,
                                                float fov_y_deg, const orc_block_ref* blocks,
                                                uint32_t nblocks, float fmin, float fmax,
                                                float threshold_rel, float quality,
                                                float alpha_cutoff, float acc_cutoff,
                                                const float* tf, uint32_t tf_n, uint32_t mode,
                                                float iso_opacity, const float* levels_rel,
                                                uint32_t level_count, float* out, uint32_t nthreads)
{
    orc_render_float_composite_mode2(view16, vp_w, vp_h, fov_y_deg, blocks, nblocks, fmin, fmax,
                                     threshold_rel, quality, alpha_cutoff, acc_cutoff, tf, tf_n, mode,
                                     iso_opacity, levels_rel, level_count, 0.f, 1.f, out, nthreads);
}

extern "C" void orc_render_float_composite(const float* view16, uint32_t vp_w, uint32_t vp_h,
                                           float fov_y_deg, const orc_block_ref* blocks,
                                           uint32_t nblocks, float fmin, float fmax,
                                           float threshold_rel, float quality, float alpha_cutoff,
                                           float acc_cutoff, const float* tf, uint32_t tf_n,
                                           float* out, uint32_t nthreads)
{
    orc_render_float_composite_mode(view16, vp_w, vp_h, fov_y_deg, blocks, nblocks, fmin, fmax,
                                    threshold_rel, quality, alpha_cutoff, acc_cutoff, tf, tf_n, kArgb,
                                    1.f, nullptr, 0, out, nthreads);
}
