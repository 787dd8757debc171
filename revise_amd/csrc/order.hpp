// Block ordering and per-block parameter setup (host C++).
//
// BackToFront3 enumerates the blocks of an index box in painter's order for
// an eye point, following the macro-cell scheme of BackToFrontOrder
// (BackToFrontOrder.hpp:61-359): each axis is split at the eye into
// (before, at, after); the 27 macro cells are visited corners first, then
// edges, faces and the centre; inside a cell blocks are swept along
// diagonals of increasing index sum, mirrored on axes where the cell lies
// after the eye.  The visit sequence must equal the reference's exactly.
#pragma once

#include "../../include/revise_hip.h"

#include <algorithm>
#include <cstdint>
#include <vector>

namespace rv {

struct Range {
    uint32_t lo = 0, hi = 0; // [lo, hi)
    bool empty() const { return lo == hi; }
};

struct AxisSplit {
    Range before;
    uint32_t at = ~0u; // ~0u: the eye is outside the range
    Range after;
};

// UniformIndexRangeSplitter (UniformIndexRangeSplitter.hpp:27-57).
class UniformSplitter {
public:
    UniformSplitter(uint32_t i0, uint32_t i1, float c0, float c1)
        : i0_(i0), i1_(i1), c0_(c0), c1_(c1), factor_((float)(i1 - i0) / (c1 - c0))
    {
    }
    AxisSplit split(float x) const
    {
        AxisSplit s;
        if (x <= c0_)
            s.after = {i0_, i1_};
        else if (x >= c1_)
            s.before = {i0_, i1_};
        else {
            const uint32_t at = (uint32_t)((x - c0_) * factor_) + i0_;
            s.before = {i0_, at};
            s.at = at;
            s.after = {at + 1, i1_};
        }
        return s;
    }

private:
    uint32_t i0_, i1_;
    float c0_, c1_, factor_;
};

// NonUniformIndexRangeSplitter (NonUniformIndexRangeSplitter.hpp:31-77).
class BoundarySplitter {
public:
    BoundarySplitter() = default;
    BoundarySplitter(uint32_t start, std::vector<float> bounds) : start_(start), b_(std::move(bounds)) {}
    AxisSplit split(float x) const
    {
        AxisSplit s;
        const uint32_t cells = (uint32_t)b_.size() - 1u;
        const auto it = std::lower_bound(b_.begin(), b_.end(), x);
        if (it == b_.begin())
            s.after = {start_, start_ + cells};
        else if (it == b_.end())
            s.before = {start_, start_ + cells};
        else {
            const uint32_t at = start_ + (uint32_t)(it - b_.begin()) - 1u;
            s.before = {start_, at};
            s.at = at;
            s.after = {at + 1, start_ + cells};
        }
        return s;
    }
    uint32_t end() const { return start_ + (uint32_t)b_.size() - 1u; }

private:
    uint32_t start_ = 0;
    std::vector<float> b_;
};

// IndexCubeSplitter<3> over [-5,5]^3 (IndexCubeSplitter.hpp:35-196): the
// 2^level cube is cut into at most split_count boxes, x first, then y, z.
struct CubeSplit {
    std::vector<uint32_t> cuts[3];
    BoundarySplitter axis[3];
    CubeSplit(uint32_t level, uint32_t split_count);
    void box(const uint32_t* pos, uint32_t* box6) const
    {
        for (int a = 0; a < 3; ++a) {
            box6[a] = cuts[a][pos[a]];
            box6[3 + a] = cuts[a][pos[a] + 1];
        }
    }
};

// Painter's-order enumerator over the three per-axis splits.
class BackToFront3 {
public:
    explicit BackToFront3(const AxisSplit* s)
    {
        for (int a = 0; a < 3; ++a)
            s_[a] = s[a];
        advance_cell();
    }
    // Writes the next block index; false when done.
    bool next(uint32_t* out)
    {
        if (done_)
            return false;
        for (int a = 0; a < 3; ++a)
            out[a] = cell_[a] == 2 ? hi_[a] - 1 + lo_[a] - p_[a] : p_[a];
        if (!step())
            advance_cell();
        return true;
    }

private:
    AxisSplit s_[3];
    int ord_ = -1;
    bool done_ = false;
    uint32_t cell_[3] = {0, 0, 0};
    uint32_t lo_[3] = {0, 0, 0}, hi_[3] = {0, 0, 0}, p_[3] = {0, 0, 0};

    Range cell_range(int a, uint32_t c) const
    {
        const AxisSplit& s = s_[a];
        return c == 0 ? s.before : (c == 1 ? Range{s.at, s.at + 1} : s.after);
    }
    bool cell_valid(int a, uint32_t c) const
    {
        const AxisSplit& s = s_[a];
        return c == 0 ? !s.before.empty() : (c == 1 ? s.at != ~0u : !s.after.empty());
    }
    void advance_cell()
    {
        // macro cells: 8 corners, 12 edges, 6 faces, centre (cell = x + 3y + 9z)
        static const uint8_t kOrder[27] = {0,  2,  6,  8,  18, 20, 24, 26, 1,  3,  5,  7,  9, 11,
                                           15, 17, 19, 21, 23, 25, 4,  10, 12, 14, 16, 22, 13};
        while (++ord_ < 27) {
            const uint32_t c = kOrder[ord_];
            cell_[0] = c % 3u;
            cell_[1] = (c / 3u) % 3u;
            cell_[2] = c / 9u;
            if (cell_valid(0, cell_[0]) && cell_valid(1, cell_[1]) && cell_valid(2, cell_[2])) {
                for (int a = 0; a < 3; ++a) {
                    const Range r = cell_range(a, cell_[a]);
                    lo_[a] = p_[a] = r.lo;
                    hi_[a] = r.hi;
                }
                return;
            }
        }
        done_ = true;
    }
    // Next position of the diagonal sweep inside the current macro cell
    // (unsigned arithmetic, BackToFrontOrder.hpp:268-309).
    bool step()
    {
        uint32_t s = p_[0] + p_[1] + p_[2];
        if (s + 3 == hi_[0] + hi_[1] + hi_[2])
            return false;
        if (p_[0] > lo_[0] && p_[1] + 1 < hi_[1]) {
            --p_[0];
            ++p_[1];
            return true;
        }
        if (p_[0] + p_[1] > lo_[0] + lo_[1] && p_[2] + 1 < hi_[2]) {
            const uint32_t z = p_[2] + 1;
            if (s - lo_[1] - p_[2] < hi_[0] + 1) {
                p_[0] = s - lo_[1] - p_[2] - 1;
                p_[1] = lo_[1];
            } else {
                p_[0] = hi_[0] - 1;
                p_[1] = s - hi_[0] - p_[2];
            }
            p_[2] = z;
            return true;
        }
        ++s;
        const uint32_t x0 = s - lo_[1] - lo_[2];
        if (x0 < hi_[0]) {
            p_[0] = x0;
            p_[1] = lo_[1];
            p_[2] = lo_[2];
            return true;
        }
        const uint32_t x1 = s - hi_[0] + 1 - lo_[2];
        if (x1 < hi_[1]) {
            p_[0] = hi_[0] - 1;
            p_[1] = x1;
            p_[2] = lo_[2];
            return true;
        }
        p_[0] = hi_[0] - 1;
        p_[1] = hi_[1] - 1;
        p_[2] = s + 2 - hi_[0] - hi_[1];
        return true;
    }
};

void eye_from_view(const float* M, float* eye);
void index_box_coords(uint32_t level, const uint32_t* bmin, const uint32_t* bmax, float* cmin,
                      float* cmax);
int block_params(const float* M, uint32_t vp_w, uint32_t vp_h, float fov_y_deg, uint32_t level,
                 const uint32_t* index, uint32_t tex_depth, float fmin, float fmax,
                 float threshold_rel, float quality, float alpha_cutoff, rv_block_params& p);
int mixed_back_to_front(const uint32_t* blocks, uint64_t n, const float* eye,
                        std::vector<uint32_t>& order);

} // namespace rv
