// Tile geometry of the brick interpolation, shared by the plan encoder (host)
// and the kernels.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define RV_HD __host__ __device__
#else
#define RV_HD
#endif

// Tile extent in y and z is 2^tlog cells for bricks of depth <= 7 and 16 cells
// for depth 8 (a finer tile would need a coarse lattice of more than 17^3
// nodes).  Selected at build time, -DRV_INTERP_TLOG=3 or 4.
#ifndef RV_INTERP_TLOG
#define RV_INTERP_TLOG 4
#endif

namespace rv {

constexpr uint32_t kPlanMagic = 0x33504e52u; // "RNP3"
constexpr uint32_t kPlanHeader = 64;          // bytes before tile_start[]

struct Geom {
    uint32_t d, n, t, T, X, tiles_x, tiles_yz, ntiles, m;
};

RV_HD inline uint32_t tile_log2_for(uint32_t d) { return d >= 8u ? 4u : (uint32_t)RV_INTERP_TLOG; }

// Tiles of X x T x T cells, X = min(2^d, 64), T = min(2^d, 2^tlog) (x spans
// whole brick rows up to 64 cells); m = cells per axis of the stride-T coarse
// lattice of the whole brick.
RV_HD inline Geom geom_of(uint32_t d)
{
    const uint32_t tlog = tile_log2_for(d);
    Geom g;
    g.d = d;
    g.n = (1u << d) + 1u;
    g.t = d < tlog ? d : tlog;
    g.T = 1u << g.t;
    g.X = d < 6u ? (1u << d) : 64u;
    g.tiles_x = (1u << d) / g.X;
    g.tiles_yz = (1u << d) / g.T;
    g.ntiles = g.tiles_x * g.tiles_yz * g.tiles_yz;
    g.m = (1u << d) >> g.t;
    return g;
}

// Index of node (x, y, z) of a brick of n nodes per axis in the micro-brick
// layout (G = floor((n + 1) / 4) + 1): micro-bricks of 4 x 4 x 4 cells, 80
// floats (16 rows of 5 x-positions; micro-brick m holds x-positions 4m ..
// 4m + 4, so neighbours share one), ordered x fastest over G^3 of them;
// along x node x is at position x + 1 (positions 0 and n .. 4G repeat the
// edge nodes), along y and z at y, z.  This is the copy in micro-brick
// (x + 1) / 4 (the other copy of a position 4m is the last of micro-brick
// m - 1).
RV_HD inline uint32_t tiled4_index(uint32_t x, uint32_t y, uint32_t z, uint32_t G)
{
    const uint32_t px = x + 1u;
    return (((z >> 2) * G + (y >> 2)) * G + (px >> 2)) * 80u + (z & 3u) * 20u + (y & 3u) * 5u + (px & 3u);
}

// LDS floats a tile needs: its (X+1)(T+1)^2 box, or the (m+1)^3 coarse
// lattice when that is larger.
RV_HD inline uint32_t tile_floats(const Geom& g)
{
    const uint32_t box = (g.X + 1u) * (g.T + 1u) * (g.T + 1u);
    const uint32_t lat = g.m >= 2u ? (g.m + 1u) * (g.m + 1u) * (g.m + 1u) : 0u;
    return box > lat ? box : lat;
}

} // namespace rv
