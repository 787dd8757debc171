// Gather plan of a subtree (host encoder).
//
// The interpolation kernel completes a brick in tiles of X x T x T cells
// (interp_geom.hpp).  The plan lists, per tile, the sparse nodes that lie in
// the tile's closed box, each with its offset in the tile's LDS array; nodes
// on faces shared by tiles appear once per tile.  A tile's list holds first
// the nodes the tile stores (its half-open box, plus the brick's far faces
// for the last tiles), then the others (the faces it shares with the next
// tiles), each part in increasing sparse index (so the kernel's reads of
// sparse values and of the plan itself are coalesced): a tile whose nodes
// are all given needs only the first part.  The plan also lists, for the
// stride-T lattice of the whole brick, the sparse index of each lattice node
// (or none), for the coarse levels.
//
// Layout (little endian, all offsets in bytes from the plan start):
//   u32 magic, u32 depth, u32 ntiles, u32 nlat,
//   u32 entries, u32 lat_off, u32 idx_off, u32 ldsoff_off,
//   u32 T, u32 X, u32 own_off, u32 reserved[5]  (64-byte header)
//   u32 tile_start[ntiles + 1]          entry range of each tile
//   u32 own_end[ntiles]                 end of each tile's stored-node part
//   u32 lat_idx[nlat]                   lattice node -> sparse index / ~0
//   u32 idx[entries]                    sparse index of each entry
//   u16 lds[entries]                    LDS offset of each entry
// A plan depends only on the subtree's n2i (like BlockTreeNodes, which the
// reference caches per subtree, BlockTreeFieldService.hpp:133-137).
#include "../../include/revise_hip.h"
#include "interp_geom.hpp"

#include <cstring>
#include <string>
#include <vector>

namespace rv {
void set_error(const std::string& msg);
}

namespace {

using rv::Geom;
using rv::geom_of;

inline uint32_t align16(uint32_t v) { return (v + 15u) & ~15u; }

// Tiles along one axis containing coordinate c (tile extent e, count n).
inline int tiles_of(uint32_t c, uint32_t e, uint32_t n, uint32_t* out)
{
    int k = 0;
    const uint32_t t = c / e;
    if (t < n)
        out[k++] = t;
    if (c % e == 0 && c > 0 && t - 1 < n)
        out[k++] = t - 1;
    return k;
}

struct Encoded {
    Geom g;
    std::vector<uint32_t> start, own_end, lat, idx;
    std::vector<uint16_t> lds;
    uint32_t own_off, lat_off, idx_off, lds_off, bytes;
};

// Whether tile-local coordinate l (0..extent) is stored by the tile: below
// its extent, or the brick's far face in the last tile.
inline bool stores(uint32_t l, uint32_t extent, bool last) { return l < extent || last; }

int encode(const uint16_t* n2i, uint64_t n, uint32_t depth, Encoded& e)
{
    if (depth > 8) {
        rv::set_error("rv_interp_plan: depth must be <= 8");
        return RV_ERR_INVALID;
    }
    if (n > 0xFFFFFFF0ull) {
        rv::set_error("rv_interp_plan: too many nodes");
        return RV_ERR_INVALID;
    }
    e.g = geom_of(depth);
    const Geom& g = e.g;
    const uint32_t npe = (1u << depth) + 1u;
    const uint32_t rowlen = g.X + 1u, tp1 = g.T + 1u;
    // A coordinate listed twice keeps its last sparse value, as the
    // reference's scatter loop does (later writes win).
    std::vector<bool> keep(n, true);
    {
        std::vector<bool> seen((uint64_t)npe * npe * npe, false);
        for (uint64_t i = n; i-- > 0;) {
            const uint32_t x = n2i[3 * i], y = n2i[3 * i + 1], z = n2i[3 * i + 2];
            if (x >= npe || y >= npe || z >= npe) {
                rv::set_error("rv_interp_plan: node coordinate outside the brick");
                return RV_ERR_INVALID;
            }
            const uint64_t o = ((uint64_t)z * npe + y) * npe + x;
            keep[i] = !seen[o];
            seen[o] = true;
        }
    }
    std::vector<uint32_t> count(g.ntiles + 1, 0), owned(g.ntiles, 0);
    // pass 1: counts
    for (uint64_t i = 0; i < n; ++i) {
        if (!keep[i])
            continue;
        const uint32_t x = n2i[3 * i], y = n2i[3 * i + 1], z = n2i[3 * i + 2];
        uint32_t ax[2], ay[2], az[2];
        const int nx = tiles_of(x, g.X, g.tiles_x, ax), ny = tiles_of(y, g.T, g.tiles_yz, ay),
                  nz = tiles_of(z, g.T, g.tiles_yz, az);
        for (int a = 0; a < nz; ++a)
            for (int b = 0; b < ny; ++b)
                for (int c = 0; c < nx; ++c) {
                    const uint32_t t = (az[a] * g.tiles_yz + ay[b]) * g.tiles_x + ax[c];
                    ++count[t + 1];
                    if (stores(x - ax[c] * g.X, g.X, ax[c] + 1 == g.tiles_x) &&
                        stores(y - ay[b] * g.T, g.T, ay[b] + 1 == g.tiles_yz) &&
                        stores(z - az[a] * g.T, g.T, az[a] + 1 == g.tiles_yz))
                        ++owned[t];
                }
    }
    e.start.assign(g.ntiles + 1, 0);
    for (uint32_t t = 0; t < g.ntiles; ++t)
        e.start[t + 1] = e.start[t] + count[t + 1];
    const uint32_t entries = e.start[g.ntiles];
    e.idx.assign(entries, 0);
    e.lds.assign(entries, 0);
    e.own_end.assign(g.ntiles, 0);
    std::vector<uint32_t> cursor(e.start.begin(), e.start.end() - 1), hcursor(g.ntiles);
    for (uint32_t t = 0; t < g.ntiles; ++t) {
        e.own_end[t] = e.start[t] + owned[t];
        hcursor[t] = e.own_end[t];
    }
    // pass 2: fill (stored nodes, then shared faces, each in increasing sparse index)
    for (uint64_t i = 0; i < n; ++i) {
        if (!keep[i])
            continue;
        const uint32_t x = n2i[3 * i], y = n2i[3 * i + 1], z = n2i[3 * i + 2];
        uint32_t ax[2], ay[2], az[2];
        const int nx = tiles_of(x, g.X, g.tiles_x, ax), ny = tiles_of(y, g.T, g.tiles_yz, ay),
                  nz = tiles_of(z, g.T, g.tiles_yz, az);
        for (int a = 0; a < nz; ++a)
            for (int b = 0; b < ny; ++b)
                for (int c = 0; c < nx; ++c) {
                    const uint32_t t = (az[a] * g.tiles_yz + ay[b]) * g.tiles_x + ax[c];
                    const uint32_t lx = x - ax[c] * g.X, ly = y - ay[b] * g.T, lz = z - az[a] * g.T;
                    const bool own = stores(lx, g.X, ax[c] + 1 == g.tiles_x) &&
                                     stores(ly, g.T, ay[b] + 1 == g.tiles_yz) &&
                                     stores(lz, g.T, az[a] + 1 == g.tiles_yz);
                    const uint32_t k = own ? cursor[t]++ : hcursor[t]++;
                    e.idx[k] = (uint32_t)i;
                    e.lds[k] = (uint16_t)((lz * tp1 + ly) * rowlen + lx);
                }
    }
    // lattice of the coarse levels
    const uint32_t mp = g.m + 1u;
    e.lat.assign(g.m >= 2u ? mp * mp * mp : 0u, 0xFFFFFFFFu);
    if (g.m >= 2u)
        for (uint64_t i = 0; i < n; ++i) {
            const uint32_t x = n2i[3 * i], y = n2i[3 * i + 1], z = n2i[3 * i + 2];
            if (x % g.T || y % g.T || z % g.T)
                continue;
            e.lat[((z / g.T) * mp + y / g.T) * mp + x / g.T] = (uint32_t)i;
        }
    const uint32_t head = rv::kPlanHeader + 4u * (g.ntiles + 1u);
    e.own_off = align16(head);
    e.lat_off = align16(e.own_off + 4u * g.ntiles);
    e.idx_off = align16(e.lat_off + 4u * (uint32_t)e.lat.size());
    e.lds_off = align16(e.idx_off + 4u * entries);
    e.bytes = align16(e.lds_off + 2u * entries);
    return RV_OK;
}

} // namespace

extern "C" uint64_t rv_interp_plan_bytes(const uint16_t* n2i, uint64_t n, uint32_t depth)
{
    Encoded e;
    if ((n && !n2i) || encode(n2i, n, depth, e) != RV_OK)
        return 0;
    return e.bytes;
}

extern "C" int rv_interp_plan_encode(const uint16_t* n2i, uint64_t n, uint32_t depth, void* plan,
                                     uint64_t bytes)
{
    if ((n && !n2i) || !plan) {
        rv::set_error("rv_interp_plan_encode: null pointer");
        return RV_ERR_INVALID;
    }
    Encoded e;
    int rc = encode(n2i, n, depth, e);
    if (rc)
        return rc;
    if (bytes < e.bytes) {
        rv::set_error("rv_interp_plan_encode: buffer too small");
        return RV_ERR_INVALID;
    }
    auto* p = static_cast<uint8_t*>(plan);
    std::memset(p, 0, e.bytes);
    const uint32_t head[11] = {rv::kPlanMagic, depth, e.g.ntiles, (uint32_t)e.lat.size(),
                               (uint32_t)e.idx.size(), e.lat_off, e.idx_off, e.lds_off,
                               e.g.T, e.g.X, e.own_off};
    std::memcpy(p, head, sizeof head);
    std::memcpy(p + rv::kPlanHeader, e.start.data(), 4 * e.start.size());
    std::memcpy(p + e.own_off, e.own_end.data(), 4 * e.own_end.size());
    if (!e.lat.empty())
        std::memcpy(p + e.lat_off, e.lat.data(), 4 * e.lat.size());
    if (!e.idx.empty()) {
        std::memcpy(p + e.idx_off, e.idx.data(), 4 * e.idx.size());
        std::memcpy(p + e.lds_off, e.lds.data(), 2 * e.lds.size());
    }
    return RV_OK;
}
