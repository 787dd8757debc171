"""Host check of the integer arithmetic the interpolation kernel does in fp32:
it splits an item number r into grid coordinates with q = trunc((r + 0.5) *
rcp(d)), where rcp is the hardware reciprocal (at most 1 ulp off).  Every
divisor and item range the kernel can meet is checked here with the
reciprocal perturbed by +-2 ulp and the product rounded to fp32.  Also
the micro-brick store's index arithmetic."""
import numpy as np
import pytest


def _divisors():
    out = set()
    # refinement boxes: X in {1..64} cells (powers of two), T <= 16; coarse lattice m <= 16
    for X, T in [(1 << a, 1 << b) for a in range(7) for b in range(5)] + [(m, m) for m in (2, 4, 8, 16)]:
        s = 2
        while s <= max(X, T):
            qx, qy = X // s, T // s
            for m in range(1, 8):
                cx = qx if m & 1 else qx + 1
                cy = qy if m & 2 else qy + 1
                cz = qy if m & 4 else qy + 1
                if cx and cy and cz:
                    out.add((cx, cx * cy, cz))
            s *= 2
    return sorted(out)


def test_fp32_reciprocal_division_is_exact():
    for cx, cxy, cz in _divisors():
        r = np.arange(cxy * cz, dtype=np.int64)
        for d, rr in ((cxy, r), (cx, r % cxy)):
            rc = np.float32(1.0) / np.float32(d)
            for ulp in (-2, -1, 0, 1, 2):
                rcp = np.float32(rc * (1 + ulp * 2.0 ** -23)) if ulp else rc
                q = np.floor((rr.astype(np.float32) + np.float32(0.5)) * rcp).astype(np.int64)
                assert np.array_equal(q, rr // d), (cx, cxy, cz, d, ulp)


# The micro-brick store of interp_tile (interp_device.hpp, job.tiled), restated
# with numpy: per tile, lane q of the loop stores the positions it owns of
# row q & 15 (5 x-positions) of micro-brick q >> 4 of the tile's range, the
# wave's first micro-brick decomposed once and the lane's offset 0..3
# carried along x.  Along x node c is at position c + 1, micro-brick m holds
# positions 4m .. 4m + 4, and positions 0 and n .. 4G repeat the edge nodes.
def _tiled_store_cover(d):
    n = (1 << d) + 1; t = min(d, 4); T = 1 << t; X = (1 << d) if d < 6 else 64
    tiles_x = (1 << d) // X; tiles_yz = (1 << d) // T
    G = ((n + 1) >> 2) + 1
    size = G**3 * 80
    written = np.zeros(size, np.int32)
    owner = -np.ones(size, np.int64)
    rowlen, tp1 = X + 1, T + 1
    for tz in range(tiles_yz):
        for ty in range(tiles_yz):
            for tx in range(tiles_x):
                x0, y0, z0 = tx * X, ty * T, tz * T
                lastx, lasty, lastz = tx + 1 == tiles_x, ty + 1 == tiles_yz, tz + 1 == tiles_yz
                ex = rowlen if lastx else X
                ey = tp1 if lasty else T
                ez = tp1 if lastz else T
                p_lo = 0 if tx == 0 else x0 + 1
                p_hi = 4 * G + 1 if lastx else x0 + ex + 1
                mx0 = 0 if p_lo == 0 else (p_lo - 1) >> 2
                my0, mz0 = y0 >> 2, z0 >> 2
                nmx = min((p_hi - 1) >> 2, G - 1) + 1 - mx0
                nmy = (G if lasty else ((y0 + ey + 3) >> 2)) - my0
                nmz = (G if lastz else ((z0 + ez + 3) >> 2)) - mz0
                total = nmx * nmy * nmz * 16
                q = np.arange(total, dtype=np.int64)
                mbw = (q >> 6) << 2; r = q & 15
                mx = mbw % nmx; my = (mbw // nmx) % nmy; mz = mbw // (nmx * nmy)
                mx = mx + ((q >> 4) & 3)
                for _ in range(4):  # the carry loop, at most 3 steps
                    w = mx >= nmx
                    mx = np.where(w, mx - nmx, mx); my = np.where(w, my + 1, my)
                    w2 = my == nmy
                    my = np.where(w2, 0, my); mz = np.where(w2, mz + 1, mz)
                assert (mx < nmx).all()
                # equals the direct decomposition
                mb = q >> 4
                assert (mx == mb % nmx).all() and (my == (mb // nmx) % nmy).all() and (mz == mb // (nmx * nmy)).all()
                mx += mx0; my += my0; mz += mz0
                y = (my << 2) | (r & 3); z = (mz << 2) | (r >> 2); pb = mx << 2
                row_ok = (y < y0 + ey) & (z < z0 + ez)
                base = ((mz * G + my) * G + mx) * 80 + r * 5
                tile_id = (tz * tiles_yz + ty) * tiles_x + tx
                for c in range(5):
                    addr = base + c
                    assert (addr >= 0).all() and (addr < size).all(), d
                    pos = pb + c
                    own = (pos >= p_lo) & (pos < p_hi)
                    node = np.clip(pos - 1, 0, n - 1)
                    valid = row_ok & own
                    # owned positions map to nodes this tile holds in LDS
                    assert ((node[own] >= x0) & (node[own] < x0 + ex)).all(), d
                    lds = ((z - z0) * tp1 + (y - y0)) * rowlen + (node - x0)
                    assert (lds[valid] < rowlen * tp1 * tp1).all() and (lds[valid] >= 0).all()
                    np.add.at(written, addr[own], 1)
                    assert (owner[addr[own]] == -1).all(), ("two tiles write one float", d)
                    owner[addr[own]] = tile_id
                    # the LDS is read exactly for the brick's real rows
                    real = (y < n) & (z < n)
                    assert (valid == (real & own)).all(), ("valid != real row", d)
    assert (written == 1).all(), d


@pytest.mark.parametrize("d", range(0, 9))
def test_tiled_store_writes_every_node_once(d):
    """Every float of every micro-brick (edge copies and padding included)
    is stored exactly once, by one tile (x-tiles share the rows at their
    seams float by float); the lanes that read LDS are exactly the brick's
    real rows, inside the tile's box; stores stay inside the brick."""
    _tiled_store_cover(d)
