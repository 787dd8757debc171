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
# with numpy: per tile, lane q of the loop stores micro-brick row q & 15 of
# micro-brick q >> 4 of the owned box, the wave's first micro-brick
# decomposed once and the lane's offset 0..3 carried along x.
def _tiled_store_cover(d):
    n = (1 << d) + 1; t = min(d, 4); T = 1 << t; X = (1 << d) if d < 6 else 64
    tiles_x = (1 << d) // X; tiles_yz = (1 << d) // T
    G = (n + 3) // 4
    written = np.zeros(G**3 * 64, np.int32)
    owner = -np.ones(G**3 * 64, np.int64)
    rowlen, tp1 = X + 1, T + 1
    for tz in range(tiles_yz):
        for ty in range(tiles_yz):
            for tx in range(tiles_x):
                x0, y0, z0 = tx * X, ty * T, tz * T
                ex = rowlen if tx + 1 == tiles_x else X
                ey = tp1 if ty + 1 == tiles_yz else T
                ez = tp1 if tz + 1 == tiles_yz else T
                mx0, my0, mz0 = x0 >> 2, y0 >> 2, z0 >> 2
                nmx = ((x0 + ex + 3) >> 2) - mx0; nmy = ((y0 + ey + 3) >> 2) - my0; nmz = ((z0 + ez + 3) >> 2) - mz0
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
                y = (my << 2) | (r & 3); z = (mz << 2) | (r >> 2); xb = mx << 2
                row_ok = (y < y0 + ey) & (z < z0 + ez)
                base = (((mz * G + my) * G + mx) << 6) + (r << 2)
                for c in range(4):
                    addr = base + c
                    assert (addr >= 0).all() and (addr < G**3 * 64).all(), d
                    valid = row_ok & (xb + c < x0 + ex)
                    lds = ((z - z0) * tp1 + (y - y0)) * rowlen + (xb - x0) + c
                    assert (lds[valid] < rowlen * tp1 * tp1).all() and (lds[valid] >= 0).all()
                    np.add.at(written, addr, 1)
                    tile_id = (tz * tiles_yz + ty) * tiles_x + tx
                    assert (owner[addr] == -1).all(), ("two tiles write one node", d)
                    owner[addr] = tile_id
                    # the node's true coordinates and validity
                    x = xb + c
                    real = (x < n) & (y < n) & (z < n)
                    assert (valid == real).all(), ("valid != real node", d)
                    # tiled index of (x,y,z) equals the address
                    ti = ((((z >> 2) * G + (y >> 2)) * G + (x >> 2)) << 6) + (((z & 3) << 4) | ((y & 3) << 2) | (x & 3))
                    assert (ti == addr).all()
    assert (written == 1).all(), d


@pytest.mark.parametrize("d", range(0, 9))
def test_tiled_store_writes_every_node_once(d):
    """Every node of every micro-brick (padding included) is stored exactly
    once, by one tile; the lanes that read LDS are exactly the brick's real
    nodes, inside the tile's box; stores stay inside the brick."""
    _tiled_store_cover(d)
