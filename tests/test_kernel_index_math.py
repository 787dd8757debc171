"""Host check of the integer arithmetic the interpolation kernel does in fp32:
it splits an item number r into grid coordinates with q = trunc((r + 0.5) *
rcp(d)), where rcp is the hardware reciprocal (at most 1 ulp off).  Every
divisor and item range the kernel can meet is checked here with the
reciprocal perturbed by +-2 ulp and the product rounded to fp32."""
import numpy as np


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
