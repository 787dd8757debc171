"""The oracle's ordering and splitting against the golden vectors of the
reference's own unit tests (tests/test_BackToFrontOrder.cpp,
tests/test_IndexRangeSplitter.cpp, tests/test_IndexCubeSplitter.cpp)."""
import json
import os

import numpy as np
import pytest

from tests.b2fo_format import format_order

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _cases():
    return json.load(open(os.path.join(GOLD, "b2fo_golden.json")))


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c["source"])
def test_b2fo_golden(orc, case):
    dim, size, eye = case["dim"], case["size"], case["eye"]
    imin, imax = [0] * dim, [size] * dim
    cmin, cmax = [0.0] * dim, [1.0] * dim
    order = orc.b2f_macro(imin, imax, cmin, cmax, eye)
    assert order.shape[0] == size ** dim
    eye_idx = tuple(orc.uniform_split(0, size, 0.0, 1.0, e)[1] for e in eye)
    got = format_order(dim, [tuple(int(v) for v in b) for b in order], eye_idx)
    assert got == case["expected"]


def test_splitter_golden(orc):
    g = json.load(open(os.path.join(GOLD, "splitter_golden.json")))
    u = g["uniform"]
    for c in u["cases"]:
        got = orc.uniform_split(u["index_range"][0], u["index_range"][1], u["coord_range"][0],
                                u["coord_range"][1], c["x"])
        assert got == (c["before"], c["at"], c["after"]), c
    nu = g["nonuniform"]
    for c in nu["cases"]:
        got = orc.nonuniform_split(nu["start"], nu["split_coords"], nu["coord_range"][0],
                                   nu["coord_range"][1], c["x"])
        assert got == (c["before"], c["at"], c["after"]), c


def _union_all(boxes):
    """Merge adjacent boxes (tests/test_IndexCubeSplitter.cpp:33-97)."""
    boxes = [b for b in boxes]
    def unite(b1, b2):
        if b2 is None:
            return True, b1
        if b1 is None:
            return True, b2
        for i in range(3):
            i1, i2 = (i + 1) % 3, (i + 2) % 3
            if not all(b1[j] == b2[j] and b1[3 + j] == b2[3 + j] for j in (i1, i2)):
                continue
            if b1[3 + i] == b2[i]:
                n = list(b1); n[3 + i] = b2[3 + i]; return True, tuple(n)
            if b1[i] == b2[3 + i]:
                n = list(b1); n[i] = b2[i]; return True, tuple(n)
        return False, b1
    while True:
        united = False
        i = 0
        while i < len(boxes):
            j = i + 1
            while j < len(boxes):
                ok, nb = unite(boxes[i], boxes[j])
                if ok:
                    boxes[i] = nb
                    del boxes[j]
                    united = True
                else:
                    j += 1
            i += 1
        if not united:
            return boxes


@pytest.mark.parametrize("level,count", [(0, 1), (0, 2), (2, 2), (2, 3), (2, 5), (2, 9), (2, 64),
                                         (2, 100), (3, 32), (3, 64), (3, 512), (4, 1024)])
def test_index_cube_split_consistency(orc, level, count):
    """tests/test_IndexCubeSplitter.cpp:148-167: the split boxes (padded
    with empty ones to split_count) unite into the whole 2^level cube."""
    res = orc.index_cube_split_all(level, count, True)
    assert res.shape[0] == count
    boxes = [None if not b.any() else tuple(int(v) for v in b) for b in res]
    u = [b for b in _union_all(boxes) if b is not None]
    assert u == [(0, 0, 0, 1 << level, 1 << level, 1 << level)]


@pytest.mark.parametrize("level,count", [(2, 2), (2, 4), (2, 8), (2, 16), (2, 32), (2, 64), (0, 1),
                                         (1, 1), (1, 2), (1, 4), (1, 8), (3, 32), (3, 64), (3, 512),
                                         (4, 1024)])
def test_index_cube_split_equal_sizes(orc, level, count):
    """tests/test_IndexCubeSplitter.cpp:169-199"""
    res = orc.index_cube_split_all(level, count, True)
    sizes = {tuple(b[3:] - b[:3]) for b in res}
    assert len(sizes) == 1


def test_b2f_is_visibility_order_bruteforce(orc):
    """Brute-force invariant on small grids: for points along rays from the
    eye, cells nearer the eye are visited later."""
    rng = np.random.default_rng(7)
    for trial in range(40):
        size = int(rng.integers(1, 6))
        eye = rng.uniform(-0.5, 1.5, size=3)
        order = orc.b2f_macro([0, 0, 0], [size] * 3, [0, 0, 0], [1, 1, 1], eye)
        rank = {tuple(int(v) for v in b): i for i, b in enumerate(order)}
        assert len(rank) == size ** 3
        for _ in range(50):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            ts = np.linspace(0.0, 4.0, 400)
            pts = eye[None, :] + ts[:, None] * d[None, :]
            inside = np.all((pts > 1e-6) & (pts < 1 - 1e-6), axis=1)
            cells = np.floor(pts[inside] * size).astype(int)
            seq = [rank[tuple(c)] for c in cells]
            # moving away from the eye the cells get farther: rank must not increase
            assert all(a >= b for a, b in zip(seq, seq[1:])), (size, eye, d)
