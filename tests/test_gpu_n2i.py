"""n2i built on the device (rv_subtree_n2i_device) equals the host list
(rv_subtree_n2i) and the oracle's (itself pinned by hand-worked trees)
exactly: random and full trees of depths 0-7, children arrays in any node
numbering, walks limited below the tree's depth; errors as the host's."""
import numpy as np
import pytest

from revise_amd import RvError, api, datasets
from tests.gpu_util import say, to_dev

pytestmark = pytest.mark.gpu


def _dev_n2i(children, depth, cap=None):
    import torch
    ch = to_dev(np.ascontiguousarray(children, dtype=np.uint32).view(np.int32))
    cap = cap if cap is not None else (2 ** depth + 1) ** 3
    out = torch.zeros((max(cap, 1), 3), dtype=torch.int16, device="cuda")
    n = api.subtree_n2i_device(ch, children.shape[0], depth, out)
    return out[:n].cpu().numpy().view(np.uint16)


def _shuffled(children, rng):
    """The same tree with its non-root nodes renumbered."""
    n = children.shape[0]
    perm = np.concatenate([[0], 1 + rng.permutation(n - 1)]).astype(np.uint32)  # old -> new
    out = np.zeros_like(children)
    mapped = np.where(children != 0, perm[children], 0)
    out[perm] = mapped
    return out


@pytest.mark.parametrize("seed,depth,p", [(1, 1, 0.5), (2, 3, 0.6), (3, 5, 0.45), (4, 6, 0.35), (5, 7, 0.25),
                                          (6, 4, 1.0), (7, 6, 1.0), (8, 0, 1.0), (9, 2, 0.0)])
def test_device_n2i_equals_host(orc, seed, depth, p):
    rng = np.random.default_rng(seed)
    ch = datasets.refined_subtree(depth, lambda l, locs: rng.random(locs.shape[0]) < p)
    d = api.subtree_depth(ch, depth)
    for walk in sorted({d, max(d - 1, 0)}):
        exp = api.subtree_n2i(ch, walk)
        assert np.array_equal(exp, orc.subtree_n2i(ch, walk))
        got = _dev_n2i(ch, walk)
        say("n2i depth", walk, "blocks", ch.shape[0], "nodes", exp.shape[0])
        assert np.array_equal(got, exp)
        assert np.array_equal(_dev_n2i(_shuffled(ch, rng), walk), exp)


def test_device_n2i_hand_worked_depth2(orc):
    """The depth-2 tree whose pre-order and breadth-first corner orders
    differ, with its hand-enumerated list."""
    from tests.test_oracle_pins import N2I_DEPTH2, _depth2_tree
    got = _dev_n2i(_depth2_tree(), 2)
    assert [tuple(int(v) for v in r) for r in got] == N2I_DEPTH2


def test_device_n2i_errors():
    ch = datasets.full_subtree(2)
    bad = ch.copy()
    bad[3, 5] = ch.shape[0] + 7
    with pytest.raises(RvError, match="out of range"):
        _dev_n2i(bad, 2)
    with pytest.raises(RvError, match="capacity"):
        _dev_n2i(ch, 2, cap=10)


def test_device_n2i_full_depth6_timing():
    import time
    import torch
    ch = datasets.full_subtree(6)
    t0 = time.perf_counter()
    exp = api.subtree_n2i(ch, 6)
    t_host = time.perf_counter() - t0
    _dev_n2i(ch, 6)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    got = _dev_n2i(ch, 6)
    t_dev = time.perf_counter() - t0
    say("n2i full depth 6 (%d blocks, %d nodes): host %.1f ms, device %.1f ms (incl. copies)"
        % (ch.shape[0], exp.shape[0], t_host * 1e3, t_dev * 1e3))
    assert np.array_equal(got, exp)
