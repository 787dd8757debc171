"""A worker fed from dataset files (rv_dataset_load_worker: .s3dmm-meta ->
subtree children -> n2i -> .s3dmm-field# values -> rv_worker_add_subtree)
renders the same shares as the oracle does from the subtrees the oracle's
writer put into those files, bit for bit."""
import os

import numpy as np
import pytest

from oracle import s3dmm_format as fmt
from revise_amd import api, datasets
from tests.gpu_util import say
from tests.test_gpu_worker import BG, CUT, FOV, Q, THR, _oracle_share

pytestmark = pytest.mark.gpu


def _dataset(tmp_path, seed, level=2, max_level=5, msd=3):
    rng = np.random.default_rng(seed)
    m = 1 << level
    blocks = []
    for z in range(m):
        for y in range(m):
            for x in range(m):
                c = rng.integers(0, 2, size=3)   # every level-2 block gets at least one child
                blocks.append((level + 1, 2 * x + int(c[0]), 2 * y + int(c[1]), 2 * z + int(c[2])))
    for _ in range(150):
        lv = int(rng.integers(level + 1, max_level + 1))
        blocks.append((lv, *[int(v) for v in rng.integers(0, 1 << lv, size=3)]))
    tree = fmt.BlockTree(blocks)
    meta, subs, size = fmt.meta_bytes(tree, msd, level, version=1, map_order_seed=seed)
    field = datasets.Field(seed, n_gauss=3, noise_octaves=1)
    vals = {s["values_pos"]: datasets.sample(field, s["level"], s["location"], s["depth"], s["n2i"])
            for s in subs}
    base = os.path.join(str(tmp_path), "run")
    meta_p, field_p, info_p = api.dataset_file_names(base, "f")
    with open(meta_p, "wb") as f:
        f.write(meta)
    with open(field_p, "wb") as f:
        f.write(fmt.field_bytes(subs, size, lambda s: vals[s["values_pos"]]))
    lv_blocks = {s["location"]: datasets.Block(level, s["location"], s["depth"], s["n2i"],
                                                vals[s["values_pos"]])
                 for s in subs if s["level"] == level}
    assert len(lv_blocks) == m ** 3
    fmin, fmax = datasets.field_range(lv_blocks.values())
    with open(info_p, "w") as f:
        f.write(fmt.fields_info_text(1, [("f", fmin, fmax, False)]))
    return meta_p, field_p, info_p, lv_blocks


@pytest.mark.parametrize("fused", [True, False])
def test_worker_from_files_bit_exact(orc, tmp_path, fused):
    level, W, H = 2, 200, 150
    meta_p, field_p, info_p, blocks = _dataset(tmp_path, 5)
    _, fields = api.read_fields_info(info_p)
    name, fmin, fmax, _ = fields[0]
    # the info file holds the range printed to 6 digits; render with the exact one
    fmin, fmax = datasets.field_range(blocks.values())
    ds = api.Dataset(meta_p)
    w = api.Worker(0)
    assert ds.load_worker(w, field_p, level, level) == len(blocks)
    tf = datasets.seeded_tf(5, 256)
    M = datasets.orbit_camera(50.0)
    w.set_state(M, W, H, FOV, Q, THR, CUT, fmin, fmax, BG, tf)
    w.set_fused(fused)
    for bmin, bmax in [((0, 0, 0), (4, 4, 4)), ((1, 2, 0), (4, 4, 3))]:
        say("worker from files, part", bmin, bmax, "fused", fused)
        rect, bits = w.render_scene_part(level, bmin, bmax)
        erect, ebits = _oracle_share(orc, blocks, level, bmin, bmax, M, W, H, fmin, fmax, tf)
        assert rect == erect
        assert np.array_equal(bits, ebits)
    w.close()
    ds.close()
