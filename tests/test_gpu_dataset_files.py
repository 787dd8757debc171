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


def _dataset(tmp_path, seed, level=2, max_level=5, msd=3, second=None):
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
    if second is not None:
        f2 = datasets.Field(second, n_gauss=4, noise_octaves=2)
        vals2 = {s["values_pos"]: datasets.sample(f2, s["level"], s["location"], s["depth"], s["n2i"])
                 for s in subs}
        with open(api.dataset_file_names(base, "g")[1], "wb") as f:
            f.write(fmt.field_bytes(subs, size, lambda s: vals2[s["values_pos"]]))
        sec = {s["location"]: vals2[s["values_pos"]] for s in subs if s["level"] == level}
    lv_blocks = {s["location"]: datasets.Block(level, s["location"], s["depth"], s["n2i"],
                                                vals[s["values_pos"]])
                 for s in subs if s["level"] == level}
    assert len(lv_blocks) == m ** 3
    fmin, fmax = datasets.field_range(lv_blocks.values())
    with open(info_p, "w") as f:
        f.write(fmt.fields_info_text(1, [("f", fmin, fmax, False)]))
    if second is not None:
        return meta_p, field_p, info_p, lv_blocks, sec
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
    w.set_state(M, W, H, FOV, Q, THR, CUT, fmin, fmax, BG, tf, exact=True)
    w.set_fused(fused)
    for bmin, bmax in [((0, 0, 0), (4, 4, 4)), ((1, 2, 0), (4, 4, 3))]:
        say("worker from files, part", bmin, bmax, "fused", fused)
        rect, bits = w.render_scene_part(level, bmin, bmax)
        erect, ebits = _oracle_share(orc, blocks, level, bmin, bmax, M, W, H, fmin, fmax, tf)
        assert rect == erect
        assert np.array_equal(bits, ebits)
    w.close()
    ds.close()


def test_worker_secondary_field_from_files(orc, tmp_path):
    """ValueOnIsosurfaces with the secondary field loaded from its own field
    file (rv_dataset_load_worker_secondary)."""
    level, W, H = 2, 180, 140
    meta_p, field_p, _, blocks, sec = _dataset(tmp_path, 8, second=9)
    fmin, fmax = datasets.field_range(blocks.values())
    lo = min(float(v.min()) for v in sec.values())
    hi = max(float(v.max()) for v in sec.values())
    ds = api.Dataset(meta_p)
    w = api.Worker(0)
    assert ds.load_worker(w, field_p, level, level) == len(blocks)
    assert ds.load_worker(w, api.dataset_file_names(meta_p[:-len(".s3dmm-meta")], "g")[1], level, level,
                          secondary=True) == len(blocks)
    tf = datasets.seeded_tf(8, 256)
    M = datasets.orbit_camera(200.0, elevation_deg=-15.0)
    levels = [0.35, 0.6]
    w.set_state(M, W, H, FOV, Q, THR, CUT, fmin, fmax, BG, tf, field_mode=api.FIELD_VALUE_ON_ISOSURFACES,
                iso_opacity=0.7, iso_levels=levels, secondary_range=(lo, hi), exact=True)
    say("worker from files, secondary field")
    rect, bits = w.render_scene_part(level, (0, 0, 0), (4, 4, 4))
    vp = np.zeros((H, W), np.uint32)
    eye = api.eye_from_view(M)
    cmin, cmax = api.index_box_coords(level, (0, 0, 0), (4, 4, 4))
    for idx in orc.b2f_macro((0, 0, 0), (4, 4, 4), cmin, cmax, eye):
        key = tuple(int(v) for v in idx)
        b = blocks[key]
        live, po = orc.block_params(M, W, H, FOV, level, b.index, b.depth, fmin, fmax, THR, Q, CUT)
        if live:
            orc.render_block_mode(po, api.FIELD_VALUE_ON_ISOSURFACES, 0.7, orc.interpolate(b.sparse, b.n2i, b.depth),
                                  tf, vp.reshape(-1), levels=levels,
                                  brick2=orc.interpolate(sec[key], b.n2i, b.depth), srange=(lo, hi))
    r = orc.bound_rect(vp)
    assert r is not None and rect == r
    x, y, ww, hh = r
    assert np.array_equal(bits, vp[y:y + hh, x:x + ww])
    w.close()
    ds.close()


def test_progressive_levels_from_files(orc, tmp_path):
    """The level-of-detail loop over a dataset read from files: levels 0..2
    (maxFullLevel 2, so every block of each level is a subtree root), each
    level's frame equals the composed frame of that level, the frame output
    block holds the last one with its number and level; a cancelled loop
    writes nothing."""
    meta_p, field_p, _, blocks = _dataset(tmp_path, 13)
    fmin, fmax = datasets.field_range(blocks.values())
    ds = api.Dataset(meta_p)
    w = api.Worker(0)
    assert ds.load_worker(w, field_p, 0, 2) == 1 + 8 + 64
    W, H = 128, 96
    tf = datasets.seeded_tf(13, 256)
    w.set_state(datasets.orbit_camera(75.0), W, H, FOV, Q, THR, CUT, fmin, fmax, BG, tf, exact=True)
    frames = [w.compose_frame(level, 3) for level in range(3)]
    for level in range(3):
        say("progressive level", level)
        n, fn, out = w.render_progressive(level, level, 3, frame_number=10 + level)
        hdr, pix = api.read_frame_output(out, W, H)
        assert (n, fn) == (1, 11 + level)
        assert hdr["frame_number"] == 10 + level and hdr["level"] == level and hdr["duration_ms"] == 0
        assert np.array_equal(pix, frames[level])
    n, fn, out = w.render_progressive(0, 2, 3, frame_number=100, measure_time=True)
    hdr, pix = api.read_frame_output(out, W, H)
    assert (n, fn) == (3, 103) and hdr["frame_number"] == 102 and hdr["level"] == 2
    assert np.array_equal(pix, frames[2])
    n, fn, out = w.render_progressive(0, 2, 3, frame_number=7, cancel=True)
    assert (n, fn) == (0, 7) and not out.any()
    w.close()
    ds.close()
