"""The controller (rv_controller_*) over a dataset read from files: its
render thread runs the level-of-detail loop, every input change cancels the
loop in progress and starts it again, and each level's composed frame lands
in the frame output segment (VsController.cpp:42-196, VsRenderer.cpp:
1061-1192, VsControllerFrameOutputRW.hpp:28-81).  The frames read from the
segment by a separate attachment equal the oracle's composed frames of the
same inputs, bit for bit (the exact sample sequence), with the field range
the dataset's info file holds."""
import os

import numpy as np
import pytest

from oracle import s3dmm_format as fmt
from revise_amd import api, datasets
from tests.gpu_util import say
from tests.test_gpu_dataset_files import _dataset
from tests.test_gpu_worker import BG, CUT, FOV, Q, THR, _oracle_frame

pytestmark = pytest.mark.gpu

W, H, K = 160, 120, 3


def test_controller_levels_into_shared_memory(orc, tmp_path):
    meta_p, field_p, _, blocks = _dataset(tmp_path, 21)
    base = meta_p[:-len(".s3dmm-meta")]
    c = api.Controller(0, split_count=K, measure_time=True)
    c.set_problem(base, 0, "f")
    name, fmin, fmax, _ = c.fields()[0]
    assert name == "f"
    tf = datasets.seeded_tf(21, 256)
    M = datasets.orbit_camera(35.0)
    c.set_input(W, H, FOV, Q, THR, CUT, BG, tf, view16=M, exact=True)
    say("controller: levels 0..2")
    c.start()
    idle, n = c.wait_idle(120)
    assert idle and n == 3 and c.status() == (0, "")
    sid, fs, w, h = c.frame_output()
    assert (fs, w, h) == (W * H * 4, W, H)
    m = api.SharedMem(shmid=sid)
    fn, _, lv, px = m.read_frame(fs)
    assert (fn, lv) == (2, 2)
    assert np.array_equal(px.reshape(H, W), _oracle_frame(orc, blocks, 2, K, M, W, H, fmin, fmax, tf))

    say("controller: new camera")
    M2 = datasets.orbit_camera(190.0, elevation_deg=-10.0)
    c.set_camera(M2)
    idle, n = c.wait_idle(120)
    assert idle and n == 6
    fn, _, lv, px = m.read_frame(fs)
    assert (fn, lv) == (5, 2)
    assert np.array_equal(px.reshape(H, W), _oracle_frame(orc, blocks, 2, K, M2, W, H, fmin, fmax, tf))

    say("controller: render level 1")
    c.set_render_level(1)
    idle, n = c.wait_idle(120)
    assert idle and n == 7
    fn, _, lv, px = m.read_frame(fs)
    assert (fn, lv) == (6, 1)
    ds = api.Dataset(meta_p)
    wk = api.Worker(0)
    ds.load_worker(wk, field_p, 0, 2)
    wk.set_state(M2, W, H, FOV, Q, THR, CUT, fmin, fmax, BG, tf, exact=True)
    assert np.array_equal(px.reshape(H, W), wk.compose_frame(1, K))
    wk.close()
    ds.close()

    say("controller: changes in quick succession")
    c.set_render_level(-1)
    for a in (10.0, 20.0, 30.0):
        c.set_camera(datasets.orbit_camera(a))
    idle, n2 = c.wait_idle(120)
    fn, _, lv, px = m.read_frame(fs)
    assert idle and n2 >= n + 3 and fn == n2 - 1 and lv == 2
    assert np.array_equal(px.reshape(H, W),
                          _oracle_frame(orc, blocks, 2, K, datasets.orbit_camera(30.0), W, H, fmin, fmax, tf))

    say("controller: new viewport size")
    c.set_input(96, 64, FOV, Q, THR, CUT, BG, tf, exact=True)
    idle, n3 = c.wait_idle(120)
    sid2, fs2, w2, h2 = c.frame_output()
    assert idle and sid2 != sid and (fs2, w2, h2) == (96 * 64 * 4, 96, 64)
    m2 = api.SharedMem(shmid=sid2)
    fn, _, lv, px = m2.read_frame(fs2)
    assert (fn, lv) == (n3 - 1, 2)
    assert np.array_equal(px.reshape(64, 96),
                          _oracle_frame(orc, blocks, 2, K, datasets.orbit_camera(30.0), 96, 64, fmin, fmax, tf))
    m2.close()
    m.close()
    c.close()


def test_controller_reports_a_missing_field_file(tmp_path):
    meta_p, field_p, _, _ = _dataset(tmp_path, 22)
    os.remove(field_p)
    c = api.Controller(0)
    c.set_problem(meta_p[:-len(".s3dmm-meta")], 0, "f")
    c.set_input(64, 48, FOV, Q, THR, CUT, BG, datasets.seeded_tf(22, 64))
    say("controller: missing field file")
    c.start()
    idle, n = c.wait_idle(60)
    rc, msg = c.status()
    assert idle and n == 0 and rc != 0 and msg
    c.close()


def _write_tree_dataset(frame_base, seed, values_seed, level=2):
    """A dataset's meta and field#f files at <frame_base>; returns the level's
    blocks {location: Block} and the subtrees."""
    rng = np.random.default_rng(seed)
    m = 1 << level
    blocks = [(level + 1, 2 * x + int(c[0]), 2 * y + int(c[1]), 2 * z + int(c[2]))
              for z in range(m) for y in range(m) for x in range(m)
              for c in [rng.integers(0, 2, size=3)]]
    for _ in range(100):
        lv = int(rng.integers(level + 1, 6))
        blocks.append((lv, *[int(v) for v in rng.integers(0, 1 << lv, size=3)]))
    meta, subs, size = fmt.meta_bytes(fmt.BlockTree(blocks), 3, level, version=1, map_order_seed=seed)
    field = datasets.Field(values_seed, n_gauss=3, noise_octaves=1)
    vals = {s["values_pos"]: datasets.sample(field, s["level"], s["location"], s["depth"], s["n2i"]) for s in subs}
    with open(frame_base + ".s3dmm-meta", "wb") as f:
        f.write(meta)
    with open(frame_base + ".s3dmm-field#f", "wb") as f:
        f.write(fmt.field_bytes(subs, size, lambda s: vals[s["values_pos"]]))
    return {s["location"]: datasets.Block(level, s["location"], s["depth"], s["n2i"], vals[s["values_pos"]])
            for s in subs if s["level"] == level}


def test_controller_time_frames(orc, tmp_path):
    """A time-dependent dataset (<dir>/<stem>/<stem>_tNNNNNN<ext> frames, the
    info file <dir>/<stem>/<stem><ext>.s3dmm-fields, filename_util.hpp:36-94):
    the controller renders the time frame it is given, with the field range
    of the info file."""
    d = tmp_path / "flow"
    d.mkdir()
    frames = [_write_tree_dataset(str(d / ("flow_t%06d.tec" % k)), 31, 40 + k) for k in range(2)]
    lo = min(datasets.field_range(b.values())[0] for b in frames)
    hi = max(datasets.field_range(b.values())[1] for b in frames)
    with open(str(d / "flow.tec.s3dmm-fields"), "w") as f:
        f.write(fmt.fields_info_text(2, [("f", lo, hi, True)]))
    c = api.Controller(0, split_count=2)
    c.set_problem(str(tmp_path / "flow.tec"), 1, "f")
    assert c.time_frame_count() == 2
    name, fmin, fmax, animated = c.fields()[0]
    assert name == "f" and animated
    tf = datasets.seeded_tf(31, 256)
    M = datasets.orbit_camera(120.0)
    c.set_input(W, H, FOV, Q, THR, CUT, BG, tf, view16=M, exact=True)
    c.set_render_level(2)
    say("controller: time frame 1")
    c.start()
    idle, n = c.wait_idle(120)
    assert idle and n == 1 and c.status() == (0, "")
    sid, fs, _, _ = c.frame_output()
    m = api.SharedMem(shmid=sid)
    fn, _, lv, px = m.read_frame(fs)
    assert (fn, lv) == (0, 2)
    exp1 = _oracle_frame(orc, frames[1], 2, 2, M, W, H, fmin, fmax, tf)
    assert np.array_equal(px.reshape(H, W), exp1)
    assert not np.array_equal(exp1, _oracle_frame(orc, frames[0], 2, 2, M, W, H, fmin, fmax, tf))
    say("controller: time frame 0")
    c.set_problem(str(tmp_path / "flow.tec"), 0, "f")
    idle, n = c.wait_idle(120)
    fn, _, lv, px = m.read_frame(fs)
    assert idle and (fn, lv) == (1, 2)
    assert np.array_equal(px.reshape(H, W), _oracle_frame(orc, frames[0], 2, 2, M, W, H, fmin, fmax, tf))
    m.close()
    c.close()


def test_controller_secondary_field(tmp_path):
    """ValueOnIsosurfaces through the controller: the primary field's
    isosurfaces coloured by the secondary field, both ranges from the info
    file; the frame equals a worker's composed frame over the same files."""
    meta_p, field_p, info_p, blocks, sec = _dataset(tmp_path, 23, second=24)
    fmin, fmax = datasets.field_range(blocks.values())
    lo = min(float(v.min()) for v in sec.values())
    hi = max(float(v.max()) for v in sec.values())
    with open(info_p, "w") as f:
        f.write(fmt.fields_info_text(1, [("f", fmin, fmax, False), ("g", lo, hi, False)]))
    base = meta_p[:-len(".s3dmm-meta")]
    c = api.Controller(0, split_count=3)
    c.set_problem(base, 0, "f", "g")
    (_, f0, f1, _), (_, g0, g1, _) = c.fields()
    tf = datasets.seeded_tf(23, 256)
    M = datasets.orbit_camera(300.0, elevation_deg=-20.0)
    levels = [0.3, 0.55, 0.8]
    c.set_input(W, H, FOV, Q, THR, CUT, BG, tf, view16=M, field_mode=api.FIELD_VALUE_ON_ISOSURFACES,
                iso_opacity=0.6, iso_levels=levels, exact=True)
    c.set_render_level(2)
    say("controller: secondary field")
    c.start()
    idle, n = c.wait_idle(120)
    assert idle and n == 1 and c.status() == (0, "")
    sid, fs, _, _ = c.frame_output()
    m = api.SharedMem(shmid=sid)
    _, _, lv, px = m.read_frame(fs)
    ds = api.Dataset(meta_p)
    wk = api.Worker(0)
    ds.load_worker(wk, field_p, 0, 2)
    ds.load_worker(wk, api.dataset_file_names(base, "g")[1], 0, 2, secondary=True)
    wk.set_state(M, W, H, FOV, Q, THR, CUT, f0, f1, BG, tf, field_mode=api.FIELD_VALUE_ON_ISOSURFACES,
                 iso_opacity=0.6, iso_levels=levels, secondary_range=(g0, g1), exact=True)
    exp = wk.compose_frame(2, 3)
    assert lv == 2 and np.array_equal(px.reshape(H, W), exp)
    assert (exp >> 24).any()
    wk.close()
    ds.close()
    m.close()
    c.close()
