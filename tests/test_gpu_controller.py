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
