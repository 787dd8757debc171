"""Frame output through System V shared memory (rv_shm_*, rv_frame_output_*_shm)
and the controller's host side, on the CPU.

The segment layout and lock protocol are the reference's
(ipc/SharedMemory.h:30-280): user size (8 bytes), control word (int32),
user bytes at offset 12; readers count the word up while it is >= 0, a
writer swaps 0 for -1000.  The frame output block is the 24-byte
VsControllerFrameOutputHeader then the pixels (VsControllerFrameOutputRW.hpp:
41-75).  The checks read and write the segments through libc's shmat, not
through the library, so a layout mistake on either side shows."""
import ctypes as C
import os
import subprocess
import sys
import threading
import time

import numpy as np
import pytest

from oracle import s3dmm_format as fmt
from revise_amd import RvError, api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_libc = C.CDLL("libc.so.6", use_errno=True)
_libc.shmat.restype = C.c_void_p
_libc.shmat.argtypes = [C.c_int, C.c_void_p, C.c_int]
_libc.shmdt.argtypes = [C.c_void_p]
_libc.shmget.argtypes = [C.c_int, C.c_size_t, C.c_int]
_libc.shmctl.argtypes = [C.c_int, C.c_int, C.c_void_p]
IPC_PRIVATE, IPC_CREAT, IPC_RMID = 0, 0o1000, 0


class Raw:
    """A segment mapped with libc, viewed as bytes."""

    def __init__(self, shmid):
        self.p = _libc.shmat(shmid, None, 0)
        assert self.p not in (None, C.c_void_p(-1).value), "shmat failed"

    def bytes(self, start, n):
        return C.string_at(self.p + start, n)

    def put(self, start, data):
        C.memmove(self.p + start, data, len(data))

    def control(self):
        return C.c_int32.from_address(self.p + 8)

    def close(self):
        _libc.shmdt(self.p)


def test_segment_layout_and_frame_block():
    W, H = 5, 3
    fs = W * H * 4
    m = api.SharedMem(24 + fs)
    raw = Raw(m.shmid)
    try:
        assert np.frombuffer(raw.bytes(0, 8), np.uint64)[0] == 24 + fs
        assert raw.control().value == 0
        assert raw.bytes(12, 24 + fs) == bytes(24 + fs)  # a new segment is zero
        px = np.arange(W * H, dtype=np.uint32) * 0x01010101
        m.write_frame(fs, 41, 17, 3, px)
        hdr = np.frombuffer(raw.bytes(12, 24), np.uint64)
        assert hdr[0] == 41 and hdr[1] == 17 and (int(hdr[2]) & 0xFFFFFFFF) == 3
        assert np.array_equal(np.frombuffer(raw.bytes(36, fs), np.uint32), px)
        assert raw.control().value == 0  # the writer released its hold
        # a header-only write keeps the pixels
        m.write_frame(fs, 42, 0, 4)
        n, d, lv, got = m.read_frame(fs)
        assert (n, d, lv) == (42, 0, 4) and np.array_equal(got, px)
        n, _, _, none = m.read_frame(fs, pixels=False)
        assert n == 42 and none is None
    finally:
        raw.close()
        m.close()


def test_reads_a_segment_laid_out_by_hand():
    """A segment made with libc and filled in the reference's layout reads
    back through rv_shm_attach + rv_frame_output_read_shm."""
    fs = 8
    sid = _libc.shmget(IPC_PRIVATE, 12 + 24 + fs, IPC_CREAT | 0o600)
    assert sid >= 0
    raw = Raw(sid)
    try:
        raw.put(0, np.uint64(24 + fs).tobytes())
        raw.put(8, np.int32(0).tobytes())
        raw.put(12, np.array([~np.uint64(0), 5], np.uint64).tobytes() + np.array([0xFFFFFFFF, 0], np.uint32).tobytes())
        raw.put(36, np.array([7, 9], np.uint32).tobytes())
        m = api.SharedMem(shmid=sid)
        assert m.size == 24 + fs
        n, d, lv, px = m.read_frame(fs)
        assert n == 2 ** 64 - 1 and d == 5 and lv == 0xFFFFFFFF and list(px) == [7, 9]
        m.close()
    finally:
        raw.close()
        _libc.shmctl(sid, IPC_RMID, None)


def _blocked(fn):
    """Runs fn on a thread; returns (thread, done-event)."""
    done = threading.Event()

    def run():
        fn()
        done.set()

    t = threading.Thread(target=run, daemon=True)
    t.start()
    return t, done


def test_lock_protocol():
    m = api.SharedMem(64)
    raw = Raw(m.shmid)
    try:
        ctl = raw.control()
        # a writer holds the word: readers wait until it adds 1000 back.  The
        # spinning reader adds and takes back 1, and this store is not an
        # atomic add, so the word may end one below 0 here.
        ctl.value = -1000
        t, done = _blocked(lambda: m.read(0, 8))
        time.sleep(0.2)
        assert not done.is_set()
        ctl.value = 0
        t.join(5)
        assert done.is_set() and ctl.value in (-1, 0)
        ctl.value = 0
        # a reader holds the word: a writer waits for it to leave
        ctl.value = 1
        t, done = _blocked(lambda: m.write(0, b"abcd"))
        time.sleep(0.2)
        assert not done.is_set()
        ctl.value -= 1
        t.join(5)
        assert done.is_set() and ctl.value == 0
        assert raw.bytes(12, 4) == b"abcd"
    finally:
        raw.close()
        m.close()


def test_another_process_reads_the_frame():
    fs = 16
    m = api.SharedMem(24 + fs)
    m.write_frame(fs, 9, 1, 2, np.array([10, 20, 30, 40], np.uint32))
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from revise_amd import api\n"
            "m = api.SharedMem(shmid=%d)\n"
            "n, d, lv, px = m.read_frame(%d)\n"
            "print(n, d, lv, *px.tolist())\n" % (ROOT, m.shmid, fs))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split() == ["9", "1", "2", "10", "20", "30", "40"]
    m.close()


def test_creator_removes_the_segment_and_bounds():
    m = api.SharedMem(32)
    sid = m.shmid
    with pytest.raises(RvError):
        m.read(30, 4)
    with pytest.raises(RvError):
        m.write(33, b"x")
    with pytest.raises(RvError):
        m.read_frame(16)  # 24 + 16 > 32
    m.close()
    with pytest.raises(RvError):
        api.SharedMem(shmid=sid)


def _write_dataset(base, fields, steps=1):
    """The field info file of a dataset (the meta and field files are not
    read before rendering)."""
    with open(base + ".s3dmm-fields", "w") as f:
        f.write(fmt.fields_info_text(steps, fields))


def test_controller_problem_paths(tmp_path):
    c = api.Controller(0)
    try:
        with pytest.raises(RvError):
            c.set_problem(str(tmp_path / "absent"))
        # plain dataset: <base>.s3dmm-meta is a file
        base = str(tmp_path / "run")
        open(base + ".s3dmm-meta", "wb").close()
        _write_dataset(base, [("p", -1.5, 2.25, False), ("q", 0.0, 1.0, True)])
        c.set_problem(base, 0, "q")
        assert c.fields() == [("p", -1.5, 2.25, False), ("q", 0.0, 1.0, True)]
        assert c.time_frame_count() == 1
        # time steps: <dir>/<stem> is a directory of frame files
        d = tmp_path / "ts"
        (d / "flow").mkdir(parents=True)
        open(str(d / "flow" / "flow_t000000.tec.s3dmm-meta"), "wb").close()
        with open(str(d / "flow" / "flow.tec.s3dmm-fields"), "w") as f:
            f.write(fmt.fields_info_text(3, [("u", 0.0, 4.0, True)]))
        c.set_problem(str(d / "flow.tec"), 2, "nope")
        assert c.fields() == [("u", 0.0, 4.0, True)] and c.time_frame_count() == 3
        with pytest.raises(RvError):
            c.set_problem(str(d / "flow.tec"), 3)
        # never started: nothing rendered, no failure
        assert c.wait_idle(1.0) == (True, 0)
        assert c.status()[0] == 0
    finally:
        c.close()


def test_controller_camera_and_frame_output():
    c = api.Controller(0)
    try:
        # CameraController::resetCameraTransform: eye along (2, 1, 3) at the
        # default distance, looking at the origin
        from revise_amd import datasets
        M = c.camera()
        assert np.allclose(M, datasets.default_camera(30.0), atol=1e-5)
        assert np.allclose(api.eye_from_view(M) / np.linalg.norm(api.eye_from_view(M)),
                           np.array([2, 1, 3]) / np.sqrt(14), atol=1e-6)
        M2 = datasets.orbit_camera(80.0)
        c.set_camera(M2)
        assert np.array_equal(c.camera(), M2)
        c.reset_camera()
        assert np.allclose(c.camera(), M, atol=1e-6)
        tf = datasets.seeded_tf(1, 16)
        c.set_input(64, 48, 30.0, 2.0, 0.5, 0.99, (0, 0, 0), tf)
        sid, fsz, w, h = c.frame_output()
        assert (fsz, w, h) == (64 * 48 * 4, 64, 48)
        m = api.SharedMem(shmid=sid)
        n, _, lv, _ = m.read_frame(fsz)  # maybeInitFrameOutput: "no frame"
        assert n == 2 ** 64 - 1 and lv == 0xFFFFFFFF
        assert c.frame_output()[0] == sid  # same size: same segment
        c.set_input(32, 32, 30.0, 2.0, 0.5, 0.99, (0, 0, 0), tf)
        sid2, fsz2, _, _ = c.frame_output()
        assert sid2 != sid and fsz2 == 32 * 32 * 4
        m.close()
    finally:
        c.close()
