"""The dataset file reader (rv_dataset_*, rv_fields_info_read): host-only
calls, so these run without a GPU.

Pinned by a hand-assembled metadata file of a 4-block tree (every byte laid
out from the reference's writer: Metadata.hpp:401-462, PrefixWriter.hpp,
BlockTree_io.hpp:64-80, binary_io.hpp) and the subtree roots, depths and n2i
counts worked out by hand for it; then the oracle's writer (checked against
the same bytes) against the library's reader on random trees, both format
versions and shuffled subtree maps; and the coverage invariant of the
subtree roots (every block lies in the subtree of its closest root, within
that subtree's depth)."""
import os
import struct

import numpy as np
import pytest

from oracle import oracle as orc
from oracle import s3dmm_format as fmt
from revise_amd import RvError, api

# Blocks (level, x, y, z): the root, its children 0 and 3, and child 7 of child 3.
TINY_BLOCKS = [(1, 0, 0, 0), (1, 1, 1, 0), (2, 3, 3, 1)]
TINY_BC = (-1.0, -2.0, -3.0, 4.0)

TINY_META = bytes.fromhex(
    "7333646d 6d2d6d65 74616461 7461"          # "s3dmm-metadata"
    "01000000"                                  # format version 1
    "0c00000000000000"                          # header size 12
    "03000000" "01000000" "00000000"            # N = 3, maxSubtreeDepth 1, maxFullLevel 0
    "2e00000000000000"                          # position prefix: 46, where the tree starts
    "000080bf" "000000c0" "000040c0" "00008040"  # bounding cube min (-1, -2, -3), size 4
    "2000000000000000"                          # dataBitCount 32
    "0400000000000000" "90000100"               # 4 node bytes: root 0b10010000, -, 0b00000001, -
    "0200000000000000"                          # 2 subtree levels
    "0100000000000000"                          # level 0: 1 subtree
    "00000000" "00000000" "00000000" "00000000"  # location (0,0,0), tree node 0
    "0000000000000000" "0100000000000000"       # values at 0, depth 1
    "0100000000000000"                          # level 1: 1 subtree
    "01000000" "01000000" "00000000" "02000000"  # location (1,1,0), tree node 2
    "5800000000000000" "0100000000000000"       # values at 88 = (2 + 20) * 4, depth 1
    "9c00000000000000"                          # field file size 156 = 88 + (2 + 15) * 4
    .replace(" ", ""))


def _write(tmp_path, name, data, mode="wb"):
    p = os.path.join(str(tmp_path), name)
    with open(p, mode) as f:
        f.write(data)
    return p


def test_oracle_writer_matches_hand_assembled_file():
    tree = fmt.BlockTree(TINY_BLOCKS)
    assert tree.masks == [0x90, 0x00, 0x01, 0x00]
    data, subs, size = fmt.meta_bytes(tree, 1, 0, TINY_BC[:3], TINY_BC[3])
    assert data == TINY_META
    assert [s["n"] for s in subs] == [20, 15] and size == 156


def test_reader_on_hand_assembled_file(tmp_path):
    ds = api.Dataset(_write(tmp_path, "tiny.s3dmm-meta", TINY_META))
    inf = ds.info
    assert (inf["format_version"], inf["max_subtree_depth"], inf["max_full_level"]) == (1, 1, 0)
    assert inf["bc_min"] == (-1.0, -2.0, -3.0) and inf["bc_size"] == 4.0
    assert (inf["node_count"], inf["subtree_count"], inf["level_count"], inf["field_file_size"]) == \
        (4, 2, 2, 156)
    assert ds.subtree(0) == dict(level=0, location=(0, 0, 0), node=0, depth=1, values_pos=0)
    assert ds.subtree(1) == dict(level=1, location=(1, 1, 0), node=2, depth=1, values_pos=88)
    # depth 1 below the root: children 0 and 3 only (child 3's child is below)
    ch = ds.children(0)
    assert ch.shape == (3, 8) and list(ch[0]) == [1, 0, 0, 2, 0, 0, 0, 0] and not ch[1:].any()
    assert list(ds.children(1)[0]) == [0, 0, 0, 0, 0, 0, 0, 1]
    n0, n1 = ds.n2i(0), ds.n2i(1)
    assert len(n0) == 20 and len(n1) == 15
    # the root's corners first, x fastest; then child 0's 7 new corners
    assert [tuple(v) for v in n0[:9]] == [(0, 0, 0), (2, 0, 0), (0, 2, 0), (2, 2, 0), (0, 0, 2), (2, 0, 2),
                                          (0, 2, 2), (2, 2, 2), (1, 0, 0)]
    assert [tuple(v) for v in n1[8:]] == [(1, 1, 1), (2, 1, 1), (1, 2, 1), (2, 2, 1), (1, 1, 2), (2, 1, 2),
                                          (1, 2, 2)]
    assert ds.find_subtree(1, (1, 1, 0)) == 1
    with pytest.raises(RvError):
        ds.find_subtree(1, (0, 0, 0))
    assert ds.closest_subtree(2, (3, 3, 1)) == 1   # below the level-1 root
    assert ds.closest_subtree(2, (0, 0, 0)) == 0   # no level-1 root above it
    assert ds.closest_subtree(7, (127, 127, 63)) == 1   # climbs from level 7
    assert ds.closest_subtree(7, (127, 127, 64)) == 0
    ds.close()


def _random_blocks(rng, max_level, count):
    out = []
    for _ in range(count):
        lv = int(rng.integers(1, max_level + 1))
        out.append((lv, *[int(v) for v in rng.integers(0, 1 << lv, size=3)]))
    return out


CASES = [  # (seed, max_level, blocks, maxSubtreeDepth, maxFullLevel, version)
    (1, 4, 30, 2, 0, 1),
    (2, 5, 60, 3, 1, 1),
    (3, 5, 40, 2, 2, 0),
    (4, 6, 80, 4, 1, 1),
    (5, 3, 200, 1, 1, 1),
    (6, 6, 25, 6, 0, 0),
]


@pytest.mark.parametrize("seed,max_level,nblocks,msd,mfl,version", CASES)
def test_reader_against_oracle_writer(tmp_path, seed, max_level, nblocks, msd, mfl, version):
    rng = np.random.default_rng(seed)
    tree = fmt.BlockTree(_random_blocks(rng, max_level, nblocks))
    data, subs, size = fmt.meta_bytes(tree, msd, mfl, (0.5, -1.0, 2.0), 3.0, version=version,
                                      map_order_seed=seed)
    info, entries = fmt.read_meta(data)       # the oracle's own reader agrees with its writer
    assert len(entries) == len(subs) and info["field_file_size"] == size
    ds = api.Dataset(_write(tmp_path, "d.s3dmm-meta", data))
    assert ds.info["node_count"] == len(tree.nodes) and len(ds) == len(subs)
    assert ds.info["format_version"] == version and ds.info["field_file_size"] == size
    vals = {}
    for s in subs:
        v = rng.standard_normal(s["n"]).astype(np.float32)
        v[rng.random(s["n"]) < 0.1] = fmt.NO_FIELD
        vals[s["values_pos"]] = v
    fpath = _write(tmp_path, "d.s3dmm-field#f", fmt.field_bytes(subs, size, lambda s: vals[s["values_pos"]]))
    for i, s in enumerate(subs):
        got = ds.subtree(i)
        want_depth = s["depth"] if version >= 1 else msd
        assert got == dict(level=s["level"], location=s["location"], node=s["node"], depth=want_depth,
                           values_pos=s["values_pos"])
        if version == 0:
            continue   # version 0 subtrees take maxSubtreeDepth, whatever their depth
        np.testing.assert_array_equal(ds.children(i), tree.subtree_children(s["node"], s["depth"]))
        n2i = ds.n2i(i)
        np.testing.assert_array_equal(n2i, s["n2i"])
        rng_, v = ds.read_field(fpath, i, len(n2i))
        np.testing.assert_array_equal(v, vals[s["values_pos"]])
        np.testing.assert_array_equal(rng_, fmt.field_range(v))
        assert ds.find_subtree(s["level"], s["location"]) == i


@pytest.mark.parametrize("seed,max_level,nblocks,msd,mfl,version", CASES)
def test_subtrees_cover_every_block(tmp_path, seed, max_level, nblocks, msd, mfl, version):
    """Each block lies in the subtree of its closest root, at most that
    subtree's depth below the root; every block at a level <= maxFullLevel is
    a root."""
    rng = np.random.default_rng(seed)
    tree = fmt.BlockTree(_random_blocks(rng, max_level, nblocks))
    data, subs, _ = fmt.meta_bytes(tree, msd, mfl, version=1)
    ds = api.Dataset(_write(tmp_path, "c.s3dmm-meta", data))
    for (lv, x, y, z) in tree.nodes:
        i = ds.closest_subtree(lv, (x, y, z))
        s = ds.subtree(i)
        assert s["level"] <= lv <= s["level"] + s["depth"]
        assert all(c >> (lv - s["level"]) == r for c, r in zip((x, y, z), s["location"]))
        if lv <= mfl:
            assert s["level"] == lv
        # the block is in the subtree's n2i walk: its corners are all present
        if version >= 1:
            d, sh = s["depth"], s["depth"] - (lv - s["level"])
            mask = (1 << (lv - s["level"])) - 1
            base = [((c & mask) << sh) for c in (x, y, z)]
            have = {tuple(v) for v in ds.n2i(i)}
            for k in range(8):
                assert (base[0] + ((k & 1) << sh), base[1] + (((k >> 1) & 1) << sh),
                        base[2] + (((k >> 2) & 1) << sh)) in have
                assert max(base) + (1 << sh) <= (1 << d)


def test_walk_subtrees_brute_force():
    """The oracle's walkSubtrees against a direct reading of its definition:
    a block is a root iff its level <= maxFullLevel, or its parent is a root
    whose blocks reach deeper than maxSubtreeDepth and its own subtree is
    deeper than maxSubtreeDepth - 1."""
    rng = np.random.default_rng(11)
    for trial in range(20):
        tree = fmt.BlockTree(_random_blocks(rng, 5, int(rng.integers(5, 60))))
        msd, mfl = int(rng.integers(1, 4)), int(rng.integers(0, 3))
        depth_of = {}

        def depth(n):
            if n not in depth_of:
                depth_of[n] = 1 + max((depth(c) for c in tree.child[n] if c), default=-1)
            return depth_of[n]
        roots = {0}
        for n in range(len(tree.nodes)):     # breadth-first: parents come first
            lv = tree.nodes[n][0]
            for c in tree.child[n]:
                if c and n in roots and (lv < mfl or (depth(n) > msd and depth(c) > msd - 1)):
                    roots.add(c)
        got = tree.walk_subtrees(mfl, msd)
        assert sorted(got) == sorted(roots) and len(got) == len(set(got))
        for n in got:
            assert tree.depth_up_to(n, msd) == min(depth(n), msd)


def test_field_range_rule():
    assert list(fmt.field_range([3.0, 1e30, -2.0, 5.0])) == [-2.0, 5.0]
    assert list(fmt.field_range([1e30, 1e30])) == [0.0, -1.0]


def _status(fn):
    try:
        fn()
    except RvError as e:
        return str(e)
    return None


def test_malformed_metadata_is_rejected(tmp_path):
    # every truncation fails cleanly
    for cut in range(0, len(TINY_META), 3):
        p = _write(tmp_path, "t%d.s3dmm-meta" % cut, TINY_META[:cut])
        assert _status(lambda: api.Dataset(p)) is not None, cut
    bad = bytearray(TINY_META)
    bad[14] = 2                                  # format version 2
    assert "version" in _status(lambda: api.Dataset(_write(tmp_path, "v.s3dmm-meta", bytes(bad))))
    bad = bytearray(TINY_META)
    bad[26] = 2                                  # dimension 2
    assert "dimension" in _status(lambda: api.Dataset(_write(tmp_path, "n.s3dmm-meta", bytes(bad))))
    bad = bytearray(TINY_META)
    bad[72] = 0x98                               # an extra root child: counts no longer add up
    assert _status(lambda: api.Dataset(_write(tmp_path, "c.s3dmm-meta", bytes(bad)))) is not None
    bad = bytearray(TINY_META)
    off = TINY_META.index(bytes.fromhex("01000000010000000000000002000000"))
    bad[off + 12] = 1                            # level-1 root says tree node 1, which is (1,0,0,0)
    assert "does not match" in _status(lambda: api.Dataset(_write(tmp_path, "r.s3dmm-meta", bytes(bad))))
    assert _status(lambda: api.Dataset(os.path.join(str(tmp_path), "absent.s3dmm-meta"))) is not None


def test_short_field_file_is_rejected(tmp_path):
    ds = api.Dataset(_write(tmp_path, "tiny.s3dmm-meta", TINY_META))
    p = _write(tmp_path, "tiny.s3dmm-field#f", b"\0" * 155)
    assert ds.read_field(p, 0, 20)[1].shape == (20,)
    with pytest.raises(RvError):
        ds.read_field(p, 1, 15)


def test_version0_file(tmp_path):
    tree = fmt.BlockTree(TINY_BLOCKS)
    data, _, _ = fmt.meta_bytes(tree, 1, 0, TINY_BC[:3], TINY_BC[3], version=0)
    # no header: the position prefix (8) opens the file; maxSubtreeDepth
    # follows the tree; entries carry no depth
    assert data[:8] == struct.pack("<Q", 8) and data[8:24] == TINY_META[46:62]
    ds = api.Dataset(_write(tmp_path, "v0.s3dmm-meta", data))
    assert (ds.info["format_version"], ds.info["max_full_level"], ds.info["max_subtree_depth"]) == (0, 0, 1)
    assert [ds.subtree(i)["values_pos"] for i in range(2)] == [0, 88]


def test_fields_info(tmp_path):
    text = fmt.fields_info_text(3, [("rho", 0.25, 1.5, True), ("p", -1e-3, 2e5, False)])
    assert text.startswith("time_steps\n3\n\nfield\tmin\tmax\tanimated\nrho\t0.25\t1.5\t1\n")
    steps, fields = api.read_fields_info(_write(tmp_path, "a.s3dmm-fields", text, "w"))
    assert steps == 3
    assert [(n, a) for n, _, _, a in fields] == [("rho", True), ("p", False)]
    np.testing.assert_allclose([f[1:3] for f in fields], [[0.25, 1.5], [-1e-3, 2e5]], rtol=1e-6)
    # without a trailing newline
    steps, fields = api.read_fields_info(_write(tmp_path, "b.s3dmm-fields", text.rstrip("\n"), "w"))
    assert len(fields) == 2
    # a field without a range ("-") and a wrong column header are errors
    for bad in (text.replace("1.5", "-"), text.replace("field\tmin", "field min"),
                text.replace("time_steps", "frames"), text + "\n"):
        p = _write(tmp_path, "bad.s3dmm-fields", bad, "w")
        with pytest.raises(RvError):
            api.read_fields_info(p)


def test_dataset_file_names():
    assert api.dataset_file_names("/d/run.tec", "rho") == \
        ("/d/run.tec.s3dmm-meta", "/d/run.tec.s3dmm-field#rho", "/d/run.tec.s3dmm-fields")
    assert api.dataset_file_names("/d/run.tec", "rho", frame=12, has_time_steps=True) == \
        ("/d/run/run_t000012.tec.s3dmm-meta", "/d/run/run_t000012.tec.s3dmm-field#rho",
         "/d/run/run.tec.s3dmm-fields")


def test_oracle_n2i_is_used_by_writer():
    tree = fmt.BlockTree(TINY_BLOCKS)
    subs, _ = fmt.subtrees_of(tree, 1, 0)
    np.testing.assert_array_equal(subs[1]["n2i"], orc.subtree_n2i(tree.subtree_children(2, 1), 1))
