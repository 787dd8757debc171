"""The reference's dataset files, written and read by a CPU restatement.

TEST INFRASTRUCTURE: only tests/ import this module, to write datasets the
library's reader must read back, and as an independent reader to check the
library's one against.

Restated from:
* the block tree: nodes in breadth-first order, one byte per node, bit
  (7 - c) for local child c (x = bit 0, y = bit 1, z = bit 2), children
  numbered consecutively in node order (BlockTreeCompressor.hpp:34-63,
  CompressedBlockTree.hpp:392-437);
* subtree roots: CompressedBlockTree::walkSubtrees with
  S3DMM_BLOCKTREE_COUNT_EACH_SUBTREE_DEPTH defined (CompressedBlockTree.hpp:
  158-244, BlockTree_defs.hpp:22) and depthUpTo (:290-300);
* the metadata file: Metadata::generate, writeHeader and the version-0 and
  version-1 reader (Metadata.hpp:182-214, 401-499), PrefixWriter
  (PrefixWriter.hpp:28-60), the binary rules (binary_io.hpp:109-250;
  size_t counts, unordered_map items as key/value pairs, MultiIndex as its
  elements) and SubtreeData's fields (Metadata.hpp:58-115);
* field files: generateSeparatedField (BlockTreeFieldProvider.hpp:216-262)
  with computeFieldRange (computeFieldRange.hpp:27-46) and the no-field
  value 1e30 (BlockTreeFieldProvider.hpp:137-139);
* the field info file as s3dmm_prep writes it (processMesh.cpp:356-369) and
  BlockTreeFieldService reads it (BlockTreeFieldService.hpp:177-228).
"""
import struct

import numpy as np

from oracle import oracle as orc

META_ID = b"s3dmm-metadata"
NO_FIELD = np.float32(1e30)


class BlockTree:
    """A block tree from a set of blocks (level, x, y, z); every block's
    parents are added.  nodes[i] = (level, x, y, z) in breadth-first order,
    masks[i] its children byte, child[i] its 8 child node indices (0 = none)."""

    def __init__(self, blocks):
        have = set()
        for (lv, x, y, z) in blocks:
            while True:
                if (lv, x, y, z) in have:
                    break
                have.add((lv, x, y, z))
                if lv == 0:
                    break
                lv, x, y, z = lv - 1, x >> 1, y >> 1, z >> 1
        have.add((0, 0, 0, 0))
        self.nodes = [(0, 0, 0, 0)]
        self.masks = []
        self.child = []
        i = 0
        while i < len(self.nodes):
            lv, x, y, z = self.nodes[i]
            m = 0
            row = [0] * 8
            for c in range(8):
                ch = (lv + 1, 2 * x + (c & 1), 2 * y + ((c >> 1) & 1), 2 * z + ((c >> 2) & 1))
                if ch in have:
                    m |= 0x80 >> c
                    row[c] = len(self.nodes)
                    self.nodes.append(ch)
            self.masks.append(m)
            self.child.append(row)
            i += 1

    def deeper_than(self, node, max_depth):
        """walk(root, maxDepth, cb)'s result: a block max_depth levels below
        node has children (CompressedBlockTree.hpp:158-181)."""
        if max_depth == 0:
            return self.masks[node] != 0
        return any(self.deeper_than(c, max_depth - 1) for c in self.child[node] if c)

    def depth_up_to(self, node, max_depth):
        """depthUpTo (CompressedBlockTree.hpp:290-300) = min(depth, max_depth)."""
        d, frontier = 0, [node]
        while d < max_depth:
            nxt = [c for n in frontier for c in self.child[n] if c]
            if not nxt:
                break
            frontier, d = nxt, d + 1
        return d

    def walk_subtrees(self, min_depth, max_depth, node=0, out=None):
        """Subtree roots in pre-order (CompressedBlockTree.hpp:189-244)."""
        if out is None:
            out = []
        out.append(node)
        if min_depth > 0 or self.deeper_than(node, max_depth):
            for c in self.child[node]:
                if c and (min_depth > 0 or self.deeper_than(c, max_depth - 1)):
                    self.walk_subtrees(max(min_depth - 1, 0), max_depth, c, out)
        return out

    def subtree_children(self, node, depth):
        """Children table (root = row 0) of the blocks down to depth below node."""
        rows, level = [[0] * 8], [0]
        tree = [node]
        r = 0
        while r < len(tree):
            if level[r] < depth:
                for c in range(8):
                    t = self.child[tree[r]][c]
                    if t:
                        rows[r][c] = len(tree)
                        tree.append(t)
                        level.append(level[r] + 1)
                        rows.append([0] * 8)
            r += 1
        return np.array(rows, dtype=np.uint32)


def subtrees_of(tree, max_subtree_depth, max_full_level):
    """Metadata::generate's subtree list: dicts with level, location, node,
    depth, values_pos, n (n2i count) and n2i (Metadata.hpp:401-437)."""
    out, vpos = [], 0
    for node in tree.walk_subtrees(max_full_level, max_subtree_depth):
        depth = tree.depth_up_to(node, max_subtree_depth)
        n2i = orc.subtree_n2i(tree.subtree_children(node, depth), depth)
        lv, x, y, z = tree.nodes[node]
        out.append(dict(level=lv, location=(x, y, z), node=node, depth=depth, values_pos=vpos,
                        n=len(n2i), n2i=n2i))
        vpos += (2 + len(n2i)) * 4
    return out, vpos


def meta_bytes(tree, max_subtree_depth, max_full_level, bc_min=(0.0, 0.0, 0.0), bc_size=1.0,
               version=1, map_order_seed=None):
    """The .s3dmm-meta bytes.  map_order_seed shuffles the entries within
    each level, as an unordered_map's iteration order may."""
    subs, field_size = subtrees_of(tree, max_subtree_depth, max_full_level)
    b = bytearray()
    if version >= 1:
        b += META_ID + struct.pack("<IQ", version, 12) + struct.pack("<III", 3, max_subtree_depth,
                                                                     max_full_level)
    b += struct.pack("<Q", len(b) + 8)          # position prefix: the tree follows
    b += struct.pack("<4f", *bc_min, bc_size)
    b += struct.pack("<QQ", 8 * len(tree.masks), len(tree.masks)) + bytes(tree.masks)
    if version == 0:
        b += struct.pack("<I", max_subtree_depth)
    nlevels = max(s["level"] for s in subs) + 1
    b += struct.pack("<Q", nlevels)
    rng = np.random.default_rng(map_order_seed) if map_order_seed is not None else None
    for lv in range(nlevels):
        items = [s for s in subs if s["level"] == lv]
        if rng is not None:
            items = [items[k] for k in rng.permutation(len(items))]
        b += struct.pack("<Q", len(items))
        for s in items:
            b += struct.pack("<3II", *s["location"], s["node"])
            b += struct.pack("<Q", s["values_pos"])
            if version >= 1:
                b += struct.pack("<Q", s["depth"])
    b += struct.pack("<Q", field_size)
    return bytes(b), subs, field_size


def field_range(values):
    """computeFieldRange: (min, max) over values != 1e30, (0, -1) if none."""
    v = np.asarray(values, dtype=np.float32)
    v = v[v != NO_FIELD]
    if v.size == 0:
        return np.array([0.0, -1.0], dtype=np.float32)
    return np.array([v.min(), v.max()], dtype=np.float32)


def field_bytes(subs, field_size, values_of):
    """A .s3dmm-field# file: per subtree (min, max, values) at its position;
    values_of(subtree dict) gives its n values (float32)."""
    buf = bytearray(field_size)
    for s in subs:
        v = np.asarray(values_of(s), dtype=np.float32)
        assert v.shape == (s["n"],)
        rec = field_range(v).tobytes() + v.tobytes()
        buf[s["values_pos"]:s["values_pos"] + len(rec)] = rec
    return bytes(buf)


def fields_info_text(frame_count, fields):
    """fields: (name, min, max, animated) tuples; numbers as an ostream
    prints a float by default (6 significant digits)."""
    lines = ["time_steps", str(frame_count), "", "field\tmin\tmax\tanimated"]
    for (name, mn, mx, animated) in fields:
        lines.append("%s\t%.6g\t%.6g\t%d" % (name, mn, mx, 1 if animated else 0))
    return "\n".join(lines) + "\n"


def read_meta(data):
    """Independent reader: (info dict, {(level, x, y, z): (node, values_pos, depth)})."""
    pos = 0

    def get(fmt):
        nonlocal pos
        v = struct.unpack_from("<" + fmt, data, pos)
        pos += struct.calcsize("<" + fmt)
        return v if len(v) > 1 else v[0]

    info = {}
    if data[:len(META_ID)] == META_ID:
        pos = len(META_ID)
        info["version"] = get("I")
        assert info["version"] <= 1
        hsize = get("Q")
        h0 = pos
        assert get("I") == 3
        info["max_subtree_depth"], info["max_full_level"] = get("II")
        assert pos - h0 == hsize
    else:
        info["version"], info["max_full_level"] = 0, 0
    pos = get("Q")
    info["bc"] = get("4f")
    bits, nbytes = get("QQ")
    assert bits == 8 * nbytes
    info["masks"] = list(data[pos:pos + nbytes])
    pos += nbytes
    if info["version"] == 0:
        info["max_subtree_depth"] = get("I")
    entries = {}
    for lv in range(get("Q")):
        for _ in range(get("Q")):
            x, y, z, node = get("3II")
            vpos = get("Q")
            depth = get("Q") if info["version"] >= 1 else info["max_subtree_depth"]
            entries[(lv, x, y, z)] = (node, vpos, depth)
    info["field_file_size"] = get("Q")
    assert pos == len(data)
    return info, entries
