"""Formatter of block orders in the layout of the reference's BlockOrderPrinter
(tests/test_BackToFrontOrder.cpp:66-281), used to compare against the golden
strings that test holds.  Input: list of block positions in visit order."""


def _row(entries, eye_idx, decorators):
    # entries: dict pos -> ord
    out = []
    pos = None
    for p in sorted(entries):
        if pos is None:
            pos = p
        else:
            pos += 1
            while pos != p:
                out.append(_cell(pos, None, eye_idx, decorators))
                pos += 1
        out.append(_cell(p, entries[p], eye_idx, decorators))
    return "".join(out)


def _cell(pos, ordv, eye_idx, decorators):
    dec = decorators[1 if pos == eye_idx else 0]
    s = " " + dec + " "
    s += "---" if ordv is None else "%3d" % ordv
    s += " " + dec + " "
    return s


def format_order(dim, blocks, eye_idx):
    """blocks: list of tuples (dim ints) in order; eye_idx: tuple with None for
    axes where the eye is outside."""
    if dim == 1:
        return _row({b[0]: i for i, b in enumerate(blocks)}, eye_idx[0], " *")
    if dim == 2:
        return _fmt2({(b[0], b[1]): i for i, b in enumerate(blocks)}, eye_idx, " *#")
    planes = {}
    for i, b in enumerate(blocks):
        planes.setdefault(b[2], {})[(b[0], b[1])] = i
    s = ""
    for z in sorted(planes):
        dec = " *#@"[1 if z == eye_idx[2] else 0:]
        s += "Plane %d:\n" % z
        s += _fmt2(planes[z], eye_idx, dec) + "\n"
    return s


def _fmt2(cells, eye_idx, decorators):
    rows = {}
    for (x, y), o in cells.items():
        rows.setdefault(y, {})[x] = o
    s = ""
    row = None
    for y in sorted(rows):
        if row is None:
            row = y
        else:
            row += 1
            while row != y:
                s += "\n"
                row += 1
        dec = decorators[1 if y == eye_idx[1] else 0:]
        s += _row(rows[y], eye_idx[0], dec) + "\n"
    return s
