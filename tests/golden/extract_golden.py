"""Write the golden-vector fixtures of the reference's own unit tests as JSON.

Reads (as text) the reference's gtest files and extracts the expected values
of their EXPECT_EQ assertions:

* tests/test_BackToFrontOrder.cpp:346-683  -> b2fo_golden.json
  (format_b2fo_result<N>(size, eye) == "<expected string>")
* tests/test_IndexRangeSplitter.cpp:30-98  -> splitter_golden.json
  (splitter(x) == IndexRangeSplit(before, at, after))

Only inputs and expected outputs are written; no reference code is copied.
Run from the repository root:  python tests/golden/extract_golden.py
"""
import json
import os
import re
import sys

REF = os.environ.get("REVISE_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def _c_string_concat(block):
    """Concatenate adjacent C string literals, decoding \\n escapes."""
    parts = re.findall(r'"((?:[^"\\]|\\.)*)"', block)
    return "".join(p.encode().decode("unicode_escape") for p in parts)


def extract_b2fo():
    src = open(os.path.join(REF, "tests/test_BackToFrontOrder.cpp")).read()
    cases = []
    # EXPECT_EQ(format_b2fo_result<N>(size, eye), "..." "...");
    pat = re.compile(
        r"EXPECT_EQ\(\s*format_b2fo_result<(\d)>\((\d+),\s*(\{[^}]*\}|[-\d.]+)\),\s*((?:\s*\"(?:[^\"\\]|\\.)*\")+)\s*\);",
        re.S)
    for m in pat.finditer(src):
        dim = int(m.group(1))
        size = int(m.group(2))
        eye_txt = m.group(3).strip("{} ")
        eye = [float(v) for v in eye_txt.split(",")]
        expected = _c_string_concat(m.group(4))
        line = src[: m.start()].count("\n") + 1
        cases.append({"dim": dim, "size": size, "eye": eye, "expected": expected,
                      "source": "tests/test_BackToFrontOrder.cpp:%d" % line})
    return cases


def _parse_ir(tok):
    tok = tok.strip()
    if tok == "empty_r":
        return None
    m = re.match(r"IR\{\s*(\d+)\s*,\s*(\d+)\s*\}", tok)
    assert m, tok
    return [int(m.group(1)), int(m.group(2))]


def _parse_at(tok):
    tok = tok.strip()
    return None if tok == "~0u" else int(tok.rstrip("u"))


def extract_splitters():
    src = open(os.path.join(REF, "tests/test_IndexRangeSplitter.cpp")).read()
    out = {"uniform": {"index_range": [10, 20], "coord_range": [100.0, 101.0], "cases": []},
           "nonuniform": {"start": 5, "split_coords": [102.0, 105.0, 109.0],
                          "coord_range": [100.0, 110.0], "cases": []}}
    pat = re.compile(
        r"EXPECT_EQ\(splitter\(\s*([\d.]+)\s*\),\s*IndexRangeSplit\(\s*(empty_r|IR\{[^}]*\})\s*,\s*(~0u|\d+u)\s*,\s*(empty_r|IR\{[^}]*\})\s*\)\s*\);")
    # the uniform test comes first in the file, the non-uniform second
    split_at = src.index("NonUniformIndexRangeSplitterTest")
    for m in pat.finditer(src):
        rec = {"x": float(m.group(1)), "before": _parse_ir(m.group(2)),
               "at": _parse_at(m.group(3)), "after": _parse_ir(m.group(4)),
               "source": "tests/test_IndexRangeSplitter.cpp:%d" % (src[: m.start()].count("\n") + 1)}
        key = "uniform" if m.start() < split_at else "nonuniform"
        out[key]["cases"].append(rec)
    return out


def main():
    b2fo = extract_b2fo()
    spl = extract_splitters()
    assert len(b2fo) == 33, len(b2fo)
    assert len(spl["uniform"]["cases"]) == 28, len(spl["uniform"]["cases"])
    assert len(spl["nonuniform"]["cases"]) == 18, len(spl["nonuniform"]["cases"])
    with open(os.path.join(HERE, "b2fo_golden.json"), "w") as f:
        json.dump(b2fo, f, indent=1)
    with open(os.path.join(HERE, "splitter_golden.json"), "w") as f:
        json.dump(spl, f, indent=1)
    print("wrote %d b2fo cases, %d+%d splitter cases" % (
        len(b2fo), len(spl["uniform"]["cases"]), len(spl["nonuniform"]["cases"])))
    return 0


if __name__ == "__main__":
    sys.exit(main())
