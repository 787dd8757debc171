"""The pow replacement's constants come from their generator: the degree-6
exp2 polynomial of pow01 (render_common.hpp) and orc_powf (oracle.cpp) is
what tests/golden/fit_exp2.py computes, and the log2 series coefficients are
float32 1/3, 1/5, 1/7, 1/9 and 2/ln 2."""
import importlib.util
import math
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _generator():
    spec = importlib.util.spec_from_file_location("fit_exp2", os.path.join(ROOT, "tests", "golden", "fit_exp2.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _hex_literals(path):
    """The hex-float literals of the exp2 Horner chain, in source order."""
    text = open(os.path.join(ROOT, path)).read()
    return [float.fromhex(x[:-1]) for x in re.findall(r"0x1\.[0-9a-f]+p-\d+f", text)]


def test_exp2_coefficients_are_the_generators():
    gen = _generator()
    c = gen.fit()  # c1 .. c6
    expect = list(reversed(c))  # the sources' Horner order: c6 first
    for path in ("revise_amd/csrc/render_common.hpp", "oracle/oracle.cpp"):
        got = _hex_literals(path)
        assert got == expect, (path, [v.hex() for v in got], [v.hex() for v in expect])
    # the literal printer round-trips
    for v in c:
        assert float.fromhex(gen.c_literal(v)[:-1]) == v


def test_exp2_fit_error():
    """The float32 polynomial with float32 Horner steps stays within 2e-7
    relative of 2^f on [0, 1] (the fit alone: about 2e-9)."""
    c = _generator().fit()
    f = np.linspace(0.0, 1.0, 20001).astype(np.float32)
    p = np.float32(c[5])
    for k in range(4, -1, -1):
        p = (p.astype(np.float64) * f + np.float32(c[k])).astype(np.float32)
    p = (p.astype(np.float64) * f + 1.0).astype(np.float32)
    assert np.max(np.abs(p / np.exp2(f.astype(np.float64)) - 1.0)) < 2e-7


def test_log2_series_constants():
    for path in ("revise_amd/csrc/render_common.hpp", "oracle/oracle.cpp"):
        text = open(os.path.join(ROOT, path)).read()
        for lit, val in (("0.111111111f", 1 / 9), ("0.142857143f", 1 / 7), ("0.333333333f", 1 / 3),
                         ("2.88539008f", 2 / math.log(2))):
            assert lit in text, (path, lit)
            assert np.float32(float(lit[:-1])) == np.float32(val), lit
