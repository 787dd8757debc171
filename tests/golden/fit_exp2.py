"""Generator of the polynomial P(f) ~ 2^f on [0, 1] of the pow replacement
(pow01 in revise_amd/csrc/render_common.hpp, orc_powf in oracle/oracle.cpp):

    P(f) = 1 + c1 f + c2 f^2 + ... + c6 f^6

with P(0) = 1 exactly and c1..c6 a relative-error minimax fit by Lawson's
algorithm (Lawson 1961; iteratively reweighted least squares, each point's
weight multiplied by its current error, then normalised) on 4097 Chebyshev
points of [0, 1], 3000 passes in float64, rounded to float32.  The log2 half
of the pow replacement needs no fit: 2 atanh(s) / ln 2 with the series
coefficients 1, 1/3, 1/5, 1/7, 1/9 and 2 / ln 2.

Run: python tests/golden/fit_exp2.py  (prints the C literals, c6 first, in
the Horner order of the sources)."""
import numpy as np

POINTS = 4097
PASSES = 3000


def fit():
    k = np.arange(POINTS)
    f = 0.5 - 0.5 * np.cos(np.pi * k / (POINTS - 1))
    y = np.exp2(f)
    A = np.stack([f ** i for i in range(1, 7)], axis=1)
    b = y - 1.0
    w = np.full(POINTS, 1.0 / POINTS)
    c = None
    for _ in range(PASSES):
        s = np.sqrt(w) / y  # relative error: residuals divided by 2^f
        c, *_ = np.linalg.lstsq(A * s[:, None], b * s, rcond=None)
        e = np.abs(A @ c - b) / y
        w = w * e
        w /= w.sum()
    return [float(np.float32(v)) for v in c]  # c1 .. c6


def c_literal(v):
    """A float32 value as a C hex-float literal (0x1.xxxxxxp-N f)."""
    m, e = np.frexp(np.float32(v))  # v = m 2^e, 0.5 <= m < 1
    frac = int(round((float(m) * 2.0 - 1.0) * (1 << 23)))
    return "0x1.%06xp%df" % (frac << 1, int(e) - 1)


if __name__ == "__main__":
    for v in reversed(fit()):
        print(c_literal(v))
