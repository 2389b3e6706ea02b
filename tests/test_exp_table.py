"""The constants of the FP64 kernels' table-driven exp (exp_tab in
ks_amd/csrc/gsks_kernels.h), read from the source and checked against 60-digit
arithmetic: the 64 table points 2^(j/64) as head + tail (head correctly
rounded, tail the rounded remainder), 64/ln 2, and ln2/64 split so that its
first part has at most 36 significant bits (n times it is then exact for the
|n| < 2^17 the clamped argument allows).  The same steps, restated with exact
fused multiply-adds, stay within 0.76 ulp of exp over sampled arguments; the
device function itself is checked on the GPU against numpy's exp
(test_gpu_dgsks.py::test_gaussian_exp_ulps)."""
import math
import os
import re
from fractions import Fraction

import pytest

mpmath = pytest.importorskip("mpmath")

SRC = os.path.join(os.path.dirname(__file__), "..", "ks_amd", "csrc", "gsks_kernels.h")
HEX = r"-?0x[0-9a-fA-F.]+p[-+]?\d+"


def _constants():
    text = open(SRC).read()
    const = {name: float.fromhex(val) for name, val in
             re.findall(r"constexpr double (kExp\w+) = (" + HEX + ");", text)}
    body = text[text.index("kExpTable[64] = {"):]
    body = body[:body.index("};")]
    pairs = re.findall(r"\{(" + HEX + r"), (" + HEX + r")\}", body)
    table = [(float.fromhex(h), float.fromhex(t)) for h, t in pairs]
    return const, table


def test_table_and_reduction_constants():
    mpmath.mp.dps = 60
    const, table = _constants()
    assert len(table) == 64
    for j, (head, tail) in enumerate(table):
        exact = mpmath.power(2, mpmath.mpf(j) / 64)
        # head correctly rounded: no neighbouring double is closer
        err = abs(mpmath.mpf(head) - exact)
        assert err <= abs(mpmath.mpf(math.nextafter(head, math.inf)) - exact), j
        assert err <= abs(mpmath.mpf(math.nextafter(head, -math.inf)) - exact), j
        assert tail == float(exact - mpmath.mpf(head)), j
    ln2_64 = mpmath.log(2) / 64
    assert const["kExpC"] == float(64 / mpmath.log(2))
    l1, l2 = const["kExpL1"], const["kExpL2"]
    mant = Fraction(l1) / Fraction(2) ** (math.frexp(l1)[1] - 36)
    assert mant.denominator == 1  # at most 36 significant bits
    assert abs(mpmath.mpf(l1) - ln2_64) < mpmath.mpf(2) ** -40
    assert l2 == float(ln2_64 - mpmath.mpf(l1))


def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _exp_restated(x, const, table):
    """exp_tab's steps with exact fused multiply-adds (test-only restatement)."""
    magic = float.fromhex("0x1.8p52")
    x = min(max(x, -746.0), 710.0)
    s = _fma(x, const["kExpC"], magic)
    nf = s - magic
    n = int(nf)
    r = _fma(-nf, const["kExpL1"], x)
    r = _fma(-nf, const["kExpL2"], r)
    q = _fma(r, 1.0 / 120.0, 1.0 / 24.0)
    q = _fma(q, r, 1.0 / 6.0)
    q = _fma(q, r, 0.5)
    q *= r
    p = _fma(q, r, r)
    head, tail = table[n & 63]
    y = head + _fma(head, p, tail)
    try:
        return math.ldexp(y, n >> 6)
    except OverflowError:
        return math.inf


def test_restated_steps_within_bound():
    import random
    mpmath.mp.dps = 40
    const, table = _constants()
    rng = random.Random(7)
    xs = [0.0, -1e-300, -0.5, -math.log(2), -708.4, -745.1, 1.0, 709.7]
    xs += [-rng.uniform(0, 1) for _ in range(300)] + [-rng.uniform(0, 745) for _ in range(300)]
    xs += [rng.uniform(0, 709) for _ in range(100)]
    tiny = 2.2250738585072014e-308
    worst = 0.0
    for x in xs:
        got = _exp_restated(x, const, table)
        want = mpmath.exp(mpmath.mpf(x))
        if float(want) >= tiny:
            worst = max(worst, float(abs(mpmath.mpf(got) - want) / math.ulp(float(want))))
        else:
            assert abs(mpmath.mpf(got) - want) <= 5e-324
    assert worst <= 0.76, worst
    assert _exp_restated(0.0, const, table) == 1.0
    assert _exp_restated(-745.2, const, table) == 0.0
    assert _exp_restated(710.0, const, table) == math.inf
