"""Constants compiled into the kernels (CPU only).

kExp2Tab (psim_device.hpp) must hold 2^(j/64), j = 0..63, each the double
nearest the exact value: checked against a 60-digit decimal evaluation.  The
table-driven exponential's reduction constant -64/ln 2 is checked the same
way, and the exponential's steps, replayed in Python double arithmetic with
the source's constants (fused multiply-adds evaluated exactly, then rounded
once), stay within 2 ulp of a 50-digit e^{-a}."""
import math
import os
import re
from decimal import Decimal, getcontext

import numpy as np

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "proteinsim2015_amd", "csrc")


def _source():
    with open(os.path.join(CSRC, "psim_device.hpp")) as f:
        return f.read()


def test_exp2_table_is_correctly_rounded():
    src = _source()
    body = re.search(r"kExp2Tab\[64\]\s*=\s*\{([^}]*)\}", src).group(1)
    vals = [float.fromhex(v.strip()) for v in body.split(",") if v.strip()]
    assert len(vals) == 64
    getcontext().prec = 60
    for j, v in enumerate(vals):
        assert v == float(Decimal(2) ** (Decimal(j) / Decimal(64))), j


def test_exp_reduction_constant():
    src = _source()
    inv = float(re.search(r"const double inv = ([0-9.eE+-]+);\s*// -64 / ln 2", src).group(1))
    getcontext().prec = 60
    assert inv == float(-Decimal(64) / Decimal(2).ln())


def _fma(a, b, c):
    return float(Decimal(a) * Decimal(b) + Decimal(c))


def test_table_exponential_within_two_ulp():
    src = _source()
    body = re.search(r"kExp2Tab\[64\]\s*=\s*\{([^}]*)\}", src).group(1)
    tab = [float.fromhex(v.strip()) for v in body.split(",") if v.strip()]
    fn = src[src.index("double exp_neg_tab("):]
    fn = fn[:fn.index("\n}")]
    inv = float(re.search(r"inv = ([0-9.eE+-]+);", fn).group(1))
    hi = float(re.search(r"ln2_hi = ([0-9.eE+-]+) / 64\.0;", fn).group(1)) / 64.0
    lo = float(re.search(r"ln2_lo = ([0-9.eE+-]+) / 64\.0;", fn).group(1)) / 64.0
    q0 = 1.0 / float(re.search(r"double q = 1\.0 / ([0-9.]+);", fn).group(1))
    coef = [1.0 / float(d) for d in re.findall(r"fma\(q, x, 1\.0 / ([0-9.]+)\)", fn)]
    assert [round(1.0 / c) for c in [q0] + coef] == [120, 24, 6], "degree-5 Taylor series"
    coef += [0.5, 1.0, 1.0]

    def exp_neg_tab(a):
        k = float(np.rint(a * inv))
        f = _fma(k, lo, _fma(k, hi, a))
        x = -f
        q = q0
        for c in coef:
            q = _fma(q, x, c)
        ki = int(k)
        return math.ldexp(q * tab[ki & 63], ki >> 6)

    getcontext().prec = 50
    rng = np.random.default_rng(7)
    pts = np.concatenate([rng.uniform(0.0, 40.0, 4000), rng.uniform(0.0, 0.05, 1000),
                          [0.0, math.log(2) / 128, 15.0, 699.0]])
    for a in pts:
        a = float(a)
        exact = Decimal(-a).exp()
        err = abs((Decimal(exp_neg_tab(a)) - exact) / Decimal(math.ulp(float(exact))))
        assert err < 2, (a, float(err))
