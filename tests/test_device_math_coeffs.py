"""The fitted polynomials of hammurabix_amd/csrc/hamx_math.h, checked on the CPU.

The device functions sincos_series, log_pos and exp_fast evaluate polynomial
fits whose coefficients are typed into the header.  This test reads those
coefficients from the source, evaluates each function's exact operation
sequence in double precision (every fused multiply-add rounded once, through
60-digit mpmath arithmetic), and compares with mpmath's sin, cos, log and exp.
A mistyped or dropped coefficient, a swapped pair or a wrong evaluation order
moves the worst error far beyond the bounds asserted here.
"""
import math
import os
import random
import re

import pytest

mp = pytest.importorskip("mpmath")
mp.mp.dps = 60

HEADER = os.path.join(os.path.dirname(__file__), "..", "hammurabix_amd", "csrc", "hamx_math.h")
NUM = r"[-+]?\d+\.\d+e[-+]\d+"


def _body(name):
    src = open(HEADER).read()
    i = src.index(f"inline double {name}(") if name != "sincos_series" else \
        src.index("inline void sincos_series(")
    j = src.index("\n}\n", i)
    return src[i:j]


def fma(a, b, c):
    return float(mp.mpf(a) * mp.mpf(b) + mp.mpf(c))


def ulp(v):
    v = abs(mp.mpf(v))
    return mp.mpf(2) ** (mp.floor(mp.log(v, 2)) - 52)


def test_coefficients_are_read():
    sc = re.findall(NUM, _body("sincos_series"))
    lg = re.findall(NUM, _body("log_pos"))
    ex = re.findall(NUM, _body("exp_fast"))
    assert len(sc) == 12  # 6 for ps, 6 for pc
    assert len(lg) == 7 + 2  # the fit, then ln 2 in two parts (the frexp threshold has no exponent)
    assert len(ex) == 2 + 9  # ln 2 in two parts, the fit (1/ln 2 is read on its own)


def test_sincos_series_within_one_ulp():
    c = [float(v) for v in re.findall(NUM, _body("sincos_series"))]
    ps, pc = c[:6], c[6:]

    def sincos(x):
        x2 = x * x
        p = fma(x2, ps[0], ps[1])
        for k in ps[2:]:
            p = fma(x2, p, k)
        s = fma(x * x2, p, x)
        q = fma(x2, pc[0], pc[1])
        for k in pc[2:]:
            q = fma(x2, q, k)
        return s, fma(x2, fma(x2, q, -0.5), 1.0)

    rng = random.Random(5)
    xs = [0.8 * i / 1500 for i in range(1, 1501)] + [rng.uniform(-0.8, 0.8) for _ in range(500)]
    ws = wc = 0.0
    for x in xs:
        s, co = sincos(x)
        X = mp.mpf(x)
        ws = max(ws, float(abs(s - mp.sin(X)) / ulp(mp.sin(X))))
        wc = max(wc, float(abs(co - mp.cos(X)) / ulp(mp.cos(X))))
    print("sincos_series worst ulp: sin %.3f cos %.3f" % (ws, wc))
    assert ws <= 1.0 and wc <= 1.0


def test_log_pos_within_two_ulp():
    c = [float(v) for v in re.findall(NUM, _body("log_pos"))]
    fit, ln2_hi, ln2_lo = c[:7], c[7], c[8]
    assert ln2_hi + ln2_lo == math.log(2) and abs(mp.mpf(ln2_hi) + mp.mpf(ln2_lo) - mp.log(2)) < 1e-32

    def log_pos(x):
        m, e = math.frexp(x)
        if m < 0.70710678118654752:
            m *= 2.0
            e -= 1
        f = (m - 1.0) * (1.0 / (m + 1.0))  # the device's Newton-refined v_rcp_f64: within 1 ulp
        s = f * f
        p = fma(s, fit[0], fit[1])
        for k in fit[2:]:
            p = fma(s, p, k)
        lnm = fma(2.0 * f, s * p, 2.0 * f)
        ed = float(e)
        return fma(ed, ln2_hi, fma(ed, ln2_lo, lnm))

    rng = random.Random(7)
    xs = [10 ** rng.uniform(-30, 30) for _ in range(1200)] + \
         [1 + rng.uniform(-0.29, 0.41) for _ in range(800)] + [2 ** 0.5 * (1 + rng.uniform(-1e-4, 1e-4)) for _ in range(200)]
    w = 0.0
    for x in xs:
        ref = mp.log(mp.mpf(x))
        if abs(ref) < 1e-12:
            continue
        w = max(w, float(abs(log_pos(x) - ref) / ulp(ref)))
    print("log_pos worst ulp %.3f" % w)
    assert w <= 2.5


def test_exp_fast_within_one_ulp():
    body = _body("exp_fast")
    inv_ln2 = float(re.search(r"rint\(x \* ([0-9.]+)\)", body).group(1))
    assert abs(inv_ln2 - 1 / math.log(2)) < 1e-16
    c = [float(v) for v in re.findall(NUM, body)]
    ln2_hi, ln2_lo, fit = c[0], c[1], c[2:]

    def exp_fast(x):
        k = float(round(x * inv_ln2))  # rint: ties never occur on these samples
        r = fma(-k, ln2_hi, x)
        r = fma(-k, ln2_lo, r)
        p = fit[0]
        for q in fit[1:]:
            p = fma(p, r, q)
        p = fma(p, r, 0.5)
        p = fma(p, r, 1.0)
        p = fma(p, r, 1.0)
        return math.ldexp(p, int(k))

    rng = random.Random(11)
    xs = [rng.uniform(-700, 700) for _ in range(1000)] + [rng.uniform(-2, 2) for _ in range(1000)]
    w = 0.0
    for x in xs:
        ref = mp.exp(mp.mpf(x))
        w = max(w, float(abs(exp_fast(x) - ref) / ulp(ref)))
    print("exp_fast worst ulp %.3f" % w)
    assert w <= 1.0
