"""The replay harness's own exp/expm1, pinned against libm (numpy float64 exp/expm1)
by ulp error over dense samples, plus the special values the kernels rely on."""
import math

import numpy as np

from oracle import orc


def _ulp32(x):
    x = np.float32(abs(x))
    return float(np.spacing(x)) if x > 0 else float(np.spacing(np.float32(0)))


def _max_ulp_err(f, ref, xs):
    worst = 0.0
    for x in xs:
        got = float(f(float(x)))
        want = ref(float(np.float32(x)))
        if want == 0.0 and got == 0.0:
            continue
        err = abs(got - want) / _ulp32(np.float32(want))
        worst = max(worst, err)
    return worst


def test_expm1f_within_2ulp():
    xs = np.concatenate([np.linspace(-17.4, 0.0, 4001), -np.logspace(-12, 1.2, 2001),
                         np.linspace(0, 80, 801)]).astype(np.float32)
    assert _max_ulp_err(orc.expm1f, math.expm1, xs) <= 2.0


def test_expf_within_2ulp():
    xs = np.concatenate([np.linspace(-86.9, 0.0, 6001), np.linspace(0, 80, 801)]).astype(np.float32)
    assert _max_ulp_err(orc.expf, math.exp, xs) <= 2.0


def test_exp64_within_2ulp():
    xs = np.concatenate([np.linspace(-707.9, 0, 20001), np.linspace(0, 700, 2001),
                         -np.logspace(-15, 2, 2001)])
    worst = 0.0
    for x in xs:
        got, want = orc.exp64(float(x)), math.exp(float(x))
        worst = max(worst, abs(got - want) / np.spacing(want))
    assert worst <= 2.0


def test_special_values():
    assert orc.expm1f(0.0) == 0.0
    assert orc.expm1f(-20.0) == -1.0            # e^x below 2^-25
    assert orc.expf(0.0) == 1.0
    assert orc.expf(-87.5) == 0.0               # flushed: results below FLT_MIN
    assert orc.exp64(0.0) == 1.0
    assert orc.exp64(-709.0) == 0.0             # flushed: results below DBL_MIN
    assert math.isinf(orc.exp64(710.0))
    # tiny arguments: expm1(x) == x to float32 precision (x^2/2 far below half an ulp)
    for x in (1e-30, 1e-10, 1e-9):
        assert orc.expm1f(-x) == np.float32(-x)
    # at 1e-7, x^2/2 is 0.7 ulp of x: within one ulp
    assert abs(float(orc.expm1f(-1e-7)) - math.expm1(float(np.float32(-1e-7)))) <= _ulp32(1e-7)


def test_fire_threshold_is_monotone_for_every_float():
    # The kernel skips the nearest-hole search of an electron whose draw fails
    # against its old (closer) hole: exact only if T(x) never decreases as x
    # grows.  Every float in [0, 30] (x >= 17.5 already gives p = 1).
    bad, worst = orc.thr_monotone_scan(0.0, 30.0)
    assert (bad, worst) == (0, 0)


def test_fire_threshold_is_zero_below_two_to_minus_32():
    # Rows whose x are all below 2^-32 skip their draws: T must be 0 there.  With
    # T monotone (above), the largest float below 2^-32 settles every smaller x.
    below = np.nextafter(np.float32(2.0 ** -32), np.float32(0.0))
    assert orc.fire_threshold(float(below)) == 0
    assert orc.fire_threshold(2.0 ** -31) > 0
