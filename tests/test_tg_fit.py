"""The glass-transition fits of the cooling scan (core/src/mechanics.cpp:1218-1379, host code of the library) against
data whose answer is known: points on two exact lines give the hinge / crossing, the slopes and a zero bootstrap error;
noisy points give a Tg near the true one with a finite error; the order of the points does not matter; and the fits say
why when they cannot give a Tg."""
import numpy as np

import caps_amd as ca

T = np.arange(200.0, 520.0, 20.0)


def _two_lines(tg, lo, hi, v0=1.0):
    return v0 + lo * np.minimum(T - tg, 0) + hi * np.maximum(T - tg, 0)


def test_bilinear_recovers_an_exact_hinge():
    y = _two_lines(357.5, 2e-4, 6e-4)
    f = ca.fit_tg(T, y)
    assert f.ok and f.note == ""
    assert abs(f.tg - 357.5) < 1e-6 and abs(f.value_at_tg - 1.0) < 1e-12
    assert abs(f.slope_low - 2e-4) < 1e-12 and abs(f.slope_high - 6e-4) < 1e-12
    assert abs(f.alpha_low - 2e-4 / f.value_at_tg) < 1e-15 and f.tg_err < 1e-6 and f.rss < 1e-20


def test_two_ranges_cross_where_the_lines_cross():
    # glassy points on v = 1 + 2e-4 (T - 340), rubbery ones on v = 1 + 6e-4 (T - 340); the middle third is left out and
    # holds other values on purpose
    y = _two_lines(340.0, 2e-4, 6e-4)
    mid = (T > 200 + 320 / 3) & (T < 500 - 320 / 3)
    y[mid] += 0.05
    f = ca.fit_tg(T, y, fit="ranges")
    assert f.ok and abs(f.tg - 340.0) < 1e-6 and f.tg_err < 1e-6
    assert abs(f.slope_low - 2e-4) < 1e-12 and abs(f.slope_high - 6e-4) < 1e-12
    # explicit ranges
    f2 = ca.fit_tg(T, _two_lines(340.0, 2e-4, 6e-4), fit="ranges", glassy_max=300, rubbery_min=400)
    assert f2.ok and abs(f2.tg - 340.0) < 1e-6


def test_order_does_not_matter_and_noise_gives_an_error():
    rng = np.random.default_rng(3)
    y = _two_lines(360.0, 2e-4, 6e-4) + rng.normal(0, 2e-4, T.size)
    f = ca.fit_tg(T, y)
    p = rng.permutation(T.size)
    g = ca.fit_tg(T[p], y[p])
    assert f.ok and abs(f.tg - 360.0) < 25 and 0 < f.tg_err < 50
    assert f.tg == g.tg and f.tg_err == g.tg_err


def test_the_fits_say_why_they_give_no_tg():
    assert ca.fit_tg(T[:4], _two_lines(300, 2e-4, 6e-4)[:4]).note.startswith("too few")
    f = ca.fit_tg(T, 1.0 + 2e-4 * T)              # one straight line: slopes within 10 %
    assert "10 %" in f.note
    f = ca.fit_tg(T, 2.0 * T, fit="ranges")           # exactly parallel lines
    assert not f.ok and "parallel" in f.note
    f = ca.fit_tg(T, 1.0 + 2e-4 * T, fit="ranges")   # parallel up to rounding: a crossing far outside the gap
    assert not f.ok and "outside the gap" in f.note
    f = ca.fit_tg(T, _two_lines(300, 2e-4, 6e-4), fit="ranges", glassy_max=400, rubbery_min=300)
    assert not f.ok and "glassy range" in f.note
