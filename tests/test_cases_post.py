"""Post-processing of the configuration runs (luma_amd/cases.py) on signals whose answer is known
in closed form: the Strouhal number from a lift history and the filament objectives from a tip
trajectory.  CPU only."""
import numpy as np
import pytest

from luma_amd import cases

D, U0, DSTEP = 40.0, 0.04 / 1.5, 20.0


def _lift(period, n=5000, amp=0.7, offset=0.05, phase=0.3):
    t = np.arange(n) * DSTEP
    return t, amp * np.sin(2.0 * np.pi * t / period + phase) + offset


@pytest.mark.parametrize("period", [2000.0, 1937.0, 2411.7, 1500.3])
def test_strouhal_of_a_pure_tone(period):
    # St = f D / U0 for a sinusoid of period P steps sampled every DSTEP steps; whole and
    # fractional numbers of periods in the analysed half (peak interpolation bias < 1e-3)
    _, cl = _lift(period)
    st, per = cases.strouhal(cl, DSTEP, D, U0)
    assert per == pytest.approx(period, rel=2e-3)
    assert st == pytest.approx(D / (period * U0), rel=2e-3)


def test_strouhal_reads_only_the_last_half():
    # a louder transient at another frequency in the first half must not set the answer
    t, late = _lift(1800.0)
    cl = np.where(t < t[len(t) // 2], 3.0 * np.sin(2.0 * np.pi * t / 700.0), late)
    st, per = cases.strouhal(cl, DSTEP, D, U0)
    assert per == pytest.approx(1800.0, rel=2e-3)


def test_strouhal_of_a_flat_or_short_signal_is_none():
    assert cases.strouhal(np.full(400, 0.3), DSTEP, D, U0) == (None, None)
    assert cases.strouhal(np.sin(np.arange(20.0)), DSTEP, D, U0) == (None, None)


def test_filament_objectives_closed_form():
    # tip y = y0 + dh (a + b sin(w s)) with a > b > 0 over whole periods of the second half:
    # mean |deflection| = mean deflection = a cells; with dy = dh (da + db sin(w s)) both
    # derivatives are da
    n, per, dh, y0 = 4000, 250, 0.5, 3.0
    s = np.arange(n)
    a, b, da, db = 0.25, 0.1, 2.0, -7.0
    tips = y0 + dh * (a + b * np.sin(2.0 * np.pi * s / per))
    dtips = dh * (da + db * np.sin(2.0 * np.pi * s / per))
    out = cases.c5_objectives(tips, dtips, y0, dh)
    assert out["J_mean_amplitude_cells"] == pytest.approx(a, abs=1e-12)
    assert out["J_mean_deflection_cells"] == pytest.approx(a, abs=1e-12)
    assert out["dJ_mean_amplitude"] == pytest.approx(da, abs=1e-12)
    assert out["dJ_mean_deflection"] == pytest.approx(da, abs=1e-12)


def test_filament_objectives_amplitude_takes_the_magnitude():
    # a deflection that changes sign: mean |a sin| over whole periods = 2a/pi, mean deflection 0,
    # d(mean |dev|) = mean(sign(dev) * ddev) = 2 da / pi for ddev = da sin
    n, per, dh, y0 = 4000, 400, 1.0, 0.0
    s = np.arange(n) + 0.5                     # no sample at a zero crossing
    a, da = 0.3, 1.5
    tips = y0 + dh * a * np.sin(2.0 * np.pi * s / per)
    dtips = dh * da * np.sin(2.0 * np.pi * s / per)
    out = cases.c5_objectives(tips, dtips, y0, dh)
    assert out["J_mean_amplitude_cells"] == pytest.approx(2.0 * a / np.pi, rel=1e-4)
    assert abs(out["J_mean_deflection_cells"]) < 1e-12
    assert out["dJ_mean_amplitude"] == pytest.approx(2.0 * da / np.pi, rel=1e-4)
    assert cases.c5_objectives(tips, None, y0, dh).keys() == {"J_mean_amplitude_cells", "J_mean_deflection_cells"}
