"""The IBI table update (V += alpha k_B T ln(g / g_target)) on the CPU: what the formula fixes."""
import numpy as np

from ibisco_amd.ibi import R_KJMOL, ibi_correction, rdf_distance

DR, NBIN, T = 0.01, 150, 500.0
RB = (np.arange(NBIN) + 0.5) * DR
R_TAB = np.arange(751) * 0.002


def _target():
    g = 1.0 + 0.6 * np.exp(-((RB - 0.52) / 0.06) ** 2) - 0.2 * np.exp(-((RB - 0.8) / 0.1) ** 2)
    g[RB < 0.35] = 0.0          # the excluded core
    return g


def test_no_change_at_the_target():
    gt = _target()
    assert np.all(ibi_correction(R_TAB, gt, gt, DR, T, alpha=0.7) == 0.0)
    assert rdf_distance(gt, gt) == 0.0


def test_known_log_ratio_is_recovered():
    """g = g_t exp(c(r) / kT) for a smooth c vanishing before the cutoff: the correction is alpha * c at the bin
    centres (no smoothing), and the table update makes V larger where the system is too structured."""
    gt = _target()
    kt = R_KJMOL * T
    c = 0.8 * np.exp(-((RB - 0.6) / 0.08) ** 2)
    g = gt * np.exp(c / kt)
    got = ibi_correction(RB, g, gt, DR, T, alpha=0.5, smooth=0)
    valid = gt > 1e-2
    assert np.allclose(got[valid], 0.5 * c[valid] - 0.5 * c[-1], rtol=0, atol=1e-12)
    assert got[np.argmin(np.abs(RB - 0.6))] > 0.3 * 0.8


def test_core_carries_the_innermost_correction_and_zero_at_cutoff():
    gt = _target()
    g = gt * 1.1
    got = ibi_correction(R_TAB, g, gt, DR, T, alpha=1.0, smooth=0)
    assert got[-1] == 0.0
    core = R_TAB < 0.3
    assert np.all(got[core] == got[0])          # constant through the core
    # a uniform ratio is a constant shift, removed by the cutoff shift
    assert np.allclose(got, 0.0, atol=1e-12)


def test_smoothing_preserves_a_constant_and_the_distance_measures_rms():
    gt = _target()
    g = gt.copy()
    g[60] += 0.1
    d = rdf_distance(g, gt)
    m = gt > 1e-2
    assert abs(d - 0.1 / np.sqrt(m.sum())) < 1e-12
