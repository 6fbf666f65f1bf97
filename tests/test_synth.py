"""The synthetic clouds' noise-free patterns (used to grade ridge quality in the bench): samples
drawn without noise or outliers lie on them."""

import numpy as np

from ridgedetection_amd import synth


def _nearest(a, b):
    """Largest distance from a point of a to its nearest point of b (float64, 32 rows of a at a
    time)."""
    worst = 0.0
    for i in range(0, len(a), 32):
        d = ((a[i:i + 32, None, :] - b[None, :, :]) ** 2).sum(-1)
        worst = max(worst, float(np.sqrt(d.min(axis=1)).max()))
    return worst


def test_bezier_union_curves_hold_the_noise_free_samples():
    P = synth.bezier_union(400, dim=6, curves=4, seed=6, sigma=0.0, outlier_frac=0.0, dtype=np.float64)
    C = synth.bezier_union_curves(dim=6, curves=4, seed=6, per_curve=20000)
    assert C.shape == (80000, 6)
    # 20000 parameters per curve: neighbouring curve points are well under 1e-3 apart
    assert _nearest(P, C) < 1e-3
    # a different seed draws other control points
    assert _nearest(P, synth.bezier_union_curves(dim=6, curves=4, seed=7, per_curve=2000)) > 0.05


def test_helix_curve_holds_the_noise_free_samples():
    P = synth.helix3d(300, seed=3, sigma=0.0, dtype=np.float64)
    C = synth.helix_curve(np.linspace(0.0, 2.0 * np.pi * 5.0, 50000))
    assert _nearest(P, C) < 1e-3
