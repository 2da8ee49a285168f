"""The reference's structured-grid error norms (sgh/io_utils/compute_L2_error.py:436-530),
pinned against SciPy's trapezoid rule and closed forms, on NumPy and torch inputs."""

import numpy as np
import pytest
from scipy.integrate import trapezoid

from spectral_galerkin_heat_amd.io_utils import compute_L2_structured, trapezoidal_weights


def test_weights_uniform_and_nonuniform():
    x = np.linspace(0.0, 2.0, 5)
    np.testing.assert_allclose(trapezoidal_weights(x), [0.25, 0.5, 0.5, 0.5, 0.25])
    xn = np.array([0.0, 0.1, 0.4, 1.0])
    w = trapezoidal_weights(xn)
    np.testing.assert_allclose(w, [0.05, 0.2, 0.45, 0.3])
    # integrates any linear function exactly, on any spacing
    f = 3.0 * xn - 1.0
    assert abs(np.dot(w, f) - (1.5 * 1.0 - 1.0)) < 1e-14
    assert trapezoidal_weights([1.0]).tolist() == [0.0]


def _nested_trapezoid(F, x, y, z):
    return trapezoid(trapezoid(trapezoid(F, x, axis=2), y, axis=1), z, axis=0)


def test_norms_equal_nested_scipy_trapezoid():
    rng = np.random.default_rng(4)
    x = np.sort(rng.uniform(0, 1e-3, 9))
    y = np.linspace(0, 5e-4, 7)
    z = np.sort(rng.uniform(0, 2e-4, 5))
    Ta = 300 + 100 * rng.random((5, 7, 9))
    Tb = Ta + rng.normal(0, 0.5, Ta.shape)
    out = compute_L2_structured(Ta, Tb, x, y, z)
    e2 = _nested_trapezoid((Ta - Tb) ** 2, x, y, z)
    r2 = _nested_trapezoid(Ta ** 2, x, y, z)
    assert abs(out["L2_abs"] - e2 ** 0.5) < 1e-12 * e2 ** 0.5
    assert abs(out["L2_rel"] - (e2 / r2) ** 0.5) < 1e-12
    assert out["Linf"] == pytest.approx(np.abs(Ta - Tb).max())
    assert out["volume"] == pytest.approx((x[-1] - x[0]) * (y[-1] - y[0]) * (z[-1] - z[0]))


def test_torch_inputs_match_numpy_and_bad_fields_raise():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    x, y, z = np.linspace(0, 1, 6), np.linspace(0, 2, 4), np.linspace(0, 3, 3)
    Ta = rng.random((3, 4, 6))
    Tb = Ta * (1 + 1e-3 * rng.random(Ta.shape))
    ref = compute_L2_structured(Ta, Tb, x, y, z)
    got = compute_L2_structured(torch.tensor(Ta), torch.tensor(Tb, dtype=torch.float32), x, y, z)
    assert got["L2_rel"] == pytest.approx(
        compute_L2_structured(Ta, Tb.astype(np.float32), x, y, z)["L2_rel"], rel=1e-12)
    assert ref["L2_rel"] > 0
    Tb_bad = Tb.copy()
    Tb_bad[0, 0, 0] = np.nan
    with pytest.raises(ValueError):
        compute_L2_structured(Ta, Tb_bad, x, y, z)
    with pytest.raises(ValueError):
        compute_L2_structured(Ta[:, :, :5], Tb[:, :, :5], x, y, z)
