"""The utils package's array helper on the host: the reference's own
known-answer cases (tests/unit/test_sample_cosine.py:9-73) on NumPy arrays
and on torch tensors, and agreement with the oracle's bubble profile."""
import numpy as np
import pytest
import torch

import oracle as O
from pyminiweather_amd.utils import TimedCodeBlock, sample_ellipse_cosine


def _mesh(kind):
    p = O.make_params(200, 100)
    x, z = O.mesh_int_ext(p)
    if kind == "torch":
        x, z = torch.from_numpy(x), torch.from_numpy(z)
    return p, x, z


def _np(a):
    return a.numpy() if isinstance(a, torch.Tensor) else a


@pytest.mark.parametrize("kind", ["numpy", "torch"])
def test_single_point_bubble(kind):
    p, x, z = _mesh(kind)
    d = _np(sample_ellipse_cosine(x, z, 1.0, p["xlen"] / 2, p["zlen"] / 2, 1e-4, 1e-4))
    assert d.shape == tuple(z.shape) and d.dtype == np.float64
    assert np.count_nonzero(d == 0.0) == d.size - 1


@pytest.mark.parametrize("kind", ["numpy", "torch"])
def test_profile_matches_closed_form_and_oracle(kind):
    p, x, z = _mesh(kind)
    x0, z0, r = p["xlen"] / 2, p["zlen"] / 2, 1e4
    d = _np(sample_ellipse_cosine(x, z, 1.0, x0, z0, r, r))
    xn, zn = _np(x), _np(z)
    rr = np.hypot(xn - x0, zn - z0) / r
    expect = np.where(rr <= 1.0, np.cos(rr * np.pi / 2) ** 2, 0.0)
    assert np.allclose(d, expect, rtol=0, atol=1e-14)
    assert np.count_nonzero(d) == np.count_nonzero(rr <= 1.0)
    ref = O.sample_ellipse_cosine(xn, zn, 1.0, x0, z0, r, r)
    np.testing.assert_allclose(d, ref, rtol=0, atol=1e-15)


def test_elliptic_amplitude_and_outside_zero():
    p, x, z = _mesh("numpy")
    args = (3.0, 0.3 * p["xlen"], 0.25 * p["zlen"], 4000.0, 1500.0)
    a = sample_ellipse_cosine(x, z, *args)
    b = _np(sample_ellipse_cosine(torch.from_numpy(x), torch.from_numpy(z), *args))
    ref = O.sample_ellipse_cosine(x, z, *args)
    np.testing.assert_array_equal(a, ref)
    np.testing.assert_allclose(b, ref, rtol=0, atol=1e-15)
    assert 0.0 < a.max() <= 3.0 and a.min() == 0.0


def test_timed_code_block_logs_label_and_seconds(caplog):
    import logging
    with caplog.at_level(logging.INFO):
        with TimedCodeBlock("host block"):
            sum(range(1000))
    assert any("host block" in r.getMessage() for r in caplog.records)
