"""The package's data module against the reference's own known answers
(tests/unit/test_constants.py:7-64): constant values, the quadrature rule
and its immutability, and the field shapes of initialize_fields
(data/fields.py:58-123) on the host."""
import numpy as np
import pytest
import torch

from pyminiweather_amd.data import Constants, Quadrature, initialize_fields


def test_constants_known_values():
    assert Constants.hv_beta.value == 0.05 and Constants.p0.value == 1.0e5
    assert Constants.C0.value == 27.5629410929725921310572974482
    assert Constants.gamma.value == 1.40027894002789400278940027894
    assert Constants.grav.value == 9.8 and Constants.cp.value == 1004.0
    assert Constants.cv.value == 717.0 and Constants.rd.value == 287
    assert Constants.pi.value == 3.14159265358979323846264338327
    assert Constants.theta0.value == 300.0 and Constants.exner0.value == 1.0
    # gamma = cp / cv and C0 = rd^gamma p0^(-rd/cv), the closed forms behind
    # the tabulated digits
    assert abs(Constants.gamma.value - 1004.0 / 717.0) < 1e-15
    assert abs(Constants.C0.value - 287.0 ** Constants.gamma.value * 1e5 ** (-287.0 / 717.0)) \
        < 1e-12


def test_quadrature_known_values():
    assert Quadrature.npoints == 3
    assert np.allclose(Quadrature.qpoints, [0.112701665379258311482073460022, 0.5,
                                            0.887298334620741688517926539980])
    assert np.allclose(Quadrature.qweights, [0.277777777777777777777777777779,
                                             0.444444444444444444444444444444,
                                             0.277777777777777777777777777779])
    np.testing.assert_array_equal(Quadrature.qweights_outer,
                                  np.outer(Quadrature.qweights, Quadrature.qweights))
    gx, gz = np.meshgrid(Quadrature.qpoints, Quadrature.qpoints)
    np.testing.assert_array_equal(Quadrature.qpoints_grid_x, gx)
    np.testing.assert_array_equal(Quadrature.qpoints_grid_z, gz)


def test_quadrature_immutable():
    for name, value in (("npoints", 4), ("qweights", 1), ("qpoints_grid_x", 3),
                        ("qpoints_grid_z", 1), ("qweights_grid_x", 3), ("qweights_grid_z", 1)):
        with pytest.raises(AttributeError):
            setattr(Quadrature, name, value)
    with pytest.raises(ValueError):
        Quadrature.qpoints[0] = 0.0


def test_initialize_fields_shapes_on_the_host():
    p = {"nx": 12, "nz": 7, "hs": 2, "s": 4, "ic_type": "thermal"}
    f = initialize_fields(p, device="cpu")
    assert f.state.shape == f.state_tmp.shape == (4, 11, 16)
    assert f.flux.shape == (4, 8, 13) and f.tend.shape == (4, 7, 12)
    assert f.vals_x.shape == f.d3_vals_x.shape == (4, 7, 13)
    assert f.vals_z.shape == f.d3_vals_z.shape == (4, 8, 12)
    assert f.hy_dens_cell.shape == f.hy_dens_theta_cell.shape == (11,)
    for k in ("hy_dens_int", "hy_dens_theta_int", "hy_pressure_int"):
        assert getattr(f, k).shape == (8,)
    assert f.state.dtype == torch.float64 and f.state.is_contiguous()
    assert not torch.any(f.state)
    with pytest.raises(AssertionError):
        initialize_fields(dict(p, hs=3), device="cpu")
