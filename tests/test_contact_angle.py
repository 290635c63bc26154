"""The contact-angle measurement of taxila-lbm_amd/contact_angle.py on
synthetic spherical caps of known angle (CPU)."""
import numpy as np
import pytest

import contact_angle as ca


@pytest.mark.parametrize("theta", [30.0, 45.0, 90.0, 120.0, 150.0])
def test_measure_recovers_cap_angle(theta):
    shape = ca.SHAPE
    nx, ny, nz = shape
    # the cap that holds the volume of the initial hemisphere (radius 20)
    c = np.cos(np.radians(theta))
    R = (2.0 * 20.0 ** 3 / (2 - 3 * c + c ** 3)) ** (1 / 3)
    zw = 0.5
    zc = zw - R * np.cos(np.radians(theta))
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    dist = np.sqrt((x - nx // 2) ** 2 + (y - ny // 2) ** 2 + (z - zc) ** 2) - R
    a = 0.5 - 0.47 * np.tanh(dist / 2.0)  # diffuse interface about 4 nodes wide
    a[0] = 0.0
    rho = np.stack([a, 1.0 - a], axis=-1).reshape(-1)
    got, parts = ca.measure(rho, shape, zw)
    assert abs(got - theta) < 1.0, (got, parts)
