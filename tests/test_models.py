"""Host logic: the synthetic workloads have the reference's index space and
the values their descriptions state."""
import numpy as np

from fwant_amd import models


def test_two_layer_model():
    m = models.two_layer_100x100x50()
    assert (m.nx, m.ny, m.nz) == (106, 106, 56)
    assert m.lam.shape == (106, 106, 56) and m.lam.flags["F_CONTIGUOUS"]
    # free surface at r = 6371 km, 1 km spacing
    assert m.z[m.nk2] == models.R_EARTH and np.allclose(np.diff(m.z), 1000.0)
    vs = np.sqrt(m.mu / m.rho)
    vp = np.sqrt((m.lam + 2 * m.mu) / m.rho)
    depth = models.R_EARTH - m.z
    crust = depth < 30000 - 1e-6
    np.testing.assert_allclose(vs[50, 50, crust], 3460.0, rtol=1e-6)
    np.testing.assert_allclose(vp[50, 50, ~crust], 8000.0, rtol=1e-6)
    np.testing.assert_allclose(m.rho[50, 50, crust], 2700.0)
    assert m.abs_number == (12, 12, 12, 12, 12, 0)
    # vertical force at the surface, 16 surface receivers inside the PML-free zone
    assert m.force.n == 1 and m.force.indx[0, 2] == m.nk2 + 1
    assert len(m.rec_indx) == 16 and np.all(m.rec_indx[:, 2] == m.nk2 + 1)
    assert np.all(m.rec_indx[:, :2] > 3 + 12 + 1) and np.all(m.rec_indx[:, :2] < 3 + 100 - 12)


def test_spherical_grid_equatorial_patch():
    x, y, z = models.spherical_grid(100, 100, 50, 1000.0)
    d = 1000.0 / models.R_EARTH
    assert abs(x[3 + 49] - (np.pi / 2 - 0.5 * d)) < 1e-12
    np.testing.assert_allclose(np.diff(x), d)
    np.testing.assert_allclose(np.diff(y), d)


def test_gradient_model_small():
    m = models.gradient_400x400x200(ni=40, nj=36, nk=20, h=500.0)
    vs = np.sqrt(m.mu / m.rho)
    vp = np.sqrt((m.lam + 2 * m.mu) / m.rho)
    np.testing.assert_allclose(vp / vs, 1.73, rtol=1e-5)
    top = vs[3:-3, 3:-3, m.nk2]
    assert 3200 * 0.949 < top.min() and top.max() < 3200 * 1.051
    np.testing.assert_allclose(m.rho, 1741.0 * (vp / 1000.0) ** 0.25, rtol=1e-5)
