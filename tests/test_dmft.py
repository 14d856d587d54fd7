"""DMFT caller (cmpy/dmft/utils.py, cmpy/dmft/twosite.py): CPU checks of the
self-consistency arithmetic against closed forms; the GPU test runs the loop
with the device impurity solvers and compares it with the closed-form loop."""

import numpy as np
import pytest
from numpy.testing import assert_allclose

from cmpy_amd import dmft

Z_AXIS = np.linspace(-10, 10, 20001) + 1e-3j


@pytest.mark.parametrize("u", [1.0, 2.5, 4.0])
def test_twosite_closed_form_loop_gives_potthoff_weight(u):
    """Two-site DMFT on the Bethe lattice: V^2 = z_qp = 1 - (U / 6)^2 for U < 6
    (M. Potthoff, cond-mat/0107502)."""
    s = dmft.twosite_dmft_half_filling(Z_AXIS, u, verbose=False, ref=True, vtol=1e-10)
    assert s.stats.success
    assert abs(float(s.v[0]) ** 2 - (1 - u**2 / 36)) < 5e-6


def test_twosite_insulator_stops_at_zero_weight():
    s = dmft.twosite_dmft_half_filling(Z_AXIS, 7.0, verbose=False, ref=True, max_iter=2000)
    assert float(s.v[0]) < 1e-3 or s.stats.status.startswith("Quasiparticle")


def test_bethe_gf_closed_form_and_sum_rule():
    t = 0.7
    w = np.linspace(-1.3, 1.3, 7)
    g = dmft.bethe_gf_omega(w + 1e-12j, t)
    assert_allclose(g, (w - 1j * np.sqrt(4 * t**2 - w**2)) / (2 * t**2), atol=1e-9)
    x = np.linspace(-2 * t, 2 * t, 400001)
    dos = -dmft.bethe_gf_omega(x + 1e-14j, t).imag / np.pi
    assert abs(np.trapezoid(dos, x) - 1.0) < 1e-6
    zb = np.array([50.0 + 3j, -80j])
    assert_allclose(dmft.bethe_gf_omega(zb, t) * zb, 1.0, atol=1e-3)


def test_quasiparticle_weight_of_linear_self_energy():
    w = np.linspace(-2, 2, 401)
    zq = 0.37
    sigma = (1 - 1 / zq) * w + 0.1j
    assert abs(dmft.quasiparticle_weight(w, sigma) - zq) < 1e-12
    assert dmft.quasiparticle_weight(w, 1e9 * w) == 0


def test_helpers():
    assert dmft.mix_values(1.0, 3.0) == 3.0
    assert dmft.mix_values(1.0, 3.0, 0.25) == 1.5
    with pytest.raises(ValueError):
        dmft.mix_values(1.0, 3.0, 1.5)
    assert_allclose(dmft.self_energy(0.5, 0.25), 2.0 - 4.0)
    st = dmft.IterationStats("a", "b")
    st.append(1e-3, 2e-4)
    st.set_maxiter_status(10)
    assert not st and st.iteration == 1 and "a=1.00e-03" in st.errorstr()
    st.set_parameter_converged("V", 0.5)
    assert st and "converged" in st.message


@pytest.mark.gpu
@pytest.mark.parametrize("u", [2.0, 4.0])
def test_twosite_loop_with_device_solvers(gpu, u):
    """ref=False solves the impurity on the device every iteration (T = 0:
    ground-sector Lehmann sum; beta = 400: the grand-canonical sum); both
    reach the closed-form loop's hybridisation."""
    z = np.linspace(-8, 8, 4001) + 1e-3j
    s_ref = dmft.twosite_dmft_half_filling(z, u, verbose=False, ref=True, vtol=1e-10)
    s_dev = dmft.twosite_dmft_half_filling(z, u, verbose=False, ref=False, vtol=1e-10)
    assert abs(float(s_dev.v[0]) - float(s_ref.v[0])) < 1e-8
    s_t = dmft.twosite_dmft_half_filling(z, u, beta=400.0, verbose=False, ref=False, vtol=1e-10)
    assert abs(float(s_t.v[0]) - float(s_ref.v[0])) < 1e-8
    g_ref = dmft.compute_lattice_greens_function(z, s_ref, 1.0, ref=True)
    g_dev = dmft.compute_lattice_greens_function(z, s_ref, 1.0, ref=False)
    assert np.abs(g_dev - g_ref).max() < 1e-8
