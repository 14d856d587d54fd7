"""GPU parity: real-time Green's functions (cmpy/exactdiag.py:248-307).

The device path (one Lanczos run from c+psi0 / c psi0, the Gauss quadrature
of its spectral measure on the device) is compared with the oracle's dense
restatement, with the reference's own procedure (scipy's expm_multiply and the
running phase product) and with the two-site closed form.

Tolerance: |g_gpu - g_ref| <= 1e-10 (times the part's weight |phi|^2 <= 1).
"""

import numpy as np
import pytest
from numpy.testing import assert_allclose

from oracle import ed_oracle as O
from cmpy_amd.basis import DN

pytestmark = pytest.mark.gpu

G_ATOL = 1e-10


@pytest.mark.parametrize("u,v", [(2.0, 1.0), (5.0, 0.4)])
def test_gf_tevo_two_site_closed_form(gpu, u, v):
    from cmpy_amd import SingleImpurityAndersonModel, gf_tevo

    m = SingleImpurityAndersonModel(u, 0.0, [0.0], [v], mu=u / 2)
    times, g = gf_tevo(m, 0.0, 30.0, 301)
    assert_allclose(times, np.linspace(0.0, 30.0, 301))
    assert_allclose(g, O.tevo_two_site_ref(times, u, v), atol=G_ATOL)


@pytest.mark.parametrize("sign_mode", ["reference", "fermion"])
@pytest.mark.parametrize("seed", [0, 3])
def test_greater_lesser_match_dense_oracle_siam(gpu, seed, sign_mode):
    from cmpy_amd import SingleImpurityAndersonModel, gf_greater, gf_lesser, groundstate

    rng = np.random.default_rng(seed)
    eps = rng.uniform(-2, 2, 5)
    v = rng.uniform(0.2, 1.0, 5)
    model = SingleImpurityAndersonModel(4.0, -2.0, eps, v, mu=0.0, sign_mode=sign_mode)
    width = 0 if sign_mode == "reference" else 6
    mode = "reference" if sign_mode == "reference" else "jordan_wigner"
    mk = lambda nu, nd: O.siam_operator(nu, nd, 4.0, -2.0, eps, v, 0.0, width=width)
    gs = groundstate(model, 3, 3)
    e0, psi = O.groundstate(mk(3, 3))
    t, gg = gf_greater(model, gs, -5.0, 25.0, 401)
    cd = O.apply_cdag(psi, 6, 3, 3, 0, O.UP, mode)
    assert_allclose(gg, O.tevo_part(mk(4, 3), cd, e0, t, +1), atol=G_ATOL)
    t, gl = gf_lesser(model, gs, -5.0, 25.0, 401, sigma=DN)
    cc = O.apply_c(psi, 6, 3, 3, 0, O.DN, mode)
    assert_allclose(gl, O.tevo_part(mk(3, 2), cc, e0, t, -1), atol=G_ATOL)


def test_greater_matches_expm_multiply_hubbard_ring(gpu):
    """Interior site, dn spin, Jordan-Wigner signs: against the reference's own
    expm_multiply procedure on a 6-site ring."""
    from cmpy_amd import HubbardModel, gf_greater, groundstate

    ns, bonds = 6, O.chain_bonds(6, periodic=True)
    model = HubbardModel(ns, bonds, inter=3.0, eps=0.0, hop=1.0, mu=1.5)
    gs = groundstate(model, 3, 3)
    op = O.hubbard_operator(ns, 3, 3, bonds, 3.0, 0.0, 1.0, mu=1.5)
    e0, psi = O.groundstate(op)
    opp = O.hubbard_operator(ns, 3, 4, bonds, 3.0, 0.0, 1.0, mu=1.5)
    cd = O.apply_cdag(psi, ns, 3, 3, 2, O.DN, "jordan_wigner")
    times, ov = O.tevo_expm_multiply(opp, cd, e0, -2.0, 8.0, 101, +1)
    t, g = gf_greater(model, gs, -2.0, 8.0, 101, pos=2, sigma=DN, sign_mode="jordan_wigner")
    assert_allclose(t, times)
    assert_allclose(g, ov, atol=1e-9)


def test_long_times_mid_sector_vs_expm_multiply(gpu):
    """1+9-site impurity model, sector (5,5) -> (6,5) with D = 52,920: the
    default step count reaches the expm_multiply result out to t = 30."""
    from cmpy_amd import SingleImpurityAndersonModel, gf_greater, groundstate

    rng = np.random.default_rng(11)
    eps = rng.uniform(-3, 3, 9)
    v = rng.uniform(0.1, 1.2, 9)
    model = SingleImpurityAndersonModel(3.0, -1.5, eps, v, mu=0.0)
    gs = groundstate(model, 5, 5)
    mk = lambda nu, nd: O.siam_operator(nu, nd, 3.0, -1.5, eps, v, 0.0)
    cd = O.apply_cdag(gs.state, 10, 5, 5, 0, O.UP)
    times, ov = O.tevo_expm_multiply(mk(6, 5), cd, gs.energy, 0.0, 30.0, 61, +1)
    t, g = gf_greater(model, gs, 0.0, 30.0, 61)
    assert_allclose(g, ov, atol=1e-9)
    # more steps change nothing: the default count has converged; and the
    # reorthogonalised recurrence gives the same quadrature as the default
    # plain one
    _, g2 = gf_greater(model, gs, 0.0, 30.0, 61, nsteps=800)
    assert np.abs(g2 - g).max() <= 1e-11
    _, g3 = gf_greater(model, gs, 0.0, 30.0, 61, reorth=True)
    assert np.abs(g3 - g).max() <= 1e-11


def test_missing_sector_returns_zeros(gpu):
    from cmpy_amd import EigenState, HubbardModel, gf_greater, gf_lesser

    model = HubbardModel(2, [(0, 1)], inter=2.0, eps=0.0, hop=1.0)
    full = EigenState(4.0, np.ones(1), 2, 2)  # no upper sector
    t, g = gf_greater(model, full, 0.0, 1.0, 11)
    assert g.dtype == np.float64 and not g.any() and len(t) == 11
    empty = EigenState(0.0, np.ones(1), 0, 0)  # no lower sector
    t, g = gf_lesser(model, empty, 0.0, 1.0, 11)
    assert g.dtype == np.float64 and not g.any()


def test_fourier_t2z_device_matches_oracle(gpu):
    """The device transform against the oracle's piecewise-linear rule on a
    non-uniform grid of random complex samples (same rule, same interval
    order: agreement to rounding), and the default eta."""
    from cmpy_amd import fourier_t2z

    rng = np.random.default_rng(12)
    t = np.sort(np.concatenate([[0.0, 9.0], rng.uniform(0, 9, 400)]))
    g = rng.standard_normal(t.size) + 1j * rng.standard_normal(t.size)
    w = np.linspace(-8, 8, 257)
    z, gw = fourier_t2z(t, g, w, delta=1e-3)
    z_ref, g_ref = O.fourier_t2z(t, g, w, delta=1e-3)
    assert_allclose(z, z_ref, rtol=0, atol=0)
    assert np.abs(gw - g_ref).max() <= 1e-12 * np.abs(g_ref).max()


def test_gf_tevo_to_frequencies_two_site(gpu):
    """gf_tevo then fourier_t2z, as the reference's script does: the
    two-site transform of -2i sum_k a_k cos(e_k t) over [0, T]."""
    from cmpy_amd import SingleImpurityAndersonModel, fourier_t2z, gf_tevo

    u, v = 2.0, 0.5
    m = SingleImpurityAndersonModel(u, 0.0, [0.0], [v], mu=u / 2)
    times, g = gf_tevo(m, 0.0, 40.0, 8001)
    w = np.linspace(-3, 3, 41)
    z, gw = fourier_t2z(times, g, w, delta=1e-4)
    a1, a2, e1, e2 = O.impurity_params_ref(u, v)
    T = times[-1]
    ref = np.zeros_like(z)
    for a, e in ((a1, e1), (a2, e2)):
        ref -= a * ((np.exp(1j * (z - e) * T) - 1) / (z - e) + (np.exp(1j * (z + e) * T) - 1) / (z + e))
    assert np.abs(gw - ref).max() <= 1e-5
