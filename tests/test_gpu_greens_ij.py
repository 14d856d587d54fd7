"""GPU parity: off-diagonal Green's functions G_ij (lehmann_full.py:19-97).

``greens_lehmann`` (finite-T, all sectors) against the oracle's dense
``lehmann_ij``; ``greens_cf_ij`` (T = 0, continued fractions by polarisation)
against ``lehmann_t0_ij``; both against the U = 0 single-particle resolvent.
Tolerance: G within 1e-8 (absolute).
"""

import numpy as np
import pytest
from numpy.testing import assert_allclose

from oracle import ed_oracle as O

pytestmark = pytest.mark.gpu

G_ATOL = 1e-8


def _tb(ns, bonds, hop, eps):
    h = np.diag(np.full(ns, float(eps)))
    for i, j in bonds:
        h[i, j] += hop
        h[j, i] += hop
    return h


@pytest.mark.parametrize("ij", [(1, 3), (0, 2), (4, 4)])
def test_greens_lehmann_matches_oracle_ring(gpu, ij):
    from cmpy_amd import HubbardModel, greens_lehmann

    ns, u, beta = 5, 2.5, 4.0
    bonds = O.chain_bonds(ns, periodic=True)
    model = HubbardModel(ns, bonds, inter=u, eps=0.0, hop=1.0, mu=u / 2)
    z = np.linspace(-6, 6, 301) + 0.1j
    r = greens_lehmann(model, z, beta, *ij)
    mk = lambda nu, nd: O.hubbard_operator(ns, nu, nd, bonds, u, 0.0, 1.0, mu=u / 2)
    assert_allclose(r.gf, O.lehmann_ij(mk, ns, z, beta, *ij), atol=G_ATOL)


@pytest.mark.parametrize("ij", [(1, 4), (2, 3)])
def test_greens_lehmann_non_interacting_resolvent(gpu, ij):
    from cmpy_amd import HubbardModel, greens_lehmann

    ns = 6
    bonds = O.chain_bonds(ns)
    model = HubbardModel(ns, bonds, inter=0.0, eps=0.2, hop=1.0, mu=0.5)
    z = np.linspace(-4, 4, 201) + 0.05j
    r = greens_lehmann(model, z, 7.0, *ij)
    ref = O.resolvent_ij(_tb(ns, bonds, 1.0, 0.2 - 0.5), z)[:, ij[0], ij[1]]
    assert_allclose(r.gf, ref, atol=G_ATOL)


def test_greens_lehmann_impurity_reference_mode(gpu):
    """Impurity model in the reference sign convention: the string-free
    ladder operators of the reference (its hopping has no signs either)."""
    from cmpy_amd import SingleImpurityAndersonModel, greens_lehmann

    eps, v = [-1.0, 0.5, 1.5], [0.6, 0.4, 0.8]
    model = SingleImpurityAndersonModel(3.0, -1.5, eps, v, mu=0.0)
    z = np.linspace(-6, 6, 201) + 0.1j
    r = greens_lehmann(model, z, 5.0, 0, 2)
    mk = lambda nu, nd: O.siam_operator(nu, nd, 3.0, -1.5, eps, v, 0.0)
    assert_allclose(r.gf, O.lehmann_ij(mk, 4, z, 5.0, 0, 2, sign_mode="reference"), atol=G_ATOL)
    with pytest.raises(ValueError):
        greens_lehmann(model, z, np.inf, 0, 2)


@pytest.mark.parametrize("ij", [(1, 4), (0, 3), (2, 2)])
@pytest.mark.parametrize("u", [0.0, 3.0])
def test_greens_cf_ij_matches_t0_lehmann(gpu, ij, u):
    from cmpy_amd import HubbardModel, greens_cf_ij, groundstate

    ns = 6
    bonds = O.chain_bonds(ns, periodic=True)
    model = HubbardModel(ns, bonds, inter=u, eps=0.0, hop=1.0, mu=u / 2)
    z = np.linspace(-5, 5, 401) + 0.1j
    gs = groundstate(model, 3, 3)
    g = greens_cf_ij(model, z, 3, 3, *ij, nsteps=400, gs=gs)
    mk = lambda nu, nd: O.hubbard_operator(ns, nu, nd, bonds, u, 0.0, 1.0, mu=u / 2)
    e0, psi = O.groundstate(mk(3, 3))
    lad = lambda f, p: f(psi, ns, 3, 3, p, O.UP, "jordan_wigner")
    ref = O.lehmann_t0_ij(mk(4, 3), mk(2, 3), e0, lad(O.apply_cdag, ij[0]), lad(O.apply_cdag, ij[1]),
                          lad(O.apply_c, ij[0]), lad(O.apply_c, ij[1]), z)
    assert_allclose(g, ref, atol=G_ATOL)
    if u == 0.0:
        assert_allclose(g, O.resolvent_ij(_tb(ns, bonds, 1.0, 0.0), z)[:, ij[0], ij[1]], atol=G_ATOL)
