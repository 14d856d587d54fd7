"""GPU parity: the observable and accumulation functions of cmpy/exactdiag.py:60-134
with the reference's signatures (state lists, eigenvalues, eigenvectors as
columns), against the oracle's occupation / double_occupation / _accumulate_sum
(pinned on the CPU by the atomic limit, the U = 0 Fermi sums and Wick's
factorisation, and through gf_lehmann).

Tolerances: occupations within 1e-12 relative, G within 1e-10.
"""

import numpy as np
import pytest
import scipy.sparse.linalg as sla
from numpy.testing import assert_allclose

from oracle import ed_oracle as O

pytestmark = pytest.mark.gpu


def _sector(ns, nu, nd, u=2.5, seed=0):
    rng = np.random.default_rng(seed)
    eps = rng.uniform(-1, 1, ns)
    bonds = O.chain_bonds(ns, periodic=True)
    op = O.SectorOperator(ns, nu, nd, bonds, [1.0] * len(bonds), eps, np.full(ns, u), ns)
    w, v = np.linalg.eigh(op.csr().toarray())
    return op, w, v


@pytest.mark.parametrize("sec", [(5, 2, 3), (6, 3, 3), (6, 4, 1)])
@pytest.mark.parametrize("beta", [0.5, 3.0])
def test_occupations_match_oracle(gpu, sec, beta):
    from cmpy_amd import double_occupation, occupation, occupation_dn, occupation_up

    ns, nu, nd = sec
    op, w, v = _sector(ns, nu, nd)
    up, dn = op.up.astype(np.int64), op.dn.astype(np.int64)
    emin = float(w[0]) - 0.3
    for pos in range(ns):
        ref_u = O.occupation(up, dn, w, v, beta, emin, pos, O.UP)
        ref_d = O.occupation(up, dn, w, v, beta, emin, pos, O.DN)
        ref_2 = O.double_occupation(up, dn, w, v, beta, emin, pos)
        assert occupation_up(up, dn, w, v, beta, emin, pos) == pytest.approx(ref_u, rel=1e-12, abs=1e-14)
        assert occupation_dn(up, dn, w, v, beta, emin, pos) == pytest.approx(ref_d, rel=1e-12, abs=1e-14)
        assert occupation(up, dn, w, v, beta, emin, pos, sigma=2) == pytest.approx(ref_d, rel=1e-12, abs=1e-14)
        assert double_occupation(up, dn, w, v, beta, emin, pos) == pytest.approx(ref_2, rel=1e-12, abs=1e-14)
    # a subset of eigenvectors, and a single vector
    k = 5
    assert occupation_up(up, dn, w[:k], v[:, :k], beta, emin, 1) == pytest.approx(
        O.occupation(up, dn, w[:k], v[:, :k], beta, emin, 1, O.UP), rel=1e-12)
    assert double_occupation(up, dn, w[:1], v[:, 0], beta, emin, 0) == pytest.approx(
        O.double_occupation(up, dn, w[:1], v[:, :1], beta, emin, 0), rel=1e-12)


def test_occupations_unsorted_state_lists(gpu):
    """The functions take any state lists (the reference's arguments), not
    only the sector's sorted ones: a permuted basis with permuted vectors."""
    from cmpy_amd import double_occupation, occupation_up

    op, w, v = _sector(5, 2, 2)
    rng = np.random.default_rng(3)
    pu, pd = rng.permutation(op.num_up), rng.permutation(op.num_dn)
    up, dn = op.up[pu], op.dn[pd]
    idx = (pu[:, None] * op.num_dn + pd[None, :]).ravel()
    vp = v[idx]
    for pos in range(5):
        assert occupation_up(up, dn, w, vp, 1.3, w[0], pos) == pytest.approx(
            O.occupation(op.up, op.dn, w, v, 1.3, w[0], pos, O.UP), rel=1e-12)
        assert double_occupation(up, dn, w, vp, 1.3, w[0], pos) == pytest.approx(
            O.double_occupation(op.up, op.dn, w, v, 1.3, w[0], pos), rel=1e-12)


def test_occupation_rejects_bad_shapes(gpu):
    from cmpy_amd import occupation_up

    op, w, v = _sector(5, 2, 2)
    with pytest.raises(ValueError):
        occupation_up(op.up, op.dn, w, v[:-1], 1.0, 0.0, 0)
    with pytest.raises(ValueError):
        occupation_up(op.up, op.dn, w[:-1], v, 1.0, 0.0, 0)
    with pytest.raises(ValueError):
        occupation_up(op.up, op.dn, w, v, 1.0, 0.0, 40)


@pytest.mark.parametrize("beta", [2.0, np.inf])
@pytest.mark.parametrize("generic", [False, True])
def test_accumulate_gf_matches_oracle(gpu, beta, generic):
    """accumulate_gf adds one sector pair's Lehmann sum into gf in place, with
    our CreationOperator (device ladder) or any operator with matmat."""
    from cmpy_amd import Basis, CreationOperator, accumulate_gf

    ns, nu, nd = 5, 2, 2
    op, w, v = _sector(ns, nu, nd)
    op1, w1, v1 = _sector(ns, nu + 1, nd)
    z = np.linspace(-5, 5, 201) + 0.1j
    basis = Basis(ns)
    sec, sec1 = basis.get_sector(nu, nd), basis.get_sector(nu + 1, nd)
    cdag = CreationOperator(sec, sec1, pos=1, sigma=1)
    if generic:
        dense = np.column_stack([O.apply_cdag(e, ns, nu, nd, 1, O.UP) for e in np.eye(op.size)])
        cdag = sla.aslinearoperator(dense)
    emin = float(w[0]) if np.isfinite(beta) else 0.0
    gf = np.full(z.shape, 0.25 + 0.5j)
    ret = accumulate_gf(gf, z, cdag, w, v, w1, v1, beta, emin)
    assert ret is None
    ref = np.full(z.shape, 0.25 + 0.5j)
    ce = np.column_stack([O.apply_cdag(v[:, k], ns, nu, nd, 1, O.UP) for k in range(v.shape[1])])
    O._accumulate_sum(ref, z, w, w1, v1, ce, beta, emin)
    assert_allclose(gf, ref, atol=1e-10 * np.abs(ref).max())


def test_accumulate_gf_builds_gf_lehmann(gpu):
    """Summing accumulate_gf and the occupations over every sector with the
    reference's visiting order and Z reproduces the oracle's gf_lehmann."""
    from cmpy_amd import Basis, CreationOperator, accumulate_gf, double_occupation, occupation

    ns, u, beta = 4, 3.0, 2.0
    basis = Basis(ns)
    mk = lambda nu, nd: O.hubbard_operator(ns, nu, nd, O.chain_bonds(ns), u, 0.0, 1.0, mu=u / 2)
    z = np.linspace(-6, 6, 121) + 0.2j
    eig = {}
    for nu in range(ns + 1):
        for nd in range(ns + 1):
            eig[(nu, nd)] = np.linalg.eigh(mk(nu, nd).csr().toarray())
    emin = min(w[0] for w, _ in eig.values())
    gf = np.zeros_like(z)
    part = occ = dbl = 0.0
    for nu in range(ns + 1):
        for nd in range(ns + 1):
            if nu == ns:
                continue
            w, v = eig[(nu, nd)]
            w1, v1 = eig[(nu + 1, nd)]
            sec = basis.get_sector(nu, nd)
            cdag = CreationOperator(sec, basis.get_sector(nu + 1, nd), pos=0, sigma=1)
            accumulate_gf(gf, z, cdag, w, v, w1, v1, beta, emin)
            part += np.exp(-beta * (w - emin)).sum()
            occ += occupation(sec.up_states, sec.dn_states, w, v, beta, emin, 0)
            dbl += double_occupation(sec.up_states, sec.dn_states, w, v, beta, emin, 0)
    ref = O.gf_lehmann(mk, ns, z, beta, pos=0, occ=True)
    assert_allclose(gf / part, ref.gf, atol=1e-10)
    assert occ / part == pytest.approx(ref.occ, abs=1e-12)
    assert dbl / part == pytest.approx(ref.occ_double, abs=1e-12)
