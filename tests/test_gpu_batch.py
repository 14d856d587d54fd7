"""Batched Anderson impurity solves against the per-model path and the oracle.

Tolerances (the ground-state and Green's-function bars of the hot path):
E0 within 1e-10 relative, G(i w_n) within 1e-8 of max |G| per model.
Small models are checked against the oracle's dense T = 0 Lehmann sum (the
reference's _accumulate_sum restricted to the ground state,
cmpy/exactdiag.py:110-129); larger ones against the single-model GPU path
(ground state + continued fractions), which is itself pinned to the oracle.
"""
import numpy as np
import pytest

from oracle import ed_oracle as O

pytestmark = pytest.mark.gpu

E0_RTOL = 1e-10
G_TOL = 1e-8


def _batch(**kw):
    from cmpy_amd.batch import AndersonBatch

    return AndersonBatch(**kw)


def _oracle_model(u, eps_imp, eb, v, mu, n_up, n_dn, wn, width=0):
    nb = len(eb)
    ns = nb + 1
    op = O.siam_operator(n_up, n_dn, u, eps_imp, eb, v, mu, width)
    e0, psi = O.groundstate(op, tol=1e-13)
    mode = "reference" if width == 0 else "jordan_wigner"
    opp = O.siam_operator(n_up + 1, n_dn, u, eps_imp, eb, v, mu, width) if n_up + 1 <= ns else None
    opm = O.siam_operator(n_up - 1, n_dn, u, eps_imp, eb, v, mu, width) if n_up >= 1 else None
    cd = O.apply_cdag(psi, ns, n_up, n_dn, 0, O.UP, mode) if opp is not None else None
    c = O.apply_c(psi, ns, n_up, n_dn, 0, O.UP, mode) if opm is not None else None
    g = O.lehmann_t0(op, opp, opm, psi, e0, cd, c, 1j * wn)
    return e0, g


def _params(seed, m, nb_range):
    from cmpy_amd.batch import random_batch

    return random_batch(m, seed, nb_range=nb_range)


def _check(res, e0_ref, g_ref, i):
    assert abs(res.e0[i] - e0_ref) <= E0_RTOL * abs(e0_ref), (i, res.e0[i], e0_ref)
    scale = np.abs(g_ref).max()
    assert np.abs(res.g[i] - g_ref).max() <= G_TOL * scale, (i, np.abs(res.g[i] - g_ref).max())


@pytest.mark.parametrize("shape", ["auto", "block", "wave"])
def test_batch_matches_oracle_small(gpu, shape):
    """Twenty models with 2..6 bath sites (D up to 1225) against the oracle."""
    from cmpy_amd.batch import matsubara

    p = _params(11, 20, (2, 6))
    b = _batch(**p)
    if shape == "block":
        b.config(-1, 8, 0)
    elif shape == "wave":
        b.config(1 << 40, 2, 2)
    wn = matsubara(50.0, 64)
    res = b.solve(wn)
    assert np.all(res.ritz_residual <= 1e-10) and np.all(res.residual <= 1e-9), res.residual
    for i in range(b.num_models):
        e0, g = _oracle_model(p["u"][i], p["eps_imp"][i], p["eps_bath"][i], p["v"][i], p["mu"][i],
                              int(b.n_up[i]), int(b.n_dn[i]), wn)
        _check(res, e0, g, i)


def test_batch_matches_single_model_path(gpu):
    """The batch sizes of the headline workload (7..9 bath sites, D up to
    63,504) against greens_cf of SingleImpurityAndersonModel, one by one."""
    from cmpy_amd.batch import matsubara
    from cmpy_amd.exactdiag import greens_cf, groundstate
    from cmpy_amd.models import SingleImpurityAndersonModel

    p = _params(5, 6, (7, 9))
    b = _batch(**p)
    wn = matsubara(50.0, 256)
    res = b.solve(wn)
    for i in range(b.num_models):
        m = SingleImpurityAndersonModel(u=p["u"][i], eps_imp=p["eps_imp"][i], eps_bath=p["eps_bath"][i],
                                        v=p["v"][i], mu=p["mu"][i])
        nu, nd = int(b.n_up[i]), int(b.n_dn[i])
        gs = groundstate(m, nu, nd, tol=1e-12)
        cf = greens_cf(m, 1j * wn, nu, nd, nsteps=120, gs=gs)
        _check(res, gs.energy, cf.gf, i)


def test_batch_replay_is_deterministic(gpu):
    """The captured graph is replayed: every replay gives bit-identical output."""
    from cmpy_amd.batch import matsubara

    b = _batch(**_params(3, 40, (3, 7)))
    wn = matsubara(50.0, 32)
    r1 = b.solve(wn)
    r2 = b.solve(wn)
    r3 = b.solve(wn)
    for r in (r2, r3):
        np.testing.assert_array_equal(r.e0, r1.e0)
        np.testing.assert_array_equal(r.g, r1.g)
        np.testing.assert_array_equal(r.iterations, r1.iterations)


def test_batch_empty_particle_or_hole_sector(gpu):
    """n_up = 0 has no hole part and n_up = num_sites no particle part."""
    from cmpy_amd.batch import matsubara

    rng = np.random.default_rng(2)
    eb = [rng.uniform(-2, 2, 3), rng.uniform(-2, 2, 4)]
    v = [rng.uniform(0.2, 1.0, 3), rng.uniform(0.2, 1.0, 4)]
    u = np.array([3.0, 2.0])
    n_up = np.array([0, 5])
    n_dn = np.array([2, 2])
    b = _batch(u=u, eps_imp=-u / 2, eps_bath=eb, v=v, mu=0.0, n_up=n_up, n_dn=n_dn)
    wn = matsubara(20.0, 16)
    res = b.solve(wn)
    for i in range(2):
        e0, g = _oracle_model(u[i], -u[i] / 2, eb[i], v[i], 0.0, int(n_up[i]), int(n_dn[i]), wn)
        _check(res, e0, g, i)


def test_batch_fermion_signs(gpu):
    """sign_mode="fermion" against the oracle with the Jordan-Wigner width."""
    from cmpy_amd.batch import matsubara

    p = _params(9, 6, (3, 6))
    b = _batch(sign_mode="fermion", **p)
    wn = matsubara(30.0, 24)
    res = b.solve(wn)
    for i in range(b.num_models):
        nb = len(p["eps_bath"][i])
        e0, g = _oracle_model(p["u"][i], p["eps_imp"][i], p["eps_bath"][i], p["v"][i], p["mu"][i],
                              int(b.n_up[i]), int(b.n_dn[i]), wn, width=nb + 1)
        _check(res, e0, g, i)


def test_block_mode_staged_sweep_matches_plain_sweep(gpu, monkeypatch):
    """Block mode with the row-staged H.v sweep (LDS-cached hop lists) against
    the plain element-wise sweep, at the batch's largest size (9 bath sites,
    D = 63,504) and a small one, both sign modes; the oracle pins the small
    ones."""
    from cmpy_amd.batch import matsubara

    p = _params(21, 6, (4, 9))
    wn = matsubara(50.0, 64)
    out = {}
    for staged in ("1", "0"):
        monkeypatch.setenv("CMPY_AMD_BATCH_STAGED", staged)
        for mode in ("reference", "fermion"):
            b = _batch(sign_mode=mode, **p)
            b.config(-1, 16, 0)
            out[staged, mode] = b.solve(wn)
    for mode in ("reference", "fermion"):
        a, c = out["1", mode], out["0", mode]
        assert np.all(a.residual <= 1e-9)
        np.testing.assert_allclose(a.e0, c.e0, rtol=E0_RTOL, atol=0)
        for i in range(len(a.e0)):
            assert np.abs(a.g[i] - c.g[i]).max() <= G_TOL * np.abs(c.g[i]).max()
    b = _batch(**p)
    for i in range(b.num_models):
        if len(p["eps_bath"][i]) <= 5:
            e0, g = _oracle_model(p["u"][i], p["eps_imp"][i], p["eps_bath"][i], p["v"][i], p["mu"][i],
                                  int(b.n_up[i]), int(b.n_dn[i]), wn)
            _check(out["1", "reference"], e0, g, i)


def test_batch_invalid_arguments(gpu):
    from cmpy_amd import _lib as L

    with pytest.raises(L.EDInvalidArgument, match="num_bath"):
        _batch(u=1.0, eps_imp=0.0, eps_bath=[np.zeros(16)], v=[np.ones(16)])
    with pytest.raises(L.EDInvalidArgument, match="finite"):
        _batch(u=float("nan"), eps_imp=0.0, eps_bath=[np.zeros(3)], v=[np.ones(3)])
    with pytest.raises(ValueError):
        _batch(u=1.0, eps_imp=0.0, eps_bath=[np.zeros(3)], v=[np.ones(2)])
