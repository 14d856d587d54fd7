"""GPU parity on the benchmark configurations themselves, against the oracle.

* 12-site open chain, (6,6), U in {2, 4, 8}: the 200-step continued fraction
  on 2048 real frequencies at Im z = 0.05, fully reorthogonalised, against the
  oracle's reorthogonalised Lanczos + continued fraction started from the same
  c+psi0 / c psi0 (the GPU's ground state).  A 200-step fraction at
  Im z = 0.05 is not converged, and its value is ill-conditioned in the start
  vector: the oracle itself moves by delta when its start vector is perturbed
  by 1e-15 (relative, random).  The tolerance is max(1e-8, 100 delta) (a
  Lanczos run's rounding acts like a start-vector perturbation of up to
  ~1e-13), and 1e-10 at Im z = 1, where the fraction is well conditioned.
  Measured: delta = 7.1e-6 / 1.2e-8 / 6.8e-9 and |G_gpu - G_oracle| =
  1.6e-5 / 1.2e-9 / 1.4e-9 at U = 2 / 4 / 8.
* Batched impurity solves (512 models, n_b in 5..9, beta = 50): the largest
  models (D = 63,504 and 15,876) against the oracle's eigsh E0 (1e-10
  relative) and its reorthogonalised-Lanczos G(i w_n) (1e-8).
* compute_groundstate (minimum over every sector, cmpy/exactdiag.py:24-43)
  on a 1+7-site impurity model, whose sectors reach D = 4,900 and so take the
  Lanczos branch, against a brute-force minimum over the oracle's sectors.
"""

import numpy as np
import pytest

from oracle import ed_oracle as O

pytestmark = pytest.mark.gpu


def _oracle_cf(op_t, phi, e0, z, sign, nsteps=200):
    h = op_t.csr()
    a, b, n2, _ = O.lanczos_normalized(lambda x: h @ x, phi, nsteps, reorth=True)
    return O.cf_greens(a, b, n2, z, e0, sign)


@pytest.mark.parametrize("u", [2.0, 4.0, 8.0])
def test_chain12_cf_at_config_broadening(gpu, u):
    from cmpy_amd import HubbardModel, chain_neighbors, greens_cf, groundstate

    ns = 12
    model = HubbardModel(ns, chain_neighbors(ns), inter=u, hop=1.0, mu=u / 2)
    gs = groundstate(model, 6, 6, tol=1e-10, device=True)
    z = np.linspace(-10, 10, 2048) + 0.05j
    z1 = np.linspace(-10, 10, 2048) + 1.0j
    r = greens_cf(model, z, 6, 6, nsteps=200, gs=gs, reorth="full")
    r1 = greens_cf(model, z1, 6, 6, nsteps=200, gs=gs, reorth="full")
    assert (r.reorth_p, r.reorth_h) == ("full", "full")
    psi = gs.state.cpu().numpy()
    e0 = gs.energy
    bonds = O.chain_bonds(ns)
    mk = lambda nu, nd: O.hubbard_operator(ns, nu, nd, bonds, u, 0.0, 1.0, mu=u / 2)
    op_p, op_m = mk(7, 6), mk(5, 6)
    rng = np.random.default_rng(11)
    g_ref = np.zeros_like(z)
    g_pert = np.zeros_like(z)
    g1_ref = np.zeros_like(z1)
    for op_t, lad, sign in ((op_p, O.apply_cdag, +1), (op_m, O.apply_c, -1)):
        phi = lad(psi, ns, 6, 6, 0, O.UP)
        h = op_t.csr()
        a, b, n2, _ = O.lanczos_normalized(lambda x: h @ x, phi, 200, reorth=True)
        g_ref += O.cf_greens(a, b, n2, z, e0, sign)
        g1_ref += O.cf_greens(a, b, n2, z1, e0, sign)
        phi2 = phi * (1.0 + 1e-15 * rng.standard_normal(phi.size))
        a2, b2, n22, _ = O.lanczos_normalized(lambda x: h @ x, phi2, 200, reorth=True)
        g_pert += O.cf_greens(a2, b2, n22, z, e0, sign)
    delta = float(np.abs(g_pert - g_ref).max())
    err = float(np.abs(r.gf - g_ref).max())
    err1 = float(np.abs(r1.gf - g1_ref).max())
    print(f"U={u}: |G_gpu - G_oracle| = {err:.2e} at eta=0.05 (1e-15 start perturbation moves the oracle "
          f"by {delta:.2e}); {err1:.2e} at eta=1")
    assert err1 <= 1e-10
    assert err <= max(1e-8, 100.0 * delta), (err, delta)


def _batch_large_models():
    from cmpy_amd.batch import random_batch

    spec = random_batch(512, 2024)
    nb = np.array([len(e) for e in spec["eps_bath"]])
    sel = list(np.nonzero(nb == 9)[0][:2]) + list(np.nonzero(nb == 8)[0][:2])
    return spec, sel


def test_batch_largest_models_against_oracle(gpu):
    import scipy.sparse.linalg as sla

    from cmpy_amd.batch import AndersonBatch, matsubara

    spec, sel = _batch_large_models()
    wn = matsubara(50.0, 256)
    res = AndersonBatch(**spec).solve(wn)
    z = 1j * wn
    for i in sel:
        u, eimp = spec["u"][i], spec["eps_imp"][i]
        eb, v = spec["eps_bath"][i], spec["v"][i]
        nb = len(eb)
        ns, n = nb + 1, (nb + 1) // 2
        mk = lambda nu, nd: O.siam_operator(nu, nd, u, eimp, eb, v, 0.0)
        op = mk(n, n)
        assert op.size in (63504, 15876)
        w, vec = sla.eigsh(op.csr(), k=1, which="SA", tol=1e-14)
        e0, psi = float(w[0]), vec[:, 0]
        assert abs(res.e0[i] - e0) <= 1e-10 * abs(e0), (i, res.e0[i], e0)
        g = _oracle_cf(mk(n + 1, n), O.apply_cdag(psi, ns, n, n, 0, O.UP), e0, z, +1, 400)
        g += _oracle_cf(mk(n - 1, n), O.apply_c(psi, ns, n, n, 0, O.UP), e0, z, -1, 400)
        err = np.abs(res.g[i] - g).max()
        print(f"model {i}: D = {op.size}, |dE0| = {abs(res.e0[i] - e0):.1e}, |dG| = {err:.1e}")
        assert err <= 1e-8, (i, err)


def test_compute_groundstate_lanczos_branch(gpu):
    from cmpy_amd import SingleImpurityAndersonModel, compute_groundstate

    rng = np.random.default_rng(8)
    eb, v = rng.uniform(-2, 2, 7), rng.uniform(0.2, 1.0, 7)
    u, eimp, mu = 3.0, -1.0, 0.4
    model = SingleImpurityAndersonModel(u, eimp, eb, v, mu=mu)
    gs = compute_groundstate(model)
    energies = {}
    for nu in range(9):
        for nd in range(9):
            energies[(nu, nd)] = O.groundstate(O.siam_operator(nu, nd, u, eimp, eb, v, mu), tol=1e-14)[0]
    e_min = min(energies.values())
    assert max(O.binom(8, gs.n_up) * O.binom(8, gs.n_dn), 0) > 50  # the Lanczos branch chose it
    assert abs(gs.energy - e_min) <= 1e-10 * abs(e_min)
    assert abs(energies[(gs.n_up, gs.n_dn)] - e_min) <= 1e-10 * abs(e_min)
    op = O.siam_operator(gs.n_up, gs.n_dn, u, eimp, eb, v, mu)
    psi = np.asarray(gs.state)
    assert np.linalg.norm(op.matvec(psi) - gs.energy * psi) <= 1e-9
