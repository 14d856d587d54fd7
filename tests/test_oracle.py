"""Pins the CPU oracle (oracle/ed_oracle.py) before anything is checked against it.

Each check is independent of the oracle's own formulas: reference-held fixtures,
closed forms, brute force and invariants chosen so that a dropped term, a wrong
sign/index or a transposed operand fails at least one of them.
"""

import itertools
import math

import numpy as np
import pytest
from numpy.testing import assert_allclose, assert_array_equal

from oracle import ed_oracle as O


# -- basis ------------------------------------------------------------------------


def test_state_tables(ref_fixtures):
    for case in ref_fixtures["basis_get_states"]["cases"]:
        ns, n = case["num_sites"], case["n"]
        got = O.generate_states_permutations(ns, n)
        assert [format(int(s), f"0{ns}b") for s in got] == case["states"]
        got2 = O.sector_states(ns, n)
        assert [format(int(s), f"0{ns}b") for s in got2] == case["states"]


@pytest.mark.parametrize("ns", range(1, 9))
def test_sector_states_match_permutation_enumeration(ns):
    for n in range(ns + 1):
        assert_array_equal(np.asarray(O.generate_states_permutations(ns, n)), O.sector_states(ns, n))


@pytest.mark.parametrize("ns", range(1, 13))
def test_combinadic_rank_is_position(ns):
    for n in range(ns + 1):
        st = O.sector_states(ns, n)
        assert len(st) == math.comb(ns, n)
        assert_array_equal(O.combinadic_rank(st), np.arange(len(st)))


def test_upper_lower_sector_tables():
    # cmpy/tests/test_basis.py:91-111
    ns = 15
    for n_up, n_dn in itertools.product(range(15), range(15)):
        assert O.upper_sector(n_up, n_dn, O.UP, ns) == (None if n_up == ns else (n_up + 1, n_dn))
        assert O.upper_sector(n_up, n_dn, O.DN, ns) == (None if n_dn == ns else (n_up, n_dn + 1))
        assert O.lower_sector(n_up, n_dn, O.UP) == (None if n_up == 0 else (n_up - 1, n_dn))
        assert O.lower_sector(n_up, n_dn, O.DN) == (None if n_dn == 0 else (n_up, n_dn - 1))


def test_state_labels(ref_fixtures):
    for case in ref_fixtures["state_labels"]["cases"]:
        ns = case["num_sites"]
        up = O.sector_states(ns, case["n_up"])
        dn = O.sector_states(ns, case["n_dn"])
        labels = [O.state_label(int(u), int(d), ns) for u in up for d in dn]
        assert labels == case["labels"]


# -- signs and projectors ---------------------------------------------------------


def test_hopping_sign_hand_cases():
    # (state, width, site1, site2) -> sign, counted by hand
    cases = [
        (0b0000, 4, 0, 3, 1), (0b0110, 4, 0, 3, 1), (0b0010, 4, 0, 3, -1), (0b0100, 4, 0, 3, -1),
        (0b1111, 4, 0, 1, 1), (0b0111, 4, 0, 2, -1), (0b0111, 0, 0, 2, 1), (0b0111, 1, 0, 2, 1),
        (0b0111, 2, 0, 2, -1), (0b11110, 5, 0, 4, -1), (0b01010, 5, 0, 4, 1),
        (0b10110, 5, 1, 4, -1), (0b10110, 5, 0, 3, 1), (0b10110, 3, 0, 4, 1),
        (0b11110, 3, 0, 4, 1), (0b11110, 4, 0, 4, -1),
    ]
    for s, w, i, j, sign in cases:
        assert O.hopping_sign(s, w, i, j) == sign, (bin(s), w, i, j)


def test_hopping_terms_vectorised_matches_loop():
    # the vectorised hopping_terms against a direct loop over cmpy's per-state rule
    for ns, n, width in [(5, 2, 5), (6, 3, 6), (6, 3, 0), (7, 3, 4)]:
        st = O.sector_states(ns, n)
        for i, j in itertools.combinations(range(ns), 2):
            o, t, a = O.hopping_terms(st, width, i, j, 0.7)
            ref = []
            for k, s in enumerate(st):
                s = int(s)
                if bool(s >> i & 1) != bool(s >> j & 1):
                    new = s ^ (1 << i) ^ (1 << j)
                    ref.append((k, int(np.searchsorted(st, new)), O.hopping_sign(s, width, i, j) * 0.7))
            assert [(int(x), int(y), float(z)) for x, y, z in zip(o, t, a)] == ref


def _dense(triples, size):
    return O.coo_dense(size, *triples)


def test_docstring_matrices(ref_fixtures):
    f = ref_fixtures["hubbard_inter_2site"]
    up = dn = O.sector_states(2, 1)
    assert_array_equal(_dense(O.hubbard_inter_triples(up, dn, f["u"]), 4), f["matrix"])
    f = ref_fixtures["onsite_energy_2site"]
    assert_array_equal(_dense(O.onsite_energy_triples(up, dn, f["eps"]), 4), f["matrix"])
    f = ref_fixtures["hopping_2site"]
    assert_array_equal(_dense(O.hopping_triples(up, dn, 2, 0, 1, f["hop"]), 4), f["matrix"])


def test_hubbard_sector_1_1(ref_fixtures):
    f = ref_fixtures["hubbard_sector_1_1"]
    up = dn = O.sector_states(2, 1)
    tr = O.hubbard_triples(up, dn, 2, f["neighbors"], f["inter"], f["eps"], f["hop"])
    assert_array_equal(_dense(tr, 4), f["matrix"])


def test_project_up_dn(ref_fixtures):
    f = ref_fixtures["project_up"]
    for up_idx, exp in f["cases"]:
        assert_array_equal(up_idx * f["num_dn"] + np.asarray(f["dn_indices"]), exp)
    f = ref_fixtures["project_dn"]
    for dn_idx, exp in f["cases"]:
        assert_array_equal(np.asarray(f["up_indices"]) * f["num_dn"] + dn_idx, exp)


@pytest.mark.parametrize("ns", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("u", [0.0, 1.0, 2.0])
@pytest.mark.parametrize("periodic", [False, True])
def test_hubbard_hermitian(ns, u, periodic):
    # cmpy/tests/test_models_hubbard.py:28-48
    bonds = O.chain_bonds(ns, periodic)
    for n_up, n_dn in itertools.product(range(ns + 1), repeat=2):
        up, dn = O.sector_states(ns, n_up), O.sector_states(ns, n_dn)
        h = _dense(O.hubbard_triples(up, dn, ns, bonds, u, 0.0, 1.0, mu=u / 2), len(up) * len(dn))
        assert_allclose(h, h.T, atol=0)


@pytest.mark.parametrize("geom", ["chain6p", "ring5", "sq4x2", "sq2x2"])
@pytest.mark.parametrize("filling", [(1, 2), (2, 2), (3, 2)])
def test_free_fermion_closed_form(geom, filling):
    """U = 0: E0 = sum of the lowest single-particle levels.  Only correct hopping
    signs (mask between the sites, width = num_sites) reproduce it on rings/tori."""
    ns, bonds = {
        "chain6p": (6, O.chain_bonds(6, True)), "ring5": (5, O.chain_bonds(5, True)),
        "sq4x2": (8, O.square_bonds(4, 2, True)), "sq2x2": (4, O.square_bonds(2, 2, False)),
    }[geom]
    n_up, n_dn = filling
    op = O.hubbard_operator(ns, n_up, n_dn, bonds, 0.0, 0.3, 1.0)
    e0, _ = O.groundstate(op)
    assert abs(e0 - O.free_fermion_energy(ns, bonds, 1.0, 0.3, n_up, n_dn)) < 1e-10


@pytest.mark.parametrize("n", [2, 5, 12])
def test_eigvalsh_chain_closed_form(n):
    """Open chain: eps + 2 t cos(k pi / (N + 1)); the free-fermion energy of
    an open chain at U = 0 is the sum of the lowest levels."""
    eps, t = 0.3, -0.7
    k = np.arange(1, n + 1)
    assert_allclose(O.eigvalsh_chain(n, eps, t), np.sort(eps + 2 * t * np.cos(k * np.pi / (n + 1))), atol=1e-13)
    w = O.eigvalsh_chain(n, 0.0, 1.0)
    nu = n // 2
    assert abs(O.free_fermion_energy(n, O.chain_bonds(n), 1.0, 0.0, nu, nu) - 2 * w[:nu].sum()) < 1e-12


def test_sign_width_matters_on_a_ring():
    # Dropping the sign (width 0) breaks the closed form on the 4x2 torus at (3,3).
    ns, bonds = 8, O.square_bonds(4, 2, True)
    bad = O.SectorOperator(ns, 3, 3, bonds, [1.0] * len(bonds), np.zeros(ns), np.zeros(ns), 0)
    e_bad, _ = O.groundstate(bad)
    assert abs(e_bad - O.free_fermion_energy(ns, bonds, 1.0, 0.0, 3, 3)) > 1e-3


def test_atomic_limit_closed_form():
    # t = 0: E0 = min over configurations; at half filling with mu = U/2 it is -U/2 * N
    ns, u = 4, 3.0
    op = O.hubbard_operator(ns, 2, 2, [], u, 0.0, 0.0, mu=u / 2)
    e0, _ = O.groundstate(op)
    assert abs(e0 - (-u / 2 * ns)) < 1e-12


@pytest.mark.parametrize("case", [("hub", 5, 2, 3), ("hub", 4, 2, 2), ("siam", 0, 3, 3), ("siam", 0, 2, 3)])
def test_factorised_operator_matches_coo(case):
    kind, ns, n_up, n_dn = case
    rng = np.random.default_rng(7)
    if kind == "hub":
        bonds = O.chain_bonds(ns, True)
        op = O.hubbard_operator(ns, n_up, n_dn, bonds, 2.0, 0.3, 1.0, mu=1.0)
        up, dn = op.up, op.dn
        tr = O.hubbard_triples(up, dn, ns, bonds, 2.0, 0.3, 1.0, mu=1.0)
    else:
        eps, v = rng.uniform(-2, 2, 5), rng.uniform(0.2, 1.0, 5)
        op = O.siam_operator(n_up, n_dn, 4.0, -2.0, eps, v, 0.0)
        tr = O.siam_triples(op.up, op.dn, 4.0, -2.0, eps, v, 0.0)
    h = _dense(tr, op.size)
    assert_allclose(op.csr().toarray(), h, atol=1e-14)
    x = rng.standard_normal(op.size)
    assert_allclose(op.matvec(x), O.coo_matvec(op.size, *tr, x), atol=1e-12)
    assert_allclose(op.diagonal().sum(), O.coo_trace(*tr), atol=1e-12)


# -- ground state and Lanczos ------------------------------------------------------


def _cfg1(seed):
    rng = np.random.default_rng(seed)
    eps = rng.uniform(-2, 2, 5)
    v = rng.uniform(0.2, 1.0, 5)
    return eps, v


def test_siam_reference_convention_energies():
    """Both sign conventions of the 1+5-site SIAM at cfg-1 parameters (seed 0):
    the reference's width-0 signs and fermionic signs give different E0."""
    eps, v = _cfg1(0)
    e_ref, _ = O.groundstate(O.siam_operator(3, 3, 4.0, -2.0, eps, v, 0.0, width=0))
    e_fer, _ = O.groundstate(O.siam_operator(3, 3, 4.0, -2.0, eps, v, 0.0, width=6))
    assert abs(e_ref - (-12.287544399997)) < 1e-9
    assert abs(e_fer - (-12.234494553501)) < 1e-9


def test_siam_u0_fermion_signs_closed_form():
    eps, v = _cfg1(3)
    op = O.siam_operator(3, 3, 0.0, 0.0, eps, v, 0.0, width=6)
    e0, _ = O.groundstate(op)
    h = np.diag(np.append(0.0, eps))
    h[0, 1:] = h[1:, 0] = v
    assert abs(e0 - 2 * np.linalg.eigvalsh(h)[:3].sum()) < 1e-10


def test_lanczos_normalized_matches_reference_recurrence():
    eps, v = _cfg1(1)
    op = O.siam_operator(3, 3, 4.0, -2.0, eps, v, 0.0)
    psi = np.random.default_rng(2).uniform(0, 1, op.size)
    a_ref, b_ref = O.lanczos_coeffs_reference(op.matvec, psi, 25)
    a, b, n2, _ = O.lanczos_normalized(op.matvec, psi, 25)
    assert_allclose(a, a_ref, rtol=1e-9, atol=1e-9)
    assert_allclose(b, b_ref, rtol=1e-9, atol=1e-9)
    assert abs(n2 - psi @ psi) < 1e-9 * n2


def test_lanczos_tridiagonal_gives_e0():
    eps, v = _cfg1(2)
    op = O.siam_operator(3, 3, 4.0, -2.0, eps, v, 0.0)
    e_exact = np.linalg.eigvalsh(op.csr().toarray())[0]
    a, b, _, _ = O.lanczos_normalized(op.matvec, np.ones(op.size), 120, reorth=True)
    e, _ = O.lanczos_ground_state(a, b)
    assert abs(e - e_exact) < 1e-10


# -- Green's functions ---------------------------------------------------------------


@pytest.mark.parametrize("u,v", [(2.0, 1.0), (4.0, 0.5), (6.0, 0.25)])
def test_lange_two_site_kat(u, v):
    """T = 0 Lehmann of the 2-site SIAM equals the closed form (dmft/twosite.py:25-81)."""
    opn = O.siam_operator(1, 1, u, 0.0, [0.0], [v], u / 2)
    e0, psi = O.groundstate(opn)
    opp = O.siam_operator(2, 1, u, 0.0, [0.0], [v], u / 2)
    opm = O.siam_operator(0, 1, u, 0.0, [0.0], [v], u / 2)
    cd = O.apply_cdag(psi, 2, 1, 1, 0, O.UP)
    cc = O.apply_c(psi, 2, 1, 1, 0, O.UP)
    z = np.linspace(-5, 5, 201) + 0.05j
    g = O.lehmann_t0(opn, opp, opm, psi, e0, cd, cc, z)
    assert_allclose(g, O.impurity_gf_ref(z, u, v), atol=1e-12)


def _siam_g_pieces(seed, width=0, nsteps=300):
    eps, v = _cfg1(seed)
    mk = lambda nu, nd: O.siam_operator(nu, nd, 4.0, -2.0, eps, v, 0.0, width=width)
    op = mk(3, 3)
    e0, psi = O.groundstate(op)
    mode = "jordan_wigner" if width else "reference"
    cd = O.apply_cdag(psi, 6, 3, 3, 0, O.UP, mode)
    cc = O.apply_c(psi, 6, 3, 3, 0, O.UP, mode)
    return mk, op, e0, psi, cd, cc


@pytest.mark.parametrize("width", [0, 6])
def test_cf_equals_lehmann_small(width):
    mk, op, e0, psi, cd, cc = _siam_g_pieces(4, width)
    z = np.linspace(-8, 8, 1000) + 0.05j
    g_l = O.lehmann_t0(op, mk(4, 3), mk(2, 3), psi, e0, cd, cc, z)
    ap, bp, np_, _ = O.lanczos_normalized(mk(4, 3).matvec, cd, 300, reorth=True)
    ah, bh, nh, _ = O.lanczos_normalized(mk(2, 3).matvec, cc, 300, reorth=True)
    g_cf = O.cf_greens(ap, bp, np_, z, e0, +1) + O.cf_greens(ah, bh, nh, z, e0, -1)
    assert_allclose(g_cf, g_l, atol=1e-8)


def test_lehmann_sum_rule_and_tail():
    mk, op, e0, psi, cd, cc = _siam_g_pieces(5)
    # total weight of c and c+ is <{c, c+}> = 1 at pos 0
    assert abs(cd @ cd + cc @ cc - 1.0) < 1e-12
    z = np.array([1e6j])
    g = O.lehmann_t0(op, mk(4, 3), mk(2, 3), psi, e0, cd, cc, z)
    assert abs(g[0] * z[0] - 1.0) < 1e-6


@pytest.mark.parametrize("u,v", [(2.0, 1.0), (5.0, 0.4)])
def test_tevo_two_site_closed_form(u, v):
    """g>(t) - g<(t) of the two-site SIAM equals -2i sum_k a_k cos(e_k t) with
    the Lange poles; at t = 0 it is -i <{c, c+}> = -i."""
    opn = O.siam_operator(1, 1, u, 0.0, [0.0], [v], u / 2)
    e0, psi = O.groundstate(opn)
    opp = O.siam_operator(2, 1, u, 0.0, [0.0], [v], u / 2)
    opm = O.siam_operator(0, 1, u, 0.0, [0.0], [v], u / 2)
    cd = O.apply_cdag(psi, 2, 1, 1, 0, O.UP)
    cc = O.apply_c(psi, 2, 1, 1, 0, O.UP)
    t = np.linspace(0.0, 30.0, 301)
    g = O.tevo_part(opp, cd, e0, t, +1) - O.tevo_part(opm, cc, e0, t, -1)
    ref = O.tevo_two_site_ref(t, u, v)
    assert_allclose(g, ref, atol=1e-12)
    assert abs(ref[0] + 1j) < 1e-14
    # the hole part alone carries +i e^{+i e_k t}: a flipped sign fails this
    a1, a2, e1, e2 = O.impurity_params_ref(u, v)
    assert_allclose(O.tevo_part(opm, cc, e0, t, -1), 1j * (a1 * np.exp(1j * e1 * t) + a2 * np.exp(1j * e2 * t)),
                    atol=1e-12)


@pytest.mark.parametrize("width", [0, 6])
def test_tevo_dense_matches_expm_multiply(width):
    """The dense-eigh restatement equals the reference's own procedure
    (scipy's expm_multiply interval algorithm + running phase product)."""
    mk, op, e0, psi, cd, cc = _siam_g_pieces(7, width)
    for sign, opx, phi in ((+1, mk(4, 3), cd), (-1, mk(2, 3), cc)):
        times, ov = O.tevo_expm_multiply(opx, phi, e0, 0.0, 12.0, 121, sign)
        assert_allclose(O.tevo_part(opx, phi, e0, times, sign), ov, atol=1e-9)


def test_tevo_dense_matches_expm_multiply_hubbard():
    ns, bonds = 6, O.chain_bonds(6, periodic=True)
    op = O.hubbard_operator(ns, 3, 3, bonds, 3.0, 0.0, 1.0, mu=1.5)
    e0, psi = O.groundstate(op)
    opp = O.hubbard_operator(ns, 3, 4, bonds, 3.0, 0.0, 1.0, mu=1.5)
    cd = O.apply_cdag(psi, ns, 3, 3, 2, O.DN, "jordan_wigner")
    times, ov = O.tevo_expm_multiply(opp, cd, e0, -2.0, 8.0, 101, +1)
    assert_allclose(O.tevo_part(opp, cd, e0, times, +1), ov, atol=1e-9)


@pytest.mark.parametrize("ns", [2, 3, 4])
def test_cdag_c_adjoint(ns):
    # cmpy/tests/test_operator.py:31-42
    for n_up, n_dn in itertools.product(range(ns + 1), repeat=2):
        for sigma in (O.UP, O.DN):
            p1 = O.upper_sector(n_up, n_dn, sigma, ns)
            if p1 is None:
                continue
            D = math.comb(ns, n_up) * math.comb(ns, n_dn)
            D1 = math.comb(ns, p1[0]) * math.comb(ns, p1[1])
            for pos in range(ns):
                for mode in ("reference", "jordan_wigner"):
                    cd = np.column_stack([O.apply_cdag(e, ns, n_up, n_dn, pos, sigma, mode) for e in np.eye(D)])
                    c = np.column_stack([O.apply_c(e, ns, *p1, pos, sigma, mode) for e in np.eye(D1)])
                    assert_array_equal(c, cd.T)


def test_jordan_wigner_anticommutation():
    """With JW signs {c_i, c+_j} = delta_ij on the up species (fails without them)."""
    ns = 3
    for i, j in itertools.product(range(ns), repeat=2):
        for n_up in range(1, ns):
            n_dn = 1
            D = math.comb(ns, n_up) * math.comb(ns, n_dn)
            M = np.zeros((D, D))
            for k, e in enumerate(np.eye(D)):
                a = O.apply_c(O.apply_cdag(e, ns, n_up, n_dn, j, O.UP, "jordan_wigner"), ns, n_up + 1, n_dn, i, O.UP, "jordan_wigner")
                b = O.apply_cdag(O.apply_c(e, ns, n_up, n_dn, i, O.UP, "jordan_wigner"), ns, n_up - 1, n_dn, j, O.UP, "jordan_wigner")
                M[:, k] = a + b
            assert_allclose(M, np.eye(D) * (i == j), atol=0)


# -- the reference's Lehmann tests (test_exactdiag.py), restated on the oracle ---------


def _chain_tb(ns, eps=0.0, hop=1.0):
    h = np.diag(np.full(ns, eps))
    for i, j in O.chain_bonds(ns):
        h[i, j] = h[j, i] = hop
    return h


@pytest.mark.parametrize("ns", [2, 3, 4])
def test_gf_lehmann_non_interacting(ns, ref_fixtures):
    f = ref_fixtures["gf_lehmann_tests"]["non_interacting"]
    z = np.linspace(f["z_min"], f["z_max"], f["nz"]) + 1j * f["eta"]
    mk = lambda nu, nd: O.hubbard_operator(ns, nu, nd, O.chain_bonds(ns), 0.0, 0.0, 1.0)
    g = O.gf_lehmann(mk, ns, z, f["beta"], pos=0, occ=False).gf
    g0 = O.gf0_lehmann(_chain_tb(ns), z)[:, 0]
    assert_allclose(g, g0, rtol=f["rtol"], atol=f["atol"])
    # gf0_lehmann equals the directly inverted resolvent on chains
    h = _chain_tb(ns)
    direct = np.array([np.linalg.inv(zz * np.eye(ns) - h)[0, 0] for zz in z[::50]])
    assert_allclose(g0[::50], direct, atol=1e-12)


def _tb_matrix(ns, bonds, hop, eps):
    h = np.diag(np.full(ns, float(eps)))
    for i, j in bonds:
        h[i, j] += hop
        h[j, i] += hop
    return h


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("ij", [(1, 3), (0, 4), (2, 2)])
def test_lehmann_ij_non_interacting_resolvent(periodic, ij):
    """At U = 0 the grand-canonical G_ij is the single-particle resolvent at any
    temperature; interior sites need the Jordan-Wigner string."""
    ns = 6
    bonds = O.chain_bonds(ns, periodic=periodic)
    mk = lambda nu, nd: O.hubbard_operator(ns, nu, nd, bonds, 0.0, 0.4, 1.0, mu=0.3)
    z = np.linspace(-4, 4, 41) + 0.1j
    g = O.lehmann_ij(mk, ns, z, 3.0, *ij)
    ref = O.resolvent_ij(_tb_matrix(ns, bonds, 1.0, 0.4 - 0.3), z)[:, ij[0], ij[1]]
    assert_allclose(g, ref, atol=1e-11)
    direct = np.linalg.inv(z[7] * np.eye(ns) - _tb_matrix(ns, bonds, 1.0, 0.1))[ij[0], ij[1]]
    assert abs(ref[7] - direct) < 1e-12
    if ij[0] != ij[1] and periodic:
        # without the string the off-diagonal sum is not the resolvent
        bad = O.lehmann_ij(mk, ns, z, 3.0, *ij, sign_mode="reference")
        assert np.abs(bad - ref).max() > 1e-3


@pytest.mark.parametrize("ij", [(1, 4), (0, 3), (2, 2)])
def test_lehmann_t0_ij_closed_shell_resolvent(ij):
    """6-site ring at half filling (closed shell): T = 0 G_ij at U = 0 is the
    resolvent; with U the i = j case equals lehmann_t0."""
    ns = 6
    bonds = O.chain_bonds(ns, periodic=True)
    z = np.linspace(-4, 4, 41) + 0.1j
    for u in (0.0, 3.0):
        mk = lambda nu, nd: O.hubbard_operator(ns, nu, nd, bonds, u, 0.0, 1.0, mu=u / 2)
        op = mk(3, 3)
        e0, psi = O.groundstate(op)
        lad = lambda f, p: f(psi, ns, 3, 3, p, O.UP, "jordan_wigner")
        g = O.lehmann_t0_ij(mk(4, 3), mk(2, 3), e0, lad(O.apply_cdag, ij[0]), lad(O.apply_cdag, ij[1]),
                            lad(O.apply_c, ij[0]), lad(O.apply_c, ij[1]), z)
        if u == 0.0:
            ref = O.resolvent_ij(_tb_matrix(ns, bonds, 1.0, 0.0), z)[:, ij[0], ij[1]]
            assert_allclose(g, ref, atol=1e-11)
        elif ij[0] == ij[1]:
            ref = O.lehmann_t0(op, mk(4, 3), mk(2, 3), psi, e0, lad(O.apply_cdag, ij[0]), lad(O.apply_c, ij[0]), z)
            assert_allclose(g, ref, atol=1e-12)


def test_lehmann_ij_diagonal_and_symmetry():
    ns = 5
    bonds = O.chain_bonds(ns, periodic=True)
    mk = lambda nu, nd: O.hubbard_operator(ns, nu, nd, bonds, 2.5, 0.0, 1.0, mu=1.25)
    z = np.linspace(-6, 6, 31) + 0.2j
    # i = j is gf_lehmann with the same ladder operators (site 0 carries no
    # string, so both sign modes agree there; at an interior site the
    # reference's string-free c+ is a different operator), up to the
    # partition function: gf_lehmann sums Z only over sectors that have an
    # upper sector (cmpy/exactdiag.py:228-243), lehmann_ij over all of them
    beta = 4.0
    spec = {(nu, nd): np.linalg.eigvalsh(mk(nu, nd).csr().toarray())
            for nu in range(ns + 1) for nd in range(ns + 1)}
    emin = min(w[0] for w in spec.values())
    z_all = sum(np.exp(-beta * (w - emin)).sum() for w in spec.values())
    z_ref = sum(np.exp(-beta * (w - emin)).sum() for (nu, nd), w in spec.items() if nu < ns)
    for pos, mode in ((0, "jordan_wigner"), (2, "reference")):
        g_ref = O.gf_lehmann(mk, ns, z, beta, pos=pos, occ=False).gf
        assert_allclose(O.lehmann_ij(mk, ns, z, beta, pos, pos, sign_mode=mode), g_ref * z_ref / z_all, atol=1e-12)
    assert_allclose(O.lehmann_ij(mk, ns, z, beta, 1, 3), O.lehmann_ij(mk, ns, z, beta, 3, 1), atol=1e-12)


@pytest.mark.parametrize("ns", [2, 3])
@pytest.mark.parametrize("u", [1.0, 4.0])
def test_gf_lehmann_atomic_and_occupation(ns, u, ref_fixtures):
    f = ref_fixtures["gf_lehmann_tests"]["atomic_limit"]
    z = np.linspace(f["z_min"], f["z_max"], f["nz"]) + 1j * f["eta"]
    mk = lambda nu, nd: O.hubbard_operator(ns, nu, nd, O.chain_bonds(ns), u, 0.0, 0.0, mu=u / 2)
    for pos in range(ns):
        g = O.gf_lehmann(mk, ns, z, f["beta"], pos=pos, occ=False).gf
        c = len(z) // 2
        e0 = z.real[:c][np.argmax(-g.imag[:c])]
        e1 = z.real[c:][np.argmax(-g.imag[c:])]
        assert abs(e0 + u / 2) < f["peak_tol"] and abs(e1 - u / 2) < f["peak_tol"]
    mk = lambda nu, nd: O.hubbard_operator(ns, nu, nd, O.chain_bonds(ns), u, 0.0, 1.0, mu=u / 2)
    m = O.gf_lehmann(mk, ns, z[:5], 10.0, pos=0, occ=True)
    assert abs(m.occ - 0.5) < 1e-3


@pytest.mark.parametrize("nb,width", [(3, 0), (4, 0), (4, 5)])
def test_bathfit_grad_dense_matches_finite_differences(nb, width):
    """The CPU reference gradient of the bath-fit loss against central
    differences (h = 1e-5) of the loss itself."""
    rng = np.random.default_rng(40 + nb + width)
    u, eimp = 3.0, -1.5
    n_up = n_dn = (nb + 1) // 2
    eb, v = rng.uniform(-3, 3, nb), rng.uniform(0.1, 1.2, nb)
    wn = (2 * np.arange(32) + 1) * np.pi / 100
    target = O.siam_gf_t0(u, eimp, rng.uniform(-3, 3, nb), rng.uniform(0.1, 1.2, nb), 0.0, n_up, n_dn, wn, width)
    loss, ge, gv = O.bathfit_grad_dense(u, eimp, eb, v, 0.0, n_up, n_dn, wn, target, width)
    assert loss == pytest.approx(np.sum(np.abs(O.siam_gf_t0(u, eimp, eb, v, 0.0, n_up, n_dn, wn, width) - target) ** 2))
    h = 1e-5

    def L(e, vv):
        return np.sum(np.abs(O.siam_gf_t0(u, eimp, e, vv, 0.0, n_up, n_dn, wn, width) - target) ** 2)

    fd = []
    for kind in ("eps", "v"):
        for k in range(nb):
            d = np.zeros(nb)
            d[k] = h
            if kind == "eps":
                fd.append((L(eb + d, v) - L(eb - d, v)) / (2 * h))
            else:
                fd.append((L(eb, v + d) - L(eb, v - d)) / (2 * h))
    fd = np.array(fd)
    got = np.concatenate([ge, gv])
    assert np.max(np.abs(got - fd)) <= 1e-6 * np.max(np.abs(fd))


# -- occupations, double occupation, hybridisation function -------------------------


def _sector_spectra(make_op, ns):
    out = {}
    for nu in range(ns + 1):
        for nd in range(ns + 1):
            op = make_op(nu, nd)
            w, v = np.linalg.eigh(op.csr().toarray())
            out[(nu, nd)] = (op, w, v)
    return out


def _thermal(make_op, ns, beta, pos, dbl=None):
    """Grand-canonical <n_up>, <n_dn>, <n_up n_dn> at `pos`: the per-sector sums
    of occupation / double_occupation over every sector, divided by Z."""
    dbl = dbl or O.double_occupation
    spec = _sector_spectra(make_op, ns)
    emin = min(w[0] for _, w, _ in spec.values())
    z = nu_ = nd_ = d_ = 0.0
    for op, w, v in spec.values():
        z += np.exp(-beta * (w - emin)).sum()
        nu_ += O.occupation(op.up, op.dn, w, v, beta, emin, pos, O.UP)
        nd_ += O.occupation(op.up, op.dn, w, v, beta, emin, pos, O.DN)
        d_ += dbl(op.up, op.dn, w, v, beta, emin, pos)
    return nu_ / z, nd_ / z, d_ / z


_ATOMIC_EPS = np.array([0.3, -0.7, 1.1])
_ATOMIC_U = np.array([2.0, 4.0, 0.5])


def _atomic_op(nu, nd):
    return O.SectorOperator(3, nu, nd, [], [], _ATOMIC_EPS, _ATOMIC_U, 3)


@pytest.mark.parametrize("beta", [0.7, 2.5])
def test_occupations_atomic_limit(beta):
    """t = 0: every site is an independent four-state atom with levels
    0, eps, eps, 2 eps + U, so <n_s> = (e^{-b eps} + e^{-b(2eps+U)})/Z_i and
    <n_up n_dn> = e^{-b(2eps+U)}/Z_i, site by site."""
    for pos in range(3):
        e, u = _ATOMIC_EPS[pos], _ATOMIC_U[pos]
        w1, w2 = np.exp(-beta * e), np.exp(-beta * (2 * e + u))
        zi = 1 + 2 * w1 + w2
        n_up, n_dn, d = _thermal(_atomic_op, 3, beta, pos)
        assert abs(n_up - (w1 + w2) / zi) < 1e-13
        assert abs(n_dn - (w1 + w2) / zi) < 1e-13
        assert abs(d - w2 / zi) < 1e-13


def _u0_chain_op(nu, nd):
    eps = np.array([0.4, -0.3, 0.9, -1.2])
    bonds = O.chain_bonds(4)
    return O.SectorOperator(4, nu, nd, bonds, [1.0, 0.6, 1.3], eps, np.zeros(4), 4)


def _u0_chain_occ(beta, pos):
    h = np.diag([0.4, -0.3, 0.9, -1.2])
    for (i, j), t in zip(O.chain_bonds(4), [1.0, 0.6, 1.3]):
        h[i, j] = h[j, i] = t
    w, v = np.linalg.eigh(h)
    return float(np.sum(v[pos] ** 2 / (1 + np.exp(beta * w))))


@pytest.mark.parametrize("pos", [0, 1, 3])
def test_occupations_u0_wick(pos):
    """U = 0 (fermionic signs, uneven levels and hoppings): <n_s> at a site is
    the Fermi sum over single-particle orbitals, and <n_up n_dn> factorises
    into <n_up><n_dn> (the two spins are independent)."""
    beta = 1.7
    n_ref = _u0_chain_occ(beta, pos)
    n_up, n_dn, d = _thermal(_u0_chain_op, 4, beta, pos)
    assert abs(n_up - n_ref) < 1e-12 and abs(n_dn - n_ref) < 1e-12
    assert abs(d - n_ref * n_ref) < 1e-12


def _dbl_running_idx_reset(up_states, dn_states, evals, evecs, beta, emin, pos):
    # planted mistake: the running index restarts for every up state
    boltz = np.exp(-beta * (evals - emin))
    occ = 0.0
    for ui, up in enumerate(up_states):
        for di, dn in enumerate(dn_states):
            if up & dn & (1 << pos):
                occ += np.sum(boltz * evecs[di, :] ** 2)
    return occ


def _dbl_dn_major(up_states, dn_states, evals, evecs, beta, emin, pos):
    # planted mistake: the flat index is taken dn-major
    boltz = np.exp(-beta * (evals - emin))
    occ = 0.0
    for ui, up in enumerate(up_states):
        for di, dn in enumerate(dn_states):
            if up & dn & (1 << pos):
                occ += np.sum(boltz * evecs[di * len(up_states) + ui, :] ** 2)
    return occ


def _dbl_skip_increment(up_states, dn_states, evals, evecs, beta, emin, pos):
    # planted mistake (the shape of the F9 race): the running index only
    # advances on the selected states
    boltz = np.exp(-beta * (evals - emin))
    occ, idx = 0.0, 0
    for up in up_states:
        for dn in dn_states:
            if up & dn & (1 << pos):
                occ += np.sum(boltz * evecs[idx, :] ** 2)
                idx += 1
    return occ


@pytest.mark.parametrize("bad", [_dbl_running_idx_reset, _dbl_dn_major, _dbl_skip_increment])
def test_double_occupation_pins_catch_index_mistakes(bad):
    """Each planted indexing mistake fails the U = 0 Wick pin."""
    beta, pos = 1.7, 1
    n_ref = _u0_chain_occ(beta, pos)
    _, _, d = _thermal(_u0_chain_op, 4, beta, pos, dbl=bad)
    assert abs(d - n_ref * n_ref) > 1e-4


def _siam_u0(eps_imp, mu, eb, v):
    ns = len(eb) + 1
    return lambda nu, nd: O.siam_operator(nu, nd, 0.0, eps_imp, eb, v, mu, width=ns)


_HYB_EB = np.array([-1.3, 0.4, 1.9])
_HYB_V = np.array([0.7, -1.1, 0.45])


def test_hybridization_func_is_the_u0_impurity_self_energy():
    """U = 0 impurity model with fermionic signs: the grand-canonical impurity
    G (lehmann_ij, pinned to the resolvent) equals 1/(z - (eps_imp - mu) - Delta(z))
    with Delta(z) = sum_k |V_k|^2/(z - eps_k); and Delta equals the bath block
    of the resolvent computed directly."""
    z = np.linspace(-4, 4, 33) + 0.15j
    eps_imp, mu = 0.6, 0.25
    g = O.lehmann_ij(_siam_u0(eps_imp, mu, _HYB_EB, _HYB_V), 4, z, 2.0, 0, 0)
    delta = O.hybridization_func(z, _HYB_V, _HYB_EB)
    assert_allclose(g, 1 / (z - (eps_imp - mu) - delta), atol=1e-11)
    direct = np.array([_HYB_V @ np.linalg.solve(zz * np.eye(3) - np.diag(_HYB_EB), _HYB_V) for zz in z])
    assert_allclose(delta, direct, atol=1e-13)
    # a |V| (not squared) restatement would fail
    bad = np.sum(np.abs(_HYB_V) / (z[:, None] - _HYB_EB), axis=-1)
    assert np.max(np.abs(bad - delta)) > 1e-2


def test_impurity_gf0_convention():
    """anderson.py:160-161 writes G0 = 1/(z + mu + eps_imp - Delta).  At
    eps_imp = 0 that is the U = 0 impurity G of the model's own Hamiltonian
    (on-site level eps_imp - mu); for eps_imp != 0 it is the U = 0 G of the
    model with the impurity level reflected, eps_imp -> -eps_imp.  A flipped
    mu sign or a physical-sign restatement fails one of the two checks."""
    z = np.linspace(-4, 4, 33) + 0.15j
    mu = 0.35
    g_zero = O.lehmann_ij(_siam_u0(0.0, mu, _HYB_EB, _HYB_V), 4, z, 2.0, 0, 0)
    assert_allclose(O.impurity_gf0(z, 3.0, 0.0, _HYB_V, _HYB_EB, mu), g_zero, atol=1e-11)
    e = 0.8
    g_refl = O.lehmann_ij(_siam_u0(-e, mu, _HYB_EB, _HYB_V), 4, z, 2.0, 0, 0)
    assert_allclose(O.impurity_gf0(z, 3.0, e, _HYB_V, _HYB_EB, mu), g_refl, atol=1e-11)
    g_phys = O.lehmann_ij(_siam_u0(e, mu, _HYB_EB, _HYB_V), 4, z, 2.0, 0, 0)
    assert np.max(np.abs(O.impurity_gf0(z, 3.0, e, _HYB_V, _HYB_EB, mu) - g_phys)) > 1e-2
    assert np.max(np.abs(O.impurity_gf0(z, 3.0, 0.0, _HYB_V, _HYB_EB, -mu) - g_zero)) > 1e-2


def test_every_oracle_function_is_in_the_parity_table():
    import inspect

    named = set()
    for key in O.PARITY:
        named.update(k.strip() for k in key.split("/"))
    public = {n for n, f in vars(O).items()
              if (inspect.isfunction(f) or inspect.isclass(f)) and f.__module__ == O.__name__}
    assert public - named == set()


@pytest.mark.parametrize("nx,ny", [(4, 4), (4, 2), (3, 3)])
def test_lattice_bonds_and_free_fermion_energy_closed_form(nx, ny):
    """Periodic square lattice: every site has four neighbours (fewer where a
    side has length 2) and the tight-binding levels are 2t(cos kx + cos ky);
    the free-fermion energy is the sum of the lowest of those levels, and the
    periodic chain's levels are 2t cos(2 pi k/N)."""
    t, ns = 0.8, nx * ny
    bonds = O.square_bonds(nx, ny, True)
    assert all(i < j for i, j in bonds) and len(set(bonds)) == len(bonds)
    per_dir = lambda n: 2 if n > 2 else 1
    assert len(bonds) == ns * (per_dir(nx) + per_dir(ny)) // 2
    kx, ky = np.meshgrid(2 * np.pi * np.arange(nx) / nx, 2 * np.pi * np.arange(ny) / ny)
    lv = np.sort((2 * t * (np.cos(kx) * (nx > 2 or 0.5) + np.cos(ky) * (ny > 2 or 0.5))).ravel())
    n = ns // 2
    assert abs(O.free_fermion_energy(ns, bonds, t, 0.0, n, n - 1) - (lv[:n].sum() + lv[:n - 1].sum())) < 1e-12
    k = np.arange(7)
    lv = np.sort(2 * t * np.cos(2 * np.pi * k / 7))
    assert abs(O.free_fermion_energy(7, O.chain_bonds(7, True), t, 0.0, 3, 2) - (lv[:3].sum() + lv[:2].sum())) < 1e-12


def test_lehmann_ij_many_equals_single_pairs():
    ns = 4
    bonds = O.chain_bonds(ns, periodic=True)
    mk = lambda nu, nd: O.hubbard_operator(ns, nu, nd, bonds, 2.0, 0.0, 1.0, mu=1.0)
    z = np.linspace(-4, 4, 9) + 0.3j
    pairs = [(0, 0), (0, 2), (1, 3)]
    many = O.lehmann_ij_many(mk, ns, z, 3.0, pairs)
    for (i, j), g in zip(pairs, many):
        assert_allclose(g, O.lehmann_ij(mk, ns, z, 3.0, i, j), atol=1e-14)


@pytest.mark.parametrize("mode", ["reference", "jordan_wigner"])
@pytest.mark.parametrize("sigma", [O.UP, O.DN])
def test_ladder_block_equals_columns(mode, sigma):
    ns, nu, nd = 5, 2, 3
    rng = np.random.default_rng(4)
    X = rng.standard_normal((O.binom(ns, nu) * O.binom(ns, nd), 6))
    for fn in (O.apply_cdag, O.apply_c):
        for pos in range(ns):
            blk = fn(X, ns, nu, nd, pos, sigma, mode)
            cols = np.column_stack([fn(X[:, k], ns, nu, nd, pos, sigma, mode) for k in range(6)])
            assert_array_equal(blk, cols)


# --- fourier_t2z (cmpy/exactdiag.py:311-316) ---------------------------------


def _t2z_exact_linear(t0, t1, c0, c1, z):
    """int_{t0}^{t1} e^{izt} (c0 + c1 t) dt in closed form."""
    a = 1j * z
    F = lambda t: np.exp(a * t) * (c0 / a + c1 * (t / a - 1.0 / a**2))
    return F(t1) - F(t0)


def test_fourier_t2z_exact_on_piecewise_linear():
    rng = np.random.default_rng(4)
    t = np.sort(np.concatenate([[0.0, 7.0], rng.uniform(0, 7, 60)]))
    w = np.linspace(-4, 4, 33)
    # one line through every sample: exact
    c0, c1 = 0.4 - 0.9j, -0.3 + 0.2j
    z, gw = O.fourier_t2z(t, c0 + c1 * t, w, eta=0.3)
    assert np.abs(gw - _t2z_exact_linear(t[0], t[-1], c0, c1, z)).max() <= 1e-12
    # a kink at a sample: exact piece by piece
    tm = t[30]
    g = np.where(t <= tm, c0 + c1 * t, (c0 + c1 * tm) - 2.0 * (t - tm))
    z, gw = O.fourier_t2z(t, g, w, eta=0.3)
    ref = _t2z_exact_linear(t[0], tm, c0, c1, z) + _t2z_exact_linear(tm, t[-1], c0 + c1 * tm + 2.0 * tm, -2.0, z)
    assert np.abs(gw - ref).max() <= 1e-12


def test_fourier_t2z_second_order_on_exponentials():
    e, T = 1.7, 6.0
    w = np.linspace(-3, 3, 25)
    errs = []
    for n in (301, 601):
        t = np.linspace(0, T, n)
        z, gw = O.fourier_t2z(t, np.exp(-1j * e * t), w, eta=0.5)
        exact = (np.exp(1j * (z - e) * T) - 1) / (1j * (z - e))
        errs.append(np.abs(gw - exact).max())
    assert 3.8 < errs[0] / errs[1] < 4.2  # O(h^2)
    assert errs[1] < 1e-4


def test_fourier_t2z_two_site_closed_form():
    """Transform of g> - g< of the two-site SIAM, -2i sum_k a_k cos(e_k t):
    G(z) = -sum_k a_k [(e^{i(z-e_k)T} - 1)/(z - e_k) + (e^{i(z+e_k)T} - 1)/(z + e_k)],
    which tends to the Lange G(z) = sum_k a_k [1/(z - e_k) + 1/(z + e_k)] as T grows."""
    u, v = 2.0, 0.5
    a1, a2, e1, e2 = O.impurity_params_ref(u, v)
    T = 40.0
    t = np.linspace(0, T, 8001)
    w = np.linspace(-3, 3, 41)
    z, gw = O.fourier_t2z(t, O.tevo_two_site_ref(t, u, v), w, delta=1e-4)
    assert abs(z[0].imag - (-np.log(1e-4) / T)) <= 1e-15
    ref = np.zeros_like(z)
    for a, e in ((a1, e1), (a2, e2)):
        ref -= a * ((np.exp(1j * (z - e) * T) - 1) / (z - e) + (np.exp(1j * (z + e) * T) - 1) / (z + e))
    assert np.abs(gw - ref).max() <= 1e-5
    lange = sum(a * (1 / (z - e) + 1 / (z + e)) for a, e in ((a1, e1), (a2, e2)))
    assert np.abs(gw - lange).max() <= 1e-3  # e^{-eta T} = delta = 1e-4 relative tail


def test_fourier_t2z_series_and_direct_agree():
    x = 0.1 * np.exp(1j * np.linspace(0, 2 * np.pi, 17)) * 0.999
    e1 = np.expm1(x) / x
    e2 = (x * np.exp(x) - np.expm1(x)) / x**2
    assert np.abs(O._series(x, 1) - e1).max() <= 1e-14
    assert np.abs(O._series(x, 2) - e2).max() <= 1e-13
