"""GPU parity: basis, projector tables, H.v and ladder operators through the C ABI.

Integer work (state lists, ranks, hop targets, signs) must be bit-exact against
the oracle; FP64 matrix elements exact (they are sums of the same terms) and
H.v within 1e-13 relative.
"""

import ctypes as C
import itertools
import math

import numpy as np
import pytest
import torch
from numpy.testing import assert_allclose, assert_array_equal

from oracle import ed_oracle as O

pytestmark = pytest.mark.gpu


def _lib():
    from cmpy_amd import _lib

    return _lib


@pytest.mark.parametrize("ns", [1, 4, 8, 12, 16, 20])
def test_device_states_and_ranks(gpu, ns):
    L = _lib()
    for n in range(ns + 1):
        cnt = math.comb(ns, n)
        st = torch.empty(cnt, dtype=torch.int32, device="cuda")
        L.call("ed_sector_states", ns, n, L.dptr(st), L.stream_ptr())
        got = st.cpu().numpy().view(np.uint32).astype(np.int64)
        assert_array_equal(got, O.sector_states(ns, n))
        rk = torch.empty(cnt, dtype=torch.int64, device="cuda")
        L.call("ed_sector_rank", ns, n, L.dptr(st), cnt, L.dptr(rk), L.stream_ptr())
        assert_array_equal(rk.cpu().numpy(), np.arange(cnt))
    # a wrong popcount ranks to -1
    bad = torch.tensor([0b111, 0b1], dtype=torch.int32, device="cuda")
    rk = torch.empty(2, dtype=torch.int64, device="cuda")
    L.call("ed_sector_rank", 4, 2, L.dptr(bad), 2, L.dptr(rk), L.stream_ptr())
    assert rk.cpu().tolist() == [-1, -1]


@pytest.mark.parametrize("ns,n", [(4, 2), (6, 3), (8, 4), (10, 3), (12, 6), (16, 8)])
@pytest.mark.parametrize("width", ["ns", 0, 3])
def test_hop_terms_bit_exact(gpu, ns, n, width):
    from cmpy_amd.operators import hopping_terms

    w = ns if width == "ns" else width
    st = O.sector_states(ns, n)
    pairs = list(itertools.combinations(range(ns), 2))
    if len(pairs) > 40:
        pairs = pairs[::3]
    for i, j in pairs:
        o, t, a = hopping_terms(ns, n, w, i, j, -0.8)
        ro, rt, ra = O.hopping_terms(st, w, i, j, -0.8)
        assert_array_equal(o, ro)
        assert_array_equal(t, rt)
        assert_array_equal(a, ra)


def test_reference_fixture_matrices(gpu, ref_fixtures):
    from cmpy_amd import HubbardModel, SectorHamiltonOperator

    f = ref_fixtures["hubbard_sector_1_1"]
    m = HubbardModel(f["num_sites"], f["neighbors"], inter=f["inter"], eps=f["eps"], hop=f["hop"])
    assert_array_equal(m.hamiltonian(1, 1), f["matrix"])
    assert_array_equal(m.coo_operator(1, 1).toarray(), f["matrix"])
    f = ref_fixtures["hubbard_inter_2site"]
    op = SectorHamiltonOperator(2, 1, 1, [], [], [0.0, 0.0], f["u"], 2)
    assert_array_equal(op.toarray(), f["matrix"])
    f = ref_fixtures["onsite_energy_2site"]
    op = SectorHamiltonOperator(2, 1, 1, [], [], f["eps"], [0.0, 0.0], 2)
    assert_array_equal(op.toarray(), f["matrix"])
    f = ref_fixtures["hopping_2site"]
    op = SectorHamiltonOperator(2, 1, 1, [(0, 1)], [f["hop"]], [0.0, 0.0], [0.0, 0.0], 2)
    assert_array_equal(op.toarray(), f["matrix"])


def _models():
    from cmpy_amd import HubbardModel, SingleImpurityAndersonModel, chain_neighbors, square_neighbors

    rng = np.random.default_rng(5)
    eps, v = rng.uniform(-2, 2, 5), rng.uniform(0.2, 1.0, 5)
    return {
        "chain5": (HubbardModel(5, chain_neighbors(5), inter=2.0, eps=0.3, hop=1.0, mu=1.0),
                   lambda nu, nd: O.hubbard_operator(5, nu, nd, O.chain_bonds(5), 2.0, 0.3, 1.0, mu=1.0)),
        "ring6": (HubbardModel(6, chain_neighbors(6, True), inter=4.0, hop=-1.0),
                  lambda nu, nd: O.hubbard_operator(6, nu, nd, O.chain_bonds(6, True), 4.0, 0.0, -1.0)),
        "sq4x2": (HubbardModel(8, square_neighbors(4, 2), inter=3.0, hop=1.0, mu=1.5),
                  lambda nu, nd: O.hubbard_operator(8, nu, nd, O.square_bonds(4, 2), 3.0, 0.0, 1.0, mu=1.5)),
        "siam": (SingleImpurityAndersonModel(4.0, -2.0, eps, v, mu=0.0),
                 lambda nu, nd: O.siam_operator(nu, nd, 4.0, -2.0, eps, v, 0.0, width=0)),
        "siam_fermion": (SingleImpurityAndersonModel(4.0, -2.0, eps, v, mu=0.0, sign_mode="fermion"),
                         lambda nu, nd: O.siam_operator(nu, nd, 4.0, -2.0, eps, v, 0.0, width=6)),
    }


@pytest.mark.parametrize("name", ["chain5", "ring6", "sq4x2", "siam", "siam_fermion"])
def test_matvec_dense_diag_vs_oracle(gpu, name):
    model, mk = _models()[name]
    ns = model.num_sites
    rng = np.random.default_rng(1)
    for n_up, n_dn in [(ns // 2, ns // 2), (ns // 2 + 1, ns // 2), (1, ns - 1), (0, 2)]:
        op = model.hamilton_operator(n_up=n_up, n_dn=n_dn)
        ref = mk(n_up, n_dn)
        h_ref = ref.csr().toarray()
        if op.size <= 4096:
            assert_allclose(op.toarray(), h_ref, rtol=0, atol=1e-14)
        x = rng.standard_normal((op.size, 3))
        y = op.matmat(x)
        assert_allclose(y, ref.csr() @ x, rtol=1e-13, atol=1e-12)
        assert_allclose(op.diagonal(), ref.diagonal(), rtol=0, atol=1e-14)
        assert abs(op.trace() - ref.diagonal().sum()) < 1e-10


@pytest.mark.parametrize("name", ["chain5", "siam"])
def test_hamiltonian_data_triples_bit_exact(gpu, name):
    """The product's cmpy-style triple generator equals the oracle's restatement."""
    model, mk = _models()[name]
    ns = model.num_sites
    sec = model.get_sector(ns // 2, ns // 2)
    data, idx = model.hamiltonian_data(sec.up_states, sec.dn_states)
    up, dn = O.sector_states(ns, ns // 2), O.sector_states(ns, ns // 2)
    if name == "chain5":
        r, c, v = O.hubbard_triples(up, dn, ns, O.chain_bonds(5), 2.0, 0.3, 1.0, mu=1.0)
    else:
        r, c, v = O.siam_triples(up, dn, 4.0, -2.0, model.eps_bath, model.v, 0.0)
    assert_array_equal(idx[0], r)
    assert_array_equal(idx[1], c)
    assert_array_equal(np.asarray(data), v)
    x = np.random.default_rng(3).standard_normal(sec.size)
    coo = model.coo_operator(sector=sec)
    assert_allclose(coo.matvec(x), O.coo_matvec(sec.size, r, c, v, x), rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("ns", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("periodic", [False, True])
def test_hermitian_all_fillings(gpu, ns, periodic):
    # cmpy/tests/test_models_hubbard.py:28-48
    from cmpy_amd import HubbardModel, chain_neighbors

    for u in (0.0, 2.0):
        m = HubbardModel(ns, chain_neighbors(ns, periodic), inter=u, mu=u / 2, hop=1.0)
        for n_up, n_dn in m.basis.iter_fillings():
            h = m.hamiltonian(n_up, n_dn)
            assert_array_equal(h, h.T)


def test_hop_table_export(gpu):
    model, _ = _models()["sq4x2"]
    op = model.hamilton_operator(n_up=4, n_dn=3)
    cnt, col, val = op.hop_table(1)
    st = O.sector_states(8, 4)
    bonds = O.square_bonds(4, 2)
    for k in range(len(st)):
        exp = []
        for (i, j) in bonds:
            o, t, a = O.hopping_terms(st[k:k + 1], 8, i, j, 1.0)
            if len(o):
                exp.append((int(np.searchsorted(st, int(st[k]) ^ (1 << i) ^ (1 << j))), float(a[0])))
        assert cnt[k] == len(exp)
        assert [(int(c), float(v)) for c, v in zip(col[k, :cnt[k]], val[k, :cnt[k]])] == exp


@pytest.mark.parametrize("ns", [2, 3, 4, 5])
@pytest.mark.parametrize("sigma", [1, 2])
def test_ladder_operators(gpu, ns, sigma):
    from cmpy_amd import Basis, CreationOperator, AnnihilationOperator

    basis = Basis(ns)
    rng = np.random.default_rng(ns)
    for n_up, n_dn in basis.iter_fillings():
        sec = basis.get_sector(n_up, n_dn)
        p1 = basis.upper_sector(n_up, n_dn, sigma)
        if p1 is None:
            continue
        for pos in range(ns):
            for mode in ("reference", "jordan_wigner"):
                cd = CreationOperator(sec, p1, pos, sigma, sign_mode=mode)
                c = AnnihilationOperator(p1, sec, pos, sigma, sign_mode=mode)
                # cmpy/tests/test_operator.py:31-42
                assert_array_equal(c.toarray(), cd.toarray().T.conj())
                x = rng.standard_normal(sec.size)
                assert_array_equal(cd.matvec(x), O.apply_cdag(x, ns, n_up, n_dn, pos, sigma, mode))
                xx = rng.standard_normal(p1.size)
                assert_array_equal(c.matvec(xx), O.apply_c(xx, ns, p1.n_up, p1.n_dn, pos, sigma, mode))


@pytest.mark.parametrize("geom", ["chain12", "torus4x3", "siam10", "torus4x4_1_6"])
def test_large_sector_matvec_both_kernels(gpu, geom, monkeypatch):
    """Sectors large enough for the two-pass sweep; both kernel paths against the oracle."""
    from cmpy_amd import SectorHamiltonOperator

    if geom == "chain12":
        ns, nu, nd, bonds, w = 12, 6, 6, O.chain_bonds(12), 12
        eps, u = np.full(ns, -0.5), np.full(ns, 4.0)
    elif geom == "torus4x3":
        ns, nu, nd, bonds, w = 12, 6, 5, O.square_bonds(4, 3), 12
        eps, u = np.zeros(ns), np.full(ns, 2.0)
    elif geom == "torus4x4_1_6":
        # 8008 dn states: one-row blocks with the bank-ordered half-width tables
        ns, nu, nd, bonds, w = 16, 1, 6, O.square_bonds(4, 4), 16
        eps, u = np.linspace(-1, 1, ns), np.full(ns, 3.0)
    else:
        ns, nu, nd, w = 10, 5, 5, 0
        bonds = [(0, k) for k in range(1, ns)]
        rng = np.random.default_rng(9)
        eps, u = np.append(-1.5, rng.uniform(-3, 3, 9)), np.append(3.0, np.zeros(9))
    hops = np.linspace(0.3, 1.1, len(bonds))
    ref = O.SectorOperator(ns, nu, nd, bonds, hops, eps, u, w)
    x = np.random.default_rng(2).standard_normal(ref.size)
    y_ref = ref.matvec(x)
    scale = np.abs(y_ref).max()
    # fused kernel; two-pass with sign-split tables (uniform hopping only) or
    # general packed tables, column tiles of 16/32/64 columns; the pair-list
    # column passes (32- and 64-column tiles): run first and written into y
    # (the default; quad lists with 16-byte loads, or pair lists), or after the
    # row pass
    base = dict(MODE="1", SPLIT="1", TW="32", HALF="1", PAIR="1", PTW="32", FIRST="1", QUAD="1", HELD="1", OCT="0",
                PRCH="16")
    combos = (dict(MODE="0"), dict(PAIR="0", TW="16"), dict(PAIR="0"), dict(PAIR="0", TW="64"), {}, dict(HELD="0"),
              dict(OCT="1"), dict(OCT="1", PRCH="8"), dict(QUAD="0"), dict(QUAD="0", PTW="64"),
              dict(HALF="0"), dict(FIRST="0"), dict(FIRST="0", PTW="64"), dict(SPLIT="0"),
              dict(SPLIT="0", TW="16"))
    for over in combos:
        case = {**base, **over}
        for k, v in case.items():
            monkeypatch.setenv("CMPY_AMD_HV_" + k, v)
        op = SectorHamiltonOperator(ns, nu, nd, bonds, hops, eps, u, w)
        y = op.matvec(x)
        assert np.abs(y - y_ref).max() <= 1e-13 * scale, case
        # rows padded to ld: padding of x never read (NaN), of y written as 0
        nrow, ncol = op.num_up, op.num_dn
        for ld in (ncol + (-ncol) % 16 or ncol + 16, ncol + 3):
            xp = np.full((nrow, ld), np.nan)
            xp[:, :ncol] = x.reshape(nrow, ncol)
            yp = torch.full((nrow, ld), float("nan"), dtype=torch.float64, device="cuda")
            op.matvec_device_ld(torch.from_numpy(xp).cuda(), ld, out=yp)
            yp = yp.cpu().numpy()
            assert np.all(yp[:, ncol:] == 0.0), (case, ld)
            assert np.abs(yp[:, :ncol].ravel() - y).max() <= 1e-13 * scale, (case, ld)
        hops_u = np.full(len(bonds), 0.7)
        op_u = SectorHamiltonOperator(ns, nu, nd, bonds, hops_u, eps, u, w)
        ref_u = O.SectorOperator(ns, nu, nd, bonds, hops_u, eps, u, w)
        y_u = ref_u.matvec(x)
        assert np.abs(op_u.matvec(x) - y_u).max() <= 1e-13 * np.abs(y_u).max(), case


def test_hv_4x4_matches_cpu_twin(gpu):
    """Full 4x4 (8,8) sector, D = 165,636,900: GPU H.v against the oracle's C twin."""
    from oracle import twin
    from cmpy_amd import SectorHamiltonOperator

    ns, bonds = 16, O.square_bonds(4, 4)
    args = (ns, 8, 8, bonds, [1.0] * len(bonds), np.zeros(ns), np.full(ns, 4.0), ns)
    op = SectorHamiltonOperator(*args)
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    x = torch.rand(op.size, dtype=torch.float64, device="cuda", generator=g) - 0.5
    y = op.matvec_device(x)
    tw = twin.TwinOperator(*args)
    y_ref = tw.matvec(x.cpu().numpy())
    diff = float(np.abs(y.cpu().numpy() - y_ref).max())
    assert diff <= 1e-13 * float(np.abs(y_ref).max()), diff
    # Hermiticity: <a, H b> = <H a, b>
    b = torch.rand(op.size, dtype=torch.float64, device="cuda", generator=g) - 0.5
    hb = op.matvec_device(b)
    lhs, rhs = float(torch.dot(x, hb)), float(torch.dot(y, b))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_padded_rows_default_stride(gpu, monkeypatch):
    """The library's own vector stride: 128-byte rows where the two-pass sweep
    runs (num_dn not a multiple of 16), contiguous otherwise.  The 12-site
    chain's (6,6) sector (D = 853,776) takes the fused single pass by default
    (below 2^21 states) and the two passes with CMPY_AMD_HV_MODE=1."""
    from cmpy_amd import SectorHamiltonOperator

    ns, bonds = 12, O.chain_bonds(12)
    args = (ns, 6, 6, bonds, [1.0] * 11, np.zeros(ns), np.full(ns, 4.0), ns)
    fused = SectorHamiltonOperator(*args)
    assert fused.num_dn == 924 and fused.vector_ld == 924
    small = SectorHamiltonOperator(6, 3, 3, O.chain_bonds(6), [1.0] * 5, np.zeros(6), np.full(6, 4.0), 6)
    assert small.vector_ld == small.num_dn
    monkeypatch.setenv("CMPY_AMD_HV_MODE", "1")
    big = SectorHamiltonOperator(*args)
    assert big.num_dn == 924 and big.vector_ld == 928
    from cmpy_amd import _lib as L

    x = torch.zeros((big.num_up, 928), dtype=torch.float64, device="cuda")
    with pytest.raises(L.EDInvalidArgument, match="ld="):
        L.call("ed_op_matvec_ld", big.handle, L.dptr(x), L.dptr(x.clone()), 900, 1, L.stream_ptr())


def test_invalid_arguments_raise(gpu):
    from cmpy_amd import SectorHamiltonOperator, _lib as L

    with pytest.raises(L.EDInvalidArgument, match="bond"):
        SectorHamiltonOperator(4, 2, 2, [(2, 1)], [1.0], np.zeros(4), np.zeros(4), 4)
    with pytest.raises(L.EDInvalidArgument, match="bond"):
        SectorHamiltonOperator(4, 2, 2, [(0, 4)], [1.0], np.zeros(4), np.zeros(4), 4)
    with pytest.raises(L.EDInvalidArgument, match="sector"):
        SectorHamiltonOperator(4, 5, 2, [(0, 1)], [1.0], np.zeros(4), np.zeros(4), 4)
    with pytest.raises(L.EDInvalidArgument, match="finite"):
        SectorHamiltonOperator(4, 2, 2, [(0, 1)], [float("nan")], np.zeros(4), np.zeros(4), 4)
    op = SectorHamiltonOperator(4, 2, 2, [(0, 1)], [1.0], np.zeros(4), np.zeros(4), 4)
    x = torch.zeros(op.size, dtype=torch.float64, device="cuda")
    with pytest.raises(L.EDInvalidArgument, match="overlap"):
        L.call("ed_op_matvec", op.handle, L.dptr(x), L.dptr(x), 1, L.stream_ptr())
    y = torch.zeros(36, dtype=torch.float64, device="cuda")
    with pytest.raises(L.EDInvalidArgument, match="does not exist"):
        L.call("ed_apply_cdag", 4, 4, 2, 0, 1, 0, L.dptr(x), L.dptr(y), 1, L.stream_ptr())
    with pytest.raises(L.EDInvalidArgument):
        L.call("ed_hop_terms", 4, 2, 4, 1, 1, 1.0, L.dptr(x), L.dptr(y), L.stream_ptr())
